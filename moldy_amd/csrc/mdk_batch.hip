// mdk_batch.hip -- many independent periodic Fe cells in one persistent launch.
//
// Energies, forces and the 3x3 stress of thousands of cells of uneven atom
// counts (the force() evaluation of src/dynamics_m.F90:135-401 applied to each
// cell, with the Fe-Fe potential of src/fetransmetal.F90).  One persistent
// kernel: every workgroup takes whole cells from an atomic work queue (largest
// cells first) and runs, per cell, the neighbour-list build, the density pass
// and the force pass separated by workgroup barriers, then reduces the cell's
// energy and virial itself -- no cross-workgroup reduction.  The launch (queue
// reset + kernel) is captured once into a hipGraph and replayed.  Block size
// and workgroups per CU are chosen by mdk_batch_tune.
//
// Pair arithmetic is FP32 (or FP64), with FP64 accumulation of densities,
// forces, energies and virials; positions are FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/moldy_hip.h"
#include "mdk_hostpot.h"
#include "mdk_internal.h"

namespace {

thread_local char b_err[512] = "";
int berr(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b_err, sizeof(b_err), fmt, ap);
  va_end(ap);
  return code;
}
#define BCK(call)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return berr(MDK_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                  __FILE__, __LINE__);                                                   \
  } while (0)

#define NB_IDX_MASK 0x07FFFFFF
#define NB_CODE_SHIFT 27

struct BatchArgs {
  int ncell;
  int maxnb;
  const int *order;      // cells, largest first
  const int *cell_off;   // ncell+1 atom offsets
  const double *box;     // 9 per cell, column-major
  const int *nsub;       // 3 per cell (1, or >= 3)
  const int *sub_off;    // ncell+1 offsets into sub_start
  const int *sub_start;  // per cell: nsub+1 local atom offsets
  const int *subc;       // per atom: local sub-cell index
  const double *s;       // 3 per atom: fractional, wrapped into [0,1), sorted by sub-cell
  const int *perm;       // sorted atom -> caller's atom index
  double4 *pos4;         // Cartesian position, F'
  int *nbr;              // per cell region: nbr[a0*maxnb + k*nc + il]
  int *numn;
  double *rho;
  double *out_e;         // per cell energy
  double *out_w;         // per cell virial W (6: xx yy zz xy xz yz), W = sum_pairs r r^T fp
  double *out_f;         // 3 per caller atom, Cartesian forces
  int *counter;
  int *flags;            // [0] max neighbour count seen when over capacity
  double rc2;
};

__device__ __forceinline__ double bwave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void r_inv(double r2, double &r, double &ri) {
  double y = __builtin_amdgcn_rsq(r2);
  const double h = 0.5 * r2;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  ri = y;
  r = r2 * y;
}
__device__ __forceinline__ void r_inv(float r2, float &r, float &ri) {
  float y = __builtin_amdgcn_rsqf(r2);
  y = y * (1.5f - 0.5f * r2 * y * y);
  ri = y;
  r = r2 * y;
}

__global__ void k_batch_reset(int *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

template <typename T, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_batch_eval(BatchArgs A, PotDev P) {
  __shared__ double S[32][3];
  __shared__ double red[BLOCK / 64][7];
  __shared__ int sh_cell;
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) sh_cell = atomicAdd(A.counter, 1);
    __syncthreads();
    const int q = sh_cell;
    __syncthreads();
    if (q >= A.ncell) break;  // uniform over the workgroup: every wave exits
    const int c = A.order[q];
    const int a0 = A.cell_off[c], nc = A.cell_off[c + 1] - a0;
    const double *B = A.box + 9 * (size_t)c;
    double b[9];
#pragma unroll
    for (int k = 0; k < 9; k++) b[k] = B[k];
    const int nx = A.nsub[3 * c], ny = A.nsub[3 * c + 1], nz = A.nsub[3 * c + 2];
    const int *sst = A.sub_start + A.sub_off[c];
    int *nbr = A.nbr + (size_t)a0 * A.maxnb;
    // image shift table of this cell; codes 27..31 never occur here
    for (int k = t; k < 32; k += BLOCK) {
      if (k >= 27) {
        S[k][0] = 1.0e6; S[k][1] = 0.0; S[k][2] = 0.0;
      } else {
        const double sx = (double)(k % 3 - 1), sy = (double)((k / 3) % 3 - 1), sz = (double)(k / 9 - 1);
        S[k][0] = b[0] * sx + b[3] * sy + b[6] * sz;
        S[k][1] = b[1] * sx + b[4] * sy + b[7] * sz;
        S[k][2] = b[2] * sx + b[5] * sy + b[8] * sz;
      }
    }
    // phase 1: Cartesian positions
    for (int il = t; il < nc; il += BLOCK) {
      const size_t a = (size_t)a0 + il;
      const double sx = A.s[3 * a], sy = A.s[3 * a + 1], sz = A.s[3 * a + 2];
      A.pos4[a] = make_double4(b[0] * sx + b[3] * sy + b[6] * sz, b[1] * sx + b[4] * sy + b[7] * sz,
                               b[2] * sx + b[5] * sy + b[8] * sz, 0.0);
    }
    __syncthreads();
    // phase 2: full neighbour list within the cell and its periodic images
    const int rx_ = nx >= 3 ? 1 : 0, ry_ = ny >= 3 ? 1 : 0, rz_ = nz >= 3 ? 1 : 0;
    for (int il = t; il < nc; il += BLOCK) {
      const size_t a = (size_t)a0 + il;
      const int sc = A.subc[a];
      const int cx = sc % nx, cy = (sc / nx) % ny, cz = sc / (nx * ny);
      const double4 pi = A.pos4[a];
      const double sxi = A.s[3 * a], syi = A.s[3 * a + 1], szi = A.s[3 * a + 2];
      int cnt = 0;
      for (int dz = -rz_; dz <= rz_; dz++) {
        int kz = cz + dz, wz = 0;
        if (kz < 0) { kz += nz; wz = 1; } else if (kz >= nz) { kz -= nz; wz = -1; }
        for (int dy = -ry_; dy <= ry_; dy++) {
          int ky = cy + dy, wy = 0;
          if (ky < 0) { ky += ny; wy = 1; } else if (ky >= ny) { ky -= ny; wy = -1; }
          for (int dx = -rx_; dx <= rx_; dx++) {
            int kx = cx + dx, wx = 0;
            if (kx < 0) { kx += nx; wx = 1; } else if (kx >= nx) { kx -= nx; wx = -1; }
            const int cj = kx + nx * (ky + ny * kz);
            for (int jl = sst[cj]; jl < sst[cj + 1]; jl++) {
              if (jl == il) continue;
              const size_t aj = (size_t)a0 + jl;
              int tx = wx, ty = wy, tz = wz;
              if (!rx_) tx = -(int)round(sxi - A.s[3 * aj]);
              if (!ry_) ty = -(int)round(syi - A.s[3 * aj + 1]);
              if (!rz_) tz = -(int)round(szi - A.s[3 * aj + 2]);
              const int code = (tx + 1) + 3 * (ty + 1) + 9 * (tz + 1);
              const double4 pj = A.pos4[aj];
              const double ex = pi.x - pj.x + S[code][0], ey = pi.y - pj.y + S[code][1],
                           ez = pi.z - pj.z + S[code][2];
              if (ex * ex + ey * ey + ez * ez < A.rc2) {
                if (cnt < A.maxnb) nbr[(size_t)cnt * nc + il] = jl | (code << NB_CODE_SHIFT);
                cnt++;
              }
            }
          }
        }
      }
      A.numn[a] = min(cnt, A.maxnb);
      if (cnt > A.maxnb) atomicMax(&A.flags[0], cnt);
    }
    __syncthreads();
    // phase 3: densities and embedding (src/dynamics_m.F90:497-579, src/potential_m.F90:148-169)
    double e_emb = 0.0;
    for (int il = t; il < nc; il += BLOCK) {
      const size_t a = (size_t)a0 + il;
      const double4 pi = A.pos4[a];
      const int nn = A.numn[a];
      double acc = 0.0;
      for (int k = 0; k < nn; k++) {
        const unsigned e = (unsigned)nbr[(size_t)k * nc + il];
        const double4 pj = A.pos4[a0 + (e & NB_IDX_MASK)];
        const int code = (int)(e >> NB_CODE_SHIFT);
        const T dx = (T)(pi.x - pj.x + S[code][0]), dy = (T)(pi.y - pj.y + S[code][1]),
                dz = (T)(pi.z - pj.z + S[code][2]);
        T r, ri;
        r_inv(dx * dx + dy * dy + dz * dz, r, ri);
        acc += (double)mdk_phi<T>(P, r);
      }
      A.rho[a] = acc;
      double f = 0.0, df = 0.0;
      if (acc > 1e-40) {
        mdk_embed_f(P, acc, f, df);
        e_emb += f;
      }
      A.pos4[a].w = df;
    }
    __syncthreads();
    // phase 4: forces and virial (src/dynamics_m.F90:235-346)
    double e_pair = 0.0, w[6] = {0, 0, 0, 0, 0, 0};
    for (int il = t; il < nc; il += BLOCK) {
      const size_t a = (size_t)a0 + il;
      const double4 pi = A.pos4[a];
      const T dfi = (T)pi.w;
      const int nn = A.numn[a];
      double fx = 0.0, fy = 0.0, fz = 0.0;
      for (int k = 0; k < nn; k++) {
        const unsigned e = (unsigned)nbr[(size_t)k * nc + il];
        const double4 pj = A.pos4[a0 + (e & NB_IDX_MASK)];
        const int code = (int)(e >> NB_CODE_SHIFT);
        const T dx = (T)(pi.x - pj.x + S[code][0]), dy = (T)(pi.y - pj.y + S[code][1]),
                dz = (T)(pi.z - pj.z + S[code][2]);
        T r, rr;
        r_inv(dx * dx + dy * dy + dz * dz, r, rr);
        T vv, dv;
        mdk_vee<T>(P, r, vv, dv);
        const T dph = mdk_dphi<T>(P, r);
        const T fp = -dv * rr + (-T(1) * (dph * dfi + dph * (T)pj.w) * rr);
        const double fpd = (double)fp, ddx = (double)dx, ddy = (double)dy, ddz = (double)dz;
        e_pair += 0.5 * (double)vv;
        fx += ddx * fpd;
        fy += ddy * fpd;
        fz += ddz * fpd;
        const double h = 0.5 * fpd;
        w[0] += ddx * ddx * h; w[1] += ddy * ddy * h; w[2] += ddz * ddz * h;
        w[3] += ddx * ddy * h; w[4] += ddx * ddz * h; w[5] += ddy * ddz * h;
      }
      const size_t u = (size_t)A.perm[a];
      A.out_f[3 * u] = fx;
      A.out_f[3 * u + 1] = fy;
      A.out_f[3 * u + 2] = fz;
    }
    // per-cell reduction (fixed order: lanes, then waves)
    double v[7] = {e_emb + e_pair, w[0], w[1], w[2], w[3], w[4], w[5]};
#pragma unroll
    for (int k = 0; k < 7; k++) v[k] = bwave_sum(v[k]);
    if ((t & 63) == 0) {
#pragma unroll
      for (int k = 0; k < 7; k++) red[t >> 6][k] = v[k];
    }
    __syncthreads();
    if (t < 7) {
      double sum = red[0][t];
      for (int q2 = 1; q2 < BLOCK / 64; q2++) sum += red[q2][t];
      if (t == 0) A.out_e[c] = sum;
      else A.out_w[6 * (size_t)c + t - 1] = sum;
    }
    __syncthreads();
  }
}

}  // namespace

struct mdk_batch {
  int ncell, natom, precision, maxnb;
  hipStream_t stream;
  int device, ncu;
  int *order, *cell_off, *nsub, *sub_off, *sub_start, *subc, *perm, *numn, *nbr, *counter, *flags;
  double *box, *s, *rho, *out_e, *out_w, *out_f;
  double4 *pos4;
  std::vector<int> h_cell_off;
  std::vector<double> h_box;
  PotDev pot;
  hostpot::Knots knots;
  bool pot_set;
  int block, wg_per_cu;
  bool use_graph;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  bool graph_valid;
  double rc;
  double min_width;  // narrowest perpendicular cell width: the minimum image needs 2 rc below it
};

namespace {

template <typename T>
void launch_eval(int block, int grid, hipStream_t st, const BatchArgs &A, const PotDev &P) {
  switch (block) {
    case 64: k_batch_eval<T, 64><<<grid, 64, 0, st>>>(A, P); break;
    case 128: k_batch_eval<T, 128><<<grid, 128, 0, st>>>(A, P); break;
    case 512: k_batch_eval<T, 512><<<grid, 512, 0, st>>>(A, P); break;
    default: k_batch_eval<T, 256><<<grid, 256, 0, st>>>(A, P); break;
  }
}

BatchArgs args_of(mdk_batch *b) {
  BatchArgs A;
  A.ncell = b->ncell; A.maxnb = b->maxnb; A.order = b->order; A.cell_off = b->cell_off;
  A.box = b->box; A.nsub = b->nsub; A.sub_off = b->sub_off; A.sub_start = b->sub_start;
  A.subc = b->subc; A.s = b->s; A.perm = b->perm; A.pos4 = b->pos4; A.nbr = b->nbr;
  A.numn = b->numn; A.rho = b->rho; A.out_e = b->out_e; A.out_w = b->out_w; A.out_f = b->out_f;
  A.counter = b->counter; A.flags = b->flags; A.rc2 = b->rc * b->rc;
  return A;
}

int enqueue(mdk_batch *b) {
  k_batch_reset<<<1, 64, 0, b->stream>>>(b->counter);
  const int grid = std::max(1, std::min(b->ncell, b->ncu * b->wg_per_cu));
  if (b->precision == MDK_PREC_FP64)
    launch_eval<double>(b->block, grid, b->stream, args_of(b), b->pot);
  else
    launch_eval<float>(b->block, grid, b->stream, args_of(b), b->pot);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return berr(MDK_ERR_HIP, "batch launch: %s", hipGetErrorString(e));
  return MDK_OK;
}

void drop_graph(mdk_batch *b) {
  if (b->graph_valid) {
    hipGraphExecDestroy(b->gexec);
    hipGraphDestroy(b->graph);
    b->graph_valid = false;
  }
}

int capture(mdk_batch *b) {
  drop_graph(b);
  BCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != MDK_OK) return rc;
  if (e != hipSuccess) return berr(MDK_ERR_HIP, "capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return berr(MDK_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  b->graph = g;
  b->graph_valid = true;
  return MDK_OK;
}

}  // namespace

extern "C" const char *mdk_batch_last_error(void) { return b_err; }

extern "C" int mdk_batch_destroy(void *h) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b) return MDK_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graph(b);
  void *p[] = {b->order, b->cell_off, b->nsub, b->sub_off, b->sub_start, b->subc, b->perm,
               b->numn, b->nbr, b->counter, b->flags, b->box, b->s, b->rho, b->out_e, b->out_w,
               b->out_f, b->pos4};
  for (void *q : p)
    if (q) hipFree(q);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return MDK_OK;
}

// Cells: natoms[c] atoms each, boxes 9 doubles per cell (column-major, columns
// are cell vectors), s = 3 fractional coordinates per atom, cells concatenated.
extern "C" int mdk_batch_create(void **handle, int ncell, const int *natoms, const double *boxes,
                                const double *s, int precision) {
  if (!handle || ncell < 1 || !natoms || !boxes || !s) return berr(MDK_ERR_BAD_ARG, "bad batch arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return berr(MDK_ERR_NO_DEVICE, "no HIP device");
  mdk_batch *b = new mdk_batch();
  b->ncell = ncell;
  b->precision = precision;
  b->maxnb = 48;
  b->block = 256;
  b->wg_per_cu = 4;
  b->use_graph = true;
  hipGetDevice(&b->device);
  hipDeviceProp_t prop;
  hipGetDeviceProperties(&prop, b->device);
  b->ncu = prop.multiProcessorCount;
  // host preprocessing: wrap, sub-cell grid, sort atoms by sub-cell per cell
  b->h_cell_off.assign(ncell + 1, 0);
  for (int c = 0; c < ncell; c++) {
    if (natoms[c] < 1) {
      delete b;
      return berr(MDK_ERR_BAD_ARG, "cell %d has no atoms", c);
    }
    b->h_cell_off[c + 1] = b->h_cell_off[c] + natoms[c];
  }
  const int natom = b->h_cell_off[ncell];
  b->natom = natom;
  b->h_box.assign(boxes, boxes + 9 * (size_t)ncell);
  b->rc = 3.72645;  // replaced by the potential's outermost knot in set_potential
  std::vector<double> hs(3 * (size_t)natom);
  std::vector<int> hperm(natom), hsubc(natom), hnsub(3 * ncell), hsuboff(ncell + 1, 0), hsubstart;
  // sub-cell widths use a conservative cutoff so a later potential with knots
  // up to 4.5 A still fits (the kernel tests r < rc exactly)
  const double rgrid = 4.5;
  b->min_width = 1e300;
  for (int c = 0; c < ncell; c++) {
    const double *B = boxes + 9 * (size_t)c;
    double vol = fabs(TT(B, 1, 1) * (TT(B, 2, 2) * TT(B, 3, 3) - TT(B, 2, 3) * TT(B, 3, 2)) +
                      TT(B, 2, 1) * (TT(B, 3, 2) * TT(B, 1, 3) - TT(B, 1, 2) * TT(B, 3, 3)) +
                      TT(B, 3, 1) * (TT(B, 1, 2) * TT(B, 2, 3) - TT(B, 2, 2) * TT(B, 1, 3)));
    int g[3];
    for (int d = 0; d < 3; d++) {
      const double *u = B + 3 * ((d + 1) % 3), *w = B + 3 * ((d + 2) % 3);
      double cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
      double width = vol / sqrt(cx * cx + cy * cy + cz * cz);
      b->min_width = std::min(b->min_width, width);
      int nc_ = (int)floor(width / rgrid);
      g[d] = nc_ < 3 ? 1 : std::min(nc_, 64);
    }
    hnsub[3 * c] = g[0]; hnsub[3 * c + 1] = g[1]; hnsub[3 * c + 2] = g[2];
    const int nsc = g[0] * g[1] * g[2];
    hsuboff[c + 1] = hsuboff[c] + nsc + 1;
    const int a0 = b->h_cell_off[c], nc = natoms[c];
    std::vector<int> key(nc), idx(nc);
    std::vector<double> w3(3 * (size_t)nc);
    for (int i = 0; i < nc; i++) {
      for (int d = 0; d < 3; d++) {
        double v = s[3 * ((size_t)a0 + i) + d];
        v = v - floor(v);
        if (v >= 1.0) v = 0.0;
        w3[3 * i + d] = v;
      }
      int cx = std::min((int)(w3[3 * i] * g[0]), g[0] - 1);
      int cy = std::min((int)(w3[3 * i + 1] * g[1]), g[1] - 1);
      int cz = std::min((int)(w3[3 * i + 2] * g[2]), g[2] - 1);
      key[i] = cx + g[0] * (cy + g[1] * cz);
      idx[i] = i;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key[x] < key[y]; });
    std::vector<int> start(nsc + 1, 0);
    for (int i = 0; i < nc; i++) start[key[i] + 1]++;
    for (int k = 0; k < nsc; k++) start[k + 1] += start[k];
    hsubstart.insert(hsubstart.end(), start.begin(), start.end());
    for (int k = 0; k < nc; k++) {
      const int i = idx[k];
      for (int d = 0; d < 3; d++) hs[3 * ((size_t)a0 + k) + d] = w3[3 * i + d];
      hperm[a0 + k] = a0 + i;
      hsubc[a0 + k] = key[i];
    }
  }
  std::vector<int> horder(ncell);
  std::iota(horder.begin(), horder.end(), 0);
  std::stable_sort(horder.begin(), horder.end(), [&](int x, int y) { return natoms[x] > natoms[y]; });
  int rc = MDK_OK;
#define BALLOC(p, bytes)                                                                \
  do {                                                                                  \
    if (hipMalloc((void **)&(p), (bytes)) != hipSuccess) {                              \
      rc = berr(MDK_ERR_HIP, "hipMalloc %zu bytes failed", (size_t)(bytes));            \
      goto fail;                                                                        \
    }                                                                                   \
  } while (0)
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    rc = berr(MDK_ERR_HIP, "stream create failed");
    goto fail;
  }
  BALLOC(b->order, sizeof(int) * ncell);
  BALLOC(b->cell_off, sizeof(int) * (ncell + 1));
  BALLOC(b->box, sizeof(double) * 9 * (size_t)ncell);
  BALLOC(b->nsub, sizeof(int) * 3 * (size_t)ncell);
  BALLOC(b->sub_off, sizeof(int) * (ncell + 1));
  BALLOC(b->sub_start, sizeof(int) * hsubstart.size());
  BALLOC(b->subc, sizeof(int) * natom);
  BALLOC(b->perm, sizeof(int) * natom);
  BALLOC(b->s, sizeof(double) * 3 * (size_t)natom);
  BALLOC(b->pos4, sizeof(double4) * natom);
  BALLOC(b->nbr, sizeof(int) * (size_t)natom * b->maxnb);
  BALLOC(b->numn, sizeof(int) * natom);
  BALLOC(b->rho, sizeof(double) * natom);
  BALLOC(b->out_e, sizeof(double) * ncell);
  BALLOC(b->out_w, sizeof(double) * 6 * (size_t)ncell);
  BALLOC(b->out_f, sizeof(double) * 3 * (size_t)natom);
  BALLOC(b->counter, sizeof(int));
  BALLOC(b->flags, sizeof(int) * 4);
  hipMemcpy(b->order, horder.data(), sizeof(int) * ncell, hipMemcpyHostToDevice);
  hipMemcpy(b->cell_off, b->h_cell_off.data(), sizeof(int) * (ncell + 1), hipMemcpyHostToDevice);
  hipMemcpy(b->box, boxes, sizeof(double) * 9 * (size_t)ncell, hipMemcpyHostToDevice);
  hipMemcpy(b->nsub, hnsub.data(), sizeof(int) * 3 * (size_t)ncell, hipMemcpyHostToDevice);
  hipMemcpy(b->sub_off, hsuboff.data(), sizeof(int) * (ncell + 1), hipMemcpyHostToDevice);
  hipMemcpy(b->sub_start, hsubstart.data(), sizeof(int) * hsubstart.size(), hipMemcpyHostToDevice);
  hipMemcpy(b->subc, hsubc.data(), sizeof(int) * natom, hipMemcpyHostToDevice);
  hipMemcpy(b->perm, hperm.data(), sizeof(int) * natom, hipMemcpyHostToDevice);
  if (hipMemcpy(b->s, hs.data(), sizeof(double) * 3 * (size_t)natom, hipMemcpyHostToDevice) != hipSuccess) {
    rc = berr(MDK_ERR_HIP, "upload failed");
    goto fail;
  }
  hipMemset(b->flags, 0, sizeof(int) * 4);
  *handle = b;
  return MDK_OK;
fail:
  mdk_batch_destroy(b);
  return rc;
#undef BALLOC
}

extern "C" int mdk_batch_set_potential(void *h, int nv, const double *av, const double *rv, int np,
                                       const double *ap, const double *rp) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || nv < 1 || nv > MDK_MAXV || np < 1 || np > MDK_MAXP) return berr(MDK_ERR_BAD_ARG, "bad potential");
  hostpot::Knots &k = b->knots;
  memset(&k, 0, sizeof(k));
  k.nv = nv;
  k.np = np;
  for (int i = 0; i < nv; i++) { k.av[i] = av[i]; k.rv[i] = rv[i]; }
  for (int i = 0; i < np; i++) { k.ap[i] = ap[i]; k.rp[i] = rp[i]; }
  hostpot::set_up(k);
  const double rint = hostpot::fill_potdev(b->pot, k);
  if (rint > 4.5) return berr(MDK_ERR_BAD_ARG, "knots beyond 4.5 A are not supported by the batch grid");
  if (2.0 * rint >= b->min_width)
    return berr(MDK_ERR_BAD_ARG, "a cell is narrower (%.3f A) than twice the interaction range", b->min_width);
  b->rc = rint;
  b->pot_set = true;
  drop_graph(b);  // the potential is a kernel argument of the captured launch
  return MDK_OK;
}

extern "C" int mdk_batch_config(void *h, int block, int wg_per_cu, int use_graph) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b) return berr(MDK_ERR_BAD_ARG, "null handle");
  if (block != 64 && block != 128 && block != 256 && block != 512) return berr(MDK_ERR_BAD_ARG, "block must be 64..512");
  if (wg_per_cu < 1 || wg_per_cu > 32) return berr(MDK_ERR_BAD_ARG, "wg_per_cu out of range");
  b->block = block;
  b->wg_per_cu = wg_per_cu;
  b->use_graph = use_graph != 0;
  drop_graph(b);
  return MDK_OK;
}

extern "C" int mdk_batch_evaluate(void *h) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || !b->pot_set) return berr(MDK_ERR_STATE, "batch potential not set");
  if (b->use_graph) {
    if (!b->graph_valid) {
      int rc = capture(b);
      if (rc != MDK_OK) return rc;
    }
    BCK(hipGraphLaunch(b->gexec, b->stream));
  } else {
    int rc = enqueue(b);
    if (rc != MDK_OK) return rc;
  }
  return MDK_OK;
}

extern "C" int mdk_batch_synchronize(void *h) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b) return berr(MDK_ERR_BAD_ARG, "null handle");
  BCK(hipStreamSynchronize(b->stream));
  int fl[4];
  BCK(hipMemcpy(fl, b->flags, sizeof(fl), hipMemcpyDeviceToHost));
  if (fl[0]) return berr(MDK_ERR_NBR_OVERFLOW, "batch neighbour capacity %d exceeded (%d)", b->maxnb, fl[0]);
  return MDK_OK;
}

// energy[ncell] (eV), virial[6*ncell] = sum over pairs of r r^T fp (xx yy zz xy xz yz;
// stress = -virial/volume), forces[3*natom] Cartesian (eV/A), rho[natom]; any may be NULL.
extern "C" int mdk_batch_results(void *h, double *energy, double *virial, double *forces, double *rho) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b) return berr(MDK_ERR_BAD_ARG, "null handle");
  int rc = mdk_batch_synchronize(h);
  if (rc != MDK_OK) return rc;
  if (energy) BCK(hipMemcpy(energy, b->out_e, sizeof(double) * b->ncell, hipMemcpyDeviceToHost));
  if (virial) BCK(hipMemcpy(virial, b->out_w, sizeof(double) * 6 * (size_t)b->ncell, hipMemcpyDeviceToHost));
  if (forces) BCK(hipMemcpy(forces, b->out_f, sizeof(double) * 3 * (size_t)b->natom, hipMemcpyDeviceToHost));
  if (rho) {
    std::vector<double> r(b->natom);
    std::vector<int> perm(b->natom);
    BCK(hipMemcpy(r.data(), b->rho, sizeof(double) * b->natom, hipMemcpyDeviceToHost));
    BCK(hipMemcpy(perm.data(), b->perm, sizeof(int) * b->natom, hipMemcpyDeviceToHost));
    for (int a = 0; a < b->natom; a++) rho[perm[a]] = r[a];
  }
  return MDK_OK;
}

// Autotuner: times every (block, workgroups per CU) candidate with HIP events
// (best of `reps` replays) and keeps the fastest; returns the choice and its time.
extern "C" int mdk_batch_tune(void *h, int reps, int *best_block, int *best_wgpc, double *best_ms) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || !b->pot_set) return berr(MDK_ERR_STATE, "batch potential not set");
  const int blocks[] = {128, 256, 512};
  const int wgs[] = {1, 2, 4, 8};
  hipEvent_t e0, e1;
  BCK(hipEventCreate(&e0));
  BCK(hipEventCreate(&e1));
  double best = 1e30;
  int bb = b->block, bw = b->wg_per_cu;
  const bool g = b->use_graph;
  for (int bl : blocks)
    for (int wg : wgs) {
      if (bl * wg > 2048) continue;
      b->block = bl;
      b->wg_per_cu = wg;
      drop_graph(b);
      b->use_graph = g;
      int rc = mdk_batch_evaluate(h);  // warm-up (and capture)
      if (rc != MDK_OK) return rc;
      double t = 1e30;
      for (int r = 0; r < std::max(1, reps); r++) {
        BCK(hipEventRecord(e0, b->stream));
        rc = mdk_batch_evaluate(h);
        if (rc != MDK_OK) return rc;
        BCK(hipEventRecord(e1, b->stream));
        BCK(hipEventSynchronize(e1));
        float ms = 0.f;
        BCK(hipEventElapsedTime(&ms, e0, e1));
        t = std::min(t, (double)ms);
      }
      if (t < best) {
        best = t;
        bb = bl;
        bw = wg;
      }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b->block = bb;
  b->wg_per_cu = bw;
  drop_graph(b);
  if (best_block) *best_block = bb;
  if (best_wgpc) *best_wgpc = bw;
  if (best_ms) *best_ms = best;
  return mdk_batch_synchronize(h);
}

// Average time of `reps` replays (HIP events on the batch stream), ms.
extern "C" int mdk_batch_time(void *h, int reps, double *ms_avg) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || reps < 1) return berr(MDK_ERR_BAD_ARG, "bad arguments");
  hipEvent_t e0, e1;
  BCK(hipEventCreate(&e0));
  BCK(hipEventCreate(&e1));
  int rc = mdk_batch_evaluate(h);
  if (rc != MDK_OK) return rc;
  BCK(hipEventRecord(e0, b->stream));
  for (int r = 0; r < reps; r++) {
    rc = mdk_batch_evaluate(h);
    if (rc != MDK_OK) return rc;
  }
  BCK(hipEventRecord(e1, b->stream));
  BCK(hipEventSynchronize(e1));
  float ms = 0.f;
  BCK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ms_avg) *ms_avg = ms / reps;
  return mdk_batch_synchronize(h);
}

extern "C" int mdk_batch_info(void *h, int *natom, int *block, int *wg_per_cu, long *pairs) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b) return berr(MDK_ERR_BAD_ARG, "null handle");
  if (natom) *natom = b->natom;
  if (block) *block = b->block;
  if (wg_per_cu) *wg_per_cu = b->wg_per_cu;
  if (pairs) {
    std::vector<int> nn(b->natom);
    BCK(hipMemcpy(nn.data(), b->numn, sizeof(int) * b->natom, hipMemcpyDeviceToHost));
    long s = 0;
    for (int v : nn) s += v;
    *pairs = s;
  }
  return MDK_OK;
}
