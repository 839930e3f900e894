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
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/moldy_hip.h"
#include "mdk_hostpot.h"
#include "mdk_internal.h"
#include "mdk_wave.h"

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
  int ncell, maxnb, ngrad;
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
  float4 *p4f;           // FP32 form: position from its sub-cell's centre (FP32), F'
  int *nbr;              // per cell region: nbr[a0*maxnb + k*nc + il]
  int *numn;
  double *rho;
  double *out_e;         // per cell energy
  double *out_w;         // per cell virial W (6: xx yy zz xy xz yz), W = sum_pairs r r^T fp
  double *out_f;         // 3 per caller atom, Cartesian forces
  int *counter;
  int *flags;            // [0] max neighbour count seen when over capacity
  double rc2;
  float rc2f;            // FP32 form's list radius squared, 1e-4 A past the outermost knot
  const float *fetab;    // FP32 form: V, V', phi' interval table of the knots (fe_ok)
  int fe_ok;
  // adjoint
  const double *e_in;    // per cell: dL/dE_c (vector-Jacobian product) or target E_c (loss)
  const double *f_in;    // 3 per caller atom: dL/dF_i (product) or target F_i (loss)
  double *gvec;          // 3 per sorted atom: the force seed dL/dF_i
  double *out_g;         // NGRAD_SLOTS per cell: gradient partials
  double *out_loss;      // per cell loss
  double f_weight;       // loss weight of sum_i |F_i - F_t,i|^2 (0.1 / total atoms)
};

// gradient slots per cell: a_k, r_k (MDK_MAXV each), A_k, R_k (MDK_MAXP each)
#define NGRAD_SLOTS (2 * MDK_MAXV + 2 * MDK_MAXP)


// Fixed-order sum over the workgroup (lanes, then waves); every thread gets the
// total.  red must hold BLOCK/64 doubles; the call contains two barriers.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int q = 1; q < BLOCK / 64; q++) s += red[q];
  __syncthreads();
  return s;
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

// One cell as seen by its workgroup.
struct CellView {
  int c, a0, nc, nx, ny, nz;
  const int *sst;
  int *nbr;
  double b[9];
};

// Takes the next cell from the queue (largest first); returns false when the
// queue is exhausted (uniform over the workgroup).
__device__ __forceinline__ bool next_cell(const BatchArgs &A, int *sh_cell, CellView &V) {
  if (threadIdx.x == 0) *sh_cell = atomicAdd(A.counter, 1);
  __syncthreads();
  const int q = *sh_cell;
  __syncthreads();
  if (q >= A.ncell) return false;
  V.c = A.order[q];
  V.a0 = A.cell_off[V.c];
  V.nc = A.cell_off[V.c + 1] - V.a0;
  const double *B = A.box + 9 * (size_t)V.c;
#pragma unroll
  for (int k = 0; k < 9; k++) V.b[k] = B[k];
  V.nx = A.nsub[3 * V.c];
  V.ny = A.nsub[3 * V.c + 1];
  V.nz = A.nsub[3 * V.c + 2];
  V.sst = A.sub_start + A.sub_off[V.c];
  V.nbr = A.nbr + (size_t)V.a0 * A.maxnb;
  return true;
}

// Per-workgroup shift tables.  FP64 form: S[code] = B t, the image shift of
// the pair (positions are absolute).  FP32 form: positions are taken from the
// centre of the atom's sub-cell (of the cell's centre along a dimension with
// one sub-cell), so |R| stays below a sub-cell width and FP32 keeps it to
// 1e-6 A; the pair's shift is B delta with delta = q / n per dimension (q the
// sub-cell offset of j from i, negated) or q (the rounding image where there
// is one sub-cell), held as FP32 hi and lo parts: r = ((R_i - R_j) + Sf) + Sl.
struct Shifts {
  double S[32][3];
  float4 Sf[32], Sl[32];
};

// shift tables and packed positions of the cell; the caller places a barrier
// after it (codes 27..31 never occur)
template <typename T, int BLOCK>
__device__ void phase_positions(const BatchArgs &A, const CellView &V, Shifts &SH) {
  const int t = threadIdx.x;
  for (int k = t; k < 32; k += BLOCK) {
    if (k >= 27) {
      SH.S[k][0] = 1.0e6; SH.S[k][1] = 0.0; SH.S[k][2] = 0.0;
      SH.Sf[k] = make_float4(1.0e6f, 0.f, 0.f, 0.f);
      SH.Sl[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else if constexpr (sizeof(T) == 8) {
      const double sx = (double)(k % 3 - 1), sy = (double)((k / 3) % 3 - 1), sz = (double)(k / 9 - 1);
      SH.S[k][0] = V.b[0] * sx + V.b[3] * sy + V.b[6] * sz;
      SH.S[k][1] = V.b[1] * sx + V.b[4] * sy + V.b[7] * sz;
      SH.S[k][2] = V.b[2] * sx + V.b[5] * sy + V.b[8] * sz;
    } else {
      const int qx = k % 3 - 1, qy = (k / 3) % 3 - 1, qz = k / 9 - 1;
      const double sx = V.nx >= 3 ? (double)qx / V.nx : (double)qx;
      const double sy = V.ny >= 3 ? (double)qy / V.ny : (double)qy;
      const double sz = V.nz >= 3 ? (double)qz / V.nz : (double)qz;
      const double dx = V.b[0] * sx + V.b[3] * sy + V.b[6] * sz, dy = V.b[1] * sx + V.b[4] * sy + V.b[7] * sz,
                   dz = V.b[2] * sx + V.b[5] * sy + V.b[8] * sz;
      const float hx = (float)dx, hy = (float)dy, hz = (float)dz;
      SH.Sf[k] = make_float4(hx, hy, hz, 0.f);
      SH.Sl[k] = make_float4((float)(dx - (double)hx), (float)(dy - (double)hy), (float)(dz - (double)hz), 0.f);
    }
  }
  for (int il = t; il < V.nc; il += BLOCK) {
    const size_t a = (size_t)V.a0 + il;
    double sx = A.s[3 * a], sy = A.s[3 * a + 1], sz = A.s[3 * a + 2];
    if constexpr (sizeof(T) == 8) {
      A.pos4[a] = make_double4(V.b[0] * sx + V.b[3] * sy + V.b[6] * sz,
                               V.b[1] * sx + V.b[4] * sy + V.b[7] * sz,
                               V.b[2] * sx + V.b[5] * sy + V.b[8] * sz, 0.0);
    } else {
      const int sc = A.subc[a], nx = V.nx, ny = V.ny, nz = V.nz;
      const int cx = sc % nx, cy = (sc / nx) % ny, cz = sc / (nx * ny);
      sx -= nx >= 3 ? (cx + 0.5) / nx : 0.5;
      sy -= ny >= 3 ? (cy + 0.5) / ny : 0.5;
      sz -= nz >= 3 ? (cz + 0.5) / nz : 0.5;
      A.p4f[a] = make_float4((float)(V.b[0] * sx + V.b[3] * sy + V.b[6] * sz),
                             (float)(V.b[1] * sx + V.b[4] * sy + V.b[7] * sz),
                             (float)(V.b[2] * sx + V.b[5] * sy + V.b[8] * sz), 0.f);
    }
  }
}

__device__ __forceinline__ float3 sep32(const float4 &pi, const float4 &pj, const Shifts &SH, int code) {
  const float4 d = SH.Sf[code], l = SH.Sl[code];
  return make_float3(((pi.x - pj.x) + d.x) + l.x, ((pi.y - pj.y) + d.y) + l.y, ((pi.z - pj.z) + d.z) + l.z);
}

// full neighbour list within the cell and its periodic images: the 3^3
// sub-cell stencil in dimensions with >= 3 sub-cells (image from the wrap),
// rounding of the fractional separation in dimensions with one sub-cell.
// FP64 form: the code is the image shift, the test r < rc in FP64.  FP32
// form: the code is the packed-shift offset q, the test on the FP32
// separation against (rc + 1e-4)^2 (pairs past the outermost knot add exact
// zeros).  The density pass rides along: each listed partner's phi is added
// to the atom's FP64 density in list order, from the separation the test
// used (the one the force pass recomputes), and the atom's embedding
// (src/dynamics_m.F90:497-579, src/potential_m.F90:148-169) follows its
// walk: rho, F' into pos4.w (and the FP32 record's w).  Returns this thread's
// embedding energy.
template <typename T, int BLOCK>
__device__ double phase_list(const BatchArgs &A, const PotDev &P, const CellView &V, const Shifts &SH,
                             bool store_rho) {
  const int nx = V.nx, ny = V.ny, nz = V.nz, nc = V.nc;
  const int rx_ = nx >= 3 ? 1 : 0, ry_ = ny >= 3 ? 1 : 0, rz_ = nz >= 3 ? 1 : 0;
  double e_emb = 0.0;
  for (int il = threadIdx.x; il < nc; il += BLOCK) {
    const size_t a = (size_t)V.a0 + il;
    const int sc = A.subc[a];
    const int cx = sc % nx, cy = (sc / nx) % ny, cz = sc / (nx * ny);
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (sizeof(T) == 8) pi = A.pos4[a];
    else pf = A.p4f[a];
    int cnt = 0;
    double acc = 0.0;
    if (sizeof(T) == 4 && rx_ && ry_ && rz_) {
      // FP32 form, stencil in every dimension (uniform branch): the shift of
      // a sub-cell's candidates is one code, loaded once per sub-cell
      const float4 *pc = A.p4f + V.a0;
      int *nb = V.nbr + il;  // entry k at nb[k * nc]
      for (int dz = -1; dz <= 1; dz++) {
        int kz = cz + dz;
        kz = kz < 0 ? kz + nz : (kz >= nz ? kz - nz : kz);
        for (int dy = -1; dy <= 1; dy++) {
          int ky = cy + dy;
          ky = ky < 0 ? ky + ny : (ky >= ny ? ky - ny : ky);
          for (int dx = -1; dx <= 1; dx++) {
            int kx = cx + dx;
            kx = kx < 0 ? kx + nx : (kx >= nx ? kx - nx : kx);
            const int cj = kx + nx * (ky + ny * kz);
            const int code = (1 - dx) + 3 * (1 - dy) + 9 * (1 - dz);
            const float4 d = SH.Sf[code], l = SH.Sl[code];
            const int tag = code << NB_CODE_SHIFT;
            const int j1 = V.sst[cj + 1];
            auto test = [&](const float4 pj, int jl) {
              const float ex = ((pf.x - pj.x) + d.x) + l.x, ey = ((pf.y - pj.y) + d.y) + l.y,
                          ez = ((pf.z - pj.z) + d.z) + l.z;
              const float r2 = ex * ex + ey * ey + ez * ez;
              if (r2 < A.rc2f && jl != il) {
                if (cnt < A.maxnb) nb[cnt * nc] = jl | tag;
                cnt++;
                float r, ri;
                r_inv(r2, r, ri);
                acc += (double)mdk_phi<float>(P, r);
              }
            };
            // four candidates per trip: their loads are in flight together;
            // the tests (and the density sum) run in candidate order
            int jl = V.sst[cj];
            for (; jl + 4 <= j1; jl += 4) {
              const float4 q0 = pc[jl], q1 = pc[jl + 1], q2 = pc[jl + 2], q3 = pc[jl + 3];
              test(q0, jl);
              test(q1, jl + 1);
              test(q2, jl + 2);
              test(q3, jl + 3);
            }
            for (; jl < j1; jl++) test(pc[jl], jl);
          }
        }
      }
    } else {
      const double sxi = A.s[3 * a], syi = A.s[3 * a + 1], szi = A.s[3 * a + 2];
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
            for (int jl = V.sst[cj]; jl < V.sst[cj + 1]; jl++) {
              if (jl == il) continue;
              const size_t aj = (size_t)V.a0 + jl;
              int tx = rx_ ? wx : -(int)round(sxi - A.s[3 * aj]);
              int ty = ry_ ? wy : -(int)round(syi - A.s[3 * aj + 1]);
              int tz = rz_ ? wz : -(int)round(szi - A.s[3 * aj + 2]);
              bool in;
              int code;
              T r2;
              if constexpr (sizeof(T) == 8) {
                code = (tx + 1) + 3 * (ty + 1) + 9 * (tz + 1);
                const double4 pj = A.pos4[aj];
                const double ex = pi.x - pj.x + SH.S[code][0], ey = pi.y - pj.y + SH.S[code][1],
                             ez = pi.z - pj.z + SH.S[code][2];
                r2 = ex * ex + ey * ey + ez * ez;
                in = r2 < A.rc2;
              } else {
                if (rx_) tx = -dx;
                if (ry_) ty = -dy;
                if (rz_) tz = -dz;
                code = (tx + 1) + 3 * (ty + 1) + 9 * (tz + 1);
                const float3 e = sep32(pf, A.p4f[aj], SH, code);
                r2 = e.x * e.x + e.y * e.y + e.z * e.z;
                in = r2 < A.rc2f;
              }
              if (in) {
                if (cnt < A.maxnb) V.nbr[(size_t)cnt * nc + il] = jl | (code << NB_CODE_SHIFT);
                cnt++;
                T r, ri;
                r_inv(r2, r, ri);
                acc += (double)mdk_phi<T>(P, r);
              }
            }
          }
        }
      }
    }
    A.numn[a] = min(cnt, A.maxnb);
    if (cnt > A.maxnb) atomicMax(&A.flags[0], cnt);
    if (store_rho) A.rho[a] = acc;
    double f = 0.0, df = 0.0;
    if (acc > 1e-40) {
      mdk_embed_f(P, acc, f, df);
      e_emb += f;
    }
    A.pos4[a].w = df;
    if constexpr (sizeof(T) == 4) A.p4f[a].w = (float)df;
  }
  return e_emb;
}

// separation of the k-th neighbour of local atom il, and the partner's F'
// (FP64 form: from pos4; FP32 form: from the FP32 records)
template <typename T>
__device__ __forceinline__ T nb_sep(const BatchArgs &A, const CellView &V, const Shifts &SH, const double4 &pi,
                                    const float4 &pf, int il, int k, T &dx, T &dy, T &dz, int &jl) {
  const unsigned e = (unsigned)V.nbr[(size_t)k * V.nc + il];
  jl = (int)(e & NB_IDX_MASK);
  const int code = (int)(e >> NB_CODE_SHIFT);
  if constexpr (sizeof(T) == 8) {
    const double4 pj = A.pos4[V.a0 + jl];
    dx = (T)(pi.x - pj.x + SH.S[code][0]);
    dy = (T)(pi.y - pj.y + SH.S[code][1]);
    dz = (T)(pi.z - pj.z + SH.S[code][2]);
    return (T)pj.w;
  } else {
    const float4 pj = A.p4f[V.a0 + jl];
    const float3 r = sep32(pf, pj, SH, code);
    dx = r.x;
    dy = r.y;
    dz = r.z;
    return pj.w;
  }
}

// densities and embedding over an existing list (the VJP's backward pass;
// the forward passes fold them into phase_list): rho (FP64 sum), F' into
// pos4.w (and the FP32 record's w); returns this thread's embedding energy
template <typename T, int BLOCK>
__device__ double phase_density(const BatchArgs &A, const PotDev &P, const CellView &V, const Shifts &SH,
                                bool store_rho) {
  double e_emb = 0.0;
  for (int il = threadIdx.x; il < V.nc; il += BLOCK) {
    const size_t a = (size_t)V.a0 + il;
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (sizeof(T) == 8) pi = A.pos4[a];
    else pf = A.p4f[a];
    const int nn = A.numn[a];
    double acc = 0.0;
    for (int k = 0; k < nn; k++) {
      T dx, dy, dz, r, ri;
      int jl;
      nb_sep<T>(A, V, SH, pi, pf, il, k, dx, dy, dz, jl);
      r_inv(dx * dx + dy * dy + dz * dz, r, ri);
      acc += (double)mdk_phi<T>(P, r);
    }
    if (store_rho) A.rho[a] = acc;
    double f = 0.0, df = 0.0;
    if (acc > 1e-40) {
      mdk_embed_f(P, acc, f, df);
      e_emb += f;
    }
    A.pos4[a].w = df;
    if constexpr (sizeof(T) == 4) A.p4f[a].w = (float)df;
  }
  return e_emb;
}

// the FP32 interval table (bins, then interval rows) into LDS
__device__ __forceinline__ void load_fe_table(const BatchArgs &A, float4 *fbin, float *fiv) {
  if (!A.fe_ok) return;
  const float4 *src = (const float4 *)A.fetab;
  for (int t = threadIdx.x; t < MDK_FE_NB + 2 * MDK_FE_NI; t += blockDim.x) {
    if (t < MDK_FE_NB) fbin[t] = src[t];
    else ((float4 *)fiv)[t - MDK_FE_NB] = src[t];
  }
}

// forces and virial (src/dynamics_m.F90:235-346).  With LOSS, also the force
// seed dL/dF_i = 2 f_weight (F_i - F_t,i) into gvec and this thread's share of
// sum |F_i - F_t,i|^2 in *fsq.  FP32 form: each atom's force, pair energy
// and virial terms are summed in FP32 over its partners, then added to the
// thread's FP64 sums; FP64 form: every term goes to the FP64 sums.
template <typename T, int BLOCK, bool LOSS>
__device__ void phase_force(const BatchArgs &A, const PotDev &P, const CellView &V, const Shifts &SH,
                            const float4 *fbin, const float *fiv, double &e_pair, double *w, double *fsq) {
  for (int il = threadIdx.x; il < V.nc; il += BLOCK) {
    const size_t a = (size_t)V.a0 + il;
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (sizeof(T) == 8) pi = A.pos4[a];
    else pf = A.p4f[a];
    const T dfi = sizeof(T) == 8 ? (T)pi.w : (T)pf.w;
    const int nn = A.numn[a];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if constexpr (sizeof(T) == 8) {
      for (int k = 0; k < nn; k++) {
        T dx, dy, dz, r, rr;
        int jl;
        const T dfj = nb_sep<T>(A, V, SH, pi, pf, il, k, dx, dy, dz, jl);
        r_inv(dx * dx + dy * dy + dz * dz, r, rr);
        T vv, dv;
        mdk_vee<T>(P, r, vv, dv);
        const T dph = mdk_dphi<T>(P, r);
        const T fp = -dv * rr + (-T(1) * (dph * dfi + dph * dfj) * rr);
        const double fpd = (double)fp, ddx = (double)dx, ddy = (double)dy, ddz = (double)dz;
        e_pair += 0.5 * (double)vv;
        fx += ddx * fpd;
        fy += ddy * fpd;
        fz += ddz * fpd;
        const double h = 0.5 * fpd;
        w[0] += ddx * ddx * h; w[1] += ddy * ddy * h; w[2] += ddz * ddz * h;
        w[3] += ddx * ddy * h; w[4] += ddx * ddz * h; w[5] += ddy * ddz * h;
      }
    } else {
      float sx = 0.f, sy = 0.f, sz = 0.f, se = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f;
      auto pair = [&](const float4 pj, unsigned e) {
        const float3 d = sep32(pf, pj, SH, (int)(e >> NB_CODE_SHIFT));
        const float dx = d.x, dy = d.y, dz = d.z, dfj = pj.w;
        float r, rr;
        r_inv(dx * dx + dy * dy + dz * dz, r, rr);
        float vv, dv, dph;
        if (A.fe_ok) {  // uniform: the knots admit the interval table
          mdk_fe_terms(P, fbin, fiv, r, vv, dv, dph);
        } else {
          mdk_vee<float>(P, r, vv, dv);
          dph = mdk_dphi<float>(P, r);
        }
        const float fp = -dv * rr + (-1.f * (dph * dfi + dph * dfj) * rr);
        const float gx = dx * fp, gy = dy * fp, gz = dz * fp;
        se += vv;
        sx += gx; sy += gy; sz += gz;
        s0 += dx * gx; s1 += dy * gy; s2 += dz * gz;
        s3 += dx * gy; s4 += dx * gz; s5 += dy * gz;
      };
      const unsigned *nb = (const unsigned *)V.nbr + il;  // entry k at nb[k * nc]
      const float4 *pc = A.p4f + V.a0;
      int k = 0;
      if constexpr (LOSS) {
        // four list entries per trip: their partner loads are in flight
        // together (loss + gradient 2.5 % faster).  The evaluation kernel
        // keeps one per trip: at its forced occupancy the four records spill
        // (1.86 -> 2.03 ms).  Either way the pairs are summed in list order.
        for (; k + 4 <= nn; k += 4) {
          unsigned e[4];
          float4 q[4];
#pragma unroll
          for (int u = 0; u < 4; u++) e[u] = nb[(size_t)(k + u) * V.nc];
#pragma unroll
          for (int u = 0; u < 4; u++) q[u] = pc[e[u] & NB_IDX_MASK];
#pragma unroll
          for (int u = 0; u < 4; u++) pair(q[u], e[u]);
        }
      }
      for (; k < nn; k++) {
        const unsigned e = nb[(size_t)k * V.nc];
        pair(pc[e & NB_IDX_MASK], e);
      }
      e_pair += 0.5 * (double)se;
      fx = sx; fy = sy; fz = sz;
      w[0] += 0.5 * (double)s0; w[1] += 0.5 * (double)s1; w[2] += 0.5 * (double)s2;
      w[3] += 0.5 * (double)s3; w[4] += 0.5 * (double)s4; w[5] += 0.5 * (double)s5;
    }
    const size_t u = (size_t)A.perm[a];
    A.out_f[3 * u] = fx;
    A.out_f[3 * u + 1] = fy;
    A.out_f[3 * u + 2] = fz;
    if (LOSS) {
      const double ex = fx - A.f_in[3 * u], ey = fy - A.f_in[3 * u + 1], ez = fz - A.f_in[3 * u + 2];
      *fsq += ex * ex + ey * ey + ez * ez;
      const double g = 2.0 * A.f_weight;
      A.gvec[3 * a] = g * ex;
      A.gvec[3 * a + 1] = g * ey;
      A.gvec[3 * a + 2] = g * ez;
    }
  }
}

// Adjoint of one cell.  With seeds e_c = dL/dE_c and G_i = dL/dF_i, F_i the
// force on atom i, r_ij = R_i - R_j and u_ij = (G_i - G_j) . r_ij / r:
//   dL/dtheta = sum_i sum_j { e_c [ V_theta/2 + F'_i phi_theta ]
//                             - [ V'_theta + (F'_i + F'_j) phi'_theta ] u_ij / 2
//                             - lambda_i phi_theta },
//   lambda_i = F''(rho_i) Q_i,  Q_i = sum_j phi'(r_ij) u_ij,  F'' = -2 F'^3,
// over the full list (every pair from both ends).  The lambda term is
// accumulated as lambda_i d(rho_i)/dtheta after the neighbour loop, so one
// pass over the list suffices.  F' comes from pos4.w (FP32 form: the
// partners' from the FP32 records), recomputed by the
// caller's density phase; G from gvec.
template <typename T, int BLOCK, int NV, int NP>
__device__ void phase_adjoint(const BatchArgs &A, const PotDev &P, const GradDev &D,
                              const CellView &V, const Shifts &SH, double ec, double *ga,
                              double *gr, double *gA, double *gR) {
  PotView<T> Q(P);
  for (int il = threadIdx.x; il < V.nc; il += BLOCK) {
    const size_t a = (size_t)V.a0 + il;
    const double4 pi = A.pos4[a];   // its w: F'_i in FP64 (both forms)
    float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (sizeof(T) == 4) pf = A.p4f[a];
    const double gix = A.gvec[3 * a], giy = A.gvec[3 * a + 1], giz = A.gvec[3 * a + 2];
    const T fpi = (T)pi.w;
    const int nn = A.numn[a];
    double q = 0.0, drA[NP], drR[NP];
#pragma unroll
    for (int m = 0; m < NP; m++) { drA[m] = 0.0; drR[m] = 0.0; }
    for (int k = 0; k < nn; k++) {
      T dx, dy, dz, r, ri;
      int jl;
      const T fpj = nb_sep<T>(A, V, SH, pi, pf, il, k, dx, dy, dz, jl);
      r_inv(dx * dx + dy * dy + dz * dz, r, ri);
      const size_t aj = (size_t)V.a0 + jl;
      const T u = ((T)(gix - A.gvec[3 * aj]) * dx + (T)(giy - A.gvec[3 * aj + 1]) * dy +
                   (T)(giz - A.gvec[3 * aj + 2]) * dz) * ri;
      const T dph = mdk_dphi<T>(P, r);
      q += (double)(dph * u);
      const T cv = T(0.5) * (T)ec, cvp = T(-0.5) * u;
      const T cp = (T)ec * fpi, cpp = T(-0.5) * (fpi + fpj) * u;
      // pair parameters
      if (r > Q.bc2()) {
#pragma unroll
        for (int m = 0; m < NV; m++) {
          if (m < P.nv) {
            const T d = Q.rv(m) - r;
            if (d > T(0)) {
              const T d2 = d * d;
              ga[m] += (double)(cv * d2 * d - T(3) * cvp * d2);
              gr[m] += (double)(Q.av(m) * (T(3) * cv * d2 - T(6) * cvp * d));
            }
          }
        }
      } else if (r > Q.bc1()) {
        // log-space Hermite join: only x2 and v2 depend on the knots
        const double h = P.bc2 - P.bc1, t = ((double)r - P.bc1) / h;
        const double t2 = t * t, t3 = t2 * t;
        const double sp = (2 * t3 - 3 * t2 + 1) * P.jx1 + (t3 - 2 * t2 + t) * P.jv1 * h +
                          (-2 * t3 + 3 * t2) * P.jx2 + (t3 - t2) * P.jv2 * h;
        const double ds = ((6 * t2 - 6 * t) * P.jx1 + (3 * t2 - 4 * t + 1) * P.jv1 * h +
                           (-6 * t2 + 6 * t) * P.jx2 + (3 * t2 - 2 * t) * P.jv2 * h) / h;
        const double ve = exp(sp), vpe = ds * ve;
        const double h01 = -2 * t3 + 3 * t2, h11 = (t3 - t2) * h;
        const double d01 = (-6 * t2 + 6 * t) / h, d11 = 3 * t2 - 2 * t;
#pragma unroll
        for (int m = 0; m < NV; m++) {
          if (m < P.nv) {
            const double sa = h01 * D.xa[m] + h11 * D.ya[m], sr = h01 * D.xr[m] + h11 * D.yr[m];
            const double pa = d01 * D.xa[m] + d11 * D.ya[m], pr = d01 * D.xr[m] + d11 * D.yr[m];
            ga[m] += (double)cv * (ve * sa) + (double)cvp * (pa * ve + vpe * sa);
            gr[m] += (double)cv * (ve * sr) + (double)cvp * (pr * ve + vpe * sr);
          }
        }
      }
      // density parameters
      if (r < Q.pcr()) {
#pragma unroll
        for (int m = 0; m < NP; m++) {
          if (m < P.np) {
            gA[m] += (double)cp * D.cpa[m];
            gR[m] += (double)cp * D.cpr[m];
            drA[m] += D.cpa[m];
            drR[m] += D.cpr[m];
          }
        }
      } else {
#pragma unroll
        for (int m = 0; m < NP; m++) {
          if (m < P.np) {
            const T d = Q.dp(m, r);
            if (d > T(0)) {
              const T d2 = d * d, pa = d2 * d, pr = T(3) * Q.ap(m) * d2;
              gA[m] += (double)(cp * pa - T(3) * cpp * d2);
              gR[m] += (double)(cp * pr - T(6) * cpp * Q.ap(m) * d);
              drA[m] += (double)pa;
              drR[m] += (double)pr;
            }
          }
        }
      }
    }
    const double lam = -2.0 * pi.w * pi.w * pi.w * q;  // F''(rho_i) Q_i
#pragma unroll
    for (int m = 0; m < NP; m++) {
      gA[m] -= lam * drA[m];
      gR[m] -= lam * drR[m];
    }
  }
}

// cell energy/virial reduction and output (fixed order: lanes, then waves);
// returns the cell energy to every thread
template <int BLOCK>
__device__ double store_cell(const BatchArgs &A, const CellView &V, double e, const double *w,
                             double (*red)[7]) {
  const int t = threadIdx.x;
  double v[7] = {e, w[0], w[1], w[2], w[3], w[4], w[5]};
#pragma unroll
  for (int k = 0; k < 7; k++) v[k] = wave_sum(v[k]);
  if ((t & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 7; k++) red[t >> 6][k] = v[k];
  }
  __syncthreads();
  double etot = red[0][0];
  for (int q2 = 1; q2 < BLOCK / 64; q2++) etot += red[q2][0];
  if (t < 7) {
    double sum = red[0][t];
    for (int q2 = 1; q2 < BLOCK / 64; q2++) sum += red[q2][t];
    if (t == 0) A.out_e[V.c] = sum;
    else A.out_w[6 * (size_t)V.c + t - 1] = sum;
  }
  __syncthreads();
  return etot;
}

// Forward: energies, forces and virials of every cell.  The kernel is bound
// by latency: six waves per SIMD (80 VGPRs, some spilled to scratch) run the
// 4096-cell batch in 2.1 ms against 3.7 ms at the three that 131 VGPRs allow.
template <typename T, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_batch_eval(BatchArgs A, PotDev P) {
  __shared__ Shifts S;
  __shared__ double red[BLOCK / 64][7];
  __shared__ int sh_cell;
  __shared__ float4 fbin[MDK_FE_NB];
  __shared__ __align__(16) float fiv[8 * MDK_FE_NI];
  load_fe_table(A, fbin, fiv);  // ordered before its use by next_cell's barriers
  CellView V;
  while (next_cell(A, &sh_cell, V)) {
    phase_positions<T, BLOCK>(A, V, S);
    __syncthreads();
    const double e_emb = phase_list<T, BLOCK>(A, P, V, S, true);
    __syncthreads();
    double e_pair = 0.0, w[6] = {0, 0, 0, 0, 0, 0};
    phase_force<T, BLOCK, false>(A, P, V, S, fbin, fiv, e_pair, w, nullptr);
    store_cell<BLOCK>(A, V, e_emb + e_pair, w, red);
  }
}

// per-cell gradient partials (fixed order) into out_g[c*NGRAD_SLOTS + slot],
// slots packed as a_0..a_{nv-1}, r_0.., A_0..A_{np-1}, R_0..
template <int BLOCK, int NV, int NP>
__device__ void store_grad(const BatchArgs &A, const PotDev &P, const CellView &V, const double *ga,
                           const double *gr, const double *gA, const double *gR, double *red) {
  double *out = A.out_g + (size_t)V.c * NGRAD_SLOTS;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    if (m < P.nv) {
      const double sa = block_sum<BLOCK>(ga[m], red);
      const double sr = block_sum<BLOCK>(gr[m], red);
      if (threadIdx.x == 0) {
        out[m] = sa;
        out[P.nv + m] = sr;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < NP; m++) {
    if (m < P.np) {
      const double sa = block_sum<BLOCK>(gA[m], red);
      const double sr = block_sum<BLOCK>(gR[m], red);
      if (threadIdx.x == 0) {
        out[2 * P.nv + m] = sa;
        out[2 * P.nv + P.np + m] = sr;
      }
    }
  }
}

// Loss and its gradient in one pass per cell:
//   L_c = (E_c - E_t,c)^2 / N_c^2 + f_weight sum_i |F_i - F_t,i|^2,
// forward phases, seeds e_c = 2 (E_c - E_t,c) / N_c^2 and G_i, then the
// adjoint phase on the forward's list and F'.
// VJP=true instead takes the seeds from e_in / f_in, reuses the neighbour
// list of the last forward launch and recomputes the densities on it.
// Four waves per SIMD (128 VGPRs and scratch spills) against the two that
// its 225 VGPRs allow: 3.7 against 5.4 ms on the 4096-cell batch.
// NV, NP: pair and density knot counts the per-thread gradient sums hold
// (6 and 2 for the Fe potential; the maxima for any other knot set).
template <typename T, int BLOCK, bool VJP, int NV, int NP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_batch_grad(BatchArgs A, PotDev P, GradDev D) {
  __shared__ Shifts S;
  __shared__ double red[BLOCK / 64][7];
  __shared__ double red1[BLOCK / 64];
  __shared__ int sh_cell;
  __shared__ float4 fbin[MDK_FE_NB];
  __shared__ __align__(16) float fiv[8 * MDK_FE_NI];
  if (!VJP) load_fe_table(A, fbin, fiv);  // ordered before its use by next_cell's barriers
  CellView V;
  while (next_cell(A, &sh_cell, V)) {
    phase_positions<T, BLOCK>(A, V, S);
    double ec;
    if (VJP) {
      for (int il = threadIdx.x; il < V.nc; il += BLOCK) {
        const size_t a = (size_t)V.a0 + il, u = (size_t)A.perm[a];
        A.gvec[3 * a] = A.f_in[3 * u];
        A.gvec[3 * a + 1] = A.f_in[3 * u + 1];
        A.gvec[3 * a + 2] = A.f_in[3 * u + 2];
      }
      ec = A.e_in[V.c];
      __syncthreads();
    } else {
      __syncthreads();
      const double e_emb = phase_list<T, BLOCK>(A, P, V, S, true);
      __syncthreads();
      double e_pair = 0.0, w[6] = {0, 0, 0, 0, 0, 0}, fsq = 0.0;
      phase_force<T, BLOCK, true>(A, P, V, S, fbin, fiv, e_pair, w, &fsq);
      const double fs = block_sum<BLOCK>(fsq, red1);
      const double ecell = store_cell<BLOCK>(A, V, e_emb + e_pair, w, red);
      const double inv_n2 = 1.0 / ((double)V.nc * (double)V.nc);
      const double de = ecell - A.e_in[V.c];
      ec = 2.0 * de * inv_n2;
      if (threadIdx.x == 0) A.out_loss[V.c] = de * de * inv_n2 + A.f_weight * fs;
    }
    // backward: the VJP recomputes the densities on the last forward's list
    // (F' into pos4.w); the loss path still holds them from its forward
    // phases.  The barrier also publishes the force seeds G for the partners.
    if (VJP) phase_density<T, BLOCK>(A, P, V, S, false);
    __syncthreads();
    double ga[NV], gr[NV], gA[NP], gR[NP];
#pragma unroll
    for (int m = 0; m < NV; m++) { ga[m] = 0.0; gr[m] = 0.0; }
#pragma unroll
    for (int m = 0; m < NP; m++) { gA[m] = 0.0; gR[m] = 0.0; }
    phase_adjoint<T, BLOCK, NV, NP>(A, P, D, V, S, ec, ga, gr, gA, gR);
    store_grad<BLOCK, NV, NP>(A, P, V, ga, gr, gA, gR, red1);
  }
}

// Sums the per-cell partials over cells in a fixed order: block k reduces
// slot k (k < nslot) of out_g, block nslot the per-cell losses.
// Block nslot (launched only for the loss) writes *loss.
__global__ void __launch_bounds__(256) k_grad_final(const double *out_g, const double *out_loss,
                                                    int ncell, int nslot, double *res, double *loss) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int c = threadIdx.x; c < ncell; c += 256)
    s += k < nslot ? out_g[(size_t)c * NGRAD_SLOTS + k] : out_loss[c];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    if (k < nslot) res[k] = s;
    else *loss = s;
  }
}
}  // namespace

// Scratch of mdk_batch_loss_grad_noalloc, one copy per GPU, held until the
// process ends: the work-queue counter (first 256 bytes), then the per-cell
// gradient partials (NGRAD_SLOTS per cell) and per-cell losses of at most
// MDK_NOALLOC_MAX_CELLS cells -- 1.6 MB.  Host code uses the address that
// hipGetSymbolAddress returns for the current device (scratch_base), never
// the array's name.
#define NOALLOC_HEAD 32
static __device__ __align__(256) double scratch[NOALLOC_HEAD + (NGRAD_SLOTS + 1) * (size_t)MDK_NOALLOC_MAX_CELLS];

namespace {
std::mutex scratch_mu;
double *scratch_base[64];  // per device ordinal: that device's copy of scratch

// Looks up the current device's copy of scratch, once per device.
int scratch_setup(int dev) {
  if (dev < 0 || dev >= 64) return berr(MDK_ERR_BAD_ARG, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_base[dev]) return MDK_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return berr(MDK_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
  scratch_base[dev] = (double *)p;
  return MDK_OK;
}
double *scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  return (dev >= 0 && dev < 64) ? scratch_base[dev] : nullptr;
}
}  // namespace

// One captured launch sequence (forward, or loss + gradient).
struct GraphSlot {
  hipGraph_t graph;
  hipGraphExec_t gexec;
  bool valid;
};

struct mdk_batch {
  int ncell, natom, precision, maxnb;
  hipStream_t stream;
  int device, ncu;
  int *order, *cell_off, *nsub, *sub_off, *sub_start, *subc, *perm, *numn, *nbr, *counter, *flags;
  double *box, *s, *rho, *out_e, *out_w, *out_f;
  double4 *pos4;
  float4 *p4f;
  float *fetab;  // FP32 interval table of the current knots (when they admit one)
  bool fe_ok;
  // adjoint buffers: targets, product seeds, force seeds (sorted order),
  // per-cell gradient and loss partials, reduced results
  double *e_tgt, *f_tgt, *e_seed, *f_seed, *gvec, *out_g, *out_loss, *res;
  std::vector<int> h_cell_off;
  std::vector<double> h_box;
  PotDev pot;
  GradDev gd;
  hostpot::Knots knots;
  bool pot_set, targets_set, list_valid;
  int block, wg_per_cu;
  bool use_graph;
  GraphSlot g_eval, g_grad;
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

template <typename T, bool VJP>
void launch_grad(int block, int grid, hipStream_t st, const BatchArgs &A, const PotDev &P,
                 const GradDev &D) {
  if (P.nv == MDK_FE_NV && P.np == MDK_FE_NP) {
    switch (block) {
      case 64: k_batch_grad<T, 64, VJP, MDK_FE_NV, MDK_FE_NP><<<grid, 64, 0, st>>>(A, P, D); break;
      case 128: k_batch_grad<T, 128, VJP, MDK_FE_NV, MDK_FE_NP><<<grid, 128, 0, st>>>(A, P, D); break;
      case 512: k_batch_grad<T, 512, VJP, MDK_FE_NV, MDK_FE_NP><<<grid, 512, 0, st>>>(A, P, D); break;
      default: k_batch_grad<T, 256, VJP, MDK_FE_NV, MDK_FE_NP><<<grid, 256, 0, st>>>(A, P, D); break;
    }
    return;
  }
  switch (block) {
    case 64: k_batch_grad<T, 64, VJP, MDK_MAXV, MDK_MAXP><<<grid, 64, 0, st>>>(A, P, D); break;
    case 128: k_batch_grad<T, 128, VJP, MDK_MAXV, MDK_MAXP><<<grid, 128, 0, st>>>(A, P, D); break;
    case 512: k_batch_grad<T, 512, VJP, MDK_MAXV, MDK_MAXP><<<grid, 512, 0, st>>>(A, P, D); break;
    default: k_batch_grad<T, 256, VJP, MDK_MAXV, MDK_MAXP><<<grid, 256, 0, st>>>(A, P, D); break;
  }
}

BatchArgs args_of(mdk_batch *b) {
  BatchArgs A;
  A.ncell = b->ncell; A.maxnb = b->maxnb; A.order = b->order; A.cell_off = b->cell_off;
  A.box = b->box; A.nsub = b->nsub; A.sub_off = b->sub_off; A.sub_start = b->sub_start;
  A.subc = b->subc; A.s = b->s; A.perm = b->perm; A.pos4 = b->pos4; A.p4f = b->p4f; A.nbr = b->nbr;
  A.numn = b->numn; A.rho = b->rho; A.out_e = b->out_e; A.out_w = b->out_w; A.out_f = b->out_f;
  A.counter = b->counter; A.flags = b->flags; A.rc2 = b->rc * b->rc;
  A.rc2f = (float)((b->rc + 1e-4) * (b->rc + 1e-4));
  A.fetab = b->fetab;
  A.fe_ok = (b->fe_ok && b->precision == MDK_PREC_FP32) ? 1 : 0;
  A.ngrad = 2 * b->knots.nv + 2 * b->knots.np;
  A.e_in = b->e_tgt; A.f_in = b->f_tgt; A.gvec = b->gvec; A.out_g = b->out_g;
  A.out_loss = b->out_loss; A.f_weight = 0.1 / (double)b->natom;
  return A;
}

int grid_of(const mdk_batch *b) { return std::max(1, std::min(b->ncell, b->ncu * b->wg_per_cu)); }

int launch_status() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return berr(MDK_ERR_HIP, "batch launch: %s", hipGetErrorString(e));
  return MDK_OK;
}

int enqueue(mdk_batch *b) {
  k_batch_reset<<<1, 64, 0, b->stream>>>(b->counter);
  if (b->precision == MDK_PREC_FP64)
    launch_eval<double>(b->block, grid_of(b), b->stream, args_of(b), b->pot);
  else
    launch_eval<float>(b->block, grid_of(b), b->stream, args_of(b), b->pot);
  return launch_status();
}

// loss + gradient (vjp = false, from the targets) or the vector-Jacobian
// product with the seeds e_seed/f_seed (vjp = true), then the fixed-order
// reduction over cells into res[0..ngrad) (and res[ngrad] = loss)
int enqueue_grad(mdk_batch *b, bool vjp) {
  BatchArgs A = args_of(b);
  if (vjp) {
    A.e_in = b->e_seed;
    A.f_in = b->f_seed;
  }
  k_batch_reset<<<1, 64, 0, b->stream>>>(b->counter);
  const int grid = grid_of(b);
  if (b->precision == MDK_PREC_FP64) {
    if (vjp) launch_grad<double, true>(b->block, grid, b->stream, A, b->pot, b->gd);
    else launch_grad<double, false>(b->block, grid, b->stream, A, b->pot, b->gd);
  } else {
    if (vjp) launch_grad<float, true>(b->block, grid, b->stream, A, b->pot, b->gd);
    else launch_grad<float, false>(b->block, grid, b->stream, A, b->pot, b->gd);
  }
  k_grad_final<<<A.ngrad + (vjp ? 0 : 1), 256, 0, b->stream>>>(b->out_g, b->out_loss, b->ncell,
                                                               A.ngrad, b->res, b->res + A.ngrad);
  return launch_status();
}

void drop_slot(GraphSlot &g) {
  if (g.valid) {
    hipGraphExecDestroy(g.gexec);
    hipGraphDestroy(g.graph);
    g.valid = false;
  }
}

void drop_graph(mdk_batch *b) {
  drop_slot(b->g_eval);
  drop_slot(b->g_grad);
}

// Captures the launches of enqueue (grad = false) or of the loss + gradient
// (grad = true) into a graph slot.
int capture(mdk_batch *b, bool grad) {
  GraphSlot &gs = grad ? b->g_grad : b->g_eval;
  drop_slot(gs);
  BCK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = grad ? enqueue_grad(b, false) : enqueue(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != MDK_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return berr(MDK_ERR_HIP, "capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&gs.gexec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return berr(MDK_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  gs.graph = g;
  gs.valid = true;
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
               b->out_f, b->pos4, b->p4f, b->fetab, b->e_tgt, b->f_tgt, b->e_seed, b->f_seed, b->gvec, b->out_g,
               b->out_loss, b->res};
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
  {
    const int rc = scratch_setup(b->device);  // mdk_batch_loss_grad_noalloc's scratch on this device
    if (rc != MDK_OK) {
      delete b;
      return rc;
    }
  }
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
  BALLOC(b->p4f, sizeof(float4) * natom);
  BALLOC(b->fetab, sizeof(float) * MDK_FE_TAB);
  BALLOC(b->nbr, sizeof(int) * (size_t)natom * b->maxnb);
  BALLOC(b->numn, sizeof(int) * natom);
  BALLOC(b->rho, sizeof(double) * natom);
  BALLOC(b->out_e, sizeof(double) * ncell);
  BALLOC(b->out_w, sizeof(double) * 6 * (size_t)ncell);
  BALLOC(b->out_f, sizeof(double) * 3 * (size_t)natom);
  BALLOC(b->e_tgt, sizeof(double) * ncell);
  BALLOC(b->f_tgt, sizeof(double) * 3 * (size_t)natom);
  BALLOC(b->e_seed, sizeof(double) * ncell);
  BALLOC(b->f_seed, sizeof(double) * 3 * (size_t)natom);
  BALLOC(b->gvec, sizeof(double) * 3 * (size_t)natom);
  BALLOC(b->out_g, sizeof(double) * NGRAD_SLOTS * (size_t)ncell);
  BALLOC(b->out_loss, sizeof(double) * ncell);
  BALLOC(b->res, sizeof(double) * (NGRAD_SLOTS + 1));
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
  hostpot::grad_consts(k, b->gd);
  // the FP32 form's interval table of V, V', phi' (knot sums where the knots
  // do not admit one); launches in flight still read the previous table
  float tab[MDK_FE_TAB];
  b->fe_ok = nv == MDK_FE_NV && np == MDK_FE_NP && hostpot::build_fe_table(k, tab, b->pot.fe_binh, b->pot.fe_binc);
  if (b->fe_ok) {
    if (hipStreamSynchronize(b->stream) != hipSuccess ||
        hipMemcpyAsync(b->fetab, tab, sizeof(tab), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return berr(MDK_ERR_HIP, "interval table upload failed");
  }
  if (rint > 4.5) return berr(MDK_ERR_BAD_ARG, "knots beyond 4.5 A are not supported by the batch grid");
  if (2.0 * rint >= b->min_width)
    return berr(MDK_ERR_BAD_ARG, "a cell is narrower (%.3f A) than twice the interaction range", b->min_width);
  b->rc = rint;
  b->pot_set = true;
  b->list_valid = false;
  drop_graph(b);  // the potential is a kernel argument of the captured launches
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
    if (!b->g_eval.valid) {
      int rc = capture(b, false);
      if (rc != MDK_OK) return rc;
    }
    BCK(hipGraphLaunch(b->g_eval.gexec, b->stream));
  } else {
    int rc = enqueue(b);
    if (rc != MDK_OK) return rc;
  }
  b->list_valid = true;
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

// ---------------------------------------------------------------------------
// Adjoint: d(loss)/d(knot parameters).  Gradient order: a_0..a_{nv-1},
// r_0..r_{nv-1}, A_0..A_{np-1}, R_0..R_{np-1} (2 nv + 2 np values).
// ---------------------------------------------------------------------------

// Targets of the loss: e_target[ncell] (eV) and f_target[3*natom] (Cartesian
// eV/A, the caller's atom order).
extern "C" int mdk_batch_set_targets(void *h, const double *e_target, const double *f_target) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || !e_target || !f_target) return berr(MDK_ERR_BAD_ARG, "bad arguments");
  BCK(hipMemcpyAsync(b->e_tgt, e_target, sizeof(double) * b->ncell, hipMemcpyHostToDevice, b->stream));
  BCK(hipMemcpyAsync(b->f_tgt, f_target, sizeof(double) * 3 * (size_t)b->natom, hipMemcpyHostToDevice,
                     b->stream));
  BCK(hipStreamSynchronize(b->stream));
  b->targets_set = true;
  return MDK_OK;
}

namespace {
int read_res(mdk_batch *b, int n, double *out) {
  int rc = mdk_batch_synchronize(b);
  if (rc != MDK_OK) return rc;
  BCK(hipMemcpy(out, b->res, sizeof(double) * n, hipMemcpyDeviceToHost));
  return MDK_OK;
}
int enqueue_loss_grad(mdk_batch *b) {
  if (!b->pot_set || !b->targets_set) return berr(MDK_ERR_STATE, "batch potential or targets not set");
  if (b->use_graph) {
    if (!b->g_grad.valid) {
      int rc = capture(b, true);
      if (rc != MDK_OK) return rc;
    }
    BCK(hipGraphLaunch(b->g_grad.gexec, b->stream));
  } else {
    int rc = enqueue_grad(b, false);
    if (rc != MDK_OK) return rc;
  }
  b->list_valid = true;
  return MDK_OK;
}
}  // namespace

// loss = sum_c (E_c - E_t,c)^2 / N_c^2 + (0.1 / natom) sum_i |F_i - F_t,i|^2 and
// grad = d(loss)/d(knot parameters), in one persistent launch (forward, seeds,
// densities recomputed, adjoint) plus a fixed-order reduction over cells.
// Energies and forces of the forward part are left for mdk_batch_results.
extern "C" int mdk_batch_loss_grad(void *h, double *loss, double *grad) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b) return berr(MDK_ERR_BAD_ARG, "null handle");
  int rc = enqueue_loss_grad(b);
  if (rc != MDK_OK) return rc;
  const int ng = 2 * b->knots.nv + 2 * b->knots.np;
  double r[NGRAD_SLOTS + 1];
  rc = read_res(b, ng + 1, r);
  if (rc != MDK_OK) return rc;
  if (grad) memcpy(grad, r, sizeof(double) * ng);
  if (loss) *loss = r[ng];
  return MDK_OK;
}

// Vector-Jacobian product: grad = sum_c dE[c] dE_c/dtheta + sum_i dF_i . dF_i/dtheta
// for seeds dE[ncell] and dF[3*natom] (Cartesian, caller's order), at the
// positions and potential of the last forward evaluation (mdk_batch_evaluate or
// mdk_batch_loss_grad; its neighbour list is reused, densities are recomputed).
extern "C" int mdk_batch_vjp(void *h, const double *dE, const double *dF, double *grad) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || !dE || !dF || !grad) return berr(MDK_ERR_BAD_ARG, "bad arguments");
  if (!b->pot_set || !b->list_valid)
    return berr(MDK_ERR_STATE, "mdk_batch_vjp needs a forward evaluation with the current potential");
  BCK(hipMemcpyAsync(b->e_seed, dE, sizeof(double) * b->ncell, hipMemcpyHostToDevice, b->stream));
  BCK(hipMemcpyAsync(b->f_seed, dF, sizeof(double) * 3 * (size_t)b->natom, hipMemcpyHostToDevice,
                     b->stream));
  int rc = enqueue_grad(b, true);
  if (rc != MDK_OK) return rc;
  return read_res(b, 2 * b->knots.nv + 2 * b->knots.np, grad);
}

// mdk_batch_loss_grad into caller device memory without allocating: the
// work-queue counter and the per-cell partials are in this device's copy of
// scratch; the kernels are those of mdk_batch_loss_grad (no graph), so the
// results agree bit for bit.
extern "C" int mdk_batch_loss_grad_noalloc(void *h, double *loss, double *grad) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || !loss || !grad) return berr(MDK_ERR_BAD_ARG, "bad arguments");
  if (!b->pot_set || !b->targets_set) return berr(MDK_ERR_STATE, "batch potential or targets not set");
  if (b->ncell > MDK_NOALLOC_MAX_CELLS)
    return berr(MDK_ERR_BAD_ARG, "%d cells exceed the no-allocation limit of %d cells", b->ncell,
                MDK_NOALLOC_MAX_CELLS);
  double *base = scratch_of(b->device);
  if (!base) return berr(MDK_ERR_STATE, "no scratch address for device %d", b->device);
  BatchArgs A = args_of(b);
  A.counter = (int *)base;
  A.out_g = base + NOALLOC_HEAD;
  A.out_loss = A.out_g + (size_t)NGRAD_SLOTS * b->ncell;
  k_batch_reset<<<1, 64, 0, b->stream>>>(A.counter);  // a kernel, not a memset: capture-safe
  const int grid = grid_of(b);
  if (b->precision == MDK_PREC_FP64)
    launch_grad<double, false>(b->block, grid, b->stream, A, b->pot, b->gd);
  else
    launch_grad<float, false>(b->block, grid, b->stream, A, b->pot, b->gd);
  k_grad_final<<<A.ngrad + 1, 256, 0, b->stream>>>(A.out_g, A.out_loss, b->ncell, A.ngrad, grad, loss);
  b->list_valid = true;
  return launch_status();
}

// Average time of `reps` loss + gradient launches (HIP events on the batch stream), ms.
extern "C" int mdk_batch_time_loss_grad(void *h, int reps, double *ms_avg) {
  mdk_batch *b = (mdk_batch *)h;
  if (!b || reps < 1) return berr(MDK_ERR_BAD_ARG, "bad arguments");
  hipEvent_t e0, e1;
  BCK(hipEventCreate(&e0));
  BCK(hipEventCreate(&e1));
  int rc = enqueue_loss_grad(b);
  if (rc != MDK_OK) return rc;
  BCK(hipEventRecord(e0, b->stream));
  for (int r = 0; r < reps; r++) {
    rc = enqueue_loss_grad(b);
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
