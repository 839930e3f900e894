// mdk_core.hip -- libmoldy_hip: MOLDY's Fe force hot path on MI355X (gfx950).
//
// Device design (one system per handle, one HIP stream):
//  * per-atom state is FP64 SoA, kept SORTED BY CELL of a fine grid (cell
//    width >= interaction range + skin) so that neighbour gathers are local;
//    the sort is redone at every full-list rebuild.  slot_of/orig map between
//    the reference's atom order and the device order.
//  * the force path uses a FULL neighbour list (both directions) pruned at
//    r_int + skin, r_int = 3.72645 A being the outermost knot beyond which
//    V, V', phi and phi' are identically zero; per-atom sums are gathers (no
//    atomics, deterministic).  It is rebuilt on the reference's ntcm cadence
//    and whenever an atom may have moved more than skin/2.
//  * the reference half list at rnear (NLIST/NUMN) is built on request by a
//    separate kernel that reproduces the reference's link-cell grid, stencil
//    and bit-exact distance test.
//  * reductions (energies, virial, kinetic tensor, thermodynamic sums) are
//    per-block partials reduced in a fixed order by one finishing block.
//  * box/thermo tensors live on the device (BoxState, double-buffered per step)
//    so an MD step needs no host round trip.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/moldy_hip.h"
#include "mdk_internal.h"
#include "mdk_wave.h"

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";
static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}
extern "C" const char *mdk_last_error(void) { return g_err; }

#define HIPCK(call)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err(MDK_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                              \
  } while (0)
#define RCK(expr)          \
  do {                     \
    int r_ = (expr);       \
    if (r_ != MDK_OK) return r_; \
  } while (0)

#include "mdk_hostpot.h"

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
enum {
  F_X0 = 0, F_Y0, F_Z0, F_X1, F_Y1, F_Z1, F_X2, F_Y2, F_Z2, F_FX, F_FY, F_FZ,
  F_RHO, F_AFRHO, F_DAFRHO, F_MASS, F_DEL2, F_X3, F_Y3, F_Z3, F_N
};

struct TimedEv {
  int cls;
  hipEvent_t a, b;
};

// monitor sample: [0] max squared displacement since the list was built,
// [1] max squared displacement of the next drift, then B0, b1 + b2 (the next
// drift's box change) and Bref (B0 at the list build), 9 each
#define MON_STRIDE 32
#define MON_B0 2
#define MON_DB 11
#define MON_BREF 20
#define MDK_NSPEC 128  // species (atomic numbers) covered by the mass/del2 table

// list grid for cell-relative packing (gx == 0: absolute packing)
struct PackGeo {
  int gx, gy, gz;
  double ix, iy, iz;  // 1/g
};

// Arguments of the half kick's finishing (k_finalize_kick, or deferred into
// the next force's finishing kernel).
struct FinKick {
  BoxState *bs;
  BoxState *bout;  // the advanced copy of a fused drift: receives the same sums
  double *trace_row, *mon;
  double dt, bdel2, press;
};

struct mdk_ctx {
  mdk_params p;
  int n;
  int pbc;                 // periodic dims mask from ivol (bit0 x, bit1 y, bit2 z)
  int device;
  hipStream_t stream;
  // per-atom state (sorted order): F_N doubles per atom, field-major
  double *atoms, *atoms_alt;
  int *zat, *zat_alt, *orig, *orig_alt, *slot_of;
  float *sref;             // fractional positions at the last full-list build
  double4 *pos4;           // packed positions (absolute or cell-relative) and F' for the pair kernels
  int *img;                // integer image floor(s) fixed at the last full-list build
  bool pos4_valid;         // pos4 (and p4f) match the positions and box
  bool p4f_valid;          // p4f matches them (the FP32 lean passes read only p4f)
  bool masschk;            // some atom has mass <= 0.05: keep the per-pair mass test
  // mass and del2 as functions of the species when every atom of a species
  // carries the same values: spec_tab[2 z], spec_tab[2 z + 1] (device), used
  // by the kick and kinten instead of the per-atom fields; host copies of the
  // uploads (reference order) decide it
  double *spec_tab;
  bool spec_ok;
  double *h_mass, *h_del2;
  int *h_z;
  // reference link cells (1-based cell per atom, sorted order)
  int *ic;
  // parity half list in reference layout: numn_ref[nm] and nlist_ref[nm*nnbrs] (orig order)
  int *numn_ref, *nlist_ref;
  // full list
  int maxn;
  int *nbr, *numn;
  int *nact;  // per atom: list slots the force pass visits (set by the density pass)
  bool rel;     // positions packed cell-relative (block list) rather than absolute
  float4 *p4f;  // FP32 copy of the packed positions (FP32 form only)
  bool part_due;  // the next density pass reorders the new list (in-range partners first)
  int *cell_of, *cell_count, *cell_start, *cell_fill, *members;
  int ncell_cap;
  int grid[3];
  int *dflags;             // [0] full-list overflow, [1] half-list overflow, [2] grid too coarse
  // list-kernel flags copied to host-mapped memory after a deferred build,
  // and the event that marks them complete
  int *flags_host, *flags_map;
  hipEvent_t list_ev;
  bool list_pending;       // a deferred build whose flags the host has not checked yet
  // 2x2x2-cell blocks of the list build (Morton key / 8), for grid lblock_g
  int *lblocks, nlblocks, lblocks_cap, lblock_g[3];
  long rebuilds;
  bool list_valid;
  bool half_valid;         // numn_ref/nlist_ref hold a reference half list
  bool x3_live;            // X3..Z3 may be non-zero (uploaded, or the predictor-corrector ran)
  // a half kick whose sums the next force's finishing kernel completes
  bool kick_pending;
  FinKick kick_pend;
  int kick_slot;
  double *kstage;          // first-stage sums of the deferred kick
  // reference link-cell chains (src/linkcell_m.F90:122-274) on the host in
  // reference atom order, 1-based: L0 per cell, LINK, IC; chains_exact is
  // false after relinks inside mdk_md_steps, which update only the device ic
  int *h_l0, *h_link, *h_ic, h_ncell;
  bool chains_exact;
  int *ref_key, *rank_orig, *rank_slot;  // device: cell keys and chain ranks
  int last_rebuild_step;
  // box state: 2 copies
  BoxState *box;
  int cur;
  // partials
  // per-block partials, value-major: force() (NPART_F values: force pass,
  // kinten, embedding energy); half kick
  double *part_f, *part_k;
  double *stage_f, *stage_k;  // first-stage reductions of force()'s and scalev's partials
  unsigned *fin_cnt;          // arrival counters of k_finalize_force and k_fold, zeroed
  int nblk;
  // host-mapped monitor ring, MON_STRIDE doubles per step slot (4 slots):
  // [0] maxdisp2, [1] maxstep2, [2..10] B0 after the step, [11] full-list
  // overflow flag, [12] grid-too-coarse flag
  double *mon_host, *mon_dev;
  hipEvent_t mon_ev[4];
  double b0_lag[9];        // B0 of a completed step (from the monitor ring)
  long b0_lag_step;        // its step index, -1 when none
  bool drifted;            // the next step's drift was fused into the last kick
  bool kin_fresh;          // kinten's partial sums match the velocities and box of the next force
  long step_counter;
  // trace buffer on device
  double *trace_dev;
  int trace_cap;
  // potential
  PotDev pot;
  bool pot_set;
  double *tabs;
  float *fetab;   // interval table of the FP32 Fe-Fe pair terms (MDK_FE_TAB floats)
  bool fe_ok;     // fetab holds the current knots' table
  hostpot::Knots knots;
  int material;            // MDK_MAT_*: fetransmetal (knots) or iron_carbon (fec)
  int iverlet;             // md_steps integrator: 1 velocity Verlet, 0 predictor-corrector
  bool pc_primed;          // a correc or velocityverlet has run (predic is live)
  double snhv;             // Nose-Hoover velocity scaling accumulated by scalev
  double *red;             // 4 doubles: results of the scalev folds
  hostpot::FecCoef fec;
  double r_int_exact;      // interaction range of the exact functions
  // reference list threshold: r < rnear  <=>  r^2 < rnear2_exact (see near_threshold)
  double rnear2_exact;
  double r_list, skin;
  int ntc;
  // timing
  int timing_mask;          // kernel classes recorded with HIP events
  int timing_every;         // record one launch in this many per class
  long t_seen[TC_NCLASS];
  std::vector<TimedEv> evs;
  std::vector<hipEvent_t> evpool;
  double t_ms[TC_NCLASS];
  long t_cnt[TC_NCLASS];
  bool linked;
};

static inline double *fld(double *base, int f, int n) { return base + (size_t)f * n; }
static inline PackGeo pack_geo(const mdk_ctx *c) {
  PackGeo pg = {0, 0, 0, 0.0, 0.0, 0.0};
  if (c->rel) {
    pg.gx = c->grid[0]; pg.gy = c->grid[1]; pg.gz = c->grid[2];
    pg.ix = 1.0 / pg.gx; pg.iy = 1.0 / pg.gy; pg.iz = 1.0 / pg.gz;
  }
  return pg;
}
static inline float4 *p4f_of(const mdk_ctx *c) { return c->p.precision == MDK_PREC_FP32 ? c->p4f : nullptr; }
// the (mass, del2) species table when it holds, else the per-atom fields
static inline const double *spec_of(const mdk_ctx *c) { return c->spec_ok ? c->spec_tab : nullptr; }

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// Block-reduce K sums (and the last KM entries as maxima) into
// part[k*gridDim.x + blockIdx.x] (value-major, so the finishing block's reads
// of one value across all blocks are contiguous).
template <int K, int KM, int BLOCK = MDK_BLOCK>
__device__ __forceinline__ void block_store(double (&v)[K], double *part) {
  __shared__ double sh[BLOCK / 64][K];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  wave_reduce_all<K, KM>(v);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) sh[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    double s = sh[0][k];
    for (int q = 1; q < BLOCK / 64; q++) s = (k < K - KM) ? s + sh[q][k] : fmax(s, sh[q][k]);
    part[(size_t)k * gridDim.x + blockIdx.x] = s;
  }
}

// kinten's terms of one atom (src/thermostat_m.F90:127-177): tk(a,b) += m
// (B0 x1)_a x1_b, column-major in v[0..8], with B0 the box of the force
// evaluation the velocities x1 belong to
__device__ __forceinline__ void kin_terms(const double *b, double a1, double a2, double a3, double mi,
                                          double (&v)[9]) {
#pragma clang fp contract(off)  // the same rounding in every kernel that calls it
  const double mvx = (b[0] * a1 + b[3] * a2 + b[6] * a3) * mi;
  const double mvy = (b[1] * a1 + b[4] * a2 + b[7] * a3) * mi;
  const double mvz = (b[2] * a1 + b[5] * a2 + b[8] * a3) * mi;
  v[0] = a1 * mvx; v[1] = a1 * mvy; v[2] = a1 * mvz;
  v[3] = a2 * mvx; v[4] = a2 * mvy; v[5] = a2 * mvz;
  v[6] = a3 * mvx; v[7] = a3 * mvy; v[8] = a3 * mvz;
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2; remapping physical block b to logical tile
// start(b%8) + b/8 gives every XCD one contiguous range of (cell-sorted)
// atoms, so its neighbour gathers stay inside its own L2.
#define MDK_NXCD 8
__device__ __forceinline__ int xcd_tile() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb / MDK_NXCD, r = nb % MDK_NXCD, x = b % MDK_NXCD;
  return x * q + min(x, r) + b / MDK_NXCD;
}

__device__ __forceinline__ void min_image(int pbc, double &dx, double &dy, double &dz) {
  if (pbc & 1) dx = dx - round(dx);
  if (pbc & 2) dy = dy - round(dy);
  if (pbc & 4) dz = dz - round(dz);
}

// Reference-order Cartesian separation squared (src/parrinellorahman_m.F90:167-172),
// no FMA: bit-identical to the reference's r*r before the square root.
__device__ __forceinline__ double sep2_exact(const double *b, double dx, double dy, double dz) {
#pragma clang fp contract(off)
  const double rx = b[0] * dx + b[3] * dy + b[6] * dz;
  const double ry = b[1] * dx + b[4] * dy + b[7] * dz;
  const double rz = b[2] * dx + b[5] * dy + b[8] * dz;
  return rx * rx + ry * ry + rz * rz;
}

// 3x3 helpers on column-major arrays (device, one thread)
__device__ void dev_invert(const double *a, double *b, int *status) {
#pragma clang fp contract(off)
  double det = TT(a, 1, 1) * (TT(a, 2, 2) * TT(a, 3, 3) - TT(a, 2, 3) * TT(a, 3, 2));
  det = det + TT(a, 2, 1) * (TT(a, 3, 2) * TT(a, 1, 3) - TT(a, 1, 2) * TT(a, 3, 3));
  det = det + TT(a, 3, 1) * (TT(a, 1, 2) * TT(a, 2, 3) - TT(a, 1, 3) * TT(a, 2, 2));
  TT(b, 1, 1) = TT(a, 2, 2) * TT(a, 3, 3) - TT(a, 2, 3) * TT(a, 3, 2);
  TT(b, 1, 2) = TT(a, 3, 1) * TT(a, 2, 3) - TT(a, 2, 1) * TT(a, 3, 3);
  TT(b, 1, 3) = TT(a, 2, 1) * TT(a, 3, 2) - TT(a, 3, 1) * TT(a, 2, 2);
  TT(b, 2, 1) = TT(a, 3, 2) * TT(a, 1, 3) - TT(a, 1, 2) * TT(a, 3, 3);
  TT(b, 2, 2) = TT(a, 1, 1) * TT(a, 3, 3) - TT(a, 1, 3) * TT(a, 3, 1);
  TT(b, 2, 3) = TT(a, 3, 1) * TT(a, 1, 2) - TT(a, 1, 1) * TT(a, 3, 2);
  TT(b, 3, 1) = TT(a, 1, 2) * TT(a, 2, 3) - TT(a, 1, 3) * TT(a, 2, 2);
  TT(b, 3, 2) = TT(a, 2, 1) * TT(a, 1, 3) - TT(a, 2, 3) * TT(a, 1, 1);
  TT(b, 3, 3) = TT(a, 1, 1) * TT(a, 2, 2) - TT(a, 1, 2) * TT(a, 2, 1);
  if (fabs(det) == 0.0) {
    *status = MDK_ERR_SINGULAR_BOX;
    return;
  }
  for (int k = 0; k < 9; k++) b[k] = b[k] / det;
}
__device__ void dev_matmul(const double *a, const double *b, double *c) {
#pragma clang fp contract(off)
  for (int i = 1; i <= 3; i++) {
    for (int j = 1; j <= 3; j++) TT(c, i, j) = 0.0;
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++) TT(c, i, j) = TT(c, i, j) + TT(a, i, k) * TT(b, k, j);
  }
}
__device__ double dev_volume(const double *b) {
#pragma clang fp contract(off)
  return TT(b, 1, 1) * (TT(b, 2, 2) * TT(b, 3, 3) - TT(b, 2, 3) * TT(b, 3, 2)) +
         TT(b, 2, 1) * (TT(b, 3, 2) * TT(b, 1, 3) - TT(b, 1, 2) * TT(b, 3, 3)) +
         TT(b, 3, 1) * (TT(b, 1, 2) * TT(b, 2, 3) - TT(b, 2, 2) * TT(b, 1, 3));
}


// A neighbour entry of nbr[] is j | code << 27.  The pair separation is
// r_ij = R_i - R_j + D[code] with R the packed positions (pos4):
//  * absolute packing (legacy list, any box): R = B0 (s - img) with the integer
//    image img fixed at the last rebuild, code = image shift (sx+1) + 3 (sy+1)
//    + 9 (sz+1) with s in {-1,0,1}^3, D = B0 (sx, sy, sz);
//  * cell-relative packing (block list): R = B0 (s - img - f) with f the
//    fractional centre of the atom's list cell, code = offset (dx+1) + 3 (dy+1)
//    + 9 (dz+1) of j's list cell (unwrapped) from i's, D = -B0 (dx/gx, dy/gy,
//    dz/gz).  |R| stays below a cell width, so an FP32 copy of R loses little
//    and one 16-byte gather carries a partner in the FP32 form.
// Codes 27..31 place the partner 1e6 A away: null entries that pad quads.
#define NBR_IDX_MASK 0x07FFFFFF
#define NBR_CODE_SHIFT 27
// Neighbour k of atom i lives at nbr[((k/4) n + i) 4 + k%4]: four consecutive
// entries of one atom form one 16-byte load, and consecutive atoms' quads are
// contiguous (coalesced dwordx4 loads).
__device__ __forceinline__ size_t nbr_at(int k, int i, int n) {
  return (((size_t)(k >> 2) * n + i) << 2) + (k & 3);
}
#define NBR_NULL_CODE 31
__device__ __forceinline__ void shift_table(const double *b, PackGeo pg, double (*S)[3]) {
  for (int c = threadIdx.x; c < 32; c += blockDim.x) {
    if (c >= 27) {
      S[c][0] = 1.0e6;
      S[c][1] = 0.0;
      S[c][2] = 0.0;
      continue;
    }
    double sx = (double)(c % 3 - 1), sy = (double)((c / 3) % 3 - 1), sz = (double)(c / 9 - 1);
    if (pg.gx) {
      sx = -sx * pg.ix;
      sy = -sy * pg.iy;
      sz = -sz * pg.iz;
    }
    S[c][0] = b[0] * sx + b[3] * sy + b[6] * sz;
    S[c][1] = b[1] * sx + b[4] * sy + b[7] * sz;
    S[c][2] = b[2] * sx + b[5] * sy + b[8] * sz;
  }
}
// inverse of a 3x3 column-major matrix (cofactors / determinant)
__device__ __forceinline__ void inv3(const double *b, double *bi) {
  const double c00 = TT(b, 2, 2) * TT(b, 3, 3) - TT(b, 2, 3) * TT(b, 3, 2);
  const double c01 = TT(b, 2, 3) * TT(b, 3, 1) - TT(b, 2, 1) * TT(b, 3, 3);
  const double c02 = TT(b, 2, 1) * TT(b, 3, 2) - TT(b, 2, 2) * TT(b, 3, 1);
  const double det = TT(b, 1, 1) * c00 + TT(b, 1, 2) * c01 + TT(b, 1, 3) * c02;
  const double id = 1.0 / det;
  TT(bi, 1, 1) = c00 * id;
  TT(bi, 2, 1) = c01 * id;
  TT(bi, 3, 1) = c02 * id;
  TT(bi, 1, 2) = (TT(b, 1, 3) * TT(b, 3, 2) - TT(b, 1, 2) * TT(b, 3, 3)) * id;
  TT(bi, 2, 2) = (TT(b, 1, 1) * TT(b, 3, 3) - TT(b, 1, 3) * TT(b, 3, 1)) * id;
  TT(bi, 3, 2) = (TT(b, 1, 2) * TT(b, 3, 1) - TT(b, 1, 1) * TT(b, 3, 2)) * id;
  TT(bi, 1, 3) = (TT(b, 1, 2) * TT(b, 2, 3) - TT(b, 1, 3) * TT(b, 2, 2)) * id;
  TT(bi, 2, 3) = (TT(b, 1, 3) * TT(b, 2, 1) - TT(b, 1, 1) * TT(b, 2, 3)) * id;
  TT(bi, 3, 3) = (TT(b, 1, 1) * TT(b, 2, 2) - TT(b, 1, 2) * TT(b, 2, 1)) * id;
}

// ---------------------------------------------------------------------------
// permutation / transfer kernels
// ---------------------------------------------------------------------------
__global__ void k_scatter_d(double *dst, const double *src, const int *slot_of, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[slot_of[i]] = src[i];
}
__global__ void k_gather_d(double *dst, const double *src, const int *slot_of, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[slot_of[i]];
}
__global__ void k_scatter_i(int *dst, const int *src, const int *slot_of, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[slot_of[i]] = src[i];
}
__global__ void k_gather_i(int *dst, const int *src, const int *slot_of, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[slot_of[i]];
}
__global__ void k_iota(int *a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
// new slot k takes old slot members[k]
// the per-atom fields a re-sort moves (the others are rewritten before any read)
struct FieldSet {
  int f[F_N];
  int cnt;
};
__global__ void k_permute(double *dst, const double *src, const int *members, int n, FieldSet fs) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * fs.cnt) return;
  const int f = fs.f[t / n], k = (int)(t % n);
  dst[(size_t)f * n + k] = src[(size_t)f * n + members[k]];
}
__global__ void k_permute_int(int *zdst, const int *zsrc, int *odst, const int *osrc, int *slot_of,
                              const int *members, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int o = members[k];
  zdst[k] = zsrc[o];
  int orig = osrc[o];
  odst[k] = orig;
  slot_of[orig] = k;
}

// ---------------------------------------------------------------------------
// cells: counting sort helpers
// ---------------------------------------------------------------------------
// Runs of equal keys within a wave (atoms arrive nearly sorted by cell): every
// lane learns its run's head lane and length, and only heads touch the
// counters, once per run.  Lanes past n form their own runs and are skipped.
struct KeyRun {
  int key, head, len;
  bool is_head;
};
__device__ __forceinline__ KeyRun key_run(const int *key, int n, int i) {
  const int lane = threadIdx.x & 63;
  const int k = i < n ? key[i] : -1;
  const int prev = __shfl_up(k, 1, 64);
  const bool head = lane == 0 || prev != k;
  const unsigned long long heads = __ballot(head);
  const unsigned long long above = lane == 63 ? 0ull : heads >> (lane + 1);
  const unsigned long long upto = lane == 63 ? heads : heads & ((2ull << lane) - 1ull);
  KeyRun r;
  r.key = k;
  r.is_head = head;
  r.len = above ? __builtin_ctzll(above) + 1 : 64 - lane;
  r.head = 63 - __builtin_clzll(upto);
  return r;
}
__global__ void k_count(const int *key, int n, int *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const KeyRun r = key_run(key, n, i);
  if (r.is_head && r.key >= 0) atomicAdd(&count[r.key], r.len);
}
// exclusive scan of count[0..m) into start[0..m] (start[m] = total) in three
// passes: per-block scans of SCAN_CHUNK elements, a scan of the block totals,
// and the add-back.
#define SCAN_CHUNK 2048
__device__ __forceinline__ int block_exclusive_scan_1024(int v, int *sh, int &total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int u = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += u;
    __syncthreads();
  }
  total = sh[1023];
  int ex = sh[t] - v;
  __syncthreads();
  return ex;
}
__global__ void __launch_bounds__(1024) k_scan_a(const int *count, int *start, int *btot, int m) {
  __shared__ int sh[1024];
  const int base = blockIdx.x * SCAN_CHUNK + 2 * threadIdx.x;
  const int c0 = base < m ? count[base] : 0, c1 = base + 1 < m ? count[base + 1] : 0;
  int total;
  int ex = block_exclusive_scan_1024(c0 + c1, sh, total);
  if (base < m) start[base] = ex;
  if (base + 1 < m) start[base + 1] = ex + c0;
  if (threadIdx.x == 0) btot[blockIdx.x] = total;
}
__global__ void __launch_bounds__(1024) k_scan_b(int *btot, int nb, int *start, int m) {
  __shared__ int sh[1024];
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    int t = b0 + threadIdx.x;
    int v = t < nb ? btot[t] : 0;
    int total;
    int ex = block_exclusive_scan_1024(v, sh, total);
    if (t < nb) btot[t] = run + ex;
    run += total;
  }
  if (threadIdx.x == 0) start[m] = run;
}
__global__ void k_scan_c(int *start, const int *btot, int m) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) start[k] += btot[k / SCAN_CHUNK];
}
__global__ void k_fill(const int *key, int n, const int *start, int *fill, int *members) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const KeyRun r = key_run(key, n, i);
  int base = 0;
  if (r.is_head && r.key >= 0) base = start[r.key] + atomicAdd(&fill[r.key], r.len);
  base = __shfl(base, r.head, 64);  // the run's slots follow its head's reservation
  if (i < n) members[base + ((int)(threadIdx.x & 63) - r.head)] = i;
}
// deterministic order inside each cell: ascending sort key (key[members] or members)
__global__ void k_sort_cells(const int *start, int m, int *members, const int *key) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  int b = start[c], e = start[c + 1];
  for (int i = b + 1; i < e; i++) {
    int v = members[i];
    int kv = key ? key[v] : v;
    int j = i - 1;
    while (j >= b) {
      int u = members[j];
      int ku = key ? key[u] : u;
      if (ku <= kv) break;
      members[j + 1] = u;
      j--;
    }
    members[j + 1] = v;
  }
}

// reference link cell of each atom (src/linkcell_m.F90:149-153), 1-based
__global__ void k_ref_cells(const double *x0, const double *y0, const double *z0, int n, int nlcx,
                            int nlcy, int nlcz, int *ic) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hm = 0.4999999999;
  double x = x0[i], y = y0[i], z = z0[i];
  int ix = (int)((x - round(x) + hm) * (double)nlcx);
  int iy = (int)((y - round(y) + hm) * (double)nlcy);
  int iz = (int)((z - round(z) + hm) * (double)nlcz);
  // the reference's range is [0, nlc-1] for finite positions; clamp guards NaN
  ix = min(max(ix, 0), nlcx - 1);
  iy = min(max(iy, 0), nlcy - 1);
  iz = min(max(iz, 0), nlcz - 1);
  ic[i] = ix + nlcx * iy + nlcx * nlcy * iz + 1;
}
__global__ void k_minus1(const int *in, int *out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] - 1;
}

// Reference half list (src/neighbourlist_m.F90:143-255): atom i lists j when
// j lies in i's own link cell after it in the cell's chain from L0 (rank =
// position in that chain) or in one of the 13 forward cells, and r < rnear
// with the reference's arithmetic; cells in the reference's stencil order and
// members in chain order (bin_atoms sorts them by rank), so NLIST(:, i) is the
// reference's, entry for entry.  Output in reference layout NLIST(nnbrs, nm),
// 1-based.
__global__ void k_half_list(const double *x0, const double *y0, const double *z0, const int *ic,
                            const int *orig, const int *rank, const int *cstart, const int *members, int n,
                            int nlcx, int nlcy, int nlcz, const BoxState *bs, double rn2,
                            int pbc, int nnbrs, int *numn_ref, int *nlist_ref, int *flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *b = bs->b0;
  const int c = ic[i] - 1;
  const int cx = c % nlcx, cy = (c / nlcx) % nlcy, cz = c / (nlcx * nlcy);
  const double xi = x0[i], yi = y0[i], zi = z0[i];
  const int oi = orig[i], ri = rank[i];
  int cnt = 0;
  for (int dz = -1; dz <= 1; dz++) {
    int jz = (cz + dz + nlcz) % nlcz;
    for (int dy = -1; dy <= 1; dy++) {
      int jy = (cy + dy + nlcy) % nlcy;
      for (int dx = -1; dx <= 1; dx++) {
        int jx = (cx + dx + nlcx) % nlcx;
        // forward stencil of src/neighbourlist_m.F90:163-191
        bool fwd = (dz == 1) || (dz == 0 && dy == 1) || (dz == 0 && dy == 0 && dx == 1);
        bool own = (dz == 0 && dy == 0 && dx == 0);
        if (!fwd && !own) continue;
        int cj = jx + nlcx * (jy + nlcy * jz);
        for (int q = cstart[cj]; q < cstart[cj + 1]; q++) {
          int j = members[q];
          if (own && rank[j] <= ri) continue;
          double ddx = xi - x0[j], ddy = yi - y0[j], ddz = zi - z0[j];
          min_image(pbc, ddx, ddy, ddz);
          double r2 = sep2_exact(b, ddx, ddy, ddz);
          if (r2 < rn2) {
            cnt++;
            if (cnt <= nnbrs) nlist_ref[(size_t)oi * nnbrs + cnt - 1] = orig[j] + 1;
          }
        }
      }
    }
  }
  numn_ref[oi] = cnt;
  if (cnt > nnbrs) atomicOr(&flags[1], 1);
}

// Morton (Z-order) key of a fine-grid cell: consecutive keys form compact 3-D
// blocks, so atoms sorted by key have spatially local neighbourhoods.
__device__ __forceinline__ unsigned spread3(unsigned v) {
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000FF;
  v = (v | (v << 8)) & 0x0300F00F;
  v = (v | (v << 4)) & 0x030C30C3;
  v = (v | (v << 2)) & 0x09249249;
  return v;
}
__device__ __forceinline__ unsigned compact3(unsigned v) {
  v &= 0x09249249;
  v = (v | (v >> 2)) & 0x030C30C3;
  v = (v | (v >> 4)) & 0x0300F00F;
  v = (v | (v >> 8)) & 0x030000FF;
  v = (v | (v >> 16)) & 0x000003FF;
  return v;
}
__device__ __forceinline__ int morton3(int x, int y, int z) {
  return (int)(spread3((unsigned)x) | (spread3((unsigned)y) << 1) | (spread3((unsigned)z) << 2));
}

// Packed position of atom i from fractional s - img (see the entry format).
// k: the atom's cell key (read only when pg.gx is set)
__device__ __forceinline__ double3 packed_pos_key(const double *b, PackGeo pg, unsigned k, double fx, double fy,
                                                  double fz) {
#pragma clang fp contract(off)  // the same packed positions from every kernel that writes them
  if (pg.gx) {
    fx -= ((double)compact3(k) + 0.5) * pg.ix;
    fy -= ((double)compact3(k >> 1) + 0.5) * pg.iy;
    fz -= ((double)compact3(k >> 2) + 0.5) * pg.iz;
  }
  return make_double3(b[0] * fx + b[3] * fy + b[6] * fz, b[1] * fx + b[4] * fy + b[7] * fz,
                      b[2] * fx + b[5] * fy + b[8] * fz);
}
__device__ __forceinline__ double3 packed_pos(const double *b, PackGeo pg, const int *cell, int i, double fx,
                                              double fy, double fz) {
  return packed_pos_key(b, pg, pg.gx ? (unsigned)cell[i] : 0u, fx, fy, fz);
}
__device__ __forceinline__ void store_packed(double4 *pos4, float4 *p4f, int i, double3 r, double w) {
  if (pos4) pos4[i] = make_double4(r.x, r.y, r.z, w);
  if (p4f) p4f[i] = make_float4((float)r.x, (float)r.y, (float)r.z, (float)w);
}


// fine-grid cell of each atom (fractional, folded into [0,1))
__global__ void k_fast_cells(const double *x0, const double *y0, const double *z0, int n, int gx,
                             int gy, int gz, int *cell) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double fx = x0[i] - floor(x0[i]), fy = y0[i] - floor(y0[i]), fz = z0[i] - floor(z0[i]);
  int cx = min(max((int)(fx * gx), 0), gx - 1);
  int cy = min(max((int)(fy * gy), 0), gy - 1);
  int cz = min(max((int)(fz * gz), 0), gz - 1);
  cell[i] = morton3(cx, cy, cz);
}

// Flags (flags[2]) a grid whose cells are narrower than the list radius in
// the current box (perpendicular widths) in any dimension the stencil spans.
__device__ void check_cell_width(const double *b, int gx, int gy, int gz, double rl2, int *flags) {
  const int gg[3] = {gx, gy, gz};
  const double vol = fabs(dev_volume(b));
  for (int d = 0; d < 3; d++) {
    const double *u = b + 3 * ((d + 1) % 3), *w = b + 3 * ((d + 2) % 3);
    const double cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
    const double width = vol / sqrt(cx * cx + cy * cy + cz * cz);
    if (gg[d] >= 3 && width / gg[d] < sqrt(rl2)) atomicMax(&flags[2], 1);
  }
}

// Neighbour list of atom i over its 27 stencil cells, gathering candidates
// from global memory (any grid; rounding images in dimensions with < 3 cells).
__device__ void list_one_atom(int i, const double4 *__restrict__ pos4, const double *atoms, const int *img,
                              const int *cell, const int *cstart, int n, int gx, int gy, int gz,
                              const double (*S)[3], double rl2, int pbc, bool rel, int maxn, int *nbr,
                              int *numn, float *sref, int *flags) {
  const unsigned c = (unsigned)cell[i];
  const int cx = (int)compact3(c), cy = (int)compact3(c >> 1), cz = (int)compact3(c >> 2);
  const int sxr = gx >= 3 ? 1 : 0, syr = gy >= 3 ? 1 : 0, szr = gz >= 3 ? 1 : 0;
  const double4 pi = pos4[i];
  const double sxi = atoms[i] - img[3 * i], syi = atoms[(size_t)n + i] - img[3 * i + 1],
               szi = atoms[2 * (size_t)n + i] - img[3 * i + 2];
  int cnt = 0;
  for (int dz = -szr; dz <= szr; dz++) {
    int kz = cz + dz, wz = 0;
    if (kz < 0) { kz += gz; wz = 1; } else if (kz >= gz) { kz -= gz; wz = -1; }
    for (int dy = -syr; dy <= syr; dy++) {
      int ky = cy + dy, wy = 0;
      if (ky < 0) { ky += gy; wy = 1; } else if (ky >= gy) { ky -= gy; wy = -1; }
      for (int dx = -sxr; dx <= sxr; dx++) {
        int kx = cx + dx, wx = 0;
        if (kx < 0) { kx += gx; wx = 1; } else if (kx >= gx) { kx -= gx; wx = -1; }
        const int cj = morton3(kx, ky, kz);
        for (int j = cstart[cj]; j < cstart[cj + 1]; j++) {
          if (j == i) continue;
          int tx = wx, ty = wy, tz = wz;
          if (!sxr && (pbc & 1)) tx = -(int)round(sxi - (atoms[j] - img[3 * j]));
          if (!syr && (pbc & 2)) ty = -(int)round(syi - (atoms[(size_t)n + j] - img[3 * j + 1]));
          if (!szr && (pbc & 4)) tz = -(int)round(szi - (atoms[2 * (size_t)n + j] - img[3 * j + 2]));
          // open axes take the raw fractional difference: positions are folded
          // by img, so the shift restores img_i - img_j, which is -1, 0 or 1
          // for any pair closer than half the box
          if (!(pbc & 1)) tx = img[3 * i] - img[3 * j];
          if (!(pbc & 2)) ty = img[3 * i + 1] - img[3 * j + 1];
          if (!(pbc & 4)) tz = img[3 * i + 2] - img[3 * j + 2];
          if (tx < -1 || tx > 1 || ty < -1 || ty > 1 || tz < -1 || tz > 1) continue;
          // cell-relative packing: the cell offset (>= 3 cells per dimension)
          const int code = rel ? (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1) : (tx + 1) + 3 * (ty + 1) + 9 * (tz + 1);
          const double4 pj = pos4[j];
          const double rx = pi.x - pj.x + S[code][0];
          const double ry = pi.y - pj.y + S[code][1];
          const double rz = pi.z - pj.z + S[code][2];
          const double r2 = rx * rx + ry * ry + rz * rz;
          if (r2 < rl2) {
            if (cnt < maxn) nbr[nbr_at(cnt, i, n)] = j | (code << NBR_CODE_SHIFT);
            cnt++;
          }
        }
      }
    }
  }
  numn[i] = min(cnt, maxn);
  if (cnt > maxn) atomicMax(&flags[0], cnt);
  for (int k = cnt; k < ((cnt + 3) & ~3) && k < maxn; k++)  // pad the last quad with null entries
    nbr[nbr_at(k, i, n)] = i | (NBR_NULL_CODE << NBR_CODE_SHIFT);
  sref[3 * i + 0] = (float)atoms[i];
  sref[3 * i + 1] = (float)atoms[(size_t)n + i];
  sref[3 * i + 2] = (float)atoms[2 * (size_t)n + i];
}

// Full neighbour list at r_list over the fine grid; atoms are already sorted by
// cell so cell c holds slots [cstart[c], cstart[c+1]).  Positions are the
// image-fixed Cartesian ones of pos4; the image shift of a candidate follows
// from the wrap of its cell (or from rounding in a dimension with one cell).
// nbr is column-major nbr[k*n + i] (index | shift code << 27).
__global__ void __launch_bounds__(MDK_BLOCK) k_full_list(
    const double4 *__restrict__ pos4, const double *atoms, const int *img, const int *cell,
    const int *cstart, int n, int gx, int gy, int gz, const BoxState *bs, double rl2, int pbc,
    int maxn, int *nbr, int *numn, float *sref, int *flags) {
  __shared__ double S[32][3];
  double b[9];
#pragma unroll
  for (int k = 0; k < 9; k++) b[k] = bs->b0[k];
  const PackGeo absolute = {0, 0, 0, 0.0, 0.0, 0.0};
  shift_table(b, absolute, S);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) check_cell_width(b, gx, gy, gz, rl2, flags);
  int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  list_one_atom(i, pos4, atoms, img, cell, cstart, n, gx, gy, gz, S, rl2, pbc, false, maxn, nbr, numn, sref,
                flags);
}

// Full neighbour list by 2x2x2-cell blocks (8 consecutive Morton keys, so one
// contiguous range of atoms), one wave per block, for cell-relative packing.
// The wave stages the atoms of the 4x4x4 cells around the block in LDS --
// positions relative to the centre of the block's first cell, in FP32 (the
// 1e-6 A rounding is immaterial against the skin) -- instead of gathering them
// from global memory.  Each lane then walks the 27 staged cells around its own
// atom's cell (about 155 candidates instead of all 368 staged ones), row by
// row; a partner's code is the offset of its cell.
// Fully periodic boxes with >= 4 cells per dimension (no staged cell occurs
// twice); a block whose neighbourhood exceeds LIST_CAP atoms uses the
// global-gather loop.  blocks[] lists the Morton key / 8 of every block that
// holds grid cells.
#define LIST_CAP 512
__global__ void __launch_bounds__(64) k_full_list_blk(
    const double4 *__restrict__ pos4, const double *atoms, const int *img, const int *cell,
    const int *__restrict__ cstart, const int *blocks, int n, PackGeo pg, const BoxState *bs,
    double rl2, int maxn, int *nbr, int *numn, float *sref, int *flags) {
  __shared__ double S[32][3];
  __shared__ float4 cpos[LIST_CAP];
  __shared__ int hoff[64], hcnt[64];  // staged cell h: candidates [hoff, hoff + hcnt)
  const int lane = threadIdx.x, gx = pg.gx, gy = pg.gy, gz = pg.gz;
  const double *b = bs->b0;
  shift_table(b, pg, S);
  if (blockIdx.x == 0 && lane == 0) check_cell_width(b, gx, gy, gz, rl2, flags);
  const unsigned base = (unsigned)blocks[xcd_tile()] << 3;
  const int x0 = 2 * (int)compact3(base >> 3), y0 = 2 * (int)compact3(base >> 4), z0 = 2 * (int)compact3(base >> 5);
  const int a0 = cstart[base], a1 = cstart[base + 8];
  if (a1 <= a0) return;  // uniform: no atoms in this block
  // staged cell h = hx + 4 hy + 16 hz at (x0-1+hx, y0-1+hy, z0-1+hz): one per lane
  const int hx = lane & 3, hy = (lane >> 2) & 3, hz = lane >> 4;
  int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
  if (x < 0) x += gx; else if (x >= gx) x -= gx;
  if (y < 0) y += gy; else if (y >= gy) y -= gy;
  if (z < 0) z += gz; else if (z >= gz) z -= gz;
  const int key = morton3(x, y, z);
  const int cs = cstart[key], cnt = cstart[key + 1] - cs;
  int off = cnt;  // exclusive prefix sum over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(off, o, 64);
    if (lane >= o) off += u;
  }
  const int total = __shfl(off, 63, 64);
  off -= cnt;
  hoff[lane] = off;
  hcnt[lane] = cnt;
  __syncthreads();  // shift table, cell ranges
  if (total > LIST_CAP) {
    for (int i = a0 + lane; i < a1; i += 64)
      list_one_atom(i, pos4, atoms, img, cell, cstart, n, gx, gy, gz, S, rl2, 7, true, maxn, nbr, numn, sref,
                    flags);
    return;
  }
  // centre of the staged cell from the centre of cell (x0,y0,z0): B0 (h - 1) / g
  const double fx = (hx - 1) * pg.ix, fy = (hy - 1) * pg.iy, fz = (hz - 1) * pg.iz;
  const double vx = b[0] * fx + b[3] * fy + b[6] * fz, vy = b[1] * fx + b[4] * fy + b[7] * fz,
               vz = b[2] * fx + b[5] * fy + b[8] * fz;
  for (int k = 0; k < cnt; k++) {
    const double4 p = pos4[cs + k];
    cpos[off + k] = make_float4((float)(p.x + vx), (float)(p.y + vy), (float)(p.z + vz), __int_as_float(cs + k));
  }
  __syncthreads();
  const float rl2f = (float)rl2;
  for (int i = a0 + lane; i < a1; i += 64) {
    const unsigned ck = (unsigned)cell[i];
    const int ax = (int)compact3(ck) - x0, ay = (int)compact3(ck >> 1) - y0, az = (int)compact3(ck >> 2) - z0;
    const double4 pd = pos4[i];
    const double ex = ax * pg.ix, ey = ay * pg.iy, ez = az * pg.iz;
    const float px = (float)(pd.x + (b[0] * ex + b[3] * ey + b[6] * ez)),
                py = (float)(pd.y + (b[1] * ex + b[4] * ey + b[7] * ez)),
                pz = (float)(pd.z + (b[2] * ex + b[5] * ey + b[8] * ez));
    const int h0 = ax + 4 * ay + 16 * az;  // staged cell at offset (-1,-1,-1) from the atom's
    int m = 0;
    // the three cells of a row along x are consecutive staged cells, so one
    // loop walks their candidates (about 17, instead of three loops of about
    // 6 whose lengths differ from lane to lane); a partner's cell offset, and
    // so its code, follows from where it lies in the row
    for (int dz = 0; dz < 3; dz++)
      for (int dy = 0; dy < 3; dy++) {
        const int h = h0 + 4 * dy + 16 * dz;  // the row's cell at dx = 0
        const int lo = hoff[h], m1 = hoff[h + 1], m2 = hoff[h + 2], hi = m2 + hcnt[h + 2];
        const int tagb = 3 * dy + 9 * dz;
        auto test = [&](const float4 q, int c) {
          const float rx = px - q.x, ry = py - q.y, rz = pz - q.z;
          const int j = __float_as_int(q.w);
          if (rx * rx + ry * ry + rz * rz < rl2f && j != i) {
            const int code = tagb + (c >= m1 ? 1 : 0) + (c >= m2 ? 1 : 0);
            if (m < maxn) nbr[nbr_at(m, i, n)] = j | (code << NBR_CODE_SHIFT);
            m++;
          }
        };
        // four candidates per trip: their LDS reads are in flight together;
        // the tests run in candidate order, so the list is the same
        int c = lo;
        for (; c + 4 <= hi; c += 4) {
          const float4 q0 = cpos[c], q1 = cpos[c + 1], q2 = cpos[c + 2], q3 = cpos[c + 3];
          test(q0, c);
          test(q1, c + 1);
          test(q2, c + 2);
          test(q3, c + 3);
        }
        for (; c < hi; c++) test(cpos[c], c);
      }
    numn[i] = min(m, maxn);
    if (m > maxn) atomicMax(&flags[0], m);
    for (int k = m; k < ((m + 3) & ~3) && k < maxn; k++)  // pad the last quad with null entries
      nbr[nbr_at(k, i, n)] = i | (NBR_NULL_CODE << NBR_CODE_SHIFT);
    sref[3 * i + 0] = (float)atoms[i];
    sref[3 * i + 1] = (float)atoms[(size_t)n + i];
    sref[3 * i + 2] = (float)atoms[2 * (size_t)n + i];
  }
}

// ---------------------------------------------------------------------------
// MD-step kernels
// ---------------------------------------------------------------------------
// Box part of the drift, by one workgroup: copies the box state into the other
// buffer with b0 += b1 + b2 and b1 += b2 (src/quench_m.F90:574-589; one field
// per thread).
__device__ void drift_box(const BoxState *bin, BoxState *bout, int ivol) {
#pragma clang fp contract(off)
  const int nw = (int)(sizeof(BoxState) / sizeof(double));
  const double *src = (const double *)bin;
  double *dst = (double *)bout;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    const int f = w - (int)(offsetof(BoxState, b0) / sizeof(double));
    const int g = w - (int)(offsetof(BoxState, b1) / sizeof(double));
    double val = src[w];
    if (f >= 0 && f < 9) {  // b0 += b1 + b2 (component-wise, as the reference's two updates)
      const bool upd = ivol < 1 || (ivol == 4 && f == 8);
      if (upd) val = bin->b0[f] + (bin->b1[f] + bin->b2[f]);
    } else if (g >= 0 && g < 9) {
      const bool upd = ivol < 1 || (ivol == 4 && g == 8);
      if (upd) val = bin->b1[g] + bin->b2[g];
    }
    dst[w] = val;
  }
}

// velocityverlet drift (src/quench_m.F90:574-589): x0 += x1 + x2 (also into the
// packed copy); block 0 copies the box state into the other buffer, advancing
// b1 and b0 there (one field per thread).
__global__ void k_drift(double *__restrict__ atoms, double4 *__restrict__ pos4, float4 *__restrict__ p4f,
                        const int *__restrict__ img, const int *__restrict__ cell, PackGeo pg, int n,
                        const BoxState *__restrict__ bin, BoxState *__restrict__ bout, int ivol,
                        const int *__restrict__ zat, const double *__restrict__ spec, double *__restrict__ kin) {
#pragma clang fp contract(off)
  int i = xcd_tile() * blockDim.x + threadIdx.x;
  double kv[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    double *x0 = atoms, *y0 = atoms + n, *z0 = atoms + 2 * (size_t)n;
    const double *x1 = atoms + 3 * (size_t)n, *y1 = atoms + 4 * (size_t)n, *z1 = atoms + 5 * (size_t)n;
    const double *x2 = atoms + 6 * (size_t)n, *y2 = atoms + 7 * (size_t)n, *z2 = atoms + 8 * (size_t)n;
    const double xn = x0[i] + (x1[i] + x2[i]);
    const double yn = y0[i] + (y1[i] + y2[i]);
    const double zn = z0[i] + (z1[i] + z2[i]);
    x0[i] = xn;
    y0[i] = yn;
    z0[i] = zn;
    // the box after this step's update (computed redundantly per thread)
    double b[9];
    for (int k = 0; k < 9; k++) {
      const bool upd = ivol < 1 || (ivol == 4 && k == 8);
      b[k] = upd ? bin->b0[k] + (bin->b1[k] + bin->b2[k]) : bin->b0[k];
    }
    const double3 r = packed_pos(b, pg, cell, i, xn - img[3 * i], yn - img[3 * i + 1], zn - img[3 * i + 2]);
    store_packed(pos4, p4f, i, r, 0.0);  // F' is stored by the density pass, which always runs next
    const int z = zat[i];
    if (z != 0) kin_terms(b, x1[i], y1[i], z1[i], spec ? spec[2 * z] : atoms[(size_t)F_MASS * n + i], kv);
  }
  if (blockIdx.x == 0) drift_box(bin, bout, ivol);
  block_store<9, 0>(kv, kin);
}

// predic (src/quench_m.F90:636-671): x0 += x1 + x2 + x3, x1 += 2 x2 + 3 x3,
// x2 += 3 x3, the box likewise for ivol 0 (all) or 4 (z only), packed
// positions with the predicted box; block 0 writes the predicted box state
// into bout.  primed == 0 leaves everything as it is (see mdk_predic).
__device__ __forceinline__ void predict_box(const BoxState *bin, double *b0, double *b1, double *b2, int ivol,
                                            int primed, int k) {
#pragma clang fp contract(off)
  const bool upd = primed && (ivol == 0 || (ivol == 4 && k == 8));
  b0[k] = upd ? bin->b0[k] + bin->b1[k] + bin->b2[k] + bin->b3[k] : bin->b0[k];
  b1[k] = upd ? bin->b1[k] + 2.0 * bin->b2[k] + 3.0 * bin->b3[k] : bin->b1[k];
  b2[k] = upd ? bin->b2[k] + 3.0 * bin->b3[k] : bin->b2[k];
}
__global__ void k_predic(double *__restrict__ atoms, double4 *__restrict__ pos4, float4 *__restrict__ p4f,
                         const int *__restrict__ img, const int *__restrict__ cell, PackGeo pg, int n,
                         const BoxState *__restrict__ bin, BoxState *__restrict__ bout, int ivol, int primed,
                         const int *__restrict__ zat, const double *__restrict__ spec, double *__restrict__ kin) {
#pragma clang fp contract(off)
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  double kv[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    double *p = atoms;
    double x[3], v[3];
    for (int d = 0; d < 3; d++) {
      double *q0 = p + (size_t)(F_X0 + d) * n, *q1 = p + (size_t)(F_X1 + d) * n, *q2 = p + (size_t)(F_X2 + d) * n;
      const double *q3 = p + (size_t)(F_X3 + d) * n;
      double v0 = q0[i];
      if (primed) {
        const double v1 = q1[i], v2 = q2[i], v3 = q3[i];
        v0 = v0 + v1 + v2 + v3;
        q0[i] = v0;
        q1[i] = v1 + 2.0 * v2 + 3.0 * v3;
        q2[i] = v2 + 3.0 * v3;
      }
      x[d] = v0;
      v[d] = q1[i];
    }
    double b[9], bd1[9], bd2[9];
    for (int k = 0; k < 9; k++) predict_box(bin, b, bd1, bd2, ivol, primed, k);
    const double3 r = packed_pos(b, pg, cell, i, x[0] - img[3 * i], x[1] - img[3 * i + 1], x[2] - img[3 * i + 2]);
    store_packed(pos4, p4f, i, r, 0.0);  // F' is stored by the density pass, which always runs next
    const int z = zat[i];
    if (z != 0) kin_terms(b, v[0], v[1], v[2], spec ? spec[2 * z] : atoms[(size_t)F_MASS * n + i], kv);
  }
  if (blockIdx.x == 0) {  // the whole box state into bout, one field per thread
    const int nw = (int)(sizeof(BoxState) / sizeof(double));
    const double *src = (const double *)bin;
    double *dst = (double *)bout;
    for (int w = threadIdx.x; w < nw; w += blockDim.x) dst[w] = src[w];
    __syncthreads();
    if (threadIdx.x < 9) predict_box(bin, bout->b0, bout->b1, bout->b2, ivol, primed, threadIdx.x);
  }
  block_store<9, 0>(kv, kin);
}

// kinten's partial sums for the velocities and box as they are (a force
// evaluation that no drift, predictor or fused half kick has prepared)
__global__ void __launch_bounds__(MDK_BLOCK) k_kin(const double *__restrict__ atoms, const int *__restrict__ zat,
                                                   int n, const BoxState *__restrict__ bs,
                                                   const double *__restrict__ spec, double *__restrict__ kin) {
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  double kv[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const int z = zat[i];
    double b[9];
    for (int k = 0; k < 9; k++) b[k] = bs->b0[k];
    if (z != 0)
      kin_terms(b, atoms[(size_t)F_X1 * n + i], atoms[(size_t)F_Y1 * n + i], atoms[(size_t)F_Z1 * n + i],
                spec ? spec[2 * z] : atoms[(size_t)F_MASS * n + i], kv);
  }
  block_store<9, 0>(kv, kin);
}

// correc (src/quench_m.F90:682-759).  The coefficients are the reference's
// REAL(4) parameters (no d0 suffix): 1/6, 5/6, 1, 1/3 rounded to float.
// cc = del2 f - tgid x1 / 2 - x2 per atom (vacancies skipped); the box (ivol
// < 1: all, 4: z only) with cc = bdel2 fb - b2, in place by nine threads of
// block 0 (no other thread reads the box's b fields).
__global__ void k_correc(double *__restrict__ atoms, const int *__restrict__ zat, int n, BoxState *bs, int ivol,
                         double bdel2, const double *__restrict__ spec) {
#pragma clang fp contract(off)
  const double ct0 = (double)0.166666666666666667f, ct1 = (double)0.833333333333333333f;
  const double ct2 = (double)1.000000000000000000f, ct3 = (double)0.333333333333333333f;
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  if (i < n && zat[i] != 0) {
    const double *g = bs->tgid;
    const int z = zat[i];
    const double d2 = spec ? spec[2 * z + 1] : atoms[(size_t)F_DEL2 * n + i];
    const double x1 = atoms[(size_t)F_X1 * n + i], y1 = atoms[(size_t)F_Y1 * n + i], z1 = atoms[(size_t)F_Z1 * n + i];
    const double s[3] = {TT(g, 1, 1) * x1 + TT(g, 1, 2) * y1 + TT(g, 1, 3) * z1,
                         TT(g, 2, 1) * x1 + TT(g, 2, 2) * y1 + TT(g, 2, 3) * z1,
                         TT(g, 3, 1) * x1 + TT(g, 3, 2) * y1 + TT(g, 3, 3) * z1};
    for (int d = 0; d < 3; d++) {
      double *q0 = atoms + (size_t)(F_X0 + d) * n, *q1 = atoms + (size_t)(F_X1 + d) * n;
      double *q2 = atoms + (size_t)(F_X2 + d) * n, *q3 = atoms + (size_t)(F_X3 + d) * n;
      const double f = atoms[(size_t)(F_FX + d) * n + i];
      const double cc = d2 * f - 0.5 * s[d] - q2[i];
      q0[i] = q0[i] + cc * ct0;
      q1[i] = q1[i] + cc * ct1;
      q2[i] = q2[i] + cc * ct2;
      q3[i] = q3[i] + cc * ct3;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 9) {
    const int k = threadIdx.x;
    if (ivol < 1 || (ivol == 4 && k == 8)) {
      const double cc = bdel2 * bs->fb[k] - bs->b2[k];
      bs->b0[k] = bs->b0[k] + ct0 * cc;
      bs->b1[k] = bs->b1[k] + ct1 * cc;
      bs->b2[k] = bs->b2[k] + ct2 * cc;
      bs->b3[k] = bs->b3[k] + ct3 * cc;
    }
  }
}

// geometry of pair (i,j): fractional separation (minimum image) and Cartesian
template <bool EXACT>
__device__ __forceinline__ double pair_geom(const double *b, int pbc, double xi, double yi, double zi,
                                            double xj, double yj, double zj, double &dx, double &dy,
                                            double &dz, double &rx, double &ry, double &rz) {
  dx = xi - xj;
  dy = yi - yj;
  dz = zi - zj;
  min_image(pbc, dx, dy, dz);
  rx = b[0] * dx + b[3] * dy + b[6] * dz;
  ry = b[1] * dx + b[4] * dy + b[7] * dz;
  rz = b[2] * dx + b[5] * dy + b[8] * dz;
  return rx * rx + ry * ry + rz * rz;
}

// r and 1/r from r^2: hardware reciprocal square root refined by Newton steps
// (two in FP64, one in FP32) instead of a square root followed by a division.
__device__ __forceinline__ void r_and_inv(double r2, double &r, double &ri) {
  double y = __builtin_amdgcn_rsq(r2);
  const double h = 0.5 * r2;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  ri = y;
  r = r2 * y;
}
__device__ __forceinline__ void r_and_inv(float r2, float &r, float &ri) {
  float y = __builtin_amdgcn_rsqf(r2);
  y = y * (1.5f - 0.5f * r2 * y * y);
  ri = y;
  r = r2 * y;
}

// Pair geometry: Cartesian minimum-image separation R'_i - R'_j + B0 shift (FP64)
// and its square in the arithmetic type T (the FP32 form rounds the components).
template <typename T>
__device__ __forceinline__ T pair_r2(const double (*S)[3], const double4 &pi, const double4 &pj,
                                     int code, T &rx, T &ry, T &rz) {
  rx = (T)(pi.x - pj.x + S[code][0]);
  ry = (T)(pi.y - pj.y + S[code][1]);
  rz = (T)(pi.z - pj.z + S[code][2]);
  return rx * rx + ry * ry + rz * rz;
}

// The same with the FP64 components kept (for FP64 accumulation of forces and
// virial); the arithmetic type T sees them rounded.
template <typename T>
__device__ __forceinline__ T pair_sep(const double (*S)[3], const double4 &pi, const double4 &pj,
                                      int code, double &dx, double &dy, double &dz, T &rx, T &ry,
                                      T &rz) {
  dx = pi.x - pj.x + S[code][0];
  dy = pi.y - pj.y + S[code][1];
  dz = pi.z - pj.z + S[code][2];
  rx = (T)dx;
  ry = (T)dy;
  rz = (T)dz;
  return rx * rx + ry * ry + rz * rz;
}


// Packed positions R (absolute or cell-relative, see the entry format) and the
// current F', for one 32-byte gather per neighbour (16 bytes for the FP32
// copy p4f, when given).  At a rebuild (with_img) it also fixes img = floor(s)
// and the list cell of each atom on the (gx, gy, gz) grid, as k_fast_cells.
__global__ void k_pack(const double *atoms, int *img, const BoxState *bs, int n, int with_img,
                       double4 *pos4, float4 *p4f, int *cell, int gx, int gy, int gz, PackGeo pg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = atoms[i], y = atoms[(size_t)n + i], z = atoms[2 * (size_t)n + i];
  if (with_img) {
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    img[3 * i] = (int)fx;
    img[3 * i + 1] = (int)fy;
    img[3 * i + 2] = (int)fz;
    const int cx = min(max((int)((x - fx) * gx), 0), gx - 1);
    const int cy = min(max((int)((y - fy) * gy), 0), gy - 1);
    const int cz = min(max((int)((z - fz) * gz), 0), gz - 1);
    cell[i] = morton3(cx, cy, cz);
  }
  const double3 r = packed_pos(bs->b0, pg, cell, i, x - img[3 * i], y - img[3 * i + 1], z - img[3 * i + 2]);
  store_packed(pos4, p4f, i, r, atoms[(size_t)F_DAFRHO * n + i]);
}

// Neighbour j's packed position and F': the FP64 record, or (P32) the FP32
// copy of a cell-relative packing, widened exactly.
template <bool P32>
__device__ __forceinline__ double4 partner(const double4 *__restrict__ pos4, const float4 *__restrict__ p4f,
                                           unsigned j) {
  if constexpr (P32) {
    const float4 q = p4f[j];
    return make_double4(q.x, q.y, q.z, q.w);
  } else {
    return pos4[j];
  }
}

// Per-atom end of the density pass: rho, and with do_embed the embedding
// energy F(rho), F'(rho) (also into the packed positions for the force pass);
// returns atom i's embedding energy (0 without do_embed).
__device__ __forceinline__ double density_atom(double *atoms, double4 *pos4, float4 *p4f, int n, int i,
                                               double acc, const PotDev &P, int do_embed) {
  double *rho = atoms + (size_t)F_RHO * n, *af = atoms + (size_t)F_AFRHO * n, *daf = atoms + (size_t)F_DAFRHO * n;
  rho[i] = acc;
  if (!do_embed) return 0.0;
  double dfi = daf[i], pe = 0.0;
  if (acc > 1e-40) {  // src/potential_m.F90:161 leaves stale values below
    double f;
    mdk_embed_f(P, acc, f, dfi);
    af[i] = f;
    daf[i] = dfi;
    pe = f;
  }
  if (pos4) pos4[i].w = dfi;
  if (p4f) p4f[i].w = (float)dfi;
  return pe;
}
// ... and the block's embedding-energy partial
__device__ __forceinline__ void density_finish(double *atoms, double4 *pos4, float4 *p4f, int n, int i,
                                               double acc, const PotDev &P, int do_embed, double *part) {
  double pe[1] = {i < n ? density_atom(atoms, pos4, p4f, n, i, acc, P, do_embed) : 0.0};
  if (do_embed) block_store<1, 0>(pe, part);
}

// Appends entry e as slot m of atom i's list through a register quad; every
// fourth entry stores the quad with one 16-byte write.
__device__ __forceinline__ void quad_put(int *nbr, int i, int n, int e, int &m, int4 &qd) {
  const int sl = m & 3;
  qd.x = sl == 0 ? e : qd.x;
  qd.y = sl == 1 ? e : qd.y;
  qd.z = sl == 2 ? e : qd.z;
  qd.w = sl == 3 ? e : qd.w;
  if (sl == 3) *(int4 *)(nbr + nbr_at(m - 3, i, n)) = qd;
  m++;
}

// Density pass (rhoset, src/dynamics_m.F90:497-579) fused with embed
// (src/potential_m.F90:148-169): rho_i = sum_j phi(r_ij) over the full list
// one lane per atom; F, F' per atom; per-block embedding energy.  F' is
// also stored in pos4[i].w for the force pass.
template <typename T, int POTM, bool P32>
__global__ void __launch_bounds__(MDK_BLOCK) k_density(double *atoms, double4 *pos4, float4 *p4f, int n,
                                                       const int *__restrict__ nbr,
                                                       const int *__restrict__ numn,
                                                       const BoxState *bs, PackGeo pg, PotDev P,
                                                       int do_embed, double *part, int *nact) {
  __shared__ double S[32][3];
  shift_table(bs->b0, pg, S);
  __syncthreads();
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  double acc = 0.0;  // per-atom density accumulated in FP64 in both forms
  // list slots up to the end of the last quad holding a partner in range: with
  // a margin of 1e-4 A the force pass's own r of every later pair is beyond the
  // outermost knot, where all pair terms are exactly zero
  const T rc = (T)P.r_int + (T)1e-4, rc2 = rc * rc;
  int na = 0;
  if (i < n) {
    const double4 pi = pos4[i];
    const int nn = numn[i];
    // one 16-byte index load and four independent gathers per quad (quads
    // are padded with null entries)
    for (int q = 0; 4 * q < nn; q++) {
      const int4 e4 = *(const int4 *)(nbr + (((size_t)q * n + i) << 2));
      const unsigned e[4] = {(unsigned)e4.x, (unsigned)e4.y, (unsigned)e4.z, (unsigned)e4.w};
      double4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = partner<P32>(pos4, p4f, e[u] & NBR_IDX_MASK);
      T mn = rc2;  // one min chain per quad keeps the register footprint of the loop
#pragma unroll
      for (int u = 0; u < 4; u++) {
        T rx, ry, rz;
        const T r2 = pair_r2<T>(S, pi, pj[u], (int)(e[u] >> NBR_CODE_SHIFT), rx, ry, rz);
        T r, ri;
        r_and_inv(r2, r, ri);
        acc += (double)mdk_pair_phi<T, POTM>(P, r);
        mn = fmin(mn, r2);
      }
      na = mn < rc2 ? 4 * q + 4 : na;
    }
    nact[i] = na;
  }
  density_finish(atoms, pos4, p4f, n, i, acc, P, do_embed, part);
}

// Density pass of the first step on a new list: as k_density, and it also
// reorders each atom's list in place -- the partners in range now (r^2 < rc2,
// the test k_density uses for nact) first, in list order, then the others --
// so that the force pass of this and the following steps stops after the
// in-range part.  Out-of-range entries wait in LDS (DENS_OBUF per lane, more
// stay in place, which costs only efficiency).  Each completed quad is written
// after its slots have been read (an atom's list belongs to its lane alone).
#define DENS_OBUF 16
template <typename T, int POTM, bool P32>
__global__ void __launch_bounds__(MDK_BLOCK) k_density_part(double *atoms, double4 *pos4, float4 *p4f, int n,
                                                            int *nbr, const int *__restrict__ numn,
                                                            const BoxState *bs, PackGeo pg, PotDev P,
                                                            int do_embed, double *part, int *nact) {
  __shared__ double S[32][3];
  __shared__ int obuf[DENS_OBUF][MDK_BLOCK];
  shift_table(bs->b0, pg, S);
  __syncthreads();
  const T rc = (T)P.r_int + (T)1e-4, rc2 = rc * rc;
  const int t = threadIdx.x, i = xcd_tile() * blockDim.x + t;
  double acc = 0.0;
  if (i < n) {
    const double4 pi = pos4[i];
    const int nn = numn[i];
    int m = 0, no = 0, last = 0;  // entries written back, entries held, end of the last in-range one
    int4 qd = make_int4(0, 0, 0, 0);
    for (int q = 0; 4 * q < nn; q++) {
      const int4 e4 = *(const int4 *)(nbr + (((size_t)q * n + i) << 2));
      const int e[4] = {e4.x, e4.y, e4.z, e4.w};
      double4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = partner<P32>(pos4, p4f, (unsigned)e[u] & NBR_IDX_MASK);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        T rx, ry, rz;
        const T r2 = pair_r2<T>(S, pi, pj[u], (int)((unsigned)e[u] >> NBR_CODE_SHIFT), rx, ry, rz);
        T r, ri;
        r_and_inv(r2, r, ri);
        acc += (double)mdk_pair_phi<T, POTM>(P, r);
        if (4 * q + u < nn) {
          const bool in = r2 < rc2;
          if (in || no == DENS_OBUF) {
            quad_put(nbr, i, n, e[u], m, qd);
            if (in) last = m;
          } else {
            obuf[no++][t] = e[u];
          }
        }
      }
    }
    for (int u = 0; u < no; u++) quad_put(nbr, i, n, obuf[u][t], m, qd);
    if (m & 3) {  // the last quad's null entries
      const int nul = i | (NBR_NULL_CODE << NBR_CODE_SHIFT), sl = m & 3;
      qd.y = sl <= 1 ? nul : qd.y;
      qd.z = sl <= 2 ? nul : qd.z;
      qd.w = nul;
      *(int4 *)(nbr + nbr_at(m & ~3, i, n)) = qd;
    }
    nact[i] = (last + 3) & ~3;
  }
  density_finish(atoms, pos4, p4f, n, i, acc, P, do_embed, part);
}

enum { FIN_FORCE = 0, FIN_VV = 1, FIN_VV_THERM = 2, FIN_INIT = 3 };

// update_therm_avs box part (src/dynamics_m.F90:672-700)
__device__ void therm_box(BoxState &s, int ivol) {
#pragma clang fp contract(off)
  for (int k = 0; k < 9; k++) { s.tg[k] = 0.0; s.tgdot[k] = 0.0; }
  for (int i = 1; i <= 3; i++)
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++) {
        TT(s.tg, i, j) = TT(s.tg, i, j) + TT(s.b0, k, i) * TT(s.b0, k, j);
        if (ivol != 1)
          TT(s.tgdot, i, j) = TT(s.tgdot, i, j) + TT(s.b1, k, i) * TT(s.b0, k, j) + TT(s.b0, k, i) * TT(s.b1, k, j);
      }
  dev_invert(s.tg, s.tginv, &s.status);
  dev_matmul(s.tginv, s.tgdot, s.tgid);
  s.vol = dev_volume(s.b0);
  const double *b = s.b0, v = s.vol;
  TT(s.binv, 1, 1) = (TT(b, 2, 2) * TT(b, 3, 3) - TT(b, 2, 3) * TT(b, 3, 2)) / v;  // xx
  TT(s.binv, 1, 2) = (TT(b, 2, 3) * TT(b, 3, 1) - TT(b, 3, 3) * TT(b, 2, 1)) / v;  // xy
  TT(s.binv, 1, 3) = (TT(b, 2, 1) * TT(b, 3, 2) - TT(b, 3, 1) * TT(b, 2, 2)) / v;  // xz
  TT(s.binv, 2, 1) = (TT(b, 3, 2) * TT(b, 1, 3) - TT(b, 1, 2) * TT(b, 3, 3)) / v;  // yx
  TT(s.binv, 2, 2) = (TT(b, 1, 1) * TT(b, 3, 3) - TT(b, 1, 3) * TT(b, 3, 1)) / v;  // yy
  TT(s.binv, 2, 3) = (TT(b, 3, 1) * TT(b, 1, 2) - TT(b, 1, 1) * TT(b, 3, 2)) / v;  // yz
  TT(s.binv, 3, 1) = (TT(b, 1, 2) * TT(b, 2, 3) - TT(b, 2, 2) * TT(b, 1, 3)) / v;  // zx
  TT(s.binv, 3, 2) = (TT(b, 1, 3) * TT(b, 2, 1) - TT(b, 1, 1) * TT(b, 2, 3)) / v;  // zy
  TT(s.binv, 3, 3) = (TT(b, 1, 1) * TT(b, 2, 2) - TT(b, 1, 2) * TT(b, 2, 1)) / v;  // zz
}

// Stage-and-finish reduction for the finishing kernels, launched as p =
// ceil(nb / 64) blocks of one wave (MDK_FIN_BLOCK).  Block b folds per-block
// partials [64 b, 64 b + 64) lane-wise and across the wave in a fixed order;
// lane 0 writes the K results to stage[k*p + b] and increments the arrival
// counter with one agent-scope release/acquire, so each block makes a single
// L2 write-back and invalidate.  The last block to arrive folds the p stage
// columns the same way into out[] (every lane receives them), resets the
// counter and returns true; the others return false.
#define MDK_FIN_BLOCK 64
template <int K, int KM>
__device__ __forceinline__ void fold_columns(const double *part, int nb, int b0, int b1, double (&v)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = (k < K - KM) ? 0.0 : -HUGE_VAL;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += MDK_FIN_BLOCK) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double x = part[(size_t)k * nb + b];
      v[k] = (k < K - KM) ? v[k] + x : fmax(v[k], x);
    }
  }
  wave_reduce_all<K, KM>(v);
}

template <int K, int KM>
__device__ bool stage_and_finish(const double *part, int nb, double *stage, unsigned *cnt, double (&out)[K]) {
  const int p = gridDim.x, b0 = blockIdx.x * 64;
  double v[K];
  fold_columns<K, KM>(part, nb, b0, min(nb, b0 + 64), v);
  int last = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) stage[(size_t)k * p + blockIdx.x] = v[k];
    const unsigned prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = prev + 1 == (unsigned)p;
  }
  // one wave: lane 0's verdict to all lanes, which are past lane 0's acquire
  if (!__builtin_amdgcn_readfirstlane(last)) return false;
  fold_columns<K, KM>(stage, p, 0, p, out);
  if (threadIdx.x == 0) __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// stage_and_finish over two partial arrays of nb blocks in one pass (part2 may
// be null): the deferred half-kick sums ride on force()'s arrival counter.
template <int K, int KM, int K2, int KM2>
__device__ bool stage_and_finish2(const double *part, const double *part2, int nb, double *stage,
                                  double *stage2, unsigned *cnt, double (&out)[K], double (&out2)[K2]) {
  const int p = gridDim.x, b0 = blockIdx.x * 64;
  double v[K], v2[K2];
  fold_columns<K, KM>(part, nb, b0, min(nb, b0 + 64), v);
  if (part2) fold_columns<K2, KM2>(part2, nb, b0, min(nb, b0 + 64), v2);
  int last = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) stage[(size_t)k * p + blockIdx.x] = v[k];
    if (part2) {
#pragma unroll
      for (int k = 0; k < K2; k++) stage2[(size_t)k * p + blockIdx.x] = v2[k];
    }
    const unsigned prev = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = prev + 1 == (unsigned)p;
  }
  if (!__builtin_amdgcn_readfirstlane(last)) return false;
  fold_columns<K, KM>(stage, p, 0, p, out);
  if (part2) fold_columns<K2, KM2>(stage2, p, 0, p, out2);
  if (threadIdx.x == 0) __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// One-workgroup reduction of the per-block partials for the finishing
// kernels (MDK_FOLD_BLOCK threads): thread t folds columns t, t + 1024, ...
// in order, then the lanes of each wave and the waves in index order; thread
// 0 receives the K totals.  No atomics and a single memory round trip.
#define MDK_FOLD_BLOCK 1024
template <int K, int KM>
__device__ __forceinline__ void fold_all(const double *part, int nb, double (&out)[K]) {
  __shared__ double sh[MDK_FOLD_BLOCK / 64][K];
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = (k < K - KM) ? 0.0 : -HUGE_VAL;
  for (int b = (int)threadIdx.x; b < nb; b += MDK_FOLD_BLOCK) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double x = part[(size_t)k * nb + b];
      v[k] = (k < K - KM) ? v[k] + x : fmax(v[k], x);
    }
  }
  wave_reduce_all<K, KM>(v);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) sh[threadIdx.x >> 6][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double s = sh[0][k];
      for (int q = 1; q < MDK_FOLD_BLOCK / 64; q++) s = (k < K - KM) ? s + sh[q][k] : fmax(s, sh[q][k]);
      out[k] = s;
    }
  }
}

// Arguments of the finishing kernels of force() and of the half kick.
struct FinForce {
  double *stage;   // ceil(nb / 64) x NPART_F first-stage sums
  unsigned *cnt;   // arrival counter, zero between launches
  BoxState *bs;
  int mode, ivol;
  double press, bdel2, dt;
  // the previous half kick's sums, when its finishing was deferred to this
  // kernel (the MD loop): its partials, their first-stage sums, its finish
  const double *kpart;
  double *kstage;
  FinKick kick;
};
__device__ void finish_kick(const double (&v)[NPART_KICK], const FinKick &F);

// Finishing kernel of force(), ceil(nb / 64) blocks of one wave: the NPART_F
// columns of per-block partials (force pass, kinten, embedding energy) are
// reduced by stage_and_finish (spread over many CUs: seventeen columns are
// too much data for one), then the last block completes force(): pe, tp, tk,
// tc, fb (dynamics_m.F90:365-394), then per mode the velocity-Verlet box kick (quench_m.F90:595-604), the initial
// tgid/b2 (moldin.F90:452-477) or the update_therm_avs box part
// (dynamics_m.F90:672-700).  Threads 0..8 each own one tensor element (i,j);
// phases are separated by barriers; only changed fields are written back.
__global__ void __launch_bounds__(MDK_FIN_BLOCK) k_finalize_force(const double *part, int nb, FinForce F) {
#pragma clang fp contract(off)  // therm_box's rounding: the step API and the fused step agree bit for bit
  // small matrices live in LDS: the nine tensor threads index them with their
  // own (ii,jj), which private arrays would turn into scratch accesses.  The
  // box fields are fetched before the partial sums so the two memory round
  // trips overlap.
  __shared__ double b0[9], b1[9], b2[9], tgid0[9], tg[9], tgd[9], tgi[9], sv[NPART_F], sbi[9];
  BoxState *bs = F.bs;
  const int mode = F.mode, ivol = F.ivol;
  const double press = F.press, bdel2 = F.bdel2, dt = F.dt;
  const int t = threadIdx.x;
  const int ii = t % 3 + 1, jj = t / 3 + 1;  // element (ii,jj) of thread t < 9
  if (t < 9) {
    b0[t] = bs->b0[t];
    b1[t] = bs->b1[t];
    b2[t] = bs->b2[t];
    tgid0[t] = bs->tgid[t];
  }
  double v[NPART_F], kv[NPART_KICK];
  if (!stage_and_finish2<NPART_F, 0, NPART_KICK, 2>(part, F.kpart, nb, F.stage, F.kstage, F.cnt, v, kv)) return;
  // the previous step's kick sums (its box copy is untouched until this
  // step's kick): only fields that this kernel does not write
  if (F.kpart && t == 0) finish_kick(kv, F.kick);
  if (t == 0)
    for (int k = 0; k < NPART_F; k++) sv[k] = v[k];
  __syncthreads();
  if (t == 0) {
    bs->pe_emb = v[PART_EMB];
    bs->pe = v[PART_EMB] + v[0];
    bs->vol = dev_volume(b0);
    inv3(b0, sbi);
  }
  __syncthreads();
  if (t < 9) {
    // W = 1/2 sum r r^T fp (symmetric; sv[1..6] = xx yy zz xy xz yz) ->
    // tp(a,c) = sum_k W(a,k) B0^-1(c,k)
    double tpv = 0.0;
    for (int k = 1; k <= 3; k++) {
      const int w = ii == k ? ii : (ii + k == 3 ? 4 : (ii + k == 4 ? 5 : 6));
      tpv += sv[w] * TT(sbi, jj, k);
    }
    const double tkv = sv[PART_KIN + t] / (dt * dt);
    // tc = cofactor matrix of b0 (dynamics_m.F90:365-373)
    const int i1 = ii % 3 + 1, i2 = (ii + 1) % 3 + 1, j1 = jj % 3 + 1, j2 = (jj + 1) % 3 + 1;
    const double tcv = TT(b0, i1, j1) * TT(b0, i2, j2) - TT(b0, i1, j2) * TT(b0, i2, j1);  // cofactor C(ii,jj)
    double tgv = 0.0, tgdv = 0.0;
    for (int k = 1; k <= 3; k++) {
      tgv += TT(b0, k, ii) * TT(b0, k, jj);
      tgdv = tgdv + TT(b0, k, ii) * TT(b1, k, jj) + TT(b1, k, ii) * TT(b0, k, jj);
    }
    const double fbv = tkv + tpv - press * tcv;
    bs->tp[t] = tpv;
    bs->tk[t] = tkv;
    bs->tc[t] = tcv;
    bs->fb[t] = fbv;
    tg[t] = tgv;
    tgd[t] = tgdv;
    if (mode == FIN_FORCE || mode == FIN_INIT) {
      bs->tg[t] = tgv;
      bs->tgdot[t] = tgdv;  // kinten's tgdot
    }
    if (mode == FIN_VV || mode == FIN_VV_THERM) {
      bs->tgid_lag[t] = tgid0[t];
      const bool upd = ivol < 1 || (ivol == 4 && t == 8);
      double b2v = b2[t];
      if (upd) b2v = bdel2 * fbv;
      bs->b2[t] = b2v;
      if (upd) b1[t] = b1[t] + b2v;
      bs->b1[t] = b1[t];
    } else if (mode == FIN_INIT) {
      bs->b2[t] = bdel2 * fbv;
    }
  }
  __syncthreads();
  if (mode == FIN_VV_THERM || mode == FIN_INIT) {
    // update_therm_avs box part (VV_THERM) or moldin's first tgid (INIT)
    if (t < 9 && mode == FIN_VV_THERM) {
      double tgdv = 0.0;
      if (ivol != 1)
        for (int k = 1; k <= 3; k++) tgdv = tgdv + TT(b1, k, ii) * TT(b0, k, jj) + TT(b0, k, ii) * TT(b1, k, jj);
      tgd[t] = tgdv;
      bs->tg[t] = tg[t];
      bs->tgdot[t] = tgdv;
    }
    if (t < 9) {
      // inverse of tg, element (ii,jj): cofactor(jj,ii)/det
      const int a1 = jj % 3 + 1, a2 = (jj + 1) % 3 + 1, c1 = ii % 3 + 1, c2 = (ii + 1) % 3 + 1;
      const double det = TT(tg, 1, 1) * (TT(tg, 2, 2) * TT(tg, 3, 3) - TT(tg, 2, 3) * TT(tg, 3, 2)) +
                         TT(tg, 2, 1) * (TT(tg, 3, 2) * TT(tg, 1, 3) - TT(tg, 1, 2) * TT(tg, 3, 3)) +
                         TT(tg, 3, 1) * (TT(tg, 1, 2) * TT(tg, 2, 3) - TT(tg, 2, 2) * TT(tg, 1, 3));
      if (fabs(det) == 0.0) bs->status = MDK_ERR_SINGULAR_BOX;
      tgi[t] = (TT(tg, a1, c1) * TT(tg, a2, c2) - TT(tg, a1, c2) * TT(tg, a2, c1)) / det;
      bs->tginv[t] = tgi[t];
    }
    __syncthreads();
    if (t < 9) {
      double g = 0.0;
      for (int k = 1; k <= 3; k++) g += TT(tgi, ii, k) * TT(tgd, k, jj);
      bs->tgid[t] = g;
      if (mode == FIN_VV_THERM) {
        // cofactor(b0)/vol arranged as update_therm_avs's b0_XXdet ... (dynamics_m.F90:692-700)
        const double vol = dev_volume(b0);
        const int r1 = ii % 3 + 1, r2 = (ii + 1) % 3 + 1, q1 = jj % 3 + 1, q2 = (jj + 1) % 3 + 1;
        bs->binv[t] = (TT(b0, r1, q1) * TT(b0, r2, q2) - TT(b0, r1, q2) * TT(b0, r2, q1)) / vol;
      }
    }
  }
}

// Completion of the half kick's sums: s1,s2,s3, KE, box KE, TE, H, TH
// (moldin.F90:763-790), the trace row and the host-visible skin monitor
// (thread 0 of the finishing kernel's last block).
__device__ void finish_kick(const double (&v)[NPART_KICK], const FinKick &F) {
#pragma clang fp contract(off)
  BoxState *bs = F.bs;
  const double dt = F.dt, bdel2 = F.bdel2, press = F.press;
  double *trace_row = F.trace_row, *mon = F.mon;
  // only the fields this function changes are read and written
  const double pe = bs->pe, vol = bs->vol;
  double b0[9], b1[9], b2[9], br[9];
  for (int k = 0; k < 9; k++) {  // every load issued before the first use
    b0[k] = bs->b0[k];
    b1[k] = bs->b1[k];
    b2[k] = bs->b2[k];
    br[k] = bs->bref[k];
  }
  const double ke = v[2] / (dt * dt) / 2.0;
  double bk = 0.0;
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++) bk = bk + (TT(b1, i, j) * TT(b1, i, j)) / bdel2;
  const double te = ke + pe, hh = te + press * vol, th = hh + bk;
  for (BoxState *o : {bs, F.bout}) {  // the fused drift's copy was taken before these sums
    if (!o) continue;
    o->s1 = v[0];
    o->s2 = v[1];
    o->s3 = v[2];
    o->maxdisp2 = v[3];
    o->maxstep2 = v[4];
    o->ke = ke;
    o->boxke = bk;
    o->te = te;
    o->h = hh;
    o->th = th;
  }
  if (trace_row) {
    trace_row[0] = pe; trace_row[1] = ke; trace_row[2] = te;
    trace_row[3] = hh; trace_row[4] = th; trace_row[5] = vol;
  }
  if (mon) {  // the host forms the box strains from B0, b1 + b2 and Bref
    mon[0] = v[3];
    mon[1] = v[4];
    for (int k = 0; k < 9; k++) {
      mon[MON_B0 + k] = b0[k];
      mon[MON_DB + k] = b1[k] + b2[k];
      mon[MON_BREF + k] = br[k];
    }
  }
}

// Finishing kernel of the half kick, one workgroup of MDK_FOLD_BLOCK threads.
__global__ void __launch_bounds__(MDK_FOLD_BLOCK) k_finalize_kick(const double *part, int nb, FinKick F) {
  double v[NPART_KICK];
  fold_all<NPART_KICK, 2>(part, nb, v);
  if (threadIdx.x == 0) finish_kick(v, F);
}

// Force pass (src/dynamics_m.F90:235-346); kinten's partial sums come from
// the kernel that last set the velocities (kin_terms).
// fp = -(V' + phi'(F'_i + F'_j))/r;  F_i = sum_j r_ij fp (Cartesian) and
// f_i = B0^-1 F_i;  the virial is accumulated as W = 1/2 sum r r^T fp (tp =
// W B0^-T in the finishing block, i.e. sum r_a s_b fp) and the pair energy as
// 1/2 sum V.  One lane per atom; MASSCHK keeps the
// reference's exclusion of pairs of immobile atoms from the virial.
template <typename T, bool MASSCHK, int POTM, bool P32>
__global__ void __launch_bounds__(MDK_BLOCK)
k_force(double *atoms, const double4 *__restrict__ pos4, const float4 *__restrict__ p4f, const int *zat,
        int n, const int *__restrict__ nbr, const int *__restrict__ nvisit, const BoxState *bs, PackGeo pg,
        PotDev P, double *part, const double *spec, const float *fetab) {
  __shared__ double S[32][3];
  // interval table of the FP32 Fe-Fe pair terms: bin records, then intervals
  __shared__ float4 fbin[MDK_FE_NB];
  __shared__ __align__(16) float fiv[8 * MDK_FE_NI];
  constexpr bool CUBIC = sizeof(T) == 4 && POTM == POT_FE;
  double b[9];
#pragma unroll
  for (int k = 0; k < 9; k++) b[k] = bs->b0[k];
  __shared__ double bi[9];  // B0^-1 for the fractional forces
  shift_table(b, pg, S);
  if (CUBIC) {
    const float4 *src = (const float4 *)fetab;
    if (threadIdx.x < MDK_FE_NB) fbin[threadIdx.x] = src[threadIdx.x];
    else if (threadIdx.x < MDK_FE_NB + 2 * MDK_FE_NI)
      ((float4 *)fiv)[threadIdx.x - MDK_FE_NB] = src[threadIdx.x];
  }
  if (threadIdx.x == 64) inv3(b, bi);
  __syncthreads();
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  const double *mass = atoms + (size_t)F_MASS * n;
  double v[NPART_FORCE];
#pragma unroll
  for (int k = 0; k < NPART_FORCE; k++) v[k] = 0.0;
  double fxi = 0.0, fyi = 0.0, fzi = 0.0;  // Cartesian force on i
  if (i < n) {
    const double4 pi = pos4[i];
    const T dfi = (T)pi.w;
    const double mi = MASSCHK ? mass[i] : 0.0;
    const int nn = nvisit[i];  // list slots that can hold a partner in range
    // sums of V and of r r^T fp; the factor 1/2 of both is applied once below
    double pe2 = 0.0, wxx = 0, wyy = 0, wzz = 0, wxy = 0, wxz = 0, wyz = 0;
    for (int q = 0; 4 * q < nn; q++) {
      const int4 e4 = *(const int4 *)(nbr + (((size_t)q * n + i) << 2));
      const unsigned e[4] = {(unsigned)e4.x, (unsigned)e4.y, (unsigned)e4.z, (unsigned)e4.w};
      double4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = partner<P32>(pos4, p4f, e[u] & NBR_IDX_MASK);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const unsigned eu = e[u];
        const int j = (int)(eu & NBR_IDX_MASK);
        double dx, dy, dz;
        T rx, ry, rz;
        const T r2 = pair_sep<T>(S, pi, pj[u], (int)(eu >> NBR_CODE_SHIFT), dx, dy, dz, rx, ry, rz);
        T r, rr;
        r_and_inv(r2, r, rr);
        T vv, dv, dph;
        if constexpr (CUBIC) {
          mdk_fe_terms(P, fbin, fiv, r, vv, dv, dph);
        } else {
          mdk_pair_vdv<T, POTM>(P, r, vv, dv);
          dph = mdk_pair_dphi<T, POTM>(P, r);
        }
        const T fp = -dv * rr + (-T(1) * (dph * dfi + dph * (T)pj[u].w) * rr);
        const double fpd = (double)fp;
        // FP32 Fe-Fe form: the force and virial sums take the FP32 separation
        // (widened exactly), which frees the FP64 one early: 93 VGPRs and five
        // waves per SIMD instead of 106 and four; its rounding (6e-8 relative)
        // is far below the form's force error
        if constexpr (CUBIC) { dx = (double)rx; dy = (double)ry; dz = (double)rz; }
        const double gx = dx * fpd, gy = dy * fpd, gz = dz * fpd;
        pe2 += (double)vv;
        fxi += gx;
        fyi += gy;
        fzi += gz;
        if (!MASSCHK || mi + mass[j] > (double)0.1f) {
          wxx += dx * gx; wyy += dy * gy; wzz += dz * gz;
          wxy += dx * gy; wxz += dx * gz; wyz += dy * gz;
        }
      }
    }
    v[0] = 0.5 * pe2;
    v[1] = 0.5 * wxx; v[2] = 0.5 * wyy; v[3] = 0.5 * wzz;
    v[4] = 0.5 * wxy; v[5] = 0.5 * wxz; v[6] = 0.5 * wyz;
  }
  if (i < n) {
    // fractional force f = B0^-1 F (the reference's fx = sum s fp, Cartesian = B0 f)
    atoms[(size_t)F_FX * n + i] = bi[0] * fxi + bi[3] * fyi + bi[6] * fzi;
    atoms[(size_t)F_FY * n + i] = bi[1] * fxi + bi[4] * fyi + bi[7] * fzi;
    atoms[(size_t)F_FZ * n + i] = bi[2] * fxi + bi[5] * fyi + bi[8] * fzi;
  }
  block_store<NPART_FORCE, 0>(v, part);
}

// ---------------------------------------------------------------------------
// FP32 Fe-Fe pair passes on cell-relative block lists (the FP32 form's main
// path).  Both atoms of a pair come from the 16-byte FP32 packed records and
// the separation (R_i - R_j) + D[code] is formed in FP32; r and 1/r come from
// the hardware reciprocal square root with one Newton step.  One lane per atom.
// ---------------------------------------------------------------------------
// D[code] as an unevaluated FP32 sum hi + lo: a rounded D would bias every
// separation of one cell offset the same way (pairs (i,j) and (j,i) alike),
// and that bias adds up in the energy; ((R_i - R_j) + hi) + lo does not.
__device__ __forceinline__ void shift_table_f32(const double (*S)[3], float4 *Sf, float4 *Sl) {
  for (int c = threadIdx.x; c < 32; c += blockDim.x) {
    const float hx = (float)S[c][0], hy = (float)S[c][1], hz = (float)S[c][2];
    Sf[c] = make_float4(hx, hy, hz, 0.f);
    Sl[c] = make_float4((float)(S[c][0] - hx), (float)(S[c][1] - hy), (float)(S[c][2] - hz), 0.f);
  }
}
__device__ __forceinline__ float pair_r2_f32(const float4 &pi, const float4 &pj, const float4 *Sf,
                                             const float4 *Sl, unsigned e, float &rx, float &ry, float &rz) {
  const float4 d = Sf[e >> NBR_CODE_SHIFT], dl = Sl[e >> NBR_CODE_SHIFT];
  rx = ((pi.x - pj.x) + d.x) + dl.x;
  ry = ((pi.y - pj.y) + d.y) + dl.y;
  rz = ((pi.z - pj.z) + d.z) + dl.z;
  return rx * rx + ry * ry + rz * rz;
}

// Density pass (rhoset + embed, src/dynamics_m.F90:497-579,
// src/potential_m.F90:148-169): phi in FP32, each atom's density summed in
// FP64; nact[i] = list slots up to the end of the last quad holding a partner
// within r_int + 1e-4 A (the force pass's visit length).
__global__ void __launch_bounds__(MDK_BLOCK) k_density_f32(double *atoms, double4 *pos4, float4 *p4f, int n,
                                                           const int *__restrict__ nbr,
                                                           const int *__restrict__ numn, const BoxState *bs,
                                                           PackGeo pg, PotDev P, int do_embed, double *part,
                                                           int *nact) {
  __shared__ double S[32][3];
  __shared__ float4 Sf[32], Sl[32];
  shift_table(bs->b0, pg, S);
  __syncthreads();
  shift_table_f32(S, Sf, Sl);
  __syncthreads();
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  const float rc = (float)P.r_int + 1e-4f, rc2 = rc * rc;
  double acc = 0.0;
  if (i < n) {
    const float4 pi = p4f[i];
    const int nn = numn[i];
    int na = 0;
    int4 e4 = make_int4(0, 0, 0, 0);
    if (nn > 0) e4 = *(const int4 *)(nbr + ((size_t)i << 2));
    for (int q = 0; 4 * q < nn; q++) {
      const unsigned e[4] = {(unsigned)e4.x, (unsigned)e4.y, (unsigned)e4.z, (unsigned)e4.w};
      float4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = p4f[e[u] & NBR_IDX_MASK];
      // the next quad's indices load while this quad's partners arrive
      if (4 * (q + 1) < nn) e4 = *(const int4 *)(nbr + (((size_t)(q + 1) * n + i) << 2));
      float mn = rc2;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        float rx, ry, rz;
        const float r2 = pair_r2_f32(pi, pj[u], Sf, Sl, e[u], rx, ry, rz);
        if (r2 < rc2) {  // phi is zero beyond r_int; the list's tail is out of range
          float r, ri;
          r_and_inv(r2, r, ri);
          acc += (double)mdk_pair_phi<float, POT_FE>(P, r);
        }
        mn = fminf(mn, r2);
      }
      na = mn < rc2 ? 4 * q + 4 : na;
    }
    nact[i] = na;
  }
  density_finish(atoms, pos4, p4f, n, i, acc, P, do_embed, part);
}

// The lean path's density pass on a new list: k_density_f32's arithmetic (so
// the densities are those of every later pass on this list), and each atom's
// list reordered in place as k_density_part does -- partners within rc first,
// in list order, then the others (DENS_OBUF of them held in LDS, more stay in
// place) -- so that the force pass stops after the in-range part.
__global__ void __launch_bounds__(MDK_BLOCK) k_density_part_f32(double *atoms, float4 *p4f, int n, int *nbr,
                                                                const int *__restrict__ numn, const BoxState *bs,
                                                                PackGeo pg, PotDev P, int do_embed, double *part,
                                                                int *nact) {
  __shared__ double S[32][3];
  __shared__ float4 Sf[32], Sl[32];
  __shared__ int obuf[DENS_OBUF][MDK_BLOCK];
  shift_table(bs->b0, pg, S);
  __syncthreads();
  shift_table_f32(S, Sf, Sl);
  __syncthreads();
  const int t = threadIdx.x, i = xcd_tile() * blockDim.x + t;
  const float rc = (float)P.r_int + 1e-4f, rc2 = rc * rc;
  double acc = 0.0;
  if (i < n) {
    const float4 pi = p4f[i];
    const int nn = numn[i];
    int m = 0, no = 0, last = 0;  // entries written back, entries held, end of the last in-range one
    int4 qd = make_int4(0, 0, 0, 0);
    for (int q = 0; 4 * q < nn; q++) {
      const int4 e4 = *(const int4 *)(nbr + (((size_t)q * n + i) << 2));
      const int e[4] = {e4.x, e4.y, e4.z, e4.w};
      float4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = p4f[(unsigned)e[u] & NBR_IDX_MASK];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        float rx, ry, rz;
        const float r2 = pair_r2_f32(pi, pj[u], Sf, Sl, (unsigned)e[u], rx, ry, rz);
        if (r2 < rc2) {
          float r, ri;
          r_and_inv(r2, r, ri);
          acc += (double)mdk_pair_phi<float, POT_FE>(P, r);
        }
        if (4 * q + u < nn) {
          const bool in = r2 < rc2;
          if (in || no == DENS_OBUF) {
            quad_put(nbr, i, n, e[u], m, qd);
            if (in) last = m;
          } else {
            obuf[no++][t] = e[u];
          }
        }
      }
    }
    for (int u = 0; u < no; u++) quad_put(nbr, i, n, obuf[u][t], m, qd);
    if (m & 3) {  // the last quad's null entries
      const int nul = i | (NBR_NULL_CODE << NBR_CODE_SHIFT), sl = m & 3;
      qd.y = sl <= 1 ? nul : qd.y;
      qd.z = sl <= 2 ? nul : qd.z;
      qd.w = nul;
      *(int4 *)(nbr + nbr_at(m & ~3, i, n)) = qd;
    }
    nact[i] = (last + 3) & ~3;
  }
  density_finish(atoms, nullptr, p4f, n, i, acc, P, do_embed, part);
}

// Force pass (src/dynamics_m.F90:235-346): V, V' and phi' from the FP32 interval table, fp = -(V' + phi'(F'_i +
// F'_j))/r, F_i = sum r fp.  Each atom's force and the lane's pair energy and
// virial W = 1/2 sum r r^T fp are FP32 sums over the atom's few tens of terms;
// the sums over lanes, waves and workgroups are FP64.
__global__ void __launch_bounds__(MDK_BLOCK) k_force_f32(double *atoms, const float4 *__restrict__ p4f,
                                                         const int *zat, int n, const int *__restrict__ nbr,
                                                         const int *__restrict__ nvisit, const BoxState *bs,
                                                         PackGeo pg, PotDev P, double *part, const double *spec,
                                                         const float *fetab) {
  __shared__ double S[32][3];
  __shared__ float4 Sf[32], Sl[32];
  __shared__ float4 fbin[MDK_FE_NB];
  __shared__ __align__(16) float fiv[8 * MDK_FE_NI];
  __shared__ double bi[9];
  double b[9];
#pragma unroll
  for (int k = 0; k < 9; k++) b[k] = bs->b0[k];
  shift_table(b, pg, S);
  {
    const float4 *src = (const float4 *)fetab;
    if (threadIdx.x < MDK_FE_NB) fbin[threadIdx.x] = src[threadIdx.x];
    else if (threadIdx.x < MDK_FE_NB + 2 * MDK_FE_NI) ((float4 *)fiv)[threadIdx.x - MDK_FE_NB] = src[threadIdx.x];
  }
  if (threadIdx.x == 128) inv3(b, bi);
  __syncthreads();
  shift_table_f32(S, Sf, Sl);
  __syncthreads();
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  double v[NPART_FORCE];
#pragma unroll
  for (int k = 0; k < NPART_FORCE; k++) v[k] = 0.0;
  if (i < n) {
    const float4 pi = p4f[i];
    const int nn = nvisit[i];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    // sums of V and of r r^T fp (the factor 1/2 of both is applied once below)
    float pe2 = 0.f, wxx = 0.f, wyy = 0.f, wzz = 0.f, wxy = 0.f, wxz = 0.f, wyz = 0.f;
    int4 e4 = make_int4(0, 0, 0, 0);
    if (nn > 0) e4 = *(const int4 *)(nbr + ((size_t)i << 2));
    for (int q = 0; 4 * q < nn; q++) {
      const unsigned e[4] = {(unsigned)e4.x, (unsigned)e4.y, (unsigned)e4.z, (unsigned)e4.w};
      float4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) pj[u] = p4f[e[u] & NBR_IDX_MASK];
      // the next quad's indices load while this quad's partners arrive
      if (4 * (q + 1) < nn) e4 = *(const int4 *)(nbr + (((size_t)(q + 1) * n + i) << 2));
#pragma unroll
      for (int u = 0; u < 4; u++) {
        float rx, ry, rz;
        const float r2 = pair_r2_f32(pi, pj[u], Sf, Sl, e[u], rx, ry, rz);
        float r, ri;
        r_and_inv(r2, r, ri);
        float vv, dv, dph;
        mdk_fe_terms(P, fbin, fiv, r, vv, dv, dph);
        const float fp = -(dv + dph * (pi.w + pj[u].w)) * ri;
        const float gx = rx * fp, gy = ry * fp, gz = rz * fp;
        fx += gx;
        fy += gy;
        fz += gz;
        pe2 += vv;
        wxx += rx * gx; wyy += ry * gy; wzz += rz * gz;
        wxy += rx * gy; wxz += rx * gz; wyz += ry * gz;
      }
    }
    v[0] = 0.5 * (double)pe2;
    v[1] = 0.5 * (double)wxx; v[2] = 0.5 * (double)wyy; v[3] = 0.5 * (double)wzz;
    v[4] = 0.5 * (double)wxy; v[5] = 0.5 * (double)wxz; v[6] = 0.5 * (double)wyz;
    // fractional force f = B0^-1 F (the reference's fx = sum s fp, Cartesian = B0 f)
    const double Fx = fx, Fy = fy, Fz = fz;
    atoms[(size_t)F_FX * n + i] = bi[0] * Fx + bi[3] * Fy + bi[6] * Fz;
    atoms[(size_t)F_FY * n + i] = bi[1] * Fx + bi[4] * Fy + bi[7] * Fz;
    atoms[(size_t)F_FZ * n + i] = bi[2] * Fx + bi[5] * Fy + bi[8] * Fz;
  }
  block_store<NPART_FORCE, 0>(v, part);
}

// scalev (src/thermostat_m.F90:317-395), Nose-Hoover velocity scaling, as
// three passes with deterministic block partials: (1) sums of x1 over all
// atoms; (2) moving atoms (Z != 0, mass >= 0.1) lose the mean, partial
// kinetic sums m |B0 x1|^2; (3) moving atoms' x1 scaled by 1 + snhv (the
// host forms ke, snhf and snhv between (2) and (3)).
__global__ void __launch_bounds__(MDK_BLOCK) k_vsum(const double *__restrict__ atoms, int n, double *part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < n)
    for (int d = 0; d < 3; d++) v[d] = atoms[(size_t)(F_X1 + d) * n + i];
  block_store<3, 0>(v, part);
}
__device__ __forceinline__ bool moving_atom(const double *atoms, const int *zat, const double *spec, int n, int i) {
  const int z = zat[i];
  const double m = spec ? spec[2 * z] : atoms[(size_t)F_MASS * n + i];
  return !(z == 0 || m < (double)0.1f);
}
__global__ void __launch_bounds__(MDK_BLOCK) k_vsub_ke(double *__restrict__ atoms, const int *__restrict__ zat, int n,
                                                       const double *__restrict__ mean, const BoxState *bs,
                                                       const double *__restrict__ spec, double *part) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (i < n && moving_atom(atoms, zat, spec, n, i)) {
    double x[3];
    for (int d = 0; d < 3; d++) {
      double *q = atoms + (size_t)(F_X1 + d) * n + i;
      x[d] = *q - mean[d];
      *q = x[d];
    }
    const double *b = bs->b0;
    const double rx = b[0] * x[0] + b[3] * x[1] + b[6] * x[2];
    const double ry = b[1] * x[0] + b[4] * x[1] + b[7] * x[2];
    const double rz = b[2] * x[0] + b[5] * x[1] + b[8] * x[2];
    const double m = spec ? spec[2 * zat[i]] : atoms[(size_t)F_MASS * n + i];
    v[0] = m * (rx * rx + ry * ry + rz * rz);
  }
  block_store<1, 0>(v, part);
}
__global__ void __launch_bounds__(MDK_BLOCK) k_vscale(double *__restrict__ atoms, const int *__restrict__ zat, int n,
                                                      const double *__restrict__ spec, double f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && moving_atom(atoms, zat, spec, n, i))
    for (int d = 0; d < 3; d++) atoms[(size_t)(F_X1 + d) * n + i] *= f;
}
// fixed-order fold of K per-block partials into out[0..K-1]
template <int K>
__global__ void __launch_bounds__(MDK_FIN_BLOCK) k_fold(const double *part, int nb, double *stage, unsigned *cnt,
                                                        double *out) {
  double v[K];
  if (!stage_and_finish<K, 0>(part, nb, stage, cnt, v)) return;
  if (threadIdx.x == 0)
    for (int k = 0; k < K; k++) out[k] = v[k];
}

// the box of a new neighbour list (one thread)
__global__ void k_set_bref(BoxState *bs) {
  if (threadIdx.x < 9) bs->bref[threadIdx.x] = bs->b0[threadIdx.x];
}
// the list kernel's flags into host-mapped memory
__global__ void k_copy_flags(const int *dflags, int *out) {
  if (threadIdx.x < 4) out[threadIdx.x] = dflags[threadIdx.x];
}

// update_therm_avs box part only (stand-alone entry point)
__global__ void k_therm_box(BoxState *bs, int ivol) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    BoxState s = *bs;
    therm_box(s, ivol);
    *bs = s;
  }
}

// Every pointer is restrict-qualified: the box state is only read, so its
// uniform fields come through scalar loads instead of one vector load per lane.
// Velocity half-kick (src/quench_m.F90:604-622) fused with the per-atom sums of
// update_therm_avs (src/dynamics_m.F90:702-732) and the list-skin monitor.
// mode 0: VV kick + sums; 1: sums only; 2: initial accelerations (moldin.F90:461-472)
// DRIFT (mode 0 only) fuses the next step's drift: x0 += x1 + x2 with the new
// velocities, packed positions with the advanced box, and block 0 writes the
// advanced box state into bout -- what k_drift would do at the start of the
// next step.
template <bool DRIFT>
__global__ void __launch_bounds__(MDK_BLOCK) k_kick(double *__restrict__ atoms, const int *__restrict__ zat, int n,
                                                    const BoxState *__restrict__ bs, int ivol, int mode,
                                                    const float *__restrict__ sref, double *__restrict__ part,
                                                    double4 *__restrict__ pos4, float4 *__restrict__ p4f,
                                                    const int *__restrict__ img, const int *__restrict__ cell,
                                                    PackGeo pg, BoxState *__restrict__ bout,
                                                    const double *__restrict__ spec, double *__restrict__ kin) {
#pragma clang fp contract(off)
  const int i = xcd_tile() * blockDim.x + threadIdx.x;
  double v[NPART_KICK] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double kv[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    double *x0p = atoms, *y0p = atoms + n, *z0p = atoms + 2 * (size_t)n;
    double *x1p = atoms + (size_t)F_X1 * n, *y1p = atoms + (size_t)F_Y1 * n, *z1p = atoms + (size_t)F_Z1 * n;
    double *x2p = atoms + (size_t)F_X2 * n, *y2p = atoms + (size_t)F_Y2 * n, *z2p = atoms + (size_t)F_Z2 * n;
    const double *fxp = atoms + (size_t)F_FX * n, *fyp = atoms + (size_t)F_FY * n, *fzp = atoms + (size_t)F_FZ * n;
    // Every per-atom load is issued here, before the first store: the fields
    // share one allocation, so a load written after a store to another field
    // could not be moved ahead of it, and each would cost a round trip of its own.
    const int z = zat[i];
    double x1 = x1p[i], y1 = y1p[i], z1 = z1p[i];
    double x2 = x2p[i], y2 = y2p[i], z2 = z2p[i];
    const double fx = fxp[i], fy = fyp[i], fz = fzp[i];
    const double x0 = x0p[i], y0 = y0p[i], z0 = z0p[i];
    int im0 = 0, im1 = 0, im2 = 0;
    unsigned ck = 0u;
    if (DRIFT) {
      im0 = img[3 * i];
      im1 = img[3 * i + 1];
      im2 = img[3 * i + 2];
      if (pg.gx) ck = (unsigned)cell[i];
    }
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    if (sref) {
      r0 = sref[3 * i];
      r1 = sref[3 * i + 1];
      r2 = sref[3 * i + 2];
    }
    // mass and del2 from the species table when they are functions of the
    // species: one 16-byte load of the pair (the table is hipMalloc-aligned)
    const double2 md = spec ? *reinterpret_cast<const double2 *>(spec + 2 * z)
                            : make_double2(atoms[(size_t)F_MASS * n + i], atoms[(size_t)F_DEL2 * n + i]);
    const double m = md.x, d2 = md.y;
    if (mode == 0) {
      const double *g = bs->tgid_lag;
      const double xd = x1 + x2, yd = y1 + y2, zd = z1 + z2;  // x1_dt2
      if (!(z == 0 || m < (double)0.1f)) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (ivol == 0) {
          sx = TT(g, 1, 1) * x1 + TT(g, 1, 2) * y1 + TT(g, 1, 3) * z1;
          sy = TT(g, 2, 1) * x1 + TT(g, 2, 2) * y1 + TT(g, 2, 3) * z1;
          sz = TT(g, 3, 1) * x1 + TT(g, 3, 2) * y1 + TT(g, 3, 3) * z1;
        }
        if (ivol == 4) sz = TT(g, 3, 3) * z1;
        x2 = d2 * fx - 0.5 * sx;
        y2 = d2 * fy - 0.5 * sy;
        z2 = d2 * fz - 0.5 * sz;
        x2p[i] = x2; y2p[i] = y2; z2p[i] = z2;
      }
      x1 = xd + x2; y1 = yd + y2; z1 = zd + z2;
      x1p[i] = x1; y1p[i] = y1; z1p[i] = z1;
    } else if (mode == 2) {
      if (z != 0) {
        const double *g = bs->tgid;
        double a1 = TT(g, 1, 1) * x1 + TT(g, 1, 2) * y1 + TT(g, 1, 3) * z1;
        double a2 = TT(g, 2, 1) * x1 + TT(g, 2, 2) * y1 + TT(g, 2, 3) * z1;
        double a3 = TT(g, 3, 1) * x1 + TT(g, 3, 2) * y1 + TT(g, 3, 3) * z1;
        x2 = d2 * fx - 0.5 * a1;
        y2 = d2 * fy - 0.5 * a2;
        z2 = d2 * fz - 0.5 * a3;
        x2p[i] = x2; y2p[i] = y2; z2p[i] = z2;
      }
    }
    const double *b = bs->b0;
    if (mode != 2 && z != 0) {
      const double *g = bs->tgid, *bi = bs->binv;
      double sx = TT(g, 1, 1) * x1 + TT(g, 1, 2) * y1 + TT(g, 1, 3) * z1;
      double sy = TT(g, 2, 1) * x1 + TT(g, 2, 2) * y1 + TT(g, 2, 3) * z1;
      double sz = TT(g, 3, 1) * x1 + TT(g, 3, 2) * y1 + TT(g, 3, 3) * z1;
      double fxq = 2.0 * x2 + sx, fyq = 2.0 * y2 + sy, fzq = 2.0 * z2 + sz;
      double fxi = TT(bi, 1, 1) * fxq + TT(bi, 1, 2) * fyq + TT(bi, 1, 3) * fzq;
      double fyi = TT(bi, 2, 1) * fxq + TT(bi, 2, 2) * fyq + TT(bi, 2, 3) * fzq;
      double fzi = TT(bi, 3, 1) * fxq + TT(bi, 3, 2) * fyq + TT(bi, 3, 3) * fzq;
      double f2 = fxi * fxi + fyi * fyi + fzi * fzi;
      v[0] = f2;
      v[1] = f2 * f2;
      double rx = TT(b, 1, 1) * x1 + TT(b, 1, 2) * y1 + TT(b, 1, 3) * z1;
      double ry = TT(b, 2, 1) * x1 + TT(b, 2, 2) * y1 + TT(b, 2, 3) * z1;
      double rz = TT(b, 3, 1) * x1 + TT(b, 3, 2) * y1 + TT(b, 3, 3) * z1;
      v[2] = m * (rx * rx + ry * ry + rz * rz);
    }
    if (sref) {
      double ex = x0 - (double)r0, ey = y0 - (double)r1, ez = z0 - (double)r2;
      double qx = b[0] * ex + b[3] * ey + b[6] * ez, qy = b[1] * ex + b[4] * ey + b[7] * ez,
             qz = b[2] * ex + b[5] * ey + b[8] * ez;
      v[3] = qx * qx + qy * qy + qz * qz;
      double sx = x1 + x2, sy = y1 + y2, sz = z1 + z2;
      qx = b[0] * sx + b[3] * sy + b[6] * sz;
      qy = b[1] * sx + b[4] * sy + b[7] * sz;
      qz = b[2] * sx + b[5] * sy + b[8] * sz;
      v[4] = qx * qx + qy * qy + qz * qz;
    }
    if (DRIFT) {
      const double xn = x0 + (x1 + x2), yn = y0 + (y1 + y2), zn = z0 + (z1 + z2);
      x0p[i] = xn;
      y0p[i] = yn;
      z0p[i] = zn;
      double bn[9];
      for (int k = 0; k < 9; k++) {
        const bool upd = ivol < 1 || (ivol == 4 && k == 8);
        bn[k] = upd ? bs->b0[k] + (bs->b1[k] + bs->b2[k]) : bs->b0[k];
      }
      const double3 r = packed_pos_key(bn, pg, ck, xn - im0, yn - im1, zn - im2);
      store_packed(pos4, p4f, i, r, 0.0);  // F' is stored by the density pass, which always runs next
      if (z != 0) kin_terms(bn, x1, y1, z1, m, kv);  // kinten of the next force evaluation
    }
  }
  if (DRIFT && blockIdx.x == 0) drift_box(bs, bout, ivol);
  block_store<NPART_KICK, 2>(v, part);
  if (DRIFT) block_store<9, 0>(kv, kin);
}

// energy_calc (src/dynamics_m.F90:410-470): en_i = F(rho_i) + 1/2 sum_j V(r_ij)
// over atom i's full list (each pair once from each end, as the reference's
// half list adds V/2 to both ends).  The pair function is the exact one even
// when the force path uses tables: the reference calls vee_src there.
template <int POTM>
__global__ void __launch_bounds__(MDK_BLOCK) k_energy_calc(const double *atoms, const double4 *__restrict__ pos4,
                                                           int n, const int *__restrict__ nbr,
                                                           const int *__restrict__ numn, const BoxState *bs,
                                                           PackGeo pg, PotDev P, double *en) {
  __shared__ double S[32][3];
  shift_table(bs->b0, pg, S);
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos4[i];
  double s = 0.0;
  for (int k = 0; k < numn[i]; k++) {
    const unsigned e = (unsigned)nbr[nbr_at(k, i, n)];
    const double4 pj = pos4[e & NBR_IDX_MASK];
    double dx, dy, dz, rx, ry, rz;
    const double r = sqrt(pair_sep<double>(S, pi, pj, (int)(e >> NBR_CODE_SHIFT), dx, dy, dz, rx, ry, rz));
    double v, dv;
    mdk_pair_vdv<double, POTM>(P, r, v, dv);
    s += v / 2.0;
  }
  en[i] = atoms[(size_t)F_AFRHO * n + i] + s;
}

// Per-atom diagonal stress of the reference's force loop
// (src/dynamics_m.F90:279-283): stress_x(i) = sum of rx_ij fp_ij over atom i's
// own half-list partners j only (the partner gets nothing from that pair),
// with fp = -V'/r - (phi' F'_i + phi' F'_j)/r as the force loop forms it.
// One thread per atom in reference order; positions at the atom's current slot.
template <int POTM>
__global__ void __launch_bounds__(MDK_BLOCK) k_atomic_stress(const double *atoms, const int *slot_of, int n,
                                                             const int *numn_ref, const int *nlist_ref,
                                                             int nnbrs, const BoxState *bs, int pbc, PotDev P,
                                                             double *sx, double *sy, double *sz) {
#pragma clang fp contract(off)
  const int oi = blockIdx.x * blockDim.x + threadIdx.x;
  if (oi >= n) return;
  const double *b = bs->b0;
  const int i = slot_of[oi];
  const double *x0 = atoms, *y0 = atoms + n, *z0 = atoms + 2 * (size_t)n;
  const double *df = atoms + (size_t)F_DAFRHO * n;
  const double xi = x0[i], yi = y0[i], zi = z0[i], dfi = df[i];
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int cnt = min(numn_ref[oi], nnbrs);
  for (int q = 0; q < cnt; q++) {
    const int j = slot_of[nlist_ref[(size_t)oi * nnbrs + q] - 1];
    double dx = xi - x0[j], dy = yi - y0[j], dz = zi - z0[j];
    min_image(pbc, dx, dy, dz);
    const double r = sqrt(sep2_exact(b, dx, dy, dz));
    const double rr = 1.0 / r;
    double v, dv;
    mdk_pair_vdv<double, POTM>(P, r, v, dv);
    const double dph = mdk_pair_dphi<double, POTM>(P, r);
    const double fp = -dv * rr + (-1.0 * (dph * dfi + dph * df[j]) * rr);
    ax += (b[0] * dx + b[3] * dy + b[6] * dz) * fp;
    ay += (b[1] * dx + b[4] * dy + b[7] * dz) * fp;
    az += (b[2] * dx + b[5] * dy + b[8] * dz) * fp;
  }
  sx[oi] = ax;
  sy[oi] = ay;
  sz[oi] = az;
}

// potential evaluation for mdk_eval_potential
__global__ void k_eval_pot(PotDev P, int which, const double *x, double *y, int cnt, const float *fetab) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  if (which >= 6) {  // the FP32 force pass's V, V', phi' (interval table)
    float v32, dv32, dph32;
    mdk_fe_terms(P, (const float4 *)fetab, fetab + 4 * MDK_FE_NB, (float)x[k], v32, dv32, dph32);
    y[k] = which == 6 ? v32 : (which == 7 ? dv32 : dph32);
    return;
  }
  double r = x[k], v = 0.0, dv = 0.0;
  const bool lk = P.use_lookup != 0, fec = P.material == MDK_MAT_IRON_CARBON;
  switch (which) {
    case 0:
    case 1:
      if (lk) mdk_pair_vdv<double, POT_TABLE>(P, r, v, dv);
      else if (fec) mdk_pair_vdv<double, POT_FEC>(P, r, v, dv);
      else mdk_pair_vdv<double>(P, r, v, dv);
      y[k] = which == 0 ? v : dv;
      break;
    case 2:
      y[k] = lk ? mdk_pair_phi<double, POT_TABLE>(P, r)
                : (fec ? mdk_pair_phi<double, POT_FEC>(P, r) : mdk_pair_phi<double>(P, r));
      break;
    case 3:
      y[k] = lk ? mdk_pair_dphi<double, POT_TABLE>(P, r)
                : (fec ? mdk_pair_dphi<double, POT_FEC>(P, r) : mdk_pair_dphi<double>(P, r));
      break;
    case 4: mdk_embed_f(P, r, v, dv); y[k] = v; break;
    case 5: mdk_embed_f(P, r, v, dv); y[k] = dv; break;
  }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
static inline int nblocks(long n) { return (int)((n + MDK_BLOCK - 1) / MDK_BLOCK); }

static hipEvent_t pooled_event(mdk_ctx *c) {
  hipEvent_t e;
  if (!c->evpool.empty()) {
    e = c->evpool.back();
    c->evpool.pop_back();
  } else {
    hipEventCreate(&e);
  }
  return e;
}
static void tbegin(mdk_ctx *c, int cls, TimedEv &te) {
  te.cls = -1;
  if (!((c->timing_mask >> cls) & 1)) return;
  if (c->timing_every > 1 && (c->t_seen[cls]++ % c->timing_every) != 0) return;
  te.cls = cls;
  te.a = pooled_event(c);
  te.b = pooled_event(c);
  hipEventRecord(te.a, c->stream);
}
// Events filled in by the kernel dispatch itself (hipExtLaunchKernel): the
// interval is the kernel's own execution, with no gap before or after it.
static void tbegin_ext(mdk_ctx *c, int cls, TimedEv &te) {
  te.cls = -1;
  if (!((c->timing_mask >> cls) & 1)) return;
  if (c->timing_every > 1 && (c->t_seen[cls]++ % c->timing_every) != 0) return;
  te.cls = cls;
  te.a = pooled_event(c);
  te.b = pooled_event(c);
}
static void tpush(mdk_ctx *c, TimedEv &te);
static void tend(mdk_ctx *c, TimedEv &te) {
  if (te.cls < 0) return;
  hipEventRecord(te.b, c->stream);
  tpush(c, te);
}
static void tpush(mdk_ctx *c, TimedEv &te) {
  if (te.cls < 0) return;
  c->evs.push_back(te);
  if (c->evs.size() > 4096) {  // resolve old ones to bound memory
    hipEventSynchronize(te.b);
    for (auto &e : c->evs) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, e.a, e.b);
      c->t_ms[e.cls] += ms;
      c->t_cnt[e.cls] += 1;
      c->evpool.push_back(e.a);
      c->evpool.push_back(e.b);
    }
    c->evs.clear();
  }
}

static int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(MDK_ERR_HIP, "launch %s: %s", what, hipGetErrorString(e));
  return MDK_OK;
}

static double near_threshold(double rn) {
  // smallest double q with sqrt(q) >= rn (host sqrt is correctly rounded), so
  // that sqrt(q) < rn <=> q < threshold without a device square root.
  double q = rn * rn;
  while (sqrt(q) >= rn) q = nextafter(q, 0.0);
  while (sqrt(q) < rn) q = nextafter(q, INFINITY);
  return q;
}

static int read_box(mdk_ctx *c, BoxState *out) {
  HIPCK(hipMemcpyAsync(out, c->box + c->cur, sizeof(BoxState), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

// Fine grid from the current box: cells at least r_list wide (perpendicular width)
static void plan_grid(mdk_ctx *c, const double *b, int g[3], double margin) {
  double vol = fabs(TT(b, 1, 1) * (TT(b, 2, 2) * TT(b, 3, 3) - TT(b, 2, 3) * TT(b, 3, 2)) +
                    TT(b, 2, 1) * (TT(b, 3, 2) * TT(b, 1, 3) - TT(b, 1, 2) * TT(b, 3, 3)) +
                    TT(b, 3, 1) * (TT(b, 1, 2) * TT(b, 2, 3) - TT(b, 2, 2) * TT(b, 1, 3)));
  for (int d = 0; d < 3; d++) {
    int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
    const double *u = b + 3 * e1, *w = b + 3 * e2;
    double cx = u[1] * w[2] - u[2] * w[1], cy = u[2] * w[0] - u[0] * w[2], cz = u[0] * w[1] - u[1] * w[0];
    double width = vol / sqrt(cx * cx + cy * cy + cz * cz);
    int nc = (int)floor(width / (c->r_list * margin));
    if (nc < 3) nc = 1;  // too few cells for a distinct 27-cell stencil: one cell in this dim
    if (nc > 256) nc = 256;
    g[d] = nc;
  }
}

static int ensure_cells(mdk_ctx *c, int m) {
  if (m <= c->ncell_cap) return MDK_OK;
  if (c->cell_count) hipFree(c->cell_count);
  if (c->cell_start) hipFree(c->cell_start);
  if (c->cell_fill) hipFree(c->cell_fill);
  int cap = std::max(m, 64);
  HIPCK(hipMalloc(&c->cell_count, sizeof(int) * cap));
  HIPCK(hipMalloc(&c->cell_start, sizeof(int) * (cap + 1)));
  HIPCK(hipMalloc(&c->cell_fill, sizeof(int) * cap));
  c->ncell_cap = cap;
  return MDK_OK;
}

// counting sort of atoms by key[] into cell_start/members with a deterministic
// in-cell order (ascending sortkey, or slot when sortkey is null)
static int bin_atoms(mdk_ctx *c, const int *key, int m, const int *sortkey) {
  const int n = c->n;
  RCK(ensure_cells(c, m));
  HIPCK(hipMemsetAsync(c->cell_count, 0, sizeof(int) * m, c->stream));
  k_count<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(key, n, c->cell_count);
  const int nb = (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
  k_scan_a<<<nb, 1024, 0, c->stream>>>(c->cell_count, c->cell_start, c->cell_fill, m);
  k_scan_b<<<1, 1024, 0, c->stream>>>(c->cell_fill, nb, c->cell_start, m);
  k_scan_c<<<nblocks(m), MDK_BLOCK, 0, c->stream>>>(c->cell_start, c->cell_fill, m);
  HIPCK(hipMemsetAsync(c->cell_fill, 0, sizeof(int) * m, c->stream));
  k_fill<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(key, n, c->cell_start, c->cell_fill, c->members);
  k_sort_cells<<<nblocks(m), MDK_BLOCK, 0, c->stream>>>(c->cell_start, m, c->members, sortkey);
  return check_launch("bin_atoms");
}

static int ensure_pos4(mdk_ctx *c);

// Full-list rebuild: fine grid from the current box, re-sort atoms by cell,
// build the list; grows the per-atom capacity and retries on overflow.
// Morton-ordered list of the 2x2x2-cell blocks that hold cells of grid g
// (recomputed when g changes).
static int plan_list_blocks(mdk_ctx *c, const int g[3]) {
  if (c->lblocks && g[0] == c->lblock_g[0] && g[1] == c->lblock_g[1] && g[2] == c->lblock_g[2]) return MDK_OK;
  const int bx = (g[0] + 1) / 2, by = (g[1] + 1) / 2, bz = (g[2] + 1) / 2;
  std::vector<int> keys;
  keys.reserve((size_t)bx * by * bz);
  for (int z = 0; z < bz; z++)
    for (int y = 0; y < by; y++)
      for (int x = 0; x < bx; x++) {
        unsigned k = 0;
        for (int bit = 0; bit < 10; bit++) {
          k |= (unsigned)((x >> bit) & 1) << (3 * bit);
          k |= (unsigned)((y >> bit) & 1) << (3 * bit + 1);
          k |= (unsigned)((z >> bit) & 1) << (3 * bit + 2);
        }
        keys.push_back((int)k);
      }
  std::sort(keys.begin(), keys.end());
  const int nb = (int)keys.size();
  if (nb > c->lblocks_cap) {
    if (c->lblocks) HIPCK(hipFree(c->lblocks));
    HIPCK(hipMalloc(&c->lblocks, sizeof(int) * nb));
    c->lblocks_cap = nb;
  }
  HIPCK(hipMemcpy(c->lblocks, keys.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
  c->nlblocks = nb;
  c->lblock_g[0] = g[0]; c->lblock_g[1] = g[1]; c->lblock_g[2] = g[2];
  return MDK_OK;
}

// defer: one attempt whose flags the caller checks after queueing the next
// kernels (confirm_list); otherwise every attempt is checked here.
// Fields a re-sort must carry.  Inside force (force_follows) the density and
// force passes rewrite RHO and FX..FZ of every atom before anything reads
// them; X3..Z3 are zero in both buffers until the predictor-corrector or an
// upload sets them (both buffers are cleared at create, and the scratch uses
// of the alternate buffer stay in its X0..Z0 columns).
static FieldSet resort_fields(const mdk_ctx *c, bool force_follows) {
  FieldSet fs;
  fs.cnt = 0;
  for (int f = 0; f < F_N; f++) {
    const bool out = f == F_FX || f == F_FY || f == F_FZ || f == F_RHO;
    const bool third = f == F_X3 || f == F_Y3 || f == F_Z3;
    if ((out && force_follows) || (third && !c->x3_live)) continue;
    fs.f[fs.cnt++] = f;
  }
  return fs;
}

static int rebuild_full(mdk_ctx *c, bool defer = false, bool force_follows = false) {
  const int n = c->n;
  c->list_valid = false;
  TimedEv te;
  tbegin(c, TC_REBUILD, te);
  // Grid from the box.  Inside the MD loop the box of a completed step comes
  // from the monitor ring (no host round trip before the build; cells planned
  // 1 % wider than the list radius cover the box change of a few steps);
  // otherwise it is read synchronously.  The host reads the list kernel's
  // flags before any force uses the list (here, or in confirm_list after the
  // density pass is queued): a list that overflowed grows its capacity and is
  // built again, and a grid found too narrow for the current box is planned
  // again from the exact box -- so no flagged list is ever used.
  bool lagged = c->rebuilds > 0 && c->b0_lag_step >= 0 && c->step_counter - c->b0_lag_step <= 4;
  for (int plan = 0; plan < 2; plan++) {
    double b0[9];
    if (lagged) {
      memcpy(b0, c->b0_lag, sizeof(b0));
    } else {
      BoxState bs;
      RCK(read_box(c, &bs));
      memcpy(b0, bs.b0, sizeof(b0));
    }
    int g[3];
    plan_grid(c, b0, g, lagged ? 1.01 : 1.0);
    int p2 = 1;
    while (p2 < std::max(g[0], std::max(g[1], g[2]))) p2 <<= 1;
    const int m = p2 * p2 * p2;  // Morton key space
    c->grid[0] = g[0]; c->grid[1] = g[1]; c->grid[2] = g[2];
    k_fast_cells<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->atoms, c->atoms + n, c->atoms + 2 * (size_t)n,
                                                         n, g[0], g[1], g[2], c->cell_of);
    RCK(bin_atoms(c, c->cell_of, m, nullptr));
    // physical re-sort of every per-atom array into cell order
    const FieldSet fs = resort_fields(c, force_follows);
    k_permute<<<nblocks((long)n * fs.cnt), MDK_BLOCK, 0, c->stream>>>(c->atoms_alt, c->atoms, c->members, n, fs);
    k_permute_int<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->zat_alt, c->zat, c->orig_alt, c->orig,
                                                          c->slot_of, c->members, n);
    std::swap(c->atoms, c->atoms_alt);
    std::swap(c->zat, c->zat_alt);
    std::swap(c->orig, c->orig_alt);
    // reference cells move with the atoms
    k_gather_i<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->zat_alt, c->ic, c->members, n);
    std::swap(c->ic, c->zat_alt);  // zat_alt is free scratch after the swap above
    const double rl2 = c->r_list * c->r_list;
    // staged 2x2x2-cell blocks need a fully periodic box with >= 4 cells per
    // dimension (and 26-bit atom indices); they pack positions cell-relative
    const bool blk = c->pbc == 7 && g[0] >= 4 && g[1] >= 4 && g[2] >= 4 && n < (1 << 26);
    c->rel = blk;
    // images and list cells fixed for this list's lifetime; packed positions
    k_pack<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->atoms, c->img, c->box + c->cur, n, 1, c->pos4, p4f_of(c),
                                                   c->cell_of, g[0], g[1], g[2], pack_geo(c));
    c->pos4_valid = c->p4f_valid = true;
    if (blk) RCK(plan_list_blocks(c, g));
    bool replan = false;
    for (int attempt = 0; attempt < 4; attempt++) {
      HIPCK(hipMemsetAsync(c->dflags, 0, sizeof(int) * 4, c->stream));
      if (blk) {
        k_full_list_blk<<<c->nlblocks, 64, 0, c->stream>>>(
            c->pos4, c->atoms, c->img, c->cell_of, c->cell_start, c->lblocks, n, pack_geo(c), c->box + c->cur,
            rl2, c->maxn, c->nbr, c->numn, c->sref, c->dflags);
        RCK(check_launch("k_full_list_blk"));
      } else {
        k_full_list<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(
            c->pos4, c->atoms, c->img, c->cell_of, c->cell_start, n, g[0], g[1], g[2], c->box + c->cur, rl2,
            c->pbc, c->maxn, c->nbr, c->numn, c->sref, c->dflags);
        RCK(check_launch("k_full_list"));
      }
      if (defer) {
        k_copy_flags<<<1, 64, 0, c->stream>>>(c->dflags, c->flags_map);
        HIPCK(hipEventRecord(c->list_ev, c->stream));
        k_set_bref<<<1, 64, 0, c->stream>>>(c->box + c->cur);
        RCK(check_launch("k_set_bref"));
        c->list_valid = true;
        c->list_pending = true;
        c->part_due = true;
        c->rebuilds++;
        c->last_rebuild_step = (int)c->step_counter;  // skin samples before this step predate the list
        tend(c, te);
        return MDK_OK;
      }
      int fl[4];
      HIPCK(hipMemcpyAsync(fl, c->dflags, sizeof(fl), hipMemcpyDeviceToHost, c->stream));
      HIPCK(hipStreamSynchronize(c->stream));
      if (fl[2] != 0) {
        if (!lagged) {
          tend(c, te);
          return set_err(MDK_ERR_STATE, "list cells narrower than the list radius");
        }
        replan = true;  // planned from a stale box: plan again from the exact one
        break;
      }
      if (fl[0] == 0) {
        k_set_bref<<<1, 64, 0, c->stream>>>(c->box + c->cur);
        RCK(check_launch("k_set_bref"));
        c->list_valid = true;
        c->part_due = true;
        c->rebuilds++;
        c->last_rebuild_step = (int)c->step_counter;  // skin samples before this step predate the list
        tend(c, te);
        return MDK_OK;
      }
      // grow capacity to the observed maximum (+25%) and retry
      const int newmax = ((int)(fl[0] * 1.25) + 8 + 3) & ~3;
      HIPCK(hipFree(c->nbr));
      c->nbr = nullptr;
      HIPCK(hipMalloc(&c->nbr, sizeof(int) * (size_t)newmax * n));
      c->maxn = newmax;
    }
    if (!replan) break;
    lagged = false;
  }
  tend(c, te);
  HIPCK(hipMemsetAsync(c->numn, 0, sizeof(int) * n, c->stream));
  return set_err(MDK_ERR_NBR_OVERFLOW, "full neighbour list overflow (max %d)", c->maxn);
}

// The FP32 form gathers the 16-byte FP32 copy of cell-relative packed
// positions (P32); absolute packing and the FP64 form gather the FP64 record.
static int pot_mode(const mdk_ctx *c);
// The FP32 form's lean kernels: Fe-Fe interval table, cell-relative lists, no
// immobile atoms (the per-pair mass test).
static bool lean_path(const mdk_ctx *c) {
  return c->rel && c->p.precision == MDK_PREC_FP32 && pot_mode(c) == POT_FE && !c->masschk;
}
// workgroups of the pair passes (and per-block partials per value)
static int block_count(mdk_ctx *c) { return nblocks(c->n); }
// kinten's column of force()'s partials (written with c->nblk == nblocks(n) blocks)
static double *kin_part(mdk_ctx *c) { return c->part_f + (size_t)PART_KIN * c->nblk; }

static void lean_density(mdk_ctx *c, int do_embed) {
  double *pe = c->part_f + (size_t)PART_EMB * block_count(c);
  // F' goes to the FP32 records only: the FP64 copy is not kept on this path
  if (c->part_due)
    k_density_part_f32<<<block_count(c), MDK_BLOCK, 0, c->stream>>>(c->atoms, c->p4f, c->n, c->nbr, c->numn,
                                                                    c->box + c->cur, pack_geo(c), c->pot, do_embed,
                                                                    pe, c->nact);
  else
    k_density_f32<<<block_count(c), MDK_BLOCK, 0, c->stream>>>(c->atoms, nullptr, c->p4f, c->n, c->nbr, c->numn,
                                                               c->box + c->cur, pack_geo(c), c->pot, do_embed, pe,
                                                               c->nact);
  c->pos4_valid = false;
  c->part_due = false;
}
static void lean_force(mdk_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  hipExtLaunchKernelGGL(k_force_f32, dim3(block_count(c)), dim3(MDK_BLOCK), 0, c->stream, ea, eb, 0, c->atoms,
                        (const float4 *)c->p4f, (const int *)c->zat, c->n, (const int *)c->nbr,
                        (const int *)c->nact, (const BoxState *)(c->box + c->cur), pack_geo(c), c->pot, c->part_f,
                        spec_of(c), (const float *)c->fetab);
}

template <typename T, int POTM, bool P32>
static void density_dispatch_p(mdk_ctx *c, int do_embed) {
  // embedding-energy partials are value PART_EMB of force()'s array
  double *pe = c->part_f + (size_t)PART_EMB * nblocks(c->n);
  if (c->part_due)
    k_density_part<T, POTM, P32><<<nblocks(c->n), MDK_BLOCK, 0, c->stream>>>(
        c->atoms, c->pos4, p4f_of(c), c->n, c->nbr, c->numn, c->box + c->cur, pack_geo(c), c->pot, do_embed, pe,
        c->nact);
  else
    k_density<T, POTM, P32><<<nblocks(c->n), MDK_BLOCK, 0, c->stream>>>(
        c->atoms, c->pos4, p4f_of(c), c->n, c->nbr, c->numn, c->box + c->cur, pack_geo(c), c->pot, do_embed, pe,
        c->nact);
  c->part_due = false;
}
template <typename T, int POTM>
static void density_dispatch(mdk_ctx *c, int do_embed) {
  if constexpr (sizeof(T) == 4) {
    if (c->rel) return density_dispatch_p<T, POTM, true>(c, do_embed);
  }
  density_dispatch_p<T, POTM, false>(c, do_embed);
}
template <typename T, bool MC, int POTM, bool P32>
static void force_dispatch_p(mdk_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  // tabulated pair terms interpolate across the range edge: those use the whole list
  hipExtLaunchKernelGGL(k_force<T, MC, POTM, P32>, dim3(nblocks(c->n)), dim3(MDK_BLOCK), 0, c->stream, ea, eb, 0,
                        c->atoms, (const double4 *)c->pos4, (const float4 *)p4f_of(c), (const int *)c->zat, c->n,
                        (const int *)c->nbr, (const int *)(POTM == POT_TABLE ? c->numn : c->nact),
                        (const BoxState *)(c->box + c->cur), pack_geo(c), c->pot, c->part_f, spec_of(c),
                        (const float *)c->fetab);
}
template <typename T, bool MC, int POTM>
static void force_dispatch(mdk_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  if constexpr (sizeof(T) == 4) {
    if (c->rel) return force_dispatch_p<T, MC, POTM, true>(c, ea, eb);
  }
  force_dispatch_p<T, MC, POTM, false>(c, ea, eb);
}
// lookup tables, iron_carbon (FP64 only: set-up refuses the FP32 form), the
// fetransmetal Fe-Fe knot counts, or any knot count
static int pot_mode(const mdk_ctx *c) {
  if (c->pot.use_lookup) return POT_TABLE;
  if (c->material == MDK_MAT_IRON_CARBON) return POT_FEC;
  if (c->pot.nv != MDK_FE_NV || c->pot.np != MDK_FE_NP) return POT_KNOTS;
  return (c->p.precision == MDK_PREC_FP32 && !c->fe_ok) ? POT_KNOTS : POT_FE;
}
template <typename T>
static void density_pot(mdk_ctx *c, int do_embed) {
  switch (pot_mode(c)) {
    case POT_TABLE: density_dispatch<T, POT_TABLE>(c, do_embed); break;
    case POT_FE: density_dispatch<T, POT_FE>(c, do_embed); break;
    case POT_FEC:
      if constexpr (sizeof(T) == 8) density_dispatch<T, POT_FEC>(c, do_embed);
      break;
    default: density_dispatch<T, POT_KNOTS>(c, do_embed); break;
  }
}
template <typename T, bool MC>
static void force_pot(mdk_ctx *c, hipEvent_t ea, hipEvent_t eb) {
  switch (pot_mode(c)) {
    case POT_TABLE: force_dispatch<T, MC, POT_TABLE>(c, ea, eb); break;
    case POT_FE: force_dispatch<T, MC, POT_FE>(c, ea, eb); break;
    case POT_FEC:
      if constexpr (sizeof(T) == 8) force_dispatch<T, MC, POT_FEC>(c, ea, eb);
      break;
    default: force_dispatch<T, MC, POT_KNOTS>(c, ea, eb); break;
  }
}

static int ensure_pos4(mdk_ctx *c) {
  if (c->pos4_valid) return MDK_OK;
  c->p4f_valid = true;
  k_pack<<<nblocks(c->n), MDK_BLOCK, 0, c->stream>>>(c->atoms, c->img, c->box + c->cur, c->n, 0, c->pos4,
                                                     p4f_of(c), c->cell_of, 0, 0, 0, pack_geo(c));
  c->pos4_valid = true;
  return check_launch("k_pack");
}

static int launch_density(mdk_ctx *c, int do_embed) {
  if (!(lean_path(c) && c->p4f_valid)) RCK(ensure_pos4(c));  // the lean passes read only the FP32 records
  TimedEv te;
  tbegin(c, TC_DENSITY, te);
  if (lean_path(c)) {
    lean_density(c, do_embed);
  } else if (c->p.precision == MDK_PREC_FP32) {
    density_pot<float>(c, do_embed);
  } else {
    density_pot<double>(c, do_embed);
  }
  tend(c, te);
  return check_launch("k_density");
}

static int launch_force(mdk_ctx *c, int fin_mode) {
  if (!c->kin_fresh) {  // no drift, predictor or fused kick prepared kinten's sums
    k_kin<<<c->nblk, MDK_BLOCK, 0, c->stream>>>(c->atoms, c->zat, c->n, c->box + c->cur, spec_of(c),
                                                kin_part(c));
    RCK(check_launch("k_kin"));
  }
  c->kin_fresh = false;  // consumed by this force evaluation
  TimedEv te;
  tbegin_ext(c, TC_FORCE, te);  // the force kernel's own start and end
  hipEvent_t ea = te.cls >= 0 ? te.a : nullptr, eb = te.cls >= 0 ? te.b : nullptr;
  const bool f32 = c->p.precision == MDK_PREC_FP32, mc = c->masschk;
  if (lean_path(c)) {
    lean_force(c, ea, eb);
  } else if (f32) {
    if (mc) force_pot<float, true>(c, ea, eb); else force_pot<float, false>(c, ea, eb);
  } else {
    if (mc) force_pot<double, true>(c, ea, eb); else force_pot<double, false>(c, ea, eb);
  }
  tpush(c, te);
  RCK(check_launch("k_force"));
  FinForce F = {c->stage_f, c->fin_cnt, c->box + c->cur, fin_mode, c->p.ivol, c->p.press, c->p.bdel2,
                c->p.deltat, nullptr, c->kstage, FinKick{}};
  const bool pend = c->kick_pending;  // the previous step's kick sums finish here
  if (pend) {
    F.kpart = c->part_k;
    F.kick = c->kick_pend;
  }
  k_finalize_force<<<(c->nblk + 63) / 64, MDK_FIN_BLOCK, 0, c->stream>>>(c->part_f, c->nblk, F);
  RCK(check_launch("k_finalize_force"));
  if (pend) {
    c->kick_pending = false;
    HIPCK(hipEventRecord(c->mon_ev[c->kick_slot], c->stream));  // that step's monitor sample is written
  }
  return MDK_OK;
}

// Finishes a deferred half kick on its own (the last step of an MD call).
static int flush_kick(mdk_ctx *c) {
  if (!c->kick_pending) return MDK_OK;
  c->kick_pending = false;
  k_finalize_kick<<<1, MDK_FOLD_BLOCK, 0, c->stream>>>(c->part_k, c->nblk, c->kick_pend);
  RCK(check_launch("k_finalize_kick"));
  HIPCK(hipEventRecord(c->mon_ev[c->kick_slot], c->stream));
  return MDK_OK;
}

// Checks the flags of a deferred build (the density pass on it is queued and
// may run meanwhile; a flagged list yields wrong values but valid indices).
// A flagged build is discarded: the list is built again with every attempt
// checked (grown capacity, grid from the exact box) and the density pass
// repeated, before any force uses it.
static int confirm_list(mdk_ctx *c) {
  if (!c->list_pending) return MDK_OK;
  c->list_pending = false;
  HIPCK(hipEventSynchronize(c->list_ev));
  const volatile int *fl = c->flags_host;
  if (fl[0] == 0 && fl[2] == 0) return MDK_OK;
  HIPCK(hipStreamSynchronize(c->stream));
  c->rebuilds--;
  c->b0_lag_step = -1;
  RCK(rebuild_full(c, false, true));
  return launch_density(c, 1);
}

// force() bookkeeping of src/dynamics_m.F90:186-198 plus the skin criterion
static int force_body(mdk_ctx *c, int fin_mode, bool skin_due) {
  if (!c->pot_set) return set_err(MDK_ERR_STATE, "potential not set");
  c->ntc = c->ntc + 1;
  bool ref_due = (c->ntc == 1);
  if (c->ntc == c->p.ntcm) c->ntc = 0;
  bool due = !c->list_valid || skin_due || (ref_due && c->p.rebuild == 0);
  if (due) RCK(rebuild_full(c, true, true));
  RCK(launch_density(c, 1));
  RCK(confirm_list(c));
  return launch_force(c, fin_mode);
}

// Half kick (mode 0), sums only (1) or initial accelerations (2).  With
// fuse_drift (mode 0) the kernel also performs the next step's drift; the
// finishing block then completes this step's sums on the current box copy and
// the next md_step skips its drift.
// defer (mode 0): the sums' finishing is left to the next force's finishing
// kernel (or flush_kick); the caller records the monitor event then.
static int launch_kick(mdk_ctx *c, int mode, double *trace_row_dev, double *mon, bool fuse_drift = false,
                       bool defer = false) {
  TimedEv te;
  tbegin(c, TC_KICK, te);
  const int nxt = 1 - c->cur;
  const bool fused = fuse_drift && mode == 0;
  if (fused)
    k_kick<true><<<c->nblk, MDK_BLOCK, 0, c->stream>>>(
        c->atoms, c->zat, c->n, c->box + c->cur, c->p.ivol, mode, c->sref, c->part_k,
        lean_path(c) ? nullptr : c->pos4, p4f_of(c), c->img, c->cell_of, pack_geo(c), c->box + nxt, spec_of(c),
        kin_part(c));
  else
    k_kick<false><<<c->nblk, MDK_BLOCK, 0, c->stream>>>(
        c->atoms, c->zat, c->n, c->box + c->cur, c->p.ivol, mode, c->sref, c->part_k, nullptr, nullptr,
        nullptr, nullptr, pack_geo(c), nullptr, spec_of(c), nullptr);
  tend(c, te);
  RCK(check_launch("k_kick"));
  if (mode != 2) {  // initial accelerations have no sums
    const FinKick F = {c->box + c->cur, fused ? c->box + nxt : nullptr, trace_row_dev,
                       mon, c->p.deltat, c->p.bdel2, c->p.press};
    if (defer && mode == 0) {
      c->kick_pend = F;
      c->kick_pending = true;
    } else {
      k_finalize_kick<<<1, MDK_FOLD_BLOCK, 0, c->stream>>>(c->part_k, c->nblk, F);
      RCK(check_launch("k_finalize_kick"));
    }
  }
  if (fused) {
    c->pos4_valid = !lean_path(c);
    c->p4f_valid = true;
    c->cur = nxt;
    c->drifted = true;
    c->kin_fresh = true;
  }
  return MDK_OK;
}

static int launch_drift(mdk_ctx *c) {
  if (c->drifted) {  // performed by the previous step's fused kick
    c->drifted = false;
    return MDK_OK;
  }
  TimedEv te;
  tbegin(c, TC_DRIFT, te);
  const int nxt = 1 - c->cur;
  k_drift<<<c->nblk, MDK_BLOCK, 0, c->stream>>>(c->atoms, lean_path(c) ? nullptr : c->pos4, p4f_of(c), c->img,
                                                c->cell_of, pack_geo(c), c->n, c->box + c->cur, c->box + nxt,
                                                c->p.ivol, c->zat, spec_of(c), kin_part(c));
  c->pos4_valid = !lean_path(c);
  c->p4f_valid = true;
  c->kin_fresh = true;
  tend(c, te);
  c->cur = nxt;
  return check_launch("k_drift");
}

// Reads the monitor sample of step s-3 (its kernels have completed or the
// host waits for them): the box for the next grid plan and the skin bound.
// In the MD loop a step's sample is finished by the next step's force
// finishing kernel, so the host stays about two steps ahead of the GPU.
static const double *poll_monitor(mdk_ctx *c) {
  const long s = c->step_counter;
  if (s < 3) return nullptr;
  const long probe = s - 3;
  const int slot = (int)(probe & 3);
  hipEventSynchronize(c->mon_ev[slot]);
  const double *m = c->mon_host + MON_STRIDE * slot;
  for (int k = 0; k < 9; k++) c->b0_lag[k] = m[MON_B0 + k];
  c->b0_lag_step = probe;
  return m;
}

// Frobenius norm of A Bref^-1 - s I for the monitor's Bref (column-major 3x3)
static double strain_norm(const double *a, const double *bref, double s) {
  const double det = bref[0] * (bref[4] * bref[8] - bref[7] * bref[5]) -
                     bref[3] * (bref[1] * bref[8] - bref[7] * bref[2]) +
                     bref[6] * (bref[1] * bref[5] - bref[4] * bref[2]);
  double inv[9];  // cofactors / determinant, column-major
  inv[0] = (bref[4] * bref[8] - bref[7] * bref[5]) / det;
  inv[3] = -(bref[3] * bref[8] - bref[6] * bref[5]) / det;
  inv[6] = (bref[3] * bref[7] - bref[6] * bref[4]) / det;
  inv[1] = -(bref[1] * bref[8] - bref[7] * bref[2]) / det;
  inv[4] = (bref[0] * bref[8] - bref[6] * bref[2]) / det;
  inv[7] = -(bref[0] * bref[7] - bref[6] * bref[1]) / det;
  inv[2] = (bref[1] * bref[5] - bref[4] * bref[2]) / det;
  inv[5] = -(bref[0] * bref[5] - bref[3] * bref[2]) / det;
  inv[8] = (bref[0] * bref[4] - bref[3] * bref[1]) / det;
  double e2 = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double x = (i == j) ? -s : 0.0;
      for (int k = 0; k < 3; k++) x += a[i + 3 * k] * inv[k + 3 * j];
      e2 += x * x;
    }
  return sqrt(e2);
}
// box strain since the list was built (B0 Bref^-1 - I) and that of the next
// drift's box change ((b1 + b2) Bref^-1)
static double strain_since_build(const double *m) { return strain_norm(m + MON_B0, m + MON_BREF, 1.0); }
static double strain_of_step(const double *m) { return strain_norm(m + MON_DB, m + MON_BREF, 0.0); }

// The list a step uses holds every pair within the potential's range while
// each atom has moved less than d and the box has strained by less than e
// (monitor sample: d^2 = m[0], e = strain_since_build) since the list was
// built, with 2 d + e r_list < skin: no pair separation has then changed by
// the skin.
static double skin_used(const mdk_ctx *c, const double *m) {
  return sqrt(fmax(m[0], 0.0)) + 0.5 * strain_since_build(m) * c->r_list;  // against skin / 2
}
static int skin_violation(const mdk_ctx *c, const double *m, long step) {
  return set_err(MDK_ERR_SKIN,
                 "step %ld: an atom moved %.4g A and the box strained %.3g since the list was built "
                 "(bound: skin/2 = %.4g A); pairs may have been missed",
                 step, sqrt(fmax(m[0], 0.0)), strain_since_build(m), 0.5 * c->skin);
}

// Checks the completed sample of step s-3 against the skin bound (an error if
// it was exceeded), then decides whether the next force evaluation could see
// it exceeded: that sample's displacement and strain plus four times its
// step's (m[1]: squared displacement of the next drift, strain_of_step: that
// of the next drift's box change) -- three drifts since the sample and one
// to spare.
static int skin_check(mdk_ctx *c, bool *due) {
  const double *m = poll_monitor(c);
  *due = !c->list_valid;
  if (!m) return MDK_OK;
  if (skin_used(c, m) >= 0.5 * c->skin) return skin_violation(c, m, c->step_counter - 3);
  if (c->step_counter - c->last_rebuild_step < 4) return MDK_OK;  // the sample predates the list
  const double stp = sqrt(fmax(m[1], 0.0)) + 0.5 * strain_of_step(m) * c->r_list;
  *due = *due || skin_used(c, m) + 4.0 * stp >= 0.5 * c->skin;
  return MDK_OK;
}

// One moldin step.  allow_fuse lets the closing kick also perform the next
// step's drift, except when relink follows this step (it must see the
// positions of this step).
static int launch_predic(mdk_ctx *c) {
  c->x3_live = true;
  TimedEv te;
  tbegin(c, TC_DRIFT, te);
  const int nxt = 1 - c->cur;
  k_predic<<<c->nblk, MDK_BLOCK, 0, c->stream>>>(c->atoms, c->pos4, p4f_of(c), c->img, c->cell_of, pack_geo(c),
                                                 c->n, c->box + c->cur, c->box + nxt, c->p.ivol,
                                                 c->pc_primed ? 1 : 0, c->zat, spec_of(c), kin_part(c));
  tend(c, te);
  c->pos4_valid = c->p4f_valid = true;
  c->kin_fresh = true;
  c->cur = nxt;
  return check_launch("k_predic");
}

static int launch_correc(mdk_ctx *c) {
  c->x3_live = true;
  TimedEv te;
  tbegin(c, TC_KICK, te);
  k_correc<<<c->nblk, MDK_BLOCK, 0, c->stream>>>(c->atoms, c->zat, c->n, c->box + c->cur, c->p.ivol,
                                                 c->p.bdel2, spec_of(c));
  tend(c, te);
  c->pos4_valid = c->p4f_valid = false;  // positions and box moved
  c->pc_primed = true;
  return check_launch("k_correc");
}

// One moldin step with iverlet = 0 (src/moldin.F90:754-758): PREDIC, FORCE,
// CORREC, then update_therm_avs (box part, then the sums with the monitor).
static int md_step_gear(mdk_ctx *c, double *trace_row_dev) {
  bool skin_due;
  RCK(skin_check(c, &skin_due));
  RCK(launch_predic(c));
  RCK(force_body(c, FIN_FORCE, skin_due));
  RCK(launch_correc(c));
  k_therm_box<<<1, 64, 0, c->stream>>>(c->box + c->cur, c->p.ivol);
  RCK(check_launch("k_therm_box"));
  int slot = (int)(c->step_counter & 3);
  RCK(launch_kick(c, 1, trace_row_dev, c->mon_dev + MON_STRIDE * slot));
  HIPCK(hipEventRecord(c->mon_ev[slot], c->stream));
  c->step_counter++;
  return MDK_OK;
}

// The half kick's sums are finished by the next step's force finishing kernel
// (one launch fewer per step); mdk_md_steps finishes the last one itself.
static int md_step(mdk_ctx *c, double *trace_row_dev, bool fused_therm, bool allow_fuse = false) {
  if (c->iverlet == 0) return md_step_gear(c, trace_row_dev);
  bool skin_due;
  RCK(skin_check(c, &skin_due));
  c->pc_primed = true;  // velocityverlet refreshes quench_m's parameter copy
  RCK(launch_drift(c));
  RCK(force_body(c, fused_therm ? FIN_VV_THERM : FIN_VV, skin_due));
  int slot = (int)(c->step_counter & 3);
  RCK(launch_kick(c, 0, trace_row_dev, c->mon_dev + MON_STRIDE * slot, allow_fuse && c->ntc != 0, true));
  c->kick_slot = slot;
  c->step_counter++;
  return MDK_OK;
}

static int check_device_status(mdk_ctx *c) {
  BoxState bs;
  RCK(read_box(c, &bs));
  if (bs.status != 0) return set_err(bs.status, "singular box matrix in matrix_invert");
  return MDK_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static mdk_ctx *H(void *h) { return (mdk_ctx *)h; }

extern "C" int mdk_create(void **handle, const mdk_params *p) {
  if (!handle || !p) return set_err(MDK_ERR_BAD_ARG, "null argument");
  if (p->nm < 1 || p->nnbrs < 1 || p->ntcm < 1 || p->nlcx < 3 || p->nlcy < 3 || p->nlcz < 3 ||
      p->deltat <= 0.0 || p->rnear <= 0.0)
    return set_err(MDK_ERR_BAD_ARG, "bad parameters (nm=%d nnbrs=%d ntcm=%d)", p->nm, p->nnbrs, p->ntcm);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(MDK_ERR_NO_DEVICE, "no HIP device");
  void *mem = ::operator new(sizeof(mdk_ctx));
  memset(mem, 0, sizeof(mdk_ctx));
  mdk_ctx *c = (mdk_ctx *)mem;
  new (&c->evs) std::vector<TimedEv>();
  new (&c->evpool) std::vector<hipEvent_t>();
  c->p = *p;
  c->n = p->nm;
  c->iverlet = 1;  // velocity Verlet unless mdk_set_integrator selects the predictor-corrector
  const int n = c->n;
  switch (p->ivol) {
    case 0: case 1: case 4: c->pbc = 7; break;
    case 2: c->pbc = 3; break;
    case 5: c->pbc = 4; break;
    default: c->pbc = 0; break;
  }
  hipGetDevice(&c->device);
  c->maxn = ((p->max_full > 0 ? p->max_full : 64) + 3) & ~3;

  c->skin = p->rskin > 0.0 ? p->rskin : 0.3;
  c->rnear2_exact = near_threshold(p->rnear);
  c->nblk = nblocks(n);
  int rc = MDK_OK;
#define ALLOC(ptr, bytes)                                                          \
  do {                                                                             \
    if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {                       \
      rc = set_err(MDK_ERR_HIP, "hipMalloc %zu bytes failed", (size_t)(bytes));    \
      goto fail;                                                                   \
    }                                                                              \
  } while (0)
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    rc = set_err(MDK_ERR_HIP, "stream create failed");
    goto fail;
  }
  ALLOC(c->atoms, sizeof(double) * (size_t)F_N * n);
  ALLOC(c->atoms_alt, sizeof(double) * (size_t)F_N * n);
  ALLOC(c->zat, sizeof(int) * n);
  ALLOC(c->zat_alt, sizeof(int) * n);
  ALLOC(c->orig, sizeof(int) * n);
  ALLOC(c->orig_alt, sizeof(int) * n);
  ALLOC(c->slot_of, sizeof(int) * n);
  ALLOC(c->sref, sizeof(float) * 3 * (size_t)n);
  ALLOC(c->ic, sizeof(int) * n);
  ALLOC(c->numn_ref, sizeof(int) * n);
  ALLOC(c->nlist_ref, sizeof(int) * (size_t)n * p->nnbrs);
  ALLOC(c->nbr, sizeof(int) * (size_t)c->maxn * n);
  ALLOC(c->numn, sizeof(int) * n);
  ALLOC(c->nact, sizeof(int) * n);
  ALLOC(c->cell_of, sizeof(int) * n);
  ALLOC(c->ref_key, sizeof(int) * n);
  ALLOC(c->rank_orig, sizeof(int) * n);
  ALLOC(c->rank_slot, sizeof(int) * n);
  ALLOC(c->members, sizeof(int) * n);
  ALLOC(c->dflags, sizeof(int) * 4);
  ALLOC(c->box, sizeof(BoxState) * 2);
  ALLOC(c->pos4, sizeof(double4) * (size_t)n);
  ALLOC(c->p4f, sizeof(float4) * (size_t)n);
  ALLOC(c->img, sizeof(int) * 3 * (size_t)n);
  ALLOC(c->part_f, sizeof(double) * (size_t)c->nblk * NPART_F);
  ALLOC(c->part_k, sizeof(double) * (size_t)c->nblk * NPART_KICK);
  ALLOC(c->stage_f, sizeof(double) * (size_t)(c->nblk / 64 + 1) * NPART_F);
  ALLOC(c->stage_k, sizeof(double) * (size_t)(c->nblk / 64 + 1) * NPART_KICK);
  ALLOC(c->kstage, sizeof(double) * (size_t)(c->nblk / 64 + 1) * NPART_KICK);
  ALLOC(c->fin_cnt, sizeof(unsigned) * 2);
  ALLOC(c->red, sizeof(double) * 4);
  ALLOC(c->spec_tab, sizeof(double) * 2 * MDK_NSPEC);
  if (hipMemset(c->fin_cnt, 0, sizeof(unsigned) * 2) != hipSuccess) {
    rc = set_err(MDK_ERR_HIP, "counter initialisation failed");
    goto fail;
  }
  if (hipHostMalloc((void **)&c->mon_host, sizeof(double) * 4 * MON_STRIDE, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->mon_dev, c->mon_host, 0) != hipSuccess) {
    rc = set_err(MDK_ERR_HIP, "mapped host buffer failed");
    goto fail;
  }
  for (int k = 0; k < 4; k++) hipEventCreateWithFlags(&c->mon_ev[k], hipEventDisableTiming);
  if (hipHostMalloc((void **)&c->flags_host, sizeof(int) * 4, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->flags_map, c->flags_host, 0) != hipSuccess ||
      hipEventCreateWithFlags(&c->list_ev, hipEventDisableTiming) != hipSuccess) {
    rc = set_err(MDK_ERR_HIP, "mapped flag buffer failed");
    goto fail;
  }
  hipMemsetAsync(c->atoms, 0, sizeof(double) * (size_t)F_N * n, c->stream);
  // both buffers start at zero: a re-sort leaves X3..Z3 unmoved while they
  // are zero in both (x3_live false; scratch uses of atoms_alt stay in X0..Z0)
  hipMemsetAsync(c->atoms_alt, 0, sizeof(double) * (size_t)F_N * n, c->stream);
  hipMemsetAsync(c->zat, 0, sizeof(int) * n, c->stream);
  hipMemsetAsync(c->box, 0, sizeof(BoxState) * 2, c->stream);
  hipMemsetAsync(c->ic, 0, sizeof(int) * n, c->stream);
  hipMemsetAsync(c->img, 0, sizeof(int) * 3 * (size_t)n, c->stream);
  hipMemsetAsync(c->numn_ref, 0, sizeof(int) * n, c->stream);
  hipMemsetAsync(c->numn, 0, sizeof(int) * n, c->stream);
  hipMemsetAsync(c->nact, 0, sizeof(int) * n, c->stream);
  k_iota<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->orig, n);
  k_iota<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->slot_of, n);
  memset(c->mon_host, 0, sizeof(double) * 4 * MON_STRIDE);
  c->b0_lag_step = -1;
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    rc = set_err(MDK_ERR_HIP, "init sync failed");
    goto fail;
  }
  *handle = c;
  return MDK_OK;
fail:
  mdk_destroy(c);
  return rc;
#undef ALLOC
}

extern "C" int mdk_destroy(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return MDK_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->pos4, c->p4f, c->img, c->atoms, c->atoms_alt, c->zat, c->zat_alt, c->orig, c->orig_alt, c->slot_of,
                  c->sref, c->ic, c->numn_ref, c->nlist_ref, c->nbr, c->numn, c->nact, c->cell_of,
                  c->members, c->dflags, c->box, c->part_f, c->part_k, c->stage_f, c->stage_k, c->fin_cnt, c->cell_count,
                  c->cell_start, c->cell_fill, c->tabs, c->trace_dev, c->lblocks, c->spec_tab, c->red, c->fetab,
                  c->ref_key, c->rank_orig, c->rank_slot, c->kstage};
  for (void *q : ptrs)
    if (q) hipFree(q);
  if (c->mon_host) hipHostFree(c->mon_host);
  if (c->flags_host) hipHostFree(c->flags_host);
  if (c->list_ev) hipEventDestroy(c->list_ev);
  free(c->h_mass);
  free(c->h_del2);
  free(c->h_z);
  free(c->h_l0);
  free(c->h_link);
  free(c->h_ic);
  for (int k = 0; k < 4; k++)
    if (c->mon_ev[k]) hipEventDestroy(c->mon_ev[k]);
  for (auto &e : c->evs) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  for (auto e : c->evpool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  c->evs.~vector();
  c->evpool.~vector();
  ::operator delete(c);
  return MDK_OK;
}

extern "C" int mdk_set_potential(void *h, int nv, const double *av, const double *rv, int np,
                                 const double *ap, const double *rp) {
  mdk_ctx *c = H(h);
  if (!c || nv < 1 || nv > MDK_MAXV || np < 1 || np > MDK_MAXP || !av || !rv || !ap || !rp)
    return set_err(MDK_ERR_BAD_ARG, "bad potential coefficients");
  hostpot::Knots &k = c->knots;
  memset(&k, 0, sizeof(k));
  k.nv = nv;
  k.np = np;
  for (int i = 0; i < nv; i++) { k.av[i] = av[i]; k.rv[i] = rv[i]; }
  for (int i = 0; i < np; i++) { k.ap[i] = ap[i]; k.rp[i] = rp[i]; }
  hostpot::set_up(k);
  PotDev &P = c->pot;
  const int keep_lookup = P.use_lookup, keep_n = P.nlookup;
  const double rint = hostpot::fill_potdev(P, k);
  float fetab[MDK_FE_TAB];
  c->fe_ok = hostpot::build_fe_table(k, fetab, P.fe_binh, P.fe_binc);
  if (c->fe_ok) {
    if (!c->fetab) HIPCK(hipMalloc(&c->fetab, sizeof(fetab)));
    HIPCK(hipMemcpy(c->fetab, fetab, sizeof(fetab), hipMemcpyHostToDevice));
  }
  c->material = MDK_MAT_FETRANSMETAL;
  c->r_int_exact = rint;
  c->r_list = rint + c->skin;
  c->pot_set = true;
  c->list_valid = false;
  if (keep_lookup) return mdk_set_lookup(h, keep_n > 0 ? keep_n : 5000);
  return MDK_OK;
}

extern "C" int mdk_set_iron_carbon(void *h, const double *iron_in) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  if (c->p.precision == MDK_PREC_FP32)
    return set_err(MDK_ERR_BAD_ARG, "the FP32 form is built for the fetransmetal potential");
  PotDev &P = c->pot;
  const int keep_lookup = P.use_lookup, keep_n = P.nlookup;
  hostpot::fec_coef(c->fec, iron_in);
  memset(&c->knots, 0, sizeof(c->knots));
  c->material = MDK_MAT_IRON_CARBON;
  const double rint = hostpot::fill_potdev_fec(P, c->fec);
  c->r_int_exact = rint;
  c->r_list = rint + c->skin;
  c->pot_set = true;
  c->list_valid = false;
  if (keep_lookup) return mdk_set_lookup(h, keep_n > 0 ? keep_n : 5000);
  return MDK_OK;
}

extern "C" int mdk_set_lookup(void *h, int nlookup) {
  mdk_ctx *c = H(h);
  if (!c || !c->pot_set) return set_err(MDK_ERR_STATE, "set the potential before the lookup tables");
  PotDev &P = c->pot;
  if (nlookup <= 0) {
    P.use_lookup = 0;
    P.r_int = c->r_int_exact;
    c->r_list = P.r_int + c->skin;
    c->list_valid = false;
    return MDK_OK;
  }
  hostpot::Material mat;
  mat.id = c->material;
  mat.k = c->knots;
  mat.f = c->fec;
  // src/lookup_m.F90:161-251 with the material's range (src/fetransmetal.F90:589-593,
  // src/iron_carbon.F90:93 through get_supported_potential_range) and potential_m's rho range (:185)
  P.nlookup = nlookup;
  P.r_min = 0.0; P.r_max = mat.range();
  P.vp_range = P.r_max - P.r_min;
  P.vp_reciprange = 1.0 / P.vp_range;
  P.vp_recipnlkp = 1.0 / (double)nlookup;
  P.vp_recipstep = (double)nlookup / P.vp_range;
  P.rho_min = 0.0; P.rho_max = 10000.0;
  P.e_range = P.rho_max - P.rho_min;
  P.e_reciprange = 1.0 / P.e_range;
  P.e_recipnlkp = 1.0 / (double)nlookup;
  P.e_recipstep = (double)nlookup / P.e_range;
  const size_t m = (size_t)nlookup + 1;
  std::vector<double> t(6 * m);
  size_t last = 0;  // last node where a pair table is nonzero
  for (size_t i = 0; i < m; i++) {
    double r = P.r_min + (double)i * P.vp_range * P.vp_recipnlkp;
    t[0 * m + i] = mat.vee(r);
    t[1 * m + i] = mat.dvee(r);
    t[2 * m + i] = mat.phi(r);
    t[3 * m + i] = mat.dphi(r);
    if (t[i] != 0.0 || t[m + i] != 0.0 || t[2 * m + i] != 0.0 || t[3 * m + i] != 0.0) last = i;
    double rh = P.rho_min + (double)i * P.e_range * P.e_recipnlkp;
    t[4 * m + i] = mat.emb(rh);
    t[5 * m + i] = mat.demb(rh);
  }
  // interpolation reaches one node beyond the last nonzero one: the list must
  // hold those pairs (two nodes of margin for the rounding of the index)
  P.r_int = std::max(c->r_int_exact, std::min(P.r_max, P.r_min + (double)(last + 2) * P.vp_range * P.vp_recipnlkp));
  c->r_list = P.r_int + c->skin;
  c->list_valid = false;
  if (c->tabs) hipFree(c->tabs);
  HIPCK(hipMalloc(&c->tabs, sizeof(double) * 6 * m));
  HIPCK(hipMemcpy(c->tabs, t.data(), sizeof(double) * 6 * m, hipMemcpyHostToDevice));
  P.tab_v = c->tabs; P.tab_dv = c->tabs + m; P.tab_p = c->tabs + 2 * m;
  P.tab_dp = c->tabs + 3 * m; P.tab_e = c->tabs + 4 * m; P.tab_de = c->tabs + 5 * m;
  P.use_lookup = 1;
  return MDK_OK;
}

extern "C" int mdk_potential_constants(void *h, double *out) {
  mdk_ctx *c = H(h);
  if (!c || !c->pot_set || !out) return set_err(MDK_ERR_STATE, "potential not set");
  if (c->material != MDK_MAT_FETRANSMETAL) return set_err(MDK_ERR_STATE, "set_up constants belong to fetransmetal");
  out[0] = c->knots.bc1; out[1] = c->knots.bc2;
  for (int i = 0; i < 4; i++) out[2 + i] = c->knots.sc[i];
  out[6] = c->knots.pcr; out[7] = c->knots.pcv;
  return MDK_OK;
}

extern "C" int mdk_eval_potential(void *h, int which, const double *x, double *y, int count) {
  mdk_ctx *c = H(h);
  if (!c || !c->pot_set) return set_err(MDK_ERR_STATE, "potential not set");
  if (which < 0 || which > 8 || count < 0 || (count > 0 && (!x || !y))) return set_err(MDK_ERR_BAD_ARG, "bad eval args");
  if (which >= 6 && (!c->fe_ok || c->material != MDK_MAT_FETRANSMETAL))
    return set_err(MDK_ERR_BAD_ARG, "no FP32 interval table for this potential");
  if (count == 0) return MDK_OK;
  double *d = nullptr;
  HIPCK(hipMalloc(&d, sizeof(double) * 2 * (size_t)count));
  HIPCK(hipMemcpyAsync(d, x, sizeof(double) * count, hipMemcpyHostToDevice, c->stream));
  k_eval_pot<<<nblocks(count), MDK_BLOCK, 0, c->stream>>>(c->pot, which, d, d + count, count, c->fetab);
  int rc = check_launch("k_eval_pot");
  hipMemcpyAsync(y, d + count, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  hipFree(d);
  return rc;
}

extern "C" int mdk_energy_calc(void *h, double *en_atom) {
  mdk_ctx *c = H(h);
  if (!c || !en_atom) return set_err(MDK_ERR_BAD_ARG, "bad energy_calc arguments");
  if (!c->pot_set || !c->list_valid)
    return set_err(MDK_ERR_STATE, "energy_calc needs the neighbour list and densities of a force evaluation");
  const int n = c->n;
  RCK(ensure_pos4(c));
  double *en = fld(c->atoms_alt, 1, n), *out = fld(c->atoms_alt, 0, n);
  const dim3 g(nblocks(n)), b(MDK_BLOCK);
  const BoxState *bs = c->box + c->cur;
  if (c->material == MDK_MAT_IRON_CARBON)
    k_energy_calc<POT_FEC><<<g, b, 0, c->stream>>>(c->atoms, c->pos4, n, c->nbr, c->numn, bs, pack_geo(c), c->pot, en);
  else if (c->pot.nv == MDK_FE_NV && c->pot.np == MDK_FE_NP)
    k_energy_calc<POT_FE><<<g, b, 0, c->stream>>>(c->atoms, c->pos4, n, c->nbr, c->numn, bs, pack_geo(c), c->pot, en);
  else
    k_energy_calc<POT_KNOTS><<<g, b, 0, c->stream>>>(c->atoms, c->pos4, n, c->nbr, c->numn, bs, pack_geo(c), c->pot, en);
  RCK(check_launch("k_energy_calc"));
  k_gather_d<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(out, en, c->slot_of, n);
  RCK(check_launch("k_gather_d"));
  HIPCK(hipMemcpyAsync(en_atom, out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_atomic_stress(void *h, double *sx, double *sy, double *sz) {
  mdk_ctx *c = H(h);
  if (!c || !sx || !sy || !sz) return set_err(MDK_ERR_BAD_ARG, "bad atomic_stress arguments");
  if (!c->pot_set || !c->half_valid)
    return set_err(MDK_ERR_STATE, "atomic_stress needs the reference half list and densities of a force evaluation");
  const int n = c->n;
  double *o = fld(c->atoms_alt, 0, n);  // three scratch columns
  const dim3 g(nblocks(n)), b(MDK_BLOCK);
  const BoxState *bs = c->box + c->cur;
#define MDK_STRESS_ARGS c->atoms, c->slot_of, n, c->numn_ref, c->nlist_ref, c->p.nnbrs, bs, c->pbc, c->pot, o, o + n, o + 2 * (size_t)n
  if (c->pot.use_lookup)
    k_atomic_stress<POT_TABLE><<<g, b, 0, c->stream>>>(MDK_STRESS_ARGS);
  else if (c->material == MDK_MAT_IRON_CARBON)
    k_atomic_stress<POT_FEC><<<g, b, 0, c->stream>>>(MDK_STRESS_ARGS);
  else if (c->pot.nv == MDK_FE_NV && c->pot.np == MDK_FE_NP)
    k_atomic_stress<POT_FE><<<g, b, 0, c->stream>>>(MDK_STRESS_ARGS);
  else
    k_atomic_stress<POT_KNOTS><<<g, b, 0, c->stream>>>(MDK_STRESS_ARGS);
#undef MDK_STRESS_ARGS
  RCK(check_launch("k_atomic_stress"));
  HIPCK(hipMemcpyAsync(sx, o, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipMemcpyAsync(sy, o + n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipMemcpyAsync(sz, o + 2 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

static const int field_map[MDK_NFIELDS] = {F_X0,  F_Y0,  F_Z0,   F_X1,     F_Y1,      F_Z1,   F_X2,
                                           F_Y2,  F_Z2,  F_FX,   F_FY,     F_FZ,      F_RHO,  F_AFRHO,
                                           F_DAFRHO, F_MASS, F_DEL2, F_X3, F_Y3, F_Z3};

// Species table of (mass, del2) when both are functions of the species.
static int keep_host_copy(double **dst, const double *src, int n) {
  if (!*dst) *dst = (double *)malloc(sizeof(double) * n);
  if (!*dst) return set_err(MDK_ERR_HIP, "host copy failed");
  memcpy(*dst, src, sizeof(double) * n);
  return MDK_OK;
}
static int update_species_table(mdk_ctx *c) {
  c->spec_ok = false;
  if (!c->h_mass || !c->h_del2 || !c->h_z) return MDK_OK;
  double tab[2 * MDK_NSPEC];
  bool seen[MDK_NSPEC] = {};
  for (int i = 0; i < c->n; i++) {
    const int z = c->h_z[i];
    if (z < 0 || z >= MDK_NSPEC) return MDK_OK;
    if (!seen[z]) {
      seen[z] = true;
      tab[2 * z] = c->h_mass[i];
      tab[2 * z + 1] = c->h_del2[i];
    } else if (memcmp(&tab[2 * z], &c->h_mass[i], sizeof(double)) != 0 ||
               memcmp(&tab[2 * z + 1], &c->h_del2[i], sizeof(double)) != 0) {
      return MDK_OK;  // per-atom values differ within a species
    }
  }
  for (int z = 0; z < MDK_NSPEC; z++)
    if (!seen[z]) tab[2 * z] = tab[2 * z + 1] = 0.0;
  HIPCK(hipMemcpy(c->spec_tab, tab, sizeof(tab), hipMemcpyHostToDevice));
  c->spec_ok = true;
  return MDK_OK;
}
extern "C" int mdk_upload(void *h, int field, const double *data) {
  mdk_ctx *c = H(h);
  if (!c || field < 0 || field >= MDK_NFIELDS || !data) return set_err(MDK_ERR_BAD_ARG, "bad upload");
  c->kin_fresh = false;  // velocities, masses or box may change: kinten's sums are recomputed
  const int n = c->n;
  double *tmp = fld(c->atoms_alt, 0, n);  // staging in the alternate buffer
  HIPCK(hipMemcpyAsync(tmp, data, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  k_scatter_d<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(fld(c->atoms, field_map[field], n), tmp, c->slot_of, n);
  RCK(check_launch("k_scatter_d"));
  HIPCK(hipStreamSynchronize(c->stream));
  if (field <= MDK_Z0) {  // positions replaced: rebuild (from the exact box) before the next force
    c->list_valid = false;
    c->b0_lag_step = -1;
  }
  if (field <= MDK_Z0 || field == MDK_DAFRHO) c->pos4_valid = c->p4f_valid = false;
  if (field == MDK_X3 || field == MDK_Y3 || field == MDK_Z3) c->x3_live = true;
  if (field == MDK_MASS) {
    // the per-pair test m_i + m_j > 0.1 (REAL(4) literal) can only fail with a mass <= 0.05
    bool any = false;
    for (int i = 0; i < n; i++) any = any || !(data[i] > 0.05);
    c->masschk = any;
  }
  if (field == MDK_MASS || field == MDK_DEL2) {
    RCK(keep_host_copy(field == MDK_MASS ? &c->h_mass : &c->h_del2, data, n));
    RCK(update_species_table(c));
  }
  return MDK_OK;
}

extern "C" int mdk_download(void *h, int field, double *data) {
  mdk_ctx *c = H(h);
  if (!c || field < 0 || field >= MDK_NFIELDS || !data) return set_err(MDK_ERR_BAD_ARG, "bad download");
  const int n = c->n;
  double *tmp = fld(c->atoms_alt, 0, n);
  k_gather_d<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(tmp, fld(c->atoms, field_map[field], n), c->slot_of, n);
  RCK(check_launch("k_gather_d"));
  HIPCK(hipMemcpyAsync(data, tmp, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_upload_species(void *h, const int *zn) {
  mdk_ctx *c = H(h);
  if (!c || !zn) return set_err(MDK_ERR_BAD_ARG, "bad species upload");
  c->kin_fresh = false;  // velocities, masses or box may change: kinten's sums are recomputed
  const int n = c->n;
  // the pair kernels evaluate the Fe-Fe functions: Fe atoms and vacancies only
  for (int i = 0; i < n; i++)
    if (zn[i] != 26 && zn[i] != 0) return set_err(MDK_ERR_BAD_ARG, "only Fe (Z=26) and vacancies (Z=0) are supported");
  HIPCK(hipMemcpyAsync(c->zat_alt, zn, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  k_scatter_i<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->zat, c->zat_alt, c->slot_of, n);
  RCK(check_launch("k_scatter_i"));
  HIPCK(hipStreamSynchronize(c->stream));
  if (!c->h_z) c->h_z = (int *)malloc(sizeof(int) * n);
  if (!c->h_z) return set_err(MDK_ERR_HIP, "host copy failed");
  memcpy(c->h_z, zn, sizeof(int) * n);
  return update_species_table(c);
}

static double *box_field(BoxState &s, int which) {
  switch (which) {
    case MDK_B0: return s.b0;
    case MDK_B1: return s.b1;
    case MDK_B2: return s.b2;
    case MDK_TK: return s.tk;
    case MDK_TC: return s.tc;
    case MDK_TG: return s.tg;
    case MDK_TP: return s.tp;
    case MDK_TGINV: return s.tginv;
    case MDK_TGDOT: return s.tgdot;
    case MDK_TGID: return s.tgid;
    case MDK_FB: return s.fb;
    case MDK_B3: return s.b3;
    default: return nullptr;
  }
}

extern "C" int mdk_set_tensor(void *h, int which, const double *t9) {
  mdk_ctx *c = H(h);
  if (!c || !t9 || which < 0 || which >= MDK_NTENSORS) return set_err(MDK_ERR_BAD_ARG, "bad tensor");
  c->kin_fresh = false;  // velocities, masses or box may change: kinten's sums are recomputed
  BoxState s;
  RCK(read_box(c, &s));
  memcpy(box_field(s, which), t9, 9 * sizeof(double));
  if (which == MDK_B0) {
    const double *b = s.b0;
    s.vol = TT(b, 1, 1) * (TT(b, 2, 2) * TT(b, 3, 3) - TT(b, 2, 3) * TT(b, 3, 2)) +
            TT(b, 2, 1) * (TT(b, 3, 2) * TT(b, 1, 3) - TT(b, 1, 2) * TT(b, 3, 3)) +
            TT(b, 3, 1) * (TT(b, 1, 2) * TT(b, 2, 3) - TT(b, 2, 2) * TT(b, 1, 3));
    c->list_valid = false;
    c->pos4_valid = c->p4f_valid = false;
    c->b0_lag_step = -1;  // the next list is planned from this box, not a lagged one
  }
  HIPCK(hipMemcpyAsync(c->box + c->cur, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_get_tensor(void *h, int which, double *t9) {
  mdk_ctx *c = H(h);
  if (!c || !t9 || which < 0 || which >= MDK_NTENSORS) return set_err(MDK_ERR_BAD_ARG, "bad tensor");
  BoxState s;
  RCK(read_box(c, &s));
  memcpy(t9, box_field(s, which), 9 * sizeof(double));
  return MDK_OK;
}

extern "C" int mdk_get_scalars(void *h, double *out) {
  mdk_ctx *c = H(h);
  if (!c || !out) return set_err(MDK_ERR_BAD_ARG, "bad scalars");
  BoxState s;
  RCK(read_box(c, &s));
  out[MDK_S_PE] = s.pe; out[MDK_S_KE] = s.ke; out[MDK_S_VOL] = s.vol;
  out[MDK_S_S1] = s.s1; out[MDK_S_S2] = s.s2; out[MDK_S_S3] = s.s3;
  out[MDK_S_TE] = s.te; out[MDK_S_H] = s.h; out[MDK_S_TH] = s.th; out[MDK_S_BOXKE] = s.boxke;
  out[MDK_S_MAXDISP] = sqrt(fmax(s.maxdisp2, 0.0));
  return MDK_OK;
}

extern "C" int mdk_set_ntc(void *h, int ntc) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  c->ntc = ntc;
  return MDK_OK;
}
extern "C" int mdk_get_ntc(void *h, int *ntc) {
  mdk_ctx *c = H(h);
  if (!c || !ntc) return set_err(MDK_ERR_BAD_ARG, "null handle");
  *ntc = c->ntc;
  return MDK_OK;
}

// Reference cells of every atom in reference order (host), from the device ic.
static int ic_to_host(mdk_ctx *c, int *ico) {
  const int n = c->n;
  k_gather_i<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->ref_key, c->ic, c->slot_of, n);
  RCK(check_launch("k_gather_i"));
  HIPCK(hipMemcpyAsync(ico, c->ref_key, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

static int chains_reserve(mdk_ctx *c) {
  const int m = c->p.nlcx * c->p.nlcy * c->p.nlcz;
  if (c->h_l0 && c->h_ncell == m) return MDK_OK;
  free(c->h_l0);
  c->h_l0 = (int *)calloc((size_t)m + 1, sizeof(int));
  if (!c->h_link) c->h_link = (int *)calloc((size_t)c->n + 1, sizeof(int));
  if (!c->h_ic) c->h_ic = (int *)calloc((size_t)c->n + 1, sizeof(int));
  if (!c->h_l0 || !c->h_link || !c->h_ic) return set_err(MDK_ERR_STATE, "out of host memory for the link-cell chains");
  c->h_ncell = m;
  return MDK_OK;
}

// linkup's chains (src/linkcell_m.F90:140-182): each cell's atoms in
// ascending order, the last linked back to the first
static void chains_linkup(mdk_ctx *c, const int *ico) {
  const int n = c->n, m = c->h_ncell;
  int *l0 = c->h_l0, *link = c->h_link;
  std::vector<int> last(m + 1, 0);
  for (int k = 0; k <= m; k++) l0[k] = 0;
  for (int i = 1; i <= n; i++) {
    const int ip = ico[i - 1];
    c->h_ic[i] = ip;
    if (last[ip]) link[last[ip]] = i; else l0[ip] = i;
    last[ip] = i;
  }
  for (int ip = 1; ip <= m; ip++)
    if (last[ip]) link[last[ip]] = l0[ip];
}

// relink's chain surgery (src/linkcell_m.F90:224-271), atoms in reference
// order: an atom whose cell changed leaves its old chain (the atom after it
// becomes that cell's starter) and enters its new cell right after the
// starter, becoming the starter itself.
static int chains_relink(mdk_ctx *c, const int *ico) {
  int *l0 = c->h_l0, *link = c->h_link, *icr = c->h_ic;
  for (int i = 1; i <= c->n; i++) {
    const int ip = ico[i - 1];
    if (ip == icr[i]) continue;
    const int me = icr[i];
    int j = l0[me], hops = 0;
    while (j > 0 && link[j] != i && hops++ < c->n) j = link[j];
    if (j <= 0 || link[j] != i) return set_err(MDK_ERR_STATE, "link-cell chains inconsistent at atom %d", i);
    link[j] = link[i];
    l0[me] = link[j];
    if (i == j) l0[me] = 0;
    j = l0[ip] == 0 ? i : l0[ip];
    link[i] = link[j];
    link[j] = i;
    icr[i] = ip;
    l0[ip] = i;
  }
  return MDK_OK;
}

// each atom's position in its cell's chain from L0, uploaded in reference order
static int chains_rank(mdk_ctx *c) {
  std::vector<int> rank(c->n, 0);
  for (int ip = 1; ip <= c->h_ncell; ip++) {
    const int s0 = c->h_l0[ip];
    if (!s0) continue;
    int i = s0, r = 0;
    do {
      rank[i - 1] = r++;
      i = c->h_link[i];
    } while (i != s0 && r <= c->n);
  }
  HIPCK(hipMemcpyAsync(c->rank_orig, rank.data(), sizeof(int) * c->n, hipMemcpyHostToDevice, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

static int launch_ref_cells(mdk_ctx *c) {
  const int n = c->n;
  k_ref_cells<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->atoms, c->atoms + n, c->atoms + 2 * (size_t)n, n,
                                                      c->p.nlcx, c->p.nlcy, c->p.nlcz, c->ic);
  return check_launch("k_ref_cells");
}

extern "C" int mdk_linkup(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  RCK(launch_ref_cells(c));
  RCK(chains_reserve(c));
  std::vector<int> ico(c->n);
  RCK(ic_to_host(c, ico.data()));
  chains_linkup(c, ico.data());
  RCK(chains_rank(c));
  c->linked = true;
  c->chains_exact = true;
  return MDK_OK;
}

extern "C" int mdk_relink(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  if (!c->linked || !c->chains_exact) return mdk_linkup(h);
  RCK(launch_ref_cells(c));
  std::vector<int> ico(c->n);
  RCK(ic_to_host(c, ico.data()));
  RCK(chains_relink(c, ico.data()));
  return chains_rank(c);
}

// The relink at the wrap of ntc inside mdk_md_steps: the device ic only (no
// host round trip in the loop).  The chain order it would give is not
// tracked; the next chain query or half list starts from linkup's chains.
static int relink_in_loop(mdk_ctx *c) {
  RCK(launch_ref_cells(c));
  c->chains_exact = false;
  return MDK_OK;
}

extern "C" int mdk_get_linkcells(void *h, int *l0, int *link, int *ic) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  if (!c->linked) return set_err(MDK_ERR_STATE, "link cells not set (call linkup)");
  if (!c->chains_exact) RCK(mdk_linkup(h));
  if (ic) memcpy(ic, c->h_ic + 1, sizeof(int) * c->n);
  if (l0) memcpy(l0, c->h_l0 + 1, sizeof(int) * c->h_ncell);
  if (link) memcpy(link, c->h_link + 1, sizeof(int) * c->n);
  return MDK_OK;
}

extern "C" int mdk_update_neighbourlist(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  if (!c->linked || !c->chains_exact) RCK(mdk_linkup(h));
  const int n = c->n, m = c->p.nlcx * c->p.nlcy * c->p.nlcz;
  // bin by reference cell (0-based key), in-cell order by chain rank; the
  // device list's own cell keys (cell_of) are left alone
  k_minus1<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->ic, c->ref_key, n);
  k_gather_i<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(c->rank_slot, c->rank_orig, c->orig, n);
  RCK(check_launch("k_gather_i"));
  RCK(bin_atoms(c, c->ref_key, m, c->rank_slot));
  HIPCK(hipMemsetAsync(c->dflags, 0, sizeof(int) * 4, c->stream));
  k_half_list<<<nblocks(n), MDK_BLOCK, 0, c->stream>>>(
      c->atoms, c->atoms + n, c->atoms + 2 * (size_t)n, c->ic, c->orig, c->rank_slot, c->cell_start, c->members, n,
      c->p.nlcx, c->p.nlcy, c->p.nlcz, c->box + c->cur, c->rnear2_exact, c->pbc, c->p.nnbrs,
      c->numn_ref, c->nlist_ref, c->dflags);
  RCK(check_launch("k_half_list"));
  int fl[4];
  HIPCK(hipMemcpyAsync(fl, c->dflags, sizeof(fl), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  if (fl[1]) return set_err(MDK_ERR_NBR_OVERFLOW, "NEIGHBOURLIST: EXCESS NEIGHBOURS! (nnbrs=%d)", c->p.nnbrs);
  c->half_valid = true;
  return MDK_OK;
}

extern "C" int mdk_get_neighbourlist(void *h, int *numn, int *nlist) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  if (numn) HIPCK(hipMemcpy(numn, c->numn_ref, sizeof(int) * c->n, hipMemcpyDeviceToHost));
  if (nlist) HIPCK(hipMemcpy(nlist, c->nlist_ref, sizeof(int) * (size_t)c->n * c->p.nnbrs, hipMemcpyDeviceToHost));
  return MDK_OK;
}

extern "C" int mdk_get_pairs(void *h, long *count, int *pi, int *pj, long cap) {
  mdk_ctx *c = H(h);
  if (!c || !count) return set_err(MDK_ERR_BAD_ARG, "null argument");
  const int n = c->n, nn = c->p.nnbrs;
  std::vector<int> numn(n), nl((size_t)n * nn);
  RCK(mdk_get_neighbourlist(h, numn.data(), nl.data()));
  long k = 0;
  for (int i = 0; i < n; i++)
    for (int q = 0; q < std::min(numn[i], nn); q++) {
      if (pi && pj && k < cap) {
        pi[k] = i;
        pj[k] = nl[(size_t)i * nn + q] - 1;
      }
      k++;
    }
  *count = k;
  return MDK_OK;
}

extern "C" int mdk_force(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  RCK(force_body(c, FIN_FORCE, false));
  HIPCK(hipStreamSynchronize(c->stream));
  return check_device_status(c);
}

extern "C" int mdk_rhoset(void *h) {
  mdk_ctx *c = H(h);
  if (!c || !c->pot_set) return set_err(MDK_ERR_STATE, "potential not set");
  if (!c->list_valid) RCK(rebuild_full(c));
  RCK(launch_density(c, 0));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_embed(void *h, double *pe_embed) {
  mdk_ctx *c = H(h);
  if (!c || !c->pot_set) return set_err(MDK_ERR_STATE, "potential not set");
  if (!c->list_valid) RCK(rebuild_full(c));
  RCK(launch_density(c, 1));
  if (pe_embed) {
    const int nb = block_count(c);
    std::vector<double> part(nb);
    HIPCK(hipMemcpyAsync(part.data(), c->part_f + (size_t)PART_EMB * nb, sizeof(double) * nb,
                         hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (double v : part) s += v;
    *pe_embed = s;
  }
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_kinten(void *h) { return mdk_force(h); }

extern "C" int mdk_md_init(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  RCK(force_body(c, FIN_INIT, false));
  RCK(launch_kick(c, 2, nullptr, nullptr));
  HIPCK(hipStreamSynchronize(c->stream));
  return check_device_status(c);
}

extern "C" int mdk_velocityverlet(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  bool skin_due;
  RCK(skin_check(c, &skin_due));
  c->pc_primed = true;
  RCK(launch_drift(c));
  RCK(force_body(c, FIN_VV, skin_due));
  // half kick without the thermodynamic sums' box part (tgid stays lagged)
  int slot = (int)(c->step_counter & 3);
  RCK(launch_kick(c, 0, nullptr, c->mon_dev + MON_STRIDE * slot));
  HIPCK(hipEventRecord(c->mon_ev[slot], c->stream));
  c->step_counter++;
  HIPCK(hipStreamSynchronize(c->stream));
  const double *m = c->mon_host + MON_STRIDE * slot;  // this step's sample
  if (skin_used(c, m) >= 0.5 * c->skin) return skin_violation(c, m, c->step_counter - 1);
  return check_device_status(c);
}

extern "C" int mdk_scalev(void *h, double nose, double rqke) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  c->kin_fresh = false;  // velocities, masses or box may change: kinten's sums are recomputed
  if (!(nose > 0.0)) return MDK_OK;  // only a positive nose acts (src/thermostat_m.F90:329)
  const int n = c->n, nb = nblocks(n);
  if (nb > c->nblk) return set_err(MDK_ERR_STATE, "partial buffers too small");
  const unsigned fb = (unsigned)((nb + 63) / 64);
  k_vsum<<<nb, MDK_BLOCK, 0, c->stream>>>(c->atoms, n, c->part_k);
  RCK(check_launch("k_vsum"));
  k_fold<3><<<fb, MDK_FIN_BLOCK, 0, c->stream>>>(c->part_k, nb, c->stage_k, c->fin_cnt + 1, c->red);
  RCK(check_launch("k_fold"));
  // mean over all atoms: sx / float(nm), a REAL(4) conversion of nm
  double s3[4];
  HIPCK(hipMemcpyAsync(s3, c->red, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  const double fn = (double)(float)n;
  const double mean[3] = {s3[0] / fn, s3[1] / fn, s3[2] / fn};
  HIPCK(hipMemcpyAsync(c->red, mean, sizeof(mean), hipMemcpyHostToDevice, c->stream));
  k_vsub_ke<<<nb, MDK_BLOCK, 0, c->stream>>>(c->atoms, c->zat, n, c->red, c->box + c->cur, spec_of(c), c->part_k);
  RCK(check_launch("k_vsub_ke"));
  k_fold<1><<<fb, MDK_FIN_BLOCK, 0, c->stream>>>(c->part_k, nb, c->stage_k, c->fin_cnt + 1, c->red + 3);
  RCK(check_launch("k_fold"));
  HIPCK(hipMemcpyAsync(s3 + 3, c->red + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  const double ke = s3[3] / (c->p.deltat * c->p.deltat) / 2.0;
  double snhf = 0.0;
  if (ke > 0.0) snhf = (ke - rqke) / (nose * (double)n);
  c->snhv = c->snhv - snhf;
  k_vscale<<<nb, MDK_BLOCK, 0, c->stream>>>(c->atoms, c->zat, n, spec_of(c), 1.0 + c->snhv);
  RCK(check_launch("k_vscale"));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_get_snhv(void *h, double *v) {
  mdk_ctx *c = H(h);
  if (!c || !v) return set_err(MDK_ERR_BAD_ARG, "bad snhv arguments");
  *v = c->snhv;
  return MDK_OK;
}

extern "C" int mdk_set_snhv(void *h, double v) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  c->snhv = v;
  return MDK_OK;
}

extern "C" int mdk_set_pressure(void *h, double press) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  c->p.press = press;
  return MDK_OK;
}

extern "C" int mdk_set_integrator(void *h, int iverlet) {
  mdk_ctx *c = H(h);
  if (!c || (iverlet != 0 && iverlet != 1)) return set_err(MDK_ERR_BAD_ARG, "iverlet must be 0 or 1");
  c->iverlet = iverlet;
  return MDK_OK;
}

extern "C" int mdk_predic(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  RCK(launch_predic(c));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_correc(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  c->kin_fresh = false;  // velocities, masses or box may change: kinten's sums are recomputed
  RCK(launch_correc(c));
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}

extern "C" int mdk_update_therm_avs(void *h, double *s1, double *s2, double *s3) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  k_therm_box<<<1, 64, 0, c->stream>>>(c->box + c->cur, c->p.ivol);
  RCK(check_launch("k_therm_box"));
  TimedEv te;
  tbegin(c, TC_KICK, te);
  k_kick<false><<<c->nblk, MDK_BLOCK, 0, c->stream>>>(
      c->atoms, c->zat, c->n, c->box + c->cur, c->p.ivol, 1, nullptr, c->part_k, nullptr, nullptr, nullptr,
      nullptr, pack_geo(c), nullptr, spec_of(c), nullptr);
  tend(c, te);
  RCK(check_launch("k_kick"));
  const FinKick F = {c->box + c->cur, nullptr, nullptr, nullptr,
                     c->p.deltat, c->p.bdel2, c->p.press};
  k_finalize_kick<<<1, MDK_FOLD_BLOCK, 0, c->stream>>>(c->part_k, c->nblk, F);
  RCK(check_launch("k_finalize_kick"));
  BoxState s;
  RCK(read_box(c, &s));
  if (s1) *s1 = s.s1;
  if (s2) *s2 = s.s2;
  if (s3) *s3 = s.s3;
  if (s.status) return set_err(s.status, "singular box matrix in matrix_invert");
  return MDK_OK;
}

extern "C" int mdk_md_steps(void *h, int nsteps, double *trace) {
  mdk_ctx *c = H(h);
  if (!c || nsteps < 0) return set_err(MDK_ERR_BAD_ARG, "bad md_steps arguments");
  if (trace && nsteps > c->trace_cap) {
    if (c->trace_dev) hipFree(c->trace_dev);
    HIPCK(hipMalloc(&c->trace_dev, sizeof(double) * 6 * (size_t)nsteps));
    c->trace_cap = nsteps;
  }
  int rc = MDK_OK;
  for (int s = 0; s < nsteps && rc == MDK_OK; s++) {
    rc = md_step(c, trace ? c->trace_dev + 6 * (size_t)s : nullptr, true, s + 1 < nsteps);
    // relink when the reference's ntc counter wraps (moldin.F90:791)
    if (rc == MDK_OK && c->ntc == 0) rc = relink_in_loop(c);
  }
  const int rf = flush_kick(c);  // the last step's sums (also after a failed step)
  if (rc != MDK_OK) return rc;
  RCK(rf);
  if (trace && nsteps > 0)
    HIPCK(hipMemcpyAsync(trace, c->trace_dev, sizeof(double) * 6 * (size_t)nsteps, hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  // the samples of the last three steps, not yet polled
  for (long st = std::max(c->step_counter - 3, c->step_counter - nsteps); st < c->step_counter; st++) {
    const double *m = c->mon_host + MON_STRIDE * (int)(st & 3);
    if (skin_used(c, m) >= 0.5 * c->skin) return skin_violation(c, m, st);
  }
  return check_device_status(c);
}

extern "C" int mdk_stats(void *h, long *rebuilds, long *full_pairs) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  if (rebuilds) *rebuilds = c->rebuilds;
  if (full_pairs) {
    std::vector<int> nn(c->n);
    HIPCK(hipMemcpy(nn.data(), c->numn, sizeof(int) * c->n, hipMemcpyDeviceToHost));
    long s = 0;
    for (int v : nn) s += v;
    *full_pairs = s;
  }
  return MDK_OK;
}

extern "C" int mdk_timing(void *h, int mask) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  c->timing_mask = mask & ((1 << TC_NCLASS) - 1);
  for (int k = 0; k < TC_NCLASS; k++) c->t_seen[k] = 0;
  if (c->timing_mask) {
    for (int k = 0; k < TC_NCLASS; k++) { c->t_ms[k] = 0.0; c->t_cnt[k] = 0; }
    hipStreamSynchronize(c->stream);
    for (auto &e : c->evs) { c->evpool.push_back(e.a); c->evpool.push_back(e.b); }
    c->evs.clear();
  }
  return MDK_OK;
}

extern "C" int mdk_timing_every(void *h, int every) {
  mdk_ctx *c = H(h);
  if (!c || every < 1) return set_err(MDK_ERR_BAD_ARG, "bad timing stride");
  c->timing_every = every;
  return MDK_OK;
}

extern "C" int mdk_timing_read(void *h, double *ms, long *launches, int nclasses) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  HIPCK(hipStreamSynchronize(c->stream));
  for (auto &e : c->evs) {
    float t = 0.f;
    hipEventElapsedTime(&t, e.a, e.b);
    c->t_ms[e.cls] += t;
    c->t_cnt[e.cls] += 1;
    c->evpool.push_back(e.a);
    c->evpool.push_back(e.b);
  }
  c->evs.clear();
  for (int k = 0; k < nclasses && k < TC_NCLASS; k++) {
    if (ms) ms[k] = c->t_ms[k];
    if (launches) launches[k] = c->t_cnt[k];
  }
  return MDK_OK;
}

extern "C" int mdk_synchronize(void *h) {
  mdk_ctx *c = H(h);
  if (!c) return set_err(MDK_ERR_BAD_ARG, "null handle");
  HIPCK(hipStreamSynchronize(c->stream));
  return MDK_OK;
}
