// mdk_wave.h -- wave64 reductions built from DPP row shifts and broadcasts.
//
// A shuffle butterfly (__shfl_xor) goes through ds_bpermute, an LDS-pipe
// instruction with a round trip of ~100 cycles per step; the kernels' block
// reductions chain six of them per value for up to twenty values.  DPP moves
// are VALU operands with a few cycles of latency.  The order of the additions
// is fixed (a prefix over each 16-lane row, then rows 0+1 and 2+3, then all
// four), so the result is deterministic run to run.  Every lane of the wave
// must be active when these are called.
#pragma once
#include <hip/hip_runtime.h>

// GFX9 DPP controls: row_shr:n = 0x110 + n, row_bcast:15 = 0x142 (lane 15 of
// each row to the whole next row), row_bcast:31 = 0x143 (lane 31 to rows 2, 3)
#define MDK_DPP_ROW_SHR(n) (0x110 + (n))
#define MDK_DPP_ROW_BCAST15 0x142
#define MDK_DPP_ROW_BCAST31 0x143

// The operand of lane i as DPP control CTRL routes it, in lanes of rows enabled
// by ROW_MASK.  Sums: a missing source or a disabled row reads 0.0.  Maxima: a
// lane without a source reads its own value (fmax(v, v) = v).
template <int CTRL, int ROW_MASK, bool MAXOP>
__device__ __forceinline__ double dpp_route(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)b, hi = (int)(b >> 32);
  const int rlo = __builtin_amdgcn_update_dpp(MAXOP ? lo : 0, lo, CTRL, ROW_MASK, 0xf, !MAXOP);
  const int rhi = __builtin_amdgcn_update_dpp(MAXOP ? hi : 0, hi, CTRL, ROW_MASK, 0xf, !MAXOP);
  return __longlong_as_double(((long long)rhi << 32) | (unsigned int)rlo);
}

template <bool MAXOP>
__device__ __forceinline__ double dpp_op(double a, double b) {
  return MAXOP ? fmax(a, b) : a + b;
}

// Reduction over the 64 lanes; every lane receives the result (lane 63's).
template <bool MAXOP>
__device__ __forceinline__ double wave_reduce(double v) {
  v = dpp_op<MAXOP>(v, dpp_route<MDK_DPP_ROW_SHR(1), 0xf, MAXOP>(v));
  v = dpp_op<MAXOP>(v, dpp_route<MDK_DPP_ROW_SHR(2), 0xf, MAXOP>(v));
  v = dpp_op<MAXOP>(v, dpp_route<MDK_DPP_ROW_SHR(4), 0xf, MAXOP>(v));
  v = dpp_op<MAXOP>(v, dpp_route<MDK_DPP_ROW_SHR(8), 0xf, MAXOP>(v));  // lane 15 of a row: the row's total
  v = dpp_op<MAXOP>(v, dpp_route<MDK_DPP_ROW_BCAST15, 0xa, MAXOP>(v));  // lanes 31, 63: rows 0+1, 2+3
  v = dpp_op<MAXOP>(v, dpp_route<MDK_DPP_ROW_BCAST31, 0xc, MAXOP>(v));  // lane 63: all rows
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double wave_sum(double v) { return wave_reduce<false>(v); }
__device__ __forceinline__ double wave_max(double v) { return wave_reduce<true>(v); }

// ---------------------------------------------------------------------------
// Reduction of K values at once (K <= 20; values k >= K - KM are maxima, the
// others sums), as a reduce-scatter: the exchange across the wave halves
// (v_permlane32_swap) combines value pairs so that each half keeps one value
// of every pair, the exchange across row pairs (v_permlane16_swap) halves the
// set again, and each 16-lane row then sums (or maxes) the values it kept with
// four DPP steps.  About 7K VALU operations instead of 18K for K separate wave
// reductions.  The order is fixed, so results are deterministic.  out[] is
// uniform (every lane receives every total).  Every lane must be active.

namespace wave_multi {
__device__ __forceinline__ double join2(unsigned lo, unsigned hi) {
  return __longlong_as_double(((long long)hi << 32) | lo);
}
// (a, b) -> (a lanes 0-31 | b lanes 0-31, a lanes 32-63 | b lanes 32-63)
__device__ __forceinline__ void swap32(double &a, double &b) {
  const long long ai = __double_as_longlong(a), bi = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ai, (unsigned)bi, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ai >> 32), (unsigned)(bi >> 32), false, false);
  a = join2(lo[0], hi[0]);
  b = join2(lo[1], hi[1]);
}
// (a, b) -> (rows a0 b0 a2 b2, rows a1 b1 a3 b3)
__device__ __forceinline__ void swap16(double &a, double &b) {
  const long long ai = __double_as_longlong(a), bi = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ai, (unsigned)bi, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ai >> 32), (unsigned)(bi >> 32), false, false);
  a = join2(lo[0], hi[0]);
  b = join2(lo[1], hi[1]);
}
template <bool MAXOP>
__device__ __forceinline__ double comb(double a, double b) {
  return MAXOP ? fmax(a, b) : a + b;
}
// combine with the operation of value index ka in lanes where sel is false and
// of kb where it is true
template <int K, int KM>
__device__ __forceinline__ double comb_sel(double a, double b, int ka, int kb, bool sel) {
  const bool ma = ka >= K - KM, mb = kb >= K - KM;
  if (ma == mb) return ma ? fmax(a, b) : a + b;
  const double s = a + b, m = fmax(a, b);
  return (sel ? mb : ma) ? m : s;
}
// row reduction: lane 15 of every row receives the row's total
template <int K, int KM>
__device__ __forceinline__ double row_reduce(double v, const int (&kr)[4], int row) {
  const bool anymax = (kr[0] >= K - KM) || (kr[1] >= K - KM) || (kr[2] >= K - KM) || (kr[3] >= K - KM);
  const bool allmax = (kr[0] >= K - KM) && (kr[1] >= K - KM) && (kr[2] >= K - KM) && (kr[3] >= K - KM);
  if (!anymax) {
    v = v + dpp_route<MDK_DPP_ROW_SHR(1), 0xf, false>(v);
    v = v + dpp_route<MDK_DPP_ROW_SHR(2), 0xf, false>(v);
    v = v + dpp_route<MDK_DPP_ROW_SHR(4), 0xf, false>(v);
    v = v + dpp_route<MDK_DPP_ROW_SHR(8), 0xf, false>(v);
    return v;
  }
  if (allmax) {
    v = fmax(v, dpp_route<MDK_DPP_ROW_SHR(1), 0xf, true>(v));
    v = fmax(v, dpp_route<MDK_DPP_ROW_SHR(2), 0xf, true>(v));
    v = fmax(v, dpp_route<MDK_DPP_ROW_SHR(4), 0xf, true>(v));
    v = fmax(v, dpp_route<MDK_DPP_ROW_SHR(8), 0xf, true>(v));
    return v;
  }
  // mixed rows: maxima in rows whose value is a maximum.  A row shift reads
  // only its own row; max rows take the lane's own value for a missing source
  // (idempotent), sum rows zero.
  const bool m0 = kr[0] >= K - KM, m1 = kr[1] >= K - KM, m2 = kr[2] >= K - KM, m3 = kr[3] >= K - KM;
  const bool mx = row == 0 ? m0 : row == 1 ? m1 : row == 2 ? m2 : m3;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    double ts, tm;
    if (s == 0) { ts = dpp_route<MDK_DPP_ROW_SHR(1), 0xf, false>(v); tm = dpp_route<MDK_DPP_ROW_SHR(1), 0xf, true>(v); }
    if (s == 1) { ts = dpp_route<MDK_DPP_ROW_SHR(2), 0xf, false>(v); tm = dpp_route<MDK_DPP_ROW_SHR(2), 0xf, true>(v); }
    if (s == 2) { ts = dpp_route<MDK_DPP_ROW_SHR(4), 0xf, false>(v); tm = dpp_route<MDK_DPP_ROW_SHR(4), 0xf, true>(v); }
    if (s == 3) { ts = dpp_route<MDK_DPP_ROW_SHR(8), 0xf, false>(v); tm = dpp_route<MDK_DPP_ROW_SHR(8), 0xf, true>(v); }
    v = mx ? fmax(v, tm) : v + ts;
  }
  return v;
}
__device__ __forceinline__ double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  return join2((unsigned)__builtin_amdgcn_readlane((int)b, l), (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l));
}
}  // namespace wave_multi

template <int K, int KM>
__device__ __forceinline__ void wave_reduce_multi(const double (&v)[K], double (&out)[K]) {
  using namespace wave_multi;
  static_assert(K >= 1 && K <= 20, "wave_reduce_multi: 1..20 values");
  constexpr int NA = (K + 1) / 2;   // pairs of the half exchange
  constexpr int NB = (NA + 1) / 2;  // pairs of the row exchange
  const int lane = (int)__lane_id();
  const bool upper = lane >= 32, oddrow = (lane >> 4) & 1;
  // stage A: w[p] holds value 2p (lanes 0-31) / 2p+1 (lanes 32-63), halves combined
  double w[2 * NB];
#pragma unroll
  for (int p = 0; p < NA; p++) {
    double a = v[2 * p], b = 2 * p + 1 < K ? v[2 * p + 1] : 0.0;
    swap32(a, b);
    w[p] = comb_sel<K, KM>(a, b, 2 * p, 2 * p + 1 < K ? 2 * p + 1 : 2 * p, upper);
  }
  if (NA < 2 * NB) w[2 * NB - 1] = 0.0;
  // stage B: x[q] holds, by row r, value idx(r, q) = 4q + {0, 2, 1, 3}[r]
#pragma unroll
  for (int q = 0; q < NB; q++) {
    double a = w[2 * q], b = w[2 * q + 1];
    swap16(a, b);
    int kr[4] = {4 * q, 4 * q + 2, 4 * q + 1, 4 * q + 3};
#pragma unroll
    for (int r = 0; r < 4; r++) kr[r] = kr[r] < K ? kr[r] : 4 * q;  // padding slots follow slot 4q's op
    const int row = lane >> 4;
    // lanes keep w[2q] in rows 0, 2 and w[2q+1] in rows 1, 3
    const int ka = upper ? kr[2] : kr[0], kb = upper ? kr[3] : kr[1];
    double x = comb_sel<K, KM>(a, b, ka, kb, oddrow);
    x = row_reduce<K, KM>(x, kr, row);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int k = 4 * q + (r == 0 ? 0 : r == 1 ? 2 : r == 2 ? 1 : 3);
      if (k < K) out[k] = read_lane(x, 16 * r + 15);
    }
  }
}

// Every lane's v[k] replaced by the wave total (sums, or maxima for k >= K-KM).
template <int K, int KM>
__device__ __forceinline__ void wave_reduce_all(double (&v)[K]) {
  if constexpr (K >= 3) {
    double t[K];
    wave_reduce_multi<K, KM>(v, t);
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = t[k];
  } else {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = (k < K - KM) ? wave_sum(v[k]) : wave_max(v[k]);
  }
}
