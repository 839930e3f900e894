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
