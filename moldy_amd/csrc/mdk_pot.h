// mdk_pot.h -- Fe-Fe Finnis-Sinclair potentials on the device (gfx950).
//
// Two material modules of the reference (chosen at compile time there,
// src/potential_m.F90:49-87; a run-time choice here): fetransmetal and
// iron_carbon (Fe-Fe functions only; see the iron_carbon section below).
//
// Functions of src/fetransmetal.F90 for the Fe-Fe pair (find_coeff factors are
// exactly 1 for Fe-Fe, :831,937-940):
//   r > 1.9 A         V  = sum_k a_k (r_k - r)^3 H(r_k - r)          (:133-139)
//   0.9 < r <= 1.9 A  V  = exp(cubic Hermite of log V)  joining       (:511-551)
//                          Biersack (0.9-1e-12) to the knot form (1.9+1e-12)
//   r < 0.9 A         V  = Biersack-Ziegler universal screening        (:414-454)
//   phi = sum_k A_k (R_k - r)^3 H(R_k - r) above its maximum r*, flat below (:256-265)
//   F(rho) = -sqrt(max(rho,0)), F'(rho) = -1/(2 sqrt(rho)) for rho > 0 (:375-404)
// and the uniform-grid linear-interpolation tables of src/lookup_m.F90:307-394.
//
// Everything is templated on the arithmetic type: double for the parity form,
// float for the reduced-precision form (the caller accumulates in double).
#pragma once
#include <hip/hip_runtime.h>

#define MDK_MAXV 8
#define MDK_MAXP 4
#define MDK_MAT_FETRANSMETAL 0
#define MDK_MAT_IRON_CARBON 1
#define MDK_FEC_NV 13
#define MDK_FEC_NP 3

struct PotDev {
  int nv, np, use_lookup, nlookup;
  double av[MDK_MAXV], rv[MDK_MAXV], ap[MDK_MAXP], rp[MDK_MAXP];
  double bc1, bc2;          // Biersack and knot-form join radii (REAL(4) 0.9 and 1.9)
  double sc[4];             // spline_coeff(1:4): V, V' at bc1-eps and at bc2+eps
  double phi_cut_r, phi_cut_v;
  double bier_a, bier_zed;  // screening length (a REAL(4) value) and Z1 Z2 e^2/(4 pi eps0)
  double r_int;             // outermost knot: V, V', phi, phi' vanish beyond it
  // lookup tables (src/lookup_m.F90): [0,6] A for V,V',phi,phi'; rho in [0,1e4]
  double r_min, r_max, vp_range, vp_reciprange, vp_recipnlkp, vp_recipstep;
  double rho_min, rho_max, e_range, e_reciprange, e_recipnlkp, e_recipstep;
  const double *tab_v, *tab_dv, *tab_p, *tab_dp, *tab_e, *tab_de;
  // float copies of the knot data for the reduced-precision form
  float avf[MDK_MAXV], rvf[MDK_MAXV], apf[MDK_MAXP], rpf[MDK_MAXP];
  // rp - rpf: the density knots' FP32 rounding (up to 3e-8 A) biases every
  // pair's phi the same way (+8e-7 eV/atom in bcc Fe), so the FP32 form
  // carries each density radius as rpf + rplf
  float rplf[MDK_MAXP];
  float bc1f, bc2f, phi_cut_rf, phi_cut_vf, r_intf;
  // pre-derived join constants: log(V) and V'/V at both ends
  double jx1, jv1, jx2, jv2;
  float jx1f, jv1f, jx2f, jv2f, bier_af, bier_zedf;
  // FP32 Fe-Fe pair terms (the force pass of the FP32 form): bin of r in the
  // interval table is (int)(r * fe_binh + fe_binc), fe_binh = 1/h
  float fe_binh, fe_binc;
  // material (MDK_MAT_*); iron_carbon Fe-Fe data (src/iron_carbon.F90:44-74):
  // pair knots, amplitudes and 3x amplitudes of the 13-knot sum, the cubic in
  // the join's exponent, density knots with the formula's signs (+, -, +),
  // embedding cludge amplitude/threshold, rho cap and F at the cap
  int material;
  double fv_r[MDK_FEC_NV], fv_a[MDK_FEC_NV], fv_a3[MDK_FEC_NV], fjoin[4];
  double fp_r[MDK_FEC_NP], fp_a[MDK_FEC_NP], fp_a3[MDK_FEC_NP];
  double f_acl, f_rcl, f_rx, f_cap;
};

// Knot-parameter sensitivities of the set_up constants (filled on the host):
// join end values x2 = log V(1.9+eps), v2 = V'/V there, per pair parameter;
// flat-cap value phi(r*) per density parameter.
struct GradDev {
  double xa[MDK_MAXV], xr[MDK_MAXV], ya[MDK_MAXV], yr[MDK_MAXV];
  double cpa[MDK_MAXP], cpr[MDK_MAXP];
};

template <typename T> __device__ __forceinline__ T mdk_exp(T x);
template <> __device__ __forceinline__ double mdk_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float mdk_exp<float>(float x) { return __expf(x); }

template <typename T> struct PotView;
template <> struct PotView<double> {
  const PotDev &P;
  __device__ PotView(const PotDev &p) : P(p) {}
  __device__ double av(int k) const { return P.av[k]; }
  __device__ double rv(int k) const { return P.rv[k]; }
  __device__ double ap(int k) const { return P.ap[k]; }
  __device__ double rp(int k) const { return P.rp[k]; }
  __device__ double dp(int k, double r) const { return P.rp[k] - r; }
  __device__ double bc1() const { return P.bc1; }
  __device__ double bc2() const { return P.bc2; }
  __device__ double pcr() const { return P.phi_cut_r; }
  __device__ double pcv() const { return P.phi_cut_v; }
  __device__ double jx1() const { return P.jx1; }
  __device__ double jv1() const { return P.jv1; }
  __device__ double jx2() const { return P.jx2; }
  __device__ double jv2() const { return P.jv2; }
  __device__ double ba() const { return P.bier_a; }
  __device__ double bz() const { return P.bier_zed; }
};
template <> struct PotView<float> {
  const PotDev &P;
  __device__ PotView(const PotDev &p) : P(p) {}
  __device__ float av(int k) const { return P.avf[k]; }
  __device__ float rv(int k) const { return P.rvf[k]; }
  __device__ float ap(int k) const { return P.apf[k]; }
  __device__ float rp(int k) const { return P.rpf[k]; }
  // R_k - r with the radius's low part added after the (exact near R_k) difference
  __device__ float dp(int k, float r) const { return (P.rpf[k] - r) + P.rplf[k]; }
  __device__ float bc1() const { return P.bc1f; }
  __device__ float bc2() const { return P.bc2f; }
  __device__ float pcr() const { return P.phi_cut_rf; }
  __device__ float pcv() const { return P.phi_cut_vf; }
  __device__ float jx1() const { return P.jx1f; }
  __device__ float jv1() const { return P.jv1f; }
  __device__ float jx2() const { return P.jx2f; }
  __device__ float jv2() const { return P.jv2f; }
  __device__ float ba() const { return P.bier_af; }
  __device__ float bz() const { return P.bier_zedf; }
};

// Biersack-Ziegler V and V' (src/fetransmetal.F90:414-454).  The rare short-range
// paths take scalars by value so the potential struct never has its address
// taken (which would force a per-thread scratch copy of it).
template <typename T> struct VdV {
  T v, dv;
};

template <typename T>
__device__ __noinline__ VdV<T> mdk_biersack(T a, T z, T r) {
  const T x = r / a, ir = T(1) / r;
  const T e1 = mdk_exp<T>(-T(3.2) * x), e2 = mdk_exp<T>(-T(0.9423) * x);
  const T e3 = mdk_exp<T>(-T(0.4029) * x), e4 = mdk_exp<T>(-T(0.2016) * x);
  // The reference evaluates the value with REAL(4)-rounded literals.
  const T ph = T((double)0.1818f) * mdk_exp<T>(-T((double)3.2f) * x) +
               T((double)0.5099f) * mdk_exp<T>(-T((double)0.9423f) * x) +
               T((double)0.2802f) * mdk_exp<T>(-T((double)0.4029f) * x) +
               T((double)0.02817f) * mdk_exp<T>(-T((double)0.2016f) * x);
  const T s = (ir + T(3.2) / a) * T(0.1818) * e1 + (ir + T(0.9423) / a) * T(0.5099) * e2 +
              (ir + T(0.4029) / a) * T(0.2802) * e3 + (ir + T(0.2016) / a) * T(0.02817) * e4;
  return VdV<T>{z * ph / r, -(z * ir) * s};
}

// log-space cubic Hermite join between a1 and a2 (src/fetransmetal.F90:511-551);
// x1 = log V(a1), v1 = V'/V(a1), x2 = log V(a2), v2 = V'/V(a2)
template <typename T>
__device__ __noinline__ VdV<T> mdk_join(T a1, T a2, T x1, T v1, T x2, T v2, T r) {
  const T h = a2 - a1;
  const T t = (r - a1) / h;
  const T t2 = t * t, t3 = t2 * t;
  const T sp = (T(2) * t3 - T(3) * t2 + T(1)) * x1 + (t3 - T(2) * t2 + t) * v1 * h +
               (-(T(2) * t3) + T(3) * t2) * x2 + (t3 - t2) * v2 * h;
  T ds = (T(6) * t2 - T(6) * t) * x1 + (T(3) * t2 - T(4) * t + T(1)) * v1 * h +
         (-(T(6) * t2) + T(6) * t) * x2 + (T(3) * t2 - T(2) * t) * v2 * h;
  ds = ds / h;
  const T e = mdk_exp<T>(sp);
  return VdV<T>{e, ds * e};
}

// Knot counts.  The pair kernels are compiled for the Fe-Fe potential's 6 pair
// and 2 density knots (POT_FE) and for the padded capacity (POT_KNOTS, any
// count up to MDK_MAXV/MDK_MAXP: unused knots are zero, so they add nothing
// for r > 0); POT_TABLE selects the lookup tables.
enum { POT_TABLE = 0, POT_KNOTS = 1, POT_FE = 2, POT_FEC = 3 };
#define MDK_FE_NV 6
#define MDK_FE_NP 2
template <int POTM> struct KnotCount {
  static constexpr int nv = POTM == POT_FE ? MDK_FE_NV : MDK_MAXV;
  static constexpr int np = POTM == POT_FE ? MDK_FE_NP : MDK_MAXP;
};

// (r_k - r) H(r_k - r) as max(r_k - r, 0)
template <typename T> __device__ __forceinline__ T mdk_pos(T d) { return d > T(0) ? d : T(0); }

// s + a*b with the product rounded first in FP64 (the reference's density knot
// sums are not contracted: at the density maximum the sum cancels exactly), a
// fused multiply-add in FP32.
__device__ __forceinline__ double knot_acc(double s, double a, double b) {
#pragma clang fp contract(off)
  return s + a * b;
}
__device__ __forceinline__ float knot_acc(float s, float a, float b) { return s + a * b; }

// Pair function V and V' (exact mode) over NV knots: the knot form is
// evaluated branch-free (6 VALU operations per knot, in the operation order
// of the reference's knot sum); the short-range forms
// (r <= 1.9 A, rare) are a divergent call.
template <typename T, int NV>
__device__ __forceinline__ void mdk_vee_n(const PotDev &P, T r, T &v, T &dv) {
  PotView<T> Q(P);
  T s = T(0), d1 = T(0);
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const T d = mdk_pos<T>(Q.rv(k) - r);
    const T ad2 = Q.av(k) * (d * d);
    s = s + ad2 * d;
    d1 = d1 - T(3) * ad2;
  }
  v = s;
  dv = d1;
  if (!(r > Q.bc2())) {
    if (r > Q.bc1()) {
      VdV<T> q = mdk_join<T>(Q.bc1(), Q.bc2(), Q.jx1(), Q.jv1(), Q.jx2(), Q.jv2(), r);
      v = q.v;
      dv = q.dv;
    } else if (r < Q.bc1()) {
      VdV<T> q = mdk_biersack<T>(Q.ba(), Q.bz(), r);
      v = q.v;
      dv = q.dv;
    } else {
      v = T(0);
      dv = T(0);
    }
  }
}

// Density function phi (exact mode), flat below its maximum r*.
template <typename T, int NP>
__device__ __forceinline__ T mdk_phi_n(const PotDev &P, T r) {
  PotView<T> Q(P);
  T s = T(0);
#pragma unroll
  for (int k = 0; k < NP; k++) {
    const T d = mdk_pos<T>(Q.dp(k, r));
    s = knot_acc(s, Q.ap(k), d * d * d);
  }
  return r < Q.pcr() ? Q.pcv() : s;
}

// Density derivative phi' (exact mode), zero below r*.
template <typename T, int NP>
__device__ __forceinline__ T mdk_dphi_n(const PotDev &P, T r) {
  PotView<T> Q(P);
  T s = T(0);
#pragma unroll
  for (int k = 0; k < NP; k++) {
    const T d = mdk_pos<T>(Q.dp(k, r));
    s = knot_acc(s, T(-3) * Q.ap(k), d * d);
  }
  return r < Q.pcr() ? T(0) : s;
}

// FP32 pair terms V, V' and phi' of a 6+2-knot set (POT_FE) for the force
// pass.  The breakpoints -- the 6 pair knots, the 2 density knots and the
// density maximum r* (phi' = 0 below it) -- split r into MDK_FE_NI intervals;
// interval m holds, in the local variable x = u - r (u its midpoint),
//   V = c0 + x (c1 + x (c2 + x c3)) = sum a_k (r_k - r)^3 over the knots beyond r,
//   phi' = q0 + x (q1 + x q2) = -3 sum A_k (R_k - r)^2,
// coefficients formed in FP64 on the host; V' = -dV/dx comes from the same
// Horner steps (p2 = c2 + x c3, p1 = c1 + x p2, dV/dx = p1 + x (p2 + x c3)).  The interval comes from a uniform
// grid of MDK_FE_NB bins of width h (a power of 2 below the smallest
// breakpoint spacing, so a bin holds at most one breakpoint): a bin record is
// (split, m below it, m at or above it), m as the float offset of the
// interval's 8-float record (u, c0, c1, c2 | c3, q0, q1, q2).  One bin read and one compare replace a compare
// per breakpoint; phi' costs two fused multiply-adds instead of a knot sum.
// The short-range forms (r <= 1.9 A, rare) are a divergent call, as in the
// knot form.
#define MDK_FE_NB 64
#define MDK_FE_NI 10
#define MDK_FE_TAB (4 * MDK_FE_NB + 8 * MDK_FE_NI)
__device__ __forceinline__ void mdk_fe_terms(const PotDev &P, const float4 *bins, const float *iv, float r, float &v,
                                             float &dv, float &dph) {
  const int b = min(max((int)fmaf(r, P.fe_binh, P.fe_binc), 0), MDK_FE_NB - 1);
  const float4 rec = bins[b];
  const float *q = iv + __float_as_int(r >= rec.x ? rec.z : rec.y);
  const float4 a = *(const float4 *)q, c = *(const float4 *)(q + 4);
  const float x = a.x - r;
  const float p2 = a.w + x * c.x, p1 = a.z + x * p2;
  v = a.y + x * p1;
  dv = -(p1 + x * (p2 + x * c.x));
  dph = c.y + x * (c.z + x * c.w);
  if (!(r > P.bc2f)) {
    if (r > P.bc1f) {
      VdV<float> s = mdk_join<float>(P.bc1f, P.bc2f, P.jx1f, P.jv1f, P.jx2f, P.jv2f, r);
      v = s.v;
      dv = s.dv;
    } else if (r < P.bc1f) {
      VdV<float> s = mdk_biersack<float>(P.bier_af, P.bier_zedf, r);
      v = s.v;
      dv = s.dv;
    } else {
      v = 0.f;
      dv = 0.f;
    }
  }
}

// ---------------------------------------------------------------------------
// iron_carbon Fe-Fe (POT_FEC, FP64), src/iron_carbon.F90:
//   r > 2.05 A (REAL(4) literal)  V = sum_13 a_k (r_k - r)^3 H(r_k - r)     (:128-153)
//   1 <= r <= 2.05 A              V = exp(j0 + j1 r + j2 r^2 + j3 r^3)     (:154-157)
//   r < 1 A                       V = Biersack-Ziegler, Z = 26 (exponents folded) (:119-126)
//   phi = A1 (R1-r)^3 H - A2 (R2-r)^3 H + A3 (R3-r)^3 H (no flat cap)      (:439-449)
//   F(rho) = -rho^1/2 - c2 rho^2 + c4 rho^4 (+ cludge above rcl), constant
//   above the cap rx; F' = 0 for rho <= 0 or rho > rx                      (:599-668)
// Sums follow the reference's left-to-right order.
#define MDK_FEC_C2 6.7314115586063e-4
#define MDK_FEC_C4 7.6514905604792e-8

// short range (r <= 2.05 A): scalars only, so the struct is never spilled
template <typename T>
__device__ __noinline__ VdV<T> mdk_fec_short(T j0, T j1, T j2, T j3, T r) {
  if (r < T(1)) {
    const T x1 = T(-28.616172142695042), x2 = T(-8.42656844064423), x3 = T(-3.6029549238411973),
            x4 = T(-1.8028188449897875);
    const T e1 = T(0.1818) * mdk_exp<T>(x1 * r), e2 = T(0.5099) * mdk_exp<T>(x2 * r),
            e3 = T(0.2802) * mdk_exp<T>(x3 * r), e4 = T(0.02817) * mdk_exp<T>(x4 * r);
    const T zr = T(9734.159682000001) / r, ir = T(1) / r;
    const T v = zr * (e1 + e2 + e3 + e4);
    const T dv = -zr * ((ir - x1) * e1 + (ir - x2) * e2 + (ir - x3) * e3 + (ir - x4) * e4);
    return VdV<T>{v, dv};
  }
  const T e = mdk_exp<T>(j0 + j1 * r + j2 * r * r + j3 * r * r * r);
  return VdV<T>{e, (j1 + T(2) * j2 * r + T(3) * j3 * r * r) * e};
}

template <typename T>
__device__ __forceinline__ void mdk_fec_vee(const PotDev &P, T r, T &v, T &dv) {
  T s = T(0), d1 = T(0);
#pragma unroll
  for (int k = 0; k < MDK_FEC_NV; k++) {
    const T d = mdk_pos<T>((T)P.fv_r[k] - r);
    const T d2 = d * d;
    s = s + (T)P.fv_a[k] * (d2 * d);
    d1 = d1 + (T)P.fv_a3[k] * d2;
  }
  v = s;
  dv = -d1;
  if (!(r > T((double)2.05f))) {
    VdV<T> q = mdk_fec_short<T>((T)P.fjoin[0], (T)P.fjoin[1], (T)P.fjoin[2], (T)P.fjoin[3], r);
    v = q.v;
    dv = q.dv;
  }
}
template <typename T>
__device__ __forceinline__ T mdk_fec_phi(const PotDev &P, T r) {
  T t[MDK_FEC_NP];
#pragma unroll
  for (int k = 0; k < MDK_FEC_NP; k++) {
    const T d = mdk_pos<T>((T)P.fp_r[k] - r);
    t[k] = (T)P.fp_a[k] * (d * d * d);
  }
  return t[0] - t[1] + t[2];
}
template <typename T>
__device__ __forceinline__ T mdk_fec_dphi(const PotDev &P, T r) {
  T t[MDK_FEC_NP];
#pragma unroll
  for (int k = 0; k < MDK_FEC_NP; k++) {
    const T d = mdk_pos<T>((T)P.fp_r[k] - r);
    t[k] = (T)P.fp_a3[k] * (d * d);
  }
  return -(t[0] - t[1] + t[2]);
}
__device__ __forceinline__ void mdk_fec_embed(const PotDev &P, double rho, double &f, double &df) {
  f = -1.0 * fabs(sqrt(rho)) - MDK_FEC_C2 * (rho * rho) + MDK_FEC_C4 * (rho * rho * rho * rho);
  df = 0.0;
  if (rho > 0.0 && !(rho > P.f_rx))
    df = 0.5 * -1.0 / sqrt(rho) - 2.0 * MDK_FEC_C2 * rho + 4.0 * MDK_FEC_C4 * (rho * rho * rho);
  if (rho > P.f_rcl) {
    const double d = rho - P.f_rcl;
    f = f + P.f_acl * (d * d * d);
    if (!(rho > P.f_rx)) df = df + 3.0 * P.f_acl * (d * d);
  }
  if (rho > P.f_rx) f = P.f_cap;
}

// Padded-capacity forms (any knot count up to MDK_MAXV / MDK_MAXP).
template <typename T>
__device__ __forceinline__ void mdk_vee(const PotDev &P, T r, T &v, T &dv) {
  mdk_vee_n<T, MDK_MAXV>(P, r, v, dv);
}
template <typename T>
__device__ __forceinline__ T mdk_phi(const PotDev &P, T r) { return mdk_phi_n<T, MDK_MAXP>(P, r); }
template <typename T>
__device__ __forceinline__ T mdk_dphi(const PotDev &P, T r) { return mdk_dphi_n<T, MDK_MAXP>(P, r); }

// Linear-interpolation lookups in the reference's index arithmetic
// (src/lookup_m.F90:307-394); no FMA so the result is the reference's.
__device__ __forceinline__ double mdk_lk_vp(const PotDev &P, const double *t, double r) {
#pragma clang fp contract(off)
  if (r >= P.r_max) return t[P.nlookup];
  int lo = (int)((double)P.nlookup * (r - P.r_min) * P.vp_reciprange);
  lo = min(max(lo, 0), P.nlookup - 1);  // r < r_min (or NaN) would index outside the table
  const double rl = P.r_min + P.vp_range * (double)lo * P.vp_recipnlkp;
  const double dr = (r - rl) * P.vp_recipstep;
  return t[lo] * (1.0 - dr) + t[lo + 1] * dr;
}
__device__ __forceinline__ double mdk_lk_e(const PotDev &P, const double *t, double rho) {
#pragma clang fp contract(off)
  if (rho >= P.rho_max) return t[P.nlookup];
  int lo = (int)((double)P.nlookup * (rho - P.rho_min) * P.e_reciprange);
  lo = min(max(lo, 0), P.nlookup - 1);
  const double rl = P.rho_min + P.e_range * (double)lo * P.e_recipnlkp;
  const double dr = (rho - rl) * P.e_recipstep;
  return t[lo] * (1.0 - dr) + t[lo + 1] * dr;
}

// Dispatchers (potential_m.F90:288-417): tables or knots, a compile-time
// choice in the pair kernels.
template <typename T, int POTM = POT_KNOTS>
__device__ __forceinline__ void mdk_pair_vdv(const PotDev &P, T r, T &v, T &dv) {
  if (POTM == POT_TABLE) {
    v = (T)mdk_lk_vp(P, P.tab_v, (double)r);
    dv = (T)mdk_lk_vp(P, P.tab_dv, (double)r);
  } else if (POTM == POT_FEC) {
    mdk_fec_vee<T>(P, r, v, dv);
  } else {
    mdk_vee_n<T, KnotCount<POTM>::nv>(P, r, v, dv);
  }
}
template <typename T, int POTM = POT_KNOTS>
__device__ __forceinline__ T mdk_pair_phi(const PotDev &P, T r) {
  if (POTM == POT_TABLE) return (T)mdk_lk_vp(P, P.tab_p, (double)r);
  if (POTM == POT_FEC) return mdk_fec_phi<T>(P, r);
  return mdk_phi_n<T, KnotCount<POTM>::np>(P, r);
}
template <typename T, int POTM = POT_KNOTS>
__device__ __forceinline__ T mdk_pair_dphi(const PotDev &P, T r) {
  if (POTM == POT_TABLE) return (T)mdk_lk_vp(P, P.tab_dp, (double)r);
  if (POTM == POT_FEC) return mdk_fec_dphi<T>(P, r);
  return mdk_dphi_n<T, KnotCount<POTM>::np>(P, r);
}
// Embedding F and F' in FP64 (per atom, cheap).
__device__ __forceinline__ void mdk_embed_f(const PotDev &P, double rho, double &f, double &df) {
  if (P.use_lookup) {
    f = mdk_lk_e(P, P.tab_e, rho);
    df = mdk_lk_e(P, P.tab_de, rho);
  } else if (P.material == MDK_MAT_IRON_CARBON) {
    mdk_fec_embed(P, rho, f, df);
  } else {
    f = -1.0 * sqrt(rho > 0.0 ? rho : 0.0);
    df = 0.0;
    if (rho > 0.0) df = -0.5 / sqrt(rho);
  }
}
