// mdk_hostpot.h -- host-side potential set-up shared by the HIP translation units.
#pragma once
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mdk_pot.h"

// host-side potential set-up (src/fetransmetal.F90:603-954), evaluated with
// the host libm exactly in the reference's operation order so that the
// derived constants and lookup tables are the reference's bit for bit.
// ---------------------------------------------------------------------------
namespace hostpot {
struct Knots {
  int nv, np;
  double av[MDK_MAXV], rv[MDK_MAXV], ap[MDK_MAXP], rp[MDK_MAXP];
  double bc1, bc2, sc[4], pcr, pcv;
};
static double h_step(double x) { return x < 0.0 ? 0.0 : 1.0; }
static double knot_v(const Knots &k, double r) {
  double s = 0.0;
  for (int i = 0; i < k.nv; i++) {
    double d = k.rv[i] - r;
    s = s + k.av[i] * (d * d * d) * h_step(d);
  }
  return s;
}
static double knot_dv(const Knots &k, double r) {
  double s = 0.0;
  for (int i = 0; i < k.nv; i++) {
    double d = k.rv[i] - r;
    s = s - 3.0 * k.av[i] * (d * d) * h_step(d);
  }
  return s;
}
static double knot_p(const Knots &k, double r) {
  double s = 0.0;
  for (int i = 0; i < k.np; i++) {
    double d = k.rp[i] - r;
    s = s + k.ap[i] * (d * d * d) * h_step(d);
  }
  return s;
}
static double knot_dp(const Knots &k, double r) {
  double s = 0.0;
  for (int i = 0; i < k.np; i++) {
    double d = k.rp[i] - r;
    s = s - 3.0 * k.ap[i] * (d * d) * h_step(d);
  }
  return s;
}
// screening length: every literal of src/fetransmetal.F90:423 is REAL(4)
static double screening_length() {
  volatile float c = 0.8854f * 0.529f;
  float t = powf(26.0f, 0.666666667f);
  volatile float s = t + t;
  float den = powf(s, 0.5f);
  return (double)(c / den);
}
static double zed() { return (double)(26 * 26) * 8.98755178e9 * 1.602176487e-19 / 1.e-10; }
static double bier(double r) {
  double a = screening_length(), x = r / a;
  double ph = (double)0.1818f * exp(-((double)3.2f * x)) + (double)0.5099f * exp(-((double)0.9423f * x)) +
              (double)0.2802f * exp(-((double)0.4029f * x)) + (double)0.02817f * exp(-((double)0.2016f * x));
  return zed() * ph / r;
}
static double dbier(double r) {
  double a = screening_length(), x = r / a;
  double s = (1.0 / r + 3.2 / a) * 0.1818 * exp(-(3.2 * x)) + (1.0 / r + 0.9423 / a) * 0.5099 * exp(-(0.9423 * x)) +
             (1.0 / r + 0.4029 / a) * 0.2802 * exp(-(0.4029 * x)) + (1.0 / r + 0.2016 / a) * 0.02817 * exp(-(0.2016 * x));
  return -(zed() / r) * s;
}
static double join(const Knots &k, double r, int deriv) {
  double a1 = k.bc1, a2 = k.bc2, x1 = k.sc[0], v1 = k.sc[1], x2 = k.sc[2], v2 = k.sc[3];
  double vv1 = v1 / x1, xx1 = log(x1), vv2 = v2 / x2, xx2 = log(x2);
  double t = (r - a1) / (a2 - a1);
  double sp = (2.0 * t * t * t - 3.0 * t * t + 1.0) * xx1 + (t * t * t - 2.0 * t * t + t) * vv1 * (a2 - a1) +
              (-(2.0 * t * t * t) + 3.0 * t * t) * xx2 + (t * t * t - t * t) * vv2 * (a2 - a1);
  if (!deriv) return exp(sp);
  double ds = (6.0 * t * t - 6.0 * t) * xx1 + (3.0 * t * t - 4.0 * t + 1.0) * vv1 * (a2 - a1) +
              (-(6.0 * t * t) + 6.0 * t) * xx2 + (3.0 * t * t - 2.0 * t) * vv2 * (a2 - a1);
  ds = ds / (a2 - a1);
  return ds * exp(sp);
}
static double vee(const Knots &k, double r) {
  if (r > k.bc2) return knot_v(k, r);
  if (r > k.bc1) return join(k, r, 0);
  if (r < k.bc1) return bier(r);
  return 0.0;
}
static double dvee(const Knots &k, double r) {
  if (r > k.bc2) return knot_dv(k, r);
  if (r > k.bc1) return join(k, r, 1);
  if (r < k.bc1) return dbier(r);
  return 0.0;
}
static double phi(const Knots &k, double r) { return r < k.pcr ? k.pcv : knot_p(k, r); }
static double dphi(const Knots &k, double r) { return r < k.pcr ? 0.0 : knot_dp(k, r); }
static double emb(double rho) { return -1.0 * sqrt(rho > 0.0 ? rho : 0.0); }
static double demb(double rho) { return rho > 0.0 ? -0.5 / sqrt(rho) : 0.0; }
// set_up_m: join constants and the maximum of phi by 500-step bisection
static void set_up(Knots &k) {
  k.bc1 = (double)0.9f;
  k.bc2 = (double)1.9f;
  const double eps = (double)0.000000000001f;
  k.sc[0] = bier(k.bc1 - eps);
  k.sc[1] = dbier(k.bc1 - eps);
  k.sc[2] = vee(k, k.bc2 + eps);
  k.sc[3] = dvee(k, k.bc2 + eps);
  k.pcr = 0.0;
  k.pcv = 0.0;
  double lo = 0.0, hi = 3.0;
  for (int i = 0; i < 500; i++) {
    double mid = (hi + lo) / 2.0;
    if (dphi(k, mid) > 0.0) lo = mid; else hi = mid;
  }
  double rs = (hi + lo) / 2.0;
  k.pcr = rs;
  k.pcv = phi(k, rs);
}

// Fills the device potential struct from the set-up knots (exact mode);
// returns the interaction range (outermost knot).
static double fill_potdev(PotDev &P, const Knots &k) {
  memset(&P, 0, sizeof(P));
  P.nv = k.nv;
  P.np = k.np;
  double rint = 0.0;
  for (int i = 0; i < k.nv; i++) {
    P.av[i] = k.av[i]; P.rv[i] = k.rv[i]; P.avf[i] = (float)k.av[i]; P.rvf[i] = (float)k.rv[i];
    rint = std::max(rint, k.rv[i]);
  }
  for (int i = 0; i < k.np; i++) {
    P.ap[i] = k.ap[i]; P.rp[i] = k.rp[i]; P.apf[i] = (float)k.ap[i]; P.rpf[i] = (float)k.rp[i];
    P.rplf[i] = (float)(k.rp[i] - (double)P.rpf[i]);
    rint = std::max(rint, k.rp[i]);
  }
  P.bc1 = k.bc1; P.bc2 = k.bc2;
  for (int i = 0; i < 4; i++) P.sc[i] = k.sc[i];
  P.phi_cut_r = k.pcr; P.phi_cut_v = k.pcv;
  P.bier_a = screening_length();
  P.bier_zed = zed();
  P.r_int = rint;
  P.jx1 = log(k.sc[0]); P.jv1 = k.sc[1] / k.sc[0];
  P.jx2 = log(k.sc[2]); P.jv2 = k.sc[3] / k.sc[2];
  P.bc1f = (float)P.bc1; P.bc2f = (float)P.bc2; P.phi_cut_rf = (float)P.phi_cut_r;
  P.phi_cut_vf = (float)P.phi_cut_v; P.r_intf = (float)rint;
  P.jx1f = (float)P.jx1; P.jv1f = (float)P.jv1; P.jx2f = (float)P.jx2; P.jv2f = (float)P.jv2;
  P.bier_af = (float)P.bier_a; P.bier_zedf = (float)P.bier_zed;
  P.fe_binh = 0.0f;
  P.fe_binc = 0.0f;
  return rint;
}

// Interval table of the FP32 Fe-Fe pair terms (mdk_fe_terms): breakpoints are
// the pair knots, the density knots and the density maximum r*, in descending
// order B[0] > B[1] > ...; interval m is B[m] <= r < B[m-1] (m = 0: r >= B[0],
// where every term vanishes; the last: r < B[last]).  Record of interval m,
// with u its expansion point and d = knot - u over the knots at or above the
// interval's upper end:
//   u, c0, c1, c2, c3, q0, q1, q2
// c0..c3 = sum a d^3, 3 sum a d^2, 3 sum a d, sum a (pair knots); q0..q2 =
// -3 sum A d^2, -6 sum A d, -3 sum A (density knots, zero when u <= r*).  The
// bin grid: h = 2^-k (k = 4..8), the first with every breakpoint spacing
// above h, every breakpoint at least 1e-6 A from a bin boundary, and all
// breakpoints inside MDK_FE_NB bins from r_lo = floor(B_min / h) h.  Returns
// false when no such grid exists (then the FP32 form uses the knot sums).
// tab: MDK_FE_NB bin records (split, m below, m at or above, 0; m as the
// float offset 8 m, stored as int bits) then the MDK_FE_NI interval records.
inline bool build_fe_table(const Knots &k, float *tab, float &binh, float &binc) {
  if (k.nv != MDK_FE_NV || k.np != MDK_FE_NP) return false;
  std::vector<double> B;
  for (int i = 0; i < k.nv; i++) B.push_back(k.rv[i]);
  for (int i = 0; i < k.np; i++) B.push_back(k.rp[i]);
  if (k.pcr > 0.0) B.push_back(k.pcr);
  std::sort(B.begin(), B.end(), [](double a, double b) { return a > b; });
  const int nb = (int)B.size();
  if (nb + 1 > MDK_FE_NI) return false;
  for (int i = 1; i < nb; i++)
    if (!(B[i - 1] - B[i] > 0.0)) return false;
  memset(tab, 0, sizeof(float) * MDK_FE_TAB);
  float *iv = tab + 4 * MDK_FE_NB;
  for (int m = 0; m < MDK_FE_NI; m++) {
    // expansion point: the interval's midpoint (its upper end for the two
    // unbounded ones), which halves |x| and the rounding of the Horner steps;
    // rounded to FP32 first, so the device's x = u - r and these expansions
    // share one u (a rounded u would shift every pair of the interval alike)
    const int mm = std::min(m, nb);
    const double u = (double)(float)(mm == 0 ? B[0] : (mm < nb ? 0.5 * (B[mm - 1] + B[mm]) : B[mm - 1]));
    double c[4] = {0, 0, 0, 0}, q[3] = {0, 0, 0};
    if (m > 0) {
      const double top = B[mm - 1];
      for (int i = 0; i < k.nv; i++) {
        if (k.rv[i] < top) continue;
        const double a = k.av[i], d = k.rv[i] - u;
        c[0] += a * d * d * d; c[1] += 3.0 * a * d * d; c[2] += 3.0 * a * d; c[3] += a;
      }
      if (top > k.pcr) {
        for (int i = 0; i < k.np; i++) {
          if (k.rp[i] < top) continue;
          const double a = k.ap[i], d = k.rp[i] - u;
          q[0] += -3.0 * a * d * d; q[1] += -6.0 * a * d; q[2] += -3.0 * a;
        }
      }
    }
    const double row[8] = {u, c[0], c[1], c[2], c[3], q[0], q[1], q[2]};
    for (int t = 0; t < 8; t++) iv[8 * m + t] = (float)row[t];
  }
  for (int e = 4; e <= 8; e++) {
    const double h = ldexp(1.0, -e);
    bool ok = true;
    for (int i = 1; i < nb && ok; i++) ok = B[i - 1] - B[i] > h;
    const double lo = floor(B[nb - 1] / h) * h;
    ok = ok && (B[0] - lo) / h < MDK_FE_NB - 1;
    for (int i = 0; i < nb && ok; i++) {
      const double f = (B[i] - lo) / h;
      ok = std::min(f - floor(f), ceil(f) - f) * h > 1e-6;
    }
    if (!ok) continue;
    for (int b = 0; b < MDK_FE_NB; b++) {
      const double blo = lo + b * h, bhi = blo + h;
      int above = 0, split = -1;
      for (int i = 0; i < nb; i++) {
        if (B[i] >= bhi) above++;
        else if (B[i] >= blo) split = i;
      }
      float *r = tab + 4 * b;
      int mlo, mge;
      if (split >= 0) {
        r[0] = (float)B[split];
        mge = split;        // breakpoints strictly above the split
        mlo = split + 1;
      } else {
        r[0] = INFINITY;
        mge = mlo = above;
      }
      const int olo = 8 * mlo, oge = 8 * mge;
      memcpy(&r[1], &olo, sizeof(int));
      memcpy(&r[2], &oge, sizeof(int));
      r[3] = 0.0f;
    }
    binh = (float)(1.0 / h);
    binc = (float)(-lo / h);
    return true;
  }
  return false;
}

// Sensitivities of the set_up-derived constants to the knot parameters, for
// the adjoint.  The join (0.9 < r <= 1.9 A) depends on the knots only through
// x2 = log V(b2) and v2 = V'(b2)/V(b2) at b2 = 1.9 + eps: X = dx2/dtheta and
// Y = dv2/dtheta for every pair parameter (a_k, r_k).  Below the maximum r* of
// phi the density is the constant phi(r*); since phi'(r*) = 0 its derivative is
// the partial derivative of the knot sum at r* (cpa = d/dA_k, cpr = d/dR_k).
static void grad_consts(const Knots &k, GradDev &D) {
  memset(&D, 0, sizeof(D));
  const double eps = (double)0.000000000001f;
  const double b = k.bc2 + eps;
  const double V = knot_v(k, b), Vp = knot_dv(k, b);
  for (int m = 0; m < k.nv; m++) {
    const double d = k.rv[m] - b, H = h_step(d);
    const double dva = d * d * d * H, dvr = 3.0 * k.av[m] * d * d * H;
    const double dpa = -3.0 * d * d * H, dpr = -6.0 * k.av[m] * d * H;
    D.xa[m] = dva / V;
    D.xr[m] = dvr / V;
    D.ya[m] = dpa / V - Vp * dva / (V * V);
    D.yr[m] = dpr / V - Vp * dvr / (V * V);
  }
  for (int m = 0; m < k.np; m++) {
    const double d = k.rp[m] - k.pcr, H = h_step(d);
    D.cpa[m] = d * d * d * H;
    D.cpr[m] = 3.0 * k.ap[m] * d * d * H;
  }
}

// ---------------------------------------------------------------------------
// iron_carbon material, Fe-Fe (src/iron_carbon.F90).  Host functions in the
// reference's operation order (sums left to right, (xH0*a)*d**3, integer
// powers as left-to-right products, x**0.5 as sqrt and x**(-0.5) as pow), so
// that lookup tables built from them are the reference's tables bit for bit.
struct FecCoef {
  double vcut[MDK_FEC_NV], vpar[MDK_FEC_NV], join[4], pcut[MDK_FEC_NP], ppar[MDK_FEC_NP];
  double acl, rcl, rx;
};
// the module's coefficients (src/iron_carbon.F90:44-74), or the 38 values of
// an "iron.in" in the order check_supported_atomic_numbers reads them (:788-795)
static void fec_coef(FecCoef &f, const double *iron_in) {
  static const double vcut[MDK_FEC_NV] = {5.3, 4.7, 4.2, 3.7, 3.3, 3.0, 2.8, 2.7, 2.6, 2.5, 2.4, 2.3, 2.2};
  static const double vpar[MDK_FEC_NV] = {-0.0030458824556234, -0.058531821042271, 0.35018615891957,
                                          -1.0260416933564,   2.657740612828,     -2.3194358924605,
                                          0.80656414937789,   -0.77361294129713,  4.2099676494795,
                                          -2.4989799053251,   2.2077118733936,    15.738054058489,
                                          -27.444805994228};
  static const double join[4] = {7.411957380630116, -0.6409989885339922, -2.6049191728604733,
                                 0.6263733940718166};
  static const double pcut[MDK_FEC_NP] = {4.2, 3.2, 2.4};
  static const double ppar[MDK_FEC_NP] = {0.47193527075943, 0.01471074009883, 11.68685940797};
  const double *src = iron_in;
  memcpy(f.vcut, src ? src : vcut, sizeof(f.vcut));
  memcpy(f.vpar, src ? src + 13 : vpar, sizeof(f.vpar));
  memcpy(f.join, src ? src + 26 : join, sizeof(f.join));
  memcpy(f.pcut, src ? src + 30 : pcut, sizeof(f.pcut));
  memcpy(f.ppar, src ? src + 33 : ppar, sizeof(f.ppar));
  f.acl = src ? src[36] : 0.0;
  f.rcl = src ? src[37] : 28.5;
  f.rx = 257.0;
}
static const double fec_bfe = 9734.159682000001;
static const double fec_b[4] = {0.1818, 0.5099, 0.2802, 0.02817};
static const double fec_bx[4] = {-28.616172142695042, -8.42656844064423, -3.6029549238411973,
                                 -1.8028188449897875};
static double fec_vee(const FecCoef &f, double r) {
  if (r < 1.0)
    return (fec_bfe / r) * (fec_b[0] * exp(fec_bx[0] * r) + fec_b[1] * exp(fec_bx[1] * r) +
                            fec_b[2] * exp(fec_bx[2] * r) + fec_b[3] * exp(fec_bx[3] * r));
  if (r > (double)2.05f) {
    double s = 0.0;
    for (int k = 0; k < MDK_FEC_NV; k++) {
      const double d = f.vcut[k] - r, t = (h_step(d) * f.vpar[k]) * (d * d * d);
      s = k == 0 ? t : s + t;
    }
    return s;
  }
  return exp(f.join[0] + f.join[1] * r + f.join[2] * r * r + f.join[3] * r * r * r);
}
static double fec_dvee(const FecCoef &f, double r) {
  if (r < 1.0)
    return -(fec_bfe / r) * ((1.0 / r - fec_bx[0]) * fec_b[0] * exp(fec_bx[0] * r) +
                             (1.0 / r - fec_bx[1]) * fec_b[1] * exp(fec_bx[1] * r) +
                             (1.0 / r - fec_bx[2]) * fec_b[2] * exp(fec_bx[2] * r) +
                             (1.0 / r - fec_bx[3]) * fec_b[3] * exp(fec_bx[3] * r));
  if (r > (double)2.05f) {
    double s = 0.0;
    for (int k = 0; k < MDK_FEC_NV; k++) {
      const double d = f.vcut[k] - r, t = ((3.0 * h_step(d)) * f.vpar[k]) * (d * d);
      s = k == 0 ? t : s + t;
    }
    return -s;
  }
  const double *j = f.join;
  return (j[1] + 2.0 * j[2] * r + 3.0 * j[3] * r * r) * exp(j[0] + j[1] * r + j[2] * r * r + j[3] * r * r * r);
}
static double fec_phi(const FecCoef &f, double r) {
  double t[MDK_FEC_NP];
  for (int k = 0; k < MDK_FEC_NP; k++) {
    const double d = f.pcut[k] - r;
    t[k] = (h_step(d) * f.ppar[k]) * (d * d * d);
  }
  return t[0] - t[1] + t[2];
}
static double fec_dphi(const FecCoef &f, double r) {
  double t[MDK_FEC_NP];
  for (int k = 0; k < MDK_FEC_NP; k++) {
    const double d = f.pcut[k] - r;
    t[k] = ((3.0 * h_step(d)) * f.ppar[k]) * (d * d);
  }
  return -(t[0] - t[1] + t[2]);
}
static double fec_emb(const FecCoef &f, double rho) {
  double e = -1.0 * fabs(sqrt(rho)) - MDK_FEC_C2 * (rho * rho) + MDK_FEC_C4 * (rho * rho * rho * rho);
  if (rho > f.rcl) {
    const double d = rho - f.rcl;
    e = e + f.acl * (d * d * d);
  }
  if (rho > f.rx) {
    const double x = f.rx, d = f.rx - f.rcl;
    e = -1.0 * fabs(sqrt(x)) - MDK_FEC_C2 * (x * x) + MDK_FEC_C4 * (x * x * x * x) + f.acl * (d * d * d);
  }
  return e;
}
static double fec_demb(const FecCoef &f, double rho) {
  if (rho <= 0.0 || rho > f.rx) return 0.0;
  double d = 0.5 * -1.0 * pow(rho, -0.5) - 2.0 * MDK_FEC_C2 * rho + 4.0 * MDK_FEC_C4 * (rho * rho * rho);
  if (rho > f.rcl) {
    const double x = rho - f.rcl;
    d = d + (double)3.0f * f.acl * (x * x);
  }
  return d;
}
// device struct for iron_carbon; returns the interaction range (outermost knot)
static double fill_potdev_fec(PotDev &P, const FecCoef &f) {
  memset(&P, 0, sizeof(P));
  P.material = MDK_MAT_IRON_CARBON;
  double rint = 0.0;
  for (int k = 0; k < MDK_FEC_NV; k++) {
    P.fv_r[k] = f.vcut[k];
    P.fv_a[k] = f.vpar[k];
    P.fv_a3[k] = (3.0 * 1.0) * f.vpar[k];
    rint = std::max(rint, f.vcut[k]);
  }
  for (int k = 0; k < 4; k++) P.fjoin[k] = f.join[k];
  for (int k = 0; k < MDK_FEC_NP; k++) {
    P.fp_r[k] = f.pcut[k];
    P.fp_a[k] = f.ppar[k];
    P.fp_a3[k] = (3.0 * 1.0) * f.ppar[k];
    rint = std::max(rint, f.pcut[k]);
  }
  P.f_acl = f.acl;
  P.f_rcl = f.rcl;
  P.f_rx = f.rx;
  P.f_cap = fec_emb(f, f.rx + 1.0);
  P.r_int = rint;
  P.r_intf = (float)rint;
  return rint;
}

// The material set on a handle, for the host-side table construction.
struct Material {
  int id;  // MDK_MAT_*
  Knots k;
  FecCoef f;
  double vee(double r) const { return id == MDK_MAT_IRON_CARBON ? fec_vee(f, r) : hostpot::vee(k, r); }
  double dvee(double r) const { return id == MDK_MAT_IRON_CARBON ? fec_dvee(f, r) : hostpot::dvee(k, r); }
  double phi(double r) const { return id == MDK_MAT_IRON_CARBON ? fec_phi(f, r) : hostpot::phi(k, r); }
  double dphi(double r) const { return id == MDK_MAT_IRON_CARBON ? fec_dphi(f, r) : hostpot::dphi(k, r); }
  double emb(double x) const { return id == MDK_MAT_IRON_CARBON ? fec_emb(f, x) : hostpot::emb(x); }
  double demb(double x) const { return id == MDK_MAT_IRON_CARBON ? fec_demb(f, x) : hostpot::demb(x); }
  // get_supported_potential_range's rmax (src/fetransmetal.F90:589-593, src/iron_carbon.F90:93)
  double range() const { return id == MDK_MAT_IRON_CARBON ? 5.4 : 6.0; }
};
}  // namespace hostpot

