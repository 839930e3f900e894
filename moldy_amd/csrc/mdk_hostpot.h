// mdk_hostpot.h -- host-side potential set-up shared by the HIP translation units.
#pragma once
#include <math.h>
#include <string.h>

#include <algorithm>

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
  // interval cubics of the pair function (used by the FP32 form of 6-knot sets)
  int ord[MDK_MAXV];
  for (int i = 0; i < k.nv; i++) ord[i] = i;
  std::sort(ord, ord + k.nv, [&](int a, int b) { return k.rv[a] > k.rv[b]; });
  for (int i = 0; i < 8; i++) P.vknf[i] = i < k.nv ? (float)k.rv[ord[i]] : -1.0f;
  for (int j = 0; j <= k.nv && j < 8; j++) {
    const double u = j > 0 ? k.rv[ord[j - 1]] : (k.nv > 0 ? k.rv[ord[0]] : 0.0);
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    for (int i = 0; i < j; i++) {
      const double a = k.av[ord[i]], d = k.rv[ord[i]] - u;
      c0 += a * d * d * d;
      c1 += 3.0 * a * d * d;
      c2 += 3.0 * a * d;
      c3 += a;
    }
    const double row[8] = {u, c0, c1, c2, c3, 2.0 * c2, 3.0 * c3, 0.0};
    for (int m = 0; m < 8; m++) P.vcub[j][m] = (float)row[m];
  }
  return rint;
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
}  // namespace hostpot

