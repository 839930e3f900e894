/*
 * moldy_oracle.c -- CPU restatement of MOLDY's Fe-Fe force hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): the product library never links it.
 * Every routine cites the reference file:line it restates.  Arithmetic is kept
 * in the reference's operation order and literal precision (several Fortran
 * literals are default REAL(4); they are reproduced with float literals), and
 * the file must be compiled with -ffp-contract=off so that no FMA is formed.
 */
#include "moldy_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define T(a, i, j) (a)[((j)-1) * 3 + ((i)-1)] /* Fortran-style 1-based (i,j) */

/* src/constants_m.F90:84-113 (unit system: eV, Angstrom, amu) */
static const double k_electron = 1.602176487e-19;
static const double k_angstrom = 1.e-10;
static const double k_boltzmann = 1.3806504e-23;
static double k_fs_to_time(void) {
  const double time_to_sec = 1.018050530893617e-014;
  const double time_to_fs = time_to_sec / 1.e-15;
  return 1.0 / time_to_fs;
}
static double k_gpa_to_press(void) {
  const double a3 = k_angstrom * k_angstrom * k_angstrom;
  const double press_to_gpa = (k_electron / a3) / 1.0e9;
  return 1.0 / press_to_gpa;
}

static int g_threads = 1;
void orc_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }

/* Fortran NINT: round half away from zero (C round()). */
static inline double f_nint(double x) { return round(x); }

/* ======================================================================
 * Potential: src/fetransmetal.F90 (Fe-Fe only: find_coeff factors are 1,
 * SolSol_(:,5) = 1 at :831,937-940, so r*p == r and *dp is exact)
 * ====================================================================== */

/* Heaviside xH0, src/fetransmetal.F90:569-579 */
static inline double xh0(double x) { return (x < 0.0) ? 0.0 : 1.0; }

/* Knot sums, src/fetransmetal.F90:133-139 (vee), :186-190 (dvee),
 * :259-265 (phi), :331-336 (dphi).  (x)**3 is evaluated as (x*x)*x. */
static double knot_v(const orc_pot *p, double r) {
  double s = 0.0;
  for (int i = 0; i < p->nv; i++) {
    double d = p->rv[i] - r;
    s = s + p->av[i] * (d * d * d) * xh0(d);
  }
  return s * 1.0;
}
static double knot_dv(const orc_pot *p, double r) {
  double s = 0.0;
  for (int i = 0; i < p->nv; i++) {
    double d = p->rv[i] - r;
    s = s - 1.0 * 3.0 * p->av[i] * (d * d) * xh0(d);
  }
  return s * 1.0;
}
static double knot_p(const orc_pot *p, double r) {
  double s = 0.0;
  for (int i = 0; i < p->np; i++) {
    double d = p->rp[i] - r;
    s = s + p->ap[i] * (d * d * d) * xh0(d);
  }
  return s * 1.0;
}
static double knot_dp(const orc_pot *p, double r) {
  double s = 0.0;
  for (int i = 0; i < p->np; i++) {
    double d = p->rp[i] - r;
    s = s - 3.0 * 1.0 * p->ap[i] * (d * d) * xh0(d);
  }
  return s * 1.0;
}

/* Biersack-Ziegler screening length for Z=26: src/fetransmetal.F90:423,445.
 * All literals there are default REAL(4): the length is a float. */
static double bier_a(void) {
  volatile float c = 0.8854f * 0.529f;
  float t = powf(26.0f, 0.666666667f);
  volatile float s = t + t;
  float den = powf(s, 0.5f);
  float a = c / den;
  return (double)a;
}
static double bier_zed(void) {
  /* na1*na2 * 8.98755178d9 * electron / angstrom, :420 */
  return (double)(26 * 26) * 8.98755178e9 * k_electron / k_angstrom;
}

/* biersack, src/fetransmetal.F90:414-431 (REAL(4) literals promoted) */
static double biersack(double r) {
  double zed = bier_zed();
  double a = bier_a();
  double x = r / a;
  double ph = (double)0.1818f * exp(-((double)3.2f * x)) +
              (double)0.5099f * exp(-((double)0.9423f * x)) +
              (double)0.2802f * exp(-((double)0.4029f * x)) +
              (double)0.02817f * exp(-((double)0.2016f * x));
  return zed * ph / r;
}
/* dvee_biersack, src/fetransmetal.F90:434-454 (REAL(8) literals) */
static double dvee_biersack(double r) {
  double zed = bier_zed();
  double a = bier_a();
  double x = r / a;
  double s = (1.0 / r + 3.2 / a) * 0.1818 * exp(-(3.2 * x)) +
             (1.0 / r + 0.9423 / a) * 0.5099 * exp(-(0.9423 * x)) +
             (1.0 / r + 0.4029 / a) * 0.2802 * exp(-(0.4029 * x)) +
             (1.0 / r + 0.2016 / a) * 0.02817 * exp(-(0.2016 * x));
  return -(zed / r) * s;
}

/* expjoin / dexpjoin, src/fetransmetal.F90:511-551 */
static double expjoin(double a1, double x1, double v1, double a2, double x2, double v2,
                      double r) {
  double vv1 = v1 / x1, xx1 = log(x1), vv2 = v2 / x2, xx2 = log(x2);
  double t = (r - a1) / (a2 - a1);
  double sp = (2.0 * t * t * t - 3.0 * t * t + 1.0) * xx1 +
              (t * t * t - 2.0 * t * t + t) * vv1 * (a2 - a1) +
              (-(2.0 * t * t * t) + 3.0 * t * t) * xx2 +
              (t * t * t - t * t) * vv2 * (a2 - a1);
  return exp(sp);
}
static double dexpjoin(double a1, double x1, double v1, double a2, double x2, double v2,
                       double r) {
  double vv1 = v1 / x1, xx1 = log(x1), vv2 = v2 / x2, xx2 = log(x2);
  double t = (r - a1) / (a2 - a1);
  double ds = (6.0 * t * t - 6.0 * t) * xx1 + (3.0 * t * t - 4.0 * t + 1.0) * vv1 * (a2 - a1) +
              (-(6.0 * t * t) + 6.0 * t) * xx2 + (3.0 * t * t - 2.0 * t) * vv2 * (a2 - a1);
  ds = ds / (a2 - a1);
  double sp = (2.0 * t * t * t - 3.0 * t * t + 1.0) * xx1 +
              (t * t * t - 2.0 * t * t + t) * vv1 * (a2 - a1) +
              (-(2.0 * t * t * t) + 3.0 * t * t) * xx2 +
              (t * t * t - t * t) * vv2 * (a2 - a1);
  return ds * exp(sp);
}

/* ======================================================================
 * Potential: src/iron_carbon.F90, the Fe-Fe pair and the Fe embedding.
 * Each Fortran expression is kept in its order: sums left to right,
 * xH0(d)*a*d**3 as (xH0*a)*(d*d*d), integer powers as a left-to-right product
 * (x**4 = ((x*x)*x)*x), x**0.5d0 as |sqrt(x)| and x**(-0.5d0) as pow(); the
 * forms were fixed against the reference build's own outputs.
 * ====================================================================== */
static const double fec_bierfe = 9734.159682000001; /* :49-53 */
static const double fec_bier[4] = {0.1818, 0.5099, 0.2802, 0.02817};
static const double fec_bierx[4] = {-28.616172142695042, -8.42656844064423,
                                    -3.6029549238411973, -1.8028188449897875};
static const double fec_c2 = 6.7314115586063e-4, fec_c4 = 7.6514905604792e-8; /* :606-609 */

/* vee_src Fe-Fe, src/iron_carbon.F90:118-158; the join/knot boundary is the
 * REAL(4) literal 2.05 */
static double fec_vee(const orc_pot *p, double r) {
  if (r < 1.0) {
    return (fec_bierfe / r) * (fec_bier[0] * exp(fec_bierx[0] * r) + fec_bier[1] * exp(fec_bierx[1] * r) +
                               fec_bier[2] * exp(fec_bierx[2] * r) + fec_bier[3] * exp(fec_bierx[3] * r));
  } else if (r > (double)2.05f) {
    double s = 0.0;
    for (int k = 0; k < 13; k++) {
      const double d = p->vcut[k] - r;
      const double t = (xh0(d) * p->vpar[k]) * (d * d * d);
      s = k == 0 ? t : s + t;
    }
    return s;
  }
  const double *j = p->join;
  return exp(j[0] + j[1] * r + j[2] * r * r + j[3] * r * r * r);
}
/* dvee_src Fe-Fe, src/iron_carbon.F90:274-314 */
static double fec_dvee(const orc_pot *p, double r) {
  if (r < 1.0) {
    return -(fec_bierfe / r) *
           ((1.0 / r - fec_bierx[0]) * fec_bier[0] * exp(fec_bierx[0] * r) +
            (1.0 / r - fec_bierx[1]) * fec_bier[1] * exp(fec_bierx[1] * r) +
            (1.0 / r - fec_bierx[2]) * fec_bier[2] * exp(fec_bierx[2] * r) +
            (1.0 / r - fec_bierx[3]) * fec_bier[3] * exp(fec_bierx[3] * r));
  } else if (r > (double)2.05f) {
    double s = 0.0;
    for (int k = 0; k < 13; k++) {
      const double d = p->vcut[k] - r;
      const double t = ((3.0 * xh0(d)) * p->vpar[k]) * (d * d);
      s = k == 0 ? t : s + t;
    }
    return -s;
  }
  const double *j = p->join;
  return (j[1] + 2.0 * j[2] * r + 3.0 * j[3] * r * r) * exp(j[0] + j[1] * r + j[2] * r * r + j[3] * r * r * r);
}
/* phi_src Fe-Fe, src/iron_carbon.F90:431-449 (no flat cap) */
static double fec_phi(const orc_pot *p, double r) {
  double t[3];
  for (int k = 0; k < 3; k++) {
    const double d = p->pcut[k] - r;
    t[k] = (xh0(d) * p->ppar[k]) * (d * d * d);
  }
  return t[0] - t[1] + t[2];
}
/* dphi_src Fe-Fe, src/iron_carbon.F90:515-532 */
static double fec_dphi(const orc_pot *p, double r) {
  double t[3];
  for (int k = 0; k < 3; k++) {
    const double d = p->pcut[k] - r;
    t[k] = ((3.0 * xh0(d)) * p->ppar[k]) * (d * d);
  }
  return -(t[0] - t[1] + t[2]);
}
/* emb_src Fe, src/iron_carbon.F90:599-615 */
static double fec_emb(const orc_pot *p, double rho) {
  double e = -1.0 * fabs(sqrt(rho)) - fec_c2 * (rho * rho) + fec_c4 * (rho * rho * rho * rho);
  if (rho > p->rcludge) {
    const double d = rho - p->rcludge;
    e = e + p->acludge * (d * d * d);
  }
  if (rho > p->rx) {
    const double x = p->rx, d = p->rx - p->rcludge;
    e = -1.0 * fabs(sqrt(x)) - fec_c2 * (x * x) + fec_c4 * (x * x * x * x) + p->acludge * (d * d * d);
  }
  return e;
}
/* demb_src Fe, src/iron_carbon.F90:651-668 */
static double fec_demb(const orc_pot *p, double rho) {
  if (rho <= 0.0 || rho > p->rx) return 0.0;
  double d = 0.5 * -1.0 * pow(rho, -0.5) - 2.0 * fec_c2 * rho + 4.0 * fec_c4 * (rho * rho * rho);
  if (rho > p->rcludge) {
    const double x = rho - p->rcludge;
    d = d + (double)3.0f * p->acludge * (x * x);
  }
  return d;
}

/* Coefficients of src/iron_carbon.F90:44-74 */
void orc_pot_init_iron_carbon(orc_pot *p, const double *iron_in) {
  static const double vcut[13] = {5.3, 4.7, 4.2, 3.7, 3.3, 3.0, 2.8, 2.7, 2.6, 2.5, 2.4, 2.3, 2.2};
  static const double vpar[13] = {-0.0030458824556234, -0.058531821042271, 0.35018615891957,
                                  -1.0260416933564,   2.657740612828,     -2.3194358924605,
                                  0.80656414937789,   -0.77361294129713,  4.2099676494795,
                                  -2.4989799053251,   2.2077118733936,    15.738054058489,
                                  -27.444805994228};
  static const double join[4] = {7.411957380630116, -0.6409989885339922, -2.6049191728604733,
                                 0.6263733940718166};
  static const double pcut[3] = {4.2, 3.2, 2.4};
  static const double ppar[3] = {0.47193527075943, 0.01471074009883, 11.68685940797};
  memset(p, 0, sizeof(*p));
  p->material = ORC_MAT_IRON_CARBON;
  if (iron_in) {
    memcpy(p->vcut, iron_in, sizeof(p->vcut));
    memcpy(p->vpar, iron_in + 13, sizeof(p->vpar));
    memcpy(p->join, iron_in + 26, sizeof(p->join));
    memcpy(p->pcut, iron_in + 30, sizeof(p->pcut));
    memcpy(p->ppar, iron_in + 33, sizeof(p->ppar));
    p->acludge = iron_in[36];
    p->rcludge = iron_in[37];
  } else {
    memcpy(p->vcut, vcut, sizeof(vcut));
    memcpy(p->vpar, vpar, sizeof(vpar));
    memcpy(p->join, join, sizeof(join));
    memcpy(p->pcut, pcut, sizeof(pcut));
    memcpy(p->ppar, ppar, sizeof(ppar));
    p->acludge = 0.0;
    p->rcludge = 28.5;
  }
  p->rx = 257.0;
}

/* pot_rmax of the material (src/fetransmetal.F90:589-593, src/iron_carbon.F90:93) */
double orc_pot_range(const orc_pot *p) { return p->material == ORC_MAT_IRON_CARBON ? 5.4 : 6.0; }

/* vee_src, src/fetransmetal.F90:100-159.  r == bier_cut1 exactly gives 0. */
double orc_vee_src(const orc_pot *p, double r) {
  if (p->material == ORC_MAT_IRON_CARBON) return fec_vee(p, r);
  if (r > p->bier_cut2) return knot_v(p, r);
  if (r > p->bier_cut1)
    return expjoin(p->bier_cut1, p->sc[0], p->sc[1], p->bier_cut2, p->sc[2], p->sc[3], r);
  if (r < p->bier_cut1) return biersack(r);
  return 0.0;
}
/* dvee_src, src/fetransmetal.F90:167-219 */
double orc_dvee_src(const orc_pot *p, double r) {
  if (p->material == ORC_MAT_IRON_CARBON) return fec_dvee(p, r);
  if (r > p->bier_cut2) return knot_dv(p, r);
  if (r > p->bier_cut1)
    return dexpjoin(p->bier_cut1, p->sc[0], p->sc[1], p->bier_cut2, p->sc[2], p->sc[3], r);
  if (r < p->bier_cut1) return dvee_biersack(r);
  return 0.0;
}
/* phi_src, src/fetransmetal.F90:227-285 (flat below phi_cut(1)) */
double orc_phi_src(const orc_pot *p, double r) {
  if (p->material == ORC_MAT_IRON_CARBON) return fec_phi(p, r);
  if (r < p->phi_cut_r) return p->phi_cut_v;
  return knot_p(p, r);
}
/* dphi_src, src/fetransmetal.F90:293-366 */
double orc_dphi_src(const orc_pot *p, double r) {
  if (p->material == ORC_MAT_IRON_CARBON) return fec_dphi(p, r);
  if (r < p->phi_cut_r) return 0.0;
  return knot_dp(p, r);
}
/* emb_src / demb_src, src/fetransmetal.F90:375-404 */
static double emb_src(const orc_pot *p, double rho) {
  if (p->material == ORC_MAT_IRON_CARBON) return fec_emb(p, rho);
  return -1.0 * sqrt(rho > 0.0 ? rho : 0.0);
}
static double demb_src(const orc_pot *p, double rho) {
  if (p->material == ORC_MAT_IRON_CARBON) return fec_demb(p, rho);
  double d = 0.0;
  if (rho > 0.0) d = -0.5 / sqrt(rho);
  return d;
}

/* phi_turn, src/fetransmetal.F90:896-926: 500-step bisection on dphi */
static double phi_turn(const orc_pot *p) {
  double lo = 0.0, hi = 3.0;
  for (int i = 0; i < 500; i++) {
    double mid = (hi + lo) / 2.0;
    double g = orc_dphi_src(p, mid);
    if (g > 0.0)
      lo = mid;
    else
      hi = mid;
  }
  return (hi + lo) / 2.0;
}

/* check_supported_atomic_numbers (:671-672 cut values, REAL(4) literals)
 * and set_up_m (:768-894) for the Fe-Fe pair. */
void orc_pot_init(orc_pot *p, int nv, const double *av, const double *rv, int np,
                  const double *ap, const double *rp) {
  memset(p, 0, sizeof(*p));
  p->nv = nv;
  p->np = np;
  for (int i = 0; i < nv; i++) { p->av[i] = av[i]; p->rv[i] = rv[i]; }
  for (int i = 0; i < np; i++) { p->ap[i] = ap[i]; p->rp[i] = rp[i]; }
  p->bier_cut1 = (double)0.9f;
  p->bier_cut2 = (double)1.9f;
  const double eps = (double)0.000000000001f;
  p->sc[0] = biersack(p->bier_cut1 - eps);
  p->sc[1] = dvee_biersack(p->bier_cut1 - eps);
  p->sc[2] = orc_vee_src(p, p->bier_cut2 + eps);
  p->sc[3] = orc_dvee_src(p, p->bier_cut2 + eps);
  p->phi_cut_r = 0.0; /* not yet set while bisecting */
  p->phi_cut_v = 0.0;
  double rs = phi_turn(p);
  p->phi_cut_r = rs;
  p->phi_cut_v = orc_phi_src(p, rs);
}

/* calculate_lookup_table_veephi/_emb, src/lookup_m.F90:161-251;
 * ranges from src/fetransmetal.F90:589-593 and src/potential_m.F90:185 */
int orc_pot_build_lookup(orc_pot *p, int nlookup) {
  p->nlookup = nlookup;
  p->r_min = 0.0;
  p->r_max = orc_pot_range(p);
  p->vp_range = p->r_max - p->r_min;
  p->vp_reciprange = 1.0 / p->vp_range;
  p->vp_recipnlkp = 1.0 / (double)nlookup;
  p->vp_recipstep = (double)nlookup / p->vp_range;
  p->rho_min = 0.0;
  p->rho_max = 10000.0;
  p->e_range = p->rho_max - p->rho_min;
  p->e_reciprange = 1.0 / p->e_range;
  p->e_recipnlkp = 1.0 / (double)nlookup;
  p->e_recipstep = (double)nlookup / p->e_range;
  for (int k = 0; k < 6; k++) {
    p->tab[k] = (double *)malloc(sizeof(double) * (size_t)(nlookup + 1));
    if (!p->tab[k]) return -1;
  }
  for (int i = 0; i <= nlookup; i++) {
    double r = p->r_min + (double)i * p->vp_range * p->vp_recipnlkp;
    p->tab[0][i] = orc_vee_src(p, r);
    p->tab[1][i] = orc_dvee_src(p, r);
    p->tab[2][i] = orc_phi_src(p, r);
    p->tab[3][i] = orc_dphi_src(p, r);
    double rh = p->rho_min + (double)i * p->e_range * p->e_recipnlkp;
    p->tab[4][i] = emb_src(p, rh);
    p->tab[5][i] = demb_src(p, rh);
  }
  p->use_lookup = 1;
  return 0;
}
void orc_pot_free(orc_pot *p) {
  for (int k = 0; k < 6; k++) { free(p->tab[k]); p->tab[k] = NULL; }
  p->use_lookup = 0;
}

/* get_value_from_lookup_veephi, src/lookup_m.F90:307-346 */
static double lk_vp(const orc_pot *p, const double *t, double r) {
  if (r >= p->r_max) return t[p->nlookup];
  int lo = (int)((double)p->nlookup * (r - p->r_min) * p->vp_reciprange);
  double rl = p->r_min + p->vp_range * (double)lo * p->vp_recipnlkp;
  double dr = (r - rl) * p->vp_recipstep;
  return t[lo] * (1.0 - dr) + t[lo + 1] * dr;
}
/* get_value_from_lookup_emb, src/lookup_m.F90:356-394 */
static double lk_e(const orc_pot *p, const double *t, double rho) {
  if (rho >= p->rho_max) return t[p->nlookup];
  int lo = (int)((double)p->nlookup * (rho - p->rho_min) * p->e_reciprange);
  double rl = p->rho_min + p->e_range * (double)lo * p->e_recipnlkp;
  double dr = (rho - rl) * p->e_recipstep;
  return t[lo] * (1.0 - dr) + t[lo + 1] * dr;
}

/* potential_m dispatch, src/potential_m.F90:288-417 */
double orc_vee(const orc_pot *p, double r) { return p->use_lookup ? lk_vp(p, p->tab[0], r) : orc_vee_src(p, r); }
double orc_dvee(const orc_pot *p, double r) { return p->use_lookup ? lk_vp(p, p->tab[1], r) : orc_dvee_src(p, r); }
double orc_phi(const orc_pot *p, double r) { return p->use_lookup ? lk_vp(p, p->tab[2], r) : orc_phi_src(p, r); }
double orc_dphi(const orc_pot *p, double r) { return p->use_lookup ? lk_vp(p, p->tab[3], r) : orc_dphi_src(p, r); }
double orc_emb(const orc_pot *p, double rho) { return p->use_lookup ? lk_e(p, p->tab[4], rho) : emb_src(p, rho); }
double orc_demb(const orc_pot *p, double rho) { return p->use_lookup ? lk_e(p, p->tab[5], rho) : demb_src(p, rho); }

/* ======================================================================
 * System state (src/particles_m.F90:52-66, src/parrinellorahman_m.F90:102-115)
 * ====================================================================== */
struct orc_sys {
  int n, ivol, nnbrs, ntcm, ntc;
  int nlcx, nlcy, nlcz;
  double deltat, rpad, rcut, rnear, press, bmass, bdel2;
  double *x0, *y0, *z0, *x1, *y1, *z1, *x2, *y2, *z2, *x3, *y3, *z3;
  double *x1dt2, *y1dt2, *z1dt2;
  int iverlet; /* 1 velocity Verlet, 0 Gear predictor-corrector */
  int pc_primed; /* quench_m's parameter copy is set (by correc or velocityverlet) */
  double snhv;   /* Nose-Hoover velocity scaling, thermostat_m's module variable */
  double *mass, *del2, *rho, *afrho, *dafrho, *fx, *fy, *fz;
  int *zat;
  int *l0, *link, *ic, *numn, *nlist;
  double b0[9], b1[9], b2[9], b3[9], tk[9], tc[9], tg[9], tp[9], tginv[9], tgdot[9], tgid[9], fb[9];
  double vol, pe, ke, s1, s2, s3, te, h, th, boxke;
  orc_pot pot;
};

static double *dalloc(int n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* pr_get_metric_tensor, src/parrinellorahman_m.F90:223-227 */
static void pr_volume(orc_sys *s) {
  const double *b = s->b0;
  s->vol = T(b, 1, 1) * (T(b, 2, 2) * T(b, 3, 3) - T(b, 2, 3) * T(b, 3, 2)) +
           T(b, 2, 1) * (T(b, 3, 2) * T(b, 1, 3) - T(b, 1, 2) * T(b, 3, 3)) +
           T(b, 3, 1) * (T(b, 1, 2) * T(b, 2, 3) - T(b, 2, 2) * T(b, 1, 3));
}

/* pr_get_realsep_from_dx, src/parrinellorahman_m.F90:149-173 (dx updated in place) */
static inline double realsep(const double *b, int ivol, double *dx, double *dy, double *dz) {
  if (ivol == 0 || ivol == 1 || ivol == 4) {
    *dx = *dx - f_nint(*dx);
    *dy = *dy - f_nint(*dy);
    *dz = *dz - f_nint(*dz);
  } else if (ivol == 2) {
    *dx = *dx - f_nint(*dx);
    *dy = *dy - f_nint(*dy);
  } else if (ivol == 5) {
    *dz = *dz - f_nint(*dz);
  }
  double rx = T(b, 1, 1) * *dx + T(b, 1, 2) * *dy + T(b, 1, 3) * *dz;
  double ry = T(b, 2, 1) * *dx + T(b, 2, 2) * *dy + T(b, 2, 3) * *dz;
  double rz = T(b, 3, 1) * *dx + T(b, 3, 2) * *dy + T(b, 3, 3) * *dz;
  return sqrt(rx * rx + ry * ry + rz * rz);
}

orc_sys *orc_create(int n, const double *b0, const double *x0, const double *y0,
                    const double *z0, const double *x1, const double *y1,
                    const double *z1, const double *mass, const int *z, int ivol,
                    double deltat_fs, double rpad, double rcut, int nnbrs, int ntcm,
                    double press_gpa, double boxmass, const orc_pot *pot) {
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  if (!s) return NULL;
  s->n = n;
  s->ivol = ivol;
  s->nnbrs = nnbrs;
  s->ntcm = ntcm;
  s->ntc = 0;
  /* src/params_m.F90:412 (fs -> model time), :454 (GPa -> model), :207-208 */
  s->deltat = deltat_fs * k_fs_to_time();
  s->press = press_gpa * k_gpa_to_press();
  s->bmass = (double)n * boxmass;
  s->bdel2 = s->deltat * s->deltat / (2.0 * s->bmass);
  s->rpad = rpad;
  /* moldin.F90:226 with the material's range (fetransmetal.F90:589-593, iron_carbon.F90:93) */
  s->rcut = rcut > orc_pot_range(pot) ? rcut : orc_pot_range(pot);
  s->rnear = s->rcut + s->rpad;      /* moldin.F90:248 */
  memcpy(s->b0, b0, sizeof(s->b0));
  /* moldin.F90:252-257: link cells from the diagonal, at least 3, never recomputed */
  s->nlcx = (int)(T(b0, 1, 1) / s->rnear);
  s->nlcy = (int)(T(b0, 2, 2) / s->rnear);
  s->nlcz = (int)(T(b0, 3, 3) / s->rnear);
  if (s->nlcx < 3) s->nlcx = 3;
  if (s->nlcy < 3) s->nlcy = 3;
  if (s->nlcz < 3) s->nlcz = 3;
  s->x0 = dalloc(n); s->y0 = dalloc(n); s->z0 = dalloc(n);
  s->x1 = dalloc(n); s->y1 = dalloc(n); s->z1 = dalloc(n);
  s->x2 = dalloc(n); s->y2 = dalloc(n); s->z2 = dalloc(n);
  s->x3 = dalloc(n); s->y3 = dalloc(n); s->z3 = dalloc(n);
  s->iverlet = 1;
  s->x1dt2 = dalloc(n); s->y1dt2 = dalloc(n); s->z1dt2 = dalloc(n);
  s->mass = dalloc(n); s->del2 = dalloc(n);
  s->rho = dalloc(n); s->afrho = dalloc(n); s->dafrho = dalloc(n);
  s->fx = dalloc(n); s->fy = dalloc(n); s->fz = dalloc(n);
  s->zat = (int *)calloc((size_t)n, sizeof(int));
  int ncell = s->nlcx * s->nlcy * s->nlcz;
  s->l0 = (int *)calloc((size_t)ncell, sizeof(int));
  s->link = (int *)calloc((size_t)n + 1, sizeof(int));
  s->ic = (int *)calloc((size_t)n + 1, sizeof(int));
  s->numn = (int *)calloc((size_t)n, sizeof(int));
  s->nlist = (int *)calloc((size_t)n * (size_t)nnbrs, sizeof(int));
  if (!s->nlist || !s->zat) { orc_destroy(s); return NULL; }
  for (int i = 0; i < n; i++) {
    s->x0[i] = x0[i]; s->y0[i] = y0[i]; s->z0[i] = z0[i];
    s->x1[i] = x1 ? x1[i] : 0.0; s->y1[i] = y1 ? y1[i] : 0.0; s->z1[i] = z1 ? z1[i] : 0.0;
    s->mass[i] = mass[i];
    s->zat[i] = z[i];
    /* moldin.F90:294-300 (2.0 is an exact REAL(4) literal) */
    s->del2[i] = (mass[i] > (double)0.01f) ? s->deltat * s->deltat / (2.0 * mass[i]) : 0.0;
  }
  s->pot = *pot;
  pr_volume(s);
  return s;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  double *d[] = {s->x0, s->y0, s->z0, s->x1, s->y1, s->z1, s->x2, s->y2, s->z2, s->x3, s->y3, s->z3,
                 s->x1dt2, s->y1dt2, s->z1dt2, s->mass, s->del2, s->rho, s->afrho,
                 s->dafrho, s->fx, s->fy, s->fz};
  for (size_t k = 0; k < sizeof(d) / sizeof(d[0]); k++) free(d[k]);
  free(s->zat); free(s->l0); free(s->link); free(s->ic); free(s->numn); free(s->nlist);
  free(s);
}

/* ======================================================================
 * Link cells, src/linkcell_m.F90 (1-based atom and cell indices inside)
 * ====================================================================== */
static const double hmeps = 0.4999999999;

static inline int cell_of(const orc_sys *s, int i /*0-based*/) {
  int ix = (int)((s->x0[i] - f_nint(s->x0[i]) + hmeps) * (double)s->nlcx);
  int iy = (int)((s->y0[i] - f_nint(s->y0[i]) + hmeps) * (double)s->nlcy);
  int iz = (int)((s->z0[i] - f_nint(s->z0[i]) + hmeps) * (double)s->nlcz);
  return ix + s->nlcx * iy + s->nlcx * s->nlcy * iz + 1;
}

/* linkup, src/linkcell_m.F90:122-198 */
int orc_linkup(orc_sys *s) {
  int ncell = s->nlcx * s->nlcy * s->nlcz;
  int *l1 = (int *)calloc((size_t)ncell + 1, sizeof(int));
  if (!l1) return -1;
  for (int c = 0; c < ncell; c++) s->l0[c] = 0;
  for (int i = 1; i <= s->n; i++) {
    int ip = cell_of(s, i - 1);
    s->ic[i] = ip;
    int me = l1[ip - 1];
    if (me != 0) {
      s->link[me] = i;
      l1[ip - 1] = i;
    } else {
      s->l0[ip - 1] = i;
      l1[ip - 1] = i;
    }
  }
  for (int ip = 1; ip <= ncell; ip++) {
    int me = l1[ip - 1];
    if (me == 0) continue;
    s->link[me] = s->l0[ip - 1];
  }
  free(l1);
  return 0;
}

/* relink, src/linkcell_m.F90:215-274 */
int orc_relink(orc_sys *s) {
  int *l0 = s->l0, *link = s->link, *ic = s->ic;
  for (int i = 1; i <= s->n; i++) {
    int ip = cell_of(s, i - 1);
    if (ip == ic[i]) continue;
    int me = ic[i];
    int j = l0[me - 1];
    for (;;) {
      if (link[j] == i) break;
      j = link[j];
    }
    link[j] = link[i];
    l0[me - 1] = link[j];
    if (i == j) l0[me - 1] = 0;
    j = l0[ip - 1];
    if (l0[ip - 1] == 0) j = i;
    link[i] = link[j];
    link[j] = i;
    ic[i] = ip;
    l0[ip - 1] = i;
  }
  return 0;
}

/* update_neighbourlist, src/neighbourlist_m.F90:103-276: half list over the
 * own-cell forward chain and 13 forward cells; NLIST(k,i) stored at
 * nlist[(i-1)*nnbrs + k-1] (column-major NLIST(nnbrs,nm)). */
int orc_update_neighbourlist(orc_sys *s) {
  const int nlcx = s->nlcx, nlcy = s->nlcy, nlcz = s->nlcz;
  const int *l0 = s->l0, *link = s->link;
  for (int i = 0; i < s->n; i++) s->numn[i] = 0;
  int overflow = 0;
  for (int ip = 1; ip <= nlcx * nlcy * nlcz && !overflow; ip++) {
    int ix = (ip - 1) % nlcx + 1;
    int iy = ((ip - 1) / nlcx) % nlcy + 1;
    int iz = ((ip - 1) / (nlcx * nlcy)) % nlcz + 1;
    int i = l0[ip - 1];
    if (i == 0) continue;
    for (int kz = 2; kz <= 3; kz++) {
      int jz = iz + kz - 2;
      if (jz - nlcz - 1 == 0) jz = 1;
      int kk = (kz == 3) ? 1 : 2;
      int jzn = (jz - 1) * nlcx * nlcy;
      for (int ky = kk; ky <= 3; ky++) {
        int jy = iy + ky - 2;
        if (jy == 0) jy = nlcy;
        if (jy - nlcy - 1 == 0) jy = 1;
        int kl = (kz == 2 && ky == 2) ? 2 : 1;
        int jyzn = jzn + (jy - 1) * nlcx;
        for (int kx = kl; kx <= 3; kx++) {
          int jx = ix + kx - 2;
          if (jx == 0) jx = nlcx;
          if (jx - nlcx - 1 == 0) jx = 1;
          int jp = jx + jyzn;
          for (;;) { /* linkloop */
            int j;
            if (ip == jp) {
              j = link[i];
              if (j == l0[ip - 1]) { i = l0[ip - 1]; goto next_kx; }
            } else {
              j = l0[jp - 1];
              if (j == 0) goto next_kx;
            }
            for (;;) { /* testcandidates */
              double dx = s->x0[i - 1] - s->x0[j - 1];
              double dy = s->y0[i - 1] - s->y0[j - 1];
              double dz = s->z0[i - 1] - s->z0[j - 1];
              double r = realsep(s->b0, s->ivol, &dx, &dy, &dz);
              if (r < s->rnear) {
                s->numn[i - 1] += 1;
                if (s->numn[i - 1] > s->nnbrs) { overflow = 1; goto done; }
                s->nlist[(size_t)(i - 1) * s->nnbrs + s->numn[i - 1] - 1] = j;
              }
              j = link[j];
              if (j == l0[jp - 1]) break;
            }
            i = link[i];
            if (i == l0[ip - 1]) goto next_kx;
          }
        next_kx:;
        }
      }
    }
  }
done:
  return overflow ? 1 : 0;
}

/* rhoset, src/dynamics_m.F90:497-579 (serial order: scatter to j, gather to i) */
void orc_rhoset(orc_sys *s) {
  const int n = s->n, nn = s->nnbrs;
  for (int i = 0; i < n; i++) s->rho[i] = 0.0;
#ifdef _OPENMP
  if (g_threads > 1) {
    int nt = g_threads;
    double *acc = (double *)calloc((size_t)n * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
      double *my = acc + (size_t)omp_get_thread_num() * n;
#pragma omp for schedule(static)
      for (int i = 0; i < n; i++) {
        double rt = 0.0;
        for (int k = 0; k < s->numn[i]; k++) {
          int j = s->nlist[(size_t)i * nn + k] - 1;
          double dx = s->x0[i] - s->x0[j], dy = s->y0[i] - s->y0[j], dz = s->z0[i] - s->z0[j];
          double r = realsep(s->b0, s->ivol, &dx, &dy, &dz);
          my[j] = my[j] + orc_phi(&s->pot, r);
          rt = rt + orc_phi(&s->pot, r);
        }
        my[i] = my[i] + rt;
      }
#pragma omp for schedule(static)
      for (int i = 0; i < n; i++) {
        double v = 0.0;
        for (int t = 0; t < nt; t++) v += acc[(size_t)t * n + i];
        s->rho[i] = v;
      }
    }
    free(acc);
    return;
  }
#endif
  for (int i = 0; i < n; i++) {
    double rt = 0.0;
    for (int k = 0; k < s->numn[i]; k++) {
      int j = s->nlist[(size_t)i * nn + k] - 1;
      double dx = s->x0[i] - s->x0[j], dy = s->y0[i] - s->y0[j], dz = s->z0[i] - s->z0[j];
      double r = realsep(s->b0, s->ivol, &dx, &dy, &dz);
      s->rho[j] = s->rho[j] + orc_phi(&s->pot, r);
      rt = rt + orc_phi(&s->pot, r);
    }
    s->rho[i] = s->rho[i] + rt;
  }
}

/* embed, src/potential_m.F90:148-169 (rho <= 1e-40 leaves stale values) */
void orc_embed(orc_sys *s) {
  double pe = 0.0;
  for (int i = 0; i < s->n; i++) {
    if (s->rho[i] <= 1e-40) continue;
    s->afrho[i] = orc_emb(&s->pot, s->rho[i]);
    pe = pe + s->afrho[i];
    s->dafrho[i] = orc_demb(&s->pot, s->rho[i]);
  }
  s->pe = pe;
}

/* pair pass of force, src/dynamics_m.F90:235-346 */
static void force_pair_loop(orc_sys *s, int i, double *fxs, double *fys, double *fzs,
                            double *pe, double *tp) {
  const double *b = s->b0;
  const int nn = s->nnbrs;
  double fxi = 0.0, fyi = 0.0, fzi = 0.0;
  for (int k = 0; k < s->numn[i]; k++) {
    int j = s->nlist[(size_t)i * nn + k] - 1;
    double dx = s->x0[i] - s->x0[j], dy = s->y0[i] - s->y0[j], dz = s->z0[i] - s->z0[j];
    double r = realsep(b, s->ivol, &dx, &dy, &dz);
    double rr = 1.0 / r;
    double pp = orc_vee(&s->pot, r);
    double fpp = -orc_dvee(&s->pot, r) * rr;
    *pe = *pe + pp;
    double fcp = -1.0 * (orc_dphi(&s->pot, r) * s->dafrho[i] + orc_dphi(&s->pot, r) * s->dafrho[j]) * rr;
    double fp = fpp + fcp;
    double rx = T(b, 1, 1) * dx + T(b, 1, 2) * dy + T(b, 1, 3) * dz;
    double ry = T(b, 2, 1) * dx + T(b, 2, 2) * dy + T(b, 2, 3) * dz;
    double rz = T(b, 3, 1) * dx + T(b, 3, 2) * dy + T(b, 3, 3) * dz;
    if ((s->mass[i] + s->mass[j]) > (double)0.1f) {
      T(tp, 1, 1) = T(tp, 1, 1) + dx * rx * fp;
      T(tp, 2, 1) = T(tp, 2, 1) + dx * ry * fp;
      T(tp, 3, 1) = T(tp, 3, 1) + dx * rz * fp;
      T(tp, 1, 2) = T(tp, 1, 2) + dy * rx * fp;
      T(tp, 2, 2) = T(tp, 2, 2) + dy * ry * fp;
      T(tp, 3, 2) = T(tp, 3, 2) + dy * rz * fp;
      T(tp, 1, 3) = T(tp, 1, 3) + dz * rx * fp;
      T(tp, 2, 3) = T(tp, 2, 3) + dz * ry * fp;
      T(tp, 3, 3) = T(tp, 3, 3) + dz * rz * fp;
    }
    double ddfx = -dx * fp, ddfy = -dy * fp, ddfz = -dz * fp;
    fxs[j] = fxs[j] + ddfx;
    fys[j] = fys[j] + ddfy;
    fzs[j] = fzs[j] + ddfz;
    fxi = fxi + ddfx;
    fyi = fyi + ddfy;
    fzi = fzi + ddfz;
  }
  fxs[i] = fxs[i] - fxi;
  fys[i] = fys[i] - fyi;
  fzs[i] = fzs[i] - fzi;
}

/* kinten, src/thermostat_m.F90:127-177 */
void orc_kinten(orc_sys *s) {
  const double *b = s->b0;
  for (int k = 0; k < 9; k++) s->tk[k] = 0.0;
  for (int i = 0; i < s->n; i++) {
    if (s->zat[i] == 0) continue;
    double x1 = s->x1[i], y1 = s->y1[i], z1 = s->z1[i], m = s->mass[i];
    double mvx = (T(b, 1, 1) * x1 + T(b, 1, 2) * y1 + T(b, 1, 3) * z1) * m;
    double mvy = (T(b, 2, 1) * x1 + T(b, 2, 2) * y1 + T(b, 2, 3) * z1) * m;
    double mvz = (T(b, 3, 1) * x1 + T(b, 3, 2) * y1 + T(b, 3, 3) * z1) * m;
    T(s->tk, 1, 1) += x1 * mvx; T(s->tk, 2, 1) += x1 * mvy; T(s->tk, 3, 1) += x1 * mvz;
    T(s->tk, 1, 2) += y1 * mvx; T(s->tk, 2, 2) += y1 * mvy; T(s->tk, 3, 2) += y1 * mvz;
    T(s->tk, 1, 3) += z1 * mvx; T(s->tk, 2, 3) += z1 * mvy; T(s->tk, 3, 3) += z1 * mvz;
  }
  for (int k = 0; k < 9; k++) s->tk[k] = s->tk[k] / (s->deltat * s->deltat);
  for (int k = 0; k < 9; k++) s->tgdot[k] = 0.0;
  for (int i = 1; i <= 3; i++)
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++)
        T(s->tgdot, i, j) = T(s->tgdot, i, j) + T(b, k, i) * T(s->b1, k, j) + T(s->b1, k, i) * T(b, k, j);
}

/* remainder of force after the pair loop: tc, kinten, fb (dynamics_m.F90:365-394) */
static void force_tail(orc_sys *s) {
  const double *b = s->b0;
  double *tc = s->tc;
  T(tc, 1, 1) = T(b, 2, 2) * T(b, 3, 3) - T(b, 2, 3) * T(b, 3, 2);
  T(tc, 2, 1) = T(b, 3, 2) * T(b, 1, 3) - T(b, 1, 2) * T(b, 3, 3);
  T(tc, 3, 1) = T(b, 1, 2) * T(b, 2, 3) - T(b, 2, 2) * T(b, 1, 3);
  T(tc, 1, 2) = T(b, 2, 3) * T(b, 3, 1) - T(b, 3, 3) * T(b, 2, 1);
  T(tc, 2, 2) = T(b, 3, 3) * T(b, 1, 1) - T(b, 1, 3) * T(b, 3, 1);
  T(tc, 3, 2) = T(b, 1, 3) * T(b, 2, 1) - T(b, 1, 1) * T(b, 2, 3);
  T(tc, 1, 3) = T(b, 2, 1) * T(b, 3, 2) - T(b, 3, 1) * T(b, 2, 2);
  T(tc, 2, 3) = T(b, 3, 1) * T(b, 1, 2) - T(b, 1, 1) * T(b, 3, 2);
  T(tc, 3, 3) = T(b, 1, 1) * T(b, 2, 2) - T(b, 2, 1) * T(b, 1, 2);
  orc_kinten(s);
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++)
      T(s->fb, i, j) = T(s->tk, i, j) + T(s->tp, i, j) - s->press * T(s->tc, i, j);
}

void orc_force_pairs(orc_sys *s) {
  const int n = s->n;
  for (int k = 0; k < 9; k++) { s->tp[k] = 0.0; s->tg[k] = 0.0; }
  for (int i = 0; i < n; i++) { s->fx[i] = 0.0; s->fy[i] = 0.0; s->fz[i] = 0.0; }
  pr_volume(s);
  orc_rhoset(s);
  orc_embed(s);
  for (int i = 1; i <= 3; i++)
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++) T(s->tg, i, j) = T(s->tg, i, j) + T(s->b0, k, i) * T(s->b0, k, j);
#ifdef _OPENMP
  if (g_threads > 1) {
    int nt = g_threads;
    double *acc = (double *)calloc((size_t)3 * n * nt, sizeof(double));
    double *pes = (double *)calloc((size_t)nt, sizeof(double));
    double *tps = (double *)calloc((size_t)9 * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
      int t = omp_get_thread_num();
      double *mx = acc + (size_t)3 * n * t, *my = mx + n, *mz = my + n;
#pragma omp for schedule(static)
      for (int i = 0; i < n; i++) force_pair_loop(s, i, mx, my, mz, &pes[t], &tps[9 * t]);
#pragma omp for schedule(static)
      for (int i = 0; i < n; i++) {
        double ax = 0, ay = 0, az = 0;
        for (int q = 0; q < nt; q++) {
          ax += acc[(size_t)3 * n * q + i];
          ay += acc[(size_t)3 * n * q + n + i];
          az += acc[(size_t)3 * n * q + 2 * n + i];
        }
        s->fx[i] = ax; s->fy[i] = ay; s->fz[i] = az;
      }
    }
    for (int q = 0; q < nt; q++) {
      s->pe += pes[q];
      for (int k = 0; k < 9; k++) s->tp[k] += tps[9 * q + k];
    }
    free(acc); free(pes); free(tps);
    force_tail(s);
    return;
  }
#endif
  for (int i = 0; i < n; i++) force_pair_loop(s, i, s->fx, s->fy, s->fz, &s->pe, s->tp);
  force_tail(s);
}

/* force, src/dynamics_m.F90:135-401 (ntc bookkeeping :186-198) */
int orc_force(orc_sys *s) {
  int st = 0;
  s->ntc = s->ntc + 1;
  if (s->ntc == 1) st = orc_update_neighbourlist(s);
  if (s->ntc == s->ntcm) s->ntc = 0;
  if (st) return st;
  orc_force_pairs(s);
  return 0;
}

/* matrix_invert / matrix_multiply, src/matrix_m.F90:60-119 */
static void mat_invert(const double *a, double *b) {
  double det = T(a, 1, 1) * (T(a, 2, 2) * T(a, 3, 3) - T(a, 2, 3) * T(a, 3, 2));
  det = det + T(a, 2, 1) * (T(a, 3, 2) * T(a, 1, 3) - T(a, 1, 2) * T(a, 3, 3));
  det = det + T(a, 3, 1) * (T(a, 1, 2) * T(a, 2, 3) - T(a, 1, 3) * T(a, 2, 2));
  T(b, 1, 1) = T(a, 2, 2) * T(a, 3, 3) - T(a, 2, 3) * T(a, 3, 2);
  T(b, 1, 2) = T(a, 3, 1) * T(a, 2, 3) - T(a, 2, 1) * T(a, 3, 3);
  T(b, 1, 3) = T(a, 2, 1) * T(a, 3, 2) - T(a, 3, 1) * T(a, 2, 2);
  T(b, 2, 1) = T(a, 3, 2) * T(a, 1, 3) - T(a, 1, 2) * T(a, 3, 3);
  T(b, 2, 2) = T(a, 1, 1) * T(a, 3, 3) - T(a, 1, 3) * T(a, 3, 1);
  T(b, 2, 3) = T(a, 3, 1) * T(a, 1, 2) - T(a, 1, 1) * T(a, 3, 2);
  T(b, 3, 1) = T(a, 1, 2) * T(a, 2, 3) - T(a, 1, 3) * T(a, 2, 2);
  T(b, 3, 2) = T(a, 2, 1) * T(a, 1, 3) - T(a, 2, 3) * T(a, 1, 1);
  T(b, 3, 3) = T(a, 1, 1) * T(a, 2, 2) - T(a, 1, 2) * T(a, 2, 1);
  for (int k = 0; k < 9; k++) b[k] = b[k] / det;
}
static void mat_mul(const double *a, const double *b, double *c) {
  for (int i = 1; i <= 3; i++) {
    for (int j = 1; j <= 3; j++) T(c, i, j) = 0.0;
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++) T(c, i, j) = T(c, i, j) + T(a, i, k) * T(b, k, j);
  }
}

/* moldin.F90:446-477: first force, tgid, initial accelerations and box acceleration */
int orc_md_init(orc_sys *s) {
  int st = orc_force(s);
  if (st) return st;
  mat_invert(s->tg, s->tginv);
  mat_mul(s->tginv, s->tgdot, s->tgid);
  const double *g = s->tgid;
  for (int i = 0; i < s->n; i++) {
    if (s->zat[i] == 0) continue; /* atomic_index==0 marks a vacancy */
    double a1 = T(g, 1, 1) * s->x1[i] + T(g, 1, 2) * s->y1[i] + T(g, 1, 3) * s->z1[i];
    double a2 = T(g, 2, 1) * s->x1[i] + T(g, 2, 2) * s->y1[i] + T(g, 2, 3) * s->z1[i];
    double a3 = T(g, 3, 1) * s->x1[i] + T(g, 3, 2) * s->y1[i] + T(g, 3, 3) * s->z1[i];
    s->x2[i] = s->del2[i] * s->fx[i] - 0.5 * a1;
    s->y2[i] = s->del2[i] * s->fy[i] - 0.5 * a2;
    s->z2[i] = s->del2[i] * s->fz[i] - 0.5 * a3;
  }
  for (int k = 0; k < 9; k++) s->b2[k] = s->bdel2 * s->fb[k];
  return 0;
}

/* velocityverlet, src/quench_m.F90:562-624 */
static int velocity_verlet(orc_sys *s) {
  const int n = s->n;
  s->pc_primed = 1; /* simparam=get_params(), :569 */
  for (int i = 0; i < n; i++) {
    s->x1dt2[i] = s->x1[i] + s->x2[i];
    s->y1dt2[i] = s->y1[i] + s->y2[i];
    s->z1dt2[i] = s->z1[i] + s->z2[i];
  }
  for (int i = 0; i < n; i++) {
    s->x0[i] = s->x0[i] + s->x1dt2[i];
    s->y0[i] = s->y0[i] + s->y1dt2[i];
    s->z0[i] = s->z0[i] + s->z1dt2[i];
  }
  if (s->ivol == 4) {
    T(s->b1, 3, 3) = T(s->b1, 3, 3) + T(s->b2, 3, 3);
    T(s->b0, 3, 3) = T(s->b0, 3, 3) + T(s->b1, 3, 3);
  }
  if (s->ivol < 1) {
    for (int k = 0; k < 9; k++) s->b1[k] = s->b1[k] + s->b2[k];
    for (int k = 0; k < 9; k++) s->b0[k] = s->b0[k] + s->b1[k];
  }
  int st = orc_force(s);
  if (st) return st;
  double g[9];
  memcpy(g, s->tgid, sizeof(g)); /* lagged tgid, quench_m.F90:595 */
  if (s->ivol < 1) {
    for (int k = 0; k < 9; k++) s->b2[k] = s->bdel2 * s->fb[k];
    for (int k = 0; k < 9; k++) s->b1[k] = s->b1[k] + s->b2[k];
  }
  if (s->ivol == 4) {
    T(s->b2, 3, 3) = s->bdel2 * T(s->fb, 3, 3);
    for (int k = 0; k < 9; k++) s->b1[k] = s->b1[k] + s->b2[k];
  }
  double sxi = 0.0, syi = 0.0, szi = 0.0;
  for (int i = 0; i < n; i++) {
    if (s->zat[i] == 0 || s->mass[i] < (double)0.1f) continue;
    if (s->ivol == 0) {
      sxi = T(g, 1, 1) * s->x1[i] + T(g, 1, 2) * s->y1[i] + T(g, 1, 3) * s->z1[i];
      syi = T(g, 2, 1) * s->x1[i] + T(g, 2, 2) * s->y1[i] + T(g, 2, 3) * s->z1[i];
      szi = T(g, 3, 1) * s->x1[i] + T(g, 3, 2) * s->y1[i] + T(g, 3, 3) * s->z1[i];
    }
    if (s->ivol == 4) szi = T(g, 3, 3) * s->z1[i];
    s->x2[i] = s->del2[i] * s->fx[i] - 0.5 * sxi;
    s->y2[i] = s->del2[i] * s->fy[i] - 0.5 * syi;
    s->z2[i] = s->del2[i] * s->fz[i] - 0.5 * szi;
  }
  for (int i = 0; i < n; i++) {
    s->x1[i] = s->x1dt2[i] + s->x2[i];
    s->y1[i] = s->y1dt2[i] + s->y2[i];
    s->z1[i] = s->z1dt2[i] + s->z2[i];
  }
  return 0;
}

/* update_therm_avs, src/dynamics_m.F90:649-734 */
static void update_therm_avs(orc_sys *s) {
  const double *b0 = s->b0, *b1 = s->b1;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int k = 0; k < 9; k++) { s->tg[k] = 0.0; s->tgdot[k] = 0.0; }
  for (int i = 1; i <= 3; i++)
    for (int k = 1; k <= 3; k++)
      for (int j = 1; j <= 3; j++) {
        T(s->tg, i, j) = T(s->tg, i, j) + T(b0, k, i) * T(b0, k, j);
        if (s->ivol != 1)
          T(s->tgdot, i, j) = T(s->tgdot, i, j) + T(b1, k, i) * T(b0, k, j) + T(b0, k, i) * T(b1, k, j);
      }
  mat_invert(s->tg, s->tginv);          /* pr_get_tgid */
  mat_mul(s->tginv, s->tgdot, s->tgid);
  pr_volume(s);
  const double v = s->vol;
  double xx = (T(b0, 2, 2) * T(b0, 3, 3) - T(b0, 2, 3) * T(b0, 3, 2)) / v;
  double xy = (T(b0, 2, 3) * T(b0, 3, 1) - T(b0, 3, 3) * T(b0, 2, 1)) / v;
  double xz = (T(b0, 2, 1) * T(b0, 3, 2) - T(b0, 3, 1) * T(b0, 2, 2)) / v;
  double yx = (T(b0, 3, 2) * T(b0, 1, 3) - T(b0, 1, 2) * T(b0, 3, 3)) / v;
  double yy = (T(b0, 1, 1) * T(b0, 3, 3) - T(b0, 1, 3) * T(b0, 3, 1)) / v;
  double yz = (T(b0, 3, 1) * T(b0, 1, 2) - T(b0, 1, 1) * T(b0, 3, 2)) / v;
  double zx = (T(b0, 1, 2) * T(b0, 2, 3) - T(b0, 2, 2) * T(b0, 1, 3)) / v;
  double zy = (T(b0, 1, 3) * T(b0, 2, 1) - T(b0, 1, 1) * T(b0, 2, 3)) / v;
  double zz = (T(b0, 1, 1) * T(b0, 2, 2) - T(b0, 1, 2) * T(b0, 2, 1)) / v;
  const double *g = s->tgid;
  for (int i = 0; i < s->n; i++) {
    if (s->zat[i] == 0) continue;
    double sxi = T(g, 1, 1) * s->x1[i] + T(g, 1, 2) * s->y1[i] + T(g, 1, 3) * s->z1[i];
    double syi = T(g, 2, 1) * s->x1[i] + T(g, 2, 2) * s->y1[i] + T(g, 2, 3) * s->z1[i];
    double szi = T(g, 3, 1) * s->x1[i] + T(g, 3, 2) * s->y1[i] + T(g, 3, 3) * s->z1[i];
    double fxp = 2.0 * s->x2[i] + sxi, fyp = 2.0 * s->y2[i] + syi, fzp = 2.0 * s->z2[i] + szi;
    double fxi = xx * fxp + xy * fyp + xz * fzp;
    double fyi = yx * fxp + yy * fyp + yz * fzp;
    double fzi = zx * fxp + zy * fyp + zz * fzp;
    double f2 = fxi * fxi + fyi * fyi + fzi * fzi;
    s1 = s1 + f2;
    s2 = s2 + f2 * f2;
    double rx = T(b0, 1, 1) * s->x1[i] + T(b0, 1, 2) * s->y1[i] + T(b0, 1, 3) * s->z1[i];
    double ry = T(b0, 2, 1) * s->x1[i] + T(b0, 2, 2) * s->y1[i] + T(b0, 2, 3) * s->z1[i];
    double rz = T(b0, 3, 1) * s->x1[i] + T(b0, 3, 2) * s->y1[i] + T(b0, 3, 3) * s->z1[i];
    s3 = s3 + s->mass[i] * (rx * rx + ry * ry + rz * rz);
  }
  s->s1 = s1;
  s->s2 = s2;
  s->s3 = s3;
}

/* predic, src/quench_m.F90:636-671.  predic reads quench_m's module copy of
 * the parameters, which only correc (:690) and velocityverlet (:569) refresh
 * (its own get_params is commented out, :640) and which moldin never sets in an
 * MD run (init_quench_m only for quenches, moldin.F90:566-575): until then it
 * holds nm = 0 and ivol = -1 (src/params_m.F90:81,47), so predic changes
 * nothing. */
void orc_predic(orc_sys *s) {
  if (!s->pc_primed) return;
  for (int i = 0; i < s->n; i++) {
    s->x0[i] = s->x0[i] + s->x1[i] + s->x2[i] + s->x3[i];
    s->y0[i] = s->y0[i] + s->y1[i] + s->y2[i] + s->y3[i];
    s->z0[i] = s->z0[i] + s->z1[i] + s->z2[i] + s->z3[i];
    s->x1[i] = s->x1[i] + 2.0 * s->x2[i] + 3.0 * s->x3[i];
    s->y1[i] = s->y1[i] + 2.0 * s->y2[i] + 3.0 * s->y3[i];
    s->z1[i] = s->z1[i] + 2.0 * s->z2[i] + 3.0 * s->z3[i];
    s->x2[i] = s->x2[i] + 3.0 * s->x3[i];
    s->y2[i] = s->y2[i] + 3.0 * s->y3[i];
    s->z2[i] = s->z2[i] + 3.0 * s->z3[i];
  }
  for (int k = 0; k < 9; k++) {
    if (s->ivol == 0 || (s->ivol == 4 && k == 8)) {
      s->b0[k] = s->b0[k] + s->b1[k] + s->b2[k] + s->b3[k];
      s->b1[k] = s->b1[k] + 2.0 * s->b2[k] + 3.0 * s->b3[k];
      s->b2[k] = s->b2[k] + 3.0 * s->b3[k];
    }
  }
}

/* correc, src/quench_m.F90:682-759.  The four coefficients are REAL(4)
 * literals (the parameters carry no d0 suffix): 1/6, 5/6, 1, 1/3 in float. */
void orc_correc(orc_sys *s) {
  s->pc_primed = 1;
  const double ct0 = (double)0.166666666666666667f, ct1 = (double)0.833333333333333333f;
  const double ct2 = (double)1.000000000000000000f, ct3 = (double)0.333333333333333333f;
  for (int k = 0; k < 9; k++) {
    if (s->ivol < 1 || (s->ivol == 4 && k == 8)) {
      const double cc = s->bdel2 * s->fb[k] - s->b2[k];
      s->b0[k] = s->b0[k] + ct0 * cc;
      s->b1[k] = s->b1[k] + ct1 * cc;
      s->b2[k] = s->b2[k] + ct2 * cc;
      s->b3[k] = s->b3[k] + ct3 * cc;
    }
  }
  const double *g = s->tgid; /* get_tgid(): the value before this step's update_therm_avs */
  for (int i = 0; i < s->n; i++) {
    if (s->zat[i] == 0) continue;
    const double sxi = T(g, 1, 1) * s->x1[i] + T(g, 1, 2) * s->y1[i] + T(g, 1, 3) * s->z1[i];
    const double syi = T(g, 2, 1) * s->x1[i] + T(g, 2, 2) * s->y1[i] + T(g, 2, 3) * s->z1[i];
    const double szi = T(g, 3, 1) * s->x1[i] + T(g, 3, 2) * s->y1[i] + T(g, 3, 3) * s->z1[i];
    const double ccx = s->del2[i] * s->fx[i] - 0.5 * sxi - s->x2[i];
    const double ccy = s->del2[i] * s->fy[i] - 0.5 * syi - s->y2[i];
    const double ccz = s->del2[i] * s->fz[i] - 0.5 * szi - s->z2[i];
    s->x0[i] = s->x0[i] + ccx * ct0;
    s->y0[i] = s->y0[i] + ccy * ct0;
    s->z0[i] = s->z0[i] + ccz * ct0;
    s->x1[i] = s->x1[i] + ccx * ct1;
    s->y1[i] = s->y1[i] + ccy * ct1;
    s->z1[i] = s->z1[i] + ccz * ct1;
    s->x2[i] = s->x2[i] + ccx * ct2;
    s->y2[i] = s->y2[i] + ccy * ct2;
    s->z2[i] = s->z2[i] + ccz * ct2;
    s->x3[i] = s->x3[i] + ccx * ct3;
    s->y3[i] = s->y3[i] + ccy * ct3;
    s->z3[i] = s->z3[i] + ccz * ct3;
  }
}

void orc_set_integrator(orc_sys *s, int iverlet) { s->iverlet = iverlet; }

/* scalev, src/thermostat_m.F90:317-395 (Nose-Hoover, Euler): remove the mean
 * of x1 over all atoms (divided by float(nm), a REAL(4) conversion) from the
 * moving atoms, KE of the moving atoms, snhv -= (ke - rqke)/(nose nm), moving
 * atoms' x1 scaled by 1 + snhv.  snhf keeps its value when ke <= 0 (a local
 * of the reference that is then undefined; 0 here). */
void orc_scalev(orc_sys *s, double nose, double rqke) {
  if (!(nose > 0.0)) return;
  const int n = s->n;
  double sx = 0.0, sy = 0.0, sz = 0.0, s3 = 0.0, snhf = 0.0;
  for (int i = 0; i < n; i++) {
    sx = sx + s->x1[i];
    sy = sy + s->y1[i];
    sz = sz + s->z1[i];
  }
  sx = sx / (double)(float)n;
  sy = sy / (double)(float)n;
  sz = sz / (double)(float)n;
  const double *b = s->b0;
  for (int i = 0; i < n; i++) {
    if (s->zat[i] == 0 || s->mass[i] < (double)0.1f) continue;
    s->x1[i] = s->x1[i] - sx;
    s->y1[i] = s->y1[i] - sy;
    s->z1[i] = s->z1[i] - sz;
    const double rx = T(b, 1, 1) * s->x1[i] + T(b, 1, 2) * s->y1[i] + T(b, 1, 3) * s->z1[i];
    const double ry = T(b, 2, 1) * s->x1[i] + T(b, 2, 2) * s->y1[i] + T(b, 2, 3) * s->z1[i];
    const double rz = T(b, 3, 1) * s->x1[i] + T(b, 3, 2) * s->y1[i] + T(b, 3, 3) * s->z1[i];
    s3 = s3 + s->mass[i] * (rx * rx + ry * ry + rz * rz);
  }
  const double ke = s3 / (s->deltat * s->deltat) / 2.0;
  if (ke > 0.0) snhf = (ke - rqke) / (nose * (double)n);
  s->snhv = s->snhv - snhf;
  for (int i = 0; i < n; i++) {
    if (s->zat[i] == 0 || s->mass[i] < (double)0.1f) continue;
    s->x1[i] = s->x1[i] * (1.0 + s->snhv);
    s->y1[i] = s->y1[i] * (1.0 + s->snhv);
    s->z1[i] = s->z1[i] * (1.0 + s->snhv);
  }
}

/* one MD step, src/moldin.F90:752-791: iverlet=1 velocityverlet, iverlet=0
 * PREDIC, FORCE, CORREC; then update_therm_avs and the energies */
int orc_md_step(orc_sys *s) {
  int st;
  if (s->iverlet == 0) {
    orc_predic(s);
    st = orc_force(s);
    if (st) return st;
    orc_correc(s);
  } else {
    st = velocity_verlet(s);
  }
  if (st) return st;
  update_therm_avs(s);
  s->ke = s->s3 / (s->deltat * s->deltat) / 2.0; /* deltat**2 */
  double bk = 0.0;
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++) bk = bk + (T(s->b1, i, j) * T(s->b1, i, j)) / s->bdel2;
  s->boxke = bk;
  s->te = s->ke + s->pe;
  s->h = s->te + s->press * s->vol;
  s->th = s->h + s->boxke;
  if (s->ntc == 0) orc_relink(s);
  return 0;
}

/* energy_calc, src/dynamics_m.F90:410-470 (serial order; vee_src even in lookup mode) */
void orc_energy_calc(orc_sys *s, double *en) {
  for (int i = 0; i < s->n; i++) en[i] = s->afrho[i];
  for (int i = 0; i < s->n; i++) {
    for (int k = 0; k < s->numn[i]; k++) {
      int j = s->nlist[(size_t)i * s->nnbrs + k] - 1;
      double dx = s->x0[i] - s->x0[j], dy = s->y0[i] - s->y0[j], dz = s->z0[i] - s->z0[j];
      double r = realsep(s->b0, s->ivol, &dx, &dy, &dz);
      double pp = orc_vee_src(&s->pot, r);
      en[i] = en[i] + pp / 2.0;
      en[j] = en[j] + pp / 2.0;
    }
  }
}

/* ======================================================================
 * Initial state as moldin builds it (random_m, read_system, setvel)
 * ====================================================================== */

/* rstart, src/random_m.F90:110-143 (the 24-bit constants are exact in REAL(4)) */
void orc_rstart(orc_rng *g, int iw, int ix, int iy, int iz) {
  int i = iw, j = ix, k = iy, l = iz;
  for (int ii = 0; ii < 97; ii++) {
    double sum = 0.0, t = 0.5;
    for (int jj = 0; jj < 24; jj++) {
      int m = ((i * j) % 179) * k % 179;
      i = j;
      j = k;
      k = m;
      l = (53 * l + 1) % 169;
      if ((l * m) % 64 >= 32) sum = sum + t;
      t = 0.5 * t;
    }
    g->u[ii] = sum;
  }
  g->c = (double)(362436.0f / 16777216.0f);
  g->cd = (double)(7654321.0f / 16777216.0f);
  g->cm = (double)(16777213.0f / 16777216.0f);
  g->iu = 97;
  g->ju = 33;
}

/* rand1, src/random_m.F90:73-95 */
double orc_rand1(orc_rng *g) {
  double uni = g->u[g->iu - 1] - g->u[g->ju - 1];
  if (uni < 0.0) uni = uni + 1.0;
  g->u[g->iu - 1] = uni;
  if (--g->iu == 0) g->iu = 97;
  if (--g->ju == 0) g->ju = 97;
  g->c = g->c - g->cd;
  if (g->c < 0.0) g->c = g->c + g->cm;
  uni = uni - g->c;
  if (uni < 0.0) uni = uni + 1.0;
  return uni;
}

/* read_system's replication loop, src/io_m.F90:216-245 */
void orc_replicate(orc_rng *g, int nbasis, const double *xb, const double *yb, const double *zb,
                   const int *nt, const double *b0, double dsp, double *x0, double *y0, double *z0) {
  const double rt[3] = {1.0 / nt[0], 1.0 / nt[1], 1.0 / nt[2]};
  long l = 0;
  for (int i = 0; i < nt[0]; i++)
    for (int j = 0; j < nt[1]; j++)
      for (int k = 0; k < nt[2]; k++)
        for (int ib = 0; ib < nbasis; ib++, l++) {
          x0[l] = (orc_rand1(g) - 0.5) * dsp / T(b0, 1, 1) + rt[0] * ((double)i + xb[ib]);
          y0[l] = (orc_rand1(g) - 0.5) * dsp / T(b0, 2, 2) + rt[1] * ((double)j + yb[ib]);
          z0[l] = (orc_rand1(g) - 0.5) * dsp / T(b0, 3, 3) + rt[2] * ((double)k + zb[ib]);
        }
}

/* bk = boltzmann/electron, src/constants_m.F90:93 */
double orc_bk(void) { return k_boltzmann / k_electron; }

/* RQKE of init_thermostat_m, src/thermostat_m.F90:97 */
double orc_rqke(double temprq, int nm) { return 1.5 * orc_bk() * temprq * (double)(nm - 1); }

/* setvel, src/thermostat_m.F90:191-255 (the atom part; the routine returns
 * before the box part for ivol < 1 and ivol == 4) */
double orc_setvel(orc_rng *g, int n, const int *z, const double *mass, const double *b0,
                  double deltat, double rqke, double *x1, double *y1, double *z1) {
  int nmove = 0;
  for (int i = 0; i < n; i++) {
    if (z[i] == 0 || mass[i] < (double)0.1f) continue;
    x1[i] = orc_rand1(g) - 0.5;
    y1[i] = orc_rand1(g) - 0.5;
    z1[i] = orc_rand1(g) - 0.5;
    nmove++;
  }
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int i = 0; i < n; i++) {
    sx = sx + x1[i] * mass[i];
    sy = sy + y1[i] * mass[i];
    sz = sz + z1[i] * mass[i];
  }
  sx = sx / nmove;
  sy = sy / nmove;
  sz = sz / nmove;
  double s3 = 0.0;
  for (int i = 0; i < n; i++) {
    if (z[i] == 0 || mass[i] < (double)0.1f) continue;
    x1[i] = x1[i] - sx / mass[i];
    y1[i] = y1[i] - sy / mass[i];
    z1[i] = z1[i] - sz / mass[i];
    double rx = T(b0, 1, 1) * x1[i] + T(b0, 1, 2) * y1[i] + T(b0, 1, 3) * z1[i];
    double ry = T(b0, 2, 1) * x1[i] + T(b0, 2, 2) * y1[i] + T(b0, 2, 3) * z1[i];
    double rz = T(b0, 3, 1) * x1[i] + T(b0, 3, 2) * y1[i] + T(b0, 3, 3) * z1[i];
    s3 = s3 + mass[i] * (rx * rx + ry * ry + rz * rz);
  }
  double ke = s3 / (deltat * deltat) / 2.0;
  /* float(nmove)/float(nm) is a REAL(4) quotient and sqrt of it a REAL(4) sqrt */
  double scale = sqrt(rqke / ke) * (double)sqrtf((float)nmove / (float)n);
  for (int i = 0; i < n; i++) {
    x1[i] = x1[i] * scale;
    y1[i] = y1[i] * scale;
    z1[i] = z1[i] * scale;
  }
  return rqke;
}

/* setvel's box part (src/thermostat_m.F90:256-299), reached for ivol 1, 2, 3
   and 5 only (ivol < 1 and ivol 4 return before it): b1(i,j) = rand1 - 0.5
   drawn row by row, each row less its sum / nm, then every element scaled by
   sqrt(temprq / boxtem) with boxtem = (sum b1(i,j)^2 / bdel2) / (4.5 bk).
   b1 is column-major; it is left untouched when the box part is skipped. */
void orc_setvel_box(orc_rng *g, int ivol, int nm, double temprq, double bdel2, double *b1) {
  if (ivol < 1 || ivol == 4) return;
  double s3 = 0.0;
  for (int i = 1; i <= 3; i++) {
    double s = 0.0;
    for (int j = 1; j <= 3; j++) {
      T(b1, i, j) = orc_rand1(g) - 0.5;
      s = s + T(b1, i, j);
    }
    s = s / nm;
    for (int j = 1; j <= 3; j++) {
      T(b1, i, j) = T(b1, i, j) - s;
      s3 = s3 + T(b1, i, j) * T(b1, i, j);
    }
  }
  const double boxtem = (s3 / bdel2) / (4.5 * orc_bk());
  const double scale = sqrt(temprq / boxtem);
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++) T(b1, i, j) = T(b1, i, j) * scale;
}

/* ---------------- accessors ---------------- */
static double *arr_by_name(const orc_sys *s, const char *nm) {
  static const char *names[] = {"x0", "y0", "z0", "x1", "y1", "z1", "x2", "y2", "z2",
                                "rho", "afrho", "dafrho", "fx", "fy", "fz", "del2", "mass",
                                "x3", "y3", "z3"};
  double *ptrs[] = {s->x0, s->y0, s->z0, s->x1, s->y1, s->z1, s->x2, s->y2, s->z2,
                    s->rho, s->afrho, s->dafrho, s->fx, s->fy, s->fz, s->del2, s->mass,
                    s->x3, s->y3, s->z3};
  for (size_t k = 0; k < sizeof(names) / sizeof(names[0]); k++)
    if (!strcmp(nm, names[k])) return ptrs[k];
  return NULL;
}
int orc_get_array(const orc_sys *s, const char *name, double *out) {
  double *a = arr_by_name(s, name);
  if (!a) return -1;
  memcpy(out, a, sizeof(double) * (size_t)s->n);
  return 0;
}
int orc_set_array(orc_sys *s, const char *name, const double *in) {
  double *a = arr_by_name(s, name);
  if (!a) return -1;
  memcpy(a, in, sizeof(double) * (size_t)s->n);
  return 0;
}
static double *ten_by_name(const orc_sys *s, const char *nm) {
  orc_sys *m = (orc_sys *)s;
  if (!strcmp(nm, "b0")) return m->b0;
  if (!strcmp(nm, "b1")) return m->b1;
  if (!strcmp(nm, "b2")) return m->b2;
  if (!strcmp(nm, "b3")) return m->b3;
  if (!strcmp(nm, "tk")) return m->tk;
  if (!strcmp(nm, "tc")) return m->tc;
  if (!strcmp(nm, "tg")) return m->tg;
  if (!strcmp(nm, "tp")) return m->tp;
  if (!strcmp(nm, "tginv")) return m->tginv;
  if (!strcmp(nm, "tgdot")) return m->tgdot;
  if (!strcmp(nm, "tgid")) return m->tgid;
  if (!strcmp(nm, "fb")) return m->fb;
  return NULL;
}
int orc_get_tensor(const orc_sys *s, const char *name, double *out) {
  double *t = ten_by_name(s, name);
  if (!t) return -1;
  memcpy(out, t, 9 * sizeof(double));
  return 0;
}
int orc_set_tensor(orc_sys *s, const char *name, const double *in) {
  double *t = ten_by_name(s, name);
  if (!t) return -1;
  memcpy(t, in, 9 * sizeof(double));
  if (!strcmp(name, "b0")) pr_volume(s);
  return 0;
}
double orc_get_scalar(const orc_sys *s, int which) {
  switch (which) {
    case 0: return s->pe;
    case 1: return s->ke;
    case 2: return s->vol;
    case 3: return s->s1;
    case 4: return s->s2;
    case 5: return s->s3;
    case 6: return s->te;
    case 7: return s->h;
    case 8: return s->th;
    case 9: return s->boxke;
    case 10: return s->deltat;
    case 11: return s->bdel2;
    case 12: return s->press;
    case 13: return s->rnear;
    default: return NAN;
  }
}
int orc_get_int(const orc_sys *s, int which) {
  switch (which) {
    case 0: return s->n;
    case 1: return s->nlcx;
    case 2: return s->nlcy;
    case 3: return s->nlcz;
    case 4: return s->ntc;
    case 5: return s->ntcm;
    case 6: return s->nnbrs;
    case 7: {
      long t = 0;
      for (int i = 0; i < s->n; i++) t += s->numn[i];
      return (int)t;
    }
    default: return -1;
  }
}
void orc_set_ntc(orc_sys *s, int ntc) { s->ntc = ntc; }

long orc_get_pairs(const orc_sys *s, int *pi, int *pj, long cap) {
  long c = 0;
  for (int i = 0; i < s->n; i++)
    for (int k = 0; k < s->numn[i]; k++) {
      if (c < cap) {
        pi[c] = i;
        pj[c] = s->nlist[(size_t)i * s->nnbrs + k] - 1;
      }
      c++;
    }
  return c;
}
int orc_get_numn(const orc_sys *s, int *numn) {
  memcpy(numn, s->numn, sizeof(int) * (size_t)s->n);
  return 0;
}
int orc_get_ic(const orc_sys *s, int *ic) {
  for (int i = 0; i < s->n; i++) ic[i] = s->ic[i + 1];
  return 0;
}
/* L0 (ncell starters) and LINK (n closed chains), 1-based, as linkup and
 * relink leave them */
int orc_get_chains(const orc_sys *s, int *l0, int *link) {
  const int ncell = s->nlcx * s->nlcy * s->nlcz;
  for (int c = 0; c < ncell; c++) l0[c] = s->l0[c];
  for (int i = 0; i < s->n; i++) link[i] = s->link[i + 1];
  return 0;
}

long orc_bruteforce_pairs(int n, const double *b0, const double *x, const double *y,
                          const double *z, int ivol, double rcut, int *pi, int *pj,
                          long cap) {
  long c = 0;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
      double r = realsep(b0, ivol, &dx, &dy, &dz);
      if (r < rcut) {
        if (c < cap) { pi[c] = i; pj[c] = j; }
        c++;
      }
    }
  return c;
}

/* Brute-force single-cell evaluation with the reference's per-pair arithmetic
 * (src/dynamics_m.F90:235-346, src/potential_m.F90:148-169); used to check the
 * batched engine on cells too small for the reference's 6.3 A link cells. */
int orc_eval_cell(const orc_pot *p, int n, const double *b0, const double *s, double rcut,
                  double *pe_out, double *fcart, double *rho, double *w9) {
  double *x = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *fx = (double *)calloc((size_t)3 * n, sizeof(double));
  double *df = (double *)calloc((size_t)n, sizeof(double));
  if (!x || !fx || !df) { free(x); free(fx); free(df); return -1; }
  for (int i = 0; i < n; i++) { x[i] = s[3 * i]; x[n + i] = s[3 * i + 1]; x[2 * n + i] = s[3 * i + 2]; }
  for (int i = 0; i < n; i++) rho[i] = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double dx = x[i] - x[j], dy = x[n + i] - x[n + j], dz = x[2 * n + i] - x[2 * n + j];
      double r = realsep(b0, 1, &dx, &dy, &dz);
      if (r >= rcut) continue;
      rho[j] = rho[j] + orc_phi(p, r);
      rho[i] = rho[i] + orc_phi(p, r);
    }
  double pe = 0.0;
  for (int i = 0; i < n; i++) {
    if (rho[i] <= 1e-40) continue;
    pe = pe + orc_emb(p, rho[i]);
    df[i] = orc_demb(p, rho[i]);
  }
  for (int k = 0; k < 9; k++) w9[k] = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double dx = x[i] - x[j], dy = x[n + i] - x[n + j], dz = x[2 * n + i] - x[2 * n + j];
      double r = realsep(b0, 1, &dx, &dy, &dz);
      if (r >= rcut) continue;
      double rr = 1.0 / r;
      pe = pe + orc_vee(p, r);
      double fp = -orc_dvee(p, r) * rr + (-1.0 * (orc_dphi(p, r) * df[i] + orc_dphi(p, r) * df[j]) * rr);
      double rv[3];
      rv[0] = T(b0, 1, 1) * dx + T(b0, 1, 2) * dy + T(b0, 1, 3) * dz;
      rv[1] = T(b0, 2, 1) * dx + T(b0, 2, 2) * dy + T(b0, 2, 3) * dz;
      rv[2] = T(b0, 3, 1) * dx + T(b0, 3, 2) * dy + T(b0, 3, 3) * dz;
      for (int a = 0; a < 3; a++) {
        fx[3 * i + a] += rv[a] * fp;
        fx[3 * j + a] -= rv[a] * fp;
        for (int c = 0; c < 3; c++) w9[c * 3 + a] += rv[a] * rv[c] * fp;
      }
    }
  for (int k = 0; k < 3 * n; k++) fcart[k] = fx[k];
  *pe_out = pe;
  free(x); free(fx); free(df);
  return 0;
}
