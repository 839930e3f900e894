/*
 * moldy_oracle.h -- CPU restatement of MOLDY's force hot path (test infrastructure).
 *
 * THIS IS THE ORACLE. It exists only to check the HIP path: tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it;
 * the product library (libmoldy_hip.so) never links or calls it.
 *
 * It restates, in the reference's exact floating-point operation order
 * (serial, no FMA contraction), the Fe-Fe Finnis-Sinclair path of gjackland/moldy:
 *   potential       src/fetransmetal.F90:100-404, 414-551, 768-926
 *   lookup tables   src/lookup_m.F90:161-251, 307-394
 *   link cells      src/linkcell_m.F90:122-274
 *   neighbour list  src/neighbourlist_m.F90:103-276
 *   min. image      src/parrinellorahman_m.F90:149-173, 223-227, 254-261
 *   rhoset/embed    src/dynamics_m.F90:497-579, src/potential_m.F90:148-169
 *   force + virial  src/dynamics_m.F90:135-401
 *   kinten          src/thermostat_m.F90:127-177
 *   velocity Verlet src/quench_m.F90:562-624
 *   therm. averages src/dynamics_m.F90:649-734
 *   MD-step glue    src/moldin.F90:446-477, 752-791
 *
 * 3x3 tensors are stored column-major exactly like the Fortran arrays:
 * t[(j-1)*3 + (i-1)] == T(i,j).  b0(:,j) is cell vector j.
 */
#ifndef MOLDY_ORACLE_H
#define MOLDY_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- potential (Fe-Fe, fetransmetal) ---------------- */
typedef struct orc_pot {
  int nv, np;                 /* number of pair / density knots */
  double av[8], rv[8];        /* pair amplitudes and knot radii */
  double ap[4], rp[4];        /* density amplitudes and knot radii */
  double bier_cut1, bier_cut2;
  double sc[4];               /* spline_coeff(1:4,Fe,Fe) */
  double phi_cut_r, phi_cut_v;/* phi_cut(1:2,Fe,Fe) */
  int use_lookup;
  int nlookup;
  double *tab[6];             /* vee, dvee, phi, dphi, emb, demb tables (0..nlookup) */
  double r_min, r_max, vp_range, vp_reciprange, vp_recipnlkp, vp_recipstep;
  double rho_min, rho_max, e_range, e_reciprange, e_recipnlkp, e_recipstep;
} orc_pot;

/* Sets coefficients and runs the set_up_m sequence (spline_coeff, phi_turn). */
void orc_pot_init(orc_pot *p, int nv, const double *av, const double *rv,
                  int np, const double *ap, const double *rp);
/* potential_set_lookup with the default 5000-point tables (call after orc_pot_init). */
int  orc_pot_build_lookup(orc_pot *p, int nlookup);
void orc_pot_free(orc_pot *p);

double orc_vee(const orc_pot *p, double r);
double orc_dvee(const orc_pot *p, double r);
double orc_phi(const orc_pot *p, double r);
double orc_dphi(const orc_pot *p, double r);
double orc_emb(const orc_pot *p, double rho);
double orc_demb(const orc_pot *p, double rho);
/* exact-mode functions regardless of use_lookup */
double orc_vee_src(const orc_pot *p, double r);
double orc_dvee_src(const orc_pot *p, double r);
double orc_phi_src(const orc_pot *p, double r);
double orc_dphi_src(const orc_pot *p, double r);

/* ---------------- system ---------------- */
typedef struct orc_sys orc_sys;

/* Creates a system.  deltat in fs, press in GPa, boxmass in amu per atom.
 * Arrays are copied.  z = atomic numbers (26 or 0 for vacancy). */
orc_sys *orc_create(int n, const double *b0, const double *x0, const double *y0,
                    const double *z0, const double *x1, const double *y1,
                    const double *z1, const double *mass, const int *z,
                    int ivol, double deltat_fs, double rpad, double rcut,
                    int nnbrs, int ntcm, double press_gpa, double boxmass,
                    const orc_pot *pot);
void orc_destroy(orc_sys *s);

int  orc_linkup(orc_sys *s);
int  orc_relink(orc_sys *s);
/* returns 0, or 1 on 'NEIGHBOURLIST: EXCESS NEIGHBOURS!' */
int  orc_update_neighbourlist(orc_sys *s);
void orc_rhoset(orc_sys *s);
void orc_embed(orc_sys *s);
/* full force() including the ntc rebuild logic; returns neighbour-list status */
int  orc_force(orc_sys *s);
/* force() pieces without list maintenance (list must exist) */
void orc_force_pairs(orc_sys *s);
void orc_kinten(orc_sys *s);
/* moldin.F90:446-477: initial force, tgid and accelerations */
int  orc_md_init(orc_sys *s);
/* one MD step (moldin.F90:752-791, iverlet=1); ke returned via getters */
int  orc_md_step(orc_sys *s);

/* getters / setters; name selects the array:
 * "x0","y0","z0","x1","y1","z1","x2","y2","z2","rho","afrho","dafrho",
 * "fx","fy","fz","del2","mass" (n doubles) */
int  orc_get_array(const orc_sys *s, const char *name, double *out);
int  orc_set_array(orc_sys *s, const char *name, const double *in);
/* "b0","b1","b2","tk","tc","tg","tp","tginv","tgdot","tgid","fb" (9 doubles) */
int  orc_get_tensor(const orc_sys *s, const char *name, double *out);
int  orc_set_tensor(orc_sys *s, const char *name, const double *in);
/* scalars: 0 pe, 1 ke, 2 vol, 3 s1 (sum f^2), 4 s2, 5 s3, 6 te, 7 h, 8 th, 9 boxke,
 *          10 deltat (model units), 11 bdel2, 12 press (model), 13 rnear */
double orc_get_scalar(const orc_sys *s, int which);
/* integers: 0 n, 1 nlcx, 2 nlcy, 3 nlcz, 4 ntc, 5 ntcm, 6 nnbrs, 7 total pairs */
int  orc_get_int(const orc_sys *s, int which);
void orc_set_ntc(orc_sys *s, int ntc);

/* Half list: numn[n]; pairs written as (i,j) 0-based in list order, at most cap. */
long orc_get_pairs(const orc_sys *s, int *pi, int *pj, long cap);
int  orc_get_numn(const orc_sys *s, int *numn);
/* link cell of each atom (1-based cell index like ic) */
int  orc_get_ic(const orc_sys *s, int *ic);

/* Brute-force O(N^2) minimum-image pair enumeration, r < rcut, i<j, 0-based.
 * Independent of link cells: used to pin the cell algorithm. */
long orc_bruteforce_pairs(int n, const double *b0, const double *x, const double *y,
                          const double *z, int ivol, double rcut,
                          int *pi, int *pj, long cap);

/* One periodic cell evaluated like force() (src/dynamics_m.F90:235-346) over a
 * brute-force minimum-image half list at cutoff rcut: energy (pe), Cartesian
 * forces B0.f (3n, row-major), rho (n) and W = sum over pairs r r^T fp (9,
 * column-major) -- the batched engine's outputs.  ivol selects the periodicity. */
int orc_eval_cell(const orc_pot *p, int n, const double *b0, const double *s, double rcut,
                  double *pe, double *fcart, double *rho, double *w9);

/* Set OpenMP-free single-thread evaluation (default) or threaded pair loops
 * for the CPU baseline; threads<=1 keeps the reference's serial order. */
void orc_set_threads(int threads);

#ifdef __cplusplus
}
#endif
#endif
