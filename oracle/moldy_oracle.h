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
 *   Nose scalev     src/thermostat_m.F90:317-395
 *   predictor-corrector src/quench_m.F90:636-759
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

/* ---------------- potential (Fe-Fe) ---------------- */
/* Material module, chosen at compile time in the reference (src/potential_m.F90:49-87) */
#define ORC_MAT_FETRANSMETAL 0
#define ORC_MAT_IRON_CARBON 1

typedef struct orc_pot {
  int material;               /* ORC_MAT_* */
  /* iron_carbon Fe-Fe data (src/iron_carbon.F90:42-67): 13 pair knots, the
   * cubic in the join's exponent, 3 density knots, embedding cludge */
  double vcut[13], vpar[13], join[4], pcut[3], ppar[3], acludge, rcludge, rx;
  int nv, np;                 /* number of pair / density knots (fetransmetal) */
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
/* iron_carbon material (src/iron_carbon.F90).  iron_in == NULL keeps the
 * module's built-in coefficients; otherwise 38 values in the order
 * check_supported_atomic_numbers reads "iron.in" (:788-795): vcut1..13,
 * vpar1..13, join1..4, pcut1..3, ppar1..3, acludge, rcludge. */
void orc_pot_init_iron_carbon(orc_pot *p, const double *iron_in);
/* get_supported_potential_range's rmax: 6 (fetransmetal), 5.4 (iron_carbon) */
double orc_pot_range(const orc_pot *p);
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

/* ---------------- initial state (test inputs, as moldin makes them) ---------------- */
/* RANMAR generator of src/random_m.F90:73-143 (rand1 / rstart) */
typedef struct orc_rng {
  double u[97], c, cd, cm;
  int iu, ju;
} orc_rng;
void orc_rstart(orc_rng *g, int iw, int ix, int iy, int iz);
double orc_rand1(orc_rng *g);
/* read_system's replication loop (src/io_m.F90:216-245): nbasis basis atoms
 * (fractional in the unit volume) repeated nt[0] x nt[1] x nt[2] times, each
 * coordinate displaced by (rand1()-0.5)*dsp/b0(k,k) with b0 the whole box;
 * writes nbasis*nt[0]*nt[1]*nt[2] positions in the reference's atom order. */
void orc_replicate(orc_rng *g, int nbasis, const double *xb, const double *yb, const double *zb,
                   const int *nt, const double *b0, double dsp, double *x0, double *y0, double *z0);
/* setvel (src/thermostat_m.F90:191-306) for ivol < 1 or ivol == 4 (atom
 * velocities only; x1 = v dt / L).  deltat in model units; rqke =
 * 1.5 bk temprq (nm - 1) as init_thermostat_m sets it.  Returns ke. */
double orc_setvel(orc_rng *g, int n, const int *z, const double *mass, const double *b0,
                  double deltat, double rqke, double *x1, double *y1, double *z1);
void orc_setvel_box(orc_rng *g, int ivol, int nm, double temprq, double bdel2, double *b1);
/* 1.5 * bk * temprq * (nm - 1) with bk of src/constants_m.F90 */
double orc_rqke(double temprq, int nm);
/* Boltzmann constant in eV/K (src/constants_m.F90) */
double orc_bk(void);

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
/* energy_calc (src/dynamics_m.F90:420-470): en_atom = F(rho_i) + half of each
 * pair energy, over the current half list; needs rho/afrho from force() */
void orc_energy_calc(orc_sys *s, double *en_atom);
void orc_kinten(orc_sys *s);
/* moldin.F90:446-477: initial force, tgid and accelerations */
int  orc_md_init(orc_sys *s);
/* one MD step (moldin.F90:752-791); ke returned via getters */
int  orc_md_step(orc_sys *s);
/* integrator of orc_md_step: 1 velocity Verlet (default), 0 Gear
 * predictor-corrector (PREDIC, FORCE, CORREC; src/quench_m.F90:636-759); predic
 * changes nothing until a correc or velocityverlet has run, as in the reference */
void orc_set_integrator(orc_sys *s, int iverlet);
void orc_predic(orc_sys *s);
/* scalev (src/thermostat_m.F90:317-395) with the Nose parameter and required
 * kinetic energy; the snhv accumulator persists in the system */
void orc_scalev(orc_sys *s, double nose, double rqke);
void orc_correc(orc_sys *s);

/* getters / setters; name selects the array:
 * "x0","y0","z0","x1","y1","z1","x2","y2","z2","rho","afrho","dafrho",
 * "fx","fy","fz","del2","mass","x3","y3","z3" (n doubles) */
int  orc_get_array(const orc_sys *s, const char *name, double *out);
int  orc_set_array(orc_sys *s, const char *name, const double *in);
/* "b0","b1","b2","b3","tk","tc","tg","tp","tginv","tgdot","tgid","fb" (9 doubles) */
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
int  orc_get_chains(const orc_sys *s, int *l0, int *link);

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
