/*
 * moldy_hip.h -- C ABI of libmoldy_hip.so, the MI355X (gfx950) implementation
 * of MOLDY's force hot path (link cells, neighbour list, Finnis-Sinclair
 * density/embedding/force+virial, velocity Verlet + Parrinello-Rahman).
 *
 * Every entry point returns an int status (MDK_OK == 0).  No C++ types, no
 * torch types and no exceptions cross this boundary: plain pointers, sizes
 * and an opaque handle.  Host arrays are always in the reference's atom order
 * (index i of X0(i) in src/particles_m.F90:58) and are copied; the handle
 * owns device-resident mirrors (atoms are kept sorted by cell on the device).
 * One handle is used by one host thread and owns one HIP stream.
 *
 * Tensors (3x3) are Fortran column-major like the reference's B0(3,3):
 * t[(j-1)*3 + (i-1)] == T(i,j); column j of B0 is cell vector j.
 *
 * Reference interfaces replaced (gjackland/moldy):
 *   src/linkcell_m.F90:52-54        linkup, relink, get_l0, get_link
 *   src/neighbourlist_m.F90:52-58   update_neighbourlist, NLIST, NUMN
 *   src/potential_m.F90:100-118     set_up, potential_set_lookup/exact, embed
 *   src/dynamics_m.F90:69-83        force, rhoset, update_therm_avs, fx/fy/fz
 *   src/thermostat_m.F90:127-177    kinten (inside force)
 *   src/parrinellorahman_m.F90:91-115  B0..B2, TK, TC, TG, TP, TGINV, TGDOT, TGID, FB, vol
 *   src/quench_m.F90:59-63          velocityverlet
 *   src/moldin.F90:446-477,752-791  initial accelerations; MD-step glue
 */
#ifndef MOLDY_HIP_H
#define MOLDY_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MDK_OK 0
#define MDK_ERR_HIP 1           /* a HIP runtime call failed (message via mdk_last_error) */
#define MDK_ERR_NBR_OVERFLOW 2  /* 'NEIGHBOURLIST: EXCESS NEIGHBOURS!' (src/neighbourlist_m.F90:222-230) */
#define MDK_ERR_BAD_ARG 3       /* invalid argument / size */
#define MDK_ERR_SINGULAR_BOX 4  /* matrix_invert on a zero determinant (src/matrix_m.F90:110-113) */
#define MDK_ERR_STATE 5         /* call out of order (e.g. force before set_potential) */
#define MDK_ERR_SKIN 6          /* a step used a device list that may miss pairs: an atom moved d and
                                   the box strained e since the list was built with 2d + e*r_list >=
                                   skin (reported by mdk_md_steps / mdk_velocityverlet) */
#define MDK_ERR_NO_DEVICE 7     /* no GPU visible */

/* ---- precision forms ---- */
#define MDK_PREC_FP64 0         /* plain FP64 (parity form) */
#define MDK_PREC_FP32 1         /* FP32 pair/spline arithmetic, FP64 accumulation/integration */

/* ---- per-atom fields for mdk_upload / mdk_download (n doubles each) ---- */
enum {
  MDK_X0 = 0, MDK_Y0, MDK_Z0,   /* fractional positions */
  MDK_X1, MDK_Y1, MDK_Z1,       /* fractional velocity * deltat */
  MDK_X2, MDK_Y2, MDK_Z2,       /* fractional acceleration * deltat^2/2 */
  MDK_FX, MDK_FY, MDK_FZ,       /* forces f (Cartesian force = B0 . f) */
  MDK_RHO, MDK_AFRHO, MDK_DAFRHO,
  MDK_MASS, MDK_DEL2,
  MDK_X3, MDK_Y3, MDK_Z3,       /* third scaled derivative (predictor-corrector) */
  MDK_NFIELDS
};
/* ---- 3x3 tensors for mdk_get_tensor / mdk_set_tensor ---- */
enum {
  MDK_B0 = 0, MDK_B1, MDK_B2, MDK_TK, MDK_TC, MDK_TG, MDK_TP, MDK_TGINV, MDK_TGDOT,
  MDK_TGID, MDK_FB, MDK_B3, MDK_NTENSORS
};
/* ---- scalars for mdk_get_scalars (array of MDK_NSCALARS doubles) ---- */
enum {
  MDK_S_PE = 0, MDK_S_KE, MDK_S_VOL, MDK_S_S1, MDK_S_S2, MDK_S_S3, MDK_S_TE, MDK_S_H,
  MDK_S_TH, MDK_S_BOXKE, MDK_S_MAXDISP, MDK_NSCALARS
};

/* Simulation parameters: the subset of type(simparameters)
 * (src/params_m.F90:44-162) that the hot path reads, already in model units
 * (deltat in 10.1805 fs units, press in eV/A^3, bdel2 = deltat^2/(2*bmass*nm)). */
typedef struct mdk_params {
  int nm;          /* number of atoms */
  int ivol;        /* 0 NPT (PR), 1 NVE, 2 free surface z, 4 z-only box, 5 pillar */
  int nnbrs;       /* capacity of the reference half list per atom (NLIST) */
  int ntcm;        /* neighbour-list update interval (nbrupdate) */
  int nlcx, nlcy, nlcz; /* reference link-cell grid (moldin.F90:252-257) */
  int precision;   /* MDK_PREC_FP64 or MDK_PREC_FP32 */
  int rebuild;     /* device list rebuilds: 0 on the reference's ntcm interval and whenever the
                      skin bound could be reached; 1 only when the skin bound could be reached */
  int max_full;    /* capacity of the device full list per atom (0 = default 64) */
  double deltat;   /* model time units */
  double rnear;    /* rcut + rpad: reference list radius */
  double rskin;    /* skin of the device full list (default rpad) */
  double press;    /* external pressure, model units */
  double bdel2;    /* box time scale */
} mdk_params;

/* Handle lifecycle.  Replaces init_linkcell_m/init_neighbourlist_m/init_dynamics_m. */
int mdk_create(void **handle, const mdk_params *p);
int mdk_destroy(void *handle);
const char *mdk_last_error(void);

/* Potential (src/fetransmetal.F90:603-954 set_up; src/potential_m.F90:181-232).
 * Knot coefficients of the Fe-Fe pair and density functions; set_up constants
 * (Biersack join, phi flat cap) are derived on the host exactly as set_up_m. */
int mdk_set_potential(void *h, int nv, const double *av, const double *rv, int np,
                      const double *ap, const double *rp);
/* Selects the iron_carbon material instead (src/iron_carbon.F90, built into
 * the reference with -DIRON_CARBON, src/potential_m.F90:49-51): its Fe-Fe pair
 * function (13-knot cubic sum above 2.05 A, exp-cubic join, screened Coulomb
 * core below 1 A), density function (3 knots) and embedding
 * -rho^1/2 - c2 rho^2 + c4 rho^4.  iron_in == NULL keeps the module's
 * coefficients; otherwise 38 values in the order check_supported_atomic_numbers
 * reads "iron.in" (src/iron_carbon.F90:785-795): vcut1..13, vpar1..13,
 * join1..4, pcut1..3, ppar1..3, acludge, rcludge.  FP64 form only (MDK_ERR_BAD_ARG
 * for a handle created with MDK_PREC_FP32); Fe atoms only (the carbon
 * functions are not built).  Replaces get_potential/check_supported_atomic_numbers
 * of that build (src/potential_m.F90:656-673, src/iron_carbon.F90:745-803). */
int mdk_set_iron_carbon(void *h, const double *iron_in);
/* potential_set_lookup: nlookup-point linear tables on [0, rmax] A (rmax 6 for
 * fetransmetal, 5.4 for iron_carbon) and rho in [0,1e4]; nlookup <= 0 selects
 * exact evaluation (potential_set_exact). */
int mdk_set_lookup(void *h, int nlookup);
/* out[0..7] = bier_cut1, bier_cut2, spline_coeff(1:4), phi_cut(1:2) (fetransmetal only) */
int mdk_potential_constants(void *h, double *out);
/* Evaluate potential_m's vee,dvee,phi,dphi (which=0..3) at r[k], or emb,demb
 * (which=4,5) at rho[k], on the device, k < count.  which=6..8: V, V' and phi'
 * as the FP32 force pass evaluates them (its interval table; exact-mode
 * fetransmetal with 6 pair and 2 density knots only, MDK_ERR_BAD_ARG
 * otherwise). */
int mdk_eval_potential(void *h, int which, const double *x, double *y, int count);

/* State transfer (ownership stays with the caller; arrays in reference order). */
int mdk_upload(void *h, int field, const double *data);
int mdk_download(void *h, int field, double *data);
/* atomic numbers: 26 (Fe) or 0 (vacancy); others give MDK_ERR_BAD_ARG */
int mdk_upload_species(void *h, const int *atomic_number);
int mdk_set_tensor(void *h, int which, const double *t9);
int mdk_get_tensor(void *h, int which, double *t9);
int mdk_get_scalars(void *h, double *out);
int mdk_set_ntc(void *h, int ntc);
int mdk_get_ntc(void *h, int *ntc);

/* Link cells: linkup (src/linkcell_m.F90:122-198), relink (:215-274).
 * get_l0/get_link return chains that visit the same cell membership;
 * l0 has nlcx*nlcy*nlcz entries, link and ic have nm entries (1-based). */
int mdk_linkup(void *h);
int mdk_relink(void *h);
int mdk_get_linkcells(void *h, int *l0, int *link, int *ic);

/* update_neighbourlist (src/neighbourlist_m.F90:103-276): the reference half
 * list at rnear over the 13 forward link cells, on the device.  Returns
 * MDK_ERR_NBR_OVERFLOW when some atom has more than nnbrs entries. */
int mdk_update_neighbourlist(void *h);
/* NUMN(nm) and NLIST(nnbrs,nm) (column-major, 1-based indices like the reference). */
int mdk_get_neighbourlist(void *h, int *numn, int *nlist);
/* Unordered half-list pair count and pairs (0-based (i,j)), cap entries max. */
int mdk_get_pairs(void *h, long *count, int *pi, int *pj, long cap);

/* Force pieces (src/dynamics_m.F90). mdk_force is the whole reference force
 * routine including its ntc/ntcm list bookkeeping; rhoset computes RHO only,
 * embed computes AFRHO/DAFRHO and the embedding energy. */
int mdk_force(void *h);
int mdk_rhoset(void *h);
int mdk_embed(void *h, double *pe_embed);
int mdk_kinten(void *h);
/* energy_calc (src/dynamics_m.F90:410-470): en_atom[i] = F(rho_i) + half of the
 * exact pair energy with each neighbour (the per-atom energies system.out
 * prints), in reference atom order.  Needs the list and densities of a
 * previous force evaluation (MDK_ERR_STATE otherwise). */
int mdk_energy_calc(void *h, double *en_atom);
/* Per-atom diagonal stress of force's atomstress option
 * (src/dynamics_m.F90:279-283, arrays of src/particles_m.F90:66): sx[i] =
 * sum of rx_ij fp_ij over atom i's own reference half-list partners (NLIST of
 * the last update_neighbourlist), at the current positions and the F' of the
 * last force evaluation, in FP64 and reference atom order.  MDK_ERR_STATE
 * before the first half list. */
int mdk_atomic_stress(void *h, double *sx, double *sy, double *sz);
/* moldin.F90:446-477: first force, tgid and initial accelerations x2, b2 */
int mdk_md_init(void *h);
/* scalev (src/thermostat_m.F90:317-395): Nose-Hoover scaling of the atom
 * velocities by 1 + snhv after snhv -= (ke - rqke)/(nose nm), with the mean
 * of x1 removed first; nothing for nose <= 0.  moldin calls it about once per
 * femtosecond of simulated time (src/moldin.F90:828-835). */
int mdk_scalev(void *h, double nose, double rqke);
/* get_snhv / set_snhv (src/thermostat_m.F90:53-85): the accumulated scaling */
int mdk_get_snhv(void *h, double *snhv);
int mdk_set_snhv(void *h, double snhv);
/* External pressure (model units) from now on: simparam%press as moldin
 * changes it between strain loops (src/moldin.F90:434). */
int mdk_set_pressure(void *h, double press);
/* Integrator of mdk_md_steps (moldin.F90:752-758): 1 velocity Verlet (default),
 * 0 Gear predictor-corrector (PREDIC, FORCE, CORREC). */
int mdk_set_integrator(void *h, int iverlet);
/* predic (src/quench_m.F90:636-671): x0..x2 and the box from the scaled
 * derivatives up to x3, b3.  Like the reference, it changes nothing until a
 * correc or velocityverlet has run on the handle: predic reads quench_m's
 * parameter copy, which only those refresh (:640, :569, :690). */
int mdk_predic(void *h);
/* correc (src/quench_m.F90:682-759): Gear corrector with the force of the
 * last force evaluation, the current tgid and bdel2 fb for the box. */
int mdk_correc(void *h);
/* velocityverlet(dt) (src/quench_m.F90:562-624), including force */
int mdk_velocityverlet(void *h);
/* update_therm_avs(s1,s2,s3) (src/dynamics_m.F90:649-734) */
int mdk_update_therm_avs(void *h, double *s1, double *s2, double *s3);
/* nsteps MD steps of the moldin loop (velocityverlet, update_therm_avs,
 * KE/TE/H/TH, relink when ntc==0) without host round trips; if trace is not
 * NULL it receives per step {pe, ke, te, h, th, vol} (6*nsteps doubles). */
int mdk_md_steps(void *h, int nsteps, double *trace);
/* Number of full-list rebuilds performed so far and the current full-list
 * pair count (sum over atoms, each pair counted twice). */
int mdk_stats(void *h, long *rebuilds, long *full_pairs);

/* Timing: per-kernel-class HIP-event accumulators on the handle's stream for
 * classes {0 drift, 1 density, 2 force, 3 kick, 4 rebuild}.  mask selects the
 * classes to record (bit k = class k; MDK_TIMING_ALL every class, 0 stops);
 * each recorded launch adds two events to the stream.  mdk_timing_read
 * returns total milliseconds and launch counts per class. */
#define MDK_TIMING_ALL 0x1F
int mdk_timing(void *h, int mask);
/* record only every `every`-th launch of each class (default 1): an event pair
 * idles the GPU for a few microseconds, so sampling bounds the overhead */
int mdk_timing_every(void *h, int every);
int mdk_timing_read(void *h, double *ms, long *launches, int nclasses);
int mdk_synchronize(void *h);

/* ---------------------------------------------------------------------------
 * Batched evaluation: energies, forces and the 3x3 virial of many independent
 * periodic cells of uneven size in ONE persistent launch (replayed as a
 * hipGraph).  Each cell is evaluated like force() in src/dynamics_m.F90:135-401
 * (densities, embedding, pair + embedding forces, virial) with the Fe-Fe
 * potential; the reference evaluates one cell per program run.
 * ------------------------------------------------------------------------- */
/* natoms[ncell]; boxes 9 doubles per cell (column-major, columns = cell
 * vectors); s = 3 fractional coordinates per atom, cells concatenated. */
int mdk_batch_create(void **handle, int ncell, const int *natoms, const double *boxes,
                     const double *s, int precision);
int mdk_batch_destroy(void *handle);
const char *mdk_batch_last_error(void);
int mdk_batch_set_potential(void *h, int nv, const double *av, const double *rv, int np,
                            const double *ap, const double *rp);
/* block in {64,128,256,512}, workgroups per CU, hipGraph replay on/off */
int mdk_batch_config(void *h, int block, int wg_per_cu, int use_graph);
int mdk_batch_evaluate(void *h);
int mdk_batch_synchronize(void *h);
/* energy[ncell] eV; virial[6*ncell] = sum over pairs r r^T fp (xx yy zz xy xz yz),
 * stress = -virial/volume; forces[3*natom] Cartesian eV/A; rho[natom]; any may be NULL */
int mdk_batch_results(void *h, double *energy, double *virial, double *forces, double *rho);
/* autotuner over block size and workgroups per CU (best of reps replays) */
int mdk_batch_tune(void *h, int reps, int *best_block, int *best_wgpc, double *best_ms);
int mdk_batch_time(void *h, int reps, double *ms_avg);
int mdk_batch_info(void *h, int *natom, int *block, int *wg_per_cu, long *pairs);

/* ---------------------------------------------------------------------------
 * Adjoint of the batched evaluation with respect to the knot parameters of the
 * Fe-Fe pair and density functions (src/fetransmetal.F90:133-139, 256-265;
 * coefficients of src/potentials/pot_026_026.in): gradient order a_0..a_{nv-1},
 * r_0..r_{nv-1}, A_0..A_{np-1}, R_0..R_{np-1} (2 nv + 2 np doubles; 16 for Fe).
 * The set_up constants (log-space join at 1.9 A, flat density cap below its
 * maximum) are differentiated with the knots.  No reference counterpart: the
 * reference fits no potentials.
 * ------------------------------------------------------------------------- */
/* loss targets: e_target[ncell] eV, f_target[3*natom] Cartesian eV/A (caller order) */
int mdk_batch_set_targets(void *h, const double *e_target, const double *f_target);
/* loss = sum_c (E_c-E_t,c)^2/N_c^2 + (0.1/natom) sum_i |F_i-F_t,i|^2 and its gradient,
 * one persistent launch (forward, seeds, densities recomputed, adjoint) */
int mdk_batch_loss_grad(void *h, double *loss, double *grad);
/* vector-Jacobian product grad = sum_c dE[c] dE_c/dtheta + sum_i dF_i . dF_i/dtheta at
 * the last forward evaluation (its neighbour list is reused, densities recomputed) */
int mdk_batch_vjp(void *h, const double *dE, const double *dF, double *grad);
/* average ms of reps loss + gradient launches */
int mdk_batch_time_loss_grad(void *h, int reps, double *ms_avg);

/* mdk_batch_loss_grad that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight training loops, harnesses that forbid hipMalloc).
 * loss (1 double) and grad (2 nv + 2 np doubles) are DEVICE pointers, written
 * asynchronously on the batch's stream; mdk_batch_synchronize before reading
 * them (it also reports a neighbour-capacity overflow).  The work-queue
 * counter and the per-cell partials live in a static device array sized for
 * MDK_NOALLOC_MAX_CELLS cells (1.6 MB on each GPU the library runs on, held
 * until the process ends); a batch with more cells gives MDK_ERR_BAD_ARG.
 * All calls on one GPU share that array: calls that overlap on one GPU (two
 * streams or threads) race.  Results equal mdk_batch_loss_grad's bit for bit;
 * the batch's graph is not used. */
#define MDK_NOALLOC_MAX_CELLS 8192
int mdk_batch_loss_grad_noalloc(void *h, double *loss, double *grad);

#ifdef __cplusplus
}
#endif
#endif
