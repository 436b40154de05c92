/* flint_amd -- MI355X (gfx950) implementation of FLINT's per-cell thermochemistry hot path.
 *
 * C ABI.  Plain pointers and sizes; every function returns 0 on success or a FLINT_E* code, with
 * a message available from flint_last_error().  "dev" pointers are device (HBM) pointers,
 * "host" pointers are host memory.  Batched state arrays are structure-of-arrays with leading
 * dimension ld >= n (cell fastest): species k of cell i is at a[k*ld + i].  A stream argument is a
 * hipStream_t passed as void* (NULL = the default stream); calls are asynchronous on it unless
 * stated otherwise.
 *
 * The FLINT reference keeps one mechanism in module globals; here a mechanism is an opaque handle,
 * and every batched call on distinct handles or distinct output arrays may run concurrently.
 */
#ifndef FLINT_AMD_H
#define FLINT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flint_mech flint_mech;

enum {
  FLINT_OK = 0,
  FLINT_EARG = 1,         /* invalid argument */
  FLINT_EHIP = 2,         /* HIP runtime error */
  FLINT_EIO = 3,          /* missing or malformed input file */
  FLINT_EMECH = 4,        /* no compiled kernel for this mechanism name */
  FLINT_EUNSUPPORTED = 5  /* operation not available for this mechanism */
};

/* Per-cell status bits written by the batched calls (status may be NULL). */
#define FLINT_ST_T_KINETICS 1 /* T outside the kinetics table: the first/last bin was used */
#define FLINT_ST_T_THERMO 2   /* T outside [Tmin, Tmax]: rhs_native's F = -1 convention */

const char *flint_last_error(void);

/* Build a mechanism from tables already in memory, in FLINT's own layout.
 * Replaces read_idealgas_thermo (Load_ThermoTransport.f90:7-89), read_chemistry
 * (Load_Chemistry.f90:7-164) and Assign_Mechanism (Lib_Chemistry_wdot.f90:24-92) for callers that
 * hold the module arrays already.
 *   mech_name     FLINT mechanism name: "WD", "TSR-GP-24" and "TSR-Rich-31" have compiled kernels;
 *                 any other name gives a mechanism served by the general routine (flint_general_setup)
 *   W[ns]         molecular weights (wm_tab)
 *   kf, kb        kf_tab/kb_tab(Tk_min:Tk_max, n_arrh), column-major (T fastest)
 *   troe          [4][n_troe][nTk]: kinf_troe_tab, k0_troe_tab, kc_troe_tab, Fcent_tab (may be NULL if 0)
 *   lind          [3][n_lind][nTk]: kinf_lind_tab, k0_lind_tab, kc_lind_tab (may be NULL if 0)
 *   cp, h         cp_tab/h_tab(Th_min:Th_max, ns), column-major
 *   ne, comp      element count and species_composition(ne, ns), column-major (may be 0 / NULL)
 * Host pointers; the tables are copied to the device in the kernels' layout. */
int flint_mech_create(const char *mech_name, int ns, const double *W, int Tk_min, int Tk_max, int n_arrh,
                      const double *kf, const double *kb, int n_troe, const double *troe, int n_lind,
                      const double *lind, int Th_min, int Th_max, const double *cp, const double *h, int ne,
                      const double *comp, flint_mech **out);

/* Read FLINT's native input folder (phase.txt, composition.txt, chemistry-info.txt, thermo.dat,
 * chemistry-{Arrhenius,Troe,Lindemann}.dat, and transport.dat when present; Tecplot ASCII only,
 * .szplt is not supported) and
 * select the compiled kernels for the mechanism named in chemistry-info.txt.
 * Replaces read_idealgas_thermo(folder) + read_chemistry(folder, mech_name) + Assign_Mechanism;
 * prefix plays the role of FLINT_phase_prefix (Lib_ThermoTransport.f90:8). */
int flint_mech_load(const char *folder, const char *prefix, flint_mech **out);

void flint_mech_destroy(flint_mech *m);

/* ns, and the table ranges [Tk_min, Tk_max] (kinetics) and [Th_min, Th_max] (thermo). */
int flint_mech_query(const flint_mech *m, int *ns, int *Tk_min, int *Tk_max, int *Th_min, int *Th_max);

/* What FLINT's loaders leave in FLINT_Lib_Thermodynamic's module arrays (read_idealgas_thermo and
 * read_idealgas_composition, Load_ThermoTransport.f90:30-126), for callers that read them:
 *   flint_mech_thermo       W[ns] (wm_tab) and the thermo tables cp, h, s, dcpi as
 *                           (Th_max - Th_min + 1) x ns column-major arrays (cp_tab(Tmin:Tmax, ns)
 *                           ...); any pointer may be NULL; s and dcpi are 0 for a handle made by
 *                           flint_mech_create (it is given cp and h only);
 *   flint_mech_names        kind 0 species names, kind 1 element names (species_names,
 *                           elements_names), blank-padded into count x len characters; *count
 *                           receives the number (buf may be NULL to query it);
 *   flint_mech_composition  species_composition(ne, ns) column-major into comp[ne*ns], and the
 *                           element weights CEA_initialize_global takes from its table by symbol
 *                           (Lib_CEA_setup.f90:157-170) into el_weight[ne] (0 for an unknown
 *                           symbol); either may be NULL. */
int flint_mech_thermo(const flint_mech *m, double *W, double *cp, double *h, double *s, double *dcpi);
int flint_mech_names(const flint_mech *m, int kind, char *buf, int len, int *count);
int flint_mech_composition(const flint_mech *m, double *comp, double *el_weight);

/* chemistry_source over a batch (Lib_Chemistry_wdot.f90:6-20 procedure pointer, e.g.
 * TSR-GP-24.f90:4-415): rho_dev[ns][ld], T_dev[n] -> wdot_dev[ns][ld] in kg/m^3/s.
 * hrr_dev (may be NULL) receives the heat release rate -sum_k h_k(T) wdot_k [W/m^3].
 * The hand-written WD routine clips rho at 0 (WD.f90:18-21); the batch leaves rho_dev untouched. */
int flint_source(const flint_mech *m, long n, long ld, const double *rho_dev, const double *T_dev, double *wdot_dev,
                 double *hrr_dev, int *status_dev, void *stream);

/* The data-driven general routine (Lib_Chemistry_wdot.f90:96-221, the chemistry_source FLINT's
 * Assign_Mechanism falls back to for a mechanism without an explicit routine): its tables, as
 * chemistry-info.txt's Reaction definition block holds them (Load_Chemistry.f90:55-89), for every
 * reaction of each kind (Arrhenius and three-body, then Troe, then Lindemann, each in file order,
 * na/nt/nl matching the mechanism's rate tables): ni1, ni2, eps [na+nt+nl][ns+1] (host), the
 * last entry of a row being the third body M.  flint_source_general then evaluates it over a batch
 * as flint_source does (status bit 0: T outside the kinetics table, the row clamped).  For a
 * mechanism created without compiled kernels flint_source itself runs general, as FLINT's
 * Assign_Mechanism falls back to it (Lib_Chemistry_wdot.f90:86-89); flint_mech_load sets the
 * tables up from chemistry-info.txt's Reaction definition block when the folder has one. */
int flint_general_setup(flint_mech *m, int na, int nt, int nl, const double *ni1, const double *ni2,
                        const double *eps);
int flint_source_general(const flint_mech *m, long n, long ld, const double *rho_dev, const double *T_dev,
                         double *wdot_dev, int *status_dev, void *stream);

/* Host-memory form of flint_source_general (the general routine of Lib_Chemistry_wdot.f90:96-221
 * that FLINT's Assign_Mechanism falls back to): rho[ns][n], T[n] in, wdot[ns][n] out, staged
 * through the handle's buffer on its own stream, synchronous. */
int flint_source_general_host(flint_mech *m, long n, const double *rho, const double *T, double *wdot);

/* rhs_native over a batch (Lib_Chemistry_rhs.f90:15-52): Z_dev[ns+1][ld] = (rho_1..rho_ns, T)
 * -> F_dev[ns+1][ld] = (wdot, dT/dt) of the constant-volume reactor; F = -1 where T is outside the
 * thermo table.  With J_dev non-NULL also the analytic Jacobian dF/dZ,
 * J_dev[(r*(ns+1)+c)*ld + i] = dF_r/dZ_c (no reference counterpart: FLINT's integrator builds
 * its own; this is the exact derivative of the table-interpolated model). */
int flint_rhs(const flint_mech *m, long n, long ld, const double *Z_dev, double *F_dev, double *J_dev, int *status_dev,
              void *stream);

/* Host-memory convenience for the per-cell drop-in (Fortran chemistry_source / rhs_native
 * callers): copies in, runs the batched kernel on the library's own stream, copies out and
 * synchronises.  Arrays as in flint_source / flint_rhs with ld = n. */
int flint_source_host(flint_mech *m, long n, const double *rho, const double *T, double *wdot);
int flint_rhs_host(flint_mech *m, long n, const double *Z, double *F, double *J);

/* cons2prim over a batch (Lib_ThermoTransport.f90:130-165): cons_dev[np][ld] = (rho_1..rho_ns,
 * rho u_1..rho u_nd, rho E) with np = ns + nd + 1 and 0 <= nd <= 3, T_dev[n] the Newton start ->
 * prim_dev[np][ld] = (rho_1..rho_ns, u_1..u_nd, p).  Newton on T from the internal energy with
 * the reference's tolerance (|dE/E| < 1e-6) and limit (20 iterations), then p = rho R T.  T_out_dev
 * (the converged T), iters_dev and status_dev may be NULL; status 2 (and NaN outputs) where T
 * leaves the thermo table, which the reference does not check. */
int flint_cons2prim(const flint_mech *m, long n, long ld, int nd, const double *cons_dev, const double *T_dev,
                    double *prim_dev, double *T_out_dev, int *iters_dev, int *status_dev, void *stream);

/* prim2cons over a batch (Lib_ThermoTransport.f90:168-185), the inverse map: prim_dev -> cons_dev
 * with rho E = rho (E(p, rho_k) + |u|^2 / 2); status 2 (NaN energy) where p / (R rho) is outside
 * the thermo table. */
int flint_prim2cons(const flint_mech *m, long n, long ld, int nd, const double *prim_dev, double *cons_dev,
                    int *status_dev, void *stream);

/* Host-memory forms of flint_cons2prim / flint_prim2cons (ld = n, synchronous, as flint_rhs_host);
 * T_out and status may be NULL. */
int flint_cons2prim_host(flint_mech *m, long n, int nd, const double *cons, const double *T, double *prim,
                         double *T_out, int *status);
int flint_prim2cons_host(flint_mech *m, long n, int nd, const double *prim, double *cons);

/* Species transport tables (replaces read_idealgas_transport's fill of mi_tab / k_tab,
 * Load_ThermoTransport.f90:129-174; flint_mech_load calls it itself when the folder holds
 * transport.dat).  mu_tab, k_tab: host arrays laid out as the reference's mi_tab(Tt_min:Tt_max, ns)
 * and k_tab, temperature fastest, viscosity in Pa s and conductivity in W/(m K).  Replaces any
 * tables set before. */
int flint_transport_setup(flint_mech *m, int Tt_min, int Tt_max, const double *mu_tab, const double *k_tab);

/* Mixture viscosity and thermal conductivity by Wilke's rule over a batch (replaces
 * co_k_mi_lam_Wilke, Lib_ThermoTransport.f90:407-462, with co_fiij :350-362; f_laminarViscosity
 * :365-404 is the same mu with T = p / (Rtot rho)): rhoi_dev[ns][ld], rho_dev[n] (the density the
 * mole fractions divide by, as the reference's rho argument), T_dev[n] -> mu_dev[n], lam_dev[n].
 * status_dev (may be NULL): 0, or 2 with NaN outputs where T is outside [Tt_min, Tt_max), which
 * the reference does not check.  FLINT_EUNSUPPORTED before flint_transport_setup. */
int flint_transport(const flint_mech *m, long n, long ld, const double *rhoi_dev, const double *rho_dev,
                    const double *T_dev, double *mu_dev, double *lam_dev, int *status_dev, void *stream);

/* Host-memory form of flint_transport (ld = n, synchronous, as flint_rhs_host). */
int flint_transport_host(flint_mech *m, long n, const double *rhoi, const double *rho, const double *T, double *mu,
                         double *lam, int *status);

/* Batched 0D reactors (no reference counterpart in FLINT's own code: it advances one reactor at a
 * time through the external OSlo integrator, test-batchF.f90:81,105, with rhs_native as the
 * callback).  Advances n cells from t = 0 to t_end in one persistent launch, variable-order BDF
 * (orders 1..5) with Newton and a per-cell dense LU, rtol/atol per component.
 *   mode 0: constant volume, X_dev[(ns+1)][ld] = (rho_1..rho_ns, T) (rhs_native)
 *   mode 1: constant pressure, X_dev = (Y_1..Y_ns, T), p_dev[n] in Pa
 * X_dev is advanced in place.  stats_dev[8][ld] (may be NULL) receives per cell: accepted steps,
 * rejected steps, Newton failures, RHS evaluations, Jacobians, LU factorisations, Newton
 * iterations, final order.  status_dev[n] (may be NULL): 0 ok, 1 step underflow, 2 max_steps
 * reached, 3 T outside the tables.  max_steps <= 0 means 20000.  Needs ns + 1 <= 32. */
int flint_reactor_advance(flint_mech *m, long n, long ld, int mode, double *X_dev, const double *p_dev, double t_end,
                          double rtol, double atol, int max_steps, int *stats_dev, int *status_dev, void *stream);

/* Right-hand side and analytic Jacobian of the reactor model that flint_reactor_advance integrates
 * (mode 0: rhs_native and d/d(rho_k, T); mode 1: d(Y_k, T)/dt at fixed p_dev and d/d(Y_k, T)), at the
 * states X_dev[(ns+1)][ld]: F_dev[(ns+1)][ld], J_dev[((ns+1)*r + c)][ld] = dF_r/dX_c (J_dev may be
 * NULL: right-hand side only).  F = -1 and J = 0 for cells outside the tables.  No reference counterpart (FLINT hands rhs_native to its
 * external integrator, test-batchF.f90:81,105, which differentiates it itself). */
int flint_reactor_jacobian(flint_mech *m, long n, long ld, int mode, const double *X_dev, const double *p_dev,
                           double *F_dev, double *J_dev, void *stream);

/* As flint_reactor_advance, launched through a hipGraph that is captured on the first call and
 * replayed while the arguments stay identical. */
int flint_reactor_advance_graph(flint_mech *m, long n, long ld, int mode, double *X_dev, const double *p_dev,
                                double t_end, double rtol, double atol, int max_steps, int *stats_dev, int *status_dev,
                                void *stream);

/* No counterpart in the reference: flint_reactor_advance for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernel and results; it allocates nothing: the per-team Jacobian slots and the work
 * counter live in a static device array of the library (room for 2048 teams, 17.3 MB per GPU,
 * held until the process ends), and the handle's program was uploaded by flint_mech_create.  A
 * launch shape (flint_reactor_set_launch) of more than 2048 teams is rejected with FLINT_EARG.
 * All calls on one GPU share that array: calls that overlap in time on one GPU race. */
int flint_reactor_advance_noalloc(flint_mech *m, long n, long ld, int mode, double *X_dev, const double *p_dev,
                                  double t_end, double rtol, double atol, int max_steps, int *stats_dev,
                                  int *status_dev, void *stream);

/* Host-memory convenience for the reactor batch (the time loop of test-batchF.f90:99-108, which
 * advances one reactor through OSlo's run_odesolver with rhs_native): copies X[(ns+1)][n] and p[n] in,
 * runs flint_reactor_advance with ld = n on the library's own stream, copies X and status[n] (may
 * be NULL) back and synchronises.  p is read for mode 1 only. */
int flint_reactor_advance_host(flint_mech *m, long n, int mode, double *X, const double *p, double t_end,
                               double rtol, double atol, int max_steps, int *status);

/* Launch-shape autotuner: times the candidate (teams per block, blocks) shapes on a copy of the
 * given sample and keeps the fastest for this handle (synchronous; the sample is not modified). */
int flint_reactor_autotune(flint_mech *m, long n, long ld, int mode, const double *X_dev, const double *p_dev,
                           double t_end, double rtol, double atol, double *best_ms);
int flint_reactor_set_launch(flint_mech *m, int teams_per_block, int blocks);
int flint_reactor_get_launch(flint_mech *m, int *teams_per_block, int *blocks);

/* Newton policy of the reactor's BDF for this handle (the setup_odesolver solver choice of
 * test-batchF.f90:81, whose 'cvode' is SUNDIALS CVODE): policy 1 (the default) iterates as CVODE
 * does (rate estimate carried between steps, I - cJ reused while c changes by at most 30 %,
 * at most 3 iterations); policy 0 as scipy's BDF (rate measured within the step, I - cJ
 * refactored at every change of step or order, at most 4 iterations). */
int flint_reactor_set_newton(flint_mech *m, int policy);

/* Work order of flint_reactor_advance and flint_reactor_advance_graph for this handle (no
 * reference counterpart: FLINT advances one reactor at a time, test-batchF.f90:99-108).  Policy 0
 * (the default) hands the cells to the persistent kernel's teams in the caller's order; policy 1
 * evaluates dX/dt at the initial states and hands them out by descending |dT/dt|, so that the
 * costliest (igniting) cells start first and no late slow cell sets the launch's tail.  Each cell is
 * integrated independently: results are identical under either policy.  Policy 1 allocates a
 * right-hand-side buffer and a permutation on first use; flint_reactor_advance_noalloc keeps the
 * caller's order. */
int flint_reactor_set_order(flint_mech *m, int policy);

/* Equilibrium data of a mechanism: CEA_initialize_global (Lib_CEA_setup.f90:6-99) for the
 * mechanism's species set.  s = s_tab(Th_min:Th_max, ns) per-mass entropy at 1 bar, column-major
 * like cp/h (read_idealgas_thermo, Load_ThermoTransport.f90:30-85); el_weight[ne] = atomic weights
 * of the elements in species_composition's row order (the reference takes them from its
 * Atmwt table by symbol, Lib_CEA_setup.f90:157-170).  flint_mech_load calls this itself. */
int flint_cea_setup(flint_mech *m, const double *s, const double *el_weight);

/* Convergence tests of the equilibrium solves of this handle (flint_cea_solve, _host, _warm):
 * tol = 0 (the default) keeps the reference's (Lib_CEA_solver.f90:345-364: relative corrections
 * <= 0.5e-5, |dln T| <= 1e-4, element residual <= 1e-6 max b0), whose converged state depends on
 * the starting point by up to ~1e-4 K; tol in (0, 1e-4] applies tol to all of them (no
 * counterpart in the reference), at about one more Newton iteration for tol = 1e-10, so that any
 * start lands on the same equilibrium to round-off. */
int flint_cea_set_tolerance(flint_mech *m, double tol);

/* CEA_solve over a batch (Lib_CEA_solver.f90:6-856, with CEA_initialize_local per state):
 *   hp = 1  HP: fixed pressure p_dev[n] (bar) and the enthalpy of the input mixture at T_dev[n]
 *           (the reference with Vol = .false.);
 *   hp = 0  UV: fixed internal energy at T_dev[n] and volume 1e5/sum(rho) (Vol = .true., the
 *           reference's default; p_dev is ignored and may be NULL).
 * rho_dev[ns][ld] = input amounts (mass fractions, or densities: the reference normalises the
 * element amounts by their sum but not the energy, so HP inputs should sum to 1).  Outputs
 * T_out_dev[n], Y_out_dev[ns][ld] = n_k W_k (kg/kg), iters_dev[n] Newton iterations and
 * status_dev[n]: 0 converged, 1 no convergence in 50 iterations, 2 singular Newton matrix after
 * the reference's retries, 3 temperature outside the thermo table (T_out and Y_out are NaN
 * unless 0).  guess_dev[ns+1][ld] (may be NULL) = (T, ln n_1..ln n_ns) starts the Newton
 * iteration there instead of at the reference's default (T_in, n_j = 0.1/ns).  iters_dev and
 * status_dev may be NULL.  Needs a generated mechanism (its element matrix is compiled in). */
int flint_cea_solve(flint_mech *m, long n, long ld, int hp, const double *p_dev, const double *T_dev,
                    const double *rho_dev, const double *guess_dev, double *T_out_dev, double *Y_out_dev,
                    int *iters_dev, int *status_dev, void *stream);

/* Host-memory convenience for per-state CEA_solve(p_in, T_in, rhoi_in, T_out, rhoi_out) callers
 * (Lib_CEA_solver.f90:6-15): as flint_cea_solve with ld = n and the default initial guess, on the
 * library's own stream, synchronous.  p[n] is read for hp = 1 only; status[n] may be NULL. */
int flint_cea_solve_host(flint_mech *m, long n, int hp, const double *p, const double *T, const double *rho,
                         double *T_out, double *Y_out, int *status);

/* Learned warm start of HP equilibria (no reference counterpart): the network net (ns+2 inputs,
 * ns+1 outputs, see flint_mlp_create below) maps the standardised (Y0_k, h0, ln p) of each state
 * (h0 = sum_k Y0_k h_k(T0)) to the standardised Newton start (T, ln n_1..ln n_ns); scaling (host,
 * 4 ns + 6 doubles) = in_mean[ns+2], in_scale[ns+2], out_mean[ns+1], out_scale[ns+1].  The batch is
 * solved as flint_cea_solve with hp = 1 from those starts, held to 6 iterations (see
 * flint_cea_warm_options), and the states that do not converge within them are solved again from
 * the default start, with the reference's 50-iteration limit, in a second pass over just those
 * states (iters_dev then counts both), so the results are those of the default path wherever it
 * converges.  Asynchronous on stream, no host synchronisation.  Arguments as in
 * flint_cea_solve (rho_dev = input mass fractions, p_dev in bar); status_dev may be NULL. */
typedef struct flint_mlp flint_mlp;
int flint_cea_solve_warm(flint_mech *m, flint_mlp *net, long n, long ld, const double *p_dev, const double *T_dev,
                         const double *rho_dev, const double *scaling, double *T_out_dev, double *Y_out_dev,
                         int *iters_dev, int *status_dev, void *stream);

/* Options of flint_cea_solve_warm (no reference counterpart), held until the next
 * flint_cea_setup: max_iter (1..50, default 6) caps the warm-started pass, so that the states
 * that converge slowly from their predicted start move to the second pass (a wavefront runs as
 * many iterations as its slowest state); fused (default 1)
 * computes the features, a 3-layer network (<= 32 inputs, hidden widths 128 and 128, <= 32
 * outputs) and the Newton start in one kernel, 0 runs them as separate launches (the layered
 * forward pass of flint_mlp_forward). */
int flint_cea_warm_options(flint_mech *m, int max_iter, int fused);

/* Multilayer perceptron on MFMA (no reference counterpart: the learned reactor step and the
 * equilibrium warm start of this library).  Layers d_0 -> d_1 -> ... -> d_L with GELU on every
 * hidden layer and a linear output; BF16 weights and activations, FP32 accumulation, FP32 master
 * weights.  Parameters are a flat FP32 vector, per layer W_l [d_l][d_{l+1}] row-major then b_l.
 * Batches are row-major FP32 device arrays X [n][d_0], targets/outputs [n][d_L]. */
int flint_mlp_create(int nlayers, const int *dims, unsigned long long seed, flint_mlp **out);
void flint_mlp_destroy(flint_mlp *m);
int flint_mlp_num_params(const flint_mlp *m, long *count);
int flint_mlp_get_params(flint_mlp *m, float *host_params);        /* synchronous */
int flint_mlp_set_params(flint_mlp *m, const float *host_params);  /* synchronous; resets Adam */
int flint_mlp_forward(flint_mlp *m, long n, const float *X_dev, float *Y_dev, void *stream);
/* gradient of the mean squared error mean((Y - T)^2) over the batch (grad_dev [nparams] and the
 * loss_dev scalar may be NULL) */
int flint_mlp_gradients(flint_mlp *m, long n, const float *X_dev, const float *T_dev, float *grad_dev,
                        float *loss_dev, void *stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch's mean squared error */
int flint_mlp_train_step(flint_mlp *m, long n, const float *X_dev, const float *T_dev, float lr, float *loss_dev,
                         void *stream);

/* The FP64 judge of the learned step as a building block, for any predictor of a
 * constant-pressure step (no reference counterpart).  For each cell, a candidate state Xp_dev
 * [(ns+1)][ld] after a step of dt from X_dev (both (Y_1..Y_ns, T), p_dev [n] in Pa): the Newton
 * corrections from Xp towards the backward-Euler and the trapezoidal step, with F and J at Xp,
 * delta_BE = -(I - dt J)^-1 (Xp - X - dt F(Xp)) and
 * delta_TR = -(I - dt/2 J)^-1 (Xp - X - dt/2 (F(X) + F(Xp))).  flag_dev [n] receives 0 when
 * both stay within tol_y in every Y_k and tol_t in T, else 1 (also for T outside the tables);
 * check_dev [2][ld] (may be NULL) the larger correction's max_k |delta_k| and |delta_T|.  X_dev
 * and Xp_dev are not changed.  Needs ns + 1 <= 32. */
int flint_reactor_check_step(flint_mech *m, long n, long ld, const double *X_dev, const double *Xp_dev,
                             const double *p_dev, double dt, double tol_y, double tol_t, int *flag_dev,
                             double *check_dev, void *stream);

/* Learned constant-pressure reactor step with the batched BDF as fallback (no reference
 * counterpart).  X_dev = (Y_1..Y_ns, T) [ns+1][ld] and p_dev [n] (Pa) as in flint_reactor_advance
 * mode 1; net maps the standardised features (log max(Y_k, ymin), T, log p) [ns+2], or with
 * 2 ns + 3 inputs those and the rate features asinh(t_end dY_k/dt / yfloor), asinh(t_end dT/dt /
 * tfloor), to the standardised increments over t_end in units of the local rate,
 * o_k = dY_k / (t_end |dY_k/dt| + yfloor) and o_T = dT / (t_end |dT/dt| + tfloor) with dX/dt at
 * the old state (the right-hand side of flint_reactor_jacobian mode 1) [ns+1].  scaling (host,
 * 2 nin + 2 ns + 2 doubles, nin the input width) = in_mean[nin], in_scale[nin], out_mean[ns+1],
 * out_scale[ns+1]; el_weight (host, [ne]) =
 * atomic weights of the elements.  A cell falls back to flint_reactor_advance(mode 1, t_end,
 * rtol, atol) when its raw prediction changes an element mass fraction or the total mass by more
 * than max_residual or predicts |dT| > max_dT; an accepted prediction is projected onto the old
 * element amounts and total mass (exactly, to round-off) and T solved from the old mixture
 * enthalpy, and the cell still falls back when that T differs from the predicted T + dT by more
 * than max_T_mismatch.  Every prediction that passes is then judged in FP64 by the Newton
 * corrections over t_end from the projected prediction Xp (F and J at Xp) towards the
 * backward-Euler and the trapezoidal step,
 * delta_BE = -(I - t_end J)^-1 (Xp - X - t_end F(Xp)),
 * delta_TR = -(I - t_end/2 J)^-1 (Xp - X - t_end/2 (F(X) + F(Xp))),
 * and the cell is accepted only when |delta_k| <= max_newton_dY for every species and
 * |delta_T| <= max_newton_dT in both.  fallback_dev [ld] (may be NULL) receives 1 for the cells
 * that took the BDF, *n_fallback their number, check_dev [2][ld] (may be NULL) the larger
 * correction's max_k |delta_k| and |delta_T| of each judged cell, stats_dev/status_dev
 * as flint_reactor_advance (zero for accepted cells).  Synchronises the stream once (to size the
 * fallback launch).  Needs ns = 24 or 31. */
int flint_reactor_advance_mlp(flint_mech *m, flint_mlp *net, long n, long ld, double *X_dev, const double *p_dev,
                              double t_end, double rtol, double atol, const double *scaling, const double *el_weight,
                              double ymin, double yfloor, double tfloor, double max_dT, double max_residual,
                              double max_T_mismatch, double max_newton_dY, double max_newton_dT, int *fallback_dev,
                              long *n_fallback, double *check_dev, int *stats_dev, int *status_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
