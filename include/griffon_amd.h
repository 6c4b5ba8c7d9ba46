/*
 * griffon_amd.h — C ABI of the MI355X (gfx950) implementation of Spitfire's Griffon chemistry hot path.
 *
 * Two kinds of entry points:
 *   1. Drop-ins for the reference's Cython boundary (src/spitfire/griffon/griffon.pyx): same argument
 *      order and meaning, host pointers, synchronous; the arithmetic runs on the GPU (the library
 *      stages the arrays through device memory). Each cites the griffon.pyx line it replaces.
 *   2. Batched device entry points: device pointers, many independent states / grid points /
 *      flamelets per call, ordered on the caller's HIP stream (passed as void*).
 *
 * Conventions: every function returns GF_OK (0) or a nonzero GF_ERR_* code; gf_last_error() gives the
 * message of the calling thread's last failure. No exceptions cross this boundary. All floating-point
 * data is FP64 unless a name says f32. Jacobians are column-major ns x ns blocks (row i, column j at
 * i + j*ns), as in griffon. State vectors are [T, Y_0 .. Y_{ns-2}].
 */
#ifndef GRIFFON_AMD_H
#define GRIFFON_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define GF_OK 0
#define GF_ERR_ARG 1          /* bad argument (index, size, null pointer) */
#define GF_ERR_STATE 2        /* call not valid in the mechanism's current state */
#define GF_ERR_UNSUPPORTED 3  /* feature outside this implementation's scope */
#define GF_ERR_HIP 4          /* HIP runtime failure (no device, launch error, ...) */

typedef struct gf_mech gf_mech;

const char *gf_last_error(void);
int gf_device_count(void);

/* ---------------- mechanism construction (griffon.pyx:234-551, chemistry_setup.cpp:21-732) ----------- */
/* Species are added in order; name->index resolution and the element map stay with the caller, which
 * passes each species' molecular weight (the reference sums element weights in mechanism_add_species,
 * chemistry_setup.cpp:36-62). Reactant/product/efficiency/order lists are index arrays given in the
 * order griffon's std::map would iterate them (sorted species names). */
int gf_mech_create(gf_mech **out);
int gf_mech_destroy(gf_mech *m);
int gf_mechanism_set_gas_constant(gf_mech *m, double Ru);              /* griffon.pyx mechanism_set_gas_constant */
int gf_mechanism_set_ref_pressure(gf_mech *m, double p_ref);           /* griffon.pyx mechanism_set_ref_pressure */
int gf_mechanism_set_ref_temperature(gf_mech *m, double T_ref);        /* griffon.pyx mechanism_set_ref_temperature */
int gf_mechanism_add_species(gf_mech *m, double molecular_weight);     /* griffon.pyx mechanism_add_species */
int gf_mechanism_resize_heat_capacity_data(gf_mech *m);                /* griffon.pyx mechanism_resize_heat_capacity_data */
int gf_mechanism_add_const_cp(gf_mech *m, int species, double Tmin, double Tmax, double T0, double h0, double s0,
                              double cp);                              /* griffon.pyx mechanism_add_const_cp */
int gf_mechanism_add_nasa7_cp(gf_mech *m, int species, double Tmin, double Tmid, double Tmax,
                              const double *low_coeffs7, const double *high_coeffs7); /* mechanism_add_nasa7_cp */
/* type: 1 simple, 2 three-body, 3 Lindemann, 4 Troe (all mechanism_add_reaction_* variants of
 * griffon.pyx, with or without special orders); kfwd = {A, b, Ea/R}, kpres likewise (falloff low-pressure
 * rate), troe = {alpha, T***, T*, T**}. */
int gf_mechanism_add_reaction(gf_mech *m, int type, int reversible, int n_reactants, const int *reactant_idx,
                              const int *reactant_stoich, int n_products, const int *product_idx,
                              const int *product_stoich, const double *kfwd, int n_third_body, const int *tb_idx,
                              const double *tb_efficiency, double default_efficiency, const double *kpres,
                              const double *troe, int n_orders, const int *order_idx, const double *order_value);
int gf_mech_n_species(const gf_mech *m);
int gf_mech_n_reactions(const gf_mech *m);

/* ---------------- drop-ins for PyCombustionKernels compute methods (host pointers) ------------------- */
int gf_production_rates(gf_mech *m, double T, double rho, const double *y, double *out_w);         /* griffon.pyx:763-768 */
int gf_prod_rates_primitive_sensitivities(gf_mech *m, double rho, double T, const double *y, int rates_sensitivity_option,
                                          double *out_sens);                                        /* griffon.pyx:773-783 */
int gf_reactor_rhs_isobaric(gf_mech *m, const double *state, double pressure, double inflow_T, const double *inflow_y,
                            double tau, double fluid_T, double surf_T, double h_conv, double eps_rad,
                            double surface_area_over_volume, int heat_transfer_option, int open,
                            double *out_rhs);                                                       /* griffon.pyx:788-802 */
int gf_reactor_jac_isobaric(gf_mech *m, const double *state, double pressure, double inflow_T, const double *inflow_y,
                            double tau, double fluid_T, double surf_T, double h_conv, double eps_rad,
                            double surface_area_over_volume, int heat_transfer_option, int open,
                            int rates_sensitivity_option, int sensitivity_transform_option, double *out_rhs,
                            double *out_jac);                                                       /* griffon.pyx:807-824 */
/* Isochoric reactor, state [rho, T, Y_0 .. Y_{ns-2}]; rhs has ns+1 entries, jac (ns+1)^2 column-major. */
int gf_reactor_rhs_isochoric(gf_mech *m, const double *state, double inflow_density, double inflow_T,
                             const double *inflow_y, double tau, double fluid_T, double surf_T, double h_conv,
                             double eps_rad, double surface_area_over_volume, int heat_transfer_option, int open,
                             double *out_rhs);                                                      /* griffon.pyx:831-845 */
int gf_reactor_jac_isochoric(gf_mech *m, const double *state, double inflow_density, double inflow_T,
                             const double *inflow_y, double tau, double fluid_T, double surf_T, double h_conv,
                             double eps_rad, double surface_area_over_volume, int heat_transfer_option, int open,
                             int rates_sensitivity_option, double *out_rhs, double *out_jac);        /* griffon.pyx:850-866 */
/* Thermodynamic getters. Scalars come back through out_value. y holds all ns mass fractions. */
int gf_mixture_molecular_weight(gf_mech *m, const double *y, double *out_value);                    /* griffon.pyx:684-688 */
int gf_mole_fractions(gf_mech *m, const double *y, double *out_x);                                  /* griffon.pyx:690-694 */
int gf_ideal_gas_density(gf_mech *m, double p, double T, const double *y, double *out_value);       /* griffon.pyx:696-700 */
int gf_ideal_gas_pressure(gf_mech *m, double rho, double T, const double *y, double *out_value);    /* griffon.pyx:702-706 */
int gf_cp_mix(gf_mech *m, double T, const double *y, double *out_value);                            /* griffon.pyx:708-712 */
int gf_cv_mix(gf_mech *m, double T, const double *y, double *out_value);                            /* griffon.pyx:714-718 */
int gf_species_cp(gf_mech *m, double T, double *out_cp);                                            /* griffon.pyx:720-724 */
int gf_species_cv(gf_mech *m, double T, double *out_cv);                                            /* griffon.pyx:726-730 */
int gf_enthalpy_mix(gf_mech *m, double T, const double *y, double *out_value);                      /* griffon.pyx:732-736 */
int gf_energy_mix(gf_mech *m, double T, const double *y, double *out_value);                        /* griffon.pyx:738-742 */
int gf_species_enthalpies(gf_mech *m, double T, double *out_h);                                     /* griffon.pyx:744-748 */
int gf_species_energies(gf_mech *m, double T, double *out_e);                                       /* griffon.pyx:750-754 */
int gf_dcpdT_species(gf_mech *m, double T, const double *y, double *out_dcpdT);                     /* griffon.pyx:756-760 */
int gf_flamelet_stencils(gf_mech *m, const double *dz, int nzi, const double *dissipation_rate,
                         const double *inv_lewis, double *out_cmajor, double *out_csub, double *out_csup,
                         double *out_mcoeff, double *out_ncoeff);                                   /* griffon.pyx:556-566 */
int gf_flamelet_jac_indices(gf_mech *m, int nzi, int *out_rows, int *out_cols);                     /* griffon.pyx:571-575 */
/* heat-loss arrays in the C++ order T_conv, h_conv, T_rad, h_rad (griffon.pyx:606-609 reorders the Python
 * order T_conv, T_rad, h_conv, h_rad into this one). */
int gf_flamelet_rhs(gf_mech *m, const double *state, double pressure, const double *oxy_state, const double *fuel_state,
                    int adiabatic, const double *T_conv, const double *h_conv, const double *T_rad,
                    const double *h_rad, int nzi, const double *cmajor, const double *csub, const double *csup,
                    const double *mcoeff, const double *ncoeff, const double *chi, int include_enthalpy_flux,
                    int include_variable_cp, int use_scaled_heat_loss, double *out_rhs);           /* griffon.pyx:580-620 */
int gf_flamelet_jacobian(gf_mech *m, const double *state, double pressure, const double *oxy_state,
                         const double *fuel_state, int adiabatic, const double *T_conv, const double *h_conv,
                         const double *T_rad, const double *h_rad, int nzi, const double *cmajor,
                         const double *csub, const double *csup, const double *mcoeff, const double *ncoeff,
                         const double *chi, int compute_eigenvalues, double diffterm, int scale_and_offset,
                         double prefactor, int rates_sensitivity_option, int sensitivity_transform_option,
                         int include_enthalpy_flux, int include_variable_cp, int use_scaled_heat_loss,
                         double *out_expeig, double *out_jac);                                     /* griffon.pyx:625-679 */

/* ---------------- drop-ins for py_btddod_* (griffon.pyx:1074-1113) ------------------------------------ */
int gf_btddod_full_factorize(double *inout_values, int num_blocks, int block_size, double *out_l_values,
                             int *out_pivots);                                                      /* griffon.pyx:1074-1079 */
int gf_btddod_full_solve(const double *d_factors, const double *l_values, const int *pivots, const double *rhs,
                         int num_blocks, int block_size, double *out_solution);                    /* griffon.pyx:1084-1091 */
int gf_btddod_scale_and_add_diagonal(double *inout_values, double matrix_scale, const double *diagonal,
                                     double diag_scale, int num_blocks, int block_size);          /* griffon.pyx:1107-1113 */

/* ---------------- batched device entry points (device pointers, stream-ordered) ---------------------- */
/* Vector steps of batched implicit stage solves (no counterpart in the reference; they fuse the elementwise
 * work of the frozen-Jacobian Newton method, nonlinear.py:185-268, for many independent systems of n
 * unknowns, system s at elements s*n .. s*n+n-1; device pointers; conv is one byte per system, 0 or 1).
 *   masked_update:  x -= d for the systems with conv[s] == 0
 *   stage_residual: r = h[s] (gamma f + prior) - (x - xn); with test != 0 also conv[s] = 1 where
 *                   max_v |r_v w_v| < tol (w: n weights). Same operations, same order, no contraction. */
int gf_batch_masked_update(int R, int n, double *x, const double *d, const unsigned char *conv, void *stream);
int gf_batch_stage_residual(int R, int n, const double *h, double gamma, const double *f, const double *prior,
                            const double *x, const double *xn, const double *w, double tol, double *r,
                            unsigned char *conv, int test, void *stream);
/* One implicit stage of the discrete adjoint of batched closed adiabatic isobaric reactor integrations (no
 * counterpart in the reference; device pointers, R reactors of n = ns state variables, 2 <= n <= 64).
 * x, f, jac, aux, cpi: the stage state, RHS, column-major Jacobian, [rho, cp, dcp/dT, mixture molecular
 * weight] and species cp, as gf_batch_reactor_jac_isobaric_terms gives them; dinvm[k] = 1/M_k - 1/M_{ns-1}.
 *   stage_prep:   rhs = h (J^T v + b_i grad(qdot)), qdot = rho cp f_T; m (may be null) = I - h a_ii J^T
 *   stage_finish: wv = v + a_ii theta, wT = wv_0 + b_i rho cp, C_k = wv_{k+1}/rho (0 for the last),
 *                 C_k -= wT/(rho cp) h_k, cm_k = C_k M_k; L += h b_i rho cp f_T */
int gf_batch_adjoint_stage_prep(int R, int n, const double *x, const double *f, const double *jac, const double *aux,
                                const double *cpi, const double *dinvm, const double *v, const double *h, double b_i,
                                double a_ii, double *rhs, double *m, void *stream);
int gf_batch_adjoint_stage_finish(int R, int n, const double *v, const double *theta, const double *aux,
                                  const double *hsp, const double *mw, const double *f, const double *h, double b_i,
                                  double a_ii, double *cm, double *L, void *stream);
/* Launch shape of the batched chemistry kernels (no counterpart in the reference): -1 (the default) times the
 * feasible shapes at the first launch of each kind (RHS + Jacobian, RHS only) and batch-size class
 * [2^k, 2^(k+1)) on a mechanism and keeps the fastest; 0..4 fixes 1x128, 2x256, 4x512, 4x256 or 2x512
 * (states x threads per workgroup). All shapes give bitwise identical results. Process-wide. */
int gf_set_chem_launch_shape(int shape);
/* Closed adiabatic isobaric reactor chemistry over n independent states. Element v of state i is read
 * from x[i*x_item_stride + v*x_var_stride] (x_item_stride=1, x_var_stride=n is structure-of-arrays);
 * p[i*p_stride] is the pressure (p_stride 0 = one pressure). rhs element v of state i goes to
 * rhs[i*rhs_item_stride + v*rhs_var_stride]; the ns*ns Jacobian of state i to jac + i*jac_item_stride.
 * Replaces n calls of reactor_rhs_isobaric / reactor_jac_isobaric (griffon.pyx:788-824) with
 * heat_transfer_option 0, open false, rates_sensitivity_option 0, sensitivity_transform_option 0. */
int gf_batch_reactor_rhs_isobaric(gf_mech *m, int n, const double *x, long x_item_stride, long x_var_stride,
                                  const double *p, long p_stride, double *rhs, long rhs_item_stride,
                                  long rhs_var_stride, void *stream);
int gf_batch_reactor_jac_isobaric(gf_mech *m, int n, const double *x, long x_item_stride, long x_var_stride,
                                  const double *p, long p_stride, double *rhs, long rhs_item_stride,
                                  long rhs_var_stride, double *jac, long jac_item_stride, void *stream);
/* The same evaluation with the extra per-state terms a discrete adjoint of reactor integrations needs (no
 * counterpart in the reference; it serves gradients of integrated heat release with respect to the rate
 * constants): aux[4 i..] = {rho, cp, dcp/dT, mixture molecular weight}, cpi[ns i..] species heat capacities,
 * h[ns i..] species enthalpies per unit mass, qa[nr i..] = d(rate of progress of reaction r)/d(ln A_r), reactions
 * in griffon order (type-sorted, mechanism.py:540), A the forward (high-pressure, for falloff) pre-exponential
 * factor. rhs and every extra output may be null. */
int gf_batch_reactor_jac_isobaric_terms(gf_mech *m, int n, const double *x, long x_item_stride, long x_var_stride,
                                        const double *p, long p_stride, double *rhs, long rhs_item_stride,
                                        long rhs_var_stride, double *jac, long jac_item_stride, double *aux,
                                        double *cpi, double *h, double *qa, void *stream);

/* Reactor RHS and Jacobian over n independent states for every configuration of reactor_rhs/jac_isobaric and
 * reactor_rhs/jac_isochoric (griffon.pyx:788-866): isochoric 0 reads state i as [T, Y_0..Y_{ns-2}] with pressure
 * p[i*p_stride]; isochoric 1 reads [rho, T, Y_0..Y_{ns-2}] (p unused, inflow_density is the feed density). State i
 * starts at x + i*x_item_stride, its variables contiguous. rhs + i*neq and jac + i*neq*neq (column-major; jac may
 * be null) receive the outputs, neq = ns (+1 isochoric). heat 0/1/2, open, rates_sensitivity_option 0/1/2 and the
 * feed and heat-transfer parameters as in the drop-ins; inflow_y is a device array of ns mass fractions (used when
 * open). Device pointers, stream-ordered. Partial results (rates, sensitivities, thermo) go to a workspace the
 * mechanism owns and grows on demand (synchronous hipMalloc); calls that overlap on one mechanism race on it. */
int gf_batch_reactor_general(gf_mech *m, int n, int isochoric, const double *x, long x_item_stride, const double *p,
                             long p_stride, double inflow_density, double inflow_T, const double *inflow_y, double tau,
                             double fluid_T, double surf_T, double h_conv, double eps_rad,
                             double surface_area_over_volume, int heat_transfer_option, int open,
                             int rates_sensitivity_option, double *rhs, double *jac, void *stream);

/* gf_batch_reactor_general without any allocation (no counterpart in the reference). It is for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The
 * partial results live in a fixed 32 MiB device array of the library (one per GPU, held until the process ends)
 * instead of the mechanism's workspace: n may be at most floor(4194304 / (3 ns + 4 + (ns+1)^2)) states with a
 * Jacobian (1,362 at 53 species) or floor(4194304 / (3 ns + 4)) without; larger n returns GF_ERR_ARG and launches
 * nothing. It calls no hipMalloc/hipFree and never times launch shapes; the mechanism must already be resident on
 * the device (any earlier compute call on it), else GF_ERR_STATE. Every call on one GPU shares that array, so calls
 * that overlap on one GPU race; results are bitwise those of gf_batch_reactor_general. */
int gf_batch_reactor_general_noalloc(gf_mech *m, int n, int isochoric, const double *x, long x_item_stride,
                                     const double *p, long p_stride, double inflow_density, double inflow_T,
                                     const double *inflow_y, double tau, double fluid_T, double surf_T, double h_conv,
                                     double eps_rad, double surface_area_over_volume, int heat_transfer_option,
                                     int open, int rates_sensitivity_option, double *rhs, double *jac, void *stream);

/* Mixture properties of n states, device pointers, stream-ordered: Ty[n][ns+1] holds T then the ns mass fractions
 * of each state; p[n] and rho[n] may each be NULL (then the density-from-pressure, resp. pressure-from-density,
 * entry is computed from 0). out[n][8] = {mixture molecular weight, ideal-gas density from p, ideal-gas pressure from
 * rho, cp, cv, enthalpy, energy, d(cp)/dT} per state, by the formulas of the mixture getters above
 * (griffon.pyx:684-758). The batched form of those getters, for the property evaluators of
 * src/spitfire/chemistry/analysis.py:77-220 (which go through Cantera's SolutionArray in the reference). */
int gf_batch_mixture_properties(gf_mech *m, int n, const double *Ty, const double *p, const double *rho, double *out,
                                void *stream);

/* Batched block-tridiagonal (BTDDOD) factor/solve over nsys independent systems of possibly different
 * block counts: system s has nblocks[s] blocks of size bs; its values start at values + voff[s] (griffon
 * layout: nblocks*bs*bs block entries, then (nblocks-1)*bs sub, then (nblocks-1)*bs sup), its L blocks at
 * l_values + loff[s], pivots at pivots + poff[s], vectors at offset xoff[s]. FP64 throughout. */
int gf_batch_btddod_full_factorize(int nsys, const int *nblocks, int bs, double *values, const long *voff,
                                   double *l_values, const long *loff, int *pivots, const long *poff, void *stream);
int gf_batch_btddod_full_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                               const double *l_values, const long *loff, const int *pivots, const long *poff,
                               const double *rhs, double *x, const long *xoff, void *stream);

/* The pseudo-time limiter's eigenvalue bound for many independent blocks, as flamelet_jacobian computes it per
 * point with compute_eigenvalues (flamelet_kernels.cpp:1329-1341): out[b] = max(max Re(lambda(B_b)) - diffterm, 0)
 * for nblk column-major n x n blocks stored one after another (device pointers; n <= 64). nfail (a device int,
 * may be null) is incremented for each block whose QR iteration did not converge. m may be null. */
int gf_batch_eigenvalue_bound(gf_mech *m, int nblk, int n, const double *blocks, double diffterm, double *out,
                              int *nfail, void *stream);

/* The same systems factored for throughput (the form the batched flamelet solver uses): the block Thomas
 * recurrence of btddod_full_factorize (btddod_matrix_kernels.cpp:19-80) run from both ends of each chain at
 * once, storing the inverse of every Schur-updated diagonal block instead of LU factors and L blocks.
 * dinv + doff[s] receives nblocks[s] column-major bs x bs inverse blocks, as FP64 (precision 64, dinv is
 * double *) or FP32 (precision 32, dinv is float *). The solve replaces btddod_full_solve
 * (btddod_matrix_kernels.cpp:82-119) for these factors; right-hand sides and solutions are FP64.
 * values are not modified; bs must be in 1..64. */
int gf_batch_btddod_inverse_factorize(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                                      void *dinv, const long *doff, int precision, void *stream);
int gf_batch_btddod_inverse_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                                  const void *dinv, const long *doff, int precision, const double *rhs, double *x,
                                  const long *xoff, void *stream);

/* ---------------- many steady flamelets at once, device-resident ------------------------------------ */
/* A batch of adiabatic flamelets sharing one mechanism and pressure, solved together on the GPU by the
 * reference's steady solvers: Newton's method (Flamelet.steady_solve_newton, flamelet.py:1352-1485) and
 * pseudo-transient continuation (Flamelet.steady_solve_psitc, flamelet.py:1487-1703), per flamelet with the
 * reference's Jacobian-refresh rules, eigenvalue-limited local pseudo-time steps, line search and restarts,
 * all flamelets advancing in lock step and every refresh, factorization, solve and RHS evaluation batched
 * over the flamelets that need it; plus the batched operators of the time-stepping fallback. Together they
 * replace a loop of Flamelet.compute_steady_state calls (flamelet.py:1705-1778) over the flamelets of a
 * library (tabulation.py:286-296, one flamelet per dissipation rate).
 *
 * Flamelet f has nzi[f] >= 2 interior points; dz holds its nzi[f]+1 grid spacings and chi its nzi[f]+2
 * dissipation rates (full grid, flamelet.py:458-466), concatenated over f; oxy/fuel hold the ns-long
 * boundary states [T, Y_0..Y_{ns-2}] of every flamelet. Unity Lewis numbers. Host arrays. */
typedef struct gf_flamelet_batch gf_flamelet_batch;
typedef struct gf_psitc_options {
  double tolerance;               /* 1e-6   */
  int max_iterations;             /* 400    */
  double ds_init;                 /* 1      */
  double ds_init_decrease;        /* 4      */
  int adaptive_restart;           /* 1      */
  int max_recursion_depth;        /* 20     */
  double diffusion_factor;        /* 4      */
  int global_ds;                  /* 0      */
  double ds_safety;               /* 0.1    */
  double ds_ramp;                 /* 1.1    */
  double ds_max;                  /* 1e4    */
  double max_factor_line_search;  /* 1.5    */
  double max_allowed_residual;    /* 1e6    */
  double min_allowable_state_var; /* -1e-6  */
  int lu_precision;               /* 64: FP64 block LU; 32: FP32 block LU factors (FP64 residuals) */
  int log_rate;                   /* 0: quiet; k: one progress line on stderr every k iterations */
} gf_psitc_options;
void gf_psitc_options_default(gf_psitc_options *opts);
int gf_flamelet_batch_create(gf_mech *m, int nfl, const int *nzi, double pressure, const double *dz,
                             const double *chi, const double *oxy, const double *fuel, int include_enthalpy_flux,
                             int include_variable_cp, gf_flamelet_batch **out);
int gf_flamelet_batch_destroy(gf_flamelet_batch *b);
/* Heat loss of nonadiabatic flamelets (the T_convection, h_convection, T_radiation, h_radiation and
 * use_scaled_heat_loss arguments of flamelet_rhs / flamelet_jacobian, griffon.pyx:580-679, for every flamelet of
 * the batch): host arrays of sum(nzi) values each, flamelet after flamelet over its interior points. Four nulls
 * make the batch adiabatic again (its initial setting). */
int gf_flamelet_batch_set_heat_loss(gf_flamelet_batch *b, const double *T_conv, const double *h_conv,
                                    const double *T_rad, const double *h_rad, int use_scaled_heat_loss);
/* Inverse Lewis numbers of every flamelet (host array [nfl][ns], flamelet f's at inv_lewis + f*ns; entry l scales
 * the diffusion stencil of equation l, the invLewisNumbers argument of flamelet_stencils, griffon.pyx /
 * flamelet_kernels.cpp:31-48). Null restores unity (the batch's initial setting). */
int gf_flamelet_batch_set_lewis_numbers(gf_flamelet_batch *b, const double *inv_lewis);
/* interior states, point-major per flamelet, concatenated: sum(nzi)*ns doubles; set_state also sets the
 * initial guess the restarts return to */
int gf_flamelet_batch_set_state(gf_flamelet_batch *b, const double *interior);
int gf_flamelet_batch_get_state(gf_flamelet_batch *b, double *interior);
/* flamelet RHS of every flamelet at the current state (same layout as the state) */
int gf_flamelet_batch_rhs(gf_flamelet_batch *b, double *out);
/* Pseudo-transient continuation (Flamelet.steady_solve_psitc, flamelet.py:1487-1703) of the flamelets with
 * select[f] != 0 (every flamelet when select is null), each from its initial guess. iterations/converged/
 * residual/min_ds: per flamelet (may be null; entries of unselected flamelets are left as they are); min_ds
 * is the smallest pseudo-time step at the end, the fourth value the reference returns and the first step
 * of its time-stepping fallback (flamelet.py:1757-1767). counters (may be null): [0] iterations summed over
 * flamelets, [1] Jacobian refreshes, [2] RHS evaluations, [3] synchronizing solver steps */
int gf_flamelet_batch_steady_psitc(gf_flamelet_batch *b, const gf_psitc_options *opts, const int *select,
                                   int *iterations, int *converged, double *residual, double *min_ds,
                                   long long *counters);
/* Newton's method with the reference's line search (Flamelet.steady_solve_newton, flamelet.py:1352-1485) for
 * the selected flamelets, each from its initial guess: the matrix -J is refreshed only after a line search
 * shortened the step; no restarts. Uses tolerance, max_iterations, max_factor_line_search,
 * max_allowed_residual, min_allowable_state_var, lu_precision and log_rate of opts. Outputs as for psitc. */
int gf_flamelet_batch_steady_newton(gf_flamelet_batch *b, const gf_psitc_options *opts, const int *select,
                                    int *iterations, int *converged, double *residual, long long *counters);
/* The operators of implicit time stepping (Flamelet.integrate with block-Thomas solves, flamelet.py:741-762,
 * 1046-1213) for the flamelets with flags[f] != 0, on DEVICE vectors in the dof layout of the batch (flamelet
 * f's interior state, point-major, from offset sum_{g<f} nzi[g]*ns; NV = sum nzi*ns doubles); entries of the
 * other flamelets are neither read nor written. Synchronous: the work is complete on return.
 *   transient_rhs:   rhs = F(x); x also becomes the current state of those flamelets
 *   transient_setup: factor prefactor[f] * J(x) - I per flagged flamelet (prefactor: host, one per flamelet);
 *                    lu_precision 64 or 32, the same for every setup while factors are held
 *   transient_solve: out = (prefactor[f] * J - I)^-1 r with each flamelet's factors from its last
 *                    transient_setup (an error if it has none, or a steady solve refreshed it since) */
int gf_flamelet_batch_transient_rhs(gf_flamelet_batch *b, const int *flags, const double *x, double *rhs);
int gf_flamelet_batch_transient_setup(gf_flamelet_batch *b, const int *flags, const double *x,
                                      const double *prefactor, int lu_precision);
int gf_flamelet_batch_transient_solve(gf_flamelet_batch *b, const int *flags, const double *r, double *out);

#ifdef __cplusplus
}
#endif
#endif
