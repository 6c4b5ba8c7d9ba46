// Internal declarations shared by the griffon_amd host code and HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstdint>

namespace gfa {

constexpr int NS_MAX = 64;     // species-array stride in the wave-per-state kernels (one lane per species)
constexpr int NS_LIMIT = 63;   // species per mechanism (the Jacobian pass uses one extra lane for the last species)
constexpr int NRC_MAX = 4;     // reactants per reaction
constexpr int NPD_MAX = 4;     // products per reaction
constexpr int NNET_MAX = 8;    // net species per reaction (griffon: 2..8)
constexpr int NSP_MAX = 4;     // species with special (non-elementary) orders per reaction

enum { CP_CONST = 1, CP_NASA7 = 2 };
enum { RK_SIMPLE = 1, RK_THIRD_BODY = 2, RK_LINDEMANN = 3, RK_TROE = 4 };
enum { KF_CONSTANT = 0, KF_LINEAR, KF_QUADRATIC, KF_RECIPROCAL, KF_ARRHENIUS };
enum { TF_NONE = 0, TF_T1, TF_T2, TF_T12, TF_T3, TF_T13, TF_T23, TF_T123 };

// One reaction as the kernels read it (48 + 88 bytes). Indices are species indices; stoichiometric
// coefficients are positive for both sides. `slot` lists map each participant to its slot in the
// per-state derivative vector (-1: the last species, which enters only through uniform corrections).
struct DevRxn {
  double kf[3];
  double kp[3];
  double troe[4];
  double baseEff;
  int type, kForm, troeForm;
  int reversible, hasOrders;
  int n_rc, n_pd, n_net, n_tb, n_sp;
  int sumStoich, sumRc, sumPd;
  int tb_off, sp_off, slot_off, n_slot;
  int orig;                       // the reaction's index in the mechanism (device order groups like reactions)
  int8_t tb_last;                 // index into the tb list of the last species, or -1
  int8_t pad[3];
  uint8_t rc_idx[NRC_MAX], rc_st[NRC_MAX];
  uint8_t pd_idx[NPD_MAX], pd_st[NPD_MAX];
  uint8_t net_idx[NNET_MAX];
  int8_t net_st[NNET_MAX];
  int8_t rc_slot[NRC_MAX], pd_slot[NPD_MAX];
  int8_t sp_slot[NSP_MAX];
};

// Device view of a finalized mechanism (all pointers into one device allocation).
struct MechView {
  int ns, nr, nslot;
  int last_const;           // highest constant-cp species index (-1: none); see cp_sens_T quirk
  double Ru, invRu, pref;
  const double *spc_c;      // [ns*16] griffon-packed NASA7/const coefficients (chemistry_setup.cpp:89-130)
  const double *invmw;      // [ns]
  const double *mw;         // [ns]
  const double *dinvm;      // [ns] invmw[s] - invmw[ns-1]
  const double *Tmin, *Tmax;
  const int *cpkind;        // [ns]
  const DevRxn *rx;         // [nr]
  const int *tb_idx;        // third-body lists
  const double *tb_eff;     //   invmw*(eff - default)  (chemistry_setup.cpp:416)
  const int *tb_slot;       //   slot of each third body (-1 for the last species)
  const int *sp_idx;        // special-order lists
  const double *sp_ord;
  const double *sp_invmw;
  // sensitivity entry (species row i < ns, species column j < ns-1) <- slots: entry e = j*ns + i
  // (drop-in raw sensitivities only)
  const int *eg_off;        // [(ns-1)*ns+1]
  const int *eg_slot;
  const double *eg_coef;
  // Compact gather tables, copied into LDS by the chemistry kernel (byte image at tbl, tbl_bytes long):
  //   u32 sgx_off[ns+1]; u64 sgx[n_sgx]   species k <- reactions with k in their net stoichiometry:
  //       bits 0-15 reaction, 16-31 slot of species k in that reaction (0xFFFF: none), 32-39 net
  //       stoichiometric coefficient nu (int8); the gather coefficient is -nu * M_k
  //   f64 tbe[n_tb]; i32 tbi[n_tb]; i32 tbs[n_tb]; f64 imw[ns]   the third-body lists (tb_eff, tb_idx, tb_slot)
  //       and 1/M, read by the reactions phase (placed before exx_off, so the RHS-only kernel copies them too)
  //   u16 exx_off[(ns-1)*ns+1]; u32 exx[n_exx]   off-diagonal sensitivity entry (i, j), e = j*ns + i, <-
  //       slots of species j: bits 0-15 slot, 16-23 nu of species i (int8); coefficient -nu * M_i
  const char *tbl;
  int tbl_bytes;
  int sgx_at, exo_at, exx_at;  // byte offsets of sgx, exx_off and exx inside the table image
  int tbe_at, tbi_at, tbs_at, imw_at;  // byte offsets of the third-body lists and 1/M
};

struct Species {
  double mw = 0.;
  int cpkind = 0;
  double Tmin = 0., Tmax = 0.;
  double c[16] = {0};
};

struct HostRxn {
  int type = RK_SIMPLE;
  bool reversible = true;
  std::vector<int> rc_idx, rc_st, pd_idx, pd_st, tb_idx, sp_idx;
  std::vector<double> tb_eff_raw, sp_ord;
  double default_eff = 0.;
  double kf[3] = {0, 0, 0}, kp[3] = {0, 0, 0}, troe[4] = {0, 0, 0, 0};
};

}  // namespace gfa

struct gf_mech {
  double Ru = 8314.46261815324;
  double pref = 101325.;
  double Tref = 298.15;
  std::vector<gfa::Species> species;
  std::vector<gfa::HostRxn> rxns;
  // finalized host tables
  bool finalized = false;
  std::vector<gfa::DevRxn> drx;
  std::vector<char> blob;         // host image of the device tables
  gfa::MechView view_host;        // offsets (pointers relative to blob base)
  int device = -1;
  void *dblob = nullptr;          // device copy
  gfa::MechView view;             // device pointers
  size_t lds_per_state = 0;       // bytes of LDS the Jacobian chemistry kernel needs per state
  size_t lds_per_state_rhs = 0;   // the same for the RHS-only kernel
  int n_cu = 0;                   // compute units of the device (persistent launch width)
  size_t lds_max = 65536;         // largest LDS allocation one workgroup may use on this device
  // chemistry launch shape picked by timing, per [Jacobian or RHS-only kernel][floor(log2(states))]; -1: not
  // yet timed (launch_chem, gf_chem.hip)
  mutable signed char chem_shape[2][64] = {};
  mutable bool chem_shape_init = false;
  // staging buffers for the host-pointer drop-ins
  void *stage = nullptr;
  size_t stage_bytes = 0;
  // device workspace of the batched entry points that pass partial results between kernels
  void *work = nullptr;
  size_t work_bytes = 0;
  hipStream_t stream = nullptr;
};

namespace gfa {
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int ensure_device(gf_mech *m);                 // finalize + upload on first compute call
int ensure_stage(gf_mech *m, size_t bytes);
int ensure_work(gf_mech *m, size_t bytes);   // grows m->work (synchronous hipFree/hipMalloc)

// Inputs/outputs of one batched chemistry launch (gf_chem.hip). Element v of state i is x[i*xs_i + v*xs_v];
// p[i*ps] is its pressure (or its density when rho_given, in which case x holds [T, y_0..y_{ns-1}]).
struct ChemIO {
  int n = 0;
  const double *x = nullptr;
  long xs_i = 0, xs_v = 0;
  const double *p = nullptr;
  long ps = 0;
  int rho_given = 0;       // 1: p holds densities and x [T, y_0..y_{ns-1}]; 2: densities, x [T, y_0..y_{ns-2}]
  int notbaf = 0;          // sensitivities without third-body / falloff multiplier derivatives (option 1)
  int no_tune = 0;         // never time launch shapes: use the tuned shape if there is one, else the default
  double *rhs = nullptr;   // rhs[i*rs_i + v*rs_v]
  long rs_i = 0, rs_v = 0;
  double *jac = nullptr;   // ns*ns column-major block at jac + i*js_i
  long js_i = 0;
  double *wout = nullptr;  // net mass production rates [n][ns]
  double *sens = nullptr;  // raw option-0 sensitivities [n][(ns+1)^2]
  double *aux = nullptr;   // per state {rho, cp, dcp/dT, mixture molecular weight}
  double *cpi = nullptr;   // per state species heat capacities [n][ns]
  double *hsp = nullptr;   // per state species enthalpies per unit mass [n][ns]
  double *qa = nullptr;    // per state d(rate of progress)/d(ln A) of every reaction [n][nr] (Jacobian kernel)
  // optional: logical state i reads and writes physical row sel[i] of every array above
  const int *sel = nullptr;
  // optional: the Jacobian block of physical row q starts at jac + js_off[q] (no block if negative)
  const long *js_off = nullptr;
};
// Reactor RHS/Jacobian assembly beyond the closed adiabatic isobaric case (gf_reactor_ext.hip), from the chemistry
// kernel's per-state outputs: w [n][ns], raw sensitivities [n][(ns+1)^2], aux [n][4] = {rho, cp, dcp/dT, M},
// species cp and h [n][ns]. State s is read at state + s*state_stride: isobaric [T, Y_0..Y_{ns-2}], isochoric
// [rho, T, Y_0..Y_{ns-2}]. Outputs rhs [n][neq] and jac [n][neq*neq] column-major (neq = ns, or ns+1 isochoric).
struct ReactorExtIO {
  int n = 0;
  int isochoric = 0;
  const double *state = nullptr;
  long state_stride = 0;
  const double *w = nullptr, *sens = nullptr, *aux = nullptr, *cpi = nullptr, *hsp = nullptr;
  int heat = 0, open = 0;          // heat transfer 0 adiabatic, 1 isothermal, 2 diathermal; open: mass exchange
  double Tin = 0., tau = 1., rho_in = 0.;
  const double *yin = nullptr;     // [ns] inflow mass fractions (device)
  double Tf = 0., Ts = 0., hc = 0., eps = 0., sov = 0.;
  double *rhs = nullptr, *jac = nullptr;
};
int launch_reactor_ext(const gf_mech *m, const ReactorExtIO &a, hipStream_t st);
// gf_batch_reactor_general's work (gf_capi.cpp): doubles of workspace per state, argument checks, and the two
// launches on a given workspace
size_t batch_work_doubles(int ns, bool jac);
int check_batch_general(const gf_mech *m, int n, int isochoric, const double *x, long x_item_stride,
                        const double *p, const double *inflow_y, int heat, int open, int rates_opt, double *rhs);
int batch_general(const gf_mech *m, int n, int isochoric, const double *x, long x_item_stride, const double *p,
                  long p_stride, double inflow_density, double inflow_T, const double *inflow_y, double tau,
                  double fluid_T, double surf_T, double h_conv, double eps_rad, double sov, int heat, int open,
                  int rates_opt, double *rhs, double *jac, double *work, int no_tune, hipStream_t st);
// griffon's thermodynamic getters for in = [T, y_0..y_{ns-1}] (device): out = {M, rho(p), p(rho), cp, cv, h, e,
// dcp/dT, x[ns], cp[ns], cv[ns], h[ns], e[ns], dcp/dT[ns]}
int launch_thermo_getters(const gf_mech *m, const double *in, double p, double rho, double *out, hipStream_t st);
int launch_mixture_properties(const gf_mech *m, int n, const double *Ty, const double *p, const double *rho,
                              double *out, hipStream_t st);
int set_chem_launch_shape(int shape);  // -1: timed per launch kind and size class; 0..4: fixed
int launch_chem(const gf_mech *m, const ChemIO &io, hipStream_t st);

// flamelet kernels (gf_flamelet.hip)
struct FlameletIO {
  int nzi = 0, ns = 0;
  const double *state = nullptr;   // [nzi*ns] interior state (point-major)
  const double *oxy = nullptr, *fuel = nullptr;
  const double *cmaj = nullptr, *csub = nullptr, *csup = nullptr, *mco = nullptr, *nco = nullptr, *chi = nullptr;
  int adiabatic = 1;
  const double *Tc = nullptr, *hc = nullptr, *Tr = nullptr, *hr = nullptr;  // heat loss (C++ order)
  int incl_ef = 0, incl_vcp = 0, scaled_hl = 0;
  // per-point chemistry outputs over nzi+2 points (oxidizer, interior..., fuel)
  const double *chem_rhs = nullptr;  // [(nzi+2)*ns]
  const double *aux = nullptr;       // [(nzi+2)*4]
  const double *cpi = nullptr;       // [(nzi+2)*ns]
};
int launch_flamelet_rhs_epilogue(const FlameletIO &f, double *out, hipStream_t st);
int launch_flamelet_jac_epilogue(const gf_mech *m, const FlameletIO &f, int scale_and_offset, double prefactor,
                                 double *jac, int phase, hipStream_t st);
// per-block largest real eigenvalue part, max(0, max Re(lambda) - diffterm) per block row (gf_eig.hip)
int launch_expeig(const gf_mech *m, const double *blocks, long bstride, long nblk, int n, double diffterm, double *out,
                  int *nfail, hipStream_t st);
int launch_flamelet_stencils(int ns, const double *dz, int nzi, const double *chi, const double *invLe, double *cmaj,
                             double *csub, double *csup, double *mco, double *nco, hipStream_t st);
int launch_flamelet_jac_indices(int ns, int nzi, int *rows, int *cols, hipStream_t st);
int launch_scale_and_add_diagonal(double *v, double a, const double *dg, double b, int nb, int bs, hipStream_t st);
// Many flamelets at once (gf_flamelet.hip kernels, gf_fbatch.cpp driver). Flamelet f has nzi[f] interior
// points; its points [oxidizer, interior..., fuel] are rows pt0[f] .. pt0[f]+nzi[f]+1 of the point arrays
// (X, chem_rhs, cpi: ns per point; aux: 4 per point); its interior unknowns start at dof0[f] in the dof
// arrays (cmaj, csub, csup, R, dX, ds, expeig), its interior points at ipt0[f] (mco, nco), its full grid at
// chi0[f] (chi), and its BTDDOD Jacobian at voff[f].
struct FlameletBatchView {
  int B = 0, ns = 0, incl_ef = 0, incl_vcp = 0;
  const int *nzi = nullptr;
  const long *pt0 = nullptr, *dof0 = nullptr, *ipt0 = nullptr, *chi0 = nullptr, *voff = nullptr;
  const double *cmaj = nullptr, *csub = nullptr, *csup = nullptr, *mco = nullptr, *nco = nullptr, *chi = nullptr;
  const double *chem_rhs = nullptr, *aux = nullptr, *cpi = nullptr;
  // heat loss (nonadiabatic flamelets): hl[k * NI + ipt0[f] + i] for k = T_conv, h_conv, T_rad, h_rad
  int adiabatic = 1, scaled_hl = 0;
  const double *hl = nullptr;
  long NI = 0;          // total interior points
  long NV = 0, NJ = 0;  // total unknowns, total Jacobian entries
  long maxNJ = 0;       // Jacobian entries of the largest flamelet
};
int launch_fb_stencils(const FlameletBatchView &b, const double *dz, const long *dz0, const double *invLe,
                       double *cmaj, double *csub, double *csup, double *mco, double *nco, hipStream_t st);
int launch_fb_rhs(const FlameletBatchView &b, const double *X, const int *flag, double *R, hipStream_t st);
// mode 0: diag(1/ds) - J (pseudo-transient, vals = ds); 1: -J (Newton); 2: prefactor * J - I (implicit time
// stepping, vals = per-flamelet prefactors, or the scalar prefactor when null)
int launch_fb_matrix(const gf_mech *m, const FlameletBatchView &b, const double *X, const int *flag, int mode,
                     const double *vals, double prefactor, double *J, hipStream_t st);
// flagged flamelets' unknowns from a vector holding flamelet f's unknowns from pk[f] into the dof layout, or
// into the padded interior rows of X when interior is set
int launch_fb_unpack(const FlameletBatchView &b, const int *flag, const long *pk, const double *packed, double *dst,
                     int interior, hipStream_t st);
int launch_fb_norms(const FlameletBatchView &b, const int *flag, const double *R, const double *X, double *res,
                    double *xmin, hipStream_t st);
int launch_fb_trial(const FlameletBatchView &b, const int *flag, const double *scale, const double *X, double *Xt,
                    double *dX, hipStream_t st);
int launch_fb_copy_interior(const FlameletBatchView &b, const int *flag, const double *src, double *dst,
                            hipStream_t st);
int launch_fb_copy_dofs(const FlameletBatchView &b, const int *flag, const double *src, double *dst, hipStream_t st);
int launch_fb_ds(const FlameletBatchView &b, const int *flag, const int *global_min, const double *expeig,
                 const double *diffterm, double safety, double ramp, double dsmax, double *ds, hipStream_t st);
int launch_fb_fill_dofs(const FlameletBatchView &b, const int *flag, const double *value, double *dst,
                        hipStream_t st);
int launch_fb_clip_points(long n, double *X, hipStream_t st);
// kexp (optional, laid out like out with nout = n) carries each block's balancing exponents between launches
int launch_expeig_list(const gf_mech *m, const double *blocks, const long *boff, const long *ooff, long nblk, int n,
                       double diffterm, double *out, int *nfail, hipStream_t st, int nout = -1,
                       float *kexp = nullptr);
// block-tridiagonal kernels (gf_btd.hip)
int launch_btd_factor(int nsys, const int *nblocks, int bs, double *values, const long *voff, double *l_values,
                      const long *loff, int *pivots, const long *poff, hipStream_t st);
int launch_btd_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                     const double *l_values, const long *loff, const int *pivots, const long *poff,
                     const double *rhs, double *x, const long *xoff, hipStream_t st);
// the same with the D factors in dfac + doff[s] (FP64 in place: dfac = values, doff = voff) and factor type T
template <typename T>
int launch_btd_factor_t(int nsys, const int *nblocks, int bs, double *values, const long *voff, T *dfac,
                        const long *doff, T *l_values, const long *loff, int *pivots, const long *poff,
                        hipStream_t st);
// block Thomas with explicit diagonal-block inverses D_i^{-1} (Gauss-Jordan), stored at dinv + doff[s]
// (bs*bs per block); the values array supplies the blocks and the off-diagonals
template <typename T>
int launch_btd_inv_factor(int nsys, const int *nblocks, int bs, const double *values, const long *voff, T *dinv,
                          const long *doff, hipStream_t st);
template <typename T>
int launch_btd_inv_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                         const T *dinv, const long *doff, const double *rhs, double *x, const long *xoff,
                         hipStream_t st);
template <typename T>
int launch_btd_solve_t(int nsys, const int *nblocks, int bs, const double *values, const long *voff, const T *dfac,
                       const long *doff, const T *l_values, const long *loff, const int *pivots, const long *poff,
                       const double *rhs, double *x, const long *xoff, hipStream_t st);
}  // namespace gfa
