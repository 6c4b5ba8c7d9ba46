// extern "C" entry points of griffon_amd: the host-pointer drop-ins for the reference's Cython boundary
// (griffon.pyx) and the batched device-pointer entry points. Every compute path launches the HIP kernels
// in gf_chem.hip / gf_btd.hip / gf_flamelet.hip; there is no host arithmetic fallback.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"

#include <cstring>
#include <vector>

using namespace gfa;

namespace {

// bump allocator over the mechanism's staging buffer
struct Stage {
  char *base = nullptr;
  size_t off = 0;
  template <typename T>
  T *take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T *p = reinterpret_cast<T *>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

size_t rup(size_t b) { return (b + 255) & ~size_t(255); }

int h2d(void *d, const void *h, size_t bytes, hipStream_t st) {
  if (!bytes) return GF_OK;
  if (!h) return fail(GF_ERR_ARG, "null input array");
  return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? GF_OK
                                                                             : fail(GF_ERR_HIP, "host-to-device copy failed");
}

int d2h(void *h, const void *d, size_t bytes, hipStream_t st) {
  if (!bytes) return GF_OK;
  if (!h) return fail(GF_ERR_ARG, "null output array");
  return hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st) == hipSuccess ? GF_OK
                                                                             : fail(GF_ERR_HIP, "device-to-host copy failed");
}

int sync(hipStream_t st) {
  const hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(GF_ERR_HIP, std::string("stream synchronize failed: ") + hipGetErrorString(e));
  return GF_OK;
}

#define GF_TRY(expr)                 \
  do {                               \
    const int _rc = (expr);          \
    if (_rc != GF_OK) return _rc;    \
  } while (0)

int prep(gf_mech *m, size_t bytes, Stage &s) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  GF_TRY(ensure_device(m));
  GF_TRY(ensure_stage(m, bytes + 4096));
  s.base = static_cast<char *>(m->stage);
  s.off = 0;
  return GF_OK;
}

int check_options(const gf_mech *m, int heat, int open, int rates_opt, int transform_opt) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  if (heat < 0 || heat > 2) return fail(GF_ERR_ARG, "heat_transfer_option must be 0, 1 or 2");
  (void)open;
  if (rates_opt < 0 || rates_opt > 2) return fail(GF_ERR_ARG, "rates_sensitivity_option must be 0, 1 or 2");
  if (rates_opt == 1)
    for (const auto &r : m->drx)
      if (r.hasOrders)
        return fail(GF_ERR_UNSUPPORTED, "rates_sensitivity_option 1 with special reaction orders is not supported");
  if (transform_opt != 0) return fail(GF_ERR_UNSUPPORTED, "only sensitivity_transform_option 0 exists");
  return GF_OK;
}

// Reactor RHS (and Jacobian) through the chemistry kernel's per-state outputs and the reactor assembly kernel:
// open and diathermal terms, option-1 sensitivities and the isochoric reactor.
int reactor_general(gf_mech *m, bool isochoric, const double *state, double p_or_rho_in, double Tin, const double *yin,
                    double tau, double Tf, double Ts, double hc, double eps, double sov, int heat, int open,
                    int rates_opt, double *out_rhs, double *out_jac) {
  Stage s;
  const int ns = gf_mech_n_species(m);
  const int n = isochoric ? ns + 1 : ns;
  const size_t nsens = (size_t)(ns + 1) * (ns + 1);
  const bool J = out_jac != nullptr;
  GF_TRY(prep(m, rup((ns + 1) * 8) + rup(8) + 4 * rup(ns * 8) + rup(4 * 8) + rup(n * 8) + (J ? rup(nsens * 8) + rup((size_t)n * n * 8) : 0), s));
  double *dst = s.take<double>(ns + 1), *dp = s.take<double>(1), *dw = s.take<double>(ns), *dcp = s.take<double>(ns),
         *dh = s.take<double>(ns), *dyin = s.take<double>(ns), *daux = s.take<double>(4), *dr = s.take<double>(n);
  double *dsens = J ? s.take<double>(nsens) : nullptr, *dj = J ? s.take<double>((size_t)n * n) : nullptr;
  hipStream_t st = m->stream;
  GF_TRY(h2d(dst, state, n * 8, st));
  if (!isochoric) GF_TRY(h2d(dp, &p_or_rho_in, 8, st));
  if (open) GF_TRY(h2d(dyin, yin, ns * 8, st));
  ChemIO io;
  io.n = 1;
  if (isochoric) {  // [rho, T, Y...]: the chemistry reads [T, Y...] with the density given
    io.x = dst + 1; io.p = dst; io.rho_given = 2;
  } else {
    io.x = dst; io.p = dp;
  }
  io.xs_i = n; io.xs_v = 1; io.ps = 0;
  io.wout = dw; io.aux = daux; io.cpi = dcp; io.hsp = dh;
  io.sens = dsens;
  io.notbaf = rates_opt == 1;
  GF_TRY(launch_chem(m, io, st));
  ReactorExtIO a;
  a.n = 1; a.isochoric = isochoric; a.state = dst; a.state_stride = n;
  a.w = dw; a.sens = dsens; a.aux = daux; a.cpi = dcp; a.hsp = dh;
  a.heat = heat; a.open = open; a.Tin = Tin; a.tau = tau; a.rho_in = isochoric ? p_or_rho_in : 0.; a.yin = dyin;
  a.Tf = Tf; a.Ts = Ts; a.hc = hc; a.eps = eps; a.sov = sov;
  a.rhs = dr; a.jac = dj;
  GF_TRY(launch_reactor_ext(m, a, st));
  GF_TRY(d2h(out_rhs, dr, n * 8, st));
  if (J) GF_TRY(d2h(out_jac, dj, (size_t)n * n * 8, st));
  return sync(st);
}

// all getters of one (T, y): out = {M, rho(p), p(rho), cp, cv, h, e, dcp/dT, x, cp_i, cv_i, h_i, e_i, dcp_i/dT}
int thermo_all(gf_mech *m, double T, const double *y, double p, double rho, double *host_out) {
  Stage s;
  const int ns = gf_mech_n_species(m);
  const size_t nout = 8 + 6 * (size_t)ns;
  GF_TRY(prep(m, rup((ns + 1) * 8) + rup(nout * 8), s));
  double *din = s.take<double>(ns + 1), *dout = s.take<double>(nout);
  hipStream_t st = m->stream;
  GF_TRY(h2d(din, &T, 8, st));
  if (y) {
    GF_TRY(h2d(din + 1, y, ns * 8, st));
  } else if (hipMemsetAsync(din + 1, 0, ns * 8, st) != hipSuccess) {
    return fail(GF_ERR_HIP, "memset failed");
  }
  GF_TRY(launch_thermo_getters(m, din, p, rho, dout, st));
  GF_TRY(d2h(host_out, dout, nout * 8, st));
  return sync(st);
}

}  // namespace

namespace gfa {

// Batched general reactor: chemistry (rates, raw sensitivities, thermo) into `work`, then the reactor assembly.
// work holds n * batch_work_doubles(ns, jac) doubles.
size_t batch_work_doubles(int ns, bool jac) {
  return (size_t)ns + (jac ? (size_t)(ns + 1) * (ns + 1) : 0) + 4 + 2 * (size_t)ns;
}

int batch_general(const gf_mech *m, int n, int isochoric, const double *x, long x_item_stride, const double *p,
                  long p_stride, double inflow_density, double inflow_T, const double *inflow_y, double tau,
                  double fluid_T, double surf_T, double h_conv, double eps_rad, double sov, int heat, int open,
                  int rates_opt, double *rhs, double *jac, double *work, int no_tune, hipStream_t st) {
  const int ns = m->view.ns;
  double *w = work, *cpi = w + (size_t)n * ns, *hsp = cpi + (size_t)n * ns, *aux = hsp + (size_t)n * ns;
  double *sens = jac ? aux + (size_t)n * 4 : nullptr;
  ChemIO io;
  io.n = n;
  if (isochoric) {
    io.x = x + 1; io.p = x; io.ps = x_item_stride; io.rho_given = 2;
  } else {
    io.x = x; io.p = p; io.ps = p_stride;
  }
  io.xs_i = x_item_stride; io.xs_v = 1;
  io.wout = w; io.aux = aux; io.cpi = cpi; io.hsp = hsp; io.sens = sens;
  io.notbaf = rates_opt == 1;
  io.no_tune = no_tune;
  GF_TRY(launch_chem(m, io, st));
  ReactorExtIO a;
  a.n = n; a.isochoric = isochoric; a.state = x; a.state_stride = x_item_stride;
  a.w = w; a.sens = sens; a.aux = aux; a.cpi = cpi; a.hsp = hsp;
  a.heat = heat; a.open = open; a.Tin = inflow_T; a.tau = tau; a.rho_in = inflow_density; a.yin = inflow_y;
  a.Tf = fluid_T; a.Ts = surf_T; a.hc = h_conv; a.eps = eps_rad; a.sov = sov;
  a.rhs = rhs; a.jac = jac;
  return launch_reactor_ext(m, a, st);
}

int check_batch_general(const gf_mech *m, int n, int isochoric, const double *x, long x_item_stride,
                        const double *p, const double *inflow_y, int heat, int open, int rates_opt, double *rhs) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  if (n < 0) return fail(GF_ERR_ARG, "negative number of states");
  if (n > 0 && (!x || !rhs || (!isochoric && !p) || (open && !inflow_y)))
    return fail(GF_ERR_ARG, "null device array");
  const int neq = m->view.ns + (isochoric ? 1 : 0);
  if (x_item_stride < neq) return fail(GF_ERR_ARG, "state item stride is smaller than the state size");
  return check_options(m, heat, open, rates_opt, 0);
}

}  // namespace gfa

extern "C" {

int gf_production_rates(gf_mech *m, double T, double rho, const double *y, double *out_w) {
  Stage s;
  const int ns = gf_mech_n_species(m);
  GF_TRY(prep(m, rup((ns + 1) * 8) + rup(8) + rup(ns * 8), s));
  double *dx = s.take<double>(ns + 1), *dp = s.take<double>(1), *dw = s.take<double>(ns);
  hipStream_t st = m->stream;
  GF_TRY(h2d(dx, &T, 8, st));
  GF_TRY(h2d(dx + 1, y, ns * 8, st));
  GF_TRY(h2d(dp, &rho, 8, st));
  ChemIO io;
  io.n = 1; io.x = dx; io.xs_i = ns + 1; io.xs_v = 1; io.p = dp; io.ps = 0; io.rho_given = 1; io.wout = dw;
  GF_TRY(launch_chem(m, io, st));
  GF_TRY(d2h(out_w, dw, ns * 8, st));
  return sync(st);
}

int gf_prod_rates_primitive_sensitivities(gf_mech *m, double rho, double T, const double *y, int option,
                                          double *out_sens) {
  GF_TRY(check_options(m, 0, 0, option, 0));
  Stage s;
  const int ns = gf_mech_n_species(m);
  const size_t nsens = (size_t)(ns + 1) * (ns + 1);
  GF_TRY(prep(m, rup((ns + 1) * 8) + rup(8) + rup(ns * 8) + rup(nsens * 8), s));
  double *dx = s.take<double>(ns + 1), *dp = s.take<double>(1), *dw = s.take<double>(ns),
         *dsens = s.take<double>(nsens);
  hipStream_t st = m->stream;
  GF_TRY(h2d(dx, &T, 8, st));
  GF_TRY(h2d(dx + 1, y, ns * 8, st));
  GF_TRY(h2d(dp, &rho, 8, st));
  ChemIO io;
  io.n = 1; io.x = dx; io.xs_i = ns + 1; io.xs_v = 1; io.p = dp; io.ps = 0; io.rho_given = 1; io.wout = dw;
  io.sens = dsens;
  io.notbaf = option == 1;
  GF_TRY(launch_chem(m, io, st));
  GF_TRY(d2h(out_sens, dsens, nsens * 8, st));
  return sync(st);
}

int gf_reactor_rhs_isobaric(gf_mech *m, const double *state, double pressure, double inflow_T, const double *inflow_y,
                            double tau, double fluid_T, double surf_T, double h_conv, double eps_rad, double sov,
                            int heat, int open, double *out_rhs) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  GF_TRY(check_options(m, heat, open, 0, 0));
  if (open || heat == 2)
    return reactor_general(m, false, state, pressure, inflow_T, inflow_y, tau, fluid_T, surf_T, h_conv, eps_rad, sov,
                           heat, open, 0, out_rhs, nullptr);
  Stage s;
  const int ns = gf_mech_n_species(m);
  GF_TRY(prep(m, 2 * rup(ns * 8) + rup(8), s));
  double *dx = s.take<double>(ns), *dp = s.take<double>(1), *dr = s.take<double>(ns);
  hipStream_t st = m->stream;
  GF_TRY(h2d(dx, state, ns * 8, st));
  GF_TRY(h2d(dp, &pressure, 8, st));
  ChemIO io;
  io.n = 1; io.x = dx; io.xs_i = ns; io.xs_v = 1; io.p = dp; io.ps = 0; io.rhs = dr; io.rs_i = ns; io.rs_v = 1;
  GF_TRY(launch_chem(m, io, st));
  if (heat == 1 && hipMemsetAsync(dr, 0, 8, st) != hipSuccess) return fail(GF_ERR_HIP, "memset failed");
  GF_TRY(d2h(out_rhs, dr, ns * 8, st));
  return sync(st);
}

int gf_reactor_jac_isobaric(gf_mech *m, const double *state, double pressure, double inflow_T, const double *inflow_y,
                            double tau, double fluid_T, double surf_T, double h_conv, double eps_rad, double sov,
                            int heat, int open, int rates_opt, int transform_opt, double *out_rhs, double *out_jac) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  GF_TRY(check_options(m, heat, open, rates_opt, transform_opt));
  if (open || heat == 2 || rates_opt == 1)
    return reactor_general(m, false, state, pressure, inflow_T, inflow_y, tau, fluid_T, surf_T, h_conv, eps_rad, sov,
                           heat, open, rates_opt, out_rhs, out_jac);
  Stage s;
  const int ns = gf_mech_n_species(m);
  GF_TRY(prep(m, 2 * rup(ns * 8) + rup(8) + rup((size_t)ns * ns * 8), s));
  double *dx = s.take<double>(ns), *dp = s.take<double>(1), *dr = s.take<double>(ns), *dj = s.take<double>(ns * ns);
  hipStream_t st = m->stream;
  GF_TRY(h2d(dx, state, ns * 8, st));
  GF_TRY(h2d(dp, &pressure, 8, st));
  ChemIO io;
  io.n = 1; io.x = dx; io.xs_i = ns; io.xs_v = 1; io.p = dp; io.ps = 0; io.rhs = dr; io.rs_i = ns; io.rs_v = 1;
  io.jac = dj; io.js_i = (long)ns * ns;
  GF_TRY(launch_chem(m, io, st));
  if (heat == 1) {  // isothermal: temperature row of the primitive Jacobian and the RHS vanish
    if (hipMemsetAsync(dr, 0, 8, st) != hipSuccess ||
        hipMemset2DAsync(dj, (size_t)ns * 8, 0, 8, ns, st) != hipSuccess)
      return fail(GF_ERR_HIP, "memset failed");
  }
  GF_TRY(d2h(out_rhs, dr, ns * 8, st));
  GF_TRY(d2h(out_jac, dj, (size_t)ns * ns * 8, st));
  return sync(st);
}

int gf_reactor_rhs_isochoric(gf_mech *m, const double *state, double inflow_density, double inflow_T,
                             const double *inflow_y, double tau, double fluid_T, double surf_T, double h_conv,
                             double eps_rad, double sov, int heat, int open, double *out_rhs) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  GF_TRY(check_options(m, heat, open, 0, 0));
  return reactor_general(m, true, state, inflow_density, inflow_T, inflow_y, tau, fluid_T, surf_T, h_conv, eps_rad,
                         sov, heat, open, 0, out_rhs, nullptr);
}

int gf_reactor_jac_isochoric(gf_mech *m, const double *state, double inflow_density, double inflow_T,
                             const double *inflow_y, double tau, double fluid_T, double surf_T, double h_conv,
                             double eps_rad, double sov, int heat, int open, int rates_opt, double *out_rhs,
                             double *out_jac) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  GF_TRY(check_options(m, heat, open, rates_opt, 0));
  if (!out_jac) return fail(GF_ERR_ARG, "null output array");
  return reactor_general(m, true, state, inflow_density, inflow_T, inflow_y, tau, fluid_T, surf_T, h_conv, eps_rad,
                         sov, heat, open, rates_opt, out_rhs, out_jac);
}

#define GF_THERMO_GETTER(BODY)                                             \
  if (!m) return fail(GF_ERR_ARG, "null mechanism");                     \
  const int ns = gf_mech_n_species(m);                                   \
  std::vector<double> o(8 + 6 * (size_t)ns);                             \
  BODY

int gf_mixture_molecular_weight(gf_mech *m, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, 300., y, 0., 0., o.data())); *out_value = o[0]; return GF_OK;)
}
int gf_mole_fractions(gf_mech *m, const double *y, double *out_x) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, 300., y, 0., 0., o.data()));
                   std::memcpy(out_x, o.data() + 8, ns * 8); return GF_OK;)
}
int gf_ideal_gas_density(gf_mech *m, double p, double T, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, p, 0., o.data())); *out_value = o[1]; return GF_OK;)
}
int gf_ideal_gas_pressure(gf_mech *m, double rho, double T, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, 0., rho, o.data())); *out_value = o[2]; return GF_OK;)
}
int gf_cp_mix(gf_mech *m, double T, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, 0., 0., o.data())); *out_value = o[3]; return GF_OK;)
}
int gf_cv_mix(gf_mech *m, double T, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, 0., 0., o.data())); *out_value = o[4]; return GF_OK;)
}
int gf_species_cp(gf_mech *m, double T, double *out_cp) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, nullptr, 0., 0., o.data()));
                   std::memcpy(out_cp, o.data() + 8 + ns, ns * 8); return GF_OK;)
}
int gf_species_cv(gf_mech *m, double T, double *out_cv) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, nullptr, 0., 0., o.data()));
                   std::memcpy(out_cv, o.data() + 8 + 2 * ns, ns * 8); return GF_OK;)
}
int gf_enthalpy_mix(gf_mech *m, double T, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, 0., 0., o.data())); *out_value = o[5]; return GF_OK;)
}
int gf_energy_mix(gf_mech *m, double T, const double *y, double *out_value) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, 0., 0., o.data())); *out_value = o[6]; return GF_OK;)
}
int gf_species_enthalpies(gf_mech *m, double T, double *out_h) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, nullptr, 0., 0., o.data()));
                   std::memcpy(out_h, o.data() + 8 + 3 * ns, ns * 8); return GF_OK;)
}
int gf_species_energies(gf_mech *m, double T, double *out_e) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, nullptr, 0., 0., o.data()));
                   std::memcpy(out_e, o.data() + 8 + 4 * ns, ns * 8); return GF_OK;)
}
int gf_dcpdT_species(gf_mech *m, double T, const double *y, double *out_dcpdT) {
  GF_THERMO_GETTER(GF_TRY(thermo_all(m, T, y, 0., 0., o.data()));
                   std::memcpy(out_dcpdT, o.data() + 8 + 5 * ns, ns * 8); return GF_OK;)
}
#undef GF_THERMO_GETTER

int gf_flamelet_stencils(gf_mech *m, const double *dz, int nzi, const double *chi, const double *invLe, double *cmaj,
                         double *csub, double *csup, double *mco, double *nco) {
  if (nzi < 1) return fail(GF_ERR_ARG, "nzi must be positive");
  Stage s;
  const int ns = gf_mech_n_species(m);
  const size_t nv = (size_t)nzi * ns;
  GF_TRY(prep(m, rup((nzi + 1) * 8) + rup((nzi + 2) * 8) + rup(ns * 8) + 3 * rup(nv * 8) + 2 * rup(nzi * 8), s));
  double *ddz = s.take<double>(nzi + 1), *dchi = s.take<double>(nzi + 2), *dle = s.take<double>(ns);
  double *a = s.take<double>(nv), *b = s.take<double>(nv), *c = s.take<double>(nv), *d = s.take<double>(nzi),
         *e = s.take<double>(nzi);
  hipStream_t st = m->stream;
  GF_TRY(h2d(ddz, dz, (nzi + 1) * 8, st));
  GF_TRY(h2d(dchi, chi, (nzi + 2) * 8, st));
  GF_TRY(h2d(dle, invLe, ns * 8, st));
  GF_TRY(launch_flamelet_stencils(ns, ddz, nzi, dchi, dle, a, b, c, d, e, st));
  GF_TRY(d2h(cmaj, a, nv * 8, st));
  GF_TRY(d2h(csub, b, nv * 8, st));
  GF_TRY(d2h(csup, c, nv * 8, st));
  GF_TRY(d2h(mco, d, nzi * 8, st));
  GF_TRY(d2h(nco, e, nzi * 8, st));
  return sync(st);
}

int gf_flamelet_jac_indices(gf_mech *m, int nzi, int *rows, int *cols) {
  if (nzi < 1) return fail(GF_ERR_ARG, "nzi must be positive");
  Stage s;
  const int ns = gf_mech_n_species(m);
  const size_t nel = (size_t)ns * (nzi * ns + 2 * (nzi - 1));
  GF_TRY(prep(m, 2 * rup(nel * 4), s));
  int *r = s.take<int>(nel), *c = s.take<int>(nel);
  hipStream_t st = m->stream;
  GF_TRY(launch_flamelet_jac_indices(ns, nzi, r, c, st));
  GF_TRY(d2h(rows, r, nel * 4, st));
  GF_TRY(d2h(cols, c, nel * 4, st));
  return sync(st);
}

namespace {

// Stages the flamelet inputs and runs the chemistry over [oxidizer, interior..., fuel] (rates + aux).
int flamelet_stage(gf_mech *m, Stage &s, const double *state, double pressure, const double *oxy, const double *fuel,
                   int adiabatic, const double *Tc, const double *hc, const double *Tr, const double *hr, int nzi,
                   const double *cmaj, const double *csub, const double *csup, const double *mco, const double *nco,
                   const double *chi, size_t extra, FlameletIO &f, double *&dpts, double *&dp) {
  const int ns = gf_mech_n_species(m);
  const size_t nv = (size_t)nzi * ns, npt = (size_t)(nzi + 2);
  const size_t bytes = rup(npt * ns * 8) + rup(8) + 3 * rup(nv * 8) + 3 * rup(nzi * 8) + 4 * rup(nzi * 8) +
                       rup(npt * ns * 8) + rup(npt * 4 * 8) + rup(npt * ns * 8) + extra;
  GF_TRY(prep(m, bytes, s));
  hipStream_t st = m->stream;
  dpts = s.take<double>(npt * ns);
  dp = s.take<double>(1);
  double *dcm = s.take<double>(nv), *dcs = s.take<double>(nv), *dcu = s.take<double>(nv);
  double *dm = s.take<double>(nzi), *dn = s.take<double>(nzi), *dchi = s.take<double>(nzi);
  double *dTc = s.take<double>(nzi), *dhc = s.take<double>(nzi), *dTr = s.take<double>(nzi), *dhr = s.take<double>(nzi);
  double *drhs = s.take<double>(npt * ns), *daux = s.take<double>(npt * 4), *dcpi = s.take<double>(npt * ns);
  GF_TRY(h2d(dpts, oxy, ns * 8, st));
  GF_TRY(h2d(dpts + ns, state, nv * 8, st));
  GF_TRY(h2d(dpts + (nzi + 1) * ns, fuel, ns * 8, st));
  GF_TRY(h2d(dp, &pressure, 8, st));
  GF_TRY(h2d(dcm, cmaj, nv * 8, st));
  GF_TRY(h2d(dcs, csub, nv * 8, st));
  GF_TRY(h2d(dcu, csup, nv * 8, st));
  GF_TRY(h2d(dm, mco, nzi * 8, st));
  GF_TRY(h2d(dn, nco, nzi * 8, st));
  GF_TRY(h2d(dchi, chi, nzi * 8, st));
  if (!adiabatic) {
    GF_TRY(h2d(dTc, Tc, nzi * 8, st));
    GF_TRY(h2d(dhc, hc, nzi * 8, st));
    GF_TRY(h2d(dTr, Tr, nzi * 8, st));
    GF_TRY(h2d(dhr, hr, nzi * 8, st));
  }
  ChemIO io;
  io.n = (int)npt; io.x = dpts; io.xs_i = ns; io.xs_v = 1; io.p = dp; io.ps = 0;
  io.rhs = drhs; io.rs_i = ns; io.rs_v = 1; io.aux = daux; io.cpi = dcpi;
  GF_TRY(launch_chem(m, io, st));
  f.nzi = nzi; f.ns = ns; f.state = dpts + ns; f.oxy = dpts; f.fuel = dpts + (nzi + 1) * ns;
  f.cmaj = dcm; f.csub = dcs; f.csup = dcu; f.mco = dm; f.nco = dn; f.chi = dchi;
  f.adiabatic = adiabatic; f.Tc = dTc; f.hc = dhc; f.Tr = dTr; f.hr = dhr;
  f.chem_rhs = drhs; f.aux = daux; f.cpi = dcpi;
  return GF_OK;
}

}  // namespace

int gf_flamelet_rhs(gf_mech *m, const double *state, double pressure, const double *oxy, const double *fuel,
                    int adiabatic, const double *Tc, const double *hc, const double *Tr, const double *hr, int nzi,
                    const double *cmaj, const double *csub, const double *csup, const double *mco, const double *nco,
                    const double *chi, int incl_ef, int incl_vcp, int scaled_hl, double *out) {
  if (nzi < 2) return fail(GF_ERR_ARG, "a flamelet needs at least two interior points");
  const int ns = gf_mech_n_species(m);
  Stage s;
  FlameletIO f;
  double *dpts, *dp;
  GF_TRY(flamelet_stage(m, s, state, pressure, oxy, fuel, adiabatic, Tc, hc, Tr, hr, nzi, cmaj, csub, csup, mco, nco,
                        chi, rup((size_t)nzi * ns * 8), f, dpts, dp));
  f.incl_ef = incl_ef; f.incl_vcp = incl_vcp; f.scaled_hl = scaled_hl;
  double *dout = s.take<double>((size_t)nzi * ns);
  hipStream_t st = m->stream;
  GF_TRY(launch_flamelet_rhs_epilogue(f, dout, st));
  GF_TRY(d2h(out, dout, (size_t)nzi * ns * 8, st));
  return sync(st);
}

int gf_flamelet_jacobian(gf_mech *m, const double *state, double pressure, const double *oxy, const double *fuel,
                         int adiabatic, const double *Tc, const double *hc, const double *Tr, const double *hr, int nzi,
                         const double *cmaj, const double *csub, const double *csup, const double *mco,
                         const double *nco, const double *chi, int compute_eigenvalues, double diffterm,
                         int scale_and_offset, double prefactor, int rates_opt, int transform_opt, int incl_ef,
                         int incl_vcp, int scaled_hl, double *out_expeig, double *out_jac) {
  if (nzi < 2) return fail(GF_ERR_ARG, "a flamelet needs at least two interior points");
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  GF_TRY(check_options(m, 0, 0, rates_opt, transform_opt));
  const int ns = gf_mech_n_species(m);
  const size_t bsz = (size_t)ns * ns, nbl = (size_t)nzi * bsz, noff = (size_t)(nzi - 1) * ns;
  const size_t nel = nbl + 2 * noff;
  Stage s;
  FlameletIO f;
  double *dpts, *dp;
  GF_TRY(flamelet_stage(m, s, state, pressure, oxy, fuel, adiabatic, Tc, hc, Tr, hr, nzi, cmaj, csub, csup, mco, nco,
                        chi, rup(nel * 8) + 2 * rup((size_t)nzi * ns * 8), f, dpts, dp));
  f.incl_ef = incl_ef; f.incl_vcp = incl_vcp; f.scaled_hl = scaled_hl;
  double *dj = s.take<double>(nel), *dr = s.take<double>((size_t)nzi * ns), *dex = s.take<double>((size_t)nzi * ns);
  hipStream_t st = m->stream;
  // off-diagonal entries accumulate onto the caller's array as in the reference (flamelet_kernels.cpp:1385-1394)
  GF_TRY(h2d(dj + nbl, out_jac + nbl, 2 * noff * 8, st));
  ChemIO io;
  io.n = nzi; io.x = dpts + ns; io.xs_i = ns; io.xs_v = 1; io.p = dp; io.ps = 0;
  io.rhs = dr; io.rs_i = ns; io.rs_v = 1; io.jac = dj; io.js_i = (long)bsz;
  io.notbaf = rates_opt == 1;
  GF_TRY(launch_chem(m, io, st));
  if (compute_eigenvalues) {
    // eigenvalues of each chemistry block after the heat-loss terms, before the diffusion terms
    // (flamelet_kernels.cpp:1290-1341)
    if (!adiabatic) GF_TRY(launch_flamelet_jac_epilogue(m, f, 0, 0., dj, 1, st));
    GF_TRY(launch_expeig(m, dj, (long)bsz, nzi, ns, diffterm, dex, nullptr, st));
    GF_TRY(launch_flamelet_jac_epilogue(m, f, scale_and_offset, prefactor, dj, 2, st));
    GF_TRY(d2h(out_expeig, dex, (size_t)nzi * ns * 8, st));
  } else {
    GF_TRY(launch_flamelet_jac_epilogue(m, f, scale_and_offset, prefactor, dj, 0, st));
  }
  GF_TRY(d2h(out_jac, dj, nel * 8, st));
  return sync(st);
}

int gf_btddod_full_factorize(double *values, int nb, int bs, double *l_values, int *pivots) {
  if (nb < 1 || bs < 1) return fail(GF_ERR_ARG, "bad block counts");
  static gf_mech *ctx = nullptr;  // staging and stream for the BTDDOD drop-ins (no mechanism needed)
  if (!ctx) gf_mech_create(&ctx);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(GF_ERR_HIP, "no HIP device available");
  if (!ctx->stream && hipStreamCreate(&ctx->stream) != hipSuccess) return fail(GF_ERR_HIP, "hipStreamCreate failed");
  const size_t nb2 = (size_t)bs * bs, nel = (size_t)bs * (nb * bs + 2 * (nb - 1));
  const size_t bytes = rup(nel * 8) + rup(nb * nb2 * 8) + rup((size_t)nb * bs * 4) + 4 * rup(8);
  GF_TRY(ensure_stage(ctx, bytes + 4096));
  Stage s;
  s.base = static_cast<char *>(ctx->stage);
  double *dv = s.take<double>(nel), *dl = s.take<double>(nb * nb2);
  int *dpv = s.take<int>((size_t)nb * bs), *dnb = s.take<int>(1);
  long *doff = s.take<long>(3);
  hipStream_t st = ctx->stream;
  const long zero3[3] = {0, 0, 0};
  GF_TRY(h2d(dv, values, nel * 8, st));
  GF_TRY(h2d(dnb, &nb, 4, st));
  GF_TRY(h2d(doff, zero3, sizeof(zero3), st));
  if (hipMemsetAsync(dl, 0, nb2 * 8, st) != hipSuccess) return fail(GF_ERR_HIP, "memset failed");
  GF_TRY(launch_btd_factor(1, dnb, bs, dv, doff, dl, doff + 1, dpv, doff + 2, st));
  GF_TRY(d2h(values, dv, (size_t)nb * nb2 * 8, st));
  GF_TRY(d2h(l_values + nb2, dl + nb2, (nb - 1) * nb2 * 8, st));
  GF_TRY(d2h(pivots, dpv, (size_t)nb * bs * 4, st));
  return sync(st);
}

int gf_btddod_full_solve(const double *d_factors, const double *l_values, const int *pivots, const double *rhs, int nb,
                         int bs, double *out) {
  if (nb < 1 || bs < 1) return fail(GF_ERR_ARG, "bad block counts");
  static gf_mech *ctx = nullptr;
  if (!ctx) gf_mech_create(&ctx);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(GF_ERR_HIP, "no HIP device available");
  if (!ctx->stream && hipStreamCreate(&ctx->stream) != hipSuccess) return fail(GF_ERR_HIP, "hipStreamCreate failed");
  const size_t nb2 = (size_t)bs * bs, nel = (size_t)bs * (nb * bs + 2 * (nb - 1)), nv = (size_t)nb * bs;
  const size_t bytes = rup(nel * 8) + rup(nb * nb2 * 8) + rup(nv * 4) + 2 * rup(nv * 8) + 4 * rup(8);
  GF_TRY(ensure_stage(ctx, bytes + 4096));
  Stage s;
  s.base = static_cast<char *>(ctx->stage);
  double *dv = s.take<double>(nel), *dl = s.take<double>(nb * nb2);
  int *dpv = s.take<int>(nv), *dnb = s.take<int>(1);
  double *db = s.take<double>(nv), *dx = s.take<double>(nv);
  long *doff = s.take<long>(4);
  hipStream_t st = ctx->stream;
  const long zero4[4] = {0, 0, 0, 0};
  GF_TRY(h2d(dv, d_factors, nel * 8, st));
  if (nb > 1) GF_TRY(h2d(dl + nb2, l_values + nb2, (nb - 1) * nb2 * 8, st));
  GF_TRY(h2d(dpv, pivots, nv * 4, st));
  GF_TRY(h2d(db, rhs, nv * 8, st));
  GF_TRY(h2d(dnb, &nb, 4, st));
  GF_TRY(h2d(doff, zero4, sizeof(zero4), st));
  GF_TRY(launch_btd_solve(1, dnb, bs, dv, doff, dl, doff + 1, dpv, doff + 2, db, dx, doff + 3, st));
  GF_TRY(d2h(out, dx, nv * 8, st));
  return sync(st);
}

int gf_btddod_scale_and_add_diagonal(double *values, double a, const double *diag, double b, int nb, int bs) {
  if (nb < 1 || bs < 1) return fail(GF_ERR_ARG, "bad block counts");
  static gf_mech *ctx = nullptr;
  if (!ctx) gf_mech_create(&ctx);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(GF_ERR_HIP, "no HIP device available");
  if (!ctx->stream && hipStreamCreate(&ctx->stream) != hipSuccess) return fail(GF_ERR_HIP, "hipStreamCreate failed");
  const size_t nel = (size_t)bs * (nb * bs + 2 * (nb - 1)), nv = (size_t)nb * bs;
  GF_TRY(ensure_stage(ctx, rup(nel * 8) + rup(nv * 8) + 4096));
  Stage s;
  s.base = static_cast<char *>(ctx->stage);
  double *dv = s.take<double>(nel), *dd = s.take<double>(nv);
  hipStream_t st = ctx->stream;
  GF_TRY(h2d(dv, values, nel * 8, st));
  GF_TRY(h2d(dd, diag, nv * 8, st));
  GF_TRY(launch_scale_and_add_diagonal(dv, a, dd, b, nb, bs, st));
  GF_TRY(d2h(values, dv, nel * 8, st));
  return sync(st);
}

// ---------------- batched device entry points ----------------

int gf_set_chem_launch_shape(int shape) { return set_chem_launch_shape(shape); }

int gf_batch_reactor_rhs_isobaric(gf_mech *m, int n, const double *x, long xs_i, long xs_v, const double *p, long ps,
                                  double *rhs, long rs_i, long rs_v, void *stream) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  if (n < 0 || !x || !p || !rhs) return fail(GF_ERR_ARG, "bad batch arguments");
  GF_TRY(ensure_device(m));
  ChemIO io;
  io.n = n; io.x = x; io.xs_i = xs_i; io.xs_v = xs_v; io.p = p; io.ps = ps; io.rhs = rhs; io.rs_i = rs_i; io.rs_v = rs_v;
  return launch_chem(m, io, static_cast<hipStream_t>(stream));
}

int gf_batch_reactor_jac_isobaric(gf_mech *m, int n, const double *x, long xs_i, long xs_v, const double *p, long ps,
                                  double *rhs, long rs_i, long rs_v, double *jac, long js_i, void *stream) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  if (n < 0 || !x || !p || !jac) return fail(GF_ERR_ARG, "bad batch arguments");
  GF_TRY(ensure_device(m));
  ChemIO io;
  io.n = n; io.x = x; io.xs_i = xs_i; io.xs_v = xs_v; io.p = p; io.ps = ps; io.rhs = rhs; io.rs_i = rs_i; io.rs_v = rs_v;
  io.jac = jac; io.js_i = js_i;
  return launch_chem(m, io, static_cast<hipStream_t>(stream));
}

int gf_batch_reactor_jac_isobaric_terms(gf_mech *m, int n, const double *x, long xs_i, long xs_v, const double *p,
                                        long ps, double *rhs, long rs_i, long rs_v, double *jac, long js_i, double *aux,
                                        double *cpi, double *h, double *qa, void *stream) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  if (n < 0 || !x || !p || !jac) return fail(GF_ERR_ARG, "bad batch arguments");
  GF_TRY(ensure_device(m));
  ChemIO io;
  io.n = n; io.x = x; io.xs_i = xs_i; io.xs_v = xs_v; io.p = p; io.ps = ps; io.rhs = rhs; io.rs_i = rs_i; io.rs_v = rs_v;
  io.jac = jac; io.js_i = js_i; io.aux = aux; io.cpi = cpi; io.hsp = h; io.qa = qa;
  return launch_chem(m, io, static_cast<hipStream_t>(stream));
}

int gf_batch_reactor_general(gf_mech *m, int n, int isochoric, const double *x, long x_item_stride, const double *p,
                             long p_stride, double inflow_density, double inflow_T, const double *inflow_y, double tau,
                             double fluid_T, double surf_T, double h_conv, double eps_rad, double sov, int heat,
                             int open, int rates_opt, double *rhs, double *jac, void *stream) {
  GF_TRY(check_batch_general(m, n, isochoric, x, x_item_stride, p, inflow_y, heat, open, rates_opt, rhs));
  if (n == 0) return GF_OK;
  GF_TRY(ensure_device(m));
  GF_TRY(ensure_work(m, (size_t)n * batch_work_doubles(m->view.ns, jac != nullptr) * sizeof(double)));
  return batch_general(m, n, isochoric, x, x_item_stride, p, p_stride, inflow_density, inflow_T, inflow_y, tau,
                       fluid_T, surf_T, h_conv, eps_rad, sov, heat, open, rates_opt, rhs, jac,
                       static_cast<double *>(m->work), 0, static_cast<hipStream_t>(stream));
}

int gf_batch_mixture_properties(gf_mech *m, int n, const double *Ty, const double *p, const double *rho,
                                double *out, void *stream) {
  if (!m) return fail(GF_ERR_ARG, "null mechanism");
  if (n < 0 || (n > 0 && (!Ty || !out))) return fail(GF_ERR_ARG, "bad mixture property arguments");
  if (n == 0) return GF_OK;
  GF_TRY(ensure_device(m));
  return launch_mixture_properties(m, n, Ty, p, rho, out, static_cast<hipStream_t>(stream));
}

int gf_batch_btddod_full_factorize(int nsys, const int *nblocks, int bs, double *values, const long *voff,
                                   double *l_values, const long *loff, int *pivots, const long *poff, void *stream) {
  return launch_btd_factor(nsys, nblocks, bs, values, voff, l_values, loff, pivots, poff,
                           static_cast<hipStream_t>(stream));
}

int gf_batch_btddod_full_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                               const double *l_values, const long *loff, const int *pivots, const long *poff,
                               const double *rhs, double *x, const long *xoff, void *stream) {
  return launch_btd_solve(nsys, nblocks, bs, values, voff, l_values, loff, pivots, poff, rhs, x, xoff,
                          static_cast<hipStream_t>(stream));
}

int gf_batch_eigenvalue_bound(gf_mech *m, int nblk, int n, const double *blocks, double diffterm, double *out,
                              int *nfail, void *stream) {
  if (nblk < 0 || !blocks || !out) return fail(GF_ERR_ARG, "bad eigenvalue batch arguments");
  if (m) GF_TRY(ensure_device(m));
  return launch_expeig_list(m, blocks, nullptr, nullptr, nblk, n, diffterm, out, nfail, static_cast<hipStream_t>(stream),
                            1);
}

int gf_batch_btddod_inverse_factorize(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                                      void *dinv, const long *doff, int precision, void *stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == 64)
    return launch_btd_inv_factor<double>(nsys, nblocks, bs, values, voff, static_cast<double *>(dinv), doff, st);
  if (precision == 32)
    return launch_btd_inv_factor<float>(nsys, nblocks, bs, values, voff, static_cast<float *>(dinv), doff, st);
  return fail(GF_ERR_ARG, "precision must be 64 or 32");
}

int gf_batch_btddod_inverse_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                                  const void *dinv, const long *doff, int precision, const double *rhs, double *x,
                                  const long *xoff, void *stream) {
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == 64)
    return launch_btd_inv_solve<double>(nsys, nblocks, bs, values, voff, static_cast<const double *>(dinv), doff,
                                        rhs, x, xoff, st);
  if (precision == 32)
    return launch_btd_inv_solve<float>(nsys, nblocks, bs, values, voff, static_cast<const float *>(dinv), doff, rhs,
                                       x, xoff, st);
  return fail(GF_ERR_ARG, "precision must be 64 or 32");
}

}  // extern "C"
