/*
 * TEST INFRASTRUCTURE ONLY — C entry points over the reference's own griffon::CombustionKernels,
 * compiled together with the unmodified reference sources into oracle/_ref/ (see oracle/Makefile).
 * Used to pin the oracle restatement (oracle/griffon_oracle.cpp); never part of the product.
 *
 * The API mirrors oracle/griffon_oracle.cpp's gfo_* functions so the same test code drives both.
 * Species are named "S000", "S001", ... so std::map ordering equals index ordering, and each species
 * is built from one pseudo-element "X" of unit mass with count = molecular weight (mw exact).
 */
#include "combustion_kernels.h"
#include "btddod_matrix_kernels.h"
#include <cstdio>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using griffon::CombustionKernels;

namespace {
std::string sname(int i) {
  char b[16];
  std::snprintf(b, sizeof(b), "S%03d", i);
  return std::string(b);
}
}  // namespace

extern "C" {

void *gfr_mech_create(int ns, const double *mw, double Ru, double pref, double Tref) {
  auto *k = new CombustionKernels();
  k->mechanism_set_gas_constant(Ru);
  k->mechanism_set_ref_pressure(pref);
  k->mechanism_set_ref_temperature(Tref);
  std::map<std::string, double> emw;
  emw["X"] = 1.0;
  k->mechanism_set_element_mw_map(emw);
  k->mechanism_add_element("X");
  for (int i = 0; i < ns; ++i) {
    std::map<std::string, double> atoms;
    atoms["X"] = mw[i];
    k->mechanism_add_species(sname(i), atoms);
  }
  k->mechanism_resize_heat_capacity_data();
  return k;
}

void gfr_mech_destroy(void *h) { delete static_cast<CombustionKernels *>(h); }

int gfr_mech_set_nasa7(void *h, int i, double Tmin, double Tmid, double Tmax, const double *low, const double *high) {
  std::vector<double> lo(low, low + 7), hi(high, high + 7);
  static_cast<CombustionKernels *>(h)->mechanism_add_nasa7_cp(sname(i), Tmin, Tmid, Tmax, lo, hi);
  return 0;
}

int gfr_mech_set_const_cp(void *h, int i, double Tmin, double Tmax, double T0, double h0, double s0, double cp) {
  static_cast<CombustionKernels *>(h)->mechanism_add_const_cp(sname(i), Tmin, Tmax, T0, h0, s0, cp);
  return 0;
}

int gfr_mech_add_reaction(void *h, int type, int reversible, int n_r, const int *r_idx, const int *r_st, int n_p,
                          const int *p_idx, const int *p_st, const double *kfwd, int n_tb, const int *tb_idx,
                          const double *tb_eff, double default_eff, const double *kpres, const double *troe,
                          int n_ord, const int *ord_idx, const double *ord_val) {
  auto *k = static_cast<CombustionKernels *>(h);
  std::map<std::string, int> rc, pd;
  for (int i = 0; i < n_r; ++i) rc[sname(r_idx[i])] = r_st[i];
  for (int i = 0; i < n_p; ++i) pd[sname(p_idx[i])] = p_st[i];
  std::map<std::string, double> tb, ord;
  for (int i = 0; i < n_tb; ++i) tb[sname(tb_idx[i])] = tb_eff[i];
  for (int i = 0; i < n_ord; ++i) ord[sname(ord_idx[i])] = ord_val[i];
  std::vector<double> tp;
  if (troe) tp.assign(troe, troe + 4);
  const bool rev = reversible != 0;
  try {
    if (n_ord == 0) {
      switch (type) {
        case 1: k->mechanism_add_reaction_simple(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2]); break;
        case 2: k->mechanism_add_reaction_three_body(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], tb, default_eff); break;
        case 3: k->mechanism_add_reaction_Lindemann(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], tb, default_eff, kpres[0], kpres[1], kpres[2]); break;
        case 4: k->mechanism_add_reaction_Troe(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], tb, default_eff, kpres[0], kpres[1], kpres[2], tp); break;
        default: return 1;
      }
    } else {
      switch (type) {
        case 1: k->mechanism_add_reaction_simple_with_special_orders(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], ord); break;
        case 2: k->mechanism_add_reaction_three_body_with_special_orders(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], tb, default_eff, ord); break;
        case 3: k->mechanism_add_reaction_Lindemann_with_special_orders(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], tb, default_eff, kpres[0], kpres[1], kpres[2], ord); break;
        case 4: k->mechanism_add_reaction_Troe_with_special_orders(rc, pd, rev, kfwd[0], kfwd[1], kfwd[2], tb, default_eff, kpres[0], kpres[1], kpres[2], tp, ord); break;
        default: return 1;
      }
    }
  } catch (...) {
    return 2;
  }
  return 0;
}

int gfr_production_rates(void *h, double T, double rho, const double *y, double *w) {
  static_cast<CombustionKernels *>(h)->production_rates(T, rho, y, w);
  return 0;
}

int gfr_prod_rates_sens(void *h, double rho, double T, const double *y, double *w, double *ws) {
  (void)w;
  static_cast<CombustionKernels *>(h)->prod_rates_primitive_sensitivities(rho, T, y, 0, ws);
  return 0;
}

int gfr_thermo(void *h, double T, const double *y, double *sc, double *cpi, double *hi, double *cpis) {
  auto *k = static_cast<CombustionKernels *>(h);
  double cps;
  sc[0] = k->mixture_molecular_weight(y);
  sc[1] = k->cp_mix(T, y);
  k->cp_sens_T(T, y, &cps, cpis);
  sc[2] = cps;
  sc[3] = k->enthalpy_mix(T, y);
  k->species_cp(T, cpi);
  k->species_enthalpies(T, hi);
  return 0;
}

int gfr_reactor_rhs_isobaric(void *h, const double *state, double p, double Tin, const double *yin, double tau,
                             double Tf, double Ts, double hc, double eps, double sov, int heat, int open, double *rhs) {
  static_cast<CombustionKernels *>(h)->reactor_rhs_isobaric(state, p, Tin, yin, tau, Tf, Ts, hc, eps, sov, heat, open != 0, rhs);
  return 0;
}

int gfr_reactor_jac_isobaric(void *h, const double *state, double p, double Tin, const double *yin, double tau,
                             double Tf, double Ts, double hc, double eps, double sov, int heat, int open,
                             double *rhs, double *jac) {
  static_cast<CombustionKernels *>(h)->reactor_jac_isobaric(state, p, Tin, yin, tau, Tf, Ts, hc, eps, sov, heat, open != 0, 0, 0, rhs, jac);
  return 0;
}

int gfr_prod_rates_sens_opt(void *h, double rho, double T, const double *y, int option, double *ws) {
  try { static_cast<CombustionKernels *>(h)->prod_rates_primitive_sensitivities(rho, T, y, option, ws); }
  catch (...) { return 2; }
  return 0;
}

int gfr_reactor_jac_isobaric_opt(void *h, const double *state, double p, double Tin, const double *yin, double tau,
                                 double Tf, double Ts, double hc, double eps, double sov, int heat, int open,
                                 int rates_opt, double *rhs, double *jac) {
  static_cast<CombustionKernels *>(h)->reactor_jac_isobaric(state, p, Tin, yin, tau, Tf, Ts, hc, eps, sov, heat,
                                                            open != 0, rates_opt, 0, rhs, jac);
  return 0;
}

int gfr_reactor_rhs_isochoric(void *h, const double *state, double rhoin, double Tin, const double *yin, double tau,
                              double Tf, double Ts, double hc, double eps, double sov, int heat, int open, double *rhs) {
  static_cast<CombustionKernels *>(h)->reactor_rhs_isochoric(state, rhoin, Tin, yin, tau, Tf, Ts, hc, eps, sov, heat,
                                                             open != 0, rhs);
  return 0;
}

int gfr_reactor_jac_isochoric(void *h, const double *state, double rhoin, double Tin, const double *yin, double tau,
                              double Tf, double Ts, double hc, double eps, double sov, int heat, int open,
                              int rates_opt, double *rhs, double *jac) {
  static_cast<CombustionKernels *>(h)->reactor_jac_isochoric(state, rhoin, Tin, yin, tau, Tf, Ts, hc, eps, sov, heat,
                                                             open != 0, rates_opt, rhs, jac);
  return 0;
}

int gfr_thermo_getters(void *h, double T, const double *y, double p, double rho, double *sc, double *x, double *cpi,
                       double *cvi, double *hi, double *ei, double *cpis) {
  auto *k = static_cast<CombustionKernels *>(h);
  double cps;
  sc[0] = k->mixture_molecular_weight(y);
  k->mole_fractions(y, x);
  sc[1] = k->ideal_gas_density(p, T, y);
  sc[2] = k->ideal_gas_pressure(rho, T, y);
  sc[3] = k->cp_mix(T, y);
  sc[4] = k->cv_mix(T, y);
  sc[5] = k->enthalpy_mix(T, y);
  sc[6] = k->energy_mix(T, y);
  k->species_cp(T, cpi);
  k->species_cv(T, cvi);
  k->species_enthalpies(T, hi);
  k->species_energies(T, ei);
  k->cp_sens_T(T, y, &cps, cpis);
  sc[7] = cps;
  return 0;
}

// OpenMP over independent states (griffon's compute methods are const and use only stack storage).
int gfr_batch_rhs_jac(void *h, int N, const double *stateSoA, const double *p, double *rhs, double *jac, int nthreads) {
  auto *k = static_cast<CombustionKernels *>(h);
  const int n = k->get_mechanism_data().phaseData.nSpecies;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) if (nthreads != 1)
#endif
  for (int i = 0; i < N; ++i) {
    std::vector<double> st(n), yin(n, 0.);
    for (int v = 0; v < n; ++v) st[v] = stateSoA[(size_t)v * N + i];
    k->reactor_jac_isobaric(st.data(), p[i], 0., yin.data(), 0., 0., 0., 0., 0., 0., 0, false, 0, 0,
                            &rhs[(size_t)i * n], &jac[(size_t)i * n * n]);
  }
  return 0;
}

int gfr_flamelet_stencils(void *h, const double *dz, int nzi, const double *chi, const double *invLe, double *cmaj,
                          double *csub, double *csup, double *mco, double *nco) {
  static_cast<CombustionKernels *>(h)->flamelet_stencils(dz, nzi, chi, invLe, cmaj, csub, csup, mco, nco);
  return 0;
}

int gfr_flamelet_jac_indices(void *h, int nzi, int *rows, int *cols) {
  static_cast<CombustionKernels *>(h)->flamelet_jac_indices(nzi, rows, cols);
  return 0;
}

int gfr_flamelet_rhs(void *h, const double *state, double p, const double *oxy, const double *fuel, int adiabatic,
                     const double *Tconv, const double *hconv, const double *Trad, const double *hrad, int nzi,
                     const double *cmaj, const double *csub, const double *csup, const double *mco, const double *nco,
                     const double *chi, int incl_ef, int incl_vcp, int scaled_hl, double *out) {
  static_cast<CombustionKernels *>(h)->flamelet_rhs(state, p, oxy, fuel, adiabatic != 0, Tconv, hconv, Trad, hrad, nzi,
                                                    cmaj, csub, csup, mco, nco, chi, incl_ef != 0, incl_vcp != 0,
                                                    scaled_hl != 0, out);
  return 0;
}

int gfr_flamelet_jacobian(void *h, const double *state, double p, const double *oxy, const double *fuel, int adiabatic,
                          const double *Tconv, const double *hconv, const double *Trad, const double *hrad, int nzi,
                          const double *cmaj, const double *csub, const double *csup, const double *mco,
                          const double *nco, const double *chi, int compute_eig, double diffterm, int scale_and_offset,
                          double prefactor, int incl_ef, int incl_vcp, int scaled_hl, double *expeig, double *out) {
  static_cast<CombustionKernels *>(h)->flamelet_jacobian(state, p, oxy, fuel, adiabatic != 0, Tconv, hconv, Trad, hrad,
                                                         nzi, cmaj, csub, csup, mco, nco, chi, compute_eig != 0,
                                                         diffterm, scale_and_offset != 0, prefactor, 0, 0,
                                                         incl_ef != 0, incl_vcp != 0, scaled_hl != 0, expeig, out);
  return 0;
}

int gfr_btddod_full_factorize(double *d, int nb, int bs, double *l, int *piv) {
  griffon::btddod::btddod_full_factorize(d, nb, bs, l, piv);
  return 0;
}

int gfr_btddod_full_solve(const double *d, const double *l, const int *piv, const double *rhs, int nb, int bs, double *x) {
  griffon::btddod::btddod_full_solve(d, l, piv, rhs, nb, bs, x);
  return 0;
}

int gfr_btddod_scale_and_add_diagonal(double *v, double a, const double *dg, double b, int nb, int bs) {
  griffon::btddod::btddod_scale_and_add_diagonal(v, a, dg, b, nb, bs);
  return 0;
}

}  // extern "C"
