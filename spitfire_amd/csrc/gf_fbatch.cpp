// Device-resident batched steady flamelet solvers: the reference's Newton method and pseudo-transient
// continuation (Flamelet.steady_solve_newton, flamelet.py:1352-1485; Flamelet.steady_solve_psitc,
// flamelet.py:1487-1703) run for many flamelets at once, plus the batched right-hand sides and implicit-step
// linear systems that the time-stepping fallback of compute_steady_state (flamelet.py:1759-1778) needs. The
// host keeps each flamelet's small control state (iteration count, Jacobian age, line-search factor, restart
// depth) and launches every heavy step batched over the flamelets that need it: chemistry Jacobians (+ per-block
// eigenvalues and pseudo-time steps for psitc) + the Newton-type matrix + block-Thomas factorization for the
// flamelets due a refresh, block-Thomas solves, trial states, chemistry RHS and residual norms. One host
// synchronization per line-search round reads back the per-flamelet residual norms.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace gfa;

namespace {

#define GF_TRY(expr)              \
  do {                            \
    const int _rc = (expr);       \
    if (_rc != GF_OK) return _rc; \
  } while (0)

template <typename T>
int up(T *d, const std::vector<T> &h, hipStream_t st) {
  if (h.empty()) return GF_OK;
  return hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st) == hipSuccess
             ? GF_OK
             : fail(GF_ERR_HIP, "host-to-device copy failed");
}

template <typename T>
int gfa_down(std::vector<T> &h, const T *d, size_t n, hipStream_t st) {
  h.resize(n);
  if (!n) return GF_OK;
  if (hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st) != hipSuccess)
    return fail(GF_ERR_HIP, "device-to-host copy failed");
  return hipStreamSynchronize(st) == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stream synchronize failed");
}

}  // namespace

struct gf_flamelet_batch {
  gf_mech *m = nullptr;
  int B = 0, ns = 0;
  double pressure = 0.;
  std::vector<int> nzi;
  std::vector<long> pt0, dof0, ipt0, chi0, voff, loff, poff, dz0;
  std::vector<double> diffmax;  // largest dissipation rate of each flamelet
  long NP = 0, NV = 0, NI = 0, NJ = 0, NL = 0, NPV = 0;
  void *dmem = nullptr;
  // device arrays
  double *X = nullptr, *Xt = nullptr, *X0 = nullptr, *R = nullptr, *Rt = nullptr, *dX = nullptr, *ds = nullptr,
         *eig = nullptr, *J = nullptr, *L = nullptr, *chem_rhs = nullptr, *aux = nullptr, *cpi = nullptr,
         *cmaj = nullptr, *csub = nullptr, *csup = nullptr, *mco = nullptr, *nco = nullptr, *chi = nullptr,
         *dz = nullptr, *dp = nullptr, *fval = nullptr, *fres = nullptr, *fmin = nullptr, *fdiff = nullptr;
  float *D32 = nullptr;  // FP32 inverse diagonal blocks (mixed-precision path); L holds the FP64 ones
  float *ekx = nullptr;  // [NV] balancing exponents of each point's chemistry block, carried between refreshes
  double *hl = nullptr;  // heat-loss parameters [4][NI] (own allocation), null while adiabatic
  double *ile = nullptr; // inverse Lewis numbers [B][ns] (own allocation), null for unity
  int *piv = nullptr, *dnzi = nullptr, *flag = nullptr, *gmin = nullptr, *sel = nullptr, *sys_nb = nullptr;
  long *dpt0 = nullptr, *ddof0 = nullptr, *dipt0 = nullptr, *dchi0 = nullptr, *dvoff = nullptr, *ddz0 = nullptr,
       *js_off = nullptr, *eb = nullptr, *eo = nullptr, *sys_v = nullptr, *sys_l = nullptr, *sys_p = nullptr,
       *sys_x = nullptr;
  int tr_prec = 64;          // precision of the transient_setup factorizations
  std::vector<char> tr_set;  // flamelet f holds transient_setup factors (and its values in J are theirs)
  FlameletBatchView view;
  // pinned host staging for the solver's small uploads; reused only after a stream synchronization
  char *hpin = nullptr;
  size_t hpin_bytes = 0;
};

namespace {

struct Carve {
  char *base;
  size_t off = 0;
  template <typename T>
  T *take(size_t n) {
    off = (off + 255) & ~size_t(255);
    T *p = reinterpret_cast<T *>(base + off);
    off += std::max<size_t>(n, 1) * sizeof(T);
    return p;
  }
};

size_t carve_all(gf_flamelet_batch *b, char *base) {
  Carve c{base};
  const int B = b->B, ns = b->ns;
  b->X = c.take<double>(b->NP * ns);
  b->Xt = c.take<double>(b->NP * ns);
  b->X0 = c.take<double>(b->NP * ns);
  b->R = c.take<double>(b->NV);
  b->Rt = c.take<double>(b->NV);
  b->dX = c.take<double>(b->NV);
  b->ds = c.take<double>(b->NV);
  b->eig = c.take<double>(b->NV);
  b->ekx = c.take<float>(b->NV);
  b->J = c.take<double>(b->NJ);
  b->L = c.take<double>(b->NL);
  b->D32 = c.take<float>(b->NL);
  b->chem_rhs = c.take<double>(b->NP * ns);
  b->aux = c.take<double>(b->NP * 4);
  b->cpi = c.take<double>(b->NP * ns);
  b->cmaj = c.take<double>(b->NV);
  b->csub = c.take<double>(b->NV);
  b->csup = c.take<double>(b->NV);
  b->mco = c.take<double>(b->NI);
  b->nco = c.take<double>(b->NI);
  b->chi = c.take<double>(b->NP);
  b->dz = c.take<double>(b->NP);
  b->dp = c.take<double>(1);
  b->fval = c.take<double>(B);
  b->fres = c.take<double>(B);
  b->fmin = c.take<double>(B);
  b->fdiff = c.take<double>(B);
  b->piv = c.take<int>(b->NPV);
  b->dnzi = c.take<int>(B);
  b->flag = c.take<int>(B);
  b->gmin = c.take<int>(B);
  b->sel = c.take<int>(b->NP);
  b->sys_nb = c.take<int>(B);
  b->dpt0 = c.take<long>(B);
  b->ddof0 = c.take<long>(B);
  b->dipt0 = c.take<long>(B);
  b->dchi0 = c.take<long>(B);
  b->dvoff = c.take<long>(B);
  b->ddz0 = c.take<long>(B);
  b->js_off = c.take<long>(b->NP);
  b->eb = c.take<long>(b->NI);
  b->eo = c.take<long>(b->NI);
  b->sys_v = c.take<long>(B);
  b->sys_l = c.take<long>(B);
  b->sys_p = c.take<long>(B);
  b->sys_x = c.take<long>(B);
  return c.off + 256;
}

// the solver for one call of gf_flamelet_batch_steady_psitc or gf_flamelet_batch_steady_newton; its helpers
// (uploads, chemistry, systems) also serve the transient entry points
struct Psitc {
  gf_flamelet_batch *b;
  gf_psitc_options o;
  hipStream_t st;
  int newton = 0;         // 1: the reference Newton method instead of pseudo-transient continuation
  const int *pick = nullptr;  // flamelets to solve (select[f] != 0); null: all
  long long nit = 0, njac = 0, nrhs = 0, nsync = 0;
  size_t cur = 0;  // staging cursor into b->hpin

  template <typename T>
  int up(T *d, const std::vector<T> &h) {
    const size_t bytes = h.size() * sizeof(T);
    if (!bytes) return GF_OK;
    cur = (cur + 255) & ~size_t(255);
    if (cur + bytes > b->hpin_bytes) {  // staging full: wait for the copies in flight, then reuse it
      if (hipStreamSynchronize(st) != hipSuccess) return fail(GF_ERR_HIP, "stream synchronize failed");
      cur = 0;
      if (bytes > b->hpin_bytes) return fail(GF_ERR_STATE, "flamelet batch staging buffer too small");
    }
    std::memcpy(b->hpin + cur, h.data(), bytes);
    const hipError_t e = hipMemcpyAsync(d, b->hpin + cur, bytes, hipMemcpyHostToDevice, st);
    cur += bytes;
    return e == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "host-to-device copy failed");
  }

  template <typename T>
  int down(std::vector<T> &h, const T *d, size_t n) {
    const int rc = gfa_down(h, d, n, st);
    cur = 0;  // synchronized: the staging buffer is free again
    return rc;
  }

  std::vector<int> flags_of(const std::vector<int> &fl) const {
    std::vector<int> f(b->B, 0);
    for (int x : fl) f[x] = 1;
    return f;
  }

  int chem(const double *X, const std::vector<int> &fl, bool jac) {
    std::vector<int> sel;
    for (int f : fl) {
      const long p0 = b->pt0[f];
      if (jac)
        for (int i = 1; i <= b->nzi[f]; ++i) sel.push_back((int)(p0 + i));
      else
        for (int i = 0; i < b->nzi[f] + 2; ++i) sel.push_back((int)(p0 + i));
    }
    if (sel.empty()) return GF_OK;
    GF_TRY(up(b->sel, sel));
    ChemIO io;
    io.n = (int)sel.size();
    io.x = X;
    io.xs_i = b->ns;
    io.xs_v = 1;
    io.p = b->dp;
    io.ps = 0;
    io.rhs = b->chem_rhs;
    io.rs_i = b->ns;
    io.rs_v = 1;
    io.aux = b->aux;
    io.cpi = b->cpi;
    io.sel = b->sel;
    if (jac) {
      io.jac = b->J;
      io.js_off = b->js_off;
    }
    return launch_chem(b->m, io, st);
  }

  // flamelet RHS at X for the flamelets in fl into Rout; residual norms (and the smallest state value) to
  // res/xmin on the host
  int rhs(const double *X, const std::vector<int> &fl, double *Rout, std::vector<double> &res,
          std::vector<double> *xmin) {
    if (fl.empty()) return GF_OK;
    nrhs += (long long)fl.size();
    GF_TRY(chem(X, fl, false));
    GF_TRY(up(b->flag, flags_of(fl)));
    GF_TRY(launch_fb_rhs(b->view, X, b->flag, Rout, st));
    GF_TRY(launch_fb_norms(b->view, b->flag, Rout, xmin ? X : nullptr, b->fres, b->fmin, st));
    ++nsync;
    GF_TRY(down(res, b->fres, b->B));
    if (xmin) GF_TRY(down(*xmin, b->fmin, b->B));
    return GF_OK;
  }

  int systems(const std::vector<int> &fl) {
    std::vector<int> nb;
    std::vector<long> v, l, p, x;
    for (int f : fl) {
      nb.push_back(b->nzi[f]);
      v.push_back(b->voff[f]);
      l.push_back(b->loff[f]);
      p.push_back(b->poff[f]);
      x.push_back(b->dof0[f]);
    }
    GF_TRY(up(b->sys_nb, nb));
    GF_TRY(up(b->sys_v, v));
    GF_TRY(up(b->sys_l, l));
    GF_TRY(up(b->sys_p, p));
    return up(b->sys_x, x);
  }

  int factor(const std::vector<int> &fl, int precision) {
    GF_TRY(systems(fl));
    if (precision == 32)
      return launch_btd_inv_factor<float>((int)fl.size(), b->sys_nb, b->ns, b->J, b->sys_v, b->D32, b->sys_l, st);
    return launch_btd_inv_factor<double>((int)fl.size(), b->sys_nb, b->ns, b->J, b->sys_v, b->L, b->sys_l, st);
  }

  // Jacobian refresh for the flamelets in fl at X. Psitc: chemistry blocks, eigenvalue limiter, local
  // pseudo-time steps, diag(1/ds) - J, block-Thomas factorization (flamelet.py:1597-1608). Newton: -J,
  // factored (flamelet.py:1421-1427).
  int refresh(const std::vector<int> &fl, const std::vector<int> &first) {
    if (fl.empty()) return GF_OK;
    njac += (long long)fl.size();
    for (int f : fl) b->tr_set[f] = 0;  // their values and factors are overwritten
    GF_TRY(chem(b->X, fl, true));
    if (newton) {
      GF_TRY(up(b->flag, flags_of(fl)));
      GF_TRY(launch_fb_matrix(b->m, b->view, b->X, b->flag, 1, nullptr, 0., b->J, st));
      return factor(fl, o.lu_precision);
    }
    std::vector<long> eb, eo;
    const long bsz = (long)b->ns * b->ns;
    for (int f : fl)
      for (int i = 0; i < b->nzi[f]; ++i) {
        eb.push_back(b->voff[f] + i * bsz);
        eo.push_back(b->dof0[f] + (long)i * b->ns);
      }
    GF_TRY(up(b->eb, eb));
    GF_TRY(up(b->eo, eo));
    // max(0, max Re(lambda)); the per-flamelet diffusion shift is applied with the step update. Each point's
    // balancing starts from the scaling its block had at the previous refresh (the same eigenvalues, fewer sweeps).
    GF_TRY(launch_expeig_list(b->m, b->J, b->eb, b->eo, (long)eb.size(), b->ns, 0., b->eig, nullptr, st, b->ns,
                              b->ekx));
    GF_TRY(up(b->flag, flags_of(fl)));
    std::vector<int> g(b->B, 0);
    for (size_t k = 0; k < fl.size(); ++k) g[fl[k]] = first[k];
    GF_TRY(up(b->gmin, g));
    GF_TRY(launch_fb_ds(b->view, b->flag, b->gmin, b->eig, b->fdiff, o.ds_safety, o.ds_ramp, o.ds_max, b->ds, st));
    GF_TRY(launch_fb_matrix(b->m, b->view, b->X, b->flag, 0, b->ds, 0., b->J, st));
    return factor(fl, o.lu_precision);
  }

  int solve(const std::vector<int> &fl, int precision) {
    if (fl.empty()) return GF_OK;
    GF_TRY(systems(fl));
    if (precision == 32)
      return launch_btd_inv_solve<float>((int)fl.size(), b->sys_nb, b->ns, b->J, b->sys_v, b->D32, b->sys_l, b->R,
                                         b->dX, b->sys_x, st);
    return launch_btd_inv_solve<double>((int)fl.size(), b->sys_nb, b->ns, b->J, b->sys_v, b->L, b->sys_l, b->R, b->dX,
                                        b->sys_x, st);
  }

  int run(int *iterations, int *converged, double *residual, double *min_ds) {
    const int B = b->B;
    std::vector<int> active(B, 1), it(B, 0), age(B, 0), evalJ(B, 1), depth(B, 0), conv(B, 0);
    std::vector<double> dsinit(B, o.ds_init), res(B, 0.), nrm(B, 0.), alpha(B, 1.), nold(B, 0.), tmp, tmin;

    // a fresh attempt from the initial guess (flamelet.py:1573-1590)
    auto start = [&](const std::vector<int> &fl) -> int {
      if (fl.empty()) return GF_OK;
      GF_TRY(up(b->flag, flags_of(fl)));
      GF_TRY(launch_fb_copy_interior(b->view, b->flag, b->X0, b->X, st));
      if (!newton) {
        std::vector<double> v(B, 0.);
        for (int f : fl) v[f] = dsinit[f];
        GF_TRY(up(b->fval, v));
        GF_TRY(launch_fb_fill_dofs(b->view, b->flag, b->fval, b->ds, st));
      }
      GF_TRY(rhs(b->X, fl, b->R, tmp, nullptr));
      for (int f : fl) {
        nrm[f] = tmp[f];
        res[f] = o.tolerance + 1.;
        it[f] = 0;
        age[f] = 0;
        evalJ[f] = 1;
      }
      return GF_OK;
    };
    // psitc: restart with a smaller initial pseudo-time step, or give up (flamelet.py:1563-1571); Newton
    // gives up at once (flamelet.py:1435-1468)
    auto restart = [&](const std::vector<int> &fl) -> int {
      std::vector<int> again;
      for (int f : fl) {
        if (newton || depth[f] > o.max_recursion_depth || !o.adaptive_restart) {
          active[f] = 0;
          conv[f] = 0;
        } else {
          depth[f] += 1;
          dsinit[f] /= o.ds_init_decrease;
          again.push_back(f);
        }
      }
      return start(again);
    };

    std::vector<int> all;
    for (int f = 0; f < B; ++f) {
      if (!pick || pick[f]) all.push_back(f);
      else active[f] = 0;
    }
    GF_TRY(start(all));
    const auto t0 = std::chrono::steady_clock::now();
    long sweep = 0;

    for (;;) {
      std::vector<int> step;
      for (int f = 0; f < B; ++f) {
        if (!active[f]) continue;
        if (!(res[f] > o.tolerance)) {
          active[f] = 0;
          conv[f] = 1;
        } else if (it[f] >= o.max_iterations) {
          active[f] = 0;  // iteration limit: failure, no restart (flamelet.py:1693-1698)
          conv[f] = 0;
        } else {
          step.push_back(f);
        }
      }
      if (step.empty()) break;
      if (o.log_rate > 0 && (sweep++ % o.log_rate) == 0) {
        double rmax = 0.;
        for (int f : step) rmax = std::max(rmax, res[f]);
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "%s sweep %ld: %zu of %d flamelets active, max residual %.3e, %.2f s\n",
                     newton ? "newton" : "psitc", sweep - 1, step.size(), B, rmax, el);
        std::fflush(stderr);
      }
      std::vector<int> ref, first;
      for (int f : step) {
        it[f] += 1;
        ++nit;
        if (evalJ[f]) {
          ref.push_back(f);
          first.push_back(o.global_ds || it[f] == 1);
          age[f] = 0;
        } else {
          age[f] += 1;
        }
        // psitc refreshes every 8 iterations and while the residual is large (flamelet.py:1615); Newton keeps
        // its Jacobian until a line search cuts the step (flamelet.py:1420-1449)
        evalJ[f] = newton ? 0 : ((age[f] == 8) || res[f] > 1.e-2);
      }
      GF_TRY(refresh(ref, first));
      GF_TRY(solve(step, o.lu_precision));
      // full step, then the line search on the scaled residual (flamelet.py:1623-1650)
      GF_TRY(up(b->flag, flags_of(step)));
      GF_TRY(launch_fb_trial(b->view, b->flag, nullptr, b->X, b->Xt, b->dX, st));
      GF_TRY(rhs(b->Xt, step, b->Rt, tmp, &tmin));
      std::vector<int> bad, search;
      for (int f : step) {
        nold[f] = nrm[f];
        alpha[f] = 1.;
        if (!std::isfinite(tmp[f])) bad.push_back(f);
        else if (tmp[f] > o.max_factor_line_search * nold[f] && alpha[f] > 0.001) search.push_back(f);
      }
      std::vector<double> tres = tmp, txmin = tmin;
      while (!search.empty()) {
        std::vector<double> sc(B, 1.);
        for (int f : search) {
          alpha[f] *= 0.5;
          sc[f] = alpha[f];  // the step shrinks by the running product of the factors, as in the reference
          evalJ[f] = 1;
        }
        GF_TRY(up(b->fval, sc));
        GF_TRY(up(b->flag, flags_of(search)));
        GF_TRY(launch_fb_trial(b->view, b->flag, b->fval, b->X, b->Xt, b->dX, st));
        GF_TRY(rhs(b->Xt, search, b->Rt, tmp, &tmin));
        std::vector<int> next;
        for (int f : search) {
          tres[f] = tmp[f];
          txmin[f] = tmin[f];
          if (tmp[f] > o.max_factor_line_search * nold[f] && alpha[f] > 0.001) next.push_back(f);
        }
        search.swap(next);
      }
      std::vector<int> acc;
      for (int f : step)
        if (std::isfinite(tres[f])) acc.push_back(f);
      if (!acc.empty()) {
        GF_TRY(up(b->flag, flags_of(acc)));
        GF_TRY(launch_fb_copy_interior(b->view, b->flag, b->Xt, b->X, st));
        GF_TRY(launch_fb_copy_dofs(b->view, b->flag, b->Rt, b->R, st));
      }
      for (int f : acc) {
        nrm[f] = tres[f];
        res[f] = tres[f];
        if (res[f] > o.max_allowed_residual || !(txmin[f] >= o.min_allowable_state_var)) bad.push_back(f);
      }
      GF_TRY(restart(bad));
    }
    GF_TRY(launch_fb_clip_points(b->NP * b->ns, b->X, st));
    GF_TRY(hipStreamSynchronize(st) == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stream synchronize failed"));
    std::vector<double> dsh;
    if (min_ds && !newton) GF_TRY(down(dsh, b->ds, b->NV));
    for (int f : all) {
      if (iterations) iterations[f] = it[f];
      if (converged) converged[f] = conv[f];
      if (residual) residual[f] = res[f];
      if (min_ds && !newton) {  // np.min(ds) as the reference returns it (flamelet.py:1693-1703)
        double mn = INFINITY;
        for (long e = 0; e < (long)b->nzi[f] * b->ns; ++e) mn = std::min(mn, dsh[b->dof0[f] + e]);
        min_ds[f] = mn;
      }
    }
    return GF_OK;
  }

  // the flamelets with flags[f] != 0, uploaded as the device flags
  int flagged(const int *flags, std::vector<int> &fl) {
    for (int f = 0; f < b->B; ++f)
      if (flags[f]) fl.push_back(f);
    return fl.empty() ? GF_OK : up(b->flag, flags_of(fl));
  }

  int sync() {
    cur = 0;
    return hipStreamSynchronize(st) == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stream synchronize failed");
  }
};

}  // namespace

extern "C" {

void gf_psitc_options_default(gf_psitc_options *o) {
  if (!o) return;
  o->tolerance = 1.e-6;
  o->max_iterations = 400;
  o->ds_init = 1.;
  o->ds_init_decrease = 4.;
  o->adaptive_restart = 1;
  o->max_recursion_depth = 20;
  o->diffusion_factor = 4.;
  o->global_ds = 0;
  o->ds_safety = 0.1;
  o->ds_ramp = 1.1;
  o->ds_max = 1.e4;
  o->max_factor_line_search = 1.5;
  o->max_allowed_residual = 1.e6;
  o->min_allowable_state_var = -1.e-6;
  o->lu_precision = 64;
  o->log_rate = 0;
}

int gf_flamelet_batch_create(gf_mech *m, int nfl, const int *nzi, double pressure, const double *dz,
                             const double *chi, const double *oxy, const double *fuel, int include_enthalpy_flux,
                             int include_variable_cp, gf_flamelet_batch **out) {
  if (!m || !out || !nzi || !dz || !chi || !oxy || !fuel || nfl < 1) return fail(GF_ERR_ARG, "bad flamelet batch arguments");
  *out = nullptr;
  GF_TRY(ensure_device(m));
  const int ns = gf_mech_n_species(m);
  auto *b = new gf_flamelet_batch();
  b->m = m;
  b->B = nfl;
  b->tr_set.assign(nfl, 0);
  b->ns = ns;
  b->pressure = pressure;
  long P = 0, V = 0, I = 0, C = 0, Jv = 0, Lv = 0, Pv = 0, Z = 0;
  for (int f = 0; f < nfl; ++f) {
    const int n = nzi[f];
    if (n < 2) {
      delete b;
      return fail(GF_ERR_ARG, "a flamelet needs at least two interior points");
    }
    b->nzi.push_back(n);
    b->pt0.push_back(P);
    b->dof0.push_back(V);
    b->ipt0.push_back(I);
    b->chi0.push_back(C);
    b->voff.push_back(Jv);
    b->loff.push_back(Lv);
    b->poff.push_back(Pv);
    b->dz0.push_back(Z);
    double cm = 0.;
    for (int i = 0; i < n + 2; ++i) cm = std::max(cm, chi[C + i]);
    b->diffmax.push_back(cm);
    P += n + 2;
    V += (long)n * ns;
    I += n;
    C += n + 2;
    Jv += (long)ns * ((long)n * ns + 2 * (n - 1));
    Lv += (long)n * ns * ns;
    Pv += (long)n * ns;
    Z += n + 1;
  }
  b->NP = P;
  b->NV = V;
  b->NI = I;
  b->NJ = Jv;
  b->NL = Lv;
  b->NPV = Pv;
  const size_t bytes = carve_all(b, nullptr);
  if (hipMalloc(&b->dmem, bytes) != hipSuccess) {
    delete b;
    return fail(GF_ERR_HIP, "hipMalloc failed for the flamelet batch");
  }
  carve_all(b, static_cast<char *>(b->dmem));
  b->hpin_bytes = ((size_t)P * 4 + (size_t)I * 16 + (size_t)nfl * 128 + 65536) * 2;
  if (hipHostMalloc((void **)&b->hpin, b->hpin_bytes, hipHostMallocDefault) != hipSuccess) {
    hipFree(b->dmem);
    delete b;
    return fail(GF_ERR_HIP, "hipHostMalloc failed for the flamelet batch");
  }
  hipStream_t st = m->stream;
  // padded point states: boundary rows hold the streams, interior rows are set by set_state
  std::vector<double> X((size_t)P * ns, 0.);
  std::vector<long> js(P, -1);
  for (int f = 0; f < nfl; ++f) {
    std::memcpy(&X[(size_t)b->pt0[f] * ns], oxy + (size_t)f * ns, ns * 8);
    std::memcpy(&X[(size_t)(b->pt0[f] + nzi[f] + 1) * ns], fuel + (size_t)f * ns, ns * 8);
    for (int i = 0; i < nzi[f]; ++i) js[b->pt0[f] + 1 + i] = b->voff[f] + (long)i * ns * ns;
  }
  std::vector<double> hdz(dz, dz + Z), hchi(chi, chi + C), pr(1, pressure);
  int rc = GF_OK;
  auto T = [&](int r) { if (rc == GF_OK) rc = r; };
  T(up(b->X, X, st));
  T(up(b->X0, X, st));
  T(up(b->Xt, X, st));
  T(up(b->js_off, js, st));
  T(up(b->dz, hdz, st));
  T(up(b->chi, hchi, st));
  T(up(b->dp, pr, st));
  T(up(b->dnzi, b->nzi, st));
  T(up(b->dpt0, b->pt0, st));
  T(up(b->ddof0, b->dof0, st));
  T(up(b->dipt0, b->ipt0, st));
  T(up(b->dchi0, b->chi0, st));
  T(up(b->dvoff, b->voff, st));
  T(up(b->ddz0, b->dz0, st));
  T(up(b->ekx, std::vector<float>((size_t)V, 0.f), st));
  FlameletBatchView &v = b->view;
  v.B = nfl;
  v.ns = ns;
  v.incl_ef = include_enthalpy_flux;
  v.incl_vcp = include_variable_cp;
  v.nzi = b->dnzi;
  v.pt0 = b->dpt0;
  v.dof0 = b->ddof0;
  v.ipt0 = b->dipt0;
  v.chi0 = b->dchi0;
  v.voff = b->dvoff;
  v.cmaj = b->cmaj;
  v.csub = b->csub;
  v.csup = b->csup;
  v.mco = b->mco;
  v.nco = b->nco;
  v.chi = b->chi;
  v.chem_rhs = b->chem_rhs;
  v.aux = b->aux;
  v.cpi = b->cpi;
  v.NI = b->NI;
  v.NV = V;
  v.NJ = Jv;
  for (int f = 0; f < nfl; ++f) v.maxNJ = std::max(v.maxNJ, (long)ns * ((long)nzi[f] * ns + 2 * (nzi[f] - 1)));
  T(launch_fb_stencils(v, b->dz, b->ddz0, nullptr, b->cmaj, b->csub, b->csup, b->mco, b->nco, st));
  if (rc == GF_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(GF_ERR_HIP, "stream synchronize failed");
  if (rc != GF_OK) {
    hipFree(b->dmem);
    hipHostFree(b->hpin);
    delete b;
    return rc;
  }
  *out = b;
  return GF_OK;
}

int gf_flamelet_batch_destroy(gf_flamelet_batch *b) {
  if (!b) return GF_OK;
  if (b->dmem) hipFree(b->dmem);
  if (b->hl) hipFree(b->hl);
  if (b->ile) hipFree(b->ile);
  if (b->hpin) hipHostFree(b->hpin);
  delete b;
  return GF_OK;
}

int gf_flamelet_batch_set_heat_loss(gf_flamelet_batch *b, const double *T_conv, const double *h_conv,
                                    const double *T_rad, const double *h_rad, int use_scaled_heat_loss) {
  if (!b) return fail(GF_ERR_ARG, "null flamelet batch");
  const hipStream_t st = b->m->stream;
  if (!T_conv && !h_conv && !T_rad && !h_rad) {
    b->view.adiabatic = 1;
    b->view.hl = nullptr;
    return GF_OK;
  }
  if (!T_conv || !h_conv || !T_rad || !h_rad) return fail(GF_ERR_ARG, "all four heat-loss arrays are needed");
  const size_t n = (size_t)b->NI;
  if (!b->hl && hipMalloc(&b->hl, 4 * n * sizeof(double)) != hipSuccess)
    return fail(GF_ERR_HIP, "hipMalloc failed for the heat-loss parameters");
  const double *src[4] = {T_conv, h_conv, T_rad, h_rad};
  for (int k = 0; k < 4; ++k)
    if (hipMemcpyAsync(b->hl + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(GF_ERR_HIP, "host-to-device copy failed");
  if (hipStreamSynchronize(st) != hipSuccess) return fail(GF_ERR_HIP, "stream synchronize failed");
  b->view.hl = b->hl;
  b->view.adiabatic = 0;
  b->view.scaled_hl = use_scaled_heat_loss ? 1 : 0;
  return GF_OK;
}

int gf_flamelet_batch_set_lewis_numbers(gf_flamelet_batch *b, const double *inv_lewis) {
  if (!b) return fail(GF_ERR_ARG, "null flamelet batch");
  const hipStream_t st = b->m->stream;
  const double *dle = nullptr;
  if (inv_lewis) {
    const size_t n = (size_t)b->B * b->ns;
    if (!b->ile && hipMalloc(&b->ile, n * sizeof(double)) != hipSuccess)
      return fail(GF_ERR_HIP, "hipMalloc failed for the Lewis numbers");
    if (hipMemcpyAsync(b->ile, inv_lewis, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(GF_ERR_HIP, "host-to-device copy failed");
    dle = b->ile;
  }
  GF_TRY(launch_fb_stencils(b->view, b->dz, b->ddz0, dle, b->cmaj, b->csub, b->csup, b->mco, b->nco, st));
  return hipStreamSynchronize(st) == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stream synchronize failed");
}

int gf_flamelet_batch_set_state(gf_flamelet_batch *b, const double *interior) {
  if (!b || !interior) return fail(GF_ERR_ARG, "null argument");
  hipStream_t st = b->m->stream;
  for (int f = 0; f < b->B; ++f) {
    const size_t n = (size_t)b->nzi[f] * b->ns * 8;
    const double *src = interior + b->dof0[f];
    if (hipMemcpyAsync(b->X + (b->pt0[f] + 1) * b->ns, src, n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(b->X0 + (b->pt0[f] + 1) * b->ns, src, n, hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(GF_ERR_HIP, "host-to-device copy failed");
  }
  return hipStreamSynchronize(st) == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stream synchronize failed");
}

int gf_flamelet_batch_get_state(gf_flamelet_batch *b, double *interior) {
  if (!b || !interior) return fail(GF_ERR_ARG, "null argument");
  hipStream_t st = b->m->stream;
  for (int f = 0; f < b->B; ++f)
    if (hipMemcpyAsync(interior + b->dof0[f], b->X + (b->pt0[f] + 1) * b->ns, (size_t)b->nzi[f] * b->ns * 8,
                       hipMemcpyDeviceToHost, st) != hipSuccess)
      return fail(GF_ERR_HIP, "device-to-host copy failed");
  return hipStreamSynchronize(st) == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stream synchronize failed");
}

int gf_flamelet_batch_rhs(gf_flamelet_batch *b, double *out) {
  if (!b || !out) return fail(GF_ERR_ARG, "null argument");
  Psitc s{b, {}, b->m->stream};
  std::vector<int> all;
  for (int f = 0; f < b->B; ++f) all.push_back(f);
  std::vector<double> res;
  GF_TRY(s.rhs(b->X, all, b->R, res, nullptr));
  std::vector<double> h;
  GF_TRY(gfa_down(h, b->R, b->NV, b->m->stream));
  std::memcpy(out, h.data(), h.size() * 8);
  return GF_OK;
}

namespace {
int steady(gf_flamelet_batch *b, const gf_psitc_options *opts, int newton, const int *select, int *iterations,
           int *converged, double *residual, double *min_ds, long long *counters) {
  if (!b) return fail(GF_ERR_ARG, "null flamelet batch");
  gf_psitc_options o;
  gf_psitc_options_default(&o);
  if (opts) o = *opts;
  if (o.lu_precision != 64 && o.lu_precision != 32) return fail(GF_ERR_ARG, "lu_precision must be 64 or 32");
  std::vector<double> d(b->B);
  for (int f = 0; f < b->B; ++f) d[f] = o.diffusion_factor * b->diffmax[f];
  GF_TRY(up(b->fdiff, d, b->m->stream));
  Psitc s{b, o, b->m->stream};
  s.newton = newton;
  s.pick = select;
  GF_TRY(s.run(iterations, converged, residual, min_ds));
  if (counters) {
    counters[0] = s.nit;
    counters[1] = s.njac;
    counters[2] = s.nrhs;
    counters[3] = s.nsync;
  }
  return GF_OK;
}
}  // namespace

int gf_flamelet_batch_steady_psitc(gf_flamelet_batch *b, const gf_psitc_options *opts, const int *select,
                                   int *iterations, int *converged, double *residual, double *min_ds,
                                   long long *counters) {
  return steady(b, opts, 0, select, iterations, converged, residual, min_ds, counters);
}

int gf_flamelet_batch_steady_newton(gf_flamelet_batch *b, const gf_psitc_options *opts, const int *select,
                                    int *iterations, int *converged, double *residual, long long *counters) {
  return steady(b, opts, 1, select, iterations, converged, residual, nullptr, counters);
}

int gf_flamelet_batch_transient_rhs(gf_flamelet_batch *b, const int *flags, const double *x, double *rhs) {
  if (!b || !flags || !x || !rhs) return fail(GF_ERR_ARG, "null argument");
  Psitc s{b, {}, b->m->stream};
  std::vector<int> fl;
  GF_TRY(s.flagged(flags, fl));
  if (fl.empty()) return GF_OK;
  GF_TRY(launch_fb_unpack(b->view, b->flag, b->ddof0, x, b->X, 1, s.st));
  GF_TRY(s.chem(b->X, fl, false));
  GF_TRY(launch_fb_rhs(b->view, b->X, b->flag, rhs, s.st));
  return s.sync();
}

int gf_flamelet_batch_transient_setup(gf_flamelet_batch *b, const int *flags, const double *x,
                                      const double *prefactor, int lu_precision) {
  if (!b || !flags || !x || !prefactor) return fail(GF_ERR_ARG, "null argument");
  if (lu_precision != 64 && lu_precision != 32) return fail(GF_ERR_ARG, "lu_precision must be 64 or 32");
  if (b->tr_prec != lu_precision && std::any_of(b->tr_set.begin(), b->tr_set.end(), [](char c) { return c != 0; }))
    return fail(GF_ERR_STATE, "transient_setup precision differs from the factors already held");
  Psitc s{b, {}, b->m->stream};
  std::vector<int> fl;
  GF_TRY(s.flagged(flags, fl));
  if (fl.empty()) return GF_OK;
  GF_TRY(launch_fb_unpack(b->view, b->flag, b->ddof0, x, b->X, 1, s.st));
  GF_TRY(s.chem(b->X, fl, true));
  GF_TRY(s.up(b->fval, std::vector<double>(prefactor, prefactor + b->B)));
  GF_TRY(launch_fb_matrix(b->m, b->view, b->X, b->flag, 2, b->fval, 0., b->J, s.st));
  GF_TRY(s.factor(fl, lu_precision));
  b->tr_prec = lu_precision;
  for (int f : fl) b->tr_set[f] = 1;
  return s.sync();
}

int gf_flamelet_batch_transient_solve(gf_flamelet_batch *b, const int *flags, const double *r, double *out) {
  if (!b || !flags || !r || !out) return fail(GF_ERR_ARG, "null argument");
  Psitc s{b, {}, b->m->stream};
  std::vector<int> fl;
  for (int f = 0; f < b->B; ++f)
    if (flags[f]) {
      if (!b->tr_set[f]) return fail(GF_ERR_STATE, "transient_solve before transient_setup for a flamelet");
      fl.push_back(f);
    }
  if (fl.empty()) return GF_OK;
  GF_TRY(s.systems(fl));
  if (b->tr_prec == 32)
    GF_TRY(launch_btd_inv_solve<float>((int)fl.size(), b->sys_nb, b->ns, b->J, b->sys_v, b->D32, b->sys_l, r, out,
                                       b->sys_x, s.st));
  else
    GF_TRY(launch_btd_inv_solve<double>((int)fl.size(), b->sys_nb, b->ns, b->J, b->sys_v, b->L, b->sys_l, r, out,
                                        b->sys_x, s.st));
  return s.sync();
}

}  // extern "C"
