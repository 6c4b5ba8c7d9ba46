"""Flamelet libraries (tabulated chemistry), mirroring src/spitfire/chemistry/tabulation.py.

build_adiabatic_slfm_library (tabulation.py:229-336) solves a steady flamelet for each dissipation rate and stores
temperature, pressure and mass fractions over (mixture fraction, stoichiometric dissipation rate). The reference
walks the dissipation rates in order, each solve (compute_steady_state with pseudo-transient continuation at
tolerance 1e-6) starting from the previous converged state, and stops at the first extinguished flamelet (maximum of
T - T_linear below 10 K).

On the MI355X the flamelets of a library are solved together (FlameletBatch, one set of batched kernels per
iteration) in one of two ways, selected by `continuation`:
  * 'sequential': the reference's order exactly, one flamelet per solve, each from its predecessor's state;
  * 'batched' (default): every dissipation rate after the first is solved in one batch from the first converged
    state. Each result is then checked in the reference's order: a flamelet that did not converge, or that is
    extinguished while its predecessor burns, is solved again from its predecessor's state as the reference would;
    only if that one is extinguished too does the library stop there. Both paths converge every flamelet to the
    reference's tolerance on the reference's continuation branch, so their tables agree with the reference's to
    that tolerance.

The first flamelet starts from the Burke-Schumann state: the reference's 'equilibrium' start needs Cantera's
equilibrium solver, which this implementation does not provide (build_adiabatic_eq_library and
build_nonadiabatic_defect_eq_library raise for the same reason).

Nonadiabatic (enthalpy-defect) libraries (tabulation.py:116-226, 339-838) add a stoichiometric enthalpy-defect
dimension to the SLFM library:
  * Burke-Schumann: the Burke-Schumann composition at enthalpies lowered by a triangular defect profile;
  * steady: for each dissipation rate, steady flamelets with convective heat loss raised step by step (adaptive
    steps) until extinction. The dissipation rates are independent, so each step solves the still-burning
    flamelets of all of them in one batch; each flamelet's own sequence of heat-loss values is the reference's;
  * transient: for each dissipation rate, the adiabatic flamelet integrated in time with heat loss until its
    temperature falls to the boundary temperatures (Flamelet.integrate_for_heat_loss, one flamelet at a time on
    the GPU kernels);
then the history is subsampled every h_stoich_spacing of stoichiometric enthalpy and interpolated (piecewise
linearly, constant outside the range) onto n_defect_st stoichiometric defects spanning all of it.
num_procs (the reference's process pool) is accepted and not used: the GPU batches instead.
"""
import copy
import time

import numpy as np

from . import analysis as sca

from .flamelet import Flamelet, FlameletBatch, FlameletSpec, compute_dissipation_rate
from .library import Dimension, Library

_mixture_fraction_name = 'mixture_fraction'
_dissipation_rate_name = 'dissipation_rate'
_enthalpy_defect_name = 'enthalpy_defect'
_stoich_suffix = '_stoich'


def _spec_copy(flamelet_specs):
    return FlameletSpec(**flamelet_specs) if isinstance(flamelet_specs, dict) else copy.copy(flamelet_specs)


def _header(kind, mech, fuel, oxy, verbose):
    if verbose:
        print('-' * 82)
        print('building %s library' % kind)
        print('-' * 82)
        print('- %d species, %d reactions' % (mech.n_species, mech.n_reactions))
        print('- stoichiometric mixture fraction: %.3f' % mech.stoich_mixture_fraction(fuel, oxy))
        print('-' * 82)
    return time.perf_counter()


def _footer(t0, verbose):
    if verbose:
        print('-' * 82)
        print('library built in %6.2f s' % (time.perf_counter() - t0))
        print('-' * 82, flush=True)


def build_unreacted_library(flamelet_specs, verbose=True):
    """linear enthalpy and mass fraction profiles over mixture fraction (tabulation.py:55-72)"""
    fs = _spec_copy(flamelet_specs)
    fs.initial_condition = 'unreacted'
    f = Flamelet(fs)
    return f.make_library_from_interior_state(f.initial_interior_state)


def build_adiabatic_bs_library(flamelet_specs, verbose=True):
    """the Burke-Schumann (complete one-step combustion) state over mixture fraction (tabulation.py:96-113)"""
    fs = _spec_copy(flamelet_specs)
    fs.initial_condition = 'Burke-Schumann'
    f = Flamelet(fs)
    return f.make_library_from_interior_state(f.initial_interior_state)


def build_adiabatic_eq_library(flamelet_specs, verbose=True):
    """tabulation.py:75-93 (Cantera equilibrium over mixture fraction): not available"""
    raise NotImplementedError('equilibrium libraries need Cantera\'s equilibrium solver, which this implementation '
                              'does not provide')


def _extinguished(f):
    return np.max(f.current_temperature - f.linear_temperature) < 10.


def _solve(specs, lu_precision, tolerance, solver_verbose):
    """compute_steady_state of the given flamelets in one batch; returns the Flamelets and their converged flags"""
    fl = [Flamelet(s) for s in specs]
    fb = FlameletBatch(fl)
    _, conv, _ = fb.compute_steady_state(tolerance=tolerance, lu_precision=lu_precision, raise_on_failure=False,
                                         log_rate=1 if solver_verbose else 0)
    return fl, np.asarray(conv, dtype=bool)


def build_adiabatic_slfm_library(flamelet_specs, diss_rate_values=np.logspace(-3, 2, 16),
                                 diss_rate_ref='stoichiometric', verbose=True, solver_verbose=False,
                                 _return_intermediates=False, include_extinguished=False, diss_rate_log_scaled=True,
                                 continuation='batched', lu_precision=64, tolerance=1.e-6):
    """Adiabatic strained laminar flamelet library (tabulation.py:229-336): temperature, pressure and mass fractions
    over (mixture_fraction, dissipation_rate_stoich), with the mechanism in extra_attributes['mech_spec'].

    diss_rate_values are stoichiometric (diss_rate_ref='stoichiometric') or maximum ('maximum') dissipation rates in
    increasing order; the table stops before the first extinguished flamelet unless include_extinguished is set.
    continuation: 'batched' or 'sequential' (see the module docstring); lu_precision: 64, or 32 for the mixed
    precision block factorization (Newton still converges the FP64 residual to the tolerance)."""
    if continuation not in ('batched', 'sequential'):
        raise ValueError('continuation must be "batched" or "sequential"')
    if diss_rate_ref not in ('stoichiometric', 'maximum'):
        raise ValueError('diss_rate_ref must be "stoichiometric" or "maximum"')
    base = _spec_copy(flamelet_specs)
    m, fuel, oxy = base.mech_spec, base.fuel_stream, base.oxy_stream
    base.initial_condition = 'Burke-Schumann'
    chis = np.asarray(diss_rate_values, dtype=np.float64).ravel()
    t00 = _header('adiabatic SLFM', m, fuel, oxy, verbose)

    def spec_at(chi, ic):
        s = copy.copy(base)
        s.initial_condition = ic
        s.dissipation_rate = None
        if diss_rate_ref == 'maximum':
            s.max_dissipation_rate, s.stoich_dissipation_rate = float(chi), None
        else:
            s.stoich_dissipation_rate, s.max_dissipation_rate = float(chi), None
        return s

    def one(chi, ic):
        fl, conv = _solve([spec_at(chi, ic)], lu_precision, tolerance, solver_verbose)
        return fl[0], bool(conv[0])

    nchi = chis.size
    results = [None] * nchi
    f0, ok0 = one(chis[0], 'Burke-Schumann')
    results[0] = (f0, ok0)
    if continuation == 'batched' and nchi > 1:
        ic = f0.current_interior_state
        fl, conv = _solve([spec_at(c, ic) for c in chis[1:]], lu_precision, tolerance, solver_verbose)
        for k in range(1, nchi):
            results[k] = (fl[k - 1], bool(conv[k - 1]))

    accepted = []
    for k in range(nchi):
        t0 = time.perf_counter()
        if verbose:
            print('%4d/%4d (chi%s = %8.1e 1/s) ' % (k + 1, nchi, _stoich_suffix if diss_rate_ref == 'stoichiometric'
                                                   else '_max', chis[k]), end='', flush=True)
        if k > 0:
            prev = accepted[-1]
            f, ok = results[k] if results[k] is not None else (None, False)
            if f is None or not ok or (_extinguished(f) and not _extinguished(prev)):
                f, ok = one(chis[k], prev.current_interior_state)
        else:
            f, ok = results[0]
        if not ok:
            raise RuntimeError('the flamelet at dissipation rate %g did not converge' % chis[k])
        if _extinguished(f) and not include_extinguished:
            if verbose:
                print(' extinction detected, stopping. The extinguished state will not be included in the table.')
            break
        if verbose:
            print(' converged in %6.2f s, T_max = %6.1f' % (time.perf_counter() - t0, np.max(f.current_temperature)))
        accepted.append(f)

    if not accepted:
        raise RuntimeError('the flamelet at the first dissipation rate is extinguished; no library to build')
    zst = m.stoich_mixture_fraction(fuel, oxy)
    x_values = np.array([compute_dissipation_rate(np.array([zst]), f._max_dissipation_rate,
                                                  f._dissipation_rate_form)[0] for f in accepted])
    tables = [f.make_library_from_interior_state(f.current_interior_state) for f in accepted]
    z = accepted[0].mixfrac_grid
    if _return_intermediates:
        table_dict = {}
        for x, f, t in zip(x_values, accepted, tables):
            table_dict[x] = {k: t[k].ravel() for k in t.props}
            table_dict[x]['adiabatic_state'] = np.copy(f.current_interior_state)
        _footer(t00, verbose)
        return table_dict, z, x_values
    out = Library(Dimension(_mixture_fraction_name, z),
                  Dimension(_dissipation_rate_name + _stoich_suffix, x_values, diss_rate_log_scaled))
    out.extra_attributes['mech_spec'] = m
    for q in tables[0].props:
        out[q] = out.get_empty_dataset()
        for ix, t in enumerate(tables):
            out[q][:, ix] = t[q]
    _footer(t00, verbose)
    return out


# ---------------- nonadiabatic (enthalpy defect) libraries ----------------
def _enthalpy(mech, lib):
    return sca.compute_specific_enthalpy(mech, lib)['enthalpy']


def _build_nonadiabatic_defect_unstrained_library(initialization, flamelet_specs, n_defect_st=16, verbose=True):
    """tabulation.py:116-186: the initial state's composition at enthalpies h_ad + g f(z), f the triangle that is 1 at
    the stoichiometric mixture fraction and 0 at the streams, g from 0 down to the defect that brings the
    stoichiometric point to the linear temperature profile; the temperature from each enthalpy at the composition
    held fixed (the reference's Cantera HPY state)"""
    if initialization == 'equilibrium':
        raise NotImplementedError('equilibrium libraries need Cantera\'s equilibrium solver, which this '
                                  'implementation does not provide')
    fs = _spec_copy(flamelet_specs)
    m, fuel, oxy = fs.mech_spec, fs.fuel_stream, fs.oxy_stream
    zst = m.stoich_mixture_fraction(fuel, oxy)
    fs.initial_condition = initialization
    f = Flamelet(fs)
    ns = m.n_species
    state_ad = f.initial_interior_state
    ad_lib = f.make_library_from_interior_state(state_ad)
    h_ad = _enthalpy(m, ad_lib)
    z = f.mixfrac_grid
    state_ce = state_ad.copy()
    state_ce[::ns] = z[1:-1] * fuel.T + (1. - z[1:-1]) * oxy.T
    h_ce = _enthalpy(m, f.make_library_from_interior_state(state_ce))
    defect_ext = np.interp(zst, z, h_ad) - np.interp(zst, z, h_ce)
    defect_range = np.linspace(-defect_ext, 0., n_defect_st)[::-1]
    out = Library(Dimension(_mixture_fraction_name, z), Dimension(_enthalpy_defect_name + _stoich_suffix,
                                                                   defect_range))
    out.extra_attributes['mech_spec'] = m
    for q in ad_lib.props + ['enthalpy_defect', 'enthalpy_cons', 'enthalpy', _mixture_fraction_name]:
        out[q] = out.get_empty_dataset()
    fz = np.where(z <= zst, z / zst, (1. - z) / (1. - zst))
    names = m.species_names
    Y = np.stack([ad_lib['mass fraction ' + s] for s in names], axis=1)
    Yi = np.maximum(Y[1:-1], 0.)
    Yi /= Yi.sum(axis=1, keepdims=True)  # as a Cantera HPY state normalizes them
    g = f.make_library_from_interior_state(state_ad)
    for ig in range(n_defect_st):
        hd = h_ad + defect_range[ig] * fz
        g['temperature'][1:-1] = m.temperatures_from_enthalpies(hd[1:-1], Yi, T0=g['temperature'][1:-1])
        for k, s in enumerate(names):
            g['mass fraction ' + s][1:-1] = Yi[:, k]
        for q in g.props:
            out[q][:, ig] = g[q]
        out['enthalpy_defect'][:, ig] = hd - h_ad
        out['enthalpy_cons'][:, ig] = h_ad
        out['enthalpy'][:, ig] = hd
        out[_mixture_fraction_name][:, ig] = z
    return out


def build_nonadiabatic_defect_eq_library(flamelet_specs, n_defect_st=16, verbose=True):
    """tabulation.py:189-206 (equilibrium with a triangular enthalpy defect): not available"""
    return _build_nonadiabatic_defect_unstrained_library('equilibrium', flamelet_specs, n_defect_st, verbose)


def build_nonadiabatic_defect_bs_library(flamelet_specs, n_defect_st=16, verbose=True):
    """Burke-Schumann library with heat loss through a triangular enthalpy defect (tabulation.py:209-226)"""
    return _build_nonadiabatic_defect_unstrained_library('Burke-Schumann', flamelet_specs, n_defect_st, verbose)


def _subsample_defects(mech, hist_lib, z, zst, time_major, h_stoich_spacing, managed, chi_st, last_always):
    """entries (chi_st, g_st) of the unstructured table: every state of a heat-loss history whose stoichiometric
    enthalpy fell by more than h_stoich_spacing since the last one kept (tabulation.py:381-404, 489-512)"""
    h = _enthalpy(mech, hist_lib)
    if not time_major:
        h = h.T
    h_ad = h[0]
    nt = h.shape[0]
    last = np.interp(zst, z, h_ad)
    idx = [0]
    for i in range(nt if last_always else nt - 1):
        this = np.interp(zst, z, h[i])
        if last - this > h_stoich_spacing:
            idx.append(i)
            last = this
    if last_always and nt - 1 not in idx:
        idx.append(nt - 1)
    for i in idx:
        defect = h[i] - h_ad
        gst = float(np.interp(zst, z, defect))
        d = {'enthalpy_defect': np.copy(defect), 'enthalpy_cons': np.copy(h_ad), 'enthalpy': np.copy(h[i]),
             _mixture_fraction_name: z}
        for q in hist_lib.props:
            d[q] = np.copy(hist_lib[q][i, :] if time_major else hist_lib[q][:, i])
        managed[(chi_st, gst)] = d


def _expand_transient(chi_st, managed, base, table_dict, h_stoich_spacing, verbose, integration_args,
                      solver_verbose):
    """tabulation.py:339-405: the adiabatic flamelet integrated with heat loss (scaled by the temperature range and
    the dissipation rate, linear reference temperature, convection coefficient 1e7 unless given, no radiation)
    until it is close to the boundary temperatures; the integration is retried with a 100x tighter tolerance when
    it fails"""
    fs = copy.copy(base)
    fs.initial_condition = table_dict[chi_st]['adiabatic_state']
    fs.stoich_dissipation_rate, fs.max_dissipation_rate, fs.dissipation_rate = chi_st, None, None
    fs.heat_transfer = 'nonadiabatic'
    fs.scale_heat_loss_by_temp_range = True
    fs.scale_convection_by_dissipation = True
    fs.use_linear_ref_temp_profile = True
    fs.convection_coefficient = fs.convection_coefficient if fs.convection_coefficient is not None else 1.e7
    fs.radiative_emissivity = 0.
    args = {'first_time_step': 1.e-9, 'max_time_step': 1.e-1, 'write_log': solver_verbose, 'log_rate': 100,
            'print_exception_on_failure': False}
    args.update(integration_args or {})
    args.setdefault('transient_tolerance', 1.e-8)
    t0 = time.perf_counter()
    hist, f = None, None
    while hist is None and args['transient_tolerance'] > 1.e-15:
        try:
            f = Flamelet(fs)
            hist = f.integrate_for_heat_loss(**args)
        except Exception:
            args['transient_tolerance'] *= 1.e-2
    if hist is None:
        raise RuntimeError('heat-loss integration failed at chi_st = %g' % chi_st)
    m = fs.mech_spec
    zst = m.stoich_mixture_fraction(f.fuel_stream, f.oxy_stream)
    _subsample_defects(m, hist, f.mixfrac_grid, zst, True, h_stoich_spacing, managed, chi_st, True)
    if verbose:
        print('chi_st = %8.1e 1/s converged in %6.2f s' % (chi_st, time.perf_counter() - t0), flush=True)


def _expand_steady(chis, managed, base, table_dict, h_stoich_spacing, verbose, lu_precision, solver_verbose):
    """tabulation.py:408-519 for every dissipation rate at once: steady flamelets with convective heat loss
    h_conv = hval (unscaled, linear reference temperature, no radiation), hval raised by adaptive steps
    dh *= min(max(sqrt(0.1 / diff), 0.1), 2) with diff the largest relative change of the state, until the
    flamelet is extinguished (max(state) < max(T_oxy, T_fuel) + 10) or hval reaches 1e10. Each round solves the
    next heat-loss value of every dissipation rate still running in one batch."""
    t0 = time.perf_counter()
    m = base.mech_spec
    run = {}
    for c in chis:
        st = table_dict[c]['adiabatic_state']
        run[c] = dict(hval=0., dh=1.e-1, state_old=np.copy(st), current=np.copy(st), hvalues=[0.],
                      solutions=[{p: v for p, v in table_dict[c].items() if p != 'adiabatic_state'}])
    active = list(chis)
    f_last = {}
    while active:
        specs = []
        for c in active:
            r = run[c]
            r['hval'] += r['dh']
            fs = copy.copy(base)
            fs.stoich_dissipation_rate, fs.max_dissipation_rate, fs.dissipation_rate = c, None, None
            fs.heat_transfer = 'nonadiabatic'
            fs.scale_heat_loss_by_temp_range = False
            fs.scale_convection_by_dissipation = False
            fs.use_linear_ref_temp_profile = True
            fs.radiative_emissivity = 0.
            fs.convection_coefficient = float(r['hval'])
            fs.initial_condition = r['current']
            specs.append(fs)
        fl = [Flamelet(s) for s in specs]
        FlameletBatch(fl).compute_steady_state(tolerance=1.e-6, lu_precision=lu_precision, raise_on_failure=False,
                                               log_rate=1 if solver_verbose else 0)
        still = []
        for c, f in zip(active, fl):
            r = run[c]
            cur = f.current_interior_state
            maxT = np.max(cur)
            diff = np.max(np.abs(cur - r['state_old']) / (np.abs(cur) + 1.e-4))
            extinguished = maxT < max(f.oxy_stream.T, f.fuel_stream.T) + 10.
            r['current'] = cur
            r['state_old'] = np.copy(cur)
            r['dh'] *= min(max(np.sqrt(1.e-1 / diff), 0.1), 2.)
            r['hvalues'].append(r['hval'])
            lib = f.make_library_from_interior_state(cur)
            r['solutions'].append({p: lib[p].ravel() for p in lib.props})
            f_last[c] = f
            if not extinguished and r['hval'] < 1.e10:
                still.append(c)
        active = still
    for c in chis:
        r = run[c]
        f = f_last[c]
        z = f.mixfrac_grid
        lib = Library(Dimension(_mixture_fraction_name, z),
                      Dimension(_enthalpy_defect_name + _stoich_suffix, np.array(r['hvalues'])))
        lib.extra_attributes['mech_spec'] = m
        for p in r['solutions'][0]:
            lib[p] = lib.get_empty_dataset()
            for ig, sol in enumerate(r['solutions']):
                lib[p][:, ig] = np.ravel(sol[p])
        zst = m.stoich_mixture_fraction(f.fuel_stream, f.oxy_stream)
        _subsample_defects(m, lib, z, zst, False, h_stoich_spacing, managed, c, False)
    if verbose:
        print('enthalpy defect dimension expanded in %6.2f s' % (time.perf_counter() - t0), flush=True)


def _interpolate_to_structured_defect_dimension(unstructured, n_defect_stoich, extend=False):
    """tabulation.py:594-665: each quantity of each dissipation rate as a function of the stoichiometric defect,
    piecewise linear and constant outside its range (linear extrapolation for the enthalpies when extend is set),
    at n_defect_stoich defects spanning all dissipation rates (two more, below, when extend is set). Returns the
    table keyed (chi_st, g_st), the sorted dissipation rates and the defects from largest to smallest."""
    gs = np.array([g for _, g in unstructured])
    gmin, gmax = float(np.min(gs)), float(np.max(gs))
    grid = np.linspace(gmin, gmax, n_defect_stoich)
    if extend:
        sp = abs(grid[1] - grid[0])
        grid = np.linspace(gmin - 2 * sp, gmax, n_defect_stoich + 2)
    by_chi = {}
    for (c, g) in unstructured:
        by_chi.setdefault(c, set()).add(g)
    out = {}
    for c, gset in by_chi.items():
        gsorted = np.sort(np.array(list(gset)))
        first = unstructured[(c, gsorted[0])]
        for q in first:
            Yq = np.stack([np.asarray(unstructured[(c, g)][q], dtype=np.float64) for g in gsorted], axis=0)
            if extend and q in ('enthalpy', 'enthalpy_defect') and gsorted.size > 1:
                lo = (Yq[1] - Yq[0]) / (gsorted[1] - gsorted[0])
                hi = (Yq[-1] - Yq[-2]) / (gsorted[-1] - gsorted[-2])
            for gs_ in grid:
                if gsorted.size == 1:
                    v = Yq[0].copy()
                else:
                    k = np.clip(np.searchsorted(gsorted, gs_) - 1, 0, gsorted.size - 2)
                    w = (gs_ - gsorted[k]) / (gsorted[k + 1] - gsorted[k])
                    if extend and q in ('enthalpy', 'enthalpy_defect'):
                        v = Yq[k] + w * (Yq[k + 1] - Yq[k])
                        v = np.where(gs_ < gsorted[0], Yq[0] + (gs_ - gsorted[0]) * lo, v)
                        v = np.where(gs_ > gsorted[-1], Yq[-1] + (gs_ - gsorted[-1]) * hi, v)
                    else:
                        w = min(max(w, 0.), 1.)
                        v = Yq[k] + w * (Yq[k + 1] - Yq[k])
                if q == 'density' and np.any(v < 1.e-14):
                    raise ValueError('density < 1.e-14 detected!')
                if q == 'temperature' and np.any(v < 1.e-14):
                    raise ValueError('temperature < 1.e-14 detected!')
                out.setdefault((c, gs_), {})[q] = v
    return out, np.array(sorted(by_chi)), grid[::-1]


def _build_nonadiabatic_defect_slfm_library(flamelet_specs, heat_loss_expansion, diss_rate_values, diss_rate_ref,
                                            verbose, solver_verbose, h_stoich_spacing, num_procs, integration_args,
                                            n_defect_st, extend_defect_dim, diss_rate_log_scaled, lu_precision):
    """tabulation.py:668-724"""
    base = _spec_copy(flamelet_specs)
    m, fuel, oxy = base.mech_spec, base.fuel_stream, base.oxy_stream
    t00 = _header('nonadiabatic (defect) SLFM', m, fuel, oxy, verbose)
    table_dict, _, _ = build_adiabatic_slfm_library(base, diss_rate_values, diss_rate_ref, verbose, solver_verbose,
                                                    _return_intermediates=True, lu_precision=lu_precision)
    if verbose:
        print('expanding (%s) enthalpy defect dimension ...' % heat_loss_expansion, flush=True)
    unstructured = {}
    base.initial_condition = 'Burke-Schumann'
    if heat_loss_expansion == 'transient':
        for c in table_dict:
            _expand_transient(c, unstructured, base, table_dict, h_stoich_spacing, verbose, integration_args,
                              solver_verbose)
    else:
        _expand_steady(list(table_dict), unstructured, base, table_dict, h_stoich_spacing, verbose, lu_precision,
                       solver_verbose)
    table, x_values, g_values = _interpolate_to_structured_defect_dimension(unstructured, n_defect_st,
                                                                            extend=extend_defect_dim)
    key0 = next(iter(table))
    z = table[key0][_mixture_fraction_name]
    out = Library(Dimension(_mixture_fraction_name, z),
                  Dimension(_dissipation_rate_name + _stoich_suffix, x_values, diss_rate_log_scaled),
                  Dimension(_enthalpy_defect_name + _stoich_suffix, g_values))
    out.extra_attributes['mech_spec'] = m
    for q in table[key0]:
        vals = out.get_empty_dataset()
        for ix, x in enumerate(x_values):
            for ig, g in enumerate(g_values):
                vals[:, ix, ig] = table[(x, g)][q]
        out[q] = vals
    _footer(t00, verbose)
    return out


def build_nonadiabatic_defect_transient_slfm_library(flamelet_specs, diss_rate_values=np.logspace(-3, 2, 16),
                                                     diss_rate_ref='stoichiometric', verbose=True,
                                                     solver_verbose=False, h_stoich_spacing=10.e3, num_procs=1,
                                                     integration_args=None, n_defect_st=32, extend_defect_dim=False,
                                                     diss_rate_log_scaled=True, lu_precision=64):
    """SLFM library with an enthalpy-defect dimension from transient heat-loss extinction (tabulation.py:727-783)"""
    return _build_nonadiabatic_defect_slfm_library(flamelet_specs, 'transient', diss_rate_values, diss_rate_ref,
                                                   verbose, solver_verbose, h_stoich_spacing, num_procs,
                                                   integration_args, n_defect_st, extend_defect_dim,
                                                   diss_rate_log_scaled, lu_precision)


def build_nonadiabatic_defect_steady_slfm_library(flamelet_specs, diss_rate_values=np.logspace(-3, 2, 16),
                                                  diss_rate_ref='stoichiometric', verbose=True, solver_verbose=False,
                                                  h_stoich_spacing=10.e3, num_procs=1, integration_args=None,
                                                  n_defect_st=32, extend_defect_dim=False, diss_rate_log_scaled=True,
                                                  lu_precision=64):
    """SLFM library with an enthalpy-defect dimension from quasisteady heat-loss extinction (tabulation.py:786-838)"""
    return _build_nonadiabatic_defect_slfm_library(flamelet_specs, 'steady', diss_rate_values, diss_rate_ref,
                                                   verbose, solver_verbose, h_stoich_spacing, num_procs,
                                                   integration_args, n_defect_st, extend_defect_dim,
                                                   diss_rate_log_scaled, lu_precision)
