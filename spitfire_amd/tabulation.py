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
equilibrium solver, which this implementation does not provide (build_adiabatic_eq_library raises for the same
reason).
"""
import copy
import time

import numpy as np

from .flamelet import Flamelet, FlameletBatch, FlameletSpec, compute_dissipation_rate
from .library import Dimension, Library

_mixture_fraction_name = 'mixture_fraction'
_dissipation_rate_name = 'dissipation_rate'
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
