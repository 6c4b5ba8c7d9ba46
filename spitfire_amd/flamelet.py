"""Steady nonpremixed flamelets: FlameletSpec / Flamelet mirroring src/spitfire/chemistry/flamelet.py.

Grid (flamelet.py:283-341), dissipation rate (240-266), initial conditions (486-597), the griffon RHS /
Jacobian wrappers (642-762), block-Thomas setup and solve (756-762, 911-915), steady_solve_newton
(1352-1485), steady_solve_psitc (1487-1703) and compute_steady_state (1705-1778), running on the MI355X
griffon drop-in (spitfire_amd.griffon).

Differences from the reference, all documented here:
  * streams are spitfire_amd.mechanism.Stream objects (no Cantera); the 'equilibrium' initial condition
    is not available (it needs Cantera's equilibrate);
  * the pseudo-transient eigenvalue estimate `expeig` (flamelet.py:1597-1604; LAPACK dgeev per block in
    the reference) comes from the GPU eigenvalue kernel (balancing + Householder Hessenberg reduction +
    Francis double-shift QR), which agrees with dgeev to rounding, not bitwise;
  * compute_steady_state falls back from Newton to psitc; the ESDIRK fallback is not built.
"""
import numpy as np
from scipy.special import erfinv


def compute_dissipation_rate(mixture_fraction, max_dissipation_rate, form='Peters'):
    """chi(Z) = chi_max exp(-2 erfinv(2Z-1)^2) (Peters), or constant (flamelet.py:240-266)."""
    if form in ('Peters', 'peters'):
        return max_dissipation_rate * np.exp(-2. * (erfinv(2. * mixture_fraction - 1.)) ** 2)
    x = np.zeros_like(mixture_fraction)
    x[:] = max_dissipation_rate
    return x


def uniform_grid(grid_points):
    z = np.linspace(0., 1., grid_points)
    return z, z[1:] - z[:-1]


def clustered_grid(grid_points, grid_cluster_point, grid_cluster_intensity=6.):
    """Anderson's sinh clustering about a point (flamelet.py:300-341)."""
    if grid_cluster_intensity < 1.e-16:
        raise ValueError('cluster_coeff must be strictly positive! Given value: ' + str(grid_cluster_intensity))
    if grid_cluster_point < 0. or grid_cluster_point > 1.:
        raise ValueError('z_cluster must be between 0 and 1! Given value: ' + str(grid_cluster_point))
    b = grid_cluster_intensity
    z = np.linspace(0., 1., grid_points)
    zo = 1.0 / (2.0 * b) * np.log((1. + (np.exp(b) - 1.) * grid_cluster_point) /
                                  (1. + (np.exp(-b) - 1.) * grid_cluster_point))
    a = np.sinh(b * zo)
    for i in range(grid_points):
        z[i] = grid_cluster_point / a * (np.sinh(b * (z[i] - zo)) + a)
    z[-1] = 1.
    return z, z[1:] - z[:-1]


class FlameletSpec:
    """Flamelet specification (flamelet.py:35-237), adiabatic flamelets."""

    def __init__(self, mech_spec=None, initial_condition='Burke-Schumann', oxy_stream=None, fuel_stream=None,
                 max_dissipation_rate=None, stoich_dissipation_rate=None, dissipation_rate_form='Peters',
                 dissipation_rate=None, grid=None, grid_points=None, grid_type='clustered',
                 grid_cluster_intensity=4., grid_cluster_point='stoichiometric', heat_transfer='adiabatic',
                 include_enthalpy_flux=False, include_variable_cp=False, rates_sensitivity_type='dense',
                 sensitivity_transform_type='exact'):
        if heat_transfer != 'adiabatic':
            raise NotImplementedError('nonadiabatic flamelets are not supported by this implementation')
        self.mech_spec = mech_spec
        self.initial_condition = initial_condition
        self.oxy_stream = oxy_stream
        self.fuel_stream = fuel_stream
        self.max_dissipation_rate = max_dissipation_rate
        self.stoich_dissipation_rate = stoich_dissipation_rate
        self.dissipation_rate_form = dissipation_rate_form
        self.dissipation_rate = dissipation_rate
        self.grid = grid
        self.grid_points = grid_points
        self.grid_type = grid_type
        self.grid_cluster_intensity = grid_cluster_intensity
        self.grid_cluster_point = grid_cluster_point
        self.heat_transfer = heat_transfer
        self.include_enthalpy_flux = include_enthalpy_flux
        self.include_variable_cp = include_variable_cp
        self.rates_sensitivity_type = rates_sensitivity_type
        self.sensitivity_transform_type = sensitivity_transform_type


class FlameletHistory(dict):
    """Fields of a transient flamelet run, each an array [time, mixture fraction] including the boundary
    streams (the reference returns a spitfire Library with these dimensions, flamelet.py:1187-1211)."""

    def __init__(self, t, z):
        super().__init__()
        self.time_values = np.asarray(t)
        self.mixture_fraction_values = np.asarray(z)


class Flamelet:
    """An adiabatic flamelet on the griffon drop-in: steady solvers and transient integration
    (flamelet.py:269-1778)."""

    def __init__(self, flamelet_specs=None, **kwargs):
        fs = flamelet_specs if isinstance(flamelet_specs, FlameletSpec) else FlameletSpec(**(flamelet_specs or kwargs))
        self._mechanism = fs.mech_spec
        self._oxy_stream, self._fuel_stream = fs.oxy_stream, fs.fuel_stream
        if abs(self._fuel_stream.P - self._oxy_stream.P) / self._oxy_stream.P > 1.e-6:
            raise ValueError('The pressure of the fuel and oxidizer streams must be the same.')
        self._pressure = float(self._oxy_stream.P)
        self._n_species = self._mechanism.n_species
        self._n_equations = self._n_species
        self._state_fuel = self._fuel_stream.state
        self._state_oxy = self._oxy_stream.state
        mech = self._mechanism
        if fs.grid is not None:
            self._z = np.array(fs.grid, dtype=np.float64)
            self._dz = self._z[1:] - self._z[:-1]
        elif fs.grid_type == 'uniform':
            self._z, self._dz = uniform_grid(fs.grid_points)
        else:
            zc = fs.grid_cluster_point
            if zc == 'stoichiometric':
                zc = mech.stoich_mixture_fraction(self._fuel_stream, self._oxy_stream)
            self._z, self._dz = clustered_grid(fs.grid_points, zc, fs.grid_cluster_intensity)
        self._nz_interior = self._z.size - 2
        self._n_dof = self._n_equations * self._nz_interior
        if fs.dissipation_rate is not None:
            self._x = np.array(fs.dissipation_rate, dtype=np.float64)
            self._max_dissipation_rate = np.max(self._x)
        else:
            if fs.max_dissipation_rate is not None:
                self._max_dissipation_rate = fs.max_dissipation_rate
            elif fs.dissipation_rate_form in ('peters', 'Peters'):
                zst = mech.stoich_mixture_fraction(self._fuel_stream, self._oxy_stream)
                self._max_dissipation_rate = fs.stoich_dissipation_rate / np.exp(-2. * (erfinv(2. * zst - 1.)) ** 2)
            else:
                self._max_dissipation_rate = fs.stoich_dissipation_rate
            self._x = compute_dissipation_rate(self._z, self._max_dissipation_rate, fs.dissipation_rate_form)
        self._lewis_numbers = np.ones(self._n_species)
        self._include_enthalpy_flux = fs.include_enthalpy_flux
        self._include_variable_cp = fs.include_variable_cp
        self._rsopt = {'dense': 0, 'no-TBAF': 1, 'sparse': 2}[fs.rates_sensitivity_type]
        self._stopt = {'exact': 0}[fs.sensitivity_transform_type]
        self._initial_state = self._make_initial_state(fs.initial_condition)
        self._current_state = np.copy(self._initial_state)
        self._variable_scales = np.ones(self._n_dof)
        self._variable_scales[::self._n_equations] = 1.e3
        self._griffon = mech.griffon
        n, nzi = self._n_species, self._nz_interior
        self._maj_coeff_griffon = np.zeros(self._n_dof)
        self._sub_coeff_griffon = np.zeros(self._n_dof)
        self._sup_coeff_griffon = np.zeros(self._n_dof)
        self._mcoeff_griffon = np.zeros(nzi)
        self._ncoeff_griffon = np.zeros(nzi)
        self._setup_stencils()
        self._jac_nelements_griffon = int(n * (nzi * n + 2 * (nzi - 1)))
        self._block_thomas_l_values = np.zeros(n * n * nzi)
        self._block_thomas_d_pivots = np.zeros(self._jac_nelements_griffon, dtype=np.int32)
        self._iteration_count = None
        self._current_time = 0.
        self._jacobian_values = None

    def _setup_stencils(self):
        self._griffon.flamelet_stencils(self._dz, self._nz_interior, self._x, 1. / self._lewis_numbers,
                                        self._maj_coeff_griffon, self._sub_coeff_griffon, self._sup_coeff_griffon,
                                        self._mcoeff_griffon, self._ncoeff_griffon)

    # ---------------- initial conditions (flamelet.py:486-597) ----------------
    def _make_initial_state(self, ic):
        mech, nzi, neq = self._mechanism, self._nz_interior, self._n_equations
        if isinstance(ic, np.ndarray):
            if ic.size != self._n_dof:
                raise ValueError('size of initial condition is incorrect!')
            return np.copy(ic)
        st = np.ndarray((nzi, neq))
        zi = self._z[1:-1]
        if ic == 'linear-TY':
            for i in range(neq):
                st[:, i] = self._state_oxy[i] + (self._state_fuel[i] - self._state_oxy[i]) * zi
        elif ic == 'unreacted':
            Yo, Yf = self._oxy_stream.Y, self._fuel_stream.Y
            Y = Yo[None, :] + zi[:, None] * (Yf - Yo)[None, :]
            h = (1. - zi) * self._oxy_stream.H + zi * self._fuel_stream.H
            st[:, 0] = mech.temperatures_from_enthalpies(h, Y, T0=self._oxy_stream.T)
            st[:, 1:] = Y[:, :-1]
        elif ic == 'Burke-Schumann':
            zst = mech.stoich_mixture_fraction(self._fuel_stream, self._oxy_stream)
            stmix = mech.mix_streams([(self._oxy_stream, 1 - zst), (self._fuel_stream, zst)], 'mass', 'HP')
            stat = mech._get_atoms_in_stream(stmix, [a for a in ('H', 'C', 'O', 'N') if a in
                                                     mech.mech_data['elements']] )
            comp = 'N2: %.17g H2O: %.17g' % (stat.get('N', 0.) / 2, 0.5 * stat.get('H', 0.))
            if 'CO2' in mech.species_names:
                comp += ' CO2: %.17g' % stat.get('C', 0.)
            sw = mech.stream('HPX', (stmix.H, self._pressure, comp))
            Yst, Yo, Yf = sw.Y, self._oxy_stream.Y, self._fuel_stream.Y
            lean = zi <= zst
            Y = np.where(lean[:, None], Yo[None, :] + (zi / zst)[:, None] * (Yst - Yo)[None, :],
                         Yst[None, :] + ((zi - zst) / (1. - zst))[:, None] * (Yf - Yst)[None, :])
            # mass-basis mixing at constant HP (mix_streams): the enthalpy is linear in Z
            h = (1. - zi) * self._oxy_stream.H + zi * self._fuel_stream.H
            st[:, 0] = mech.temperatures_from_enthalpies(h, Y)
            st[:, 1:] = Y[:, :-1]
        else:
            raise ValueError('unsupported initial condition ' + str(ic))
        return st.ravel().copy()

    # ---------------- griffon wrappers (flamelet.py:642-762, 911-915) ----------------
    def _adiabatic_rhs(self, t, state_interior):
        rhs = np.zeros(self._n_dof)
        null = np.zeros(1)
        self._griffon.flamelet_rhs(state_interior, self._pressure, self._state_oxy, self._state_fuel, True, null, null,
                                   null, null, self._nz_interior, self._maj_coeff_griffon, self._sub_coeff_griffon,
                                   self._sup_coeff_griffon, self._mcoeff_griffon, self._ncoeff_griffon, self._x,
                                   self._include_enthalpy_flux, self._include_variable_cp, False, rhs)
        return rhs

    def _adiabatic_jac(self, state_interior, scale_and_offset=False, prefactor=0.):
        values = np.zeros(self._jac_nelements_griffon)
        null = np.zeros(1)
        self._griffon.flamelet_jacobian(state_interior, self._pressure, self._state_oxy, self._state_fuel, True, null,
                                        null, null, null, self._nz_interior, self._maj_coeff_griffon,
                                        self._sub_coeff_griffon, self._sup_coeff_griffon, self._mcoeff_griffon,
                                        self._ncoeff_griffon, self._x, False, 0., scale_and_offset, prefactor,
                                        self._rsopt, self._stopt, self._include_enthalpy_flux,
                                        self._include_variable_cp, False, null, values)
        return values

    def rhs(self, t, state_interior):
        return self._adiabatic_rhs(t, state_interior)

    def jac(self, state_interior):
        return self._adiabatic_jac(state_interior)

    def _jac_and_expeig(self, state_interior, diffterm):
        """Jacobian values and the per-point explosive-eigenvalue estimate max(0, max Re(lambda) - diffterm)
        of each chemistry block (flamelet.py:1597-1600 with compute_eigenvalues=True)."""
        values = np.zeros(self._jac_nelements_griffon)
        expeig = np.zeros(self._n_dof)
        null = np.zeros(1)
        self._griffon.flamelet_jacobian(state_interior, self._pressure, self._state_oxy, self._state_fuel, True, null,
                                        null, null, null, self._nz_interior, self._maj_coeff_griffon,
                                        self._sub_coeff_griffon, self._sup_coeff_griffon, self._mcoeff_griffon,
                                        self._ncoeff_griffon, self._x, True, diffterm, False, 0.,
                                        self._rsopt, self._stopt, self._include_enthalpy_flux,
                                        self._include_variable_cp, False, expeig, values)
        return values, expeig

    def eigenvalue_bound(self, state_interior, diffterm):
        return self._jac_and_expeig(state_interior, diffterm)[1]

    # ---------------- linear algebra (griffon py_btddod_*) ----------------
    def _factorize(self, values):
        from .griffon import py_btddod_full_factorize
        py_btddod_full_factorize(values, self._nz_interior, self._n_equations, self._block_thomas_l_values,
                                 self._block_thomas_d_pivots)

    def _solve(self, values, rhs):
        from .griffon import py_btddod_full_solve
        out = np.zeros(self._n_dof)
        py_btddod_full_solve(values, self._block_thomas_l_values, self._block_thomas_d_pivots, rhs,
                             self._nz_interior, self._n_equations, out)
        return out

    def _scale_add_diag(self, values, a, diag, b):
        from .griffon import py_btddod_scale_and_add_diagonal
        py_btddod_scale_and_add_diagonal(values, a, diag, b, self._nz_interior, self._n_equations)

    # ---------------- steady solvers (flamelet.py:1352-1778) ----------------
    def steady_solve_newton(self, initial_guess=None, tolerance=1.e-6, max_iterations=10, max_factor_line_search=1.5,
                            max_allowed_residual=1.e6, min_allowable_state_var=-1.e-6, norm_order=np.inf,
                            log_rate=100000, verbose=True):
        state = np.copy(self._initial_state if initial_guess is None else initial_guess)
        inv_dofscales = 1. / self._variable_scales
        iteration_count = 0
        res = tolerance + 1.
        rhs = self.rhs(0., state)
        evaluate_jacobian = True
        J = None
        while res > tolerance and iteration_count < max_iterations:
            iteration_count += 1
            if evaluate_jacobian:
                J = -self.jac(state)
                self._factorize(J)
                evaluate_jacobian = False
            dstate = self._solve(J, rhs)
            if np.any(~np.isfinite(dstate)):
                return None, None, False
            norm_rhs_old = np.linalg.norm(rhs * inv_dofscales, ord=norm_order)
            alpha = 1.
            rhs = self.rhs(0., state + dstate)
            if np.any(~np.isfinite(rhs)):
                return None, None, False
            while np.linalg.norm(rhs * inv_dofscales, ord=norm_order) > max_factor_line_search * norm_rhs_old \
                    and alpha > 0.001:
                alpha *= 0.5
                dstate *= alpha  # as in the reference, the step shrinks by the running product of alphas
                rhs = self.rhs(0., state + dstate)
                evaluate_jacobian = True
            state += dstate
            res = np.linalg.norm(rhs * inv_dofscales, ord=norm_order)
            if res > max_allowed_residual or np.min(state) < min_allowable_state_var:
                return None, None, False
        state[state < 0] = 0.
        if res > tolerance:
            return None, None, False
        self._current_state = np.copy(state)
        self._iteration_count = iteration_count
        return state, iteration_count, True

    def steady_solve_psitc(self, initial_guess=None, tolerance=1.e-6, max_iterations=400,
                           min_allowable_state_var=-1.e-6, ds_init=1., ds_init_decrease=4., adaptive_restart=True,
                           diffusion_factor=4., global_ds=False, ds_safety=0.1, ds_ramp=1.1, ds_max=1.e4,
                           max_factor_line_search=1.5, max_allowed_residual=1.e6, log_rate=100000,
                           norm_order=np.inf, _recursion_depth=0, max_recursion_depth=20, verbose=True):
        args = dict(locals())
        del args['self']

        def restart():
            if _recursion_depth > max_recursion_depth or not adaptive_restart:
                return None, None, False, np.min(ds)
            args['ds_init'] = ds_init / ds_init_decrease
            args['_recursion_depth'] = _recursion_depth + 1
            return self.steady_solve_psitc(**args)

        state = np.copy(self._initial_state if initial_guess is None else initial_guess)
        inv_dofscales = 1. / self._variable_scales
        diffterm = diffusion_factor * self._max_dissipation_rate
        ds = np.full(self._n_dof, float(ds_init))
        iteration_count = 0
        jac_age = 0
        res = tolerance + 1.
        rhs = self.rhs(0., state)
        evaluate_jacobian = True
        J = None
        while res > tolerance and iteration_count < max_iterations:
            iteration_count += 1
            if evaluate_jacobian:
                J, expeig = self._jac_and_expeig(state, diffterm)
                ds = np.min(np.vstack([ds_safety / (expeig + 1.e-16), ds_ramp * ds, ds_max * np.ones(self._n_dof)]),
                            axis=0)
                if global_ds or iteration_count == 1:
                    ds[:] = np.min(ds)
                self._scale_add_diag(J, -1., 1. / ds, 1.)
                self._factorize(J)
                jac_age = 0
            else:
                jac_age += 1
            evaluate_jacobian = (jac_age == 8) or res > 1.e-2
            dstate = self._solve(J, rhs)
            if np.any(~np.isfinite(dstate)):
                return restart()
            norm_rhs_old = np.linalg.norm(rhs * inv_dofscales, ord=norm_order)
            alpha = 1.
            rhs = self.rhs(0., state + dstate)
            if np.any(~np.isfinite(rhs)):
                return restart()
            while np.linalg.norm(rhs * inv_dofscales, ord=norm_order) > max_factor_line_search * norm_rhs_old \
                    and alpha > 0.001:
                alpha *= 0.5
                dstate *= alpha
                rhs = self.rhs(0., state + dstate)
                evaluate_jacobian = True
            state += dstate
            res = np.linalg.norm(rhs * inv_dofscales, ord=norm_order)
            if res > max_allowed_residual or np.min(state) < min_allowable_state_var:
                return restart()
        state[state < 0] = 0.
        if res > tolerance:
            return None, None, False, np.min(ds)
        self._iteration_count = iteration_count
        self._current_state = np.copy(state)
        return state, iteration_count, True, np.min(ds)

    def compute_steady_state(self, tolerance=1.e-6, verbose=False, use_psitc=True, newton_args=None,
                             psitc_args=None, transient_args=None):
        """Newton, then pseudo-transient continuation, then ESDIRK64 time stepping to steady state
        (flamelet.py:1705-1778); returns the converged interior state."""
        nargs = {'tolerance': tolerance, 'verbose': verbose, 'max_iterations': 38}
        nargs.update(newton_args or {})
        state, it, conv = self.steady_solve_newton(**nargs)
        if conv:
            return state
        mds = 1.e-6
        if use_psitc:
            pargs = {'tolerance': tolerance, 'verbose': verbose, 'max_iterations': 400}
            pargs.update(psitc_args or {})
            state, it, conv, mds = self.steady_solve_psitc(**pargs)
            if conv:
                return state
        targs = {'steady_tolerance': tolerance, 'transient_tolerance': 1.e-8, 'max_time_step': 1.e4,
                 'first_time_step': 1.e-2 * mds, 'maximum_steps_per_jacobian': 10, 'save_first_and_last_only': True}
        targs.update(transient_args or {})
        self.integrate_to_steady(**targs)
        return self.current_interior_state

    # ---------------- transient integration (flamelet.py:1046-1237) ----------------
    def _setup_block_thomas(self, t, state_interior, prefactor):
        """prefactor * J - I, block-Thomas factored in place (flamelet.py:756-762)"""
        self._jacobian_values = self._adiabatic_jac(state_interior, scale_and_offset=True, prefactor=prefactor)
        self._factorize(self._jacobian_values)

    def _solve_block_thomas(self, residual):
        return self._solve(self._jacobian_values, residual)

    def integrate(self, stop_at_time=None, stop_at_steady=None, stop_criteria=None, first_time_step=1.e-6,
                  max_time_step=1.e-3, minimum_time_step_count=40, transient_tolerance=1.e-10,
                  maximum_steps_per_jacobian=10, nonlinear_solve_tolerance=1.e-12, save_first_and_last_only=False,
                  return_info=False):
        """ESDIRK64 with a frozen-Jacobian Newton solver, PI step control and block-Thomas linear solves of
        dt*gamma*J - I (flamelet.py:1046-1213, linear_solver='block thomas'); negative entries are clipped after
        every step. Starts from the current state and time and advances them. Returns a FlameletHistory with
        the time values and, per field, an array [time, mixture fraction] that includes the boundary streams."""
        from .time_integration import KennedyCarpenterS6P4Q3, PIController, SimpleNewtonSolver, odesolve
        w = 1. / self._variable_scales
        newton = SimpleNewtonSolver(evaluate_jacobian_every_iter=False, norm_weighting=w,
                                    tolerance=nonlinear_solve_tolerance)
        stepper = KennedyCarpenterS6P4Q3(newton, norm_weighting=w)
        control = PIController(first_step=first_time_step, max_step=max_time_step, target_error=transient_tolerance)

        def post_step(t, state, *args):
            state[state < 0.] = 0.
            return state

        out = odesolve(self._adiabatic_rhs, self._current_state, initial_time=self._current_time,
                       stop_at_time=stop_at_time, stop_at_steady=stop_at_steady, stop_criteria=stop_criteria,
                       minimum_time_step_count=minimum_time_step_count, post_step_callback=post_step,
                       method=stepper, step_size=control, linear_setup=self._setup_block_thomas,
                       linear_solve=self._solve_block_thomas, linear_setup_rate=maximum_steps_per_jacobian,
                       norm_weighting=w, save_each_step=not save_first_and_last_only, return_info=True)
        if save_first_and_last_only:
            state, t, _, info = out
            ts, states = np.array([t]), state[None, :]
        else:
            ts, states, info = out
        self._current_state = np.copy(states[-1])
        self._current_time = float(ts[-1])
        self.info = info
        hist = FlameletHistory(ts, self._z)
        ne = self._n_equations
        names = self._mechanism.species_names
        T = np.empty((ts.size, self._z.size))
        T[:, 0], T[:, -1] = self._oxy_stream.T, self._fuel_stream.T
        T[:, 1:-1] = states[:, ::ne]
        hist['temperature'] = T
        hist['pressure'] = np.full_like(T, self._pressure)
        last = np.ones_like(T)
        for i, sp in enumerate(names[:-1]):
            Y = np.empty_like(T)
            Y[:, 0], Y[:, -1] = self._oxy_stream.Y[i], self._fuel_stream.Y[i]
            Y[:, 1:-1] = states[:, 1 + i::ne]
            hist['mass fraction ' + sp] = Y
            last -= Y
        hist['mass fraction ' + names[-1]] = last
        return (hist, info) if return_info else hist

    def integrate_to_steady(self, steady_tolerance=1.e-4, **kwargs):
        return self.integrate(stop_at_steady=steady_tolerance, **kwargs)

    def integrate_to_time(self, final_time, **kwargs):
        return self.integrate(stop_at_time=final_time, **kwargs)

    # ---------------- accessors ----------------
    @property
    def mixture_fraction_grid(self):
        return self._z.copy()

    @property
    def dissipation_rate(self):
        return self._x.copy()

    @property
    def initial_interior_state(self):
        return self._initial_state.copy()

    @property
    def current_interior_state(self):
        return self._current_state.copy()

    @property
    def current_time(self):
        return self._current_time

    @property
    def iteration_count(self):
        return self._iteration_count

    @property
    def n_interior_points(self):
        return self._nz_interior

    @property
    def oxy_stream(self):
        return self._oxy_stream

    @property
    def fuel_stream(self):
        return self._fuel_stream

    @property
    def pressure(self):
        return self._pressure


class FlameletBatch:
    """Many adiabatic flamelets of one mechanism and pressure solved together on the GPU.

    The flamelets keep their own grids, dissipation rates, boundary streams and initial conditions; the
    steady solve is the reference's pseudo-transient continuation (Flamelet.steady_solve_psitc,
    flamelet.py:1487-1703) applied to every flamelet with device-resident state and batched kernels
    (gf_flamelet_batch_* in include/griffon_amd.h). Converged states are written back to each Flamelet
    (current_interior_state), as compute_steady_state does for one flamelet.
    """

    def __init__(self, flamelets):
        from ._native import lib, check, dptr, iptr
        import ctypes
        self._flamelets = list(flamelets)
        if not self._flamelets:
            raise ValueError('empty flamelet batch')
        f0 = self._flamelets[0]
        mech = f0._mechanism
        for f in self._flamelets:
            if f._mechanism is not mech or f._pressure != f0._pressure:
                raise ValueError('flamelets in a batch share one mechanism and one pressure')
            if (f._include_enthalpy_flux, f._include_variable_cp) != (f0._include_enthalpy_flux,
                                                                      f0._include_variable_cp):
                raise ValueError('flamelets in a batch share the enthalpy-flux and variable-cp options')
        self._lib = lib()
        self._mech = mech
        self._ns = mech.n_species
        self._nzi = np.array([f._nz_interior for f in self._flamelets], dtype=np.int32)
        self._offsets = np.concatenate([[0], np.cumsum(self._nzi.astype(np.int64) * self._ns)])
        dz = np.ascontiguousarray(np.concatenate([f._dz for f in self._flamelets]))
        chi = np.ascontiguousarray(np.concatenate([f._x for f in self._flamelets]))
        oxy = np.ascontiguousarray(np.concatenate([f._state_oxy for f in self._flamelets]))
        fuel = np.ascontiguousarray(np.concatenate([f._state_fuel for f in self._flamelets]))
        h = ctypes.c_void_p()
        check(self._lib.gf_flamelet_batch_create(mech.griffon.handle, len(self._flamelets), iptr(self._nzi),
                                                 float(f0._pressure), dptr(dz), dptr(chi), dptr(oxy), dptr(fuel),
                                                 int(f0._include_enthalpy_flux), int(f0._include_variable_cp),
                                                 ctypes.byref(h)))
        self._h = h
        self.set_state(np.concatenate([f._initial_state for f in self._flamelets]))

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self._lib.gf_flamelet_batch_destroy(h)
            self._h = None

    def __len__(self):
        return len(self._flamelets)

    @property
    def n_unknowns(self):
        return int(self._offsets[-1])

    def set_state(self, interior):
        from ._native import check, dptr
        x = np.ascontiguousarray(interior, dtype=np.float64)
        if x.size != self.n_unknowns:
            raise ValueError('state size %d, expected %d' % (x.size, self.n_unknowns))
        check(self._lib.gf_flamelet_batch_set_state(self._h, dptr(x)))

    def get_state(self):
        from ._native import check, dptr
        x = np.zeros(self.n_unknowns)
        check(self._lib.gf_flamelet_batch_get_state(self._h, dptr(x)))
        return x

    def split(self, x):
        return [x[self._offsets[k]:self._offsets[k + 1]].copy() for k in range(len(self._flamelets))]

    def rhs(self):
        """Flamelet RHS of every flamelet at the current state, concatenated."""
        from ._native import check, dptr
        out = np.zeros(self.n_unknowns)
        check(self._lib.gf_flamelet_batch_rhs(self._h, dptr(out)))
        return out

    def compute_steady_state(self, tolerance=1.e-6, lu_precision=64, **psitc_args):
        """Pseudo-transient continuation of every flamelet from its current state (psitc_args: the keyword
        arguments of Flamelet.steady_solve_psitc). Returns (states, converged, iterations); raises if any
        flamelet fails, unless psitc_args['raise_on_failure'] is False."""
        from ._native import check, iptr, dptr, PsitcOptions
        import ctypes
        raise_on_failure = psitc_args.pop('raise_on_failure', True)
        o = PsitcOptions()
        self._lib.gf_psitc_options_default(ctypes.byref(o))
        o.tolerance = tolerance
        o.lu_precision = int(lu_precision)
        names = {'max_iterations': 'max_iterations', 'ds_init': 'ds_init', 'ds_init_decrease': 'ds_init_decrease',
                 'adaptive_restart': 'adaptive_restart', 'max_recursion_depth': 'max_recursion_depth',
                 'diffusion_factor': 'diffusion_factor', 'global_ds': 'global_ds', 'ds_safety': 'ds_safety',
                 'ds_ramp': 'ds_ramp', 'ds_max': 'ds_max', 'max_factor_line_search': 'max_factor_line_search',
                 'max_allowed_residual': 'max_allowed_residual',
                 'min_allowable_state_var': 'min_allowable_state_var'}
        for k, v in psitc_args.items():
            if k == 'log_rate':
                o.log_rate = int(v)
                continue
            if k in ('verbose', 'norm_order'):
                continue
            if k not in names:
                raise TypeError('unknown psitc argument ' + k)
            setattr(o, names[k], type(getattr(o, names[k]))(v))
        B = len(self._flamelets)
        it = np.zeros(B, dtype=np.int32)
        conv = np.zeros(B, dtype=np.int32)
        res = np.zeros(B)
        cnt = (ctypes.c_longlong * 4)()
        check(self._lib.gf_flamelet_batch_steady_psitc(self._h, ctypes.byref(o), iptr(it), iptr(conv), dptr(res),
                                                       cnt))
        self.counters = {'iterations': cnt[0], 'jacobian_refreshes': cnt[1], 'rhs_evaluations': cnt[2],
                         'synchronizing_steps': cnt[3]}
        states = self.split(self.get_state())
        for f, s, c, n in zip(self._flamelets, states, conv, it):
            if c:
                f._current_state = s.copy()
                f._iteration_count = int(n)
        self.residuals = res
        if raise_on_failure and not np.all(conv):
            raise RuntimeError('steady flamelet solve did not converge for %d of %d flamelets'
                               % (int(B - conv.sum()), B))
        return states, conv.astype(bool), it
