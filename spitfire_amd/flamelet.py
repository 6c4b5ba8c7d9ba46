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
  * none in compute_steady_state: Newton, then psitc, then ESDIRK64 time stepping, as the reference; the
    FlameletBatch form runs each stage batched over the flamelets still unconverged.
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
    """Flamelet specification (flamelet.py:35-237). Nonadiabatic flamelets lose heat by convection,
    h_conv (T - T_conv), and radiation, eps sigma (T^4 - T_rad^4), per interior point (flamelet_kernels.cpp:1121-1138,
    1290-1320)."""

    def __init__(self, mech_spec=None, initial_condition='Burke-Schumann', oxy_stream=None, fuel_stream=None,
                 max_dissipation_rate=None, stoich_dissipation_rate=None, dissipation_rate_form='Peters',
                 dissipation_rate=None, grid=None, grid_points=None, grid_type='clustered',
                 grid_cluster_intensity=4., grid_cluster_point='stoichiometric', heat_transfer='adiabatic',
                 convection_temperature=None, radiation_temperature=None, convection_coefficient=None,
                 radiative_emissivity=None, scale_heat_loss_by_temp_range=False, scale_convection_by_dissipation=False,
                 use_linear_ref_temp_profile=False, include_enthalpy_flux=True, include_variable_cp=True,
                 rates_sensitivity_type='dense', sensitivity_transform_type='exact', library_slice=None):
        if heat_transfer not in ('adiabatic', 'nonadiabatic'):
            raise ValueError('Flamelet specifications: Bad heat_transfer argument detected: ' + str(heat_transfer) +
                             "; acceptable values: ['adiabatic', 'nonadiabatic']")
        self.convection_temperature = convection_temperature
        self.radiation_temperature = radiation_temperature
        self.convection_coefficient = convection_coefficient
        self.radiative_emissivity = radiative_emissivity
        self.scale_heat_loss_by_temp_range = scale_heat_loss_by_temp_range
        self.scale_convection_by_dissipation = scale_convection_by_dissipation
        self.use_linear_ref_temp_profile = use_linear_ref_temp_profile
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
        if library_slice is not None:
            self._from_library_slice(library_slice)

    def _from_library_slice(self, lib):
        """mechanism, streams, grid and initial state from a one-dimensional library over mixture fraction (trivial
        extra dimensions are squeezed; flamelet.py:135-180)"""
        from .library import Library
        shape = lib.shape
        if len(shape) < 1 or any(n > 1 for n in shape[1:]):
            raise ValueError('Error in Flamelet construction from library_slice. The library must be one-dimensional, '
                             'with mixture fraction as the first dimension (shape = %s).' % (shape,))
        if len(shape) > 1:
            lib = Library.squeeze(lib)
        if 'mixture_fraction' not in lib.dims[0].name:
            raise ValueError('Error in Flamelet construction from library_slice. The library provided does not appear '
                             'to have mixture fraction as the first dimension.')
        m = lib.extra_attributes['mech_spec']
        self.mech_spec = m
        names = m.species_names
        p = lib['pressure'][0]
        self.oxy_stream = m.stream('TPY', (lib['temperature'][0], p, [lib['mass fraction ' + s][0] for s in names]))
        self.fuel_stream = m.stream('TPY', (lib['temperature'][-1], p, [lib['mass fraction ' + s][-1] for s in names]))
        self.grid = lib.mixture_fraction_values
        self.grid_points = self.grid_type = self.grid_cluster_intensity = self.grid_cluster_point = None
        ic = np.zeros((self.grid.size - 2, m.n_species))
        ic[:, 0] = lib['temperature'][1:-1]
        for i, s in enumerate(names[:-1]):
            ic[:, 1 + i] = lib['mass fraction ' + s][1:-1]
        self.initial_condition = ic.ravel()


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
        # dissipation rate (flamelet.py:427-468): given values, or a Peters/constant profile from the maximum or
        # stoichiometric value, or zero when neither is given or the form is not one of those
        if fs.dissipation_rate is not None:
            self._x = np.array(fs.dissipation_rate, dtype=np.float64)
            self._max_dissipation_rate = np.max(self._x)
            self._dissipation_rate_form = 'custom'
        elif fs.dissipation_rate_form not in ('peters', 'Peters', 'constant') or \
                (fs.max_dissipation_rate is None and fs.stoich_dissipation_rate is None):
            self._x = np.zeros_like(self._z)
            self._max_dissipation_rate = 0.
            self._dissipation_rate_form = 'unspecified-set-to-0'
        else:
            if fs.max_dissipation_rate is not None:
                self._max_dissipation_rate = fs.max_dissipation_rate
            elif fs.dissipation_rate_form in ('peters', 'Peters'):
                zst = mech.stoich_mixture_fraction(self._fuel_stream, self._oxy_stream)
                self._max_dissipation_rate = fs.stoich_dissipation_rate / np.exp(-2. * (erfinv(2. * zst - 1.)) ** 2)
            else:
                self._max_dissipation_rate = fs.stoich_dissipation_rate
            self._dissipation_rate_form = fs.dissipation_rate_form
            self._x = compute_dissipation_rate(self._z, self._max_dissipation_rate, fs.dissipation_rate_form)
        self._lewis_numbers = np.ones(self._n_species)
        self._setup_heat_transfer(fs)
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
        self._initial_time = 0.
        self._current_time = 0.
        self._solution_times = []
        self._jacobian_values = None

    def _heat_transfer_array(self, value, name):
        """a float -> that value at every interior point; an array -> a copy (flamelet.py:354-364)"""
        if value is None:
            raise ValueError('Flamelet specifications: Nonadiabatic heat transfer was selected but no ' + name +
                             ' argument was given.')
        if isinstance(value, float):
            return value + np.zeros(self._nz_interior)
        if isinstance(value, np.ndarray):
            return np.array(value, dtype=np.float64)
        raise ValueError(name + ' was not given as a float (constant) or numpy array')

    def _setup_heat_transfer(self, fs):
        """convection and radiation parameters at the interior points (flamelet.py:470-519)"""
        self._heat_transfer = fs.heat_transfer
        self._scale_heat_loss_by_temp_range = bool(getattr(fs, 'scale_heat_loss_by_temp_range', False))
        self._T_conv = self._T_rad = self._h_conv = self._h_rad = None
        if self._heat_transfer == 'adiabatic':
            return
        self._h_conv = self._heat_transfer_array(fs.convection_coefficient, 'convection_coefficient')
        self._h_rad = self._heat_transfer_array(fs.radiative_emissivity, 'radiative_emissivity')
        if fs.use_linear_ref_temp_profile:
            self._T_conv = self._oxy_stream.T + self._z[1:-1] * (self._fuel_stream.T - self._oxy_stream.T)
            self._T_rad = self._T_conv.copy()
        else:
            self._T_rad = self._heat_transfer_array(fs.radiation_temperature, 'radiation_temperature')
            self._T_conv = self._heat_transfer_array(fs.convection_temperature, 'convection_temperature')
        if fs.scale_convection_by_dissipation:
            zst = self._mechanism.stoich_mixture_fraction(self._fuel_stream, self._oxy_stream)
            factor = np.max(self._x) / (1. - zst) / zst
            self._h_conv = self._h_conv * factor
            self._h_rad = self._h_rad * factor
        for a in (self._T_conv, self._T_rad, self._h_conv, self._h_rad):
            if a.size != self._nz_interior:
                raise ValueError('heat transfer parameters must have one value per interior grid point')

    def _heat_loss_args(self):
        """(adiabatic, T_convection, T_radiation, h_convection, h_radiation) in the griffon argument order"""
        if self._heat_transfer == 'adiabatic':
            null = np.zeros(1)
            return True, null, null, null, null
        return False, self._T_conv, self._T_rad, self._h_conv, self._h_rad

    @property
    def heat_transfer(self):
        return self._heat_transfer

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
    # the reference's _adiabatic_* and _nonadiabatic_* wrappers in one: the heat-loss arguments decide
    def _flamelet_rhs(self, t, state_interior):
        rhs = np.zeros(self._n_dof)
        self._griffon.flamelet_rhs(state_interior, self._pressure, self._state_oxy, self._state_fuel,
                                   *self._heat_loss_args(), self._nz_interior, self._maj_coeff_griffon,
                                   self._sub_coeff_griffon, self._sup_coeff_griffon, self._mcoeff_griffon,
                                   self._ncoeff_griffon, self._x, self._include_enthalpy_flux, self._include_variable_cp,
                                   self._scale_heat_loss_by_temp_range, rhs)
        return rhs

    def _flamelet_jac(self, state_interior, scale_and_offset=False, prefactor=0.):
        # the reference passes use_scaled_heat_loss=False to the adiabatic Jacobian wrappers, where it has no effect
        values = np.zeros(self._jac_nelements_griffon)
        null = np.zeros(1)
        self._griffon.flamelet_jacobian(state_interior, self._pressure, self._state_oxy, self._state_fuel,
                                        *self._heat_loss_args(), self._nz_interior, self._maj_coeff_griffon,
                                        self._sub_coeff_griffon, self._sup_coeff_griffon, self._mcoeff_griffon,
                                        self._ncoeff_griffon, self._x, False, 0., scale_and_offset, prefactor,
                                        self._rsopt, self._stopt, self._include_enthalpy_flux,
                                        self._include_variable_cp, self._scale_heat_loss_by_temp_range, null, values)
        return values

    def rhs(self, t, state_interior):
        return self._flamelet_rhs(t, state_interior)

    def jac(self, state_interior):
        return self._flamelet_jac(state_interior)

    def _jac_indices(self):
        if getattr(self, '_jac_indices_griffon', None) is None:
            rows = np.zeros(self._jac_nelements_griffon, dtype=np.int32)
            cols = np.zeros(self._jac_nelements_griffon, dtype=np.int32)
            self._griffon.flamelet_jac_indices(self._nz_interior, rows, cols)
            self._jac_indices_griffon = (rows, cols)
        return self._jac_indices_griffon

    def jac_csc(self, state_interior):
        """the Jacobian as a scipy CSC matrix (flamelet.py:897-898)"""
        from scipy.sparse import csc_matrix
        return csc_matrix((self._flamelet_jac(state_interior), self._jac_indices()))

    def jac_and_eig(self, state_interior, diffterm=0.):
        """Jacobian values and the per-point explosive-eigenvalue estimate (flamelet.py:900-901; the reference's
        method passes no diffterm to its helper)"""
        return self._jac_and_expeig(state_interior, diffterm)

    @classmethod
    def _uniform_grid(cls, grid_points):
        return uniform_grid(grid_points)

    @classmethod
    def _clustered_grid(cls, grid_points, grid_cluster_point, grid_cluster_intensity=6.):
        return clustered_grid(grid_points, grid_cluster_point, grid_cluster_intensity)

    @classmethod
    def make_clustered_grid(cls, grid_points, grid_cluster_point, grid_cluster_intensity=6.):
        return clustered_grid(grid_points, grid_cluster_point, grid_cluster_intensity)[0]

    @classmethod
    def _compute_dissipation_rate(cls, mixture_fraction, max_dissipation_rate, form='Peters'):
        return compute_dissipation_rate(mixture_fraction, max_dissipation_rate, form)

    def _jac_and_expeig(self, state_interior, diffterm):
        """Jacobian values and the per-point explosive-eigenvalue estimate max(0, max Re(lambda) - diffterm)
        of each chemistry block (flamelet.py:1597-1600 with compute_eigenvalues=True)."""
        values = np.zeros(self._jac_nelements_griffon)
        expeig = np.zeros(self._n_dof)
        self._griffon.flamelet_jacobian(state_interior, self._pressure, self._state_oxy, self._state_fuel,
                                        *self._heat_loss_args(), self._nz_interior, self._maj_coeff_griffon,
                                        self._sub_coeff_griffon, self._sup_coeff_griffon, self._mcoeff_griffon,
                                        self._ncoeff_griffon, self._x, True, diffterm, False, 0.,
                                        self._rsopt, self._stopt, self._include_enthalpy_flux,
                                        self._include_variable_cp, self._scale_heat_loss_by_temp_range, expeig, values)
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
        return self.make_library_from_interior_state(state), iteration_count, True

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
        return self.make_library_from_interior_state(state), iteration_count, True, np.min(ds)

    def compute_steady_state(self, tolerance=1.e-6, verbose=False, use_psitc=True, newton_args=None,
                             psitc_args=None, transient_args=None):
        """Newton, then pseudo-transient continuation, then ESDIRK64 time stepping to steady state
        (flamelet.py:1705-1778); returns a Library of temperature, pressure and mass fractions over the mixture
        fraction grid (boundary streams included), and leaves the state in current_interior_state."""
        nargs = {'tolerance': tolerance, 'verbose': verbose, 'max_iterations': 38}
        nargs.update(newton_args or {})
        lib, it, conv = self.steady_solve_newton(**nargs)
        if conv:
            return lib
        mds = 1.e-6
        if use_psitc:
            pargs = {'tolerance': tolerance, 'verbose': verbose, 'max_iterations': 400}
            pargs.update(psitc_args or {})
            lib, it, conv, mds = self.steady_solve_psitc(**pargs)
            if conv:
                return lib
        targs = {'steady_tolerance': tolerance, 'transient_tolerance': 1.e-8, 'max_time_step': 1.e4,
                 'first_time_step': 1.e-2 * mds, 'maximum_steps_per_jacobian': 10, 'save_first_and_last_only': True}
        targs.update(transient_args or {})
        self.integrate_to_steady(**targs)
        return self.make_library_from_interior_state(self.current_interior_state)

    def make_library_from_interior_state(self, state_in):
        """Library over the mixture fraction grid with temperature, pressure and every mass fraction, the
        boundary streams at the ends (flamelet.py:1330-1350)."""
        return self._fields_library(np.asarray(state_in)[None, :], None)

    def _fields_library(self, states, t):
        """temperature, pressure and mass fractions (the last species by difference) for interior states
        [n, n_dof], over (time, mixture fraction) when t is given and over mixture fraction alone otherwise"""
        from .library import Dimension, Library
        zdim = Dimension('mixture_fraction', self._z)
        lib = Library(Dimension('time', t), zdim) if t is not None else Library(zdim)
        lib.extra_attributes['mech_spec'] = self._mechanism
        ne = self._n_equations
        names = self._mechanism.species_names

        def field(interior, left, right):
            a = np.empty((states.shape[0], self._z.size))
            a[:, 0], a[:, -1] = left, right
            a[:, 1:-1] = interior
            return a if t is not None else a[0]

        T = field(states[:, ::ne], self._oxy_stream.T, self._fuel_stream.T)
        lib['temperature'] = T
        lib['pressure'] = np.zeros_like(T) + self._pressure
        last = np.ones_like(T)
        for i, sp in enumerate(names[:-1]):
            Y = field(states[:, 1 + i::ne], self._oxy_stream.Y[i], self._fuel_stream.Y[i])
            lib['mass fraction ' + sp] = Y
            last = last - Y
        lib['mass fraction ' + names[-1]] = last
        return lib

    # ---------------- transient integration (flamelet.py:1046-1237) ----------------
    def _setup_block_thomas(self, t, state_interior, prefactor):
        """prefactor * J - I, block-Thomas factored in place (flamelet.py:756-762)"""
        self._jacobian_values = self._flamelet_jac(state_interior, scale_and_offset=True, prefactor=prefactor)
        self._factorize(self._jacobian_values)

    def _setup_superlu(self, t, state_interior, prefactor):
        """prefactor * J - I as a CSC matrix, factored by SuperLU (flamelet.py:750-754)"""
        from scipy.sparse import csc_matrix
        from scipy.sparse.linalg import splu
        jac = csc_matrix((self._flamelet_jac(state_interior, scale_and_offset=True, prefactor=prefactor),
                          self._jac_indices()))
        jac.eliminate_zeros()
        self._linear_inverse_operator = splu(jac)

    def _solve_superlu(self, residual):
        return self._linear_inverse_operator.solve(residual), 1, True

    def _solve_block_thomas(self, residual):
        return self._solve(self._jacobian_values, residual)

    def integrate(self, stop_at_time=None, stop_at_steady=None, stop_criteria=None, first_time_step=1.e-6,
                  max_time_step=1.e-3, minimum_time_step_count=40, transient_tolerance=1.e-10,
                  maximum_steps_per_jacobian=10, nonlinear_solve_tolerance=1.e-12, save_first_and_last_only=False,
                  return_info=False, write_log=False, log_rate=100, linear_solver='block thomas',
                  print_exception_on_failure=False):
        """ESDIRK64 with a frozen-Jacobian Newton solver, PI step control and block-Thomas linear solves of
        dt*gamma*J - I (flamelet.py:1046-1213, linear_solver='block thomas'); negative entries are clipped after
        every step. Starts from the current state and time and advances them. Returns a Library over (time,
        mixture fraction) of temperature, pressure and mass fractions, boundary streams included
        (flamelet.py:1187-1211)."""
        from .time_integration import KennedyCarpenterS6P4Q3, PIController, SimpleNewtonSolver, odesolve
        if linear_solver.lower() == 'superlu':
            setup, solve = self._setup_superlu, self._solve_superlu
        elif linear_solver.lower() == 'block thomas':
            setup, solve = self._setup_block_thomas, self._solve_block_thomas
        else:
            raise ValueError('linear solver ' + linear_solver + ' is invalid, must be superlu or block thomas')
        w = 1. / self._variable_scales
        newton = SimpleNewtonSolver(evaluate_jacobian_every_iter=False, norm_weighting=w,
                                    tolerance=nonlinear_solve_tolerance)
        stepper = KennedyCarpenterS6P4Q3(newton, norm_weighting=w)
        control = PIController(first_step=first_time_step, max_step=max_time_step, target_error=transient_tolerance)

        def post_step(t, state, *args):
            state[state < 0.] = 0.
            return state

        out = odesolve(self._flamelet_rhs, self._current_state, initial_time=self._current_time,
                       stop_at_time=stop_at_time, stop_at_steady=stop_at_steady, stop_criteria=stop_criteria,
                       minimum_time_step_count=minimum_time_step_count, post_step_callback=post_step,
                       method=stepper, step_size=control, linear_setup=setup,
                       linear_solve=solve, linear_setup_rate=maximum_steps_per_jacobian,
                       norm_weighting=w, save_each_step=not save_first_and_last_only, return_info=True)
        if save_first_and_last_only:
            state, t, _, info = out
            ts, states = np.array([t]), state[None, :]
        else:
            ts, states, info = out
        self._current_state = np.copy(states[-1])
        self._current_time = float(ts[-1])
        self._solution_times = list(ts)
        self.info = info
        hist = self._fields_library(states, ts)
        return (hist, info) if return_info else hist

    def integrate_to_steady(self, steady_tolerance=1.e-4, **kwargs):
        return self.integrate(stop_at_steady=steady_tolerance, **kwargs)

    def integrate_to_time(self, final_time, **kwargs):
        return self.integrate(stop_at_time=final_time, **kwargs)

    def _has_ignited(self, state, delta_temperature_ignition):
        ne = self._n_equations
        return np.max(state[::ne] - self._initial_state[::ne]) > delta_temperature_ignition

    def integrate_for_heat_loss(self, temperature_tolerance=0.05, steady_tolerance=1.e-4, **kwargs):
        """integrate until max(state) <= (1 + temperature_tolerance) max(T_oxy, T_fuel) or steady
        (flamelet.py:1263-1288; the reference tests the largest entry of the whole state vector)"""
        T_bc_max = max(self._oxy_stream.T, self._fuel_stream.T)

        def stop(t, state, residual, *args, **kw):
            return np.max(state) < (1. + temperature_tolerance) * T_bc_max or residual < steady_tolerance

        return self.integrate(stop_criteria=stop, **kwargs)

    def integrate_to_steady_after_ignition(self, steady_tolerance=1.e-4, delta_temperature_ignition=400., **kwargs):
        """integrate until steady after the temperature rose by delta_temperature_ignition somewhere
        (flamelet.py:1239-1261)"""
        def stop(t, state, residual, *args, **kw):
            return self._has_ignited(state, delta_temperature_ignition) and residual < steady_tolerance

        return self.integrate(stop_criteria=stop, **kwargs)

    def compute_ignition_delay(self, delta_temperature_ignition=400., minimum_allowable_residual=1.e-12,
                               return_solution=False, **kwargs):
        """time at which the temperature first rises by delta_temperature_ignition; ValueError if the residual
        drops below minimum_allowable_residual first (flamelet.py:1290-1328)"""
        def stop(t, state, residual, *args, **kw):
            if residual > minimum_allowable_residual:
                return self._has_ignited(state, delta_temperature_ignition)
            raise ValueError('From compute_ignition_delay(): residual < minimum allowable value (%g), suggesting that '
                             'the flamelet will not ignite.' % minimum_allowable_residual)

        lib = self.integrate(stop_criteria=stop, save_first_and_last_only=not return_solution, **kwargs)
        tau = lib.time_values[-1]
        return (tau, lib) if return_solution else tau

    # ---------------- accessors ----------------
    @property
    def mixture_fraction_grid(self):
        return self._z.copy()

    @property
    def linear_temperature(self):
        """T_oxy + (T_fuel - T_oxy) z on the full grid (flamelet.py:971-976)"""
        To, Tf = self._oxy_stream.T, self._fuel_stream.T
        return To + (Tf - To) * self._z

    @property
    def current_temperature(self):
        """temperature of the current state on the full grid, boundary streams included (flamelet.py:1007-1010)"""
        return np.hstack((self._state_oxy[0], self._current_state[::self._n_equations], self._state_fuel[0]))

    @property
    def dissipation_rate(self):
        return self._x.copy()

    @property
    def mechanism(self):
        return self._mechanism

    @property
    def mixfrac_grid(self):
        return self._z

    def _state_2d_with_bcs(self, state):
        """interior state with the oxidizer state prepended and the fuel state appended (flamelet.py:935-939)"""
        return np.vstack((self._state_oxy, np.asarray(state).reshape((-1, self._n_equations)),
                          self._state_fuel)).ravel()

    def _mass_fraction_with_bcs(self, key, state):
        """a mass fraction over the full grid from a full state; the last species by difference
        (flamelet.py:923-933)"""
        if isinstance(key, str):
            key = self._mechanism.species_index(key)
        if key == self._n_species - 1:
            yn = np.ones(self._nz_interior + 2)
            for i in range(1, self._n_equations):
                yn -= state[i::self._n_equations]
            return yn
        return state[key + 1::self._n_equations]

    @property
    def initial_state(self):
        return self._state_2d_with_bcs(self._initial_state)

    @property
    def initial_temperature(self):
        return np.hstack((self._state_oxy[0], self._initial_state[::self._n_equations], self._state_fuel[0]))

    def initial_mass_fraction(self, key):
        return self._mass_fraction_with_bcs(key, self.initial_state)

    @property
    def current_state(self):
        return self._state_2d_with_bcs(self._current_state)

    def current_mass_fraction(self, key):
        return self._mass_fraction_with_bcs(key, self.current_state)

    @property
    def solution_times(self):
        """times of the last integrate() call (the reference keeps an always-empty list here, flamelet.py:615)"""
        return np.array(self._solution_times)

    def offset_time(self, delta):
        self._initial_time += delta
        self._current_time += delta

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

    The flamelets keep their own grids, dissipation rates, boundary streams and initial conditions. The steady
    solvers are the reference's, applied to every flamelet with device-resident state and batched kernels
    (gf_flamelet_batch_* in include/griffon_amd.h): Newton (Flamelet.steady_solve_newton,
    flamelet.py:1352-1485), pseudo-transient continuation (steady_solve_psitc, flamelet.py:1487-1703),
    ESDIRK64 time stepping to steady state (integrate_to_steady, flamelet.py:1215-1237), and the cascade of
    the three (compute_steady_state, flamelet.py:1705-1778). Converged states are written back to each
    Flamelet (current_interior_state), as compute_steady_state does for one flamelet.
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
            if (f._heat_transfer, f._scale_heat_loss_by_temp_range) != (f0._heat_transfer,
                                                                        f0._scale_heat_loss_by_temp_range):
                raise ValueError('flamelets in a batch share the heat transfer setting and its scaling option')
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
        ile = np.ascontiguousarray(np.concatenate([1. / f._lewis_numbers for f in self._flamelets]))
        if np.any(ile != 1.):
            check(self._lib.gf_flamelet_batch_set_lewis_numbers(h, dptr(ile)))
        if f0._heat_transfer == 'nonadiabatic':
            hl = [np.ascontiguousarray(np.concatenate([getattr(f, a) for f in self._flamelets]))
                  for a in ('_T_conv', '_h_conv', '_T_rad', '_h_rad')]
            check(self._lib.gf_flamelet_batch_set_heat_loss(h, *[dptr(a) for a in hl],
                                                            int(f0._scale_heat_loss_by_temp_range)))
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
        """current state and initial guess of every flamelet (interior states, concatenated)"""
        from ._native import check, dptr
        x = np.ascontiguousarray(interior, dtype=np.float64)
        if x.size != self.n_unknowns:
            raise ValueError('state size %d, expected %d' % (x.size, self.n_unknowns))
        check(self._lib.gf_flamelet_batch_set_state(self._h, dptr(x)))
        self._x0 = x.copy()

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

    _OPTION_NAMES = ('max_iterations', 'ds_init', 'ds_init_decrease', 'adaptive_restart', 'max_recursion_depth',
                     'diffusion_factor', 'global_ds', 'ds_safety', 'ds_ramp', 'ds_max', 'max_factor_line_search',
                     'max_allowed_residual', 'min_allowable_state_var')

    def _options(self, tolerance, lu_precision, args, allowed):
        from ._native import PsitcOptions
        import ctypes
        o = PsitcOptions()
        self._lib.gf_psitc_options_default(ctypes.byref(o))
        o.tolerance = tolerance
        o.lu_precision = int(lu_precision)
        for k, v in args.items():
            if k == 'log_rate':
                o.log_rate = int(v)
            elif k in ('verbose', 'norm_order', 'initial_guess'):
                if k == 'initial_guess' and v is not None:
                    raise TypeError('the batch solvers start from the initial guesses given to set_state')
            elif k in allowed:
                setattr(o, k, type(getattr(o, k))(v))
            else:
                raise TypeError('unknown solver argument ' + k)
        return o

    def _select(self, select):
        B = len(self._flamelets)
        if select is None:
            return np.ones(B, dtype=np.int32)
        s = np.ascontiguousarray(np.asarray(select).astype(bool).astype(np.int32))
        if s.size != B:
            raise ValueError('select has %d entries for %d flamelets' % (s.size, B))
        return s

    def _finish(self, sel, conv, it, raise_on_failure, what):
        """write converged states back to their Flamelets; (states, converged, iterations) over the batch"""
        states = self.split(self.get_state())
        for k in np.flatnonzero(sel):
            if conv[k]:
                self._flamelets[k]._current_state = states[k].copy()
                self._flamelets[k]._iteration_count = int(it[k])
        if raise_on_failure and not np.all(conv[sel.astype(bool)]):
            raise RuntimeError('%s did not converge for %d of %d flamelets'
                               % (what, int(sel.sum() - conv[sel.astype(bool)].sum()), int(sel.sum())))
        return states, conv.astype(bool), it

    def _steady(self, newton, tolerance, lu_precision, select, args):
        from ._native import check, iptr, dptr
        import ctypes
        raise_on_failure = args.pop('raise_on_failure', True)
        allowed = (('max_iterations', 'max_factor_line_search', 'max_allowed_residual', 'min_allowable_state_var')
                   if newton else self._OPTION_NAMES)
        if newton:
            args.setdefault('max_iterations', 10)  # steady_solve_newton's default (flamelet.py:1355)
        o = self._options(tolerance, lu_precision, args, allowed)
        sel = self._select(select)
        B = len(self._flamelets)
        it = np.zeros(B, dtype=np.int32)
        conv = np.zeros(B, dtype=np.int32)
        res = np.full(B, np.nan)
        cnt = (ctypes.c_longlong * 4)()
        if newton:
            check(self._lib.gf_flamelet_batch_steady_newton(self._h, ctypes.byref(o), iptr(sel), iptr(it), iptr(conv),
                                                            dptr(res), cnt))
        else:
            self.min_ds = np.full(B, np.nan)
            check(self._lib.gf_flamelet_batch_steady_psitc(self._h, ctypes.byref(o), iptr(sel), iptr(it), iptr(conv),
                                                           dptr(res), dptr(self.min_ds), cnt))
        self.counters = {'iterations': cnt[0], 'jacobian_refreshes': cnt[1], 'rhs_evaluations': cnt[2],
                         'synchronizing_steps': cnt[3]}
        self.residuals = res
        return self._finish(sel, conv, it, raise_on_failure,
                            'steady flamelet Newton solve' if newton else 'steady flamelet solve')

    def steady_solve_newton(self, tolerance=1.e-6, lu_precision=64, select=None, **newton_args):
        """Flamelet.steady_solve_newton (flamelet.py:1352-1485) for the selected flamelets (all by default, or
        a boolean mask), each from its initial guess; newton_args: max_iterations (10), max_factor_line_search,
        max_allowed_residual, min_allowable_state_var, log_rate, raise_on_failure (True). Returns (states,
        converged, iterations) over the whole batch; unselected flamelets report not converged."""
        return self._steady(True, tolerance, lu_precision, select, dict(newton_args))

    def steady_solve_psitc(self, tolerance=1.e-6, lu_precision=64, select=None, **psitc_args):
        """Pseudo-transient continuation (Flamelet.steady_solve_psitc, flamelet.py:1487-1703) of the selected
        flamelets from their initial guesses (psitc_args: its keyword arguments, plus raise_on_failure, True
        by default). Returns (states, converged, iterations); the smallest final pseudo-time step of each
        flamelet is left in min_ds."""
        return self._steady(False, tolerance, lu_precision, select, dict(psitc_args))

    def integrate_to_steady(self, steady_tolerance=1.e-4, lu_precision=64, select=None, first_time_step=1.e-6,
                            max_time_step=1.e-3, minimum_time_step_count=40, transient_tolerance=1.e-10,
                            maximum_steps_per_jacobian=10, nonlinear_solve_tolerance=1.e-12, max_nonlinear_iter=20,
                            maximum_time_step_count=None, stall_window=None, log_rate=0, raise_on_failure=True,
                            **unused):
        """Flamelet.integrate_to_steady (flamelet.py:1215-1237; Flamelet.integrate with block-Thomas solves,
        flamelet.py:1046-1213; same defaults) for the selected flamelets, each from its initial guess, all on the
        GPU at once. Every flamelet keeps its own time, step size, stage Newton convergence, Jacobian-refresh
        state and steadiness test, exactly as the reference's step loop (integrator.py:100-130, 229-713) runs
        for one flamelet: ESDIRK64 stages solved by the frozen-Jacobian Newton method (nonlinear.py:185-268) with
        block-Thomas solves of dt*gamma*J - I, PI step control (stepcontrol.py:50-117), negative entries clipped
        after every step, and a stop when |dx * 1/scale|_inf / dt < steady_tolerance. first_time_step: a scalar
        or one value per flamelet. Returns (states, converged, steps); converged where that test stopped the
        flamelet (not the step limit).

        Two optional bounds, neither in the reference (whose loop has no limit): maximum_time_step_count stops a
        flamelet after that many steps, and stall_window stops one whose steadiness measure has not halved within
        that many steps. They end the two ways a flamelet can fail to reach a steady tolerance that lies below the
        rounding floor of its residual: the steadiness measure hovers at that floor, or the stage Newton solves
        cannot reach nonlinear_solve_tolerance at the current step size, every step counts as slow (step size
        x0.8) and the step size collapses. transient_stop[f] records why each flamelet stopped ('steady',
        'step limit', 'stalled' or None if not selected)."""
        import ctypes
        import sys
        import torch
        from ._native import check, iptr
        from .time_integration import KennedyCarpenterS6P4Q3, SimpleNewtonSolver
        for k in unused:
            if k not in ('write_log', 'save_first_and_last_only', 'stop_at_time', 'stop_criteria'):
                raise TypeError('unknown transient argument ' + k)
        sel = self._select(select).astype(bool)
        B = len(self._flamelets)
        dev = torch.device('cuda:0')
        lib, h = self._lib, self._h
        tab = KennedyCarpenterS6P4Q3(SimpleNewtonSolver())
        A, bw, bh, g = tab.A, tab.b, tab.bh, tab.gamma
        seg = torch.tensor(np.repeat(np.arange(B), np.diff(self._offsets)), device=dev)
        w = 1. / torch.tensor(np.concatenate([f._variable_scales for f in self._flamelets]), device=dev)
        x = torch.tensor(self._x0, device=dev)
        maxsteps = np.inf if maximum_time_step_count is None else int(maximum_time_step_count)
        dt = np.broadcast_to(np.asarray(first_time_step, dtype=np.float64), (B,)).copy()
        t = np.zeros(B)
        nsteps = np.zeros(B, dtype=np.int64)
        setup_count = np.zeros(B, dtype=np.int64)
        eval_setup = np.ones(B, dtype=bool)
        running = sel.copy()
        steady = np.zeros(B, dtype=bool)
        resid = np.full(B, np.nan)
        best = np.full(B, np.inf)        # steadiness at the last halving, and the step count then
        best_step = np.zeros(B, dtype=np.int64)
        stalled = np.zeros(B, dtype=bool)
        stats = {'steps': 0, 'nonlinear_iter': 0, 'projector_setups': 0, 'failed_steps': 0, 'sweeps': 0}
        mask = lambda m: np.ascontiguousarray(m, dtype=np.int32)
        on_dofs = lambda m: torch.from_numpy(np.asarray(m, dtype=bool)).to(dev)[seg]

        def fnorm(v):
            """per flamelet max |v * 1/scale|, NaN where any entry is not finite"""
            m = torch.zeros(B, dtype=torch.float64, device=dev).scatter_reduce_(0, seg, (v * w).abs(), 'amax')
            bad = torch.zeros(B, dtype=torch.float64, device=dev).scatter_reduce_(
                0, seg, (~torch.isfinite(v)).to(torch.float64), 'amax')
            return torch.where(bad > 0, torch.full_like(m, np.nan), m).cpu().numpy()

        def rhs(m, xv, out):
            torch.cuda.current_stream().synchronize()
            check(lib.gf_flamelet_batch_transient_rhs(h, iptr(mask(m)), xv.data_ptr(), out.data_ptr()))

        def setup(m, xv):
            torch.cuda.current_stream().synchronize()
            pre = np.ascontiguousarray(dt * g)
            check(lib.gf_flamelet_batch_transient_setup(h, iptr(mask(m)), xv.data_ptr(),
                                                        pre.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                        int(lu_precision)))
            stats['projector_setups'] += int(m.sum())

        def solve(m, r, out):
            torch.cuda.current_stream().synchronize()
            check(lib.gf_flamelet_batch_transient_solve(h, iptr(mask(m)), r.data_ptr(), out.data_ptr()))

        d = torch.zeros_like(x)
        while True:
            run = running & ~steady
            if not run.any():
                break
            stats['sweeps'] += 1
            need = run & eval_setup
            if need.any():
                setup(need, x)
            setup_count[run] += 1
            hd = torch.from_numpy(np.where(run, dt, 0.)).to(dev)[seg]
            xn = x.clone()
            f0 = torch.zeros_like(x)
            rhs(run, xn, f0)
            ks = [f0]
            xi = xn.clone()
            converged = run.copy()
            for s in range(1, 6):
                prior = sum(A[s, j] * ks[j] for j in range(s))
                fi = ks[-1].clone()
                r = hd * (g * fi + prior) - (xi - xn)
                done = ~run
                for it in range(max_nonlinear_iter):
                    act = ~done
                    solve(act, r, d)
                    xi = torch.where(on_dofs(act), xi - d, xi)
                    rhs(act, xi, fi)
                    r = hd * (g * fi + prior) - (xi - xn)
                    stats['nonlinear_iter'] += int(act.sum())
                    done = done | (act & (fnorm(r) < nonlinear_solve_tolerance))
                    if done.all():
                        break
                converged &= done  # a stage left unconverged also marks the step's solve slow
                ks.append(fi.clone())
            dstate = hd * sum(bw[j] * ks[j] for j in range(6))
            err = fnorm(dstate - hd * sum(bh[j] * ks[j] for j in range(6)))
            with np.errstate(invalid='ignore', divide='ignore'):
                res = fnorm(dstate) / dt
            acc = run & np.isfinite(res)
            rej = run & ~np.isfinite(res)
            xnew = x + dstate
            xnew = torch.where(xnew < 0., torch.zeros_like(xnew), xnew)
            x = torch.where(on_dofs(acc), xnew, x)
            recent = dt.copy()
            t[acc] += dt[acc]
            nsteps[acc] += 1
            with np.errstate(divide='ignore', over='ignore'):
                ramp = np.minimum(1.1, (transient_tolerance / err) ** 0.1333333333)
            dn = np.where(err < 1.e-16, dt * 1.1, dt * ramp)  # PI control (its history ratio is 1 here)
            dn = np.minimum(dn, max_time_step)
            # Jacobian refresh rule (integrator.py:100-130)
            rate = acc & (setup_count == maximum_steps_per_jacobian)
            slow = acc & ~rate & ~converged
            grow = acc & ~rate & ~slow & (dn > recent * 1.05)
            shrink = acc & ~rate & ~slow & ~grow & (dn < recent * 0.9)
            setup_count[rate] = 0
            dn = np.where(slow, dn * 0.8, dn)
            dt = np.where(acc, dn, dt)
            eval_setup = np.where(acc, rate | slow | grow | shrink, eval_setup)
            dt = np.where(rej, dt * 0.8, dt)
            eval_setup = eval_setup | rej
            stats['failed_steps'] += int(rej.sum())
            resid = np.where(run, res, resid)
            stop_max = run & (nsteps > maxsteps)
            steady = steady | (run & ~stop_max & (res < steady_tolerance))
            if stall_window is not None:
                halved = acc & (res < 0.5 * best)
                best = np.where(halved, res, best)
                best_step = np.where(halved, nsteps, best_step)
                stalled = stalled | (run & ~steady & ~stop_max & (nsteps - best_step > stall_window))
            running = running & ~stop_max & ~stalled
            if log_rate and stats['sweeps'] % log_rate == 0:
                left = running & ~steady
                print('transient sweep %d: %d of %d flamelets stepping, max steps %d, min t %.3e, '
                      'max steady residual %.3e' % (stats['sweeps'], left.sum(), sel.sum(), nsteps[sel].max(),
                                                    t[left].min() if left.any() else t[sel].min(),
                                                    np.nanmax(resid[left]) if left.any() else 0.),
                      file=sys.stderr, flush=True)
        stats['steps'] = int(nsteps[sel].max()) if sel.any() else 0
        fin = torch.zeros_like(x)
        if sel.any():
            rhs(sel, x, fin)  # also makes the final states the batch's current states
        self.residuals = np.where(sel, fnorm(fin), np.nan)
        self.transient_info = stats
        self.transient_times = t
        self.transient_stop = np.array([('steady' if steady[f] else 'stalled' if stalled[f] else 'step limit')
                                        if sel[f] else None for f in range(B)], dtype=object)
        conv = (sel & steady).astype(np.int32)
        return self._finish(sel.astype(np.int32), conv, nsteps.astype(np.int32), raise_on_failure,
                            'steady flamelet time stepping')

    def compute_steady_state(self, tolerance=1.e-6, lu_precision=64, use_psitc=True, newton_args=None,
                             psitc_args=None, transient_args=None, raise_on_failure=True, log_rate=0,
                             use_time_stepping=True):
        """Flamelet.compute_steady_state (flamelet.py:1705-1778) for every flamelet: Newton (38 iterations) from
        the initial guess; for those it leaves unconverged, pseudo-transient continuation (400 iterations)
        from the initial guess; for those still unconverged, ESDIRK64 time stepping to steady state from the
        initial guess with a first step of 1e-2 times psitc's final smallest pseudo-time step (skipped when
        use_time_stepping is False). Each stage is batched over its flamelets. Returns (states, converged,
        iterations);
        method[f] names the stage that converged flamelet f ('newton', 'psitc', 'transient' or None).
        Unlike the reference, the time-stepping stage is bounded by default: at most 5000 steps, and a flamelet
        whose steadiness measure has not halved in 1000 steps is stopped as stalled (integrate_to_steady's
        maximum_time_step_count and stall_window; transient_args can change or remove both)."""
        B = len(self._flamelets)
        self.residuals = np.full(B, np.nan)
        self.method = np.array([None] * B, dtype=object)
        self.stage_counters = {}
        nargs = {'max_iterations': 38, 'log_rate': log_rate}
        nargs.update(newton_args or {})
        nargs['raise_on_failure'] = False
        states, conv, it = self.steady_solve_newton(tolerance, lu_precision, **nargs)
        res = self.residuals.copy()
        self.method[conv] = 'newton'
        self.stage_counters['newton'] = self.counters
        todo = ~conv
        mds = np.full(B, 1.e-6)
        if use_psitc and todo.any():
            pargs = {'max_iterations': 400, 'log_rate': log_rate}
            pargs.update(psitc_args or {})
            pargs['raise_on_failure'] = False
            _, c2, it2 = self.steady_solve_psitc(tolerance, lu_precision, select=todo, **pargs)
            mds[todo] = self.min_ds[todo]
            res[todo] = self.residuals[todo]
            now = todo & c2
            self.method[now] = 'psitc'
            it[todo] = it2[todo]
            conv = conv | now
            todo = ~conv
            self.stage_counters['psitc'] = self.counters
        if use_time_stepping and todo.any():
            targs = {'steady_tolerance': tolerance, 'transient_tolerance': 1.e-8, 'max_time_step': 1.e4,
                     'first_time_step': 1.e-2 * mds, 'maximum_steps_per_jacobian': 10, 'log_rate': log_rate,
                     'maximum_time_step_count': 5000, 'stall_window': 1000}
            targs.update(transient_args or {})
            targs['raise_on_failure'] = False
            _, c3, st3 = self.integrate_to_steady(lu_precision=lu_precision, select=todo, **targs)
            res[todo] = self.residuals[todo]
            now = todo & c3
            self.method[now] = 'transient'
            it[todo] = st3[todo]
            conv = conv | now
            self.stage_counters['transient'] = dict(self.transient_info)
        self.residuals = res
        states = self.split(self.get_state())
        if raise_on_failure and not np.all(conv):
            raise RuntimeError('steady flamelet solve did not converge for %d of %d flamelets'
                               % (int(B - conv.sum()), B))
        return states, conv, it
