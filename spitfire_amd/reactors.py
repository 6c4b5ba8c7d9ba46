"""HomogeneousReactor: zero-dimensional reactors integrated with ESDIRK64 + Newton + PI control on the griffon drop-in
(src/spitfire/chemistry/reactors.py:48-800).

Configurations: isobaric (state [T, Y_0 .. Y_{ns-2}]) or isochoric (state [rho, T, Y_0 .. Y_{ns-2}]); heat transfer
adiabatic, isothermal or diathermal (convection and radiation through the reactor's surface, the surface-to-volume
ratio from a shape and a characteristic length or volume); closed or open (mass exchange with a feed stream over a
residence time). Diathermal and feed parameters may be constants or functions of time. Each Newton setup evaluates
the analytic Jacobian on the GPU, forms dt*gamma*J - I and factors it with LAPACK (reactors.py:340-356), or with
SuperLU for linear_solver='superlu'; negative mass fractions are clipped after every step (reactors.py:505-507).
The reference returns a spitfire Library; here integrate returns a ReactorOutput with the same fields (time values,
temperature, pressure or density, 'mass fraction <name>').
"""
import numpy as np
from scipy.linalg import lu_factor, lu_solve

from .time_integration import KennedyCarpenterS6P4Q3, PIController, SimpleNewtonSolver, odesolve


class ReactorOutput(dict):
    def __init__(self, t):
        super().__init__()
        self.time_values = np.asarray(t)


def _cube_root(v):
    return np.power(v, 1. / 3.)


# surface area / volume of each shape from its characteristic length a (edge, or radius for the sphere and the
# capsule, whose cylinder is 2a long) or from its volume v
_SHAPES = {
    'cube': (lambda a: 6. / a, lambda v: 6. / _cube_root(v)),
    'sphere': (lambda a: 3. / a, lambda v: 3. / _cube_root(3. * v / (4. * np.pi))),
    'capsule': (lambda a: 12. / (5. * a), lambda v: 12. / (5. * _cube_root(3. * v / (10. * np.pi)))),
    'tetrahedron': (lambda a: 6. * np.sqrt(6.) / a, lambda v: 6. * np.sqrt(6.) / _cube_root(12. * v / np.sqrt(2.))),
    'octahedron': (lambda a: 3. * np.sqrt(6.) / a, lambda v: 3. * np.sqrt(6.) / _cube_root(3. * v / np.sqrt(2.))),
    'icosahedron': (lambda a: 12. * np.sqrt(3.) / ((3. + np.sqrt(5.)) * a),
                    lambda v: 12. * np.sqrt(3.) / ((3. + np.sqrt(5.)) * _cube_root(12. * v / 5. / (3. + np.sqrt(5.))))),
}


class HomogeneousReactor:
    _configurations = {'constant pressure': 'isobaric', 'isobaric': 'isobaric',
                       'constant volume': 'isochoric', 'isochoric': 'isochoric'}
    _heat_transfers = ('adiabatic', 'isothermal', 'diathermal')
    _mass_transfers = ('closed', 'open')

    @classmethod
    def get_supported_reactor_shapes(cls):
        return _SHAPES.keys()

    def __init__(self, mech_spec, initial_mixture, configuration='isobaric', heat_transfer='adiabatic',
                 mass_transfer='closed', convection_temperature=None, radiation_temperature=None,
                 convection_coefficient=None, radiative_emissivity=None, shape_dimension_dict=None, mixing_tau=None,
                 feed_temperature=None, feed_mass_fractions=None, feed_density=None, rates_sensitivity_type='dense',
                 sensitivity_transform_type='exact', initial_time=0.):
        def pick(value, options, what):
            if not isinstance(value, str) or value.lower() not in options:
                raise ValueError('Bad %s argument %r; acceptable values: %s' % (what, value, list(options)))
            return value.lower()

        self._configuration = self._configurations[pick(configuration, self._configurations, 'configuration')]
        self._heat_transfer = pick(heat_transfer, self._heat_transfers, 'heat transfer')
        self._mass_transfer = pick(mass_transfer, self._mass_transfers, 'mass transfer')

        def need(value, name, why):
            if value is None:
                raise ValueError('The %s argument is needed but was unspecified: %s' % (name, why))
            return value

        self._params = {k: 0. for k in ('Tc', 'Tr', 'hc', 'eps', 'tau', 'Tfeed', 'rhofeed')}
        self._params['yfeed'] = np.zeros(mech_spec.n_species)
        self._surface_area_to_volume = 0.
        if self._heat_transfer == 'diathermal':
            why = 'heat transfer is diathermal'
            self._params['Tc'] = need(convection_temperature, 'convection_temperature', why)
            self._params['Tr'] = need(radiation_temperature, 'radiation_temperature', why)
            self._params['hc'] = need(convection_coefficient, 'convection_coefficient', why)
            self._params['eps'] = need(radiative_emissivity, 'radiative_emissivity', why)
            sd = need(shape_dimension_dict, 'shape_dimension_dict', why)
            if 'shape' not in sd:
                raise ValueError("The shape_dimension_dict argument did not have the required 'shape' key")
            shape = pick(sd['shape'], _SHAPES, 'shape')
            if ('char. length' in sd) == ('volume' in sd):
                raise ValueError("shape_dimension_dict needs exactly one of 'char. length' and 'volume'")
            l2sov, v2sov = _SHAPES[shape]
            self._surface_area_to_volume = float(l2sov(sd['char. length']) if 'char. length' in sd
                                                 else v2sov(sd['volume']))
        if self._mass_transfer == 'open':
            why = 'mass transfer is open'
            self._params['tau'] = np.inf if mixing_tau is None else mixing_tau
            self._params['Tfeed'] = need(feed_temperature, 'feed_temperature', why)
            self._params['yfeed'] = need(feed_mass_fractions, 'feed_mass_fractions', why)
            if self._configuration == 'isochoric':
                self._params['rhofeed'] = need(feed_density, 'feed_density', why)
            elif feed_density is not None:
                self._params['rhofeed'] = feed_density
        self._time_functions = {k for k, v in self._params.items() if callable(v)}

        self._rates_sensitivity_option = {'dense': 0, 'no-TBAF': 1, 'sparse': 2}[rates_sensitivity_type]
        self._sensitivity_transform_option = {'exact': 0}[sensitivity_transform_type]
        self._heat_transfer_option = {'adiabatic': 0, 'isothermal': 1, 'diathermal': 2}[self._heat_transfer]
        self._is_open = self._mass_transfer == 'open'
        self._mechanism = mech_spec
        self._griffon = mech_spec.griffon
        self._n_species = mech_spec.n_species
        self._n_reactions = mech_spec.n_reactions
        self._initial_pressure = float(initial_mixture.P)
        self._initial_temperature = float(initial_mixture.T)
        self._initial_mass_fractions = np.copy(initial_mixture.Y)
        self._current_pressure = self._initial_pressure
        self._current_temperature = self._initial_temperature
        self._current_mass_fractions = np.copy(self._initial_mass_fractions)
        if self._configuration == 'isobaric':
            self._n_equations = self._n_species
            self._temperature_index = 0
            self._initial_state = np.hstack((self._initial_temperature, self._initial_mass_fractions[:-1]))
        else:
            self._n_equations = self._n_species + 1
            self._temperature_index = 1
            self._initial_state = np.hstack((initial_mixture.density, self._initial_temperature,
                                             self._initial_mass_fractions[:-1]))
        self._current_state = np.copy(self._initial_state)
        self._initial_time = float(initial_time)
        self._current_time = float(initial_time)
        self._variable_scales = np.ones(self._n_equations)
        self._variable_scales[self._temperature_index] = 1.e3
        self._lhs = None

    # ---------------- griffon calls (reactors.py:539-600) ----------------
    def _values(self, t):
        v = {k: (f(t) if k in self._time_functions else f) for k, f in self._params.items()}
        v['yfeed'] = np.ascontiguousarray(v['yfeed'], dtype=np.float64)
        if v['yfeed'].size != self._n_species:
            v['yfeed'] = np.zeros(self._n_species)
        return v

    def _griffon_args(self, t):
        v = self._values(t)
        first = self._initial_pressure if self._configuration == 'isobaric' else float(v['rhofeed'])
        return (first, float(v['Tfeed']), v['yfeed'], float(v['tau']), float(v['Tc']), float(v['Tr']),
                float(v['hc']), float(v['eps']), self._surface_area_to_volume, self._heat_transfer_option,
                self._is_open)

    def _rhs(self, t, state):
        k = np.zeros(self._n_equations)
        st = np.ascontiguousarray(state, dtype=np.float64)
        if self._configuration == 'isobaric':
            self._griffon.reactor_rhs_isobaric(st, *self._griffon_args(t), k)
        else:
            self._griffon.reactor_rhs_isochoric(st, *self._griffon_args(t), k)
        return k

    def _jac(self, state, t=None):
        n = self._n_equations
        k, j = np.zeros(n), np.zeros(n * n)
        st = np.ascontiguousarray(state, dtype=np.float64)
        args = self._griffon_args(self._current_time if t is None else t)
        if self._configuration == 'isobaric':
            self._griffon.reactor_jac_isobaric(st, *args, self._rates_sensitivity_option,
                                               self._sensitivity_transform_option, k, j)
        else:
            self._griffon.reactor_jac_isochoric(st, *args, self._rates_sensitivity_option, k, j)
        return j.reshape((n, n), order='F')

    def _lapack_setup(self, t, state, prefactor):
        j = self._jac(state, t) * prefactor
        j[np.diag_indices(self._n_equations)] -= 1.
        self._lhs = lu_factor(j)

    def _lapack_solve(self, residual):
        return lu_solve(self._lhs, residual)

    def _superlu_setup(self, t, state, prefactor):
        from scipy.sparse import csc_matrix
        from scipy.sparse.linalg import splu
        j = self._jac(state, t) * prefactor
        j[np.diag_indices(self._n_equations)] -= 1.
        m = csc_matrix(j)
        m.eliminate_zeros()
        self._lhs = splu(m)

    def _superlu_solve(self, residual):
        return self._lhs.solve(residual)

    def integrate(self, stop_at_time=None, stop_at_steady=None, stop_criteria=None, first_time_step=1.e-6,
                  max_time_step=1.e6, minimum_time_step_count=40, transient_tolerance=1.e-10, write_log=False,
                  log_rate=100, maximum_steps_per_jacobian=1, nonlinear_solve_tolerance=1.e-12, linear_solver='lapack',
                  plot=None, stepper_type=KennedyCarpenterS6P4Q3, nlsolver_type=SimpleNewtonSolver,
                  stepcontrol_type=PIController, extra_integrator_args=None, extra_stepper_args=None,
                  extra_nlsolver_args=None, extra_stepcontrol_args=None, save_first_and_last_only=False,
                  return_info=False):
        """reactors.py:429-678: ESDIRK64/Newton/PI defaults, LAPACK (or SuperLU) linear solves."""
        if linear_solver not in ('lapack', 'superlu'):
            raise ValueError("linear_solver must be 'lapack' or 'superlu'")
        weights = 1. / self._variable_scales
        newton = nlsolver_type(**dict({'evaluate_jacobian_every_iter': False, 'norm_weighting': weights,
                                       'tolerance': nonlinear_solve_tolerance}, **(extra_nlsolver_args or {})))
        stepper = stepper_type(**dict({'nonlinear_solver': newton, 'norm_weighting': weights},
                                      **(extra_stepper_args or {})))
        control = stepcontrol_type(**dict({'first_step': first_time_step, 'max_step': max_time_step,
                                           'target_error': transient_tolerance}, **(extra_stepcontrol_args or {})))
        setup = self._lapack_setup if linear_solver == 'lapack' else self._superlu_setup
        solve = self._lapack_solve if linear_solver == 'lapack' else self._superlu_solve

        def post_step(t, state, *args):
            state[state < 0.] = 0.
            return state

        args = dict(initial_time=self._current_time, stop_at_time=stop_at_time, stop_at_steady=stop_at_steady,
                    stop_criteria=stop_criteria, minimum_time_step_count=minimum_time_step_count,
                    post_step_callback=post_step, method=stepper, step_size=control, linear_setup=setup,
                    linear_solve=solve, linear_setup_rate=maximum_steps_per_jacobian, norm_weighting=weights,
                    save_each_step=not save_first_and_last_only, return_info=True)
        args.update(extra_integrator_args or {})
        out = odesolve(self._rhs, self._current_state, **args)
        if save_first_and_last_only:
            state, t, _, info = out
            ts, states = np.array([t]), state[None, :]
        else:
            ts, states, info = out
        self._current_state = np.copy(states[-1])
        self._current_time = float(ts[-1])
        self.info = info
        lib = ReactorOutput(ts)
        ti = self._temperature_index
        ynm1 = self._current_state[ti + 1:]
        self._current_mass_fractions = np.hstack((ynm1, 1. - np.sum(ynm1)))
        self._current_temperature = float(self._current_state[ti])
        if self._configuration == 'isobaric':
            self._current_pressure = self._initial_pressure
            lib['temperature'] = states[:, 0]
            lib['pressure'] = self._current_pressure + np.zeros_like(ts, dtype=np.float64)
        else:
            mw = self._mechanism.molecular_weights
            self._current_pressure = float(self._current_state[0] * self._current_state[1] *
                                           self._mechanism.gas_constant * np.sum(self._current_mass_fractions / mw))
            lib['density'] = states[:, 0]
            lib['temperature'] = states[:, 1]
        names = self._mechanism.species_names
        last = np.ones_like(ts, dtype=np.float64)
        for i, s in enumerate(names[:-1]):
            lib['mass fraction ' + s] = states[:, ti + 1 + i]
            last = last - states[:, ti + 1 + i]
        lib['mass fraction ' + names[-1]] = last
        return (lib, info) if return_info else lib

    def integrate_to_time(self, final_time, **kwargs):
        return self.integrate(stop_at_time=final_time, **kwargs)

    def integrate_to_steady(self, steady_tolerance=1.e-6, **kwargs):
        return self.integrate(stop_at_steady=steady_tolerance, **kwargs)

    def integrate_to_steady_after_ignition(self, steady_tolerance=1.e-4, delta_temperature_ignition=400., **kwargs):
        """Steady state reached after the temperature has risen by delta_temperature_ignition (reactors.py:712-739)."""
        if self._heat_transfer == 'isothermal':
            raise ValueError('integrate_to_steady_after_ignition() is not supported for an isothermal reactor')
        T0, ti = self._initial_temperature, self._temperature_index

        def stop(t, state, residual, *args, **kw):
            return state[ti] > T0 + delta_temperature_ignition and residual < steady_tolerance
        return self.integrate(stop_criteria=stop, **kwargs)

    def compute_ignition_delay(self, delta_temperature_ignition=400., minimum_allowable_residual=1.e-12,
                               return_solution=False, **kwargs):
        """Time at which the temperature has risen by delta_temperature_ignition (reactors.py:741-790); raises when
        the residual falls below minimum_allowable_residual first (no ignition)."""
        if self._heat_transfer == 'isothermal':
            raise ValueError('compute_ignition_delay() is not supported for an isothermal reactor')
        T0, ti = self._initial_temperature, self._temperature_index

        def stop(t, state, residual, *args, **kw):
            if residual <= minimum_allowable_residual:
                raise ValueError('residual < minimum allowable value (%g): the reactor does not ignite'
                                 % minimum_allowable_residual)
            return state[ti] > T0 + delta_temperature_ignition
        lib = self.integrate(stop_criteria=stop, **kwargs)
        tau = float(lib.time_values[-1])
        return (tau, lib) if return_solution else tau

    # ---------------- properties (reactors.py:360-427) ----------------
    @property
    def initial_state(self):
        return np.copy(self._initial_state)

    @property
    def current_state(self):
        return np.copy(self._current_state)

    @property
    def initial_temperature(self):
        return self._initial_temperature

    @property
    def current_temperature(self):
        return float(self._current_state[self._temperature_index])

    @property
    def initial_pressure(self):
        return self._initial_pressure

    @property
    def current_pressure(self):
        return self._current_pressure

    @property
    def initial_mass_fractions(self):
        return np.copy(self._initial_mass_fractions)

    @property
    def current_mass_fractions(self):
        y = self._current_state[self._temperature_index + 1:]
        return np.hstack((y, 1. - np.sum(y)))

    @property
    def initial_time(self):
        return self._initial_time

    @property
    def current_time(self):
        return self._current_time

    @property
    def n_species(self):
        return self._n_species

    @property
    def n_reactions(self):
        return self._n_reactions
