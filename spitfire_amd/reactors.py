"""HomogeneousReactor: closed adiabatic (or isothermal) isobaric reactors integrated with ESDIRK64 + Newton +
PI control on the griffon drop-in (src/spitfire/chemistry/reactors.py:254-710).

The state is [T, Y_0 .. Y_{ns-2}] (reactors.py:306-312); each Newton setup evaluates the analytic Jacobian,
forms dt*gamma*J - I and factors it with LAPACK (reactors.py:340-356); negative mass fractions are clipped
after every step (reactors.py:505-507). The reference returns a spitfire Library; here integrate returns a
ReactorOutput with the same fields (time values, temperature, pressure, 'mass fraction <name>').
"""
import numpy as np
from scipy.linalg import lu_factor, lu_solve

from .time_integration import KennedyCarpenterS6P4Q3, PIController, SimpleNewtonSolver, odesolve


class ReactorOutput(dict):
    def __init__(self, t):
        super().__init__()
        self.time_values = np.asarray(t)


class HomogeneousReactor:
    def __init__(self, mech_spec, initial_mixture, configuration='isobaric', heat_transfer='adiabatic',
                 mass_transfer='closed', rates_sensitivity_type='dense', sensitivity_transform_type='exact',
                 initial_time=0.):
        if configuration != 'isobaric':
            raise NotImplementedError('isochoric reactors are not supported by this implementation')
        if mass_transfer != 'closed':
            raise NotImplementedError('open reactors are not supported by this implementation')
        if heat_transfer not in ('adiabatic', 'isothermal'):
            raise NotImplementedError('diathermal reactors are not supported by this implementation')
        self._mechanism = mech_spec
        self._griffon = mech_spec.griffon
        self._configuration = configuration
        self._heat_transfer_option = {'adiabatic': 0, 'isothermal': 1}[heat_transfer]
        self._rates_sensitivity_option = {'dense': 0, 'no-TBAF': 1, 'sparse': 2}[rates_sensitivity_type]
        self._sensitivity_transform_option = {'exact': 0}[sensitivity_transform_type]
        self._initial_pressure = float(initial_mixture.P)
        self._initial_temperature = float(initial_mixture.T)
        self._initial_mass_fractions = np.copy(initial_mixture.Y)
        self._n_species = mech_spec.n_species
        self._n_equations = self._n_species
        self._current_state = np.hstack((self._initial_temperature, self._initial_mass_fractions[:-1]))
        self._initial_state = np.copy(self._current_state)
        self._current_time = float(initial_time)
        self._variable_scales = np.ones(self._n_equations)
        self._variable_scales[0] = 1.e3
        self._lhs = None

    # griffon wrappers (reactors.py:539-562): closed reactor, no heat-transfer or feed parameters
    def _rhs(self, t, state):
        k = np.zeros(self._n_equations)
        null = np.zeros(1)
        self._griffon.reactor_rhs_isobaric(state, self._initial_pressure, 0., null, 0., 0., 0., 0., 0., 0.,
                                           self._heat_transfer_option, False, k)
        return k

    def _jac(self, state):
        k = np.zeros(self._n_equations)
        j = np.zeros(self._n_equations * self._n_equations)
        null = np.zeros(1)
        self._griffon.reactor_jac_isobaric(state, self._initial_pressure, 0., null, 0., 0., 0., 0., 0., 0.,
                                           self._heat_transfer_option, False, self._rates_sensitivity_option,
                                           self._sensitivity_transform_option, k, j)
        return j.reshape((self._n_equations, self._n_equations), order='F')

    def _lapack_setup(self, t, state, prefactor):
        j = self._jac(state) * prefactor
        j[np.diag_indices(self._n_equations)] -= 1.
        self._lhs = lu_factor(j)

    def _lapack_solve(self, residual):
        return lu_solve(self._lhs, residual)

    def integrate(self, stop_at_time=None, stop_at_steady=None, stop_criteria=None, first_time_step=1.e-6,
                  max_time_step=1.e6, minimum_time_step_count=40, transient_tolerance=1.e-10,
                  maximum_steps_per_jacobian=1, nonlinear_solve_tolerance=1.e-12, save_first_and_last_only=False,
                  return_info=False):
        """reactors.py:429-678 with the LAPACK linear solver and the ESDIRK64/Newton/PI defaults"""
        newton = SimpleNewtonSolver(evaluate_jacobian_every_iter=False, norm_weighting=1. / self._variable_scales,
                                    tolerance=nonlinear_solve_tolerance)
        stepper = KennedyCarpenterS6P4Q3(newton, norm_weighting=1. / self._variable_scales)
        control = PIController(first_step=first_time_step, max_step=max_time_step, target_error=transient_tolerance)

        def post_step(t, state, *args):
            state[state < 0.] = 0.
            return state

        out = odesolve(self._rhs, self._current_state, initial_time=self._current_time, stop_at_time=stop_at_time,
                       stop_at_steady=stop_at_steady, stop_criteria=stop_criteria,
                       minimum_time_step_count=minimum_time_step_count, post_step_callback=post_step, method=stepper,
                       step_size=control, linear_setup=self._lapack_setup, linear_solve=self._lapack_solve,
                       linear_setup_rate=maximum_steps_per_jacobian, norm_weighting=1. / self._variable_scales,
                       save_each_step=not save_first_and_last_only, return_info=True)
        if save_first_and_last_only:
            state, t, _, info = out
            ts, states = np.array([t]), state[None, :]
        else:
            ts, states, info = out
        self._current_state = np.copy(states[-1])
        self._current_time = float(ts[-1])
        self.info = info
        lib = ReactorOutput(ts)
        lib['temperature'] = states[:, 0]
        lib['pressure'] = self._initial_pressure + np.zeros_like(ts, dtype=np.float64)
        names = self._mechanism.species_names
        last = np.ones_like(ts, dtype=np.float64)
        for i, s in enumerate(names[:-1]):
            lib['mass fraction ' + s] = states[:, 1 + i]
            last = last - states[:, 1 + i]
        lib['mass fraction ' + names[-1]] = last
        return (lib, info) if return_info else lib

    def integrate_to_time(self, final_time, **kwargs):
        return self.integrate(stop_at_time=final_time, **kwargs)

    def integrate_to_steady(self, steady_tolerance=1.e-6, **kwargs):
        return self.integrate(stop_at_steady=steady_tolerance, **kwargs)

    @property
    def current_state(self):
        return np.copy(self._current_state)

    @property
    def current_temperature(self):
        return float(self._current_state[0])

    @property
    def current_mass_fractions(self):
        y = self._current_state[1:]
        return np.hstack((y, 1. - np.sum(y)))

    @property
    def current_time(self):
        return self._current_time
