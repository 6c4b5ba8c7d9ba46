"""Adaptive implicit time integration for the reactor path: the Kennedy & Carpenter ESDIRK64 stepper, a
frozen-Jacobian Newton solver, the PI step controller and the step loop, restated from
src/spitfire/time/methods.py:441-612 (KennedyCarpenterS6P4Q3), nonlinear.py:143-268 (SimpleNewtonSolver),
stepcontrol.py:50-117 (PIController) and integrator.py:100-130, 229-713 (odesolve), reduced to the options the
reactor drivers use.

The ESDIRK tableau is the implicit part of Kennedy & Carpenter's ARK4(3)6L[2]SA (Appl. Numer. Math. 44,
2003), gamma = 1/4, with the reference's rational coefficients; the embedded third-order solution gives the
error estimate.
"""
import numpy as np

_INF = np.inf


class SolverOutput:
    __slots__ = ('solution', 'rhs_at_converged', 'iter', 'converged', 'slow_convergence', 'projector_setups')

    def __init__(self, solution, rhs_at_converged, iter, converged, slow_convergence, projector_setups):
        self.solution = solution
        self.rhs_at_converged = rhs_at_converged
        self.iter = iter
        self.converged = converged
        self.slow_convergence = slow_convergence
        self.projector_setups = projector_setups


class SimpleNewtonSolver:
    """Newton iterations x <- x - solve(r(x)) with the linear operator set up by the step loop; converged when
    ||r * weighting||_ord < tolerance (nonlinear.py:185-268)."""

    def __init__(self, evaluate_jacobian_every_iter=False, max_nonlinear_iter=20, slowness_detection_iter=_INF,
                 must_converge=False, norm_weighting=1., norm_order=_INF, tolerance=1.e-12, raise_naninf=False):
        self.evaluate_jacobian_every_iter = evaluate_jacobian_every_iter
        self.max_nonlinear_iter = max_nonlinear_iter
        self.slowness_detection_iter = slowness_detection_iter
        self.must_converge = must_converge
        self.norm_weighting = norm_weighting
        self.norm_order = norm_order
        self.tolerance = tolerance
        self.raise_naninf = raise_naninf

    @property
    def setup_projector_in_governor(self):
        return not self.evaluate_jacobian_every_iter

    def _check(self, x, what):
        if self.raise_naninf and np.any(~np.isfinite(x)):
            raise ValueError('NaN or Inf detected in Simple Newton Solve: ' + what)

    def __call__(self, residual_method, setup_method, solve_method, initial_guess, initial_rhs):
        solution = np.copy(initial_guess)
        self._check(solution, 'in the initial solution!')
        scalar = solution.size == 1
        nrm = abs if scalar else (lambda x: np.linalg.norm(x, ord=self.norm_order))
        residual, rhs = residual_method(solution, existing_rhs=initial_rhs, evaluate_new_rhs=False)
        setups = 0
        for it in range(1, self.max_nonlinear_iter + 1):
            if self.evaluate_jacobian_every_iter:
                setup_method(solution)
                setups += 1
            dstate = solve_method(residual)
            if isinstance(dstate, tuple):  # the reference's solve methods return (update, iterations, converged)
                dstate = dstate[0]
            self._check(dstate, 'solution update check!')
            solution -= dstate
            residual, rhs = residual_method(solution, evaluate_new_rhs=True)
            if nrm(residual * self.norm_weighting) < self.tolerance:
                return SolverOutput(solution, rhs, it, True, it > self.slowness_detection_iter, setups)
        if self.must_converge:
            raise ValueError('Simple Newton method did not converge and must_converge=True!')
        return SolverOutput(solution, rhs, self.max_nonlinear_iter, False, True, setups)


class StepOutput:
    __slots__ = ('solution_update', 'temporal_error', 'nonlinear_iter', 'nonlinear_converged',
                 'slow_nonlinear_convergence', 'projector_setups')

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)


class BackwardEulerS1P1Q1:
    """Backward Euler, x_{n+1} = x_n + dt f(t + dt, x_{n+1}), with the error estimate of methods.py:339-398: the
    difference between a one-iteration Newton predictor and the converged solution. As in the reference, the
    second solve starts from x_n but takes the predictor's right-hand side as its initial one."""
    name = 'backward Euler'
    order = 1
    is_implicit = True
    is_adaptive = True
    implicit_coefficient = 1.

    def __init__(self, nonlinear_solver, norm_weighting=1., norm_order=_INF):
        self.nonlinear_solver = nonlinear_solver
        self.norm_weighting = norm_weighting
        self.norm_order = norm_order

    def norm(self, x):
        return np.linalg.norm(x * self.norm_weighting, ord=self.norm_order)

    def single_step(self, state, t, dt, rhs, lhs_setup, lhs_solve):
        x_n = np.copy(state)

        def residual(xa, existing_rhs=None, evaluate_new_rhs=True):
            f = rhs(t + dt, xa) if evaluate_new_rhs else existing_rhs
            return dt * f - (xa - x_n), f

        ns = self.nonlinear_solver
        cap = ns.max_nonlinear_iter
        ns.max_nonlinear_iter = 1
        try:
            pred = ns(residual_method=residual, setup_method=lambda xa: lhs_setup(t + dt, xa), solve_method=lhs_solve,
                      initial_guess=state, initial_rhs=rhs(t + dt, state))
        finally:
            ns.max_nonlinear_iter = cap
        out = ns(residual_method=residual, setup_method=lambda xa: lhs_setup(t + dt, xa), solve_method=lhs_solve,
                 initial_guess=state, initial_rhs=pred.rhs_at_converged)
        return StepOutput(solution_update=out.solution - state, temporal_error=self.norm(out.solution - pred.solution),
                          nonlinear_iter=out.iter, nonlinear_converged=out.converged,
                          slow_nonlinear_convergence=out.slow_convergence, projector_setups=out.projector_setups)


class KennedyCarpenterS6P4Q3:
    """Six-stage, fourth-order ESDIRK with an embedded third-order error estimate (methods.py:441-612)."""
    name = 'Kennedy/Carpenter ESDIRK64'
    order = 4
    is_implicit = True
    is_adaptive = True
    implicit_coefficient = 0.25

    def __init__(self, nonlinear_solver, norm_weighting=1., norm_order=_INF):
        self.nonlinear_solver = nonlinear_solver
        self.norm_weighting = norm_weighting
        self.norm_order = norm_order
        g = 0.25
        self.gamma = g
        A = np.zeros((6, 6))
        A[1, :2] = [0.25, g]
        A[2, :3] = [8611. / 62500., -1743. / 31250., g]
        A[3, :4] = [5012029. / 34652500., -654441. / 2922500., 174375. / 388108., g]
        A[4, :5] = [15267082809. / 155376265600., -71443401. / 120774400., 730878875. / 902184768.,
                    2285395. / 8070912., g]
        A[5, :6] = [82889. / 524892., 0., 15625. / 83664., 69875. / 102672., -2260. / 8211., g]
        self.A = A
        self.b = A[5].copy()
        self.bh = np.array([4586570599. / 29645900160., 0., 178811875. / 945068544., 814220225. / 1159782912.,
                            -3700637. / 11593932., 61727. / 225920.])
        self.c = np.sum(A, axis=1)

    def norm(self, x):
        return np.linalg.norm(x * self.norm_weighting, ord=self.norm_order)

    def single_step(self, state, t, dt, rhs, lhs_setup, lhs_solve):
        state_n = np.copy(state)
        g = self.gamma
        ks = [rhs(t, state)]
        nl_iter, converged, slow, setups = 0, True, False, 0
        x = state
        for s in range(1, 6):
            prior = sum(self.A[s, j] * ks[j] for j in range(s))
            cs = self.c[s]

            def residual(xa, existing_rhs=None, evaluate_new_rhs=True, prior=prior, cs=cs):
                f = rhs(t + cs * dt, xa) if evaluate_new_rhs else existing_rhs
                return dt * (g * f + prior) - (xa - state_n), f

            out = self.nonlinear_solver(residual_method=residual,
                                        setup_method=lambda xa, cs=cs: lhs_setup(t + cs * dt, xa),
                                        solve_method=lhs_solve, initial_guess=x, initial_rhs=ks[-1])
            x = out.solution
            ks.append(out.rhs_at_converged)
            nl_iter += out.iter
            converged = converged and out.converged
            slow = slow or out.slow_convergence
            setups += out.projector_setups
        dstate = dt * sum(self.b[j] * ks[j] for j in range(6))
        dstate_h = dt * sum(self.bh[j] * ks[j] for j in range(6))
        return StepOutput(solution_update=dstate, temporal_error=self.norm(dstate - dstate_h), nonlinear_iter=nl_iter,
                          nonlinear_converged=converged, slow_nonlinear_convergence=slow, projector_setups=setups)


class PIController:
    """PI control of the step size on the embedded error estimate e (stepcontrol.py:50-117):

        dt_new = min(dt * min(max_ramp, (target / e)^ki * (e_rec / e)^kp), max_step)

    where e_rec is the most recently recorded error. The reference records the current error before it forms
    the proportional factor, so e_rec = e and that factor is exactly 1; this controller records the same way.
    An error below 1e-16 ramps the step by max_ramp without recording it."""

    def __init__(self, kp=0.06666666667, ki=0.1333333333, target_error=1.e-4, max_step=1.e4, max_ramp=1.1,
                 first_step=1.e-3):
        self._kp, self._ki = kp, ki
        self._target_error = target_error
        self._max_step = max_step
        self._max_ramp = max_ramp
        self._first_step = first_step
        self._recorded_error = None
        self._recorded_step = None

    def __call__(self, step_count, step, step_output, *args, **kwargs):
        e = step_output.temporal_error
        if e < 1.e-16:
            return min(step * self._max_ramp, self._max_step)
        self._recorded_error, self._recorded_step = e, step
        factor = (self._target_error / e) ** self._ki
        if step_count > 0:
            factor = factor * (self._recorded_error / e) ** self._kp
        return min(step * min(self._max_ramp, factor), self._max_step)

    def first_step_size(self):
        return self._first_step

    def last_step_size(self):
        return self._recorded_step

    def target_error(self):
        return self._target_error

    def step_size_is_constant(self):
        return False


def odesolve(right_hand_side, initial_state, initial_time=0., stop_at_time=None, stop_at_steady=None,
             stop_criteria=None, minimum_time_step_count=0, maximum_time_step_count=_INF, post_step_callback=None,
             method=None, step_size=None, linear_setup=None, linear_solve=None, linear_setup_rate=1,
             norm_weighting=1., save_each_step=False, time_step_reduction_factor_on_failure=0.8,
             time_step_reduction_factor_on_slow_solve=0.8, time_step_increase_factor_to_force_jacobian=1.05,
             time_step_decrease_factor_to_force_jacobian=0.9, return_info=False):
    """The step loop of integrator.py:229-713 for an implicit adaptive method with a user linear operator
    (linear_setup(t, state, dt * gamma) builds dt*gamma*J - I; linear_solve(r) returns its solve). A step is
    rejected only for NaN/Inf updates (the reference's default, non-strict error control)."""
    if method is None:
        method = KennedyCarpenterS6P4Q3(SimpleNewtonSolver())
    if step_size is None:
        step_size = PIController()
    if stop_at_time is None and stop_at_steady is None and stop_criteria is None:
        raise ValueError('odesolve needs stop_at_time, stop_at_steady or stop_criteria')
    steady_tol = None if stop_at_steady is None else (1.e-4 if isinstance(stop_at_steady, bool) else stop_at_steady)
    state = np.copy(initial_state)
    t = float(initial_time)
    dt = step_size.first_step_size()
    gamma = method.implicit_coefficient
    eval_setup = True
    setup_count = 0
    nsteps = 0
    stats = {'steps': 0, 'nonlinear_iter': 0, 'projector_setups': 0, 'failed_steps': 0}
    ts, xs = [t], [np.copy(state)]
    go = True
    while go:
        if stop_at_time is not None and t + dt > stop_at_time:
            dt = stop_at_time - t
        if method.nonlinear_solver.setup_projector_in_governor and eval_setup:
            linear_setup(t, state, dt * gamma)
            stats['projector_setups'] += 1
        setup_count += 1
        out = method.single_step(state, t, dt, right_hand_side, lambda tt, x, dt=dt: linear_setup(tt, x, dt * gamma),
                                 linear_solve)
        dstate = out.solution_update
        stats['projector_setups'] += out.projector_setups
        residual = np.linalg.norm(dstate * norm_weighting, ord=np.inf) / dt
        recent_dt = dt
        if np.all(np.isfinite(dstate)):
            state = state + dstate
            t += dt
            nsteps += 1
            stats['nonlinear_iter'] += out.nonlinear_iter
            if post_step_callback is not None:
                r = post_step_callback(t, state, residual, nsteps)
                state = state if r is None else np.copy(r)
            if save_each_step:
                ts.append(t)
                xs.append(np.copy(state))
            dt = step_size(nsteps, dt, out)
            # projector refresh rule (integrator.py:100-130)
            if setup_count == linear_setup_rate:
                setup_count = 0
                eval_setup = True
            elif out.slow_nonlinear_convergence:
                dt *= time_step_reduction_factor_on_slow_solve
                eval_setup = True
            elif not out.nonlinear_converged:
                dt *= time_step_reduction_factor_on_failure
                eval_setup = True
            elif dt > recent_dt * time_step_increase_factor_to_force_jacobian:
                eval_setup = True
            elif dt < recent_dt * time_step_decrease_factor_to_force_jacobian:
                eval_setup = True
            else:
                eval_setup = False
        else:
            dt *= time_step_reduction_factor_on_failure
            eval_setup = True
            stats['failed_steps'] += 1
        if stop_criteria is not None:
            go = not stop_criteria(t, state, residual, nsteps)
        if nsteps < minimum_time_step_count:
            go = True
        if nsteps > maximum_time_step_count:
            go = False
        elif stop_at_time is not None and t >= stop_at_time:
            go = False
        elif steady_tol is not None and residual < steady_tol:
            go = False
    stats['steps'] = nsteps
    result = (np.array(ts), np.array(xs)) if save_each_step else (state, t, dt)
    return (*result, stats) if return_info else result
