"""TEST INFRASTRUCTURE ONLY — a HomogeneousReactor whose RHS and Jacobian run on the CPU oracle
(oracle/griffon_oracle.cpp) instead of the GPU: the checker for reactor integrations. The integrator
(ESDIRK64 + Newton + PI control) is the product's spitfire_amd.reactors.HomogeneousReactor, unchanged."""
from oracle.oracle import CpuGriffon
from spitfire_amd.mechdata import CompiledMechanism
from spitfire_amd.reactors import HomogeneousReactor


class OracleReactor(HomogeneousReactor):
    def __init__(self, mech_spec, initial_mixture, cpu=None, **kwargs):
        super().__init__(mech_spec, initial_mixture, **kwargs)
        self._cpu = cpu if cpu is not None else CpuGriffon(CompiledMechanism(mech_spec.mech_data, order='name'))

    def _rhs(self, t, state):
        return self._cpu.reactor_rhs_isobaric(state, self._initial_pressure, heat=self._heat_transfer_option)

    def _jac(self, state):
        n = self._n_equations
        return self._cpu.reactor_jac_isobaric(state, self._initial_pressure,
                                              heat=self._heat_transfer_option)[1].reshape((n, n), order='F')
