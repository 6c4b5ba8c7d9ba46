"""The reference's own gold trajectories through the GPU drop-ins.

The h2-burke closed reactors of tests/reactor/closed_reactors (phi = 1 H2/air at 1200 K, 1 atm; isobaric and
isochoric, adiabatic and isothermal; integrated to steady state) run through HomogeneousReactor on the gfx950
kernels, with the mechanism read from the reference's YAML file by this package's loader. Tolerances are the
reference test's: t, T and Y at rtol 1e-4, atol 1e-8 (test.py:23-28).
"""
import pytest

from tests.test_reference_gold import CLOSED_REACTORS, _closed_reactor_setup, check_closed_reactor_gold

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('key', list(CLOSED_REACTORS))
def test_closed_reactor_gold_on_gpu(key):
    from spitfire_amd.reactors import HomogeneousReactor
    print('closed reactor gold on the GPU drop-ins:', key, flush=True)
    m, mix = _closed_reactor_setup()
    conf, heat = CLOSED_REACTORS[key]
    r = HomogeneousReactor(m, mix, configuration=conf, heat_transfer=heat, mass_transfer='closed')
    lib = r.integrate_to_steady()
    check_closed_reactor_gold(key, lib, m)
