"""Host-side pieces of the nonadiabatic library builders on the CPU: the interpolation of the unstructured
(dissipation rate, stoichiometric defect) table onto the structured defect grid, against scipy's
InterpolatedUnivariateSpline(k=1, ext='const') and interp1d(fill_value='extrapolate'), the interpolants the
reference uses (tabulation.py:594-665)."""
import numpy as np
import pytest
from scipy.interpolate import InterpolatedUnivariateSpline, interp1d

from spitfire_amd.tabulation import _interpolate_to_structured_defect_dimension


@pytest.mark.parametrize('extend', [False, True])
def test_structured_defect_interpolation_matches_scipy(extend):
    rng = np.random.default_rng(3)
    nz = 7
    table = {}
    for c in (1., 3., 10.):
        gs = -np.sort(rng.uniform(0., 4.e5, int(rng.integers(3, 9))))
        gs[0] = 0.
        for g in gs:
            table[(c, g)] = {'temperature': 300. + 2000. * rng.random(nz), 'enthalpy': g + rng.random(nz),
                             'mass fraction X': rng.random(nz)}
    out, x, grid = _interpolate_to_structured_defect_dimension(table, 9, extend=extend)
    assert np.array_equal(x, [1., 3., 10.]) and grid[0] >= grid[-1]
    allg = np.array([g for _, g in table])
    n = 9 + (2 if extend else 0)
    assert grid.size == n and grid[0] == allg.max()
    for c in x:
        gsorted = np.sort([g for (cc, g) in table if cc == c])
        for q in ('temperature', 'enthalpy', 'mass fraction X'):
            for iz in range(nz):
                y = np.array([table[(c, g)][q][iz] for g in gsorted])
                if extend and q == 'enthalpy':
                    f = interp1d(gsorted, y, kind='linear', fill_value='extrapolate')
                else:
                    f = InterpolatedUnivariateSpline(gsorted, y, k=1, ext='const')
                for g in grid:
                    assert abs(out[(c, g)][q][iz] - float(f(g))) <= 1e-9 * max(1., abs(float(f(g))))
