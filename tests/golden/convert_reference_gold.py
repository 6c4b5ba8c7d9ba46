"""Write the reference's own gold vectors as small fixtures (run in a container that has the reference checkout).

Sources (numpy-only pickles under the reference's tests/, read with the data-only reader in safe_pickle.py, which
executes nothing from the file):
  tests/time_integration/chemistry_abc_implicit/gold.pkl  -> ref_abc_gold.npz
      odesolve trajectories (t, c) of the A -> B, A + B -> 2C problem for five implicit steppers and three
      linear solvers (the reference's rebless.py:115-139)
  tests/reactor/closed_reactors/gold.pkl                  -> ref_closed_reactors_gold.npz
      h2-burke reactors integrated to steady state: (t, T, Y) for isobaric/isochoric x adiabatic/isothermal
  tests/tabulation/adiabatic_slfm/gold.pkl                -> ref_slfm_gold.npz + ref_h2_burke_mech_data.json
      the adiabatic SLFM library (34-point grid x 8 dissipation rates) and the mechanism data the reference
      extracted from h2-burke.yaml (its Library pickles the ChemicalMechanismSpec state, mechanism.py:99-103)

    python tests/golden/convert_reference_gold.py [/root/reference]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from spitfire_amd.safe_pickle import Stub, load  # noqa: E402


def _plain(v):
    if isinstance(v, dict):
        return {str(k): _plain(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_plain(x) for x in v]
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    return v


def main(ref):
    t = os.path.join(ref, 'tests')
    abc = load(os.path.join(t, 'time_integration', 'chemistry_abc_implicit', 'gold.pkl'))
    out = {}
    for key, (tt, sol) in abc.items():
        out['t|' + key] = tt
        out['sol|' + key] = sol
    np.savez_compressed(os.path.join(HERE, 'ref_abc_gold.npz'), **out)

    cr = load(os.path.join(t, 'reactor', 'closed_reactors', 'gold.pkl'))
    out = {}
    for key, (tt, T, Y) in cr.items():
        out['t|' + key], out['T|' + key], out['Y|' + key] = tt, T, Y
    np.savez_compressed(os.path.join(HERE, 'ref_closed_reactors_gold.npz'), **out)

    lib = load(os.path.join(t, 'tabulation', 'adiabatic_slfm', 'gold.pkl'),
               stub_classes=[('spitfire.chemistry.library', 'Library'),
                             ('spitfire.chemistry.mechanism', 'ChemicalMechanismSpec')])
    st = lib.state
    out = {'dim|' + d['name']: d['values'] for d in st['dimensions'].values()}
    for k, v in st['properties'].items():
        out['prop|' + k] = v
    np.savez_compressed(os.path.join(HERE, 'ref_slfm_gold.npz'), **out)
    mech = st['extra_attributes']['mech_spec']
    assert isinstance(mech, Stub)
    with open(os.path.join(HERE, 'ref_h2_burke_mech_data.json'), 'w') as f:
        json.dump(_plain(mech.state['mech_data']), f, indent=0)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
