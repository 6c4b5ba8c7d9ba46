"""Mechanism data -> griffon-ready index tables.

Turns the reference's extracted-mechanism schema (species dict + reaction list,
src/spitfire/chemistry/mechanism.py:387-541) into the index/coefficient arrays that griffon's
setters consume (chemistry_setup.cpp:36-345), resolving species names to indices on the host
exactly as griffon does:

* molecular weights are sum(element_mw * count) over the atom map in sorted element-name order
  (std::map iteration in mechanism_add_species, chemistry_setup.cpp:52-56);
* reactants, products, third-body efficiencies and special orders are listed in sorted
  species-name order (std::map<std::string,...> iteration, chemistry_setup.cpp:379-395) unless
  `order='index'` is requested;
* reactions are stable-sorted by type: simple, three-body, Lindemann, Troe (mechanism.py:540);
* activation energies are divided by the gas constant (mechanism.py:185).
"""
import numpy as np

RATE_TYPE = {'simple': 1, 'three-body': 2, 'Lindemann': 3, 'Troe': 4}
_TYPE_RANK = {'simple': 0, 'three-body': 1, 'Lindemann': 2, 'Troe': 3}


def species_molecular_weight(atoms, element_mw_map):
    mw = 0.0
    for e in sorted(atoms):
        mw += element_mw_map[e] * atoms[e]
    return mw


def sorted_reactions(reactions):
    """Stable sort by reaction type, returning (sorted list, original indices)."""
    order = sorted(range(len(reactions)), key=lambda i: _TYPE_RANK[reactions[i]['type']])
    return [reactions[i] for i in order], order


class CompiledMechanism:
    """Index-based mechanism tables in griffon's order (one object per ChemicalMechanismSpec)."""

    def __init__(self, mech, order='name'):
        self.names = list(mech['species_names'])
        self.ns = len(self.names)
        self.index = {n: i for i, n in enumerate(self.names)}
        self.gas_constant = float(mech['gas_constant'])
        self.ref_pressure = float(mech['ref_pressure'])
        self.ref_temperature = float(mech['ref_temperature'])
        emw = mech['element_mw_map']
        self.mw = np.array([species_molecular_weight(mech['species'][n]['atoms'], emw) for n in self.names])
        self.thermo = []
        for n in self.names:
            cp = mech['species'][n]['heat-capacity']
            if cp['type'] == 'NASA7':
                self.thermo.append(('NASA7', float(cp['Tmin']), float(cp['Tmid']), float(cp['Tmax']),
                                    np.asarray(cp['low-coeffs'], dtype=np.float64),
                                    np.asarray(cp['high-coeffs'], dtype=np.float64)))
            elif cp['type'] == 'constant':
                self.thermo.append(('constant', float(cp['Tmin']), float(cp['Tmax']), float(cp['T0']),
                                    float(cp['h0']), float(cp['s0']), float(cp['cp'])))
            else:
                raise ValueError('unsupported heat capacity type ' + str(cp['type']))
        rx_sorted, self.file_order = sorted_reactions(mech['reactions'])
        self.reactions = [self._compile(r, order) for r in rx_sorted]
        self.nr = len(self.reactions)

    def _keys(self, d, order):
        keys = list(d.keys())
        if order == 'name':
            keys.sort()
        else:
            keys.sort(key=lambda k: self.index[k])
        return keys

    def _compile(self, rx, order):
        R = self.gas_constant
        t = rx['type']
        out = {'type': RATE_TYPE[t], 'reversible': bool(rx['reversible'])}
        for side in ('reactants', 'products'):
            keys = self._keys(rx[side], order)
            st = [rx[side][k] for k in keys]
            if any(abs(s - round(s)) > 0 for s in st):
                raise ValueError('non-integer stoichiometric coefficients are not supported')
            out[side] = (np.array([self.index[k] for k in keys], dtype=np.int32),
                         np.array([int(round(s)) for s in st], dtype=np.int32))
        if t in ('simple', 'three-body'):
            out['kfwd'] = np.array([rx['A'], rx['b'], rx['Ea'] / R])
        else:
            out['kfwd'] = np.array([rx['fwd-A'], rx['fwd-b'], rx['fwd-Ea'] / R])
            out['kpres'] = np.array([rx['flf-A'], rx['flf-b'], rx['flf-Ea'] / R])
        if t != 'simple':
            eff = {k: v for k, v in rx.get('efficiencies', {}).items() if k in self.index}
            keys = self._keys(eff, order)
            out['tb'] = (np.array([self.index[k] for k in keys], dtype=np.int32),
                         np.array([float(eff[k]) for k in keys]))
            out['default_eff'] = float(rx.get('default-eff', 1.0))
        else:
            out['tb'] = (np.zeros(0, dtype=np.int32), np.zeros(0))
            out['default_eff'] = 0.0
        if t == 'Troe':
            tp = np.asarray(rx['Troe-params'], dtype=np.float64)
            out['troe'] = np.concatenate([tp, np.zeros(4 - tp.size)])
        if 'orders' in rx and rx['orders']:
            keys = self._keys(rx['orders'], order)
            out['orders'] = (np.array([self.index[k] for k in keys], dtype=np.int32),
                             np.array([float(rx['orders'][k]) for k in keys]))
        else:
            out['orders'] = (np.zeros(0, dtype=np.int32), np.zeros(0))
        out.setdefault('kpres', np.zeros(3))
        out.setdefault('troe', np.zeros(4))
        return out
