"""Mechanism ingestion without Cantera: Cantera-format YAML files and the reference's pickled mechanism data,
both turned into the mechanism-data schema `ChemicalMechanismSpec` consumes.

The schema is the one the reference's `_extract_cantera_mechanism_data` returns (mechanism.py:387-541): a species
dict {name: {'atoms', 'heat-capacity'}} in phase order and a reaction list of dicts with 'type' in simple /
three-body / Lindemann / Troe, SI-kmol units (A in m, kmol, s; Ea in J/kmol), plus the element list, the element
molecular-weight map, the gas constant and the reference pressure and temperature.

YAML semantics follow Cantera's documented input format:
  - units: the file's `units:` block (length, quantity, time, activation-energy, energy, pressure), and unit strings
    on individual values ('1.5286e+04 cal/mol'); defaults m, kmol, s, J/kmol, as Cantera's;
  - A is converted with the reaction order: a rate constant of order n (sum of reactant coefficients, +1 for a
    three-body reaction and for the low-pressure limit of a falloff reaction) has units (length^3/quantity)^(n-1)/time;
  - elementary, three-body (`+ M`) and falloff reactions (`(+ M)` or `(+ species)`, Lindemann or Troe), reversible
    (`<=>`, `=`) or irreversible (`=>`), `duplicate`, `efficiencies`, `default-efficiency`, `orders`;
  - NASA7 (one or two temperature ranges) and constant-cp thermo.
Pressure-dependent Arrhenius (PLOG), Chebyshev, chemically activated, SRI and NASA9 data are rejected with an error:
the griffon kernels this package replaces support none of them.
"""
import re

import numpy as np

GAS_CONSTANT = 8314.46261815324  # J/(kmol K)
ONE_ATM = 101325.
REF_TEMPERATURE = 298.15

# conventional atomic weights (kg/kmol) of Cantera's element table
ATOMIC_WEIGHTS = {
    'H': 1.008, 'D': 2.01410178, 'T': 3.0160492, 'He': 4.002602, 'Li': 6.94, 'Be': 9.0121831, 'B': 10.81,
    'C': 12.011, 'N': 14.007, 'O': 15.999, 'F': 18.998403163, 'Ne': 20.1797, 'Na': 22.98976928, 'Mg': 24.305,
    'Al': 26.9815385, 'Si': 28.085, 'P': 30.973761998, 'S': 32.06, 'Cl': 35.45, 'Ar': 39.95, 'K': 39.0983,
    'Ca': 40.078, 'Ti': 47.867, 'Cr': 51.9961, 'Mn': 54.938044, 'Fe': 55.845, 'Ni': 58.6934, 'Cu': 63.546,
    'Zn': 65.38, 'Br': 79.904, 'Kr': 83.798, 'I': 126.90447, 'Xe': 131.293, 'U': 238.02891, 'E': 0.000548579909,
}

_LENGTH = {'m': 1., 'cm': 1.e-2, 'mm': 1.e-3, 'km': 1.e3}
_QUANTITY = {'kmol': 1., 'mol': 1.e-3, 'molec': 1. / 6.02214076e26, 'molecule': 1. / 6.02214076e26}
_TIME = {'s': 1., 'ms': 1.e-3, 'us': 1.e-6, 'min': 60., 'h': 3600.}
_ENERGY = {'J': 1., 'kJ': 1.e3, 'cal': 4.184, 'kcal': 4184., 'erg': 1.e-7, 'eV': 1.602176634e-19}


def _energy_per_quantity(unit):
    """factor from `unit` (e.g. 'cal/mol', 'kJ/mol', 'J/kmol', 'K', 'eV') to J/kmol"""
    unit = unit.strip()
    if unit == 'K':
        return GAS_CONSTANT
    if unit == 'eV':
        return _ENERGY['eV'] * 6.02214076e26
    if '/' not in unit:
        raise ValueError('unsupported activation-energy unit ' + unit)
    e, q = unit.split('/', 1)
    return _ENERGY[e.strip()] / _QUANTITY[q.strip()]


class _Units:
    def __init__(self, block=None):
        b = dict(block or {})
        self.length = _LENGTH[b.get('length', 'm')]
        self.quantity = _QUANTITY[b.get('quantity', 'kmol')]
        self.time = _TIME[b.get('time', 's')]
        if 'activation-energy' in b:
            self.ea = _energy_per_quantity(b['activation-energy'])
        else:
            self.ea = _ENERGY[b.get('energy', 'J')] / self.quantity

    def rate_factor(self, order):
        """SI-kmol value of one unit of an order-`order` rate constant"""
        return (self.length ** 3 / self.quantity) ** (order - 1) / self.time

    def ea_value(self, v):
        if isinstance(v, str):
            num, unit = v.strip().split(None, 1)
            return float(num) * _energy_per_quantity(unit)
        return float(v) * self.ea


def _side(text, species):
    """'2 OH + H2' -> {'OH': 2.0, 'H2': 1.0}; 'M' terms are dropped (returned count)"""
    out, m = {}, 0
    for tok in [t.strip() for t in re.split(r'\s\+\s', ' ' + text.strip() + ' ') if t.strip()]:
        mm = re.match(r'^([0-9]*\.?[0-9]+)\s+(\S+)$', tok) or re.match(r'^([0-9]*\.?[0-9]+)(\S+)$', tok)
        coef, name = (float(mm.group(1)), mm.group(2)) if mm and mm.group(2) in species else (1., tok)
        if name == 'M':
            m += 1
            continue
        if name not in species:
            raise ValueError('unknown species %r in reaction side %r' % (name, text))
        out[name] = out.get(name, 0.) + coef
    return out, m


def _parse_equation(eq, species):
    """-> reactants, products, reversible, third-body marker: None, 'M' (three-body), or the falloff collider
    ('M' or a species name) with a flag"""
    if '<=>' in eq:
        lhs, rhs = eq.split('<=>')
        rev = True
    elif '=>' in eq:
        lhs, rhs = eq.split('=>')
        rev = False
    elif '=' in eq:
        lhs, rhs = eq.split('=')
        rev = True
    else:
        raise ValueError('no reaction arrow in ' + eq)
    falloff = None
    pat = re.compile(r'\(\s*\+\s*([^()\s]+)\s*\)')
    fl, fr = pat.findall(lhs), pat.findall(rhs)
    if fl or fr:
        if fl != fr or len(fl) != 1:
            raise ValueError('unbalanced falloff collider in ' + eq)
        falloff = fl[0]
        lhs, rhs = pat.sub(' ', lhs), pat.sub(' ', rhs)
    reac, ml = _side(lhs, species)
    prod, mr = _side(rhs, species)
    if ml != mr or ml > 1:
        raise ValueError('unbalanced third body in ' + eq)
    return reac, prod, rev, falloff, ml == 1


def _rate(d, units, order):
    if not isinstance(d, dict):
        raise ValueError('rate constants must be given as {A, b, Ea} maps')
    A = d['A']
    if isinstance(A, str):
        raise ValueError('rate constants with unit strings on A are not supported')
    return float(A) * units.rate_factor(order), float(d.get('b', 0.)), units.ea_value(d.get('Ea', 0.))


def _thermo(th, name):
    model = th.get('model')
    if model == 'NASA7':
        tr = [float(t) for t in th['temperature-ranges']]
        data = [np.asarray(c, dtype=np.float64) for c in th['data']]
        if len(tr) == 2:
            return {'type': 'NASA7', 'Tmin': tr[0], 'Tmid': tr[1], 'Tmax': tr[1], 'low-coeffs': data[0],
                    'high-coeffs': data[0].copy()}
        if len(tr) != 3 or len(data) != 2:
            raise ValueError('species %s: NASA7 needs two or three temperature bounds' % name)
        return {'type': 'NASA7', 'Tmin': tr[0], 'Tmid': tr[1], 'Tmax': tr[2], 'low-coeffs': data[0],
                'high-coeffs': data[1]}
    if model == 'constant-cp':
        def q(key, default, factor):
            v = th.get(key, default)
            if isinstance(v, str):
                num, unit = v.split(None, 1)
                return float(num) * (_energy_per_quantity(unit) if '/' in unit and 'K' not in unit else 1.)
            return float(v) * factor
        return {'type': 'constant', 'Tmin': float(th.get('T-min', 0.)), 'Tmax': float(th.get('T-max', 1.e30)),
                'T0': q('T0', REF_TEMPERATURE, 1.), 'h0': q('h0', 0., 1.), 's0': q('s0', 0., 1.),
                'cp': q('cp0', 0., 1.)}
    raise ValueError('species %s: thermo model %r is not supported' % (name, model))


def load_cantera_yaml(path, group_name=None):
    """Mechanism data from a Cantera YAML file, for the phase `group_name` (default: the first phase)."""
    import yaml
    with open(path) as f:
        doc = yaml.safe_load(f)
    phases = doc.get('phases', [])
    if not phases:
        raise ValueError('no phases in ' + path)
    if group_name in (None, 'gas') and not any(p.get('name') == group_name for p in phases):
        phase = phases[0]
    else:
        matches = [p for p in phases if p.get('name') == group_name]
        if not matches:
            raise ValueError('phase %r not found in %s' % (group_name, path))
        phase = matches[0]
    units = _Units(doc.get('units'))
    user_elements = {e['symbol']: float(e['atomic-weight']) for e in doc.get('elements', []) or []}
    elements = list(phase['elements'])
    emw = {}
    for e in elements:
        if e in user_elements:
            emw[e] = user_elements[e]
        elif e in ATOMIC_WEIGHTS:
            emw[e] = ATOMIC_WEIGHTS[e]
        else:
            raise ValueError('no atomic weight for element ' + e)

    all_species = {s['name']: s for s in doc.get('species', [])}
    sp_decl = phase.get('species', 'all')
    if sp_decl == 'all':
        names = [s['name'] for s in doc.get('species', [])]
    else:
        names = []
        for item in sp_decl:
            if isinstance(item, str):
                names.append(item)
            elif isinstance(item, dict):
                for section, lst in item.items():
                    sec = {s['name']: s for s in doc.get(section, [])}
                    all_species.update(sec)
                    names += list(sec) if lst == 'all' else list(lst)
    species = {}
    for n in names:
        s = all_species[n]
        species[n] = {'atoms': {k: float(v) for k, v in (s.get('composition') or {}).items()},
                      'heat-capacity': _thermo(s['thermo'], n)}
    known = set(names)

    rx_decl = phase.get('reactions', 'all' if 'reactions' in doc else 'none')
    if rx_decl == 'all':
        rx_src = list(doc.get('reactions', []))
    elif rx_decl in ('none', None):
        rx_src = []
    else:
        rx_src = []
        for section in rx_decl:
            rx_src += list(doc.get(section if isinstance(section, str) else list(section)[0], []))
    reactions = []
    for r in rx_src:
        reactions.append(_reaction(r, units, known))
    return {'species_names': names, 'species': species, 'reactions': reactions, 'elements': elements,
            'element_mw_map': emw, 'gas_constant': GAS_CONSTANT, 'ref_pressure': ONE_ATM,
            'ref_temperature': REF_TEMPERATURE}


def _troe_params(t):
    """[A, T3, T1, T2] as Cantera reports them: T3 and T1 pass through their reciprocals (Cantera stores 1/T3 and
    1/T1), T2 is 0 when absent"""
    a = float(t['A'])
    t3, t1 = float(t['T3']), float(t['T1'])
    t2 = float(t.get('T2', 0.))
    return np.array([a, 1. / (1. / t3) if t3 != 0. else 0., 1. / (1. / t1) if t1 != 0. else 0., t2])


def _reaction(r, units, species):
    eq = r['equation']
    reac, prod, rev, falloff, three_body = _parse_equation(eq, species)
    rtype = r.get('type', 'elementary')
    order = sum(reac.values())
    out = {'reversible': rev, 'reactants': reac, 'products': prod}
    if r.get('orders'):
        out['orders'] = {k: float(v) for k, v in r['orders'].items()}
        order = sum(out['orders'].get(k, v) for k, v in reac.items()) + \
            sum(v for k, v in out['orders'].items() if k not in reac)
    if rtype in ('elementary', 'reaction') and not three_body and falloff is None:
        A, b, Ea = _rate(r['rate-constant'], units, order)
        out.update({'type': 'simple', 'A': A, 'b': b, 'Ea': Ea})
    elif rtype == 'three-body' or (rtype in ('elementary', 'reaction') and three_body):
        A, b, Ea = _rate(r['rate-constant'], units, order + 1)
        out.update({'type': 'three-body', 'A': A, 'b': b, 'Ea': Ea,
                    'efficiencies': {k: float(v) for k, v in (r.get('efficiencies') or {}).items()},
                    'default-eff': float(r.get('default-efficiency', 1.))})
    elif rtype == 'falloff':
        if 'SRI' in r or 'Tsang' in r:
            raise ValueError('falloff form of %r is not supported' % eq)
        hi = _rate(r['high-P-rate-constant'], units, order)
        lo = _rate(r['low-P-rate-constant'], units, order + 1)
        if falloff not in (None, 'M'):
            eff, deff = {falloff: 1.}, 0.
        else:
            eff = {k: float(v) for k, v in (r.get('efficiencies') or {}).items()}
            deff = float(r.get('default-efficiency', 1.))
        out.update({'fwd-A': hi[0], 'fwd-b': hi[1], 'fwd-Ea': hi[2], 'flf-A': lo[0], 'flf-b': lo[1],
                    'flf-Ea': lo[2], 'efficiencies': eff, 'default-eff': deff})
        if 'Troe' in r:
            out.update({'type': 'Troe', 'Troe-params': _troe_params(r['Troe'])})
        else:
            out['type'] = 'Lindemann'
    else:
        raise ValueError('reaction type %r of %r is not supported' % (rtype, eq))
    return out


def from_reference_mech_data(md):
    """Mechanism data in the schema above from the reference's stored form (`ChemicalMechanismSpec.mech_data`,
    mechanism.py:122-258: species {name: {'atom_map', 'cp': (kind, ...)}} and type-sorted reaction tuples)."""
    names = list(md['species'].keys())
    species = {}
    for n in names:
        s = md['species'][n]
        cp = s['cp']
        if cp[0] == 'NASA7':
            hc = {'type': 'NASA7', 'Tmin': float(cp[1]), 'Tmid': float(cp[2]), 'Tmax': float(cp[3]),
                  'low-coeffs': np.asarray(cp[4], dtype=np.float64), 'high-coeffs': np.asarray(cp[5], dtype=np.float64)}
        elif cp[0] == 'constant':
            hc = {'type': 'constant', 'Tmin': float(cp[1]), 'Tmax': float(cp[2]), 'T0': float(cp[3]),
                  'h0': float(cp[4]), 's0': float(cp[5]), 'cp': float(cp[6])}
        else:
            raise ValueError('species %s: heat capacity %r is not supported' % (n, cp[0]))
        species[n] = {'atoms': {k: float(v) for k, v in s['atom_map'].items()}, 'heat-capacity': hc}
    reactions = []
    for rx in md['reactions']:
        kind = rx[0]
        base = kind.replace('-special', '')
        out = {'type': base, 'reactants': dict(rx[1]), 'products': dict(rx[2]), 'reversible': bool(rx[3])}
        if base in ('simple', 'three-body'):
            out.update({'A': float(rx[4]), 'b': float(rx[5]), 'Ea': float(rx[6])})
            if base == 'three-body':
                out.update({'efficiencies': dict(rx[7]), 'default-eff': float(rx[8])})
        elif base in ('Lindemann', 'Troe'):
            out.update({'fwd-A': float(rx[4]), 'fwd-b': float(rx[5]), 'fwd-Ea': float(rx[6]),
                        'efficiencies': dict(rx[7]), 'default-eff': float(rx[8]),
                        'flf-A': float(rx[9]), 'flf-b': float(rx[10]), 'flf-Ea': float(rx[11])})
            if base == 'Troe':
                out['Troe-params'] = np.asarray(rx[12], dtype=np.float64)
        else:
            raise ValueError('reaction kind %r is not supported' % kind)
        if kind.endswith('-special'):
            out['orders'] = dict(rx[-1])
        reactions.append(out)
    return {'species_names': names, 'species': species, 'reactions': reactions, 'elements': list(md['elements']),
            'element_mw_map': {k: float(v) for k, v in md['element_mw_map'].items()},
            'gas_constant': float(md.get('gas_constant', GAS_CONSTANT)), 'ref_pressure': float(md['ref_pressure']),
            'ref_temperature': float(md['ref_temperature'])}


def to_reference_mech_data(data):
    """The reference's stored form of mechanism data in the schema above (reactions type-sorted, stable)."""
    from .mechdata import sorted_reactions
    species = {}
    for n in data['species_names']:
        hc = data['species'][n]['heat-capacity']
        if hc['type'] == 'NASA7':
            cp = ('NASA7', hc['Tmin'], hc['Tmid'], hc['Tmax'], list(map(float, hc['low-coeffs'])),
                  list(map(float, hc['high-coeffs'])))
        else:
            cp = ('constant', hc['Tmin'], hc['Tmax'], hc['T0'], hc['h0'], hc['s0'], hc['cp'])
        species[n] = {'atom_map': dict(data['species'][n]['atoms']), 'cp': cp}
    rxs = []
    for rx in sorted_reactions(data['reactions'])[0]:
        t = rx['type']
        head = (t + ('-special' if rx.get('orders') else ''), rx['reactants'], rx['products'], rx['reversible'])
        if t == 'simple':
            tail = (rx['A'], rx['b'], rx['Ea'])
        elif t == 'three-body':
            tail = (rx['A'], rx['b'], rx['Ea'], rx['efficiencies'], rx['default-eff'])
        else:
            tail = (rx['fwd-A'], rx['fwd-b'], rx['fwd-Ea'], rx['efficiencies'], rx['default-eff'], rx['flf-A'],
                    rx['flf-b'], rx['flf-Ea'])
            if t == 'Troe':
                tail = tail + (list(map(float, rx['Troe-params'])),)
        if rx.get('orders'):
            tail = tail + (rx['orders'],)
        rxs.append(head + tail)
    return {'ref_pressure': data['ref_pressure'], 'ref_temperature': data['ref_temperature'],
            'elements': list(data['elements']), 'species': species, 'reactions': rxs,
            'gas_constant': data['gas_constant'], 'element_mw_map': dict(data['element_mw_map'])}
