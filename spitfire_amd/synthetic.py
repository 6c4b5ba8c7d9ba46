"""Seeded synthetic H/O/N chemical mechanisms in the reference's mechanism-data schema.

The reference builds griffon's tables from Cantera objects
(src/spitfire/chemistry/mechanism.py:387-541 `_extract_cantera_mechanism_data`); Cantera is not
available here, so workloads are generated directly in the same schema: a species dictionary
(atoms + NASA7 heat-capacity data) and a reaction list with keys 'type', 'reactants', 'products',
'reversible', 'A'/'b'/'Ea' (or 'fwd-*'/'flf-*'), 'efficiencies', 'default-eff', 'Troe-params'.

Units follow Cantera's SI-kmol convention that the reference passes straight into griffon:
pre-exponentials in (m^3/kmol)^(order-1)/s, activation energies in J/kmol (divided by the gas
constant in ChemicalMechanismSpec, mechanism.py:185).

Rate parameters follow the workload definitions: A log-uniform in [1e6, 1e14] in cgs units
(converted with A_SI = A_cgs * 1e-3^(order-1)), b uniform in [-1.5, 1.5], Ea uniform in
[0, 200] kJ/mol; reverse rates come from equilibrium constants. Heat capacities are smooth,
monotone seeded cp(T) curves, least-squares fitted by NASA7 polynomials on [300, 1000] K and
[1000, 3500] K with cp, h and s continuous at 1000 K. Formation enthalpies of the named species are
approximate tabulated values; the generic species follow a seeded bond-class model (atomization enthalpies
minus H-X and heavy-heavy bond energies in textbook ranges), so H2/O2 mixtures release realistic heat.
Every reaction conserves elements, so mass production rates sum to zero, and is written in its exothermic
direction (recombinations as A + B (+M) -> C), so reverse rates from equilibrium constants stay bounded.
"""
import itertools
import numpy as np

GAS_CONSTANT = 8314.46261815324  # J/(kmol K), Cantera's ct.gas_constant
ONE_ATM = 101325.0
ELEMENT_MW = {'H': 1.008, 'O': 15.999, 'N': 14.007}

# named H/O/N species (composition h, o, n); the remaining slots are filled with generic names
_NAMED = [('H2', (2, 0, 0)), ('O2', (0, 2, 0)), ('H2O', (2, 1, 0)), ('H', (1, 0, 0)), ('O', (0, 1, 0)),
          ('OH', (1, 1, 0)), ('HO2', (1, 2, 0)), ('H2O2', (2, 2, 0)), ('NO', (0, 1, 1)), ('N', (0, 0, 1)),
          ('NH', (1, 0, 1)), ('NH2', (2, 0, 1)), ('NH3', (3, 0, 1)), ('NO2', (0, 2, 1)), ('N2O', (0, 1, 2)),
          ('HNO', (1, 1, 1)), ('NNH', (1, 0, 2)), ('HONO', (1, 2, 1)), ('N2H2', (2, 0, 2)), ('N2H3', (3, 0, 2)),
          ('N2H4', (4, 0, 2)), ('H2NO', (2, 1, 1)), ('HNO2', (1, 2, 1)), ('O3', (0, 3, 0)), ('NO3', (0, 3, 1)),
          ('HNO3', (1, 3, 1)), ('H2N2O', (2, 1, 2)), ('NH2OH', (3, 1, 1))]
_ATOMIZATION = {'H': 218.0, 'O': 249.2, 'N': 472.7}  # kJ/mol, elements -> gaseous atoms
# approximate tabulated standard enthalpies of formation at 298.15 K (kJ/mol) of the named species, so that
# H2/air burns with a realistic heat release
_H298_NAMED = {'H2O': -241.8, 'H': 218.0, 'O': 249.2, 'OH': 37.3, 'HO2': 12.0, 'H2O2': -136.1, 'NO': 91.3,
               'N': 472.7, 'NH': 358.0, 'NH2': 189.0, 'NH3': -45.9, 'NO2': 33.1, 'N2O': 82.1, 'HNO': 106.8,
               'NNH': 249.5, 'HONO': -78.6, 'N2H2': 212.9, 'N2H3': 222.0, 'N2H4': 95.4, 'H2NO': 66.0,
               'HNO2': -44.0, 'O3': 142.7, 'NO3': 73.7, 'HNO3': -134.3, 'H2N2O': 80.0, 'NH2OH': -40.0}
_STANDARD = {(2, 0, 0), (0, 2, 0), (0, 0, 2)}

# A hydrogen-oxidation core with approximate literature rate parameters (A in cgs mol-cm-s units, b,
# Ea in kJ/mol) that every mechanism starts from, so that H2/air flames ignite and burn; the remaining
# reactions are random. Three-body efficiencies: H2 2.5, H2O 12.
_CORE_SIMPLE = [(('H', 'O2'), ('O', 'OH'), 1.04e14, 0.0, 63.96),
                (('O', 'H2'), ('H', 'OH'), 3.82e12, 0.0, 33.26),
                (('H2', 'OH'), ('H2O', 'H'), 2.16e8, 1.51, 14.35),
                (('OH', 'OH'), ('O', 'H2O'), 3.34e4, 2.42, -8.07),
                (('HO2', 'H'), ('H2', 'O2'), 2.75e6, 2.09, -6.0),
                (('HO2', 'H'), ('OH', 'OH'), 7.08e13, 0.0, 1.23),
                (('HO2', 'OH'), ('H2O', 'O2'), 2.89e13, 0.0, -2.09),
                (('HO2', 'HO2'), ('H2O2', 'O2'), 4.2e14, 0.0, 50.1),
                (('H2O2', 'H'), ('H2O', 'OH'), 2.41e13, 0.0, 16.6)]
_CORE_THREE_BODY = [(('H', 'H'), ('H2',), 1.0e18, -1.0, 0.0),
                    (('O', 'O'), ('O2',), 1.2e17, -1.0, 0.0),
                    (('H', 'OH'), ('H2O',), 2.2e22, -2.0, 0.0),
                    (('O', 'H'), ('OH',), 4.71e18, -1.0, 0.0)]
_CORE_TROE = [(('H', 'O2'), ('HO2',), (4.65e12, 0.44, 0.0), (6.37e20, -1.72, 2.2), (0.5, 1.0e-30, 1.0e30, 0.0)),
              (('OH', 'OH'), ('H2O2',), (7.4e13, -0.37, 0.0), (2.3e18, -0.9, -7.1), (0.7346, 94.0, 1756.0, 5182.0))]


def _species_set(n_species, rng):
    """Pick n_species-1 distinct compositions (named ones first) plus inert N2 last."""
    comps, names, seen = [], [], set()
    for name, c in _NAMED:
        if len(comps) == n_species - 1:
            break
        if c in seen:
            continue
        comps.append(c)
        names.append(name)
        seen.add(c)
    extra = [c for c in itertools.product(range(6), range(4), range(3))
             if 1 <= sum(c) <= 7 and c not in seen and c != (0, 0, 2)]
    rng.shuffle(extra)
    for c in extra:
        if len(comps) == n_species - 1:
            break
        comps.append(c)
        names.append('X%dH%dO%dN' % c)
        seen.add(c)
    if len(comps) != n_species - 1:
        raise ValueError('cannot build %d H/O/N species' % n_species)
    comps.append((0, 0, 2))
    names.append('N2')
    return names, comps


def _nasa7_fit(rng, natoms, h298, s298):
    """Fit NASA7 coefficients to a seeded monotone cp(T)/R curve; returns (low7, high7)."""
    if natoms == 1:
        cp0, cpinf = 2.5, 2.5 + rng.uniform(0.0, 0.3)
    else:
        cp0 = 3.5 + 0.5 * (natoms - 2) + rng.uniform(0.0, 0.5)
        cpinf = cp0 + 0.8 * max(1.0, 3 * natoms - 5 + rng.uniform(0.0, 1.0))
    theta = rng.uniform(400.0, 1500.0)

    def cp_r(T):
        return cp0 + (cpinf - cp0) * (1.0 - np.exp(-(T - 200.0) / theta))

    Tl = np.linspace(300.0, 1000.0, 64)
    Th = np.linspace(1000.0, 3500.0, 128)
    Vl = np.vander(Tl, 5, increasing=True)
    low5 = np.linalg.lstsq(Vl, cp_r(Tl), rcond=None)[0]
    cpm = float(np.polyval(low5[::-1], 1000.0))
    # high range passes through the low fit at 1000 K: cp = cpm + sum c_k (T^k - 1000^k)
    Vh = np.stack([Th ** k - 1000.0 ** k for k in range(1, 5)], axis=1)
    ck = np.linalg.lstsq(Vh, cp_r(Th) - cpm, rcond=None)[0]
    high5 = np.concatenate([[cpm - sum(ck[k - 1] * 1000.0 ** k for k in range(1, 5))], ck])

    def h_int(a, T):  # integral of cp/R dT (no constant)
        return a[0] * T + a[1] * T ** 2 / 2 + a[2] * T ** 3 / 3 + a[3] * T ** 4 / 4 + a[4] * T ** 5 / 5

    def s_int(a, T):  # integral of cp/(R T) dT (no constant)
        return a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4

    Tr = 298.15
    a6l = h298 * 1.0e6 / GAS_CONSTANT - h_int(low5, Tr)  # h298 in kJ/mol = 1e6 J/kmol
    a7l = s298 - s_int(low5, Tr)
    a6h = a6l + h_int(low5, 1000.0) - h_int(high5, 1000.0)
    a7h = a7l + s_int(low5, 1000.0) - s_int(high5, 1000.0)
    return np.concatenate([low5, [a6l, a7l]]), np.concatenate([high5, [a6h, a7h]])


def _ordered(d):
    return {k: d[k] for k in sorted(d)}


def make_mechanism(n_species=53, n_reactions=325, n_three_body=None, n_troe=None, seed=0):
    """Return a seeded synthetic mechanism in the reference's extracted-data schema.

    Reactions are returned in 'file order' (types interleaved); ChemicalMechanismSpec stable-sorts
    them by type exactly as the reference does (mechanism.py:540).
    """
    rng = np.random.default_rng(seed)
    if n_three_body is None:
        n_three_body = int(round(0.10 * n_reactions))
    if n_troe is None:
        n_troe = int(round(0.08 * n_reactions))
    n_simple = n_reactions - n_three_body - n_troe
    names, comps = _species_set(n_species, rng)
    ns = n_species
    mw = [sum(c[j] * ELEMENT_MW[e] for j, e in enumerate('HON')) for c in comps]

    species = {}
    h298_of = []
    for i, (nm, c) in enumerate(zip(names, comps)):
        natoms = sum(c)
        if c in _STANDARD:
            h298 = 0.0
        elif nm in _H298_NAMED:
            h298 = _H298_NAMED[nm]
        else:
            # bond-class model: H atoms bonded to heavy atoms (360-460 kJ/mol per bond, H-H 436 when there is no
            # heavy atom), heavy-heavy single/multiple bonds 150-350 kJ/mol (textbook ranges)
            atomization = sum(c[j] * _ATOMIZATION[e] for j, e in enumerate('HON'))
            nh, nheavy = c[0], c[1] + c[2]
            if nheavy == 0:
                bonds = 436.0 * (nh // 2)
            else:
                bonds = sum(rng.uniform(360.0, 460.0) for _ in range(nh))
                bonds += sum(rng.uniform(150.0, 350.0) for _ in range(nheavy - 1))
            h298 = atomization - bonds
        h298_of.append(h298)
        s298 = 14.0 + 4.0 * (natoms - 1) + rng.uniform(-1.0, 3.0)
        low, high = _nasa7_fit(rng, natoms, h298, s298)
        species[nm] = {'atoms': _ordered({e: float(c[j]) for j, e in enumerate('HON') if c[j] > 0}),
                       'heat-capacity': {'type': 'NASA7', 'Tmin': 300.0, 'Tmid': 1000.0, 'Tmax': 3500.0,
                                         'low-coeffs': low, 'high-coeffs': high}}

    active = list(range(ns - 1))  # the last species (N2) is inert: never a reactant or product
    by_comp = {comps[i]: i for i in active}

    def add(a, b):
        return tuple(x + y for x, y in zip(comps[a], comps[b]))

    def cgs_A(order):
        return 10.0 ** rng.uniform(6.0, 14.0) * (1.0e-3) ** (order - 1)

    def arr(order):
        return cgs_A(order), rng.uniform(-1.5, 1.5), rng.uniform(0.0, 200.0) * 1.0e6  # Ea J/kmol

    def exothermic(reac, prod):
        """Write the reaction in its exothermic direction, so that kr = kf/Kc stays bounded by the
        forward Arrhenius form (Ea - dG >= 0)."""
        dH = sum(h298_of[i] for i in prod) - sum(h298_of[i] for i in reac)
        return (reac, prod) if dH <= 0. else (prod, reac)

    def stoich_map(idxs):
        m = {}
        for i in idxs:
            m[names[i]] = m.get(names[i], 0) + 1
        return _ordered(m)

    recomb_pairs = [(a, b) for a in active for b in active if a <= b and add(a, b) in by_comp]
    # falloff and three-body reactions are written as recombinations A + B (+M) -> C, so only exothermic ones
    exo_pairs = [(a, b) for a, b in recomb_pairs if h298_of[by_comp[add(a, b)]] <= h298_of[a] + h298_of[b]]
    rxns = []
    index = {nm: i for i, nm in enumerate(names)}
    core_eff = _ordered({'H2': 2.5, 'H2O': 12.0})

    def core_arr(A, bexp, Ea, order):
        return A * (1.0e-3) ** (order - 1), bexp, Ea * 1.0e6

    def cmap(nms):
        return stoich_map([index[n] for n in nms])

    core_ok = all(n in index for n in ('H2', 'O2', 'H2O', 'H', 'O', 'OH', 'HO2', 'H2O2'))
    if core_ok:
        for reac, prod, A, bexp, Ea in _CORE_SIMPLE[:n_simple]:
            A, bexp, Ea = core_arr(A, bexp, Ea, 2)
            rxns.append({'type': 'simple', 'reversible': True, 'reactants': cmap(reac), 'products': cmap(prod),
                         'A': A, 'b': bexp, 'Ea': Ea})
        for reac, prod, A, bexp, Ea in _CORE_THREE_BODY[:n_three_body]:
            A, bexp, Ea = core_arr(A, bexp, Ea, 3)
            rxns.append({'type': 'three-body', 'reversible': True, 'reactants': cmap(reac), 'products': cmap(prod),
                         'A': A, 'b': bexp, 'Ea': Ea, 'efficiencies': core_eff, 'default-eff': 1.0})
        for reac, prod, ki, k0, troe in _CORE_TROE[:n_troe]:
            Ai, bi, Ei = core_arr(*ki, 2)
            A0, b0, E0 = core_arr(*k0, 3)
            rxns.append({'type': 'Troe', 'reversible': True, 'reactants': cmap(reac), 'products': cmap(prod),
                         'fwd-A': Ai, 'fwd-b': bi, 'fwd-Ea': Ei, 'flf-A': A0, 'flf-b': b0, 'flf-Ea': E0,
                         'efficiencies': core_eff, 'default-eff': 1.0, 'Troe-params': np.array(troe)})
    n_core_tb = sum(1 for r in rxns if r['type'] == 'three-body')
    n_core_troe = sum(1 for r in rxns if r['type'] == 'Troe')
    # simple reversible elementary reactions
    tries = 0
    while sum(1 for r in rxns if r['type'] == 'simple') < n_simple:
        tries += 1
        if tries > 200000:
            raise RuntimeError('synthetic mechanism generation failed')
        pattern = rng.uniform()
        if pattern < 0.7:  # A + B <=> C + D
            a, b = rng.choice(active, 2)
            s = add(a, b)
            splits = [(c, d) for c in active for d in active if c <= d
                      and tuple(x - y for x, y in zip(s, comps[c])) == comps[d]
                      and sorted((c, d)) != sorted((a, b))]
            if not splits:
                continue
            c, d = splits[rng.integers(len(splits))]
            reac, prod = [a, b], [c, d]
        elif pattern < 0.85:  # A + B <=> C
            if not recomb_pairs:
                continue
            a, b = recomb_pairs[rng.integers(len(recomb_pairs))]
            reac, prod = [a, b], [by_comp[add(a, b)]]
        else:  # A <=> B + C
            if not recomb_pairs:
                continue
            b, c = recomb_pairs[rng.integers(len(recomb_pairs))]
            reac, prod = [by_comp[add(b, c)]], [b, c]
        reac, prod = exothermic(reac, prod)
        A, bexp, Ea = arr(len(reac))
        rxns.append({'type': 'simple', 'reversible': True, 'reactants': stoich_map(reac),
                     'products': stoich_map(prod), 'A': A, 'b': bexp, 'Ea': Ea})

    def efficiencies():
        k = int(rng.integers(0, 6))
        cand = rng.choice(ns, size=k, replace=False) if k else []
        eff = {names[i]: float(np.round(rng.uniform(0.5, 5.0), 3)) for i in cand}
        return _ordered(eff)

    for _ in range(n_three_body - n_core_tb):
        a, b = exo_pairs[rng.integers(len(exo_pairs))]
        A, bexp, Ea = arr(3)
        rxns.append({'type': 'three-body', 'reversible': True, 'reactants': stoich_map([a, b]),
                     'products': stoich_map([by_comp[add(a, b)]]), 'A': A, 'b': bexp, 'Ea': Ea,
                     'efficiencies': efficiencies(), 'default-eff': 1.0})
    for _ in range(n_troe - n_core_troe):
        a, b = exo_pairs[rng.integers(len(exo_pairs))]
        Ai, bi, Ei = arr(2)
        A0, b0, E0 = arr(3)
        troe = [float(rng.uniform(0.05, 0.95)), float(rng.uniform(50.0, 600.0)), float(rng.uniform(500.0, 6000.0)),
                float(rng.uniform(2000.0, 9000.0)) if rng.uniform() < 0.7 else 0.0]
        rxns.append({'type': 'Troe', 'reversible': True, 'reactants': stoich_map([a, b]),
                     'products': stoich_map([by_comp[add(a, b)]]),
                     'fwd-A': Ai, 'fwd-b': bi, 'fwd-Ea': Ei, 'flf-A': A0, 'flf-b': b0, 'flf-Ea': E0,
                     'efficiencies': efficiencies(), 'default-eff': 1.0, 'Troe-params': np.array(troe)})
    order = rng.permutation(len(rxns))
    rxns = [rxns[i] for i in order]
    return {'element_mw_map': dict(ELEMENT_MW), 'elements': ['H', 'O', 'N'], 'ref_temperature': 298.15,
            'ref_pressure': ONE_ATM, 'gas_constant': GAS_CONSTANT, 'species_names': names,
            'species': species, 'reactions': rxns, 'seed': seed}


def mechanism_c1(seed=1):
    """10 species / 30 reactions: 24 reversible elementary, 4 three-body, 2 Troe falloff."""
    return make_mechanism(10, 30, n_three_body=4, n_troe=2, seed=seed)


def mechanism_53(seed=0):
    """53 species / 325 reactions (about 10% three-body, 8% Troe falloff)."""
    return make_mechanism(53, 325, seed=seed)


def random_states(ns, n, seed=0, T_range=(300.0, 2500.0), p_range=(0.1 * ONE_ATM, 10.0 * ONE_ATM), alpha=0.5):
    """Independent states: T uniform, p log-uniform, Y ~ Dirichlet(alpha).

    Returns (state SoA array [ns, n] with rows T, Y_0..Y_{ns-2}; pressures [n]; full Y [n, ns]).
    """
    rng = np.random.default_rng(seed)
    T = rng.uniform(T_range[0], T_range[1], n)
    p = np.exp(rng.uniform(np.log(p_range[0]), np.log(p_range[1]), n))
    Y = rng.dirichlet(alpha * np.ones(ns), n)
    state = np.empty((ns, n))
    state[0] = T
    state[1:] = Y[:, :-1].T
    return state, p, Y
