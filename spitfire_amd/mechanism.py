"""ChemicalMechanismSpec: mechanism container that owns a griffon drop-in (mirrors
src/spitfire/chemistry/mechanism.py:53-757).

The reference builds from Cantera (`cantera_input=` / `cantera_solution=`). This implementation reads Cantera-format
YAML files itself (`spitfire_amd.mechanism_io`, no Cantera), accepts the reference's stored mechanism data (the
`mech_data` a pickled reference ChemicalMechanismSpec carries) through `from_mech_data`, and takes mechanism data in
the extracted schema of `_extract_cantera_mechanism_data` (mechanism.py:387-541) directly, e.g. from
`spitfire_amd.synthetic`. Population of the griffon object follows `_populate_griffon_mechanism_data`
(mechanism.py:122-258): reactions stable-sorted by type, activation energies divided by the gas constant.
"""
import numpy as np

from .griffon import PyCombustionKernels
from .mechdata import sorted_reactions, species_molecular_weight


class Stream:
    """A gas mixture at (T, P, Y): the role Cantera Quantity objects play for the reference's streams
    (mechanism.py:602-646). Enthalpy and mixing use the mechanism's NASA7 / constant-cp data."""

    def __init__(self, mech, T, P, Y):
        self._mech = mech
        self.T = float(T)
        self.P = float(P)
        Y = np.asarray(Y, dtype=np.float64)
        self.Y = Y / np.sum(Y)

    @property
    def TP(self):
        return self.T, self.P

    @TP.setter
    def TP(self, tp):
        self.T, self.P = float(tp[0]), float(tp[1])

    @property
    def TPY(self):
        return self.T, self.P, self.Y.copy()

    @TPY.setter
    def TPY(self, v):
        self.T, self.P = float(v[0]), float(v[1])
        self.Y = self._mech._composition(v[2], 'Y')

    @property
    def TPX(self):
        return self.T, self.P, self.X

    @TPX.setter
    def TPX(self, v):
        self.T, self.P = float(v[0]), float(v[1])
        self.Y = self._mech._composition(v[2], 'X')

    @property
    def density(self):
        """ideal-gas density p M / (R T) (kg/m^3)"""
        return self.P * self.mean_molecular_weight / (self._mech.gas_constant * self.T)

    @property
    def species_names(self):
        return self._mech.species_names

    @property
    def mean_molecular_weight(self):
        return 1. / np.sum(self.Y / self._mech.molecular_weights)

    @property
    def X(self):
        return self.Y / self._mech.molecular_weights * self.mean_molecular_weight

    @property
    def H(self):
        """mixture enthalpy per unit mass (J/kg)"""
        return float(np.dot(self._mech.species_enthalpies(self.T), self.Y))

    @property
    def state(self):
        """[T, Y_0 .. Y_{ns-2}], the reactor / flamelet state vector"""
        return np.hstack([self.T, self.Y[:-1]])

    def n_atoms(self, species, element):
        return self._mech.mech_data['species'][species]['atoms'].get(element, 0.)


class ChemicalMechanismSpec:
    """Mechanism container: ChemicalMechanismSpec(cantera_input='mech.yaml', group_name='gas'), or
    ChemicalMechanismSpec(mech_data) with mechanism data in the extracted schema (see the module docstring)."""

    def __init__(self, cantera_input=None, group_name='gas', cantera_solution=None, cantera_xml=None,
                 mech_data=None):
        if cantera_solution is not None:
            raise NotImplementedError('Cantera solutions are not supported: give a Cantera YAML file or mechanism data')
        if isinstance(cantera_input, dict):  # mechanism data given positionally
            mech_data, cantera_input = cantera_input, None
        if cantera_input is None:
            cantera_input = cantera_xml
        if mech_data is None:
            if cantera_input is None:
                raise ValueError('ChemicalMechanismSpec needs a Cantera YAML file or mechanism data')
            from .mechanism_io import load_cantera_yaml
            mech_data = load_cantera_yaml(cantera_input, group_name)
            self._mech_file_path = cantera_input
        else:
            self._mech_file_path = None
        self._group_name = group_name
        self._element_stoichiometry = dict(self._default_element_stoichiometry)
        self._mech_data = mech_data
        self._species_names = list(mech_data['species_names'])
        self._species_index = {n: i for i, n in enumerate(self._species_names)}
        self._reactions, self._reaction_file_order = sorted_reactions(mech_data['reactions'])
        emw = mech_data['element_mw_map']
        self._mw = np.array([species_molecular_weight(mech_data['species'][n]['atoms'], emw)
                             for n in self._species_names])
        self._griffon = PyCombustionKernels()
        self._populate_griffon()

    @classmethod
    def from_mech_data(cls, mech_data):
        """From the reference's stored mechanism data (a pickled reference ChemicalMechanismSpec's
        'mech_data', mechanism.py:99-103) or from data in this package's extracted schema."""
        if 'species_names' not in mech_data:
            from .mechanism_io import from_reference_mech_data
            mech_data = from_reference_mech_data(mech_data)
        return cls(mech_data=mech_data)

    def reference_mech_data(self):
        """This mechanism in the reference's stored form (type-sorted reaction tuples, mechanism.py:122-258)."""
        from .mechanism_io import to_reference_mech_data
        return to_reference_mech_data(self._mech_data)

    @property
    def mech_file_path(self):
        return self._mech_file_path

    @property
    def group_name(self):
        return self._group_name

    @property
    def element_stoichiometry(self):
        return self._element_stoichiometry

    @element_stoichiometry.setter
    def element_stoichiometry(self, custom):
        if custom is not None:
            self._element_stoichiometry = custom

    def molecular_weight(self, ni):
        if isinstance(ni, str):
            return float(self._mw[self._species_index[ni]])
        if isinstance(ni, (int, np.integer)):
            return float(self._mw[ni])
        raise TypeError('ChemicalMechanismSpec.molecular_weight(ni) takes a string or integer, given ' + str(ni))

    @classmethod
    def from_synthetic(cls, n_species=53, n_reactions=325, seed=0, **kwargs):
        from .synthetic import make_mechanism
        return cls(make_mechanism(n_species, n_reactions, seed=seed, **kwargs))

    def _populate_griffon(self):
        g, md = self._griffon, self._mech_data
        R = md['gas_constant']
        g.mechanism_set_ref_pressure(md['ref_pressure'])
        g.mechanism_set_ref_temperature(md['ref_temperature'])
        g.mechanism_set_gas_constant(R)
        g.mechanism_set_element_mw_map(md['element_mw_map'])
        for e in md['elements']:
            g.mechanism_add_element(e)
        for s in self._species_names:
            g.mechanism_add_species(s, md['species'][s]['atoms'])
        g.mechanism_resize_heat_capacity_data()
        for s in self._species_names:
            cp = md['species'][s]['heat-capacity']
            if cp['type'] == 'constant':
                g.mechanism_add_const_cp(s, cp['Tmin'], cp['Tmax'], cp['T0'], cp['h0'], cp['s0'], cp['cp'])
            elif cp['type'] == 'NASA7':
                g.mechanism_add_nasa7_cp(s, cp['Tmin'], cp['Tmid'], cp['Tmax'], list(cp['low-coeffs']),
                                         list(cp['high-coeffs']))
            else:
                g.mechanism_add_nasa9_cp(s, cp['Tmin'], cp['Tmax'], cp['coeffs'])
        for rx in self._reactions:
            t = rx['type']
            orders = rx.get('orders')
            if t in ('simple', 'three-body'):
                args = (rx['reactants'], rx['products'], rx['reversible'], rx['A'], rx['b'], rx['Ea'] / R)
            else:
                args = (rx['reactants'], rx['products'], rx['reversible'], rx['fwd-A'], rx['fwd-b'],
                        rx['fwd-Ea'] / R)
            if t == 'simple':
                if orders:
                    g.mechanism_add_reaction_simple_with_special_orders(*args, orders)
                else:
                    g.mechanism_add_reaction_simple(*args)
            elif t == 'three-body':
                extra = (rx['efficiencies'], rx['default-eff'])
                if orders:
                    g.mechanism_add_reaction_three_body_with_special_orders(*args, *extra, orders)
                else:
                    g.mechanism_add_reaction_three_body(*args, *extra)
            elif t == 'Lindemann':
                extra = (rx['efficiencies'], rx['default-eff'], rx['flf-A'], rx['flf-b'], rx['flf-Ea'] / R)
                if orders:
                    g.mechanism_add_reaction_Lindemann_with_special_orders(*args, *extra, orders)
                else:
                    g.mechanism_add_reaction_Lindemann(*args, *extra)
            elif t == 'Troe':
                extra = (rx['efficiencies'], rx['default-eff'], rx['flf-A'], rx['flf-b'], rx['flf-Ea'] / R,
                         list(np.asarray(rx['Troe-params'], dtype=float)))
                if orders:
                    g.mechanism_add_reaction_Troe_with_special_orders(*args, *extra, orders)
                else:
                    g.mechanism_add_reaction_Troe(*args, *extra)
            else:
                raise ValueError('unknown reaction type ' + str(t))

    @property
    def griffon(self):
        return self._griffon

    @property
    def mech_data(self):
        return self._mech_data

    @property
    def n_species(self):
        return len(self._species_names)

    @property
    def n_reactions(self):
        return len(self._reactions)

    @property
    def species_names(self):
        return list(self._species_names)

    @property
    def molecular_weights(self):
        return self._mw.copy()

    @property
    def gas_constant(self):
        return self._mech_data['gas_constant']

    @property
    def reaction_file_order(self):
        """griffon reaction index -> index in the input reaction list (the type sort of mechanism.py:540)."""
        return list(self._reaction_file_order)

    def species_index(self, name):
        return self._species_index[name]

    # ---------------- streams (mechanism.py:602-757) ----------------
    _default_element_stoichiometry = {'O': -1.0, 'H': 0.5, 'C': 2.0, 'Al': 1.5, 'U': 1.0, 'Ar': 0.0, 'N': 0.0,
                                      'He': 0.0}

    def _composition(self, spec, basis):
        y = np.zeros(self.n_species)
        if isinstance(spec, str):
            import re
            for name, val in re.findall(r'([^\s,:]+)\s*:\s*([-+0-9.eE]+)', spec):
                matches = [n for n in self._species_names if n.lower() == name.lower()]
                if not matches:
                    raise KeyError('unknown species ' + name)
                y[self._species_index[matches[0]]] += float(val)
        else:
            y[:] = np.asarray(spec, dtype=np.float64)
        if basis == 'X':
            y = y * self._mw
        return y / np.sum(y)

    def stream(self, properties=None, values=None, stp_air=False):
        """A Stream from 'TPY', 'TPX', 'HPY', 'HPX', 'TP', 'X' or 'Y' properties (mechanism.py:602-640), or
        stp_air=True: O2:1 N2:3.74 (moles) at 300 K and 1 atm. Properties not given keep the defaults of a fresh
        Cantera phase: 300 K, 1 atm, mole fractions all on the first species."""
        if stp_air:
            return Stream(self, 300., 101325., self._composition('O2:1 N2:3.74', 'X'))
        if properties is None or values is None:
            raise ValueError('ChemicalMechanismSpec.stream() needs either stp_air=True or properties and values, '
                             "e.g. stream('TPY', (300., 101325., 'O2:1, N2:1'))")
        y0 = np.zeros(self.n_species)
        y0[0] = 1.
        if properties in ('X', 'Y'):
            return Stream(self, 300., 101325., self._composition(values, properties))
        if properties == 'TP':
            return Stream(self, values[0], values[1], y0)
        if properties not in ('TPY', 'TPX', 'HPY', 'HPX'):
            raise ValueError('supported stream properties: TPY, TPX, HPY, HPX, TP, X, Y')
        a, p, comp = values
        Y = self._composition(comp, properties[2])
        if properties[0] == 'T':
            return Stream(self, a, p, Y)
        return Stream(self, self.temperature_from_enthalpy(a, Y), p, Y)

    def copy_stream(self, stream):
        return Stream(self, stream.T, stream.P, stream.Y.copy())

    def mix_for_equivalence_ratio(self, phi, fuel, oxy):
        """Fuel and oxidizer mixed (by moles) to equivalence ratio phi, at 300 K and 1 atm (the state of the
        reference's phase object). The molar fuel/oxidizer ratio is phi times the stoichiometric one, where
        oxygen required = C + S + H/4 and oxygen present = O/2 per mole of each stream, as in Cantera's
        set_equivalence_ratio (mechanism.py:685-688)."""
        def o2_net(stream):
            X = stream.X
            req = pres = 0.
            for i, n in enumerate(self._species_names):
                atoms = self._mech_data['species'][n]['atoms']
                req += X[i] * (atoms.get('C', 0.) + atoms.get('S', 0.) + 0.25 * atoms.get('H', 0.))
                pres += X[i] * atoms.get('O', 0.)
            return req - 0.5 * pres
        of, oo = o2_net(fuel), o2_net(oxy)
        if of <= 0. or oo > 0.:
            raise ValueError('fuel and oxidizer streams do not define an equivalence ratio')
        X = phi * (-oo / of) * fuel.X + oxy.X
        return Stream(self, 300., 101325., X * self._mw)

    def mix_for_normalized_equivalence_ratio(self, normalized_phi, fuel, oxy):
        return self.mix_for_equivalence_ratio(normalized_phi / (1. - normalized_phi), fuel, oxy)

    def mix_fuels_for_stoich_mixture_fraction(self, fuel1, fuel2, zstoich, oxy):
        """Two fuels mixed (by moles) so that the blend against `oxy` has stoichiometric mixture fraction zstoich
        (mechanism.py:732-757; hydrocarbon/hydrogen oxidation)."""
        atoms = ['H', 'O', 'C']
        ca, cb, cx = (self._get_atoms_in_stream(s, atoms) for s in (fuel1, fuel2, oxy))
        dmf = fuel1.mean_molecular_weight - fuel2.mean_molecular_weight
        etax = cx['O'] - 0.5 * cx['H'] - 2. * cx['C']
        netab = -cb['O'] + 0.5 * cb['H'] + 2. * cb['C']
        netaa = -ca['O'] + 0.5 * ca['H'] + 2. * ca['C']
        lhs = oxy.mean_molecular_weight / etax * zstoich / (1. - zstoich)
        k = (fuel2.mean_molecular_weight - lhs * netab) / (lhs * (netaa - netab) - dmf)
        if k > 1. or k < 0.:
            raise ValueError('The provided fuel combination cannot reach the desired stoichiometric mixture fraction!')
        return self.mix_streams([(fuel1, k), (fuel2, 1. - k)], 'mole')

    def species_enthalpies(self, T):
        """species enthalpies per unit mass at T (thermodynamics_kernels.cpp:262-351 clamping rules)"""
        h = np.zeros(self.n_species)
        R = self.gas_constant
        for i, n in enumerate(self._species_names):
            cp = self._mech_data['species'][n]['heat-capacity']
            if cp['type'] == 'constant':
                h[i] = (cp['h0'] + cp['cp'] * (T - cp['T0'])) / self._mw[i]
                continue
            Tmid, Tmin, Tmax = cp['Tmid'], cp['Tmin'], cp['Tmax']
            Te = min(max(T, Tmin), Tmax)
            a = np.asarray(cp['low-coeffs'] if Te <= Tmid else cp['high-coeffs'])
            hRT = a[0] * Te + a[1] * Te ** 2 / 2 + a[2] * Te ** 3 / 3 + a[3] * Te ** 4 / 4 + a[4] * Te ** 5 / 5 + a[5]
            cpe = a[0] + Te * (a[1] + Te * (a[2] + Te * (a[3] + Te * a[4])))
            h[i] = R * (hRT + cpe * (T - Te)) / self._mw[i]
        return h

    def species_cp(self, T):
        cpv = np.zeros(self.n_species)
        R = self.gas_constant
        for i, n in enumerate(self._species_names):
            cp = self._mech_data['species'][n]['heat-capacity']
            if cp['type'] == 'constant':
                cpv[i] = cp['cp'] / self._mw[i]
                continue
            Te = min(max(T, cp['Tmin']), cp['Tmax'])
            a = np.asarray(cp['low-coeffs'] if Te <= cp['Tmid'] else cp['high-coeffs'])
            cpv[i] = R * (a[0] + Te * (a[1] + Te * (a[2] + Te * (a[3] + Te * a[4])))) / self._mw[i]
        return cpv

    def _thermo_tables(self):
        """NASA7 / constant-cp coefficients as arrays, for the vectorized thermo below"""
        if getattr(self, '_tt', None) is None:
            ns = self.n_species
            lo, hi = np.zeros((ns, 7)), np.zeros((ns, 7))
            tmid, tmin, tmax = np.zeros(ns), np.zeros(ns), np.zeros(ns)
            const = np.zeros(ns, dtype=bool)
            c0 = np.zeros((ns, 3))  # T0, h0, cp of constant-cp species
            for i, n in enumerate(self._species_names):
                cp = self._mech_data['species'][n]['heat-capacity']
                if cp['type'] == 'constant':
                    const[i] = True
                    c0[i] = cp['T0'], cp['h0'], cp['cp']
                    continue
                lo[i], hi[i] = cp['low-coeffs'], cp['high-coeffs']
                tmid[i], tmin[i], tmax[i] = cp['Tmid'], cp['Tmin'], cp['Tmax']
            self._tt = (lo, hi, tmid, tmin, tmax, const, c0)
        return self._tt

    def species_enthalpies_array(self, T):
        """species enthalpies per unit mass [len(T), ns] for an array of temperatures (species_enthalpies
        applied row by row)"""
        lo, hi, tmid, tmin, tmax, const, c0 = self._thermo_tables()
        T = np.asarray(T, dtype=np.float64)[:, None]
        Te = np.minimum(np.maximum(T, tmin[None, :]), tmax[None, :])
        a = np.where((Te <= tmid[None, :])[:, :, None], lo[None, :, :], hi[None, :, :])
        hRT = (a[..., 0] * Te + a[..., 1] * Te ** 2 / 2 + a[..., 2] * Te ** 3 / 3 + a[..., 3] * Te ** 4 / 4 +
               a[..., 4] * Te ** 5 / 5 + a[..., 5])
        cpe = a[..., 0] + Te * (a[..., 1] + Te * (a[..., 2] + Te * (a[..., 3] + Te * a[..., 4])))
        h = self.gas_constant * (hRT + cpe * (T - Te)) / self._mw[None, :]
        if const.any():
            h[:, const] = (c0[const, 1] + c0[const, 2] * (T - c0[const, 0])) / self._mw[const]
        return h

    def species_cp_array(self, T):
        lo, hi, tmid, tmin, tmax, const, c0 = self._thermo_tables()
        T = np.asarray(T, dtype=np.float64)[:, None]
        Te = np.minimum(np.maximum(T, tmin[None, :]), tmax[None, :])
        a = np.where((Te <= tmid[None, :])[:, :, None], lo[None, :, :], hi[None, :, :])
        cp = self.gas_constant * (a[..., 0] + Te * (a[..., 1] + Te * (a[..., 2] + Te * (a[..., 3] + Te * a[..., 4]))))
        cp = cp / self._mw[None, :]
        if const.any():
            cp[:, const] = c0[const, 2] / self._mw[const]
        return cp

    def temperatures_from_enthalpies(self, h, Y, T0=1000., tol=1e-12):
        """temperature_from_enthalpy for many mixtures at once (h [n], Y [n, ns])"""
        h = np.asarray(h, dtype=np.float64)
        T = np.full(h.shape, float(T0)) if np.ndim(T0) == 0 else np.array(T0, dtype=np.float64)
        done = np.zeros(h.shape, dtype=bool)
        for _ in range(100):
            f = np.sum(self.species_enthalpies_array(T) * Y, axis=1) - h
            dT = f / np.sum(self.species_cp_array(T) * Y, axis=1)
            T = np.where(done, T, T - dT)
            done |= np.abs(dT) < tol * T
            if done.all():
                break
        return T

    def temperature_from_enthalpy(self, h, Y, T0=1000., tol=1e-12):
        """T such that h_mix(T, Y) = h, by Newton iteration in FP64."""
        T = T0
        for _ in range(100):
            f = np.dot(self.species_enthalpies(T), Y) - h
            dT = f / np.dot(self.species_cp(T), Y)
            T -= dT
            if abs(dT) < tol * T:
                break
        return float(T)

    @staticmethod
    def mix_streams(streams, basis, constant='HP'):
        """Mix [(stream, amount)] by mass (or moles) at constant HP or TP (mechanism.py:648-684)."""
        mech = streams[0][0]._mech
        m = np.array([amt if basis == 'mass' else amt * s.mean_molecular_weight for s, amt in streams])
        Y = sum(mi * s.Y for mi, (s, _) in zip(m, streams)) / np.sum(m)
        P = streams[0][0].P
        if constant == 'HP':
            H = sum(mi * s.H for mi, (s, _) in zip(m, streams)) / np.sum(m)
            return Stream(mech, mech.temperature_from_enthalpy(H, Y, T0=streams[0][0].T), P, Y)
        if constant == 'TP':
            return Stream(mech, streams[0][0].T, P, Y)
        raise ValueError('constant must be HP or TP')

    def _get_atoms_in_stream(self, stream, atom_names):
        out = {a: 0. for a in atom_names}
        X = stream.X
        for i, n in enumerate(self._species_names):
            for a in atom_names:
                out[a] += X[i] * self._mech_data['species'][n]['atoms'].get(a, 0.)
        return out

    def stoich_molar_fuel_to_oxy_ratio(self, fuel_stream, oxy_stream):
        for a in self._mech_data['elements']:
            if a not in self._element_stoichiometry:
                raise KeyError('atom %r of the mechanism is not in the element stoichiometry map %r'
                               % (a, self._element_stoichiometry))
        atoms = [a for a in self._element_stoichiometry if a in self._mech_data['elements']]
        fa = self._get_atoms_in_stream(fuel_stream, atoms)
        oa = self._get_atoms_in_stream(oxy_stream, atoms)
        es = self._element_stoichiometry
        return -sum(es[a] * oa[a] for a in atoms) / sum(es[a] * fa[a] for a in atoms)

    def stoich_mass_fuel_to_oxy_ratio(self, fuel_stream, oxy_stream):
        return (fuel_stream.mean_molecular_weight / oxy_stream.mean_molecular_weight *
                self.stoich_molar_fuel_to_oxy_ratio(fuel_stream, oxy_stream))

    def stoich_mixture_fraction(self, fuel_stream, oxy_stream):
        eta = self.stoich_mass_fuel_to_oxy_ratio(fuel_stream, oxy_stream)
        return eta / (1. + eta)

    def __getstate__(self):
        return {'mech_data': self._mech_data, 'element_stoichiometry': self._element_stoichiometry}

    def __setstate__(self, state):
        self.__init__(mech_data=state['mech_data'])
        if 'element_stoichiometry' in state:
            self._element_stoichiometry = state['element_stoichiometry']
