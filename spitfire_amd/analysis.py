"""Property evaluators and chemical explosive mode analysis over Library states, mirroring
src/spitfire/chemistry/analysis.py.

The reference evaluates the properties through a Cantera SolutionArray (analysis.py:24-220). Here they come from
one batched kernel on the MI355X (gf_batch_mixture_properties: the NASA7 mixture getters of griffon.pyx:684-758 over
every state of the library at once), with the state set as Cantera's SolutionArray.TPY / TDY would set it: negative
mass fractions are clipped to zero and the mass fractions renormalised to sum to one; the pressure is used when the
library holds one, else the density.

compute_viscosity needs mixture-averaged transport (Cantera's collision-integral fits and Wilke mixing), which this
implementation does not provide: it raises NotImplementedError.
"""
import numpy as np

from .library import Library
from .mechanism import ChemicalMechanismSpec


def _check(mechanism, library, who):
    if not isinstance(mechanism, ChemicalMechanismSpec):
        raise TypeError('Input argument "mechanism" to %s() must be a ChemicalMechanismSpec' % who)
    if not isinstance(library, Library):
        raise TypeError('Input argument "output_library" to %s() must be a Library' % who)


def _library_states(mechanism, library):
    """(shape, T [n], Y [n, ns] as Cantera sets them, p [n] or None, rho [n] or None), analysis.py:24-74"""
    required = ['temperature'] + ['mass fraction ' + s for s in mechanism.species_names]
    if any(r not in library for r in required) or ('density' not in library and 'pressure' not in library):
        raise ValueError('Library argument must contain temperature, all mass fractions, and pressure or density')
    shape = library['temperature'].shape
    T = np.ascontiguousarray(library['temperature'].ravel(), dtype=np.float64)
    Y = np.empty((T.size, mechanism.n_species))
    for i, name in enumerate(mechanism.species_names):
        Y[:, i] = library['mass fraction ' + name].ravel()
    Y = np.maximum(Y, 0.)
    Y /= Y.sum(axis=1, keepdims=True)
    if 'pressure' in library:
        return shape, T, Y, np.ascontiguousarray(library['pressure'].ravel(), dtype=np.float64), None
    return shape, T, Y, None, np.ascontiguousarray(library['density'].ravel(), dtype=np.float64)


def mixture_properties(mechanism, library):
    """Every state of the library through the batched mixture-property kernel. Returns (shape, dict) with
    'molecular weight', 'density', 'pressure', 'cp', 'cv', 'enthalpy', 'energy' as flat arrays."""
    import torch
    shape, T, Y, p, rho = _library_states(mechanism, library)
    n = T.size
    dev = torch.device('cuda')
    Ty = torch.from_numpy(np.ascontiguousarray(np.hstack((T[:, None], Y)))).to(dev)
    pd = torch.from_numpy(p).to(dev) if p is not None else None
    rd = torch.from_numpy(rho).to(dev) if rho is not None else None
    out = torch.empty((n, 8), dtype=torch.float64, device=dev)
    mechanism.griffon.batch_mixture_properties(Ty, out, p=pd, rho=rd)
    o = out.cpu().numpy()
    props = {'molecular weight': o[:, 0], 'cp': o[:, 3], 'cv': o[:, 4], 'enthalpy': o[:, 5], 'energy': o[:, 6]}
    props['density'] = o[:, 1] if p is not None else rho
    props['pressure'] = p if p is not None else o[:, 2]
    return shape, props


def _add(mechanism, output_library, who, prop, name):
    _check(mechanism, output_library, who)
    shape, props = mixture_properties(mechanism, output_library)
    output_library[name] = props[prop].reshape(shape)
    return output_library


def compute_specific_enthalpy(mechanism, output_library):
    """adds 'enthalpy', the mixture's specific enthalpy in J/kg (analysis.py:77-98)"""
    return _add(mechanism, output_library, 'compute_specific_enthalpy', 'enthalpy', 'enthalpy')


def compute_density(mechanism, output_library):
    """adds 'density' in kg/m^3 (analysis.py:101-122)"""
    return _add(mechanism, output_library, 'compute_density', 'density', 'density')


def compute_pressure(mechanism, output_library):
    """adds 'pressure' in Pa (analysis.py:125-146)"""
    return _add(mechanism, output_library, 'compute_pressure', 'pressure', 'pressure')


def compute_isobaric_specific_heat(mechanism, output_library):
    """adds 'heat capacity cp' in J/kg/K (analysis.py:173-194)"""
    return _add(mechanism, output_library, 'compute_isobaric_specific_heat', 'cp', 'heat capacity cp')


def compute_isochoric_specific_heat(mechanism, output_library):
    """adds 'heat capacity cv' in J/kg/K (analysis.py:197-218)"""
    return _add(mechanism, output_library, 'compute_isochoric_specific_heat', 'cv', 'heat capacity cv')


def compute_viscosity(mechanism, output_library):
    """analysis.py:149-170 (Cantera mixture-averaged viscosity): not available without transport data fits"""
    raise NotImplementedError('compute_viscosity needs mixture-averaged transport properties, which this '
                              'implementation does not provide')


def explosive_mode_analysis(mechanism, output_library, configuration='isobaric', heat_transfer='adiabatic',
                            compute_explosion_indices=False, compute_participation_indices=False,
                            include_secondary_mode=False):
    """Chemical explosive mode analysis of every state of a library (analysis.py:221-383): the reactor Jacobians of
    all states come from one batched launch (gf_batch_reactor_general; the reference calls reactor_jac_isobaric /
    reactor_jac_isochoric per state), and their eigen-decompositions from LAPACK on the host as in the reference.
    Adds 'cema-lexp1' (the eigenvalue with the largest real part, real parts within 1e-4 of zero excluded), and
    on request 'cema-lexp2', and the explosion indices 'cema-ei1 T', 'cema-ei1 <species>' ('cema-ei2 ...').
    Participation indices need the reference's Cantera stoichiometry and enthalpy evaluation and raise
    NotImplementedError."""
    import torch
    from scipy.linalg import eig, eigvals
    _check(mechanism, output_library, 'explosive_mode_analysis')
    if compute_participation_indices:
        raise NotImplementedError('CEMA participation indices are not available in this implementation')
    if configuration not in ('isobaric', 'isochoric'):
        raise ValueError('configuration must be "isobaric" or "isochoric"')
    if heat_transfer not in ('adiabatic', 'diathermal', 'isothermal'):
        raise ValueError('heat_transfer must be "adiabatic", "diathermal" or "isothermal"')
    ns = mechanism.n_species
    iso_v = configuration == 'isochoric'
    ne = ns + 1 if iso_v else ns
    Tidx = 1 if iso_v else 0
    shape = output_library['temperature'].shape
    T = output_library['temperature'].ravel()
    n = T.size
    x = np.empty((n, ne))
    if iso_v:
        x[:, 0] = output_library['density'].ravel()
    x[:, Tidx] = T
    for i, name in enumerate(mechanism.species_names[:-1]):
        x[:, Tidx + 1 + i] = output_library['mass fraction ' + name].ravel()
    dev = torch.device('cuda')
    xd = torch.from_numpy(x).to(dev)
    rhs = torch.empty((n, ne), dtype=torch.float64, device=dev)
    jac = torch.empty((n, ne * ne), dtype=torch.float64, device=dev)
    p = None if iso_v else torch.from_numpy(np.ascontiguousarray(output_library['pressure'].ravel(),
                                                                  dtype=np.float64)).to(dev)
    # the reference's adiabatic and diathermal cases both pass heat-transfer option 0 (analysis.py:325-343)
    mechanism.griffon.batch_reactor_general(xd, rhs, jac, p=p, isochoric=iso_v,
                                            heat_transfer_option=1 if heat_transfer == 'isothermal' else 0)
    J = jac.cpu().numpy()
    want_vectors = compute_explosion_indices
    lexp1, lexp2 = np.zeros(n), np.zeros(n)
    ei1, ei2 = np.zeros((ns, n)), np.zeros((ns, n))
    for i in range(n):
        Ji = J[i].reshape((ne, ne), order='F')
        if want_vectors:
            w, vl, vr = eig(Ji, left=True, right=True)
        else:
            w = eigvals(Ji)
        re = np.real(w)
        re[(re > -1.e-4) & (re < 1.e-4)] = -np.inf
        k1 = np.argmax(re)
        k2 = np.argmax(np.delete(re, k1))
        lexp1[i] = re[k1]
        lexp2[i] = re[k2]
        if want_vectors:
            ep1 = np.abs(np.real(vl[:, k1]) * np.real(vr[:, k1]))
            ei1[:, i] = ep1[Tidx:] / np.sum(np.abs(ep1))
            if include_secondary_mode:
                # as the reference, the secondary vectors use the index into the reduced list (analysis.py:304-318)
                ep2 = np.abs(np.real(vl[:, k2]) * np.real(vr[:, k2]))
                s2 = np.sum(np.abs(ep2))
                ei2[:, i] = ep2[Tidx:] / (1. if s2 < 1.e-16 else s2)
    output_library['cema-lexp1'] = lexp1.reshape(shape)
    names = mechanism.species_names
    if compute_explosion_indices:
        output_library['cema-ei1 T'] = ei1[0].reshape(shape)
        for k in range(ns - 1):
            output_library['cema-ei1 ' + names[k]] = ei1[1 + k].reshape(shape)
    if include_secondary_mode:
        output_library['cema-lexp2'] = lexp2.reshape(shape)
        if compute_explosion_indices:
            output_library['cema-ei2 T'] = ei2[0].reshape(shape)
            for k in range(ns - 1):
                output_library['cema-ei2 ' + names[k]] = ei2[1 + k].reshape(shape)
    return output_library
