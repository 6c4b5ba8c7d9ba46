"""Drop-in for the reference's `spitfire.griffon.griffon` Cython module (src/spitfire/griffon/griffon.pyx).

`PyCombustionKernels` keeps the reference's method names, positional argument order and in-place output
convention (contiguous float64 / int32 NumPy arrays written in place, methods return None), including the
reference's quirks:
  * production_rates(T, rho, y, out): the Cython parameter names say (density, temperature) but the values
    are forwarded to C++ as (T, rho) and every caller passes T first (griffon.pyx:763-768);
  * flamelet_rhs/flamelet_jacobian take the heat-loss arrays in the order T_convection, T_radiation,
    h_convection, h_radiation and forward them to C++ as T_conv, h_conv, T_rad, h_rad (griffon.pyx:580-679).
Errors raise NativeError instead of terminating the process.

Every compute method runs the gfx950 HIP kernels of libgriffon_amd.so on the current HIP device. Batched
device entry points operate on torch tensors that already live on the GPU (`batch_*` methods).
"""
import numpy as np

from . import _native
from ._native import check, dptr, iptr, lib

_RTYPE = {'simple': 1, 'three-body': 2, 'Lindemann': 3, 'Troe': 4}


def _f64(a, name='array'):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or a.ndim != 1 or not a.flags['C_CONTIGUOUS']:
        raise ValueError(name + ' must be a contiguous 1-D float64 numpy array')
    return a


def _i32(a, name='array'):
    if not isinstance(a, np.ndarray) or a.dtype != np.int32 or a.ndim != 1 or not a.flags['C_CONTIGUOUS']:
        raise ValueError(name + ' must be a contiguous 1-D int32 numpy array')
    return a


class PyCombustionKernels:
    """Python-side interface to the MI355X griffon kernels (mirrors griffon.pyx:220-987)."""

    def __init__(self):
        self._lib = lib()
        h = _native._V()
        check(self._lib.gf_mech_create(h))
        self._h = h
        self._element_mw = {}
        self._elements = []
        self._species = []
        self._index = {}
        self._ns = 0

    def __del__(self):
        try:
            self._lib.gf_mech_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def n_species(self):
        return self._ns

    # ---------------- mechanism setters (griffon.pyx:234-551) ----------------
    def mechanism_set_element_mw_map(self, element_mw_map):
        self._element_mw = dict(element_mw_map)

    def mechanism_add_element(self, element_name):
        if element_name not in self._elements:
            self._elements.append(element_name)

    def mechanism_add_species(self, species_name, atom_map):
        if species_name in self._index:
            raise ValueError('species ' + species_name + ' cannot be added twice')
        for a in atom_map:
            if a not in self._elements:
                raise ValueError('cannot find atom ' + a)
        mw = 0.0
        for a in sorted(atom_map):  # std::map order (chemistry_setup.cpp:52-56)
            mw += self._element_mw[a] * atom_map[a]
        check(self._lib.gf_mechanism_add_species(self._h, mw))
        self._index[species_name] = self._ns
        self._species.append(species_name)
        self._ns += 1

    def mechanism_set_ref_pressure(self, p):
        check(self._lib.gf_mechanism_set_ref_pressure(self._h, float(p)))

    def mechanism_set_ref_temperature(self, T):
        check(self._lib.gf_mechanism_set_ref_temperature(self._h, float(T)))

    def mechanism_set_gas_constant(self, Ru):
        check(self._lib.gf_mechanism_set_gas_constant(self._h, float(Ru)))

    def mechanism_resize_heat_capacity_data(self):
        check(self._lib.gf_mechanism_resize_heat_capacity_data(self._h))

    def mechanism_add_const_cp(self, spec_name, Tmin, Tmax, T0, h0, s0, cp):
        check(self._lib.gf_mechanism_add_const_cp(self._h, self._index[spec_name], Tmin, Tmax, T0, h0, s0, cp))

    def mechanism_add_nasa7_cp(self, spec_name, Tmin, Tmid, Tmax, low_coeffs, high_coeffs):
        lo = np.ascontiguousarray(low_coeffs, dtype=np.float64)
        hi = np.ascontiguousarray(high_coeffs, dtype=np.float64)
        if lo.size != 7 or hi.size != 7:
            raise ValueError('NASA7 data needs seven coefficients per range')
        check(self._lib.gf_mechanism_add_nasa7_cp(self._h, self._index[spec_name], Tmin, Tmid, Tmax, dptr(lo),
                                                  dptr(hi)))

    def mechanism_add_nasa9_cp(self, spec_name, Tmin, Tmax, coeffs):
        raise NotImplementedError('NASA9 heat capacities are not supported by the MI355X implementation')

    def _idx_list(self, d, cast):
        keys = sorted(d)  # std::map<std::string, ...> iteration order
        idx = np.array([self._index[k] for k in keys], dtype=np.int32)
        val = np.array([cast(d[k]) for k in keys], dtype=np.int32 if cast is int else np.float64)
        return idx, val

    def _add(self, rtype, reactants, products, reversible, A, b, Ea, efficiencies=None, default_eff=0.0,
             flf=None, troe=None, orders=None):
        ri, rs = self._idx_list(reactants, int)
        pi, ps = self._idx_list(products, int)
        kf = np.array([A, b, Ea], dtype=np.float64)
        if efficiencies:
            eff = {k: v for k, v in efficiencies.items() if k in self._index}
            ti, te = self._idx_list(eff, float)
        else:
            ti, te = np.zeros(0, dtype=np.int32), np.zeros(0)
        kp = np.array(flf if flf is not None else [0., 0., 0.], dtype=np.float64)
        tp = np.zeros(4)
        if troe is not None:
            t = np.asarray(troe, dtype=np.float64)
            tp[:t.size] = t
        if orders:
            oi, ov = self._idx_list(orders, float)
        else:
            oi, ov = np.zeros(0, dtype=np.int32), np.zeros(0)
        check(self._lib.gf_mechanism_add_reaction(self._h, rtype, int(bool(reversible)), len(ri), iptr(ri), iptr(rs),
                                                  len(pi), iptr(pi), iptr(ps), dptr(kf), len(ti), iptr(ti), dptr(te),
                                                  float(default_eff), dptr(kp), dptr(tp), len(oi), iptr(oi),
                                                  dptr(ov)))

    def mechanism_add_reaction_simple(self, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value,
                                      fwd_temp_exponent, fwd_act_energy):
        self._add(1, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy)

    def mechanism_add_reaction_three_body(self, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value,
                                          fwd_temp_exponent, fwd_act_energy, three_body_efficiencies,
                                          default_efficiency):
        self._add(2, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, three_body_efficiencies, default_efficiency)

    def mechanism_add_reaction_Lindemann(self, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value,
                                         fwd_temp_exponent, fwd_act_energy, three_body_efficiencies,
                                         default_efficiency, flf_pre_exp_value, flf_temp_exponent, flf_act_energy):
        self._add(3, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, three_body_efficiencies, default_efficiency,
                  [flf_pre_exp_value, flf_temp_exponent, flf_act_energy])

    def mechanism_add_reaction_Troe(self, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value,
                                    fwd_temp_exponent, fwd_act_energy, three_body_efficiencies, default_efficiency,
                                    flf_pre_exp_value, flf_temp_exponent, flf_act_energy, troe_parameters):
        self._add(4, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, three_body_efficiencies, default_efficiency,
                  [flf_pre_exp_value, flf_temp_exponent, flf_act_energy], troe_parameters)

    def mechanism_add_reaction_simple_with_special_orders(self, reactants_stoich, products_stoich, reversible,
                                                          fwd_pre_exp_value, fwd_temp_exponent, fwd_act_energy,
                                                          special_orders):
        self._add(1, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, orders=special_orders)

    def mechanism_add_reaction_three_body_with_special_orders(self, reactants_stoich, products_stoich, reversible,
                                                              fwd_pre_exp_value, fwd_temp_exponent, fwd_act_energy,
                                                              three_body_efficiencies, default_efficiency,
                                                              special_orders):
        self._add(2, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, three_body_efficiencies, default_efficiency, orders=special_orders)

    def mechanism_add_reaction_Lindemann_with_special_orders(self, reactants_stoich, products_stoich, reversible,
                                                             fwd_pre_exp_value, fwd_temp_exponent, fwd_act_energy,
                                                             three_body_efficiencies, default_efficiency,
                                                             flf_pre_exp_value, flf_temp_exponent, flf_act_energy,
                                                             special_orders):
        self._add(3, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, three_body_efficiencies, default_efficiency,
                  [flf_pre_exp_value, flf_temp_exponent, flf_act_energy], orders=special_orders)

    def mechanism_add_reaction_Troe_with_special_orders(self, reactants_stoich, products_stoich, reversible,
                                                        fwd_pre_exp_value, fwd_temp_exponent, fwd_act_energy,
                                                        three_body_efficiencies, default_efficiency,
                                                        flf_pre_exp_value, flf_temp_exponent, flf_act_energy,
                                                        troe_parameters, special_orders):
        self._add(4, reactants_stoich, products_stoich, reversible, fwd_pre_exp_value, fwd_temp_exponent,
                  fwd_act_energy, three_body_efficiencies, default_efficiency,
                  [flf_pre_exp_value, flf_temp_exponent, flf_act_energy], troe_parameters, special_orders)

    # ---------------- flamelet methods (griffon.pyx:556-679) ----------------
    def flamelet_stencils(self, dz, nzi, dissipationRate, invLewisNumbers, out_cmajor, out_csub, out_csup,
                          out_mcoeff, out_ncoeff):
        for a in (dz, dissipationRate, invLewisNumbers, out_cmajor, out_csub, out_csup, out_mcoeff, out_ncoeff):
            _f64(a)
        check(self._lib.gf_flamelet_stencils(self._h, dptr(dz), int(nzi), dptr(dissipationRate),
                                             dptr(invLewisNumbers), dptr(out_cmajor), dptr(out_csub),
                                             dptr(out_csup), dptr(out_mcoeff), dptr(out_ncoeff)))

    def flamelet_jac_indices(self, nzi, out_row_indices, out_col_indices):
        _i32(out_row_indices)
        _i32(out_col_indices)
        check(self._lib.gf_flamelet_jac_indices(self._h, int(nzi), iptr(out_row_indices), iptr(out_col_indices)))

    def flamelet_rhs(self, state, pressure, oxyState, fuelState, adiabatic, T_convection, T_radiation, h_convection,
                     h_radiation, nzi, cmajor, csub, csup, mcoeff, ncoeff, chi, include_enthalpy_flux,
                     include_variable_cp, use_scaled_heat_loss, out_rhs):
        for a in (state, oxyState, fuelState, T_convection, T_radiation, h_convection, h_radiation, cmajor, csub,
                  csup, mcoeff, ncoeff, chi, out_rhs):
            _f64(a)
        check(self._lib.gf_flamelet_rhs(self._h, dptr(state), float(pressure), dptr(oxyState), dptr(fuelState),
                                        int(bool(adiabatic)), dptr(T_convection), dptr(h_convection),
                                        dptr(T_radiation), dptr(h_radiation), int(nzi), dptr(cmajor), dptr(csub),
                                        dptr(csup), dptr(mcoeff), dptr(ncoeff), dptr(chi),
                                        int(bool(include_enthalpy_flux)), int(bool(include_variable_cp)),
                                        int(bool(use_scaled_heat_loss)), dptr(out_rhs)))

    def flamelet_jacobian(self, state, pressure, oxyState, fuelState, adiabatic, T_convection, T_radiation,
                          h_convection, h_radiation, nzi, cmajor, csub, csup, mcoeff, ncoeff, chi,
                          compute_eigenvalues, diffterm, scale_and_offset, prefactor, rates_sensitivity_option,
                          sensitivity_transform_option, include_enthalpy_flux, include_variable_cp,
                          use_scaled_heat_loss, out_expeig, out_jac):
        for a in (state, oxyState, fuelState, T_convection, T_radiation, h_convection, h_radiation, cmajor, csub,
                  csup, mcoeff, ncoeff, chi, out_expeig, out_jac):
            _f64(a)
        check(self._lib.gf_flamelet_jacobian(self._h, dptr(state), float(pressure), dptr(oxyState), dptr(fuelState),
                                             int(bool(adiabatic)), dptr(T_convection), dptr(h_convection),
                                             dptr(T_radiation), dptr(h_radiation), int(nzi), dptr(cmajor),
                                             dptr(csub), dptr(csup), dptr(mcoeff), dptr(ncoeff), dptr(chi),
                                             int(bool(compute_eigenvalues)), float(diffterm),
                                             int(bool(scale_and_offset)), float(prefactor),
                                             int(rates_sensitivity_option), int(sensitivity_transform_option),
                                             int(bool(include_enthalpy_flux)), int(bool(include_variable_cp)),
                                             int(bool(use_scaled_heat_loss)), dptr(out_expeig), dptr(out_jac)))

    # ---------------- kinetics and reactors (griffon.pyx:763-824) ----------------
    def production_rates(self, density, temperature, mass_fractions, prodrates):
        # forwarded positionally as (T, rho): callers pass temperature first (see module docstring)
        _f64(mass_fractions)
        _f64(prodrates)
        check(self._lib.gf_production_rates(self._h, float(density), float(temperature), dptr(mass_fractions),
                                            dptr(prodrates)))

    def prod_rates_primitive_sensitivities(self, density, temperature, mass_fractions, rates_sensitivity_option,
                                           prod_rates_sens):
        _f64(mass_fractions)
        _f64(prod_rates_sens)
        check(self._lib.gf_prod_rates_primitive_sensitivities(self._h, float(density), float(temperature),
                                                              dptr(mass_fractions), int(rates_sensitivity_option),
                                                              dptr(prod_rates_sens)))

    def reactor_rhs_isobaric(self, state, pressure, inflowTemperature, inflowY, tau, fluidTemperature,
                             surfTemperature, hConv, epsRad, surfaceAreaOverVolume, heatTransferOption, open,
                             rightHandSide):
        _f64(state)
        _f64(inflowY)
        _f64(rightHandSide)
        check(self._lib.gf_reactor_rhs_isobaric(self._h, dptr(state), float(pressure), float(inflowTemperature),
                                                dptr(inflowY), float(tau), float(fluidTemperature),
                                                float(surfTemperature), float(hConv), float(epsRad),
                                                float(surfaceAreaOverVolume), int(heatTransferOption),
                                                int(bool(open)), dptr(rightHandSide)))

    def reactor_jac_isobaric(self, state, pressure, inflowTemperature, inflowY, tau, fluidTemperature, surfTemperature,
                             hConv, epsRad, surfaceAreaOverVolume, heatTransferOption, open, rates_sensitivity_option,
                             sensitivity_transform_option, rightHandSide, jacobian):
        _f64(state)
        _f64(inflowY)
        _f64(rightHandSide)
        _f64(jacobian)
        check(self._lib.gf_reactor_jac_isobaric(self._h, dptr(state), float(pressure), float(inflowTemperature),
                                                dptr(inflowY), float(tau), float(fluidTemperature),
                                                float(surfTemperature), float(hConv), float(epsRad),
                                                float(surfaceAreaOverVolume), int(heatTransferOption),
                                                int(bool(open)), int(rates_sensitivity_option),
                                                int(sensitivity_transform_option), dptr(rightHandSide),
                                                dptr(jacobian)))

    def reactor_rhs_isochoric(self, state, inflowDensity, inflowTemperature, inflowY, tau, fluidTemperature,
                              surfTemperature, hConv, epsRad, surfaceAreaOverVolume, heatTransferOption, open,
                              rightHandSide):
        """state [rho, T, Y_0..Y_{ns-2}] (griffon.pyx:831-845)"""
        for a in (state, inflowY, rightHandSide):
            _f64(a)
        check(self._lib.gf_reactor_rhs_isochoric(self._h, dptr(state), float(inflowDensity), float(inflowTemperature),
                                                 dptr(inflowY), float(tau), float(fluidTemperature),
                                                 float(surfTemperature), float(hConv), float(epsRad),
                                                 float(surfaceAreaOverVolume), int(heatTransferOption),
                                                 int(bool(open)), dptr(rightHandSide)))

    def reactor_jac_isochoric(self, state, inflowDensity, inflowTemperature, inflowY, tau, fluidTemperature,
                              surfTemperature, hConv, epsRad, surfaceAreaOverVolume, heatTransferOption, open,
                              rates_sensitivity_option, rightHandSide, jacobian):
        """(ns+1) x (ns+1) column-major Jacobian in (rho, T, Y) (griffon.pyx:850-866)"""
        for a in (state, inflowY, rightHandSide, jacobian):
            _f64(a)
        check(self._lib.gf_reactor_jac_isochoric(self._h, dptr(state), float(inflowDensity), float(inflowTemperature),
                                                 dptr(inflowY), float(tau), float(fluidTemperature),
                                                 float(surfTemperature), float(hConv), float(epsRad),
                                                 float(surfaceAreaOverVolume), int(heatTransferOption),
                                                 int(bool(open)), int(rates_sensitivity_option), dptr(rightHandSide),
                                                 dptr(jacobian)))

    # ---------------- thermodynamics (griffon.pyx:684-760) ----------------
    def _scalar(self, fn, *args):
        out = np.zeros(1)
        check(fn(self._h, *args, dptr(out)))
        return float(out[0])

    def mixture_molecular_weight(self, y):
        return self._scalar(self._lib.gf_mixture_molecular_weight, dptr(_f64(y)))

    def mole_fractions(self, y, x):
        check(self._lib.gf_mole_fractions(self._h, dptr(_f64(y)), dptr(_f64(x))))

    def ideal_gas_density(self, p, T, y):
        return self._scalar(self._lib.gf_ideal_gas_density, float(p), float(T), dptr(_f64(y)))

    def ideal_gas_pressure(self, rho, T, y):
        return self._scalar(self._lib.gf_ideal_gas_pressure, float(rho), float(T), dptr(_f64(y)))

    def cp_mix(self, T, y):
        return self._scalar(self._lib.gf_cp_mix, float(T), dptr(_f64(y)))

    def cv_mix(self, T, y):
        return self._scalar(self._lib.gf_cv_mix, float(T), dptr(_f64(y)))

    def species_cp(self, T, cpspecies):
        check(self._lib.gf_species_cp(self._h, float(T), dptr(_f64(cpspecies))))

    def species_cv(self, T, cvspecies):
        check(self._lib.gf_species_cv(self._h, float(T), dptr(_f64(cvspecies))))

    def enthalpy_mix(self, T, y):
        return self._scalar(self._lib.gf_enthalpy_mix, float(T), dptr(_f64(y)))

    def energy_mix(self, T, y):
        return self._scalar(self._lib.gf_energy_mix, float(T), dptr(_f64(y)))

    def species_enthalpies(self, T, enthalpyspecies):
        check(self._lib.gf_species_enthalpies(self._h, float(T), dptr(_f64(enthalpyspecies))))

    def species_energies(self, T, energyspecies):
        check(self._lib.gf_species_energies(self._h, float(T), dptr(_f64(energyspecies))))

    def dcpdT_species(self, T, mass_fractions, dcpdT_spec):
        check(self._lib.gf_dcpdT_species(self._h, float(T), dptr(_f64(mass_fractions)), dptr(_f64(dcpdT_spec))))

    # ---------------- batched device entry points (torch tensors on the GPU) ----------------
    def batch_mixture_properties(self, Ty, out, p=None, rho=None, stream=None):
        """Ty: [n, ns+1] float64 CUDA tensor (T then the ns mass fractions); p, rho: None or [n]; out: [n, 8]
        receives {molecular weight, density from p, pressure from rho, cp, cv, h, e, dcp/dT} per state."""
        n = Ty.shape[0]
        if Ty.shape != (n, self._ns + 1) or out.shape != (n, 8) or not Ty.is_contiguous() or not out.is_contiguous():
            raise ValueError('Ty must be a contiguous [n, %d] tensor and out a contiguous [n, 8] tensor' % (self._ns + 1))
        for a in (p, rho):
            if a is not None and (a.shape != (n,) or not a.is_contiguous()):
                raise ValueError('p and rho must be contiguous [n] tensors')
        for a in (Ty, out, p, rho):
            if a is not None and (not a.is_cuda or str(a.dtype) != 'torch.float64'):
                raise ValueError('mixture property arrays must be float64 CUDA tensors')
        check(self._lib.gf_batch_mixture_properties(self._h, n, Ty.data_ptr(),
                                                    p.data_ptr() if p is not None else None,
                                                    rho.data_ptr() if rho is not None else None, out.data_ptr(),
                                                    _stream(stream)))

    def batch_reactor_rhs_isobaric(self, x, p, rhs, layout='soa', stream=None):
        """x, rhs: [ns, n] (layout 'soa') or [n, ns] ('aos') float64 CUDA tensors; p: [n] or [1]."""
        n, xs = _strides(x, self._ns, layout)
        _, rs = _strides(rhs, self._ns, layout)
        check(self._lib.gf_batch_reactor_rhs_isobaric(self._h, n, x.data_ptr(), xs[0], xs[1], p.data_ptr(),
                                                      1 if p.numel() > 1 else 0, rhs.data_ptr(), rs[0], rs[1],
                                                      _stream(stream)))

    def batch_reactor_general(self, x, rhs, jac=None, p=None, isochoric=False, inflow_density=0., inflow_T=0.,
                              inflow_y=None, tau=1., fluid_T=0., surf_T=0., h_conv=0., eps_rad=0.,
                              surface_area_over_volume=0., heat_transfer_option=0, open=False,
                              rates_sensitivity_option=0, stream=None):
        """Any reactor configuration over n states (gf_batch_reactor_general). x: [n, neq] float64 CUDA tensor
        ([T, Y..] isobaric with p: [n] or [1] pressures; [rho, T, Y..] isochoric); rhs: [n, neq]; jac: None or
        [n, neq*neq] (column-major blocks); inflow_y: [ns] CUDA tensor when open."""
        self._general(self._lib.gf_batch_reactor_general, x, rhs, jac, p, isochoric, inflow_density, inflow_T,
                      inflow_y, tau, fluid_T, surf_T, h_conv, eps_rad, surface_area_over_volume, heat_transfer_option,
                      open, rates_sensitivity_option, stream)

    def batch_reactor_general_noalloc(self, x, rhs, jac=None, p=None, isochoric=False, inflow_density=0.,
                                      inflow_T=0., inflow_y=None, tau=1., fluid_T=0., surf_T=0., h_conv=0.,
                                      eps_rad=0., surface_area_over_volume=0., heat_transfer_option=0, open=False,
                                      rates_sensitivity_option=0, stream=None):
        """batch_reactor_general on the library's fixed device scratch: no allocation, bounded n, calls that overlap
        on one GPU race (gf_batch_reactor_general_noalloc)."""
        self._general(self._lib.gf_batch_reactor_general_noalloc, x, rhs, jac, p, isochoric, inflow_density,
                      inflow_T, inflow_y, tau, fluid_T, surf_T, h_conv, eps_rad, surface_area_over_volume,
                      heat_transfer_option, open, rates_sensitivity_option, stream)

    def _general(self, fn, x, rhs, jac, p, isochoric, inflow_density, inflow_T, inflow_y, tau, fluid_T, surf_T,
                 h_conv, eps_rad, surface_area_over_volume, heat_transfer_option, open, rates_sensitivity_option,
                 stream):
        n, neq = self._general_shapes(x, rhs, jac, isochoric)
        if not isochoric and (p is None or not p.is_cuda):
            raise ValueError('isobaric reactors need a CUDA pressure tensor p')
        check(fn(
            self._h, n, int(bool(isochoric)), x.data_ptr(), neq, p.data_ptr() if p is not None else None,
            (1 if p.numel() > 1 else 0) if p is not None else 0, float(inflow_density), float(inflow_T),
            inflow_y.data_ptr() if inflow_y is not None else None, float(tau), float(fluid_T), float(surf_T),
            float(h_conv), float(eps_rad), float(surface_area_over_volume), int(heat_transfer_option),
            int(bool(open)), int(rates_sensitivity_option), rhs.data_ptr(), jac.data_ptr() if jac is not None else None,
            _stream(stream)))

    def _general_shapes(self, x, rhs, jac, isochoric):
        import torch
        neq = self._ns + (1 if isochoric else 0)
        for t in (x, rhs) + ((jac,) if jac is not None else ()):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or t.dim() != 2:
                raise ValueError('expected contiguous 2-D float64 CUDA tensors')
        n = x.shape[0]
        if x.shape[1] != neq or tuple(rhs.shape) != (n, neq) or (jac is not None and tuple(jac.shape) != (n, neq * neq)):
            raise ValueError('x and rhs must be [n, %d] and jac [n, %d]' % (neq, neq * neq))
        return n, neq

    def batch_reactor_jac_isobaric(self, x, p, rhs, jac, layout='soa', stream=None):
        """As batch_reactor_rhs_isobaric; jac: [n, ns*ns] float64 (column-major ns x ns block per state)."""
        n, xs = _strides(x, self._ns, layout)
        _, rs = _strides(rhs, self._ns, layout)
        if jac.dim() != 2 or jac.shape[0] != n or jac.shape[1] != self._ns * self._ns or not jac.is_contiguous():
            raise ValueError('jac must be a contiguous [n, ns*ns] tensor')
        check(self._lib.gf_batch_reactor_jac_isobaric(self._h, n, x.data_ptr(), xs[0], xs[1], p.data_ptr(),
                                                      1 if p.numel() > 1 else 0, rhs.data_ptr(), rs[0], rs[1],
                                                      jac.data_ptr(), self._ns * self._ns, _stream(stream)))


    def batch_reactor_jac_isobaric_terms(self, x, p, rhs, jac, aux, cpi, h, qa, stream=None):
        """As batch_reactor_jac_isobaric with [n, ns] ('aos') states, plus per state aux [n, 4] = {rho, cp,
        dcp/dT, mixture molecular weight}, cpi and h [n, ns] (species heat capacities and enthalpies per unit
        mass) and qa [n, nr] = d(rate of progress)/d(ln A) per reaction in griffon order."""
        n, xs = _strides(x, self._ns, 'aos')
        for t, w in ((aux, 4), (cpi, self._ns), (h, self._ns), (qa, None)):
            if t.dim() != 2 or t.shape[0] != n or (w is not None and t.shape[1] != w) or not t.is_contiguous():
                raise ValueError('per-state outputs must be contiguous [n, k] tensors')
        check(self._lib.gf_batch_reactor_jac_isobaric_terms(self._h, n, x.data_ptr(), xs[0], xs[1], p.data_ptr(),
                                                            1 if p.numel() > 1 else 0, rhs.data_ptr(), self._ns, 1,
                                                            jac.data_ptr(), self._ns * self._ns, aux.data_ptr(),
                                                            cpi.data_ptr(), h.data_ptr(), qa.data_ptr(),
                                                            _stream(stream)))


def _strides(t, ns, layout):
    import torch
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or t.dim() != 2:
        raise ValueError('expected a contiguous 2-D float64 CUDA tensor')
    if layout == 'soa':
        if t.shape[0] != ns:
            raise ValueError('soa tensors are [ns, n]')
        return t.shape[1], (1, t.shape[1])
    if t.shape[1] != ns:
        raise ValueError('aos tensors are [n, ns]')
    return t.shape[0], (ns, 1)


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


# ---------------- block-tridiagonal, diagonal off-diagonal (griffon.pyx:1009-1113) ----------------
def py_btddod_full_factorize(out_d_factors, num_blocks, block_size, out_l_values, out_d_pivots):
    _f64(out_d_factors)
    _f64(out_l_values)
    _i32(out_d_pivots)
    check(lib().gf_btddod_full_factorize(dptr(out_d_factors), int(num_blocks), int(block_size), dptr(out_l_values),
                                         iptr(out_d_pivots)))


def py_btddod_full_solve(d_factors, l_values, d_pivots, rhs, num_blocks, block_size, out_solution):
    _f64(d_factors)
    _f64(l_values)
    _i32(d_pivots)
    _f64(rhs)
    _f64(out_solution)
    check(lib().gf_btddod_full_solve(dptr(d_factors), dptr(l_values), iptr(d_pivots), dptr(rhs), int(num_blocks),
                                     int(block_size), dptr(out_solution)))


def py_btddod_scale_and_add_diagonal(in_out_matrix_values, matrix_scale, diagonal, diag_scale, num_blocks,
                                     block_size):
    _f64(in_out_matrix_values)
    _f64(diagonal)
    check(lib().gf_btddod_scale_and_add_diagonal(dptr(in_out_matrix_values), float(matrix_scale), dptr(diagonal),
                                                 float(diag_scale), int(num_blocks), int(block_size)))
