// Reactor configurations beyond the closed adiabatic isobaric fast path, and griffon's thermodynamic getters.
//
// reactor_ext_kernel assembles, per state, the RHS and Jacobian of
//   - the isobaric reactor with open (mass exchange) and diathermal (convection + radiation) terms, isothermal
//     operation, and the (rho, T, Y) -> (T, Y) transform (isobaric_reactor_kernels.cpp:19-343);
//   - the isochoric reactor in (rho, T, Y) with the same options (isochoric_reactor_kernels.cpp:18-333);
// from the chemistry kernel's per-state production rates, raw option-0/1 sensitivities (ns+1)^2, mixture scalars
// {rho, cp, dcp/dT, M} and species cp and h. One 64-thread workgroup per state; thread t owns primitive-Jacobian
// columns t, t+64, ...; inner products run serially in the reference's summation order.
//
// thermo_getters_kernel evaluates the getters of griffon.pyx:684-758 for one (T, y): mixture molecular weight, mole
// fractions, ideal-gas density and pressure, cp/cv/h/e of the mixture and of the species, and dcp/dT of the species
// (thermodynamics_kernels.cpp:18-44, 131-181, 183-260, 262-383). Species values come from one lane each; mixture sums
// are taken serially by lane 0 in the reference's order.
#include "gf_internal.h"
#include "gf_thermo.h"
#include "../../include/griffon_amd.h"

#include <mutex>

namespace gfa {

namespace {

constexpr double kStefanBoltzmann = 5.67e-8;  // the constant griffon uses (isobaric_reactor_kernels.cpp:36)

__device__ inline double dotp(int n, const double *a, const double *b) {
  double d = 0.;
  for (int i = 0; i < n; ++i) d += a[i] * b[i];
  return d;
}

template <bool ISOCHORIC>
__global__ void __launch_bounds__(64) reactor_ext_kernel(MechView mv, ReactorExtIO a) {
  const int ns = mv.ns, nm = ns - 1;
  const int np = ns + 1;                   // columns of the primitive Jacobian (rho, T, Y_0..Y_{ns-2})
  const int n = ISOCHORIC ? ns + 1 : ns;   // equations
  const long s = blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double y[NS_MAX], ue[NS_MAX], cvi[NS_MAX], w[NS_MAX], uin[NS_MAX], yin[NS_MAX], extra[8];
  extern __shared__ double pj[];  // isobaric: ns x (ns+1) primitive Jacobian; isochoric: unused

  const double *st = a.state + s * a.state_stride;
  const double *ws = a.sens + s * (long)np * np;
  const double rho = a.aux[s * 4 + 0], cp = a.aux[s * 4 + 1], cpsT = a.aux[s * 4 + 2], mmw = a.aux[s * 4 + 3];
  const double T = ISOCHORIC ? st[1] : st[0];
  const double *ystate = ISOCHORIC ? st + 2 : st + 1;
  const double Ru = mv.Ru;
  if (t < ns) {
    // species heat capacity (cv for the isochoric reactor) and enthalpy (energy), at T and at the inflow temperature
    const double c = a.cpi[s * ns + t], h = a.hsp[s * ns + t];
    cvi[t] = ISOCHORIC ? c - Ru * mv.invmw[t] : c;
    ue[t] = ISOCHORIC ? h - Ru * T * mv.invmw[t] : h;
    w[t] = a.w[s * ns + t];
    yin[t] = a.open ? a.yin[t] : 0.;
    uin[t] = 0.;
    if (a.open) {
      const double Ti = a.Tin;
      double c2, cs2, h2, g2, d2;
      species_thermo(mv, t, Ti, log(Ti), 1. / Ti, c2, cs2, h2, g2, d2);
      uin[t] = ISOCHORIC ? h2 - Ru * Ti * mv.invmw[t] : h2;
    }
  }
  if (t == 0) {  // y_{ns-1} = 1 - sum of the others, in extract_y's order (combustion_kernels.h:526-530)
    double last = 1.;
    for (int j = 0; j < nm; ++j) {
      y[j] = ystate[j];
      last -= y[j];
    }
    y[nm] = last;
  }
  __syncthreads();

  const double cm = ISOCHORIC ? cp - Ru / mmw : cp;  // cv or cp of the mixture
  const double iR = 1. / rho, iC = 1. / cm, iRC = 1. / (rho * cm);
  if (t == 0) {
    // chemistry RHS of the energy equation, then the mass-exchange and heat-transfer terms
    const double r_e = -dotp(ns, ISOCHORIC ? ue : w, ISOCHORIC ? w : ue) / (rho * cm);
    double m_e = 0., m_rho = 0.;
    if (a.open) {
      m_e = (uin[nm] - ue[nm]) * yin[nm];
      for (int i = 0; i < nm; ++i) m_e += (uin[i] - ue[i]) * yin[i];
      if (ISOCHORIC) {
        m_rho = (a.rho_in - rho) / a.tau;
        m_e = m_e / cm * (1. / a.tau * a.rho_in * iR);
      } else {
        m_e = m_e / cm * (1. / a.tau);
      }
    }
    double q = 0.;
    if (a.heat == 2) q = a.sov / (rho * cm) * (a.hc * (a.Tf - T) + a.eps * kStefanBoltzmann * (a.Ts * a.Ts * a.Ts * a.Ts - T * T * T * T));
    extra[0] = r_e;  // chemistry part (enters the Jacobian's energy row)
    extra[1] = m_e;
    extra[2] = m_rho;
    extra[3] = q;
  }
  __syncthreads();
  const double r_e = extra[0], m_e = extra[1], m_rho = extra[2], q = extra[3];
  const double it = a.open ? 1. / a.tau : 0.;
  const double mscale = ISOCHORIC ? it * a.rho_in * iR : it;  // factor of the species mass-exchange terms

  // RHS
  if (a.rhs) {
    double *out = a.rhs + s * n;
    const int e0 = ISOCHORIC ? 1 : 0;  // energy equation index
    for (int i = t; i < n; i += 64) {
      double v;
      if (ISOCHORIC && i == 0) {
        v = 0. + (a.open ? m_rho : 0.);
      } else if (i == e0) {
        v = (a.heat == 1) ? 0. : r_e + (a.open ? m_e : 0.) + q;
      } else {
        const int k = i - e0 - 1;
        v = w[k] * iR;
        if (a.open) v += (yin[k] - y[k]) * mscale;
      }
      out[i] = v;
    }
  }
  if (!a.jac) return;

  if (!ISOCHORIC) {
    // primitive Jacobian n x (n+1) column-major (isobaric_reactor_kernels.cpp:58-168), then the transform
    for (int c = t; c < np; c += 64) {
      double *col = pj + (long)c * ns;
      const double *wc = ws + (long)c * np;
      if (c == 0) {
        col[0] = -iRC * dotp(ns, wc, ue) - iR * r_e;
        for (int i = 0; i < nm; ++i) col[1 + i] = iR * (wc[i] - iR * w[i]);
        if (a.heat == 2) col[0] += -q / rho;
      } else if (c == 1) {
        col[0] = -iRC * (dotp(ns, wc, ue) + dotp(ns, w, cvi)) - r_e * cpsT * iC;
        for (int i = 0; i < nm; ++i) col[1 + i] = wc[i] * iR;
        if (a.open) col[0] += -iC * (cpsT * m_e + it * dotp(ns, yin, cvi));
        if (a.heat == 2)
          col[0] += -iC * cpsT * q - a.sov * iRC * (a.hc + 4. * a.eps * kStefanBoltzmann * T * T * T);
      } else {
        const int k = c - 2;
        col[0] = -iRC * dotp(ns, wc, ue) - r_e * (cvi[k] - cvi[nm]) * iC;
        for (int i = 0; i < nm; ++i) col[1 + i] = iR * wc[i];
        if (a.open) {
          col[0] += -m_e * (cvi[k] - cvi[nm]) * iC;
          col[1 + k] += -it;
        }
        if (a.heat == 2) col[0] += iC * q * (cvi[nm] - cvi[k]);
      }
      if (a.heat == 1) col[0] = 0.;
    }
    __syncthreads();
    // (rho, T, Y) -> (T, Y) at constant pressure (isobaric_reactor_kernels.cpp:319-343)
    double *J = a.jac + s * (long)ns * ns;
    const double roT = rho / T, nrm = -rho * mmw;
    for (int e = t; e < ns * ns; e += 64) {
      const int c = e / ns, i = e - c * ns;
      J[e] = (c == 0) ? pj[ns + i] - roT * pj[i] : pj[(1 + c) * ns + i] + nrm * mv.dinvm[c - 1] * pj[i];
    }
    return;
  }

  // isochoric: (ns+1) x (ns+1) column-major in (rho, T, Y) (isochoric_reactor_kernels.cpp:60-333)
  double *J = a.jac + s * (long)np * np;
  for (int c = t; c < np; c += 64) {
    double *col = J + (long)c * np;
    const double *wc = ws + (long)c * np;
    col[0] = 0.;
    if (c == 0) {
      col[1] = -iRC * dotp(ns, ue, wc) - iR * r_e;
      for (int i = 0; i < nm; ++i) col[2 + i] = iR * (wc[i] - iR * w[i]);
      if (a.open) {
        col[0] = -it;
        col[1] += -iR * m_e;
        for (int i = 0; i < nm; ++i) col[2 + i] += -iR * ((yin[i] - y[i]) * mscale);
      }
      if (a.heat == 2) col[1] += -q / rho;
    } else if (c == 1) {
      col[1] = -iC * (iR * (dotp(ns, ue, wc) + dotp(ns, cvi, w)) + r_e * cpsT);
      for (int i = 0; i < nm; ++i) col[2 + i] = iR * wc[i];
      if (a.open) col[1] += -iC * (it * iR * (a.rho_in * dotp(ns, cvi, yin)) + cpsT * m_e);
      if (a.heat == 2) col[1] += -iC * cpsT * q - a.sov * iRC * (a.hc + 4. * a.eps * kStefanBoltzmann * T * T * T);
    } else {
      const int k = c - 2;
      col[1] = -iRC * dotp(ns, ue, wc) - r_e * (cvi[k] - cvi[nm]) * iC;
      for (int i = 0; i < nm; ++i) col[2 + i] = iR * wc[i];
      if (a.open) {
        col[1] += -iC * m_e * (cvi[k] - cvi[nm]);
        col[2 + k] += -it * a.rho_in / rho;
      }
      if (a.heat == 2) col[1] += iC * q * (cvi[nm] - cvi[k]);
    }
    if (a.heat == 1) col[1] = 0.;
  }
}

// one (T, y) per launch, one 64-thread workgroup: out = {mmw, density(p), pressure(rho), cp, cv, h, e, dcp/dT,
// x[ns], cp[ns], cv[ns], h[ns], e[ns], dcp/dT[ns]}
// One state per workgroup: state b reads in[b * in_stride] = {T, y_0 .. y_{ns-1}} and, when given, p[b] and rho[b]
// (else the scalars p0 and rho0). Mixture values go to out[b * out_stride + 0..7]; species values to out + 8 .. 8 + 6 ns
// when with_species is set.
__global__ void __launch_bounds__(64) thermo_getters_kernel(MechView mv, const double *in, long in_stride,
                                                            const double *pv, const double *rhov, double p0,
                                                            double rho0, double *out, long out_stride,
                                                            int with_species) {
  const int ns = mv.ns;
  const int t = threadIdx.x;
  const long b = blockIdx.x;
  in += b * in_stride;
  out += b * out_stride;
  const double T = in[0];
  const double *y = in + 1;
  const double p = pv ? pv[b] : p0, rho = rhov ? rhov[b] : rho0;
  const double Ru = mv.Ru;
  __shared__ double mm, scp[64], sh[64], se[64], scs[64];
  if (t == 0) {  // 1 / (sum y_i / M_i), combustion_kernels.h:381-387
    double d = 0.;
    for (int i = 0; i < ns; ++i) d += y[i] * mv.invmw[i];
    mm = 1. / d;
  }
  __syncthreads();
  if (t < ns) {
    double c, cs, h, g, d;
    species_thermo(mv, t, T, log(T), 1. / T, c, cs, h, g, d);
    // cp_sens_T: zero outside [Tmin, Tmax] and for constant cp (thermodynamics_kernels.cpp:183-260)
    scp[t] = c;
    sh[t] = h;
    se[t] = h - Ru * T * mv.invmw[t];
    scs[t] = cs;
    if (with_species) {
      double *x = out + 8, *cpo = x + ns, *cvo = cpo + ns, *ho = cvo + ns, *eo = ho + ns, *cso = eo + ns;
      x[t] = y[t] * mm * mv.invmw[t];
      cpo[t] = c;
      cvo[t] = c - Ru * mv.invmw[t];
      ho[t] = h;
      eo[t] = se[t];
      cso[t] = cs;
    }
  }
  __syncthreads();
  if (t == 0) {
    double cp = 0., h = 0., e = 0., cs = 0.;
    for (int i = 0; i < ns; ++i) {
      cp += y[i] * scp[i];
      h += sh[i] * y[i];
      e += se[i] * y[i];
      // a constant-cp species resets the accumulated derivative (thermodynamics_kernels.cpp:200-203)
      cs = mv.cpkind[i] == CP_CONST ? 0. : cs + y[i] * scs[i];
    }
    out[0] = mm;
    out[1] = p * mm / (T * Ru);
    out[2] = rho * T * Ru / mm;
    out[3] = cp;
    out[4] = cp - Ru / mm;
    out[5] = h;
    out[6] = e;
    out[7] = cs;
  }
}

}  // namespace

int launch_reactor_ext(const gf_mech *m, const ReactorExtIO &a, hipStream_t st) {
  if (a.n <= 0) return GF_OK;
  const int ns = m->view.ns;
  if (a.isochoric) {
    hipLaunchKernelGGL(reactor_ext_kernel<true>, dim3((unsigned)a.n), dim3(64), 0, st, m->view, a);
  } else {
    const size_t lds = (size_t)ns * (ns + 1) * sizeof(double);
    hipLaunchKernelGGL(reactor_ext_kernel<false>, dim3((unsigned)a.n), dim3(64), lds, st, m->view, a);
  }
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "reactor kernel launch failed");
  return GF_OK;
}

int launch_thermo_getters(const gf_mech *m, const double *in, double p, double rho, double *out, hipStream_t st) {
  hipLaunchKernelGGL(thermo_getters_kernel, dim3(1), dim3(64), 0, st, m->view, in, 0L, nullptr, nullptr, p, rho, out,
                     0L, 1);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "thermodynamics kernel launch failed");
  return GF_OK;
}

int launch_mixture_properties(const gf_mech *m, int n, const double *Ty, const double *p, const double *rho,
                              double *out, hipStream_t st) {
  if (n <= 0) return GF_OK;
  hipLaunchKernelGGL(thermo_getters_kernel, dim3((unsigned)n), dim3(64), 0, st, m->view, Ty, (long)m->view.ns + 1,
                     p, rho, 0., 0., out, 8L, 0);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "mixture properties kernel launch failed");
  return GF_OK;
}

// ---------------- allocation-free variant of gf_batch_reactor_general ----------------
// Its workspace is this file-scope device array: each GPU holds its own copy for the life of the process. Host code
// never passes `scratch` itself to a kernel: its device address on the current GPU comes from hipGetSymbolAddress,
// looked up once per device on the first call and cached.
namespace {
constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per device
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kScratchDoubles];
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int scratch_on_current_device(double **ptr) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(GF_ERR_HIP, "no HIP device available");
  if (dev < 0 || dev >= kMaxDevices) return fail(GF_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GF_ERR_HIP, std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return GF_OK;
}
}  // namespace

}  // namespace gfa

extern "C" int gf_batch_reactor_general_noalloc(gf_mech *m, int n, int isochoric, const double *x, long x_item_stride,
                                                const double *p, long p_stride, double inflow_density,
                                                double inflow_T, const double *inflow_y, double tau, double fluid_T,
                                                double surf_T, double h_conv, double eps_rad, double sov, int heat,
                                                int open, int rates_opt, double *rhs, double *jac, void *stream) {
  using namespace gfa;
  const int rc = check_batch_general(m, n, isochoric, x, x_item_stride, p, inflow_y, heat, open, rates_opt, rhs);
  if (rc != GF_OK) return rc;
  if (n == 0) return GF_OK;
  if (!m->finalized || !m->dblob)
    return fail(GF_ERR_STATE, "the mechanism is not on the device yet; make one call through another entry point "
                              "first (this entry point allocates nothing, the mechanism tables included)");
  const size_t per_state = batch_work_doubles(m->view.ns, jac != nullptr);
  const size_t max_n = kScratchDoubles / per_state;
  if ((size_t)n > max_n)
    return fail(GF_ERR_ARG, "n = " + std::to_string(n) + " exceeds the allocation-free scratch (" +
                                std::to_string(max_n) + " states of this mechanism " +
                                (jac ? "with" : "without") + " Jacobians fit in 32 MiB)");
  double *work = nullptr;
  const int r2 = scratch_on_current_device(&work);
  if (r2 != GF_OK) return r2;
  return batch_general(m, n, isochoric, x, x_item_stride, p, p_stride, inflow_density, inflow_T, inflow_y, tau,
                       fluid_T, surf_T, h_conv, eps_rad, sov, heat, open, rates_opt, rhs, jac, work, 1,
                       static_cast<hipStream_t>(stream));
}

