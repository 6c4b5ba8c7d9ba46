// Flamelet terms around the batched chemistry kernel: mixture-fraction stencils, the flamelet RHS
// (heat loss, enthalpy flux, variable cp, diffusion) and the flamelet Jacobian epilogue (heat-loss rows,
// diffusion diagonal, enthalpy-flux correction, sub/super diagonals, scale-and-offset), all FP64.
//
// Restated from flamelet_kernels.cpp:31-90 (stencils, Jacobian sparsity), 1039-1218 (RHS) and 1220-1409
// (Jacobian). As in the reference, the dissipation rate of interior point i is read as chi[i] in the
// enthalpy-flux and variable-cp terms while the stencils use chi[1+i].
//
// Per-point chemistry (production rates, cp, species cp, density) comes from the chemistry kernel run
// over nzi+2 points laid out [oxidizer, interior 0..nzi-1, fuel], so point i of the interior is entry i+1.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"

namespace gfa {

namespace {

__global__ void stencils_kernel(int ns, const double *dz, int nzi, const double *chi, const double *invLe,
                                double *cmaj, double *csub, double *csup, double *mco, double *nco) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nzi * ns) return;
  const int i = e / ns, l = e % ns;
  const double dzt = dz[i] + dz[i + 1];
  cmaj[e] = -chi[1 + i] / (dz[i] * dz[i + 1]) * invLe[l];
  csub[e] = chi[1 + i] / (dzt * dz[i]) * invLe[l];
  csup[e] = chi[1 + i] / (dzt * dz[i + 1]) * invLe[l];
  if (l == 0) {
    nco[i] = 1 / (dz[i] + dz[i + 1]);
    mco[i] = -nco[i];
  }
}

__global__ void jac_indices_kernel(int ns, int nzi, int *rows, int *cols) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nbl = (long)nzi * ns * ns;
  const long noff = (long)(nzi - 1) * ns;
  if (e < nbl) {
    const int iz = (int)(e / (ns * ns));
    const int r = (int)(e % (ns * ns));
    const int iq = r / ns, jq = r % ns;  // block entry (row jq, column iq), column-major
    rows[e] = iz * ns + jq;
    cols[e] = iz * ns + iq;
  } else if (e < nbl + noff) {
    const long q = e - nbl;  // subdiagonal, block rows 1..nzi-1
    const int idx = (int)(q + ns);
    rows[e] = idx;
    cols[e] = idx - ns;
  } else if (e < nbl + 2 * noff) {
    const long q = e - nbl - noff;  // superdiagonal, block rows 0..nzi-2
    rows[e] = (int)q;
    cols[e] = (int)q + ns;
  }
}

__device__ __forceinline__ double max_interior_T(const FlameletIO &f) {
  double m = -1.;
  for (int i = 0; i < f.nzi; ++i) m = fmax(m, f.state[(long)i * f.ns]);
  return m;
}

// one thread per interior (point, variable)
__global__ void rhs_epilogue_kernel(FlameletIO f, double *out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ns = f.ns, nzi = f.nzi;
  if (e >= (long)nzi * ns) return;
  const int i = (int)(e / ns), l = (int)(e % ns);
  const long pt = i + 1;  // chemistry point index
  const double *sm = (i == 0) ? f.oxy : f.state + (long)(i - 1) * ns;
  const double *sp = (i == nzi - 1) ? f.fuel : f.state + (long)(i + 1) * ns;
  double v = f.chem_rhs[pt * ns + l];
  if (l == 0) {
    const double T = f.state[(long)i * ns];
    const double rho = f.aux[pt * 4 + 0], cp = f.aux[pt * 4 + 1];
    if (!f.adiabatic) {  // flamelet_kernels.cpp:1121-1138
      const double hc = f.hc[i], hr = f.hr[i], Tc = f.Tc[i], Tr = f.Tr[i];
      double q;
      if (f.scaled_hl) {
        const double maxT = max_interior_T(f);
        const double maxT4 = maxT * maxT * maxT * maxT;
        const double Tr4 = Tr * Tr * Tr * Tr;
        q = hc * (Tc - T) / (maxT - Tc) + hr * 5.67e-8 * (Tr4 - T * T * T * T) / (maxT4 - Tr4);
      } else {
        q = hc * (Tc - T) + hr * 5.67e-8 * (Tr * Tr * Tr * Tr - T * T * T * T);
      }
      v += q / (rho * cp);
    }
    const double dTdZ = f.mco[i] * sm[0] + f.nco[i] * sp[0];
    double cpz = 0.;
    if (f.incl_vcp) {  // flamelet_kernels.cpp:1062-1092 (boundary points use the stream states' cp)
      cpz = f.mco[i] * f.aux[(pt - 1) * 4 + 1] + f.nco[i] * f.aux[(pt + 1) * 4 + 1];
    }
    if (f.incl_ef) {  // flamelet_kernels.cpp:1140-1180
      const double *cpi = f.cpi + pt * ns;
      const double cpn = cpi[ns - 1];
      double dY = 0.;
      for (int j = 0; j < ns - 1; ++j) dY += (cpi[j] - cpn) * (f.mco[i] * sm[1 + j] + f.nco[i] * sp[1 + j]);
      v += 0.5 * f.chi[i] / cp * dTdZ * dY;
      if (f.incl_vcp) v += 0.5 * f.chi[i] * cpz / cp * dTdZ;
    }
    if (f.incl_vcp && !f.incl_ef) v += 0.5 * f.chi[i] * cpz / cp * dTdZ;
  }
  // diffusion stencil (flamelet_kernels.cpp:1207-1217)
  v += f.cmaj[e] * f.state[e] + f.csub[e] * sm[l] + f.csup[e] * sp[l];
  out[e] = v;
}

// One thread per entry of the full BTDDOD array: nzi blocks, then (nzi-1)*ns sub, then (nzi-1)*ns sup.
// The chemistry kernel has already written each point's transformed isobaric block.
// phase 0: every term; 1: only the heat-loss terms (which the reference adds before its eigenvalue
// computation); 2: every term but the heat-loss ones
__global__ void jac_epilogue_kernel(MechView mv, FlameletIO f, int scale_and_offset, double prefactor, double *jac,
                                    int phase) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ns = f.ns, nzi = f.nzi;
  const long bsz = (long)ns * ns;
  const long nbl = (long)nzi * bsz;
  const long noff = (long)(nzi - 1) * ns;
  if (e >= nbl + 2 * noff) return;
  double v = jac[e];
  bool diag = false;
  if (e < nbl) {
    const int iz = (int)(e / bsz);
    const int r = (int)(e % bsz);
    const int col = r / ns, row = r % ns;
    const long pt = iz + 1;
    if (row == 0 && !f.adiabatic && phase != 2) {  // flamelet_kernels.cpp:1290-1320 on the primitive Jacobian, then transformed
      const double T = f.state[(long)iz * ns];
      const double rho = f.aux[pt * 4 + 0], cp = f.aux[pt * 4 + 1], cps = f.aux[pt * 4 + 2],
                   mmw = f.aux[pt * 4 + 3];
      const double *cpi = f.cpi + pt * ns;
      const double iRC = 1. / (rho * cp), iC = 1. / cp;
      const double Tc = f.Tc[iz], Tr = f.Tr[iz], hcv = f.hc[iz], hrd = f.hr[iz];
      double q, dT;
      if (f.scaled_hl) {
        const double maxT = max_interior_T(f);
        const double maxT4 = maxT * maxT * maxT * maxT;
        const double Tr4 = Tr * Tr * Tr * Tr;
        q = (hcv * (Tc - T) / (maxT - Tc) + hrd * 5.67e-8 * (Tr4 - T * T * T * T) / (maxT4 - Tr4)) * iRC;
        dT = -(iC * cps * q + iRC * (hcv / (maxT - Tc) + 4. * hrd / (maxT4 - Tr4) * 5.67e-8 * T * T * T));
      } else {
        q = (hcv * (Tc - T) + hrd * 5.67e-8 * (Tr * Tr * Tr * Tr - T * T * T * T)) * iRC;
        dT = -(iC * cps * q + iRC * (hcv + 4. * hrd * 5.67e-8 * T * T * T));
      }
      const double drho = -q / rho;
      if (col == 0) {
        v += dT - rho / T * drho;
      } else {
        const int k = col - 1;
        v += iC * q * (cpi[ns - 1] - cpi[k]) + (-rho * mmw * mv.dinvm[k]) * drho;
      }
    }
    if (phase == 1) {
      jac[e] = v;
      return;
    }
    diag = (row == col);
    if (diag) v += f.cmaj[(long)iz * ns + row];  // flamelet_kernels.cpp:1343-1346
    if (row == 0 && col == 0 && f.incl_ef) {      // flamelet_kernels.cpp:1348-1381
      const double cp = f.aux[pt * 4 + 1], cps = f.aux[pt * 4 + 2];
      const double *sm = (iz == 0) ? f.oxy : f.state + (long)(iz - 1) * ns;
      const double *sp = (iz == nzi - 1) ? f.fuel : f.state + (long)(iz + 1) * ns;
      const double dTdZ = f.mco[iz] * sm[0] + f.nco[iz] * sp[0];
      const double dcpdZ = f.mco[iz] * f.aux[(pt - 1) * 4 + 1] + f.nco[iz] * f.aux[(pt + 1) * 4 + 1];
      const double f1 = 0.5 * f.chi[iz] / cp * dTdZ * dcpdZ;
      v -= f1 / cp * cps;
    }
  } else if (phase == 1) {
    return;
  } else if (e < nbl + noff) {
    const long q = e - nbl;  // subdiagonal entry (iz-1)*ns + iq gets csub[iz*ns + iq] (flamelet_kernels.cpp:1385-1394)
    v += f.csub[q + ns];
  } else {
    const long q = e - nbl - noff;  // superdiagonal entry (iz-1)*ns + iq gets csup[(iz-1)*ns + iq]
    v += f.csup[q];
  }
  if (scale_and_offset) {  // flamelet_kernels.cpp:1395-1408
    v *= prefactor;
    if (diag) v -= 1.;
  }
  jac[e] = v;
}

__global__ void scale_add_diag_kernel(double *val, double a, const double *dg, double b, int nb, int bs) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long bsz = (long)bs * bs;
  const long nel = (long)bs * (nb * bs + 2 * (nb - 1));
  if (e >= nel) return;
  double v = val[e] * a;
  if (e < (long)nb * bsz) {
    const int iz = (int)(e / bsz);
    const int r = (int)(e % bsz);
    if (r / bs == r % bs) v += b * dg[(long)iz * bs + r % bs];
  }
  val[e] = v;
}

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int launch_flamelet_stencils(int ns, const double *dz, int nzi, const double *chi, const double *invLe, double *cmaj,
                             double *csub, double *csup, double *mco, double *nco, hipStream_t st) {
  if (nzi <= 0) return GF_OK;
  hipLaunchKernelGGL(stencils_kernel, dim3(nblk((long)nzi * ns, 256)), dim3(256), 0, st, ns, dz, nzi, chi, invLe, cmaj,
                     csub, csup, mco, nco);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stencil kernel launch failed");
}

int launch_flamelet_jac_indices(int ns, int nzi, int *rows, int *cols, hipStream_t st) {
  if (nzi <= 0) return GF_OK;
  const long nel = (long)ns * (nzi * ns + 2 * (nzi - 1));
  hipLaunchKernelGGL(jac_indices_kernel, dim3(nblk(nel, 256)), dim3(256), 0, st, ns, nzi, rows, cols);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "jac index kernel launch failed");
}

int launch_flamelet_rhs_epilogue(const FlameletIO &f, double *out, hipStream_t st) {
  hipLaunchKernelGGL(rhs_epilogue_kernel, dim3(nblk((long)f.nzi * f.ns, 256)), dim3(256), 0, st, f, out);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "flamelet rhs kernel launch failed");
}

int launch_flamelet_jac_epilogue(const gf_mech *m, const FlameletIO &f, int scale_and_offset, double prefactor,
                                 double *jac, int phase, hipStream_t st) {
  const long nel = (long)f.ns * (f.nzi * f.ns + 2 * (f.nzi - 1));
  hipLaunchKernelGGL(jac_epilogue_kernel, dim3(nblk(nel, 256)), dim3(256), 0, st, m->view, f, scale_and_offset,
                     prefactor, jac, phase);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "flamelet jacobian kernel launch failed");
}

int launch_scale_and_add_diagonal(double *v, double a, const double *dg, double b, int nb, int bs, hipStream_t st) {
  const long nel = (long)bs * (nb * bs + 2 * (nb - 1));
  hipLaunchKernelGGL(scale_add_diag_kernel, dim3(nblk(nel, 256)), dim3(256), 0, st, v, a, dg, b, nb, bs);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "scale_and_add_diagonal launch failed");
}

}  // namespace gfa
