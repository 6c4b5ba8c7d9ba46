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

// flamelet RHS entry e = (interior point, variable)
__device__ void rhs_entry(const FlameletIO &f, long e, double *out) {
  const int ns = f.ns, nzi = f.nzi;
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

// one thread per interior (point, variable)
__global__ void rhs_epilogue_kernel(FlameletIO f, double *out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)f.nzi * f.ns) return;
  rhs_entry(f, e, out);
}

// Entry e of the flamelet Jacobian (BTDDOD layout: nzi blocks, then (nzi-1)*ns sub, then (nzi-1)*ns sup) from v, the value the chemistry kernel left there (the off-diagonal entries
// accumulate onto v). Returns false when the entry is left unchanged (phase 1 outside the heat-loss rows).
// diag_row is the row of a block-diagonal entry on the diagonal, else -1.
// phase 0: every term; 1: only the heat-loss terms (which the reference adds before its eigenvalue
// computation); 2: every term but the heat-loss ones
__device__ bool jac_entry(const MechView &mv, const FlameletIO &f, long e, int scale_and_offset, double prefactor,
                          int phase, double &v, int &diag_row) {
  const int ns = f.ns, nzi = f.nzi;
  const long bsz = (long)ns * ns;
  const long nbl = (long)nzi * bsz;
  const long noff = (long)(nzi - 1) * ns;
  bool diag = false;
  diag_row = -1;
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
    if (phase == 1) return true;
    diag = (row == col);
    if (diag) diag_row = iz * ns + row;
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
    return false;
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
  return true;
}

// One thread per entry of the full BTDDOD array: nzi blocks, then (nzi-1)*ns sub, then (nzi-1)*ns sup.
// The chemistry kernel has already written each point's transformed isobaric block.
__global__ void jac_epilogue_kernel(MechView mv, FlameletIO f, int scale_and_offset, double prefactor, double *jac,
                                    int phase) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)f.ns * (f.nzi * f.ns + 2 * (f.nzi - 1))) return;
  double v = jac[e];
  int dr;
  if (jac_entry(mv, f, e, scale_and_offset, prefactor, phase, v, dr)) jac[e] = v;
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

// ---------------- many flamelets at once ----------------
// flamelet owning global index e: the last f with off[f] <= e
__device__ __forceinline__ int fb_find(const long *off, int B, long e) {
  int lo = 0, hi = B - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ FlameletIO fb_io(const FlameletBatchView &b, int f, const double *X) {
  FlameletIO io;
  const int nzi = b.nzi[f], ns = b.ns;
  const long p0 = b.pt0[f], d0 = b.dof0[f], i0 = b.ipt0[f];
  io.nzi = nzi;
  io.ns = ns;
  io.state = X + (p0 + 1) * ns;
  io.oxy = X + p0 * ns;
  io.fuel = X + (p0 + nzi + 1) * ns;
  io.cmaj = b.cmaj + d0;
  io.csub = b.csub + d0;
  io.csup = b.csup + d0;
  io.mco = b.mco + i0;
  io.nco = b.nco + i0;
  io.chi = b.chi + b.chi0[f];
  io.adiabatic = b.adiabatic;
  if (!b.adiabatic) {
    io.Tc = b.hl + i0;
    io.hc = b.hl + b.NI + i0;
    io.Tr = b.hl + 2 * b.NI + i0;
    io.hr = b.hl + 3 * b.NI + i0;
    io.scaled_hl = b.scaled_hl;
  }
  io.incl_ef = b.incl_ef;
  io.incl_vcp = b.incl_vcp;
  io.chem_rhs = b.chem_rhs + p0 * ns;
  io.aux = b.aux + p0 * 4;
  io.cpi = b.cpi + p0 * ns;
  return io;
}

// mixture-fraction stencils of every flamelet, dz of flamelet f at dz + dz0[f]; equation l of flamelet f is
// scaled by invLe[f * ns + l] (flamelet_kernels.cpp:31-48), unity when invLe is null
__global__ void fb_stencils_kernel(FlameletBatchView b, const double *dz, const long *dz0, const double *invLe,
                                   double *cmaj, double *csub, double *csup, double *mco, double *nco) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  const long le = e - b.dof0[f];
  const int i = (int)(le / b.ns), l = (int)(le % b.ns);
  const double *d = dz + dz0[f];
  const double *chi = b.chi + b.chi0[f];
  const double dzt = d[i] + d[i + 1];
  const double ile = invLe ? invLe[(long)f * b.ns + l] : 1.;
  cmaj[e] = -chi[1 + i] / (d[i] * d[i + 1]) * ile;
  csub[e] = chi[1 + i] / (dzt * d[i]) * ile;
  csup[e] = chi[1 + i] / (dzt * d[i + 1]) * ile;
  if (l == 0) {
    nco[b.ipt0[f] + i] = 1 / (d[i] + d[i + 1]);
    mco[b.ipt0[f] + i] = -nco[b.ipt0[f] + i];
  }
}

__global__ void fb_rhs_kernel(FlameletBatchView b, const double *X, const int *flag, double *R) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  if (flag && !flag[f]) return;
  const FlameletIO io = fb_io(b, f, X);
  rhs_entry(io, e - b.dof0[f], R + b.dof0[f]);
}

// The Newton-type matrix of each flagged flamelet from a fresh Jacobian (the chemistry blocks are already in J):
//   mode 0  diag(1/ds) - J     pseudo-transient continuation (flamelet.py:1597-1606:
//                              btddod_scale_and_add_diagonal(J, -1, 1/ds, 1))
//   mode 1  -J                 Newton (flamelet.py:1422)
//   mode 2  prefactor * J - I  implicit time stepping (flamelet.py:756-762, flamelet_kernels.cpp:1395-1408), with
//                              the prefactor of flamelet f in vals[f] (or the scalar prefactor when vals is null)
// vals: mode 0 the local pseudo-time steps (dof layout), mode 2 per-flamelet prefactors
// Grid: y = flamelet, x over that flamelet's entries (blocks past its end, or of an unflagged flamelet, exit).
__global__ void fb_matrix_kernel(MechView mv, FlameletBatchView b, const double *X, const int *flag, int mode,
                                 const double *vals, double prefactor, double *J) {
  const int f = blockIdx.y;
  if (!flag[f]) return;
  const long el = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= (f + 1 < b.B ? b.voff[f + 1] : b.NJ) - b.voff[f]) return;
  const long e = b.voff[f] + el;
  const FlameletIO io = fb_io(b, f, X);
  double v = el < (long)io.nzi * io.ns * io.ns ? J[e] : 0.;
  int dr;
  jac_entry(mv, io, el, mode == 2, mode == 2 && vals ? vals[f] : prefactor, 0, v, dr);
  double out = mode == 2 ? v : -v;
  if (mode == 0 && dr >= 0) out += 1. / vals[b.dof0[f] + dr];
  J[e] = out;
}

// Copies the flagged flamelets' unknowns from a vector holding flamelet f's unknowns from pk[f] into the batch
// layout: the dof layout (unknowns at dof0[f]) or, when interior is set, the padded interior rows of X
__global__ void fb_unpack_kernel(FlameletBatchView b, const int *flag, const long *pk, const double *packed,
                                 double *dst, int interior) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  if (!flag[f]) return;
  const long el = e - b.dof0[f];
  dst[interior ? (b.pt0[f] + 1) * b.ns + el : e] = packed[pk[f] + el];
}

// per flagged flamelet: max |R * 1/scale| (T scaled by 1e3, flamelet.py variable scales; NaN if any entry is
// not finite) and the smallest interior state value
__global__ void __launch_bounds__(256) fb_norm_kernel(FlameletBatchView b, const int *flag, const double *R,
                                                      const double *X, double *res, double *xmin) {
  const int f = blockIdx.x;
  if (flag && !flag[f]) return;
  const long n = (long)b.nzi[f] * b.ns, d0 = b.dof0[f];
  const double *x = X ? X + (b.pt0[f] + 1) * b.ns : nullptr;
  double m = 0., mn = INFINITY;
  int bad = 0;
  for (long e = threadIdx.x; e < n; e += 256) {
    const double r = R[d0 + e];
    const double a = fabs(r * ((e % b.ns) == 0 ? 1. / 1.e3 : 1.));
    if (!isfinite(r)) bad = 1;
    m = fmax(m, a);
    if (x) {
      const double xv = x[e];
      if (!isfinite(xv)) bad = 1;
      mn = fmin(mn, xv);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    m = fmax(m, __shfl_xor(m, o, 64));
    mn = fmin(mn, __shfl_xor(mn, o, 64));
    bad |= __shfl_xor(bad, o, 64);
  }
  __shared__ double sm[4], smn[4];
  __shared__ int sb[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sm[w] = m; smn[w] = mn; sb[w] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) { m = fmax(m, sm[k]); mn = fmin(mn, smn[k]); bad |= sb[k]; }
    res[f] = bad ? NAN : m;
    if (xmin) xmin[f] = bad ? NAN : mn;
  }
}

// Xt = X + dX over the interior of each flagged flamelet, first scaling dX by scale[f] when given
__global__ void fb_trial_kernel(FlameletBatchView b, const int *flag, const double *scale, const double *X,
                                double *Xt, double *dX) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  if (!flag[f]) return;
  double d = dX[e];
  if (scale) {
    d *= scale[f];
    dX[e] = d;
  }
  const long q = (b.pt0[f] + 1) * b.ns + (e - b.dof0[f]);
  Xt[q] = X[q] + d;
}

__global__ void fb_copy_interior_kernel(FlameletBatchView b, const int *flag, const double *src, double *dst) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  if (flag && !flag[f]) return;
  const long q = (b.pt0[f] + 1) * b.ns + (e - b.dof0[f]);
  dst[q] = src[q];
}

__global__ void fb_copy_dofs_kernel(FlameletBatchView b, const int *flag, const double *src, double *dst) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  if (flag && !flag[f]) return;
  dst[e] = src[e];
}

__global__ void fb_fill_dofs_kernel(FlameletBatchView b, const int *flag, const double *value, double *dst) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= b.NV) return;
  const int f = fb_find(b.dof0, b.B, e);
  if (flag && !flag[f]) return;
  dst[e] = value[f];
}

// ds <- min(safety / (expeig + 1e-16), ramp * ds, dsmax), then the flamelet minimum everywhere when
// global_min[f] (flamelet.py:1599-1604); expeig = max(0, eig - diffterm[f]) from eig = max(0, max Re(lambda))
__global__ void __launch_bounds__(256) fb_ds_kernel(FlameletBatchView b, const int *flag, const int *global_min,
                                                    const double *eig, const double *diffterm, double safety,
                                                    double ramp, double dsmax, double *ds) {
  const int f = blockIdx.x;
  if (!flag[f]) return;
  const long n = (long)b.nzi[f] * b.ns, d0 = b.dof0[f];
  double mn = INFINITY;
  for (long e = threadIdx.x; e < n; e += 256) {
    const double ex = fmax(eig[d0 + e] - diffterm[f], 0.);
    const double v = fmin(fmin(safety / (ex + 1.e-16), ramp * ds[d0 + e]), dsmax);
    ds[d0 + e] = v;
    mn = fmin(mn, v);
  }
  if (!global_min[f]) return;
  for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o, 64));
  __shared__ double smn[4];
  if ((threadIdx.x & 63) == 0) smn[threadIdx.x >> 6] = mn;
  __syncthreads();
  mn = fmin(fmin(smn[0], smn[1]), fmin(smn[2], smn[3]));
  for (long e = threadIdx.x; e < n; e += 256) ds[d0 + e] = mn;
}

__global__ void fb_clip_kernel(long n, double *X) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n && X[e] < 0.) X[e] = 0.;
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

#define FB_LAUNCH(name, n, ...)                                                                 \
  do {                                                                                          \
    if ((n) > 0) hipLaunchKernelGGL(name, dim3(nblk((n), 256)), dim3(256), 0, st, __VA_ARGS__); \
    if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, #name " launch failed");        \
    return GF_OK;                                                                               \
  } while (0)

int launch_fb_stencils(const FlameletBatchView &b, const double *dz, const long *dz0, const double *invLe,
                       double *cmaj, double *csub, double *csup, double *mco, double *nco, hipStream_t st) {
  FB_LAUNCH(fb_stencils_kernel, b.NV, b, dz, dz0, invLe, cmaj, csub, csup, mco, nco);
}
int launch_fb_rhs(const FlameletBatchView &b, const double *X, const int *flag, double *R, hipStream_t st) {
  FB_LAUNCH(fb_rhs_kernel, b.NV, b, X, flag, R);
}
int launch_fb_matrix(const gf_mech *m, const FlameletBatchView &b, const double *X, const int *flag, int mode,
                     const double *vals, double prefactor, double *J, hipStream_t st) {
  if (b.maxNJ > 0)
    hipLaunchKernelGGL(fb_matrix_kernel, dim3(nblk(b.maxNJ, 256), (unsigned)b.B), dim3(256), 0, st, m->view, b, X, flag,
                       mode, vals, prefactor, J);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "fb_matrix_kernel launch failed");
  return GF_OK;
}
int launch_fb_unpack(const FlameletBatchView &b, const int *flag, const long *pk, const double *packed, double *dst,
                     int interior, hipStream_t st) {
  FB_LAUNCH(fb_unpack_kernel, b.NV, b, flag, pk, packed, dst, interior);
}

int launch_fb_trial(const FlameletBatchView &b, const int *flag, const double *scale, const double *X, double *Xt,
                    double *dX, hipStream_t st) {
  FB_LAUNCH(fb_trial_kernel, b.NV, b, flag, scale, X, Xt, dX);
}
int launch_fb_copy_interior(const FlameletBatchView &b, const int *flag, const double *src, double *dst,
                            hipStream_t st) {
  FB_LAUNCH(fb_copy_interior_kernel, b.NV, b, flag, src, dst);
}
int launch_fb_copy_dofs(const FlameletBatchView &b, const int *flag, const double *src, double *dst, hipStream_t st) {
  FB_LAUNCH(fb_copy_dofs_kernel, b.NV, b, flag, src, dst);
}
int launch_fb_fill_dofs(const FlameletBatchView &b, const int *flag, const double *value, double *dst,
                        hipStream_t st) {
  FB_LAUNCH(fb_fill_dofs_kernel, b.NV, b, flag, value, dst);
}
int launch_fb_norms(const FlameletBatchView &b, const int *flag, const double *R, const double *X, double *res,
                    double *xmin, hipStream_t st) {
  if (b.B <= 0) return GF_OK;
  hipLaunchKernelGGL(fb_norm_kernel, dim3(b.B), dim3(256), 0, st, b, flag, R, X, res, xmin);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "fb_norm_kernel launch failed");
}
int launch_fb_ds(const FlameletBatchView &b, const int *flag, const int *global_min, const double *expeig,
                 const double *diffterm, double safety, double ramp, double dsmax, double *ds, hipStream_t st) {
  if (b.B <= 0) return GF_OK;
  hipLaunchKernelGGL(fb_ds_kernel, dim3(b.B), dim3(256), 0, st, b, flag, global_min, expeig, diffterm, safety, ramp,
                     dsmax, ds);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "fb_ds_kernel launch failed");
}
int launch_fb_clip_points(long n, double *X, hipStream_t st) { FB_LAUNCH(fb_clip_kernel, n, n, X); }

}  // namespace gfa
