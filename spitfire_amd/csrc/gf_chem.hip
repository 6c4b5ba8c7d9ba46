// Batched FP64 isobaric-reactor chemistry on gfx950: net production rates, RHS, and the exact analytic
// Jacobian (griffon's option-0 sensitivities + isobaric transform), one workgroup per S states.
//
// Math restated from the reference: thermodynamics_kernels.cpp:45-351 (NASA7 cp, dcp/dT, h),
// chemistry_kernels.cpp:35-463 (production rates), rates_sensitivities_exact.cpp:33-1028 (exact
// sensitivities), isobaric_reactor_kernels.cpp:19-98, 319-343 (RHS, primitive Jacobian, transform).
//
// Structure: persistent workgroups (LDS-limited), each owning a contiguous range of states processed S at
// a time. The compact gather tables of the mechanism are copied into LDS once per workgroup; all
// per-state intermediates live in LDS; no atomics, so results are deterministic:
//   P0  one wave per state, lane = species: mixture thermo; wave-shuffle reductions for sum(y/M), cp, dcp/dT
//   P1  lane = (reaction, state), a reaction's S states in adjacent lanes: kf, Kc, falloff factor, mass
//       action and all partial derivatives.
//       dq_r/dY_s = sum_p v_{r,p} delta(s,p) + a_r (1/M_s - 1/M_last) + b_r, so each reaction stores five
//       scalars (q, dq/drho, dq/dT, a, b) and one value v per participant slot.
//   P2  one wave per state, lane = species k: gathers over the reactions with k in their net
//       stoichiometry give w_k, dw_k/drho, dw_k/dT, the rank-1 coefficients and the diagonal sensitivity
//       entry of column k; wave reductions give the enthalpy-weighted sums; the RHS is written here.
//   P3  one wave per Jacobian column, lane = row (plus one lane for the last species), looping over the S
//       states: off-diagonal sparse gather + rank-1 part give the species rows (chem_jac_isobaric + the
//       (rho,T,Y)->(T,Y) transform at constant pressure); the temperature row is the enthalpy-weighted
//       wave sum of the column; each column of each state is one contiguous 53-double store.
#include "gf_internal.h"
#include "gf_thermo.h"
#include "../../include/griffon_amd.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace gfa {

namespace {

constexpr int SP_STRIDE = NS_MAX;   // per-species LDS arrays
constexpr int N_SPEC_ARR = 12;      // y h cpi G dBdT w ws0 ws1 A B pr DG
constexpr int N_SPEC_ARR_RHS = 4;   // the RHS-only kernel keeps y h cpi G
constexpr int N_SCAL = 16;

enum { Y_ = 0, H_, CPI_, G_, DBDT_, W_, WS0_, WS1_, A_, B_, PR_, DG_ };
enum { S_T = 0, S_RHO, S_MMW, S_CP, S_CPST, S_INVT, S_LOGT, S_CT, S_INVM, S_LOGPORT, S_RHS0, S_HS0, S_HS1, S_WCP,
       S_HA, S_HB };

using ChemArgs = ChemIO;

// one DPP step of the wave scan: v + (the value of lane l - shift), 0 where no such lane (old value 0)
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_add(double v) {
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
  return v + __hiloint2double(hi, lo);
}

// Sum over the 64 lanes, the same value returned to every lane: an inclusive scan by DPP row shifts and row
// broadcasts (no LDS traffic), read from lane 63. The summation order is fixed, so results are reproducible.
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0x111, 0xF>(v);  // row_shr:1
  v = dpp_add<0x112, 0xF>(v);  // row_shr:2
  v = dpp_add<0x114, 0xF>(v);  // row_shr:4
  v = dpp_add<0x118, 0xF>(v);  // row_shr:8
  v = dpp_add<0x142, 0xA>(v);  // row_bcast:15
  v = dpp_add<0x143, 0xC>(v);  // row_bcast:31
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Swaps of 32-bit halves between two lane groups, applied to both words of a double: with 32, the upper 32
// lanes of a with the lower 32 lanes of b; with 16, the odd 16-lane rows of a with the even rows of b.
template <int W>
__device__ __forceinline__ void lane_swap(double &a, double &b) {
  const auto hi = W == 32 ? __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false)
                          : __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
  const auto lo = W == 32 ? __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false)
                          : __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
  a = __hiloint2double((int)hi[0], (int)lo[0]);
  b = __hiloint2double((int)hi[1], (int)lo[1]);
}

// Sums over the 64 lanes of four vectors at once, returned to every lane. Lanes l and l + 32 are added
// (two vectors per register after a 32-lane swap), then rows r and r + 1 (four vectors per register after a
// 16-lane swap), then the 16 lanes of each row by a DPP scan. Each vector is summed by the same tree whatever
// its position (every pairwise sum is commutative), so a vector's sum does not depend on what the other three
// hold; about a third of the instructions of four separate wave sums.
__device__ __forceinline__ void wave_sum4(const double (&v)[4], double (&out)[4]) {
  double a0 = v[0], b0 = v[1], a1 = v[2], b1 = v[3];
  lane_swap<32>(a0, b0);  // a0 = [v0 lanes 0-31 | v1 lanes 0-31], b0 = [v0 lanes 32-63 | v1 lanes 32-63]
  lane_swap<32>(a1, b1);
  double p = a0 + b0, q = a1 + b1;  // p: v0 in lanes 0-31, v1 in lanes 32-63; q likewise v2, v3
  lane_swap<16>(p, q);              // p = rows (p0, q0, p2, q2), q = rows (p1, q1, p3, q3)
  double t = p + q;                 // rows hold v0, v2, v1, v3
  t = dpp_add<0x111, 0xF>(t);       // row_shr:1
  t = dpp_add<0x112, 0xF>(t);       // row_shr:2
  t = dpp_add<0x114, 0xF>(t);       // row_shr:4
  t = dpp_add<0x118, 0xF>(t);       // row_shr:8: lane 15 of each row holds the row's sum
  const int src[4] = {15, 47, 31, 63};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    out[k] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(t), src[k]),
                              __builtin_amdgcn_readlane(__double2loint(t), src[k]));
}

__device__ __forceinline__ double arrh(const double *k, double logT, double invT) {
  return k[0] * exp(k[1] * logT - k[2] * invT);
}
__device__ __forceinline__ double arrh_sens_over_k(const double *k, double invT) { return invT * (k[1] + k[2] * invT); }

__device__ __forceinline__ double kf_of(const DevRxn &R, double T, double logT, double invT) {
  switch (R.kForm) {
    case KF_CONSTANT: return R.kf[0];
    case KF_LINEAR: return R.kf[0] * T;
    case KF_QUADRATIC: return R.kf[0] * T * T;
    case KF_RECIPROCAL: return R.kf[0] * invT;
    default: return arrh(R.kf, logT, invT);
  }
}

__device__ __forceinline__ double ipow3(double c, int st) {  // c^st for st in 1..3
  return st == 1 ? c : (st == 2 ? c * c : c * c * c);
}

__device__ __forceinline__ double troe_fcent(const DevRxn &R, double T, double invT, double &dfdT) {
  const double *t = R.troe;
  double f = 0., d = 0.;
  const int tf = R.troeForm;
  const bool has1 = (tf == TF_T1 || tf == TF_T12 || tf == TF_T13 || tf == TF_T123);
  const bool has2 = (tf == TF_T2 || tf == TF_T12 || tf == TF_T23 || tf == TF_T123);
  const bool has3 = (tf == TF_T3 || tf == TF_T13 || tf == TF_T23 || tf == TF_T123);
  if (has1) { const double e = exp(-T / t[1]); f += (1 - t[0]) * e; d += (t[0] - 1) / t[1] * e; }
  if (has2) { const double e = exp(-T / t[2]); f += t[0] * e; d += -t[0] / t[2] * e; }
  if (has3) { const double e = exp(-invT * t[3]); f += e; d += e * t[3] * invT * invT; }
  dfdT = d;
  return f;
}

template <int S, int THREADS, bool JAC>
__global__ void __launch_bounds__(THREADS) chem_kernel(MechView mv, ChemArgs a, long groups_per_wg) {
  static_assert(S >= 1 && S <= 4, "the temperature-row sums take at most four states at once");
  extern __shared__ double lds[];
  const int ns = mv.ns, nr = mv.nr, nm = ns - 1;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int NSA = JAC ? N_SPEC_ARR : N_SPEC_ARR_RHS;
  const long stride = NSA * SP_STRIDE + N_SCAL + (JAC ? 5L * nr + mv.nslot : (long)nr);
  const long ngroups = (a.n + S - 1) / S;
  // each workgroup owns a contiguous range of state groups (neighbouring states share SoA input lines)
  const long g0 = (long)blockIdx.x * groups_per_wg;
  const long g1 = min(ngroups, g0 + groups_per_wg);

  // gather tables -> LDS (the RHS-only kernel needs only the species gather)
  const int tcopy = JAC ? mv.tbl_bytes : ((mv.exo_at + 15) & ~15);
  for (int q = threadIdx.x; q < tcopy / 16; q += THREADS)
    reinterpret_cast<uint4 *>(lds)[q] = reinterpret_cast<const uint4 *>(mv.tbl)[q];
  const char *tb = reinterpret_cast<const char *>(lds);
  const uint32_t *sgx_off = reinterpret_cast<const uint32_t *>(tb);
  const uint64_t *sgx = reinterpret_cast<const uint64_t *>(tb + mv.sgx_at);
  const uint16_t *exo = reinterpret_cast<const uint16_t *>(tb + mv.exo_at);
  const uint32_t *exx = reinterpret_cast<const uint32_t *>(tb + mv.exx_at);
  const double *tbe = reinterpret_cast<const double *>(tb + mv.tbe_at);  // third-body lists and 1/M in LDS
  const int *tbi = reinterpret_cast<const int *>(tb + mv.tbi_at);
  const int *tbsl = reinterpret_cast<const int *>(tb + mv.tbs_at);
  const double *imw = reinterpret_cast<const double *>(tb + mv.imw_at);
  double *sbase = lds + tcopy / 8;
  __syncthreads();

  for (long group = g0; group < g1; ++group) {
    const long first = group * S;

    // ---------------- P0: mixture thermo (wave per state, lane = species) ----------------
    if (wave < S) {
      const long il = first + wave;
      double *sp = sbase + wave * stride;
      double *sc = sp + NSA * SP_STRIDE;
      if (il < a.n) {
        const long i = a.sel ? (long)a.sel[il] : il;
        const double *xi = a.x + i * a.xs_i;
        const double T = xi[0];
        double yk = 0.;
        if (a.rho_given == 1) {
          if (lane < ns) yk = xi[(1 + lane) * a.xs_v];
        } else {
          if (lane < nm) yk = xi[(1 + lane) * a.xs_v];
          const double sy = wave_sum(yk);
          if (lane == nm) yk = 1. - sy;
        }
        const double mm = wave_sum(lane < ns ? yk * mv.invmw[lane] : 0.);
        const double mmw = 1. / mm;
        const double invT = 1. / T, logT = log(T);
        const double rho = a.rho_given ? a.p[i * a.ps] : a.p[i * a.ps] * mmw / (T * mv.Ru);
        double cp = 0., cps = 0., h = 0., Gk = 0., dB = 0.;
        if (lane < ns) species_thermo(mv, lane, T, logT, invT, cp, cps, h, Gk, dB);
        if (lane < ns) {
          sp[Y_ * SP_STRIDE + lane] = yk;
          sp[H_ * SP_STRIDE + lane] = h;
          sp[CPI_ * SP_STRIDE + lane] = cp;
          sp[G_ * SP_STRIDE + lane] = Gk;
          if (JAC) sp[DBDT_ * SP_STRIDE + lane] = dB;
        }
        const double cpm = wave_sum(lane < ns ? yk * cp : 0.);
        // a constant-cp species resets the dcp/dT accumulation (thermodynamics_kernels.cpp:200-203)
        const double cpsm = wave_sum((lane < ns && lane > mv.last_const) ? yk * cps : 0.);
        if (a.aux && lane < 4) a.aux[i * 4 + lane] = lane == 0 ? rho : (lane == 1 ? cpm : (lane == 2 ? cpsm : mmw));
        if (a.cpi && lane < ns) a.cpi[i * ns + lane] = cp;
        if (a.hsp && lane < ns) a.hsp[i * ns + lane] = h;
        if (lane == 0) {
          sc[S_T] = T; sc[S_RHO] = rho; sc[S_MMW] = mmw; sc[S_CP] = cpm; sc[S_CPST] = cpsm;
          sc[S_INVT] = invT; sc[S_LOGT] = logT; sc[S_CT] = rho * mm; sc[S_INVM] = mm;
          sc[S_LOGPORT] = log(mv.pref * invT * mv.invRu);
        }
      }
    }
    __syncthreads();

    // ---------------- P1: reactions (lane = (reaction, state); a reaction's S states are adjacent lanes) -------
    for (int it = threadIdx.x; it < S * nr; it += THREADS) {
      const int r = it / S, s = it - r * S;
      const long i = first + s;
      if (i >= a.n) continue;
      double *sp = sbase + s * stride;
      const double *sc = sp + NSA * SP_STRIDE;
      double *rq = sp + NSA * SP_STRIDE + N_SCAL;
      double *v = rq + 5 * nr;
      const double *y = sp + Y_ * SP_STRIDE;
      const double T = sc[S_T], rho = sc[S_RHO], invT = sc[S_INVT], logT = sc[S_LOGT], ct = sc[S_CT],
                   invM = sc[S_INVM];
      const DevRxn &R = mv.rx[r];
      const double kf = kf_of(R, T, logT, invT);

      // third-body / falloff multiplier
      double tbsum = 0.;
      for (int q = 0; q < R.n_tb; ++q) tbsum += tbe[R.tb_off + q] * y[tbi[R.tb_off + q]];

      if (!JAC) {
        // production_rates formulation (chemistry_kernels.cpp:140-461)
        double k = kf, kr = 0.;
        const double M = R.baseEff * ct + rho * tbsum;
        if (R.type == RK_THIRD_BODY) {
          k *= M;
        } else if (R.type == RK_LINDEMANN) {
          k /= (1 + k / (arrh(R.kp, logT, invT) * M));
        } else if (R.type == RK_TROE) {
          const double pr = arrh(R.kp, logT, invT) / k * M;
          double dd;
          const double logFc = log10(troe_fcent(R, T, invT, dd));
          const double logPrC = log10(fmax(pr, 1.e-300)) + (-0.4 - 0.67 * logFc);
          const double F1 = logPrC / ((0.75 - 1.27 * logFc) - 0.14 * logPrC);
          k = k * pow(10., logFc / (1 + F1 * F1)) * pr / (1 + pr);
        }
        if (R.hasOrders) {
          for (int q = 0; q < R.n_sp; ++q) {
            const int sidx = mv.sp_idx[R.sp_off + q];
            k *= pow(fmax(y[sidx] * rho * mv.sp_invmw[R.sp_off + q], 0.), mv.sp_ord[R.sp_off + q]);
          }
        } else {
          if (R.reversible) {
            double sg = 0.;
            for (int q = 0; q < R.n_net; ++q) sg += R.net_st[q] * sp[G_ * SP_STRIDE + R.net_idx[q]];
            kr = k * exp(R.sumStoich * sc[S_LOGPORT] - invT * mv.invRu * sg);
          }
          for (int j = 0; j < R.n_rc; ++j) {
            const int sidx = R.rc_idx[j];
            k *= ipow3(y[sidx] * rho * imw[sidx], R.rc_st[j]);
          }
          if (R.reversible)
            for (int j = 0; j < R.n_pd; ++j) {
              const int sidx = R.pd_idx[j];
              kr *= ipow3(y[sidx] * rho * imw[sidx], R.pd_st[j]);
            }
        }
        rq[r] = k - kr;
        continue;
      }

      double Ctbaf = 1., dCdrho = 0., dCdT = 0., cdense = 0., tbs = 0.;
      double qaC = 0.;  // d(q)/d(ln A) = Rnet * qaC; A is the forward (high-pressure) pre-exponential
      if (R.type == RK_THIRD_BODY) {
        Ctbaf = R.baseEff * ct + rho * tbsum;
        dCdrho = R.baseEff * invM + tbsum;
        cdense = rho * R.baseEff;
        tbs = rho;
      } else if (R.type == RK_LINDEMANN || R.type == RK_TROE) {
        const double flf = R.baseEff * ct + rho * tbsum;
        const double kpk = arrh(R.kp, logT, invT) / kf;
        const double pr = kpk * flf;
        const double dlnk = arrh_sens_over_k(R.kp, invT) - arrh_sens_over_k(R.kf, invT);
        double nsT;
        if (R.type == RK_LINDEMANN) {
          Ctbaf = pr / (1. + pr);
          dCdT = Ctbaf / (1. + pr) * dlnk;
          nsT = kpk / ((1. + pr) * (1. + pr));
        } else {  // rates_sensitivities_exact.cpp:905-995
          double dfc;
          const double fc = troe_fcent(R, T, invT, dfc);
          const double l10pr = log10(fmax(pr, 1.e-300));
          const double l10fc = log10(fmax(fc, 1.e-300));
          const double lfc = log(fmax(fc, 1.e-300));
          const double ln10 = 2.302585092994045684;  // log(10)
          const double aT = l10pr - 0.67 * l10fc - 0.4;
          const double bT = -0.14 * l10pr - 1.1762 * l10fc + 0.806;
          const double gT = 1 / (1 + (aT / bT) * (aT / bT));
          const double fT = pow(fc, gT);
          Ctbaf = fT * pr / (1 + pr);
          const double dfTdT = fT * (gT / fc * dfc + lfc * (-2.0 * gT * gT / ln10 * aT / (bT * bT * bT) *
                                                              ((bT + 0.14 * aT) * dlnk - (0.67 * bT - 1.1762 * aT) * dfc / fc)));
          dCdT = 1. / (1. + 1. / pr) * dfTdT + fT * pr / ((1. + pr) * (1. + pr)) * dlnk;
          nsT = kpk * (-2.0 / (1. + pr) * fT * lfc * gT * gT / ln10 * aT / (bT * bT * bT) * (bT + 0.14 * aT) +
                       fT / ((1. + pr) * (1 + pr)));
        }
        dCdrho = nsT * R.baseEff * invM + nsT * tbsum;
        cdense = nsT * rho * R.baseEff;
        tbs = nsT * rho;
        // Pr = k0 [M] / kinf: d ln(q) / d ln(kinf) = 1 - (Pr / Ctbaf) dCtbaf/dPr, and nsT = (k0 / kinf) dCtbaf/dPr
        qaC = Ctbaf - flf * nsT;
      }

      if (a.notbaf) {  // rates_sensitivity_option 1: no derivatives of the third-body / falloff multiplier
        dCdrho = 0.;
        dCdT = 0.;
        cdense = 0.;
        tbs = 0.;
      }
      double *vr = v + R.slot_off;
      for (int q = 0; q < R.n_slot; ++q) vr[q] = 0.;
      double Rnet, dRdrho, dRdT, U = 0.;
      const double ksens = arrh_sens_over_k(R.kf, invT);
      if (R.hasOrders) {  // rates_sensitivities_exact.cpp:198-281
        double C[NSP_MAX], ordv[NSP_MAX];
        int sidx[NSP_MAX];
        Rnet = kf;
        double sumO = 0.;
        for (int q = 0; q < R.n_sp; ++q) {
          sidx[q] = mv.sp_idx[R.sp_off + q];
          ordv[q] = mv.sp_ord[R.sp_off + q];
          C[q] = y[sidx[q]] * rho * mv.sp_invmw[R.sp_off + q];
          Rnet *= pow(fmax(C[q], 0.), ordv[q]);
          sumO += ordv[q];
        }
        dRdrho = Rnet / ct * invM * sumO;
        dRdT = Rnet * ksens;
        for (int j = 0; j < R.n_sp; ++j) {
          const double im = mv.sp_invmw[R.sp_off + j];
          const bool last = (sidx[j] == ns - 1);
          double m = kf;
          for (int l = 0; l < R.n_sp; ++l) {
            if (l != j) {
              m *= last ? pow(C[l], ordv[l]) : pow(fmax(C[l], 0.), ordv[l]);
            } else {
              if (last || ordv[l] > 1) m *= ordv[l] * rho * im * pow(fmax(C[l], 1.e-16), ordv[l] - 1.);
              else m *= ordv[l] * rho * im / pow(fmax(C[l], 1.e-16), 1. - ordv[l]);
            }
          }
          if (last) U -= m;
          else vr[R.sp_slot[j]] += Ctbaf * m;
        }
      } else {
        double C[NRC_MAX];
        Rnet = kf;
        for (int j = 0; j < R.n_rc; ++j) {
          C[j] = y[R.rc_idx[j]] * rho * imw[R.rc_idx[j]];
          Rnet *= ipow3(C[j], R.rc_st[j]);
        }
        dRdrho = Rnet / ct * invM * R.sumRc;
        dRdT = Rnet * ksens;
        for (int j = 0; j < R.n_rc; ++j) {
          double d = kf * rho * imw[R.rc_idx[j]] * R.rc_st[j] * (R.rc_st[j] == 1 ? 1. : ipow3(C[j], R.rc_st[j] - 1));
          for (int l = 0; l < R.n_rc; ++l)
            if (l != j) d *= ipow3(C[l], R.rc_st[l]);
          if (R.rc_slot[j] >= 0) vr[R.rc_slot[j]] += Ctbaf * d;
          else U -= d;
        }
        if (R.reversible) {
          double sg = 0., sb = 0.;
          for (int q = 0; q < R.n_net; ++q) {
            sg += R.net_st[q] * sp[G_ * SP_STRIDE + R.net_idx[q]];
            sb += R.net_st[q] * sp[DBDT_ * SP_STRIDE + R.net_idx[q]];
          }
          const double Kc = exp(-(R.sumStoich * sc[S_LOGPORT] - invT * mv.invRu * sg));
          const double kr = kf / Kc;
          double Cp[NPD_MAX];
          double Rr = kr;
          for (int j = 0; j < R.n_pd; ++j) {
            Cp[j] = y[R.pd_idx[j]] * rho * imw[R.pd_idx[j]];
            Rr *= ipow3(Cp[j], R.pd_st[j]);
          }
          Rnet -= Rr;
          dRdrho -= Rr / ct * invM * R.sumPd;
          dRdT -= Rr * (ksens + sb);
          for (int j = 0; j < R.n_pd; ++j) {
            double d = kr * rho * imw[R.pd_idx[j]] * R.pd_st[j] * (R.pd_st[j] == 1 ? 1. : ipow3(Cp[j], R.pd_st[j] - 1));
            for (int l = 0; l < R.n_pd; ++l)
              if (l != j) d *= ipow3(Cp[l], R.pd_st[l]);
            if (R.pd_slot[j] >= 0) vr[R.pd_slot[j]] -= Ctbaf * d;
            else U += d;
          }
        }
      }
      double Elast = 0.;
      for (int q = 0; q < R.n_tb; ++q) {
        const int slot = tbsl[R.tb_off + q];
        const double e = tbe[R.tb_off + q];
        if (slot >= 0) vr[slot] += Rnet * tbs * e;
        else if (q == R.tb_last) Elast = tbs * e;
      }
      rq[r] = Rnet * Ctbaf;
      if (a.qa) {
        const long q = a.sel ? (long)a.sel[i] : i;
        a.qa[q * nr + R.orig] = Rnet * ((R.type == RK_LINDEMANN || R.type == RK_TROE) ? qaC : Ctbaf);
      }
      rq[nr + r] = dRdrho * Ctbaf + dCdrho * Rnet;
      rq[2 * nr + r] = dRdT * Ctbaf + dCdT * Rnet;
      rq[3 * nr + r] = Rnet * cdense;
      rq[4 * nr + r] = Ctbaf * U - Rnet * Elast;
    }
    __syncthreads();

    // ---------------- P2: species gathers + reductions (wave per state, lane = species) ----------------
    if (wave < S) {
      const long il = first + wave;
      double *sp = sbase + wave * stride;
      double *sc = sp + NSA * SP_STRIDE;
      const double *rq = sc + N_SCAL;
      const double *v = rq + 5 * nr;
      if (il < a.n) {
        const long i = a.sel ? (long)a.sel[il] : il;
        double w = 0., w0 = 0., w1 = 0., A = 0., B = 0., DG = 0.;
        if (lane < ns) {
          const double Mk = mv.mw[lane];
          const uint32_t e0 = sgx_off[lane], e1 = sgx_off[lane + 1];
          uint32_t e = e0;
          for (; e + 1 < e1; e += 2) {  // two reactions per trip: their LDS loads overlap (same summation order)
            const uint64_t ra = sgx[e], rb = sgx[e + 1];
            const int r0 = (int)(ra & 0xFFFFu), r1 = (int)(rb & 0xFFFFu);
            const double c0 = -(double)(int8_t)(ra >> 32) * Mk, c1 = -(double)(int8_t)(rb >> 32) * Mk;
            const double q0 = rq[r0], q1 = rq[r1];
            if (JAC) {
              const int d0 = (int)((ra >> 16) & 0xFFFFu), d1 = (int)((rb >> 16) & 0xFFFFu);
              const double a00 = rq[nr + r0], a01 = rq[nr + r1], a10 = rq[2 * nr + r0], a11 = rq[2 * nr + r1];
              const double a20 = rq[3 * nr + r0], a21 = rq[3 * nr + r1], a30 = rq[4 * nr + r0], a31 = rq[4 * nr + r1];
              const double v0 = d0 != 0xFFFF ? v[d0] : 0., v1 = d1 != 0xFFFF ? v[d1] : 0.;
              w += c0 * q0; w += c1 * q1;
              w0 += c0 * a00; w0 += c1 * a01;
              w1 += c0 * a10; w1 += c1 * a11;
              A += c0 * a20; A += c1 * a21;
              B += c0 * a30; B += c1 * a31;
              if (d0 != 0xFFFF) DG += c0 * v0;
              if (d1 != 0xFFFF) DG += c1 * v1;
            } else {
              w += c0 * q0; w += c1 * q1;
            }
          }
          for (; e < e1; ++e) {
            const uint64_t rec = sgx[e];
            const int r = (int)(rec & 0xFFFFu);
            const double c = -(double)(int8_t)(rec >> 32) * Mk;  // -nu * M_k (rates_sensitivities_exact.cpp:1018)
            w += c * rq[r];
            if (JAC) {
              w0 += c * rq[nr + r];
              w1 += c * rq[2 * nr + r];
              A += c * rq[3 * nr + r];
              B += c * rq[4 * nr + r];
              const int ds = (int)((rec >> 16) & 0xFFFFu);
              if (ds != 0xFFFF) DG += c * v[ds];
            }
          }
        }
        const double h = lane < ns ? sp[H_ * SP_STRIDE + lane] : 0.;
        const double rho = sc[S_RHO], cp = sc[S_CP];
        const double HW = wave_sum(h * w);
        const double rhs0 = -HW / (rho * cp);
        const double iR = 1. / rho;
        if (a.rhs) {
          double *ro = a.rhs + i * a.rs_i;
          if (lane == 0) ro[0] = rhs0;
          if (lane < nm) ro[(1 + lane) * a.rs_v] = w * iR;
        }
        if (a.wout && lane < ns) a.wout[i * ns + lane] = w;
        if (JAC) {
          const double cpi = lane < ns ? sp[CPI_ * SP_STRIDE + lane] : 0.;
          const double HS0 = wave_sum(h * w0), HS1 = wave_sum(h * w1), WCP = wave_sum(w * cpi);
          if (lane < ns) {
            sp[W_ * SP_STRIDE + lane] = w;
            sp[WS0_ * SP_STRIDE + lane] = w0;
            sp[WS1_ * SP_STRIDE + lane] = w1;
            sp[A_ * SP_STRIDE + lane] = A;
            sp[B_ * SP_STRIDE + lane] = B;
            sp[DG_ * SP_STRIDE + lane] = DG;
          }
          const double iRC = 1. / (rho * cp);
          if (lane < nm) sp[PR_ * SP_STRIDE + 1 + lane] = iR * (w0 - iR * w);
          if (lane == 0) {
            sp[PR_ * SP_STRIDE] = -iRC * HS0 - iR * rhs0;
            sc[S_RHS0] = rhs0; sc[S_HS0] = HS0; sc[S_HS1] = HS1; sc[S_WCP] = WCP;
          }
        }
      }
    }
    if (JAC) {
      __syncthreads();
      // ---------------- P3: Jacobian columns (wave = column, lane = row; lane ns = the last species) -------
      // Every multiply-add here is written as an explicit fma or a rounded product (__dmul_rn), so the compiler's
      // contraction choices cannot differ between launch shapes and every shape gives bitwise the same Jacobian.
      int nact = 0;
      double iR[S], iRC[S], ric[S], rmw[S];  // 1/rho, 1/(rho cp), rhs0/cp, -rho M of each state
      double *jb[S];  // Jacobian block of each state (nullptr: none)
      for (int s = 0; s < S; ++s) {
        jb[s] = nullptr;
        if (first + s < a.n) {
          nact = s + 1;
          const long q = a.sel ? (long)a.sel[first + s] : first + s;
          if (a.jac) {
            const long off = a.js_off ? a.js_off[q] : q * a.js_i;
            if (off >= 0) jb[s] = a.jac + off;
          }
        }
        const double *sc = sbase + s * stride + NSA * SP_STRIDE;
        iR[s] = 1. / sc[S_RHO];
        iRC[s] = 1. / (sc[S_RHO] * sc[S_CP]);
        ric[s] = sc[S_RHS0] / sc[S_CP];
        rmw[s] = -sc[S_RHO] * sc[S_MMW];
      }
      if (a.jac) {
        constexpr int NW = THREADS / 64;
        // per-row values of every state, loaded once for all the columns this wave assembles
        const int iiL = lane - 1;
        const bool rowL = lane >= 1 && lane <= ns;
        double rA[S], rB[S], rH[S], rP[S];
        // the row's molecular weight (loaded here: in the column loop the Jacobian stores could alias it)
        const double MiL = rowL ? mv.mw[iiL] : 0.;
        for (int s = 0; s < S; ++s) {
          const double *sp = sbase + s * stride;
          rA[s] = rowL ? sp[A_ * SP_STRIDE + iiL] : 0.;
          rB[s] = rowL ? sp[B_ * SP_STRIDE + iiL] : 0.;
          rH[s] = rowL ? sp[H_ * SP_STRIDE + iiL] : 0.;
          rP[s] = lane < ns ? sp[PR_ * SP_STRIDE + lane] : 0.;
        }
        for (int c = wave; c < ns; c += NW) {
          double val[S];
          if (c == 0) {
            for (int s = 0; s < S; ++s) {
              const double *sp = sbase + s * stride;
              const double *sc = sp + NSA * SP_STRIDE;
              const double *pr = sp + PR_ * SP_STRIDE;
              const double roT = sc[S_RHO] / sc[S_T];
              const double T00 = fma(-iRC[s], sc[S_HS1] + sc[S_WCP], -__dmul_rn(ric[s], sc[S_CPST]));
              val[s] = (lane == 0) ? fma(-roT, pr[0], T00)
                                   : (lane < ns ? fma(-roT, pr[lane], __dmul_rn(sp[WS1_ * SP_STRIDE + lane - 1], iR[s]))
                                                : 0.);
            }
          } else {
            const int j = c - 1;
            const double dj = mv.dinvm[j];
            const int ii = lane - 1;  // species row; lane ns is the last species (temperature row only)
            const bool row = lane >= 1 && lane <= ns;
            double full[S];
            for (int s = 0; s < S; ++s) full[s] = 0.;
            if (row) {
              for (int s = 0; s < S; ++s) full[s] = (ii == j) ? sbase[s * stride + DG_ * SP_STRIDE + j] : 0.;
              const double Mi = MiL;
              const int ee = j * ns + ii;
              const int t0 = exo[ee], t1 = exo[ee + 1];
              const double *vb = sbase + NSA * SP_STRIDE + N_SCAL + 5 * nr;
              int t = t0;
              for (; t + 1 < t1; t += 2) {  // two entries per trip: their LDS loads overlap (same summation order)
                const uint32_t ra = exx[t], rb = exx[t + 1];
                const int sa = (int)(ra & 0xFFFFu), sb = (int)(rb & 0xFFFFu);
                const double ca = -(double)(int8_t)(ra >> 16) * Mi, cb = -(double)(int8_t)(rb >> 16) * Mi;
                double va[S], vbv[S];
                for (int s = 0; s < S; ++s) {
                  va[s] = vb[s * stride + sa];
                  vbv[s] = vb[s * stride + sb];
                }
                for (int s = 0; s < S; ++s) {
                  full[s] = fma(ca, va[s], full[s]);
                  full[s] = fma(cb, vbv[s], full[s]);
                }
              }
              if (t < t1) {
                const uint32_t rec = exx[t];
                const int slot = (int)(rec & 0xFFFFu);
                const double coef = -(double)(int8_t)(rec >> 16) * Mi;
                for (int s = 0; s < S; ++s) full[s] = fma(coef, vb[s * stride + slot], full[s]);
              }
              for (int s = 0; s < S; ++s) full[s] += fma(dj, rA[s], rB[s]);
            }
            // temperature row: -(1/(rho cp)) sum_k h_k dw_k/dY_j over every species (isobaric_reactor_kernels.cpp:58-98),
            // the S states' sums formed together (absent states are zero vectors)
            double hv[4], hsum[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) hv[s] = (s < S && row) ? __dmul_rn(rH[s < S ? s : 0], full[s < S ? s : 0]) : 0.;
            wave_sum4(hv, hsum);
            for (int s = 0; s < S; ++s) {
              const double *sp = sbase + s * stride;
              const double hs = hsum[s];
              const double rd = __dmul_rn(rmw[s], dj);
              val[s] = (lane == 0)
                           ? fma(-iRC[s], hs, __dmul_rn(rd, sp[PR_ * SP_STRIDE]) -
                                                  __dmul_rn(ric[s], sp[CPI_ * SP_STRIDE + j] - sp[CPI_ * SP_STRIDE + nm]))
                           : (lane < ns ? fma(iR[s], full[s], __dmul_rn(rd, rP[s])) : 0.);
            }
          }
          if (lane < ns)
            for (int s = 0; s < nact; ++s)
              if (jb[s]) jb[s][(long)c * ns + lane] = val[s];
        }
      }
      if (a.sens) {
        // raw option-0 sensitivities, (ns+1)^2 column-major with stride ns+1 (rates_sensitivities_exact.cpp:1011-1026)
        for (int s = 0; s < nact; ++s) {
          const double *sp = sbase + s * stride;
          const double *v = sp + NSA * SP_STRIDE + N_SCAL + 5 * nr;
          const long q = a.sel ? (long)a.sel[first + s] : first + s;
          double *so = a.sens + q * (long)(ns + 1) * (ns + 1);
          for (int e = threadIdx.x; e < (ns + 1) * (ns + 1); e += THREADS) {
            const int c = e / (ns + 1), k = e - c * (ns + 1);
            double val = 0.;
            if (k < ns) {
              if (c == 0) val = sp[WS0_ * SP_STRIDE + k];
              else if (c == 1) val = sp[WS1_ * SP_STRIDE + k];
              else {
                const int j = c - 2;
                const int ee = j * ns + k;
                double spv = 0.;
                for (int t = mv.eg_off[ee]; t < mv.eg_off[ee + 1]; ++t) spv += mv.eg_coef[t] * v[mv.eg_slot[t]];
                val = spv + mv.dinvm[j] * sp[A_ * SP_STRIDE + k] + sp[B_ * SP_STRIDE + k];
              }
            }
            so[e] = val;
          }
        }
      }
    }
    __syncthreads();  // the next group reuses the LDS
  }
}

template <int S, int THREADS, bool JAC>
size_t lds_bytes(const gf_mech *m) {
  const size_t t = JAC ? (size_t)m->view.tbl_bytes : (((size_t)m->view.exo_at + 15) & ~size_t(15));
  return t + S * (JAC ? m->lds_per_state : m->lds_per_state_rhs);
}

template <int S, int THREADS, bool JAC>
int launch_cfg(const gf_mech *m, const ChemArgs &a, hipStream_t st) {
  const size_t lds = lds_bytes<S, THREADS, JAC>(m);
  auto kern = chem_kernel<S, THREADS, JAC>;
  if (lds > m->lds_max) return fail(GF_ERR_UNSUPPORTED, "mechanism needs more LDS per workgroup than the device allows");
  if (lds > 65536) {
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return fail(GF_ERR_UNSUPPORTED, "mechanism needs more LDS per workgroup than the device allows");
  }
  const long groups = (a.n + S - 1) / S;
  // persistent launch: as many workgroups as fit the device at once (LDS-limited, at most 2048 threads
  // per CU), each owning a contiguous range of groups
  const long per_cu = std::max<long>(1, std::min<long>((long)(std::max<size_t>(m->lds_max, 65536) / lds), 2048 / THREADS));
  const long width = std::max<long>(1, (long)std::max(m->n_cu, 1) * per_cu);
  const long nwg = std::min(groups, width);
  const long gpw = (groups + nwg - 1) / nwg;
  const long grid = (groups + gpw - 1) / gpw;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(THREADS), lds, st, m->view, a, gpw);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "chemistry kernel launch failed");
  return GF_OK;
}

}  // namespace

// Launch shapes: states per workgroup x threads per workgroup. Every shape computes each state with the same
// lanes in the same order (per-state waves and reductions do not depend on S), so their outputs are bitwise
// identical and only their speed differs.
namespace {
constexpr int kShapes = 5;  // 0: 1x128, 1: 2x256, 2: 4x512, 3: 4x256, 4: 2x512

bool shape_fits(const gf_mech *m, int c, bool J) {
  if (!J) return true;
  const size_t lim = m->lds_max;
  if (c == 2 || c == 3) return lds_bytes<4, 512, true>(m) <= lim;
  if (c == 1 || c == 4) return lds_bytes<2, 256, true>(m) <= lim;
  return true;
}

int launch_shape(const gf_mech *m, const ChemIO &a, hipStream_t st, int c, bool J) {
  if (J) {
    switch (c) {
      case 0: return launch_cfg<1, 128, true>(m, a, st);
      case 1: return launch_cfg<2, 256, true>(m, a, st);
      case 3: return launch_cfg<4, 256, true>(m, a, st);
      case 4: return launch_cfg<2, 512, true>(m, a, st);
      default: return launch_cfg<4, 512, true>(m, a, st);
    }
  }
  switch (c) {
    case 0: return launch_cfg<1, 128, false>(m, a, st);
    case 1: return launch_cfg<2, 256, false>(m, a, st);
    case 3: return launch_cfg<4, 256, false>(m, a, st);
    case 4: return launch_cfg<2, 512, false>(m, a, st);
    default: return launch_cfg<4, 512, false>(m, a, st);
  }
}

// The fastest feasible shape for this launch: each candidate runs once to warm up, then twice between HIP
// events on the caller's stream (the outputs are the caller's, written identically by every candidate).
int autotune_shape(const gf_mech *m, const ChemIO &a, hipStream_t st, bool J, int &best) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(GF_ERR_HIP, "hipEventCreate failed");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(GF_ERR_HIP, "hipEventCreate failed");
  }
  float tbest = 0.f;
  bool have = false;
  best = 2;
  int rc = GF_OK;
  for (int c = 0; c < kShapes && rc == GF_OK; ++c) {
    if (!shape_fits(m, c, J)) continue;
    rc = launch_shape(m, a, st, c, J);
    if (rc == GF_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(GF_ERR_HIP, "hipEventRecord failed");
    for (int k = 0; k < 2 && rc == GF_OK; ++k) rc = launch_shape(m, a, st, c, J);
    if (rc == GF_OK && hipEventRecord(e1, st) != hipSuccess) rc = fail(GF_ERR_HIP, "hipEventRecord failed");
    if (rc == GF_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(GF_ERR_HIP, "hipEventSynchronize failed");
    float ms = 0.f;
    if (rc == GF_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(GF_ERR_HIP, "hipEventElapsedTime failed");
    if (rc == GF_OK && (!have || ms < tbest)) {
      have = true;
      tbest = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}
}  // namespace

// GF_CHEM_CFG=<S>x<threads> (or gf_set_chem_launch_shape) fixes the launch shape for every call. Otherwise
// the first launch of a kind
// (RHS + Jacobian, or RHS only) at a batch size in [2^k, 2^(k+1)) times the feasible shapes and the fastest is
// kept for that kind and size class on this mechanism; under stream capture the default 4x512 is used
// untimed. Concurrent first calls from several host threads on one mechanism may time twice.
namespace {
int g_fixed_shape = -2;  // -2: not read from the environment yet; -1: autotune; 0..4: that shape
int fixed_shape() {
  if (g_fixed_shape == -2) {
    g_fixed_shape = -1;
    if (const char *e = getenv("GF_CHEM_CFG")) {
      if (!strcmp(e, "1x128")) g_fixed_shape = 0;
      else if (!strcmp(e, "2x256")) g_fixed_shape = 1;
      else if (!strcmp(e, "4x512")) g_fixed_shape = 2;
      else if (!strcmp(e, "4x256")) g_fixed_shape = 3;
      else if (!strcmp(e, "2x512")) g_fixed_shape = 4;
    }
  }
  return g_fixed_shape;
}
}  // namespace

int set_chem_launch_shape(int shape) {
  if (shape < -1 || shape >= kShapes) return fail(GF_ERR_ARG, "launch shape must be -1 (timed) or 0..4");
  g_fixed_shape = shape;
  return GF_OK;
}

int launch_chem(const gf_mech *m, const ChemIO &a, hipStream_t st) {
  if (a.n <= 0) return GF_OK;
  const int fixed = fixed_shape();
  const bool J = a.jac || a.sens;
  auto fallback = [&](int c) {  // fewer states per workgroup where the LDS does not fit
    if ((c == 2 || c == 3) && !shape_fits(m, c, J)) c = 1;
    if ((c == 1 || c == 4) && !shape_fits(m, c, J)) c = 0;
    return c;
  };
  if (fixed >= 0) return launch_shape(m, a, st, fallback(fixed), J);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return launch_shape(m, a, st, fallback(2), J);
  if (!m->chem_shape_init) {
    std::memset(m->chem_shape, -1, sizeof m->chem_shape);
    m->chem_shape_init = true;
  }
  const int k = 63 - __builtin_clzll((unsigned long long)a.n);
  signed char &slot = m->chem_shape[J ? 0 : 1][k];
  if (a.no_tune) return launch_shape(m, a, st, fallback(slot >= 0 ? slot : 2), J);
  if (slot < 0) {
    int best = 2;
    const int rc = autotune_shape(m, a, st, J, best);
    if (rc != GF_OK) return rc;
    slot = (signed char)best;
  }
  return launch_shape(m, a, st, slot, J);
}

}  // namespace gfa
