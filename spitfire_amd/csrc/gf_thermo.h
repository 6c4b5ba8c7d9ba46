// Species thermodynamics on the device, shared by the chemistry and reactor kernels.
#pragma once
#include "gf_internal.h"

namespace gfa {

// NASA7 / constant-cp species thermo at T (thermodynamics_kernels.cpp:45-351; chemistry_kernels.cpp:56-97
// for the Gibbs-like term, which is not clamped to [Tmin, Tmax]; rates_sensitivities_exact.cpp:82-126 for dB/dT)
__device__ inline void species_thermo(const MechView &mv, int k, double T, double logT, double invT, double &cp, double &cps,
                               double &h, double &G, double &dB) {
  const double *c = mv.spc_c + 16 * k;
  const double im = mv.invmw[k];
  const double Ru = 1. / mv.invRu;
  if (mv.cpkind[k] == CP_NASA7) {
    const double minT = mv.Tmin[k], maxT = mv.Tmax[k];
    const bool lowr = T <= c[0];
    const int o = lowr ? 8 : 1;
    double te = T;
    if (T < minT) te = minT;
    else if (T > maxT) te = maxT;
    const int oe = (T < minT) ? 8 : ((T > maxT) ? 1 : o);
    cp = im * (c[oe] + te * (2. * c[oe + 1] + te * (6. * c[oe + 2] + te * (12. * c[oe + 3] + 20. * te * c[oe + 4]))));
    if (T < minT || T > maxT) {
      cps = 0.;
      const double Te = te;
      const int hb = (T < minT) ? 8 : 1;
      const double a6 = (T < minT) ? c[13] : c[6];
      h = im * (a6 + c[hb] * T + Te * (2. * c[hb + 1] * T + Te * (3. * 2. * c[hb + 2] * T - c[hb + 1] + Te * (4. * 3. * c[hb + 3] * T - 2. * 2. * c[hb + 2] + Te * (5. * 4. * c[hb + 4] * T - 3. * 3. * c[hb + 3] + Te * -4. * 4. * c[hb + 4])))));
    } else {
      cps = im * ((2. * c[o + 1] + T * (12. * c[o + 2] + T * (36. * c[o + 3] + 80. * T * c[o + 4]))));
      const double a6 = lowr ? c[13] : c[6];
      h = im * (a6 + T * (c[o] + T * (c[o + 1] + T * (2. * c[o + 2] + T * (3. * c[o + 3] + T * 4. * c[o + 4])))));
    }
    const double a6 = lowr ? c[13] : c[6];
    const double a7 = lowr ? c[14] : c[7];
    G = a6 + T * (c[o] - a7 - c[o] * logT - T * (c[o + 1] + T * (c[o + 2] + T * (c[o + 3] + T * c[o + 4]))));
    dB = mv.invRu * ((c[o] - Ru) * invT + c[o + 1] + T * (2 * c[o + 2] + T * (3 * c[o + 3] + T * 4 * c[o + 4])) + a6 * invT * invT);
  } else {  // constant cp
    cp = im * c[3];
    cps = 0.;
    h = im * (c[1] + c[3] * (T - c[0]));
    G = c[1] + c[3] * (T - c[0]) - T * (c[2] + c[3] * (logT - log(c[0])));
    dB = invT * (mv.mw[k] * mv.invRu * (c[3] - invT * (c[3] * c[0] - c[1])) - 1);
  }
}

}  // namespace gfa
