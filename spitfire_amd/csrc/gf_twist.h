// Twisted block factorisation with register-resident Gauss-Jordan inverses (the factor half of the
// batched flamelet solver's linear algebra). The kernel keeps one block row per lane in registers, so it is
// compiled for every block size 1..64; the instantiations are spread over gf_twist_*.hip so that they
// compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

namespace gfa {

// Launches btd_twist_factor_kernel<T, N> (one workgroup of two waves per system). Defined for N in 1..64 and
// T in {double, float} by the gf_twist_*.hip units.
template <typename T, int N>
void launch_twist_factor(int nsys, const int *nblocks, const double *values, const long *voff, T *dinv,
                         const long *doff, hipStream_t st);

}  // namespace gfa
