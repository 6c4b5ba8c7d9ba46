// Twisted block factorisation with register-resident Gauss-Jordan inverses (the factor half of the
// batched flamelet solver's linear algebra). The kernel keeps one block row per lane in registers, so it is
// compiled for every block size 1..64; the instantiations are spread over gf_twist_*.hip so that they
// compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

namespace gfa {

// Launches the twisted factorisation for block size N: with `split` (and N >= 12) the column-split kernel,
// two waves per chain, whose pivot steps are shorter; otherwise one wave per chain, which wastes no issue
// slots on barriers once the batch alone fills every SIMD. Defined for N in 1..64 and T in {double, float} by
// the gf_twist_*.hip units.
template <typename T, int N>
void launch_twist_factor(int nsys, const int *nblocks, const double *values, const long *voff, T *dinv,
                         const long *doff, bool split, hipStream_t st);

}  // namespace gfa
