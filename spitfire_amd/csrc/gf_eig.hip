// Per-block largest-real-part eigenvalue kernel (the pseudo-time limiter input of the flamelet solvers).
#include "gf_internal.h"
#include "gf_eig.h"
#include "gf_wave.h"
#include "../../include/griffon_amd.h"

#include <type_traits>

namespace gfa {

// Device policy for eig_max_real: one 64-lane wave per matrix, lane t owns row/column t.
struct WaveLanes {
  int lane;
  // one value per index, held by the index's lane
  struct LaneVal {
    float v;
  };
  __device__ float &at(LaneVal &x, int) { return x.v; }
  __device__ float get(const LaneVal &x, int i) { return wave::lane_value(x.v, i); }
  template <class F> __device__ void each(int lo, int hi, F f) {
    if (lane >= lo && lane < hi) f(lane);
    __syncthreads();
  }
  // the largest i in [lo, hi) with pred(i), or lo - 1 if none: every lane tests its own index, one ballot
  template <class F> __device__ int last_true(int lo, int hi, F pred) {
    const bool v = lane >= lo && lane < hi && pred(lane);
    const unsigned long long mask = __ballot(v);
    return mask ? 63 - __clzll((long long)mask) : lo - 1;
  }
  // 1/x from the hardware reciprocal and one Newton step (within an ulp or so; the reflector coefficients
  // need no more), keeping 1/0 = inf and 1/inf = 0
  __device__ double recip(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.);
    return isfinite(e) ? fma(r, e, r) : r;
  }
  template <class F> __device__ double sum(int lo, int hi, F f) {
    // DPP scan (no LDS permutes); the same value in every lane
    return wave::dpp_wave_sum((lane >= lo && lane < hi) ? f(lane) : 0.);
  }
};

// expeig for block b: max(0, max Re(lambda(B_b)) - diffterm), written nout times to out[b*nout ..]
// (flamelet_kernels.cpp:1329-1341, which writes it once per point unknown: nout = n). Blocks are column-major
// n x n at blocks + b*bstride. With boff/ooff given, block b starts at blocks + boff[b] and its outputs at
// out + ooff[b]. kexp, when given, is laid out like out with nout = n and carries each block's balancing
// exponents from one launch to the next.
template <int LD>
__global__ __launch_bounds__(64) void expeig_kernel(const double *__restrict__ blocks, long bstride,
                                                    const long *__restrict__ boff, const long *__restrict__ ooff,
                                                    long nblk, int n, double diffterm, double *__restrict__ out,
                                                    int nout, int *__restrict__ nfail, float *__restrict__ kexp) {
  extern __shared__ double smem[];
  const std::integral_constant<int, LD> ld{};  // row stride of the LDS matrix, odd and >= n
  double *a = smem;
  WaveLanes par{(int)threadIdx.x};
  for (long b = blockIdx.x; b < nblk; b += gridDim.x) {
    const double *src = blocks + (boff ? boff[b] : b * bstride);
    if ((int)threadIdx.x < n)  // column j of the column-major block, lane = row: coalesced, no index division
      for (int j = 0; j < n; ++j) a[threadIdx.x * ld + j] = src[(long)j * n + threadIdx.x];
    __syncthreads();
    int failed = 0;
    const double em = eig_max_real(par, a, n, ld, &failed, kexp ? kexp + (ooff ? ooff[b] : b * n) : nullptr);
    const double ex = fmax(em - diffterm, 0.);
    double *dst = out + (ooff ? ooff[b] : b * nout);
    for (int q = threadIdx.x; q < nout; q += 64) dst[q] = ex;
    if (failed && threadIdx.x == 0 && nfail) atomicAdd(nfail, 1);
    __syncthreads();
  }
}

// Row strides the kernel is compiled for (odd, so column walks are free of bank conflicts); a block of size n
// uses the smallest one not below n | 1.
constexpr int kEigStrides[] = {9, 17, 25, 33, 41, 49, 53, 57, 61, 65};
int eig_stride(int n) {
  for (int s : kEigStrides)
    if (s >= (n | 1)) return s;
  return 65;
}
size_t expeig_lds_bytes(int n) { return (size_t)n * eig_stride(n) * sizeof(double); }

template <int LD>
int launch_expeig_ld(const gf_mech *m, const double *blocks, const long *boff, const long *ooff, long nblk, int n,
                     double diffterm, double *out, int *nfail, hipStream_t st, int nout, float *kexp) {
  const size_t lds = expeig_lds_bytes(n);
  // one wave per block, persistent: exactly as many workgroups as can be resident at once (LDS-limited: 22.0 KiB
  // per block at n = 53), each looping over its share of the blocks, so no workgroup waits for a second round
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, expeig_kernel<LD>, 64, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  long grid = (long)(m && m->n_cu > 0 ? m->n_cu : 256) * per_cu;
  if (grid > nblk) grid = nblk;
  hipLaunchKernelGGL(expeig_kernel<LD>, dim3((unsigned)grid), dim3(64), lds, st, blocks, (long)n * n, boff, ooff, nblk,
                     n, diffterm, out, nout, nfail, kexp);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "eigenvalue kernel launch failed");
}

int launch_expeig_list(const gf_mech *m, const double *blocks, const long *boff, const long *ooff, long nblk, int n,
                       double diffterm, double *out, int *nfail, hipStream_t st, int nout, float *kexp) {
  if (nout < 0) nout = n;
  if (nblk <= 0) return GF_OK;
  if (n < 1 || n > 64) return fail(GF_ERR_UNSUPPORTED, "eigenvalue blocks must be 1..64 wide");
  if (kexp && nout != n) return fail(GF_ERR_ARG, "carried balancing exponents need one output per block row");
  switch (eig_stride(n)) {
    case 9: return launch_expeig_ld<9>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 17: return launch_expeig_ld<17>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 25: return launch_expeig_ld<25>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 33: return launch_expeig_ld<33>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 41: return launch_expeig_ld<41>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 49: return launch_expeig_ld<49>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 53: return launch_expeig_ld<53>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 57: return launch_expeig_ld<57>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    case 61: return launch_expeig_ld<61>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
    default: return launch_expeig_ld<65>(m, blocks, boff, ooff, nblk, n, diffterm, out, nfail, st, nout, kexp);
  }
}

int launch_expeig(const gf_mech *m, const double *blocks, long bstride, long nblk, int n, double diffterm, double *out,
                  int *nfail, hipStream_t st) {
  if (bstride != (long)n * n) return fail(GF_ERR_ARG, "eigenvalue blocks must be contiguous");
  return launch_expeig_list(m, blocks, nullptr, nullptr, nblk, n, diffterm, out, nfail, st);
}

}  // namespace gfa
