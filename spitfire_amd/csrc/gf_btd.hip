// Batched block-tridiagonal factor/solve with diagonal off-diagonal blocks (griffon's BTDDOD format), with
// FP64 factors (in place, as the reference) or FP32 factors in separate arrays (the mixed-precision path of
// the flamelet solvers: FP32 block LU, FP64 right-hand sides, solutions and residuals).
//
// Algorithm restated from btddod_matrix_kernels.cpp:19-119 (block Thomas):
//   D_0 = LU(D_0);  for i >= 1:  L_i = diag(sub_{i-1}) D_{i-1}^{-1},  D_i -= L_i diag(sup_{i-1}),  D_i = LU(D_i)
//   forward  y_0 = b_0,  y_i = b_i - L_i y_{i-1};   backward  x_{n-1} = D^{-1} y,  x_i = D_i^{-1}(y_i - sup_i * x_{i+1})
// with LAPACK dgetrf/dgetrs semantics for each dense block (partial pivoting, first maximum, 1-based pivots,
// factors stored in place in the values array, column-major blocks).
//
// Factorization: one workgroup per system, the factored diagonal block and the L block live in LDS
// (2 x bs^2 values); right-looking LU with a wave-reduced pivot search; D^{-1} by column-parallel
// forward/back substitution on the identity. Solve: one wave per system, lane = row, each block staged in LDS
// by one coalesced load, register-resident vectors, shuffles broadcast the pivot/substitution values.
//
// The batched flamelet solver uses a throughput form of the same recurrence (btd_inv_*): each chain is
// eliminated from both ends at once (twisted factorisation, halving the serial chain), every Schur-updated
// diagonal block is inverted in registers by Gauss-Jordan elimination with partial pivoting (gf_twist_impl.h)
// and stored as its inverse, so no L blocks are stored and both sweeps of a solve are matrix-vector products.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"
#include "gf_wave.h"
#include "gf_twist.h"

#include <atomic>
#include <type_traits>
#include <utility>

namespace gfa {

namespace {

using wave::lane_value;

// Compute units of the current device, looked up once per device.
int compute_units() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n <= 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

constexpr int BW = 64;
constexpr int BT_THREADS = 256;

// LU with partial pivoting of a bs x bs column-major block in LDS (LAPACK dgetf2 order of operations).
template <typename T>
__device__ void lds_lu(T *A, int bs, int *piv_lds, int *gpiv) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  __shared__ int s_p;
  for (int j = 0; j < bs; ++j) {
    if (wave == 0) {
      double v = -1.0;
      int idx = bs;
      if (lane >= j && lane < bs) { v = fabs((double)A[j * bs + lane]); idx = lane; }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
      }
      if (idx >= bs) idx = j;  // no comparable pivot candidate (NaN column): keep the diagonal
      if (lane == 0) {
        s_p = idx;
        piv_lds[j] = idx + 1;
        if (gpiv) gpiv[j] = idx + 1;
      }
    }
    __syncthreads();
    const int p = s_p;
    if (p != j)
      for (int c = t; c < bs; c += BT_THREADS) {
        const T a0 = A[c * bs + j];
        A[c * bs + j] = A[c * bs + p];
        A[c * bs + p] = a0;
      }
    __syncthreads();
    const T d = A[j * bs + j];
    if (d != T(0)) {
      const T id = T(1) / d;
      for (int k = j + 1 + t; k < bs; k += BT_THREADS) A[j * bs + k] *= id;
    }
    __syncthreads();
    const int m = bs - j - 1;
    for (int e = t; e < m * m; e += BT_THREADS) {
      const int c = j + 1 + e / m, k = j + 1 + e % m;
      A[c * bs + k] -= A[j * bs + k] * A[c * bs + j];
    }
    __syncthreads();
  }
}

// X <- A^{-1} X for bs right-hand sides, A an LU factorization with pivots in LDS (LAPACK dgetrs, 'N').
template <typename T>
__device__ void lds_lu_solve_multi(const T *A, const int *piv, int bs, T *X) {
  const int t = threadIdx.x;
  for (int j = 0; j < bs; ++j) {  // row interchanges, applied in order
    const int p = piv[j] - 1;
    if (p != j)
      for (int c = t; c < bs; c += BT_THREADS) {
        const T a0 = X[c * bs + j];
        X[c * bs + j] = X[c * bs + p];
        X[c * bs + p] = a0;
      }
    __syncthreads();
  }
  for (int j = 0; j < bs; ++j) {  // unit lower triangular
    const int m = bs - j - 1;
    for (int e = t; e < m * bs; e += BT_THREADS) {
      const int c = e / m, k = j + 1 + e % m;
      X[c * bs + k] -= A[j * bs + k] * X[c * bs + j];
    }
    __syncthreads();
  }
  for (int j = bs - 1; j >= 0; --j) {  // upper triangular
    for (int c = t; c < bs; c += BT_THREADS) X[c * bs + j] /= A[j * bs + j];
    __syncthreads();
    for (int e = t; e < j * bs; e += BT_THREADS) {
      const int c = e / j, k = e % j;
      X[c * bs + k] -= A[j * bs + k] * X[c * bs + j];
    }
    __syncthreads();
  }
}

// Factors of system sys: D blocks to dout + doff[sys] (for FP64 the values array itself, in place), L blocks to
// lout + loff[sys]; the input blocks and off-diagonals are the FP64 values array.
template <typename T>
__global__ void __launch_bounds__(BT_THREADS) btd_factor_kernel(const int *nblocks, int bs, double *values,
                                                               const long *voff, T *dout, const long *doff, T *lout,
                                                               const long *loff, int *piv, const long *poff) {
  extern __shared__ double sm_raw[];
  T *sm = reinterpret_cast<T *>(sm_raw);
  const int sys = blockIdx.x;
  const int nb = nblocks[sys];
  const int nb2 = bs * bs;
  const double *V = values + voff[sys];
  T *Dg = dout + doff[sys];
  T *L = lout + loff[sys];
  int *P = piv + poff[sys];
  const double *sub = V + (long)nb * nb2;
  const double *sup = sub + (long)(nb - 1) * bs;
  T *D = sm, *Lb = sm + nb2;                // factored diagonal block, L block
  int *pv = reinterpret_cast<int *>(sm + 2 * nb2);
  const int t = threadIdx.x;

  for (int e = t; e < nb2; e += BT_THREADS) D[e] = (T)V[e];
  __syncthreads();
  lds_lu(D, bs, pv, P);
  for (int e = t; e < nb2; e += BT_THREADS) Dg[e] = D[e];
  for (int i = 1; i < nb; ++i) {
    const long o1 = (long)i * nb2;
    for (int e = t; e < nb2; e += BT_THREADS) Lb[e] = ((e % bs) == (e / bs)) ? T(1) : T(0);
    __syncthreads();
    lds_lu_solve_multi(D, pv, bs, Lb);  // Lb = D_{i-1}^{-1}
    const int b1 = (i - 1) * bs;
    for (int e = t; e < nb2; e += BT_THREADS) {
      const int j = e / bs, k = e % bs;
      const T l = Lb[e] * (T)sub[b1 + k];
      L[o1 + e] = l;
      D[e] = (T)(V[o1 + e] + (-sup[b1 + j]) * (double)l);
    }
    __syncthreads();
    lds_lu(D, bs, pv, P + i * bs);
    for (int e = t; e < nb2; e += BT_THREADS) Dg[o1 + e] = D[e];
  }
}

// One wave per system, lane = row; each block is staged in LDS with one coalesced load before it is used.
// FP64 arithmetic; FP32 factors are widened on load.
template <typename T>
__global__ void __launch_bounds__(BW) btd_solve_kernel(int nsys, const int *nblocks, int bs, const double *values,
                                                      const long *voff, const T *dfac, const long *doff,
                                                      const T *lval, const long *loff, const int *piv,
                                                      const long *poff, const double *rhs, double *x,
                                                      const long *xoff) {
  extern __shared__ double sm_raw[];
  T *Bk = reinterpret_cast<T *>(sm_raw);
  int *pvs = reinterpret_cast<int *>(Bk + bs * bs);
  const int sys = blockIdx.x;
  if (sys >= nsys) return;
  const int lane = threadIdx.x;
  const int nb = nblocks[sys];
  const int nb2 = bs * bs;
  const double *V = values + voff[sys];
  const T *DF = dfac + doff[sys];
  const T *L = lval + loff[sys];
  const int *P = piv + poff[sys];
  const double *b = rhs + xoff[sys];
  double *X = x + xoff[sys];
  const double *sup = V + (long)nb * nb2 + (long)(nb - 1) * bs;
  const bool act = lane < bs;

  // forward: y_i = b_i - L_i y_{i-1}; y kept in X
  double y = act ? b[lane] : 0.;
  if (act) X[lane] = y;
  for (int i = 1; i < nb; ++i) {
    const T *Li = L + (long)i * nb2;
    for (int e = lane; e < nb2; e += BW) Bk[e] = Li[e];
    __syncthreads();
    double yi = act ? b[(long)i * bs + lane] : 0.;
    for (int c = 0; c < bs; ++c) {
      const double yc = __shfl(y, c, 64);
      if (act) yi = yi + (double)Bk[c * bs + lane] * (-1. * yc);
    }
    y = yi;
    if (act) X[(long)i * bs + lane] = y;
    __syncthreads();
  }
  // backward
  double xn = 0.;
  for (int i = nb - 1; i >= 0; --i) {
    const T *Dg = DF + (long)i * nb2;
    for (int e = lane; e < nb2; e += BW) Bk[e] = Dg[e];
    if (act) pvs[lane] = P[i * bs + lane];
    __syncthreads();
    double r = act ? X[(long)i * bs + lane] : 0.;
    if (i < nb - 1 && act) r = r - sup[(long)i * bs + lane] * xn;
    for (int j = 0; j < bs; ++j) {  // row interchanges
      const int p = pvs[j] - 1;
      if (p != j) {
        const double vj = __shfl(r, j, 64), vp = __shfl(r, p, 64);
        if (lane == j) r = vp;
        else if (lane == p) r = vj;
      }
    }
    for (int j = 0; j < bs; ++j) {  // unit lower
      const double rj = __shfl(r, j, 64);
      if (act && lane > j) r -= (double)Bk[j * bs + lane] * rj;
    }
    for (int j = bs - 1; j >= 0; --j) {  // upper
      if (lane == j) r /= (double)Bk[j * bs + j];
      const double rj = __shfl(r, j, 64);
      if (act && lane < j) r -= (double)Bk[j * bs + lane] * rj;
    }
    xn = r;
    if (act) X[(long)i * bs + lane] = r;
    __syncthreads();
  }
}

// ---------------- block Thomas with explicit diagonal-block inverses ----------------

// y = D v for a column-major bs x bs block and v held one entry per lane (lane = row); the entries of v are
// broadcast through scalar registers and four partial sums keep the column loads independent.
template <typename T>
__device__ __forceinline__ double wave_matvec(const T *D, int bs, double v, int lane) {
  const bool act = lane < bs;
  double m0 = 0., m1 = 0., m2 = 0., m3 = 0.;
  int c = 0;
  // Sixteen columns per trip, their loads issued together before the products (column c still goes to the
  // partial sum of c mod 4, in column order, so the result is the same as four at a time). Lanes past the block
  // read row 0 and discard it: a load under a lane condition was compiled as a branch around each load for
  // FP32 blocks, which serialised them.
  const int rl = act ? lane : 0;
  for (; c + 15 < bs; c += 16) {
    double d[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) d[q] = (double)D[(c + q) * bs + rl];
#pragma unroll
    for (int q = 0; q < 16; q += 4) {
      const double v0 = lane_value(v, c + q), v1 = lane_value(v, c + q + 1), v2 = lane_value(v, c + q + 2),
                   v3 = lane_value(v, c + q + 3);
      if (act) {
        m0 += d[q] * v0;
        m1 += d[q + 1] * v1;
        m2 += d[q + 2] * v2;
        m3 += d[q + 3] * v3;
      }
    }
  }
  for (; c + 3 < bs; c += 4) {
    const double d0 = (double)D[c * bs + rl], d1 = (double)D[(c + 1) * bs + rl], d2 = (double)D[(c + 2) * bs + rl],
                 d3 = (double)D[(c + 3) * bs + rl];
    const double v0 = lane_value(v, c), v1 = lane_value(v, c + 1), v2 = lane_value(v, c + 2),
                 v3 = lane_value(v, c + 3);
    if (act) {
      m0 += d0 * v0;
      m1 += d1 * v1;
      m2 += d2 * v2;
      m3 += d3 * v3;
    }
  }
  for (; c < bs; ++c) {
    const double d0 = (double)D[c * bs + rl];
    const double v0 = lane_value(v, c);
    if (act) m0 += d0 * v0;
  }
  return (m0 + m1) + (m2 + m3);
}

// Solve with the twisted factors, one workgroup of two waves per system, lane = row:
//   wave 0:  y_0 = b_0,  y_i = b_i - sub_{i-1} (D_{i-1}^{-1} y_{i-1}),  i < m
//   wave 1:  y_{n-1} = b_{n-1},  y_i = b_i - sup_i (E_{i+1}^{-1} y_{i+1}),  i > m
//   middle:  x_m = C_m^{-1} (b_m - sub_{m-1} (D_{m-1}^{-1} y_{m-1}) - sup_m (E_{m+1}^{-1} y_{m+1}))
//   wave 0:  x_i = D_i^{-1} (y_i - sup_i x_{i+1}),  i = m-1 .. 0;   wave 1:  x_i = E_i^{-1} (y_i - sub_{i-1} x_{i-1}),  i > m
template <typename T>
__global__ void __launch_bounds__(128) btd_twist_solve_kernel(int nsys, const int *nblocks, int bs,
                                                              const double *values, const long *voff,
                                                              const T *dinv, const long *doff, const double *rhs,
                                                              double *x, const long *xoff) {
  __shared__ double mid[2][64];
  const int sys = blockIdx.x;
  if (sys >= nsys) return;
  const int lane = threadIdx.x & 63;
  const bool top = threadIdx.x < 64;
  const bool act = lane < bs;
  const int nb = nblocks[sys];
  const int m = (nb - 1) / 2;
  const long nb2 = (long)bs * bs;
  const double *V = values + voff[sys];
  const T *DI = dinv + doff[sys];
  const double *b = rhs + xoff[sys];
  double *X = x + xoff[sys];
  const double *sub = V + (long)nb * nb2;
  const double *sup = sub + (long)(nb - 1) * bs;

  // sweep towards the middle; y_i is kept in X
  const int len = top ? m : nb - 1 - m;
  double y = 0.;
  for (int t = 0; t < len; ++t) {
    const int i = top ? t : nb - 1 - t;
    const double bi = act ? b[(long)i * bs + lane] : 0.;
    if (t == 0) {
      y = bi;
    } else {
      const int prev = top ? i - 1 : i + 1;
      const double mv = wave_matvec(DI + prev * nb2, bs, y, lane);
      const double cpl = act ? (top ? sub[(long)(i - 1) * bs + lane] : sup[(long)i * bs + lane]) : 0.;
      y = bi - cpl * mv;
    }
    if (act) X[(long)i * bs + lane] = y;
  }
  // this wave's term of the middle equation
  double term = 0.;
  if (len > 0) {
    const int prev = top ? m - 1 : m + 1;
    const double mv = wave_matvec(DI + prev * nb2, bs, y, lane);
    term = act ? (top ? sub[(long)(m - 1) * bs + lane] : sup[(long)m * bs + lane]) * mv : 0.;
  }
  mid[top ? 0 : 1][lane] = term;
  __syncthreads();
  double xm = 0.;
  {
    const double r = act ? b[(long)m * bs + lane] - mid[0][lane] - mid[1][lane] : 0.;
    xm = wave_matvec(DI + m * nb2, bs, r, lane);  // both waves form x_m (no second exchange)
  }
  if (top && act) X[(long)m * bs + lane] = xm;
  // sweep outwards
  double xn = xm;
  for (int t = len - 1; t >= 0; --t) {
    const int i = top ? t : nb - 1 - t;
    const double yi = act ? X[(long)i * bs + lane] : 0.;
    const double cpl = act ? (top ? sup[(long)i * bs + lane] : sub[(long)(i - 1) * bs + lane]) : 0.;
    xn = wave_matvec(DI + i * nb2, bs, yi - cpl * xn, lane);
    if (act) X[(long)i * bs + lane] = xn;
  }
}

}  // namespace

int launch_btd_factor(int nsys, const int *nblocks, int bs, double *values, const long *voff, double *l_values,
                      const long *loff, int *pivots, const long *poff, hipStream_t st) {
  return launch_btd_factor_t<double>(nsys, nblocks, bs, values, voff, values, voff, l_values, loff, pivots, poff, st);
}

int launch_btd_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                     const double *l_values, const long *loff, const int *pivots, const long *poff, const double *rhs,
                     double *x, const long *xoff, hipStream_t st) {
  return launch_btd_solve_t<double>(nsys, nblocks, bs, values, voff, values, voff, l_values, loff, pivots, poff, rhs,
                                    x, xoff, st);
}

template <typename T>
int launch_btd_factor_t(int nsys, const int *nblocks, int bs, double *values, const long *voff, T *dfac,
                        const long *doff, T *l_values, const long *loff, int *pivots, const long *poff,
                        hipStream_t st) {
  if (nsys <= 0) return GF_OK;
  if (bs < 1 || bs > 64) return fail(GF_ERR_UNSUPPORTED, "block size must be in 1..64");
  const size_t lds = sizeof(T) * 2 * bs * bs + sizeof(int) * 64 + 16;
  auto kern = btd_factor_kernel<T>;
  if (lds > 65536)
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return fail(GF_ERR_UNSUPPORTED, "block too large for LDS");
  hipLaunchKernelGGL(kern, dim3(nsys), dim3(BT_THREADS), lds, st, nblocks, bs, values, voff, dfac, doff, l_values,
                     loff, pivots, poff);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd factor launch failed");
  return GF_OK;
}

template <typename T>
int launch_btd_solve_t(int nsys, const int *nblocks, int bs, const double *values, const long *voff, const T *dfac,
                       const long *doff, const T *l_values, const long *loff, const int *pivots, const long *poff,
                       const double *rhs, double *x, const long *xoff, hipStream_t st) {
  if (nsys <= 0) return GF_OK;
  if (bs < 1 || bs > 64) return fail(GF_ERR_UNSUPPORTED, "block size must be in 1..64");
  const size_t lds = sizeof(T) * bs * bs + sizeof(int) * 64 + 16;
  hipLaunchKernelGGL(btd_solve_kernel<T>, dim3(nsys), dim3(BW), lds, st, nsys, nblocks, bs, values, voff, dfac, doff,
                     l_values, loff, pivots, poff, rhs, x, xoff);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd solve launch failed");
  return GF_OK;
}

template <typename T>
using TwistFactorFn = void (*)(int, const int *, const double *, const long *, T *, const long *, bool,
                               hipStream_t);

template <typename T, int... I>
constexpr TwistFactorFn<T> twist_factor_table_entry(int bs, std::integer_sequence<int, I...>) {
  constexpr TwistFactorFn<T> tab[] = {&launch_twist_factor<T, I + 1>...};
  return tab[bs - 1];
}

template <typename T>
int launch_btd_inv_factor(int nsys, const int *nblocks, int bs, const double *values, const long *voff, T *dinv,
                          const long *doff, hipStream_t st) {
  if (nsys <= 0) return GF_OK;
  if (bs < 1 || bs > 64) return fail(GF_ERR_UNSUPPORTED, "block size must be in 1..64");
  // Two waves per chain shorten every pivot step, but their barriers cost issue slots: measured on 53x53
  // blocks, the split kernel wins up to 512 systems (chains of 62..254 blocks: 7.5 against 11.2 ms) and loses
  // from 1024 systems on (single blocks: 0.105 against 0.057 ms; chains: 13.2 against 11.9 ms), i.e. once
  // the systems alone give every SIMD (four per compute unit) a chain.
  const bool split = nsys < 4 * compute_units();
  twist_factor_table_entry<T>(bs, std::make_integer_sequence<int, 64>{})(nsys, nblocks, values, voff, dinv, doff,
                                                                          split, st);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd inverse factor launch failed");
  return GF_OK;
}

template <typename T>
int launch_btd_inv_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                         const T *dinv, const long *doff, const double *rhs, double *x, const long *xoff,
                         hipStream_t st) {
  if (nsys <= 0) return GF_OK;
  if (bs < 1 || bs > 64) return fail(GF_ERR_UNSUPPORTED, "block size must be in 1..64");
  hipLaunchKernelGGL(btd_twist_solve_kernel<T>, dim3(nsys), dim3(128), 0, st, nsys, nblocks, bs, values, voff, dinv,
                     doff, rhs, x, xoff);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd inverse solve launch failed");
  return GF_OK;
}

template int launch_btd_inv_factor<double>(int, const int *, int, const double *, const long *, double *, const long *,
                                           hipStream_t);
template int launch_btd_inv_factor<float>(int, const int *, int, const double *, const long *, float *, const long *,
                                          hipStream_t);
template int launch_btd_inv_solve<double>(int, const int *, int, const double *, const long *, const double *,
                                          const long *, const double *, double *, const long *, hipStream_t);
template int launch_btd_inv_solve<float>(int, const int *, int, const double *, const long *, const float *,
                                         const long *, const double *, double *, const long *, hipStream_t);
template int launch_btd_factor_t<float>(int, const int *, int, double *, const long *, float *, const long *, float *,
                                        const long *, int *, const long *, hipStream_t);
template int launch_btd_solve_t<float>(int, const int *, int, const double *, const long *, const float *,
                                       const long *, const float *, const long *, const int *, const long *,
                                       const double *, double *, const long *, hipStream_t);

}  // namespace gfa
