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
// Factorization: one workgroup per system, the factored diagonal block and the L block live in
// LDS (2 x bs^2 doubles); right-looking LU with a wave-reduced pivot search; D^{-1} by column-parallel
// forward/back substitution on the identity. Solve: one wave per system, lane = row, register-resident
// vectors, shuffles broadcast the pivot/substitution values.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"

namespace gfa {

namespace {

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

// one wave per system
// (FP64 arithmetic; FP32 factors are widened on load)
template <typename T>
__global__ void __launch_bounds__(64) btd_solve_kernel(int nsys, const int *nblocks, int bs, const double *values,
                                                       const long *voff, const T *dfac, const long *doff,
                                                       const T *lval, const long *loff, const int *piv,
                                                       const long *poff, const double *rhs, double *x,
                                                       const long *xoff) {
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
    double yi = act ? b[(long)i * bs + lane] : 0.;
    for (int c = 0; c < bs; ++c) {
      const double yc = __shfl(y, c, 64);
      if (act) yi = yi + (double)Li[c * bs + lane] * (-1. * yc);
    }
    y = yi;
    if (act) X[(long)i * bs + lane] = y;
  }
  // backward
  double xn = 0.;
  for (int i = nb - 1; i >= 0; --i) {
    const T *D = DF + (long)i * nb2;
    const int *Pi = P + i * bs;
    double r = act ? X[(long)i * bs + lane] : 0.;
    if (i < nb - 1 && act) r = r - sup[(long)i * bs + lane] * xn;
    for (int j = 0; j < bs; ++j) {  // row interchanges
      const int p = Pi[j] - 1;
      if (p != j) {
        const double vj = __shfl(r, j, 64), vp = __shfl(r, p, 64);
        if (lane == j) r = vp;
        else if (lane == p) r = vj;
      }
    }
    for (int j = 0; j < bs; ++j) {  // unit lower
      const double rj = __shfl(r, j, 64);
      if (act && lane > j) r -= (double)D[j * bs + lane] * rj;
    }
    for (int j = bs - 1; j >= 0; --j) {  // upper
      if (lane == j) r /= (double)D[j * bs + j];
      const double rj = __shfl(r, j, 64);
      if (act && lane < j) r -= (double)D[j * bs + lane] * rj;
    }
    xn = r;
    if (act) X[(long)i * bs + lane] = r;
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
  hipLaunchKernelGGL(btd_solve_kernel<T>, dim3(nsys), dim3(64), 0, st, nsys, nblocks, bs, values, voff, dfac, doff,
                     l_values, loff, pivots, poff, rhs, x, xoff);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd solve launch failed");
  return GF_OK;
}

template int launch_btd_factor_t<float>(int, const int *, int, double *, const long *, float *, const long *, float *,
                                        const long *, int *, const long *, hipStream_t);
template int launch_btd_solve_t<float>(int, const int *, int, const double *, const long *, const float *,
                                       const long *, const float *, const long *, const int *, const long *,
                                       const double *, double *, const long *, hipStream_t);

}  // namespace gfa
