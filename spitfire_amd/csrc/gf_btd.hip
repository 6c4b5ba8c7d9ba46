// Batched block-tridiagonal factor/solve with diagonal off-diagonal blocks (griffon's BTDDOD format), FP64.
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
__device__ void lds_lu(double *A, int bs, int *piv_lds, int *gpiv) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  __shared__ int s_p;
  for (int j = 0; j < bs; ++j) {
    if (wave == 0) {
      double v = -1.0;
      int idx = bs;
      if (lane >= j && lane < bs) { v = fabs(A[j * bs + lane]); idx = lane; }
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
        const double a0 = A[c * bs + j];
        A[c * bs + j] = A[c * bs + p];
        A[c * bs + p] = a0;
      }
    __syncthreads();
    const double d = A[j * bs + j];
    if (d != 0.) {
      const double id = 1. / d;
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
__device__ void lds_lu_solve_multi(const double *A, const int *piv, int bs, double *X) {
  const int t = threadIdx.x;
  for (int j = 0; j < bs; ++j) {  // row interchanges, applied in order
    const int p = piv[j] - 1;
    if (p != j)
      for (int c = t; c < bs; c += BT_THREADS) {
        const double a0 = X[c * bs + j];
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

__global__ void __launch_bounds__(BT_THREADS) btd_factor_kernel(const int *nblocks, int bs, double *values,
                                                               const long *voff, double *lval, const long *loff,
                                                               int *piv, const long *poff) {
  extern __shared__ double sm[];
  const int sys = blockIdx.x;
  const int nb = nblocks[sys];
  const int nb2 = bs * bs;
  double *V = values + voff[sys];
  double *L = lval + loff[sys];
  int *P = piv + poff[sys];
  const double *sub = V + (long)nb * nb2;
  const double *sup = sub + (long)(nb - 1) * bs;
  double *D = sm, *Lb = sm + nb2;                // factored diagonal block, L block
  int *pv = reinterpret_cast<int *>(sm + 2 * nb2);
  const int t = threadIdx.x;

  for (int e = t; e < nb2; e += BT_THREADS) D[e] = V[e];
  __syncthreads();
  lds_lu(D, bs, pv, P);
  for (int e = t; e < nb2; e += BT_THREADS) V[e] = D[e];
  for (int i = 1; i < nb; ++i) {
    const long o1 = (long)i * nb2;
    for (int e = t; e < nb2; e += BT_THREADS) Lb[e] = ((e % bs) == (e / bs)) ? 1. : 0.;
    __syncthreads();
    lds_lu_solve_multi(D, pv, bs, Lb);  // Lb = D_{i-1}^{-1}
    const int b1 = (i - 1) * bs;
    for (int e = t; e < nb2; e += BT_THREADS) {
      const int j = e / bs, k = e % bs;
      const double l = Lb[e] * sub[b1 + k];
      L[o1 + e] = l;
      D[e] = V[o1 + e] + (-sup[b1 + j]) * l;
    }
    __syncthreads();
    lds_lu(D, bs, pv, P + i * bs);
    for (int e = t; e < nb2; e += BT_THREADS) V[o1 + e] = D[e];
  }
}

// one wave per system
__global__ void __launch_bounds__(64) btd_solve_kernel(int nsys, const int *nblocks, int bs, const double *values,
                                                       const long *voff, const double *lval, const long *loff,
                                                       const int *piv, const long *poff, const double *rhs,
                                                       double *x, const long *xoff) {
  const int sys = blockIdx.x;
  if (sys >= nsys) return;
  const int lane = threadIdx.x;
  const int nb = nblocks[sys];
  const int nb2 = bs * bs;
  const double *V = values + voff[sys];
  const double *L = lval + loff[sys];
  const int *P = piv + poff[sys];
  const double *b = rhs + xoff[sys];
  double *X = x + xoff[sys];
  const double *sup = V + (long)nb * nb2 + (long)(nb - 1) * bs;
  const bool act = lane < bs;

  // forward: y_i = b_i - L_i y_{i-1}; y kept in X
  double y = act ? b[lane] : 0.;
  if (act) X[lane] = y;
  for (int i = 1; i < nb; ++i) {
    const double *Li = L + (long)i * nb2;
    double yi = act ? b[(long)i * bs + lane] : 0.;
    for (int c = 0; c < bs; ++c) {
      const double yc = __shfl(y, c, 64);
      if (act) yi = yi + Li[c * bs + lane] * (-1. * yc);
    }
    y = yi;
    if (act) X[(long)i * bs + lane] = y;
  }
  // backward
  double xn = 0.;
  for (int i = nb - 1; i >= 0; --i) {
    const double *D = V + (long)i * nb2;
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
      if (act && lane > j) r -= D[j * bs + lane] * rj;
    }
    for (int j = bs - 1; j >= 0; --j) {  // upper
      if (lane == j) r /= D[j * bs + j];
      const double rj = __shfl(r, j, 64);
      if (act && lane < j) r -= D[j * bs + lane] * rj;
    }
    xn = r;
    if (act) X[(long)i * bs + lane] = r;
  }
}

}  // namespace

int launch_btd_factor(int nsys, const int *nblocks, int bs, double *values, const long *voff, double *l_values,
                      const long *loff, int *pivots, const long *poff, hipStream_t st) {
  if (nsys <= 0) return GF_OK;
  if (bs < 1 || bs > 64) return fail(GF_ERR_UNSUPPORTED, "block size must be in 1..64");
  const size_t lds = sizeof(double) * 2 * bs * bs + sizeof(int) * 64;
  if (lds > 65536)
    if (hipFuncSetAttribute((const void *)btd_factor_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
      return fail(GF_ERR_UNSUPPORTED, "block too large for LDS");
  hipLaunchKernelGGL(btd_factor_kernel, dim3(nsys), dim3(BT_THREADS), lds, st, nblocks, bs, values, voff, l_values,
                     loff, pivots, poff);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd factor launch failed");
  return GF_OK;
}

int launch_btd_solve(int nsys, const int *nblocks, int bs, const double *values, const long *voff,
                     const double *l_values, const long *loff, const int *pivots, const long *poff, const double *rhs,
                     double *x, const long *xoff, hipStream_t st) {
  if (nsys <= 0) return GF_OK;
  if (bs < 1 || bs > 64) return fail(GF_ERR_UNSUPPORTED, "block size must be in 1..64");
  hipLaunchKernelGGL(btd_solve_kernel, dim3(nsys), dim3(64), 0, st, nsys, nblocks, bs, values, voff, l_values, loff,
                     pivots, poff, rhs, x, xoff);
  if (hipGetLastError() != hipSuccess) return fail(GF_ERR_HIP, "btd solve launch failed");
  return GF_OK;
}

}  // namespace gfa
