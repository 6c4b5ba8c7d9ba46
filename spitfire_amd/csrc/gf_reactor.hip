// Per-system vector steps of the batched implicit stage solves (ESDIRK stages of many reactors solved by the
// frozen-Jacobian Newton method, nonlinear.py:185-268), fused so that one Newton iteration of the whole batch
// is a solve, an update, a chemistry launch and a residual launch. System s owns elements s*n .. s*n+n-1.
// The arithmetic is the Newton solver's, operation by operation with no contraction into fused multiply-adds,
// so results are bitwise those of the same steps written elementwise.
#include "gf_internal.h"
#include "../../include/griffon_amd.h"

namespace gfa {

namespace {

// x -= d for the systems not yet converged
__global__ void __launch_bounds__(256) masked_update_kernel(long total, int n, double *x, const double *d,
                                                            const bool *conv) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  if (!conv[e / n]) x[e] = x[e] - d[e];
}

// r = h_s (gamma f + prior) - (x - xn); with test set, conv[s] |= max_v |r_v w_v| < tol. One wave per system.
__global__ void __launch_bounds__(64) stage_residual_kernel(int R, int n, const double *h, double gamma,
                                                            const double *f, const double *prior, const double *x,
                                                            const double *xn, const double *w, double tol,
                                                            double *r, bool *conv, int test) {
#pragma clang fp contract(off)
  const int s = blockIdx.x;
  if (s >= R) return;
  const double hs = h[s];
  double m = 0.;
  int nan = 0;  // a NaN anywhere keeps the system unconverged, as a NaN maximum would
  for (int v = threadIdx.x; v < n; v += 64) {
    const long e = (long)s * n + v;
    const double rv = hs * (gamma * f[e] + prior[e]) - (x[e] - xn[e]);
    r[e] = rv;
    const double a = fabs(rv * w[v]);
    nan |= isnan(a);
    m = fmax(m, a);
  }
  if (!test) return;
  for (int o = 32; o > 0; o >>= 1) {
    m = fmax(m, __shfl_xor(m, o, 64));
    nan |= __shfl_xor(nan, o, 64);
  }
  if (threadIdx.x == 0 && !nan && m < tol) conv[s] = true;
}

// Discrete adjoint of one implicit stage i of one reactor (one wave, lane k = state variable; n <= 64):
//   drho_k, dcp_k: gradients of density and mixture cp w.r.t. (T, Y_0..Y_{n-2}) at constant pressure
//   gq_k = rho cp J(0, k) + f_T (cp drho_k + rho dcp_k)          (the gradient of qdot = rho cp f_T)
//   rhs_k = h (sum_row J(row, k) v_row + b_i gq_k)                 (h (J^T v + b_i grad qdot))
//   M = I - (h a_ii) J^T, column-major, when want_m
// J is column-major (row, column) at row + column*n; aux holds rho, cp, dcp/dT, mixture molecular weight.
__global__ void __launch_bounds__(64) adjoint_prep_kernel(int R, int n, const double *x, const double *f,
                                                          const double *J, const double *aux, const double *cpi,
                                                          const double *dinvm, const double *v, const double *h,
                                                          double bi, double aii, double *rhs, double *M,
                                                          int want_m) {
#pragma clang fp contract(off)
  const int r = blockIdx.x;
  if (r >= R) return;
  const int k = threadIdx.x;
  const double *Jr = J + (long)r * n * n;
  const double rho = aux[r * 4 + 0], cp = aux[r * 4 + 1], dcp = aux[r * 4 + 2], mmw = aux[r * 4 + 3];
  const double hr = h[r];
  if (k < n) {
    const double fT = f[(long)r * n];
    const double drho = k == 0 ? -rho / x[(long)r * n] : -(rho * mmw) * dinvm[k - 1];
    const double dcpv = k == 0 ? dcp : cpi[(long)r * n + k - 1] - cpi[(long)r * n + n - 1];
    const double gq = (rho * cp) * Jr[(long)k * n] + fT * (cp * drho + rho * dcpv);
    double jtv = 0.;
    for (int row = 0; row < n; ++row) jtv = jtv + Jr[(long)k * n + row] * v[(long)r * n + row];
    rhs[(long)r * n + k] = hr * (jtv + bi * gq);
  }
  if (!want_m) return;
  const double ha = hr * aii;
  for (int e = k; e < n * n; e += 64) {
    const int a = e / n, b = e - a * n;  // M[a*n + b] = (a == b) - ha J(a, b)
    M[(long)r * n * n + e] = (a == b ? 1. : 0.) - ha * Jr[(long)b * n + a];
  }
}

// The stage's share of the gradient and loss (one wave per reactor, lane k):
//   wv = v + a_ii theta;  wT = wv_0 + b_i rho cp
//   C_k = wv_{k+1} / rho (k < n-1), 0 (k = n-1);  C_k -= wT / (rho cp) * h_k;  cm_k = C_k M_k
//   L += h b_i rho cp f_T
__global__ void __launch_bounds__(64) adjoint_finish_kernel(int R, int n, const double *v, const double *theta,
                                                            const double *aux, const double *hsp, const double *mw,
                                                            const double *f, const double *h, double bi,
                                                            double aii, double *cm, double *L) {
#pragma clang fp contract(off)
  const int r = blockIdx.x;
  if (r >= R) return;
  const int k = threadIdx.x;
  const double rho = aux[r * 4 + 0], cp = aux[r * 4 + 1];
  const double wv = k < n ? v[(long)r * n + k] + aii * theta[(long)r * n + k] : 0.;
  const double wv0 = __shfl(wv, 0, 64);
  const double wnext = __shfl_down(wv, 1, 64);
  if (k < n) {
    const double wT = wv0 + bi * rho * cp;
    double C = k < n - 1 ? wnext / rho : 0.;
    C = C - (wT / (rho * cp)) * hsp[(long)r * n + k];
    cm[(long)r * n + k] = C * mw[k];
  }
  if (k == 0) L[r] = L[r] + h[r] * bi * rho * cp * f[(long)r * n];
}

}  // namespace

}  // namespace gfa

using namespace gfa;

extern "C" {

int gf_batch_masked_update(int R, int n, double *x, const double *d, const unsigned char *conv, void *stream) {
  if (R <= 0 || n <= 0) return GF_OK;
  if (!x || !d || !conv) return fail(GF_ERR_ARG, "null argument");
  const long total = (long)R * n;
  hipLaunchKernelGGL(masked_update_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     total, n, x, d, reinterpret_cast<const bool *>(conv));
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "masked update launch failed");
}

int gf_batch_stage_residual(int R, int n, const double *h, double gamma, const double *f, const double *prior,
                            const double *x, const double *xn, const double *w, double tol, double *r,
                            unsigned char *conv, int test, void *stream) {
  if (R <= 0 || n <= 0) return GF_OK;
  if (!h || !f || !prior || !x || !xn || !w || !r || (test && !conv)) return fail(GF_ERR_ARG, "null argument");
  hipLaunchKernelGGL(stage_residual_kernel, dim3((unsigned)R), dim3(64), 0, (hipStream_t)stream, R, n, h, gamma, f,
                     prior, x, xn, w, tol, r, reinterpret_cast<bool *>(conv), test);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "stage residual launch failed");
}

int gf_batch_adjoint_stage_prep(int R, int n, const double *x, const double *f, const double *jac, const double *aux,
                                const double *cpi, const double *dinvm, const double *v, const double *h, double b_i,
                                double a_ii, double *rhs, double *m, void *stream) {
  if (R <= 0) return GF_OK;
  if (n < 2 || n > 64) return fail(GF_ERR_UNSUPPORTED, "adjoint stage kernels need 2..64 state variables");
  if (!x || !f || !jac || !aux || !cpi || !dinvm || !v || !h || !rhs) return fail(GF_ERR_ARG, "null argument");
  hipLaunchKernelGGL(adjoint_prep_kernel, dim3((unsigned)R), dim3(64), 0, (hipStream_t)stream, R, n, x, f, jac, aux,
                     cpi, dinvm, v, h, b_i, a_ii, rhs, m, m ? 1 : 0);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "adjoint stage launch failed");
}

int gf_batch_adjoint_stage_finish(int R, int n, const double *v, const double *theta, const double *aux,
                                  const double *hsp, const double *mw, const double *f, const double *h, double b_i,
                                  double a_ii, double *cm, double *L, void *stream) {
  if (R <= 0) return GF_OK;
  if (n < 2 || n > 64) return fail(GF_ERR_UNSUPPORTED, "adjoint stage kernels need 2..64 state variables");
  if (!v || !theta || !aux || !hsp || !mw || !f || !h || !cm || !L) return fail(GF_ERR_ARG, "null argument");
  hipLaunchKernelGGL(adjoint_finish_kernel, dim3((unsigned)R), dim3(64), 0, (hipStream_t)stream, R, n, v, theta, aux,
                     hsp, mw, f, h, b_i, a_ii, cm, L);
  return hipGetLastError() == hipSuccess ? GF_OK : fail(GF_ERR_HIP, "adjoint stage launch failed");
}

}  // extern "C"
