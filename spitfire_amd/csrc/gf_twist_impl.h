// Kernels behind launch_twist_factor; included only by the gf_twist_*.hip units, each of which instantiates
// a range of block sizes with GF_TWIST(N).
#pragma once
#include "gf_twist.h"
#include "gf_wave.h"

namespace gfa {

namespace {

using wave::lane_value;
using wave::wave_first_argmax;

// Gauss-Jordan inverse of an N x N block held in the registers of one wave, with partial (row) pivoting.
//
// Layout: lane l holds one row, x[k] the entry in slot k. The entry of lane l, slot k is element
// (rowid of lane l, colid of slot k) of the logical matrix, where colid of slot k is held by lane k of
// `colid`. Rows are never interchanged: the pivot of step j is the unused lane with the largest |entry| in
// the pivot column (ties: lowest lane), and the exchange at (pivot lane p, slot j) is done in place. After N
// steps the block holds the inverse in the same kind of layout: the lane pivoted at step j holds row
// colid_j of the inverse, slot j holds its column rowid_{p_j}.
//
// The slots form a queue: the pivot column is always x[0]; each step shifts the other columns down by one
// and appends the transformed pivot column at x[N-1], so that all register indices are compile-time
// constants and after N steps every column is back in its slot. The pivot row is broadcast through scalar
// registers (readlane), so a step has no LDS traffic and no barrier.
//
// Row scaling is deferred: lane r keeps its row unscaled and a factor sig_r, the true row being sig_r times
// the stored one. Every row is pivoted exactly once and has sig = 1 until then, so at the step that pivots
// lane p (stored pivot value d):
//   rows r != p:   x_r[c] -= g_r x_p[c] (c != k),  x_r[k] = -g_r,  with g_r = x_r[k] / d
//   row p:         x_p[c] unchanged,  x_p[k] = 1,  sig_p = 1 / d
// which is the usual step (row p divided by its pivot) with the division of row p left to the end: the update
// of a slot is one fma, with no separate scaling of the pivot row. After N steps every row is multiplied by
// its sig.
template <typename T, int N>
__device__ __forceinline__ void reg_gj_inverse(T (&x)[N], int &rowid, int &colid, int lane) {
  bool used = lane >= N;
  T sig = T(1);
  for (int j = 0; j < N; ++j) {
    const T f = x[0];
    const T rf = T(1) / f;  // every lane's reciprocal, formed beside the pivot search instead of after it
    const int p = wave_first_argmax(fabs(f), !used);  // N - j unused lanes remain, so p >= 0
    const T ipv = lane_value(rf, p);
    const bool isp = lane == p;
    const T g = isp ? T(0) : f * ipv;  // multiplier of the broadcast pivot row (none for the pivot row)
#pragma unroll
    for (int c = 1; c < N; ++c) x[c - 1] = fma(-g, lane_value(x[c], p), x[c]);
    x[N - 1] = isp ? T(1) : -g;  // the pivot column
    if (isp) sig = ipv;
    const int cj = __builtin_amdgcn_readlane(colid, j);
    const int rp = __builtin_amdgcn_readlane(rowid, p);
    if (isp) {
      rowid = cj;
      used = true;
    }
    if (lane == j) colid = rp;
  }
#pragma unroll
  for (int c = 0; c < N; ++c) x[c] *= sig;
}

// Block i of the values in the identity layout (lane = row, slot = column).
template <typename T, int N>
__device__ __forceinline__ void reg_load_block(T (&x)[N], const double *Vi, int &rowid, int &colid, int lane) {
  rowid = lane;
  colid = lane;
#pragma unroll
  for (int k = 0; k < N; ++k) x[k] = lane < N ? (T)Vi[k * N + lane] : T(0);
}

// x <- V_i - diag(rs) X diag(cs), where x holds X (an inverse from the previous block of the chain) in its
// permuted layout; the result keeps that layout.
template <typename T, int N>
__device__ __forceinline__ void reg_schur(T (&x)[N], const double *Vi, const double *rs, const double *cs,
                                          int rowid, int colid, int lane) {
  const bool act = lane < N;
  const double r = act ? rs[rowid] : 0.;
  const double csl = act ? cs[lane] : 0.;  // lane q holds cs[q]
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int c = __builtin_amdgcn_readlane(colid, k);
    const double cc = lane_value(csl, c);
    if (act) x[k] = (T)(Vi[c * N + rowid] - r * cc * (double)x[k]);
  }
}

template <typename T, int N>
__device__ __forceinline__ void reg_store_block(const T (&x)[N], T *Di, int rowid, int colid, int lane) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int c = __builtin_amdgcn_readlane(colid, k);
    if (lane < N) Di[c * N + rowid] = x[k];
  }
}

// Twisted (burn-at-both-ends) block factorisation, one workgroup of two waves per system. Row i of the
// system is diag(sub_{i-1}) x_{i-1} + V_i x_i + diag(sup_i) x_{i+1} = b_i; with m = (n - 1) / 2
//   wave 0:  D_0 = V_0,      D_i = V_i - diag(sub_{i-1}) D_{i-1}^{-1} diag(sup_{i-1}),  i = 1 .. m-1
//   wave 1:  E_{n-1} = V_{n-1},  E_i = V_i - diag(sup_i) E_{i+1}^{-1} diag(sub_i),      i = n-2 .. m+1
//   middle:  C_m = V_m - diag(sub_{m-1}) D_{m-1}^{-1} diag(sup_{m-1}) - diag(sup_m) E_{m+1}^{-1} diag(sub_m)
// and block i of dinv receives D_i^{-1} (i < m), C_m^{-1} (i = m) or E_i^{-1} (i > m), column-major. The top
// half is the reference's forward block Thomas recurrence; running the bottom half at the same time halves
// the serial chain. The bottom wave hands its term of C_m to the top wave through LDS.
template <typename T, int N>
__global__ void __launch_bounds__(128) btd_twist_factor_kernel(const int *nblocks, const double *values,
                                                               const long *voff, T *dinv, const long *doff) {
  __shared__ double sb[N * N];
  constexpr long NB2 = (long)N * N;
  const int sys = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const bool top = threadIdx.x < 64;
  const int nb = nblocks[sys];
  const int m = (nb - 1) / 2;
  const double *V = values + voff[sys];
  const double *sub = V + nb * NB2;
  const double *sup = sub + (long)(nb - 1) * N;
  T *Dg = dinv + doff[sys];
  T x[N];
  int rowid = lane, colid = lane;
  const int len = top ? m : nb - 1 - m;
  for (int t = 0; t < len; ++t) {
    const int i = top ? t : nb - 1 - t;
    if (t == 0) reg_load_block<T, N>(x, V + i * NB2, rowid, colid, lane);
    else if (top) reg_schur<T, N>(x, V + i * NB2, sub + (i - 1) * N, sup + (i - 1) * N, rowid, colid, lane);
    else reg_schur<T, N>(x, V + i * NB2, sup + i * N, sub + i * N, rowid, colid, lane);
    reg_gj_inverse<T, N>(x, rowid, colid, lane);
    reg_store_block<T, N>(x, Dg + i * NB2, rowid, colid, lane);
  }
  const bool below = nb - 1 - m > 0;
  if (!top && below) {
    const double r = lane < N ? sup[m * N + rowid] : 0.;
    const double csl = lane < N ? sub[m * N + lane] : 0.;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const int c = __builtin_amdgcn_readlane(colid, k);
      const double cc = lane_value(csl, c);
      if (lane < N) sb[c * N + rowid] = r * (double)x[k] * cc;
    }
  }
  __syncthreads();
  if (top) {
    if (m == 0) reg_load_block<T, N>(x, V + m * NB2, rowid, colid, lane);
    else reg_schur<T, N>(x, V + m * NB2, sub + (m - 1) * N, sup + (m - 1) * N, rowid, colid, lane);
    if (below) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const int c = __builtin_amdgcn_readlane(colid, k);
        if (lane < N) x[k] = (T)((double)x[k] - sb[c * N + rowid]);
      }
    }
    reg_gj_inverse<T, N>(x, rowid, colid, lane);
    reg_store_block<T, N>(x, Dg + m * NB2, rowid, colid, lane);
  }
}

// ---- Column-split variant: W waves share each block inverse ----
//
// The single-wave inverse above is a serial chain of N pivot steps, each a few hundred instructions of one
// wave (at N = 53: 52 slot updates of two readlanes and an fma each, plus the pivot search and the reciprocal),
// and a flamelet system has only one such chain per half. Splitting the columns of the block over W waves
// divides the update work of a step by W: wave w holds slots w*C .. w*C + C - 1 (C = ceil(N / W)) of every
// lane's row. Step j is pivoted by the wave that holds slot j; it publishes the multipliers g (one per lane;
// the pivot lane's entry carries its scale 1 / d) and the pivot lane through LDS, and after one workgroup
// barrier every wave updates its own slots. The arithmetic of every entry is the one of the single-wave
// routine (same pivot, same g, same fma, same final scaling), so the inverse is bitwise the same.
//
// All waves compute a pivot candidate on their own slot of step j (only the owner publishes it): that work
// sits in the same basic block as the updates of step j - 1, where the compiler can interleave it, and the
// other waves would otherwise wait at the barrier. The multipliers are double-buffered on the step parity
// `ph` (carried across calls), so one barrier per step suffices: the buffer written at step j was last read
// at step j - 2, before every wave passed the barrier of step j - 1.
template <int N>
constexpr int twist_waves() {
  return N >= 12 ? 2 : 1;
}

template <typename T, int N, int W>
__device__ __forceinline__ void split_gj_inverse(T (&x)[(N + W - 1) / W], int &rowid, int &colid, int lane, int w,
                                                 T *gb, int *pb, int &ph) {
  constexpr int C = (N + W - 1) / W;
  bool used = lane >= N;
  T sig = T(1);
  for (int ww = 0; ww < W; ++ww) {
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int j = ww * C + k;
      if (j < N) {
        {
          const T f = x[k];
          const T rf = T(1) / f;  // every lane's reciprocal, formed beside the pivot search instead of after it
          const int p = wave_first_argmax(fabs(f), !used);  // N - j unused lanes remain, so p >= 0
          const T ipv = lane_value(rf, p);
          if (w == ww) {
            gb[ph * 64 + lane] = lane == p ? ipv : f * ipv;  // g of each row; the pivot lane carries 1 / d
            if (lane == 0) pb[ph] = p;
          }
        }
        __syncthreads();
        const int p = __builtin_amdgcn_readfirstlane(pb[ph]);
        const T gl = gb[ph * 64 + lane];
        ph ^= 1;
        const bool isp = lane == p;
        const T g = isp ? T(0) : gl;
#pragma unroll
        for (int c = 0; c < C; ++c) x[c] = fma(-g, lane_value(x[c], p), x[c]);
        if (w == ww) x[k] = isp ? T(1) : -g;  // the pivot column
        if (isp) sig = gl;
        const int cj = __builtin_amdgcn_readlane(colid, j);
        const int rp = __builtin_amdgcn_readlane(rowid, p);
        if (isp) {
          rowid = cj;
          used = true;
        }
        if (lane == j) colid = rp;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) x[c] *= sig;
}

// Slot k of wave w is logical slot w*C + k; slots at or beyond N are padding and stay zero.
template <typename T, int N, int W>
__device__ __forceinline__ void split_load_block(T (&x)[(N + W - 1) / W], const double *Vi, int &rowid, int &colid,
                                                 int lane, int w) {
  constexpr int C = (N + W - 1) / W;
  rowid = lane;
  colid = lane;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int s = w * C + k;
    x[k] = (lane < N && s < N) ? (T)Vi[s * N + lane] : T(0);
  }
}

// The identity: the block a chain inverts while the other chain inverts a real one, so that both run the same
// number of barriers.
template <typename T, int N, int W>
__device__ __forceinline__ void split_load_identity(T (&x)[(N + W - 1) / W], int &rowid, int &colid, int lane,
                                                    int w) {
  constexpr int C = (N + W - 1) / W;
  rowid = lane;
  colid = lane;
#pragma unroll
  for (int k = 0; k < C; ++k) x[k] = (lane == w * C + k && lane < N) ? T(1) : T(0);
}

template <typename T, int N, int W>
__device__ __forceinline__ void split_schur(T (&x)[(N + W - 1) / W], const double *Vi, const double *rs,
                                            const double *cs, int rowid, int colid, int lane, int w) {
  constexpr int C = (N + W - 1) / W;
  const bool act = lane < N;
  const double r = act ? rs[rowid] : 0.;
  const double csl = act ? cs[lane] : 0.;  // lane q holds cs[q]
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int s = w * C + k;
    if (s < N) {
      const int c = __builtin_amdgcn_readlane(colid, s);
      const double cc = lane_value(csl, c);
      if (act) x[k] = (T)(Vi[c * N + rowid] - r * cc * (double)x[k]);
    }
  }
}

template <typename T, int N, int W>
__device__ __forceinline__ void split_store_block(const T (&x)[(N + W - 1) / W], T *Di, int rowid, int colid,
                                                  int lane, int w) {
  constexpr int C = (N + W - 1) / W;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int s = w * C + k;
    if (s < N) {
      const int c = __builtin_amdgcn_readlane(colid, s);
      if (lane < N) Di[c * N + rowid] = x[k];
    }
  }
}

// The twisted factorisation of btd_twist_factor_kernel with W waves per chain (2W per workgroup: waves
// 0 .. W-1 run the top chain, W .. 2W-1 the bottom chain). Both chains run max(len) inverses so that every
// wave passes the same barriers: the shorter chain starts with one inverse of the identity, which it
// discards (it comes first, so the last real inverse stays in registers for the middle block). The middle
// block is inverted by the top chain alone.
template <typename T, int N, int W>
__global__ void __launch_bounds__(128 * W) __attribute__((amdgpu_waves_per_eu(2))) btd_twist_factor_split_kernel(const int *nblocks, const double *values,
                                                                         const long *voff, T *dinv,
                                                                         const long *doff) {
  constexpr int C = (N + W - 1) / W;
  constexpr long NB2 = (long)N * N;
  __shared__ double sb[N * N];
  __shared__ T gbuf[2][2 * 64];  // [chain][parity * 64 + lane]
  __shared__ int pbuf[2][2];     // [chain][parity]
  const int sys = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int chain = wave / W;
  const int w = wave - chain * W;
  const bool top = chain == 0;
  const int nb = nblocks[sys];
  const int m = (nb - 1) / 2;
  const double *V = values + voff[sys];
  const double *sub = V + nb * NB2;
  const double *sup = sub + (long)(nb - 1) * N;
  T *Dg = dinv + doff[sys];
  T *gb = gbuf[chain];
  int *pb = pbuf[chain];
  T x[C];
  int rowid = lane, colid = lane, ph = 0;
  const int len = top ? m : nb - 1 - m;
  const int iters = nb - 1 - m;  // >= m
  // Cache warming: while a block is inverted, the chain's waves read one word of every 128-byte line of the
  // block they update next (its values and the two scale vectors), so that its Schur update after this
  // inverse finds them in cache instead of waiting on HBM. The loads are volatile (not removed or moved
  // after the inverse) and their results are consumed only after it.
  auto warm = [&](int t) -> unsigned {
    int i = -1;
    const double *rs = nullptr, *cs = nullptr;
    if (t == iters) {
      if (top) {
        i = m;
        if (m > 0) { rs = sub + (m - 1) * N; cs = sup + (m - 1) * N; }
      }
    } else {
      const int tt = t - (iters - len);
      if (tt >= 0) {
        i = top ? tt : nb - 1 - tt;
        if (tt > 0) {
          if (top) { rs = sub + (i - 1) * N; cs = sup + (i - 1) * N; }
          else { rs = sup + i * N; cs = sub + i * N; }
        }
      }
    }
    if (i < 0) return 0u;
    constexpr int LB = (int)((NB2 * 8 + 127) / 128);  // lines of a block
    constexpr int LV = (N * 8 + 127) / 128;           // lines of a scale vector
    unsigned acc = 0u;
    for (int u = w * 64 + lane; u < LB + 2 * LV; u += W * 64) {
      const char *a;
      if (u < LB) a = reinterpret_cast<const char *>(V + i * NB2) + min((long)u * 128, NB2 * 8 - 4);
      else {
        const double *vec = u < LB + LV ? rs : cs;
        if (!vec) continue;
        a = reinterpret_cast<const char *>(vec) + min((u - LB - (u < LB + LV ? 0 : LV)) * 128, N * 8 - 4);
      }
      acc += *reinterpret_cast<const volatile unsigned *>(a);
    }
    return acc;
  };
  for (int t = 0; t < iters; ++t) {
    const int tt = t - (iters - len);  // < 0: the leading identity of the shorter chain
    const int i = top ? tt : nb - 1 - tt;
    if (tt < 0) split_load_identity<T, N, W>(x, rowid, colid, lane, w);
    else if (tt == 0) split_load_block<T, N, W>(x, V + i * NB2, rowid, colid, lane, w);
    else if (top) split_schur<T, N, W>(x, V + i * NB2, sub + (i - 1) * N, sup + (i - 1) * N, rowid, colid, lane, w);
    else split_schur<T, N, W>(x, V + i * NB2, sup + i * N, sub + i * N, rowid, colid, lane, w);
    const unsigned warmed = warm(t + 1);
    split_gj_inverse<T, N, W>(x, rowid, colid, lane, w, gb, pb, ph);
    asm volatile("" ::"v"(warmed));
    if (tt >= 0) split_store_block<T, N, W>(x, Dg + i * NB2, rowid, colid, lane, w);
  }
  const bool below = nb - 1 - m > 0;
  if (!top && below) {
    const double r = lane < N ? sup[m * N + rowid] : 0.;
    const double csl = lane < N ? sub[m * N + lane] : 0.;
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int s = w * C + k;
      if (s < N) {
        const int c = __builtin_amdgcn_readlane(colid, s);
        const double cc = lane_value(csl, c);
        if (lane < N) sb[c * N + rowid] = r * (double)x[k] * cc;
      }
    }
  }
  __syncthreads();
  // The bottom chain is done. Its waves end here, and the barriers of the middle inverse wait only on the
  // waves still running (s_barrier counts the surviving waves of a workgroup).
  if (!top) return;
  if (m == 0) split_load_block<T, N, W>(x, V + m * NB2, rowid, colid, lane, w);
  else split_schur<T, N, W>(x, V + m * NB2, sub + (m - 1) * N, sup + (m - 1) * N, rowid, colid, lane, w);
  if (below) {
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int s = w * C + k;
      if (s < N) {
        const int c = __builtin_amdgcn_readlane(colid, s);
        if (lane < N) x[k] = (T)((double)x[k] - sb[c * N + rowid]);
      }
    }
  }
  split_gj_inverse<T, N, W>(x, rowid, colid, lane, w, gb, pb, ph);
  split_store_block<T, N, W>(x, Dg + m * NB2, rowid, colid, lane, w);
}

}  // namespace

template <typename T, int N>
void launch_twist_factor(int nsys, const int *nblocks, const double *values, const long *voff, T *dinv,
                         const long *doff, bool split, hipStream_t st) {
  constexpr int W = twist_waves<N>();
  if (W == 1 || !split)
    hipLaunchKernelGGL((btd_twist_factor_kernel<T, N>), dim3(nsys), dim3(128), 0, st, nblocks, values, voff, dinv,
                       doff);
  else
    hipLaunchKernelGGL((btd_twist_factor_split_kernel<T, N, (W > 1 ? W : 2)>), dim3(nsys), dim3(128 * W), 0, st,
                       nblocks, values, voff, dinv, doff);
}

}  // namespace gfa

// Explicit instantiations of the launcher (and so of its kernel) for block size N, FP64 and FP32 inverses.
#define GF_TWIST(N)                                                                                        \
  template void gfa::launch_twist_factor<double, N>(int, const int *, const double *, const long *, double *, \
                                                    const long *, bool, hipStream_t);                      \
  template void gfa::launch_twist_factor<float, N>(int, const int *, const double *, const long *, float *,   \
                                                   const long *, bool, hipStream_t);
