// Largest real part of the eigenvalues of a small dense nonsymmetric matrix, for the pseudo-time limiter
// of the flamelet solvers.
//
// The reference computes every eigenvalue of each ns x ns chemistry block with LAPACK dgeev (no
// eigenvectors; blas_lapack_kernels.h:142-180, called at flamelet_kernels.cpp:1329-1341), i.e. balancing
// (dgebal), Householder reduction to upper Hessenberg form (dgehrd) and the Francis double-shift QR
// iteration (dhseqr, which for n < 75 is the small-bulge dlahqr). This file restates the same three
// published steps for one 64-lane wave per matrix:
//   - diagonal balancing by powers of two (Parlett & Reinsch, Numer. Math. 13, 1969), which is exact;
//   - Householder reduction to upper Hessenberg form;
//   - the Francis implicit double-shift QR iteration on the Hessenberg matrix (Wilkinson & Reinsch,
//     Handbook for Automatic Computation II, 1971, procedure hqr), with exceptional shifts every 10
//     iterations on one window, deflating 1x1 and 2x2 blocks and tracking only the largest real part.
// Every scalar decision is taken from values that all lanes read from the same LDS words or from
// reductions that end in a lane-uniform value, so control flow stays uniform across the wave.
//
// The algorithm is written once against a small "lanes" policy: the device policy runs a lane loop as
// one lane per index plus a barrier, the host policy as a plain loop, so the same code is unit-tested on
// the CPU.
#pragma once
#include <cmath>
#include <type_traits>

#if defined(__HIPCC__)
#define GF_HD __host__ __device__
#else
#define GF_HD
#endif

namespace gfa {

// Matrix storage: row-major a[i * ld + j], ld odd (fewer LDS bank conflicts for column walks). The stride may
// be an int or a compile-time constant (std::integral_constant), which lets the device fold row offsets into
// the LDS instructions. n is at most 64 (one lane per index); a larger n sets *failed and returns 0.
//
// kexp, when given, holds n balancing exponents (floats holding integers) carried from one call to the next
// on a similar matrix: the matrix is first scaled by them (a_tu * 2^(k_u - k_t), a similarity that is exact
// unless an entry underflows), the balancing sweeps continue from there, and the exponents of the final
// scaling are written back. Any exponents give the same eigenvalues in exact arithmetic; good ones save most
// of the sweeps.
template <class P, class LD>
GF_HD inline double eig_max_real(P &par, double *a, int n, LD ld, int *failed, float *kexp = nullptr) {
  using std::fabs;
  using std::ldexp;
  using std::sqrt;
  *failed = n > 64;
  if (n <= 0 || n > 64) return 0.;
  if (n == 1) return a[0];

  // 1) balancing: scale row i by 1/f and column i by f (f a power of 2) until row and column norms agree.
  //    The off-diagonal column and row norms of every index are summed at the start of a sweep (each index
  //    by its own lane, no reductions) and kept current as the sweep scales indices: scaling index i
  //    multiplies a_ti, a term of row t's norm, by f and a_it, a term of column t's norm, by 1/f. The norms
  //    only choose powers of 2, and every choice is an exact similarity, so neither their single precision
  //    nor the rounding of the updates can change the eigenvalues; each sweep starts again from fresh sums.
  typename P::LaneVal cn, rn, ke;
  par.each(0, n, [&](int t) { par.at(ke, t) = kexp ? kexp[t] : 0.f; });
  if (kexp)
    par.each(0, n, [&](int t) {
      const int kt = (int)par.at(ke, t);
      for (int u = 0; u < n; ++u) a[t * ld + u] = ldexp(a[t * ld + u], (int)par.get(ke, u) - kt);
    });
  for (int sweep = 0; sweep < 64; ++sweep) {
    par.each(0, n, [&](int t) {
      float c0 = 0.f, c1 = 0.f, r0 = 0.f, r1 = 0.f;
      int u = 0;
      for (; u + 1 < n; u += 2) {
        const double x0 = a[u * ld + t], x1 = a[(u + 1) * ld + t], y0 = a[t * ld + u], y1 = a[t * ld + u + 1];
        c0 += u == t ? 0.f : (float)fabs(x0);
        c1 += u + 1 == t ? 0.f : (float)fabs(x1);
        r0 += u == t ? 0.f : (float)fabs(y0);
        r1 += u + 1 == t ? 0.f : (float)fabs(y1);
      }
      if (u < n && u != t) {
        c0 += (float)fabs(a[u * ld + t]);
        r0 += (float)fabs(a[t * ld + u]);
      }
      par.at(cn, t) = c0 + c1;
      par.at(rn, t) = r0 + r1;
    });
    bool changed = false;
    for (int i = 0; i < n; ++i) {
      double c = par.get(cn, i), r = par.get(rn, i);
      if (!(c > 0.) || !(r > 0.)) continue;  // also a norm that cancellation left at or below zero
      double g = r * 0.5, f = 1.;
      const double s = c + r;
      int guard = 0, k = 0;
      while (c < g && guard++ < 2100) { f *= 2.; c *= 4.; ++k; }
      g = r * 2.;
      guard = 0;
      while (c > g && guard++ < 2100) { f *= 0.5; c *= 0.25; --k; }
      if ((c + r) / f < 0.95 * s) {
        changed = true;
        const double gi = 1. / f, fm = f - 1., gm = gi - 1.;
        par.each(0, n, [&](int t) {
          float &ct = par.at(cn, t), &rt = par.at(rn, t);
          if (t == i) {  // the diagonal entry is unchanged; the whole column grows by f, the row shrinks by it
            ct = (float)(ct * f);
            rt = (float)(rt * gi);
            par.at(ke, t) += (float)k;
          } else {
            const double ait = a[i * ld + t], ati = a[t * ld + i];
            a[i * ld + t] = ait * gi;
            a[t * ld + i] = ati * f;
            rt = (float)(rt + fabs(ati) * fm);
            ct = (float)(ct + fabs(ait) * gm);
          }
        });
      }
    }
    if (!changed) break;
  }
  // Exponents are carried only while all of them stay within +-96; otherwise the next call starts cold. Chemistry
  // blocks with entries over ~280 decades can end with exponents near 200: started from those, the next
  // refresh's slightly different (nearly reducible) block balanced to a far larger norm, and its eigenvalues lost
  // all accuracy, while no block started from exponents within +-128 did.
  if (kexp) {
    const bool keep = par.last_true(0, n, [&](int t) { return fabs(par.at(ke, t)) > 96.f; }) < 0;
    par.each(0, n, [&](int t) { kexp[t] = keep ? par.at(ke, t) : 0.f; });
  }

  // 2) Householder reduction to upper Hessenberg form: column k is reduced below its subdiagonal by
  //    H = I - beta v v^T acting on rows/columns k+1..n-1. The vector v is kept in column k itself (v0 in row
  //    k+1, the other entries are the column's own), which neither update reads or writes, and column k is set
  //    to alpha e1 after both: no separate vector buffer, so a matrix takes only its own n x ld words of LDS.
  for (int k = 0; k < n - 2; ++k) {
    const double x0 = a[(k + 1) * ld + k];
    const double tail = par.sum(k + 2, n, [&](int t) { const double x = a[t * ld + k]; return x * x; });
    if (tail == 0.) continue;
    const double nrm = sqrt(x0 * x0 + tail);
    const double alpha = x0 >= 0. ? -nrm : nrm;
    const double v0 = x0 - alpha;
    const double beta = -1. / (alpha * v0);
    par.each(k + 1, k + 2, [&](int t) { a[t * ld + k] = v0; });
    const double *vk = a + k;  // v_i = vk[i * ld] for i > k
    // left: rows k+1..n-1 of columns k+1..n-1; four rows per trip so that the loads of a trip are in flight
    // together
    par.each(k + 1, n, [&](int j) {
      double w0 = 0., w1 = 0., w2 = 0., w3 = 0.;
      int i = k + 1;
      for (; i + 3 < n; i += 4) {
        const double x0 = a[i * ld + j], x1 = a[(i + 1) * ld + j], x2 = a[(i + 2) * ld + j], x3 = a[(i + 3) * ld + j];
        w0 += vk[i * ld] * x0;
        w1 += vk[(i + 1) * ld] * x1;
        w2 += vk[(i + 2) * ld] * x2;
        w3 += vk[(i + 3) * ld] * x3;
      }
      for (; i < n; ++i) w0 += vk[i * ld] * a[i * ld + j];
      const double w = ((w0 + w1) + (w2 + w3)) * beta;
      i = k + 1;
      for (; i + 3 < n; i += 4) {
        const double x0 = a[i * ld + j], x1 = a[(i + 1) * ld + j], x2 = a[(i + 2) * ld + j], x3 = a[(i + 3) * ld + j];
        a[i * ld + j] = x0 - vk[i * ld] * w;
        a[(i + 1) * ld + j] = x1 - vk[(i + 1) * ld] * w;
        a[(i + 2) * ld + j] = x2 - vk[(i + 2) * ld] * w;
        a[(i + 3) * ld + j] = x3 - vk[(i + 3) * ld] * w;
      }
      for (; i < n; ++i) a[i * ld + j] -= vk[i * ld] * w;
    });
    // right: all rows, columns k+1..n-1 (four columns per trip)
    par.each(0, n, [&](int i) {
      double u0 = 0., u1 = 0., u2 = 0., u3 = 0.;
      double *ai = a + i * ld;
      int j = k + 1;
      for (; j + 3 < n; j += 4) {
        const double x0 = ai[j], x1 = ai[j + 1], x2 = ai[j + 2], x3 = ai[j + 3];
        u0 += x0 * vk[j * ld];
        u1 += x1 * vk[(j + 1) * ld];
        u2 += x2 * vk[(j + 2) * ld];
        u3 += x3 * vk[(j + 3) * ld];
      }
      for (; j < n; ++j) u0 += ai[j] * vk[j * ld];
      const double u = ((u0 + u1) + (u2 + u3)) * beta;
      j = k + 1;
      for (; j + 3 < n; j += 4) {
        const double x0 = ai[j], x1 = ai[j + 1], x2 = ai[j + 2], x3 = ai[j + 3];
        ai[j] = x0 - u * vk[j * ld];
        ai[j + 1] = x1 - u * vk[(j + 1) * ld];
        ai[j + 2] = x2 - u * vk[(j + 2) * ld];
        ai[j + 3] = x3 - u * vk[(j + 3) * ld];
      }
      for (; j < n; ++j) ai[j] -= u * vk[j * ld];
    });
    par.each(0, n, [&](int t) {  // column k becomes alpha e1 exactly
      if (t == k + 1) a[t * ld + k] = alpha;
      else if (t > k + 1) a[t * ld + k] = 0.;
    });
  }

  // 3) Francis double-shift QR on the Hessenberg matrix (eigenvalues only, active window [l, nn])
  const double anorm = par.sum(0, n, [&](int i) {
    double s = 0.;
    for (int j = (i > 0 ? i - 1 : 0); j < n; ++j) s += fabs(a[i * ld + j]);
    return s;
  });
  double emax = -INFINITY;
  int nn = n - 1;
  double t = 0.;  // accumulated exceptional shift
  int total = 0;
  const int total_max = 30 * (n > 10 ? n : 10);
  while (nn >= 0) {
    int its = 0, l;
    for (;;) {
      // the negligible subdiagonal element closest to the bottom of the window (0 if none): the largest l in
      // [1, nn] passing the test, evaluated for every l at once (the same l as a downward scan)
      l = par.last_true(1, nn + 1, [&](int i) {
        double s = fabs(a[(i - 1) * ld + i - 1]) + fabs(a[i * ld + i]);
        if (s == 0.) s = anorm;
        return fabs(a[i * ld + i - 1]) + s == s;
      });
      if (l < 1) l = 0;
      const double x = a[nn * ld + nn];
      if (l == nn) {  // one real root
        emax = fmax(emax, x + t);
        nn -= 1;
        break;
      }
      const double y = a[(nn - 1) * ld + nn - 1];
      const double w = a[nn * ld + nn - 1] * a[(nn - 1) * ld + nn];
      if (l == nn - 1) {  // a pair of roots from the trailing 2x2 block
        const double p = 0.5 * (y - x);
        const double q = p * p + w;
        const double z = sqrt(fabs(q));
        const double xs = x + t;
        if (q >= 0.) {
          const double zz = p + (p >= 0. ? z : -z);
          const double r1 = xs + zz;
          const double r2 = zz != 0. ? xs - w / zz : r1;
          emax = fmax(emax, fmax(r1, r2));
        } else {
          emax = fmax(emax, xs + p);
        }
        nn -= 2;
        break;
      }
      if (total >= total_max) {  // no convergence: bound the rest of the window by its diagonal
        *failed = 1;
        for (int i = l; i <= nn; ++i) emax = fmax(emax, a[i * ld + i] + t);
        nn = l - 1;
        break;
      }
      double xx = x, yy = y, ww = w;
      if (its > 0 && its % 10 == 0) {  // exceptional shift
        t += xx;
        par.each(0, nn + 1, [&](int i) { a[i * ld + i] -= xx; });
        const double s = fabs(a[nn * ld + nn - 1]) + fabs(a[(nn - 1) * ld + nn - 2]);
        xx = yy = 0.75 * s;
        ww = -0.4375 * s * s;
      }
      ++its;
      ++total;
      // look for two consecutive small subdiagonal elements: the largest m in [l+1, nn-2] passing the test
      // (evaluated for every m at once; the same m as a downward scan), else m = l; then the shift vector at m
      auto shift_vector = [&](int mm, double &pp, double &qq, double &rr3) {
        const double z = a[mm * ld + mm];
        const double rr = xx - z, ss = yy - z;
        pp = (rr * ss - ww) / a[(mm + 1) * ld + mm] + a[mm * ld + mm + 1];
        qq = a[(mm + 1) * ld + mm + 1] - z - rr - ss;
        rr3 = a[(mm + 2) * ld + mm + 1];
        const double is = par.recip(fabs(pp) + fabs(qq) + fabs(rr3));
        pp *= is;
        qq *= is;
        rr3 *= is;
      };
      int m = par.last_true(l + 1, nn - 1, [&](int i) {
        double pi, qi, ri;
        shift_vector(i, pi, qi, ri);
        const double u = fabs(a[i * ld + i - 1]) * (fabs(qi) + fabs(ri));
        const double v = fabs(pi) * (fabs(a[(i - 1) * ld + i - 1]) + fabs(a[i * ld + i]) + fabs(a[(i + 1) * ld + i + 1]));
        return u + v == v;
      });
      if (m < l + 1) m = l;
      double p = 0., q = 0., r = 0.;
      shift_vector(m, p, q, r);
      par.each(m + 2, nn + 1, [&](int i) {
        a[i * ld + i - 2] = 0.;
        if (i != m + 2) a[i * ld + i - 3] = 0.;
      });
      // the double-shift QR sweep: 3x3 reflectors chase the bulge down the window
      for (int k = m; k <= nn - 1; ++k) {
        double xk = 1.;
        if (k != m) {
          p = a[k * ld + k - 1];
          q = a[(k + 1) * ld + k - 1];
          r = k != nn - 1 ? a[(k + 2) * ld + k - 1] : 0.;
          xk = fabs(p) + fabs(q) + fabs(r);
          if (xk != 0.) {
            const double ix = par.recip(xk);
            p *= ix;
            q *= ix;
            r *= ix;
          }
        }
        const double sn = sqrt(p * p + q * q + r * r);
        const double s = p >= 0. ? sn : -sn;
        if (s == 0.) continue;
        const double is = par.recip(s), ips = par.recip(p + s);
        const double hx = (p + s) * is, hy = q * is, hz = r * is;
        const double qq = q * ips, rr = r * ips;
        const bool three = k != nn - 1;
        const double sub = (k == m) ? (l != m ? -a[k * ld + k - 1] : a[k * ld + (k > 0 ? k - 1 : 0)])
                                    : -s * xk;
        const bool write_sub = (k != m) || (l != m);
        // row transformation on rows k..k+2, columns k..nn (and the subdiagonal entry of column k-1)
        par.each(k, nn + 1, [&](int j) {
          double pp = a[k * ld + j] + qq * a[(k + 1) * ld + j];
          if (three) {
            pp += rr * a[(k + 2) * ld + j];
            a[(k + 2) * ld + j] -= pp * hz;
          }
          a[(k + 1) * ld + j] -= pp * hy;
          a[k * ld + j] -= pp * hx;
          if (j == k && write_sub && k > 0) a[k * ld + k - 1] = sub;
        });
        // column transformation on columns k..k+2, rows l..min(nn, k+3)
        const int imax = nn < k + 3 ? nn : k + 3;
        par.each(l, imax + 1, [&](int i) {
          double pp = hx * a[i * ld + k] + hy * a[i * ld + k + 1];
          if (three) {
            pp += hz * a[i * ld + k + 2];
            a[i * ld + k + 2] -= pp * rr;
          }
          a[i * ld + k + 1] -= pp * qq;
          a[i * ld + k] -= pp;
        });
      }
    }
  }
  return emax;
}

// Host policy: lane loops run sequentially (used by the CPU unit tests of this file).
struct HostLanes {
  // one value per index (the device keeps index t's in lane t's register)
  struct LaneVal {
    float v[64];
  };
  float &at(LaneVal &x, int t) { return x.v[t]; }
  float get(const LaneVal &x, int i) { return x.v[i]; }
  template <class F> void each(int lo, int hi, F f) { for (int t = lo; t < hi; ++t) f(t); }
  double recip(double x) { return 1. / x; }
  // the largest i in [lo, hi) with pred(i), or lo - 1 if none
  template <class F> int last_true(int lo, int hi, F pred) {
    for (int t = hi - 1; t >= lo; --t)
      if (pred(t)) return t;
    return lo - 1;
  }
  template <class F> double sum(int lo, int hi, F f) {
    double s = 0.;
    for (int t = lo; t < hi; ++t) s += f(t);
    return s;
  }
};

}  // namespace gfa
