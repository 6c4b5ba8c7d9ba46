// Sanitizer driver for the CPU build of the device eigenvalue routine (test infrastructure): seeded random,
// badly scaled, upper-triangular, defective (Jordan) and companion matrices of orders 1..64, each checked for a
// finite result and no reported failure. Built with -fsanitize=address,undefined by tests/test_sanitizers.py.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>
#include "../../spitfire_amd/csrc/gf_eig.h"

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double uniform() {  // xorshift64*
  g_state ^= g_state >> 12;
  g_state ^= g_state << 25;
  g_state ^= g_state >> 27;
  return (double)((g_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

static double run(const std::vector<double> &colmajor, int n, int *failed, float *kexp = nullptr) {
  const int ld = n | 1;
  std::vector<double> a((size_t)n * ld + 1);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[(size_t)i * ld + j] = colmajor[(size_t)j * n + i];
  gfa::HostLanes par;
  return gfa::eig_max_real(par, a.data(), n, ld, failed, kexp);
}

int main() {
  int bad = 0, cases = 0;
  for (int n = 1; n <= 64; n += (n < 8 ? 1 : 7)) {
    for (int kind = 0; kind < 5; ++kind) {
      std::vector<double> A((size_t)n * n, 0.);
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double x = 2. * uniform() - 1.;
          if (kind == 1) x *= std::pow(10., 16. * uniform() - 8.);  // entries over 16 decades
          if (kind == 2 && i > j) x = 0.;                           // upper triangular
          if (kind == 3) x = (i == j) ? 2. : (i + 1 == j ? 1. : 0.);  // one Jordan block
          if (kind == 4) x = (i == j + 1) ? 1. : (j == n - 1 ? -(double)(i + 1) : 0.);  // companion
          A[(size_t)j * n + i] = x;
        }
      // plain, then with carried balancing exponents: zero, the ones written back, then arbitrary ones
      std::vector<float> k((size_t)n, 0.f);
      for (int pass = 0; pass < 4; ++pass) {
        if (pass == 3)
          for (int i = 0; i < n; ++i) k[(size_t)i] = (float)(int)(80. * uniform() - 40.);
        int failed = 0;
        const double r = run(A, n, &failed, pass ? k.data() : nullptr);
        ++cases;
        if (!std::isfinite(r) || failed) {
          ++bad;
          std::printf("n=%d kind=%d pass=%d: result %g, failed %d\n", n, kind, pass, r, failed);
        }
      }
    }
  }
  std::printf("%d cases, %d bad\n", cases, bad);
  return bad ? 1 : 0;
}
