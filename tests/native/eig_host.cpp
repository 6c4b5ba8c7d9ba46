// CPU build of the device eigenvalue routine (host lane policy), for unit tests only.
#include <vector>
#include "../../spitfire_amd/csrc/gf_eig.h"

// kexp: null, or n balancing exponents carried in and written back
extern "C" double eig_max_real_host_k(const double *colmajor, int n, int *failed, float *kexp) {
  const int ld = n | 1;
  std::vector<double> a((size_t)n * ld + 1);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[(size_t)i * ld + j] = colmajor[(size_t)j * n + i];
  gfa::HostLanes par;
  return gfa::eig_max_real(par, a.data(), n, ld, failed, kexp);
}

extern "C" double eig_max_real_host(const double *colmajor, int n, int *failed) {
  return eig_max_real_host_k(colmajor, n, failed, nullptr);
}
