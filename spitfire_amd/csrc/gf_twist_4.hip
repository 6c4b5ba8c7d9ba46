// Block sizes 59..64 of the twisted block factorisation kernels.
#include "gf_twist_impl.h"

GF_TWIST(59) GF_TWIST(60) GF_TWIST(61) GF_TWIST(62) GF_TWIST(63) GF_TWIST(64)
