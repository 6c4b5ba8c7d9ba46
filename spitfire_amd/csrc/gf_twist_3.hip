// Block sizes 52..58 of the twisted block factorisation kernels.
#include "gf_twist_impl.h"

GF_TWIST(52) GF_TWIST(53) GF_TWIST(54) GF_TWIST(55) GF_TWIST(56) GF_TWIST(57) GF_TWIST(58)
