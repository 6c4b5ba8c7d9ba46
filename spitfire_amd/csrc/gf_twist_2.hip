// Block sizes 41..51 of the twisted block factorisation kernels.
#include "gf_twist_impl.h"

GF_TWIST(41) GF_TWIST(42) GF_TWIST(43) GF_TWIST(44) GF_TWIST(45) GF_TWIST(46) GF_TWIST(47) GF_TWIST(48)
GF_TWIST(49) GF_TWIST(50) GF_TWIST(51)
