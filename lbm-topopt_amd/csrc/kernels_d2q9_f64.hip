// Sweeps for D2Q9: incompressible BGK / MRT / cascaded (forward + adjoint), standard BGK (forward); storage f64.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD2Q9
#define PREC PrecF64
#define SUFFIX d2q9_f64
#define FWD_OPS X(C_BGK) X(C_MRT) X(C_CASC) X(C_BGK_STD)
#define ADJ_OPS X(C_BGK) X(C_MRT) X(C_CASC)
#include "lbmx_dispatch.inc"
