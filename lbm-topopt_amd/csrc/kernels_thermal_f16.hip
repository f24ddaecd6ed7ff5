// Sweeps for D2Q9 + D2Q5 double distribution (flow BGK / MRT with thermal MRT), forward + adjoint; storage f16.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD2Q9D2Q5
#define PREC PrecF16
#define SUFFIX thermal_f16
#define FWD_OPS X(C_BGK) X(C_MRT)
#define ADJ_OPS X(C_BGK) X(C_MRT)
#include "lbmx_dispatch.inc"
