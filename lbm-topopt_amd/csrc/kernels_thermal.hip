// D2Q9 + D2Q5 double-distribution sweeps (flow BGK / MRT with the thermal MRT), forward + adjoint.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD2Q9D2Q5
#define SUFFIX thermal
#define FWD_OPS X(C_BGK) X(C_MRT)
#define ADJ_OPS X(C_BGK) X(C_MRT)
#include "lbmx_dispatch.inc"
