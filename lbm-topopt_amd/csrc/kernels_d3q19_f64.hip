// Sweeps for D3Q19: incompressible BGK / MRT (forward + adjoint), standard MRT / BGK (forward); storage f64.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD3Q19
#define PREC PrecF64
#define SUFFIX d3q19_f64
#define FWD_OPS X(C_BGK) X(C_MRT) X(C_MRT_STD) X(C_BGK_STD)
#define ADJ_OPS X(C_BGK) X(C_MRT)
#include "lbmx_dispatch.inc"
