// Batched sweeps (every instance of a batch in one launch) for D3Q19 incompressible BGK / MRT;
// storage f32.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD3Q19
#define PREC PrecF32
#define SUFFIX d3q19_f32
#define BATCH_OPS X(C_BGK) X(C_MRT)
#include "lbmx_batch_dispatch.inc"
