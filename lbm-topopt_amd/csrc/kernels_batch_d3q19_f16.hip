// Batched sweeps (every instance of a batch in one launch) for D3Q19 incompressible BGK / MRT;
// shifted f16 storage, FP32 arithmetic.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD3Q19
#define PREC PrecF16
#define SUFFIX d3q19_f16
#define BATCH_OPS X(C_BGK) X(C_MRT)
#include "lbmx_batch_dispatch.inc"
