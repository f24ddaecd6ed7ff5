// Batched sweeps (every instance of a batch in one launch) for the thermal D2Q9 + D2Q5 lattice, incompressible BGK / MRT flow part;
// shifted bf16 storage, FP32 arithmetic.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD2Q9D2Q5
#define PREC PrecBF16
#define SUFFIX thermal_bf16
#define BATCH_OPS X(C_BGK) X(C_MRT)
#include "lbmx_batch_dispatch.inc"
