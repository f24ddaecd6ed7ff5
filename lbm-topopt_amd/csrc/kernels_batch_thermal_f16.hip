// Batched sweeps (every instance of a batch in one launch) for the thermal D2Q9 + D2Q5 lattice, incompressible BGK / MRT flow part;
// shifted f16 storage, FP32 arithmetic.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD2Q9D2Q5
#define PREC PrecF16
#define SUFFIX thermal_f16
#define BATCH_OPS X(C_BGK) X(C_MRT)
#include "lbmx_batch_dispatch.inc"
