// Batched sweeps (every instance of a batch in one launch) for D2Q9 incompressible BGK / MRT / cascaded;
// shifted f16 storage, FP32 arithmetic.
#include "lbmx_kernels.hpp"
#include "lbmx_launch.h"
#define LAT LatD2Q9
#define PREC PrecF16
#define SUFFIX d2q9_f16
#define BATCH_OPS X(C_BGK) X(C_MRT) X(C_CASC)
#include "lbmx_batch_dispatch.inc"
