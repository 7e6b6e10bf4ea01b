// FP16 instantiations of the field kernels of gx_kernels.hip, compiled as a translation unit of
// their own so the three storage types build in parallel.
#define GX_PART 2
#include "gx_kernels.hip"
