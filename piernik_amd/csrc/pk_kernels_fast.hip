// pk_kernels_fast.hip -- the kernels of pk_kernels.hip built with FMA
// contraction and fast (reciprocal-based) division and square root
// (-ffp-contract=fast -ffast-math), in namespace pk_fast; PK_RELAXED selects
// the refined hardware reciprocal and reciprocal square root of qdiv/qsqrt
// (pk_math.hpp).  Selected at run time by pk_config.fast_math; results agree
// with the exact build to about 1e-15 per operation instead of bit for bit.
#define PK_KNS pk_fast
#define PK_RELAXED 1
#include "pk_kernels.hip"
