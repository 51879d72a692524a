// pk_unsplit_fast.hip -- the unsplit kernels of pk_unsplit.hip built with FMA
// contraction and fast (reciprocal-based) division and square root
// (-ffp-contract=fast -ffast-math), in namespace pk_fast; PK_RELAXED selects
// the refined hardware reciprocal and reciprocal square root of qdiv/qsqrt
// (pk_math.hpp).  Selected at run time by pk_config.fast_math, as the sweep
// kernels are.
#define PK_KNS pk_fast
#define PK_RELAXED 1
#include "pk_unsplit.hip"
