// pk_unsplit_fast.hip -- the unsplit kernels of pk_unsplit.hip built with FMA
// contraction and fast (reciprocal-based) division and square root
// (-ffp-contract=fast -ffast-math), in namespace pk_fast.  Selected at run
// time by pk_config.fast_math, as the sweep kernels are.
#define PK_KNS pk_fast
#include "pk_unsplit.hip"
