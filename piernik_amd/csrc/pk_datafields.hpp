// pk_datafields.hpp -- output fields derived from the device state, as
// datafields_hdf5 computes them for the data files (data_hdf5.F90:345-691).
#pragma once
#include <hip/hip_runtime.h>

namespace pk {

enum DatafieldKind {
    DF_COPY = 0,  // one state plane: den, mom[xyz], ene, mag[xyz], psi
    DF_VEL = 1,   // momentum / density
    DF_PRE = 2,   // gam_1 * (E - ekin - emag_c)
    DF_ETH = 3,   // (E - ekin - emag_c) / density
    DF_MAGB = 4,  // sqrt(2 emag_c)
    DF_VTOT = 5,  // sqrt(sq_sum3(m)) / density
};

struct DatafieldArgs {
    const double *state;  // SoA planes of the block (the library's layout)
    long long sv;         // plane stride
    int sy, sz;
    int g[3], n[3];       // ghost layers, interior extents
    int kind;
    int plane;            // DF_COPY: the plane; DF_VEL: the momentum plane
    int fluid;            // first plane of the fluid (density)
    int mag;              // first B plane, -1 without a magnetic field for this field
    double gam_1;
    double *out;          // n[0]*n[1]*n[2] values, x fastest (the reference's (nx, ny, nz))
    float *out32;         // if set instead: the values rounded to nearest FP32
};

hipError_t launch_datafield(const DatafieldArgs &a, hipStream_t st);

}  // namespace pk
