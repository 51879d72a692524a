// pk_datafields.hip -- output fields of the data files computed where the
// state lives: one lane per interior cell, the cell's planes read once
// (coalesced along x), the field written in the reference's (nx, ny, nz)
// order.  The arithmetic is datafields_hdf5's (data_hdf5.F90:496-580) with
// ekin, emag and sq_sum3 of func.F90:55-88, evaluated left to right without
// contraction.
#include <hip/hip_runtime.h>
#include "pk_datafields.hpp"

namespace pk {

namespace {

__global__ __launch_bounds__(256) void k_datafield(const DatafieldArgs a)
{
    const long long nx = a.n[0], nxy = nx * a.n[1];
    const long long total = nxy * a.n[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int i = (int)(e % nx), j = (int)((e / nx) % a.n[1]), k = (int)(e / nxy);
    const long long c = (long long)(i + a.g[0]) + (long long)a.sy * (j + a.g[1]) + (long long)a.sz * (k + a.g[2]);
    const double *s = a.state;
    auto at = [&](int p) { return s[(long long)p * a.sv + c]; };
    double r;
    if (a.kind == DF_COPY) {
        r = at(a.plane);
    } else if (a.kind == DF_VEL) {
        r = at(a.plane) / at(a.fluid);
    } else {
        // emag_c = half*sq_sum3(b) with cell-centred B (cc_mag), 0 without B
        double em = 0.0;
        if (a.mag >= 0) {
            const double bx = at(a.mag), by = at(a.mag + 1), bz = at(a.mag + 2);
            em = 0.5 * ((bx * bx + by * by) + bz * bz);
        }
        if (a.kind == DF_MAGB) {
            r = sqrt(2.0 * em);
        } else {
            const double dn = at(a.fluid);
            const double mx = at(a.fluid + 1), my = at(a.fluid + 2), mz = at(a.fluid + 3);
            const double m2 = (mx * mx + my * my) + mz * mz;
            if (a.kind == DF_VTOT) {
                r = sqrt(m2) / dn;
            } else {
                const double ek = (0.5 * m2) / dn;
                const double eth = (at(a.fluid + 4) - ek) - em;
                r = (a.kind == DF_PRE) ? a.gam_1 * eth : eth / dn;
            }
        }
    }
    if (a.out32)
        a.out32[e] = (float)r;  // round to nearest even, as HDF5's FP64 -> FP32 conversion
    else
        a.out[e] = r;
}

}  // namespace

hipError_t launch_datafield(const DatafieldArgs &a, hipStream_t st)
{
    const long long total = (long long)a.n[0] * a.n[1] * a.n[2];
    const long long blocks = (total + 255) / 256;
    if (blocks < 1 || blocks > 0x7fffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_datafield, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace pk
