// pk_mass.hpp -- the mass the density floor adds (local_magic_mass,
// src/fluids/sources.F90:356-358), accumulated exactly and in any order.
//
// Every floor event adds dm = max(rho, smalld) - rho >= 0 to its cell.  The
// device sums these values as integers in units of 2^-160: a value's 53-bit
// mantissa, shifted to its place, is cut into 32-bit digits, and digit j of
// every contribution is added to the 64-bit counter acc[j], j = 0..7 (place
// values 2^(32j-160)).  Integer addition is associative, so the total does
// not depend on the order in which lanes, waves, workgroups or launches add
// their parts: every run, tile shape and kernel gives the same counters.
// Limits: bits of a value below 2^-160 are dropped, a single value must stay
// below 2^85, and a counter wraps after about 2^32 contributions of its
// largest digit (about 2^26 full waves); the counters are read and reset by
// the host between steps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pk {

constexpr int kMassDigits = 8;
constexpr int kMassLsb = -160;

// Place of x >= 0: x = (d0 + d1*2^32 + d2*2^64) * 2^(32*i - 160), digits
// below 2^32 each, truncated below 2^-160.  Returns i (-1 for x == 0).
__host__ __device__ inline int mass_place(double x, uint32_t &d0, uint32_t &d1, uint32_t &d2)
{
    d0 = d1 = d2 = 0u;
    const uint64_t bits = __builtin_bit_cast(uint64_t, x) & 0x7fffffffffffffffull;
    if (bits == 0ull) return -1;
    const int e = (int)(bits >> 52);
    uint64_t m = bits & 0xfffffffffffffull;
    int p;  // exponent of m's lowest bit relative to 2^-160
    if (e == 0) {
        p = -1074 - kMassLsb;
    } else {
        m |= 1ull << 52;
        p = e - 1075 - kMassLsb;
    }
    if (p < 0) {
        if (-p >= 53) return -1;
        m >>= -p;
        p = 0;
        if (m == 0ull) return -1;
    }
    const int i = p >> 5, o = p & 31;
    const uint64_t lo = m << o;
    const uint64_t hi = o ? (m >> (64 - o)) : 0ull;
    d0 = (uint32_t)lo;
    d1 = (uint32_t)(lo >> 32);
    d2 = (uint32_t)hi;
    return i;
}

// Sum of v over the lanes of this wave that exist (the last wave of a
// workgroup may be partial): xor butterfly, partners beyond it skipped.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
    const int lane = threadIdx.x & 63;
    const int nact = min(64, (int)blockDim.x - (int)(threadIdx.x & ~63u));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t t = (uint64_t)__shfl_xor((unsigned long long)v, o, 64);
        if ((lane ^ o) < nact) v += t;
    }
    return v;
}

// Add each lane's dm (>= 0; 0 for no contribution) to the counters.  All
// lanes of the wave must call it together; only waves where some lane has a
// contribution do any work.
__device__ __forceinline__ void mass_add(unsigned long long *acc, double dm)
{
    if (!__any(dm != 0.0)) return;
    uint32_t d0, d1, d2;
    const int i = mass_place(dm, d0, d1, d2);
#pragma unroll
    for (int j = 0; j < kMassDigits; j++) {
        uint64_t c = 0ull;
        if (j == i) c = d0;
        if (j == i + 1) c = d1;
        if (j == i + 2) c = d2;
        const uint64_t s = wave_sum_u64(c);
        if ((threadIdx.x & 63) == 0 && s != 0ull) atomicAdd(acc + j, (unsigned long long)s);
    }
}

// The counters as one FP64 value, correctly rounded: the digits are
// carried into 32-bit limbs, the 64 bits from the highest nonzero limb down
// are converted to FP64 with a sticky bit for everything below them (so the
// one conversion rounds to nearest as the exact sum would), then scaled.
// Host and device compute the same bits.
__host__ __device__ inline double mass_value(const unsigned long long *acc)
{
    uint32_t limb[kMassDigits + 2];
    uint64_t carry = 0ull;
    for (int j = 0; j < kMassDigits; j++) {
        const uint64_t lo = (acc[j] & 0xffffffffull) + (carry & 0xffffffffull);
        limb[j] = (uint32_t)lo;
        carry = (acc[j] >> 32) + (carry >> 32) + (lo >> 32);
    }
    limb[kMassDigits] = (uint32_t)carry;
    limb[kMassDigits + 1] = (uint32_t)(carry >> 32);
    int top = kMassDigits + 1;
    while (top > 0 && limb[top] == 0u) top--;
    if (top == 0) return (double)limb[0] * __builtin_ldexp(1.0, kMassLsb);
    uint64_t m = ((uint64_t)limb[top] << 32) | limb[top - 1];
    bool sticky = false;
    for (int j = top - 2; j >= 0; j--) sticky = sticky || limb[j] != 0u;
    // m has its top limb nonzero, so at least 33 significant bits; a
    // leading-zero count shifts the top bit to bit 63 with a sticky bit below
    const int lz = __builtin_clzll(m);
    int shift = 32 * (top - 1) + kMassLsb;
    if (lz > 0) {
        // bring in lz more bits from the next limb down
        const uint64_t next = top >= 2 ? limb[top - 2] : 0u;
        m = (m << lz) | (next >> (32 - lz));
        shift -= lz;
        sticky = false;
        if (top >= 2 && (next & ((1ull << (32 - lz)) - 1ull)) != 0ull) sticky = true;
        for (int j = top - 3; j >= 0; j--) sticky = sticky || limb[j] != 0u;
    }
    if (sticky) m |= 1ull;
    return __builtin_ldexp((double)m, shift);
}
}  // namespace pk
