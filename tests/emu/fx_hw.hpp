// Portable stand-ins for the gfx950 instructions and binary64 estimates of the
// device limb arithmetic (cft-fp256_amd/csrc/fx_hw.hpp), so that the same fx_*.hpp
// headers compile with g++ and the test suite can check their logic against
// the oracle without a GPU.  Test infrastructure only.
#pragma once

#include <math.h>
#include <stdint.h>

#define __device__
#define __forceinline__ inline
#define __constant__
#define __global__

namespace fx {

inline uint32_t alignbit(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31u));
}

inline void mac(uint64_t &acc, uint32_t &hi, uint32_t a, uint32_t b)
{
    const uint64_t p = (uint64_t)a * b;
    const uint64_t s = acc + p;
    hi += s < p ? 1u : 0u;
    acc = s;
}

inline void mac_first(uint64_t &acc, uint32_t &hi, uint32_t a, uint32_t b)
{
    const uint64_t p = (uint64_t)a * b;
    const uint64_t s = acc + p;
    hi = s < p ? 1u : 0u;
    acc = s;
}

inline uint32_t addc(uint32_t x, uint32_t y, uint32_t cin, uint32_t *cout)
{
    const uint64_t s = (uint64_t)x + y + cin;
    *cout = (uint32_t)(s >> 32);
    return (uint32_t)s;
}

inline uint32_t subc(uint32_t x, uint32_t y, uint32_t bin, uint32_t *bout)
{
    const uint64_t t = (uint64_t)x - y - bin;
    *bout = (uint32_t)(t >> 63);
    return (uint32_t)t;
}

// the device's recip_mid and sqrt_mid are correctly rounded for their range
inline double recip_mid(double x) { return 1.0 / x; }
inline double sqrt_mid(double x) { return sqrt(x); }

} // namespace fx
