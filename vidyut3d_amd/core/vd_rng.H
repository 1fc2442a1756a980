// Counter-based random numbers for seeded initial conditions.
//
// Each value is a pure function of (seed, stream, index): the same cell gets
// the same number on the host and on the device, in any launch order.  The
// mixer is SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014); normals use the Box-Muller transform.
#ifndef VD_RNG_H_
#define VD_RNG_H_
#include <AMReX_compat.H>
#include <cstdint>

namespace vd_rng {

VD_HD AMREX_FORCE_INLINE uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// uniform in (0, 1]
VD_HD AMREX_FORCE_INLINE double uniform(uint64_t seed, uint64_t stream, uint64_t idx, uint64_t draw)
{
    uint64_t h = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ULL + splitmix64(idx * 4ULL + draw)));
    return (double((h >> 11) + 1ULL)) * (1.0 / 9007199254740992.0);
}

// standard normal
VD_HD AMREX_FORCE_INLINE double normal(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u1 = uniform(seed, stream, idx, 0);
    double u2 = uniform(seed, stream, idx, 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

} // namespace vd_rng
#endif
