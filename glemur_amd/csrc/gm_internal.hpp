// Internal declarations shared by the host-side translation units.
#pragma once

#include "../../include/glemur.h"

#include <string>

namespace gm {

// Thread-local error text; every C entry point returns a code and stores
// the message here.
int fail(int code, const std::string &msg);
void clear_error();

inline bool starts(const char *s, const char *w)
{
    size_t n = 0;
    while (w[n]) ++n;
    for (size_t i = 0; i < n; ++i)
        if (s[i] != w[i]) return false;
    return true;
}

inline long long npad(const gm_params &p)
{
    return (long long)(p.nx + 2) * (p.ny + 2) * (p.nz + 2);
}
inline long long ninner(const gm_params &p)
{
    return (long long)p.nx * p.ny * p.nz;
}

// splitmix64 (Steele, Lea, Flood 2014): the seeded generator of the braid
// workloads; doubles take the top 53 bits.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

}  // namespace gm
