// Primitives shared by the host library and the HIP kernels: the RNG stack
// the reference draws from and the neighbourhood orders its results depend
// on.  Everything here is `CPM_HD` so host initialisation and device
// exact-replay consume the same stream with the same arithmetic.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CPM_HD __host__ __device__ __forceinline__
#else
#define CPM_HD inline
#endif

namespace cpm {

// --- Xoshiro256++ (rand_xoshiro 0.6.0; crates/cpm-core/src/rng.rs:7-14) ---
struct Xoshiro {
    uint64_t s[4];
};

CPM_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

CPM_HD uint64_t xo_next(Xoshiro& r) {
    const uint64_t out = rotl64(r.s[0] + r.s[3], 23) + r.s[0];
    const uint64_t t = r.s[1] << 17;
    r.s[2] ^= r.s[0];
    r.s[3] ^= r.s[1];
    r.s[1] ^= r.s[2];
    r.s[0] ^= r.s[3];
    r.s[2] ^= t;
    r.s[3] = rotl64(r.s[3], 45);
    return out;
}

// SeedableRng::seed_from_u64 — SplitMix64 fills the 32-byte seed.
inline Xoshiro xo_seed(uint64_t seed) {
    Xoshiro r;
    uint64_t st = seed;
    for (int i = 0; i < 4; ++i) {
        st += 0x9e3779b97f4a7c15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        r.s[i] = z ^ (z >> 31);
    }
    if ((r.s[0] | r.s[1] | r.s[2] | r.s[3]) == 0) return xo_seed(0);
    return r;
}

// Rejection zone of rand 0.8's UniformInt<u64>::sample_single_inclusive for
// a range of `range` values: (range << lz(range)) - 1.
CPM_HD uint64_t range_zone(uint64_t range) {
    return (range << __builtin_clzll(range)) - 1;
}

// One widening-multiply trial: returns true and the sample in *out when
// the draw v falls inside the zone.
CPM_HD bool range_trial(uint64_t v, uint64_t range, uint64_t zone, uint64_t* out) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t hi = __umul64hi(v, range);
#else
    const uint64_t hi = (uint64_t)(((unsigned __int128)v * range) >> 64);
#endif
    const uint64_t lo = v * range;
    *out = hi;
    return lo <= zone;
}

// rng.gen_range(0..n) for n >= 1 (rand 0.8: sample_single(0, n) ==
// sample_single_inclusive(0, n - 1)).
CPM_HD uint64_t gen_range(Xoshiro& r, uint64_t n) {
    const uint64_t zone = range_zone(n);
    uint64_t out;
    while (!range_trial(xo_next(r), n, zone, &out)) {
    }
    return out;
}

// rng.gen_range(0..=hi)
CPM_HD uint64_t gen_range_incl(Xoshiro& r, uint64_t hi) {
    const uint64_t range = hi + 1;
    if (range == 0) return xo_next(r);
    return gen_range(r, range);
}

// rng.gen::<f64>(): the 53 high bits scaled by 2^-53.
CPM_HD double gen_f64(Xoshiro& r) {
    return (double)(xo_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

// --- stencil orders (crates/cpm-core/src/neighborhood.rs:45-115) ---
enum Neighborhood : int32_t {
    NB_VON_NEUMANN = 0,
    NB_MOORE = 1,
    NB_EIGHTEEN = 2,
    NB_TWENTY_SIX = 3,
    NB_TWENTY_SIX_WEIGHTED = 4
};

// Fills `offs` (n x dim, component d = axis d) and returns n.  Full-shell
// order: code = 0..3^dim, component 0 varies fastest, origin skipped.
inline int stencil(int dim, int kind, int32_t* offs) {
    int n = 0;
    if (kind == NB_VON_NEUMANN) {
        for (int axis = 0; axis < dim; ++axis)
            for (int sgn = 1; sgn >= -1; sgn -= 2) {
                for (int d = 0; d < dim; ++d) offs[n * dim + d] = d == axis ? sgn : 0;
                ++n;
            }
        return n;
    }
    const int total = dim == 2 ? 9 : 27;
    for (int code = 0; code < total; ++code) {
        int c = code, nz = 0, o[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            o[d] = c % 3 - 1;
            c /= 3;
            nz += o[d] != 0;
        }
        if (!nz) continue;
        if (kind == NB_EIGHTEEN && nz > 2) continue;
        for (int d = 0; d < dim; ++d) offs[n * dim + d] = o[d];
        ++n;
    }
    return n;
}

}  // namespace cpm
