// splitmix64 streams and the engine's SSSP edge-weight functions, shared by
// host code (g++) and device code (hipcc).
#ifndef SG_RNG_H
#define SG_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define SG_HD __host__ __device__ __forceinline__
#else
#define SG_HD inline
#endif

// splitmix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014).
SG_HD uint64_t sg_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Counter-based stream: value k of stream (seed, id) is splitmix64 of a
// per-stream start plus k golden-ratio steps.
struct SgRng {
    uint64_t s;
    SG_HD SgRng(uint64_t seed, uint64_t id) : s(sg_splitmix64(seed ^ sg_splitmix64(id + 0x632BE59BD9B4E019ull))) {}
    SG_HD uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t x = s;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    }
};

// One R-MAT draw of `scale` levels; 32 random bits per level compared with
// the thresholds a, a + b, a + b + c scaled to 2^32 (rounded on the host).
struct SgRmatThresh {
    uint64_t ab, a, abc;
};

SG_HD void sg_rmat_draw(SgRng &rng, uint32_t scale, const SgRmatThresh &t, uint32_t &src, uint32_t &dst) {
    uint32_t s = 0, d = 0;
    uint64_t bits = 0;
    for (uint32_t lvl = 0; lvl < scale; ++lvl) {
        if ((lvl & 1) == 0) bits = rng.next();
        const uint64_t r = (lvl & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFull);
        s <<= 1;
        d <<= 1;
        if (r < t.ab) {
            if (r >= t.a) d |= 1;
        } else {
            s |= 1;
            if (r >= t.abc) d |= 1;
        }
    }
    src = s;
    dst = d;
}

// Weight modes (sgraph.h sg_weight_mode).
//   0: (u + v) % 128 + 1  — the reference's device weight, push_functor.h:306
//   1: 1 + splitmix64(seed ^ (u << 32 | v)) % 63  — seeded, in [1, 63]
//   2: 1
SG_HD uint32_t sg_weight(uint32_t mode, uint64_t seed, uint32_t u, uint32_t v) {
    if (mode == 0) return ((u + v) & 127u) + 1u;
    if (mode == 1) return 1u + (uint32_t)(sg_splitmix64(seed ^ (((uint64_t)u << 32) | v)) % 63ull);
    return 1u;
}

#endif
