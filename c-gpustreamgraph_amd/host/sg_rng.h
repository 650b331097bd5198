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
