/*
 * Test infrastructure: the seeded sequential random stream shared by the
 * oracle restatement (the C files in oracle/) and the reference harness
 * (oracle/ref_harness). Both draw their uniforms from this stream so that a
 * path traced by the oracle and by the compiled reference consume identical
 * numbers in identical order.
 *
 * The reference draws every uniform from a thread-local, non-seedable
 * std::mt19937 (source/math/random.cpp:12-19), so its renders are never
 * reproducible. The harness replaces that file at link time with a stream of
 * this generator; the oracle in "sequential" mode uses the same one.
 *
 * Generator: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014). A uniform float in [0,1) takes the top
 * 24 bits of one output.
 */
#ifndef MPT_ORACLE_SEQRNG_H
#define MPT_ORACLE_SEQRNG_H

#include <stdint.h>

static inline uint64_t seqrng_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* State of one stream; the stream for (seed, a, b) starts at seqrng_key. */
static inline uint64_t seqrng_key(uint64_t seed, uint64_t a, uint64_t b) {
    return seqrng_mix64(seed ^ seqrng_mix64(a * 0x9E3779B97F4A7C15ull + seqrng_mix64(b + 0x632BE59BD9B4E019ull)));
}

static inline float seqrng_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = seqrng_mix64(z);
    return (float)(z >> 40) * (1.0f / 16777216.0f);
}

#endif
