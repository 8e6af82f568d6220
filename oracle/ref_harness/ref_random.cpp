// Test infrastructure (reference harness only): link-time replacement for the
// reference's source/math/random.cpp. The reference seeds a thread-local
// std::mt19937 from std::random_device (random.cpp:12-19), which makes every
// render irreproducible; the harness swaps in a seeded stream so that the
// compiled reference can be driven path-by-path against the oracle.
//
// Two sources of uniforms, chosen per thread by the harness:
//   * a SplitMix64 stream (oracle/seqrng.h), re-keyed per (seed, pixel, sample);
//   * an explicit queue of values, for golden vectors of single functions
//     (BSDF / light sampling with given u).
#include "math/random.h"
#include "math/math.h"

#include "ref_random.h"

#include <cstdio>
#include <cstdlib>

extern "C" {
#include "../seqrng.h"
}

namespace refharness {

thread_local RngState g_rng;

void rng_set_stream(uint64_t key) {
    g_rng.mode  = RngState::STREAM;
    g_rng.state = key;
    g_rng.draws = 0;
}

void rng_set_queue(const float* values, std::size_t n) {
    g_rng.mode = RngState::QUEUE;
    g_rng.queue.assign(values, values + n);
    g_rng.qpos  = 0;
    g_rng.draws = 0;
}

} // namespace refharness

namespace cadise {

real Random::nextReal() {
    using refharness::g_rng;
    g_rng.draws += 1;
    if (g_rng.mode == refharness::RngState::QUEUE) {
        if (g_rng.qpos >= g_rng.queue.size()) {
            std::fprintf(stderr, "ref harness: random queue exhausted\n");
            std::abort();
        }
        return g_rng.queue[g_rng.qpos++];
    }
    return seqrng_next(&g_rng.state);
}

// Same index formula as the reference (random.cpp:21-26): the algorithm, not
// the generator, is what the replacement must keep.
std::size_t Random::nextIndex(const std::size_t minIndex, const std::size_t maxIndex) {
    const std::size_t interval    = maxIndex - minIndex;
    const std::size_t sampleIndex = static_cast<std::size_t>(minIndex + nextReal() * interval);

    return math::clamp(sampleIndex, minIndex, maxIndex - 1);
}

} // namespace cadise
