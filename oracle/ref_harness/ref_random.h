// Test infrastructure: per-thread control of the harness's Random replacement.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace refharness {

struct RngState {
    enum Mode { STREAM, QUEUE };
    Mode               mode  = STREAM;
    uint64_t           state = 0x853c49e6748fea9bull;
    std::vector<float> queue;
    std::size_t        qpos  = 0;
    uint64_t           draws = 0;
};

extern thread_local RngState g_rng;

void rng_set_stream(uint64_t key);
void rng_set_queue(const float* values, std::size_t n);

} // namespace refharness
