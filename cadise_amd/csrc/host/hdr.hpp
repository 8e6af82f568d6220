// Radiance RGBE (.hdr) images: the environment light's radiance file.
#pragma once

#include <string>
#include <vector>

namespace mpt {

// Reads a Radiance .hdr file (flat or new-style run-length scanlines) into
// width x height x 3 floats, scanlines top to bottom as stored. A pixel
// (R, G, B, E) decodes to (R, G, B) * 2^(E - 136), 0 for E == 0 (Ward's
// shared-exponent format). Returns false with a message on malformed input.
bool read_hdr(const std::string& path, std::vector<float>* rgb, int* width, int* height, std::string* err);

}  // namespace mpt
