// CRSD reader: the reference's scene text format (README.md:15-18) with the
// same grammar, keys and defaults (file-io/scene-description/sdParser.cpp,
// tokenizer.cpp, sdData.cpp, core/instantiator/*.cpp), producing a HostScene
// for the path-tracing path. Blocks of other renderers, estimators,
// materials or lights report MPT_ERR_UNSUPPORTED.
//
// Extensions beside the reference's text arrays (for generated scenes of
// millions of triangles):
//   triangle-mesh  [positions-file string "p.f32"] [normals-file ...] [uvws-file ...]
//                  raw float32, 9 per triangle, the same order as the
//                  positions / normals / uvws vector3r-arrays
//   texture-*      [type string "image-raw"] [filename string "t.f32"]
//                  [width int32 W] [height int32 H] [wrap-mode string "repeat"|"clamp"]
//                  raw float32 H x W x C (C = 3 spectrum, 1 real)
#pragma once

#include <string>

#include "scene.hpp"

namespace mpt {

struct CrsdResult {
    int         code = 0;      // MPT_OK or MPT_ERR_*
    std::string message;
};

CrsdResult parse_crsd(const std::string& text, const std::string& base_dir, HostScene* out);

}  // namespace mpt
