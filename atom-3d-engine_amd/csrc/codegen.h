// codegen.h — scene -> HIP source (the twin of the reference's WGSL code generator,
// mm3e-gpu/src/wgsl.rs:34-362): the world fold of one compiled scene as straight-line device
// code with every object's parameters, transform and bounding sphere as constants, compiled at
// run time through hipRTC into the render kernel of that scene.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "scene_compile.h"

namespace mm3e {

// Source of the scene's `GenMap` (a map functor like TableMap) plus its render kernel
// `mm3e_gen_render` for reflection depth maxb; sm selects the state-machine path.  queued adds
// the stages of the queued frame (mm3e_gen_wf_first / _cont / _shade, dev_wavefront.h); film
// adds the film and G-buffer kernels (mm3e_gen_film, mm3e_gen_gbuffer).
std::string emit_scene_source(const Blob& b, uint32_t maxb, bool sm, bool queued = false, bool film = false);

// Source of a batch module: each listed scene's map as GenMap<k> and the persistent batch kernel
// `mm3e_gen_batch` (dev_batch.h's loop; a job with gen == k renders with GenMap<k>, any other
// with the table walk), for reflection depth maxb, tables staged in LDS when staged.
std::string emit_batch_source(const std::vector<const Blob*>& maps, uint32_t maxb, bool staged);

// Whether the generated kernel reads the object tables (union runs walked through their
// bounding-sphere tree still use the table walk).
bool generated_needs_tables(const Blob& b);

// Kernel family for the generated map: the state machine once the fold is long.
bool generated_prefers_sm(const Blob& b);

// Device headers the generated source includes (embedded in the library at build time).
struct EmbeddedHeader {
  const char* name;
  const char* text;
};
extern const EmbeddedHeader kEmbeddedHeaders[];
extern const int kEmbeddedHeaderCount;

}  // namespace mm3e
