// scene_compile.h — host-side scene compiler: IR -> validated scene -> device table blob.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mm3e_scene_ir.h"
#include "dev_scene.h"

namespace mm3e {

struct HObj {
  mm3e_ir_object ir;
};

struct HVol {
  mm3e_ir_volume ir;
};

struct HostScene {
  mm3e_ir_header hdr;
  std::vector<mm3e_ir_object> objects;
  std::vector<mm3e_ir_material> materials;
  std::vector<mm3e_ir_light> lights;
  std::vector<mm3e_ir_volume> volumes;
  std::vector<float> volume_data;
  std::vector<float> gi_cubes;  // [probe][6][3] when hdr.gi_mode == MM3E_GI_BAKED
  std::vector<mm3e_ir_csg_node> csg_nodes;
  std::vector<uint32_t> csg_roots;   // expression id -> root node
};

// Parse + validate the IR (the GPU backend's checks, mm3e-gpu/src/lib.rs:138-142 and
// wgsl.rs:41-47, plus the limits of this implementation).  Returns false with a message.
bool parse_scene_ir(const void* ir, size_t len, HostScene& out, std::string& err);

struct Sphere {
  bool some;
  float c[3];
  float r;
};
// csg.rs:120-150 Expr::bound of the expression rooted at `node`.
Sphere csg_bound(const HostScene& sc, uint32_t node);
// Compile expression `id` to the postfix program evaluated by csg_distance on the device.
void csg_compile(const HostScene& sc, uint32_t id, std::vector<DCsgOp>& out);
// Object::world_bound (lib.rs:470-496).
Sphere world_bound(const HostScene& sc, const mm3e_ir_object& o);
// gi::bounds_of over all bounded objects with pad 1.5 (lib.rs:668-671, gi.rs:192-203).
bool gi_bounds(const HostScene& sc, float lo[3], float hi[3]);
bool scene_dual_safe(const HostScene& sc);

// Lay out the device blob for `sc` (+ optional extra objects/materials appended at the end,
// used for dynamic spheres).  `dev_base` is the device address the blob will be copied to;
// table pointers in the header are filled with addresses inside it.  `gi_dev` (may be 0)
// points at an already-resident [probe][6][4] GI array that the header should reference
// instead of the one carried in the blob.
struct Blob {
  std::vector<uint8_t> bytes;
  DevScene header;
  uint64_t gi_offset = 0;   // offset of the GI array in `bytes` (0 if none in blob)
  uint64_t gi_bytes = 0;
  // host copies of the world-plan tables (what the code generator folds into constants)
  std::vector<DObj> objs;
  std::vector<DSeg> segs;
  std::vector<uint32_t> unb;
};
void build_blob(const HostScene& sc, const std::vector<mm3e_ir_object>& extra_objects,
                const std::vector<mm3e_ir_material>& extra_materials, uint32_t aa, uint32_t bounces,
                uint64_t dev_base, const float* gi_dev, bool with_gi_flag, Blob& out);
// Size the blob would have (for allocation before the base address is known).
uint64_t blob_size(const HostScene& sc, const std::vector<mm3e_ir_object>& extra_objects,
                   const std::vector<mm3e_ir_material>& extra_materials);

// Segment candidate grid as laid out in the blob (DevScene::grid*).
struct SegGrid {
  float lo[3] = {0, 0, 0}, inv = 0;
  uint32_t dims[3] = {0, 0, 0}, n_words = 0;
  float m[kGridLevels] = {}, kmax = 0, eps = 0;
  float in_lo[3] = {0, 0, 0}, in_hi[3] = {0, 0, 0};  // box of the bounded segments' spheres
  std::vector<uint32_t> words;  // [cell (+1 outside)][level][word]
};
void build_seg_grid(const std::vector<DSeg>& segs, const std::vector<std::vector<uint32_t>>& members,
                    const std::vector<Sphere>& bounds, float slack, SegGrid& g);

// gi.rs:38-50 axis_dirs for the six cube faces: xyz + weight per ray.
void gi_axis_dirs(uint32_t samples, float dirs[6][5][4], uint32_t* rays_per_face);

// post.rs:78-92 gaussian_kernel.
std::vector<float> gaussian_kernel(uint32_t radius);

}  // namespace mm3e
