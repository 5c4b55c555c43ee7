// dev_scene.h — the compiled scene as it lives in HBM (one allocation per compiled scene).
//
// Plain-old-data shared by the host compiler (scene_compile.cpp) and the device code
// (dev_math.h / dev_render.h).  One `DevScene` header is followed by its tables; every table
// pointer in the header is a device address inside the same allocation, so a kernel needs
// exactly one pointer argument to reach the whole scene.  Scalar fields are read through
// that uniform pointer and land in SGPRs (s_load); per-object records are 16-byte aligned so
// a wave walking the object table in lock-step issues dwordx4 scalar/vector loads.
#pragma once
#include <stdint.h>

namespace mm3e {

// DObj::flags bit layout
enum : uint32_t {
  OBJ_KIND_MASK = 0xFu,      // mm3e_prim_kind
  OBJ_MIRROR_X = 1u << 4,
  OBJ_MIRROR_Y = 1u << 5,
  OBJ_MIRROR_Z = 1u << 6,
  OBJ_ELONGATE = 1u << 7,
  OBJ_REPEAT = 1u << 8,
  OBJ_TWIST = 1u << 9,
  OBJ_BEND = 1u << 10,
  OBJ_ROUND = 1u << 11,
  OBJ_ONION = 1u << 12,
  OBJ_BOUNDED = 1u << 13,
  OBJ_IDENT_ROT = 1u << 14,  // rotation is the identity: to_local skips the matrix product
  OBJ_UNIT_SCALE = 1u << 15, // scale == 1: to_local and Object::field skip the two multiplies
  OBJ_COMBINE_SHIFT = 16,    // bits 16-17: mm3e_combine of this object
  // The reference leaves the object unbounded (bend: lib.rs:478-480), but bc / bc_r hold a
  // sphere that provably lies below its field (see skip_bound in scene_compile.cpp): used only
  // to prove that folding the object changes nothing, never as its field value.
  OBJ_SKIP = 1u << 18,
};

// One placed object (mm3e-orchestrator/src/lib.rs:330-336) with its world bounding sphere
// (lib.rs:470-496) precomputed on the host.
struct alignas(16) DObj {
  float c0[3];  // Transform::rot columns (to_local takes dot(c_i, p - pos))
  float inv;    // 1/scale, or 0 for |scale| <= 1e-12 (vec.rs:219)
  float c1[3];
  float scale;
  float c2[3];
  uint32_t mat;
  float pos[3];
  uint32_t flags;
  float p[8];       // prim parameters (p[0] holds the volume id bits for volumes)
  float rep[3];
  float twist;
  float elo[3];
  float bend;
  float round, onion;
  float bc_r;       // bounding-sphere radius
  float k;          // smooth-union radius (Combine::Smooth)
  float bc[3];      // bounding-sphere centre
  uint32_t pad1;
};

// accel.rs:13-19 plus a parent link so the device walks the tree without a stack.
struct alignas(16) DNode {
  float c[3];
  float r;
  int32_t left, right, leaf, parent;
};

// accel.rs:122-125: Segment::One(index) or Segment::Run{tree, unbounded}.
struct alignas(16) DSeg {
  uint32_t kind;        // 0 = one object, 1 = run of plain unions walked through its tree,
                        // 2 = run of plain unions folded linearly (members in unb[])
  uint32_t index;       // one: object index; run: root node index (global)
  uint32_t has_tree;
  uint32_t unb_base;    // run: first entry in the unbounded-object index list
  uint32_t unb_count;
  uint32_t combine;     // one: the object's Combine
  float k;              // one: smooth radius
  uint32_t pad;
};

// One instruction of an object-local CSG expression (mm3e-kit/src/csg.rs:95-118) compiled to
// postfix: leaves push a distance, binary ops pop two and push one.  The evaluation point
// lives in three registers: ENTER maps it into a Transform node's frame, LEAVE scales the
// child's distance and resets the point to the object-local point, after which the ancestor
// ENTERs are replayed (bit-identical to the recursion's saved point).
enum : uint32_t {
  CSG_LEAF = 0,       // push prim_distance(prim, p, point)
  CSG_ENTER = 1,      // point = to_local(xform, point)
  CSG_LEAVE = 2,      // top *= scale; point = object-local point
  CSG_UNION = 3,
  CSG_INTERSECT = 4,
  CSG_SUBTRACT = 5,
  CSG_SMOOTH = 6,     // k = p[0]
  CSG_OFFSET = 7,     // top -= p[0]
  CSG_SHELL = 8,      // top = |top| - p[0]
};
// Deepest reflection recursion (Scene::bounces) the kernels take: 0..4 as nested inlined levels
// or the state machine, 5..16 as the state machine with a 16-level per-lane stack.
constexpr uint32_t kMaxBounces = 16;
constexpr uint32_t kCsgStack = 10;  // postfix depth of any expression of <= 1024 nodes (csg.rs:13-14)

struct alignas(16) DCsgOp {
  uint32_t op, prim;
  uint32_t swap;     // binary: operands were emitted b-first (deeper child first)
  uint32_t pad;
  float p[8];
  float c0[3], inv;  // Transform::rot columns and 1/scale (0 for |scale| <= 1e-12)
  float c1[3], scale;
  float c2[3], pad1;
  float pos[3], pad2;
};

struct alignas(16) DMat {  // mm3e-kit/src/color.rs:37-45
  float albedo[3];
  float metallic;
  float specular;
  float roughness;
  float reflectivity;
  uint32_t checker;
  float emissive[3];
  uint32_t pad;
};

struct alignas(16) DLight {  // lib.rs:503-510
  float vec[3];
  float radius;
  float color[3];
  uint32_t directional;
};

struct alignas(16) DVol {  // mm3e-kit/src/volume.rs:14-23
  uint32_t dims[3];
  uint32_t offset;   // first sample inside vol_data
  float min[3];
  uint32_t valid;    // volume.rs:47-62 has_valid_layout
  float cell[3];
  uint32_t pad;
  float hi[3];       // min + (dims-1)*cell (volume.rs:97-103)
  uint32_t pad2;
};

enum : uint32_t {
  SC_SHADOWS = 1u << 0,
  SC_AO = 1u << 1,
  SC_DUAL_SAFE = 1u << 2,
  SC_SECANT = 1u << 3,
  SC_HAS_GI = 1u << 4,
  SC_SS1 = 1u << 5,  // step_scale >= 1: enhanced sphere tracing with omega = 1.4
};

// Mask levels of the segment candidate grid (DevScene::grid_m).
constexpr uint32_t kGridLevels = 8;

struct alignas(16) DevScene {
  uint32_t n_obj, n_seg, n_mat, n_light;
  uint32_t max_steps, shadow_steps, ao_samples, bounces;
  uint32_t flags, n_vol, aa, pad0;
  // The object/segment/node/run-list tables are one contiguous byte range of the blob, so a
  // block can copy them to LDS in one sweep: stage_bytes long, starting at `obj`; the other
  // tables start st_seg / st_node / st_unb bytes into it.
  uint32_t stage_bytes, st_seg, st_node, st_unb;
  uint32_t gi_dims[3], pad1;
  float max_dist, eps, step_scale, lod;
  float subitize, normal_h, slack, fog_density;
  float sun[3], exposure;
  float ambient[3], bloom_threshold;
  float sky_ambient[3], bloom_intensity;
  float fog[3], pad2;
  float gi_min[3], pad3;
  float gi_size[3], pad4;
  // tables (device addresses inside this allocation)
  const DObj* obj;
  const DSeg* seg;
  const DNode* nodes;
  const uint32_t* unb;
  const DMat* mat;
  const DLight* light;
  const float* gi;       // [probe][6 faces][4] (RGB + pad), probe = (k*ny + j)*nx + i
  const DVol* vol;
  const float* vol_data;
  // Segment candidate grid (grid == nullptr when absent).  Cubic cells of side 1/grid_inv
  // from grid_lo; cell (x, y, z) has index (z * ny + y) * nx + x and index nx*ny*nz stands for
  // every point outside the grid.  Each cell holds kGridLevels bit masks over the segments
  // (grid_words words each, [cell][level][word]): bit s of level l is clear only if segment s's
  // members all lie farther than grid_m[l] from every point of the cell; the outside cell's level
  // 0 holds only the unbounded segments.  Built only for scenes
  // without subtract segments, where the accumulator never increases along the fold.
  float grid_lo[3], grid_inv;
  uint32_t grid_dims[3], grid_words;
  float grid_m[kGridLevels];
  float grid_kmax, grid_eps;
  // Box around the bounded segments' spheres (rounded outwards): a point outside the grid whose
  // distance to this box exceeds its threshold is farther than that from every bounded member.
  float grid_in_lo[3], grid_in_hi[3];
  const uint32_t* grid;
  // Postfix CSG programs; a CSG object's p[0] / p[1] hold its first op and op count (bits).
  const DCsgOp* csg;
};

// Per-frame camera record (mm3e-kit/src/camera.rs:39-47) plus the image size.
// Table view used by the field walk: device addresses in HBM, or the same tables staged in LDS.
struct Tables {
  const DObj* obj;
  const DSeg* seg;
  const DNode* nodes;
  const uint32_t* unb;
};

struct FrameParams {
  float eye[3], fov_scale;
  float fwd[3], pad0;
  float right[3], pad1;
  float up[3], pad2;
  uint32_t width, height;
  float aspect, pad3;
};

// Thin lens of a film render (Lens, mm3e-kit/src/camera.rs:14-35).
struct FilmParams {
  float radius, focus;
  uint32_t pad0, pad1;
};

}  // namespace mm3e
