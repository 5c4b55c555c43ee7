// mesh.h — triangle meshes on the host side of the mesh -> SDF bake (mm3e-orchestrator/src/mesh.rs):
// OBJ parsing, validation, the bake's grid layout and the nearest-triangle BVH, laid out for the
// device bake kernels (mesh_bake.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mm3e {

struct HostMesh {
  std::vector<float> pos;      // [n][3]
  std::vector<uint32_t> tris;  // [t][3]
  size_t n_pos() const { return pos.size() / 3; }
  size_t n_tris() const { return tris.size() / 3; }
};

// parse_obj (mesh.rs:57-103) followed by Mesh::validate (mesh.rs:33-52).
bool parse_obj(const char* text, size_t len, HostMesh& m, std::string& err);
// Mesh::validate (mesh.rs:33-52).
bool validate_mesh(const HostMesh& m, std::string& err);

// Grid of a bake: bake_sdf's bounds, per-axis sample counts and spacing (mesh.rs:265-279).
struct BakeLayout {
  uint32_t dims[3];
  float lo[3];
  float cell[3];
  uint64_t samples() const { return (uint64_t)dims[0] * dims[1] * dims[2]; }
};
// Validates the mesh and the padding; the layout itself is checked by check_layout, which the
// reference does when the finished volume is built (SdfVolume::new, after the BVH build).
bool bake_layout(const HostMesh& m, uint32_t resolution, float padding, BakeLayout& out, std::string& err);
bool check_layout(const BakeLayout& L, std::string& err);

// One node of the nearest-triangle BVH (TriBvh, mesh.rs:148-260), 32 bytes.  Interior nodes:
// a = left child, b = right child.  Leaves: a = -(triangle + 1), b = 0.
struct BakeNode {
  float lo[3];
  int32_t a;
  float hi[3];
  uint32_t b;
};

struct TriBvhHost {
  std::vector<BakeNode> nodes;
  uint32_t root = 0;
  uint32_t depth = 0;  // interior nodes on the longest root-to-leaf path
};
// TriBvh::build (mesh.rs:166-222): median split on the longest centroid axis, ties by triangle
// index; triangles whose doubled area squared is below 1e-16 are left out.  False when no
// triangle is left.
bool build_tri_bvh(const HostMesh& m, TriBvhHost& out);

}  // namespace mm3e
