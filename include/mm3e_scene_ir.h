/*
 * mm3e_scene_ir.h — the flat, little-endian binary scene description that crosses the
 * C ABI of libmm3e_hip.so.
 *
 * The reference GPU backend receives a Rust `mm3e_orchestrator::Scene` value
 * (mm3e-orchestrator/src/lib.rs:560-592) and bakes it into shader text
 * (mm3e-gpu/src/wgsl.rs:41-102).  A C ABI cannot take a Rust struct, and the reference's
 * own `.mm3e` text format is lossy (no CSG, volumes, GI, shadow_steps, ao_samples or
 * normal_h: mm3e-orchestrator/src/scene_io.rs:47-105), so the scene travels as this
 * IR instead.  It carries every `Scene` field the beauty path reads:
 *
 *   header                       mm3e_ir_header            (fixed size)
 *   objects[n_objects]           mm3e_ir_object            Scene::objects   (lib.rs:565)
 *   materials[n_materials]       mm3e_ir_material          Scene::materials (lib.rs:566)
 *   lights[n_lights]             mm3e_ir_light             Scene::lights    (lib.rs:567)
 *   volumes[n_volumes]           mm3e_ir_volume            Scene::volumes   (lib.rs:587)
 *   volume_data[]                f32                       SdfVolume::data  (volume.rs:22)
 *   csg_nodes[n_csg_nodes]       mm3e_ir_csg_node          Scene::csgs      (lib.rs:589)
 *   csg_roots[n_csg_exprs]       u32 (node index)
 *   gi_cubes[nx*ny*nz*6*3]       f32                       GiVolume::cubes  (gi.rs:22)
 *
 * Every array starts at the byte offset the header names, 4-byte aligned.  All floats are
 * IEEE-754 binary32 and are consumed bit-for-bit: the host that builds the IR owns every
 * rounding decision (e.g. `Vec3::normalize` of a light direction), exactly as the Rust
 * `Scene` value does.
 */
#ifndef MM3E_SCENE_IR_H
#define MM3E_SCENE_IR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM3E_IR_MAGIC "MM3EIR01"
#define MM3E_IR_VERSION 1u

/* Prim kinds — mm3e-orchestrator/src/lib.rs:157-212, in declaration order. */
enum mm3e_prim_kind {
  MM3E_PRIM_SPHERE = 0,     /* p[0]=r                              sdf.rs:33  */
  MM3E_PRIM_BOX = 1,        /* p[0..3]=half                        sdf.rs:38  */
  MM3E_PRIM_ROUND_BOX = 2,  /* p[0..3]=half p[3]=radius            sdf.rs:44  */
  MM3E_PRIM_TORUS = 3,      /* p[0]=major p[1]=minor               sdf.rs:54  */
  MM3E_PRIM_CYLINDER = 4,   /* p[0]=h p[1]=r                       sdf.rs:60  */
  MM3E_PRIM_CAPSULE = 5,    /* p[0..3]=a p[3..6]=b p[6]=r          sdf.rs:69  */
  MM3E_PRIM_CONE = 6,       /* p[0]=r1 p[1]=r2 p[2]=h (round cone) sdf.rs:78  */
  MM3E_PRIM_ELLIPSOID = 7,  /* p[0..3]=r                           sdf.rs:94  */
  MM3E_PRIM_OCTAHEDRON = 8, /* p[0]=s                              sdf.rs:104 */
  MM3E_PRIM_HEX_PRISM = 9,  /* p[0]=r p[1]=h                       sdf.rs:110 */
  MM3E_PRIM_PLANE = 10,     /* p[0..3]=n p[3]=h                    sdf.rs:49  */
  MM3E_PRIM_CSG = 11,       /* p[0]=expression id (bit pattern u32) lib.rs:201 */
  MM3E_PRIM_VOLUME = 12,    /* p[0]=volume id (bit pattern u32)     lib.rs:205 */
  MM3E_PRIM_SURFACE = 13    /* native triangle shell: always rejected by the GPU path */
};

/* Combine — lib.rs:297-304. */
enum mm3e_combine { MM3E_COMBINE_UNION = 0, MM3E_COMBINE_SMOOTH = 1, MM3E_COMBINE_SUBTRACT = 2 };

/* GI payload kinds. */
enum mm3e_gi_mode {
  MM3E_GI_NONE = 0,   /* Scene::gi == None                                         */
  MM3E_GI_BAKED = 1,  /* gi_cubes holds a baked GiVolume (gi_min/gi_size/gi_dims)  */
  MM3E_GI_BAKE = 2    /* bake request: Scene::bake_gi(gi_dims, gi_samples) at load  */
};

typedef struct mm3e_ir_header {
  char magic[8];            /* "MM3EIR01" */
  uint32_t version;         /* MM3E_IR_VERSION */
  uint32_t total_bytes;     /* size of the whole IR blob */
  uint32_t width, height;   /* Scene::width/height (the CPU renderer's image size) */
  uint32_t aa, bounces;     /* Scene::aa (0 acts as 1), Scene::bounces */
  uint32_t shadows, ao;     /* bools */
  uint32_t mode;            /* RenderMode (lib.rs:49-61): 0 Beauty, 1 Normal, 2 Depth, 3 Ao, 4 Steps,
                               5 Albedo; 1..5 render one centre sample straight to RGBA8 */
  uint32_t appearance;      /* nonzero = texture bindings present (rejected, wgsl.rs:42-44) */
  float sun_dir[3];
  float fog_density;
  float ambient[3];
  float _pad0;
  float sky_ambient[3];
  float _pad1;
  float fog[3];
  float _pad2;
  /* Marcher (mm3e-kit/src/march.rs:41-79) */
  uint32_t max_steps;
  float max_dist, eps, step_scale;
  uint32_t shadow_steps, ao_samples;
  float lod_footprint, subitize;
  uint32_t secant;
  float normal_h;
  uint32_t _pad3[2];
  /* Post (mm3e-orchestrator/src/post.rs:13-29) */
  float exposure;
  uint32_t bloom;
  float bloom_threshold, bloom_intensity;
  uint32_t bloom_radius;
  uint32_t _pad4[3];
  /* counts */
  uint32_t n_objects, n_materials, n_lights, n_volumes;
  uint32_t n_csg_exprs, n_csg_nodes, gi_mode, gi_samples;
  uint32_t gi_dims[3];
  uint32_t n_volume_floats;
  float gi_min[3];
  float _pad5;
  float gi_size[3];
  float _pad6;
  /* byte offsets from the start of the blob */
  uint32_t off_objects, off_materials, off_lights, off_volumes;
  uint32_t off_volume_data, off_csg_nodes, off_csg_roots, off_gi_cubes;
} mm3e_ir_header;

/* One placed object — lib.rs:330-336 (Prim + Transform + mat + Combine + Modifiers). */
typedef struct mm3e_ir_object {
  uint32_t prim;      /* mm3e_prim_kind */
  uint32_t mat;       /* material id (taken modulo n_materials at shading time) */
  uint32_t combine;   /* mm3e_combine */
  float k;            /* smooth-union radius when combine == SMOOTH */
  float p[8];         /* prim parameters, see mm3e_prim_kind */
  float rot[9];       /* Transform::rot as columns: cols[0].xyz, cols[1].xyz, cols[2].xyz */
  float pos[3];       /* Transform::pos */
  float scale;        /* Transform::scale */
  uint32_t mirror;    /* Modifiers::mirror bit0=x bit1=y bit2=z */
  float repeat[3];    /* Modifiers::repeat */
  float twist, bend;  /* Modifiers::twist / bend */
  float elongate[3];  /* Modifiers::elongate */
  float round, onion; /* Modifiers::round / onion */
} mm3e_ir_object;     /* 36 words */

/* Material — mm3e-kit/src/color.rs:37-45. */
typedef struct mm3e_ir_material {
  float albedo[3];
  float metallic;
  float specular;
  float roughness;
  float reflectivity;
  float emissive[3];
  uint32_t checker;
  uint32_t _pad;
} mm3e_ir_material;   /* 12 words */

/* Light — lib.rs:503-510. `vec` is the unit direction toward a directional light, or a position. */
typedef struct mm3e_ir_light {
  float vec[3];
  float radius;
  float color[3];
  uint32_t directional;
} mm3e_ir_light;      /* 8 words */

/* SdfVolume header — mm3e-kit/src/volume.rs:14-23; samples live in volume_data[offset..]. */
typedef struct mm3e_ir_volume {
  uint32_t dims[3];
  uint32_t offset;    /* first sample, in floats, inside volume_data */
  float min[3];
  float cell[3];
  uint32_t _pad[2];
} mm3e_ir_volume;     /* 12 words */

/* CSG expression node — mm3e-kit/src/csg.rs:17-91. Children are node indices. */
enum mm3e_csg_op {
  MM3E_CSG_PRIM = 0,       /* leaf: prim kind in `prim`, parameters in p[] */
  MM3E_CSG_TRANSFORM = 1,  /* a = child; rot/pos/scale */
  MM3E_CSG_UNION = 2,
  MM3E_CSG_INTERSECT = 3,
  MM3E_CSG_SUBTRACT = 4,
  MM3E_CSG_SMOOTH_UNION = 5, /* k = p[0] */
  MM3E_CSG_OFFSET = 6,       /* a = child, distance = p[0] */
  MM3E_CSG_SHELL = 7         /* a = child, thickness = p[0] */
};
typedef struct mm3e_ir_csg_node {
  uint32_t op;
  uint32_t prim;
  uint32_t a, b;
  float p[8];
  float rot[9];
  float pos[3];
  float scale;
  uint32_t _pad[3];
} mm3e_ir_csg_node;   /* 28 words */

#ifdef __cplusplus
}
#endif

#endif /* MM3E_SCENE_IR_H */
