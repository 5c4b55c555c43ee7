/*
 * mm3e_hip.h — C ABI of libmm3e_hip.so, the MI355X-native per-pixel SDF render path of MM3E.
 *
 * Drop-in boundary for the reference GPU backend `mm3e-gpu` (mm3e-gpu/src/lib.rs): compile a
 * scene once, re-upload only the camera per frame, render to RGBA8 / a framebuffer / BMP.
 * Each entry point names the reference item it replaces.  Conventions:
 *   - every function returns an int status, MM3E_OK (0) on success;
 *   - functions that can fail with a message copy it into (err, errlen) when err != NULL,
 *     NUL-terminated and truncated to errlen; the reference returns these as
 *     `Result<_, String>` (lib.rs:80-142) or panics via `expect` (lib.rs:389-392);
 *   - the caller owns every output buffer;
 *   - a renderer may be used from several host threads (its calls are serialised by an
 *     internal mutex); one compiled scene must be driven by one host thread at a time;
 *   - scenes travel as the flat IR of mm3e_scene_ir.h.
 * The library has no CPU fallback: without a usable gfx950 device every constructor fails.
 */
#ifndef MM3E_HIP_H
#define MM3E_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "mm3e_scene_ir.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MM3E_OK 0
#define MM3E_ERR_INVALID 1   /* bad argument / IR / unsupported scene                   */
#define MM3E_ERR_DEVICE 2    /* HIP runtime failure (device lost, OOM, launch failure)  */
#define MM3E_ERR_BUSY 3      /* admission refused (serving: HBM budget / queue full)    */

/* mm3e-gpu/src/lib.rs:22 */
#define MM3E_MAX_DYN 24

typedef struct mm3e_renderer mm3e_renderer; /* GpuRenderer (lib.rs:57-61) */
typedef struct mm3e_scene mm3e_scene;       /* GpuScene    (lib.rs:289-300) */

/* Camera (mm3e-kit/src/camera.rs:39-47), host-built as Camera::look_at does. */
typedef struct mm3e_camera {
  float eye[3];
  float forward[3];
  float right[3];
  float up[3];
  float fov_scale; /* tan(fov_y / 2) */
} mm3e_camera;

/* DynSphere (mm3e-gpu/src/lib.rs:27-32). */
typedef struct mm3e_dyn_sphere {
  float pos[3];
  float radius;
  float albedo[3];
  float metallic;
} mm3e_dyn_sphere;

/* ---- renderer ------------------------------------------------------------------------ */

/* GpuRenderer::new / with_adapter (lib.rs:80-118).  adapter_filter == NULL honours the
 * MM3E_GPU_ADAPTER environment variable; otherwise the first HIP device whose descriptor
 * ("<name> (<gcnArchName>)") contains the filter, case-insensitively, is opened.
 * Error: "no GPU adapter matching '<f>'; available: <a> | <b>". */
int mm3e_renderer_create(const char* adapter_filter, mm3e_renderer** out, char* err, size_t errlen);
/* GpuRenderer::adapter_name (lib.rs:120-122); valid until the renderer is destroyed. */
const char* mm3e_renderer_adapter_name(const mm3e_renderer* r);
/* GpuRenderer::wait_idle (lib.rs:126-128): block until all submitted work finished. */
int mm3e_renderer_wait_idle(mm3e_renderer* r);
void mm3e_renderer_destroy(mm3e_renderer* r);
/* list_adapters (lib.rs:73-76): writes "a | b | ..." into buf; returns the device count. */
int mm3e_list_adapters(char* buf, size_t buflen);

/* ---- scene compile ------------------------------------------------------------------- */

/* GpuRenderer::compile_checked (lib.rs:138-285).  Validates w,h in 1..=16384
 * ("GpuRenderer::compile: {w}x{h} out of range 1..=16384"), rejects native triangle
 * surfaces and texture appearance with the reference's messages (wgsl.rs:42-47), uploads
 * the scene tables to HBM, bakes a requested GI volume on the device, and allocates the
 * resident framebuffers.  Captures aa and bounces at compile time, as GpuScene does. */
int mm3e_scene_compile(mm3e_renderer* r, const void* scene_ir, size_t ir_len, uint32_t width, uint32_t height,
                       mm3e_scene** out, char* err, size_t errlen);
void mm3e_scene_destroy(mm3e_scene* s);
/* The validation half of compile_checked without a device: the same checks and messages as
 * mm3e_scene_compile (lib.rs:139-141, wgsl.rs:42-47), nothing allocated. */
int mm3e_scene_check(const void* scene_ir, size_t ir_len, uint32_t width, uint32_t height, char* err, size_t errlen);

/* ---- per-frame ----------------------------------------------------------------------- */

/* GpuScene::render_rgba / render_rgba_dyn (lib.rs:377-402): unpadded RGBA8, top-down, alpha
 * 255; out_len must be >= w*h*4.  Dynamic spheres are unioned as appended objects with the
 * default material (the CPU backend's semantics, mm3e-cpu-parallel/src/lib.rs:50-65); at most
 * MM3E_MAX_DYN are used, extra spheres are ignored as in lib.rs:306-307. */
int mm3e_scene_render_rgba(mm3e_scene* s, const mm3e_camera* cam, const mm3e_dyn_sphere* dyn, uint32_t n_dyn,
                           uint8_t* out, size_t out_len);
/* GpuScene::render_only (lib.rs:363-374): enqueue a frame, no readback, no wait. */
int mm3e_scene_render_only(mm3e_scene* s, const mm3e_camera* cam);
/* Scene-linear HDR radiance before post (render_linear, lib.rs:888-947): w*h*3 floats. */
int mm3e_scene_render_linear(mm3e_scene* s, const mm3e_camera* cam, float* out, size_t out_floats);
/* Centre-pixel primary-hit AOVs (film.rs:114-118): hit mask, material id (u32::MAX on a
 * miss), camera-forward depth (+inf on a miss), march steps.  Any pointer may be NULL. */
int mm3e_scene_render_aov(mm3e_scene* s, const mm3e_camera* cam, uint8_t* hit, uint32_t* mat, float* depth,
                          uint32_t* steps);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): one frame of the
 * compiled scene, as mm3e_scene_render_rgba renders it, written as RGBA8 (w*h*4 bytes,
 * top-down, alpha 255) into the caller's DEVICE buffer rgba_dev.  The frame's HDR radiance and
 * bloom intermediate (24 bytes per pixel) live in a fixed device array of
 * MM3E_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on until the process
 * ends; larger frames are rejected with a message (at most 2,796,202 pixels, e.g. 1920x1080).
 * Calls no allocator.  Enqueued on the renderer's stream without waiting
 * (mm3e_renderer_wait_idle waits).  All calls on one GPU share that array and it keeps the last
 * frame's contents: calls that overlap on one GPU (renderers on other streams) race.  The frame
 * is one kernel in tile index order (no tile order, no queued stages); dynamic spheres are not
 * taken; debug render modes write their image directly. */
#define MM3E_NOALLOC_SCRATCH_BYTES (64ull << 20)
int mm3e_scene_render_rgba_noalloc(mm3e_scene* s, const mm3e_camera* cam, void* rgba_dev, size_t rgba_len, char* err,
                                   size_t errlen);
/* Thin lens (Lens, mm3e-kit/src/camera.rs:14-35): the eye is the lens centre; the focus plane
 * is perpendicular to the camera's forward axis at focus_distance.  Radius 0 = pinhole. */
typedef struct mm3e_lens {
  float aperture_radius;
  float focus_distance;
} mm3e_lens;
/* render_film (mm3e-orchestrator/src/film.rs:63-198): scene-linear film of the compiled scene,
 * before any exposure or post pass, always Beauty whatever the scene's debug mode.  aa^2
 * samples per pixel integrate the pixel footprint and the aperture (a single sample uses the
 * lens centre); per pixel:
 *   beauty[3]      mean of hit radiance and primary-miss sky (equal to render_linear, bit for
 *                  bit, when the aperture radius is 0);
 *   foreground[3]  hit radiance / samples (premultiplied by coverage);
 *   alpha          fraction of samples whose primary ray hit a surface;
 *   depth, normals[3], material_ids   the centre pinhole ray's hit: camera-forward depth,
 *                  world normal, material (+inf, 0, u32::MAX on a miss).
 * Any output may be NULL.  An invalid lens fails with Lens::validate's message. */
int mm3e_scene_render_film(mm3e_scene* s, const mm3e_camera* cam, const mm3e_lens* lens, float* beauty,
                           float* foreground, float* alpha, float* depth, float* normals, uint32_t* material_ids,
                           char* err, size_t errlen);
/* Raw GI probe volume of the compiled scene (GiVolume::raw, gi.rs:133-135): min, size, dims
 * and, when cubes != NULL, the [probe][6][3] irradiance cubes.  *n_probes = 0 if no GI. */
int mm3e_scene_gi_raw(mm3e_scene* s, float* min3, float* size3, uint32_t* dims3, float* cubes, uint64_t* n_probes);
/* Compiled layout: objects, world-plan segments (accel.rs:122-125), cells and words per
 * mask of the segment candidate grid (0 when the scene has none) and LDS bytes of the
 * staged tables (0 when they are read from HBM). */
int mm3e_scene_info(const mm3e_scene* s, uint32_t* n_objects, uint32_t* n_segments, uint32_t* grid_cells,
                    uint32_t* grid_words, uint32_t* stage_bytes);
/* Name of the render kernel the plain per-frame launch of this scene dispatches (the
 * LDS-staged or the HBM-table variant, for the scene's reflection depth). */
int mm3e_scene_kernel_name(const mm3e_scene* s, char* buf, size_t buflen);
/* Last frame's device time of the render kernel and of the post kernels, in milliseconds. */
int mm3e_scene_last_timing(mm3e_scene* s, float* render_ms, float* post_ms);
/* HIP-event kernel times summed over the frames launched since the last reset (at most the
 * last 128 frames are kept): *frames, sum of render-kernel ms, sum of post-kernel ms.
 * Synchronises the renderer's stream.  reset != 0 starts a new window afterwards. */
int mm3e_scene_timing(mm3e_scene* s, int reset, uint64_t* frames, double* render_ms_sum, double* post_ms_sum);

/* ---- parity probes (the reference's test-only `probe` / `film_probe` kernels,
 *      wgsl.rs:641-646, film_depth_tests.rs:31-42) ---------------------------------------- */
int mm3e_scene_field(mm3e_scene* s, const float* pts, uint32_t n, float* dist, uint32_t* mat);
int mm3e_scene_march(mm3e_scene* s, const float* rays, uint32_t n, float* t, uint8_t* hit, uint32_t* mat,
                     float* normal, uint32_t* steps);

/* ---- host helpers -------------------------------------------------------------------- */

/* Camera::look_at (camera.rs:50-55) and orbit_camera (lib.rs:1426-1435). */
void mm3e_camera_look_at(const float eye[3], const float target[3], const float up[3], float fov_y, mm3e_camera* out);
void mm3e_orbit_camera(const float target[3], float radius, float yaw, float pitch, float fov_y, mm3e_camera* out);
/* Framebuffer::to_bmp (framebuffer.rs:55-98) of an opaque RGBA8 image.  With out == NULL
 * only *out_len is set to the required size. */
int mm3e_write_bmp(const uint8_t* rgba, uint32_t w, uint32_t h, uint8_t* out, size_t* out_len);

/* ---- scene code generation ------------------------------------------------------------
 * The twin of the reference's WGSL code generator (mm3e-gpu/src/wgsl.rs:34-362): each compiled
 * scene's world fold is emitted as HIP source with the scene's constants folded in and compiled
 * at run time (hipRTC) into a render kernel for that scene.  The generated kernel performs the
 * same floating-point operations as the table-driven kernels, so images are identical; only
 * speed differs.  Scenes of more than 2000 objects keep the table kernels (the cap bounds the
 * run-time compile: about 45 s for 1000 objects; MM3E_CODEGEN_MAX_OBJECTS overrides).  Modes:
 * OFF (table kernels only), BACKGROUND (default: compile on worker
 * threads, the scene renders with the table kernels until its kernel is ready), SYNC (compile
 * inside mm3e_scene_compile).  The environment variable MM3E_CODEGEN=off|background|sync sets
 * the default for new renderers. */
enum { MM3E_CODEGEN_OFF = 0, MM3E_CODEGEN_BACKGROUND = 1, MM3E_CODEGEN_SYNC = 2 };
enum { MM3E_CODEGEN_NONE = 0, MM3E_CODEGEN_QUEUED = 1, MM3E_CODEGEN_READY = 2, MM3E_CODEGEN_FAILED = 3 };
/* Mode for scenes compiled afterwards on this renderer. */
int mm3e_renderer_set_codegen(mm3e_renderer* r, int mode);
/* Generated modules of scenes compiled afterwards also hold the film (render_film, film.rs:63-198)
 * and G-buffer / hole (render_gbuffer / reproject_hybrid, lib.rs:951-1080) kernels, which then
 * replace the table-driven ones for those calls; default off (the module takes about twice as
 * long to compile).  Images are the same bit for bit either way.  MM3E_CODEGEN_FILM=1 sets the
 * default for new renderers (and makes mm3e_scene_emit_source include them). */
int mm3e_renderer_set_codegen_film(mm3e_renderer* r, int on);
/* Tile order for generated kernels of scenes compiled afterwards (default on; the environment
 * variable MM3E_TILE_ORDER=off sets the default off): the frame's tiles are dispatched most
 * expensive first, ordered by the previous frame's per-tile costs, so a frame whose cost sits
 * in a few long tiles does not end with them running alone.  The image is the same bit for
 * bit either way. */
int mm3e_renderer_set_tile_order(mm3e_renderer* r, int on);
/* Queued frames (no reference counterpart; the reference's frame is one kernel, wgsl.rs:562-581,
 * which this replaces for the renderer's subsequent frames): budget > 0 splits each frame into
 * a first march pass of at most `budget` field evaluations per ray, `passes` continuation
 * passes over the compacted queue of unfinished rays (the last runs them to their end; 0..7),
 * shading of the compacted hit queue and a per-pixel resolve.  blocks_per_cu sizes the
 * grid-stride stages (0: one lane per queue entry; at most 64).  budget 0 restores one kernel
 * per frame.  Generated scene kernels carry the queued stages only when the scene is compiled
 * while queued frames are on (otherwise the table kernels run them).  Every image is the same
 * bit for bit either way.  The environment variable MM3E_WAVEFRONT="budget,passes,bpc" sets
 * the renderer's default. */
int mm3e_renderer_set_wavefront(mm3e_renderer* r, uint32_t budget, uint32_t passes, uint32_t blocks_per_cu);
/* A scene's own queued-frame setting (own != 0; same arguments as above), or back to the
 * renderer's (own == 0).  No reference counterpart. */
int mm3e_scene_set_wavefront(mm3e_scene* s, int own, uint32_t budget, uint32_t passes, uint32_t blocks_per_cu);
/* Autotuner over the frame's form (no reference counterpart): renders reps frames of `cam`
 * (1..127) with one kernel and with queued frames at several step budgets, pass counts and
 * grid sizes, keeps the fastest render stage as the scene's own setting and writes one line
 * per candidate into report.  Images are the same under every candidate. */
int mm3e_scene_autotune(mm3e_scene* s, const mm3e_camera* cam, uint32_t reps, char* report, size_t report_len);
/* Device self-test (no reference counterpart): the render kernels' square root (a shortened
 * correctly rounded sequence) against the compiler's IEEE sqrtf over all 2^32 f32 inputs;
 * *mismatches counts the inputs whose results differ (0 expected). */
int mm3e_selftest_sqrt(mm3e_renderer* r, uint64_t* mismatches);
/* State of a scene's generated kernel (MM3E_CODEGEN_NONE/QUEUED/READY/FAILED); wait != 0
 * blocks until it is READY or FAILED.  On FAILED the compiler's message goes to err. */
int mm3e_scene_codegen_status(mm3e_scene* s, int wait, int* state, char* err, size_t errlen);
/* wgsl::build_shader_checked twin (wgsl.rs:34-102): the generated HIP source of a scene IR (as
 * mm3e_scene_compile would build it; with the queued-frame stages when MM3E_WAVEFRONT sets a
 * budget).  With src == NULL or *src_len too small, *src_len is set
 * to the size needed (including the terminating NUL) and MM3E_ERR_INVALID is returned. */
int mm3e_scene_emit_source(const void* scene_ir, size_t ir_len, char* src, size_t* src_len);

/* ---- the `.mm3e` text scene format (mm3e-orchestrator/src/scene_io.rs) --------------------
 * parse (scene_io.rs:200-331): text -> scene IR + camera; errors carry the reference's
 * messages ("line N: ...").  With ir == NULL or *ir_len too small, *ir_len is set to the size
 * needed and MM3E_ERR_INVALID is returned.  The reference's format drops CSG, volumes, GI, the
 * shadow/AO budgets, LOD, the march boosts and normal_h (scene_io.rs:47-62, 103-105); this
 * library extends it losslessly with directives written only when a scene needs them
 * (sun_raw, marcher_ext, gi bake|baked, volume sdfv|raw, csgnode, csgexpr, obj csg, obj volume;
 * binary payloads in base64), so IR -> text -> IR returns the same bytes and a scene inside the
 * reference's subset is written exactly in the reference's format. */
int mm3e_scene_parse_text(const char* text, size_t text_len, uint8_t* ir, size_t* ir_len, mm3e_camera* cam,
                          char* err, size_t errlen);
/* try_serialize (scene_io.rs:32-164): scene IR + camera -> text (numbers as the shortest
 * decimal that reads back to the same f32, plus the extension directives above), NUL-terminated;
 * same size protocol as above.  Texture appearance and triangle surfaces are refused with the
 * reference's messages. */
int mm3e_scene_to_text(const void* scene_ir, size_t ir_len, const mm3e_camera* cam, char* text, size_t* text_len,
                       char* err, size_t errlen);

/* ---- G-buffer frames and reprojection (mm3e-orchestrator/src/reproject.rs, lib.rs:951-1080) ---
 * A GFrame (reproject.rs:24-36) lives in device memory: display RGBA8, primary-ray hit
 * distance (+inf for sky), the index of the moving object each pixel belongs to (-1: static or
 * sky) and the camera it was rendered with.  Frame generation warps it into later cameras
 * instead of rendering them. */
typedef struct mm3e_gframe mm3e_gframe;
/* A moving object for render_gbuffer's tags: a hit within radius + 0.05 of center is tagged. */
typedef struct mm3e_mover {
  float center[3];
  float radius;
} mm3e_mover;
int mm3e_gframe_create(mm3e_renderer* r, uint32_t width, uint32_t height, mm3e_gframe** out, char* err,
                       size_t errlen);
void mm3e_gframe_destroy(mm3e_gframe* g);
/* Host <-> device copies of a GFrame: color w*h*4 bytes, depth w*h f32, obj w*h i32. */
int mm3e_gframe_upload(mm3e_gframe* g, const mm3e_camera* cam, const uint8_t* color, const float* depth,
                       const int32_t* obj);
int mm3e_gframe_read(mm3e_gframe* g, mm3e_camera* cam, uint8_t* color, float* depth, int32_t* obj);
/* render_gbuffer (lib.rs:951-1030) into g (same renderer, the scene's size): one centre sample
 * per pixel through the full shading (reflections included), displayed with exposure, ACES and
 * gamma (no bloom); movers (at most 256) tag the pixels whose hit lies within radius + 0.05. */
int mm3e_scene_render_gbuffer(mm3e_scene* s, const mm3e_camera* cam, const mm3e_mover* movers, uint32_t n_movers,
                              mm3e_gframe* g, char* err, size_t errlen);
/* reproject::warp / reproject::reproject (reproject.rs:82-170): forward-warp g into cam with
 * per-object deltas (mover_deltas[n][3], world units; at most 256), nearest point winning; with
 * fill_holes each hole takes the last filled pixel to its left on its row, then leading holes
 * the first to their right (reproject); without, holes stay (0,0,0,0) (warp).  out: w*h*4;
 * filled (optional, w*h bytes): the warp's coverage before any fill. */
int mm3e_gframe_reproject(mm3e_gframe* g, const mm3e_camera* cam, const float* mover_deltas, uint32_t n_deltas,
                          int fill_holes, uint8_t* out, uint8_t* filled, char* err, size_t errlen);
/* reproject_hybrid (lib.rs:1043-1080): the warp, then only the holes rendered for real through
 * the scene at the new camera (display colour as render_gbuffer).  *n_holes = pixels rendered. */
int mm3e_scene_reproject_hybrid(mm3e_scene* s, mm3e_gframe* g, const mm3e_camera* cam, const float* mover_deltas,
                                uint32_t n_deltas, uint8_t* out, uint32_t* n_holes, char* err, size_t errlen);

/* ---- meshes and sampled SDF volumes --------------------------------------------------------
 * mm3e-orchestrator/src/mesh.rs (OBJ import, mesh -> SDF bake) and mm3e-kit/src/volume.rs
 * (SdfVolume, the `.sdfv` codec).  A mesh travels as positions[n_positions][3] (f32) and
 * triangles[n_triangles][3] (u32 vertex indices); a volume as an mm3e_sdf_layout plus
 * dims[0]*dims[1]*dims[2] f32 samples at min + (i,j,k)*cell, index (k*ny + j)*nx + i. */
typedef struct mm3e_sdf_layout {
  uint32_t dims[3]; /* SdfVolume::dims (volume.rs:16) */
  float min[3];     /* SdfVolume::min  (volume.rs:18) */
  float cell[3];    /* SdfVolume::cell (volume.rs:20) */
} mm3e_sdf_layout;

/* parse_obj (mesh.rs:57-103, then Mesh::validate, mesh.rs:33-52): `v` and `f` lines, faces
 * fan-triangulated, `v/vt/vn` forms and negative indices accepted, other directives ignored;
 * errors carry the reference's messages ("line N: ...").  With positions and triangles both
 * NULL only the counts are returned; otherwise the capacities (in vertices / triangles) must
 * hold them. */
int mm3e_mesh_parse_obj(const char* text, size_t text_len, float* positions, size_t pos_cap, uint32_t* triangles,
                        size_t tri_cap, size_t* n_positions, size_t* n_triangles, char* err, size_t errlen);
/* The grid bake_sdf would produce (mesh.rs:265-279): bounds padded by `padding`, `resolution`
 * (clamped to 2..=512) samples along the longest axis.  Host only; same errors as the bake. */
int mm3e_mesh_sdf_layout(const float* positions, size_t n_positions, const uint32_t* triangles, size_t n_triangles,
                         uint32_t resolution, float padding, mm3e_sdf_layout* layout, char* err, size_t errlen);
/* bake_sdf (mesh.rs:265-336) on the renderer's GPU: unsigned distance to the nearest triangle,
 * negative where a +x ray from the sample crosses the mesh an odd number of times.  The samples
 * equal the reference's CPU bake bit for bit.  data_cap (floats) must hold the layout's sample
 * count (query it with mm3e_mesh_sdf_layout). */
int mm3e_mesh_bake_sdf(mm3e_renderer* r, const float* positions, size_t n_positions, const uint32_t* triangles,
                       size_t n_triangles, uint32_t resolution, float padding, float* data, size_t data_cap,
                       mm3e_sdf_layout* layout, char* err, size_t errlen);
/* The same bake as a resident object (for repeated bakes and timing): create validates, builds
 * the nearest-triangle BVH (mesh.rs:166-222) on the host and uploads it; run enqueues the two
 * bake kernels on the renderer's stream; wait returns their GPU times (HIP events); read waits,
 * copies the samples out and rejects non-finite ones as SdfVolume::new does. */
typedef struct mm3e_mesh_baker mm3e_mesh_baker;
int mm3e_mesh_baker_create(mm3e_renderer* r, const float* positions, size_t n_positions, const uint32_t* triangles,
                           size_t n_triangles, uint32_t resolution, float padding, mm3e_mesh_baker** out, char* err,
                           size_t errlen);
int mm3e_mesh_baker_layout(const mm3e_mesh_baker* b, mm3e_sdf_layout* layout);
int mm3e_mesh_baker_info(const mm3e_mesh_baker* b, uint32_t* n_triangles, uint32_t* n_nodes, uint32_t* depth);
int mm3e_mesh_baker_run(mm3e_mesh_baker* b);
int mm3e_mesh_baker_wait(mm3e_mesh_baker* b, float* parity_ms, float* nearest_ms);
int mm3e_mesh_baker_read(mm3e_mesh_baker* b, float* data, size_t data_cap, char* err, size_t errlen);
void mm3e_mesh_baker_destroy(mm3e_mesh_baker* b);

/* SdfVolume::to_sdfv_bytes (volume.rs:115-136): `MM3ESDF1`, u32 nx ny nz, f32 min xyz, f32 cell
 * xyz, then the samples, all little-endian.  The volume is validated first (volume.rs:33-85).
 * With out == NULL or *out_len too small, *out_len is set to the size needed and
 * MM3E_ERR_INVALID is returned. */
int mm3e_sdfv_encode(const mm3e_sdf_layout* layout, const float* data, size_t n_samples, uint8_t* out,
                     size_t* out_len, char* err, size_t errlen);
/* SdfVolume::from_sdfv_bytes (volume.rs:138-174), with its messages.  data == NULL: validate
 * and return the layout only; otherwise data_cap (floats) must hold the samples. */
int mm3e_sdfv_decode(const uint8_t* bytes, size_t len, mm3e_sdf_layout* layout, float* data, size_t data_cap,
                     char* err, size_t errlen);

/* ---- batched rendering ----------------------------------------------------------------
 * Many (compiled scene, camera, size) jobs rendered by one persistent launch that pulls
 * tiles from a global atomic queue, followed by one batched bloom/tonemap over all jobs;
 * the three kernels are captured once into a hipGraph and replayed.  The reference renders
 * one frame per dispatch (mm3e-gpu/src/lib.rs:328-356); every job's image equals
 * mm3e_scene_render_rgba of the same scene at the job's size, bit for bit (a job may use a
 * size other than the one its scene was compiled for).  Jobs keep their own framebuffers. */
typedef struct mm3e_batch mm3e_batch;
typedef struct mm3e_batch_job {
  mm3e_scene* scene;  /* compiled on the same renderer */
  mm3e_camera cam;
  uint32_t width, height;
  uint8_t* out;       /* RGBA8 destination for mm3e_render_batch (may be NULL) */
} mm3e_batch_job;

int mm3e_batch_create(mm3e_renderer* r, const mm3e_batch_job* jobs, uint32_t n_jobs, mm3e_batch** out, char* err,
                      size_t errlen);
/* Time the persistent launch over blocks-per-CU {2,3,4,6,8} and, when single-sample jobs are
 * present, tile shapes {16x16, 32x8, 64x4, 8x32}; keep the fastest and re-capture the graph.
 * A human-readable table of the trials is written to report. */
int mm3e_batch_autotune(mm3e_batch* b, uint32_t reps, char* report, size_t report_len);
/* Change one job's camera for the next launch (the job table is re-uploaded, the graph kept). */
int mm3e_batch_set_camera(mm3e_batch* b, uint32_t job, const mm3e_camera* cam);
/* Enqueue one replay of the batch graph; returns without waiting. */
int mm3e_batch_launch(mm3e_batch* b);
/* Wait for the last launch; *gpu_ms = its device time (queue reset + render + post). */
int mm3e_batch_wait(mm3e_batch* b, float* gpu_ms);
/* Copy job `job`'s RGBA8 image (width*height*4 bytes) to host memory. */
int mm3e_batch_read(mm3e_batch* b, uint32_t job, uint8_t* out, size_t out_len);
int mm3e_batch_info(mm3e_batch* b, uint64_t* total_pixels, uint32_t* total_tiles, uint32_t* blocks_per_cu,
                    uint32_t* tile_w, uint32_t* tile_h, uint32_t* lds_bytes);
void mm3e_batch_destroy(mm3e_batch* b);
/* One-shot: create, launch, wait, read every job with out != NULL, destroy. */
int mm3e_render_batch(mm3e_renderer* r, const mm3e_batch_job* jobs, uint32_t n_jobs, char* err, size_t errlen);

/* ---- serving: continuous batching over resident scenes ---------------------------------
 * Requests (resident compiled scene, camera, size) may be submitted from any thread.  A
 * worker thread admits them FIFO into batches while the HBM budget allows (a request needs
 * w*h*28 bytes of framebuffers, carved from one slab allocated at creation), renders each
 * batch with the persistent tile-queue kernels, and keeps up to max_inflight batches on the
 * device at once (separate default-priority streams); requests arriving meanwhile join the
 * next batch.  A request's image equals its per-frame render.  Latency = completion - submit. */
typedef struct mm3e_server mm3e_server;
typedef struct mm3e_server_config {
  uint64_t hbm_budget_bytes;    /* framebuffer slab (default 4 GiB) */
  uint32_t max_batch_requests;  /* requests per batch (default 4) */
  uint32_t max_inflight;        /* batches on the device at once, 1..8 (default 8) */
  uint32_t queue_capacity;      /* pending requests before submit returns MM3E_ERR_BUSY (4096) */
  uint32_t blocks_per_cu;       /* persistent blocks per CU (default 4) */
} mm3e_server_config;

void mm3e_server_default_config(mm3e_server_config* cfg);
int mm3e_server_create(mm3e_renderer* r, const mm3e_server_config* cfg, mm3e_server** out, char* err, size_t errlen);
/* Register a compiled scene (it stays resident; the caller keeps it alive). */
int mm3e_server_add_scene(mm3e_server* s, mm3e_scene* scene, uint32_t* scene_id);
/* Queue a request; out (w*h*4 bytes, may be NULL) is written before the request completes.
 * MM3E_ERR_BUSY when the queue is full; a request larger than the whole budget completes
 * with MM3E_ERR_BUSY. */
int mm3e_server_submit(mm3e_server* s, uint32_t scene_id, const mm3e_camera* cam, uint32_t w, uint32_t h, uint8_t* out,
                       uint64_t* request_id);
/* A request given as `.mm3e` scene text (scene_io.rs:200-333 with this library's lossless
 * extension, see mm3e_scene_parse_text): the text is parsed (scene + `cam` line), compiled once
 * and kept resident by the server under its IR bytes (later requests with the same scene skip
 * the compile, whatever their camera), then queued as mm3e_server_submit at (w, h), or at the
 * text's `size` when both are 0.  Parse and compile errors are copied into err; the compile
 * (tables, device GI bake, generated kernel) runs on the calling thread. */
int mm3e_server_submit_text(mm3e_server* s, const char* text, size_t text_len, uint32_t w, uint32_t h, uint8_t* out,
                            uint64_t* request_id, char* err, size_t errlen);
int mm3e_server_poll(mm3e_server* s, uint64_t request_id, int* done);
/* Block until the request completes; returns its status and releases it.  Each request id may
 * be waited for once: a second wait on an id (while the first is blocked, or after it returned)
 * returns MM3E_ERR_INVALID. */
int mm3e_server_wait(mm3e_server* s, uint64_t request_id, double* latency_ms);
int mm3e_server_stats(mm3e_server* s, uint64_t* batches, uint64_t* completed, uint64_t* rejected, uint64_t* max_batch,
                      uint64_t* pixels);
/* Finishes queued work, then stops the worker and frees device memory. */
void mm3e_server_destroy(mm3e_server* s);

#ifdef __cplusplus
}
#endif

#endif /* MM3E_HIP_H */
