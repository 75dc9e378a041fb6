/*
 * rz.h — C ABI of librz, the MI355X (gfx950) frame path of the rasteriser.
 *
 * The reference renderer exports exactly one frame-path symbol:
 *
 *   void draw_frame(const vector<vec3>& model_vertices, const vector<Triangle>& faces,
 *                   const vector<vec3>& model_vertnormals, const vector<vec2>& vertuvs,
 *                   vector<Light>& lights, const vector<Material>& materials,
 *                   const Args& arguments,
 *                   CImg<unsigned char>* frame_buffer, CImg<float>* depth_buffer);
 *   (headers/drawing.h:16-18, drawing.cpp:205-258)
 *
 * librz replaces it with plain pointers and sizes.  The record types below keep the
 * byte layout of the reference's own structs (Triangle, Light), so a caller holding
 * the reference's std::vectors passes .data() straight through.  Buffers keep the
 * reference layout: the frame is planar uint8 [3][H][W] (CImg<unsigned char>(W,H,1,3),
 * element (x,y,c) at x + y*W + c*W*H), the depth is float [H][W] (CImg<float>(W,H,1,1)).
 *
 * Every entry point returns an rz_status; no exception crosses the ABI.  A context is
 * bound to one HIP device and owns one HIP stream; calls on one context must not
 * overlap (one context per host thread).
 */
#ifndef RZ_H
#define RZ_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RZ_ABI_VERSION 2

typedef enum rz_status {
    RZ_OK = 0,
    RZ_ERR_INVALID_ARG = 1,   /* null pointer, zero/oversized image, bad tuning value   */
    RZ_ERR_INDEX_OOB = 2,     /* a face index outside its attribute array (reference UB, drawing.cpp:166-173) */
    RZ_ERR_HIP = 3,           /* a HIP runtime call failed; rz_last_error() has the text */
    RZ_ERR_OOM = 4,           /* device allocation failed                                 */
    RZ_ERR_NO_DEVICE = 5      /* no HIP device, or the device index is out of range       */
} rz_status;

/* Reference: headers/face.h:6-13 (struct Triangle).  Separate index streams per
 * attribute, OBJ style.  uvs[i] < 0 means "no texture coordinate" (the reference
 * skips the fetch, drawing.cpp:168-170, and interpolates a zero vec2). normals[i]
 * and material must be valid indices: the reference reads them unchecked, librz
 * rejects them with RZ_ERR_INDEX_OOB.  40 bytes, same field order. */
typedef struct rz_triangle {
    int32_t vertices[3];
    int32_t normals[3];
    int32_t uvs[3];
    int32_t material;
} rz_triangle;

/* Reference: headers/light.h:7-14 (struct Light).  direction is the world-space
 * direction the light travels; trans_dir is written by every draw call with
 * normalize(view * (direction, 0)) exactly like Light::transform
 * (geometry.cpp:124-133).  40 bytes, same field order. */
typedef struct rz_light {
    float direction[3];
    float intensity;
    float colour[3];
    float trans_dir[3];
} rz_light;

/* Reference: headers/material.h:11-25 (class Material).  With tex_channels == 0 the
 * flat diffuse_colour is the albedo (material.cpp:23-25).  Otherwise texels holds the
 * 8-bit image exactly as the file stores it (tex_height rows of tex_width texels,
 * tex_channels interleaved bytes each, 3 or 4 channels); librz applies the Material
 * constructor's CImg::normalize(0,1) (material.h:20-23: min and max taken over all
 * channels) and samples channels 0..2 with CImg::_linear_atXY (material.cpp:17-22).
 *
 * No reference counterpart: an optional tangent-space normal map (normal_map_channels
 * 3 or 4, 0 = none), 8-bit as in its file.  It is sampled at the same uv with the same
 * bilinear rule on texel/255 (no min/max normalisation), decoded t = 2*s - 1, and
 * bends the shading normal: n = normalize(T*t.x + B*t.y + N*t.z) with N the
 * interpolated normal, T the face tangent (from positions and uvs) made orthogonal
 * to N, and B = cross(N, T). */
typedef struct rz_material {
    float diffuse_colour[3];
    int32_t tex_width;
    int32_t tex_height;
    int32_t tex_channels;
    const uint8_t* texels;
    int32_t normal_map_width;
    int32_t normal_map_height;
    int32_t normal_map_channels;
    int32_t reserved;
    const uint8_t* normal_map;
} rz_material;

/* Reference: headers/arguments.h:7-21 (struct Args) — the fields draw_frame reads —
 * plus the projection the reference hard-codes (geometry.cpp:30: fovy 45 degrees,
 * near 0.1, far 6).  The view matrix stays the reference's fixed
 * translate(0,0,-3) (drawing.cpp:224).  rz_default_args() fills the reference
 * defaults (arguments.cpp:20-31). */
typedef struct rz_args {
    uint32_t image_width;
    uint32_t image_height;
    float aspect_ratio;           /* the reference sets (float)width/(float)height, arguments.cpp:39 */
    int32_t flat;                 /* 1: one normal per face, normalize(cross(p1-p0, p2-p0)) in model space */
    int32_t wind_clockwise;       /* --wind-clockwise (drawing.cpp:176-178, 131-132) */
    float scale;
    float displacement[3];
    float tait_bryan_angles[3];   /* (rx, ry, rz), composed Y·X·Z (geometry.cpp:22-25) */
    float fovy_degrees;
    float z_near;
    float z_far;
} rz_args;

/* A mesh in the reference's own representation (fileloader.cpp:79-121 output).
 * With raster_space != 0, vertices holds num_vertices raster-space vec4
 * (x_px, y_px, ndc_z, 1/clip_w) and the position transform is skipped (the
 * screen-space triangle workloads); normals still pass through the normal matrix. */
typedef struct rz_mesh_desc {
    const float* vertices;        /* num_vertices * 3 floats (num_vertices * 4 when raster_space) */
    const float* normals;         /* num_normals * 3 floats */
    const float* uvs;             /* num_uvs * 2 floats */
    const rz_triangle* faces;
    const rz_material* materials;
    int32_t num_vertices;
    int32_t num_normals;
    int32_t num_uvs;
    int32_t num_faces;
    int32_t num_materials;
    int32_t raster_space;
} rz_mesh_desc;

/* The four matrices one frame uses, column-major float[16] (GLM layout:
 * element (col, row) at [4*col + row]). */
typedef struct rz_matrices {
    float model[16];              /* transformation_matrix(scale, displacement, angles) */
    float view[16];               /* transformation_matrix(1, (0,0,-3), 0)              */
    float modelview[16];          /* view * model                                       */
    float camera[16];             /* perspective(fovy, aspect, near, far) * modelview   */
    float normal[16];             /* transpose(inverse(modelview))                      */
} rz_matrices;

/* Counters of one frame (filled when a stats pointer is passed; costs one sync). */
typedef struct rz_frame_stats {
    uint64_t triangles;           /* input faces                                             */
    uint64_t front_triangles;     /* faces passing the backface test (drawing.cpp:176-178)   */
    uint64_t bin_pairs;           /* (front face, tile) pairs binned: the face's clamped bbox,
                                     cut to the pixels its inside test can pass, meets the tile */
    uint64_t overflow_tiles;      /* tiles whose bin overflowed and went the overflow path   */
    uint64_t fragments;           /* (pixel, front face) pairs passing the inside test (drawing.cpp:111) */
    uint64_t covered_pixels;      /* pixels holding a winning fragment                       */
    float ms_vertex;              /* GPU time per pass, HIP events on the context stream;    */
    float ms_bin;                 /* zero unless rz_set_profiling(ctx, 1) (ms_vertex: batches */
                                  /* only; a single frame's bin pass does the transform)    */
    float ms_raster;
    float ms_overflow;
    float ms_shade;
    float ms_frame;
    uint64_t raster_tiles;        /* tiles that received a face: the ones the raster pass visits */
    uint64_t raster_pixels;       /* their pixels inside the image                           */
} rz_frame_stats;

typedef struct rz_ctx rz_ctx;
typedef struct rz_mesh rz_mesh;

/* rz_draw_frame flags */
#define RZ_FLAG_CLEARED        1u  /* frame holds 0 and depth holds 1.0 (the caller's clear, renderer.cpp:85-86,107-108): neither is read */
#define RZ_FLAG_DEVICE_BUFFERS 2u  /* frame/depth/ids are device pointers on the context's device */
#define RZ_FLAG_ASYNC          4u  /* with DEVICE_BUFFERS: return without waiting for the stream */
#define RZ_FLAG_MLP_SHADE      8u  /* shade with the context's learned network instead of the light loop */

/* ---- context ---------------------------------------------------------------- */

/* No reference counterpart (the reference has no device).  Creates a context on HIP
 * device `device` with its own stream (default priority). */
rz_status rz_create(int32_t device, rz_ctx** out_ctx);
void rz_destroy(rz_ctx* ctx);
const char* rz_status_string(rz_status status);
const char* rz_last_error(const rz_ctx* ctx);
int32_t rz_abi_version(void);

/* Tile edge (8, 16 or 32 pixels) and per-tile bin capacity (triangles) used by the
 * next frames.  Defaults: 16 and 0, automatic: a single frame gets room in every tile
 * for every face of its mesh when those bins fit 4 GiB and an eighth of the device
 * memory free at rz_create (no bin can overflow), else 512; batches get 512.  A tile
 * whose faces exceed the capacity takes the overflow path (same pixels, slower). */
rz_status rz_set_tuning(rz_ctx* ctx, int32_t tile_size, int32_t bin_capacity);
rz_status rz_get_tuning(const rz_ctx* ctx, int32_t* tile_size, int32_t* bin_capacity);
/* Single frames and frame graphs: 1 (default) shades each tile in the raster pass from
 * its LDS visibility words, so only tiles whose bin overflowed use the HBM visibility
 * buffer and a separate shade pass; 0 runs the deferred shade over the whole visibility
 * buffer as its own pass.  Results are identical.  Batches always use the latter. */
rz_status rz_set_fused_shade(rz_ctx* ctx, int32_t enable);
int32_t rz_get_fused_shade(const rz_ctx* ctx);
/* 1: bracket every pass with HIP events so rz_frame_stats carries GPU times. */
rz_status rz_set_profiling(rz_ctx* ctx, int32_t enable);
/* The context's hipStream_t, for callers that order their own work against it. */
void* rz_stream(rz_ctx* ctx);
/* Wait for all work queued on the context's stream. */
rz_status rz_sync(rz_ctx* ctx);

/* ---- mesh cache ---------------------------------------------------------------- */

/* Validates every index (RZ_ERR_INDEX_OOB) and copies the mesh into device memory
 * in structure-of-arrays form, textures as RGBA8 plus their normalisation table.
 * The descriptor's arrays are not referenced after the call returns. */
rz_status rz_mesh_create(rz_ctx* ctx, const rz_mesh_desc* desc, rz_mesh** out_mesh);
void rz_mesh_destroy(rz_ctx* ctx, rz_mesh* mesh);

/* ---- frame ------------------------------------------------------------------- */

/* Replaces draw_frame (headers/drawing.h:16-18, drawing.cpp:205-258) for a cached
 * mesh.  lights[i].trans_dir is written back (drawing.cpp:238 mutates the caller's
 * lights); any number of lights up to 2^20 (the first 64 travel in the kernel
 * arguments, the rest in a device array).  Image sides up to 65,535 pixels.  frame_buffer: uint8 [3][H][W]; depth_buffer: float [H][W]; both are read
 * (unless RZ_FLAG_CLEARED) and written in place like the reference, which does not
 * clear them.  id_buffer (may be NULL): int32 [H][W], the index of the face whose
 * fragment the pixel finally holds, or -1; a pixel the frame does not touch keeps
 * -1.  stats may be NULL. */
rz_status rz_draw_frame(rz_ctx* ctx, const rz_mesh* mesh,
                        rz_light* lights, int32_t num_lights,
                        const rz_args* args,
                        uint8_t* frame_buffer, float* depth_buffer, int32_t* id_buffer,
                        uint32_t flags, rz_frame_stats* stats);

/* No reference counterpart.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  Draws one
 * frame as rz_draw_frame does on device buffers (same kernels, same results), with every
 * working buffer of the frame (face records, tile counters and lists, bins, visibility
 * words, statistics) carved from a fixed device array of RZ_SCRATCH_BYTES that librz
 * holds on each GPU from the first rz_create on it until the process ends; nothing is
 * allocated or freed.  frame_device / depth_device / id_device (may be NULL) are device
 * buffers of the context's device (RZ_FLAG_DEVICE_BUFFERS is implied); RZ_FLAG_CLEARED
 * and RZ_FLAG_ASYNC apply.  With the context's automatic bin capacity the capacity is the
 * largest that fits the array (tiles beyond it take the overflow path, same pixels).
 * Limits: a frame whose working set does not fit RZ_SCRATCH_BYTES is rejected with
 * RZ_ERR_INVALID_ARG (the 81,920-face icosphere at 1920x1080 fits, with bins of about 1,150
 * faces; 128 bytes per face plus 8 bytes per pixel and the bins are the bulk), as are more
 * than 64 lights and RZ_FLAG_MLP_SHADE.  Every call on one GPU uses the same array and leaves
 * its contents behind (each call clears what must start at zero): calls that overlap on
 * one GPU (two contexts' streams, two host threads) race and must be serialised by the
 * caller. */
#define RZ_SCRATCH_BYTES (64u << 20)
rz_status rz_draw_frame_noalloc(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t num_lights,
                                const rz_args* args, uint8_t* frame_device, float* depth_device,
                                int32_t* id_device, uint32_t flags);

/* draw_frame with the reference's by-value semantics: uploads desc, draws one frame,
 * frees the upload (drawing.h:16-18 signature, host buffers). */
rz_status rz_draw_frame_desc(rz_ctx* ctx, const rz_mesh_desc* desc,
                             rz_light* lights, int32_t num_lights,
                             const rz_args* args,
                             uint8_t* frame_buffer, float* depth_buffer, uint32_t flags);

/* ---- frame graphs ------------------------------------------------------------- */

typedef struct rz_graph rz_graph;

/* No reference counterpart: the reference's spin loop (renderer.cpp:94-126) calls
 * draw_frame once per iteration.  Captures `frames` consecutive draw_frame calls
 * into one HIP graph, frame k rotated by k*ry_step about y (the spin loop's
 * tait_bryan_angles.y += dt), writing every frame into the same device buffers
 * (RZ_FLAG_DEVICE_BUFFERS required).  The graph refers to the mesh and to the
 * context's per-frame buffers: the mesh must outlive it, and a later frame that
 * grows those buffers (a larger image or mesh) makes rz_graph_launch refuse it. */
rz_status rz_graph_create(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t num_lights,
                          const rz_args* args, uint8_t* frame_device, float* depth_device,
                          int32_t* id_device, uint32_t flags, int32_t frames, float ry_step,
                          rz_graph** out_graph);
/* Enqueues all captured frames on the context stream (asynchronous). */
rz_status rz_graph_launch(rz_graph* graph);
void rz_graph_destroy(rz_graph* graph);

/* ---- the spin loop ---------------------------------------------------------------- */

typedef struct rz_spin rz_spin;

/* No reference counterpart as an object: the reference's spin loop (renderer.cpp:94-126)
 * calls draw_frame once per iteration with tait_bryan_angles.y advanced by the elapsed
 * time.  rz_spin_create captures that frame once (device buffers, RZ_FLAG_DEVICE_BUFFERS
 * required; light loop only); each rz_spin_frame sets the frame's rotation in the captured
 * kernels' arguments (the matrices built on the host exactly as rz_draw_frame builds them)
 * and replays it on the context stream, asynchronously, with no re-capture.  The results
 * are rz_draw_frame's for those angles, bit for bit.  The mesh and buffers must outlive
 * the spin object; a later frame that grows the context's buffers makes rz_spin_frame
 * refuse it. */
rz_status rz_spin_create(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t num_lights,
                         const rz_args* args, uint8_t* frame_device, float* depth_device, int32_t* id_device,
                         uint32_t flags, rz_spin** out_spin);
rz_status rz_spin_frame(rz_spin* spin, const float tait_bryan_angles[3]);
void rz_spin_destroy(rz_spin* spin);

/* Reference: renderer.cpp:114, frame_buffer.draw_text(5, 5, text, white, black), the spin
 * loop's frame rate.  Draws up to 32 characters of text into a planar device frame at
 * (x, y) on the context stream (asynchronous): 12x16-pixel cells, foreground where the
 * glyph is set and background over the rest of each cell, clipped to the frame.  Glyphs
 * are librz's own 5x7 bitmaps drawn at twice their size, for the characters a frame rate
 * prints (digits, '.', '-', "inf", "nan"; others draw as background), not CImg's font. */
rz_status rz_draw_text(rz_ctx* ctx, uint8_t* frame_device, uint32_t width, uint32_t height, int32_t x,
                       int32_t y, const char* text, const uint8_t foreground[3], const uint8_t background[3]);

/* ---- batches ----------------------------------------------------------------- */

/* One mesh of a batch: its own image size, transform and lights (args), and its own
 * device frame/depth/id buffers. */
typedef struct rz_batch_item {
    const rz_mesh* mesh;
    rz_light* lights;           /* trans_dir is written back, as for rz_draw_frame */
    int32_t num_lights;
    int32_t reserved;
    rz_args args;
    uint8_t* frame;             /* device, uint8 [3][H][W] */
    float* depth;               /* device, float [H][W] */
    int32_t* ids;               /* device, int32 [H][W], or NULL */
} rz_batch_item;

/* No reference counterpart (the reference draws one mesh per call).  Draws every
 * item as rz_draw_frame would (bit for bit).  By default a batch frame is five
 * launches, each running one pass of every item (vertex transform; binning and buffer
 * clears; raster-and-shade of the tiles of all items; the overflow loop of tiles whose
 * bins overflowed; their shade); with rz_set_batch_persistent it is one launch of a
 * resident grid that runs those passes as phases between grid barriers.  At most 1024
 * items.  Requires RZ_FLAG_DEVICE_BUFFERS; RZ_FLAG_CLEARED and RZ_FLAG_ASYNC apply to
 * all items.  stats (may be NULL) sums over the items; with rz_set_profiling its ms_*
 * are the passes' times (five launches: HIP events between them) or the phases' spans
 * inside the one launch (first start to last end; they overlap; ms_vertex 0, the bin
 * phase transforms the vertices), and ms_frame the whole frame. */
rz_status rz_draw_batch(rz_ctx* ctx, const rz_batch_item* items, int32_t num_items, uint32_t flags,
                        rz_frame_stats* stats);
/* 1: a batch frame is one persistent launch (every workgroup resident; a wait that
 * would exceed about a second gives up and the call returns RZ_ERR_HIP); 0 (default):
 * five launches, one per pass.  Same bits and statistics.  Read when a batch is
 * prepared (rz_draw_batch, rz_batch_graph_create, rz_autotune). */
rz_status rz_set_batch_persistent(rz_ctx* ctx, int32_t enable);
int32_t rz_get_batch_persistent(const rz_ctx* ctx);
/* `frames` repetitions of the batch captured in one HIP graph that owns its working
 * buffers and parameter tables (independent of later calls on the context); the
 * meshes and output buffers must outlive it. */
rz_status rz_batch_graph_create(rz_ctx* ctx, const rz_batch_item* items, int32_t num_items, uint32_t flags,
                                int32_t frames, rz_graph** out_graph);
/* Chooses the tile size (8, 16, 32) and bin capacity for this workload by measuring:
 * one probe frame per tile size gives the largest per-tile face count, the capacity
 * is the next power of two that holds it (within a 24 GiB bin budget), and the tile
 * size with the fastest timed frames wins.  The choice is stored in the context
 * (as rz_set_tuning would) and returned. */
rz_status rz_autotune(rz_ctx* ctx, const rz_batch_item* items, int32_t num_items, uint32_t flags,
                      int32_t* tile_size, int32_t* bin_capacity, float* ms_per_frame);

/* No reference counterpart.  Times each pass of one frame on the context's stream: a
 * HIP graph of `reps` back-to-back launches of the pass, timed by HIP events around its
 * second launch.  us[0..5] = vertex, bin, raster, overflow, shade, shade chain in
 * microseconds per launch (vertex = 0: the bin pass transforms each face's corners
 * itself; overflow = 0: the overflow loop runs inside the raster pass; shade = the
 * separate shade pass, or with the fused pipeline the overflow shade, 0 when no bin can
 * overflow).  "shade chain" is the pass that shades (raster+shade fused, else the separate
 * shade pass) run with neither the light loop nor the network: the per-pixel work every
 * shading mode shares, so the pass's time minus it is the light loop's (or with
 * RZ_FLAG_MLP_SHADE the network's) share.  Device buffers and RZ_FLAG_CLEARED required.
 * Leaves the frame rendered in frame_device/depth_device. */
rz_status rz_time_passes(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t num_lights,
                         const rz_args* args, uint8_t* frame_device, float* depth_device, uint32_t flags,
                         int32_t reps, float us[6]);

/* ---- learned shader ------------------------------------------------------------ */

/* No reference counterpart.  A 12 -> 64 -> 64 -> 3 MLP (ReLU, ReLU, linear) that
 * replaces the per-fragment light loop when RZ_FLAG_MLP_SHADE is set: BF16 operands on
 * the matrix cores, FP32 accumulation.  Inputs per fragment: interpolated normal (3),
 * unit view vector (3), albedo (3), decoded normal-map sample (3, (0,0,1) without a
 * map); output colour/255 (3), rounded to 8 bits.  Coverage and depth are unchanged.
 * Parameters (RZ_MLP_PARAMS floats): W1[64][12], b1[64], W2[64][64], b2[64], W3[3][64],
 * b3[3].  The network belongs to the context. */
#define RZ_MLP_PARAMS 5187
#define RZ_MLP_SAMPLE_FLOATS 15   /* training sample: 12 inputs, 3 targets (colour/255) */

rz_status rz_mlp_init(rz_ctx* ctx, uint64_t seed);                     /* He-uniform, zero biases */
rz_status rz_mlp_set_weights(rz_ctx* ctx, const float* params);
rz_status rz_mlp_get_weights(rz_ctx* ctx, float* params);
/* Renders one frame with the plain shader and appends every shaded fragment's inputs
 * and colour/255 to samples_device (capacity samples); *count is the fill level before
 * (in) and after (out). */
rz_status rz_mlp_collect(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t num_lights,
                         const rz_args* args, float* samples_device, int64_t capacity, int64_t* count);
/* `steps` Adam steps of `batch` samples (a multiple of 64) drawn at random from the n
 * device samples; learning rate lr with cosine decay to lr/100.  losses (may be NULL)
 * receives the mean squared error of every step. */
rz_status rz_mlp_train(rz_ctx* ctx, const float* samples_device, int64_t n, int32_t steps, int32_t batch,
                       float lr, uint64_t seed, float* losses);
/* The network on n input rows (12 floats each) in device memory -> n rows of 3. */
rz_status rz_mlp_infer(rz_ctx* ctx, const float* features_device, int64_t n, float* out_device);

/* ---- device memory ------------------------------------------------------------ */

/* Plain device allocations on the context's device, for callers that keep frame and
 * depth buffers in HBM (RZ_FLAG_DEVICE_BUFFERS) without an allocator of their own.
 * The copies run on the context stream and return when complete. */
rz_status rz_device_alloc(rz_ctx* ctx, size_t bytes, void** out_ptr);
rz_status rz_device_free(rz_ctx* ctx, void* ptr);
rz_status rz_copy_to_device(rz_ctx* ctx, void* dst_device, const void* src_host, size_t bytes);
rz_status rz_copy_to_host(rz_ctx* ctx, void* dst_host, const void* src_device, size_t bytes);
/* Device-to-device copy on the context's stream by a plain streaming kernel (16-byte
 * words; any unaligned remainder by hipMemcpy); waits for it. */
rz_status rz_copy_device(rz_ctx* ctx, void* dst_device, const void* src_device, size_t bytes);

/* ---- self-check ------------------------------------------------------------------- */

/* No counterpart in the reference.  The raster pass computes the barycentric quotients
 * e / area of most candidates from a per-face reciprocal: the compiler's correctly
 * rounded division with its scaling and fix-up steps left out, which are identities for
 * |area| in [2^-40, 2^40] and |e| in [2^-20, 2^40].  This runs that computation against
 * the IEEE division on about `samples` random operand pairs (every sign, exponents one
 * beyond each end of the ranges, mantissas random, all ones or zero) on the device and
 * reports the pairs in range compared and how many differed in any bit (0 expected).
 * Synchronous; allocates 16 bytes of device memory for the call. */
rz_status rz_check_quotients(rz_ctx* ctx, uint64_t samples, uint32_t seed, uint64_t* compared,
                             uint64_t* mismatches);

/* ---- host helpers -------------------------------------------------------------- */

/* Reference defaults of Args (arguments.cpp:20-31) for a width x height image, plus
 * the projection of geometry.cpp:30. */
void rz_default_args(rz_args* args, uint32_t width, uint32_t height);
/* The matrices rz_draw_frame builds for args (geometry.cpp:22-33, drawing.cpp:222-229,
 * geometry.cpp:101). */
rz_status rz_frame_matrices(const rz_args* args, rz_matrices* out);

#ifdef __cplusplus
}
#endif

#endif /* RZ_H */
