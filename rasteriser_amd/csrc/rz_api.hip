// rz_api.hip — the C ABI of librz (include/rz.h): contexts, the device mesh cache
// and the per-frame pass sequence vertex -> bin -> raster -> overflow -> shade.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rz.h"
#include "rz_device.h"
#include "rz_host_math.h"
#include "rz_mlp.h"

namespace rz {
hipError_t launch_bin(const BinParams& p, hipStream_t s);
hipError_t launch_raster(const RasterParams& p, int tile_size, hipStream_t s);
hipError_t launch_raster_shade(const RasterParams& p, const ShadeParams& sp, int tile_size, hipStream_t s);
hipError_t launch_shade_tiles(const RasterParams& p, const ShadeParams& sp, int tile_size, hipStream_t s);
hipError_t launch_shade(const ShadeParams& p, hipStream_t s);
hipError_t launch_vertex_batch(const VertexParams* items, const uint2* map, uint32_t blocks, hipStream_t s);
hipError_t launch_bin_batch(const BinParams* items, const uint2* map, uint32_t blocks, hipStream_t s);
hipError_t launch_overflow_batch(const RasterParams* items, const uint2* map, uint32_t blocks, hipStream_t s);
uint32_t raster_batch_blocks_per_cu(int tile_size, int mode);
uint32_t batch_frame_blocks_per_cu(int tile_size, int mode);
hipError_t launch_batch_frame(const BinParams* b, const RasterParams* r, const RasterParams* o,
                              const ShadeParams* sh, const BatchLists& bl0, const BatchLists& bl1,
                              const BatchSync& q, int tile_size, int mode, hipStream_t s);
hipError_t launch_batch_tiles(const RasterParams* items, const ShadeParams* sitems, const BatchLists& bl,
                              uint32_t raster_grid, uint32_t tiles_grid, int tile_size, int mode, hipStream_t s,
                              bool raster);
hipError_t launch_hold(uint64_t ticks, hipStream_t s);
hipError_t launch_check_quotients(uint32_t blocks, uint32_t per_thread, uint32_t seed, unsigned long long* counts,
                                  hipStream_t s);
hipError_t launch_text(uint8_t* frame, uint32_t W, uint32_t H, int32_t x0, int32_t y0, const char* text,
                       const uint8_t fg[3], const uint8_t bg[3], hipStream_t s);
uint32_t raster_blocks_per_cu(int tile_size, bool fused, int mode, bool narrow, bool depth_in);
hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s);
hipError_t launch_mlp_pack(const float* W, uint16_t* packed, hipStream_t s);
hipError_t launch_mlp_infer(const float* x, int64_t n, const uint16_t* packed, float* y, hipStream_t s);
hipError_t launch_mlp_step(const float* samples, int64_t n, float* W, float* m, float* v, float* gpart,
                           float* lpart, float* loss_out, int nblocks, uint64_t seed, uint32_t step, float lr,
                           float c1, float c2, hipStream_t s);
}  // namespace rz

using namespace rz;

namespace {

// bbox corners are packed in 16 bits each (largest corner W - 1, H - 1)
constexpr uint32_t kMaxImageSide = 65535;
// Lights per frame: the first kMaxLights travel in the kernel arguments, the rest in a
// device array (the reference takes any number, shading.cpp:24-29).
constexpr int32_t kMaxFrameLights = 1 << 20;
// Workgroups per CU of the overflow shade launch (persistent: it has work only when a
// bin overflowed).
constexpr int kPersistentPerCU = 4;

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct rz_mesh {
    int device = 0;
    int32_t nv = 0, nn = 0, nt = 0, nf = 0, nm = 0;
    bool raster_space = false;
    float *px = nullptr, *py = nullptr, *pz = nullptr;
    float4* rin = nullptr;
    // (x, y, z, 0) copies for the single-frame bin pass, which transforms each face's
    // corners itself: positions, vertex normals, face normals
    float4 *pos4 = nullptr, *nrm4 = nullptr, *fnrm4 = nullptr;
    float *nx = nullptr, *ny = nullptr, *nz = nullptr;      // vertex normals
    float *fnx = nullptr, *fny = nullptr, *fnz = nullptr;   // face normals (flat)
    float *ftx = nullptr, *fty = nullptr, *ftz = nullptr;   // face tangents (normal maps)
    float2* uvs = nullptr;
    // the faces' indices in binning order: sorted by the Morton code of their centroid
    // (the bin pass's workgroups then cover compact screen regions whatever the input
    // order; visibility does not depend on it)
    int4* bin_face_v = nullptr;       // (v0, v1, v2, face)
    int4* bin_face_attr = nullptr;    // (n0, n1, n2, material), (t0, t1, t2, 0)
    // every face has the same normal, uv (-1: none) and material: shading takes them
    // from here and the bin pass gathers nothing (uniform = 1)
    int32_t uniform = 0, uniform_n = 0, uniform_t = -1, uniform_m = 0;
    float uniform_nrm[3] = {0.f, 0.f, 0.f};   // model-space normal uniform_n
    DevMaterial* mats = nullptr;
    uint32_t* texels = nullptr;
    float* luts = nullptr;
    std::vector<void*> allocs;
};

struct rz_graph {
    rz_ctx* ctx = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t frames = 0;
    uint64_t buffer_gen = 0;   // ctx->buffer_gen when captured (graphs over the context's buffers)
    int32_t ntiles = 0;        // frame graphs: tiles of the captured frames (counter-line layout)
    bool owns_buffers = false; // batch graphs carry their own working buffers and tables
    DevBuf arena, tables;
    DevBuf lights_ext;         // frame graphs with more lights than the kernel arguments hold
};

struct rz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int tile_size = 16;
    int bin_cap = 0;   // 0: automatic (frame_bin_cap)
    bool fused = true;
    bool batch_persistent = false;  // a batch frame is one launch (k_batch_frame), else five
    bool profiling = false;
    uint32_t persistent_blocks = 0;   // workgroups of the overflow shade launch (a few per CU)
    int cus = 0;
    size_t auto_bin_bytes = 0;        // automatic bin capacity: the most bytes of bins (rz_create)
    hipEvent_t ev[6] = {};
    std::string err;
    DevBuf raster, cnormal, rec, tile_count, bins, vis, lists, frame, depth, ids, counters, partials;
    DevBuf batch_arena, batch_tables;
    // learned shader: FP32 parameters, Adam moments, packed BF16 network, training scratch
    DevBuf mlp_w, mlp_m, mlp_v, mlp_packed, mlp_gpart, mlp_lpart, mlp_loss, mlp_count;
    DevBuf lights_ext;     // eager frames' lights beyond the kernel arguments
    bool mlp_ready = false;
    uint32_t mlp_steps = 0;
    std::vector<uint32_t> host_partials;
    unsigned long long* host_counters = nullptr;   // pinned
    uint64_t buffer_gen = 0;                       // bumped whenever a per-frame buffer is reallocated
    // Single frames' counter lines: generation count_gen is all zero for a layout of
    // count_ntiles tiles in the tile_count buffer at count_ptr (anything else: unknown)
    const void* count_ptr = nullptr;
    int32_t count_ntiles = -1;
    int count_gen = 0;
    // this device's copy of `scratch` (rz_draw_frame_noalloc's working memory)
    unsigned char* scratch = nullptr;
};

namespace {

rz_status fail(rz_ctx* ctx, rz_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    fprintf(stderr, "librz: %s\n", buf);
    return st;
}

#define RZ_HIP(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), e_ == hipErrorOutOfMemory ? RZ_ERR_OOM : RZ_ERR_HIP, "%s: %s (%s:%d)", \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);                         \
    } while (0)

// Grows a working buffer to `bytes`.  stales_graphs: frame graphs captured over the old
// allocation refuse to launch afterwards (ctx->buffer_gen).
rz_status ensure(rz_ctx* ctx, DevBuf& b, size_t bytes, bool stales_graphs = true) {
    if (bytes == 0) bytes = 256;
    if (b.bytes >= bytes) return RZ_OK;
    if (b.ptr) {
        RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
        RZ_HIP(ctx, hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    RZ_HIP(ctx, hipMalloc(&b.ptr, bytes));
    b.bytes = bytes;
    if (stales_graphs) ++ctx->buffer_gen;
    if (&b == &ctx->tile_count) ctx->count_ntiles = -1;   // new memory: contents unknown
    return RZ_OK;
}

// rz_draw_frame_noalloc carves every working buffer of its frame from this array: one
// copy per GPU, held from the library's first context on that GPU until the process
// ends.  Host code never uses the name itself (on the host it is only a stand-in for
// the symbol): each device's address comes from hipGetSymbolAddress once, at the first
// rz_create on that device (scratch_address), and only that address reaches kernels.
static __device__ __align__(256) unsigned char scratch[RZ_SCRATCH_BYTES];

std::mutex g_scratch_mutex;
std::vector<unsigned char*> g_scratch_ptr;   // per device, nullptr until looked up

// This device's address of `scratch` (the device must be current).
rz_status scratch_address(rz_ctx* ctx) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= ctx->device) g_scratch_ptr.resize((size_t)ctx->device + 1, nullptr);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ctx, RZ_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
                        hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = static_cast<unsigned char*>(p);
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    return RZ_OK;
}

int log2_exact(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return (1 << l) == v ? l : -1;
}

template <typename T>
rz_status upload(rz_ctx* ctx, rz_mesh* m, T** dst, const void* src, size_t bytes) {
    *dst = nullptr;
    if (bytes == 0) bytes = 16;
    void* p = nullptr;
    RZ_HIP(ctx, hipMalloc(&p, bytes));
    m->allocs.push_back(p);
    if (src) RZ_HIP(ctx, hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *dst = static_cast<T*>(p);
    return RZ_OK;
}

rz_status validate_mesh(rz_ctx* ctx, const rz_mesh_desc* d) {
    if (!d) return fail(ctx, RZ_ERR_INVALID_ARG, "mesh descriptor is null");
    if (d->num_vertices < 0 || d->num_normals < 0 || d->num_uvs < 0 || d->num_faces < 0 || d->num_materials < 0)
        return fail(ctx, RZ_ERR_INVALID_ARG, "negative mesh array size");
    if ((d->num_vertices && !d->vertices) || (d->num_normals && !d->normals) || (d->num_uvs && !d->uvs) ||
        (d->num_faces && !d->faces) || (d->num_materials && !d->materials))
        return fail(ctx, RZ_ERR_INVALID_ARG, "mesh array pointer is null");
    for (int32_t f = 0; f < d->num_faces; ++f) {
        const rz_triangle& t = d->faces[f];
        for (int k = 0; k < 3; ++k) {
            if (t.vertices[k] < 0 || t.vertices[k] >= d->num_vertices)
                return fail(ctx, RZ_ERR_INDEX_OOB, "face %d: vertex index %d out of [0,%d)", f, t.vertices[k],
                            d->num_vertices);
            if (t.normals[k] < 0 || t.normals[k] >= d->num_normals)
                return fail(ctx, RZ_ERR_INDEX_OOB, "face %d: normal index %d out of [0,%d)", f, t.normals[k],
                            d->num_normals);
            if (t.uvs[k] >= d->num_uvs)
                return fail(ctx, RZ_ERR_INDEX_OOB, "face %d: uv index %d out of [0,%d)", f, t.uvs[k], d->num_uvs);
        }
        if (t.material < 0 || t.material >= d->num_materials)
            return fail(ctx, RZ_ERR_INDEX_OOB, "face %d: material %d out of [0,%d)", f, t.material,
                        d->num_materials);
    }
    for (int32_t i = 0; i < d->num_materials; ++i) {
        const rz_material& m = d->materials[i];
        if (m.normal_map_channels != 0 &&
            ((m.normal_map_channels != 3 && m.normal_map_channels != 4) || !m.normal_map ||
             m.normal_map_width <= 0 || m.normal_map_height <= 0))
            return fail(ctx, RZ_ERR_INVALID_ARG, "material %d: normal map needs 3 or 4 channels and a size", i);
        if (m.normal_map_channels != 0 && d->raster_space)
            return fail(ctx, RZ_ERR_INVALID_ARG, "material %d: normal maps need model-space positions", i);
        if (m.tex_channels == 0) continue;
        if ((m.tex_channels != 3 && m.tex_channels != 4) || !m.texels || m.tex_width <= 0 || m.tex_height <= 0)
            return fail(ctx, RZ_ERR_INVALID_ARG, "material %d: texture needs 3 or 4 channels and a size", i);
    }
    return RZ_OK;
}


// Sizes and statistics sections of one frame's working buffers.
struct FrameLayout {
    int ts = 16, tsl = 4, tiles_x = 0, tiles_y = 0, ntiles = 0;
    bool fused = false;     // raster pass shades non-overflowed tiles
    bool batch = false;     // one mesh of a batch (tile lists and raster statistics shared)
    // two generations of counter lines and no vertex pass (single frames, and the meshes
    // of a batch's persistent launch): the bin pass transforms each face's corners and
    // zeroes the other generation
    bool twogen = true;
    uint32_t raster_grid = 0, tiles_grid = 0;   // workgroups of the list-driven launches
    uint32_t W = 0, H = 0, cap = 0;
    size_t np = 0;
    int32_t nfn = 0;
    // per-workgroup statistics sections (uint32 each)
    size_t off_front = 0, off_pairs = 0, off_frag = 0, off_ovf = 0, off_cov = 0, n_part = 0;
    size_t bytes[9] = {};   // raster, cnormal, rec, tile_count, bins, vis, lists, counters, partials
};

enum { kBufRaster, kBufCnormal, kBufRec, kBufTiles, kBufBins, kBufVis, kBufLists, kBufCounters, kBufPart, kNumBufs };

// Single frame: raster_grid / tiles_grid workgroups of the list-driven raster and
// overflow-shade launches.  Batch meshes (batch = true) are always fused; the batch owns
// the tile lists and the statistics of those two launches, and each mesh keeps the
// partial sums of its overflow loop (k_overflow_batch).
FrameLayout layout_frame(const rz_mesh* mesh, const rz_args& a, int ts, int cap, bool fused, uint32_t raster_grid,
                         uint32_t tiles_grid, bool batch, int bblock = 256) {
    FrameLayout L;
    L.ts = ts;
    L.fused = fused || batch;
    L.batch = batch;
    L.twogen = !batch;
    L.raster_grid = raster_grid;
    L.tiles_grid = tiles_grid;
    L.tsl = log2_exact(ts);
    L.W = a.image_width;
    L.H = a.image_height;
    L.cap = (uint32_t)cap;
    L.tiles_x = (int)((L.W + ts - 1) / ts);
    L.tiles_y = (int)((L.H + ts - 1) / ts);
    L.ntiles = L.tiles_x * L.tiles_y;
    L.np = (size_t)L.W * L.H;
    L.nfn = a.flat ? mesh->nf : mesh->nn;
    L.off_front = 0;
    const size_t nbin = batch ? bin_blocks(mesh->nf, bblock) : frame_bin_blocks(mesh->nf);   // bin workgroups
    L.off_pairs = L.off_front + nbin;
    L.off_frag = L.off_pairs + nbin;
    L.off_ovf = L.off_frag + (batch ? 0 : (size_t)raster_grid);
    L.off_cov = L.off_ovf + (batch ? overflow_blocks(mesh->nf, bblock) : 0);   // single frames: in the raster pass
    L.n_part = L.off_cov + (batch ? 0
                                  : L.fused ? (size_t)raster_grid + tiles_grid
                                            : (size_t)shade_blocks_x(L.W) * shade_blocks_y(L.H));
    // batches: the vertex pass's outputs (raster space: read as given), one generation
    // of tile counts + the 8 list counters, and the global counters apart.  Single frames
    // transform in the bin pass and keep two generations of counter lines (count_lines):
    // each frame's bin pass zeroes the other one for the next frame.
    L.bytes[kBufRaster] = batch && !mesh->raster_space ? (size_t)mesh->nv * 16 : 0;
    L.bytes[kBufCnormal] = batch ? (size_t)L.nfn * 16 : 0;
    // the face records, then the bin pass's front mask (a bit per face)
    L.bytes[kBufRec] = (size_t)mesh->nf * sizeof(FaceRecord) + ((size_t)mesh->nf + 63) / 64 * 8;
    L.bytes[kBufTiles] = batch ? ((size_t)L.ntiles + 8) * 4 * kCountStride
                               : (size_t)2 * count_lines(L.ntiles) * 4 * kCountStride;
    L.bytes[kBufLists] = batch ? 0 : (size_t)9 * L.ntiles * 4;   // 8 lists, then the overflowed tiles
    L.bytes[kBufBins] = (size_t)L.ntiles * L.cap * 4;
    L.bytes[kBufVis] = L.np * 8;
    L.bytes[kBufCounters] = batch ? kNumCounters * 8 : 0;
    L.bytes[kBufPart] = L.n_part * 4;
    return L;
}

// One mesh of a batch's persistent launch: a single frame's buffers (two generations of
// counter lines, the bin pass transforming each face's corners) with the batch's shared
// tile lists and raster statistics; its bin and overflow work in workgroups of bblock.
FrameLayout layout_batch_sync(const rz_mesh* mesh, const rz_args& a, int ts, int cap, int bblock) {
    FrameLayout L = layout_frame(mesh, a, ts, cap, true, 0, 0, false);
    L.batch = true;
    L.twogen = true;
    const size_t nbin = bin_blocks(mesh->nf, bblock);
    L.off_front = 0;
    L.off_pairs = nbin;
    L.off_frag = 2 * nbin;
    L.off_ovf = L.off_frag;   // (raster statistics: the batch's, per workgroup)
    L.off_cov = L.off_ovf + overflow_blocks(mesh->nf, bblock);
    L.n_part = L.off_cov;
    L.bytes[kBufLists] = 0;   // the batch's lists
    L.bytes[kBufPart] = L.n_part * 4;
    return L;
}

// One frame's kernel parameters, built on the host before anything is enqueued.
struct FrameLaunch {
    VertexParams vp;
    BinParams bp;
    RasterParams rp;
    RasterParams orp;   // the overflow pass: rp with its own statistics section
    ShadeParams sp;
    FrameLayout lay;
    uint32_t* part = nullptr;
    unsigned long long* counters = nullptr;
    int gen = 0;               // single frames: the generation of counter lines used
    // lights beyond the kernel arguments (six floats each, ShadeParams::lights_ext), for
    // the caller to place in device memory
    std::vector<float> ext_lights;
};

rz_status check_frame(rz_ctx* ctx, const rz_mesh* mesh, const rz_light* lights, int32_t nl, const rz_args* args) {
    if (!mesh || !args) return fail(ctx, RZ_ERR_INVALID_ARG, "null mesh or args");
    if (nl < 0 || nl > kMaxFrameLights || (nl > 0 && !lights))
        return fail(ctx, RZ_ERR_INVALID_ARG, "light count %d not in [0,%d]", nl, kMaxFrameLights);
    const uint32_t W = args->image_width, H = args->image_height;
    if (W == 0 || H == 0 || W > kMaxImageSide || H > kMaxImageSide)
        return fail(ctx, RZ_ERR_INVALID_ARG, "image %ux%u not in [1,%u]^2", W, H, kMaxImageSide);
    if (args->flat && mesh->raster_space)
        return fail(ctx, RZ_ERR_INVALID_ARG, "flat shading needs model-space positions");
    if (mesh->device != ctx->device) return fail(ctx, RZ_ERR_INVALID_ARG, "mesh lives on another device");
    return RZ_OK;
}

// Builds matrices and light directions (written back to `lights`, as drawing.cpp:238
// does) and fills every pass's parameters over the working buffers `buf`.
// Single frames use generation `gen` of the counter lines (and zero the other one).
void fill_frame(const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args& a, uint32_t flags,
                const FrameLayout& lay, void* const buf[kNumBufs], FrameLaunch& L, int gen = 0) {
    const bool cleared = (flags & RZ_FLAG_CLEARED) != 0;
    rz_matrices mats;
    build_matrices(a, mats);
    L.lay = lay;
    std::memset(&L.sp, 0, sizeof L.sp);
    L.ext_lights.assign(nl > kMaxLights ? 6 * (size_t)(nl - kMaxLights) : 0, 0.0f);
    for (int32_t l = 0; l < nl; ++l) {
        light_direction(mats.view, lights[l].direction, lights[l].trans_dir);   // drawing.cpp:238
        float* ic = l < kMaxLights ? L.sp.light_ic[l] : &L.ext_lights[6 * (size_t)(l - kMaxLights)];
        float* nd = l < kMaxLights ? L.sp.light_nd[l] : ic + 3;
        for (int c = 0; c < 3; ++c) {
            ic[c] = lights[l].intensity * lights[l].colour[c];
            nd[c] = -lights[l].trans_dir[c];
        }
    }
    const size_t gen_words = (size_t)count_lines(lay.ntiles) * kCountStride;
    uint32_t* tile_count = (uint32_t*)buf[kBufTiles] + (lay.twogen ? gen * gen_words : 0);
    auto* counters = !lay.twogen ? (unsigned long long*)buf[kBufCounters]
                               : (unsigned long long*)(tile_count + ((size_t)lay.ntiles + 8) * kCountStride);
    auto* part = (uint32_t*)buf[kBufPart];
    L.part = part;
    L.counters = counters;
    L.gen = gen;

    VertexParams& vp = L.vp;
    std::memset(&vp, 0, sizeof vp);
    std::memcpy(vp.camera, mats.camera, 64);
    std::memcpy(vp.normal, mats.normal, 64);
    vp.px = mesh->px; vp.py = mesh->py; vp.pz = mesh->pz;
    vp.nx = a.flat ? mesh->fnx : mesh->nx;
    vp.ny = a.flat ? mesh->fny : mesh->ny;
    vp.nz = a.flat ? mesh->fnz : mesh->nz;
    vp.raster = (float4*)buf[kBufRaster];
    vp.cnormal = (float4*)buf[kBufCnormal];
    vp.tile_count = tile_count;
    vp.counters = counters;
    // (batches only: single frames transform in the bin pass)
    vp.nv = mesh->nv; vp.nn = lay.nfn; vp.ntiles = lay.ntiles + 8;   // clears the list counters too
    // raster-space vertices need no vertex stage: the bin pass reads them where they are
    if (mesh->raster_space) {
        vp.nv = 0;
        vp.raster = mesh->rin;
    }
    vp.Wf = (float)lay.W; vp.Hf = (float)lay.H;

    BinParams& bp = L.bp;
    std::memset(&bp, 0, sizeof bp);
    bp.face_v = mesh->bin_face_v;
    bp.face_attr = mesh->bin_face_attr;
    bp.raster = vp.raster;
    bp.cnormal = vp.cnormal;
    bp.uvs = mesh->uvs;
    bp.flat = a.flat != 0;
    const bool uniform = mesh->uniform && !a.flat;
    bp.gather = uniform ? 0 : 1;
    bp.rec = (FaceRecord*)buf[kBufRec];
    bp.front_mask = (unsigned long long*)(bp.rec + mesh->nf);
    bp.tile_count = vp.tile_count;
    bp.bins = (uint32_t*)buf[kBufBins];
    bp.counters = counters;
    bp.part_front = part + lay.off_front;
    bp.part_pairs = part + lay.off_pairs;
    bp.nf = mesh->nf; bp.W = lay.W; bp.H = lay.H;
    bp.tiles_x = lay.tiles_x; bp.ts_log2 = lay.tsl; bp.cap = lay.cap;
    bp.wind_cw = a.wind_clockwise != 0;
    bp.ntiles = lay.ntiles;
    bp.lists = (uint32_t*)buf[kBufLists];   // (batches: replaced by the batch's lists)
    bp.ovf_list = bp.lists ? bp.lists + (size_t)8 * lay.ntiles : nullptr;
    bp.list_count = bp.tile_count + (size_t)lay.ntiles * kCountStride;
    bp.ovf_vis = (unsigned long long*)buf[kBufVis];
    if (lay.twogen) {
        // no vertex pass: the bin pass transforms each face's corners
        bp.pos = mesh->raster_space ? nullptr : mesh->pos4;
        bp.raster = mesh->raster_space ? mesh->rin : nullptr;
        bp.nrm = a.flat ? mesh->fnrm4 : mesh->nrm4;
        bp.cnormal = nullptr;
        std::memcpy(bp.camera, mats.camera, 64);
        std::memcpy(bp.normal, mats.normal, 64);
        bp.Wf = (float)lay.W; bp.Hf = (float)lay.H;
        bp.zero_lines = (uint32_t*)buf[kBufTiles] + (1 - gen) * gen_words;
        bp.zero_count = count_lines(lay.ntiles);
        bp.zero_blocks = (bp.zero_count + kFrameBinBlock - 1) / kFrameBinBlock;
    }

    RasterParams& rp = L.rp;
    std::memset(&rp, 0, sizeof rp);
    rp.rec = bp.rec;
    rp.tile_count = bp.tile_count;
    rp.bins = bp.bins;
    rp.vis = (unsigned long long*)buf[kBufVis];
    rp.counters = counters;
    rp.part_frag = part + lay.off_frag;
    rp.nf = mesh->nf; rp.W = lay.W; rp.H = lay.H;
    rp.tiles_x = lay.tiles_x; rp.ntiles = lay.ntiles; rp.ts_log2 = lay.tsl; rp.cap = lay.cap;
    rp.lists = bp.lists;
    rp.list_count = bp.list_count;
    rp.ovf_list = bp.ovf_list;
    rp.front_mask = bp.front_mask;
    rp.face_v = mesh->bin_face_v;
    rp.max_blocks = lay.raster_grid;
    L.orp = rp;   // batches: k_overflow_batch
    L.orp.part_frag = part + lay.off_ovf;

    ShadeParams& sp = L.sp;
    sp.vis = rp.vis;
    sp.rec = bp.rec;
    sp.mats = mesh->mats; sp.texels = mesh->texels; sp.luts = mesh->luts;
    sp.uniform = uniform ? 1 : -1;
    sp.uniform_n = mesh->uniform_n;
    sp.uniform_t = mesh->uniform_t;
    sp.uniform_m = mesh->uniform_m;
    {
        // the uniform normal as the vertex transform gives it (same operations, in order)
        const float* m = mats.normal;
        const float x = mesh->uniform_nrm[0], y = mesh->uniform_nrm[1], z = mesh->uniform_nrm[2], w = 0.0f;
        sp.uniform_cn = make_float4((m[0] * x + m[4] * y) + (m[8] * z + m[12] * w),
                                    (m[1] * x + m[5] * y) + (m[9] * z + m[13] * w),
                                    (m[2] * x + m[6] * y) + (m[10] * z + m[14] * w), 0.0f);
    }
    sp.uvs = mesh->uvs;
    sp.ftx = mesh->ftx; sp.fty = mesh->fty; sp.ftz = mesh->ftz;
    std::memcpy(sp.normal_mat, mats.normal, 64);
    sp.part_cov = part + lay.off_cov;
    sp.W = lay.W; sp.H = lay.H;
    sp.cleared = cleared;
    sp.wind_cw = a.wind_clockwise != 0;
    sp.flat = a.flat != 0;
    sp.nl = nl;
    sp.nmat = mesh->nm;
    sp.mode = (flags & RZ_FLAG_MLP_SHADE) ? kShadeMlp : kShadePlain;
    const float th = std::tan(a.fovy_degrees * 0.01745329251994329576923690768489f / 2.0f);
    sp.inv_p00 = a.aspect_ratio * th;
    sp.inv_p11 = th;
}

// Frames with at least this many 16-pixel tiles raster them with 64-thread workgroups.
constexpr uint32_t kNarrowTiles = 16384;

// Bin capacity of the automatic setting when a bin per tile for every face would not fit
// the context's auto_bin_bytes (and of batches, which the autotuner sizes).
constexpr int kDefaultBinCap = 512;
// auto_bin_bytes: 4 GiB, or an eighth of the device memory free at rz_create if less.
constexpr size_t kAutoBinBytesMax = (size_t)4 << 30;

// A single frame's per-tile bin capacity: the context's, or (automatic) room in every
// tile for every face of the mesh when that fits auto_bin_bytes — bins are written only
// where faces land, so their size costs device memory (hipMalloc commits it), not
// traffic.  With a capacity of at least the face count no bin can overflow and the frame
// needs no overflow-shade launch.
int frame_bin_cap(const rz_ctx* ctx, const rz_mesh* mesh, uint32_t ntiles) {
    if (ctx->bin_cap > 0) return ctx->bin_cap;
    const size_t nf = mesh->nf > 0 ? (size_t)mesh->nf : 1;
    if (nf <= ((size_t)1 << 20) && (size_t)ntiles * nf * 4 <= ctx->auto_bin_bytes) return (int)nf;
    return kDefaultBinCap;
}

// Launch shape of a single frame's raster pass on this context: 64-thread workgroups on
// 16-pixel tiles for frames of many tiles, and the resident grid (a multiple of 8).
// 16-pixel tiles: 256-thread workgroups finish one tile soonest, and at 1080p every
// covered tile is in flight at once (1,490 tiles, 1,536 workgroups); with many more
// tiles than resident workgroups, 64-thread workgroups keep four times as many tiles
// in flight (1080p: 23 against 63 us; 4M triangles at 4096x4096: 326 against 254 us)
// flags: the frame's; without RZ_FLAG_CLEARED the raster pass stages the input depths
// in dynamic LDS, which the residency count includes.
void frame_shape(const rz_ctx* ctx, const rz_args& a, int mode, uint32_t flags, uint32_t& ntiles, bool& narrow,
                 uint32_t& grid) {
    ntiles = ((a.image_width + ctx->tile_size - 1) / ctx->tile_size) *
             ((a.image_height + ctx->tile_size - 1) / ctx->tile_size);
    narrow = ctx->tile_size == 16 && ntiles >= kNarrowTiles;
    const uint32_t per_cu =
        raster_blocks_per_cu(ctx->tile_size, ctx->fused, mode, narrow, (flags & RZ_FLAG_CLEARED) == 0);
    grid = (uint32_t)ctx->cus * per_cu / 8u * 8u;
    if (grid < 8) grid = 8;
}

// The lights beyond the kernel arguments into `buf` (grown as needed; the stream is
// drained first, since queued frames may still read the old contents).
rz_status stage_ext_lights(rz_ctx* ctx, DevBuf& buf, FrameLaunch& L, bool stales_graphs) {
    L.sp.lights_ext = nullptr;
    if (L.ext_lights.empty()) return RZ_OK;
    const size_t bytes = L.ext_lights.size() * sizeof(float);
    rz_status st = ensure(ctx, buf, bytes, stales_graphs);
    if (st != RZ_OK) return st;
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipMemcpy(buf.ptr, L.ext_lights.data(), bytes, hipMemcpyHostToDevice));
    L.sp.lights_ext = (const float*)buf.ptr;
    return RZ_OK;
}

// Single frame: validates, sizes the context's working buffers and fills L.  mode: the
// shading mode (-1: the one flags select).
rz_status prepare_frame(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args* args,
                        uint32_t flags, FrameLaunch& L, int gen, int mode = -1) {
    rz_status st = check_frame(ctx, mesh, lights, nl, args);
    if (st != RZ_OK) return st;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    if (mode < 0) mode = (flags & RZ_FLAG_MLP_SHADE) ? kShadeMlp : kShadePlain;
    uint32_t ntiles = 0, grid = 0;
    bool narrow = false;
    frame_shape(ctx, *args, mode, flags, ntiles, narrow, grid);
    const int cap = frame_bin_cap(ctx, mesh, ntiles);
    const bool can_overflow = (uint32_t)cap < (uint32_t)mesh->nf;
    const FrameLayout lay = layout_frame(mesh, *args, ctx->tile_size, cap, ctx->fused, grid,
                                         can_overflow ? ctx->persistent_blocks : 0u, false);
    DevBuf* bufs[kNumBufs] = {&ctx->raster, &ctx->cnormal, &ctx->rec,   &ctx->tile_count, &ctx->bins,
                              &ctx->vis,    &ctx->lists,   &ctx->counters, &ctx->partials};
    void* ptrs[kNumBufs];
    for (int i = 0; i < kNumBufs; ++i) {
        if ((st = ensure(ctx, *bufs[i], lay.bytes[i])) != RZ_OK) return st;
    }
    for (int i = 0; i < kNumBufs; ++i) ptrs[i] = bufs[i]->ptr;
    if ((flags & RZ_FLAG_MLP_SHADE) && !ctx->mlp_ready)
        return fail(ctx, RZ_ERR_INVALID_ARG, "RZ_FLAG_MLP_SHADE without network weights (rz_mlp_init / rz_mlp_set_weights)");
    fill_frame(mesh, lights, nl, *args, flags, lay, ptrs, L, gen);
    if ((st = stage_ext_lights(ctx, ctx->lights_ext, L, false)) != RZ_OK) return st;
    L.rp.narrow = narrow ? 1 : 0;
    L.sp.mode = mode;
    L.sp.mlp = (const uint16_t*)ctx->mlp_packed.ptr;
    L.sp.max_blocks = lay.tiles_grid;
    return RZ_OK;
}

// The overflow shade's parameters: its coverage partials follow the raster pass's.
ShadeParams tiles_params(const FrameLaunch& L) {
    ShadeParams t = L.sp;
    t.part_cov += L.lay.raster_grid;
    return t;
}

// Workgroups of the bin launch that clear buffers: single frames, and per pixel for a
// batch's meshes.
constexpr uint32_t kClearBlocks = 1024;
constexpr uint32_t kClearPixelsPerBlock = 8192;

void bind_outputs(FrameLaunch& L, bool cleared, uint8_t* frame_d, float* depth_d, int32_t* ids_d) {
    L.rp.depth_in = cleared ? nullptr : depth_d;
    L.orp.depth_in = L.rp.depth_in;
    L.sp.frame = frame_d;
    L.sp.depth = depth_d;
    L.sp.ids = ids_d;
    // fused pipeline on cleared buffers: the bin launch clears the outputs, so tiles
    // with nothing binned need no raster workgroup work at all
    const bool pre = L.lay.fused && cleared;
    // the list-driven raster visits listed tiles only: with the fused pipeline the face
    // ids of the others must still read -1, with the separate-shade pipeline their
    // visibility words must be empty
    const bool ids = L.lay.fused && ids_d;
    const bool vis = !L.lay.fused;
    L.bp.clear_frame = pre ? frame_d : nullptr;
    L.bp.clear_depth = pre ? depth_d : nullptr;
    L.bp.clear_ids = ids ? ids_d : nullptr;
    L.bp.clear_vis = vis ? L.rp.vis : nullptr;
    L.bp.clear_np = (pre || ids || vis) ? L.lay.np : 0;
    const size_t per = (L.lay.np + kClearPixelsPerBlock - 1) / kClearPixelsPerBlock;
    // (kClearBlocks workgroups of 256 threads' worth: single frames' bin launch has
    // kFrameBinBlock-thread workgroups)
    const uint32_t nclear = L.lay.batch ? (uint32_t)(per < kClearBlocks ? per : kClearBlocks)
                                        : kClearBlocks * (256u / (uint32_t)kFrameBinBlock);
    L.bp.clear_blocks = (pre || ids || vis) ? nclear : 0u;
    L.rp.pre_cleared = L.orp.pre_cleared = pre ? 1 : 0;
    L.sp.pre_cleared = pre ? 1 : 0;
}

// 200 us of the 100 MHz clock: longer than the host takes to queue one frame
constexpr uint64_t kHoldTicks = 20000;

rz_status enqueue_frame(rz_ctx* ctx, const FrameLaunch& L, hipStream_t s, bool prof) {
    if (prof) RZ_HIP(ctx, launch_hold(kHoldTicks, s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    // (no vertex pass: the bin pass transforms each face's corners)
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    RZ_HIP(ctx, launch_bin(L.bp, s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[2], s));
    if (L.lay.fused)
        RZ_HIP(ctx, launch_raster_shade(L.rp, L.sp, L.lay.ts, s));
    else
        RZ_HIP(ctx, launch_raster(L.rp, L.lay.ts, s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[3], s));
    // (the overflow loop runs inside the raster pass)
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[4], s));
    if (!L.lay.fused)
        RZ_HIP(ctx, launch_shade(L.sp, s));
    else if (L.lay.tiles_grid > 0)   // (0: no bin can overflow)
        RZ_HIP(ctx, launch_shade_tiles(L.rp, tiles_params(L), L.lay.ts, s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[5], s));
    return RZ_OK;
}

// Makes generation `gen` of the single-frame counter lines all zero for `ntiles` tiles,
// unless the context knows it is (a zeroing launch on the context's stream otherwise).
rz_status counts_ready(rz_ctx* ctx, int32_t ntiles, int gen) {
    if (ctx->count_ptr == ctx->tile_count.ptr && ctx->count_ntiles == ntiles && ctx->count_gen == gen) return RZ_OK;
    BinParams z;
    std::memset(&z, 0, sizeof z);
    z.zero_count = count_lines(ntiles);
    z.zero_lines = (uint32_t*)ctx->tile_count.ptr + (size_t)gen * z.zero_count * kCountStride;
    z.zero_blocks = (z.zero_count + kFrameBinBlock - 1) / kFrameBinBlock;
    RZ_HIP(ctx, launch_bin(z, ctx->stream));
    ctx->count_ptr = ctx->tile_count.ptr;
    ctx->count_ntiles = ntiles;
    ctx->count_gen = gen;
    return RZ_OK;
}

// A single frame on the context's stream, prepared with generation ctx->count_gen: the
// counter lines are made ready first, and the other generation is the next frame's.
rz_status run_frame(rz_ctx* ctx, const FrameLaunch& L, bool prof) {
    rz_status st = counts_ready(ctx, L.lay.ntiles, L.gen);
    if (st != RZ_OK) return st;
    ctx->count_ntiles = -1;   // unknown until the frame is enqueued whole
    if ((st = enqueue_frame(ctx, L, ctx->stream, prof)) != RZ_OK) return st;
    ctx->count_ntiles = L.lay.ntiles;
    ctx->count_gen = L.gen ^ 1;
    return RZ_OK;
}

// Statistics of one frame from its counters and per-workgroup partial sums (host copies).
void reduce_stats(const FrameLaunch& L, const unsigned long long* counters, const uint32_t* part,
                  uint64_t faces, rz_frame_stats* st) {
    auto sum = [&](size_t a, size_t b) {
        uint64_t t = 0;
        for (size_t i = a; i < b; ++i) t += part[i];
        return t;
    };
    const FrameLayout& y = L.lay;
    st->triangles += faces;
    st->front_triangles += sum(y.off_front, y.off_pairs);
    st->bin_pairs += sum(y.off_pairs, y.off_frag);
    st->overflow_tiles += counters[kCntOverflowTiles];
    // (batch meshes: that range holds only the overflow loop's partials, which the
    // one-launch form leaves unwritten for a mesh without an overflowed tile)
    if (!y.batch || counters[kCntOverflowTiles] != 0ull) st->fragments += sum(y.off_frag, y.off_cov);
    st->covered_pixels += sum(y.off_cov, y.n_part);
}

// Tiles the raster pass visits (those on the tile lists) and their pixels inside the
// image, from host copies of the 8 list counters (kCountStride apart) and the lists
// (`stride` entries apart); entries are (mesh << kListMeshShift) | tile when `lays`
// holds one layout per mesh of a batch, plain tiles otherwise.
void count_listed(const uint32_t* counts, const uint32_t* lists, size_t stride, const FrameLayout* lays,
                  size_t nlays, bool tagged, uint64_t* tiles_out, uint64_t* pixels_out) {
    for (int x = 0; x < 8; ++x) {
        const uint32_t n = counts[(size_t)x * kCountStride];
        for (uint32_t k = 0; k < n && k < stride; ++k) {
            const uint32_t e = lists[(size_t)x * stride + k];
            const uint32_t m = tagged ? e >> kListMeshShift : 0u, t = tagged ? e & kListTileMask : e;
            if (m >= nlays) continue;
            const FrameLayout& y = lays[m];
            const uint32_t tx = t % (uint32_t)y.tiles_x, ty = t / (uint32_t)y.tiles_x;
            const uint64_t w = std::min<uint64_t>((uint64_t)y.ts, (uint64_t)y.W - (uint64_t)tx * y.ts);
            const uint64_t h = std::min<uint64_t>((uint64_t)y.ts, (uint64_t)y.H - (uint64_t)ty * y.ts);
            tiles_out[tagged ? m : 0] += 1;
            pixels_out[tagged ? m : 0] += w * h;
        }
    }
}

// ---------------------------------------------------------------------------------
// Batches: every mesh's working buffers carved from one arena, the per-mesh parameters
// and the block maps of the vertex, bin and overflow passes in one table buffer.  The
// raster pass is one resident grid over tile lists that all meshes' bin workgroups fill
// (entries (mesh << kListMeshShift) | tile), raster and shade fused as for single
// frames; the overflowed tiles of every mesh are shaded by one more list-driven launch.

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

struct BatchLaunch {
    std::vector<FrameLaunch> frames;
    std::vector<uint64_t> faces;
    int ts = 16;
    int mode = kShadePlain;
    const VertexParams* d_vp = nullptr;
    const BinParams* d_bp = nullptr;
    const RasterParams* d_rp = nullptr;
    const RasterParams* d_orp = nullptr;
    const ShadeParams* d_sp = nullptr;
    const uint2* d_map[3] = {};     // vertex, bin (+ clears), overflow
    uint32_t blocks[3] = {};
    BatchLists bl = {};
    uint32_t raster_grid = 0, tiles_grid = 0;
    size_t arena_part_bytes = 0;   // statistics: first bytes of the arena (the batch's, then every frame's)
    size_t part_off = 0;           // the batch's statistics, after the barrier block
    // the persistent launch (k_batch_frame): each mesh's launch of generation 1 (frames[]
    // holds generation 0), the lists of each generation, the barriers
    bool persistent = false;
    std::vector<FrameLaunch> frames1;
    BatchLists bl1 = {};
    BatchSync q = {};
};

// The persistent launch's block at the start of the arena: 3 barriers of 8 counters (one
// per 128-byte line), the error flag and the phase clock (4 phases) on the next lines,
// then the phase-1 item cursor on a line of its own.
constexpr size_t kSyncWords = 3 * 8 * (kCountStride / 2);
constexpr size_t kCursorWord = kSyncWords + 16;   // (64-bit words from the block's start)
constexpr size_t kSyncBytes = (kCursorWord + 16) * 8;

rz_status prepare_batch(rz_ctx* ctx, const rz_batch_item* items, int32_t n, uint32_t flags, DevBuf& arena,
                        DevBuf& tables, BatchLaunch& B) {
    if (n < 1 || !items) return fail(ctx, RZ_ERR_INVALID_ARG, "empty batch");
    if (n > kMaxBatchMeshes) return fail(ctx, RZ_ERR_INVALID_ARG, "batch of %d meshes: at most %d", n, kMaxBatchMeshes);
    if (!(flags & RZ_FLAG_DEVICE_BUFFERS))
        return fail(ctx, RZ_ERR_INVALID_ARG, "a batch needs device buffers (RZ_FLAG_DEVICE_BUFFERS)");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    const bool cleared = (flags & RZ_FLAG_CLEARED) != 0;
    if ((flags & RZ_FLAG_MLP_SHADE) && !ctx->mlp_ready)
        return fail(ctx, RZ_ERR_INVALID_ARG, "RZ_FLAG_MLP_SHADE without network weights (rz_mlp_init / rz_mlp_set_weights)");
    B.ts = ctx->tile_size;
    B.mode = (flags & RZ_FLAG_MLP_SHADE) ? kShadeMlp : kShadePlain;
    B.persistent = false;
    const int bblock = 256;
    uint32_t grid = (uint32_t)ctx->cus * raster_batch_blocks_per_cu(B.ts, B.mode) / 8u * 8u;
    B.raster_grid = grid < 8 ? 8 : grid;
    B.tiles_grid = (uint32_t)ctx->persistent_blocks;
    std::vector<FrameLayout> lays(n);
    // arena: [batch statistics][statistics of every frame][tile lists][the other buffers of every frame]
    const size_t qbytes = 0;
    const size_t batch_stat_bytes = align256(((size_t)B.raster_grid * 2 + B.tiles_grid) * 4);
    size_t stat_bytes = batch_stat_bytes, work_bytes = 0, stride = 0;
    for (int32_t i = 0; i < n; ++i) {
        const rz_batch_item& it = items[i];
        rz_status st = check_frame(ctx, it.mesh, it.lights, it.num_lights, &it.args);
        if (st != RZ_OK) return st;
        if (!it.frame || !it.depth) return fail(ctx, RZ_ERR_INVALID_ARG, "batch item %d: null frame or depth", i);
        lays[i] = layout_frame(it.mesh, it.args, ctx->tile_size, ctx->bin_cap > 0 ? ctx->bin_cap : kDefaultBinCap,
                               true, 0, 0, true, bblock);
        if ((uint32_t)lays[i].ntiles > kListTileMask)
            return fail(ctx, RZ_ERR_INVALID_ARG, "batch item %d: %d tiles, at most %u", i, lays[i].ntiles, kListTileMask);
        stat_bytes += align256(lays[i].bytes[kBufCounters]) + align256(lays[i].bytes[kBufPart]);
        for (int b = 0; b < kBufCounters; ++b) work_bytes += align256(lays[i].bytes[b]);
        stride += (size_t)lays[i].ntiles;
    }
    const size_t list_bytes = align256(9 * stride * 4);   // 8 lists, then the overflowed tiles
    rz_status st = ensure(ctx, arena, stat_bytes + list_bytes + work_bytes);
    if (st != RZ_OK) return st;
    B.frames.assign(n, FrameLaunch());
    B.faces.assign(n, 0);
    B.arena_part_bytes = stat_bytes;
    B.part_off = qbytes;
    char* base = (char*)arena.ptr;
    uint32_t* batch_part = (uint32_t*)(base + qbytes);
    uint32_t* lists = (uint32_t*)(base + stat_bytes);
    size_t so = batch_stat_bytes, wo = stat_bytes + list_bytes;
    std::vector<uint2> map[3], clear_items;
    for (int32_t i = 0; i < n; ++i) {
        const rz_batch_item& it = items[i];
        void* ptrs[kNumBufs];
        for (int b = 0; b < kBufCounters; ++b) {
            ptrs[b] = base + wo;
            wo += align256(lays[i].bytes[b]);
        }
        ptrs[kBufCounters] = base + so;
        so += align256(lays[i].bytes[kBufCounters]);
        ptrs[kBufPart] = base + so;
        so += align256(lays[i].bytes[kBufPart]);
        FrameLaunch& L = B.frames[i];
        fill_frame(it.mesh, it.lights, it.num_lights, it.args, flags, lays[i], ptrs, L);
        L.sp.mlp = (const uint16_t*)ctx->mlp_packed.ptr;
        bind_outputs(L, cleared, it.frame, it.depth, it.ids);
        // the shared lists; their 8 counters are the first mesh's (cleared by its vertex pass)
        const FrameLaunch& L0 = B.frames[0];
        L.bp.lists = lists;
        L.bp.list_count = L0.bp.list_count;
        L.bp.ntiles = (int32_t)stride;
        L.bp.list_tag = (uint32_t)i << kListMeshShift;
        L.bp.batch_ovf = L0.counters + kCntBatchOverflow;
        L.bp.ovf_list = lists + (size_t)8 * stride;
        L.rp.lists = L.orp.lists = lists;
        L.rp.list_count = L.orp.list_count = L0.bp.list_count;
        L.rp.ovf_list = L.orp.ovf_list = L.bp.ovf_list;
        B.faces[i] = (uint64_t)it.mesh->nf;
        const uint32_t nb[3] = {vertex_blocks(L.vp.nv, L.vp.nn, L.vp.ntiles, bblock),
                                bin_blocks(L.bp.nf, bblock) + L.bp.clear_blocks, overflow_blocks(L.rp.nf, bblock)};
        for (int p = 0; p < 3; ++p)
            for (uint32_t b = 0; b < nb[p]; ++b)
                (p == 1 && b >= bin_blocks(L.bp.nf, bblock) ? clear_items : map[p]).push_back(make_uint2((uint32_t)i, b));
    }
    // the clearing workgroups after every mesh's bin workgroups: stores that only stream
    // run best once the bin workgroups' dependent loads and atomics are done (the
    // 256-mesh batch's bin launch 0.88 -> 0.83 ms against each mesh's clears after its
    // bins, and 0.88 ms with the clears dealt evenly between the bin workgroups)
    map[1].insert(map[1].end(), clear_items.begin(), clear_items.end());
    B.bl.lists = lists;
    B.bl.list_count = B.frames[0].bp.list_count;
    B.bl.stride = (uint32_t)stride;
    B.bl.overflow = B.frames[0].counters + kCntBatchOverflow;
    B.bl.ovf_list = lists + (size_t)8 * stride;
    B.bl.part_frag = batch_part;
    B.bl.part_cov = batch_part + B.raster_grid;
    // tables: the four parameter arrays and the three block maps
    size_t off[7], total = 0;
    const size_t sz[4] = {sizeof(VertexParams), sizeof(BinParams), sizeof(RasterParams), sizeof(ShadeParams)};
    for (int p = 0; p < 4; ++p) {
        off[p] = total;
        total = align256(total + sz[p] * (size_t)n);
    }
    const size_t off_orp = total;
    total = align256(total + sizeof(RasterParams) * (size_t)n);
    for (int p = 0; p < 3; ++p) {
        off[4 + p] = total;
        total = align256(total + map[p].size() * sizeof(uint2));
        B.blocks[p] = (uint32_t)map[p].size();
    }
    // the items' lights beyond the kernel arguments, one after the other
    const size_t off_ext = total;
    std::vector<size_t> ext_at(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        ext_at[i] = total;
        total += B.frames[i].ext_lights.size() * sizeof(float);
    }
    total = align256(total);
    if ((st = ensure(ctx, tables, total)) != RZ_OK) return st;
    for (int32_t i = 0; i < n; ++i)
        B.frames[i].sp.lights_ext =
            B.frames[i].ext_lights.empty() ? nullptr : (const float*)((char*)tables.ptr + ext_at[i]);
    std::vector<char> host(total, 0);
    for (int32_t i = 0; i < n; ++i)
        if (!B.frames[i].ext_lights.empty())
            std::memcpy(host.data() + ext_at[i], B.frames[i].ext_lights.data(),
                        B.frames[i].ext_lights.size() * sizeof(float));
    (void)off_ext;
    for (int32_t i = 0; i < n; ++i) {
        std::memcpy(host.data() + off[0] + i * sz[0], &B.frames[i].vp, sz[0]);
        std::memcpy(host.data() + off[1] + i * sz[1], &B.frames[i].bp, sz[1]);
        std::memcpy(host.data() + off[2] + i * sz[2], &B.frames[i].rp, sz[2]);
        std::memcpy(host.data() + off[3] + i * sz[3], &B.frames[i].sp, sz[3]);
        std::memcpy(host.data() + off_orp + i * sizeof(RasterParams), &B.frames[i].orp, sizeof(RasterParams));
    }
    for (int p = 0; p < 3; ++p)
        if (!map[p].empty()) std::memcpy(host.data() + off[4 + p], map[p].data(), map[p].size() * sizeof(uint2));
    // the table buffer may still be read by queued work of an earlier batch
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipMemcpy(tables.ptr, host.data(), total, hipMemcpyHostToDevice));
    char* t = (char*)tables.ptr;
    B.d_vp = (const VertexParams*)(t + off[0]);
    B.d_bp = (const BinParams*)(t + off[1]);
    B.d_rp = (const RasterParams*)(t + off[2]);
    B.d_sp = (const ShadeParams*)(t + off[3]);
    B.d_orp = (const RasterParams*)(t + off_orp);
    for (int p = 0; p < 3; ++p) B.d_map[p] = (const uint2*)(t + off[4 + p]);
    return RZ_OK;
}

// A batch for the persistent launch (k_batch_frame): every mesh laid out as a single frame
// (two generations of counter lines), the tile lists and raster statistics shared, the
// parameter tables holding each mesh's bin, raster and overflow parameters for generation 0
// and then generation 1, and the barriers at the start of the arena, zeroed.
rz_status prepare_batch_sync(rz_ctx* ctx, const rz_batch_item* items, int32_t n, uint32_t flags, DevBuf& arena,
                             DevBuf& tables, BatchLaunch& B) {
    if (n < 1 || !items) return fail(ctx, RZ_ERR_INVALID_ARG, "empty batch");
    if (n > kMaxBatchMeshes) return fail(ctx, RZ_ERR_INVALID_ARG, "batch of %d meshes: at most %d", n, kMaxBatchMeshes);
    if (!(flags & RZ_FLAG_DEVICE_BUFFERS))
        return fail(ctx, RZ_ERR_INVALID_ARG, "a batch needs device buffers (RZ_FLAG_DEVICE_BUFFERS)");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    const bool cleared = (flags & RZ_FLAG_CLEARED) != 0;
    if ((flags & RZ_FLAG_MLP_SHADE) && !ctx->mlp_ready)
        return fail(ctx, RZ_ERR_INVALID_ARG, "RZ_FLAG_MLP_SHADE without network weights (rz_mlp_init / rz_mlp_set_weights)");
    B.ts = ctx->tile_size;
    B.mode = (flags & RZ_FLAG_MLP_SHADE) ? kShadeMlp : kShadePlain;
    B.persistent = true;
    const int bblock = batch_block(B.ts);
    // every workgroup of the launch resident at once: the launch's occupancy on every CU
    uint32_t grid = (uint32_t)ctx->cus * batch_frame_blocks_per_cu(B.ts, B.mode) / 8u * 8u;
    B.raster_grid = B.tiles_grid = grid < 8 ? 8 : grid;
    const int cap = ctx->bin_cap > 0 ? ctx->bin_cap : kDefaultBinCap;
    std::vector<FrameLayout> lays(n);
    // arena: [barriers][batch statistics][statistics of every mesh][tile lists][every mesh's buffers]
    const size_t sync_bytes = align256(kSyncBytes);
    const size_t batch_stat_bytes = sync_bytes + align256(((size_t)B.raster_grid * 2 + B.tiles_grid) * 4);
    size_t stat_bytes = batch_stat_bytes, work_bytes = 0, stride = 0;
    for (int32_t i = 0; i < n; ++i) {
        const rz_batch_item& it = items[i];
        rz_status st = check_frame(ctx, it.mesh, it.lights, it.num_lights, &it.args);
        if (st != RZ_OK) return st;
        if (!it.frame || !it.depth) return fail(ctx, RZ_ERR_INVALID_ARG, "batch item %d: null frame or depth", i);
        lays[i] = layout_batch_sync(it.mesh, it.args, B.ts, cap, bblock);
        if ((uint32_t)lays[i].ntiles > kListTileMask)
            return fail(ctx, RZ_ERR_INVALID_ARG, "batch item %d: %d tiles, at most %u", i, lays[i].ntiles, kListTileMask);
        stat_bytes += align256(lays[i].bytes[kBufPart]);
        for (int b = 0; b < kNumBufs; ++b)
            if (b != kBufPart) work_bytes += align256(lays[i].bytes[b]);
        stride += (size_t)lays[i].ntiles;
    }
    const size_t list_bytes = align256(9 * stride * 4);   // 8 lists, then the overflowed tiles
    rz_status st = ensure(ctx, arena, stat_bytes + list_bytes + work_bytes);
    if (st != RZ_OK) return st;
    B.frames.assign(n, FrameLaunch());
    B.frames1.assign(n, FrameLaunch());
    B.faces.assign(n, 0);
    B.arena_part_bytes = stat_bytes;
    B.part_off = sync_bytes;
    char* base = (char*)arena.ptr;
    uint32_t* batch_part = (uint32_t*)(base + sync_bytes);
    uint32_t* lists = (uint32_t*)(base + stat_bytes);
    size_t so = batch_stat_bytes, wo = stat_bytes + list_bytes;
    std::vector<uint2> bmap, store_items;
    for (int32_t i = 0; i < n; ++i) {
        const rz_batch_item& it = items[i];
        void* ptrs[kNumBufs];
        for (int b = 0; b < kNumBufs; ++b) {
            if (b == kBufPart) continue;
            ptrs[b] = base + wo;
            wo += align256(lays[i].bytes[b]);
        }
        ptrs[kBufPart] = base + so;
        so += align256(lays[i].bytes[kBufPart]);
        for (int g = 0; g < 2; ++g) {
            FrameLaunch& L = g ? B.frames1[i] : B.frames[i];
            fill_frame(it.mesh, it.lights, it.num_lights, it.args, flags, lays[i], ptrs, L, g);
            L.sp.mlp = (const uint16_t*)ctx->mlp_packed.ptr;
            L.sp.mode = B.mode;
            bind_outputs(L, cleared, it.frame, it.depth, it.ids);
            // the shared lists; their 8 counters, and the batch's overflow count, are the
            // first mesh's (its bin work zeroes the other generation's for the next frame)
            const FrameLaunch& L0 = g ? B.frames1[0] : B.frames[0];
            L.bp.lists = lists;
            L.bp.list_count = L0.bp.list_count;
            L.bp.ntiles = (int32_t)stride;
            L.bp.list_tag = (uint32_t)i << kListMeshShift;
            L.bp.batch_ovf = L0.counters + kCntBatchOverflow;
            L.bp.ovf_list = lists + (size_t)8 * stride;   // (one frame at a time: shared by both)
            L.rp.lists = L.orp.lists = lists;
            L.rp.list_count = L.orp.list_count = L0.bp.list_count;
            L.rp.ovf_list = L.orp.ovf_list = L.bp.ovf_list;
            L.orp.part_frag = L.part + lays[i].off_ovf;
        }
        B.faces[i] = (uint64_t)it.mesh->nf;
        const FrameLaunch& L = B.frames[i];
        const uint32_t nb = bin_blocks(L.bp.nf, bblock) + (L.bp.zero_count + bblock - 1) / bblock + L.bp.clear_blocks;
        const uint32_t nbin = bin_blocks(L.bp.nf, bblock);
        for (uint32_t b = 0; b < nb; ++b) (b < nbin ? bmap : store_items).push_back(make_uint2((uint32_t)i, b));
    }
    bmap.insert(bmap.end(), store_items.begin(), store_items.end());   // (as for the bin launch)
    for (int g = 0; g < 2; ++g) {
        BatchLists& bl = g ? B.bl1 : B.bl;
        const FrameLaunch& L0 = g ? B.frames1[0] : B.frames[0];
        bl.lists = lists;
        bl.list_count = L0.bp.list_count;
        bl.stride = (uint32_t)stride;
        bl.overflow = L0.counters + kCntBatchOverflow;
        bl.ovf_list = lists + (size_t)8 * stride;
        bl.part_frag = batch_part;
        bl.part_cov = batch_part + B.raster_grid;
    }
    // tables: bin, raster and overflow parameters of both generations, shade parameters,
    // the bin items' map, the lights beyond the kernel arguments
    size_t total = 0;
    const size_t off_bp = total;
    total = align256(total + 2 * sizeof(BinParams) * (size_t)n);
    const size_t off_rp = total;
    total = align256(total + 2 * sizeof(RasterParams) * (size_t)n);
    const size_t off_orp = total;
    total = align256(total + 2 * sizeof(RasterParams) * (size_t)n);
    const size_t off_sp = total;
    total = align256(total + sizeof(ShadeParams) * (size_t)n);
    const size_t off_bmap = total;
    total = align256(total + bmap.size() * sizeof(uint2));
    std::vector<size_t> ext_at(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        ext_at[i] = total;
        total += B.frames[i].ext_lights.size() * sizeof(float);
    }
    total = align256(total);
    if ((st = ensure(ctx, tables, total)) != RZ_OK) return st;
    std::vector<char> host(total, 0);
    for (int32_t i = 0; i < n; ++i) {
        FrameLaunch& L = B.frames[i];
        L.sp.lights_ext = L.ext_lights.empty() ? nullptr : (const float*)((char*)tables.ptr + ext_at[i]);
        if (!L.ext_lights.empty())
            std::memcpy(host.data() + ext_at[i], L.ext_lights.data(), L.ext_lights.size() * sizeof(float));
        for (int g = 0; g < 2; ++g) {
            const FrameLaunch& Lg = g ? B.frames1[i] : B.frames[i];
            const size_t k = (size_t)g * n + i;
            std::memcpy(host.data() + off_bp + k * sizeof(BinParams), &Lg.bp, sizeof(BinParams));
            std::memcpy(host.data() + off_rp + k * sizeof(RasterParams), &Lg.rp, sizeof(RasterParams));
            std::memcpy(host.data() + off_orp + k * sizeof(RasterParams), &Lg.orp, sizeof(RasterParams));
        }
        std::memcpy(host.data() + off_sp + i * sizeof(ShadeParams), &L.sp, sizeof(ShadeParams));
    }
    if (!bmap.empty()) std::memcpy(host.data() + off_bmap, bmap.data(), bmap.size() * sizeof(uint2));
    // the table buffer and the arena may still be read by queued work of an earlier batch
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipMemcpy(tables.ptr, host.data(), total, hipMemcpyHostToDevice));
    char* t = (char*)tables.ptr;
    B.d_bp = (const BinParams*)(t + off_bp);
    B.d_rp = (const RasterParams*)(t + off_rp);
    B.d_orp = (const RasterParams*)(t + off_orp);
    B.d_sp = (const ShadeParams*)(t + off_sp);
    // the barriers start at zero, and so does generation 0 of every mesh's counter lines
    // (each frame's bin work zeroes the other generation for the next)
    // (on the context stream, which the launches follow; this is never inside a capture)
    RZ_HIP(ctx, hipMemsetAsync(arena.ptr, 0, sync_bytes, ctx->stream));
    for (int32_t i = 0; i < n; ++i)
        RZ_HIP(ctx, hipMemsetAsync(B.frames[i].bp.tile_count, 0, (size_t)count_lines(lays[i].ntiles) * 4 * kCountStride,
                                   ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    BatchSync& q = B.q;
    q.bar = (unsigned long long*)arena.ptr;
    q.error = (uint32_t*)((unsigned long long*)arena.ptr + kSyncWords);
    q.clock = nullptr;   // (set per frame when profiling)
    q.cursor = (uint32_t*)((unsigned long long*)arena.ptr + kCursorWord);
    q.G = B.raster_grid;
    q.nB = (uint32_t)bmap.size();
    q.bmap = (const uint2*)(t + off_bmap);
    q.nmesh = (uint32_t)n;
    return RZ_OK;
}

rz_status prepare_any_batch(rz_ctx* ctx, const rz_batch_item* items, int32_t n, uint32_t flags, DevBuf& arena,
                            DevBuf& tables, BatchLaunch& B) {
    return ctx->batch_persistent ? prepare_batch_sync(ctx, items, n, flags, arena, tables, B)
                                 : prepare_batch(ctx, items, n, flags, arena, tables, B);
}

rz_status enqueue_batch(rz_ctx* ctx, const BatchLaunch& B, hipStream_t s, bool prof) {
    if (B.persistent) {
        if (prof) RZ_HIP(ctx, launch_hold(kHoldTicks, s));
        if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[0], s));
        RZ_HIP(ctx, launch_batch_frame(B.d_bp, B.d_rp, B.d_orp, B.d_sp, B.bl, B.bl1, B.q, B.ts, B.mode, s));
        if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[5], s));
        return RZ_OK;
    }
    if (prof) RZ_HIP(ctx, launch_hold(kHoldTicks, s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    RZ_HIP(ctx, launch_vertex_batch(B.d_vp, B.d_map[0], B.blocks[0], s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    RZ_HIP(ctx, launch_bin_batch(B.d_bp, B.d_map[1], B.blocks[1], s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[2], s));
    RZ_HIP(ctx, launch_batch_tiles(B.d_rp, B.d_sp, B.bl, B.raster_grid, B.tiles_grid, B.ts, B.mode, s, true));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[3], s));
    RZ_HIP(ctx, launch_overflow_batch(B.d_orp, B.d_map[2], B.blocks[2], s));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[4], s));
    RZ_HIP(ctx, launch_batch_tiles(B.d_rp, B.d_sp, B.bl, B.raster_grid, B.tiles_grid, B.ts, B.mode, s, false));
    if (prof) RZ_HIP(ctx, hipEventRecord(ctx->ev[5], s));
    return RZ_OK;
}

rz_status event_times(rz_ctx* ctx, rz_frame_stats* stats) {
    float t[5];
    for (int i = 0; i < 5; ++i) RZ_HIP(ctx, hipEventElapsedTime(&t[i], ctx->ev[i], ctx->ev[i + 1]));
    stats->ms_vertex = t[0];
    stats->ms_bin = t[1];
    stats->ms_raster = t[2];
    stats->ms_overflow = t[3];
    stats->ms_shade = t[4];
    RZ_HIP(ctx, hipEventElapsedTime(&stats->ms_frame, ctx->ev[0], ctx->ev[5]));
    return RZ_OK;
}

// Captures `frames` repetitions of enqueue(ctx, stream) into an executable graph.
template <typename F>
rz_status capture(rz_ctx* ctx, int32_t frames, F enqueue, hipGraphExec_t* out) {
    hipStream_t s = ctx->stream;
    RZ_HIP(ctx, hipStreamSynchronize(s));
    RZ_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int32_t k = 0; k < frames; ++k) {
        rz_status st = enqueue(k, s);
        if (st != RZ_OK) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
    }
    hipGraph_t g = nullptr;
    RZ_HIP(ctx, hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(ctx, RZ_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return RZ_OK;
}

rz_status batch_stats(rz_ctx* ctx, const BatchLaunch& B, const DevBuf& arena, rz_frame_stats* stats) {
    std::vector<char> host(B.arena_part_bytes);
    RZ_HIP(ctx, hipMemcpyAsync(host.data(), arena.ptr, host.size(), hipMemcpyDeviceToHost, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::memset(stats, 0, sizeof *stats);
    // the persistent launch: the last frame's generation of counter lines, from the
    // arrivals at its final barrier
    int gen = 0;
    if (B.persistent) {
        const unsigned long long* sync = (const unsigned long long*)host.data();
        if (((const uint32_t*)(sync + kSyncWords))[0] != 0u)   // BatchSync::error
            return fail(ctx, RZ_ERR_HIP, "batch frame: a barrier wait gave up");
        unsigned long long arrived = 0;
        for (int x = 0; x < 8; ++x) arrived += sync[(size_t)(2 * 8 + x) * (kCountStride / 2)];
        const unsigned long long frames = arrived / B.q.G;
        gen = frames > 0 ? (int)((frames - 1) & 1ull) : 0;
    }
    const std::vector<FrameLaunch>& launches = gen ? B.frames1 : B.frames;
    const BatchLists& bl = gen ? B.bl1 : B.bl;
    std::vector<uint32_t> list_counts(8 * kCountStride), lists((size_t)8 * bl.stride);
    RZ_HIP(ctx, hipMemcpyAsync(list_counts.data(), bl.list_count, list_counts.size() * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
    RZ_HIP(ctx, hipMemcpyAsync(lists.data(), bl.lists, lists.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    // each mesh's global counters (in the statistics region, or in its counter lines)
    std::vector<unsigned long long> counters(launches.size() * kNumCounters);
    for (size_t i = 0; i < launches.size(); ++i)
        RZ_HIP(ctx, hipMemcpyAsync(counters.data() + i * kNumCounters, launches[i].counters, kNumCounters * 8,
                                   hipMemcpyDeviceToHost, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t* bp = (const uint32_t*)(host.data() + B.part_off);   // the batch's raster and overflow-shade partials
    for (uint32_t i = 0; i < B.raster_grid; ++i) stats->fragments += bp[i];
    for (uint32_t i = 0; i < B.raster_grid + B.tiles_grid; ++i) stats->covered_pixels += bp[B.raster_grid + i];
    for (size_t i = 0; i < launches.size(); ++i) {
        const FrameLaunch& L = launches[i];
        const size_t op = (size_t)((char*)L.part - (char*)arena.ptr);
        reduce_stats(L, counters.data() + i * kNumCounters, (const uint32_t*)(host.data() + op), B.faces[i], stats);
    }
    std::vector<FrameLayout> lays;
    for (const FrameLaunch& L : launches) lays.push_back(L.lay);
    std::vector<uint64_t> tiles(lays.size(), 0), pixels(lays.size(), 0);
    count_listed(list_counts.data(), lists.data(), bl.stride, lays.data(), lays.size(), true, tiles.data(),
                 pixels.data());
    for (size_t i = 0; i < lays.size(); ++i) {
        stats->raster_tiles += tiles[i];
        stats->raster_pixels += pixels[i];
    }
    return RZ_OK;
}

}  // namespace

extern "C" {

int32_t rz_abi_version(void) { return RZ_ABI_VERSION; }

const char* rz_status_string(rz_status s) {
    switch (s) {
        case RZ_OK: return "ok";
        case RZ_ERR_INVALID_ARG: return "invalid argument";
        case RZ_ERR_INDEX_OOB: return "index out of bounds";
        case RZ_ERR_HIP: return "HIP runtime error";
        case RZ_ERR_OOM: return "out of device memory";
        case RZ_ERR_NO_DEVICE: return "no HIP device";
    }
    return "unknown status";
}

const char* rz_last_error(const rz_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

void rz_default_args(rz_args* a, uint32_t width, uint32_t height) {
    if (!a) return;
    std::memset(a, 0, sizeof *a);
    a->image_width = width;
    a->image_height = height;
    a->aspect_ratio = height ? (float)width / (float)height : 1.0f;
    a->scale = 1.0f;
    a->fovy_degrees = 45.0f;
    a->z_near = 0.1f;
    a->z_far = 6.0f;
}

rz_status rz_frame_matrices(const rz_args* args, rz_matrices* out) {
    if (!args || !out) return RZ_ERR_INVALID_ARG;
    build_matrices(*args, *out);
    return RZ_OK;
}

rz_status rz_create(int32_t device, rz_ctx** out) {
    if (!out) return RZ_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, RZ_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(nullptr, RZ_ERR_NO_DEVICE, "device %d out of range [0,%d)", device, n);
    rz_ctx* ctx = new rz_ctx();
    ctx->device = device;
    rz_status st = RZ_OK;
    do {
        hipError_t e = hipSetDevice(device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        for (int i = 0; i < 6 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->ev[i]);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&ctx->host_counters, sizeof(unsigned long long) * kNumCounters,
                              hipHostMallocDefault);
        int cus = 0;
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if (e == hipSuccess) ctx->persistent_blocks = (uint32_t)(cus * kPersistentPerCU) / 8u * 8u;
        ctx->cus = cus;
        size_t free_bytes = 0, total_bytes = 0;
        if (e == hipSuccess) e = hipMemGetInfo(&free_bytes, &total_bytes);
        ctx->auto_bin_bytes = std::min(kAutoBinBytesMax, free_bytes / 8);
        if (e != hipSuccess) st = fail(ctx, RZ_ERR_HIP, "context setup: %s", hipGetErrorString(e));
        if (st == RZ_OK) st = scratch_address(ctx);
    } while (0);
    if (st != RZ_OK) {
        rz_destroy(ctx);
        return st;
    }
    *out = ctx;
    return RZ_OK;
}

void rz_destroy(rz_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->raster, &ctx->cnormal, &ctx->rec,   &ctx->tile_count, &ctx->bins,
                      &ctx->vis,    &ctx->lists,   &ctx->frame,   &ctx->depth, &ctx->ids,        &ctx->counters,
                      &ctx->partials, &ctx->batch_arena, &ctx->batch_tables, &ctx->mlp_w, &ctx->mlp_m,
                      &ctx->mlp_v, &ctx->mlp_packed, &ctx->mlp_gpart, &ctx->mlp_lpart, &ctx->mlp_loss,
                      &ctx->mlp_count, &ctx->lights_ext};
    for (DevBuf* b : bufs)
        if (b->ptr) (void)hipFree(b->ptr);
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->host_counters) (void)hipHostFree(ctx->host_counters);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

rz_status rz_set_tuning(rz_ctx* ctx, int32_t tile_size, int32_t bin_capacity) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (tile_size != 8 && tile_size != 16 && tile_size != 32)
        return fail(ctx, RZ_ERR_INVALID_ARG, "tile size %d not in {8,16,32}", tile_size);
    if (bin_capacity < 0 || bin_capacity > (1 << 20))
        return fail(ctx, RZ_ERR_INVALID_ARG, "bin capacity %d not in [0, 2^20] (0: automatic)", bin_capacity);
    ctx->tile_size = tile_size;
    ctx->bin_cap = bin_capacity;
    return RZ_OK;
}

rz_status rz_set_fused_shade(rz_ctx* ctx, int32_t enable) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    ctx->fused = enable != 0;
    return RZ_OK;
}

int32_t rz_get_fused_shade(const rz_ctx* ctx) { return ctx && ctx->fused ? 1 : 0; }

rz_status rz_get_tuning(const rz_ctx* ctx, int32_t* tile_size, int32_t* bin_capacity) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (tile_size) *tile_size = ctx->tile_size;
    if (bin_capacity) *bin_capacity = ctx->bin_cap;
    return RZ_OK;
}

rz_status rz_set_profiling(rz_ctx* ctx, int32_t enable) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    ctx->profiling = enable != 0;
    return RZ_OK;
}

void* rz_stream(rz_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

rz_status rz_sync(rz_ctx* ctx) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RZ_OK;
}

rz_status rz_device_alloc(rz_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return RZ_ERR_INVALID_ARG;
    *out = nullptr;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, hipMalloc(out, bytes ? bytes : 1));
    return RZ_OK;
}

rz_status rz_device_free(rz_ctx* ctx, void* ptr) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (!ptr) return RZ_OK;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipFree(ptr));
    return RZ_OK;
}

rz_status rz_copy_to_device(rz_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return RZ_ERR_INVALID_ARG;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RZ_OK;
}

rz_status rz_copy_to_host(rz_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return RZ_ERR_INVALID_ARG;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RZ_OK;
}

rz_status rz_copy_device(rz_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return RZ_ERR_INVALID_ARG;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    const bool aligned = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0);
    const size_t head = aligned ? bytes / 16 * 16 : 0;
    RZ_HIP(ctx, launch_copy(src, dst, head / 16, ctx->stream));
    if (bytes > head)
        RZ_HIP(ctx, hipMemcpyAsync((uint8_t*)dst + head, (const uint8_t*)src + head, bytes - head,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RZ_OK;
}

namespace {

// 10 bits of each coordinate interleaved: z2 y2 x2 z1 y1 x1 ...
uint32_t morton30(uint32_t x, uint32_t y, uint32_t z) {
    auto spread = [](uint32_t v) {
        v &= 0x3FFu;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

// Face processing order of the bin pass: by the Morton code of the face centroid
// (positions as given: model space, or raster x, y for raster-space meshes), ties in
// input order; the faces' indices (validated) in that order.
void bin_faces(const rz_mesh_desc* d, std::vector<BinFace>& out) {
    std::vector<int32_t> fv(3 * (size_t)d->num_faces);
    for (size_t f = 0; f < fv.size() / 3; ++f)
        for (int k = 0; k < 3; ++k) fv[3 * f + k] = d->faces[f].vertices[k];
    const size_t F = (size_t)d->num_faces;
    const int stride = d->raster_space ? 4 : 3;
    const int dims = d->raster_space ? 2 : 3;
    std::vector<float> c(3 * F, 0.f);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t f = 0; f < F; ++f)
        for (int a = 0; a < dims; ++a) {
            float v = 0.f;
            for (int k = 0; k < 3; ++k) v += d->vertices[(size_t)stride * fv[3 * f + k] + a];
            v /= 3.0f;
            c[3 * f + a] = v;
            if (std::isfinite(v)) {
                lo[a] = std::min(lo[a], v);
                hi[a] = std::max(hi[a], v);
            }
        }
    std::vector<uint64_t> key(F);
    for (size_t f = 0; f < F; ++f) {
        uint32_t q[3] = {0, 0, 0};
        for (int a = 0; a < dims; ++a) {
            const float span = hi[a] - lo[a];
            const float t = span > 0.f ? (c[3 * f + a] - lo[a]) / span : 0.f;
            q[a] = t >= 0.f ? (t < 1.f ? (uint32_t)(t * 1023.0f) : 1023u) : 0u;   // NaN -> 0
        }
        key[f] = ((uint64_t)morton30(q[0], q[1], q[2]) << 32) | (uint64_t)f;
    }
    std::sort(key.begin(), key.end());
    out.resize(F);
    for (size_t i = 0; i < F; ++i) {
        const size_t f = (size_t)(key[i] & 0xFFFFFFFFu);
        const rz_triangle& t = d->faces[f];
        BinFace& b = out[i];
        b.q0 = make_int4(t.vertices[0], t.vertices[1], t.vertices[2], (int32_t)f);
        b.q1 = make_int4(t.normals[0], t.normals[1], t.normals[2], t.material);
        b.q2 = make_int4(t.uvs[0] < 0 ? -1 : t.uvs[0], t.uvs[1] < 0 ? -1 : t.uvs[1], t.uvs[2] < 0 ? -1 : t.uvs[2], 0);
    }
}

}  // namespace

rz_status rz_mesh_create(rz_ctx* ctx, const rz_mesh_desc* d, rz_mesh** out) {
    if (!ctx || !out) return RZ_ERR_INVALID_ARG;
    *out = nullptr;
    rz_status st = validate_mesh(ctx, d);
    if (st != RZ_OK) return st;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    rz_mesh* m = new rz_mesh();
    m->device = ctx->device;
    m->nv = d->num_vertices;
    m->nn = d->num_normals;
    m->nt = d->num_uvs;
    m->nf = d->num_faces;
    m->nm = d->num_materials;
    m->raster_space = d->raster_space != 0;
    auto cleanup = [&](rz_status s) {
        rz_mesh_destroy(ctx, m);
        return s;
    };
    const size_t V = (size_t)m->nv, N = (size_t)m->nn, F = (size_t)m->nf;
    // The faces in binning order; vertices, normals and uvs are renumbered by first use
    // in that order, so that the bin pass reads them (and a batch's vertex pass writes
    // them) in address order.  Only storage moves: every value and face index is the caller's.
    std::vector<BinFace> faces;
    bin_faces(d, faces);
    std::vector<int32_t> vorder, norder, torder;   // new index -> caller's index
    {
        std::vector<int32_t> vmap(V, -1), nmap(N, -1), tmap((size_t)m->nt, -1);
        auto use = [](int32_t& idx, std::vector<int32_t>& map, std::vector<int32_t>& order) {
            if (idx < 0) return;
            if (map[(size_t)idx] < 0) {
                map[(size_t)idx] = (int32_t)order.size();
                order.push_back(idx);
            }
            idx = map[(size_t)idx];
        };
        for (BinFace& b : faces) {
            use(b.q0.x, vmap, vorder); use(b.q0.y, vmap, vorder); use(b.q0.z, vmap, vorder);
            use(b.q1.x, nmap, norder); use(b.q1.y, nmap, norder); use(b.q1.z, nmap, norder);
            use(b.q2.x, tmap, torder); use(b.q2.y, tmap, torder); use(b.q2.z, tmap, torder);
        }
        for (size_t i = 0; i < V; ++i) if (vmap[i] < 0) vorder.push_back((int32_t)i);
        for (size_t i = 0; i < N; ++i) if (nmap[i] < 0) norder.push_back((int32_t)i);
        for (size_t i = 0; i < tmap.size(); ++i) if (tmap[i] < 0) torder.push_back((int32_t)i);
    }
    // positions: (x, y, z, 0) per vertex for the single-frame bin pass, SoA for a batch's
    // coalesced vertex pass
    if (m->raster_space) {
        std::vector<float> rv(4 * V);
        for (size_t i = 0; i < V; ++i)
            for (int c = 0; c < 4; ++c) rv[4 * i + c] = d->vertices[4 * (size_t)vorder[i] + c];
        if ((st = upload(ctx, m, &m->rin, rv.data(), V * 16)) != RZ_OK) return cleanup(st);
    } else {
        std::vector<float> sx(V), sy(V), sz(V);
        for (size_t i = 0; i < V; ++i) {
            const size_t o = (size_t)vorder[i];
            sx[i] = d->vertices[3 * o];
            sy[i] = d->vertices[3 * o + 1];
            sz[i] = d->vertices[3 * o + 2];
        }
        if ((st = upload(ctx, m, &m->px, sx.data(), V * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->py, sy.data(), V * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->pz, sz.data(), V * 4)) != RZ_OK) return cleanup(st);
        std::vector<float4> p4(V);
        for (size_t i = 0; i < V; ++i) p4[i] = make_float4(sx[i], sy[i], sz[i], 0.f);
        if ((st = upload(ctx, m, &m->pos4, p4.data(), V * 16)) != RZ_OK) return cleanup(st);
        // face normals for flat shading, model space
        std::vector<float> fx(F), fy(F), fz(F);
        for (size_t f = 0; f < F; ++f) {
            const rz_triangle& t = d->faces[f];
            float nrm[3];
            face_normal(d->vertices + 3 * (size_t)t.vertices[0], d->vertices + 3 * (size_t)t.vertices[1],
                        d->vertices + 3 * (size_t)t.vertices[2], nrm);
            fx[f] = nrm[0]; fy[f] = nrm[1]; fz[f] = nrm[2];
        }
        if ((st = upload(ctx, m, &m->fnx, fx.data(), F * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->fny, fy.data(), F * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->fnz, fz.data(), F * 4)) != RZ_OK) return cleanup(st);
        std::vector<float4> f4(F);
        for (size_t f = 0; f < F; ++f) f4[f] = make_float4(fx[f], fy[f], fz[f], 0.f);
        if ((st = upload(ctx, m, &m->fnrm4, f4.data(), F * 16)) != RZ_OK) return cleanup(st);
        // face tangents for normal maps, model space
        for (size_t f = 0; f < F; ++f) {
            const rz_triangle& t = d->faces[f];
            float tc[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
            for (int k = 0; k < 3; ++k)
                if (t.uvs[k] >= 0) {
                    tc[k][0] = d->uvs[2 * (size_t)t.uvs[k]];
                    tc[k][1] = d->uvs[2 * (size_t)t.uvs[k] + 1];
                }
            float tg[3];
            face_tangent(d->vertices + 3 * (size_t)t.vertices[0], d->vertices + 3 * (size_t)t.vertices[1],
                         d->vertices + 3 * (size_t)t.vertices[2], tc[0], tc[1], tc[2], tg);
            fx[f] = tg[0]; fy[f] = tg[1]; fz[f] = tg[2];
        }
        if ((st = upload(ctx, m, &m->ftx, fx.data(), F * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->fty, fy.data(), F * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->ftz, fz.data(), F * 4)) != RZ_OK) return cleanup(st);
    }
    {
        std::vector<float> sx(N), sy(N), sz(N);
        for (size_t i = 0; i < N; ++i) {
            const size_t o = (size_t)norder[i];
            sx[i] = d->normals[3 * o];
            sy[i] = d->normals[3 * o + 1];
            sz[i] = d->normals[3 * o + 2];
        }
        if ((st = upload(ctx, m, &m->nx, sx.data(), N * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->ny, sy.data(), N * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->nz, sz.data(), N * 4)) != RZ_OK) return cleanup(st);
        std::vector<float4> n4(N);
        for (size_t i = 0; i < N; ++i) n4[i] = make_float4(sx[i], sy[i], sz[i], 0.f);
        if ((st = upload(ctx, m, &m->nrm4, n4.data(), N * 16)) != RZ_OK) return cleanup(st);
    }
    {
        std::vector<float> uv(2 * (size_t)m->nt);
        for (size_t i = 0; i < (size_t)m->nt; ++i) {
            uv[2 * i] = d->uvs[2 * (size_t)torder[i]];
            uv[2 * i + 1] = d->uvs[2 * (size_t)torder[i] + 1];
        }
        if ((st = upload(ctx, m, &m->uvs, uv.data(), (size_t)m->nt * 8)) != RZ_OK) return cleanup(st);
    }
    {
        std::vector<int4> fv(F), fa(2 * F);
        for (size_t i = 0; i < F; ++i) {
            fv[i] = faces[i].q0;
            fa[2 * i] = faces[i].q1;
            fa[2 * i + 1] = faces[i].q2;
        }
        if ((st = upload(ctx, m, &m->bin_face_v, fv.data(), 16 * F)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->bin_face_attr, fa.data(), 32 * F)) != RZ_OK) return cleanup(st);
        // (renumbered indices: a mesh whose faces all share one normal, uv and material)
        bool same = F > 0;
        for (size_t i = 0; i < F && same; ++i) {
            const int4 a = faces[i].q1, b = faces[i].q2, a0 = faces[0].q1, b0 = faces[0].q2;
            same = a.x == a0.x && a.y == a0.x && a.z == a0.x && a.w == a0.w && b.x == b0.x && b.y == b0.x &&
                   b.z == b0.x;
        }
        if (same) {
            m->uniform = 1;
            m->uniform_n = faces[0].q1.x;
            m->uniform_t = faces[0].q2.x;
            m->uniform_m = faces[0].q1.w;
            if (m->uniform_n >= 0 && (size_t)m->uniform_n < N)
                for (int c = 0; c < 3; ++c) m->uniform_nrm[c] = d->normals[3 * (size_t)norder[m->uniform_n] + c];
        }
    }
    {
        // materials: RGBA8 texel pool plus one 256-entry normalisation table each
        std::vector<DevMaterial> dm(m->nm);
        std::vector<uint32_t> pool;
        std::vector<float> luts((size_t)m->nm * 256, 0.f);
        for (int32_t i = 0; i < m->nm; ++i) {
            const rz_material& src = d->materials[i];
            DevMaterial& o = dm[i];
            std::memset(&o, 0, sizeof o);
            o.diffuse[0] = src.diffuse_colour[0];
            o.diffuse[1] = src.diffuse_colour[1];
            o.diffuse[2] = src.diffuse_colour[2];
            o.lut_offset = (uint32_t)(256 * i);
            if (src.normal_map_channels != 0) {
                const size_t n = (size_t)src.normal_map_width * src.normal_map_height;
                if (pool.size() + n > 0xFFFFFFFFull)
                    return cleanup(fail(ctx, RZ_ERR_INVALID_ARG, "texture pool above 2^32 texels"));
                o.nmap_w = src.normal_map_width;
                o.nmap_h = src.normal_map_height;
                o.nmap_offset = (uint32_t)pool.size();
                const int C = src.normal_map_channels;
                pool.resize(pool.size() + n);
                uint32_t* dst = pool.data() + o.nmap_offset;
                for (size_t t = 0; t < n; ++t) {
                    const uint8_t* s = src.normal_map + t * C;
                    dst[t] = (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16);
                }
            }
            if (src.tex_channels == 0) continue;
            const size_t n = (size_t)src.tex_width * src.tex_height;
            if (pool.size() + n > 0xFFFFFFFFull)
                return cleanup(fail(ctx, RZ_ERR_INVALID_ARG, "texture pool above 2^32 texels"));
            o.tex_w = src.tex_width;
            o.tex_h = src.tex_height;
            o.tex_offset = (uint32_t)pool.size();
            const int C = src.tex_channels;
            normalisation_table(src.texels, n * C, luts.data() + 256 * (size_t)i);
            pool.resize(pool.size() + n);
            uint32_t* dst = pool.data() + o.tex_offset;
            for (size_t t = 0; t < n; ++t) {
                const uint8_t* s = src.texels + t * C;
                dst[t] = (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16) |
                         ((uint32_t)(C == 4 ? s[3] : 0) << 24);
            }
        }
        if ((st = upload(ctx, m, &m->mats, dm.data(), dm.size() * sizeof(DevMaterial))) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->texels, pool.data(), pool.size() * 4)) != RZ_OK) return cleanup(st);
        if ((st = upload(ctx, m, &m->luts, luts.data(), luts.size() * 4)) != RZ_OK) return cleanup(st);
    }
    *out = m;
    return RZ_OK;
}

void rz_mesh_destroy(rz_ctx* ctx, rz_mesh* m) {
    if (!m) return;
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
}

rz_status rz_time_passes(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args* args,
                         uint8_t* frame_device, float* depth_device, uint32_t flags, int32_t reps, float us[6]) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (!frame_device || !depth_device || !us || reps < 1 || !(flags & RZ_FLAG_CLEARED))
        return fail(ctx, RZ_ERR_INVALID_ARG, "rz_time_passes: device buffers, reps >= 1 and RZ_FLAG_CLEARED required");
    if ((flags & RZ_FLAG_MLP_SHADE) && !ctx->mlp_ready)
        return fail(ctx, RZ_ERR_INVALID_ARG, "RZ_FLAG_MLP_SHADE without network weights");
    // the frame with each generation of counter lines (the bin launches alternate)
    FrameLaunch L2[2];
    rz_status st = RZ_OK;
    for (int g = 0; g < 2; ++g) {
        if ((st = prepare_frame(ctx, mesh, lights, nl, args, flags, L2[g], g)) != RZ_OK) return st;
        bind_outputs(L2[g], true, frame_device, depth_device, nullptr);
    }
    const int32_t ntiles = L2[0].lay.ntiles;
    hipStream_t s = ctx->stream;
    // one graph of `reps` launches of the pass(es) `enqueue_one(k)` issues; ms of the
    // second launch.  `ready` runs before each launch (outside the timed region).
    auto timed = [&](auto enqueue_one, auto ready, float& ms) -> rz_status {
        hipGraphExec_t exec = nullptr;
        rz_status r = capture(ctx, reps, [&](int32_t k, hipStream_t cs) { return enqueue_one(k, cs); }, &exec);
        if (r != RZ_OK) return r;
        hipError_t e = hipSuccess;
        r = ready();
        if (r == RZ_OK) e = hipGraphLaunch(exec, s);
        if (r == RZ_OK && e == hipSuccess) r = ready();
        if (r == RZ_OK && e == hipSuccess) e = hipEventRecord(ctx->ev[0], s);
        if (r == RZ_OK && e == hipSuccess) e = hipGraphLaunch(exec, s);
        if (r == RZ_OK && e == hipSuccess) e = hipEventRecord(ctx->ev[5], s);
        if (r == RZ_OK && e == hipSuccess) e = hipStreamSynchronize(s);
        if (r == RZ_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[5]);
        (void)hipGraphExecDestroy(exec);
        if (r != RZ_OK) return r;
        if (e != hipSuccess) return fail(ctx, RZ_ERR_HIP, "rz_time_passes: %s", hipGetErrorString(e));
        return RZ_OK;
    };
    auto nothing = [&]() -> rz_status { return RZ_OK; };
    // bin launches k = 0, 1, ... use generation k % 2, from a ready generation 0; after
    // `reps` of them generation reps % 2 is the ready one
    auto bin = [&](int32_t k, hipStream_t cs) -> rz_status {
        RZ_HIP(ctx, launch_bin(L2[k & 1].bp, cs));
        return RZ_OK;
    };
    auto bin_ready = [&]() -> rz_status {
        rz_status r = counts_ready(ctx, ntiles, 0);
        ctx->count_gen = reps & 1;   // (as it stands after the launch that follows)
        return r;
    };
    float ms_b = 0, ms_r = 0, ms_s = 0, ms_c = 0;
    if ((st = timed(bin, bin_ready, ms_b)) != RZ_OK) return st;
    // one whole frame: raster, overflow and shade are idempotent on its binned state
    const int gen = ctx->count_gen;
    const FrameLaunch& L = L2[gen];
    if ((st = run_frame(ctx, L, false)) != RZ_OK) return st;
    // the same frame's shading with neither light loop nor network (kShadeChain)
    FrameLaunch Lc;
    if ((st = prepare_frame(ctx, mesh, lights, nl, args, flags, Lc, gen, kShadeChain)) != RZ_OK) return st;
    bind_outputs(Lc, true, frame_device, depth_device, nullptr);
    auto raster = [&](int32_t, hipStream_t cs) -> rz_status {
        if (L.lay.fused)
            RZ_HIP(ctx, launch_raster_shade(L.rp, L.sp, L.lay.ts, cs));
        else
            RZ_HIP(ctx, launch_raster(L.rp, L.lay.ts, cs));
        return RZ_OK;
    };
    auto shade = [&](int32_t, hipStream_t cs) -> rz_status {
        if (L.lay.fused)
            RZ_HIP(ctx, launch_shade_tiles(L.rp, tiles_params(L), L.lay.ts, cs));
        else
            RZ_HIP(ctx, launch_shade(L.sp, cs));
        return RZ_OK;
    };
    auto chain = [&](int32_t, hipStream_t cs) -> rz_status {
        if (Lc.lay.fused)
            RZ_HIP(ctx, launch_raster_shade(Lc.rp, Lc.sp, Lc.lay.ts, cs));
        else
            RZ_HIP(ctx, launch_shade(Lc.sp, cs));
        return RZ_OK;
    };
    // (the chain first: the passes timed after it leave the frame as the caller's
    // shading draws it)
    if ((st = timed(chain, nothing, ms_c)) != RZ_OK) return st;
    if ((st = timed(raster, nothing, ms_r)) != RZ_OK) return st;
    if (L.lay.fused && L.lay.tiles_grid == 0) {
        ms_s = 0.0f;   // no bin can overflow: the frame has no overflow-shade launch
    } else if ((st = timed(shade, nothing, ms_s)) != RZ_OK) {
        return st;
    }
    const float k = 1000.0f / (float)reps;
    us[0] = 0.0f;   // no vertex pass: the bin pass transforms each face's corners
    us[1] = ms_b * k;
    us[2] = ms_r * k;
    us[3] = 0.0f;   // the overflow loop runs inside the raster pass
    us[4] = ms_s * k;
    us[5] = ms_c * k;
    return RZ_OK;
}

rz_status rz_draw_frame(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args* args,
                        uint8_t* frame_buffer, float* depth_buffer, int32_t* id_buffer, uint32_t flags,
                        rz_frame_stats* stats) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    const bool cleared = (flags & RZ_FLAG_CLEARED) != 0;
    const bool on_device = (flags & RZ_FLAG_DEVICE_BUFFERS) != 0;
    FrameLaunch L;
    rz_status st = prepare_frame(ctx, mesh, lights, nl, args, flags, L, ctx->count_gen);
    if (st != RZ_OK) return st;
    if (!frame_buffer || !depth_buffer) return fail(ctx, RZ_ERR_INVALID_ARG, "null frame or depth buffer");
    const size_t np = L.lay.np;
    uint8_t* frame_d = frame_buffer;
    float* depth_d = depth_buffer;
    int32_t* ids_d = id_buffer;
    hipStream_t s = ctx->stream;
    if (!on_device) {
        if ((st = ensure(ctx, ctx->frame, np * 3)) != RZ_OK) return st;
        if ((st = ensure(ctx, ctx->depth, np * 4)) != RZ_OK) return st;
        frame_d = (uint8_t*)ctx->frame.ptr;
        depth_d = (float*)ctx->depth.ptr;
        if (id_buffer) {
            if ((st = ensure(ctx, ctx->ids, np * 4)) != RZ_OK) return st;
            ids_d = (int32_t*)ctx->ids.ptr;
        }
        if (!cleared) {
            RZ_HIP(ctx, hipMemcpyAsync(frame_d, frame_buffer, np * 3, hipMemcpyHostToDevice, s));
            RZ_HIP(ctx, hipMemcpyAsync(depth_d, depth_buffer, np * 4, hipMemcpyHostToDevice, s));
        }
    }
    bind_outputs(L, cleared, frame_d, depth_d, ids_d);
    if ((st = run_frame(ctx, L, ctx->profiling)) != RZ_OK) return st;
    if (!on_device) {
        RZ_HIP(ctx, hipMemcpyAsync(frame_buffer, frame_d, np * 3, hipMemcpyDeviceToHost, s));
        RZ_HIP(ctx, hipMemcpyAsync(depth_buffer, depth_d, np * 4, hipMemcpyDeviceToHost, s));
        if (id_buffer) RZ_HIP(ctx, hipMemcpyAsync(id_buffer, ids_d, np * 4, hipMemcpyDeviceToHost, s));
    }
    std::vector<uint32_t> list_counts, lists;
    if (stats) {
        RZ_HIP(ctx, hipMemcpyAsync(ctx->host_counters, L.counters, kNumCounters * 8, hipMemcpyDeviceToHost, s));
        ctx->host_partials.resize(L.lay.n_part);
        RZ_HIP(ctx, hipMemcpyAsync(ctx->host_partials.data(), L.part, L.lay.n_part * 4, hipMemcpyDeviceToHost, s));
        list_counts.resize(8 * kCountStride);
        lists.resize((size_t)8 * L.lay.ntiles);
        RZ_HIP(ctx, hipMemcpyAsync(list_counts.data(), L.bp.list_count, list_counts.size() * 4, hipMemcpyDeviceToHost, s));
        RZ_HIP(ctx, hipMemcpyAsync(lists.data(), L.bp.lists, lists.size() * 4, hipMemcpyDeviceToHost, s));
    }
    if (stats || !on_device || !(flags & RZ_FLAG_ASYNC)) RZ_HIP(ctx, hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        reduce_stats(L, ctx->host_counters, ctx->host_partials.data(), (uint64_t)mesh->nf, stats);
        count_listed(list_counts.data(), lists.data(), (size_t)L.lay.ntiles, &L.lay, 1, false, &stats->raster_tiles,
                     &stats->raster_pixels);
        if (ctx->profiling && (st = event_times(ctx, stats)) != RZ_OK) return st;
    }
    return RZ_OK;
}

rz_status rz_graph_create(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args* args,
                          uint8_t* frame_d, float* depth_d, int32_t* ids_d, uint32_t flags, int32_t frames,
                          float ry_step, rz_graph** out) {
    if (!ctx || !out) return RZ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!(flags & RZ_FLAG_DEVICE_BUFFERS))
        return fail(ctx, RZ_ERR_INVALID_ARG, "a frame graph needs device buffers (RZ_FLAG_DEVICE_BUFFERS)");
    if (frames < 1 || frames > 100000) return fail(ctx, RZ_ERR_INVALID_ARG, "frames %d not in [1, 100000]", frames);
    if (!frame_d || !depth_d) return fail(ctx, RZ_ERR_INVALID_ARG, "null frame or depth buffer");
    std::vector<FrameLaunch> launches(frames);
    rz_args a = args ? *args : rz_args{};
    // frame k uses generation k % 2 of the counter lines (rz_graph_launch readies 0)
    for (int32_t k = 0; k < frames; ++k) {
        rz_status st = prepare_frame(ctx, mesh, lights, nl, args ? &a : nullptr, flags, launches[k], k & 1);
        if (st != RZ_OK) return st;
        bind_outputs(launches[k], (flags & RZ_FLAG_CLEARED) != 0, frame_d, depth_d, ids_d);
        a.tait_bryan_angles[1] += ry_step;   // the spin loop's rotation about y (renderer.cpp:124)
    }
    rz_graph* gr = new rz_graph();
    gr->ctx = ctx;
    gr->frames = frames;
    gr->ntiles = launches[0].lay.ntiles;
    // lights beyond the kernel arguments: the graph's own copy (a later frame may rewrite
    // the context's); every frame has the same ones (the view matrix is fixed)
    rz_status st = stage_ext_lights(ctx, gr->lights_ext, launches[0], false);
    for (int32_t k = 1; k < frames && st == RZ_OK; ++k) launches[k].sp.lights_ext = launches[0].sp.lights_ext;
    if (st == RZ_OK)
        st = capture(ctx, frames, [&](int32_t k, hipStream_t s) { return enqueue_frame(ctx, launches[k], s, false); },
                     &gr->exec);
    if (st != RZ_OK) {
        rz_graph_destroy(gr);
        return st;
    }
    gr->buffer_gen = ctx->buffer_gen;
    *out = gr;
    return RZ_OK;
}

// ---- the spin loop as a replayed graph ---------------------------------------------

}  // extern "C"

namespace {

// The fields of a single frame's launch that depend on the model rotation: the camera and
// normal matrices of the bin pass's vertex transform, and the shade pass's normal matrix
// (tangents) and uniform camera-space normal.  fill_frame sets them the same way.
void set_frame_rotation(FrameLaunch& L, const rz_mesh* mesh, const rz_args& a) {
    rz_matrices mats;
    build_matrices(a, mats);
    std::memcpy(L.bp.camera, mats.camera, 64);
    std::memcpy(L.bp.normal, mats.normal, 64);
    std::memcpy(L.sp.normal_mat, mats.normal, 64);
    const float* m = mats.normal;
    const float x = mesh->uniform_nrm[0], y = mesh->uniform_nrm[1], z = mesh->uniform_nrm[2], w = 0.0f;
    L.sp.uniform_cn = make_float4((m[0] * x + m[4] * y) + (m[8] * z + m[12] * w),
                                  (m[1] * x + m[5] * y) + (m[9] * z + m[13] * w),
                                  (m[2] * x + m[6] * y) + (m[10] * z + m[14] * w), 0.0f);
}

}  // namespace

struct rz_spin {
    rz_ctx* ctx = nullptr;
    const rz_mesh* mesh = nullptr;
    rz_args args{};
    FrameLaunch L[2];                 // the frame with each generation of counter lines
    hipGraph_t graph[2] = {};
    hipGraphExec_t exec[2] = {};
    // captured kernel nodes whose arguments carry the rotation: the bin pass, the pass
    // that shades (raster+shade, or the separate shade pass), the overflow shade
    hipGraphNode_t bin[2] = {}, shade[2] = {}, tiles[2] = {};
    hipKernelNodeParams kbin[2] = {}, kshade[2] = {}, ktiles[2] = {};
    int next = 0;                     // generation of the next frame
    uint64_t buffer_gen = 0;
    DevBuf lights_ext;
};

namespace {

// The node the last launch on a capturing stream added.
rz_status captured_node(rz_ctx* ctx, hipStream_t s, hipGraphNode_t* node, hipKernelNodeParams* kp) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const hipGraphNode_t* deps = nullptr;
    size_t nd = 0;
    RZ_HIP(ctx, hipStreamGetCaptureInfo_v2(s, &cs, nullptr, nullptr, &deps, &nd));
    if (cs != hipStreamCaptureStatusActive || nd != 1 || !deps)
        return fail(ctx, RZ_ERR_HIP, "spin graph: unexpected capture state (%d, %zu dependencies)", (int)cs, nd);
    *node = deps[0];
    RZ_HIP(ctx, hipGraphKernelNodeGetParams(*node, kp));
    return RZ_OK;
}

// One frame's launches on a capturing stream, noting the nodes set_frame_rotation
// concerns.
rz_status capture_spin_frame(rz_ctx* ctx, rz_spin* sp, int g, hipStream_t s) {
    const FrameLaunch& L = sp->L[g];
    RZ_HIP(ctx, launch_bin(L.bp, s));
    rz_status st = captured_node(ctx, s, &sp->bin[g], &sp->kbin[g]);
    if (st != RZ_OK) return st;
    if (L.lay.fused) {
        RZ_HIP(ctx, launch_raster_shade(L.rp, L.sp, L.lay.ts, s));
    } else {
        RZ_HIP(ctx, launch_raster(L.rp, L.lay.ts, s));
        RZ_HIP(ctx, launch_shade(L.sp, s));
    }
    if ((st = captured_node(ctx, s, &sp->shade[g], &sp->kshade[g])) != RZ_OK) return st;
    if (L.lay.fused && L.lay.tiles_grid > 0) {
        RZ_HIP(ctx, launch_shade_tiles(L.rp, tiles_params(L), L.lay.ts, s));
        if ((st = captured_node(ctx, s, &sp->tiles[g], &sp->ktiles[g])) != RZ_OK) return st;
    }
    return RZ_OK;
}

}  // namespace

extern "C" {

rz_status rz_spin_create(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args* args,
                         uint8_t* frame_d, float* depth_d, int32_t* ids_d, uint32_t flags, rz_spin** out) {
    if (!ctx || !out) return RZ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!(flags & RZ_FLAG_DEVICE_BUFFERS))
        return fail(ctx, RZ_ERR_INVALID_ARG, "a spin graph needs device buffers (RZ_FLAG_DEVICE_BUFFERS)");
    if (flags & RZ_FLAG_MLP_SHADE) return fail(ctx, RZ_ERR_INVALID_ARG, "a spin graph shades with the light loop");
    if (!frame_d || !depth_d || !args) return fail(ctx, RZ_ERR_INVALID_ARG, "null frame, depth or args");
    rz_spin* sp = new rz_spin();
    sp->ctx = ctx;
    sp->mesh = mesh;
    sp->args = *args;
    rz_status st = RZ_OK;
    for (int g = 0; g < 2 && st == RZ_OK; ++g) {
        st = prepare_frame(ctx, mesh, lights, nl, args, flags, sp->L[g], g);
        if (st == RZ_OK) bind_outputs(sp->L[g], (flags & RZ_FLAG_CLEARED) != 0, frame_d, depth_d, ids_d);
    }
    // lights beyond the kernel arguments: the spin graph's own copy
    if (st == RZ_OK) st = stage_ext_lights(ctx, sp->lights_ext, sp->L[0], false);
    if (st == RZ_OK) sp->L[1].sp.lights_ext = sp->L[0].sp.lights_ext;
    for (int g = 0; g < 2 && st == RZ_OK; ++g) {
        hipStream_t s = ctx->stream;
        hipError_t e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            st = fail(ctx, RZ_ERR_HIP, "spin capture: %s", hipGetErrorString(e));
            break;
        }
        st = capture_spin_frame(ctx, sp, g, s);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(s, &graph);
        sp->graph[g] = graph;
        if (st == RZ_OK && e != hipSuccess) st = fail(ctx, RZ_ERR_HIP, "spin capture: %s", hipGetErrorString(e));
        if (st == RZ_OK) {
            e = hipGraphInstantiate(&sp->exec[g], graph, nullptr, nullptr, 0);
            if (e != hipSuccess) st = fail(ctx, RZ_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
    }
    if (st != RZ_OK) {
        rz_spin_destroy(sp);
        return st;
    }
    sp->buffer_gen = ctx->buffer_gen;
    *out = sp;
    return RZ_OK;
}

rz_status rz_spin_frame(rz_spin* sp, const float tait_bryan_angles[3]) {
    if (!sp || !tait_bryan_angles) return RZ_ERR_INVALID_ARG;
    rz_ctx* ctx = sp->ctx;
    if (sp->buffer_gen != ctx->buffer_gen)
        return fail(ctx, RZ_ERR_INVALID_ARG, "spin graph is stale: a later frame reallocated the context's buffers");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    const int g = sp->next;
    for (int k = 0; k < 3; ++k) sp->args.tait_bryan_angles[k] = tait_bryan_angles[k];
    FrameLaunch& L = sp->L[g];
    set_frame_rotation(L, sp->mesh, sp->args);
    // the captured kernels' arguments with the new rotation
    void* bin_args[] = {&L.bp};
    hipKernelNodeParams kp = sp->kbin[g];
    kp.kernelParams = bin_args;
    kp.extra = nullptr;
    RZ_HIP(ctx, hipGraphExecKernelNodeSetParams(sp->exec[g], sp->bin[g], &kp));
    void* fused_args[] = {&L.rp, &L.sp};
    void* shade_args[] = {&L.sp};
    kp = sp->kshade[g];
    kp.kernelParams = L.lay.fused ? fused_args : shade_args;
    kp.extra = nullptr;
    RZ_HIP(ctx, hipGraphExecKernelNodeSetParams(sp->exec[g], sp->shade[g], &kp));
    if (sp->tiles[g]) {
        ShadeParams tp = tiles_params(L);
        void* tiles_args[] = {&L.rp, &tp};
        kp = sp->ktiles[g];
        kp.kernelParams = tiles_args;
        kp.extra = nullptr;
        RZ_HIP(ctx, hipGraphExecKernelNodeSetParams(sp->exec[g], sp->tiles[g], &kp));
    }
    rz_status st = counts_ready(ctx, L.lay.ntiles, g);
    if (st != RZ_OK) return st;
    ctx->count_ntiles = -1;
    RZ_HIP(ctx, hipGraphLaunch(sp->exec[g], ctx->stream));
    ctx->count_ntiles = L.lay.ntiles;
    ctx->count_gen = g ^ 1;
    sp->next = g ^ 1;
    return RZ_OK;
}

void rz_spin_destroy(rz_spin* sp) {
    if (!sp) return;
    (void)hipSetDevice(sp->ctx->device);
    (void)hipStreamSynchronize(sp->ctx->stream);
    for (int g = 0; g < 2; ++g) {
        if (sp->exec[g]) (void)hipGraphExecDestroy(sp->exec[g]);
        if (sp->graph[g]) (void)hipGraphDestroy(sp->graph[g]);
    }
    if (sp->lights_ext.ptr) (void)hipFree(sp->lights_ext.ptr);
    delete sp;
}

rz_status rz_draw_text(rz_ctx* ctx, uint8_t* frame_device, uint32_t width, uint32_t height, int32_t x, int32_t y,
                       const char* text, const uint8_t foreground[3], const uint8_t background[3]) {
    if (!ctx || !frame_device || !text || !foreground || !background) return RZ_ERR_INVALID_ARG;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, launch_text(frame_device, width, height, x, y, text, foreground, background, ctx->stream));
    return RZ_OK;
}

rz_status rz_check_quotients(rz_ctx* ctx, uint64_t samples, uint32_t seed, uint64_t* compared,
                             uint64_t* mismatches) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (!compared || !mismatches || samples == 0 || samples > (1ull << 40))
        return fail(ctx, RZ_ERR_INVALID_ARG, "rz_check_quotients: 1 <= samples <= 2^40 and two outputs required");
    constexpr uint32_t kPerThread = 1024, kThreads = 256;
    const uint64_t blocks = (samples + (uint64_t)kPerThread * kThreads - 1) / ((uint64_t)kPerThread * kThreads);
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    unsigned long long* counts = nullptr;
    RZ_HIP(ctx, hipMalloc(&counts, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemsetAsync(counts, 0, sizeof h, ctx->stream);
    // grids of at most 65,536 workgroups, each thread its own 1,024 pairs
    for (uint64_t b0 = 0; e == hipSuccess && b0 < blocks; b0 += 65536) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(65536, blocks - b0);
        e = launch_check_quotients(nb, kPerThread, seed + (uint32_t)(b0 * kThreads * 0x632BE5ABull), counts,
                                   ctx->stream);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, counts, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(counts);
    if (e != hipSuccess) return fail(ctx, RZ_ERR_HIP, "rz_check_quotients: %s", hipGetErrorString(e));
    *compared = h[0];
    *mismatches = h[1];
    return RZ_OK;
}

rz_status rz_graph_launch(rz_graph* g) {
    if (!g) return RZ_ERR_INVALID_ARG;
    rz_ctx* ctx = g->ctx;
    if (!g->owns_buffers && g->buffer_gen != ctx->buffer_gen)
        return fail(ctx, RZ_ERR_INVALID_ARG, "graph is stale: a later frame reallocated the context's buffers");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    if (!g->owns_buffers) {   // a frame graph: its first frame uses generation 0
        rz_status st = counts_ready(ctx, g->ntiles, 0);
        if (st != RZ_OK) return st;
        ctx->count_ntiles = -1;
    }
    RZ_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    if (!g->owns_buffers) {
        ctx->count_ntiles = g->ntiles;
        ctx->count_gen = g->frames & 1;
    }
    return RZ_OK;
}

void rz_graph_destroy(rz_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->arena.ptr) (void)hipFree(g->arena.ptr);
    if (g->tables.ptr) (void)hipFree(g->tables.ptr);
    if (g->lights_ext.ptr) (void)hipFree(g->lights_ext.ptr);
    delete g;
}

rz_status rz_draw_batch(rz_ctx* ctx, const rz_batch_item* items, int32_t n, uint32_t flags, rz_frame_stats* stats) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    BatchLaunch B;
    rz_status st = prepare_any_batch(ctx, items, n, flags, ctx->batch_arena, ctx->batch_tables, B);
    if (st != RZ_OK) return st;
    const bool phase_clock = ctx->profiling && B.persistent;
    if (phase_clock) {
        // per phase: first start (min, from all ones) and last end (max, from zero)
        unsigned long long init[8];
        for (int k = 0; k < 4; ++k) {
            init[2 * k] = ~0ull;
            init[2 * k + 1] = 0ull;
        }
        B.q.clock = B.q.bar + kSyncWords + 1;
        RZ_HIP(ctx, hipMemcpy(B.q.clock, init, sizeof init, hipMemcpyHostToDevice));
    }
    if ((st = enqueue_batch(ctx, B, ctx->stream, ctx->profiling)) != RZ_OK) return st;
    if (stats) {
        if ((st = batch_stats(ctx, B, ctx->batch_arena, stats)) != RZ_OK) return st;
        if (ctx->profiling) {
            if (phase_clock) {
                // each phase as the span of its work (first start to last end: the phases
                // of different workgroups overlap around the barriers)
                unsigned long long clk[8];
                RZ_HIP(ctx, hipMemcpy(clk, B.q.clock, sizeof clk, hipMemcpyDeviceToHost));
                float* ms[4] = {&stats->ms_bin, &stats->ms_raster, &stats->ms_overflow, &stats->ms_shade};
                stats->ms_vertex = 0.0f;   // (the bin work transforms each face's corners)
                for (int k = 0; k < 4; ++k)
                    *ms[k] = clk[2 * k + 1] > clk[2 * k] ? (float)(clk[2 * k + 1] - clk[2 * k]) * 1e-5f : 0.0f;
                RZ_HIP(ctx, hipEventElapsedTime(&stats->ms_frame, ctx->ev[0], ctx->ev[5]));
            } else if ((st = event_times(ctx, stats)) != RZ_OK) {
                return st;
            }
        }
    } else if (!(flags & RZ_FLAG_ASYNC)) {
        RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (B.persistent) {
            uint32_t err = 0;
            RZ_HIP(ctx, hipMemcpy(&err, B.q.error, 4, hipMemcpyDeviceToHost));
            if (err) return fail(ctx, RZ_ERR_HIP, "batch frame: a barrier wait gave up");
        }
    }
    return RZ_OK;
}

rz_status rz_set_batch_persistent(rz_ctx* ctx, int32_t enable) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    ctx->batch_persistent = enable != 0;
    return RZ_OK;
}

int32_t rz_get_batch_persistent(const rz_ctx* ctx) { return ctx && ctx->batch_persistent ? 1 : 0; }

rz_status rz_batch_graph_create(rz_ctx* ctx, const rz_batch_item* items, int32_t n, uint32_t flags, int32_t frames,
                                rz_graph** out) {
    if (!ctx || !out) return RZ_ERR_INVALID_ARG;
    *out = nullptr;
    if (frames < 1 || frames > 100000) return fail(ctx, RZ_ERR_INVALID_ARG, "frames %d not in [1, 100000]", frames);
    rz_graph* gr = new rz_graph();
    gr->ctx = ctx;
    gr->frames = frames;
    gr->owns_buffers = true;
    BatchLaunch B;
    rz_status st = prepare_any_batch(ctx, items, n, flags, gr->arena, gr->tables, B);
    if (st == RZ_OK)
        st = capture(ctx, frames, [&](int32_t, hipStream_t s) { return enqueue_batch(ctx, B, s, false); },
                     &gr->exec);
    if (st != RZ_OK) {
        rz_graph_destroy(gr);
        return st;
    }
    *out = gr;
    return RZ_OK;
}

rz_status rz_autotune(rz_ctx* ctx, const rz_batch_item* items, int32_t n, uint32_t flags, int32_t* tile_size,
                      int32_t* bin_capacity, float* ms_per_frame) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    const int32_t keep_ts = ctx->tile_size, keep_cap = ctx->bin_cap;
    // bytes of bins the tuner may choose: a twelfth of the 288 GB of HBM (bins are only
    // touched where written, so a large capacity costs address space, not traffic)
    const size_t kBinBudget = (size_t)24 << 30;
    int best_ts = keep_ts, best_cap = keep_cap;
    float best_ms = 0.f;
    rz_status st = RZ_OK;
    for (int ts : {8, 16, 32}) {
        // one probe frame: the per-tile totals give the capacity that never overflows
        ctx->tile_size = ts;
        ctx->bin_cap = 64;
        BatchLaunch B;
        if ((st = prepare_any_batch(ctx, items, n, flags, ctx->batch_arena, ctx->batch_tables, B)) != RZ_OK) break;
        if ((st = enqueue_batch(ctx, B, ctx->stream, false)) != RZ_OK) break;
        uint32_t need = 1;
        size_t tiles = 0;
        std::vector<uint32_t> counts;
        for (const FrameLaunch& L : B.frames) {
            counts.resize(L.lay.ntiles);
            RZ_HIP(ctx, hipMemcpy2DAsync(counts.data(), 4, L.vp.tile_count, 4 * kCountStride, 4, counts.size(),
                                         hipMemcpyDeviceToHost, ctx->stream));
            RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (uint32_t c : counts) need = c > need ? c : need;
            tiles += L.lay.ntiles;
        }
        uint32_t cap = 16;
        while (cap < need && cap < (1u << 20)) cap <<= 1;
        while (cap > 16 && tiles * (size_t)cap * 4 > kBinBudget) cap >>= 1;   // overflow path beyond the budget
        ctx->bin_cap = (int)cap;
        if ((st = prepare_any_batch(ctx, items, n, flags, ctx->batch_arena, ctx->batch_tables, B)) != RZ_OK) break;
        for (int w = 0; w < 3 && st == RZ_OK; ++w) st = enqueue_batch(ctx, B, ctx->stream, false);
        if (st != RZ_OK) break;
        const int reps = 10;
        RZ_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
        for (int r = 0; r < reps && st == RZ_OK; ++r) st = enqueue_batch(ctx, B, ctx->stream, false);
        if (st != RZ_OK) break;
        RZ_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
        RZ_HIP(ctx, hipEventSynchronize(ctx->ev[5]));
        float ms = 0.f;
        RZ_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[5]));
        ms /= reps;
        if (best_ms == 0.f || ms < best_ms) {
            best_ms = ms;
            best_ts = ts;
            best_cap = (int)cap;
        }
    }
    ctx->tile_size = st == RZ_OK ? best_ts : keep_ts;
    ctx->bin_cap = st == RZ_OK ? best_cap : keep_cap;
    if (st != RZ_OK) return st;
    if (tile_size) *tile_size = best_ts;
    if (bin_capacity) *bin_capacity = best_cap;
    if (ms_per_frame) *ms_per_frame = best_ms;
    return RZ_OK;
}

// ---- learned shader ----------------------------------------------------------------

namespace {

rz_status mlp_buffers(rz_ctx* ctx) {
    rz_status st;
    if ((st = ensure(ctx, ctx->mlp_w, kMlpParams * 4)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->mlp_m, kMlpParams * 4)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->mlp_v, kMlpParams * 4)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->mlp_packed, kPackedBytes)) != RZ_OK) return st;
    return RZ_OK;
}

rz_status mlp_reset_adam(rz_ctx* ctx) {
    std::vector<float> z(kMlpParams, 0.f);
    RZ_HIP(ctx, hipMemcpyAsync(ctx->mlp_m.ptr, z.data(), kMlpParams * 4, hipMemcpyHostToDevice, ctx->stream));
    RZ_HIP(ctx, hipMemcpyAsync(ctx->mlp_v.ptr, z.data(), kMlpParams * 4, hipMemcpyHostToDevice, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->mlp_steps = 0;
    return RZ_OK;
}

}  // namespace

rz_status rz_mlp_set_weights(rz_ctx* ctx, const float* params) {
    if (!ctx || !params) return RZ_ERR_INVALID_ARG;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    rz_status st = mlp_buffers(ctx);
    if (st != RZ_OK) return st;
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipMemcpy(ctx->mlp_w.ptr, params, kMlpParams * 4, hipMemcpyHostToDevice));
    RZ_HIP(ctx, launch_mlp_pack((const float*)ctx->mlp_w.ptr, (uint16_t*)ctx->mlp_packed.ptr, ctx->stream));
    if ((st = mlp_reset_adam(ctx)) != RZ_OK) return st;
    ctx->mlp_ready = true;
    return RZ_OK;
}

rz_status rz_mlp_get_weights(rz_ctx* ctx, float* params) {
    if (!ctx || !params) return RZ_ERR_INVALID_ARG;
    if (!ctx->mlp_ready) return fail(ctx, RZ_ERR_INVALID_ARG, "no network weights");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RZ_HIP(ctx, hipMemcpy(params, ctx->mlp_w.ptr, kMlpParams * 4, hipMemcpyDeviceToHost));
    return RZ_OK;
}

rz_status rz_mlp_init(rz_ctx* ctx, uint64_t seed) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    // He-uniform weights from a SplitMix64 stream, zero biases
    std::vector<float> w(kMlpParams, 0.f);
    uint64_t state = seed;
    auto uniform = [&]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const struct { int off, n, fan_in; } layers[3] = {{kOffW1, 768, kMlpIn}, {kOffW2, 4096, kMlpHidden},
                                                       {kOffW3, 192, kMlpHidden}};
    for (const auto& L : layers) {
        const double r = std::sqrt(6.0 / L.fan_in);
        for (int i = 0; i < L.n; ++i) w[L.off + i] = (float)(r * uniform());
    }
    return rz_mlp_set_weights(ctx, w.data());
}

rz_status rz_mlp_collect(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl, const rz_args* args,
                         float* samples_device, int64_t capacity, int64_t* count) {
    if (!ctx || !samples_device || capacity < 0 || !count) return RZ_ERR_INVALID_ARG;
    FrameLaunch L;
    rz_status st = prepare_frame(ctx, mesh, lights, nl, args, RZ_FLAG_CLEARED, L, ctx->count_gen);
    if (st != RZ_OK) return st;
    const size_t np = L.lay.np;
    if ((st = ensure(ctx, ctx->frame, np * 3)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->depth, np * 4)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->mlp_count, 8)) != RZ_OK) return st;
    bind_outputs(L, true, (uint8_t*)ctx->frame.ptr, (float*)ctx->depth.ptr, nullptr);
    const unsigned long long start = (unsigned long long)*count;
    RZ_HIP(ctx, hipMemcpyAsync(ctx->mlp_count.ptr, &start, 8, hipMemcpyHostToDevice, ctx->stream));
    L.sp.mode = kShadeCollect;
    L.sp.samples = samples_device;
    L.sp.sample_count = (unsigned long long*)ctx->mlp_count.ptr;
    L.sp.sample_cap = (unsigned long long)capacity;
    if ((st = run_frame(ctx, L, false)) != RZ_OK) return st;
    unsigned long long end = 0;
    RZ_HIP(ctx, hipMemcpyAsync(&end, ctx->mlp_count.ptr, 8, hipMemcpyDeviceToHost, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = (int64_t)(end < (unsigned long long)capacity ? end : (unsigned long long)capacity);
    return RZ_OK;
}

rz_status rz_mlp_train(rz_ctx* ctx, const float* samples_device, int64_t n, int32_t steps, int32_t batch, float lr,
                       uint64_t seed, float* losses) {
    if (!ctx || !samples_device || n <= 0 || steps < 1 || batch < 64) return RZ_ERR_INVALID_ARG;
    if (!ctx->mlp_ready) return fail(ctx, RZ_ERR_INVALID_ARG, "no network weights (rz_mlp_init first)");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    const int nblocks = batch / 64;
    rz_status st;
    if ((st = ensure(ctx, ctx->mlp_gpart, (size_t)nblocks * kMlpParams * 4)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->mlp_lpart, (size_t)nblocks * 4)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->mlp_loss, (size_t)steps * 4)) != RZ_OK) return st;
    for (int32_t k = 0; k < steps; ++k) {
        // cosine decay from lr to lr/100
        const float frac = (float)k / (float)steps;
        const float lr_k = lr * (0.01f + 0.99f * 0.5f * (1.0f + std::cos(3.14159265f * frac)));
        const uint32_t t = ++ctx->mlp_steps;
        const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)t));
        const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)t));
        RZ_HIP(ctx, launch_mlp_step(samples_device, n, (float*)ctx->mlp_w.ptr, (float*)ctx->mlp_m.ptr,
                                    (float*)ctx->mlp_v.ptr, (float*)ctx->mlp_gpart.ptr, (float*)ctx->mlp_lpart.ptr,
                                    (float*)ctx->mlp_loss.ptr + k, nblocks, seed, t, lr_k, c1, c2, ctx->stream));
    }
    RZ_HIP(ctx, launch_mlp_pack((const float*)ctx->mlp_w.ptr, (uint16_t*)ctx->mlp_packed.ptr, ctx->stream));
    if (losses)
        RZ_HIP(ctx, hipMemcpyAsync(losses, ctx->mlp_loss.ptr, (size_t)steps * 4, hipMemcpyDeviceToHost, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (losses)
        for (int32_t k = 0; k < steps; ++k) losses[k] /= (float)(nblocks * 64);
    return RZ_OK;
}

rz_status rz_mlp_infer(rz_ctx* ctx, const float* features_device, int64_t n, float* out_device) {
    if (!ctx || (n > 0 && (!features_device || !out_device)) || n < 0) return RZ_ERR_INVALID_ARG;
    if (!ctx->mlp_ready) return fail(ctx, RZ_ERR_INVALID_ARG, "no network weights");
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    RZ_HIP(ctx, launch_mlp_infer(features_device, n, (const uint16_t*)ctx->mlp_packed.ptr, out_device, ctx->stream));
    RZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RZ_OK;
}

rz_status rz_draw_frame_noalloc(rz_ctx* ctx, const rz_mesh* mesh, rz_light* lights, int32_t nl,
                                const rz_args* args, uint8_t* frame_device, float* depth_device, int32_t* id_device,
                                uint32_t flags) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (!ctx->scratch) return fail(ctx, RZ_ERR_HIP, "rz_draw_frame_noalloc: no scratch address for device %d", ctx->device);
    if (!frame_device || !depth_device) return fail(ctx, RZ_ERR_INVALID_ARG, "null frame or depth buffer");
    if (flags & RZ_FLAG_MLP_SHADE)
        return fail(ctx, RZ_ERR_INVALID_ARG, "rz_draw_frame_noalloc: RZ_FLAG_MLP_SHADE is not supported");
    if (nl > kMaxLights)
        return fail(ctx, RZ_ERR_INVALID_ARG,
                    "rz_draw_frame_noalloc: %d lights, at most %d (more need device memory of their own)", nl,
                    kMaxLights);
    rz_status st = check_frame(ctx, mesh, lights, nl, args);
    if (st != RZ_OK) return st;
    RZ_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t ntiles = 0, grid = 0;
    bool narrow = false;
    frame_shape(ctx, *args, kShadePlain, flags, ntiles, narrow, grid);
    // Working set in `scratch`: the layout of rz_draw_frame with a bin capacity that fits.
    // With the fused pipeline and room for every face in every bin no tile can overflow
    // and the visibility buffer is never touched, so it takes no space.
    const uint32_t nf = (uint32_t)(mesh->nf > 0 ? mesh->nf : 1);
    auto total_bytes = [&](uint32_t cap, FrameLayout& lay) {
        const bool can_overflow = cap < (uint32_t)mesh->nf;
        lay = layout_frame(mesh, *args, ctx->tile_size, (int)cap, ctx->fused, grid,
                           can_overflow ? ctx->persistent_blocks : 0u, false);
        if (ctx->fused && !can_overflow) lay.bytes[kBufVis] = 0;
        size_t t = 0;
        for (int b = 0; b < kNumBufs; ++b) t += align256(lay.bytes[b]);
        return t;
    };
    FrameLayout lay;
    uint32_t cap = ctx->bin_cap > 0 ? (uint32_t)ctx->bin_cap : nf;
    size_t need = total_bytes(cap, lay);
    if (need > (size_t)RZ_SCRATCH_BYTES && ctx->bin_cap == 0) {
        // automatic capacity: the largest that fits beside the other buffers
        const size_t rest = total_bytes(1, lay) - align256(lay.bytes[kBufBins]);
        const size_t per_entry = (size_t)lay.ntiles * 4;
        cap = rest + 256 < (size_t)RZ_SCRATCH_BYTES ? (uint32_t)std::min<size_t>(
                                                           ((size_t)RZ_SCRATCH_BYTES - rest - 256) / per_entry, nf)
                                                     : 0u;
        if (cap >= 1) need = total_bytes(cap, lay);
    }
    if (cap < 1 || need > (size_t)RZ_SCRATCH_BYTES)
        return fail(ctx, RZ_ERR_INVALID_ARG,
                    "rz_draw_frame_noalloc: a %ux%u frame of %d faces needs %zu bytes of working memory, more "
                    "than the %zu of the no-allocation path",
                    args->image_width, args->image_height, mesh->nf, need, (size_t)RZ_SCRATCH_BYTES);
    void* ptrs[kNumBufs];
    size_t at = 0;
    for (int b = 0; b < kNumBufs; ++b) {
        ptrs[b] = ctx->scratch + at;
        at += align256(lay.bytes[b]);
    }
    FrameLaunch L;
    fill_frame(mesh, lights, nl, *args, flags, lay, ptrs, L, 0);
    L.rp.narrow = narrow ? 1 : 0;
    L.sp.max_blocks = lay.tiles_grid;
    bind_outputs(L, (flags & RZ_FLAG_CLEARED) != 0, frame_device, depth_device, id_device);
    hipStream_t s = ctx->stream;
    // the array keeps the last call's contents: the frame's counter lines (generation 0)
    // start at zero, by a fill kernel
    BinParams z;
    std::memset(&z, 0, sizeof z);
    z.zero_count = count_lines(lay.ntiles);
    z.zero_lines = (uint32_t*)ptrs[kBufTiles];
    z.zero_blocks = (z.zero_count + kFrameBinBlock - 1) / kFrameBinBlock;
    RZ_HIP(ctx, launch_bin(z, s));
    if ((st = enqueue_frame(ctx, L, s, false)) != RZ_OK) return st;
    if (!(flags & RZ_FLAG_ASYNC)) RZ_HIP(ctx, hipStreamSynchronize(s));
    return RZ_OK;
}

rz_status rz_draw_frame_desc(rz_ctx* ctx, const rz_mesh_desc* desc, rz_light* lights, int32_t nl,
                             const rz_args* args, uint8_t* frame_buffer, float* depth_buffer, uint32_t flags) {
    if (!ctx) return RZ_ERR_INVALID_ARG;
    if (flags & RZ_FLAG_DEVICE_BUFFERS) flags &= ~RZ_FLAG_ASYNC;
    rz_mesh* m = nullptr;
    rz_status st = rz_mesh_create(ctx, desc, &m);
    if (st != RZ_OK) return st;
    st = rz_draw_frame(ctx, m, lights, nl, args, frame_buffer, depth_buffer, nullptr, flags, nullptr);
    rz_mesh_destroy(ctx, m);
    return st;
}

}  // extern "C"
