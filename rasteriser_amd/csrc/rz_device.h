// rz_device.h — data layouts and bit-exact arithmetic shared by librz's gfx950 kernels.
//
// Every expression below keeps the reference's operation order (cited per function)
// and the whole library is built with -ffp-contract=off, IEEE divide/sqrt and IEEE
// denormals, so each float result is bit-identical to the reference's x86-64 build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace rz {

// Lights carried in the shade pass's kernel arguments; a frame with more lights reads
// the rest from a device array (ShadeParams::lights_ext).
constexpr int kMaxLights = 64;
constexpr uint64_t kEmptyKey = ~0ull;           // visibility word of a pixel with no fragment
constexpr uint32_t kNoFace = 0xFFFFFFFFu;
constexpr int kBigTriangleTiles = 16;           // above this many tiles a face is binned by its whole wave
// Tile counters sit one per 128-byte line: device-scope atomics on one line serialise,
// so packing 32 counters per line would queue every face of 32 tiles behind each other.
constexpr uint32_t kCountStride = 32;

// Global counter slots (uint64), zero when a frame's bin pass starts.  Only the
// overflow-tile count is a device-wide atomic (it is rare and doubles as the flag
// the overflow pass tests); every other statistic is a per-workgroup partial sum
// written with a plain store and reduced only when a caller asks for statistics
// (same-address device-scope atomics serialise across the eight XCDs).
// Shading modes: the reference's light loop, the learned surrogate on the matrix cores,
// the light loop with every shaded fragment's network inputs and colour appended to a
// training buffer, and kShadeChain: the per-pixel work every mode shares (visibility
// word, face record, interpolation, texture and normal-map samples, the stores) with
// neither light loop nor network, colour 0 -- the timing reference that splits a shade
// pass's time into that chain and the light loop or the network.
enum ShadeMode : int { kShadePlain = 0, kShadeMlp = 1, kShadeCollect = 2, kShadeChain = 3 };

enum Counter : int {
    kCntOverflowTiles = 0,
    kCntBatchOverflow = 1,      // batches, in the first mesh's counters: overflowed tiles of all meshes
    kNumCounters = 8
};

// A batch's tile lists hold (mesh << kListMeshShift) | tile.
constexpr uint32_t kListMeshShift = 22;
constexpr uint32_t kListTileMask = (1u << kListMeshShift) - 1u;
constexpr int32_t kMaxBatchMeshes = 1 << (32 - kListMeshShift);

// Per-face setup record written by the bin pass (128 bytes, one per face):
//   r0 = (x0, y0, z0, area)            raster vertex 0, area = edge(v2, v0, v1)
//   r1 = (x1, y1, z1, bits(bbox_lo))   bbox_lo = tl.x | tl.y << 16
//   r2 = (x2, y2, z2, bits(bbox_hi))   bbox_hi = br.x | br.y << 16
//   r3 = (w0, w1, w2, 0)               1/clip_w per vertex
// and, for the shading of the pixels the face wins, its corners' camera-space normals
// (as the vertex transform gives them) and uvs ((0, 0) for a corner without one):
//   r4 = (n0.x, n0.y, n0.z, u0)
//   r5 = (n1.x, n1.y, n1.z, v0)
//   r6 = (n2.x, n2.y, n2.z, u1)
//   r7 = (v1, u2, v2, bits(material))
// The box is the clamped bbox cut to the pixels the inside test can pass.  A face that
// is culled, or whose box holds no pixel, has no record (its front-mask bit is clear).
struct FaceRecord {
    float4 r0, r1, r2, r3, r4, r5, r6, r7;
};

// Device copy of one material.
struct DevMaterial {
    float diffuse[3];
    int32_t tex_w;         // 0: no texture
    int32_t tex_h;
    uint32_t tex_offset;   // first texel (RGBA8 packed in a uint32) in the mesh texel pool
    uint32_t lut_offset;   // first entry of the 256-float normalisation table
    int32_t nmap_w;        // 0: no normal map
    int32_t nmap_h;
    uint32_t nmap_offset;  // first normal-map texel in the pool
    uint32_t pad[2];
};

struct VertexParams {
    float camera[16];
    float normal[16];
    const float* px; const float* py; const float* pz;   // model positions (SoA)
    const float* nx; const float* ny; const float* nz;   // normals to transform (SoA)
    float4* raster;                                      // out: (x_px, y_px, ndc_z, 1/w)
    float4* cnormal;                                     // out: camera-space normal (xyz)
    uint32_t* tile_count;
    unsigned long long* counters;
    int32_t nv, nn, ntiles;
    float Wf, Hf;
};

// A face's indices in the bin pass's order, in two arrays: (v0, v1, v2, face), and
// (n0, n1, n2, material), (t0, t1, t2, 0) with t = -1 for a corner without a uv.
struct BinFace {
    int4 q0, q1, q2;
};

struct BinParams {
    const int4* face_v;         // (v0, v1, v2, face) in binning order
    const int4* face_attr;      // (n0, n1, n2, material), (t0, t1, t2, 0) in binning order
    const float4* raster;
    const float4* cnormal;      // camera-space normals (flat: one per face)
    // Single frames have no vertex pass: each face transforms its own corners.  pos
    // (x, y, z, -): model positions (null: `raster` holds raster-space positions as
    // given); nrm (x, y, z, -): model normals (flat: one per face; null: `cnormal`).
    const float4* pos;
    const float4* nrm;
    float camera[16];
    float normal[16];
    float Wf, Hf;
    const float2* uvs;
    int32_t flat;
    int32_t gather;             // 0: every face has the same normal, uv and material (ShadeParams::uniform)
    FaceRecord* rec;
    uint32_t* tile_count;
    uint32_t* bins;
    unsigned long long* counters;
    uint32_t* part_front;       // per workgroup: front faces
    uint32_t* part_pairs;       // per workgroup: (front face, tile) pairs
    int32_t nf;
    uint32_t W, H;
    int32_t tiles_x;
    int32_t ts_log2;
    uint32_t cap;
    int32_t wind_cw;
    int32_t ntiles;
    // per-XCD lists of the tiles that received a face, `ntiles` entries apart (a batch's
    // meshes share one set: entries are list_tag | tile)
    uint32_t* lists;
    uint32_t* list_count;       // 8 counters, kCountStride apart
    uint32_t list_tag;
    // the visibility words, which the push that overflows a tile's bin clears for that
    // tile (the raster pass skips it and the overflow loop fills them)
    unsigned long long* ovf_vis;
    unsigned long long* batch_ovf;   // batches: also counts the overflowed tile here
    // the overflowed tiles (list_tag | tile), each at the index its count returned (the
    // batch's count for batches): the overflow shade visits these, not every listed tile
    uint32_t* ovf_list;
    // bit i: the face at binning position i faces the camera and its box holds a pixel
    // (one word per 64 positions; only such faces have their record written)
    unsigned long long* front_mask;
    // extra workgroups of the bin launch that clear buffers (clear_np = 0: none): the
    // visibility words (separate-shade pipeline) and, for the fused pipeline on cleared
    // buffers, frame 0, depth 1.0, ids -1
    uint8_t* clear_frame;
    float* clear_depth;
    int32_t* clear_ids;
    unsigned long long* clear_vis;
    size_t clear_np;
    uint32_t clear_blocks;
    // single frames: the other generation of the counter lines (tile counts, list counters,
    // global counters; see count_lines), zeroed by zero_blocks more workgroups for the
    // next frame while this one fills its own
    uint32_t* zero_lines;
    uint32_t zero_count;        // lines
    uint32_t zero_blocks;
};

struct RasterParams {
    const FaceRecord* rec;
    const uint32_t* tile_count;
    const uint32_t* bins;
    const float* depth_in;      // null when the depth buffer is known cleared to 1.0
    unsigned long long* vis;
    unsigned long long* counters;
    uint32_t* part_frag;        // per workgroup: fragments passing the inside test
    int32_t nf;
    uint32_t W, H;
    int32_t tiles_x, ntiles;
    int32_t ts_log2;
    uint32_t cap;
    uint32_t max_blocks;        // workgroups of the list-driven launch (a multiple of 8)
    int32_t narrow;             // 16-pixel tiles by 64-thread workgroups (frames with many tiles)
    int32_t pre_cleared;        // outputs cleared by the bin launch
    const uint32_t* lists;      // per-XCD lists of tiles with faces (see BinParams)
    const uint32_t* list_count;
    const uint32_t* ovf_list;   // overflowed tiles, counters[kCntOverflowTiles] of them
    // the overflow loop's faces: the bin pass's front mask (by binning position) and the
    // faces' indices in binning order (the face id is the fourth word)
    const unsigned long long* front_mask;
    const int4* face_v;
};

// A batch's raster and overflow-shade launches: the shared tile lists and the launches'
// per-workgroup statistics.
struct BatchLists {
    const uint32_t* lists;      // 8 lists, `stride` entries apart
    uint32_t* list_count;       // 8 counters, kCountStride apart (second word: work cursor)
    uint32_t stride;
    const unsigned long long* overflow;   // tiles overflowed over the whole batch
    const uint32_t* ovf_list;   // those tiles, (mesh << kListMeshShift) | tile
    uint32_t* part_frag;        // per raster workgroup: fragments passing the inside test
    uint32_t* part_cov;         // per raster workgroup, then per overflow-shade workgroup: covered pixels
};

struct ShadeParams {
    const unsigned long long* vis;
    const FaceRecord* rec;
    const DevMaterial* mats;
    // uniform >= 0: every face shares normal uniform_n (camera space: uniform_cn, as the
    // vertex transform gives it), uv uniform_t (-1: none) and material uniform_m, and the
    // records hold no copies
    int32_t uniform, uniform_n, uniform_t, uniform_m;
    float4 uniform_cn;
    const float2* uvs;
    const uint32_t* texels;
    const float* luts;
    const float* ftx; const float* fty; const float* ftz;   // model-space face tangents
    uint8_t* frame;             // planar [3][H][W]
    float* depth;
    int32_t* ids;               // may be null
    uint32_t* part_cov;         // per workgroup: pixels with a winning fragment
    uint32_t W, H;
    int32_t cleared;
    int32_t wind_cw;
    int32_t flat;
    int32_t nl;
    int32_t nmat;
    int32_t mode;                    // ShadeMode
    const uint16_t* mlp;             // packed network (kShadeMlp)
    float* samples;                  // training samples out (kShadeCollect)
    unsigned long long* sample_count;
    unsigned long long sample_cap;
    float inv_p00, inv_p11;          // 1 / projection[0][0], 1 / projection[1][1] (view vector)
    uint32_t max_blocks;             // persistent strip kernel: workgroups to launch (0 = one per strip)
    int32_t pre_cleared;             // empty pixels already hold 0 / 1.0 / -1 (bin launch cleared them)
    float normal_mat[16];            // inverse-transpose of the model-view (tangents)
    float light_ic[kMaxLights][3];   // intensity * colour per light
    float light_nd[kMaxLights][3];   // -trans_dir per light
    // lights kMaxLights .. nl-1 (nl > kMaxLights only): (intensity * colour, -trans_dir),
    // six floats per light, in device memory
    const float* lights_ext;
};

// Launch shapes (shared by the launchers, the batch block maps and the host's
// statistics reduction).  `block`: workgroup size of a batch's vertex, bin and overflow
// work (256, or the persistent batch launch's batch_block).
__host__ __device__ inline uint32_t vertex_blocks(int32_t nv, int32_t nn, int32_t ntiles, int block = 256) {
    int32_t n = nv > nn ? nv : nn;
    if (ntiles > n) n = ntiles;
    if (kNumCounters > n) n = kNumCounters;
    return (uint32_t)((n + block - 1) / block);
}
__host__ __device__ inline uint32_t bin_blocks(int32_t nf, int block = 256) {
    return (uint32_t)((nf + block - 1) / block);
}
// Workgroup size of a batch's persistent launch (k_batch_frame): one thread per pixel of
// an 8-pixel tile, 256 on larger tiles.
__host__ __device__ inline int batch_block(int tile_size) { return tile_size == 8 ? 64 : 256; }
// A single frame's bin launch: workgroups of kFrameBinBlock threads (batches: 256).
constexpr int kFrameBinBlock = 128;
__host__ __device__ inline uint32_t frame_bin_blocks(int32_t nf) {
    return (uint32_t)((nf + kFrameBinBlock - 1) / kFrameBinBlock);
}
// Counter lines of one generation (single frames): tile counts, then the 8 list counters,
// then one line holding the kNumCounters global counters.
__host__ __device__ inline uint32_t count_lines(int32_t ntiles) { return (uint32_t)ntiles + 9u; }
__host__ __device__ inline uint32_t overflow_blocks(int32_t nf, int block = 256) {
    const uint32_t b = bin_blocks(nf, block), most = 1024u * 256u / (uint32_t)block;
    return b > most ? most : b;
}

// The persistent batch launch (k_batch_frame): G workgroups, all resident at once (G is
// the launch's occupancy times the CUs), meet at three barriers.  A frame: phase 1, every
// mesh's bin work (each face's corners transformed in place, as a single frame's bin pass
// does), the zeroing of the other generation of counter lines (each mesh keeps two, as
// single frames do) and the buffer clears, taken one item at a time from a cursor;
// barrier 1; phase 2, raster and shade of the tiles of all meshes from per-list cursors;
// phase 3, the overflow loop of overflowed tiles (static striding); barrier 2; phase 4,
// their shade; barrier 3 (arrival only: the next frame's launch reads it).  A barrier is
// 8 counters, one per XCD (workgroup b counts on counter b % 8), each on its own 128-byte
// line; they only grow: frame f's workgroups wait for (f + 1) * G / 8 on each, and learn
// f from the third barrier's counters at the start (every workgroup of frame f - 1 has
// arrived there, and none of frame f can before all of them passed barrier 1).  The
// counters and the cursor start at zero when the batch is prepared; workgroup 0 resets
// the cursor once barrier 1 is passed, when every take of the frame is done.
struct BatchSync {
    unsigned long long* bar;     // 3 barriers x 8 counters, kCountStride / 2 words apart
    uint32_t* error;             // nonzero: a barrier wait gave up (a guard; never set when all G are resident)
    uint32_t G;                  // workgroups of the launch (a multiple of 8)
    uint32_t nB;                 // phase 1 items (bin, zero and clear workgroups of every mesh)
    const uint2* bmap;           // phase 1 item -> (mesh, workgroup of the mesh's bin launch)
    uint32_t nmesh;              // meshes (the parameter tables hold generation 0, then generation 1)
    unsigned long long* clock;   // null, or per phase (bin, raster, overflow, shade) the first start (min)
                                 // and the last end (max) of its work, 100 MHz
    uint32_t* cursor;            // phase 1's item cursor (zero at a frame's start; reset after barrier 0)
};

__host__ __device__ inline uint32_t shade_blocks_x(uint32_t W) { return (W + 63) / 64; }
__host__ __device__ inline uint32_t shade_blocks_y(uint32_t H) { return (H + 3) / 4; }

// XCD-aware placement of a 2-D grid of work units.  Workgroups are dealt to the eight
// XCDs round robin (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Units
// are grouped into 4x4 (raster tiles) or 1x16 (shade strips) groups of 16; group g runs
// on XCD g % 8, its 16 units as consecutive workgroups of that XCD, so neighbouring
// units share one L2 while neighbouring groups spread over all XCDs.
struct XcdGroups {
    uint32_t gw, gh;        // group shape in units (gw * gh == 16)
    uint32_t nx, ny;        // units
    __host__ __device__ uint32_t groups_x() const { return (nx + gw - 1) / gw; }
    __host__ __device__ uint32_t groups() const { return groups_x() * ((ny + gh - 1) / gh); }
    // workgroups to launch: every XCD gets the same number of whole groups
    __host__ __device__ uint32_t grid() const { return (groups() + 7) / 8 * 8 * 16; }
    // unit index (y * nx + x) of workgroup b, or -1 for padding
    __host__ __device__ int unit(uint32_t b) const {
        const uint32_t xcd = b & 7u, k = b >> 3, g = (k >> 4) * 8u + xcd, t = k & 15u;
        if (g >= groups()) return -1;
        const uint32_t gx = groups_x();
        const uint32_t x = (g % gx) * gw + t % gw, y = (g / gx) * gh + t / gw;
        if (x >= nx || y >= ny) return -1;
        return (int)(y * nx + x);
    }
};

// ---------------------------------------------------------------------------------
// Arithmetic (host+device so host-side unit checks can call the same code).

// GLM 0.9.7 scalar min / max (func_common.inl).
__host__ __device__ __forceinline__ float gmin(float x, float y) { return x < y ? x : y; }
__host__ __device__ __forceinline__ float gmax(float x, float y) { return x > y ? x : y; }
__host__ __device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

// (unsigned int)f as GCC compiles it for x86-64 (cvttss2si to 64 bits, low half).
__host__ __device__ __forceinline__ uint32_t cast_u32(float f) {
    if (f >= 0.0f && f < 4294967296.0f) return (uint32_t)f;
    if (!(f > -9223372036854775808.0f && f < 9223372036854775808.0f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)f;
}

// drawing.cpp:36-39: (b.x-a.x)*(p.y-a.y) - (b.y-a.y)*(p.x-a.x)
__host__ __device__ __forceinline__ float edge(float px, float py, float ax, float ay, float bx, float by) {
    return (bx - ax) * (py - ay) - (by - ay) * (px - ax);
}

// geometry.cpp:76-83, left to right.
__host__ __device__ __forceinline__ float signed_area(float x0, float y0, float x1, float y1, float x2, float y2) {
    return -0.5f * (x0 * y1 - x1 * y0 + x1 * y2 - x2 * y1 + x2 * y0 - x0 * y2);
}

// Order-preserving map of a non-NaN float to uint32 with -0 and +0 sharing a key.
__host__ __device__ __forceinline__ uint32_t depth_order(float z) {
    uint32_t b = __float_as_uint(z);
    if ((b & 0x7FFFFFFFu) == 0) b = 0;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__host__ __device__ __forceinline__ unsigned long long vis_key(float z, uint32_t face) {
    return ((unsigned long long)depth_order(z) << 32) | face;
}

// Clamped inclusive bbox of drawing.cpp:77-93 (GLM clamp, ceil, unsigned cast).
__host__ __device__ __forceinline__ void bounding_box(float x0, float y0, float x1, float y1, float x2, float y2,
                                                      uint32_t W, uint32_t H, uint32_t& tlx, uint32_t& tly,
                                                      uint32_t& brx, uint32_t& bry) {
    const float bx = (float)(W - 1), by = (float)(H - 1);
    const float mnx = gmin(gmin(x0, x1), x2), mny = gmin(gmin(y0, y1), y2);
    const float mxx = gmax(gmax(x0, x1), x2), mxy = gmax(gmax(y0, y1), y2);
    tlx = cast_u32(gclamp(mnx, 0.f, bx));
    tly = cast_u32(gclamp(mny, 0.f, by));
    brx = cast_u32(gclamp(ceilf(mxx), 0.f, bx));
    bry = cast_u32(gclamp(ceilf(mxy), 0.f, by));
}

}  // namespace rz
