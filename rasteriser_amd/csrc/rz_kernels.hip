// rz_kernels.hip — gfx950 kernels of the frame path.
//
//   k_bin       per face: the vertex transform of its corners, clip = camera*(p,1),
//               z-divide, NDC->raster (geometry.cpp:44-74, 85-95, 110-116, 135-142),
//               and of its normals (geometry.cpp:35-42, 97-108); backface test
//               (geometry.cpp:76-83, drawing.cpp:176-178), clamped bbox
//               (drawing.cpp:77-93) cut to the triangle's extent, per-face setup and
//               shading record, appends the face to the bin of every tile its cut box
//               meets (fixed capacity per tile)
//               and each tile's first face puts the tile on a per-XCD list; more
//               workgroups zero the next frame's generation of counter lines.
//               (Batches run the transform as a vertex pass of their own,
//               k_vertex_batch, which also clears the meshes' counters.)
//   k_raster    resident grid over the tile lists, one tile per workgroup at a time:
//               expands (face, box pixel) candidates with wavefront prefix scans,
//               runs the inside and depth tests of drawing.cpp:96-121 and keeps per
//               pixel the minimum 64-bit (ordered depth << 32 | face) key in LDS;
//               writes the visibility words; then its share of the faces of tiles
//               whose bin overflowed, with global 64-bit atomicMin.
//   k_raster_shade  k_raster, shading each tile from LDS (the default pipeline);
//               k_shade_tiles shades the overflowed tiles afterwards.
//   k_shade     deferred shade of the winning fragment: depth, perspective-correct
//               normal/uv, texture, Lambert (drawing.cpp:121-146, shading.cpp:20-34,
//               material.cpp:11-26, CImg _linear_atXY).
//   k_*_batch   the same passes over many meshes per launch.
//
// The packed key makes the winner the lexicographic minimum of (depth, face index)
// over covered fragments with depth below the initial depth: the same fragment the
// reference's in-order strict `<` test (drawing.cpp:119, faces in file order) leaves.
#include <algorithm>

#include "rz_device.h"
#include "rz_mlp.h"

namespace rz {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// A read the compiler issues where it is written, ahead of the caller's staging loads,
// instead of sinking it to its first use (a relaxed atomic load: no ordering, no wait).
template <typename T>
__device__ __forceinline__ T issue_now(const T* q) {
    return __hip_atomic_load(const_cast<T*>(q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// op over the wave's 64 lanes, all of them active, returned to every lane: the row
// shifts and row broadcasts of wave_scan_incl (below) with `identity` where a shift
// leaves its row, then lane 63's value.  Six DPP operations, no LDS round trip.
template <typename Op>
__device__ __forceinline__ int wave_reduce(int x, int identity, Op op) {
    x = op(x, __builtin_amdgcn_update_dpp(identity, x, 0x111, 0xf, 0xf, false));   // row_shr:1
    x = op(x, __builtin_amdgcn_update_dpp(identity, x, 0x112, 0xf, 0xf, false));   // row_shr:2
    x = op(x, __builtin_amdgcn_update_dpp(identity, x, 0x114, 0xf, 0xf, false));   // row_shr:4
    x = op(x, __builtin_amdgcn_update_dpp(identity, x, 0x118, 0xf, 0xf, false));   // row_shr:8
    x = op(x, __builtin_amdgcn_update_dpp(identity, x, 0x142, 0xa, 0xf, false));   // row_bcast:15
    x = op(x, __builtin_amdgcn_update_dpp(identity, x, 0x143, 0xc, 0xf, false));   // row_bcast:31
    return __builtin_amdgcn_readlane(x, 63);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return (uint32_t)wave_reduce((int)v, 0, [](int a, int b) { return (int)((uint32_t)a + (uint32_t)b); });
}

// Inclusive prefix sum of v over the wave's 64 lanes, all of them active: shifts within
// each row of 16 lanes (a lane the shift would take past its row's start adds 0), then
// the totals of the rows below through the row broadcasts (lane 15 into rows 1 and 3,
// lane 31 into rows 2 and 3).  Six DPP adds, no LDS round trip.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);    // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);    // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);    // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);    // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);   // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return (uint32_t)x;
}

// Sum of v over the workgroup, stored by thread 0 (no atomics).  Every thread of the
// workgroup must call it.
__device__ __forceinline__ void store_block_sum(uint32_t v, uint32_t* out, int tid, int nthreads) {
    __shared__ uint32_t s_part[16];
    v = wave_sum(v);
    if ((tid & 63) == 0) s_part[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int w = 0; w < (nthreads >> 6); ++w) t += s_part[w];
        *out = t;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------

// A model position to (x_px, y_px, ndc_z, 1/w): GLM mat4*vec4 (pairwise) with w = 1,
// perspective divide, viewport (geometry.cpp, drawing.cpp:20-40).
__device__ __forceinline__ float4 raster_vertex(const float* m, float x, float y, float z, float Wf, float Hf) {
    const float cx = (m[0] * x + m[4] * y) + (m[8] * z + m[12] * 1.0f);
    const float cy = (m[1] * x + m[5] * y) + (m[9] * z + m[13] * 1.0f);
    const float cz = (m[2] * x + m[6] * y) + (m[10] * z + m[14] * 1.0f);
    const float cw = (m[3] * x + m[7] * y) + (m[11] * z + m[15] * 1.0f);
    const float nx = cx / cw, ny = cy / cw, nz = cz / cw, nw = 1.f / cw;
    float4 r;
    r.x = 0.5f * (nx + 1.0f) * Wf;
    r.y = 0.5f * (-ny + 1.0f) * Hf;
    r.z = nz;
    r.w = nw;
    return r;
}

// A model normal to camera space: the normal matrix times (x, y, z, 0), pairwise.
__device__ __forceinline__ float4 camera_normal(const float* m, float x, float y, float z) {
    const float w = 0.0f;
    float4 n;
    n.x = (m[0] * x + m[4] * y) + (m[8] * z + m[12] * w);
    n.y = (m[1] * x + m[5] * y) + (m[9] * z + m[13] * w);
    n.z = (m[2] * x + m[6] * y) + (m[10] * z + m[14] * w);
    n.w = 0.0f;
    return n;
}

// Batches: the vertex pass of every mesh (single frames transform in the bin pass).
template <int BLOCK = 256>
__device__ __forceinline__ void vertex_body(const VertexParams& p, int blk) {
    const int i = blk * BLOCK + threadIdx.x;
    if (i < p.nv)   // (raster-space meshes have nv = 0: the bin pass reads their vertices)
        p.raster[i] = raster_vertex(p.camera, p.px[i], p.py[i], p.pz[i], p.Wf, p.Hf);
    if (i < p.nn) p.cnormal[i] = camera_normal(p.normal, p.nx[i], p.ny[i], p.nz[i]);
    if (i < p.ntiles) p.tile_count[(size_t)i * kCountStride] = 0u;
    // (the last 8 entries are the tile lists' counters; their second word is a batch's
    // work cursor)
    if (i + 8 >= p.ntiles && i < p.ntiles) p.tile_count[(size_t)i * kCountStride + 1] = 0u;
    if (i < kNumCounters) p.counters[i] = 0ull;
}

// ---------------------------------------------------------------------------------
// A tile's first face puts the tile on an XCD's list for the raster pass.  Tiles are
// dealt to XCDs by (tx + 3 ty) mod 8, so any screen region splits almost evenly (each
// XCD's list must fit its resident workgroups, or its last tiles start a round late).
__device__ __forceinline__ void list_tile(const BinParams& p, uint32_t tile) {
    if (!p.lists) return;
    const uint32_t tx = tile % (uint32_t)p.tiles_x, ty = tile / (uint32_t)p.tiles_x;
    const uint32_t x = (tx + 3u * ty) & 7u;
    const uint32_t k = atomicAdd(&p.list_count[x * kCountStride], 1u);
    p.lists[(size_t)x * p.ntiles + k] = p.list_tag | tile;
}

// Called once per tile, by the push that takes its bin past capacity.
__device__ __forceinline__ void tile_overflowed(const BinParams& p, uint32_t tile) {
    const unsigned long long k = atomicAdd(&p.counters[kCntOverflowTiles], 1ull);
    const unsigned long long kb = p.batch_ovf ? atomicAdd(p.batch_ovf, 1ull) : k;
    if (p.ovf_list) p.ovf_list[kb] = p.list_tag | tile;   // (a tile overflows once: kb < tiles)
    // the raster kernel's overflow loop merges this tile's fragments into these words
    const uint32_t ts = 1u << p.ts_log2;
    const uint32_t ox = (tile % (uint32_t)p.tiles_x) * ts, oy = (tile / (uint32_t)p.tiles_x) * ts;
    const uint32_t x1 = min(ox + ts, p.W), y1 = min(oy + ts, p.H);
    for (uint32_t y = oy; y < y1; ++y)
        for (uint32_t x = ox; x < x1; ++x) p.ovf_vis[(size_t)y * p.W + x] = kEmptyKey;
}

__device__ __forceinline__ void bin_push(const BinParams& p, uint32_t face, uint32_t tile) {
    const uint32_t slot = atomicAdd(&p.tile_count[(size_t)tile * kCountStride], 1u);
    if (slot == 0) list_tile(p, tile);
    if (slot < p.cap)
        p.bins[(size_t)tile * p.cap + slot] = face;
    else if (slot == p.cap)
        tile_overflowed(p, tile);
}

// A face's candidate rectangle is cut twice.  The bin pass cuts the face's whole box to
// the triangle's vertex extent (shrink_rect<kSlabNever>), once per face: the 256-mesh
// batch bins 14.0 M -> 11.4 M (face, tile) pairs, the 4M-triangle frame 2.65 M -> 2.0 M.
// The raster pass then cuts a rectangle (the cut box inside its tile) larger than SHRINK
// pixels to the triangle's part of the tile's slabs (shrink_rect<kSlabAlways>): every
// rectangle of two or more pixels in batches (kShrinkPixels), whose triangle soups make
// large rectangles the common case (8 and 32 measured no faster), and only rectangles
// larger than kSlabPixels in single frames (4M-triangle raster+shade pass 179 -> 166 us
// against every rectangle).  Each raster-pass cut costs about 300 instructions a wave;
// a candidate it removes costs about 113 per 64.
constexpr int kShrinkPixels = 1;
constexpr int kSlabPixels = 32;
constexpr int kSlabAlways = -1;
// The vertex-extent cut alone (the bin pass's cut of a face's whole box).
constexpr int kSlabNever = -2;

// Range of T ∩ {lo <= v <= hi} along the other axis (u), T the triangle (u_k, v_k):
// the vertices inside the slab and the edges' crossings of its two lines.  Returns false
// when the slab misses T.  The crossings use the hardware reciprocal; callers widen the
// range by a bound on that error.
__device__ __forceinline__ bool slab_range(float u0, float v0, float u1, float v1, float u2, float v2, float lo,
                                           float hi, float& umin, float& umax) {
    umin = INFINITY; umax = -INFINITY;
    auto vert = [&](float u, float v) {
        if (v >= lo && v <= hi) { umin = fminf(umin, u); umax = fmaxf(umax, u); }
    };
    auto cross = [&](float pu, float pv, float qu, float qv) {
        if (pv == qv) return;
        const float k = (qu - pu) * __builtin_amdgcn_rcpf(qv - pv);
        if ((pv - lo) * (qv - lo) <= 0.0f) { const float u = pu + (lo - pv) * k; umin = fminf(umin, u); umax = fmaxf(umax, u); }
        if ((pv - hi) * (qv - hi) <= 0.0f) { const float u = pu + (hi - pv) * k; umin = fminf(umin, u); umax = fmaxf(umax, u); }
    };
    vert(u0, v0); vert(u1, v1); vert(u2, v2);
    cross(u0, v0, u1, v1); cross(u1, v1, u2, v2); cross(u2, v2, u0, v0);
    return umin <= umax;
}

// Shrinks a face's candidate rectangle in its tile (tile-local [rx0, rx1] x
// [ry0, ry1], origin ox, oy) to the pixels that can pass the inside test; returns false
// when none can.  Bound used: the computed edge function differs from the exact one by
// at most 5.7 u L (D + 1) (u = 2^-24, L the edge's length, D the largest coordinate
// distance between a vertex and a rectangle pixel), so a passing pixel lies within
// m = 8.01 u (D + 1) of every edge's half-plane — when the area's sign is certain
// (|area| > 2 m L_max) — hence within g = m (1 + 3 D / r) of the triangle (r its
// inradius, >= |area| / 2 / perimeter): the offset triangle is the triangle scaled by
// 1 + m / r about its incentre, and no vertex is farther than 2 sqrt(2) D from it.
// (Lengths are bounded by |dx| + |dy|; the products' and the subtraction's roundings
// give 4 u (|A| |py - ay| + |B| |px - ax|) <= 5.7 u L (D + 1).)  So a passing pixel p has
// a triangle point t with |p - t| <= g: t lies in the slab of rows [Y0 - g, Y1 + g], and
// px is within g of that slab's part of the triangle.  The slabs are widened to
// w = g + d and the x (y) range taken as [ceil(min - w), floor(max + w)], where
// d = 2^-19 (C + 16), C the largest coordinate magnitude, exceeds the error of a computed
// edge crossing (pu + (lo - pv) (qu - pu) rcp(qv - pv): |error| <= 15.2 u C, the
// hardware reciprocal within 1 ulp) plus the roundings of the slab bounds and of
// min - w, max + w.  With SLAB >= 0 the rectangle is first cut to the triangle's own
// extent, [ceil(v_min - w), floor(v_max + w)] per axis (a point within g of the triangle
// lies within g of its vertices' range), and only a rectangle still larger than SLAB
// pixels takes the slab cut; kSlabNever stops after that cut, kSlabAlways skips it.
// Faces where g would exceed 4 pixels, or whose doubled area lies outside
// [2^-40, 2^40], keep their rectangle.  The rectangle may be a whole image (the bin
// pass): its corners stay below 65,536.
template <int SLAB>
__device__ __forceinline__ bool shrink_rect(const float4& a, const float4& b, const float4& c, int ox, int oy,
                                            int& rx0, int& rx1, int& ry0, int& ry1) {
    const float X0 = (float)(ox + rx0), X1 = (float)(ox + rx1), Y0 = (float)(oy + ry0), Y1 = (float)(oy + ry1);
    const float vx0 = fminf(fminf(a.x, b.x), c.x), vx1 = fmaxf(fmaxf(a.x, b.x), c.x);
    const float vy0 = fminf(fminf(a.y, b.y), c.y), vy1 = fmaxf(fmaxf(a.y, b.y), c.y);
    const float D = fmaxf(fmaxf(vx1 - X0, X1 - vx0), fmaxf(vy1 - Y0, Y1 - vy0));
    const float area = fabsf(a.w);
    // (comparisons false for NaN: such faces keep their rectangle)
    if (!(D <= 1048576.0f) || !(area >= 0x1p-40f) || !(area <= 0x1p40f)) return true;
    const float l0 = fabsf(b.x - a.x) + fabsf(b.y - a.y), l1 = fabsf(c.x - b.x) + fabsf(c.y - b.y);
    const float l2 = fabsf(a.x - c.x) + fabsf(a.y - c.y);
    const float m = 8.01f * 0x1p-24f * (D + 1.0f);
    if (!(area > 2.0f * m * fmaxf(fmaxf(l0, l1), l2) * 1.01f)) return true;
    // g = m (1 + 3 D / r) with r = area / (2 P) (P = l0 + l1 + l2 bounds the perimeter):
    // 6 D P / area through the hardware reciprocal (within 1 ulp), the factor 6.07 and the
    // final 1.01 covering its and the products' roundings
    const float g = m * (1.0f + 6.07f * D * (l0 + l1 + l2) * __builtin_amdgcn_rcpf(area)) * 1.01f;
    if (!(g <= 4.0f)) return true;
    const float C = fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(b.x), fabsf(b.y))),
                          fmaxf(fmaxf(fabsf(c.x), fabsf(c.y)), fmaxf(X1, Y1)));
    const float w = g + 0x1p-19f * (C + 16.0f);
    if constexpr (SLAB >= 0 || SLAB == kSlabNever) {
        // first the triangle's own extent (its vertices), which is all a small face needs
        rx0 = max(rx0, (int)ceilf(vx0 - w) - ox); rx1 = min(rx1, (int)floorf(vx1 + w) - ox);
        ry0 = max(ry0, (int)ceilf(vy0 - w) - oy); ry1 = min(ry1, (int)floorf(vy1 + w) - oy);
        if (rx0 > rx1 || ry0 > ry1) return false;
        if (SLAB == kSlabNever || (rx1 - rx0 + 1) * (ry1 - ry0 + 1) <= SLAB) return true;
    }
    float xmin, xmax, ymin, ymax;
    if (!slab_range(a.x, a.y, b.x, b.y, c.x, c.y, Y0 - w, Y1 + w, xmin, xmax)) return false;
    if (!slab_range(a.y, a.x, b.y, b.x, c.y, c.x, X0 - w, X1 + w, ymin, ymax)) return false;
    rx0 = max(rx0, (int)ceilf(xmin - w) - ox); rx1 = min(rx1, (int)floorf(xmax + w) - ox);
    ry0 = max(ry0, (int)ceilf(ymin - w) - oy); ry1 = min(ry1, (int)floorf(ymax + w) - oy);
    return rx0 <= rx1 && ry0 <= ry1;
}

// LDS bin_body needs (carved by the caller: kernels that run other phases share it):
// per-tile counts and bases of the union rectangle, each face's slots, the rectangle.
template <int BLOCK>
struct BinLds {
    static constexpr int kAggTiles = 4 * BLOCK;
    static constexpr int kBytes = 2 * kAggTiles * 4 + BLOCK * kBigTriangleTiles + 24;
};

template <int BLOCK>
__device__ __forceinline__ void bin_body(const BinParams& p, int blk, unsigned char* lds) {
    // faces are taken in the mesh's binning order (centroid Morton order), so the
    // workgroup's faces are screen neighbours; f is the face's own index
    const int i = blk * BLOCK + threadIdx.x;
    int f = p.nf;
    bool front = false;
    uint32_t tx0 = 1, tx1 = 0, ty0 = 1, ty1 = 0;   // empty tile rectangle
    if (i < p.nf) {
        // the face's attribute indices are read beside its vertex indices, and the shading
        // attributes beside the vertices, before the cull decides whether they are kept
        // (none for a mesh whose faces all share them): two dependent round trips, not three
        const int4 q0 = p.face_v[i];
        int4 q1 = make_int4(0, 0, 0, 0), q2 = q1;
        if (p.gather) {
            q1 = p.face_attr[2 * i];
            q2 = p.face_attr[2 * i + 1];
        }
        f = q0.w;
        // every read is issued before any transform waits on one
        const float4* vsrc = p.pos ? p.pos : p.raster;
        float4 a = vsrc[q0.x], b = vsrc[q0.y], c = vsrc[q0.z];
        float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
        float2 t0 = make_float2(0.f, 0.f), t1 = t0, t2 = t0;
        int32_t mat = 0;
        if (p.gather) {
            const float4* nsrc = p.nrm ? p.nrm : p.cnormal;
            n0 = nsrc[p.flat ? f : q1.x]; n1 = nsrc[p.flat ? f : q1.y]; n2 = nsrc[p.flat ? f : q1.z];
            if (q2.x >= 0) t0 = p.uvs[q2.x];
            if (q2.y >= 0) t1 = p.uvs[q2.y];
            if (q2.z >= 0) t2 = p.uvs[q2.z];
            mat = q1.w;
        }
        if (p.pos) {
            a = raster_vertex(p.camera, a.x, a.y, a.z, p.Wf, p.Hf);
            b = raster_vertex(p.camera, b.x, b.y, b.z, p.Wf, p.Hf);
            c = raster_vertex(p.camera, c.x, c.y, c.z, p.Wf, p.Hf);
        }
        if (p.gather && p.nrm) {
            n0 = camera_normal(p.normal, n0.x, n0.y, n0.z);
            n1 = p.flat ? n0 : camera_normal(p.normal, n1.x, n1.y, n1.z);
            n2 = p.flat ? n0 : camera_normal(p.normal, n2.x, n2.y, n2.z);
        }
        const float s = signed_area(a.x, a.y, b.x, b.y, c.x, c.y);
        front = (s > 0.0f) != (p.wind_cw != 0);
        FaceRecord* r = p.rec + f;
        uint32_t tlx = 1, tly = 1, brx = 0, bry = 0;
        if (front) bounding_box(a.x, a.y, b.x, b.y, c.x, c.y, p.W, p.H, tlx, tly, brx, bry);
        float area = 0.0f;
        if (front && tlx <= brx && tly <= bry) {
            // the box cut to the triangle's own extent, widened by shrink_rect's bound on
            // the inside test's rounding: every pixel that can pass the test stays in it,
            // and a face that can pass none keeps no pixel
            area = edge(c.x, c.y, a.x, a.y, b.x, b.y);
            int x0 = (int)tlx, x1 = (int)brx, y0 = (int)tly, y1 = (int)bry;
            if (shrink_rect<kSlabNever>(make_float4(a.x, a.y, a.z, area), b, c, 0, 0, x0, x1, y0, y1)) {
                tlx = (uint32_t)x0; brx = (uint32_t)x1; tly = (uint32_t)y0; bry = (uint32_t)y1;
            } else {
                tlx = tly = 1; brx = bry = 0;
            }
        }
        // a record only for a face with pixels to raster (a culled face, or one whose box
        // holds no pixel, is never read: the mask below leaves it out)
        if (front && tlx <= brx && tly <= bry) {
            r->r0 = make_float4(a.x, a.y, a.z, area);
            r->r1 = make_float4(b.x, b.y, b.z, __uint_as_float(tlx | (tly << 16)));
            r->r2 = make_float4(c.x, c.y, c.z, __uint_as_float(brx | (bry << 16)));
            r->r3 = make_float4(a.w, b.w, c.w, 0.0f);
            if (p.gather) {
                r->r4 = make_float4(n0.x, n0.y, n0.z, t0.x);
                r->r5 = make_float4(n1.x, n1.y, n1.z, t0.y);
                r->r6 = make_float4(n2.x, n2.y, n2.z, t1.x);
                r->r7 = make_float4(t1.y, t2.x, t2.y, __int_as_float(mat));
            }
            tx0 = tlx >> p.ts_log2; tx1 = brx >> p.ts_log2;
            ty0 = tly >> p.ts_log2; ty1 = bry >> p.ts_log2;
        }
    }
    {
        // the wave's 64 binning positions (blk * BLOCK is a multiple of 64): the faces
        // with a record
        const unsigned long long fm = __ballot(tx0 <= tx1 && ty0 <= ty1);
        if ((threadIdx.x & 63) == 0 && i < p.nf) p.front_mask[(uint32_t)i >> 6] = fm;
    }
    const uint32_t nt = (tx0 <= tx1 && ty0 <= ty1) ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0u;
    const bool big = nt > (uint32_t)kBigTriangleTiles;
    const bool small = nt > 0 && !big;

    // Small faces: aggregate the workgroup's (tile, face) pairs in LDS so that each
    // distinct tile costs one device-scope atomic per workgroup, all issued at once.
    // Faces of one workgroup are neighbours in the mesh, so their tiles usually fall in
    // a small rectangle; when the union of their tile rectangles is larger than
    // kAggTiles the pairs go to the tile counters one by one.
    constexpr int kAggTiles = BinLds<BLOCK>::kAggTiles;
    uint32_t* s_cnt = reinterpret_cast<uint32_t*>(lds);
    uint32_t* s_base = s_cnt + kAggTiles;
    uint8_t* s_slot = reinterpret_cast<uint8_t*>(s_base + kAggTiles);
    // union tile rect x0 y0 x1 y1, then the workgroup's front faces and (face, tile) pairs
    int* s_rect = reinterpret_cast<int*>(s_slot + BLOCK * kBigTriangleTiles);
    uint32_t* s_stat = reinterpret_cast<uint32_t*>(s_rect + 4);
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_rect[0] = s_rect[1] = 0x7FFFFFFF;
        s_rect[2] = s_rect[3] = -1;
        s_stat[0] = s_stat[1] = 0u;
    }
    __syncthreads();
    {
        // union of the small faces' tile rectangles and the statistics' partial sums: per
        // wave, then one LDS atomic each (one exchange for both)
        auto mn = [](int a, int b) { return min(a, b); };
        auto mx = [](int a, int b) { return max(a, b); };
        const int r0 = wave_reduce(small ? (int)tx0 : 0x7FFFFFFF, 0x7FFFFFFF, mn);
        const int r1 = wave_reduce(small ? (int)ty0 : 0x7FFFFFFF, 0x7FFFFFFF, mn);
        const int r2 = wave_reduce(small ? (int)tx1 : -1, INT_MIN, mx);
        const int r3 = wave_reduce(small ? (int)ty1 : -1, INT_MIN, mx);
        const uint32_t nfront = wave_sum(front ? 1u : 0u), npairs = wave_sum(nt);
        if ((tid & 63) == 0) {
            if (r2 >= 0) {
                atomicMin(&s_rect[0], r0); atomicMin(&s_rect[1], r1);
                atomicMax(&s_rect[2], r2); atomicMax(&s_rect[3], r3);
            }
            atomicAdd(&s_stat[0], nfront);
            atomicAdd(&s_stat[1], npairs);
        }
    }
    __syncthreads();
    if (tid == 0) {
        p.part_front[blk] = s_stat[0];
        p.part_pairs[blk] = s_stat[1];
    }
    const int ux0 = s_rect[0], uy0 = s_rect[1];
    const int uw = s_rect[2] - ux0 + 1, uh = s_rect[3] - uy0 + 1;
    const bool agg = s_rect[2] >= 0 && uw * uh <= kAggTiles;   // uniform over the workgroup
    if (agg) {
        const int area = uw * uh;
        for (int i = tid; i < area; i += BLOCK) s_cnt[i] = 0u;
        __syncthreads();
        if (small) {
            int k = 0;
            for (uint32_t ty = ty0; ty <= ty1; ++ty)
                for (uint32_t tx = tx0; tx <= tx1; ++tx, ++k)
                    s_slot[tid * kBigTriangleTiles + k] =
                        (uint8_t)atomicAdd(&s_cnt[((int)ty - uy0) * uw + ((int)tx - ux0)], 1u);
        }
        __syncthreads();
        for (int i = tid; i < area; i += BLOCK) {
            const uint32_t c = s_cnt[i];
            if (!c) continue;
            const uint32_t tile = (uint32_t)(uy0 + i / uw) * p.tiles_x + (uint32_t)(ux0 + i % uw);
            const uint32_t base = atomicAdd(&p.tile_count[(size_t)tile * kCountStride], c);
            s_base[i] = base;
            if (base == 0) list_tile(p, tile);
            if (base <= p.cap && base + c > p.cap) tile_overflowed(p, tile);
        }
        __syncthreads();
        if (small) {
            int k = 0;
            for (uint32_t ty = ty0; ty <= ty1; ++ty)
                for (uint32_t tx = tx0; tx <= tx1; ++tx, ++k) {
                    const int i = ((int)ty - uy0) * uw + ((int)tx - ux0);
                    const uint32_t slot = s_base[i] + s_slot[tid * kBigTriangleTiles + k];
                    if (slot < p.cap) p.bins[(size_t)(ty * p.tiles_x + tx) * p.cap + slot] = (uint32_t)f;
                }
        }
    } else if (small) {
        for (uint32_t ty = ty0; ty <= ty1; ++ty)
            for (uint32_t tx = tx0; tx <= tx1; ++tx) bin_push(p, (uint32_t)f, ty * p.tiles_x + tx);
    }
    // Faces meeting many tiles are binned by the whole wave, one face at a time.
    unsigned long long bm = __ballot(big);
    while (bm) {
        const int src = __ffsll((long long)bm) - 1;
        bm &= bm - 1;
        // (src is uniform: lane reads, not LDS permutes)
        const uint32_t bx0 = __builtin_amdgcn_readlane(tx0, src), bx1 = __builtin_amdgcn_readlane(tx1, src);
        const uint32_t by0 = __builtin_amdgcn_readlane(ty0, src), by1 = __builtin_amdgcn_readlane(ty1, src);
        const uint32_t bf = __builtin_amdgcn_readlane((uint32_t)f, src);
        const uint32_t w = bx1 - bx0 + 1, n = w * (by1 - by0 + 1);
        for (uint32_t k = lane_id(); k < n; k += 64) {
            const uint32_t ty = by0 + k / w, tx = bx0 + k % w;
            bin_push(p, bf, ty * p.tiles_x + tx);
        }
    }
}

// ---------------------------------------------------------------------------------
// CImg::_linear_atXY (CImg.h:13475-13491) of channels 0..2 of an RGBA8 image at
// (u*W, (1-v)*H) (material.cpp:19-22); decode(byte) gives the texel value.
template <typename Decode>
__device__ __forceinline__ void sample_bilinear(const uint32_t* __restrict__ t, int32_t Wt, int32_t Ht, float uvx,
                                                float uvy, Decode decode, float out[3]) {
    const float u = uvx * (float)Wt;
    const float v = (1.f - uvy) * (float)Ht;
    const float wm = (float)(Wt - 1), hm = (float)(Ht - 1);
    const float nfx = u < 0.0f ? 0.0f : (u > wm ? wm : u);   // cimg::cut
    const float nfy = v < 0.0f ? 0.0f : (v > hm ? hm : v);
    const uint32_t x = cast_u32(nfx), y = cast_u32(nfy);
    const float dx = nfx - (float)x, dy = nfy - (float)y;
    const uint32_t nx = dx > 0 ? x + 1 : x, ny = dy > 0 ? y + 1 : y;
    const uint32_t W = (uint32_t)Wt;
    const uint32_t tcc = t[y * W + x], tnc = t[y * W + nx], tcn = t[ny * W + x], tnn = t[ny * W + nx];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int sh = 8 * c;
        const float Icc = decode((tcc >> sh) & 255u), Inc = decode((tnc >> sh) & 255u);
        const float Icn = decode((tcn >> sh) & 255u), Inn = decode((tnn >> sh) & 255u);
        out[c] = Icc + dx * (Inc - Icc + dy * (Icc + Inn - Icn - Inc)) + dy * (Icn - Icc);
    }
}

// Albedo texel: the Material constructor's normalize(0,1) as a per-byte table L.
__device__ __forceinline__ void sample_texture(const DevMaterial& m, const uint32_t* __restrict__ texels,
                                               const float* __restrict__ L, float uvx, float uvy, float out[3]) {
    sample_bilinear(texels + m.tex_offset, m.tex_w, m.tex_h, uvx, uvy, [&](uint32_t b) { return L[b]; }, out);
}

// Tangent-space normal map: n = normalize(T*t.x + B*t.y + N*t.z), t = 2*s - 1 with s the
// bilinear sample of texel/255, T the face tangent (camera space) made orthogonal to N,
// B = cross(N, T).
__device__ __forceinline__ void normal_map_sample(const ShadeParams& p, const DevMaterial& m, float u, float v,
                                                  float ts[3]) {
    float s[3];
    sample_bilinear(p.texels + m.nmap_offset, m.nmap_w, m.nmap_h, u, v,
                    [](uint32_t b) { return (float)b / 255.0f; }, s);
    ts[0] = 2.0f * s[0] - 1.0f; ts[1] = 2.0f * s[1] - 1.0f; ts[2] = 2.0f * s[2] - 1.0f;
}

__device__ __forceinline__ void bend_normal(const ShadeParams& p, uint32_t f, const float ts[3], float& nx, float& ny,
                                            float& nz) {
    const float tx = p.ftx[f], ty = p.fty[f], tz = p.ftz[f];
    const float* M = p.normal_mat;
    // (an opaque 0: the products M[12..14] * 0 are formed here, not hoisted out of the
    // tile loop into registers that then spill)
    float w = 0.0f;
    asm volatile("" : "+v"(w));
    const float cx = (M[0] * tx + M[4] * ty) + (M[8] * tz + M[12] * w);
    const float cy = (M[1] * tx + M[5] * ty) + (M[9] * tz + M[13] * w);
    const float cz = (M[2] * tx + M[6] * ty) + (M[10] * tz + M[14] * w);
    const float d = (nx * cx + ny * cy) + nz * cz;
    float px = cx - nx * d, py = cy - ny * d, pz = cz - nz * d;
    const float rp = 1.f / sqrtf((px * px + py * py) + pz * pz);
    px = px * rp; py = py * rp; pz = pz * rp;
    const float bx = ny * pz - py * nz, by = nz * px - pz * nx, bz = nx * py - px * ny;
    const float t0 = ts[0], t1 = ts[1], t2 = ts[2];
    float qx = (px * t0 + bx * t1) + nx * t2;
    float qy = (py * t0 + by * t1) + ny * t2;
    float qz = (pz * t0 + bz * t1) + nz * t2;
    const float rq = 1.f / sqrtf((qx * qx + qy * qy) + qz * qz);
    nx = qx * rq; ny = qy * rq; nz = qz * rq;
}

// Normalisation tables of the first kLdsLuts materials, staged in LDS by the whole
// workgroup (a barrier must follow before they are read).
constexpr int kLdsLuts = 4;
__device__ __forceinline__ void stage_luts(const ShadeParams& p, float* s_lut, int tid, int nthreads) {
    const int n = (p.nmat < kLdsLuts ? p.nmat : kLdsLuts) * 256;
    for (int e = tid; e < n; e += nthreads) s_lut[e] = p.luts[e];
}

// The first kLdsMats materials, staged in LDS by the whole workgroup (a barrier must
// follow before they are read).
constexpr int kLdsMats = 8;
__device__ __forceinline__ void stage_mats(const ShadeParams& p, DevMaterial* s_mats, int tid, int nthreads) {
    const int n = (p.nmat < kLdsMats ? p.nmat : kLdsMats) * (int)(sizeof(DevMaterial) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(p.mats);
    uint32_t* dst = reinterpret_cast<uint32_t*>(s_mats);
    for (int e = tid; e < n; e += nthreads) dst[e] = src[e];
}

// stage_mats in two halves, so that other reads can be issued between the loads and the
// LDS stores (which wait for them): each thread holds up to kPasses words meanwhile.
template <int BLOCK>
struct MatStage {
    static constexpr int kWords = kLdsMats * (int)(sizeof(DevMaterial) / 4);
    static constexpr int kPasses = (kWords + BLOCK - 1) / BLOCK;
    int n;
    uint32_t v[kPasses];
    __device__ __forceinline__ MatStage(const ShadeParams& p, int tid) {
        n = (p.nmat < kLdsMats ? p.nmat : kLdsMats) * (int)(sizeof(DevMaterial) / 4);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(p.mats);
#pragma unroll
        for (int j = 0; j < kPasses; ++j) {
            const int e = tid + j * BLOCK;
            v[j] = e < n ? src[e] : 0u;
        }
    }
    __device__ __forceinline__ void store(DevMaterial* s_mats, int tid) const {
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_mats);
#pragma unroll
        for (int j = 0; j < kPasses; ++j) {
            const int e = tid + j * BLOCK;
            if (e < n) dst[e] = v[j];
        }
    }
};

// A pixel no fragment won: 0 / 1.0 when the buffers are known cleared (unless the bin
// launch already wrote them), id -1; otherwise untouched (the reference never clears).
__device__ __forceinline__ void clear_pixel(const ShadeParams& p, size_t i) {
    if (p.pre_cleared) return;
    if (p.cleared) {
        const size_t np = (size_t)p.W * p.H;
        p.frame[i] = 0; p.frame[i + np] = 0; p.frame[i + 2 * np] = 0;
        p.depth[i] = 1.0f;
    }
    if (p.ids) p.ids[i] = -1;
}

// Shades one pixel from its visibility word; every lane of the wave calls it together
// (the learned shader and sample collection work per wave).  Writes frame, depth and
// ids; returns whether the pixel has a winning fragment.
// MODE: kShadePlain (the reference's light loop), kShadeMlp (the learned surrogate on
// matrix cores), kShadeCollect (light loop, and each shaded fragment's MLP inputs and
// colour/255 appended to the sample buffer for training), kShadeChain (neither: colour 0).
template <int MODE>
__device__ __forceinline__ bool shade_pixel(const ShadeParams& p, const float* s_lut, const DevMaterial* s_mats,
                                            const uint16_t* s_net,
                                            uint32_t px, uint32_t py, bool valid, unsigned long long key, int lane,
                                            uint16_t* s_in, float* s_out) {
    const size_t np = (size_t)p.W * p.H;
    const size_t i = (size_t)py * p.W + px;
    const bool hit = valid && key != kEmptyKey;
    float feat[kMlpIn];
    uint8_t rgb[3] = {0, 0, 0};
    if (valid && !hit) clear_pixel(p, i);
    if (hit) {
        const uint32_t f = (uint32_t)key;
        const FaceRecord& r = p.rec[f];
        const float4 a = r.r0, b = r.r1, c = r.r2, w = r.r3;
        float4 g0, g1, g2, g3;   // the record's attribute words (see FaceRecord)
        if (p.uniform >= 0) {    // uniform over the launch
            const float4 n = p.uniform_cn;
            const float2 t = p.uniform_t >= 0 ? p.uvs[p.uniform_t] : make_float2(0.f, 0.f);
            g0 = make_float4(n.x, n.y, n.z, t.x);
            g1 = make_float4(n.x, n.y, n.z, t.y);
            g2 = make_float4(n.x, n.y, n.z, t.x);
            g3 = make_float4(t.y, t.x, t.y, __int_as_float(p.uniform_m));
        } else {
            g0 = r.r4; g1 = r.r5; g2 = r.r6; g3 = r.r7;
        }
        const float pxf = (float)px, pyf = (float)py;
        // barycentric and depth, exactly as the raster pass computed them
        const float e0 = edge(pxf, pyf, b.x, b.y, c.x, c.y);
        const float e1 = edge(pxf, pyf, c.x, c.y, a.x, a.y);
        const float e2 = edge(pxf, pyf, a.x, a.y, b.x, b.y);
        const float b0 = e0 / a.w, b1 = e1 / a.w, b2 = e2 / a.w;
        const float z = (a.z * b0 + b.z * b1) + c.z * b2;
        // drawing.cpp:125-135: perspective-correct interpolation
        const float i0 = w.x * b0, i1 = w.y * b1, i2 = w.z * b2;
        const float d = 1.f / (i0 + i1 + i2);
        float nx = d * (i0 * g0.x + i1 * g1.x + i2 * g2.x);
        float ny = d * (i0 * g0.y + i1 * g1.y + i2 * g2.y);
        float nz = d * (i0 * g0.z + i1 * g1.z + i2 * g2.z);
        const float rn = 1.f / sqrtf((nx * nx + ny * ny) + nz * nz);
        nx = nx * rn; ny = ny * rn; nz = nz * rn;
        if (p.wind_cw) { nx = -nx; ny = -ny; nz = -nz; }
        const int mi = __float_as_int(g3.w);
        DevMaterial m;
        if (s_mats && mi < kLdsMats)
            m = s_mats[mi];
        else
            m = p.mats[mi];
        float alb[3];
        float u = 0.f, v = 0.f;
        if (m.tex_w > 0 || m.nmap_w > 0) {
            u = d * (i0 * g0.w + i1 * g2.w + i2 * g3.y);
            v = d * (i0 * g1.w + i1 * g3.x + i2 * g3.z);
        }
        if (m.tex_w > 0) {
            if (mi < kLdsLuts)
                sample_texture(m, p.texels, s_lut + mi * 256, u, v, alb);
            else
                sample_texture(m, p.texels, p.luts + m.lut_offset, u, v, alb);
        } else {
            alb[0] = m.diffuse[0]; alb[1] = m.diffuse[1]; alb[2] = m.diffuse[2];
        }
        float ts[3] = {0.f, 0.f, 1.f};
        feat[0] = nx; feat[1] = ny; feat[2] = nz;
        if (m.nmap_w > 0) {
            normal_map_sample(p, m, u, v, ts);
            if (MODE == kShadePlain || MODE == kShadeCollect)   // the network takes the sample itself
                bend_normal(p, f, ts, nx, ny, nz);
        }
        p.depth[i] = z;
        if (p.ids) p.ids[i] = (int32_t)f;
        if (MODE == kShadePlain || MODE == kShadeCollect) {
            // shading.cpp:20-34
            constexpr float kOneOverPi = 0.318309886183790671537767526745028724f;
            float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
            const int nk = p.nl < kMaxLights ? p.nl : kMaxLights;
            for (int l = 0; l < nk; ++l) {
                const float dd = (nx * p.light_nd[l][0] + ny * p.light_nd[l][1]) + nz * p.light_nd[l][2];
                const float mm = gmax(0.f, dd);
                acc0 = acc0 + p.light_ic[l][0] * alb[0] * mm * kOneOverPi;
                acc1 = acc1 + p.light_ic[l][1] * alb[1] * mm * kOneOverPi;
                acc2 = acc2 + p.light_ic[l][2] * alb[2] * mm * kOneOverPi;
            }
            // the lights beyond the kernel arguments, in order (uniform: scalar loads; unrolled
            // so that several lights' loads are in flight at once, the sums still in order)
#pragma unroll 8
            for (int l = kMaxLights; l < p.nl; ++l) {
                const float* e = p.lights_ext + 6 * (l - kMaxLights);
                const float dd = (nx * e[3] + ny * e[4]) + nz * e[5];
                const float mm = gmax(0.f, dd);
                acc0 = acc0 + e[0] * alb[0] * mm * kOneOverPi;
                acc1 = acc1 + e[1] * alb[1] * mm * kOneOverPi;
                acc2 = acc2 + e[2] * alb[2] * mm * kOneOverPi;
            }
            rgb[0] = (uint8_t)cast_u32(gmin(acc0, 255.f));
            rgb[1] = (uint8_t)cast_u32(gmin(acc1, 255.f));
            rgb[2] = (uint8_t)cast_u32(gmin(acc2, 255.f));
        }
        if (MODE == kShadeChain) {
            // the samples every mode takes stay live (the compiler would drop their loads)
            asm volatile("" ::"v"(alb[0]), "v"(alb[1]), "v"(alb[2]), "v"(ts[0]), "v"(ts[1]), "v"(ts[2]), "v"(nx),
                         "v"(ny), "v"(nz));
        }
        if (MODE == kShadeMlp || MODE == kShadeCollect) {
            // view vector: from the surface point (eye space, depth d) to the eye
            const float xn = 2.0f * pxf / (float)p.W - 1.0f, yn = 1.0f - 2.0f * pyf / (float)p.H;
            float vx = -(xn * d) * p.inv_p00, vy = -(yn * d) * p.inv_p11, vz = d;
            const float rv = 1.f / sqrtf((vx * vx + vy * vy) + vz * vz);
            feat[3] = vx * rv; feat[4] = vy * rv; feat[5] = vz * rv;
            feat[6] = alb[0]; feat[7] = alb[1]; feat[8] = alb[2];
            feat[9] = ts[0]; feat[10] = ts[1]; feat[11] = ts[2];
        }
    }
    if (MODE == kShadeCollect) {
        // append (inputs, colour/255) of this wave's shaded fragments
        const unsigned long long bal = __ballot(hit);
        unsigned long long base = 0;
        if (lane == 0 && bal) base = atomicAdd(p.sample_count, (unsigned long long)__popcll(bal));
        base = __shfl(base, 0, 64);
        const unsigned long long slot = base + __popcll(bal & ((1ull << lane) - 1ull));
        if (hit && slot < p.sample_cap) {
            float* o = p.samples + slot * kMlpSampleFloats;
            for (int k = 0; k < kMlpIn; ++k) o[k] = feat[k];
            for (int c = 0; c < 3; ++c) o[kMlpIn + c] = (float)rgb[c] / 255.0f;
        }
    }
    if (MODE == kShadeMlp) {
#pragma unroll
        for (int k = 0; k < 16; ++k) s_in[lane * 16 + k] = to_bf16((hit && k < kMlpIn) ? feat[k] : 0.f);
        if (__ballot(hit)) {   // uniform over the wave: waves with no fragment skip the network
            __builtin_amdgcn_wave_barrier();
            mlp_wave(s_net, s_in, s_out, lane);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = s_out[lane * 4 + c] * 255.0f;
                const float yc = y > 0.f ? (y < 255.f ? y : 255.f) : 0.f;
                rgb[c] = (uint8_t)(uint32_t)(yc + 0.5f);
            }
        }
    }
    if (hit) {
        p.frame[i] = rgb[0];
        p.frame[i + np] = rgb[1];
        p.frame[i + 2 * np] = rgb[2];
    }
    return hit;
}

// Deferred shade pass: 64 x 4 pixel blocks, each wave 64 consecutive pixels of a row.
// s_lut / s_mats / s_net: the normalisation tables, materials and network, staged in LDS
// by the caller.
template <int MODE>
__device__ __forceinline__ void shade_body(const ShadeParams& p, const float* s_lut, const DevMaterial* s_mats,
                                           const uint16_t* s_net, int blk) {
    __shared__ __align__(16) uint16_t s_in[MODE == kShadeMlp ? 4 : 1][64 * 16];
    __shared__ __align__(16) float s_out[MODE == kShadeMlp ? 4 : 1][64 * 4];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int ws = MODE == kShadeMlp ? threadIdx.y : 0;
    const uint32_t nbx = shade_blocks_x(p.W);
    const uint32_t px = (blk % nbx) * 64 + threadIdx.x, py = (blk / nbx) * 4 + threadIdx.y;
    const bool valid = px < p.W && py < p.H;
    const unsigned long long key = valid ? p.vis[(size_t)py * p.W + px] : kEmptyKey;
    const bool hit = shade_pixel<MODE>(p, s_lut, s_mats, s_net, px, py, valid, key, threadIdx.x, s_in[ws], s_out[ws]);
    store_block_sum(hit ? 1u : 0u, p.part_cov + blk, tid, 256);
}

// Deferred shade of the tiles whose bin overflowed, when the raster pass shades the
// others itself (one workgroup per tile, visibility words from HBM).
template <int TS, int BLOCK, int MODE>
__device__ __forceinline__ void shade_tile_body(const RasterParams& rp, const ShadeParams& p, const float* s_lut,
                                                const DevMaterial* s_mats, const uint16_t* s_net, int tile,
                                                uint32_t& cov) {
    constexpr int NW = BLOCK / 64;
    __shared__ __align__(16) uint16_t s_in[MODE == kShadeMlp ? NW : 1][64 * 16];
    __shared__ __align__(16) float s_out[MODE == kShadeMlp ? NW : 1][64 * 4];
    const int tid = threadIdx.x, lane = tid & 63;
    const int ws = MODE == kShadeMlp ? tid >> 6 : 0;
    const int ox = (tile % rp.tiles_x) * TS, oy = (tile / rp.tiles_x) * TS;
#pragma unroll
    for (int l0 = 0; l0 < TS * TS; l0 += BLOCK) {
        const int l = l0 + tid;
        const uint32_t x = ox + (l % TS), y = oy + (l / TS);
        const bool valid = x < p.W && y < p.H;
        const unsigned long long key = valid ? p.vis[(size_t)y * p.W + x] : kEmptyKey;
        const bool hit = shade_pixel<MODE>(p, s_lut, s_mats, s_net, x, y, valid, key, lane, s_in[ws], s_out[ws]);
        cov += hit ? 1u : 0u;
    }
    __syncthreads();   // the staging buffers are reused by the next tile
}

// ---------------------------------------------------------------------------------
// Inside + depth test of one (face, pixel) candidate; returns true when inside.
// The reference tests b_i = e_i / area >= 0 (drawing.cpp:105-113).  When |area| lies in
// [2^-100, 2^100] and |e_i| >= 2^-20, the quotient is a normal number far from zero and
// infinity, so its sign is sign(e_i) * sign(area): such a candidate with a sign mismatch
// is outside without dividing.  Every other candidate takes the three divisions, so the
// result is the reference's predicate bit for bit.
__device__ __forceinline__ bool area_in_range(float area) {
    const float a = fabsf(area);
    return a >= 0x1p-100f && a <= 0x1p100f;
}

__device__ __forceinline__ bool surely_outside(float e, bool area_pos) {
    return fabsf(e) >= 0x1p-20f && ((e > 0.0f) != area_pos);
}

// The general test from the three edge values.
__device__ __forceinline__ bool key_from_edges(float e0, float e1, float e2, float z0, float z1, float z2,
                                               float area, float& z) {
    if (area_in_range(area)) {
        const bool ap = area > 0.0f;
        if (surely_outside(e0, ap) || surely_outside(e1, ap) || surely_outside(e2, ap)) return false;
    }
    const float b0 = e0 / area, b1 = e1 / area, b2 = e2 / area;
    if (!(b0 >= 0.0f && b1 >= 0.0f && b2 >= 0.0f)) return false;
    z = (z0 * b0 + z1 * b1) + z2 * b2;
    return true;
}

__device__ __forceinline__ bool fragment_key(float pxf, float pyf, float x0, float y0, float z0, float x1,
                                             float y1, float z1, float x2, float y2, float z2, float area,
                                             float& z) {
    const float e0 = edge(pxf, pyf, x1, y1, x2, y2);
    const float e1 = edge(pxf, pyf, x2, y2, x0, y0);
    const float e2 = edge(pxf, pyf, x0, y0, x1, y1);
    return key_from_edges(e0, e1, e2, z0, z1, z2, area, z);
}

// Quotients e / d sharing one divisor d, bit for bit.  The compiler's correctly rounded
// f32 division is div_scale of both operands, rcp of the scaled divisor, two Newton
// steps on the reciprocal (t = fma(-d, r, 1), r1 = fma(t, r, r)), then q = n r1,
// q = fma(fma(-d, q, n), r1, q), div_fmas(fma(-d, q, n), r1, q), div_fixup.  div_scale
// returns its operand unchanged and clears the flag div_fmas scales by unless an operand
// is zero, the exponent difference is 96 or more, d or 1 / d is not normal, the
// quotient is not normal or |n| < 2^-103; div_fixup returns the magnitude it is given
// with the quotient's sign unless an operand is zero, infinite or NaN or the result
// under- or overflows.  None of that happens for |d| in [2^-40, 2^40] and |n| in
// [2^-20, 2^40] (exponent difference at most 80, quotient in [2^-60, 2^80]), where the
// division is therefore the sequence without them: r1 depends on d alone and is computed
// once per face (refined_rcp), and each quotient costs one multiply and four FMAs instead
// of the ten-instruction sequence.  The fused raster's candidates, the overwhelming
// majority of which fall in that range, use it; the others take the general test.
__device__ __forceinline__ bool divisor_in_range(float d) {
    const float a = fabsf(d);
    return a >= 0x1p-40f && a <= 0x1p40f;
}

__device__ __forceinline__ bool dividend_in_range(float n) {
    const float a = fabsf(n);
    return (a >= 0x1p-20f) & (a <= 0x1p40f);
}

// The divisor's refined reciprocal, or 0 for a divisor outside divisor_in_range.
__device__ __forceinline__ float refined_rcp(float d) {
    if (!divisor_in_range(d)) return 0.0f;
    const float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
}

__device__ __forceinline__ float quotient_in_range(float n, float d, float r1) {
    float q = n * r1;
    q = __builtin_fmaf(__builtin_fmaf(-d, q, n), r1, q);
    return __builtin_fmaf(__builtin_fmaf(-d, q, n), r1, q);
}

// fragment_key with the face's refined_rcp(area) at hand.  In range, each quotient is a
// normal number, so the inside test is the edge values' signs against the area's.
__device__ __forceinline__ bool fragment_key_rcp(float pxf, float pyf, float x0, float y0, float z0, float x1,
                                                 float y1, float z1, float x2, float y2, float z2, float area,
                                                 float ra, float& z) {
    const float e0 = edge(pxf, pyf, x1, y1, x2, y2);
    const float e1 = edge(pxf, pyf, x2, y2, x0, y0);
    const float e2 = edge(pxf, pyf, x0, y0, x1, y1);
    // (non-short-circuit: every operand is a compare, and branches would cost more)
    if ((ra != 0.0f) & dividend_in_range(e0) & dividend_in_range(e1) & dividend_in_range(e2)) {
        const bool ap = area > 0.0f;
        if (((e0 > 0.0f) != ap) | ((e1 > 0.0f) != ap) | ((e2 > 0.0f) != ap)) return false;
        const float b0 = quotient_in_range(e0, area, ra), b1 = quotient_in_range(e1, area, ra);
        const float b2 = quotient_in_range(e2, area, ra);
        z = (z0 * b0 + z1 * b1) + z2 * b2;
        return true;
    }
    return key_from_edges(e0, e1, e2, z0, z1, z2, area, z);
}

// One tile, by one workgroup.  The tile's bin is taken BLOCK faces at a time: the faces
// whose bbox meets the tile are compacted (ballot + popcount ranks, wave totals in LDS)
// and their counts of candidate pairs (two pixels of the face's rectangle in its tile)
// prefix-summed across the workgroup.  Pairs are then handed out one per thread, up to
// 4096 per round: a bitmask in LDS marks each face's first pair, and a pair's face is
// found by counting the marks at or before it (per-word prefix counts + one popcount),
// with no search.  A pair's two pixels share the face's LDS reads and the pixel's place
// in the rectangle (a face with an odd count leaves one lane idle in its last pair).
//
// FUSED: the workgroup then shades its own pixels from the LDS visibility words (the
// deferred shade, without the round trip of the visibility buffer through HBM).  Tiles
// whose bin overflowed are skipped: the overflow loop fills their visibility words and
// k_shade_tiles shades them.
// A tile's face count and the first BLOCK entries of its bin (thread tid's one), read
// together: what raster_body needs before anything else.  An entry past the count is
// never used; every index below the capacity is inside the bin.
struct TileHead {
    uint32_t count, f_first;
};
// z: 0, or an opaque copy of 0 that keeps the count in a vector register until its use
// (a uniform load result is otherwise moved to a scalar register, and waited for, where
// it is read).
__device__ __forceinline__ TileHead tile_head(const RasterParams& p, int tile, int tid, uint32_t z = 0u) {
    TileHead h;
    h.f_first = issue_now(p.bins + (size_t)tile * p.cap + ((uint32_t)tid < p.cap ? (uint32_t)tid : 0u));
    h.count = issue_now(p.tile_count + (size_t)tile * kCountStride + z);
    return h;
}

// Bytes of raster_body's scratch LDS: face data by rank, reused by the learned shader's
// staging after the raster loop (and, in the batch-frame kernel, by its bin phase).
template <int BLOCK, int MODE>
struct RasterLds {
    static constexpr int kRaster = 14 * BLOCK * 4;
    static constexpr int kMlp = MODE == kShadeMlp ? (BLOCK / 64) * (64 * 16 * 2 + 64 * 4 * 4) : 0;
    static constexpr int kBytes = kRaster > kMlp ? kRaster : kMlp;
};

// EXT_RAW: the scratch LDS (RasterLds bytes) is the caller's `raw`, else the function's own.
// DYN_D0: the initial depths live in the launch's dynamic LDS (TS * TS floats when the
// frame has an input depth buffer, none otherwise), not in a static array.
template <int TS, int BLOCK, bool FUSED, int MODE, int SHRINK = kSlabPixels, bool EXT_RAW = false,
          int SLAB = kSlabAlways, bool DYN_D0 = false>
__device__ __forceinline__ void raster_body(const RasterParams& p, const ShadeParams* sp, const float* s_lut,
                                            const DevMaterial* s_mats, const uint16_t* s_net, int tile,
                                            const TileHead& head, uint32_t& frags, uint32_t& cov,
                                            unsigned char* raw = nullptr) {
    constexpr int NPX = TS * TS;
    constexpr int PPT = NPX / BLOCK;
    constexpr int NW = BLOCK / 64;
    constexpr int kWords = 64;                 // mask words per round: 4096 candidate pairs
    static_assert(PPT >= 1 && NPX % BLOCK == 0, "tile must cover the block");
    constexpr bool kNet = FUSED && MODE == kShadeMlp;
    __shared__ unsigned long long s_key[NPX];
    float* s_d0;
    if constexpr (DYN_D0) {
        extern __shared__ float rz_dyn_d0[];
        s_d0 = rz_dyn_d0;
    } else {
        __shared__ float s_d0_static[NPX];
        s_d0 = s_d0_static;
    }
    unsigned char* s_raw = raw;
    if constexpr (!EXT_RAW) {
        __shared__ __align__(16) unsigned char s_own[RasterLds<BLOCK, MODE>::kBytes];
        s_raw = s_own;
    }
    // x0 y0 z0 x1 y1 z1 x2 y2 z2 area refined_rcp(area)
    float (*s_v)[BLOCK] = reinterpret_cast<float (*)[BLOCK]>(s_raw);
    uint32_t* s_rect = reinterpret_cast<uint32_t*>(s_raw + 11 * BLOCK * 4);   // rx0 | ry0 << 8 | rw << 16
    uint32_t* s_face = s_rect + BLOCK;
    uint32_t* s_start = s_face + BLOCK;
    __shared__ uint32_t s_wsum[NW], s_wlive[NW];
    __shared__ unsigned long long s_mask[kWords];
    __shared__ uint32_t s_wpre[kWords];

    // an opaque copy of the thread index: values derived from it are recomputed for each
    // tile instead of being hoisted out of the kernels' tile loops into registers
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    tile = __builtin_amdgcn_readfirstlane(tile);   // (uniform; read ahead in a vector register)
    const int tx = tile % p.tiles_x, ty = tile / p.tiles_x;
    const int ox = tx * TS, oy = ty * TS;

    // the first chunk's bin entries, read beside the count (tile_head)
    const uint32_t f_first = head.f_first;
    const uint32_t count = __builtin_amdgcn_readfirstlane(head.count);
    // the bin pass cleared an overflowed tile's visibility words and the overflow loop
    // fills them: such a tile has no chunk and no stores here (uniform over the workgroup)
    const bool overflowed = count > p.cap;
    const uint32_t n = overflowed ? 0u : count;
    // the first chunk's face records, read before anything waits
    float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), b0 = a0, c0 = a0;
    if ((uint32_t)tid < n) {
        const FaceRecord* r = p.rec + f_first;
        a0 = r->r0; b0 = r->r1; c0 = r->r2;
    }
    // initial depths: the caller's depth buffer, staged in LDS, or 1 everywhere for a
    // frame without one (uniform; then nothing is staged)
    const bool has_d0 = p.depth_in != nullptr;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int l = tid + j * BLOCK;
        s_key[l] = kEmptyKey;
        if (has_d0) {
            const uint32_t x = ox + (l % TS), y = oy + (l / TS);
            float d0 = 1.0f;
            if (x < p.W && y < p.H) d0 = p.depth_in[(size_t)y * p.W + x];
            s_d0[l] = d0;
        }
    }
    __syncthreads();

    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long upto = (below << 1) | 1ull;   // bits 0..lane
    static_assert(BLOCK % 64 == 0, "candidates are dealt in whole waves");
    // A chunk's faces (entry i of the bin: face f, record a b c) into the LDS face table
    // by rank: their candidate rectangles, the prefix sums, the chunk's total of pairs.
    struct Staged {
        uint32_t cnt, start, total;
    };
    auto stage = [&](uint32_t i, uint32_t f, const float4& a, const float4& b, const float4& c) -> Staged {
        uint32_t cnt = 0, rect = 0;
        if (i < n) {
            const uint32_t lo = __float_as_uint(b.w), hi = __float_as_uint(c.w);
            int rx0 = max((int)(lo & 0xFFFFu) - ox, 0), rx1 = min((int)(hi & 0xFFFFu) - ox, TS - 1);
            int ry0 = max((int)(lo >> 16) - oy, 0), ry1 = min((int)(hi >> 16) - oy, TS - 1);
            // large rectangles: only the part the triangle can reach
            bool keep = true;
            if (rx0 <= rx1 && ry0 <= ry1 && (rx1 - rx0 + 1) * (ry1 - ry0 + 1) > SHRINK)
                keep = shrink_rect<SLAB>(a, b, c, ox, oy, rx0, rx1, ry0, ry1);
            if (keep && rx0 <= rx1 && ry0 <= ry1) {
                const uint32_t rw = rx1 - rx0 + 1, rh = ry1 - ry0 + 1;
                cnt = (rw * rh + 1u) >> 1;   // candidate pairs
                rect = (uint32_t)rx0 | ((uint32_t)ry0 << 8) | (rw << 16) | (rh << 24);
            }
        }
        // workgroup-wide exclusive prefix sums of the candidate pair counts and of the live
        // faces
        const unsigned long long live = __ballot(cnt > 0);
        const uint32_t v = wave_scan_incl(cnt);
        uint32_t wbase = 0, lbase = 0, total = 0;
        if constexpr (NW == 1) {
            total = __builtin_amdgcn_readlane(v, 63);
        } else {
            if (lane == 63) {
                s_wsum[wave] = v;
                s_wlive[wave] = (uint32_t)__popcll(live);
            }
            __syncthreads();
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint32_t sw = s_wsum[w], lw = s_wlive[w];
                if (w < wave) { wbase += sw; lbase += lw; }
                total += sw;
            }
        }
        const uint32_t start = wbase + v - cnt;
        if (cnt > 0) {
            const int rank = (int)(lbase + (uint32_t)__popcll(live & below));
            s_v[0][rank] = a.x; s_v[1][rank] = a.y; s_v[2][rank] = a.z;
            s_v[3][rank] = b.x; s_v[4][rank] = b.y; s_v[5][rank] = b.z;
            s_v[6][rank] = c.x; s_v[7][rank] = c.y; s_v[8][rank] = c.z;
            s_v[9][rank] = a.w; s_v[10][rank] = refined_rcp(a.w);
            s_rect[rank] = rect;
            s_face[rank] = f;
            s_start[rank] = start;
        }
        return Staged{cnt, start, total};
    };
    // the first chunk from the records read above (which are then dead: they are not
    // carried through the candidate loop), each later one read when its turn comes
    Staged st{0u, 0u, 0u};
    if (n > 0) st = stage((uint32_t)tid, f_first, a0, b0, c0);
    for (uint32_t base = 0; base < n; base += BLOCK) {
        if (base > 0) {
            const uint32_t i = base + tid;
            uint32_t f = 0;
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a, c = a;
            if (i < n) {
                f = p.bins[(size_t)tile * p.cap + i];
                const FaceRecord* r = p.rec + f;
                a = r->r0; b = r->r1; c = r->r2;
            }
            st = stage(i, f, a, b, c);
        }
        const uint32_t cnt = st.cnt, start = st.start, total = st.total;
        // candidate pair k of face j (rank): the inside and depth tests of the pixels 2k and
        // 2k + 1 of its rectangle (row by row; the second only if the rectangle has it),
        // which share the face's reads and the pixel's place in the rectangle
        auto test = [&](uint32_t k, int j) {
            const uint32_t rc = s_rect[j];
            const uint32_t rw = (rc >> 16) & 0xFFu, rh = rc >> 24;
            const uint32_t l = 2u * (k - s_start[j]);
            // l / rw exactly: |error| of the product < 1/(2*rw) for l < 1024
            const uint32_t dy = (uint32_t)(((float)l + 0.5f) * __builtin_amdgcn_rcpf((float)rw));
            const uint32_t dx = l - dy * rw;
            const int lx = (int)(rc & 0xFFu) + (int)dx, ly = (int)((rc >> 8) & 0xFFu) + (int)dy;
            const bool wrap = dx + 1u == rw;
            const int lx2 = wrap ? (int)(rc & 0xFFu) : lx + 1, ly2 = wrap ? ly + 1 : ly;
            const bool two = l + 1u < rw * rh;
            const float x0 = s_v[0][j], y0 = s_v[1][j], z0 = s_v[2][j], x1 = s_v[3][j], y1 = s_v[4][j];
            const float z1 = s_v[5][j], x2 = s_v[6][j], y2 = s_v[7][j], z2 = s_v[8][j], ar = s_v[9][j];
            const float ra = s_v[10][j];
            float z, zb;
            const bool ina = fragment_key_rcp((float)(ox + lx), (float)(oy + ly), x0, y0, z0, x1, y1, z1, x2, y2,
                                              z2, ar, ra, z);
            const bool inb = two && fragment_key_rcp((float)(ox + lx2), (float)(oy + ly2), x0, y0, z0, x1, y1, z1,
                                                     x2, y2, z2, ar, ra, zb);
            if (ina | inb) {
                const uint32_t fj = s_face[j];
                if (ina) {
                    ++frags;
                    const int pl = ly * TS + lx;
                    if (z < (has_d0 ? s_d0[pl] : 1.0f)) atomicMin(&s_key[pl], vis_key(z, fj));
                }
                if (inb) {
                    ++frags;
                    const int pl = ly2 * TS + lx2;
                    if (zb < (has_d0 ? s_d0[pl] : 1.0f)) atomicMin(&s_key[pl], vis_key(zb, fj));
                }
            }
        };
        uint32_t before = 0;   // live faces whose first candidate precedes the round
        if constexpr (NW == 1) {
            // One wave: each pass of 64 candidate pairs finds their faces from a 64-bit mask of
            // the face starts inside it, OR-reduced across the wave, with no LDS round trip.
            auto bor = [](int x, int y) { return x | y; };
            __syncthreads();   // (the face table above is read by other lanes)
            for (uint32_t r0 = 0; r0 < total; r0 += 64) {   // (uniform)
                const bool here = cnt > 0 && start >= r0 && start - r0 < 64u;
                const uint32_t sb = start - r0;
                const uint32_t lo = wave_reduce(here && sb < 32u ? (int)(1u << sb) : 0, 0, bor);
                const uint32_t hi = wave_reduce(here && sb >= 32u ? (int)(1u << (sb - 32u)) : 0, 0, bor);
                const unsigned long long m = ((unsigned long long)hi << 32) | lo;
                const uint32_t k = r0 + (uint32_t)lane;
                if (k < total) test(k, (int)(before + (uint32_t)__popcll(m & upto)) - 1);
                before += (uint32_t)__popcll(m);
            }
        } else {
            for (uint32_t r0 = 0; r0 < total; r0 += kWords * 64) {
                if (tid < kWords) s_mask[tid] = 0ull;
                __syncthreads();
                if (cnt > 0 && start >= r0 && start - r0 < kWords * 64u)
                    atomicOr(&s_mask[(start - r0) >> 6], 1ull << ((start - r0) & 63u));
                __syncthreads();
                if (wave == 0) {
                    const uint32_t pc = (uint32_t)__popcll(s_mask[lane]);
                    s_wpre[lane] = wave_scan_incl(pc) - pc;
                }
                __syncthreads();
                const uint32_t r1 = min(total, r0 + kWords * 64u);
                for (uint32_t k = r0 + tid; k < r1; k += BLOCK) {
                    // (r0 and BLOCK are multiples of 64: candidate k's bit in its mask word
                    // is the lane, so the mask of bits up to it is the loop-invariant upto)
                    const uint32_t wi = (k - r0) >> 6;
                    test(k, (int)(before + s_wpre[wi] + (uint32_t)__popcll(s_mask[wi] & upto)) - 1);
                }
                before += s_wpre[kWords - 1] + (uint32_t)__popcll(s_mask[kWords - 1]);
                __syncthreads();   // the mask is cleared for the next round
            }
        }
        __syncthreads();       // ranks and wave totals are rewritten by the next chunk
    }
    if (overflowed) return;

    if (FUSED && PPT > 1) {
        // Several pixels per thread: the covered pixels are compacted first, so that a
        // thread shades one of them per pass instead of running the shading chain once
        // per pixel it owns (a sparse tile then costs one chain, not PPT).
        uint16_t* s_in = reinterpret_cast<uint16_t*>(s_raw) + (kNet ? wave : 0) * 64 * 16;
        float* s_out = reinterpret_cast<float*>(s_raw + NW * 64 * 16 * 2) + (kNet ? wave : 0) * 64 * 4;
        __shared__ uint16_t s_hits[PPT > 1 ? NPX : 1];
        uint32_t nh = 0;   // uniform over the workgroup
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int l = tid + j * BLOCK;
            const uint32_t x = ox + (l % TS), y = oy + (l / TS);
            const bool valid = x < p.W && y < p.H;
            const bool hit = valid && s_key[l] != kEmptyKey;
            if (valid && !hit) clear_pixel(*sp, (size_t)y * p.W + x);
            const unsigned long long bal = __ballot(hit);
            uint32_t off = nh, tot = nh;
            if constexpr (NW == 1) {
                tot += (uint32_t)__popcll(bal);
            } else {
                if (lane == 0) s_wsum[wave] = (uint32_t)__popcll(bal);
                __syncthreads();
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const uint32_t v = s_wsum[w];
                    if (w < wave) off += v;
                    tot += v;
                }
                __syncthreads();   // s_wsum is rewritten by the next pass
            }
            if (hit) s_hits[off + (uint32_t)__popcll(bal & below)] = (uint16_t)l;
            nh = tot;
        }
        __syncthreads();
        for (uint32_t k0 = 0; k0 < nh; k0 += BLOCK) {   // uniform over the workgroup
            const uint32_t k = k0 + (uint32_t)tid;
            const bool valid = k < nh;
            const int l = valid ? (int)s_hits[k] : 0;
            const uint32_t x = ox + (l % TS), y = oy + (l / TS);
            const bool hit = shade_pixel<MODE>(*sp, s_lut, s_mats, s_net, x, y, valid,
                                               valid ? s_key[l] : kEmptyKey, lane, s_in, s_out);
            cov += hit ? 1u : 0u;
        }
    } else if (FUSED) {
        uint16_t* s_in = reinterpret_cast<uint16_t*>(s_raw) + (kNet ? wave : 0) * 64 * 16;
        float* s_out = reinterpret_cast<float*>(s_raw + NW * 64 * 16 * 2) + (kNet ? wave : 0) * 64 * 4;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int l = tid + j * BLOCK;
            const uint32_t x = ox + (l % TS), y = oy + (l / TS);
            const bool valid = x < p.W && y < p.H;
            const bool hit = shade_pixel<MODE>(*sp, s_lut, s_mats, s_net, x, y, valid, s_key[l], lane, s_in, s_out);
            cov += hit ? 1u : 0u;
        }
    } else {
        // the tile's pixels by rows: one address, then BLOCK / TS rows further per pass;
        // the bounds test only in a tile at the image's right or bottom edge (uniform)
        constexpr int kRows = BLOCK / TS;
        const uint32_t x = ox + (tid % TS), y = oy + (tid / TS);
        unsigned long long* dst = p.vis + (size_t)y * p.W + x;
        const size_t stride = (size_t)kRows * p.W;
        if ((uint32_t)ox + TS <= p.W && (uint32_t)oy + TS <= p.H) {
#pragma unroll
            for (int j = 0; j < PPT; ++j) dst[j * stride] = s_key[tid + j * BLOCK];
        } else {
#pragma unroll
            for (int j = 0; j < PPT; ++j)
                if (x < p.W && y + j * kRows < p.H) dst[j * stride] = s_key[tid + j * BLOCK];
        }
    }
    __syncthreads();   // LDS is reused by the workgroup's next tile
}

// ---------------------------------------------------------------------------------
// Every face meeting an overflowed tile, into the visibility words by global atomicMin;
// thread t of nthreads (both multiples of 64 at wave starts).  Each lane tests one face
// for an overflowed tile in its tile rectangle; the wave then takes the faces found one
// at a time, its 64 lanes spread over the face's pixels in each overflowed tile (a big
// face's pixels would otherwise be one lane's serial loop).  Returns the fragments passing
// the inside test.
__device__ __forceinline__ uint32_t overflow_faces(const RasterParams& p, uint32_t t, uint32_t nthreads) {
    const int32_t ts_log2 = p.ts_log2;
    const uint32_t ts = 1u << ts_log2;
    const uint32_t lane = t & 63u;
    uint32_t frags = 0;
    for (uint32_t base = t - lane; base < (uint32_t)p.nf; base += nthreads) {   // uniform over the wave
        // binning positions base + lane: those with a record, through their face ids
        const uint32_t i = base + lane;
        const unsigned long long fm = p.front_mask[base >> 6];
        bool meets = false;
        uint32_t f = 0;
        if (i < (uint32_t)p.nf && ((fm >> lane) & 1ull)) {
            f = (uint32_t)p.face_v[i].w;
            const FaceRecord* r = p.rec + f;
            const uint32_t lo = __float_as_uint(r->r1.w), hi = __float_as_uint(r->r2.w);
            const uint32_t bx0 = lo & 0xFFFFu, by0 = lo >> 16, bx1 = hi & 0xFFFFu, by1 = hi >> 16;
            if (bx0 <= bx1 && by0 <= by1) {
                for (uint32_t ty = by0 >> ts_log2; ty <= (by1 >> ts_log2) && !meets; ++ty)
                    for (uint32_t tx = bx0 >> ts_log2; tx <= (bx1 >> ts_log2) && !meets; ++tx)
                        meets = p.tile_count[(size_t)(ty * p.tiles_x + tx) * kCountStride] > p.cap;
            }
        }
        unsigned long long todo = __ballot(meets);
        while (todo) {
            const uint32_t g = __builtin_amdgcn_readlane(f, __ffsll((long long)todo) - 1);   // (todo uniform)
            todo &= todo - 1ull;
            const FaceRecord* r = p.rec + g;
            const float4 a = r->r0, b = r->r1, c = r->r2;
            const uint32_t lo = __float_as_uint(b.w), hi = __float_as_uint(c.w);
            const uint32_t bx0 = lo & 0xFFFFu, by0 = lo >> 16, bx1 = hi & 0xFFFFu, by1 = hi >> 16;
            for (uint32_t ty = by0 >> ts_log2; ty <= (by1 >> ts_log2); ++ty) {
                for (uint32_t tx = bx0 >> ts_log2; tx <= (bx1 >> ts_log2); ++tx) {
                    const uint32_t tile = ty * p.tiles_x + tx;
                    if (p.tile_count[(size_t)tile * kCountStride] <= p.cap) continue;
                    const uint32_t x0 = max(bx0, tx * ts), x1 = min(bx1, tx * ts + ts - 1);
                    const uint32_t y0 = max(by0, ty * ts), y1 = min(by1, ty * ts + ts - 1);
                    const uint32_t rw = x1 - x0 + 1, npx = rw * (y1 - y0 + 1);   // npx <= 1024
                    for (uint32_t l = lane; l < npx; l += 64) {
                        // l / rw exactly: |error| of the product < 1/(2*rw) for l < 1024
                        const uint32_t dy = (uint32_t)(((float)l + 0.5f) * __builtin_amdgcn_rcpf((float)rw));
                        const uint32_t x = x0 + (l - dy * rw), y = y0 + dy;
                        float z;
                        if (!fragment_key((float)x, (float)y, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, a.w, z))
                            continue;
                        ++frags;
                        const size_t pi = (size_t)y * p.W + x;
                        const float d0 = p.depth_in ? p.depth_in[pi] : 1.0f;
                        if (z < d0) atomicMin(&p.vis[pi], vis_key(z, g));
                    }
                }
            }
        }
    }
    return frags;
}

template <int BLOCK = 256>
__device__ __forceinline__ void overflow_body(const RasterParams& p, int blk, int nblk) {
    const uint32_t frags = p.counters[kCntOverflowTiles] == 0ull
                               ? 0u
                               : overflow_faces(p, (uint32_t)blk * BLOCK + threadIdx.x, (uint32_t)nblk * BLOCK);
    store_block_sum(frags, p.part_frag + blk, threadIdx.x, BLOCK);
}

// ---------------------------------------------------------------------------------
// Entry kernels.  Single frame: parameters by value.  Batch: per-mesh parameter tables
// in device memory, with a map from block to (mesh, block within that mesh's pass) or
// the shared tile lists, so one launch covers every mesh.

// Fills nbytes at base with a 32-bit pattern laid out as in a 4-byte-aligned array of
// it (bytes at a ragged head and tail, 16-byte stores between); thread t of nthreads.
__device__ __forceinline__ void fill_pattern(uint8_t* base, size_t nbytes, uint32_t pat, size_t t, size_t nthreads) {
    const uintptr_t a = (uintptr_t)base;
    const size_t mis = (16u - (a & 15u)) & 15u;
    const size_t head = mis < nbytes ? mis : nbytes;
    const size_t n16 = (nbytes - head) / 16;
    for (size_t i = t; i < head; i += nthreads) base[i] = (uint8_t)(pat >> (8u * ((a + i) & 3u)));
    uint4* body = reinterpret_cast<uint4*>(base + head);
    const uint4 v = make_uint4(pat, pat, pat, pat);
    for (size_t i = t; i < n16; i += nthreads) body[i] = v;
    for (size_t i = head + 16 * n16 + t; i < nbytes; i += nthreads) base[i] = (uint8_t)(pat >> (8u * ((a + i) & 3u)));
}

template <int BLOCK>
__device__ __forceinline__ void clear_body(const BinParams& p, uint32_t blk, uint32_t nblk) {
    const size_t t = (size_t)blk * BLOCK + threadIdx.x, nt = (size_t)nblk * BLOCK;
    if (p.clear_frame) fill_pattern(p.clear_frame, 3 * p.clear_np, 0u, t, nt);
    if (p.clear_depth)
        fill_pattern(reinterpret_cast<uint8_t*>(p.clear_depth), 4 * p.clear_np, 0x3F800000u, t, nt);   // 1.0f
    if (p.clear_ids) fill_pattern(reinterpret_cast<uint8_t*>(p.clear_ids), 4 * p.clear_np, 0xFFFFFFFFu, t, nt);
    if (p.clear_vis) fill_pattern(reinterpret_cast<uint8_t*>(p.clear_vis), 8 * p.clear_np, 0xFFFFFFFFu, t, nt);
}

// The next frame's generation of counter lines: words 0-1 of each line (a count, and a
// list counter's second word), all of the last line (the global counters).
template <int BLOCK>
__device__ __forceinline__ void zero_lines_body(const BinParams& p, uint32_t blk) {
    const uint32_t i = blk * BLOCK + threadIdx.x;
    if (i >= p.zero_count) return;
    uint32_t* line = p.zero_lines + (size_t)i * kCountStride;
    if (i + 1 < p.zero_count) {
        *reinterpret_cast<uint2*>(line) = make_uint2(0u, 0u);
    } else {
#pragma unroll
        for (int k = 0; k < 2 * kNumCounters; k += 4) *reinterpret_cast<uint4*>(line + k) = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Binning workgroups first, then the ones zeroing the next frame's counter lines, then
// (fused pipeline, cleared buffers) the clearing ones, which run on the CUs the
// latency-bound binning leaves idle.
// (single frames: kFrameBinBlock-thread workgroups)
__global__ __launch_bounds__(kFrameBinBlock) void k_bin(BinParams p) {
    __shared__ __align__(16) unsigned char s_lds[BinLds<kFrameBinBlock>::kBytes];
    const uint32_t nb = frame_bin_blocks(p.nf);
    if (blockIdx.x < nb)
        bin_body<kFrameBinBlock>(p, blockIdx.x, s_lds);
    else if (blockIdx.x < nb + p.zero_blocks)
        zero_lines_body<kFrameBinBlock>(p, blockIdx.x - nb);
    else
        clear_body<kFrameBinBlock>(p, blockIdx.x - nb - p.zero_blocks, p.clear_blocks);
}
// The raster and overflow-shade passes are driven by the per-XCD lists of tiles that
// received a face (built by the bin pass): a resident grid of workgroups, a multiple of
// 8, where workgroup b takes the tiles of list b % 8 (run on XCD b % 8) with stride
// gridDim.x / 8.  Empty tiles cost no workgroup dispatch at all: dispatching one
// workgroup per tile measured ~27 us for 8,160 tiles at 1080p, longer than the work.
// When a bin overflowed, each raster workgroup then takes its share of the faces for
// the overflow loop (the tile's visibility words were cleared by the bin pass and the
// raster loop skipped it), so the frame needs no launch of its own for it.
// Statistics are per-workgroup partial sums.
// (minimum waves per SIMD in the launch bounds: 6 for the 256-thread raster and fused
// kernels, 8 for the 64-thread raster kernel, 3 for the fused learned shader: the most
// the 16-pixel-tile kernels reach without register spills; the sample-collecting and
// timing modes and other tile sizes spill a few registers at these bounds)
// A resident workgroup's walk over its XCD's tile list: list entries k0, k0 + step, ...
// below the list's count.  The first entry is read beside the count rather than after
// it, one dependent memory round trip less before the first tile starts (an entry past
// the count is never used; k0 < ntiles keeps the read inside the list).
// Whether the bin pass overflowed a tile (the count's low word: fewer than 2^32 tiles),
// read early like the list, held as a uniform flag.
__device__ __forceinline__ bool overflowed_tiles(const RasterParams& p) {
    const uint32_t lo = issue_now(reinterpret_cast<const uint32_t*>(p.counters + kCntOverflowTiles));
    return __builtin_amdgcn_readfirstlane(lo) != 0u;
}
struct ListWalk {
    const uint32_t* list;
    uint32_t k0, step, n, first;
    uint32_t z;   // an opaque copy of 0 (tile_head)
    __device__ __forceinline__ explicit ListWalk(const RasterParams& p) {
        const uint32_t xcd = blockIdx.x & 7u;
        list = p.lists + (size_t)xcd * p.ntiles;
        k0 = blockIdx.x >> 3;
        step = gridDim.x >> 3;
        first = issue_now(list + (k0 < (uint32_t)p.ntiles ? k0 : 0u));
        n = issue_now(p.list_count + xcd * kCountStride);
        z = 0u;
        asm volatile("" : "+v"(z));
    }
    // Entry k, read now into a vector register; past the list's end, entry k0 (read only
    // where k0 < n).
    __device__ __forceinline__ uint32_t entry(uint32_t k) const { return issue_now(list + (k < n ? k : k0) + z); }
};

// A workgroup's tiles in list order, their reads pipelined: while tile k is rasterised,
// the head (count, first bin entries) of tile k + step and the list entry of tile
// k + 2 step are in flight, so a tile starts with its head in registers and its first
// wait is for its face records alone.  (The reads go out beside the current tile's own
// first reads, whose wait then covers them.)
template <typename Body>
__device__ __forceinline__ void walk_tiles(const RasterParams& p, const ListWalk& w, Body body) {
    if (w.k0 >= w.n) return;
    const int tid = threadIdx.x;
    uint32_t t = w.first;
    TileHead h = tile_head(p, (int)t, tid);
    uint32_t t1 = w.k0 + w.step < w.n ? w.entry(w.k0 + w.step) : t;
    for (uint32_t k = w.k0; k < w.n; k += w.step) {
        // (nothing is read ahead for a workgroup's last tile, the only one of most
        // workgroups in a frame of few tiles)
        uint32_t t2 = t1;
        TileHead h1 = h;
        if (k + w.step < w.n) {
            t2 = w.entry(k + 2 * w.step);
            h1 = tile_head(p, (int)t1, tid, w.z);
        }
        body((int)t, h);
        t = t1; h = h1; t1 = t2;
    }
}

template <int TS, int BLOCK>
__global__ __launch_bounds__(BLOCK, BLOCK == 256 ? 6 : 8) void k_raster(RasterParams p) {
    const ListWalk w(p);
    const bool overflow = overflowed_tiles(p);
    uint32_t frags = 0, cov = 0;
    walk_tiles(p, w, [&](int tile, const TileHead& h) {
        raster_body<TS, BLOCK, false, kShadePlain, kSlabPixels, false, kSlabAlways, true>(
            p, nullptr, nullptr, nullptr, nullptr, tile, h, frags, cov);
    });
    if (overflow)
        frags += overflow_faces(p, blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK);
    store_block_sum(frags, p.part_frag + blockIdx.x, threadIdx.x, BLOCK);
}
// Raster and shade in one pass.
// (the normalisation tables are read through the caches here: their LDS copy would cost
// a resident workgroup per CU)
template <int TS, int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK, MODE == kShadeMlp ? 3 : 6) void k_raster_shade(RasterParams p, ShadeParams sp) {
    const float* s_lut = sp.luts;
    __shared__ __align__(16) uint16_t s_net[MODE == kShadeMlp ? kPackedBytes / 2 : 8];
    __shared__ DevMaterial s_mats[kLdsMats];
    // the materials' staging reads, the list reads and the overflow count go out together:
    // one memory round trip before the first tile instead of three
    const MatStage<BLOCK> mats(sp, threadIdx.x);
    const ListWalk w(p);
    const bool overflow = overflowed_tiles(p);
    if (MODE == kShadeMlp) stage_net(sp.mlp, s_net, threadIdx.x, BLOCK);
    mats.store(s_mats, threadIdx.x);
    // (raster_body's first barrier orders these stores before any read)
    uint32_t frags = 0, cov = 0;
    walk_tiles(p, w, [&](int tile, const TileHead& h) {
        raster_body<TS, BLOCK, true, MODE, kSlabPixels, false, kSlabAlways, true>(p, &sp, s_lut, s_mats, s_net,
                                                                                  tile, h, frags, cov);
    });
    if (overflow)
        frags += overflow_faces(p, blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK);
    __syncthreads();
    store_block_sum(frags, p.part_frag + blockIdx.x, threadIdx.x, BLOCK);
    store_block_sum(cov, sp.part_cov + blockIdx.x, threadIdx.x, BLOCK);
}
template <int TS, int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK) void k_shade_tiles(RasterParams p, ShadeParams sp) {
    uint32_t cov = 0;
    if (p.counters[kCntOverflowTiles] != 0ull) {   // else the raster pass shaded every tile
        __shared__ float s_lut[kLdsLuts * 256];
        __shared__ __align__(16) uint16_t s_net[MODE == kShadeMlp ? kPackedBytes / 2 : 8];
        __shared__ DevMaterial s_mats[kLdsMats];
        stage_luts(sp, s_lut, threadIdx.x, BLOCK);
        stage_mats(sp, s_mats, threadIdx.x, BLOCK);
        if (MODE == kShadeMlp) stage_net(sp.mlp, s_net, threadIdx.x, BLOCK);
        __syncthreads();
        // (the raster pass shaded every other tile)
        const uint32_t n = (uint32_t)p.counters[kCntOverflowTiles];
        for (uint32_t k = blockIdx.x; k < n; k += gridDim.x)
            shade_tile_body<TS, BLOCK, MODE>(p, sp, s_lut, s_mats, s_net, (int)(p.ovf_list[k] & kListTileMask), cov);
    }
    store_block_sum(cov, sp.part_cov + blockIdx.x, threadIdx.x, BLOCK);
}
template <int MODE>
__global__ __launch_bounds__(256) void k_shade(ShadeParams p) {
    __shared__ float s_lut[kLdsLuts * 256];
    __shared__ __align__(16) uint16_t s_net[MODE == kShadeMlp ? kPackedBytes / 2 : 8];
    __shared__ DevMaterial s_mats[kLdsMats];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    stage_luts(p, s_lut, tid, 256);
    stage_mats(p, s_mats, tid, 256);
    if (MODE == kShadeMlp) stage_net(p.mlp, s_net, tid, 256);
    __syncthreads();
    const XcdGroups g{1, 16, shade_blocks_x(p.W), shade_blocks_y(p.H)};
    for (uint32_t u = blockIdx.x, n = g.grid(); u < n; u += gridDim.x) {
        const int blk = g.unit(u);
        if (blk >= 0) shade_body<MODE>(p, s_lut, s_mats, s_net, blk);
    }
}

__global__ __launch_bounds__(256) void k_vertex_batch(const VertexParams* __restrict__ items,
                                                      const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    vertex_body(items[m.x], (int)m.y);
}
// Batch bin launch: each mesh's binning workgroups, then its clearing ones.
__global__ __launch_bounds__(256) void k_bin_batch(const BinParams* __restrict__ items, const uint2* __restrict__ map) {
    __shared__ __align__(16) unsigned char s_lds[BinLds<256>::kBytes];
    const uint2 m = map[blockIdx.x];
    const BinParams& p = items[m.x];
    const uint32_t nb = bin_blocks(p.nf);
    if (m.y < nb)
        bin_body<256>(p, (int)m.y, s_lds);
    else
        clear_body<256>(p, m.y - nb, p.clear_blocks);
}
// A batch's raster pass: one resident grid over the tile lists the meshes share, entry
// (mesh << kListMeshShift) | tile, raster and shade of the tile with its mesh's
// parameters.  A batch's tiles differ in cost by orders of magnitude (overdraw), so
// workgroups take tiles one at a time from per-list cursors (the second word of each
// list counter's line): first list b % 8, then the others in turn.  Every cursor only
// grows, so each workgroup leaves after one failed take per list.
template <int TS, int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK, MODE == kShadeMlp ? 3 : 6) void k_raster_shade_batch(
    const RasterParams* __restrict__ items, const ShadeParams* __restrict__ sitems, BatchLists bl) {
    __shared__ __align__(16) uint16_t s_net[MODE == kShadeMlp ? kPackedBytes / 2 : 8];
    __shared__ uint32_t s_take;
    if (MODE == kShadeMlp) stage_net(sitems[0].mlp, s_net, threadIdx.x, BLOCK);   // one network per batch
    uint32_t frags = 0, cov = 0;
    for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t x = (blockIdx.x + j) & 7u;
        const uint32_t n = bl.list_count[x * kCountStride];
        uint32_t* cursor = bl.list_count + x * kCountStride + 1;
        for (;;) {
            if (threadIdx.x == 0) s_take = atomicAdd(cursor, 1u);
            __syncthreads();
            const uint32_t k = s_take;
            __syncthreads();   // (s_take is rewritten by the next take)
            if (k >= n) break;
            const uint32_t e = bl.lists[(size_t)x * bl.stride + k];
            const uint32_t m = e >> kListMeshShift;
            const int tile = (int)(e & kListTileMask);
            raster_body<TS, BLOCK, true, MODE, kShrinkPixels, false, kSlabAlways>(
                items[m], &sitems[m], sitems[m].luts, nullptr, s_net, tile, tile_head(items[m], tile, threadIdx.x),
                frags, cov);
        }
    }
    __syncthreads();
    store_block_sum(frags, bl.part_frag + blockIdx.x, threadIdx.x, BLOCK);
    store_block_sum(cov, bl.part_cov + blockIdx.x, threadIdx.x, BLOCK);
}
__global__ __launch_bounds__(256) void k_overflow_batch(const RasterParams* __restrict__ items,
                                                        const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    const RasterParams& p = items[m.x];
    overflow_body(p, (int)m.y, (int)overflow_blocks(p.nf));
}
// Shade of the batch's overflowed tiles (after the overflow loop), as k_shade_tiles;
// `part_cov` is the section after the raster launch's.
template <int TS, int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK) void k_shade_tiles_batch(const RasterParams* __restrict__ items,
                                                             const ShadeParams* __restrict__ sitems, BatchLists bl,
                                                             uint32_t* part_cov) {
    uint32_t cov = 0;
    if (*bl.overflow != 0ull) {
        __shared__ __align__(16) uint16_t s_net[MODE == kShadeMlp ? kPackedBytes / 2 : 8];
        if (MODE == kShadeMlp) stage_net(sitems[0].mlp, s_net, threadIdx.x, BLOCK);
        __syncthreads();
        const uint32_t n = (uint32_t)*bl.overflow;
        for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
            const uint32_t e = bl.ovf_list[k];
            const uint32_t m = e >> kListMeshShift;
            shade_tile_body<TS, BLOCK, MODE>(items[m], sitems[m], sitems[m].luts, nullptr, s_net,
                                             (int)(e & kListTileMask), cov);
        }
    }
    store_block_sum(cov, part_cov + blockIdx.x, threadIdx.x, BLOCK);
}

// ---------------------------------------------------------------------------------
// A batch frame as one persistent launch (BatchSync in rz_device.h).

// Spins of a barrier wait before it gives up (about a second, far beyond a batch frame):
// reached only if the launch's workgroups are not all resident at once.
constexpr uint32_t kMaxSpins = 1u << 20;
constexpr uint32_t kBarStride = kCountStride / 2;   // 64-bit words between barrier counters

// Arrives at barrier b (after releasing this workgroup's writes at device scope) and,
// with wait, waits until each of the barrier's 8 counters reaches `target`, then
// acquires what the arrivals wrote.  A wait that gives up flags q.error, after which
// every wait of the launch gives up at once.  Every thread calls it.
__device__ __forceinline__ void batch_barrier(const BatchSync& q, int b, unsigned long long target, bool wait) {
    __threadfence();
    __syncthreads();
    unsigned long long* bar = q.bar + (size_t)b * 8 * kBarStride;
    if (threadIdx.x == 0) atomicAdd(bar + (blockIdx.x & 7u) * kBarStride, 1ull);
    if (!wait) return;
    if (threadIdx.x == 0) {
        for (uint32_t x = 0; x < 8; ++x) {
            uint32_t spins = 0;
            while (__hip_atomic_load(bar + x * kBarStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++spins > kMaxSpins || __hip_atomic_load(q.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    atomicOr(q.error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
    }
    __syncthreads();
    __threadfence();   // acquire: the other workgroups' writes, from any XCD
    __syncthreads();
}

__device__ __forceinline__ void phase_clock(const BatchSync& q, int phase, bool start) {
    if (q.clock && threadIdx.x == 0) {
        const unsigned long long now = wall_clock64();
        if (start)
            atomicMin(&q.clock[2 * phase], now);
        else
            atomicMax(&q.clock[2 * phase + 1], now);
    }
}

template <int BLOCK, int MODE>
struct BatchFrameLds {
    static constexpr int kBytes =
        BinLds<BLOCK>::kBytes > RasterLds<BLOCK, MODE>::kBytes ? BinLds<BLOCK>::kBytes : RasterLds<BLOCK, MODE>::kBytes;
};

// bitems / ritems / oitems: nmesh parameter records of generation 0, then of generation 1
// (their counter lines differ); sitems: one per mesh.
template <int TS, int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK, MODE == kShadeMlp ? 3 : 4) void k_batch_frame(
    const BinParams* __restrict__ bitems, const RasterParams* __restrict__ ritems,
    const RasterParams* __restrict__ oitems, const ShadeParams* __restrict__ sitems, BatchLists bl0,
    BatchLists bl1, BatchSync q) {
    // one LDS region for the bin phase's tile aggregation and the raster phase's face data
    __shared__ __align__(16) unsigned char s_lds[BatchFrameLds<BLOCK, MODE>::kBytes];
    __shared__ __align__(16) uint16_t s_net[MODE == kShadeMlp ? kPackedBytes / 2 : 8];
    __shared__ unsigned long long s_frame;
    __shared__ uint32_t s_take;
    if (MODE == kShadeMlp) stage_net(sitems[0].mlp, s_net, threadIdx.x, BLOCK);   // one network per batch
    // this frame's number: every workgroup of the previous frames has arrived at barrier 2
    if (threadIdx.x == 0) {
        unsigned long long arrived = 0;
        for (uint32_t x = 0; x < 8; ++x)
            arrived += __hip_atomic_load(q.bar + (size_t)(2 * 8 + x) * kBarStride, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        s_frame = arrived / q.G;
    }
    __syncthreads();
    const unsigned long long f = s_frame;
    const unsigned long long target = (f + 1) * (q.G / 8u);
    const uint32_t gen = (uint32_t)(f & 1ull);
    const BatchLists& bl = gen ? bl1 : bl0;
    const uint32_t nm = q.nmesh;
    const uint32_t w = blockIdx.x;

    // phase 1: binning (with the vertex transform), the other generation's zeroing, clears.
    // Items are taken one at a time from one cursor (the 256-mesh batch's frame: 4.00 ms
    // with static striding, which left workgroups idle behind the slowest items, 3.57-3.61
    // ms this way; 3.67-3.69 ms with the next take issued as an item starts, since the
    // item's first wait for its own loads then also waits for the atomic; 3.74, 4.00 and
    // 4.29 ms taking 2, 4 or 8 consecutive items).
    phase_clock(q, 0, true);
    if (threadIdx.x == 0) s_take = atomicAdd(q.cursor, 1u);
    __syncthreads();
    uint32_t i = s_take;
    __syncthreads();   // (s_take is rewritten below)
    while (i < q.nB) {
        const uint2 m = q.bmap[i];
        const BinParams& p = bitems[gen * nm + m.x];
        const uint32_t nb = bin_blocks(p.nf, BLOCK);
        const uint32_t nz = (p.zero_count + BLOCK - 1) / BLOCK;
        if (m.y < nb)
            bin_body<BLOCK>(p, (int)m.y, s_lds);
        else if (m.y < nb + nz)
            zero_lines_body<BLOCK>(p, m.y - nb);
        else
            clear_body<BLOCK>(p, m.y - nb - nz, p.clear_blocks);
        __syncthreads();   // (s_lds is reused by the next item)
        if (threadIdx.x == 0) s_take = atomicAdd(q.cursor, 1u);
        __syncthreads();
        i = s_take;
        __syncthreads();
    }
    phase_clock(q, 0, false);
    batch_barrier(q, 0, target, true);
    // every take of this frame is done: the cursor starts the next frame at zero
    if (w == 0 && threadIdx.x == 0) __hip_atomic_store(q.cursor, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // phase 2: tiles of the shared lists, one at a time from per-list cursors (first list
    // w % 8, then the others)
    uint32_t frags = 0, cov = 0, cov_ovf = 0;
    phase_clock(q, 1, true);
    for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t x = (w + j) & 7u;
        const uint32_t n = bl.list_count[x * kCountStride];
        uint32_t* cursor = bl.list_count + x * kCountStride + 1;
        for (;;) {
            if (threadIdx.x == 0) s_take = atomicAdd(cursor, 1u);
            __syncthreads();
            const uint32_t k = s_take;
            __syncthreads();   // (s_take is rewritten by the next take)
            if (k >= n) break;
            const uint32_t e = bl.lists[(size_t)x * bl.stride + k];
            const uint32_t m = e >> kListMeshShift;
            const int tile = (int)(e & kListTileMask);
            raster_body<TS, BLOCK, true, MODE, kShrinkPixels, true, kSlabAlways>(
                ritems[gen * nm + m], &sitems[m], sitems[m].luts, nullptr, s_net, tile,
                tile_head(ritems[gen * nm + m], tile, threadIdx.x), frags, cov, s_lds);
        }
    }
    phase_clock(q, 1, false);

    // phases 3 and 4, when a bin overflowed (uniform: the count is final after barrier 0)
    const bool overflow = __builtin_amdgcn_readfirstlane(
                              (uint32_t)(__hip_atomic_load(const_cast<unsigned long long*>(bl.overflow),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull)) != 0u;
    if (overflow) {
        // only the meshes with an overflowed tile (their bits in LDS, read once): the
        // others' overflow statistics are left unwritten and not summed
        phase_clock(q, 2, true);
        __shared__ uint32_t s_ovf_mesh[(kMaxBatchMeshes + 31) / 32];
        for (uint32_t k = threadIdx.x; k < (nm + 31) / 32; k += BLOCK) s_ovf_mesh[k] = 0u;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nm; k += BLOCK)
            if (oitems[gen * nm + k].counters[kCntOverflowTiles] != 0ull) atomicOr(&s_ovf_mesh[k >> 5], 1u << (k & 31u));
        __syncthreads();
        for (uint32_t wd = 0; wd < (nm + 31) / 32; ++wd) {
            uint32_t bits = s_ovf_mesh[wd];   // (uniform)
            while (bits) {
                const uint32_t k = wd * 32u + (uint32_t)(__ffs(bits) - 1);
                bits &= bits - 1u;
                const RasterParams& p = oitems[gen * nm + k];
                const uint32_t nblk = overflow_blocks(p.nf, BLOCK);
                for (uint32_t b = w; b < nblk; b += q.G) overflow_body<BLOCK>(p, (int)b, (int)nblk);
            }
        }
        phase_clock(q, 2, false);
    }
    batch_barrier(q, 1, target, overflow);
    if (overflow) {
        phase_clock(q, 3, true);
        const uint32_t n = (uint32_t)__hip_atomic_load(const_cast<unsigned long long*>(bl.overflow),
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (uint32_t k = w; k < n; k += q.G) {
            const uint32_t e = bl.ovf_list[k];
            const uint32_t m = e >> kListMeshShift;
            shade_tile_body<TS, BLOCK, MODE>(ritems[gen * nm + m], sitems[m], sitems[m].luts, nullptr, s_net,
                                             (int)(e & kListTileMask), cov_ovf);
        }
        phase_clock(q, 3, false);
    }
    __syncthreads();
    store_block_sum(frags, bl.part_frag + w, threadIdx.x, BLOCK);
    store_block_sum(cov, bl.part_cov + w, threadIdx.x, BLOCK);
    store_block_sum(cov_ovf, bl.part_cov + q.G + w, threadIdx.x, BLOCK);
    batch_barrier(q, 2, target, false);   // the next frame's launch counts these arrivals
}

// Workgroups of `f` (BLOCK threads) that a CU holds at once, for a launch whose
// workgroups wait on each other: the runtime's occupancy, bounded by this count of the
// kernel's own resources with margin: waves by VGPRs (512 per lane per SIMD, granules of
// 8, at most 8 waves per SIMD) and workgroups by LDS within 7/8 of the CU's 160 KiB.
// (A workgroup of 7,896 LDS bytes was counted 20 to a CU, 160 KiB once rounded to
// 512-byte granules, and fewer were resident.)  A smaller grid is only slower; a grid
// beyond residency leaves a barrier waiting on workgroups that cannot start.
// lds_budget: the LDS bytes per CU counted on (the single-frame raster passes, whose
// workgroups never wait on each other, count on 156 KiB: the 4M-triangle raster-only
// pass with 7,496-byte workgroups measured 84.3, 81.9, 80.5, 79.1, 80.6 and 103.7 µs at
// 16, 18, 19, 20, 21 and 22 workgroups per CU, the last beyond residency).
// dyn_lds: the launch's dynamic LDS bytes.
uint32_t resident_blocks_per_cu(const void* f, int block, size_t lds_budget = (size_t)163840 * 7 / 8,
                                size_t dyn_lds = 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, dyn_lds) != hipSuccess || n < 1) n = 1;
    hipFuncAttributes at{};
    if (hipFuncGetAttributes(&at, f) != hipSuccess) return 1;
    const int vgprs = ((at.numRegs > 0 ? at.numRegs : 512) + 7) / 8 * 8;
    const int waves_per_simd = std::min(8, 512 / vgprs);
    const int waves_per_block = (block + 63) / 64;
    const int by_waves = 4 * waves_per_simd / waves_per_block;
    const size_t lds = ((size_t)at.sharedSizeBytes + dyn_lds + 511) / 512 * 512;
    const int by_lds = lds > 0 ? (int)(lds_budget / lds) : by_waves;
    n = std::min(n, std::min(by_waves, by_lds));
    return (uint32_t)(n < 1 ? 1 : n);
}

// Resident workgroups per CU of the persistent batch launch.
uint32_t batch_frame_blocks_per_cu(int tile_size, int mode) {
    const void* f = nullptr;
    int block = 256;
#define RZ_PICK(TS, B)                                                                               \
    block = B;                                                                                       \
    f = mode == kShadeMlp ? reinterpret_cast<const void*>(&k_batch_frame<TS, B, kShadeMlp>)           \
                          : reinterpret_cast<const void*>(&k_batch_frame<TS, B, kShadePlain>);
    switch (tile_size) {
        case 8: RZ_PICK(8, 64) break;
        case 16: RZ_PICK(16, 256) break;
        default: RZ_PICK(32, 256) break;
    }
#undef RZ_PICK
    return resident_blocks_per_cu(f, block);
}

hipError_t launch_batch_frame(const BinParams* b, const RasterParams* r, const RasterParams* o,
                              const ShadeParams* sh, const BatchLists& bl0, const BatchLists& bl1,
                              const BatchSync& q, int tile_size, int mode, hipStream_t s) {
#define RZ_LAUNCH(TS, B, M) \
    hipLaunchKernelGGL((k_batch_frame<TS, B, M>), dim3(q.G), dim3(B), 0, s, b, r, o, sh, bl0, bl1, q)
    const bool net = mode == kShadeMlp;
    if (mode != kShadePlain && mode != kShadeMlp) return hipErrorInvalidValue;
    switch (tile_size) {
        case 8: if (net) RZ_LAUNCH(8, 64, kShadeMlp); else RZ_LAUNCH(8, 64, kShadePlain); break;
        case 16: if (net) RZ_LAUNCH(16, 256, kShadeMlp); else RZ_LAUNCH(16, 256, kShadePlain); break;
        case 32: if (net) RZ_LAUNCH(32, 256, kShadeMlp); else RZ_LAUNCH(32, 256, kShadePlain); break;
        default: return hipErrorInvalidValue;
    }
#undef RZ_LAUNCH
    return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_hold(uint64_t ticks) {
    const uint64_t t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(32);
}

// ---------------------------------------------------------------------------------
// Launchers (host side).

// Device-to-device copy of n16 16-byte words (one read and one write per byte).
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s) {
    if (n16 == 0) return hipSuccess;
    const size_t blocks = (n16 + 255) / 256;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256), 0, s,
                       (const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Text into a planar frame: the spin loop's frame-rate line (renderer.cpp:114,
// CImg::draw_text(5, 5, text, white, black)).  Glyphs are this library's own 5x7 bitmaps
// (digits, '.', '-', ' ', and the letters of "inf" and "nan"), drawn at twice their size
// in 12x16 cells: foreground where a glyph bit is set, background over the rest of the
// cell; characters without a glyph draw as background.
constexpr int kTextMax = 32;
struct TextParams {
    uint8_t* frame;          // planar [3][H][W]
    uint32_t W, H;
    int32_t x0, y0;
    int32_t len;
    uint8_t fg[3], bg[3];
    char text[kTextMax];
};

__device__ __forceinline__ uint8_t glyph_row(char c, int row) {
    // rows top to bottom, bit 4 = leftmost column
    static constexpr uint8_t kDigits[10][7] = {
        {0x0E, 0x11, 0x13, 0x15, 0x19, 0x11, 0x0E}, {0x04, 0x0C, 0x04, 0x04, 0x04, 0x04, 0x0E},
        {0x0E, 0x11, 0x01, 0x02, 0x04, 0x08, 0x1F}, {0x1F, 0x02, 0x04, 0x02, 0x01, 0x11, 0x0E},
        {0x02, 0x06, 0x0A, 0x12, 0x1F, 0x02, 0x02}, {0x1F, 0x10, 0x1E, 0x01, 0x01, 0x11, 0x0E},
        {0x06, 0x08, 0x10, 0x1E, 0x11, 0x11, 0x0E}, {0x1F, 0x01, 0x02, 0x04, 0x08, 0x08, 0x08},
        {0x0E, 0x11, 0x11, 0x0E, 0x11, 0x11, 0x0E}, {0x0E, 0x11, 0x11, 0x0F, 0x01, 0x02, 0x0C}};
    static constexpr uint8_t kDot[7] = {0, 0, 0, 0, 0, 0x0C, 0x0C};
    static constexpr uint8_t kMinus[7] = {0, 0, 0, 0x1F, 0, 0, 0};
    static constexpr uint8_t kI[7] = {0x04, 0, 0x0C, 0x04, 0x04, 0x04, 0x0E};
    static constexpr uint8_t kN[7] = {0, 0, 0x16, 0x19, 0x11, 0x11, 0x11};
    static constexpr uint8_t kF[7] = {0x06, 0x09, 0x08, 0x1C, 0x08, 0x08, 0x08};
    static constexpr uint8_t kA[7] = {0, 0, 0x0E, 0x01, 0x0F, 0x11, 0x0F};
    if (c >= '0' && c <= '9') return kDigits[c - '0'][row];
    switch (c) {
        case '.': return kDot[row];
        case '-': return kMinus[row];
        case 'i': return kI[row];
        case 'n': return kN[row];
        case 'f': return kF[row];
        case 'a': return kA[row];
        default: return 0;
    }
}

// One thread per pixel of the text's cells (12 * len by 16), clipped to the frame.
__global__ __launch_bounds__(256) void k_text(TextParams p) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    const int bw = 12 * p.len;
    if (i >= bw * 16) return;
    const int dx = i % bw, dy = i / bw;
    const int64_t x = (int64_t)p.x0 + dx, y = (int64_t)p.y0 + dy;
    if (x < 0 || y < 0 || x >= (int64_t)p.W || y >= (int64_t)p.H) return;
    const int col = (dx % 12) >> 1, row = (dy >> 1) - 1;   // glyph rows 1..7 of the cell's 8
    const bool on = col < 5 && row >= 0 && row < 7 && ((glyph_row(p.text[dx / 12], row) >> (4 - col)) & 1);
    const size_t np = (size_t)p.W * p.H, q = (size_t)y * p.W + (size_t)x;
    const uint8_t* c = on ? p.fg : p.bg;
    p.frame[q] = c[0];
    p.frame[q + np] = c[1];
    p.frame[q + 2 * np] = c[2];
}

hipError_t launch_text(uint8_t* frame, uint32_t W, uint32_t H, int32_t x0, int32_t y0, const char* text,
                       const uint8_t fg[3], const uint8_t bg[3], hipStream_t s) {
    TextParams p{};
    p.frame = frame;
    p.W = W; p.H = H; p.x0 = x0; p.y0 = y0;
    int n = 0;
    while (text[n] && n < kTextMax) {
        p.text[n] = text[n];
        ++n;
    }
    p.len = n;
    for (int c = 0; c < 3; ++c) { p.fg[c] = fg[c]; p.bg[c] = bg[c]; }
    if (n == 0) return hipSuccess;
    const int threads = 12 * n * 16;
    hipLaunchKernelGGL(k_text, dim3((threads + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

// Self-check of quotient_in_range: random operands (every sign, exponents one beyond
// each end of the ranges, mantissas random or all ones or zero), each pair in range
// compared bit for bit with the IEEE division.  counts[0] += pairs compared,
// counts[1] += mismatches.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

__global__ void k_check_quotients(uint32_t per_thread, uint32_t seed, unsigned long long* counts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long tested = 0, bad = 0;
    for (uint32_t k = 0; k < per_thread; ++k) {
        const uint32_t h1 = mix32(t * 0x9E3779B9u + k * 0x85EBCA6Bu + seed);
        const uint32_t h2 = mix32(h1 ^ 0xC2B2AE35u), h3 = mix32(h2 + 0x27D4EB2Fu);
        const int ed = -41 + (int)(h3 % 83u), en = -21 + (int)((h3 >> 8) % 63u);
        uint32_t md = h1 & 0x7FFFFFu, mn = h2 & 0x7FFFFFu;
        if ((h3 >> 16) % 16u == 0) md = 0x7FFFFFu;
        if ((h3 >> 20) % 16u == 1) mn = 0x7FFFFFu;
        if ((h3 >> 24) % 16u == 2) md = 0;
        if ((h3 >> 28) % 16u == 3) mn = 0;
        const float d = __uint_as_float((h1 & 0x80000000u) | ((uint32_t)(ed + 127) << 23) | md);
        const float n = __uint_as_float((h2 & 0x80000000u) | ((uint32_t)(en + 127) << 23) | mn);
        const float r1 = refined_rcp(d);
        if (r1 == 0.0f || !dividend_in_range(n)) continue;
        ++tested;
        bad += __float_as_uint(quotient_in_range(n, d, r1)) != __float_as_uint(n / d) ? 1u : 0u;
    }
    atomicAdd(&counts[0], tested);
    if (bad) atomicAdd(&counts[1], bad);
}

hipError_t launch_check_quotients(uint32_t blocks, uint32_t per_thread, uint32_t seed, unsigned long long* counts,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_check_quotients, dim3(blocks), dim3(256), 0, s, per_thread, seed, counts);
    return hipGetLastError();
}

hipError_t launch_hold(uint64_t ticks, hipStream_t s) {
    hipLaunchKernelGGL(k_hold, dim3(1), dim3(64), 0, s, ticks);
    return hipGetLastError();
}

hipError_t launch_bin(const BinParams& p, hipStream_t s) {
    const uint32_t blocks = frame_bin_blocks(p.nf) + p.zero_blocks + (p.clear_np ? p.clear_blocks : 0u);
    if (blocks) hipLaunchKernelGGL(k_bin, dim3(blocks), dim3(kFrameBinBlock), 0, s, p);
    return hipGetLastError();
}

// Resident workgroups per CU of the list-driven raster kernel for this configuration.
// depth_in: the frame stages an input depth buffer (the kernels' dynamic LDS).
uint32_t raster_blocks_per_cu(int tile_size, bool fused, int mode, bool narrow, bool depth_in) {
    const void* f = nullptr;
    int block = 256;
#define RZ_PICK(TS, B)                                                                              \
    block = B;                                                                                      \
    if (!fused)                                                                                     \
        f = reinterpret_cast<const void*>(&k_raster<TS, B>);                                         \
    else if (mode == kShadeMlp)                                                                     \
        f = reinterpret_cast<const void*>(&k_raster_shade<TS, B, kShadeMlp>);                        \
    else if (mode == kShadeCollect)                                                                 \
        f = reinterpret_cast<const void*>(&k_raster_shade<TS, B, kShadeCollect>);                    \
    else if (mode == kShadeChain)                                                                   \
        f = reinterpret_cast<const void*>(&k_raster_shade<TS, B, kShadeChain>);                      \
    else                                                                                            \
        f = reinterpret_cast<const void*>(&k_raster_shade<TS, B, kShadePlain>);
    switch (tile_size) {
        case 8: RZ_PICK(8, 64) break;
        case 16:
            if (narrow) {
                RZ_PICK(16, 64)
            } else {
                RZ_PICK(16, 256)
            }
            break;
        default: RZ_PICK(32, 256) break;
    }
#undef RZ_PICK
    // (the walk over the tile lists is a static stride: a workgroup beyond residency
    // starts its tiles only when another workgroup ends, doubling the launch's tail)
    const uint32_t n = resident_blocks_per_cu(f, block, (size_t)163840 - 4096,
                                              depth_in ? (size_t)tile_size * tile_size * sizeof(float) : 0u);
    // One-wave workgroups: at most 20 to a CU, whatever the registers and LDS allow (the
    // 4M-triangle raster-only pass, 6,472 LDS bytes and 71 VGPRs per workgroup, measured
    // 81.7, 79.1, 80.1, 103.6, 102.8 and 99.8 µs at 18, 20, 21, 22, 23 and 24 per CU; the
    // fused pass 161.5, 155.8, 155.4, 199.2 at 18, 20, 21, 22).
    return block == 64 && n > 20u ? 20u : n;
}

// The initial depths' dynamic LDS of a frame with an input depth buffer.
__host__ __forceinline__ size_t d0_lds_bytes(const RasterParams& p, int tile_size) {
    return p.depth_in ? (size_t)tile_size * tile_size * sizeof(float) : 0u;
}

hipError_t launch_raster(const RasterParams& p, int tile_size, hipStream_t s) {
    const uint32_t grid = p.max_blocks;
    const size_t dyn = d0_lds_bytes(p, tile_size);
    switch (tile_size) {
        case 8: hipLaunchKernelGGL((k_raster<8, 64>), dim3(grid), dim3(64), dyn, s, p); break;
        case 16:
            if (p.narrow)
                hipLaunchKernelGGL((k_raster<16, 64>), dim3(grid), dim3(64), dyn, s, p);
            else
                hipLaunchKernelGGL((k_raster<16, 256>), dim3(grid), dim3(256), dyn, s, p);
            break;
        case 32: hipLaunchKernelGGL((k_raster<32, 256>), dim3(grid), dim3(256), dyn, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_raster_shade_mode(const RasterParams& p, const ShadeParams& sp, int tile_size, hipStream_t s) {
    const size_t dyn = d0_lds_bytes(p, tile_size);
    switch (tile_size) {
        case 8: {
            const uint32_t grid = p.max_blocks;
            hipLaunchKernelGGL((k_raster_shade<8, 64, MODE>), dim3(grid), dim3(64), dyn, s, p, sp);
            break;
        }
        case 16: {
            const uint32_t grid = p.max_blocks;
            if (p.narrow)
                hipLaunchKernelGGL((k_raster_shade<16, 64, MODE>), dim3(grid), dim3(64), dyn, s, p, sp);
            else
                hipLaunchKernelGGL((k_raster_shade<16, 256, MODE>), dim3(grid), dim3(256), dyn, s, p, sp);
            break;
        }
        case 32: {
            const uint32_t grid = p.max_blocks;
            hipLaunchKernelGGL((k_raster_shade<32, 256, MODE>), dim3(grid), dim3(256), dyn, s, p, sp);
            break;
        }
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_raster_shade(const RasterParams& p, const ShadeParams& sp, int tile_size, hipStream_t s) {
    switch (sp.mode) {
        case kShadePlain: return launch_raster_shade_mode<kShadePlain>(p, sp, tile_size, s);
        case kShadeMlp: return launch_raster_shade_mode<kShadeMlp>(p, sp, tile_size, s);
        case kShadeCollect: return launch_raster_shade_mode<kShadeCollect>(p, sp, tile_size, s);
        case kShadeChain: return launch_raster_shade_mode<kShadeChain>(p, sp, tile_size, s);
        default: return hipErrorInvalidValue;
    }
}

template <int MODE>
hipError_t launch_shade_tiles_mode(const RasterParams& p, const ShadeParams& sp, int tile_size, hipStream_t s) {
    const uint32_t grid = sp.max_blocks;
    switch (tile_size) {
        case 8: hipLaunchKernelGGL((k_shade_tiles<8, 64, MODE>), dim3(grid), dim3(64), 0, s, p, sp); break;
        case 16: hipLaunchKernelGGL((k_shade_tiles<16, 256, MODE>), dim3(grid), dim3(256), 0, s, p, sp); break;
        case 32: hipLaunchKernelGGL((k_shade_tiles<32, 256, MODE>), dim3(grid), dim3(256), 0, s, p, sp); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_shade_tiles(const RasterParams& p, const ShadeParams& sp, int tile_size, hipStream_t s) {
    switch (sp.mode) {
        case kShadePlain: return launch_shade_tiles_mode<kShadePlain>(p, sp, tile_size, s);
        case kShadeMlp: return launch_shade_tiles_mode<kShadeMlp>(p, sp, tile_size, s);
        case kShadeCollect: return launch_shade_tiles_mode<kShadeCollect>(p, sp, tile_size, s);
        case kShadeChain: return launch_shade_tiles_mode<kShadeChain>(p, sp, tile_size, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_shade(const ShadeParams& p, hipStream_t s) {
    const dim3 grid(XcdGroups{1, 16, shade_blocks_x(p.W), shade_blocks_y(p.H)}.grid()), block(64, 4);
    switch (p.mode) {
        case kShadePlain: hipLaunchKernelGGL(k_shade<kShadePlain>, grid, block, 0, s, p); break;
        case kShadeMlp: hipLaunchKernelGGL(k_shade<kShadeMlp>, grid, block, 0, s, p); break;
        case kShadeCollect: hipLaunchKernelGGL(k_shade<kShadeCollect>, grid, block, 0, s, p); break;
        case kShadeChain: hipLaunchKernelGGL(k_shade<kShadeChain>, grid, block, 0, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Batch launchers: `items` and `map` are device arrays, `blocks` the summed block
// count of the pass over all meshes.
hipError_t launch_vertex_batch(const VertexParams* items, const uint2* map, uint32_t blocks, hipStream_t s) {
    if (blocks) hipLaunchKernelGGL(k_vertex_batch, dim3(blocks), dim3(256), 0, s, items, map);
    return hipGetLastError();
}
hipError_t launch_bin_batch(const BinParams* items, const uint2* map, uint32_t blocks, hipStream_t s) {
    if (blocks) hipLaunchKernelGGL(k_bin_batch, dim3(blocks), dim3(256), 0, s, items, map);
    return hipGetLastError();
}
hipError_t launch_overflow_batch(const RasterParams* items, const uint2* map, uint32_t blocks, hipStream_t s) {
    if (blocks) hipLaunchKernelGGL(k_overflow_batch, dim3(blocks), dim3(256), 0, s, items, map);
    return hipGetLastError();
}

// Resident workgroups per CU of the batch raster kernel.
uint32_t raster_batch_blocks_per_cu(int tile_size, int mode) {
    const void* f = nullptr;
    int block = 256;
#define RZ_PICK(TS, B)                                                                              \
    block = B;                                                                                      \
    f = mode == kShadeMlp ? reinterpret_cast<const void*>(&k_raster_shade_batch<TS, B, kShadeMlp>)   \
                          : reinterpret_cast<const void*>(&k_raster_shade_batch<TS, B, kShadePlain>);
    switch (tile_size) {
        case 8: RZ_PICK(8, 64) break;
        case 16: RZ_PICK(16, 256) break;
        default: RZ_PICK(32, 256) break;
    }
#undef RZ_PICK
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, 0) != hipSuccess || n < 1) n = 1;
    return (uint32_t)n;
}

template <int MODE>
hipError_t launch_batch_tiles_mode(const RasterParams* items, const ShadeParams* sitems, const BatchLists& bl,
                                   uint32_t raster_grid, uint32_t tiles_grid, int tile_size, hipStream_t s,
                                   bool raster) {
    uint32_t* tiles_cov = bl.part_cov + raster_grid;
#define RZ_LAUNCH(TS, B)                                                                                          \
    if (raster)                                                                                                   \
        hipLaunchKernelGGL((k_raster_shade_batch<TS, B, MODE>), dim3(raster_grid), dim3(B), 0, s, items, sitems, bl); \
    else                                                                                                          \
        hipLaunchKernelGGL((k_shade_tiles_batch<TS, B, MODE>), dim3(tiles_grid), dim3(B), 0, s, items, sitems, bl,    \
                           tiles_cov);
    switch (tile_size) {
        case 8: RZ_LAUNCH(8, 64) break;
        case 16: RZ_LAUNCH(16, 256) break;
        case 32: RZ_LAUNCH(32, 256) break;
        default: return hipErrorInvalidValue;
    }
#undef RZ_LAUNCH
    return hipGetLastError();
}

// The batch's raster pass (raster = true) or its overflow shade; grids are multiples of 8.
hipError_t launch_batch_tiles(const RasterParams* items, const ShadeParams* sitems, const BatchLists& bl,
                              uint32_t raster_grid, uint32_t tiles_grid, int tile_size, int mode, hipStream_t s,
                              bool raster) {
    switch (mode) {
        case kShadePlain:
            return launch_batch_tiles_mode<kShadePlain>(items, sitems, bl, raster_grid, tiles_grid, tile_size, s, raster);
        case kShadeMlp:
            return launch_batch_tiles_mode<kShadeMlp>(items, sitems, bl, raster_grid, tiles_grid, tile_size, s, raster);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rz
