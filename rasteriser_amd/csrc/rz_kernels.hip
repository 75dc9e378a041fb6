// rz_kernels.hip — gfx950 kernels of the frame path.
//
//   k_vertex    SoA vertex stage: clip = camera*(p,1), z-divide, NDC->raster
//               (geometry.cpp:44-74, 85-95, 110-116, 135-142) and the normal
//               transform (geometry.cpp:35-42, 97-108); clears the per-frame bins.
//   k_bin       backface test (geometry.cpp:76-83, drawing.cpp:176-178), clamped
//               bbox (drawing.cpp:77-93), per-face setup record, and appends the face
//               to the bin of every tile its bbox meets (fixed capacity per tile).
//   k_raster    one workgroup per tile: expands (face, bbox pixel) candidates with a
//               wavefront prefix scan, runs the inside and depth tests of
//               drawing.cpp:96-121 and keeps per pixel the minimum 64-bit
//               (ordered depth << 32 | face) key in LDS; writes the visibility word.
//   k_overflow  faces of tiles whose bin overflowed, with global 64-bit atomicMin.
//   k_shade     deferred shade of the winning fragment: depth, perspective-correct
//               normal/uv, texture, Lambert (drawing.cpp:121-146, shading.cpp:20-34,
//               material.cpp:11-26, CImg _linear_atXY).
//
// The packed key makes the winner the lexicographic minimum of (depth, face index)
// over covered fragments with depth below the initial depth: the same fragment the
// reference's in-order strict `<` test (drawing.cpp:119, faces in file order) leaves.
#include "rz_device.h"
#include "rz_mlp.h"

namespace rz {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
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
__device__ __forceinline__ void vertex_body(const VertexParams& p, int blk) {
    const int i = blk * 256 + threadIdx.x;
    if (i < p.nv) {
        float4 r;
        if (p.raster_in) {
            r = p.raster_in[i];
        } else {
            const float x = p.px[i], y = p.py[i], z = p.pz[i];
            const float* m = p.camera;
            // GLM mat4*vec4 (pairwise), w = 1
            const float cx = (m[0] * x + m[4] * y) + (m[8] * z + m[12] * 1.0f);
            const float cy = (m[1] * x + m[5] * y) + (m[9] * z + m[13] * 1.0f);
            const float cz = (m[2] * x + m[6] * y) + (m[10] * z + m[14] * 1.0f);
            const float cw = (m[3] * x + m[7] * y) + (m[11] * z + m[15] * 1.0f);
            const float nx = cx / cw, ny = cy / cw, nz = cz / cw, nw = 1.f / cw;
            r.x = 0.5f * (nx + 1.0f) * p.Wf;
            r.y = 0.5f * (-ny + 1.0f) * p.Hf;
            r.z = nz;
            r.w = nw;
        }
        p.raster[i] = r;
    }
    if (i < p.nn) {
        const float x = p.nx[i], y = p.ny[i], z = p.nz[i];
        const float* m = p.normal;
        const float w = 0.0f;
        float4 n;
        n.x = (m[0] * x + m[4] * y) + (m[8] * z + m[12] * w);
        n.y = (m[1] * x + m[5] * y) + (m[9] * z + m[13] * w);
        n.z = (m[2] * x + m[6] * y) + (m[10] * z + m[14] * w);
        n.w = 0.0f;
        p.cnormal[i] = n;
    }
    if (i < p.ntiles) p.tile_count[i] = 0u;
    if (i < kNumCounters) p.counters[i] = 0ull;
}

// ---------------------------------------------------------------------------------
__device__ __forceinline__ void bin_push(const BinParams& p, uint32_t face, uint32_t tile) {
    const uint32_t slot = atomicAdd(&p.tile_count[tile], 1u);
    if (slot < p.cap)
        p.bins[(size_t)tile * p.cap + slot] = face;
    else if (slot == p.cap)
        atomicAdd(&p.counters[kCntOverflowTiles], 1ull);
}

__device__ __forceinline__ void bin_body(const BinParams& p, int blk) {
    const int f = blk * 256 + threadIdx.x;
    bool front = false;
    uint32_t tx0 = 1, tx1 = 0, ty0 = 1, ty1 = 0;   // empty tile rectangle
    if (f < p.nf) {
        const int i0 = p.fv[3 * f], i1 = p.fv[3 * f + 1], i2 = p.fv[3 * f + 2];
        const float4 a = p.raster[i0], b = p.raster[i1], c = p.raster[i2];
        const float s = signed_area(a.x, a.y, b.x, b.y, c.x, c.y);
        front = (s > 0.0f) != (p.wind_cw != 0);
        FaceRecord* r = p.rec + f;
        if (front) {
            uint32_t tlx, tly, brx, bry;
            bounding_box(a.x, a.y, b.x, b.y, c.x, c.y, p.W, p.H, tlx, tly, brx, bry);
            const float area = edge(c.x, c.y, a.x, a.y, b.x, b.y);
            r->r0 = make_float4(a.x, a.y, a.z, area);
            r->r1 = make_float4(b.x, b.y, b.z, __uint_as_float(tlx | (tly << 16)));
            r->r2 = make_float4(c.x, c.y, c.z, __uint_as_float(brx | (bry << 16)));
            r->r3 = make_float4(a.w, b.w, c.w, 0.0f);
            if (tlx <= brx && tly <= bry) {
                tx0 = tlx >> p.ts_log2; tx1 = brx >> p.ts_log2;
                ty0 = tly >> p.ts_log2; ty1 = bry >> p.ts_log2;
            }
        } else {
            r->r1 = make_float4(0.f, 0.f, 0.f, __uint_as_float(0xFFFFFFFFu));
            r->r2 = make_float4(0.f, 0.f, 0.f, __uint_as_float(0u));
        }
    }
    const uint32_t nt = (tx0 <= tx1 && ty0 <= ty1) ? (tx1 - tx0 + 1) * (ty1 - ty0 + 1) : 0u;
    store_block_sum(front ? 1u : 0u, p.part_front + blk, threadIdx.x, 256);
    store_block_sum(nt, p.part_pairs + blk, threadIdx.x, 256);
    const bool big = nt > (uint32_t)kBigTriangleTiles;
    const bool small = nt > 0 && !big;

    // Small faces: aggregate the workgroup's (tile, face) pairs in LDS so that each
    // distinct tile costs one device-scope atomic per workgroup, all issued at once.
    // Faces of one workgroup are neighbours in the mesh, so their tiles usually fall in
    // a small rectangle; when the union of their tile rectangles is larger than
    // kAggTiles the pairs go to the tile counters one by one.
    constexpr int kAggTiles = 1024;
    __shared__ int s_rect[4];                       // union tile rect: x0 y0 x1 y1
    __shared__ uint32_t s_cnt[kAggTiles];
    __shared__ uint32_t s_base[kAggTiles];
    __shared__ uint8_t s_slot[256 * kBigTriangleTiles];
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_rect[0] = s_rect[1] = 0x7FFFFFFF;
        s_rect[2] = s_rect[3] = -1;
    }
    __syncthreads();
    if (small) {
        atomicMin(&s_rect[0], (int)tx0); atomicMin(&s_rect[1], (int)ty0);
        atomicMax(&s_rect[2], (int)tx1); atomicMax(&s_rect[3], (int)ty1);
    }
    __syncthreads();
    const int ux0 = s_rect[0], uy0 = s_rect[1];
    const int uw = s_rect[2] - ux0 + 1, uh = s_rect[3] - uy0 + 1;
    const bool agg = s_rect[2] >= 0 && uw * uh <= kAggTiles;   // uniform over the workgroup
    if (agg) {
        const int area = uw * uh;
        for (int i = tid; i < area; i += 256) s_cnt[i] = 0u;
        __syncthreads();
        if (small) {
            int k = 0;
            for (uint32_t ty = ty0; ty <= ty1; ++ty)
                for (uint32_t tx = tx0; tx <= tx1; ++tx, ++k)
                    s_slot[tid * kBigTriangleTiles + k] =
                        (uint8_t)atomicAdd(&s_cnt[((int)ty - uy0) * uw + ((int)tx - ux0)], 1u);
        }
        __syncthreads();
        for (int i = tid; i < area; i += 256) {
            const uint32_t c = s_cnt[i];
            if (!c) continue;
            const uint32_t tile = (uint32_t)(uy0 + i / uw) * p.tiles_x + (uint32_t)(ux0 + i % uw);
            const uint32_t base = atomicAdd(&p.tile_count[tile], c);
            s_base[i] = base;
            if (base <= p.cap && base + c > p.cap) atomicAdd(&p.counters[kCntOverflowTiles], 1ull);
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
        const uint32_t bx0 = __shfl(tx0, src, 64), bx1 = __shfl(tx1, src, 64);
        const uint32_t by0 = __shfl(ty0, src, 64), by1 = __shfl(ty1, src, 64);
        const uint32_t bf = __shfl((uint32_t)f, src, 64);
        const uint32_t w = bx1 - bx0 + 1, n = w * (by1 - by0 + 1);
        for (uint32_t k = lane_id(); k < n; k += 64) {
            const uint32_t ty = by0 + k / w, tx = bx0 + k % w;
            bin_push(p, bf, ty * p.tiles_x + tx);
        }
    }
}

// ---------------------------------------------------------------------------------
// Inside + depth test of one (face, pixel) candidate; returns true when inside.
__device__ __forceinline__ bool fragment_key(float pxf, float pyf, float x0, float y0, float z0, float x1,
                                             float y1, float z1, float x2, float y2, float z2, float area,
                                             float& z) {
    const float e0 = edge(pxf, pyf, x1, y1, x2, y2);
    const float e1 = edge(pxf, pyf, x2, y2, x0, y0);
    const float e2 = edge(pxf, pyf, x0, y0, x1, y1);
    const float b0 = e0 / area, b1 = e1 / area, b2 = e2 / area;
    if (!(b0 >= 0.0f && b1 >= 0.0f && b2 >= 0.0f)) return false;
    z = (z0 * b0 + z1 * b1) + z2 * b2;
    return true;
}

template <int TS, int BLOCK>
__device__ __forceinline__ void raster_body(const RasterParams& p, int blk) {
    constexpr int NPX = TS * TS;
    constexpr int PPT = NPX / BLOCK;
    constexpr int NW = BLOCK / 64;
    static_assert(PPT >= 1 && NPX % BLOCK == 0, "tile must cover the block");
    __shared__ unsigned long long s_key[NPX];
    __shared__ float s_d0[NPX];
    __shared__ float s_v[10][BLOCK];   // x0 y0 z0 x1 y1 z1 x2 y2 z2 area
    __shared__ uint32_t s_rect[BLOCK]; // rx0 | ry0 << 8 | rw << 16 (tile-local)
    __shared__ uint32_t s_face[BLOCK];
    __shared__ uint32_t s_off[BLOCK];
    __shared__ uint32_t s_wsum[NW];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // Consecutive tiles go to different XCDs (workgroups are dealt round robin), which
    // spreads a screen region's heavy tiles over all eight.
    const int tile = blk;
    const int tx = tile % p.tiles_x, ty = tile / p.tiles_x;
    const int ox = tx * TS, oy = ty * TS;

#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int l = tid + j * BLOCK;
        const uint32_t x = ox + (l % TS), y = oy + (l / TS);
        float d0 = 1.0f;
        if (p.depth_in && x < p.W && y < p.H) d0 = p.depth_in[(size_t)y * p.W + x];
        s_key[l] = kEmptyKey;
        s_d0[l] = d0;
    }
    const uint32_t count = p.tile_count[tile];
    const uint32_t n = count > p.cap ? 0u : count;   // overflowed tiles go to k_overflow
    uint32_t frags = 0;
    __syncthreads();

    for (uint32_t base = 0; base < n; base += BLOCK) {
        const uint32_t i = base + tid;
        uint32_t cnt = 0;
        if (i < n) {
            const uint32_t f = p.bins[(size_t)tile * p.cap + i];
            const FaceRecord* r = p.rec + f;
            const float4 a = r->r0, b = r->r1, c = r->r2;
            const uint32_t lo = __float_as_uint(b.w), hi = __float_as_uint(c.w);
            const int rx0 = max((int)(lo & 0xFFFFu) - ox, 0), rx1 = min((int)(hi & 0xFFFFu) - ox, TS - 1);
            const int ry0 = max((int)(lo >> 16) - oy, 0), ry1 = min((int)(hi >> 16) - oy, TS - 1);
            uint32_t rw = 0;
            if (rx0 <= rx1 && ry0 <= ry1) {
                rw = rx1 - rx0 + 1;
                cnt = rw * (ry1 - ry0 + 1);
            }
            s_v[0][tid] = a.x; s_v[1][tid] = a.y; s_v[2][tid] = a.z;
            s_v[3][tid] = b.x; s_v[4][tid] = b.y; s_v[5][tid] = b.z;
            s_v[6][tid] = c.x; s_v[7][tid] = c.y; s_v[8][tid] = c.z;
            s_v[9][tid] = a.w;
            s_rect[tid] = (uint32_t)rx0 | ((uint32_t)ry0 << 8) | (rw << 16);
            s_face[tid] = f;
        }
        // block-wide exclusive prefix sum of candidate counts
        uint32_t v = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(v, d, 64);
            if (lane >= d) v += t;
        }
        if (lane == 63) s_wsum[wave] = v;
        __syncthreads();
        uint32_t wbase = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t s = s_wsum[w];
            if (w < wave) wbase += s;
            total += s;
        }
        s_off[tid] = wbase + v - cnt;
        __syncthreads();

        const int nc = (int)min(n - base, (uint32_t)BLOCK);
        for (uint32_t k = tid; k < total; k += BLOCK) {
            int lo = 0, hi = nc - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= k) lo = mid; else hi = mid - 1;
            }
            const int j = lo;
            const uint32_t rect = s_rect[j];
            const uint32_t rw = rect >> 16;
            const uint32_t l = k - s_off[j];
            // l / rw exactly: |error| of the product < 1/(2*rw) for l < 1024
            const uint32_t dy = (uint32_t)(((float)l + 0.5f) * __builtin_amdgcn_rcpf((float)rw));
            const uint32_t dx = l - dy * rw;
            const int lx = (int)(rect & 0xFFu) + (int)dx, ly = (int)((rect >> 8) & 0xFFu) + (int)dy;
            float z;
            if (fragment_key((float)(ox + lx), (float)(oy + ly), s_v[0][j], s_v[1][j], s_v[2][j], s_v[3][j],
                             s_v[4][j], s_v[5][j], s_v[6][j], s_v[7][j], s_v[8][j], s_v[9][j], z)) {
                ++frags;
                const int pl = ly * TS + lx;
                if (z < s_d0[pl]) atomicMin(&s_key[pl], vis_key(z, s_face[j]));
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int l = tid + j * BLOCK;
        const uint32_t x = ox + (l % TS), y = oy + (l / TS);
        if (x < p.W && y < p.H) p.vis[(size_t)y * p.W + x] = s_key[l];
    }
    store_block_sum(frags, p.part_frag + blk, tid, BLOCK);
}

// ---------------------------------------------------------------------------------
__device__ __forceinline__ void overflow_body(const RasterParams& p, int blk, int nblk) {
    const int32_t ts_log2 = p.ts_log2;
    if (p.counters[kCntOverflowTiles] == 0ull) {
        if (threadIdx.x == 0) p.part_frag[blk] = 0u;
        return;
    }
    const uint32_t ts = 1u << ts_log2;
    uint32_t frags = 0;
    for (int f = blk * 256 + threadIdx.x; f < p.nf; f += nblk * 256) {
        const FaceRecord* r = p.rec + f;
        const float4 b = r->r1, c = r->r2;
        const uint32_t lo = __float_as_uint(b.w), hi = __float_as_uint(c.w);
        const uint32_t bx0 = lo & 0xFFFFu, by0 = lo >> 16, bx1 = hi & 0xFFFFu, by1 = hi >> 16;
        if (bx0 > bx1 || by0 > by1) continue;
        const float4 a = r->r0;
        for (uint32_t ty = by0 >> ts_log2; ty <= (by1 >> ts_log2); ++ty) {
            for (uint32_t tx = bx0 >> ts_log2; tx <= (bx1 >> ts_log2); ++tx) {
                const uint32_t tile = ty * p.tiles_x + tx;
                if (p.tile_count[tile] <= p.cap) continue;
                const uint32_t x0 = max(bx0, tx * ts), x1 = min(bx1, tx * ts + ts - 1);
                const uint32_t y0 = max(by0, ty * ts), y1 = min(by1, ty * ts + ts - 1);
                for (uint32_t y = y0; y <= y1; ++y) {
                    for (uint32_t x = x0; x <= x1; ++x) {
                        float z;
                        if (!fragment_key((float)x, (float)y, a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, a.w, z))
                            continue;
                        ++frags;
                        const size_t pi = (size_t)y * p.W + x;
                        const float d0 = p.depth_in ? p.depth_in[pi] : 1.0f;
                        if (z < d0) atomicMin(&p.vis[pi], vis_key(z, (uint32_t)f));
                    }
                }
            }
        }
    }
    store_block_sum(frags, p.part_frag + blk, threadIdx.x, 256);
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
    const float w = 0.0f;
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

// 64 x 4 pixel blocks: each wave shades 64 consecutive pixels of one row.
// MODE: kShadePlain (the reference's light loop), kShadeMlp (the learned surrogate on
// matrix cores), kShadeCollect (light loop, and each shaded fragment's MLP inputs and
// colour/255 appended to the sample buffer for training).
template <int MODE>
__device__ __forceinline__ void shade_body(const ShadeParams& p, int blk) {
    // normalisation tables of the first kLdsLuts materials, staged in LDS
    constexpr int kLdsLuts = 4;
    __shared__ float s_lut[kLdsLuts * 256];
    {
        const int t = threadIdx.y * 64 + threadIdx.x;
        const int nl = p.nmat < kLdsLuts ? p.nmat : kLdsLuts;
#pragma unroll
        for (int j = 0; j < kLdsLuts; ++j)
            if (j < nl) s_lut[j * 256 + t] = p.luts[j * 256 + t];
        __syncthreads();
    }
    const size_t np = (size_t)p.W * p.H;
    const uint32_t nbx = shade_blocks_x(p.W);
    const uint32_t px = (blk % nbx) * 64 + threadIdx.x, py = (blk / nbx) * 4 + threadIdx.y;
    const bool valid = px < p.W && py < p.H;
    const size_t i = (size_t)py * p.W + px;
    const unsigned long long key = valid ? p.vis[i] : kEmptyKey;
    const bool hit = key != kEmptyKey;
    float feat[kMlpIn];
    uint8_t rgb[3] = {0, 0, 0};
    if (valid && !hit) {
        if (p.cleared) {
            p.frame[i] = 0; p.frame[i + np] = 0; p.frame[i + 2 * np] = 0;
            p.depth[i] = 1.0f;
        }
        if (p.ids) p.ids[i] = -1;
    }
    if (hit) {
        const uint32_t f = (uint32_t)key;
        const FaceRecord& r = p.rec[f];
        const float4 a = r.r0, b = r.r1, c = r.r2, w = r.r3;
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
        int n0, n1, n2;
        if (p.flat) {
            n0 = n1 = n2 = (int)f;
        } else {
            n0 = p.fn[3 * f]; n1 = p.fn[3 * f + 1]; n2 = p.fn[3 * f + 2];
        }
        const float4 N0 = p.cnormal[n0], N1 = p.cnormal[n1], N2 = p.cnormal[n2];
        float nx = d * (i0 * N0.x + i1 * N1.x + i2 * N2.x);
        float ny = d * (i0 * N0.y + i1 * N1.y + i2 * N2.y);
        float nz = d * (i0 * N0.z + i1 * N1.z + i2 * N2.z);
        const float rn = 1.f / sqrtf((nx * nx + ny * ny) + nz * nz);
        nx = nx * rn; ny = ny * rn; nz = nz * rn;
        if (p.wind_cw) { nx = -nx; ny = -ny; nz = -nz; }
        const int mi = p.fmat[f];
        const DevMaterial& m = p.mats[mi];
        float alb[3];
        float u = 0.f, v = 0.f;
        if (m.tex_w > 0 || m.nmap_w > 0) {
            float2 t[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int ui = p.fuv[3 * f + k];
                t[k] = ui >= 0 ? p.uvs[ui] : make_float2(0.f, 0.f);
            }
            u = d * (i0 * t[0].x + i1 * t[1].x + i2 * t[2].x);
            v = d * (i0 * t[0].y + i1 * t[1].y + i2 * t[2].y);
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
            bend_normal(p, f, ts, nx, ny, nz);
        }
        p.depth[i] = z;
        if (p.ids) p.ids[i] = (int32_t)f;
        if (MODE != kShadeMlp) {
            // shading.cpp:20-34
            constexpr float kOneOverPi = 0.318309886183790671537767526745028724f;
            float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
            for (int l = 0; l < p.nl; ++l) {
                const float dd = (nx * p.light_nd[l][0] + ny * p.light_nd[l][1]) + nz * p.light_nd[l][2];
                const float mm = gmax(0.f, dd);
                acc0 = acc0 + p.light_ic[l][0] * alb[0] * mm * kOneOverPi;
                acc1 = acc1 + p.light_ic[l][1] * alb[1] * mm * kOneOverPi;
                acc2 = acc2 + p.light_ic[l][2] * alb[2] * mm * kOneOverPi;
            }
            rgb[0] = (uint8_t)cast_u32(gmin(acc0, 255.f));
            rgb[1] = (uint8_t)cast_u32(gmin(acc1, 255.f));
            rgb[2] = (uint8_t)cast_u32(gmin(acc2, 255.f));
        }
        if (MODE != kShadePlain) {
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
        const int lane = threadIdx.x;
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
        __shared__ uint16_t s_in[4][64 * 16];
        __shared__ float s_out[4][64 * 4];
        const int wave = threadIdx.y, lane = threadIdx.x;
#pragma unroll
        for (int k = 0; k < 16; ++k) s_in[wave][lane * 16 + k] = to_bf16((hit && k < kMlpIn) ? feat[k] : 0.f);
        if (__ballot(hit)) {   // uniform over the wave: waves with no fragment skip the network
            __builtin_amdgcn_wave_barrier();
            mlp_wave(p.mlp, s_in[wave], s_out[wave], lane);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = s_out[wave][lane * 4 + c] * 255.0f;
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
    store_block_sum(hit ? 1u : 0u, p.part_cov + blk, threadIdx.y * 64 + threadIdx.x, 256);
}

// ---------------------------------------------------------------------------------
// Entry kernels.  Single frame: parameters by value, one block per unit of work.
// Batch: per-mesh parameter tables in device memory and a map from block to
// (mesh, block within that mesh's pass), so one launch covers every mesh.

__global__ __launch_bounds__(256) void k_vertex(VertexParams p) { vertex_body(p, blockIdx.x); }
__global__ __launch_bounds__(256) void k_bin(BinParams p) { bin_body(p, blockIdx.x); }
template <int TS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_raster(RasterParams p) { raster_body<TS, BLOCK>(p, blockIdx.x); }
__global__ __launch_bounds__(256) void k_overflow(RasterParams p) { overflow_body(p, blockIdx.x, gridDim.x); }
template <int MODE>
__global__ __launch_bounds__(256) void k_shade(ShadeParams p) { shade_body<MODE>(p, blockIdx.x); }

__global__ __launch_bounds__(256) void k_vertex_batch(const VertexParams* __restrict__ items,
                                                      const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    vertex_body(items[m.x], (int)m.y);
}
__global__ __launch_bounds__(256) void k_bin_batch(const BinParams* __restrict__ items, const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    bin_body(items[m.x], (int)m.y);
}
template <int TS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_raster_batch(const RasterParams* __restrict__ items,
                                                        const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    raster_body<TS, BLOCK>(items[m.x], (int)m.y);
}
__global__ __launch_bounds__(256) void k_overflow_batch(const RasterParams* __restrict__ items,
                                                        const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    const RasterParams& p = items[m.x];
    overflow_body(p, (int)m.y, (int)overflow_blocks(p.nf));
}
template <int MODE>
__global__ __launch_bounds__(256) void k_shade_batch(const ShadeParams* __restrict__ items,
                                                     const uint2* __restrict__ map) {
    const uint2 m = map[blockIdx.x];
    shade_body<MODE>(items[m.x], (int)m.y);
}

// ---------------------------------------------------------------------------------
// Launchers (host side).

hipError_t launch_vertex(const VertexParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_vertex, dim3(vertex_blocks(p.nv, p.nn, p.ntiles)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_bin(const BinParams& p, hipStream_t s) {
    if (p.nf > 0) hipLaunchKernelGGL(k_bin, dim3(bin_blocks(p.nf)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_raster(const RasterParams& p, int tile_size, hipStream_t s) {
    switch (tile_size) {
        case 8: hipLaunchKernelGGL((k_raster<8, 64>), dim3(p.ntiles), dim3(64), 0, s, p); break;
        case 16: hipLaunchKernelGGL((k_raster<16, 256>), dim3(p.ntiles), dim3(256), 0, s, p); break;
        case 32: hipLaunchKernelGGL((k_raster<32, 256>), dim3(p.ntiles), dim3(256), 0, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_overflow(const RasterParams& p, hipStream_t s) {
    if (p.nf > 0) hipLaunchKernelGGL(k_overflow, dim3(overflow_blocks(p.nf)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_shade(const ShadeParams& p, hipStream_t s) {
    const dim3 grid(shade_blocks_x(p.W) * shade_blocks_y(p.H)), block(64, 4);
    switch (p.mode) {
        case kShadePlain: hipLaunchKernelGGL(k_shade<kShadePlain>, grid, block, 0, s, p); break;
        case kShadeMlp: hipLaunchKernelGGL(k_shade<kShadeMlp>, grid, block, 0, s, p); break;
        case kShadeCollect: hipLaunchKernelGGL(k_shade<kShadeCollect>, grid, block, 0, s, p); break;
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
hipError_t launch_raster_batch(const RasterParams* items, const uint2* map, uint32_t blocks, int tile_size,
                               hipStream_t s) {
    if (!blocks) return hipSuccess;
    switch (tile_size) {
        case 8: hipLaunchKernelGGL((k_raster_batch<8, 64>), dim3(blocks), dim3(64), 0, s, items, map); break;
        case 16: hipLaunchKernelGGL((k_raster_batch<16, 256>), dim3(blocks), dim3(256), 0, s, items, map); break;
        case 32: hipLaunchKernelGGL((k_raster_batch<32, 256>), dim3(blocks), dim3(256), 0, s, items, map); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_overflow_batch(const RasterParams* items, const uint2* map, uint32_t blocks, hipStream_t s) {
    if (blocks) hipLaunchKernelGGL(k_overflow_batch, dim3(blocks), dim3(256), 0, s, items, map);
    return hipGetLastError();
}
hipError_t launch_shade_batch(const ShadeParams* items, const uint2* map, uint32_t blocks, int mode,
                              hipStream_t s) {
    if (!blocks) return hipSuccess;
    switch (mode) {
        case kShadePlain: hipLaunchKernelGGL(k_shade_batch<kShadePlain>, dim3(blocks), dim3(64, 4), 0, s, items, map); break;
        case kShadeMlp: hipLaunchKernelGGL(k_shade_batch<kShadeMlp>, dim3(blocks), dim3(64, 4), 0, s, items, map); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rz
