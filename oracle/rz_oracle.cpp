/*
 * rz_oracle.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the reference rasteriser's frame path, used as the checker
 * of librz and as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.  It is never part of the product path.
 *
 * It follows the reference's own structure and operation order, statement by
 * statement, with its own code (the reference is GPL-3.0 and is not copied):
 *
 *   draw_frame    drawing.cpp:205-258   -> oracle_draw_frame
 *   draw_triangle drawing.cpp:151-203   -> draw_triangle
 *   update_pixel  drawing.cpp:96-149    -> update_pixel
 *   barycentric   drawing.cpp:36-49     -> edge, barycentric
 *   bounding_box  drawing.cpp:77-93     -> bounding_box
 *   interpolation drawing.cpp:51-75     -> (inline in update_pixel)
 *   vertex stage  geometry.cpp:22-142   -> transformation_matrix, camera_matrix, ...
 *   shading       shading.cpp:20-34     -> shade
 *   texture       material.cpp:11-26 + CImg.h:5184-5186, 13466-13491, 26786-26794
 *
 * Tangent-space normal maps (rz_material.normal_map) have no reference counterpart;
 * they follow the definition in include/rz.h: tangent from positions and uvs,
 * Gram-Schmidt against the interpolated normal, bitangent cross(N, T), sample texel/255
 * with CImg's bilinear rule, t = 2s-1, n = normalize(T t.x + B t.y + N t.z).
 *
 * Arithmetic contract: IEEE binary32, round to nearest, no FMA contraction
 * (built with -O2 -ffp-contract=off -fno-fast-math), true divisions and sqrt.
 *
 * GLM 0.9.7.6 is not vendored by the reference and is absent here.  Where its
 * operation order fixes result bits the order used is its published one:
 *   mat4*vec4  = (m0*v.x + m1*v.y) + (m2*v.z + m3*v.w)            (pairwise)
 *   mat4*mat4  column i = ((A0*B[i].x + A1*B[i].y) + A2*B[i].z) + A3*B[i].w
 *   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   normalize  = v * (1 / sqrt(dot(v,v)))
 *   min(x,y)   = x < y ? x : y ;  max(x,y) = x > y ? x : y
 *   clamp      = min(max(x, lo), hi) ;  3-arg min/max fold left
 *   rotate/translate/scale/perspective/inverse: matrix_transform.inl / func_matrix.inl
 * Parity of this GLM boundary with the real library is unpinned (no GLM here); the
 * checks in tests/test_oracle_*.py pin everything the mathematics fixes.
 *
 * Undefined behaviour in the reference is given one defined meaning here and in
 * librz: the float -> unsigned casts (drawing.cpp:86-92, shading.cpp:33,
 * CImg.h:13480-13481) follow GCC on x86-64 without AVX-512 (cvttss2si into a
 * 64-bit register, low 32 bits kept: NaN and |f| >= 2^63 give 0).
 */
#include "../include/rz.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };
struct M4 { V4 c[4]; };   // column-major, c[col] holds rows x..w

// ---- GLM scalar semantics (see header) --------------------------------------
inline float gmin(float x, float y) { return x < y ? x : y; }
inline float gmax(float x, float y) { return x > y ? x : y; }
inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

// (unsigned int)f on x86-64 GCC: cvttss2si rax; keep eax.
inline uint32_t cast_u32(float f) {
    if (!(f > -9223372036854775808.0f && f < 9223372036854775808.0f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)f;
}

inline V4 v4add(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 v4mul(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline V4 v4mulv(V4 a, V4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }

inline float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 normalize3(V3 v) {
    float r = 1.0f / std::sqrt(dot3(v, v));
    return {v.x * r, v.y * r, v.z * r};
}

M4 identity() {
    M4 m;
    m.c[0] = {1, 0, 0, 0}; m.c[1] = {0, 1, 0, 0}; m.c[2] = {0, 0, 1, 0}; m.c[3] = {0, 0, 0, 1};
    return m;
}

// GLM mat4 * vec4, pairwise (type_mat4x4.inl operator*(mat, col)).
V4 mul_mv(const M4& m, V4 v) {
    V4 a = v4add(v4mul(m.c[0], v.x), v4mul(m.c[1], v.y));
    V4 b = v4add(v4mul(m.c[2], v.z), v4mul(m.c[3], v.w));
    return v4add(a, b);
}

// GLM mat4 * mat4, left fold per column.
M4 mul_mm(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i) {
        V4 s = v4add(v4mul(a.c[0], b.c[i].x), v4mul(a.c[1], b.c[i].y));
        s = v4add(s, v4mul(a.c[2], b.c[i].z));
        s = v4add(s, v4mul(a.c[3], b.c[i].w));
        r.c[i] = s;
    }
    return r;
}

// glm::translate(m, v): column 3 = m0*v.x + m1*v.y + m2*v.z + m3 (left fold).
M4 translate(const M4& m, V3 v) {
    M4 r = m;
    V4 s = v4add(v4mul(m.c[0], v.x), v4mul(m.c[1], v.y));
    s = v4add(s, v4mul(m.c[2], v.z));
    r.c[3] = v4add(s, m.c[3]);
    return r;
}

// glm::rotate(m, angle, axis): Rodrigues matrix with temp = (1-c)*axis, then
// columns 0..2 = m0*R[j][0] + m1*R[j][1] + m2*R[j][2], column 3 kept.
M4 rotate(const M4& m, float angle, V3 v) {
    const float c = std::cos(angle);
    const float s = std::sin(angle);
    const V3 ax = normalize3(v);
    const V3 t = {(1.0f - c) * ax.x, (1.0f - c) * ax.y, (1.0f - c) * ax.z};
    float R[3][3];
    R[0][0] = c + t.x * ax.x;
    R[0][1] = t.x * ax.y + s * ax.z;
    R[0][2] = t.x * ax.z - s * ax.y;
    R[1][0] = t.y * ax.x - s * ax.z;
    R[1][1] = c + t.y * ax.y;
    R[1][2] = t.y * ax.z + s * ax.x;
    R[2][0] = t.z * ax.x + s * ax.y;
    R[2][1] = t.z * ax.y - s * ax.x;
    R[2][2] = c + t.z * ax.z;
    M4 r;
    for (int j = 0; j < 3; ++j) {
        V4 acc = v4add(v4mul(m.c[0], R[j][0]), v4mul(m.c[1], R[j][1]));
        r.c[j] = v4add(acc, v4mul(m.c[2], R[j][2]));
    }
    r.c[3] = m.c[3];
    return r;
}

// glm::scale(m, v)
M4 scale(const M4& m, V3 v) {
    M4 r;
    r.c[0] = v4mul(m.c[0], v.x);
    r.c[1] = v4mul(m.c[1], v.y);
    r.c[2] = v4mul(m.c[2], v.z);
    r.c[3] = m.c[3];
    return r;
}

// glm::perspective (right handed, depth -1..1).
M4 perspective(float fovy, float aspect, float zn, float zf) {
    const float th = std::tan(fovy / 2.0f);
    M4 r;
    std::memset(&r, 0, sizeof r);
    r.c[0].x = 1.0f / (aspect * th);
    r.c[1].y = 1.0f / th;
    r.c[2].z = -(zf + zn) / (zf - zn);
    r.c[2].w = -1.0f;
    r.c[3].z = -(2.0f * zf * zn) / (zf - zn);
    return r;
}

// glm::radians
float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

float& at(M4& m, int col, int row) { return (&m.c[col].x)[row]; }
float at(const M4& m, int col, int row) { return (&m.c[col].x)[row]; }

// glm::inverse for mat4 (cofactor expansion of func_matrix.inl).
M4 inverse(const M4& m) {
    auto e = [&](int c, int r) { return at(m, c, r); };
    float c00 = e(2, 2) * e(3, 3) - e(3, 2) * e(2, 3);
    float c02 = e(1, 2) * e(3, 3) - e(3, 2) * e(1, 3);
    float c03 = e(1, 2) * e(2, 3) - e(2, 2) * e(1, 3);
    float c04 = e(2, 1) * e(3, 3) - e(3, 1) * e(2, 3);
    float c06 = e(1, 1) * e(3, 3) - e(3, 1) * e(1, 3);
    float c07 = e(1, 1) * e(2, 3) - e(2, 1) * e(1, 3);
    float c08 = e(2, 1) * e(3, 2) - e(3, 1) * e(2, 2);
    float c10 = e(1, 1) * e(3, 2) - e(3, 1) * e(1, 2);
    float c11 = e(1, 1) * e(2, 2) - e(2, 1) * e(1, 2);
    float c12 = e(2, 0) * e(3, 3) - e(3, 0) * e(2, 3);
    float c14 = e(1, 0) * e(3, 3) - e(3, 0) * e(1, 3);
    float c15 = e(1, 0) * e(2, 3) - e(2, 0) * e(1, 3);
    float c16 = e(2, 0) * e(3, 2) - e(3, 0) * e(2, 2);
    float c18 = e(1, 0) * e(3, 2) - e(3, 0) * e(1, 2);
    float c19 = e(1, 0) * e(2, 2) - e(2, 0) * e(1, 2);
    float c20 = e(2, 0) * e(3, 1) - e(3, 0) * e(2, 1);
    float c22 = e(1, 0) * e(3, 1) - e(3, 0) * e(1, 1);
    float c23 = e(1, 0) * e(2, 1) - e(2, 0) * e(1, 1);
    V4 f0 = {c00, c00, c02, c03}, f1 = {c04, c04, c06, c07}, f2 = {c08, c08, c10, c11};
    V4 f3 = {c12, c12, c14, c15}, f4 = {c16, c16, c18, c19}, f5 = {c20, c20, c22, c23};
    V4 a0 = {e(1, 0), e(0, 0), e(0, 0), e(0, 0)};
    V4 a1 = {e(1, 1), e(0, 1), e(0, 1), e(0, 1)};
    V4 a2 = {e(1, 2), e(0, 2), e(0, 2), e(0, 2)};
    V4 a3 = {e(1, 3), e(0, 3), e(0, 3), e(0, 3)};
    auto sub = [](V4 a, V4 b) { return V4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; };
    V4 i0 = v4add(sub(v4mulv(a1, f0), v4mulv(a2, f1)), v4mulv(a3, f2));
    V4 i1 = v4add(sub(v4mulv(a0, f0), v4mulv(a2, f3)), v4mulv(a3, f4));
    V4 i2 = v4add(sub(v4mulv(a0, f1), v4mulv(a1, f3)), v4mulv(a3, f5));
    V4 i3 = v4add(sub(v4mulv(a0, f2), v4mulv(a1, f4)), v4mulv(a2, f5));
    const V4 sa = {+1, -1, +1, -1}, sb = {-1, +1, -1, +1};
    M4 inv;
    inv.c[0] = v4mulv(i0, sa);
    inv.c[1] = v4mulv(i1, sb);
    inv.c[2] = v4mulv(i2, sa);
    inv.c[3] = v4mulv(i3, sb);
    V4 row0 = {inv.c[0].x, inv.c[1].x, inv.c[2].x, inv.c[3].x};
    V4 d0 = v4mulv(m.c[0], row0);
    float d1 = (d0.x + d0.y) + (d0.z + d0.w);
    float od = 1.0f / d1;
    M4 r;
    for (int i = 0; i < 4; ++i) r.c[i] = v4mul(inv.c[i], od);
    return r;
}

M4 transpose(const M4& m) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) at(r, c, rr) = at(m, rr, c);
    return r;
}

// geometry.cpp:22-25: translate * rotY * rotX * rotZ * scale
M4 transformation_matrix(float factor, V3 disp, V3 tb) {
    M4 I = identity();
    M4 r = mul_mm(translate(I, disp), rotate(I, tb.y, {0.f, 1.f, 0.f}));
    r = mul_mm(r, rotate(I, tb.x, {1.f, 0.f, 0.f}));
    r = mul_mm(r, rotate(I, tb.z, {0.f, 0.f, 1.f}));
    r = mul_mm(r, scale(I, {factor, factor, factor}));
    return r;
}

struct Frame {
    M4 model, view, modelview, camera, normal;
};

Frame build_frame(const rz_args& a) {
    Frame f;
    f.model = transformation_matrix(a.scale,
                                    {a.displacement[0], a.displacement[1], a.displacement[2]},
                                    {a.tait_bryan_angles[0], a.tait_bryan_angles[1], a.tait_bryan_angles[2]});
    f.view = transformation_matrix(1.f, {0.f, 0.f, -3.f}, {0.f, 0.f, 0.f});   // drawing.cpp:224
    f.modelview = mul_mm(f.view, f.model);                                     // drawing.cpp:226
    f.camera = mul_mm(perspective(radians(a.fovy_degrees), a.aspect_ratio, a.z_near, a.z_far),
                      f.modelview);                                            // geometry.cpp:27-33
    f.normal = transpose(inverse(f.modelview));                                // geometry.cpp:101
    return f;
}

void store(const M4& m, float* out) { std::memcpy(out, &m, 64); }

// geometry.cpp:35-42
V3 transform_direction(const M4& m, V3 p) {
    V4 t = mul_mv(m, {p.x, p.y, p.z, 0.f});
    return {t.x, t.y, t.z};
}

// geometry.cpp:44-74 for one vertex: clip, z-divide, ndc->raster.
V4 vertex_to_raster(const M4& camera, V3 p, int W, int H) {
    V4 clip = mul_mv(camera, {p.x, p.y, p.z, 1.f});
    V4 ndc = {clip.x / clip.w, clip.y / clip.w, clip.z / clip.w, 1.f / clip.w};
    return {0.5f * (ndc.x + 1.0f) * (float)W, 0.5f * (-ndc.y + 1.0f) * (float)H, ndc.z, ndc.w};
}

// geometry.cpp:76-83
float signed_area_2d(const V4 v[3]) {
    return -0.5f * (v[0].x * v[1].y - v[1].x * v[0].y + v[1].x * v[2].y - v[2].x * v[1].y +
                    v[2].x * v[0].y - v[0].x * v[2].y);
}

// drawing.cpp:36-39
inline float edge(V2 p, V4 a, V4 b) { return (b.x - a.x) * (p.y - a.y) - (b.y - a.y) * (p.x - a.x); }

// drawing.cpp:41-49
inline V3 barycentric(V2 p, const V4 v[3]) {
    float area = edge({v[2].x, v[2].y}, v[0], v[1]);
    return {edge(p, v[1], v[2]) / area, edge(p, v[2], v[0]) / area, edge(p, v[0], v[1]) / area};
}

// drawing.cpp:77-93
void bounding_box(uint32_t tl[2], uint32_t br[2], const V4 v[3], uint32_t W, uint32_t H) {
    const float bx = (float)(W - 1), by = (float)(H - 1);
    float mnx = gmin(gmin(v[0].x, v[1].x), v[2].x), mny = gmin(gmin(v[0].y, v[1].y), v[2].y);
    float mxx = gmax(gmax(v[0].x, v[1].x), v[2].x), mxy = gmax(gmax(v[0].y, v[1].y), v[2].y);
    tl[0] = cast_u32(gclamp(mnx, 0.f, bx));
    tl[1] = cast_u32(gclamp(mny, 0.f, by));
    br[0] = cast_u32(gclamp(std::ceil(mxx), 0.f, bx));
    br[1] = cast_u32(gclamp(std::ceil(mxy), 0.f, by));
}

// Texture of one material after Material's normalize(0,1) (material.h:20-23,
// CImg.h:26786-26794), stored planar like CImg<float>(W,H,1,C).
struct Texture {
    int W = 0, H = 0, C = 0;
    std::vector<float> data;   // [C][H][W]
};

// A normal map as floats texel/255, planar (no min/max normalisation).
Texture make_normal_map(const rz_material& m) {
    Texture t;
    if (m.normal_map_channels <= 0 || !m.normal_map) return t;
    t.W = m.normal_map_width; t.H = m.normal_map_height; t.C = m.normal_map_channels;
    const size_t n = (size_t)t.W * t.H;
    t.data.resize(n * t.C);
    for (int c = 0; c < t.C; ++c)
        for (size_t i = 0; i < n; ++i) t.data[c * n + i] = (float)m.normal_map[i * t.C + c] / 255.0f;
    return t;
}

Texture make_texture(const rz_material& m) {
    Texture t;
    if (m.tex_channels <= 0 || !m.texels) return t;
    t.W = m.tex_width; t.H = m.tex_height; t.C = m.tex_channels;
    const size_t n = (size_t)t.W * t.H;
    t.data.resize(n * t.C);
    for (int c = 0; c < t.C; ++c)
        for (size_t i = 0; i < n; ++i) t.data[c * n + i] = (float)m.texels[i * t.C + c];
    // CImg::normalize(0.f, 1.f): min/max over every channel; equal -> fill(0).
    float mn = t.data[0], mx = t.data[0];
    for (float v : t.data) { if (v < mn) mn = v; if (v > mx) mx = v; }
    const float a = 0.f, b = 1.f;
    if (mn == mx) {
        for (float& v : t.data) v = a;
    } else if (mn != a || mx != b) {
        for (float& v : t.data) v = (v - mn) / (mx - mn) * (b - a) + a;
    }
    return t;
}

// CImg::_linear_atXY (CImg.h:13475-13491) for channel c, with cimg::cut (CImg.h:5184-5186).
float linear_atXY(const Texture& t, float fx, float fy, int c) {
    const float nfx = fx < 0 ? 0.f : (fx > (float)(t.W - 1) ? (float)(t.W - 1) : fx);
    const float nfy = fy < 0 ? 0.f : (fy > (float)(t.H - 1) ? (float)(t.H - 1) : fy);
    const uint32_t x = cast_u32(nfx), y = cast_u32(nfy);
    const float dx = nfx - (float)x, dy = nfy - (float)y;
    const uint32_t nx = dx > 0 ? x + 1 : x, ny = dy > 0 ? y + 1 : y;
    const float* p = t.data.data() + (size_t)c * t.W * t.H;
    const float Icc = p[x + (size_t)y * t.W], Inc = p[nx + (size_t)y * t.W];
    const float Icn = p[x + (size_t)ny * t.W], Inn = p[nx + (size_t)ny * t.W];
    return Icc + dx * (Inc - Icc + dy * (Icc + Inn - Icn - Inc)) + dy * (Icn - Icc);
}

// material.cpp:11-26
V3 sample(const rz_material& m, const Texture& t, V2 uv) {
    if (t.C > 0) {
        float u = uv.x * (float)t.W;
        float v = (1.f - uv.y) * (float)t.H;
        return {linear_atXY(t, u, v, 0), linear_atXY(t, u, v, 1), linear_atXY(t, u, v, 2)};
    }
    return {m.diffuse_colour[0], m.diffuse_colour[1], m.diffuse_colour[2]};
}

const float kOneOverPi = 0.318309886183790671537767526745028724f;

// shading.cpp:20-34.  Per light: I*colour*albedo*max(0, n.-l)*(1/pi), summed from
// vec3(0) in light order, then min(., 255) and the unsigned cast.
void shade(V3 n, V3 albedo, const rz_light* lights, int nl, uint32_t out[3]) {
    float acc[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < nl; ++i) {
        const rz_light& L = lights[i];
        V3 nd = {-L.trans_dir[0], -L.trans_dir[1], -L.trans_dir[2]};
        float m = gmax(0.f, dot3(n, nd));
        const float alb[3] = {albedo.x, albedo.y, albedo.z};
        for (int c = 0; c < 3; ++c) {
            float v = L.intensity * L.colour[c];
            v = v * alb[c];
            v = v * m;
            v = v * kOneOverPi;
            acc[c] = acc[c] + v;
        }
    }
    for (int c = 0; c < 3; ++c) out[c] = cast_u32(gmin(acc[c], 255.f));
}

struct Counts {
    uint64_t front = 0, pairs16 = 0, fragments = 0, depth_passes = 0;
};

struct Scene {
    const rz_mesh_desc* mesh;
    std::vector<V4> raster;          // per vertex
    std::vector<V3> cnormals;        // camera-space normals (or face normals when flat)
    std::vector<V3> ctangents;       // camera-space face tangents (normal maps)
    std::vector<Texture> textures;
    std::vector<Texture> normal_maps;
    const rz_light* lights;
    int nl;
    uint32_t W, H;
    bool wind_cw, flat;
};

// drawing.cpp:96-149
void update_pixel(const Scene& s, uint32_t px, uint32_t py, const V4 rv[3], const V3 vn[3],
                  const V2 uv[3], int mat, int32_t face, uint8_t* frame, float* depth,
                  int32_t* ids, Counts& cnt) {
    V3 b = barycentric({(float)px, (float)py}, rv);
    if (!(b.x >= 0.f && b.y >= 0.f && b.z >= 0.f)) return;
    ++cnt.fragments;
    float z = rv[0].z * b.x + rv[1].z * b.y + rv[2].z * b.z;
    const size_t p = (size_t)py * s.W + px;
    if (!(z < depth[p])) return;
    ++cnt.depth_passes;
    depth[p] = z;
    if (ids) ids[p] = face;
    V3 inter = {rv[0].w * b.x, rv[1].w * b.y, rv[2].w * b.z};
    float d = 1.f / (inter.x + inter.y + inter.z);
    V3 n = {d * (inter.x * vn[0].x + inter.y * vn[1].x + inter.z * vn[2].x),
            d * (inter.x * vn[0].y + inter.y * vn[1].y + inter.z * vn[2].y),
            d * (inter.x * vn[0].z + inter.y * vn[1].z + inter.z * vn[2].z)};
    n = normalize3(n);
    if (s.wind_cw) n = {-n.x, -n.y, -n.z};
    V2 t = {d * (inter.x * uv[0].x + inter.y * uv[1].x + inter.z * uv[2].x),
            d * (inter.x * uv[0].y + inter.y * uv[1].y + inter.z * uv[2].y)};
    V3 albedo = sample(s.mesh->materials[mat], s.textures[mat], t);
    const Texture& nm = s.normal_maps[mat];
    if (nm.C > 0) {
        // N = n, T = tangent orthogonalised against N, B = N x T
        const V3 T = s.ctangents[face];
        const float dd = dot3(n, T);
        V3 P = normalize3({T.x - n.x * dd, T.y - n.y * dd, T.z - n.z * dd});
        V3 B = {n.y * P.z - P.y * n.z, n.z * P.x - P.z * n.x, n.x * P.y - P.x * n.y};
        const float u = t.x * (float)nm.W, v = (1.f - t.y) * (float)nm.H;
        float ts[3];
        for (int c = 0; c < 3; ++c) ts[c] = 2.0f * linear_atXY(nm, u, v, c) - 1.0f;
        n = normalize3({(P.x * ts[0] + B.x * ts[1]) + n.x * ts[2], (P.y * ts[0] + B.y * ts[1]) + n.y * ts[2],
                        (P.z * ts[0] + B.z * ts[1]) + n.z * ts[2]});
    }
    uint32_t px3[3];
    shade(n, albedo, s.lights, s.nl, px3);
    const size_t plane = (size_t)s.W * s.H;
    frame[p] = (uint8_t)px3[0];
    frame[p + plane] = (uint8_t)px3[1];
    frame[p + 2 * plane] = (uint8_t)px3[2];
}

// drawing.cpp:151-203
void draw_triangle(const Scene& s, int32_t f, uint8_t* frame, float* depth, int32_t* ids, Counts& cnt) {
    const rz_triangle& face = s.mesh->faces[f];
    V4 rv[3];
    V3 vn[3];
    V2 uv[3] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
    for (int k = 0; k < 3; ++k) {
        rv[k] = s.raster[face.vertices[k]];
        if (face.uvs[k] >= 0) uv[k] = {s.mesh->uvs[2 * face.uvs[k]], s.mesh->uvs[2 * face.uvs[k] + 1]};
        vn[k] = s.flat ? s.cnormals[f] : s.cnormals[face.normals[k]];
    }
    bool front = (signed_area_2d(rv) > 0) != s.wind_cw;
    if (!front) return;
    ++cnt.front;
    uint32_t tl[2], br[2];
    bounding_box(tl, br, rv, s.W, s.H);
    if (tl[0] <= br[0] && tl[1] <= br[1])
        cnt.pairs16 += (uint64_t)(br[0] / 16 - tl[0] / 16 + 1) * (br[1] / 16 - tl[1] / 16 + 1);
    for (uint32_t y = tl[1]; y <= br[1]; ++y)
        for (uint32_t x = tl[0]; x <= br[0]; ++x)
            update_pixel(s, x, y, rv, vn, uv, face.material, f, frame, depth, ids, cnt);
}

int validate(const rz_mesh_desc* m) {
    if (!m || (m->num_faces > 0 && !m->faces) || m->num_vertices < 0 || m->num_faces < 0)
        return RZ_ERR_INVALID_ARG;
    for (int32_t f = 0; f < m->num_faces; ++f) {
        const rz_triangle& t = m->faces[f];
        for (int k = 0; k < 3; ++k) {
            if (t.vertices[k] < 0 || t.vertices[k] >= m->num_vertices) return RZ_ERR_INDEX_OOB;
            if (t.normals[k] < 0 || t.normals[k] >= m->num_normals) return RZ_ERR_INDEX_OOB;
            if (t.uvs[k] >= m->num_uvs) return RZ_ERR_INDEX_OOB;
        }
        if (t.material < 0 || t.material >= m->num_materials) return RZ_ERR_INDEX_OOB;
    }
    for (int32_t i = 0; i < m->num_materials; ++i) {
        const rz_material& mt = m->materials[i];
        if (mt.normal_map_channels != 0 &&
            ((mt.normal_map_channels != 3 && mt.normal_map_channels != 4) || !mt.normal_map ||
             mt.normal_map_width <= 0 || mt.normal_map_height <= 0))
            return RZ_ERR_INVALID_ARG;
        if (mt.tex_channels == 0) continue;
        // channels 0..2 are sampled (material.cpp:22): 1- or 2-channel images would
        // read past the CImg buffer in the reference.
        if ((mt.tex_channels != 3 && mt.tex_channels != 4) || !mt.texels || mt.tex_width <= 0 ||
            mt.tex_height <= 0)
            return RZ_ERR_INVALID_ARG;
    }
    return RZ_OK;
}

}  // namespace

extern "C" {

typedef struct oracle_counts {
    uint64_t front_triangles;
    uint64_t bin_pairs16;
    uint64_t fragments;
    uint64_t depth_passes;
} oracle_counts;

int oracle_frame_matrices(const rz_args* a, rz_matrices* out) {
    if (!a || !out) return RZ_ERR_INVALID_ARG;
    Frame f = build_frame(*a);
    store(f.model, out->model);
    store(f.view, out->view);
    store(f.modelview, out->modelview);
    store(f.camera, out->camera);
    store(f.normal, out->normal);
    return RZ_OK;
}

// drawing.cpp:205-258.  frame/depth are read and written in place (no clear);
// ids (may be null) receives the face index of every depth-test pass.
int oracle_draw_frame(const rz_mesh_desc* mesh, rz_light* lights, int32_t nl, const rz_args* a,
                      uint8_t* frame, float* depth, int32_t* ids, oracle_counts* counts) {
    if (!a || !frame || !depth || nl < 0 || (nl > 0 && !lights)) return RZ_ERR_INVALID_ARG;
    if (a->image_width == 0 || a->image_height == 0) return RZ_ERR_INVALID_ARG;
    int st = validate(mesh);
    if (st != RZ_OK) return st;
    if (a->flat && mesh->raster_space) return RZ_ERR_INVALID_ARG;
    Frame fr = build_frame(*a);
    Scene s;
    s.mesh = mesh;
    s.lights = lights;
    s.nl = nl;
    s.W = a->image_width;
    s.H = a->image_height;
    s.wind_cw = a->wind_clockwise != 0;
    s.flat = a->flat != 0;
    // normals (geometry.cpp:97-108): inverse-transpose of the model-view, w = 0
    if (s.flat) {
        s.cnormals.resize(mesh->num_faces);
        for (int32_t f = 0; f < mesh->num_faces; ++f) {
            const rz_triangle& t = mesh->faces[f];
            const float* P = mesh->vertices;
            V3 p0 = {P[3 * t.vertices[0]], P[3 * t.vertices[0] + 1], P[3 * t.vertices[0] + 2]};
            V3 p1 = {P[3 * t.vertices[1]], P[3 * t.vertices[1] + 1], P[3 * t.vertices[1] + 2]};
            V3 p2 = {P[3 * t.vertices[2]], P[3 * t.vertices[2] + 1], P[3 * t.vertices[2] + 2]};
            V3 e1 = {p1.x - p0.x, p1.y - p0.y, p1.z - p0.z};
            V3 e2 = {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z};
            V3 cr = {e1.y * e2.z - e2.y * e1.z, e1.z * e2.x - e2.z * e1.x, e1.x * e2.y - e2.x * e1.y};
            s.cnormals[f] = transform_direction(fr.normal, normalize3(cr));
        }
    } else {
        s.cnormals.resize(mesh->num_normals);
        for (int32_t i = 0; i < mesh->num_normals; ++i)
            s.cnormals[i] = transform_direction(
                fr.normal, {mesh->normals[3 * i], mesh->normals[3 * i + 1], mesh->normals[3 * i + 2]});
    }
    // lights (geometry.cpp:124-133): trans_dir = normalize(view * (dir, 0))
    for (int32_t i = 0; i < nl; ++i) {
        V3 d = transform_direction(fr.view, {lights[i].direction[0], lights[i].direction[1], lights[i].direction[2]});
        V3 n = normalize3(d);
        lights[i].trans_dir[0] = n.x;
        lights[i].trans_dir[1] = n.y;
        lights[i].trans_dir[2] = n.z;
    }
    // positions (drawing.cpp:241-247)
    s.raster.resize(mesh->num_vertices);
    for (int32_t i = 0; i < mesh->num_vertices; ++i) {
        if (mesh->raster_space) {
            const float* q = mesh->vertices + 4 * (size_t)i;
            s.raster[i] = {q[0], q[1], q[2], q[3]};
        } else {
            const float* q = mesh->vertices + 3 * (size_t)i;
            s.raster[i] = vertex_to_raster(fr.camera, {q[0], q[1], q[2]}, (int)s.W, (int)s.H);
        }
    }
    s.textures.resize(mesh->num_materials);
    s.normal_maps.resize(mesh->num_materials);
    bool any_nm = false;
    for (int32_t i = 0; i < mesh->num_materials; ++i) {
        s.textures[i] = make_texture(mesh->materials[i]);
        s.normal_maps[i] = make_normal_map(mesh->materials[i]);
        any_nm = any_nm || s.normal_maps[i].C > 0;
    }
    if (any_nm) {
        if (mesh->raster_space) return RZ_ERR_INVALID_ARG;
        // face tangent: (e1*dv2 - e2*dv1)/(du1*dv2 - du2*dv1), normalised; e1 if not finite
        s.ctangents.resize(mesh->num_faces);
        for (int32_t f = 0; f < mesh->num_faces; ++f) {
            const rz_triangle& tr = mesh->faces[f];
            const float* P = mesh->vertices;
            V3 p[3];
            V2 q[3];
            for (int k = 0; k < 3; ++k) {
                p[k] = {P[3 * tr.vertices[k]], P[3 * tr.vertices[k] + 1], P[3 * tr.vertices[k] + 2]};
                q[k] = tr.uvs[k] >= 0 ? V2{mesh->uvs[2 * tr.uvs[k]], mesh->uvs[2 * tr.uvs[k] + 1]} : V2{0.f, 0.f};
            }
            const V3 e1 = {p[1].x - p[0].x, p[1].y - p[0].y, p[1].z - p[0].z};
            const V3 e2 = {p[2].x - p[0].x, p[2].y - p[0].y, p[2].z - p[0].z};
            const float du1 = q[1].x - q[0].x, dv1 = q[1].y - q[0].y;
            const float du2 = q[2].x - q[0].x, dv2 = q[2].y - q[0].y;
            const float den = du1 * dv2 - du2 * dv1;
            V3 tg = {(e1.x * dv2 - e2.x * dv1) / den, (e1.y * dv2 - e2.y * dv1) / den,
                     (e1.z * dv2 - e2.z * dv1) / den};
            const float l2 = dot3(tg, tg);
            if (!(l2 > 0.f && l2 < INFINITY)) tg = e1;
            s.ctangents[f] = transform_direction(fr.normal, normalize3(tg));
        }
    }
    Counts c;
    for (int32_t f = 0; f < mesh->num_faces; ++f) draw_triangle(s, f, frame, depth, ids, c);
    if (counts) {
        counts->front_triangles = c.front;
        counts->bin_pairs16 = c.pairs16;
        counts->fragments = c.fragments;
        counts->depth_passes = c.depth_passes;
    }
    return RZ_OK;
}

// ---- single routines, for the known-answer tests ------------------------------

float oracle_edge(float px, float py, const float a[4], const float b[4]) {
    return edge({px, py}, {a[0], a[1], a[2], a[3]}, {b[0], b[1], b[2], b[3]});
}

void oracle_barycentric(float px, float py, const float v[12], float out[3]) {
    V4 rv[3] = {{v[0], v[1], v[2], v[3]}, {v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11]}};
    V3 b = barycentric({px, py}, rv);
    out[0] = b.x; out[1] = b.y; out[2] = b.z;
}

float oracle_signed_area(const float v[12]) {
    V4 rv[3] = {{v[0], v[1], v[2], v[3]}, {v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11]}};
    return signed_area_2d(rv);
}

void oracle_bounding_box(const float v[12], uint32_t W, uint32_t H, uint32_t out[4]) {
    V4 rv[3] = {{v[0], v[1], v[2], v[3]}, {v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11]}};
    bounding_box(out, out + 2, rv, W, H);
}

void oracle_vertex_to_raster(const float camera[16], const float p[3], int W, int H, float out[4]) {
    M4 m;
    std::memcpy(&m, camera, 64);
    V4 r = vertex_to_raster(m, {p[0], p[1], p[2]}, W, H);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void oracle_shade(const float n[3], const float albedo[3], const rz_light* lights, int32_t nl, uint32_t out[3]) {
    shade({n[0], n[1], n[2]}, {albedo[0], albedo[1], albedo[2]}, lights, nl, out);
}

// Samples channels 0..2 of material m at texture coordinate (u, v).
void oracle_sample(const rz_material* m, float u, float v, float out[3]) {
    Texture t = make_texture(*m);
    V3 r = sample(*m, t, {u, v});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// CImg::_linear_atXY of channel c at raster coordinates (fx, fy) of material m's
// normalised texture.
float oracle_linear_atXY(const rz_material* m, float fx, float fy, int c) {
    Texture t = make_texture(*m);
    return linear_atXY(t, fx, fy, c);
}

// The normalised texel values of material m, planar [C][H][W].
void oracle_normalized_texture(const rz_material* m, float* out) {
    Texture t = make_texture(*m);
    std::memcpy(out, t.data.data(), t.data.size() * sizeof(float));
}

uint32_t oracle_cast_u32(float f) { return cast_u32(f); }

}  // extern "C"
