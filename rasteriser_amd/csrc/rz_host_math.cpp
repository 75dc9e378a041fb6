// rz_host_math.cpp — per-frame matrices and light directions, built on the host once
// per frame (the reference builds them once per frame too, drawing.cpp:222-238).
//
// GLM 0.9.7.6 is the reference's vector library (conanfile.txt:2) and is not in the
// image; its operation order is restated here because it fixes the bits of every
// raster vertex: mat4*vec4 sums the column products pairwise, mat4*mat4 folds them
// left to right, normalize is v * (1/sqrt(dot(v,v))) with dot = (x*x + y*y) + z*z.
#include "rz_host_math.h"

#include <cmath>
#include <cstring>

namespace rz {
namespace {

struct Vec4 { float v[4]; };
struct Mat4 { Vec4 col[4]; };

Vec4 axpy4(const Vec4& a, float s) {   // a * s
    return {{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}};
}
Vec4 add4(const Vec4& a, const Vec4& b) {
    return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}};
}
Vec4 sub4(const Vec4& a, const Vec4& b) {
    return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]}};
}
Vec4 had4(const Vec4& a, const Vec4& b) {
    return {{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}};
}

Mat4 eye() {
    Mat4 m{};
    for (int i = 0; i < 4; ++i) m.col[i].v[i] = 1.0f;
    return m;
}

Mat4 operator*(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; ++i) {
        Vec4 acc = add4(axpy4(a.col[0], b.col[i].v[0]), axpy4(a.col[1], b.col[i].v[1]));
        acc = add4(acc, axpy4(a.col[2], b.col[i].v[2]));
        r.col[i] = add4(acc, axpy4(a.col[3], b.col[i].v[3]));
    }
    return r;
}

Vec4 operator*(const Mat4& m, const Vec4& x) {
    return add4(add4(axpy4(m.col[0], x.v[0]), axpy4(m.col[1], x.v[1])),
                add4(axpy4(m.col[2], x.v[2]), axpy4(m.col[3], x.v[3])));
}

void unit3(const float in[3], float out[3]) {
    const float len2 = (in[0] * in[0] + in[1] * in[1]) + in[2] * in[2];
    const float inv = 1.0f / std::sqrt(len2);
    for (int i = 0; i < 3; ++i) out[i] = in[i] * inv;
}

Mat4 translation(const float d[3]) {
    const Mat4 I = eye();
    Mat4 r = I;
    Vec4 c = add4(axpy4(I.col[0], d[0]), axpy4(I.col[1], d[1]));
    c = add4(c, axpy4(I.col[2], d[2]));
    r.col[3] = add4(c, I.col[3]);
    return r;
}

// Rotation by `angle` about `axis`, applied to the identity (glm::rotate(mat4(1), a, axis)).
Mat4 rotation(float angle, const float axis_in[3]) {
    const float c = std::cos(angle), s = std::sin(angle);
    float k[3];
    unit3(axis_in, k);
    const float om = 1.0f - c;
    const float t[3] = {om * k[0], om * k[1], om * k[2]};
    // R[j] is the j-th column of the Rodrigues rotation
    const float R[3][3] = {
        {c + t[0] * k[0], t[0] * k[1] + s * k[2], t[0] * k[2] - s * k[1]},
        {t[1] * k[0] - s * k[2], c + t[1] * k[1], t[1] * k[2] + s * k[0]},
        {t[2] * k[0] + s * k[1], t[2] * k[1] - s * k[0], c + t[2] * k[2]},
    };
    const Mat4 I = eye();
    Mat4 r;
    for (int j = 0; j < 3; ++j)
        r.col[j] = add4(add4(axpy4(I.col[0], R[j][0]), axpy4(I.col[1], R[j][1])), axpy4(I.col[2], R[j][2]));
    r.col[3] = I.col[3];
    return r;
}

Mat4 scaling(float f) {
    const Mat4 I = eye();
    Mat4 r;
    r.col[0] = axpy4(I.col[0], f);
    r.col[1] = axpy4(I.col[1], f);
    r.col[2] = axpy4(I.col[2], f);
    r.col[3] = I.col[3];
    return r;
}

// geometry.cpp:22-25
Mat4 model_matrix(float factor, const float disp[3], const float angles[3]) {
    static const float kY[3] = {0.f, 1.f, 0.f}, kX[3] = {1.f, 0.f, 0.f}, kZ[3] = {0.f, 0.f, 1.f};
    return translation(disp) * rotation(angles[1], kY) * rotation(angles[0], kX) * rotation(angles[2], kZ) *
           scaling(factor);
}

Mat4 projection(float fovy_deg, float aspect, float zn, float zf) {
    const float fovy = fovy_deg * 0.01745329251994329576923690768489f;   // glm::radians
    const float th = std::tan(fovy / 2.0f);
    Mat4 p{};
    p.col[0].v[0] = 1.0f / (aspect * th);
    p.col[1].v[1] = 1.0f / th;
    p.col[2].v[2] = -(zf + zn) / (zf - zn);
    p.col[2].v[3] = -1.0f;
    p.col[3].v[2] = -(2.0f * zf * zn) / (zf - zn);
    return p;
}

// glm::inverse (mat4): 2x2 sub-determinants ("Coef"), cofactor vectors, signs, and
// one division by the determinant taken from row 0.
Mat4 inverted(const Mat4& m) {
    auto e = [&](int c, int r) { return m.col[c].v[r]; };
    auto coef = [&](int a, int b, int c, int d) {   // e(a,c)*e(b,d) - e(b,c)*e(a,d)
        return e(a, c) * e(b, d) - e(b, c) * e(a, d);
    };
    // sub-determinants over column pairs (2,3), (1,3), (1,2) of row pairs
    const float s00 = coef(2, 3, 2, 3), s02 = coef(1, 3, 2, 3), s03 = coef(1, 2, 2, 3);
    const float s04 = coef(2, 3, 1, 3), s06 = coef(1, 3, 1, 3), s07 = coef(1, 2, 1, 3);
    const float s08 = coef(2, 3, 1, 2), s10 = coef(1, 3, 1, 2), s11 = coef(1, 2, 1, 2);
    const float s12 = coef(2, 3, 0, 3), s14 = coef(1, 3, 0, 3), s15 = coef(1, 2, 0, 3);
    const float s16 = coef(2, 3, 0, 2), s18 = coef(1, 3, 0, 2), s19 = coef(1, 2, 0, 2);
    const float s20 = coef(2, 3, 0, 1), s22 = coef(1, 3, 0, 1), s23 = coef(1, 2, 0, 1);
    const Vec4 F0{{s00, s00, s02, s03}}, F1{{s04, s04, s06, s07}}, F2{{s08, s08, s10, s11}};
    const Vec4 F3{{s12, s12, s14, s15}}, F4{{s16, s16, s18, s19}}, F5{{s20, s20, s22, s23}};
    const Vec4 V0{{e(1, 0), e(0, 0), e(0, 0), e(0, 0)}};
    const Vec4 V1{{e(1, 1), e(0, 1), e(0, 1), e(0, 1)}};
    const Vec4 V2{{e(1, 2), e(0, 2), e(0, 2), e(0, 2)}};
    const Vec4 V3{{e(1, 3), e(0, 3), e(0, 3), e(0, 3)}};
    const Vec4 sgnA{{+1.f, -1.f, +1.f, -1.f}}, sgnB{{-1.f, +1.f, -1.f, +1.f}};
    Mat4 adj;
    adj.col[0] = had4(add4(sub4(had4(V1, F0), had4(V2, F1)), had4(V3, F2)), sgnA);
    adj.col[1] = had4(add4(sub4(had4(V0, F0), had4(V2, F3)), had4(V3, F4)), sgnB);
    adj.col[2] = had4(add4(sub4(had4(V0, F1), had4(V1, F3)), had4(V3, F5)), sgnA);
    adj.col[3] = had4(add4(sub4(had4(V0, F2), had4(V1, F4)), had4(V2, F5)), sgnB);
    const Vec4 row0{{adj.col[0].v[0], adj.col[1].v[0], adj.col[2].v[0], adj.col[3].v[0]}};
    const Vec4 d = had4(m.col[0], row0);
    const float det = (d.v[0] + d.v[1]) + (d.v[2] + d.v[3]);
    const float inv_det = 1.0f / det;
    Mat4 r;
    for (int i = 0; i < 4; ++i) r.col[i] = axpy4(adj.col[i], inv_det);
    return r;
}

Mat4 transposed(const Mat4& m) {
    Mat4 r;
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) r.col[c].v[rr] = m.col[rr].v[c];
    return r;
}

void put(const Mat4& m, float out[16]) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[4 * c + r] = m.col[c].v[r];
}

Mat4 get(const float in[16]) {
    Mat4 m;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m.col[c].v[r] = in[4 * c + r];
    return m;
}

}  // namespace

void build_matrices(const rz_args& a, rz_matrices& out) {
    static const float kViewDisp[3] = {0.f, 0.f, -3.f}, kZero[3] = {0.f, 0.f, 0.f};
    const Mat4 model = model_matrix(a.scale, a.displacement, a.tait_bryan_angles);
    const Mat4 view = model_matrix(1.f, kViewDisp, kZero);                       // drawing.cpp:224
    const Mat4 mv = view * model;                                                // drawing.cpp:226
    const Mat4 cam = projection(a.fovy_degrees, a.aspect_ratio, a.z_near, a.z_far) * mv;   // geometry.cpp:27-33
    put(model, out.model);
    put(view, out.view);
    put(mv, out.modelview);
    put(cam, out.camera);
    put(transposed(inverted(mv)), out.normal);                                   // geometry.cpp:101
}

// geometry.cpp:124-127: trans_dir = normalize(view * (direction, 0))
void light_direction(const float view[16], const float dir[3], float out[3]) {
    const Vec4 t = get(view) * Vec4{{dir[0], dir[1], dir[2], 0.0f}};
    const float xyz[3] = {t.v[0], t.v[1], t.v[2]};
    unit3(xyz, out);
}

// Flat shading normal of a model-space face: normalize(cross(p1 - p0, p2 - p0)).
void face_normal(const float p0[3], const float p1[3], const float p2[3], float out[3]) {
    const float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const float c[3] = {a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]};
    unit3(c, out);
}

void face_tangent(const float p0[3], const float p1[3], const float p2[3], const float t0[2], const float t1[2],
                  const float t2[2], float out[3]) {
    const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const float du1 = t1[0] - t0[0], dv1 = t1[1] - t0[1], du2 = t2[0] - t0[0], dv2 = t2[1] - t0[1];
    const float den = du1 * dv2 - du2 * dv1;
    float t[3];
    for (int c = 0; c < 3; ++c) t[c] = (e1[c] * dv2 - e2[c] * dv1) / den;
    const float len2 = (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
    if (!(len2 > 0.f && len2 < INFINITY)) {
        t[0] = e1[0]; t[1] = e1[1]; t[2] = e1[2];
    }
    unit3(t, out);
}

// CImg::normalize(0, 1) of an 8-bit image (CImg.h:26786-26794) as a per-byte table:
// min and max over every channel; equal min and max fill with 0; an image already
// spanning exactly [0, 1] is left as is.
void normalisation_table(const uint8_t* texels, size_t count, float lut[256]) {
    uint8_t lo = 255, hi = 0;
    for (size_t i = 0; i < count; ++i) {
        if (texels[i] < lo) lo = texels[i];
        if (texels[i] > hi) hi = texels[i];
    }
    const float fm = (float)lo, fM = (float)hi, a = 0.f, b = 1.f;
    for (int v = 0; v < 256; ++v) {
        const float x = (float)v;
        if (lo == hi)
            lut[v] = a;
        else if (fm != a || fM != b)
            lut[v] = (x - fm) / (fM - fm) * (b - a) + a;
        else
            lut[v] = x;
    }
}

}  // namespace rz
