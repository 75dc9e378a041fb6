// cimg_ref.cpp — TEST INFRASTRUCTURE ONLY.
//
// A thin C entry layer over the reference's vendored CImg 2.0.0 header
// (vendor/cimg/CImg.h, included from the reference tree at build time, never copied).
// It exposes the two CImg routines on the frame path so tests can pin the oracle's
// restatement of them against the real code:
//   CImg<float>::normalize(0,1)   CImg.h:26786-26794 (Material ctor, material.h:22)
//   CImg<float>::_linear_atXY     CImg.h:13475-13491 (Material::sample, material.cpp:22)
// and the single-frame writer of depth.png, depth_buffer.normalize(0,255).save(...)
// (renderer.cpp:93: CImg::normalize, then CImg's libpng writer CImg.h:52154-52290,
// built with cimg_use_png).
// Built by oracle/Makefile into oracle/_ref/ only when the reference tree is present.
#define cimg_display 0
#define cimg_verbosity 0
#define cimg_use_png
#include "CImg.h"

#include <cstring>

using cimg_library::CImg;

extern "C" {

// in/out: planar [C][H][W] floats, as CImg stores them.
void cimg_ref_normalize(const float* in, int W, int H, int C, float* out) {
    CImg<float> img(W, H, 1, C);
    std::memcpy(img.data(), in, sizeof(float) * (size_t)W * H * C);
    img.normalize(0.f, 1.f);
    std::memcpy(out, img.data(), sizeof(float) * (size_t)W * H * C);
}

// Samples n points (fx[i], fy[i]) of channel c of a planar [C][H][W] float image.
void cimg_ref_linear_atXY(const float* data, int W, int H, int C, const float* fx, const float* fy, int c,
                          int n, float* out) {
    CImg<float> img(W, H, 1, C);
    std::memcpy(img.data(), data, sizeof(float) * (size_t)W * H * C);
    for (int i = 0; i < n; ++i) out[i] = img._linear_atXY(fx[i], fy[i], 0, c);
}

// renderer.cpp:93 on a [H][W] float depth buffer; returns 0, or 1 if CImg threw.
int cimg_ref_save_depth_png(const char* path, const float* depth, int W, int H) {
    try {
        CImg<float> db(W, H, 1, 1);
        std::memcpy(db.data(), depth, sizeof(float) * (size_t)W * H);
        db.normalize(0, 255).save(path);
    } catch (...) {
        return 1;
    }
    return 0;
}

}  // extern "C"
