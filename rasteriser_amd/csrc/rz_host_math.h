// rz_host_math.h — host-side per-frame setup of the frame path.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/rz.h"

namespace rz {

// model, view, modelview, camera and normal matrices of one frame (geometry.cpp:22-33,
// drawing.cpp:222-229, geometry.cpp:101), column-major.
void build_matrices(const rz_args& args, rz_matrices& out);

// Light::transform (geometry.cpp:124-127).
void light_direction(const float view[16], const float dir[3], float out[3]);

// normalize(cross(p1 - p0, p2 - p0)): the flat-shading normal of a face.
void face_normal(const float p0[3], const float p1[3], const float p2[3], float out[3]);

// Unit tangent of a face along increasing u: ((e1*dv2 - e2*dv1) / (du1*dv2 - du2*dv1))
// normalised, with e1 = p1-p0, e2 = p2-p0; the edge e1 when that is not finite.
void face_tangent(const float p0[3], const float p1[3], const float p2[3], const float t0[2], const float t1[2],
                  const float t2[2], float out[3]);

// Per-byte values of CImg::normalize(0,1) over `count` bytes of texels.
void normalisation_table(const uint8_t* texels, size_t count, float lut[256]);

}  // namespace rz
