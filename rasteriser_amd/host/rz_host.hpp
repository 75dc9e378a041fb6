// rz_host.hpp — the C++ host layer over librz's C ABI, keeping the reference
// renderer's interfaces: draw_frame (headers/drawing.h:16-18), load_obj / load_lights
// (headers/fileloader.h:17-19), Args and its command-line flags
// (headers/arguments.h:7-21, arguments.cpp:10-51), Material (headers/material.h), and
// the planar CImg frame/depth buffer layouts (renderer.cpp:85-86).  Vector types are
// plain structs in place of GLM's (binary-compatible float[2]/float[3]).
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rz.h"

namespace rz {

struct Vec2 { float x, y; };
struct Vec3 { float x, y, z; };

using Triangle = rz_triangle;   // face.h:6-13 layout
using Light = rz_light;         // light.h:7-14 layout

inline Light make_light(Vec3 direction, float intensity, Vec3 colour) {
    Light l{};
    l.direction[0] = direction.x; l.direction[1] = direction.y; l.direction[2] = direction.z;
    l.intensity = intensity;
    l.colour[0] = colour.x; l.colour[1] = colour.y; l.colour[2] = colour.z;
    return l;
}

// An 8-bit image as stored in its file: rows of width texels, channels interleaved.
struct Image8 {
    int width = 0, height = 0, channels = 0;
    std::vector<uint8_t> data;
};

Image8 load_image(const std::string& path);                        // PNG (8-bit) or binary PPM/PGM
void save_png(const std::string& path, const uint8_t* interleaved, int W, int H, int C);
void save_frame_png(const std::string& path, const uint8_t* planar_rgb, int W, int H);
void save_depth_png(const std::string& path, const float* depth, int W, int H);

// headers/material.h:11-25: a flat diffuse colour, or a texture file that replaces it.
class Material {
public:
    explicit Material(const Vec3& dc) : diffuse_colour_(dc), has_texture_(false) {}
    Material(const Vec3& dc, const std::string& texture_file);    // loads the image (material.h:20-23)
    Material(const Vec3& dc, Image8 texture);
    const Vec3& diffuse_colour() const { return diffuse_colour_; }
    bool has_texture() const { return has_texture_; }
    const Image8& texture() const { return texture_; }
    rz_material as_c() const;

private:
    Vec3 diffuse_colour_;
    bool has_texture_;
    std::string texture_file_;
    Image8 texture_;
};

// headers/arguments.h:7-21 with the same flags and defaults (arguments.cpp:15-33), plus
// the projection the reference hard-codes (geometry.cpp:30) and librz's tuning.
struct Args {
    unsigned int image_width = 540;
    unsigned int image_height = 304;
    float aspect_ratio = 540.f / 304.f;
    bool spin = false;
    bool flat = false;
    bool wind_clockwise = false;
    float scale = 1.f;
    Vec3 displacement{0.f, 0.f, 0.f};
    Vec3 tait_bryan_angles{0.f, 0.f, 0.f};
    std::string obj_file;
    std::string lights_file;
    std::string materials_directory;
    // beyond the reference
    float fovy_degrees = 45.f, z_near = 0.1f, z_far = 6.f;
    int device = 0;
    int tile_size = 16;
    int bin_capacity = 0;   // 0: librz's automatic capacity
    int frames = 120;            // spin mode without a display: frames to render
    std::string out_dir = ".";
    std::string dump_mesh;       // write the loaded mesh as raw arrays (loader checks)

    Args() = default;
    Args(int argc, char** argv);   // prints "Error: ..." and exits 1 on a bad command line
    rz_args as_c() const;
};

// Planar CImg<unsigned char>(W,H,1,3) and CImg<float>(W,H,1,1) buffers.
struct FrameBuffer {
    int width = 0, height = 0;
    std::vector<uint8_t> data;   // [3][H][W]
    FrameBuffer(int w, int h, uint8_t fill = 0) : width(w), height(h), data((size_t)w * h * 3, fill) {}
    void fill(uint8_t v) { std::fill(data.begin(), data.end(), v); }
};
struct DepthBuffer {
    int width = 0, height = 0;
    std::vector<float> data;     // [H][W]
    DepthBuffer(int w, int h, float fill = 1.f) : width(w), height(h), data((size_t)w * h, fill) {}
    void fill(float v) { std::fill(data.begin(), data.end(), v); }
};

// fileloader.cpp:79-133 semantics (tinyobjloader 1.0.x for OBJ/MTL).
void load_obj(const Args& arguments, std::vector<Vec3>& vertices, std::vector<Triangle>& triangles,
              std::vector<Vec3>& vertnormals, std::vector<Vec2>& vertuvs, std::vector<Material>& materials);
void load_lights(const std::string& file, std::vector<Light>& lights);

// The reference's only frame-path entry point (drawing.cpp:205-258): transforms and
// draws every face into the buffers, which it reads and updates in place, and writes
// the lights' trans_dir.  Runs on the calling thread's librz context (device
// arguments.device); throws std::runtime_error on an error status.
void draw_frame(const std::vector<Vec3>& model_vertices, const std::vector<Triangle>& faces,
                const std::vector<Vec3>& model_vertnormals, const std::vector<Vec2>& vertuvs,
                std::vector<Light>& lights, const std::vector<Material>& materials, const Args& arguments,
                FrameBuffer* frame_buffer, DepthBuffer* depth_buffer);

// Planar frame and depth buffers in device memory (rz_device_alloc), for frames that
// stay on the device (the spin loop); download() copies them into host buffers.
struct DeviceFrame {
    rz_ctx* ctx = nullptr;
    int width = 0, height = 0;
    uint8_t* frame = nullptr;
    float* depth = nullptr;
    DeviceFrame(rz_ctx* c, int w, int h);
    ~DeviceFrame();
    DeviceFrame(const DeviceFrame&) = delete;
    DeviceFrame& operator=(const DeviceFrame&) = delete;
    void download(FrameBuffer& f, DepthBuffer& d) const;
};

// A mesh kept resident on the device between frames (the spin loop's case).
class Renderer {
public:
    explicit Renderer(int device = 0);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    void set_tuning(int tile_size, int bin_capacity);
    void upload(const std::vector<Vec3>& vertices, const std::vector<Triangle>& faces,
                const std::vector<Vec3>& vertnormals, const std::vector<Vec2>& vertuvs,
                const std::vector<Material>& materials);
    void draw(std::vector<Light>& lights, const Args& arguments, FrameBuffer* frame, DepthBuffer* depth,
              bool cleared = false, rz_frame_stats* stats = nullptr);
    rz_ctx* context() { return ctx_; }
    rz_mesh* mesh() { return mesh_; }

private:
    rz_ctx* ctx_ = nullptr;
    rz_mesh* mesh_ = nullptr;
};

void check(rz_status status, rz_ctx* ctx, const char* what);

}  // namespace rz
