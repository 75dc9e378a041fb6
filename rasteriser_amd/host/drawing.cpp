// drawing.cpp — host-side draw_frame over librz, Args parsing and Material.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "rz_host.hpp"

namespace rz {

void check(rz_status status, rz_ctx* ctx, const char* what) {
    if (status == RZ_OK) return;
    std::string msg = std::string(what) + ": " + rz_status_string(status);
    if (ctx && rz_last_error(ctx)[0]) msg += std::string(" (") + rz_last_error(ctx) + ")";
    throw std::runtime_error(msg);
}

Material::Material(const Vec3& dc, const std::string& texture_file)
    : diffuse_colour_(dc), has_texture_(true), texture_file_(texture_file), texture_(load_image(texture_file)) {
    if (texture_.channels != 3 && texture_.channels != 4)
        throw std::runtime_error(texture_file + ": a texture needs 3 or 4 channels");
    std::printf("Loaded texture %s.\n", texture_file.c_str());   // material.h:21
}

Material::Material(const Vec3& dc, Image8 texture)
    : diffuse_colour_(dc), has_texture_(true), texture_(std::move(texture)) {}

rz_material Material::as_c() const {
    rz_material m{};
    m.diffuse_colour[0] = diffuse_colour_.x;
    m.diffuse_colour[1] = diffuse_colour_.y;
    m.diffuse_colour[2] = diffuse_colour_.z;
    if (has_texture_) {
        m.tex_width = texture_.width;
        m.tex_height = texture_.height;
        m.tex_channels = texture_.channels;
        m.texels = texture_.data.data();
    }
    return m;
}

rz_args Args::as_c() const {
    rz_args a{};
    a.image_width = image_width;
    a.image_height = image_height;
    a.aspect_ratio = aspect_ratio;
    a.flat = flat;
    a.wind_clockwise = wind_clockwise;
    a.scale = scale;
    a.displacement[0] = displacement.x; a.displacement[1] = displacement.y; a.displacement[2] = displacement.z;
    a.tait_bryan_angles[0] = tait_bryan_angles.x;
    a.tait_bryan_angles[1] = tait_bryan_angles.y;
    a.tait_bryan_angles[2] = tait_bryan_angles.z;
    a.fovy_degrees = fovy_degrees;
    a.z_near = z_near;
    a.z_far = z_far;
    return a;
}

namespace {

[[noreturn]] void arg_error(const std::string& msg, const std::string& arg) {
    std::fprintf(stderr, "Error: %s for arg %s\n", msg.c_str(), arg.c_str());   // arguments.cpp:49-50
    std::exit(1);
}

const char* kUsage =
    "renderer -l <lights.csv> [-o <model.obj>] [--mats-dir <path/>] [-x <pixels>] [-y <pixels>]\n"
    "         [-s] [-f] [--wind-clockwise] [--rx <rad>] [--ry <rad>] [--rz <rad>] [--scale <f>]\n"
    "         [--dx <d>] [--dy <d>] [--dz <d>]\n"
    "         [--fovy <deg>] [--near <d>] [--far <d>] [--device <n>] [--tile 8|16|32] [--cap <n>]\n"
    "         [--frames <n>] [--out-dir <dir>] [--dump-mesh <prefix>]\n"
    "(--cap 0, the default: librz sizes the bins itself)\n";

}  // namespace

// arguments.cpp:10-51: the same short and long flags, defaults and required -l.
Args::Args(int argc, char** argv) {
    bool have_lights = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i], val;
        const size_t eq = a.find('=');
        bool has_eq = a.rfind("--", 0) == 0 && eq != std::string::npos;
        if (has_eq) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
        }
        auto value = [&]() -> std::string {
            if (has_eq) return val;
            if (i + 1 >= argc) arg_error("Missing a value for this argument!", a);
            return argv[++i];
        };
        auto num = [&](const std::string& s) -> float {
            char* end = nullptr;
            const float f = std::strtof(s.c_str(), &end);
            if (s.empty() || *end) arg_error("Couldn't read argument value from string '" + s + "'", a);
            return f;
        };
        auto uint = [&](const std::string& s) -> unsigned {
            char* end = nullptr;
            const long v = std::strtol(s.c_str(), &end, 10);
            if (s.empty() || *end || v < 0) arg_error("Couldn't read argument value from string '" + s + "'", a);
            return (unsigned)v;
        };
        if (a == "-h" || a == "--help") {
            std::printf("Render a model by rasterisation.\n%s", kUsage);
            std::exit(0);
        } else if (a == "-o" || a == "--obj") obj_file = value();
        else if (a == "-l" || a == "--lights") { lights_file = value(); have_lights = true; }
        else if (a == "--mats-dir") materials_directory = value();
        else if (a == "-x" || a == "--width") image_width = uint(value());
        else if (a == "-y" || a == "--height") image_height = uint(value());
        else if (a == "-s" || a == "--spin") spin = true;
        else if (a == "-f" || a == "--flat") flat = true;
        else if (a == "--wind-clockwise") wind_clockwise = true;
        else if (a == "--rx") tait_bryan_angles.x = num(value());
        else if (a == "--ry") tait_bryan_angles.y = num(value());
        else if (a == "--rz") tait_bryan_angles.z = num(value());
        else if (a == "--scale") scale = num(value());
        else if (a == "--dx") displacement.x = num(value());
        else if (a == "--dy") displacement.y = num(value());
        else if (a == "--dz") displacement.z = num(value());
        else if (a == "--fovy") fovy_degrees = num(value());
        else if (a == "--near") z_near = num(value());
        else if (a == "--far") z_far = num(value());
        else if (a == "--device") device = (int)uint(value());
        else if (a == "--tile") tile_size = (int)uint(value());
        else if (a == "--cap") bin_capacity = (int)uint(value());
        else if (a == "--frames") frames = (int)uint(value());
        else if (a == "--out-dir") out_dir = value();
        else if (a == "--dump-mesh") dump_mesh = value();
        else arg_error("Couldn't find match for argument", a);
    }
    if (!have_lights) arg_error("Required argument missing: lights", "-l");
    if (image_width == 0 || image_height == 0) arg_error("image size must be positive", "-x/-y");
    aspect_ratio = (float)image_width / (float)image_height;   // arguments.cpp:39
}

// ---------------------------------------------------------------------------------

namespace {

struct ThreadContext {
    rz_ctx* ctx = nullptr;
    int device = -1;
    ~ThreadContext() {
        if (ctx) rz_destroy(ctx);
    }
    rz_ctx* get(int dev) {
        if (ctx && device != dev) {
            rz_destroy(ctx);
            ctx = nullptr;
        }
        if (!ctx) {
            check(rz_create(dev, &ctx), nullptr, "rz_create");
            device = dev;
        }
        return ctx;
    }
};

thread_local ThreadContext t_ctx;

rz_mesh_desc describe(const std::vector<Vec3>& vertices, const std::vector<Triangle>& faces,
                      const std::vector<Vec3>& vertnormals, const std::vector<Vec2>& vertuvs,
                      const std::vector<rz_material>& mats) {
    rz_mesh_desc d{};
    d.vertices = vertices.empty() ? nullptr : &vertices[0].x;
    d.normals = vertnormals.empty() ? nullptr : &vertnormals[0].x;
    d.uvs = vertuvs.empty() ? nullptr : &vertuvs[0].x;
    d.faces = faces.empty() ? nullptr : faces.data();
    d.materials = mats.empty() ? nullptr : mats.data();
    d.num_vertices = (int32_t)vertices.size();
    d.num_normals = (int32_t)vertnormals.size();
    d.num_uvs = (int32_t)vertuvs.size();
    d.num_faces = (int32_t)faces.size();
    d.num_materials = (int32_t)mats.size();
    return d;
}

void check_buffers(const Args& a, FrameBuffer* f, DepthBuffer* d) {
    if (!f || !d || f->width != (int)a.image_width || f->height != (int)a.image_height ||
        d->width != (int)a.image_width || d->height != (int)a.image_height)
        throw std::runtime_error("frame/depth buffers do not match the image size");
}

}  // namespace

void draw_frame(const std::vector<Vec3>& model_vertices, const std::vector<Triangle>& faces,
                const std::vector<Vec3>& model_vertnormals, const std::vector<Vec2>& vertuvs,
                std::vector<Light>& lights, const std::vector<Material>& materials, const Args& arguments,
                FrameBuffer* frame_buffer, DepthBuffer* depth_buffer) {
    check_buffers(arguments, frame_buffer, depth_buffer);
    rz_ctx* ctx = t_ctx.get(arguments.device);
    check(rz_set_tuning(ctx, arguments.tile_size, arguments.bin_capacity), ctx, "rz_set_tuning");
    std::vector<rz_material> mats;
    for (const Material& m : materials) mats.push_back(m.as_c());
    const rz_mesh_desc d = describe(model_vertices, faces, model_vertnormals, vertuvs, mats);
    const rz_args a = arguments.as_c();
    check(rz_draw_frame_desc(ctx, &d, lights.data(), (int32_t)lights.size(), &a, frame_buffer->data.data(),
                             depth_buffer->data.data(), 0u),
          ctx, "draw_frame");
}

Renderer::Renderer(int device) { check(rz_create(device, &ctx_), nullptr, "rz_create"); }

Renderer::~Renderer() {
    if (mesh_) rz_mesh_destroy(ctx_, mesh_);
    if (ctx_) rz_destroy(ctx_);
}

void Renderer::set_tuning(int tile_size, int bin_capacity) {
    check(rz_set_tuning(ctx_, tile_size, bin_capacity), ctx_, "rz_set_tuning");
}

void Renderer::upload(const std::vector<Vec3>& vertices, const std::vector<Triangle>& faces,
                      const std::vector<Vec3>& vertnormals, const std::vector<Vec2>& vertuvs,
                      const std::vector<Material>& materials) {
    std::vector<rz_material> mats;
    for (const Material& m : materials) mats.push_back(m.as_c());
    const rz_mesh_desc d = describe(vertices, faces, vertnormals, vertuvs, mats);
    if (mesh_) {
        rz_mesh_destroy(ctx_, mesh_);
        mesh_ = nullptr;
    }
    check(rz_mesh_create(ctx_, &d, &mesh_), ctx_, "rz_mesh_create");
}

void Renderer::draw(std::vector<Light>& lights, const Args& arguments, FrameBuffer* frame, DepthBuffer* depth,
                    bool cleared, rz_frame_stats* stats) {
    if (!mesh_) throw std::runtime_error("Renderer::draw before upload");
    check_buffers(arguments, frame, depth);
    const rz_args a = arguments.as_c();
    check(rz_draw_frame(ctx_, mesh_, lights.data(), (int32_t)lights.size(), &a, frame->data.data(),
                        depth->data.data(), nullptr, cleared ? RZ_FLAG_CLEARED : 0u, stats),
          ctx_, "draw_frame");
}

}  // namespace rz
