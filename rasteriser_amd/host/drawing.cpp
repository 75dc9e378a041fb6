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

// A 64-bit hash of n bytes, continuing from h: FNV-1a over 8-byte words (tail bytes one
// at a time), a word per multiply, about 1 ms for an 82k-face mesh and its texture.
uint64_t fnv1a(const void* data, size_t n, uint64_t h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * 0x100000001B3ull;
        h ^= h >> 29;
    }
    for (; i < n; ++i) h = (h ^ p[i]) * 0x100000001B3ull;
    return h;
}

// One librz context per host thread, and the mesh it last uploaded with the hash of
// everything that mesh was made from: draw_frame re-uploads only when the arrays,
// materials or textures it is given differ from the last call's (the spin loop draws
// the same ones every frame, renderer.cpp:111).
struct ThreadContext {
    rz_ctx* ctx = nullptr;
    int device = -1;
    rz_mesh* mesh = nullptr;
    uint64_t mesh_hash = 0;
    ~ThreadContext() {
        if (mesh) rz_mesh_destroy(ctx, mesh);
        if (ctx) rz_destroy(ctx);
    }
    rz_ctx* get(int dev) {
        if (ctx && device != dev) {
            if (mesh) rz_mesh_destroy(ctx, mesh);
            mesh = nullptr;
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
    // the mesh on the device: uploaded again only when its contents changed
    uint64_t h = 0xCBF29CE484222325ull;
    const size_t sizes[5] = {model_vertices.size(), faces.size(), model_vertnormals.size(), vertuvs.size(),
                             materials.size()};
    h = fnv1a(sizes, sizeof sizes, h);
    h = fnv1a(model_vertices.data(), model_vertices.size() * sizeof(Vec3), h);
    h = fnv1a(faces.data(), faces.size() * sizeof(Triangle), h);
    h = fnv1a(model_vertnormals.data(), model_vertnormals.size() * sizeof(Vec3), h);
    h = fnv1a(vertuvs.data(), vertuvs.size() * sizeof(Vec2), h);
    for (const rz_material& m : mats) {
        h = fnv1a(m.diffuse_colour, sizeof m.diffuse_colour, h);
        const int dims[3] = {m.tex_width, m.tex_height, m.tex_channels};
        h = fnv1a(dims, sizeof dims, h);
        if (m.tex_channels) h = fnv1a(m.texels, (size_t)m.tex_width * m.tex_height * m.tex_channels, h);
    }
    if (!t_ctx.mesh || t_ctx.mesh_hash != h) {
        if (t_ctx.mesh) rz_mesh_destroy(ctx, t_ctx.mesh);
        t_ctx.mesh = nullptr;
        const rz_mesh_desc d = describe(model_vertices, faces, model_vertnormals, vertuvs, mats);
        check(rz_mesh_create(ctx, &d, &t_ctx.mesh), ctx, "rz_mesh_create");
        t_ctx.mesh_hash = h;
    }
    const rz_args a = arguments.as_c();
    check(rz_draw_frame(ctx, t_ctx.mesh, lights.data(), (int32_t)lights.size(), &a, frame_buffer->data.data(),
                        depth_buffer->data.data(), nullptr, 0u, nullptr),
          ctx, "draw_frame");
}

DeviceFrame::DeviceFrame(rz_ctx* c, int w, int h) : ctx(c), width(w), height(h) {
    void* p = nullptr;
    check(rz_device_alloc(ctx, (size_t)w * h * 3, &p), ctx, "rz_device_alloc");
    frame = static_cast<uint8_t*>(p);
    p = nullptr;
    const rz_status st = rz_device_alloc(ctx, (size_t)w * h * 4, &p);
    if (st != RZ_OK) {
        rz_device_free(ctx, frame);
        check(st, ctx, "rz_device_alloc");
    }
    depth = static_cast<float*>(p);
}

DeviceFrame::~DeviceFrame() {
    rz_device_free(ctx, frame);
    rz_device_free(ctx, depth);
}

void DeviceFrame::download(FrameBuffer& f, DepthBuffer& d) const {
    if (f.width != width || f.height != height || d.width != width || d.height != height)
        throw std::runtime_error("DeviceFrame::download: buffer sizes differ");
    check(rz_copy_to_host(ctx, f.data.data(), frame, f.data.size()), ctx, "rz_copy_to_host");
    check(rz_copy_to_host(ctx, d.data.data(), depth, d.data.size() * sizeof(float)), ctx, "rz_copy_to_host");
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
