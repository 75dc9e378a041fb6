// renderer.cpp — command-line renderer over librz, with the reference's flags
// (arguments.cpp:15-33) and program flow (renderer.cpp:52-129): load the lights and
// the model (or the default square, renderer.cpp:32-50), draw one frame into cleared
// buffers and save frame.png and depth.png, or with -s run the spin loop (rotation
// about y by the elapsed time, renderer.cpp:94-126).  There is no window on a GPU
// server: the spin loop renders --frames frames into device buffers, replaying one
// captured frame graph with each frame's rotation (rz_spin_frame), draws the previous
// frame's rate into the frame at (5, 5) like the reference (rz_draw_text), prints it,
// and saves the last frame (with its rate) and depth; it prints the last frame's angles.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <stdexcept>

#include "rz_host.hpp"

using namespace rz;

namespace {

// renderer.cpp:32-50
void add_square(std::vector<Vec3>& vertices, std::vector<Triangle>& faces, std::vector<Vec3>& vertnormals,
                std::vector<Vec2>& vertuvs, std::vector<Material>& materials) {
    vertices = {{-0.5f, -0.5f, 0.f}, {0.5f, -0.5f, 0.f}, {-0.5f, 0.5f, 0.f}, {0.5f, 0.5f, 0.f}};
    vertnormals = {{0.f, 0.f, 1.f}};
    vertuvs = {{0.f, 0.f}, {1.f, 0.f}, {0.f, 1.f}, {1.f, 1.f}};
    faces = {Triangle{{0, 1, 2}, {0, 0, 0}, {0, 1, 2}, 0}, Triangle{{2, 1, 3}, {0, 0, 0}, {2, 1, 3}, 0}};
    materials.emplace_back(Vec3{1.f, 1.f, 1.f});
}

template <typename T>
void dump(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write((const char*)v.data(), (std::streamsize)(v.size() * sizeof(T)));
}

}  // namespace

int main(int argc, char** argv) {
    Args arguments(argc, argv);
    try {
        std::vector<Vec3> vertices, vertnormals;
        std::vector<Vec2> vertuvs;
        std::vector<Triangle> faces;
        std::vector<Material> materials;
        std::vector<Light> lights;
        load_lights(arguments.lights_file, lights);
        if (!arguments.obj_file.empty()) {
            load_obj(arguments, vertices, faces, vertnormals, vertuvs, materials);
            std::printf("Loaded %zu triangles from %s.\n", faces.size(), arguments.obj_file.c_str());
        } else {
            add_square(vertices, faces, vertnormals, vertuvs, materials);
        }
        if (!arguments.dump_mesh.empty()) {
            const std::string& p = arguments.dump_mesh;
            dump(p + ".vertices.f32", vertices);
            dump(p + ".normals.f32", vertnormals);
            dump(p + ".uvs.f32", vertuvs);
            dump(p + ".faces.i32", faces);
            dump(p + ".lights.f32", lights);
            std::ofstream m(p + ".materials.txt");
            for (const Material& mt : materials)
                m << mt.diffuse_colour().x << ' ' << mt.diffuse_colour().y << ' ' << mt.diffuse_colour().z << ' '
                  << (mt.has_texture() ? mt.texture().width : 0) << ' ' << (mt.has_texture() ? mt.texture().height : 0)
                  << ' ' << (mt.has_texture() ? mt.texture().channels : 0) << '\n';
            return 0;
        }
        const int W = (int)arguments.image_width, H = (int)arguments.image_height;
        FrameBuffer frame_buffer(W, H, 0);
        DepthBuffer depth_buffer(W, H, 1.f);
        const std::string dir = arguments.out_dir.empty() ? "." : arguments.out_dir;
        if (!arguments.spin) {
            draw_frame(vertices, faces, vertnormals, vertuvs, lights, materials, arguments, &frame_buffer,
                       &depth_buffer);
        } else {
            Renderer r(arguments.device);
            r.set_tuning(arguments.tile_size, arguments.bin_capacity);
            r.upload(vertices, faces, vertnormals, vertuvs, materials);
            rz_ctx* ctx = r.context();
            DeviceFrame dev(ctx, W, H);
            const rz_args a0 = arguments.as_c();
            rz_spin* spin = nullptr;
            // every iteration clears the buffers first (renderer.cpp:107-108): RZ_FLAG_CLEARED
            check(rz_spin_create(ctx, r.mesh(), lights.data(), (int32_t)lights.size(), &a0, dev.frame, dev.depth,
                                 nullptr, RZ_FLAG_CLEARED | RZ_FLAG_DEVICE_BUFFERS, &spin),
                  ctx, "rz_spin_create");
            const uint8_t white[3] = {255, 255, 255}, black[3] = {0, 0, 0};
            std::string frame_rate;   // empty before the first frame is timed (renderer.cpp:99)
            float angles[3] = {a0.tait_bryan_angles[0], a0.tait_bryan_angles[1], a0.tait_bryan_angles[2]};
            float last_angles[3] = {angles[0], angles[1], angles[2]};
            auto last = std::chrono::steady_clock::now();
            try {
                for (int k = 0; k < arguments.frames; ++k) {
                    check(rz_spin_frame(spin, angles), ctx, "rz_spin_frame");
                    check(rz_draw_text(ctx, dev.frame, (uint32_t)W, (uint32_t)H, 5, 5, frame_rate.c_str(), white,
                                       black),
                          ctx, "rz_draw_text");
                    check(rz_sync(ctx), ctx, "rz_sync");
                    const auto now = std::chrono::steady_clock::now();
                    const float dt = std::chrono::duration<float>(now - last).count();
                    last = now;
                    frame_rate = std::to_string(1.f / dt);   // renderer.cpp:118-121
                    std::printf("frame %d: %s fps\n", k, frame_rate.c_str());
                    for (int c = 0; c < 3; ++c) last_angles[c] = angles[c];
                    angles[1] += dt;   // renderer.cpp:124
                }
            } catch (...) {
                rz_spin_destroy(spin);
                throw;
            }
            rz_spin_destroy(spin);
            dev.download(frame_buffer, depth_buffer);
            std::printf("last frame angles: %.9g %.9g %.9g\n", last_angles[0], last_angles[1], last_angles[2]);
        }
        save_frame_png(dir + "/frame.png", frame_buffer.data.data(), W, H);
        save_depth_png(dir + "/depth.png", depth_buffer.data.data(), W, H);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
