// tinyobj_ref.cpp — TEST INFRASTRUCTURE ONLY.
//
// Loads an OBJ with the reference's vendored tinyobjloader (vendor/tinyobjloader/
// tiny_obj_loader.h, included from the reference tree at build time) and flattens it
// the way the reference's load_obj does (fileloader.cpp:79-121: components to
// vec3/vec2, one Triangle per three indices of each shape, shapes in order), so the
// host layer's own OBJ loader can be checked against the real parser.
// Built by oracle/Makefile into oracle/_ref/ only when the reference tree is present.
#define TINYOBJLOADER_IMPLEMENTATION
#include "tiny_obj_loader.h"

#include <cstdio>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: tinyobj_ref model.obj mtl_basedir out_prefix\n");
        return 2;
    }
    tinyobj::attrib_t attrib;
    std::vector<tinyobj::shape_t> shapes;
    std::vector<tinyobj::material_t> materials;
    std::string err;
    const bool ok = tinyobj::LoadObj(&attrib, &shapes, &materials, &err, argv[1], argv[2], true);
    if (!ok) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    const std::string p = argv[3];
    auto dump = [&](const std::string& name, const void* data, size_t bytes) {
        FILE* f = std::fopen((p + name).c_str(), "wb");
        if (bytes) std::fwrite(data, 1, bytes, f);
        std::fclose(f);
    };
    dump(".vertices.f32", attrib.vertices.data(), attrib.vertices.size() * 4);
    dump(".normals.f32", attrib.normals.data(), attrib.normals.size() * 4);
    dump(".uvs.f32", attrib.texcoords.data(), attrib.texcoords.size() * 4);
    std::vector<int> faces;
    for (const tinyobj::shape_t& s : shapes) {
        const std::vector<tinyobj::index_t>& ix = s.mesh.indices;
        for (size_t f = 0; f < s.mesh.material_ids.size(); ++f) {
            for (int k = 0; k < 3; ++k) faces.push_back(ix[3 * f + k].vertex_index);
            for (int k = 0; k < 3; ++k) faces.push_back(ix[3 * f + k].normal_index);
            for (int k = 0; k < 3; ++k) faces.push_back(ix[3 * f + k].texcoord_index);
            faces.push_back(s.mesh.material_ids[f]);
        }
    }
    dump(".faces.i32", faces.data(), faces.size() * 4);
    FILE* m = std::fopen((p + ".materials.txt").c_str(), "w");
    for (const tinyobj::material_t& mt : materials)
        std::fprintf(m, "%.9g %.9g %.9g %s\n", mt.diffuse[0], mt.diffuse[1], mt.diffuse[2],
                     mt.diffuse_texname.empty() ? "-" : mt.diffuse_texname.c_str());
    std::fclose(m);
    return 0;
}
