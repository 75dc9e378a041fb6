// fileloader.cpp — OBJ/MTL and lights-CSV loading for the host layer.
//
// load_obj keeps the reference loader's results (fileloader.cpp:79-121 over
// tinyobjloader 1.0.x, vendor/tinyobjloader/tiny_obj_loader.h), because they fix the
// face indices and the input bits every later stage sees:
//   * numbers are converted with tinyobjloader's own decimal routine
//     (tiny_obj_loader.h:463-586, not correctly rounded): integer digits accumulate in
//     a double as m*10+d, fraction digit k adds d*10^-k (a table of double constants
//     for k < 8), an exponent e gives ldexp(m * 5^e, e), then the result is cast to
//     float; a field that fails to parse keeps its default (0);
//   * face corners are v, v/vt, v//vn or v/vt/vn; 1-based indices become 0-based,
//     0 stays 0, negative ones count back from the end, absent ones are -1
//     (tiny_obj_loader.h:414-418, 680-712);
//   * a polygon (i0, i1, ..., in) becomes the fan (i0, i[k-1], i[k]) for k = 2..n
//     (tiny_obj_loader.h:897-920);
//   * usemtl looks the name up among the materials loaded so far (-1 when unknown);
//     faces keep the material that was current when they were read; the face group is
//     flushed into the current shape on a material change and the shape is closed on
//     g/o — a shape whose last group was flushed by usemtl just before g/o is dropped,
//     as in tiny_obj_loader.h:1517-1640 — and all shapes are concatenated in file order
//     (fileloader.cpp:116-118);
//   * mtllib names are tried in order (materials_directory + name) until one opens;
//     newmtl starts a material, the first of a duplicated name keeps the lookup, the
//     last material is always appended; Kd is the diffuse colour and map_Kd's last
//     non-option token the texture, loaded from materials_directory + name
//     (fileloader.cpp:47-58).
// load_lights reads rows of 7 numbers (direction xyz, intensity, colour rgb) per line
// (fileloader.cpp:123-133), converted like an istream would; blank lines are skipped.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "rz_host.hpp"

namespace rz {
namespace {

bool is_space(char c) { return c == ' ' || c == '\t'; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

// tinyobjloader's decimal conversion (see the file comment).
bool parse_number(const char* s, const char* end, double* out) {
    if (s >= end) return false;
    const char* c = s;
    double m = 0.0;
    int sign = 1, esign = 1, expo = 0, digits = 0;
    if (*c == '+' || *c == '-') {
        sign = *c == '-' ? -1 : 1;
        ++c;
    } else if (!is_digit(*c)) {
        return false;
    }
    while (c != end && is_digit(*c)) {
        m = m * 10;
        m += (int)(*c - '0');
        ++c;
        ++digits;
    }
    if (digits == 0) return false;
    if (c != end) {
        bool exp_part = false;
        if (*c == '.') {
            ++c;
            static const double kNeg[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
            for (int k = 1; c != end && is_digit(*c); ++k, ++c)
                m += (int)(*c - '0') * (k < 8 ? kNeg[k] : std::pow(10.0, -k));
            exp_part = c != end;
        } else {
            exp_part = *c == 'e' || *c == 'E';
        }
        if (exp_part && (*c == 'e' || *c == 'E')) {
            ++c;
            if (c != end && (*c == '+' || *c == '-')) {
                esign = *c == '-' ? -1 : 1;
                ++c;
            } else if (!is_digit(*c)) {
                return false;
            }
            int ed = 0;
            while (c != end && is_digit(*c)) {
                expo = expo * 10 + (int)(*c - '0');
                ++c;
                ++ed;
            }
            expo *= esign;
            if (ed == 0) return false;
        }
    }
    *out = sign * (expo ? std::ldexp(m * std::pow(5.0, expo), expo) : m);
    return true;
}

float next_float(const char** tok, double dflt = 0.0) {
    *tok += std::strspn(*tok, " \t");
    const char* end = *tok + std::strcspn(*tok, " \t\r");
    double v = dflt;
    parse_number(*tok, end, &v);
    *tok = end;
    return (float)v;
}

int fix_index(int idx, int n) {
    if (idx > 0) return idx - 1;
    if (idx == 0) return 0;
    return n + idx;
}

struct Corner { int v = -1, vt = -1, vn = -1; };

Corner parse_corner(const char** tok, int nv, int nvn, int nvt) {
    Corner c;
    c.v = fix_index(std::atoi(*tok), nv);
    *tok += std::strcspn(*tok, "/ \t\r");
    if ((*tok)[0] != '/') return c;
    ++*tok;
    if ((*tok)[0] == '/') {
        ++*tok;
        c.vn = fix_index(std::atoi(*tok), nvn);
        *tok += std::strcspn(*tok, "/ \t\r");
        return c;
    }
    c.vt = fix_index(std::atoi(*tok), nvt);
    *tok += std::strcspn(*tok, "/ \t\r");
    if ((*tok)[0] != '/') return c;
    ++*tok;
    c.vn = fix_index(std::atoi(*tok), nvn);
    *tok += std::strcspn(*tok, "/ \t\r");
    return c;
}

std::string first_word(const char* tok) {
    tok += std::strspn(tok, " \t");
    return std::string(tok, tok + std::strcspn(tok, " \t\r"));
}

struct MtlEntry {
    std::string name;
    Vec3 kd{0.f, 0.f, 0.f};
    std::string map_kd;
};

std::string trim_line(std::string line) {
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    return line;
}

// map_Kd [options] name: options are '-' keys followed by their values.
std::string texture_name(const char* tok) {
    std::istringstream in(tok);
    std::vector<std::string> words;
    for (std::string w; in >> w;) words.push_back(w);
    static const std::map<std::string, int> kArgs = {
        {"-blendu", 1}, {"-blendv", 1}, {"-boost", 1}, {"-mm", 2}, {"-o", 3}, {"-s", 3}, {"-t", 3},
        {"-texres", 1}, {"-clamp", 1}, {"-bm", 1}, {"-imfchan", 1}, {"-type", 1}};
    std::string name;
    for (size_t i = 0; i < words.size(); ++i) {
        auto it = kArgs.find(words[i]);
        if (it == kArgs.end()) {
            name = words[i];
            break;
        }
        size_t take = 0;
        while ((int)take < it->second && i + 1 + take < words.size()) {
            const std::string& w = words[i + 1 + take];
            double d;
            const bool num = parse_number(w.c_str(), w.c_str() + w.size(), &d);
            if (!num && take > 0 && (it->first == "-o" || it->first == "-s" || it->first == "-t")) break;
            ++take;
        }
        i += take;
    }
    return name;
}

bool load_mtl(const std::string& path, std::vector<MtlEntry>& mats, std::map<std::string, int>& ids) {
    std::ifstream in(path);
    if (!in) return false;
    MtlEntry cur;
    for (std::string line; std::getline(in, line);) {
        line = trim_line(line);
        const size_t last = line.find_last_not_of(" \t");
        if (last == std::string::npos) continue;
        line.resize(last + 1);
        const char* t = line.c_str() + std::strspn(line.c_str(), " \t");
        if (!*t || *t == '#') continue;
        if (!std::strncmp(t, "newmtl", 6) && is_space(t[6])) {
            if (!cur.name.empty()) {
                ids.insert({cur.name, (int)mats.size()});
                mats.push_back(cur);
            }
            cur = MtlEntry();
            cur.name = first_word(t + 7);
        } else if (t[0] == 'K' && t[1] == 'd' && is_space(t[2])) {
            const char* p = t + 2;
            cur.kd.x = next_float(&p);
            cur.kd.y = next_float(&p);
            cur.kd.z = next_float(&p);
        } else if (!std::strncmp(t, "map_Kd", 6) && is_space(t[6])) {
            cur.map_kd = texture_name(t + 7);
        }
    }
    ids.insert({cur.name, (int)mats.size()});
    mats.push_back(cur);
    return true;
}

}  // namespace

void load_obj(const Args& args, std::vector<Vec3>& vertices, std::vector<Triangle>& triangles,
              std::vector<Vec3>& vertnormals, std::vector<Vec2>& vertuvs, std::vector<Material>& materials) {
    std::ifstream in(args.obj_file);
    if (!in) throw std::runtime_error("cannot open OBJ file " + args.obj_file);
    std::vector<float> v, vn, vt;
    std::vector<MtlEntry> mtl;
    std::map<std::string, int> mtl_ids;
    std::vector<std::vector<Corner>> group;            // faces not yet flushed
    std::vector<int> group_mat;
    std::vector<Triangle> shape;                       // current shape's triangles
    std::vector<Triangle> all;
    int material = -1;
    auto flush_group = [&]() -> bool {
        if (group.empty()) return false;
        for (size_t g = 0; g < group.size(); ++g) {
            const std::vector<Corner>& f = group[g];
            for (size_t k = 2; k < f.size(); ++k) {
                const Corner c[3] = {f[0], f[k - 1], f[k]};
                Triangle t{};
                for (int j = 0; j < 3; ++j) {
                    t.vertices[j] = c[j].v;
                    t.normals[j] = c[j].vn;
                    t.uvs[j] = c[j].vt;
                }
                t.material = group_mat[g];
                shape.push_back(t);
            }
        }
        group.clear();
        group_mat.clear();
        return true;
    };
    for (std::string line; std::getline(in, line);) {
        line = trim_line(line);
        const char* t = line.c_str() + std::strspn(line.c_str(), " \t");
        if (!*t || *t == '#') continue;
        if (t[0] == 'v' && is_space(t[1])) {
            const char* p = t + 2;
            for (int k = 0; k < 3; ++k) v.push_back(next_float(&p));
        } else if (t[0] == 'v' && t[1] == 'n' && is_space(t[2])) {
            const char* p = t + 3;
            for (int k = 0; k < 3; ++k) vn.push_back(next_float(&p));
        } else if (t[0] == 'v' && t[1] == 't' && is_space(t[2])) {
            const char* p = t + 3;
            for (int k = 0; k < 2; ++k) vt.push_back(next_float(&p));
        } else if (t[0] == 'f' && is_space(t[1])) {
            const char* p = t + 2;
            p += std::strspn(p, " \t");
            std::vector<Corner> f;
            while (*p && *p != '\r' && *p != '\n') {
                f.push_back(parse_corner(&p, (int)(v.size() / 3), (int)(vn.size() / 3), (int)(vt.size() / 2)));
                p += std::strspn(p, " \t\r");
            }
            group.push_back(f);
            group_mat.push_back(material);
        } else if (!std::strncmp(t, "usemtl", 6) && is_space(t[6])) {
            const std::string name = first_word(t + 7);
            auto it = mtl_ids.find(name);
            const int id = it == mtl_ids.end() ? -1 : it->second;
            if (id != material) {
                flush_group();
                material = id;
            }
        } else if (!std::strncmp(t, "mtllib", 6) && is_space(t[6])) {
            std::istringstream names(t + 7);
            for (std::string n; std::getline(names, n, ' ');) {
                if (n.empty()) continue;
                if (load_mtl(args.materials_directory + n, mtl, mtl_ids)) break;
            }
        } else if ((t[0] == 'g' || t[0] == 'o') && is_space(t[1])) {
            if (flush_group()) all.insert(all.end(), shape.begin(), shape.end());
            shape.clear();
        }
    }
    if (flush_group() || !shape.empty()) all.insert(all.end(), shape.begin(), shape.end());

    for (size_t i = 0; i + 2 < v.size(); i += 3) vertices.push_back({v[i], v[i + 1], v[i + 2]});
    for (size_t i = 0; i + 2 < vn.size(); i += 3) vertnormals.push_back({vn[i], vn[i + 1], vn[i + 2]});
    for (size_t i = 0; i + 1 < vt.size(); i += 2) vertuvs.push_back({vt[i], vt[i + 1]});
    for (const MtlEntry& m : mtl) {
        if (m.map_kd.empty())
            materials.emplace_back(m.kd);
        else
            materials.emplace_back(m.kd, args.materials_directory + m.map_kd);
    }
    triangles.insert(triangles.end(), all.begin(), all.end());
}

void load_lights(const std::string& file, std::vector<Light>& lights) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open lights file " + file);
    for (std::string line; std::getline(in, line);) {
        line = trim_line(line);
        if (line.find_first_not_of(" \t,") == std::string::npos) continue;
        float f[7];
        const char* p = line.c_str();
        for (int k = 0; k < 7; ++k) {
            char* end = nullptr;
            f[k] = std::strtof(p, &end);
            if (end == p) throw std::runtime_error("lights file " + file + ": expected 7 numbers per row");
            p = end + std::strspn(end, " \t");
            if (*p == ',') ++p;
        }
        lights.push_back(make_light({f[0], f[1], f[2]}, f[3], {f[4], f[5], f[6]}));
    }
}

}  // namespace rz
