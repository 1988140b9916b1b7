// pt_obj.cpp -- OBJ/MTL import and the reference-order BVH builder.
//
// The reference's input step (src/scene/obj_loader.h:120-237 load_obj,
// src/scene/obj_scene.h:8-32 obj_scene, src/bvh/bvh.h:132-171 bvh_node,
// src/flat/flat_export.h:12-123 flat_exporter) is restated here so that the
// same OBJ file, transform, aspect ratio and srand48 seed give the same flat
// arrays byte for byte:
//   * OBJ/MTL text follows the vendored tinyobjloader 2.0 (third_party/
//     tinyobjloader/tiny_obj_loader.h): its line splitting, its decimal
//     parser (tryParseDouble, :897-1028), index fixing (:819-850), quad split
//     on the shorter diagonal and ear clipping of larger polygons
//     (:1509-1930), and MTL keys and texture options (:1276-1358);
//   * materials map as load_obj maps them (slot 0 the 0.7 grey default, MTL
//     material i in slot i+1; emission, then glass, mirror, diffuse);
//   * the stage adds the floor and light spheres and the camera placement of
//     obj_scene;
//   * PTS_SPLIT_REFERENCE rebuilds the reference's hierarchy: drand48 after
//     srand48(seed) picks the axis of every bvh_node, a stable sort on the
//     boxes' minima (glibc's qsort is a merge sort) orders the objects, and
//     the halves split at n/2; the exporter then numbers nodes depth first
//     and materials and textures in first-encounter order.
// Other split methods build the same primitives with this library's own
// builder, numbering materials the same way.
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pt_scene_impl.h"

namespace pts {
std::vector<uint8_t> load_image(const std::string& path, int& width, int& height);
}

namespace {

using namespace pts;
namespace fs = std::filesystem;

// ---------------------------------------------------------------- drand48
// POSIX drand48/srand48: X' = (0x5DEECE66D X + 0xB) mod 2^48, value X'/2^48;
// srand48(s) sets X to (s mod 2^32) * 2^16 + 0x330E.
struct rand48 {
    uint64_t x;
    explicit rand48(int64_t seed) : x((uint64_t(uint32_t(seed)) << 16) | 0x330Eu) {}
    double next() {
        x = (0x5DEECE66Dull * x + 0xBull) & ((1ull << 48) - 1);
        return std::ldexp(double(x), -48);
    }
};

// ---------------------------------------------------------------- source scene
// Materials and textures as the reference's objects, before flat numbering.
struct src_texture {
    int32_t type = PT_CONSTANT;
    pt_vec3 color{0, 0, 0}, scale{1, 1, 0}, offset{0, 0, 0};
    int32_t clamp = 1;
    int32_t image = -1;             // PT_MAP: source texture index of the image
    int32_t width = 0, height = 0;  // PT_IMAGE
    std::vector<uint8_t> rgb;
};
struct src_material {
    int32_t type = PT_DIFFUSE;
    pt_vec3 albedo{0, 0, 0};
    float fuzz = 0, ref_idx = 1;
    int32_t texture = -1;
};
struct source_scene {
    std::vector<src_texture> textures;
    std::vector<src_material> materials;
    std::vector<pt_flat_triangle> triangles;  // material_id: source material index
    std::vector<pt_flat_sphere> spheres;
};

// flat_exporter's texture_id/material_id (flat_export.h:24-80): records in
// first-encounter order, an image before the map that reads it.
class flat_numbering {
public:
    flat_numbering(const source_scene& src, pts_scene& out) : source(src), scene(out) {
        tex_ids.assign(src.textures.size(), -1);
        mat_ids.assign(src.materials.size(), -1);
    }
    int32_t material(int32_t index) {
        if(index < 0 || size_t(index) >= source.materials.size()) throw std::runtime_error("cannot flatten null material");
        if(mat_ids[size_t(index)] >= 0) return mat_ids[size_t(index)];
        const src_material& m = source.materials[size_t(index)];
        pt_flat_material record = flat_material(PT_DIFFUSE, v3(0, 0, 0), 0, 1, -1);
        if(m.type == PT_DIFFUSE || m.type == PT_LIGHT) {
            record.type = m.type;
            record.texture_id = texture(m.texture);
        } else if(m.type == PT_METAL) {
            record.type = PT_METAL; record.albedo = m.albedo; record.fuzz = m.fuzz;
        } else if(m.type == PT_GLASS) {
            record.type = PT_GLASS; record.ref_idx = m.ref_idx;
        } else {
            throw std::runtime_error("unsupported flat material");
        }
        int32_t id = checked_index(scene.materials.size());
        scene.materials.push_back(record);
        mat_ids[size_t(index)] = id;
        return id;
    }

private:
    const source_scene& source;
    pts_scene& scene;
    std::vector<int32_t> tex_ids, mat_ids;

    static int32_t checked_index(size_t n) {
        if(n >= size_t(INT32_MAX)) throw std::runtime_error("flat array exceeds 32-bit index range");
        return int32_t(n);
    }
    int32_t texture(int32_t index) {
        if(index < 0 || size_t(index) >= source.textures.size()) throw std::runtime_error("cannot flatten null texture");
        if(tex_ids[size_t(index)] >= 0) return tex_ids[size_t(index)];
        const src_texture& t = source.textures[size_t(index)];
        pt_flat_texture record = constant_texture(v3(0, 0, 0));
        if(t.type == PT_CONSTANT) {
            record.color = t.color;
        } else if(t.type == PT_IMAGE) {
            if(t.width <= 0 || t.height <= 0 || t.rgb.empty()) throw std::runtime_error("cannot flatten empty image");
            uint64_t bytes = uint64_t(t.width) * uint64_t(t.height) * 3;
            if(bytes + scene.pixels.size() >= uint64_t(INT32_MAX))
                throw std::runtime_error("flat image storage exceeds 32-bit offsets");
            record.type = PT_IMAGE;
            record.width = t.width; record.height = t.height;
            record.byte_offset = checked_index(scene.pixels.size());
            scene.pixels.insert(scene.pixels.end(), t.rgb.begin(), t.rgb.begin() + std::ptrdiff_t(bytes));
        } else if(t.type == PT_MAP) {
            if(t.image < 0 || size_t(t.image) >= source.textures.size() || source.textures[size_t(t.image)].type != PT_IMAGE)
                throw std::runtime_error("flat diffuse map requires an image child");
            record.type = PT_MAP;
            record.image_id = texture(t.image);
            record.color = t.color; record.scale = t.scale; record.offset = t.offset; record.clamp = t.clamp;
        } else {
            throw std::runtime_error("unsupported flat texture");
        }
        int32_t id = checked_index(scene.textures.size());
        scene.textures.push_back(record);
        tex_ids[size_t(index)] = id;
        return id;
    }
};

// ---------------------------------------------------------------- reference-order hierarchy
// The reference's hitable tree: triangle and sphere leaves and bvh_node
// branches (left == right when a node holds one object).
struct ref_object {
    int32_t kind;       // PT_TRIANGLE, PT_SPHERE or PT_BRANCH
    int32_t index;      // source primitive
    int32_t left, right;
    box3 box;
};

class reference_tree {
public:
    explicit reference_tree(int64_t seed) : rng(seed) {}
    std::vector<ref_object> objects;

    int32_t leaf(int32_t kind, int32_t index, const box3& box) {
        objects.push_back(ref_object{kind, index, -1, -1, box});
        return int32_t(objects.size() - 1);
    }

    // bvh_node(l, n) (bvh.h:132-171) over items[begin, begin + n).
    int32_t node(std::vector<int32_t>& items, size_t begin, size_t n) {
        if(n == 0) throw std::invalid_argument("BVH requires a nonempty object array");
        const int axis = int(3 * rng.next());
        auto first = items.begin() + std::ptrdiff_t(begin);
        std::stable_sort(first, first + std::ptrdiff_t(n), [&](int32_t a, int32_t b) {
            return objects[size_t(a)].box.lo[axis] - objects[size_t(b)].box.lo[axis] < 0.0;
        });
        int32_t left, right;
        if(n == 1) {
            left = right = items[begin];
        } else if(n == 2) {
            left = items[begin]; right = items[begin + 1];
        } else {
            left = node(items, begin, n / 2);
            right = node(items, begin + n / 2, n - n / 2);
        }
        const box3& lb = objects[size_t(left)].box;
        const box3& rb = objects[size_t(right)].box;
        box3 box;
        for(int a = 0; a < 3; a++) {
            box.lo[a] = lb.lo[a] < rb.lo[a] ? lb.lo[a] : rb.lo[a];
            box.hi[a] = lb.hi[a] > rb.hi[a] ? lb.hi[a] : rb.hi[a];
        }
        objects.push_back(ref_object{PT_BRANCH, -1, left, right, box});
        return int32_t(objects.size() - 1);
    }

private:
    rand48 rng;
};

// flat_exporter::object (flat_export.h:85-123): depth-first, left before
// right, an aliased child emitted once.
class reference_export {
public:
    reference_export(const reference_tree& t, const source_scene& src, pts_scene& out)
        : tree(t), source(src), scene(out), numbering(src, out) {}
    void run(int32_t root) {
        scene.nodes.clear(); scene.triangles.clear(); scene.spheres.clear();
        scene.materials.clear(); scene.textures.clear(); scene.pixels.clear();
        scene.max_depth = 0;
        scene.root = emit(root, 1);
    }

private:
    const reference_tree& tree;
    const source_scene& source;
    pts_scene& scene;
    flat_numbering numbering;

    int32_t emit(int32_t index, int depth) {
        if(depth > 64) throw std::runtime_error("source BVH exceeds flat traversal stack capacity");
        const ref_object& o = tree.objects[size_t(index)];
        for(int a = 0; a < 3; a++)
            if(!std::isfinite(o.box.lo[a]) || !std::isfinite(o.box.hi[a])) throw std::runtime_error("nonfinite flat bounds");
        if(scene.nodes.size() >= size_t(INT32_MAX)) throw std::runtime_error("flat array exceeds 32-bit index range");
        int32_t id = int32_t(scene.nodes.size());
        pt_flat_node node{};
        node.minimum = v3(o.box.lo[0], o.box.lo[1], o.box.lo[2]);
        node.maximum = v3(o.box.hi[0], o.box.hi[1], o.box.hi[2]);
        node.type = PT_BRANCH; node.left = -1; node.right = -1; node.primitive = -1;
        scene.nodes.push_back(node);
        scene.max_depth = std::max(scene.max_depth, depth);
        if(o.kind == PT_BRANCH) {
            int32_t left = emit(o.left, depth + 1);
            int32_t right = o.left == o.right ? left : emit(o.right, depth + 1);
            scene.nodes[size_t(id)].left = left;
            scene.nodes[size_t(id)].right = right;
        } else if(o.kind == PT_TRIANGLE) {
            pt_flat_triangle t = source.triangles[size_t(o.index)];
            t.material_id = numbering.material(t.material_id);
            scene.nodes[size_t(id)].type = PT_TRIANGLE;
            scene.nodes[size_t(id)].primitive = int32_t(scene.triangles.size());
            scene.triangles.push_back(t);
        } else {
            pt_flat_sphere s = source.spheres[size_t(o.index)];
            s.material_id = numbering.material(s.material_id);
            scene.nodes[size_t(id)].type = PT_SPHERE;
            scene.nodes[size_t(id)].primitive = int32_t(scene.spheres.size());
            scene.spheres.push_back(s);
        }
        return id;
    }
};

// This library's builders over the same primitives, then first-encounter
// material numbering along the leaves in depth-first order.
void build_own(const source_scene& src, pts_scene& out, int split) {
    build_scene(out, src.triangles, src.spheres, split);
    out.materials.clear(); out.textures.clear(); out.pixels.clear();
    flat_numbering numbering(src, out);
    std::vector<int32_t> stack;
    if(out.root >= 0) stack.push_back(out.root);
    std::vector<char> seen(out.nodes.size(), 0);
    while(!stack.empty()) {
        int32_t n = stack.back();
        stack.pop_back();
        if(seen[size_t(n)]) continue;
        seen[size_t(n)] = 1;
        const pt_flat_node& node = out.nodes[size_t(n)];
        if(node.type == PT_BRANCH) {
            stack.push_back(node.right);
            stack.push_back(node.left);
        } else if(node.type == PT_TRIANGLE) {
            auto& t = out.triangles[size_t(node.primitive)];
            t.material_id = numbering.material(t.material_id);
        } else {
            auto& s = out.spheres[size_t(node.primitive)];
            s.material_id = numbering.material(s.material_id);
        }
    }
}

// ---------------------------------------------------------------- OBJ/MTL text
bool is_space(char c) { return c == ' ' || c == '\t'; }
bool is_new_line(char c) { return c == '\r' || c == '\n' || c == '\0'; }
bool is_digit(char c) { return unsigned(c - '0') < 10u; }

// tinyobjloader's decimal parser: the mantissa accumulated in double digit by
// digit, fraction digits scaled by 10^-k, then ldexp(m * 5^e, e).
bool parse_double(const char *s, const char *end, double *result) {
    if(s >= end) return false;
    double mantissa = 0.0;
    int exponent = 0;
    char sign = '+', exp_sign = '+';
    const char *c = s;
    bool leading_dot = false;
    if(*c == '+' || *c == '-') {
        sign = *c++;
        if(c != end && *c == '.') leading_dot = true;
    } else if(is_digit(*c)) {
    } else if(*c == '.') {
        leading_dot = true;
    } else {
        return false;
    }
    bool more = c != end;
    if(!leading_dot) {
        int read = 0;
        while(more && is_digit(*c)) {
            mantissa *= 10;
            mantissa += int(*c - '0');
            c++; read++;
            more = c != end;
        }
        if(read == 0) return false;
    }
    if(more) {
        bool exponent_next = false;
        if(*c == '.') {
            c++;
            int read = 1;
            more = c != end;
            static const double tenths[] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
            while(more && is_digit(*c)) {
                mantissa += int(*c - '0') * (read < 8 ? tenths[read] : std::pow(10.0, -read));
                read++; c++;
                more = c != end;
            }
            exponent_next = more;
        } else if(*c == 'e' || *c == 'E') {
            exponent_next = true;
        }
        if(exponent_next && (*c == 'e' || *c == 'E')) {
            c++;
            more = c != end;
            if(more && (*c == '+' || *c == '-')) exp_sign = *c++;
            else if(!is_digit(*c)) return false;
            int read = 0;
            more = c != end;
            while(more && is_digit(*c)) {
                if(exponent > 2147483647 / 10) return false;
                exponent = exponent * 10 + int(*c - '0');
                c++; read++;
                more = c != end;
            }
            exponent *= exp_sign == '+' ? 1 : -1;
            if(read == 0) return false;
        }
    }
    *result = (sign == '+' ? 1 : -1) * (exponent ? std::ldexp(mantissa * std::pow(5.0, exponent), exponent) : mantissa);
    return true;
}

float parse_real(const char **token, double fallback = 0.0) {
    *token += std::strspn(*token, " \t");
    const char *end = *token + std::strcspn(*token, " \t\r");
    double value = fallback;
    parse_double(*token, end, &value);
    *token = end;
    return float(value);
}
bool parse_real(const char **token, float *out) {
    *token += std::strspn(*token, " \t");
    const char *end = *token + std::strcspn(*token, " \t\r");
    double value;
    bool ok = parse_double(*token, end, &value);
    if(ok) *out = float(value);
    *token = end;
    return ok;
}
void parse_real3(const char **token, float *x, float *y, float *z, double dx = 0, double dy = 0, double dz = 0) {
    *x = parse_real(token, dx); *y = parse_real(token, dy); *z = parse_real(token, dz);
}
int parse_int(const char **token) {
    *token += std::strspn(*token, " \t");
    int value = std::atoi(*token);
    *token += std::strcspn(*token, " \t\r");
    return value;
}
std::string parse_string(const char **token) {
    *token += std::strspn(*token, " \t");
    size_t e = std::strcspn(*token, " \t\r");
    std::string s(*token, *token + e);
    *token += e;
    return s;
}
bool parse_on_off(const char **token, bool fallback) {
    *token += std::strspn(*token, " \t");
    const char *end = *token + std::strcspn(*token, " \t\r");
    bool value = fallback;
    if(std::strncmp(*token, "on", 2) == 0) value = true;
    else if(std::strncmp(*token, "off", 3) == 0) value = false;
    *token = end;
    return value;
}

// Lines split at "\n", "\r\n" or "\r", as tinyobjloader's safeGetline.
std::vector<std::string> split_lines(const std::string& text) {
    std::vector<std::string> lines;
    std::string line;
    for(size_t i = 0; i < text.size(); i++) {
        char c = text[i];
        if(c == '\n') { lines.push_back(line); line.clear(); }
        else if(c == '\r') {
            if(i + 1 < text.size() && text[i + 1] == '\n') i++;
            lines.push_back(line); line.clear();
        } else line += c;
    }
    if(!line.empty()) lines.push_back(line);
    return lines;
}

std::string read_file(const fs::path& path, bool& ok) {
    std::ifstream in(path, std::ios::binary);
    ok = bool(in);
    if(!ok) return std::string();
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

struct texture_option {
    float scale[3] = {1, 1, 1}, offset[3] = {0, 0, 0};
    bool clamp = false;
};

struct mtl_material {
    std::string name;
    float diffuse[3] = {0, 0, 0}, specular[3] = {0, 0, 0}, emission[3] = {0, 0, 0};
    float ior = 1, dissolve = 1, shininess = 1, roughness = 0, metallic = 0;
    int illum = 0;
    std::string diffuse_texname, alpha_texname, bump_texname, normal_texname, emissive_texname;
    texture_option diffuse_opt;
    fs::path directory;  // of the MTL file that declared it
};

// tinyobjloader's ParseTextureNameAndOption: options, then the rest of the
// line is the file name (which may contain spaces).
bool parse_texture(const char *token, std::string& name, texture_option& opt) {
    bool found = false;
    float ignored[3];
    while(!is_new_line(*token)) {
        token += std::strspn(token, " \t");
        auto option = [&](const char *key) {
            size_t n = std::strlen(key);
            return std::strncmp(token, key, n) == 0 && is_space(token[n]);
        };
        if(option("-blendu") || option("-blendv")) { token += 8; parse_on_off(&token, true); }
        else if(option("-clamp")) { token += 7; opt.clamp = parse_on_off(&token, true); }
        else if(option("-boost")) { token += 7; parse_real(&token, 1.0); }
        else if(option("-bm")) { token += 4; parse_real(&token, 1.0); }
        else if(option("-o")) { token += 3; parse_real3(&token, &opt.offset[0], &opt.offset[1], &opt.offset[2]); }
        else if(option("-s")) { token += 3; parse_real3(&token, &opt.scale[0], &opt.scale[1], &opt.scale[2], 1, 1, 1); }
        else if(option("-t")) { token += 3; parse_real3(&token, &ignored[0], &ignored[1], &ignored[2]); }
        else if(option("-type")) { token += 5; token += std::strspn(token, " \t"); token += std::strcspn(token, " \t\r"); }
        else if(option("-texres")) { token += 7; parse_int(&token); }
        else if(option("-imfchan")) { token += 9; token += std::strspn(token, " \t"); token += std::strcspn(token, " \t\r"); }
        else if(option("-mm")) { token += 4; parse_real(&token, 0.0); parse_real(&token, 1.0); }
        else if(option("-colorspace")) { token += 12; parse_string(&token); }
        else {
            name = std::string(token);
            token += name.size();
            found = true;
        }
    }
    return found;
}

// tinyobjloader's LoadMtl, for the keys the import reads.
void load_mtl(const std::string& text, const fs::path& directory, std::vector<mtl_material>& materials,
              std::map<std::string, int>& index, std::string& warnings) {
    mtl_material m;
    bool has_d = false, has_tr = false, has_kd = false;
    size_t line_no = 0;
    std::ostringstream tr_warnings;
    auto flush = [&]() {
        index.insert({m.name, int(materials.size())});
        m.directory = directory;
        materials.push_back(m);
    };
    for(std::string line : split_lines(text)) {
        line_no++;
        if(!line.empty()) line = line.substr(0, line.find_last_not_of(" \t") + 1);
        if(!line.empty() && line.back() == '\n') line.pop_back();
        if(!line.empty() && line.back() == '\r') line.pop_back();
        if(line.empty()) continue;
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if(t[0] == '\0' || t[0] == '#') continue;
        auto key = [&](const char *k) {
            size_t n = std::strlen(k);
            return std::strncmp(t, k, n) == 0 && is_space(t[n]);
        };
        if(key("newmtl")) {
            if(!m.name.empty()) flush();
            m = mtl_material();
            has_d = has_tr = false;
            t += 7;
            m.name = parse_string(&t);
            if(m.name.empty()) warnings += "empty material name in `newmtl`\n";
        } else if(key("Kd")) { t += 2; parse_real3(&t, &m.diffuse[0], &m.diffuse[1], &m.diffuse[2]); has_kd = true; }
        else if(key("Ks")) { t += 2; parse_real3(&t, &m.specular[0], &m.specular[1], &m.specular[2]); }
        else if(key("Ni")) { t += 2; m.ior = parse_real(&t); }
        else if(key("Ke")) { t += 2; parse_real3(&t, &m.emission[0], &m.emission[1], &m.emission[2]); }
        else if(key("Ns")) { t += 2; m.shininess = parse_real(&t); }
        else if(key("illum")) { t += 6; m.illum = parse_int(&t); }
        else if(key("d")) {
            t += 1;
            m.dissolve = parse_real(&t);
            if(has_tr)
                tr_warnings << "Both `d` and `Tr` parameters defined for \"" << m.name
                            << "\". Use the value of `d` for dissolve (line " << line_no << " in .mtl.)\n";
            has_d = true;
        } else if(key("Tr")) {
            t += 2;
            if(has_d) {
                tr_warnings << "Both `d` and `Tr` parameters defined for \"" << m.name
                            << "\". Use the value of `d` for dissolve (line " << line_no << " in .mtl.)\n";
            } else {
                m.dissolve = 1.0f - parse_real(&t);
            }
            has_tr = true;
        } else if(key("Pr")) { t += 2; m.roughness = parse_real(&t); }
        else if(key("Pm")) { t += 2; m.metallic = parse_real(&t); }
        else if(key("map_Kd")) {
            t += 7;
            parse_texture(t, m.diffuse_texname, m.diffuse_opt);
            if(!has_kd) m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = 0.6f;
        } else if(key("map_d")) {
            t += 6;
            texture_option o;
            m.alpha_texname = t;
            parse_texture(t, m.alpha_texname, o);
        } else if(key("map_bump") || key("map_Bump")) { t += 9; texture_option o; parse_texture(t, m.bump_texname, o); }
        else if(key("bump")) { t += 5; texture_option o; parse_texture(t, m.bump_texname, o); }
        else if(key("map_Ke")) { t += 7; texture_option o; parse_texture(t, m.emissive_texname, o); }
        else if(key("norm")) { t += 5; texture_option o; parse_texture(t, m.normal_texname, o); }
    }
    flush();
    warnings += tr_warnings.str();
}

struct vertex_index { int v = -1, vt = -1, vn = -1; };
struct obj_face { std::vector<vertex_index> vertices; };
struct obj_triangle { vertex_index idx[3]; int material; };
struct obj_shape { std::string name; std::vector<obj_triangle> triangles; bool lines_or_points = false; };

// tinyobjloader's fixIndex: 1-based, negative relative to the count so far.
bool fix_index(int idx, int n, int *out, bool allow_zero, size_t line, std::string& warnings) {
    if(idx > 0) { *out = idx - 1; return true; }
    if(idx == 0) {
        warnings += "A zero value index found (will have a value of -1 for normal and tex indices. Line " +
                    std::to_string(line) + ").\n";
        *out = -1;
        return allow_zero;
    }
    *out = n + idx;
    return *out >= 0;
}

bool parse_triple(const char **token, int vn, int nn, int tn, vertex_index *out, size_t line, std::string& warnings) {
    vertex_index vi;
    if(!fix_index(std::atoi(*token), vn, &vi.v, false, line, warnings)) return false;
    *token += std::strcspn(*token, "/ \t\r");
    if((*token)[0] != '/') { *out = vi; return true; }
    (*token)++;
    if((*token)[0] == '/') {
        (*token)++;
        if(!fix_index(std::atoi(*token), nn, &vi.vn, true, line, warnings)) return false;
        *token += std::strcspn(*token, "/ \t\r");
        *out = vi;
        return true;
    }
    if(!fix_index(std::atoi(*token), tn, &vi.vt, true, line, warnings)) return false;
    *token += std::strcspn(*token, "/ \t\r");
    if((*token)[0] != '/') { *out = vi; return true; }
    (*token)++;
    if(!fix_index(std::atoi(*token), nn, &vi.vn, true, line, warnings)) return false;
    *token += std::strcspn(*token, "/ \t\r");
    *out = vi;
    return true;
}

// W. R. Franklin's crossing test, as tinyobjloader's pnpoly.
bool point_in_triangle(const float *vx, const float *vy, float tx, float ty) {
    bool inside = false;
    for(int i = 0, j = 2; i < 3; j = i++) {
        if(((vy[i] > ty) != (vy[j] > ty)) && (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]))
            inside = !inside;
    }
    return inside;
}

// tinyobjloader's exportGroupsToShape for faces: triangles kept, quads split
// on the shorter diagonal, larger polygons ear-clipped.
void triangulate(const std::vector<obj_face>& faces, int material, const std::vector<float>& v, obj_shape& shape,
                 std::string& warnings) {
    auto push = [&](vertex_index a, vertex_index b, vertex_index c) {
        obj_triangle t;
        t.idx[0] = a; t.idx[1] = b; t.idx[2] = c; t.material = material;
        shape.triangles.push_back(t);
    };
    for(const obj_face& face : faces) {
        size_t n = face.vertices.size();
        if(n < 3) { warnings += "Degenerated face found\n."; continue; }
        if(n == 3) { push(face.vertices[0], face.vertices[1], face.vertices[2]); continue; }
        if(n == 4) {
            const vertex_index *q = face.vertices.data();
            bool valid = true;
            for(int k = 0; k < 4; k++) if(3 * size_t(q[k].v) + 2 >= v.size()) valid = false;
            if(!valid) { warnings += "Face with invalid vertex index found.\n"; continue; }
            auto at = [&](int k, int a) { return v[size_t(q[k].v) * 3 + size_t(a)]; };
            float e02x = at(2, 0) - at(0, 0), e02y = at(2, 1) - at(0, 1), e02z = at(2, 2) - at(0, 2);
            float e13x = at(3, 0) - at(1, 0), e13y = at(3, 1) - at(1, 1), e13z = at(3, 2) - at(1, 2);
            float sqr02 = e02x * e02x + e02y * e02y + e02z * e02z;
            float sqr13 = e13x * e13x + e13y * e13y + e13z * e13z;
            if(sqr02 < sqr13) { push(q[0], q[1], q[2]); push(q[0], q[2], q[3]); }
            else { push(q[0], q[1], q[3]); push(q[1], q[2], q[3]); }
            continue;
        }
        // Projection plane from the first non-degenerate corner.
        size_t axes[2] = {1, 2};
        for(size_t k = 0; k < n; k++) {
            size_t i0 = size_t(face.vertices[k].v), i1 = size_t(face.vertices[(k + 1) % n].v),
                   i2 = size_t(face.vertices[(k + 2) % n].v);
            if(3 * i0 + 2 >= v.size() || 3 * i1 + 2 >= v.size() || 3 * i2 + 2 >= v.size()) continue;
            float e0x = v[i1 * 3] - v[i0 * 3], e0y = v[i1 * 3 + 1] - v[i0 * 3 + 1], e0z = v[i1 * 3 + 2] - v[i0 * 3 + 2];
            float e1x = v[i2 * 3] - v[i1 * 3], e1y = v[i2 * 3 + 1] - v[i1 * 3 + 1], e1z = v[i2 * 3 + 2] - v[i1 * 3 + 2];
            float cx = std::fabs(e0y * e1z - e0z * e1y), cy = std::fabs(e0z * e1x - e0x * e1z),
                  cz = std::fabs(e0x * e1y - e0y * e1x);
            const float eps = std::numeric_limits<float>::epsilon();
            if(cx > eps || cy > eps || cz > eps) {
                if(!(cx > cy && cx > cz)) {
                    axes[0] = 0;
                    if(cz > cx && cz > cy) axes[1] = 1;
                }
                break;
            }
        }
        std::vector<vertex_index> remaining = face.vertices;
        size_t guess = 0, iterations = remaining.size(), previous = remaining.size();
        vertex_index ind[3];
        float vx[3], vy[3];
        while(remaining.size() > 3 && iterations > 0) {
            size_t m = remaining.size();
            if(guess >= m) guess -= m;
            if(previous != m) { previous = m; iterations = m; }
            else iterations--;
            for(size_t k = 0; k < 3; k++) {
                ind[k] = remaining[(guess + k) % m];
                size_t vi = size_t(ind[k].v);
                if(vi * 3 + axes[0] >= v.size() || vi * 3 + axes[1] >= v.size()) { vx[k] = 0; vy[k] = 0; }
                else { vx[k] = v[vi * 3 + axes[0]]; vy[k] = v[vi * 3 + axes[1]]; }
            }
            float e0x = vx[1] - vx[0], e0y = vy[1] - vy[0], e1x = vx[2] - vx[1], e1y = vy[2] - vy[1];
            float cross_z = e0x * e1y - e0y * e1x;
            float area = (vx[0] * vy[1] - vy[0] * vx[1]) * 0.5f;
            if(cross_z * area < 0.0f) { guess++; continue; }
            bool overlap = false;
            for(size_t other = 3; other < m; other++) {
                size_t idx = (guess + other) % m;
                if(idx >= remaining.size()) continue;
                size_t ovi = size_t(remaining[idx].v);
                if(ovi * 3 + axes[0] >= v.size() || ovi * 3 + axes[1] >= v.size()) continue;
                if(point_in_triangle(vx, vy, v[ovi * 3 + axes[0]], v[ovi * 3 + axes[1]])) { overlap = true; break; }
            }
            if(overlap) { guess++; continue; }
            push(ind[0], ind[1], ind[2]);
            size_t removed = (guess + 1) % m;
            while(removed + 1 < m) { remaining[removed] = remaining[removed + 1]; removed++; }
            remaining.pop_back();
        }
        if(remaining.size() == 3) push(remaining[0], remaining[1], remaining[2]);
    }
}

// One OBJ file as tinyobjloader returns it (triangulated shapes in file order).
struct obj_data {
    std::vector<float> v, vn, vt;
    std::vector<obj_shape> shapes;
    std::vector<mtl_material> materials;
    std::string warnings;
};

void load_obj_text(const std::string& text, const fs::path& directory, obj_data& out) {
    std::map<std::string, int> material_index;
    std::set<std::string> loaded_libraries;
    std::vector<obj_face> faces;
    bool lines_or_points = false;
    int material = -1;
    std::string name;
    obj_shape shape;
    int greatest_v = -1, greatest_vn = -1, greatest_vt = -1;
    auto flush_faces = [&]() {
        bool any = !faces.empty() || lines_or_points;
        if(!faces.empty()) {
            shape.name = name;
            triangulate(faces, material, out.v, shape, out.warnings);
        }
        if(lines_or_points) { shape.name = name; shape.lines_or_points = true; }
        return any;
    };
    size_t line_no = 0;
    for(std::string line : split_lines(text)) {
        line_no++;
        if(!line.empty() && line.back() == '\n') line.pop_back();
        if(!line.empty() && line.back() == '\r') line.pop_back();
        if(line.empty()) continue;
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if(t[0] == '\0' || t[0] == '#') continue;
        if(t[0] == 'v' && is_space(t[1])) {
            t += 2;
            float x = parse_real(&t), y = parse_real(&t), z = parse_real(&t);
            out.v.push_back(x); out.v.push_back(y); out.v.push_back(z);
            continue;
        }
        if(t[0] == 'v' && t[1] == 'n' && is_space(t[2])) {
            t += 3;
            float x, y, z;
            parse_real3(&t, &x, &y, &z);
            out.vn.push_back(x); out.vn.push_back(y); out.vn.push_back(z);
            continue;
        }
        if(t[0] == 'v' && t[1] == 't' && is_space(t[2])) {
            t += 3;
            float x = parse_real(&t), y = parse_real(&t);
            out.vt.push_back(x); out.vt.push_back(y);
            continue;
        }
        const int vn_count = int(out.v.size() / 3), nn_count = int(out.vn.size() / 3), tn_count = int(out.vt.size() / 2);
        if((t[0] == 'l' || t[0] == 'p') && is_space(t[1])) {
            const char kind = t[0];
            t += 2;
            while(!is_new_line(t[0])) {
                vertex_index vi;
                if(!parse_triple(&t, vn_count, nn_count, tn_count, &vi, line_no, out.warnings))
                    throw std::runtime_error(std::string("Failed to parse `") + kind +
                                             "' line (e.g. a zero value for vertex index. Line " +
                                             std::to_string(line_no) + ").");
                t += std::strspn(t, " \t\r");
            }
            lines_or_points = true;
            continue;
        }
        if(t[0] == 'f' && is_space(t[1])) {
            t += 2;
            t += std::strspn(t, " \t");
            obj_face face;
            while(!is_new_line(t[0])) {
                vertex_index vi;
                if(!parse_triple(&t, vn_count, nn_count, tn_count, &vi, line_no, out.warnings))
                    throw std::runtime_error("Failed to parse `f' line (e.g. a zero value for vertex index or invalid "
                                             "relative vertex index). Line " + std::to_string(line_no) + ").");
                greatest_v = std::max(greatest_v, vi.v);
                greatest_vn = std::max(greatest_vn, vi.vn);
                greatest_vt = std::max(greatest_vt, vi.vt);
                face.vertices.push_back(vi);
                t += std::strspn(t, " \t\r");
            }
            faces.push_back(face);
            continue;
        }
        if(std::strncmp(t, "usemtl", 6) == 0) {
            t += 6;
            std::string mtl = parse_string(&t);
            int next = -1;
            auto it = material_index.find(mtl);
            if(it != material_index.end()) next = it->second;
            else out.warnings += "material [ '" + mtl + "' ] not found in .mtl\n";
            if(next != material) {
                if(!faces.empty()) { shape.name = name; triangulate(faces, material, out.v, shape, out.warnings); }
                faces.clear();
                material = next;
            }
            continue;
        }
        if(std::strncmp(t, "mtllib", 6) == 0 && is_space(t[6])) {
            t += 7;
            // SplitString(' ', '\\'): backslash escapes, empty pieces dropped but a final one kept.
            std::vector<std::string> files;
            std::string piece;
            bool escaping = false;
            for(const char *c = t; *c; c++) {
                if(escaping) { escaping = false; }
                else if(*c == '\\') { escaping = true; continue; }
                else if(*c == ' ') { if(!piece.empty()) files.push_back(piece); piece.clear(); continue; }
                piece += *c;
            }
            files.push_back(piece);
            bool found = false;
            for(const std::string& file : files) {
                if(loaded_libraries.count(file)) { found = true; continue; }
                std::string clean = file;
                for(char& ch : clean) if(ch == '\\') ch = '/';
                fs::path path = directory / fs::path(clean);
                bool ok = false;
                std::string mtl_text = read_file(path, ok);
                if(!ok)
                    throw std::runtime_error("could not load MTL: " + path.string() +
                                             "\nCheck the OBJ mtllib filename and keep the referenced MTL beside the "
                                             "OBJ (or at that relative path).");
                load_mtl(mtl_text, path.parent_path(), out.materials, material_index, out.warnings);
                loaded_libraries.insert(file);
                found = true;
                break;
            }
            if(!found) out.warnings += "Failed to load material file(s). Use default material.\n";
            continue;
        }
        if((t[0] == 'g' || t[0] == 'o') && is_space(t[1])) {
            const bool group = t[0] == 'g';
            flush_faces();
            faces.clear();
            bool keep = !shape.triangles.empty() || (!group && shape.lines_or_points);
            if(keep) out.shapes.push_back(shape);
            shape = obj_shape();
            lines_or_points = false;
            if(group) {
                std::vector<std::string> names;
                while(!is_new_line(t[0])) {
                    names.push_back(parse_string(&t));
                    t += std::strspn(t, " \t\r");
                }
                if(names.size() < 2) {
                    out.warnings += "Empty group name. line: " + std::to_string(line_no) + "\n";
                    name = "";
                } else {
                    name = names[1];
                    for(size_t i = 2; i < names.size(); i++) name += " " + names[i];
                }
            } else {
                name = std::string(t + 2);
            }
            continue;
        }
    }
    if(greatest_v >= int(out.v.size() / 3))
        out.warnings += "Vertex indices out of bounds (line " + std::to_string(line_no) + ".)\n\n";
    if(greatest_vn >= int(out.vn.size() / 3))
        out.warnings += "Vertex normal indices out of bounds (line " + std::to_string(line_no) + ".)\n\n";
    if(greatest_vt >= int(out.vt.size() / 2))
        out.warnings += "Vertex texcoord indices out of bounds (line " + std::to_string(line_no) + ".)\n\n";
    bool any = flush_faces();
    if(any || !shape.triangles.empty()) out.shapes.push_back(shape);
}

// ---------------------------------------------------------------- mesh transform
// src/transforms/mesh_transform.h:10-75: scale, X/Y/Z rotation in degrees, translation.
class mesh_transform {
public:
    mesh_transform(const float scale[3], const float degrees[3], const float translation[3]) {
        for(int i = 0; i < 3; i++) {
            if(!std::isfinite(scale[i]) || scale[i] == 0 || !std::isfinite(degrees[i]) || !std::isfinite(translation[i]))
                throw std::invalid_argument("mesh transform requires finite values and nonzero scale");
            factors[i] = scale[i]; offset[i] = translation[i];
            double angle = std::remainder(double(degrees[i]), 360.0) * 3.14159265358979323846 / 180.0;
            sine[i] = float(std::sin(angle));
            cosine[i] = float(std::cos(angle));
            if(scale[i] != 1 || degrees[i] != 0 || translation[i] != 0) unchanged = false;
            if(scale[i] < 0) mirrored = !mirrored;
        }
    }
    void apply(pt_flat_triangle& f) const {
        if(unchanged) return;
        f.v0 = position(f.v0); f.v1 = position(f.v1); f.v2 = position(f.v2);
        if(f.smooth) { f.n0 = normal(f.n0); f.n1 = normal(f.n1); f.n2 = normal(f.n2); }
        if(mirrored) { std::swap(f.v1, f.v2); std::swap(f.n1, f.n2); std::swap(f.uv1, f.uv2); }
        pt_vec3 c = cross(f.v1 - f.v0, f.v2 - f.v0);
        float area = c.x * c.x + c.y * c.y + c.z * c.z;
        if(!(area > 0) || !std::isfinite(area))
            throw std::runtime_error("mesh transform collapses geometry or exceeds supported float range");
    }

private:
    float factors[3], offset[3], sine[3], cosine[3];
    bool unchanged = true, mirrored = false;

    pt_vec3 rotate(pt_vec3 v) const {
        v = v3(v.x, cosine[0] * v.y - sine[0] * v.z, sine[0] * v.y + cosine[0] * v.z);
        v = v3(cosine[1] * v.x + sine[1] * v.z, v.y, -sine[1] * v.x + cosine[1] * v.z);
        return v3(cosine[2] * v.x - sine[2] * v.y, sine[2] * v.x + cosine[2] * v.y, v.z);
    }
    pt_vec3 position(pt_vec3 p) const {
        pt_vec3 r = rotate(v3(p.x * factors[0], p.y * factors[1], p.z * factors[2]));
        r = v3(r.x + offset[0], r.y + offset[1], r.z + offset[2]);
        if(!std::isfinite(r.x) || !std::isfinite(r.y) || !std::isfinite(r.z))
            throw std::runtime_error("transformed position exceeds float range");
        return r;
    }
    pt_vec3 normal(pt_vec3 n) const {
        double x = double(n.x) / factors[0], y = double(n.y) / factors[1], z = double(n.z) / factors[2];
        double len = std::sqrt(x * x + y * y + z * z);
        if(!(len > 0) || !std::isfinite(len)) throw std::runtime_error("invalid transformed normal");
        return unit(rotate(v3(float(x / len), float(y / len), float(z / len))));
    }
};

// ---------------------------------------------------------------- load_obj
bool finite3(pt_vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

pt_vec3 attribute(const std::vector<float>& values, int index, int width) {
    if(index < 0 || size_t(index) >= values.size() / size_t(width)) throw std::runtime_error("OBJ attribute index is out of range");
    size_t o = size_t(index) * size_t(width);
    pt_vec3 r = v3(values[o], values[o + 1], width == 3 ? values[o + 2] : 0.0f);
    if(!finite3(r)) throw std::runtime_error("OBJ contains a non-finite attribute");
    return r;
}

pt_vec3 mtl_color(const float c[3], bool reflectance) {
    pt_vec3 v = v3(c[0], c[1], c[2]);
    if(!finite3(v)) throw std::runtime_error("MTL contains a non-finite color");
    float *e[3] = {&v.x, &v.y, &v.z};
    for(int i = 0; i < 3; i++) {
        *e[i] = *e[i] > 0 ? *e[i] : 0;          // ffmax(0, value)
        if(reflectance) *e[i] = 1 < *e[i] ? 1 : *e[i];  // ffmin(1, value)
    }
    return v;
}

struct obj_import {
    source_scene source;
    int32_t mesh_triangles = 0;
    int64_t source_vertices = 0, shape_count = 0;
    std::string warnings;
};

void import_obj(const std::string& filename, const mesh_transform& transform, obj_import& out) {
    std::string clean = filename;
    for(char& ch : clean) if(ch == '\\') ch = '/';
    fs::path path(clean);
    bool ok = false;
    std::string text = read_file(path, ok);
    if(!ok) throw std::runtime_error("could not load OBJ: " + filename);
    obj_data data;
    try {
        load_obj_text(text, path.parent_path(), data);
    } catch(const std::runtime_error& e) {
        throw std::runtime_error("OBJ load failed: " + filename + "\n" + e.what());
    }
    out.warnings = data.warnings;
    out.source_vertices = int64_t(data.v.size() / 3);
    out.shape_count = int64_t(data.shapes.size());
    source_scene& src = out.source;

    // Slot 0: the default grey Lambertian; MTL material i: slot i + 1.
    src_texture grey;
    grey.color = v3(0.7f, 0.7f, 0.7f);
    src.textures.push_back(grey);
    src_material fallback;
    fallback.texture = 0;
    src.materials.push_back(fallback);
    std::map<std::string, int32_t> images;
    for(const mtl_material& m : data.materials) {
        pt_vec3 diffuse = mtl_color(m.diffuse, true);
        pt_vec3 emission = mtl_color(m.emission, false);
        bool glass = m.illum == 4 || m.illum == 6 || m.illum == 7 || m.illum == 9;
        bool mirror = m.illum == 3 || m.illum == 5 || m.illum == 8;
        src_material record;
        if(emission.x * emission.x + emission.y * emission.y + emission.z * emission.z > 0) {
            src_texture c;
            c.color = emission;
            src.textures.push_back(c);
            record.type = PT_LIGHT;
            record.texture = int32_t(src.textures.size() - 1);
        } else if(glass) {
            if(!(m.ior > 0) || !std::isfinite(m.ior)) throw std::runtime_error("invalid MTL refractive index");
            record.type = PT_GLASS;
            record.ref_idx = m.ior;
        } else if(mirror) {
            record.type = PT_METAL;
            record.albedo = mtl_color(m.specular, true);
            record.fuzz = 0;
        } else if(m.diffuse_texname.empty()) {
            src_texture c;
            c.color = diffuse;
            src.textures.push_back(c);
            record.texture = int32_t(src.textures.size() - 1);
        } else {
            std::string tex = m.diffuse_texname;
            for(char& ch : tex) if(ch == '\\') ch = '/';
            std::string key = (m.directory / fs::path(tex)).lexically_normal().string();
            if(!images.count(key)) {
                src_texture image;
                image.type = PT_IMAGE;
                image.rgb = load_image(key, image.width, image.height);
                src.textures.push_back(std::move(image));
                images[key] = int32_t(src.textures.size() - 1);
            }
            src_texture map;
            map.type = PT_MAP;
            map.image = images[key];
            map.color = diffuse;
            map.scale = v3(m.diffuse_opt.scale[0], m.diffuse_opt.scale[1], 0);
            map.offset = v3(m.diffuse_opt.offset[0], m.diffuse_opt.offset[1], 0);
            if(!finite3(map.scale) || !finite3(map.offset)) throw std::runtime_error("invalid MTL UV transform");
            map.clamp = m.diffuse_opt.clamp ? 1 : 0;
            src.textures.push_back(map);
            record.texture = int32_t(src.textures.size() - 1);
        }
        src.materials.push_back(record);
        if(m.dissolve != 1 || !m.alpha_texname.empty() || !m.bump_texname.empty() || !m.normal_texname.empty() ||
           !m.emissive_texname.empty() || m.metallic != 0 || m.roughness != 0 || m.shininess != 1 ||
           m.specular[0] != 0 || m.specular[1] != 0 || m.specular[2] != 0)
            out.warnings += "MTL '" + m.name + "': simplified material mapping; see README for unsupported shading properties.\n";
    }

    size_t degenerate = 0, missing_uv = 0, flat_normals = 0;
    for(const obj_shape& shape : data.shapes) {
        if(shape.lines_or_points) out.warnings += "Ignoring OBJ lines/points in '" + shape.name + "'.\n";
        for(const obj_triangle& tri : shape.triangles) {
            pt_vec3 p[3], n[3], uv[3];
            bool has_normals = true, has_uvs = true;
            for(int k = 0; k < 3; k++) {
                const vertex_index& vi = tri.idx[k];
                p[k] = attribute(data.v, vi.v, 3);
                if(vi.vn == -1) has_normals = false;
                else {
                    n[k] = attribute(data.vn, vi.vn, 3);
                    float len = length(n[k]);
                    if(!(len > 0) || !std::isfinite(len)) has_normals = false;
                    else {
                        float k_inv = float(1.0 / double(len));  // vec3::operator/= (vec3.h:152-158)
                        n[k] = v3(n[k].x * k_inv, n[k].y * k_inv, n[k].z * k_inv);
                    }
                }
                if(vi.vt == -1) has_uvs = false;
                else uv[k] = attribute(data.vt, vi.vt, 2);
            }
            pt_vec3 c = cross(p[1] - p[0], p[2] - p[0]);
            float area = c.x * c.x + c.y * c.y + c.z * c.z;
            if(!std::isfinite(area)) throw std::runtime_error("OBJ geometry exceeds supported float range");
            if(!(area > 0)) { degenerate++; continue; }
            if(tri.material < -1 || tri.material >= int(data.materials.size())) throw std::runtime_error("invalid OBJ material index");
            pt_flat_triangle f{};
            f.v0 = p[0]; f.v1 = p[1]; f.v2 = p[2];
            f.uv0 = v3(0, 0, 0); f.uv1 = v3(1, 0, 0); f.uv2 = v3(0, 1, 0);
            f.material_id = tri.material + 1;
            if(has_normals) { f.smooth = 1; f.n0 = n[0]; f.n1 = n[1]; f.n2 = n[2]; }
            else flat_normals++;
            if(has_uvs) { f.uv0 = uv[0]; f.uv1 = uv[1]; f.uv2 = uv[2]; }
            else missing_uv++;
            transform.apply(f);
            src.triangles.push_back(f);
        }
    }
    if(src.triangles.empty()) throw std::runtime_error("OBJ contains no usable triangles: " + filename);
    if(src.triangles.size() > size_t(INT32_MAX)) throw std::invalid_argument("too many triangles for the CPU BVH");
    if(degenerate) out.warnings += "Skipped " + std::to_string(degenerate) + " degenerate triangles.\n";
    if(flat_normals)
        out.warnings += "Using face normals for " + std::to_string(flat_normals) + " triangles without complete usable normals.\n";
    if(missing_uv) out.warnings += "Using canonical UVs for " + std::to_string(missing_uv) + " triangles without complete UVs.\n";
    out.mesh_triangles = int32_t(src.triangles.size());
}

} // namespace

namespace {

// pts_build_seeded in reference mode: one bvh_node over [triangles..., spheres...].
void build_reference_flat(const source_scene& src, pts_scene& out, int64_t seed) {
    reference_tree tree(seed);
    std::vector<int32_t> items;
    items.reserve(src.triangles.size() + src.spheres.size());
    for(size_t i = 0; i < src.triangles.size(); i++) items.push_back(tree.leaf(PT_TRIANGLE, int32_t(i), triangle_bounds(src.triangles[i])));
    for(size_t i = 0; i < src.spheres.size(); i++) items.push_back(tree.leaf(PT_SPHERE, int32_t(i), sphere_bounds(src.spheres[i])));
    if(items.empty()) {
        out.nodes.clear(); out.triangles.clear(); out.spheres.clear();
        out.materials.clear(); out.textures.clear(); out.pixels.clear();
        out.root = -1; out.max_depth = 0;
        return;
    }
    int32_t root = tree.node(items, 0, items.size());
    reference_export(tree, src, out).run(root);
}

source_scene source_from_flat(const pt_flat_triangle *triangles, int64_t nt, const pt_flat_sphere *spheres, int64_t ns,
                              const pt_flat_material *materials, int64_t nm, const pt_flat_texture *textures,
                              int64_t nx, const uint8_t *pixels, int64_t np) {
    source_scene src;
    src.triangles.assign(triangles, triangles + nt);
    src.spheres.assign(spheres, spheres + ns);
    for(int64_t i = 0; i < nx; i++) {
        const pt_flat_texture& t = textures[i];
        src_texture s;
        s.type = t.type; s.color = t.color; s.scale = t.scale; s.offset = t.offset; s.clamp = t.clamp;
        s.image = t.image_id;
        if(t.type == PT_IMAGE) {
            int64_t bytes = int64_t(t.width) * int64_t(t.height) * 3;
            if(t.width <= 0 || t.height <= 0 || t.byte_offset < 0 || t.byte_offset + bytes > np)
                throw std::runtime_error("image texture outside the pixel array");
            s.width = t.width; s.height = t.height;
            s.rgb.assign(pixels + t.byte_offset, pixels + t.byte_offset + bytes);
        }
        src.textures.push_back(std::move(s));
    }
    for(int64_t i = 0; i < nm; i++) {
        const pt_flat_material& m = materials[i];
        src_material s;
        s.type = m.type; s.albedo = m.albedo; s.fuzz = m.fuzz; s.ref_idx = m.ref_idx; s.texture = m.texture_id;
        src.materials.push_back(s);
    }
    return src;
}

} // namespace

extern "C" {

pts_scene *pts_build_seeded(const pt_flat_triangle *triangles, int64_t triangle_count,
                            const pt_flat_sphere *spheres, int64_t sphere_count,
                            const pt_flat_material *materials, int64_t material_count,
                            const pt_flat_texture *textures, int64_t texture_count,
                            const uint8_t *pixels, int64_t pixel_bytes, int32_t split, int64_t seed) {
    try {
        if(triangle_count < 0 || sphere_count < 0 || material_count < 0 || texture_count < 0 || pixel_bytes < 0)
            throw std::invalid_argument("negative array size");
        if(split != PTS_SPLIT_REFERENCE)
            return pts_build(triangles, triangle_count, spheres, sphere_count, materials, material_count, textures,
                             texture_count, pixels, pixel_bytes, split);
        source_scene src = source_from_flat(triangles, triangle_count, spheres, sphere_count, materials, material_count,
                                            textures, texture_count, pixels, pixel_bytes);
        auto s = std::make_unique<pts_scene>();
        build_reference_flat(src, *s, seed);
        return s.release();
    } catch(const std::exception& e) { set_error(e.what()); return nullptr; }
}

pts_scene *pts_load_obj(const char *filename, const pts_obj_options *options, pts_obj_info *info) {
    try {
        if(!filename || !options) throw std::invalid_argument("null argument");
        if(filename[0] == '\0') throw std::invalid_argument("OBJ filename must not be empty");
        mesh_transform transform(options->scale, options->rotate, options->translate);
        obj_import imported;
        import_obj(filename, transform, imported);
        source_scene& src = imported.source;

        // obj_scene (obj_scene.h:10-31): camera target, floor and light spheres.
        box3 bounds;
        bounds.empty();
        for(const auto& t : src.triangles) bounds.grow(triangle_bounds(t));
        pt_vec3 lo = v3(bounds.lo[0], bounds.lo[1], bounds.lo[2]), hi = v3(bounds.hi[0], bounds.hi[1], bounds.hi[2]);
        pt_vec3 lookat = 0.5f * lo + 0.5f * hi;
        float half_diag = 0.5f * length(hi - lo);
        float radius = 0.01f > half_diag ? 0.01f : half_diag;
        if(!std::isfinite(radius)) throw std::runtime_error("OBJ bounds exceed camera range");
        const float aspect = options->aspect;
        float half_angle = float(std::atan(std::tan(20 * M_PI / 180) * double(1 < aspect ? 1.0f : aspect)));
        float focus_dist = 1.2f * radius / std::sin(half_angle);
        pt_vec3 lookfrom = lookat + focus_dist * unit(v3(0.6f, 0.35f, 1.0f));
        src_texture floor_color, light_color;
        floor_color.color = v3(0.55f, 0.55f, 0.55f);
        light_color.color = v3(3, 3, 3);
        src.textures.push_back(floor_color);
        src_material floor_material;
        floor_material.texture = int32_t(src.textures.size() - 1);
        src.textures.push_back(light_color);
        src_material light_material;
        light_material.type = PT_LIGHT;
        light_material.texture = int32_t(src.textures.size() - 1);
        src.materials.push_back(floor_material);
        src.materials.push_back(light_material);
        pt_flat_sphere ground{v3(lookat.x, lo.y - 100 * radius, lookat.z), 100 * radius, int32_t(src.materials.size() - 2)};
        pt_flat_sphere light{lookat + radius * v3(-1, 3, 3), 1.5f * radius, int32_t(src.materials.size() - 1)};
        src.spheres.push_back(ground);
        src.spheres.push_back(light);

        auto s = std::make_unique<pts_scene>();
        if(options->split == PTS_SPLIT_REFERENCE) {
            // triangle_mesh's bvh_node over the faces, then obj_scene's over
            // {model, ground, light}; one drand48 stream for both.
            reference_tree tree(options->seed);
            std::vector<int32_t> faces;
            faces.reserve(src.triangles.size());
            for(size_t i = 0; i < src.triangles.size(); i++)
                faces.push_back(tree.leaf(PT_TRIANGLE, int32_t(i), triangle_bounds(src.triangles[i])));
            int32_t mesh = tree.node(faces, 0, faces.size());
            std::vector<int32_t> top = {mesh, tree.leaf(PT_SPHERE, 0, sphere_bounds(ground)),
                                        tree.leaf(PT_SPHERE, 1, sphere_bounds(light))};
            int32_t root = tree.node(top, 0, top.size());
            reference_export(tree, src, *s).run(root);
        } else {
            build_own(src, *s, options->split);
        }
        s->warnings = imported.warnings;
        if(info) {
            info->lookfrom = lookfrom; info->lookat = lookat; info->focus_dist = focus_dist;
            info->source_vertices = imported.source_vertices;
            info->shape_count = imported.shape_count;
            info->mesh_triangles = imported.mesh_triangles;
        }
        return s.release();
    } catch(const std::exception& e) { set_error(e.what()); return nullptr; }
}

}
