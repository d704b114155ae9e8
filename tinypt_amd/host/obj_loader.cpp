// obj_loader.cpp — Wavefront OBJ / MTL into a TriangleMesh, without tinyobjloader.
//
// Geometry (the subset of the OBJ format meshes use): `v x y z`, `vn x y z`,
// `vt u [v]`, `f` with `v`, `v/t`, `v//n`, `v/t/n` corners (1-based, negative =
// relative to the end), `usemtl`, `mtllib`, and `o` / `g`, which start a new
// shape once the current one has faces.  Polygons are split into triangles the
// way tinyobjloader's default ("simple") triangulation does for quads: along the
// shorter diagonal; larger polygons are fanned from their first corner.  The
// submodule's commit is unrecorded, so this is not checked against it.
//
// Mapping into the scene (object.cpp:454-611): y-up vertices and normals become
// z-up (x, -z, y) (object.cpp:436); each shape is one group of faces; MTL
// materials map to BXDFs (`illum` 3/5 -> Metal(Ks), 7 -> Dielectric(1, Ni), else
// Lambertian(Kd / map_Kd) plus, with a specular colour and Ns >= EPS, Glossy(Ns)
// below Ns 1024 or Metal above, added); `Ke` / `map_Ke` is the emission,
// `map_d` / `d` the alpha (else the diffuse map's own alpha), `bump` /
// `map_bump` the bump map; faces without a material get the default material,
// appended last.
#include <algorithm>
#include <cctype>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>

#include "tinypt.h"

namespace fs = std::filesystem;

namespace tinypt {
namespace {

struct MtlRecord { // tinyobj material_t defaults
    std::string name;
    Vec3f diffuse{0, 0, 0}, specular{0, 0, 0}, emission{0, 0, 0};
    float shininess = 1, ior = 1, dissolve = 1;
    int illum = 0;
    std::string diffuse_tex, specular_tex, emissive_tex, alpha_tex, bump_tex;
};

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string lower(std::string s) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

Vec3f read_color(std::istringstream &in) {
    float r = 0, g, b;
    in >> r;
    if (!(in >> g)) return {r, r, r};
    if (!(in >> b)) b = g;
    return {r, g, b};
}

// Texture statement: options ("-bm 0.5", "-o u v w", ...) then the file name.
std::string texture_name(std::istringstream &in) {
    static const std::map<std::string, int> arity = {{"-blendu", 1}, {"-blendv", 1}, {"-boost", 1},  {"-mm", 2},
                                                     {"-o", 3},      {"-s", 3},      {"-t", 3},      {"-texres", 1},
                                                     {"-clamp", 1},  {"-bm", 1},     {"-imfchan", 1}, {"-type", 1},
                                                     {"-colorspace", 1}};
    std::vector<std::string> tok;
    for (std::string t; in >> t;) tok.push_back(t);
    size_t i = 0;
    while (i < tok.size() && tok[i].size() > 1 && tok[i][0] == '-') {
        auto it = arity.find(tok[i]);
        i += 1 + (it == arity.end() ? 0 : (size_t)it->second);
    }
    std::string name;
    for (; i < tok.size(); i++) name += (name.empty() ? "" : " ") + tok[i];
    return name;
}

void load_mtl(const fs::path &path, std::vector<MtlRecord> &mats, std::map<std::string, int> &by_name) {
    std::ifstream f(path);
    TINYPT_CHECK_EX(f.good(), std::runtime_error, "cannot open material library " + path.string());
    MtlRecord *m = nullptr;
    for (std::string line; std::getline(f, line);) {
        line = trim(line.substr(0, line.find('#')));
        if (line.empty()) continue;
        std::istringstream in(line);
        std::string key;
        in >> key;
        if (key == "newmtl") {
            std::string name = trim(line.substr(6));
            by_name[name] = (int)mats.size();
            mats.push_back(MtlRecord{});
            m = &mats.back();
            m->name = name;
            continue;
        }
        if (!m) continue;
        if (key == "Kd") m->diffuse = read_color(in);
        else if (key == "Ks") m->specular = read_color(in);
        else if (key == "Ke") m->emission = read_color(in);
        else if (key == "Ns") in >> m->shininess;
        else if (key == "Ni") in >> m->ior;
        else if (key == "d") in >> m->dissolve;
        else if (key == "Tr") {
            float tr = 0;
            in >> tr;
            m->dissolve = 1.f - tr;
        } else if (key == "illum") in >> m->illum;
        else if (key == "map_Kd") m->diffuse_tex = texture_name(in);
        else if (key == "map_Ks") m->specular_tex = texture_name(in);
        else if (key == "map_Ke") m->emissive_tex = texture_name(in);
        else if (key == "map_d") m->alpha_tex = texture_name(in);
        else if (key == "bump" || key == "map_bump" || key == "map_Bump" || key == "Bump") m->bump_tex = texture_name(in);
    }
}

// get_tex_path (object.cpp:417-433): the name relative to the OBJ directory,
// with '\' as '/'; else a case-insensitive match of the file name.
fs::path texture_path(const fs::path &dir, std::string name) {
    std::replace(name.begin(), name.end(), '\\', '/');
    fs::path p = dir / name;
    if (fs::is_regular_file(p)) return p;
    fs::path parent = p.parent_path();
    std::string want = lower(p.filename().string());
    if (fs::is_directory(parent))
        for (const auto &e : fs::directory_iterator(parent))
            if (lower(e.path().filename().string()) == want) return e.path();
    TINYPT_THROW("Cannot find file " + name + " in directory " + dir.string());
}

class ImageCache {
  public:
    explicit ImageCache(fs::path dir) : _dir(std::move(dir)) {}
    const RGBAImage &open(const std::string &name, bool grayscale) {
        std::string path = texture_path(_dir, name).string();
        auto key = std::make_pair(path, grayscale);
        auto it = _cache.find(key);
        if (it == _cache.end()) it = _cache.emplace(key, Image::open(path, grayscale)).first;
        return it->second;
    }

  private:
    fs::path _dir;
    std::map<std::pair<std::string, bool>, RGBAImage> _cache;
};

bool is_zero(const Vec3f &v) { return std::fabs(v[0]) <= 1e-5f && std::fabs(v[1]) <= 1e-5f && std::fabs(v[2]) <= 1e-5f; }

Material to_material(const MtlRecord &m, ImageCache &images) {
    RGBTexture emission = m.emissive_tex.empty() ? RGBTexture(m.emission)
                                                 : RGBTexture(images.open(m.emissive_tex, false).rgb);
    if (m.illum == 3 || m.illum == 5) return Material(std::make_shared<Metal>(RGBTexture(m.specular)), emission);
    if (m.illum == 7) return Material(std::make_shared<Dielectric>(RGBTexture(Vec3f{1, 1, 1}), m.ior), emission);
    AlphaTexture alpha;
    if (!m.alpha_tex.empty()) alpha = AlphaTexture(images.open(m.alpha_tex, true).alpha, 0);
    else if (std::fabs(m.dissolve - 1.f) >= 1e-3f) alpha = AlphaTexture(m.dissolve); // !is_close(d, 1)
    RGBTexture diffuse(m.diffuse);
    if (!m.diffuse_tex.empty()) {
        const RGBAImage &img = images.open(m.diffuse_tex, false);
        if (img.alpha && alpha.empty()) alpha = AlphaTexture(img.alpha, 0);
        diffuse = RGBTexture(img.rgb);
    }
    RGBTexture specular;
    if (!m.specular_tex.empty()) specular = RGBTexture(images.open(m.specular_tex, false).rgb);
    else if (!is_zero(m.specular)) specular = RGBTexture(m.specular);
    BumpTexture bump;
    if (!m.bump_tex.empty()) bump = BumpTexture(*images.open(m.bump_tex, true).alpha);
    if (m.shininess >= 1e-3f && !specular.empty()) {
        std::shared_ptr<BXDF> spec;
        if (m.shininess < 1024) spec = std::make_shared<Glossy>(m.shininess, specular);
        else spec = std::make_shared<Metal>(specular);
        return Material(std::make_shared<AddBXDF>(std::make_shared<Lambertian>(diffuse), spec), emission, alpha, bump);
    }
    return Material(std::make_shared<Lambertian>(diffuse), emission, alpha, bump);
}

struct Corner {
    int v, t, n;
};

// "v", "v/t", "v//n", "v/t/n" -> 0-based indices (-1 when absent)
Corner parse_corner(const std::string &tok, int nv, int nt, int nn, const std::string &where) {
    int idx[3] = {0, 0, 0};
    bool has[3] = {false, false, false};
    size_t p = 0;
    for (int k = 0; k < 3 && p <= tok.size(); k++) {
        size_t q = tok.find('/', p);
        std::string s = tok.substr(p, q == std::string::npos ? std::string::npos : q - p);
        if (!s.empty()) {
            idx[k] = std::stoi(s);
            has[k] = true;
        }
        if (q == std::string::npos) break;
        p = q + 1;
    }
    TINYPT_CHECK_EX(has[0], std::runtime_error, "face corner without a vertex index at " + where);
    auto fix = [&](int i, int count) {
        int r = i > 0 ? i - 1 : count + i;
        TINYPT_CHECK_EX(r >= 0 && r < count, std::runtime_error, "index out of range at " + where);
        return r;
    };
    return {fix(idx[0], nv), has[1] ? fix(idx[1], nt) : -1, has[2] ? fix(idx[2], nn) : -1};
}

} // namespace

std::shared_ptr<TriangleMesh> TriangleMesh::from_obj(const std::string &path, const std::shared_ptr<Material> &material) {
    std::ifstream f(path);
    TINYPT_CHECK_EX(f.good(), std::runtime_error, "Failed to parse obj file [" + path + "]");
    const fs::path dir = fs::path(path).parent_path();
    auto mesh = std::make_shared<TriangleMesh>();
    std::vector<float> raw_v, raw_n; // y-up, as read
    std::vector<MtlRecord> mtls;
    std::map<std::string, int> by_name;
    int cur_mat = -1;
    mesh->shape_offsets.push_back(0);
    int line_no = 0;
    for (std::string line; std::getline(f, line);) {
        line_no++;
        line = trim(line.substr(0, line.find('#')));
        if (line.empty()) continue;
        std::istringstream in(line);
        std::string key;
        in >> key;
        const std::string where = path + ":" + std::to_string(line_no);
        if (key == "v" || key == "vn") {
            float x = 0, y = 0, z = 0;
            TINYPT_CHECK_EX((bool)(in >> x >> y >> z), std::runtime_error, "bad " + key + " at " + where);
            auto &dst = key == "v" ? raw_v : raw_n;
            dst.insert(dst.end(), {x, y, z});
        } else if (key == "vt") {
            float u = 0, v = 0;
            TINYPT_CHECK_EX((bool)(in >> u), std::runtime_error, "bad vt at " + where);
            in >> v;
            mesh->texcoords.insert(mesh->texcoords.end(), {u, v});
        } else if (key == "f") {
            std::vector<Corner> poly;
            for (std::string tok; in >> tok;)
                poly.push_back(parse_corner(tok, (int)raw_v.size() / 3, (int)mesh->texcoords.size() / 2,
                                            (int)raw_n.size() / 3, where));
            TINYPT_CHECK_EX(poly.size() >= 3, std::runtime_error, "face with fewer than 3 corners at " + where);
            auto emit = [&](int a, int b, int c) {
                for (int k : {a, b, c}) mesh->faces.insert(mesh->faces.end(), {poly[k].v, poly[k].n, poly[k].t});
                mesh->face_material.push_back(cur_mat);
            };
            if (poly.size() == 4) {
                auto d2 = [&](int a, int b) {
                    float s = 0;
                    for (int c = 0; c < 3; c++) {
                        float e = raw_v[3 * poly[b].v + c] - raw_v[3 * poly[a].v + c];
                        s += e * e;
                    }
                    return s;
                };
                if (d2(0, 2) < d2(1, 3)) {
                    emit(0, 1, 2);
                    emit(0, 2, 3);
                } else {
                    emit(0, 1, 3);
                    emit(1, 2, 3);
                }
            } else {
                for (size_t k = 1; k + 1 < poly.size(); k++) emit(0, (int)k, (int)k + 1);
            }
        } else if (key == "usemtl") {
            auto it = by_name.find(trim(line.substr(6)));
            cur_mat = it == by_name.end() ? -1 : it->second;
        } else if (key == "mtllib") {
            for (std::string name; in >> name;) load_mtl(dir / name, mtls, by_name);
        } else if (key == "o" || key == "g") {
            if ((int)mesh->face_material.size() > mesh->shape_offsets.back())
                mesh->shape_offsets.push_back((int)mesh->face_material.size());
        }
    }
    if ((int)mesh->face_material.size() > mesh->shape_offsets.back())
        mesh->shape_offsets.push_back((int)mesh->face_material.size());
    TINYPT_CHECK_EX(!mesh->face_material.empty(), std::runtime_error, "no faces in obj file [" + path + "]");
    ImageCache images(dir);
    for (const auto &m : mtls) mesh->materials.push_back(to_material(m, images));
    mesh->materials.push_back(material ? *material
                                       : Material(std::make_shared<Lambertian>(RGBTexture(Vec3f{1, 1, 1}))));
    const int dflt = (int)mesh->materials.size() - 1;
    for (auto &id : mesh->face_material)
        if (id < 0) id = dflt;
    // yup2zup (object.cpp:436): (x, y, z) -> (x, -z, y)
    auto zup = [](const std::vector<float> &src, std::vector<float> &dst) {
        dst.resize(src.size());
        for (size_t i = 0; i < src.size(); i += 3) {
            dst[i] = src[i];
            dst[i + 1] = -src[i + 2];
            dst[i + 2] = src[i + 1];
        }
    };
    zup(raw_v, mesh->vertices);
    zup(raw_n, mesh->normals);
    return mesh;
}

} // namespace tinypt
