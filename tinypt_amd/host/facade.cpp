// facade.cpp — flattening of the host object model into a tpt_scene_desc and
// the Device / Scene::to / PathTracer::render facade for device "hip".
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "tinypt.h"

namespace tinypt {

// ---------------------------------------------------------------- helpers
Mat3f rotation_from_euler(const std::array<int, 3> &axes, const Vec3f &angles, bool is_degrees) {
    for (int i = 0; i < 3; i++)
        TINYPT_CHECK_EX(0 <= axes[i] && axes[i] <= 2, std::invalid_argument,
                        "expect all axes to be between 0 and 2, got axes[" + std::to_string(i) + "] = " + std::to_string(axes[i]));
    double a[3];
    for (int i = 0; i < 3; i++) a[i] = is_degrees ? (double)radians(angles[i]) : (double)angles[i];
    auto R = [](int ax, double ang, double M[3][3]) {
        double c = std::cos(ang), s = std::sin(ang);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) M[i][j] = i == j;
        int i1 = (ax + 1) % 3, i2 = (ax + 2) % 3;
        M[i1][i1] = c;
        M[i1][i2] = -s;
        M[i2][i1] = s;
        M[i2][i2] = c;
    };
    double Z[3][3], Y[3][3], X[3][3], T[3][3];
    R(axes[2], a[2], Z);
    R(axes[1], a[1], Y);
    R(axes[0], a[0], X);
    Mat3f out;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T[i][j] = 0;
            for (int k = 0; k < 3; k++) T[i][j] += Z[i][k] * Y[k][j];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += T[i][k] * X[k][j];
            out[i * 3 + j] = (float)s;
        }
    return out;
}

Mat3f rotation_from_direction(const Vec3f &w) {
    auto nrm = [](Vec3f v) {
        float n = v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
        if (n > 0) {
            float s = std::sqrt(n);
            for (auto &x : v) x /= s;
        }
        return v;
    };
    Vec3f u = std::fabs(w[1]) < .9f ? nrm({w[2], 0.f, -w[0]}) : nrm({0.f, -w[2], w[1]});
    Vec3f v{w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
    return {u[0], v[0], w[0], u[1], v[1], w[1], u[2], v[2], w[2]};
}

Image Image::from_f32(int w, int h, int c, const float *data) {
    TINYPT_CHECK_EX(w > 0 && h > 0 && c > 0 && c <= 4, std::invalid_argument, "invalid image shape");
    Image im;
    im._w = w;
    im._h = h;
    im._c = c;
    im._f.assign(data, data + (size_t)w * h * c);
    return im;
}
Image Image::from_u8(int w, int h, int c, const uint8_t *data) {
    TINYPT_CHECK_EX(w > 0 && h > 0 && c > 0 && c <= 4, std::invalid_argument, "invalid image shape");
    Image im;
    im._w = w;
    im._h = h;
    im._c = c;
    im._u8.assign(data, data + (size_t)w * h * c);
    return im;
}

std::vector<float> sobel_gradient(const float *hm, int w, int h, float scale) {
    // cv::Sobel(ksize=3, scale/8, BORDER_REPLICATE) for d/dx and d/dy (texture.cpp:34-44)
    scale /= 8;
    std::vector<float> g((size_t)w * h * 2);
    auto at = [&](int x, int y) {
        x = std::min(std::max(x, 0), w - 1);
        y = std::min(std::max(y, 0), h - 1);
        return hm[(size_t)y * w + x];
    };
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float gx = (at(x + 1, y - 1) - at(x - 1, y - 1)) + 2.f * (at(x + 1, y) - at(x - 1, y)) +
                       (at(x + 1, y + 1) - at(x - 1, y + 1));
            float gy = (at(x - 1, y + 1) - at(x - 1, y - 1)) + 2.f * (at(x, y + 1) - at(x, y - 1)) +
                       (at(x + 1, y + 1) - at(x + 1, y - 1));
            g[((size_t)y * w + x) * 2] = gx * scale;
            g[((size_t)y * w + x) * 2 + 1] = gy * scale;
        }
    return g;
}

BumpTexture::BumpTexture(const Image &hm, float scale) {
    if (hm.empty()) return;
    TINYPT_CHECK_EX(hm.channels() == 1, std::runtime_error,
                    "Expect height map to be grayscale, got " + std::to_string(hm.channels()) + " channels");
    std::vector<float> h((size_t)hm.width() * hm.height());
    for (int y = 0; y < hm.height(); y++)
        for (int x = 0; x < hm.width(); x++) h[(size_t)y * hm.width() + x] = hm.at(x, y, 0);
    std::vector<float> g = sobel_gradient(h.data(), hm.width(), hm.height(), scale);
    _grad = std::make_shared<Image>(Image::from_f32(hm.width(), hm.height(), 2, g.data()));
}

// ---------------------------------------------------------------- flattening
struct FlatScene {
    tpt_scene_desc desc{};
    std::vector<tpt_object> objects;
    std::vector<tpt_mesh> meshes;
    std::vector<tpt_material> materials;
    std::vector<tpt_image> images;
    std::vector<tpt_delta_light> delta;
    std::vector<tpt_object_light> object_lights;
    std::vector<std::shared_ptr<Image>> image_refs;
    std::map<const Image *, int> image_index;
    struct MeshArrays {
        std::vector<float> v, n, t;
        std::vector<int32_t> f, fm, so;
    };
    std::vector<std::unique_ptr<MeshArrays>> mesh_arrays;

    int image_of(const std::shared_ptr<Image> &img) {
        auto it = image_index.find(img.get());
        if (it != image_index.end()) return it->second;
        tpt_image ti{};
        ti.width = img->width();
        ti.height = img->height();
        ti.channels = img->channels();
        ti.format = img->is_u8() ? TPT_IMG_U8 : TPT_IMG_F32;
        ti.data = img->is_u8() ? (const void *)img->u8().data() : (const void *)img->f32().data();
        int idx = (int)images.size();
        images.push_back(ti);
        image_refs.push_back(img);
        image_index[img.get()] = idx;
        return idx;
    }
    tpt_texture rgb(const RGBTexture &t) {
        tpt_texture r{};
        r.image = -1;
        r.value[0] = t.value()[0];
        r.value[1] = t.value()[1];
        r.value[2] = t.value()[2];
        if (t.is_map()) r.image = image_of(t.map());
        return r;
    }
    tpt_texture alpha(const AlphaTexture &t) {
        tpt_texture r{};
        r.image = -1;
        r.value[0] = t.value();
        if (t.is_map()) {
            r.image = image_of(t.map());
            r.channel = t.channel();
        }
        return r;
    }
    tpt_bxdf leaf(const BXDF *b) {
        tpt_bxdf r{};
        if (auto p = dynamic_cast<const Lambertian *>(b)) {
            r.type = TPT_BXDF_LAMBERTIAN;
            r.color = rgb(p->texture);
        } else if (auto p = dynamic_cast<const Metal *>(b)) {
            r.type = TPT_BXDF_METAL;
            r.color = rgb(p->texture);
        } else if (auto p = dynamic_cast<const Dielectric *>(b)) {
            r.type = TPT_BXDF_DIELECTRIC;
            r.param = p->ior;
            r.color = rgb(p->texture);
        } else if (auto p = dynamic_cast<const Glossy *>(b)) {
            r.type = p->model == Glossy::MICRO_BLINN_PHONG ? TPT_BXDF_GLOSSY_BLINN : TPT_BXDF_GLOSSY_GGX;
            r.param = p->shininess;
            r.color = rgb(p->texture);
        } else {
            TINYPT_THROW("Unsupported BXDF for device hip (nested AddBXDF/MixBXDF or a Python subclass)");
        }
        return r;
    }
    int material(const Material &m) {
        tpt_material r{};
        const BXDF *s = m.surface().get();
        r.weight = 0.5f;
        if (auto a = dynamic_cast<const AddBXDF *>(s)) {
            r.combine = TPT_COMBINE_ADD;
            r.lobe[0] = leaf(a->bxdf1.get());
            r.lobe[1] = leaf(a->bxdf2.get());
        } else if (auto x = dynamic_cast<const MixBXDF *>(s)) {
            r.combine = TPT_COMBINE_MIX;
            r.weight = x->weight;
            r.lobe[0] = leaf(x->bxdf1.get());
            r.lobe[1] = leaf(x->bxdf2.get());
        } else {
            r.combine = TPT_COMBINE_NONE;
            r.lobe[0] = leaf(s);
        }
        r.emission = rgb(m.emission_texture());
        r.alpha = alpha(m.alpha_texture());
        r.bump_image = m.bump_texture().empty() ? -1 : image_of(m.bump_texture().map());
        materials.push_back(r);
        return (int)materials.size() - 1;
    }
    static void rot9(const Mat3f &m, float *out) { // row-major matrix -> columns u, v, w
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) out[c * 3 + r] = m[r * 3 + c];
    }
    void mesh(const TriangleMesh &tm, tpt_object &o) {
        auto A = std::make_unique<MeshArrays>();
        size_t nv = tm.vertices.size() / 3, nn = tm.normals.size() / 3;
        TINYPT_CHECK_EX(tm.vertices.size() % 3 == 0 && tm.normals.size() % 3 == 0 && tm.texcoords.size() % 2 == 0,
                        std::invalid_argument, "invalid mesh array sizes");
        TINYPT_CHECK_EX(tm.faces.size() == tm.face_material.size() * 9, std::invalid_argument, "faces must hold 9 ints per face");
        const Mat3f &R = tm.rotation;
        const Vec3f &L = tm.location;
        bool ident = R == identity3() && L == Vec3f{0, 0, 0};
        A->v.resize(tm.vertices.size());
        for (size_t i = 0; i < nv; i++) {
            const float *p = &tm.vertices[3 * i];
            for (int r = 0; r < 3; r++)
                A->v[3 * i + r] = ident ? p[r] : (R[r * 3] * p[0] + (R[r * 3 + 1] * p[1] + R[r * 3 + 2] * p[2])) + L[r];
        }
        A->n.resize(tm.normals.size());
        for (size_t i = 0; i < nn; i++) {
            const float *p = &tm.normals[3 * i];
            for (int r = 0; r < 3; r++)
                A->n[3 * i + r] = ident ? p[r] : R[r * 3] * p[0] + (R[r * 3 + 1] * p[1] + R[r * 3 + 2] * p[2]);
        }
        A->t = tm.texcoords;
        A->f = tm.faces;
        int mat_base = (int)materials.size();
        for (const Material &m : tm.materials) material(m);
        A->fm.resize(tm.face_material.size());
        for (size_t i = 0; i < tm.face_material.size(); i++) {
            int fm = tm.face_material[i];
            TINYPT_CHECK_EX(fm >= 0 && fm < (int)tm.materials.size(), std::invalid_argument, "face material out of range");
            A->fm[i] = mat_base + fm;
        }
        A->so = tm.shape_offsets.empty() ? std::vector<int32_t>{0, tm.num_faces()} : tm.shape_offsets;
        tpt_mesh me{};
        me.num_vertices = (int32_t)nv;
        me.num_normals = (int32_t)nn;
        me.num_texcoords = (int32_t)(tm.texcoords.size() / 2);
        me.num_faces = tm.num_faces();
        me.num_shapes = (int32_t)A->so.size() - 1;
        me.vertices = A->v.data();
        me.normals = A->n.data();
        me.texcoords = A->t.data();
        me.faces = A->f.data();
        me.face_material = A->fm.data();
        me.shape_offsets = A->so.data();
        o.type = TPT_OBJ_MESH;
        o.mesh = (int32_t)meshes.size();
        meshes.push_back(me);
        mesh_arrays.push_back(std::move(A));
    }
};

static std::shared_ptr<FlatScene> flatten(const Camera &cam, const std::vector<std::shared_ptr<Object>> &objects,
                                          const std::vector<std::shared_ptr<Light>> &lights,
                                          const std::shared_ptr<EnvTexture> &bg) {
    auto F = std::make_shared<FlatScene>();
    for (const auto &obj : objects) {
        TINYPT_CHECK_EX(obj != nullptr, std::invalid_argument, "null object");
        tpt_object o{};
        o.material = -1;
        o.mesh = -1;
        FlatScene::rot9(identity3(), o.rot);
        if (auto s = dynamic_cast<const Sphere *>(obj.get())) {
            o.type = TPT_OBJ_SPHERE;
            std::memcpy(o.center, s->location.data(), 12);
            o.radius = s->radius;
            o.material = F->material(s->material);
        } else if (auto r = dynamic_cast<const Rectangle *>(obj.get())) {
            o.type = TPT_OBJ_RECTANGLE;
            std::memcpy(o.center, r->location.data(), 12);
            FlatScene::rot9(r->rotation, o.rot);
            o.dim[0] = r->dimension[0];
            o.dim[1] = r->dimension[1];
            o.material = F->material(r->material);
        } else if (auto c = dynamic_cast<const Circle *>(obj.get())) {
            o.type = TPT_OBJ_CIRCLE;
            std::memcpy(o.center, c->location.data(), 12);
            FlatScene::rot9(c->rotation, o.rot);
            o.radius = c->radius;
            o.material = F->material(c->material);
        } else if (auto m = dynamic_cast<const TriangleMesh *>(obj.get())) {
            F->mesh(*m, o);
        } else {
            TINYPT_THROW("Unsupported object for device hip");
        }
        F->objects.push_back(o);
    }
    for (const auto &l : lights) {
        TINYPT_CHECK_EX(l != nullptr, std::invalid_argument, "null light");
        // non-delta lights join the object-light mixture (cpu/scene.cpp:21-27)
        if (auto ol = dynamic_cast<const ObjectLight *>(l.get())) {
            auto it = std::find(objects.begin(), objects.end(), ol->object);
            TINYPT_CHECK_EX(ol->object != nullptr && it != objects.end(), std::invalid_argument,
                            "ObjectLight: its object must be one of the scene's objects");
            F->object_lights.push_back({TPT_OBJECT_LIGHT_OBJECT, (int32_t)(it - objects.begin())});
            continue;
        }
        if (auto el = dynamic_cast<const EnvLight *>(l.get())) {
            TINYPT_CHECK_EX(el->env != nullptr && el->env == bg, std::invalid_argument,
                            "EnvLight: its environment must be the scene's background");
            F->object_lights.push_back({TPT_OBJECT_LIGHT_ENV, -1});
            continue;
        }
        tpt_delta_light d{};
        if (auto dl = dynamic_cast<const DistantLight *>(l.get())) {
            d.type = TPT_LIGHT_DISTANT;
            FlatScene::rot9(dl->rotation, d.rot);
            std::memcpy(d.color, dl->color.data(), 12);
            d.power = dl->power;
            d.radius = dl->angle();
        } else if (auto pl = dynamic_cast<const PointLight *>(l.get())) {
            d.type = TPT_LIGHT_POINT;
            FlatScene::rot9(identity3(), d.rot);
            std::memcpy(d.center, pl->location.data(), 12);
            std::memcpy(d.color, pl->color.data(), 12);
            d.power = pl->power;
            d.radius = pl->radius;
        } else {
            TINYPT_THROW("Unsupported light for device hip");
        }
        F->delta.push_back(d);
    }
    tpt_scene_desc &D = F->desc;
    std::memcpy(D.camera.center, cam.location.data(), 12);
    FlatScene::rot9(cam.rotation, D.camera.rot);
    D.camera.width = cam.width;
    D.camera.height = cam.height;
    D.camera.fov = cam.fov;
    D.camera.aperture = cam.aperture;
    D.camera.focus_dist = cam.focus_dist;
    EnvTexture env = bg ? *bg : EnvTexture(RGBTexture(Vec3f{0, 0, 0}));
    D.background.rgb = F->rgb(env.rgb_texture());
    FlatScene::rot9(env.rotation(), D.background.rot);
    D.background.importance = env.importance() ? 1 : 0;
    D.num_objects = (int32_t)F->objects.size();
    D.num_meshes = (int32_t)F->meshes.size();
    D.num_materials = (int32_t)F->materials.size();
    D.num_images = (int32_t)F->images.size();
    D.num_delta_lights = (int32_t)F->delta.size();
    D.objects = F->objects.data();
    D.meshes = F->meshes.data();
    D.materials = F->materials.data();
    D.images = F->images.data();
    D.delta_lights = F->delta.data();
    D.num_object_lights = (int32_t)F->object_lights.size();
    D.object_lights = F->object_lights.data();
    return F;
}

// ---------------------------------------------------------------- Scene / PathTracer
Scene::Scene() : Scene(Camera{}, {}) {}

Scene::Scene(const Camera &camera, std::vector<std::shared_ptr<Object>> objects,
             std::vector<std::shared_ptr<Light>> lights, std::shared_ptr<EnvTexture> background)
    : _device(Device::DEVICE_CPU), _camera(camera), _flat(flatten(camera, objects, lights, background)) {}

const tpt_scene_desc *Scene::desc() const { return &_flat->desc; }

static void throw_code(int code, const char *file, int line) {
    std::string msg = tpt_hip_last_error();
    if (code == TPT_ERR_INVALID_ARGUMENT) throw_at<std::invalid_argument>(file, line, msg);
    throw_at<std::runtime_error>(file, line, msg);
}

Scene Scene::to(Device device) const { // scene.h:24-37
    if (_device == device) return *this;
    if (device.is_hip() && _device.is_cpu()) {
        if (tpt_hip_device_count() <= 0) TINYPT_THROW("hip is not available: no HIP device visible");
        tpt_hip_scene *h = nullptr;
        int rc = tpt_hip_scene_create(&_flat->desc, &h);
        if (rc != TPT_OK) throw_code(rc, __FILE__, __LINE__);
        Scene s = *this;
        s._device = Device::DEVICE_HIP;
        s._hip = std::shared_ptr<tpt_hip_scene>(h, tpt_hip_scene_destroy);
        return s;
    }
    if (device.is_cuda()) TINYPT_THROW("CUDA is not enabled at compile time");
    TINYPT_THROW("Not implemented");
}

PathTracer::PathTracer(Device d) : _device(d) { // path_tracer.h:14-25
    if (_device.is_cuda()) TINYPT_THROW("CUDA is not enabled at compile time");
    if (_device.is_cpu()) TINYPT_THROW("device cpu renders in the reference build; this library implements device hip");
}

Image PathTracer::render(const Scene &scene, int num_samples, const RenderOptions &o) const { // path_tracer.h:26-44
    TINYPT_CHECK_EX(scene.is_hip(), std::runtime_error, "Expect scene on hip, got " + scene.device().str() + " scene");
    TINYPT_CHECK_EX(num_samples > 0, std::invalid_argument, "num_samples must be positive");
    tpt_render_params p{};
    p.spp = num_samples;
    p.max_depth = o.max_depth;
    p.rr_depth = o.rr_depth;
    p.shadow_bounded = o.shadow_bounded ? 1 : 0;
    p.seed = o.seed;
    p.output = o.hdr ? TPT_OUT_HDR_MEAN : TPT_OUT_LDR_CLIPPED;
    p.accum_fp64 = o.fp64 ? 1 : 0;
    p.queue_capacity = o.queue_capacity;
    p.block_size = o.block_size;
    p.count_traffic = o.count_traffic ? 1 : 0;
    Image img(scene.camera().width, scene.camera().height, 3);
    int rc = tpt_hip_render(scene.hip_handle(), &p, img.f32().data(), &_stats);
    if (rc != TPT_OK) throw_code(rc, __FILE__, __LINE__);
    return img;
}

std::vector<Image> PathTracer::render_batch(const std::vector<Scene> &scenes, const std::vector<int> &num_samples,
                                            const std::vector<uint64_t> &seeds, const RenderOptions &o) const {
    TINYPT_CHECK_EX(!scenes.empty(), std::invalid_argument, "empty batch");
    TINYPT_CHECK_EX(num_samples.size() == scenes.size() && seeds.size() == scenes.size(), std::invalid_argument,
                    "expect one sample count and one seed per scene");
    std::vector<tpt_hip_scene *> handles;
    std::vector<tpt_batch_item> items;
    std::vector<Image> imgs;
    std::vector<float *> outs;
    imgs.reserve(scenes.size());
    for (size_t i = 0; i < scenes.size(); i++) {
        TINYPT_CHECK_EX(scenes[i].is_hip(), std::runtime_error,
                        "Expect scene on hip, got " + scenes[i].device().str() + " scene");
        TINYPT_CHECK_EX(num_samples[i] > 0, std::invalid_argument, "num_samples must be positive");
        handles.push_back(scenes[i].hip_handle());
        tpt_batch_item it{};
        it.spp = num_samples[i];
        it.seed = seeds[i];
        items.push_back(it);
        imgs.emplace_back(scenes[i].camera().width, scenes[i].camera().height, 3);
    }
    for (auto &im : imgs) outs.push_back(im.f32().data());
    tpt_render_params p{};
    p.max_depth = o.max_depth;
    p.rr_depth = o.rr_depth;
    p.shadow_bounded = o.shadow_bounded ? 1 : 0;
    p.output = o.hdr ? TPT_OUT_HDR_MEAN : TPT_OUT_LDR_CLIPPED;
    p.accum_fp64 = o.fp64 ? 1 : 0;
    p.queue_capacity = o.queue_capacity;
    p.block_size = o.block_size;
    _batch_scene_ms.assign(scenes.size(), 0.0);
    int rc = tpt_hip_render_batch(handles.data(), items.data(), (int32_t)handles.size(), &p, outs.data(), &_batch_stats,
                                  _batch_scene_ms.data());
    if (rc != TPT_OK) throw_code(rc, __FILE__, __LINE__);
    return imgs;
}

} // namespace tinypt
