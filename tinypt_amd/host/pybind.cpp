// pybind.cpp — the `_C` Python module of the MI355X backend.  Same class and
// property names as li-plus/tinypt src/tinypt/pybind.cpp:67-181, plus what the
// reference never exposed (Glossy/GGX, AddBXDF, MixBXDF, PointLight radius,
// BumpTexture from a height map, meshes from arrays, render options).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstring>

#include "tinypt.h"

namespace py = pybind11;
using namespace tinypt;

namespace {

Mat3f to_mat3(py::array_t<float, py::array::c_style | py::array::forcecast> a) {
    if (a.ndim() != 2 || a.shape(0) != 3 || a.shape(1) != 3) throw std::invalid_argument("expect a 3x3 rotation matrix");
    Mat3f m;
    std::memcpy(m.data(), a.data(), 36);
    return m;
}
py::array_t<float> from_mat3(const Mat3f &m) {
    py::array_t<float> a({3, 3});
    std::memcpy(a.mutable_data(), m.data(), 36);
    return a;
}

// reference Image(array): HxWx3 sRGB-ish values -> pow(2.2), flipped so row 0 is the bottom (pybind.cpp:77-90)
std::shared_ptr<Image> image_from_display(py::array_t<float, py::array::c_style | py::array::forcecast> arr) {
    if (arr.size() == 0) return std::make_shared<Image>();
    if (!(arr.ndim() == 3 && arr.shape(2) == 3)) throw std::invalid_argument("Invalid image, expect an RGB image");
    int h = (int)arr.shape(0), w = (int)arr.shape(1);
    std::vector<float> buf((size_t)w * h * 3);
    const float *src = arr.data();
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w * 3; x++)
            buf[(size_t)(h - 1 - y) * w * 3 + x] = std::pow(src[(size_t)y * w * 3 + x], SRGB_GAMMA);
    return std::make_shared<Image>(Image::from_f32(w, h, 3, buf.data()));
}

// linear data, rows as given (row 0 = bottom unless flip), f32 or u8, HxW or HxWxC
std::shared_ptr<Image> image_from_linear(py::array arr, bool flip) {
    if (arr.ndim() != 2 && arr.ndim() != 3) throw std::invalid_argument("expect an HxW or HxWxC array");
    int h = (int)arr.shape(0), w = (int)arr.shape(1), c = arr.ndim() == 3 ? (int)arr.shape(2) : 1;
    size_t row = (size_t)w * c;
    if (arr.dtype().is(py::dtype::of<uint8_t>())) {
        auto a = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>::ensure(arr);
        std::vector<uint8_t> buf((size_t)h * row);
        for (int y = 0; y < h; y++) std::memcpy(&buf[(size_t)(flip ? h - 1 - y : y) * row], a.data() + (size_t)y * row, row);
        return std::make_shared<Image>(Image::from_u8(w, h, c, buf.data()));
    }
    auto a = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(arr);
    std::vector<float> buf((size_t)h * row);
    for (int y = 0; y < h; y++)
        std::memcpy(&buf[(size_t)(flip ? h - 1 - y : y) * row], a.data() + (size_t)y * row, row * 4);
    return std::make_shared<Image>(Image::from_f32(w, h, c, buf.data()));
}

py::array_t<float> image_raw(const Image &im) {
    py::array_t<float> a({im.height(), im.width(), im.channels()});
    float *d = a.mutable_data();
    for (int y = 0; y < im.height(); y++)
        for (int x = 0; x < im.width(); x++)
            for (int c = 0; c < im.channels(); c++) d[((size_t)y * im.width() + x) * im.channels() + c] = im.at(x, y, c);
    return a;
}

py::dict stats_dict(const tpt_render_stats &s) {
    py::dict d;
    d["render_ms"] = s.render_ms;
    d["extend_ms"] = s.extend_ms;
    d["shade_ms"] = s.shade_ms;
    d["shadow_ms"] = s.shadow_ms;
    d["paths"] = s.paths;
    d["extend_launches"] = s.extend_launches;
    d["nodes_visited"] = s.nodes_visited;
    d["tris_tested"] = s.tris_tested;
    d["rays_closest"] = s.rays_closest;
    d["rays_shadow"] = s.rays_shadow;
    return d;
}

py::array_t<float> rows_f32(const std::vector<float> &v, int cols) {
    py::array_t<float> a({(py::ssize_t)(v.size() / cols), (py::ssize_t)cols});
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * 4);
    return a;
}

std::shared_ptr<TriangleMesh> mesh_from_arrays(py::array_t<float, py::array::c_style | py::array::forcecast> vertices,
                                               py::array_t<int32_t, py::array::c_style | py::array::forcecast> faces,
                                               std::vector<Material> materials, py::object face_material,
                                               py::object normals, py::object texcoords, py::object shape_offsets) {
    auto m = std::make_shared<TriangleMesh>();
    if (vertices.ndim() != 2 || vertices.shape(1) != 3) throw std::invalid_argument("vertices must be (N, 3)");
    if (faces.ndim() != 2 || (faces.shape(1) != 3 && faces.shape(1) != 9)) throw std::invalid_argument("faces must be (F, 3) or (F, 9)");
    m->vertices.assign(vertices.data(), vertices.data() + vertices.size());
    bool has_n = !normals.is_none(), has_t = !texcoords.is_none();
    if (has_n) {
        auto n = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(normals);
        if (!n || n.ndim() != 2 || n.shape(1) != 3) throw std::invalid_argument("normals must be (N, 3)");
        m->normals.assign(n.data(), n.data() + n.size());
    }
    if (has_t) {
        auto t = py::array_t<float, py::array::c_style | py::array::forcecast>::ensure(texcoords);
        if (!t || t.ndim() != 2 || t.shape(1) != 2) throw std::invalid_argument("texcoords must be (N, 2)");
        m->texcoords.assign(t.data(), t.data() + t.size());
    }
    int F = (int)faces.shape(0);
    m->faces.resize((size_t)F * 9);
    const int32_t *f = faces.data();
    for (int i = 0; i < F; i++) {
        for (int c = 0; c < 3; c++) {
            if (faces.shape(1) == 9) {
                for (int q = 0; q < 3; q++) m->faces[(size_t)i * 9 + 3 * c + q] = f[(size_t)i * 9 + 3 * c + q];
            } else {
                int v = f[(size_t)i * 3 + c];
                m->faces[(size_t)i * 9 + 3 * c] = v;
                m->faces[(size_t)i * 9 + 3 * c + 1] = has_n ? v : -1;
                m->faces[(size_t)i * 9 + 3 * c + 2] = has_t ? v : -1;
            }
        }
    }
    if (materials.empty()) materials.push_back(Material(std::make_shared<Lambertian>(RGBTexture(Vec3f{1, 1, 1}))));
    m->materials = std::move(materials);
    if (face_material.is_none()) {
        m->face_material.assign(F, 0);
    } else {
        auto fm = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(face_material);
        if (!fm || fm.size() != F) throw std::invalid_argument("face_material must have one entry per face");
        m->face_material.assign(fm.data(), fm.data() + F);
    }
    if (!shape_offsets.is_none()) {
        auto so = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(shape_offsets);
        m->shape_offsets.assign(so.data(), so.data() + so.size());
    }
    return m;
}

} // namespace

PYBIND11_MODULE(_C, m) {
    m.doc() = "tinypt MI355X (device=hip) python binding";
    m.attr("ABI_VERSION") = tpt_hip_abi_version();
    m.def("hip_device_count", &tpt_hip_device_count);

    // integrator (pybind.cpp:71-73)
    py::class_<PathTracer>(m, "PathTracer")
        .def(py::init([](const std::string &d) { return std::make_unique<PathTracer>(Device(d)); }))
        .def(
            "render",
            [](const PathTracer &self, const Scene &scene, int num_samples, int max_depth, int rr_depth, bool shadow_bounded,
               uint64_t seed, bool hdr, bool fp64, int queue_capacity, int block_size, bool count_traffic) {
                RenderOptions o;
                o.max_depth = max_depth;
                o.rr_depth = rr_depth;
                o.shadow_bounded = shadow_bounded;
                o.seed = seed;
                o.hdr = hdr;
                o.fp64 = fp64;
                o.queue_capacity = queue_capacity;
                o.block_size = block_size;
                o.count_traffic = count_traffic;
                py::gil_scoped_release nogil;
                return self.render(scene, num_samples, o);
            },
            py::arg("scene"), py::arg("num_samples"), py::arg("max_depth") = 5, py::arg("rr_depth") = -1,
            py::arg("shadow_bounded") = false, py::arg("seed") = 0, py::arg("hdr") = false, py::arg("fp64") = false,
            py::arg("queue_capacity") = 0, py::arg("block_size") = 0, py::arg("count_traffic") = false)
        .def("last_stats", [](const PathTracer &self) { return stats_dict(self.last_stats()); })
        .def(
            "render_batch",
            [](const PathTracer &self, const std::vector<Scene> &scenes, const std::vector<int> &num_samples,
               const std::vector<uint64_t> &seeds, int max_depth, int rr_depth, bool shadow_bounded, bool hdr, bool fp64,
               int queue_capacity, int block_size) {
                RenderOptions o;
                o.max_depth = max_depth;
                o.rr_depth = rr_depth;
                o.shadow_bounded = shadow_bounded;
                o.hdr = hdr;
                o.fp64 = fp64;
                o.queue_capacity = queue_capacity;
                o.block_size = block_size;
                py::gil_scoped_release nogil;
                return self.render_batch(scenes, num_samples, seeds, o);
            },
            py::arg("scenes"), py::arg("num_samples"), py::arg("seeds"), py::arg("max_depth") = 5,
            py::arg("rr_depth") = -1, py::arg("shadow_bounded") = false, py::arg("hdr") = false, py::arg("fp64") = false,
            py::arg("queue_capacity") = 0, py::arg("block_size") = 0)
        .def("last_batch_stats", [](const PathTracer &self) {
            const tpt_batch_stats &s = self.last_batch_stats();
            py::dict d;
            d["render_ms"] = s.render_ms;
            d["tune_ms"] = s.tune_ms;
            d["paths"] = s.paths;
            d["queue_capacity"] = s.queue_capacity;
            d["block_size"] = s.block_size;
            d["waves"] = s.waves;
            d["graphs"] = s.graphs;
            d["scene_ms"] = self.last_batch_scene_ms();
            return d;
        });

    // image (pybind.cpp:76-100)
    py::class_<Image, std::shared_ptr<Image>>(m, "Image")
        .def(py::init(&image_from_display))
        .def_static("from_linear", &image_from_linear, py::arg("array"), py::arg("flip") = false)
        .def_property_readonly("width", &Image::width)
        .def_property_readonly("height", &Image::height)
        .def_property_readonly("channels", &Image::channels)
        .def("raw", &image_raw)
        .def_static("open", [](const std::string &path, bool grayscale) {
            RGBAImage r = Image::open(path, grayscale);
            return py::make_tuple(r.rgb ? py::cast(r.rgb) : py::none(), r.alpha ? py::cast(r.alpha) : py::none());
        }, py::arg("path"), py::arg("grayscale") = false,
           "(rgb, alpha) of an image file, as cpu::Image::open reads it: linear RGB, row 0 at the bottom")
        .def("save", &Image::save, py::arg("path"))
        .def("numpy", [](const Image &self) {
            py::array_t<float> a({self.height(), self.width(), self.channels()});
            float *d = a.mutable_data();
            for (int y = 0; y < self.height(); y++)
                for (int x = 0; x < self.width(); x++)
                    for (int c = 0; c < self.channels(); c++)
                        d[((size_t)(self.height() - 1 - y) * self.width() + x) * self.channels() + c] =
                            std::pow(self.at(x, y, c), 1.f / SRGB_GAMMA);
            return a;
        });

    py::class_<Camera>(m, "Camera")
        .def(py::init())
        .def_readwrite("location", &Camera::location)
        .def_property("rotation", [](const Camera &c) { return from_mat3(c.rotation); },
                      [](Camera &c, py::array_t<float> r) { c.rotation = to_mat3(r); })
        .def_readwrite("width", &Camera::width)
        .def_readwrite("height", &Camera::height)
        .def_readwrite("fov", &Camera::fov)
        .def_readwrite("aperture", &Camera::aperture)
        .def_readwrite("focus_dist", &Camera::focus_dist)
        .def("set_lens", &Camera::set_lens)
        .def("set_resolution", &Camera::set_resolution);

    py::class_<EnvTexture, std::shared_ptr<EnvTexture>>(m, "EnvTexture")
        .def(py::init<RGBTexture>())
        .def(py::init([](RGBTexture t, bool importance) { return std::make_shared<EnvTexture>(t, identity3(), importance); }),
             py::arg("texture"), py::arg("importance"))
        .def_property("rotation", [](const EnvTexture &e) { return from_mat3(e.rotation()); },
                      [](EnvTexture &e, py::array_t<float> r) { e.set_rotation(to_mat3(r)); })
        .def_property("importance", &EnvTexture::importance, &EnvTexture::set_importance);

    py::class_<Scene>(m, "Scene")
        .def(py::init<Camera, std::vector<std::shared_ptr<Object>>, std::vector<std::shared_ptr<Light>>,
                      std::shared_ptr<EnvTexture>>(),
             py::arg("camera"), py::arg("objects"), py::arg("lights") = std::vector<std::shared_ptr<Light>>{},
             py::arg("background") = nullptr)
        .def("to", [](const Scene &self, const std::string &d) { return self.to(Device(d)); })
        .def_property_readonly("device", [](const Scene &s) { return s.device().str(); })
        .def("_desc_ptr", [](const Scene &s) { return (uintptr_t)s.desc(); })
        .def("_hip_ptr", [](const Scene &s) { return (uintptr_t)s.hip_handle(); })
        .def("camera", [](const Scene &s) { return s.camera(); });

    // objects (pybind.cpp:115-135)
    py::class_<Object, std::shared_ptr<Object>>(m, "Object3d");
    py::class_<Sphere, Object, std::shared_ptr<Sphere>>(m, "Sphere")
        .def(py::init())
        .def_readwrite("location", &Sphere::location)
        .def_readwrite("radius", &Sphere::radius)
        .def_readwrite("material", &Sphere::material);
    py::class_<Rectangle, Object, std::shared_ptr<Rectangle>>(m, "Rectangle")
        .def(py::init())
        .def_readwrite("location", &Rectangle::location)
        .def_property("rotation", [](const Rectangle &o) { return from_mat3(o.rotation); },
                      [](Rectangle &o, py::array_t<float> r) { o.rotation = to_mat3(r); })
        .def_readwrite("dimension", &Rectangle::dimension)
        .def_readwrite("material", &Rectangle::material);
    py::class_<Circle, Object, std::shared_ptr<Circle>>(m, "Circle")
        .def(py::init())
        .def_readwrite("location", &Circle::location)
        .def_property("rotation", [](const Circle &o) { return from_mat3(o.rotation); },
                      [](Circle &o, py::array_t<float> r) { o.rotation = to_mat3(r); })
        .def_readwrite("radius", &Circle::radius)
        .def_readwrite("material", &Circle::material);
    py::class_<TriangleMesh, Object, std::shared_ptr<TriangleMesh>>(m, "TriangleMesh")
        .def(py::init())
        .def_static("from_obj", &TriangleMesh::from_obj, py::arg("path"), py::arg("material") = nullptr)
        .def_static("from_arrays", &mesh_from_arrays, py::arg("vertices"), py::arg("faces"),
                    py::arg("materials") = std::vector<Material>{}, py::arg("face_material") = py::none(),
                    py::arg("normals") = py::none(), py::arg("texcoords") = py::none(),
                    py::arg("shape_offsets") = py::none())
        .def_readwrite("location", &TriangleMesh::location)
        .def_property("rotation", [](const TriangleMesh &o) { return from_mat3(o.rotation); },
                      [](TriangleMesh &o, py::array_t<float> r) { o.rotation = to_mat3(r); })
        .def_property_readonly("num_faces", &TriangleMesh::num_faces)
        // read-only views of the mesh arrays (faces: F x 9 = (vertex, normal, texcoord) per corner, -1 when absent)
        .def_property_readonly("vertices", [](const TriangleMesh &t) { return rows_f32(t.vertices, 3); })
        .def_property_readonly("normals", [](const TriangleMesh &t) { return rows_f32(t.normals, 3); })
        .def_property_readonly("texcoords", [](const TriangleMesh &t) { return rows_f32(t.texcoords, 2); })
        .def_property_readonly("faces", [](const TriangleMesh &t) {
            py::array_t<int32_t> a({(py::ssize_t)t.num_faces(), (py::ssize_t)9});
            std::memcpy(a.mutable_data(), t.faces.data(), t.faces.size() * 4);
            return a;
        })
        .def_property_readonly("face_material", [](const TriangleMesh &t) { return py::array_t<int32_t>(t.face_material.size(), t.face_material.data()); })
        .def_property_readonly("shape_offsets", [](const TriangleMesh &t) { return py::array_t<int32_t>(t.shape_offsets.size(), t.shape_offsets.data()); })
        .def_property_readonly("materials", [](const TriangleMesh &t) { return t.materials; });

    // lights (pybind.cpp:138-152)
    py::class_<Light, std::shared_ptr<Light>>(m, "Light");
    py::class_<DistantLight, Light, std::shared_ptr<DistantLight>>(m, "DistantLight")
        .def(py::init())
        .def_property("rotation", [](const DistantLight &o) { return from_mat3(o.rotation); },
                      [](DistantLight &o, py::array_t<float> r) { o.rotation = to_mat3(r); })
        .def_readwrite("color", &DistantLight::color)
        .def_readwrite("power", &DistantLight::power)
        .def_property("angle", &DistantLight::angle, &DistantLight::set_angle);
    py::class_<PointLight, Light, std::shared_ptr<PointLight>>(m, "PointLight")
        .def(py::init())
        .def_readwrite("location", &PointLight::location)
        .def_readwrite("color", &PointLight::color)
        .def_readwrite("power", &PointLight::power)
        .def_readwrite("radius", &PointLight::radius);
    py::class_<ObjectLight, Light, std::shared_ptr<ObjectLight>>(m, "ObjectLight")
        .def(py::init<std::shared_ptr<Object>>(), py::arg("object"));
    py::class_<EnvLight, Light, std::shared_ptr<EnvLight>>(m, "EnvLight")
        .def(py::init())
        .def(py::init<std::shared_ptr<EnvTexture>>(), py::arg("env"));

    // bxdf (pybind.cpp:155-158)
    py::class_<BXDF, std::shared_ptr<BXDF>>(m, "BXDF");
    py::class_<Lambertian, BXDF, std::shared_ptr<Lambertian>>(m, "Lambertian")
        .def(py::init<RGBTexture>())
        .def_readonly("texture", &Lambertian::texture);
    py::class_<Metal, BXDF, std::shared_ptr<Metal>>(m, "Metal")
        .def(py::init<RGBTexture>())
        .def_readonly("texture", &Metal::texture);
    py::class_<Dielectric, BXDF, std::shared_ptr<Dielectric>>(m, "Dielectric")
        .def(py::init<RGBTexture, float>())
        .def_readonly("texture", &Dielectric::texture)
        .def_readonly("ior", &Dielectric::ior);
    py::class_<Glossy, BXDF, std::shared_ptr<Glossy>> glossy(m, "Glossy");
    py::enum_<Glossy::MicrofacetModelType>(glossy, "MicrofacetModelType")
        .value("MICRO_BLINN_PHONG", Glossy::MICRO_BLINN_PHONG)
        .value("MICRO_TROWBRIDGE_REITZ", Glossy::MICRO_TROWBRIDGE_REITZ)
        .export_values();
    glossy.def(py::init<float, RGBTexture, Glossy::MicrofacetModelType>(), py::arg("shininess"), py::arg("texture"),
               py::arg("micro_type") = Glossy::MICRO_BLINN_PHONG)
        .def_readonly("shininess", &Glossy::shininess)
        .def_readonly("texture", &Glossy::texture);
    py::class_<AddBXDF, BXDF, std::shared_ptr<AddBXDF>>(m, "AddBXDF")
        .def(py::init<std::shared_ptr<BXDF>, std::shared_ptr<BXDF>>())
        .def_readonly("bxdf1", &AddBXDF::bxdf1)
        .def_readonly("bxdf2", &AddBXDF::bxdf2);
    py::class_<MixBXDF, BXDF, std::shared_ptr<MixBXDF>>(m, "MixBXDF")
        .def(py::init<std::shared_ptr<BXDF>, std::shared_ptr<BXDF>, float>(), py::arg("bxdf1"), py::arg("bxdf2"),
             py::arg("weight") = 0.5f);

    // material (pybind.cpp:161-166)
    py::class_<Material, std::shared_ptr<Material>>(m, "Material")
        .def(py::init<std::shared_ptr<BXDF>>())
        .def_property_readonly("surface", &Material::surface)
        .def_property("emission_texture", &Material::emission_texture, &Material::set_emission_texture)
        .def_property("alpha_texture", &Material::alpha_texture, &Material::set_alpha_texture)
        .def_property("bump_texture", &Material::bump_texture, &Material::set_bump_texture);

    // textures (pybind.cpp:169-180)
    py::class_<RGBTexture, std::shared_ptr<RGBTexture>>(m, "RGBTexture")
        .def(py::init())
        .def(py::init<Vec3f>())
        .def(py::init<std::shared_ptr<Image>>())
        .def_property("value", &RGBTexture::value, &RGBTexture::set_value)
        .def_property("map", &RGBTexture::map, &RGBTexture::set_map);
    py::class_<AlphaTexture, std::shared_ptr<AlphaTexture>>(m, "AlphaTexture")
        .def(py::init())
        .def(py::init<float>())
        .def(py::init<std::shared_ptr<Image>, int>(), py::arg("map"), py::arg("channel") = 0)
        .def_property("value", &AlphaTexture::value, &AlphaTexture::set_value)
        .def_property("map", &AlphaTexture::map, &AlphaTexture::set_map);
    py::class_<BumpTexture, std::shared_ptr<BumpTexture>>(m, "BumpTexture")
        .def(py::init())
        .def(py::init([](const std::shared_ptr<Image> &hm, float scale) { return BumpTexture(*hm, scale); }),
             py::arg("height_map"), py::arg("scale") = 64.f)
        .def_property_readonly("map", &BumpTexture::map);

    m.def("rotation_from_euler", [](std::array<int, 3> axes, Vec3f angles, bool deg) {
        return from_mat3(rotation_from_euler(axes, angles, deg));
    }, py::arg("axes"), py::arg("angles"), py::arg("is_degrees") = false);
    m.def("rotation_from_direction", [](Vec3f w) { return from_mat3(rotation_from_direction(w)); });
    m.def("sobel_gradient", [](py::array_t<float, py::array::c_style | py::array::forcecast> h, float scale) {
        if (h.ndim() != 2) throw std::invalid_argument("height map must be 2-D");
        auto g = sobel_gradient(h.data(), (int)h.shape(1), (int)h.shape(0), scale);
        py::array_t<float> out({h.shape(0), h.shape(1), (py::ssize_t)2});
        std::memcpy(out.mutable_data(), g.data(), g.size() * 4);
        return out;
    }, py::arg("height"), py::arg("scale") = 64.f);
}
