// tinypt.h — host facade of the MI355X backend, mirroring tinypt's public C++
// API (li-plus/tinypt src/tinypt/{defs,scene,path_tracer}.h and the cpu::
// object model) so that `device = "hip"` is a drop-in beside "cpu" and "cuda".
//
// The objects here are scene DESCRIPTIONS: they carry the same attributes as
// the reference's cpu:: classes and are flattened into a tpt_scene_desc, which
// the C ABI (include/tinypt_hip.h) turns into a device arena.  Rendering for
// device "hip" goes through libtinypt_hip.so; "cpu" and "cuda" are recognised
// device strings whose renderers live in the reference build.
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tinypt_hip.h"

namespace tinypt {

// ---------------------------------------------------------------- errors (cpu/defs.h:43-62)
template <typename E>
[[noreturn]] inline void throw_at(const char *file, int line, const std::string &msg) {
    std::ostringstream os;
    os << file << ':' << line << ' ' << msg;
    throw E(os.str());
}
#define TINYPT_THROW_EX(E, msg) ::tinypt::throw_at<E>(__FILE__, __LINE__, (msg))
#define TINYPT_THROW(msg) TINYPT_THROW_EX(std::runtime_error, msg)
#define TINYPT_CHECK_EX(cond, E, msg)                                                                                  \
    do {                                                                                                               \
        if (!(cond)) TINYPT_THROW_EX(E, std::string("Check failed: " #cond " ") + (msg));                              \
    } while (0)

using Vec2f = std::array<float, 2>;
using Vec3f = std::array<float, 3>;
// 3x3 rotation, row-major storage m[r*3+c]; its columns are the frame u, v, w (cpu/mapping.h:14-41)
using Mat3f = std::array<float, 9>;

static constexpr float INF = 1e20f;
static constexpr float EPS = 1e-3f;
static constexpr float SRGB_GAMMA = 2.2f;

inline Mat3f identity3() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }
inline float radians(float deg) { return deg * (3.14159265358979323846f / 180.f); }
inline float degrees(float rad) { return rad * (180.f / 3.14159265358979323846f); }
// Rotation::from_euler (cpu/mapping.cpp:32-48): Rz * Ry * Rx about the given axes
Mat3f rotation_from_euler(const std::array<int, 3> &axes, const Vec3f &angles, bool is_degrees = false);
// Rotation::from_direction (cpu/mapping.cpp:20-30)
Mat3f rotation_from_direction(const Vec3f &w);

// ---------------------------------------------------------------- Device (defs.h:62-86)
class Device {
  public:
    enum DeviceType { DEVICE_CPU, DEVICE_CUDA, DEVICE_HIP };
    Device() : _device(DEVICE_CPU) {}
    Device(DeviceType d) : _device(d) {}
    Device(const std::string &s) {
        if (s == "cpu") _device = DEVICE_CPU;
        else if (s == "cuda") _device = DEVICE_CUDA;
        else if (s == "hip") _device = DEVICE_HIP;
        else TINYPT_THROW_EX(std::invalid_argument, "Unknown device " + s);
    }
    Device(const char *s) : Device(std::string(s)) {}
    bool operator==(const Device &o) const { return _device == o._device; }
    bool operator!=(const Device &o) const { return !(*this == o); }
    bool is_cpu() const { return _device == DEVICE_CPU; }
    bool is_cuda() const { return _device == DEVICE_CUDA; }
    bool is_hip() const { return _device == DEVICE_HIP; }
    std::string str() const { return is_cpu() ? "cpu" : (is_cuda() ? "cuda" : "hip"); }

  private:
    DeviceType _device;
};

// ---------------------------------------------------------------- Image (cpu/image.h:13-94)
// Row-major, row 0 at the bottom (y up), like cpu::Image.  F32 or U8 texels.
struct RGBAImage;
class Image {
  public:
    Image() = default;
    Image(int width, int height, int channels) : _w(width), _h(height), _c(channels), _f((size_t)width * height * channels, 0.f) {}
    static Image from_f32(int w, int h, int c, const float *data);
    static Image from_u8(int w, int h, int c, const uint8_t *data);
    // Image files (image_io.cpp), with the conventions of cpu::Image::open / save
    // (cpu/image.cpp:6-72): linear RGB (pow 2.2) and alpha, row 0 at the bottom.
    static RGBAImage open(const std::string &path, bool grayscale = false);
    void save(const std::string &path) const;
    bool empty() const { return _w == 0 || _h == 0; }
    int width() const { return _w; }
    int height() const { return _h; }
    int channels() const { return _c; }
    bool is_u8() const { return !_u8.empty(); }
    const std::vector<float> &f32() const { return _f; }
    std::vector<float> &f32() { return _f; }
    const std::vector<uint8_t> &u8() const { return _u8; }
    float at(int x, int y, int c) const {
        size_t i = ((size_t)y * _w + x) * _c + c;
        return is_u8() ? (float)_u8[i] / 255.f : _f[i];
    }

  private:
    int _w = 0, _h = 0, _c = 0;
    std::vector<float> _f;
    std::vector<uint8_t> _u8;
};

// rgb: 3 channels (or null), alpha: 1 channel (or null) (cpu/image.h:95-120)
struct RGBAImage {
    std::shared_ptr<Image> rgb, alpha;
};

// ---------------------------------------------------------------- textures (cpu/texture.h)
class RGBTexture {
  public:
    RGBTexture() : _value{0, 0, 0} {}
    RGBTexture(const Vec3f &v) : _value(v) {}
    RGBTexture(std::shared_ptr<Image> map) : _value{0, 0, 0}, _map(std::move(map)) {
        TINYPT_CHECK_EX(_map && (_map->channels() == 3 || _map->channels() == 4), std::runtime_error, "invalid rgb texture map");
    }
    bool empty() const { return !is_map() && std::fabs(_value[0]) <= 1e-5f && std::fabs(_value[1]) <= 1e-5f && std::fabs(_value[2]) <= 1e-5f; }
    bool is_value() const { return !is_map(); }
    bool is_map() const { return _map && !_map->empty(); }
    const Vec3f &value() const { return _value; }
    void set_value(const Vec3f &v) { _value = v; }
    const std::shared_ptr<Image> &map() const { return _map; }
    void set_map(std::shared_ptr<Image> m) { _map = std::move(m); }

  private:
    Vec3f _value;
    std::shared_ptr<Image> _map;
};

class AlphaTexture {
  public:
    AlphaTexture() : _value(1) {}
    AlphaTexture(float v) : _value(v) {}
    // channel: which image channel holds alpha (0 for a grayscale map, 3 for RGBA)
    AlphaTexture(std::shared_ptr<Image> map, int channel = 0) : _value(1), _map(std::move(map)), _channel(channel) {
        if (_map && !_map->empty())
            TINYPT_CHECK_EX(channel >= 0 && channel < _map->channels(), std::invalid_argument,
                            "expect alpha channel inside the map, got " + std::to_string(_map->channels()) + " channels");
    }
    bool empty() const { return !is_map() && _value == 1; }
    bool is_map() const { return _map && !_map->empty(); }
    float value() const { return _value; }
    void set_value(float v) { _value = v; }
    const std::shared_ptr<Image> &map() const { return _map; }
    void set_map(std::shared_ptr<Image> m) { _map = std::move(m); _channel = 0; }
    int channel() const { return _channel; }

  private:
    float _value;
    std::shared_ptr<Image> _map;
    int _channel = 0;
};

// BumpTexture (texture.cpp:34-44): the 3x3 Sobel gradient of a height map,
// scale/8, replicated border, stored as a 2-channel image.
class BumpTexture {
  public:
    BumpTexture() = default;
    BumpTexture(const Image &height_map, float scale = 64);
    bool empty() const { return !_grad || _grad->empty(); }
    const std::shared_ptr<Image> &map() const { return _grad; }

  private:
    std::shared_ptr<Image> _grad;
};

class EnvTexture {
  public:
    EnvTexture() : _rot(identity3()) {}
    EnvTexture(RGBTexture env, const Mat3f &rotation = identity3(), bool importance = false)
        : _env(std::move(env)), _rot(rotation), _importance(importance) {}
    const RGBTexture &rgb_texture() const { return _env; }
    void set_rgb_texture(RGBTexture t) { _env = std::move(t); }
    const Mat3f &rotation() const { return _rot; }
    void set_rotation(const Mat3f &r) { _rot = r; }
    // false: the reference's effective behaviour (no CDF, uniform uv, pdf 1/(4 pi));
    // true: luminance (max-channel) CDF importance sampling
    bool importance() const { return _importance; }
    void set_importance(bool b) { _importance = b; }

  private:
    RGBTexture _env;
    Mat3f _rot;
    bool _importance = false;
};

// ---------------------------------------------------------------- BXDFs (cpu/material.h)
class BXDF {
  public:
    virtual ~BXDF() = default;
    virtual bool is_specular() const = 0;
};
class Lambertian : public BXDF {
  public:
    Lambertian() = default;
    explicit Lambertian(RGBTexture t) : texture(std::move(t)) {}
    bool is_specular() const override { return false; }
    RGBTexture texture;
};
class Metal : public BXDF {
  public:
    explicit Metal(RGBTexture t) : texture(std::move(t)) {}
    bool is_specular() const override { return true; }
    RGBTexture texture;
};
class Dielectric : public BXDF {
  public:
    Dielectric(RGBTexture t, float ior_) : texture(std::move(t)), ior(ior_) {}
    bool is_specular() const override { return true; }
    RGBTexture texture;
    float ior;
};
class Glossy : public BXDF {
  public:
    enum MicrofacetModelType { MICRO_BLINN_PHONG, MICRO_TROWBRIDGE_REITZ };
    Glossy(float shininess_, RGBTexture t, MicrofacetModelType m = MICRO_BLINN_PHONG)
        : shininess(shininess_), texture(std::move(t)), model(m) {}
    bool is_specular() const override { return false; }
    float shininess;
    RGBTexture texture;
    MicrofacetModelType model;
};
class AddBXDF : public BXDF {
  public:
    AddBXDF(std::shared_ptr<BXDF> a, std::shared_ptr<BXDF> b) : bxdf1(std::move(a)), bxdf2(std::move(b)) {
        TINYPT_CHECK_EX(bxdf1 && bxdf2, std::invalid_argument, "null BXDF");
    }
    bool is_specular() const override { return bxdf1->is_specular() || bxdf2->is_specular(); }
    std::shared_ptr<BXDF> bxdf1, bxdf2;
};
class MixBXDF : public BXDF {
  public:
    MixBXDF(std::shared_ptr<BXDF> a, std::shared_ptr<BXDF> b, float w = 0.5f)
        : bxdf1(std::move(a)), bxdf2(std::move(b)), weight(w) {
        TINYPT_CHECK_EX(bxdf1 && bxdf2, std::invalid_argument, "null BXDF");
        TINYPT_CHECK_EX(0 <= w && w <= 1, std::invalid_argument, "[MixBXDF] invalid weight " + std::to_string(w));
    }
    bool is_specular() const override { return bxdf1->is_specular() || bxdf2->is_specular(); }
    std::shared_ptr<BXDF> bxdf1, bxdf2;
    float weight;
};

class Material { // material.h:194-216
  public:
    Material() : _surface(std::make_shared<Lambertian>()) {}
    explicit Material(std::shared_ptr<BXDF> s, RGBTexture e = {}, AlphaTexture a = {}, BumpTexture b = {})
        : _surface(std::move(s)), _emission(std::move(e)), _alpha(std::move(a)), _bump(std::move(b)) {
        TINYPT_CHECK_EX(_surface, std::invalid_argument, "null surface");
    }
    const std::shared_ptr<BXDF> &surface() const { return _surface; }
    void set_surface(std::shared_ptr<BXDF> s) { _surface = std::move(s); }
    const RGBTexture &emission_texture() const { return _emission; }
    void set_emission_texture(RGBTexture t) { _emission = std::move(t); }
    const AlphaTexture &alpha_texture() const { return _alpha; }
    void set_alpha_texture(AlphaTexture t) { _alpha = std::move(t); }
    const BumpTexture &bump_texture() const { return _bump; }
    void set_bump_texture(BumpTexture t) { _bump = std::move(t); }

  private:
    std::shared_ptr<BXDF> _surface;
    RGBTexture _emission;
    AlphaTexture _alpha;
    BumpTexture _bump;
};

// ---------------------------------------------------------------- objects (cpu/object.h)
class Object {
  public:
    virtual ~Object() = default;
    virtual int kind() const = 0; // TPT_OBJ_*
};
class Sphere : public Object {
  public:
    int kind() const override { return TPT_OBJ_SPHERE; }
    Vec3f location{0, 0, 0};
    float radius = 1;
    Material material;
};
class Rectangle : public Object {
  public:
    int kind() const override { return TPT_OBJ_RECTANGLE; }
    Vec3f location{0, 0, 0};
    Mat3f rotation = identity3();
    Vec2f dimension{1, 1};
    Material material;
};
class Circle : public Object {
  public:
    int kind() const override { return TPT_OBJ_CIRCLE; }
    Vec3f location{0, 0, 0};
    Mat3f rotation = identity3();
    float radius = 1;
    Material material;
};
// TriangleMesh: vertices / normals / texcoords with per-corner (v, n, t) face
// indices and one material per face, as TriangleMesh::from_obj fills them
// (object.cpp:454-611).  Location / rotation (the ObjectGroup Mapping) are
// baked into world-space vertices when the scene is flattened.
class TriangleMesh : public Object {
  public:
    int kind() const override { return TPT_OBJ_MESH; }
    Vec3f location{0, 0, 0};
    Mat3f rotation = identity3();
    std::vector<float> vertices, normals, texcoords; // xyz / xyz / uv
    std::vector<int32_t> faces;                       // 9 per face
    std::vector<int32_t> face_material;               // index into materials
    std::vector<int32_t> shape_offsets;               // num_shapes + 1
    std::vector<Material> materials;
    int num_faces() const { return (int)face_material.size(); }
    // Wavefront OBJ + MTL (obj_loader.cpp), as TriangleMesh::from_obj maps them
    // (object.cpp:454-611); `material` (or a white Lambertian) serves faces
    // without a material.
    static std::shared_ptr<TriangleMesh> from_obj(const std::string &path,
                                                  const std::shared_ptr<Material> &material = nullptr);
};

// ---------------------------------------------------------------- lights (cpu/light.h)
class Light {
  public:
    virtual ~Light() = default;
    virtual bool is_delta() const = 0;
};
class DistantLight : public Light {
  public:
    bool is_delta() const override { return true; }
    Mat3f rotation = identity3();
    Vec3f color{1, 1, 1};
    float power = 1;
    float angle() const { return _angle; }
    void set_angle(float a) {
        TINYPT_CHECK_EX(0 <= a && a <= 3.14159265358979323846f, std::invalid_argument, "angle out of range: " + std::to_string(a));
        _angle = a;
    }

  private:
    float _angle = 0;
};
class PointLight : public Light {
  public:
    bool is_delta() const override { return true; }
    Vec3f location{0, 0, 0};
    Vec3f color{1, 1, 1};
    float power = 1;
    float radius = 0;
};

// ObjectLight (cpu/light.h:30-43): importance sampling towards an object; passed
// in Scene()'s lights, it joins the object-light mixture after the scene's
// emissive objects.  Its object must be one of the scene's objects.
class ObjectLight : public Light {
  public:
    ObjectLight() = default;
    explicit ObjectLight(std::shared_ptr<Object> obj) : object(std::move(obj)) {}
    bool is_delta() const override { return false; }
    std::shared_ptr<Object> object;
};
// EnvLight (cpu/light.h:113-124): importance sampling of the environment; passed
// in Scene()'s lights, it joins the mixture as ObjectLight does.  Its environment
// must be the scene's background.
class EnvLight : public Light {
  public:
    EnvLight() = default;
    explicit EnvLight(std::shared_ptr<EnvTexture> e) : env(std::move(e)) {}
    bool is_delta() const override { return false; }
    std::shared_ptr<EnvTexture> env;
};

// ---------------------------------------------------------------- camera (cpu/camera.h)
class Camera {
  public:
    Vec3f location{0, 0, 0};
    Mat3f rotation = identity3();
    int width = 1920, height = 1080;
    float fov = 45;
    float aperture = 0, focus_dist = 1;
    void set_resolution(int w, int h) {
        width = w;
        height = h;
    }
    void set_lens(float a, float f) {
        aperture = a;
        focus_dist = f;
    }
};

// ---------------------------------------------------------------- flattened description
struct FlatScene; // owns a tpt_scene_desc and every array it points at

struct RenderOptions {
    int max_depth = 5;         // cpu/path_tracer.cpp:103
    int rr_depth = -1;         // no Russian roulette (reference)
    bool shadow_bounded = false;
    uint64_t seed = 0;
    bool hdr = false;
    bool fp64 = false;
    int queue_capacity = 0;
    int block_size = 0;
    bool count_traffic = false;
};

// ---------------------------------------------------------------- Scene (scene.h:12-52)
class Scene {
    friend class PathTracer;

  public:
    Scene();
    Scene(const Camera &camera, std::vector<std::shared_ptr<Object>> objects,
          std::vector<std::shared_ptr<Light>> lights = {}, std::shared_ptr<EnvTexture> background = nullptr);
    bool is_cpu() const { return _device.is_cpu(); }
    bool is_cuda() const { return _device.is_cuda(); }
    bool is_hip() const { return _device.is_hip(); }
    const Device &device() const { return _device; }
    Scene to(Device device) const;
    const tpt_scene_desc *desc() const; // the flat description handed to the C ABI
    tpt_hip_scene *hip_handle() const { return _hip.get(); }
    const Camera &camera() const { return _camera; }

  private:
    Device _device;
    Camera _camera;
    std::shared_ptr<FlatScene> _flat;
    std::shared_ptr<tpt_hip_scene> _hip;
};

// ---------------------------------------------------------------- PathTracer (path_tracer.h:12-51)
class PathTracer {
  public:
    PathTracer(Device device = {});
    Image render(const Scene &scene, int num_samples) const { return render(scene, num_samples, RenderOptions{}); }
    Image render(const Scene &scene, int num_samples, const RenderOptions &opt) const;
    const tpt_render_stats &last_stats() const { return _stats; }
    // Many independent renders through one shared pipeline (no counterpart in the
    // reference): scene i with num_samples[i] and seeds[i]; `opt` gives the shared
    // depth / RR / output settings and the queue capacity and workgroup size (0 =
    // autotuned).  Each image equals render(scenes[i], num_samples[i]) with that seed.
    std::vector<Image> render_batch(const std::vector<Scene> &scenes, const std::vector<int> &num_samples,
                                    const std::vector<uint64_t> &seeds, const RenderOptions &opt) const;
    const tpt_batch_stats &last_batch_stats() const { return _batch_stats; }
    const std::vector<double> &last_batch_scene_ms() const { return _batch_scene_ms; }

  private:
    Device _device;
    mutable tpt_render_stats _stats{};
    mutable tpt_batch_stats _batch_stats{};
    mutable std::vector<double> _batch_scene_ms;
};

// Sobel gradient used by BumpTexture (exposed for tests)
std::vector<float> sobel_gradient(const float *height, int w, int h, float scale);

} // namespace tinypt
