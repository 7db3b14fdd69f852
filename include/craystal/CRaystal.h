// CRaystal.h - C++ host API over libcrayhip.so, mirroring the reference's
// public classes so reference-style host code ports by changing includes:
//   CRay::Float2/Float3/UInt2   Core/Vec.h:8-63 (Vec.h: .x/.y/.z, .r/.g/.b, [])
//   CRay::Image                 Utils/Image.h:22-95 (Image.h)
//   CRay::Spectrum              Core/Spectrum.h:237-276 (RGB, as the reference uses it)
//   CRay::RuntimeError          Core/Error.h:59-66
//   CRay::MaterialData          Core/Material.h:34-66 (+ finalize)
//   CRay::MeshData / SceneData  Scene/Triangle.h:36-42, Scene/Scene.h:40-47
//   CRay::BVH                   Core/BVH.h:48-61
//   CRay::Sensor                Core/Sensor.h:31-70
//   CRay::Camera                Core/Camera.h:12-50
//   CRay::Scene                 Scene/Scene.h:49-77
//   CRay::PathTraceIntegrator   Core/Integrator.h:10-38
// Header-only; link with -lcrayhip -lz.  Every nonzero C-ABI status throws
// CRay::RuntimeError carrying crh_last_error().  Ownership follows the
// reference: host objects own device memory, copy is deleted, move is
// allowed, Camera/Sensor/Scene are shared through ::Ref aliases.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../crayhip.h"
#include "Image.h"
#include "Vec.h"

namespace CRay {

// The reference transports RGB (using Spectrum = RGBSpectrum,
// Core/Spectrum.h:326): three floats with the Spectrum helpers host code
// calls.
struct Spectrum : Float3 {
    using Float3::Float3;
    Spectrum() = default;
    Spectrum(const Float3& v) : Float3(v) {}
    static Spectrum fromRGB(Float3 rgb, bool isIllum = false) {
        (void)isIllum;
        return Spectrum(rgb);
    }
    static void initialize() {}
    Float3 toRGB() const { return *this; }
    float maxValue() const { return std::max(std::max(x, y), z); }
    float averageValue() const { return (x + y + z) / 3.0f; }
};

// Logger::init (Utils/Logger.cpp): this library logs nothing itself.
struct Logger {
    static void init() {}
};

namespace detail {
inline void check(int status) {
    if (status != CRH_OK) throw RuntimeError(status, crh_last_error());
}

// One context per device, created on first use (the reference relies on the
// implicit CUDA runtime context).
inline crh_context* context(int device = 0) {
    struct Holder {
        crh_context* ctx = nullptr;
        ~Holder() {
            if (ctx) crh_context_destroy(ctx);
        }
    };
    static Holder h[16];
    if (device < 0 || device >= 16) throw RuntimeError(CRH_ERR_INVALID_ARG, "device index");
    if (!h[device].ctx) check(crh_context_create(device, &h[device].ctx));
    return h[device].ctx;
}
}  // namespace detail

enum class MaterialType : uint32_t { FullDiffuse = 0, Principled = 1, Phong = 2 };
enum class MaterialFlags : uint32_t { IsEmissive = 1, IsDiffuse = 2, hasDiffuseTexture = 4 };

struct MaterialData {
    Spectrum diffuseRefl{0, 0, 0};
    uint32_t diffuseHandle = 0;
    Spectrum specularRefl{0, 0, 0};
    Spectrum transmittance{0, 0, 0};
    Spectrum emission{0, 0, 0};
    Float shininess = 0;
    Float IoR = 1;
    MaterialType type = MaterialType::FullDiffuse;
    uint32_t flags = 0;

    bool isEmissive() const { return flags & uint32_t(MaterialFlags::IsEmissive); }
    bool isDiffuse() const { return flags & uint32_t(MaterialFlags::IsDiffuse); }

    crh_material toC() const {
        crh_material m{};
        for (int k = 0; k < 3; ++k) {
            m.diffuse[k] = diffuseRefl[k];
            m.specular[k] = specularRefl[k];
            m.transmittance[k] = transmittance[k];
            m.emission[k] = emission[k];
        }
        m.diffuse_tex = diffuseHandle;
        m.shininess = shininess;
        m.ior = IoR;
        m.type = uint32_t(type);
        m.flags = flags;
        return m;
    }
    void finalize() {
        crh_material m = toC();
        detail::check(crh_material_finalize(&m));
        flags = m.flags;
    }
};

struct MeshData {
    std::vector<Float3> position;
    std::vector<Float3> normal;
    std::vector<Float2> texCrd;
    std::vector<uint32_t> index;
    uint32_t materialID = 0;
};

// SphereData (Scene/Sphere.h).
struct SphereData {
    Float3 center{0, 0, 0};
    float radius = 1.f;
};

struct SceneData {
    std::vector<SphereData> spheres;
    std::vector<MeshData> meshes;
    std::vector<MaterialData> materials;
    std::vector<Image> textureImages;  // diffuse textures, by MaterialData::diffuseHandle
    std::vector<uint32_t> emissiveIndex;
    Float3 envRadiance{0, 0, 0};  // constant environment; zero = reference
    crh_bvh_params bvhParams{32, 4, 20, 0};
    // Per-sphere materials; empty = the reference's rule (a sphere takes the
    // material of the triangle with its index).
    std::vector<uint32_t> sphereMaterials;
};

class BVH {
   public:
    using Ref = std::shared_ptr<BVH>;
    BVH() = default;
    BVH(const BVH&) = delete;
    BVH& operator=(const BVH&) = delete;
    ~BVH() {
        if (mBvh) crh_bvh_destroy(mBvh);
    }

    // BVH::build (Core/BVH.cu:88-116) over the concatenated meshes.
    void build(const SceneData& data) {
        std::vector<float> pos;
        std::vector<uint32_t> idx;
        for (const auto& m : data.meshes) {
            uint32_t off = uint32_t(pos.size() / 3);
            for (const auto& p : m.position) pos.insert(pos.end(), {p.x, p.y, p.z});
            for (uint32_t i : m.index) idx.push_back(i + off);
        }
        if (mBvh) crh_bvh_destroy(mBvh);
        mBvh = nullptr;
        detail::check(crh_bvh_build(pos.data(), uint32_t(pos.size() / 3), idx.data(),
                                    uint32_t(idx.size() / 3), &data.bvhParams, &mBvh));
    }
    crh_bvh_stats stats() const {
        crh_bvh_stats s{};
        detail::check(crh_bvh_get_stats(mBvh, &s));
        return s;
    }
    void exportNodes(std::vector<crh_bvh_node_ref>& nodes, std::vector<uint32_t>& prims) const {
        crh_bvh_stats s = stats();
        nodes.resize(s.node_count);
        prims.resize(s.prim_count);
        detail::check(crh_bvh_export(mBvh, nodes.data(), prims.data()));
    }

   private:
    crh_bvh* mBvh = nullptr;
};

class Sensor {
   public:
    using Ref = std::shared_ptr<Sensor>;
    explicit Sensor(UInt2 size, uint32_t spp = 1u, int device = 0) : mSize(size), mSpp(spp) {
        detail::check(crh_film_create(detail::context(device), size.x, size.y, &mFilm));
        mData.resize(size_t(size.x) * size.y * 3);
        mWeight.resize(size_t(size.x) * size.y);
    }
    Sensor(const Sensor&) = delete;
    Sensor& operator=(const Sensor&) = delete;
    Sensor(Sensor&& o) noexcept
        : mFilm(std::exchange(o.mFilm, nullptr)), mSize(o.mSize), mSpp(o.mSpp),
          mData(std::move(o.mData)), mWeight(std::move(o.mWeight)) {}
    ~Sensor() {
        if (mFilm) crh_film_destroy(mFilm);
    }

    void setSPP(uint32_t v) { mSpp = v; }
    uint32_t getSPP() const { return mSpp; }
    UInt2 getSize() const { return mSize; }
    void clear() { detail::check(crh_film_clear(mFilm)); }
    void readbackDeviceData() {
        detail::check(crh_film_readback(mFilm, mData.data(), mWeight.data()));
    }
    // Sensor::createImage (Core/Sensor.cu:93-112): a 3-channel linear image
    // of rgb = sum / weight from the last readback.
    Image::Ref createImage() const {
        auto img = std::make_shared<Image>(mSize.x, mSize.y, 3, ColorSpace::Linear);
        for (uint32_t y = 0; y < mSize.y; ++y)
            for (uint32_t x = 0; x < mSize.x; ++x) {
                const size_t i = size_t(mSize.x) * y + x;
                for (int c = 0; c < 3; ++c)
                    img->setPixel(int(x), int(y), c, mData[3 * i + c] / mWeight[i]);
            }
        return img;
    }
    // The film's sums (3 floats per pixel, row-major) and weights of the last readback.
    const std::vector<float>& sums() const { return mData; }
    const std::vector<float>& weights() const { return mWeight; }
    crh_film* handle() const { return mFilm; }

   private:
    crh_film* mFilm = nullptr;
    UInt2 mSize;
    uint32_t mSpp;
    std::vector<float> mData, mWeight;
};

class Camera {
   public:
    using Ref = std::shared_ptr<Camera>;
    Camera() {
        mData = crh_camera{};
        mData.fov_y = 60.f;
        mData.target[2] = -1.f;
        mData.near_z = 0.01f;
        mData.up[1] = 1.f;
        mData.far_z = 10000.f;
        mData.u[2] = 1.f;
        mData.v[1] = 1.f;
        mData.w[0] = 1.f;
    }
    void setFovY(Float v) { mData.fov_y = v; }
    void setWorldPosition(const Float3& p) { copy3(mData.pos, p); }
    void setTarget(const Float3& p) { copy3(mData.target, p); }
    void setUp(const Float3& p) { copy3(mData.up, p); }
    void setSensor(const Sensor::Ref& s) { mpSensor = s; }
    Sensor::Ref getSensor() { return mpSensor; }
    void calculateCameraData() { detail::check(crh_camera_update(&mData)); }
    const crh_camera& getData() const { return mData; }

   private:
    static void copy3(float* d, const Float3& s) {
        d[0] = s[0];
        d[1] = s[1];
        d[2] = s[2];
    }
    crh_camera mData;
    Sensor::Ref mpSensor;
};

class Scene {
   public:
    using Ref = std::shared_ptr<Scene>;
    // Scene::Scene (Scene/Scene.cu:61-78): builds the BVH and uploads.
    explicit Scene(SceneData&& data, int device = 0) {
        std::vector<crh_mesh> meshes(data.meshes.size());
        std::vector<std::vector<float>> nrmFill(data.meshes.size());
        for (size_t i = 0; i < data.meshes.size(); ++i) {
            const MeshData& m = data.meshes[i];
            crh_mesh& c = meshes[i];
            c.position = m.position.empty() ? nullptr : &m.position[0].x;
            if (m.normal.size() == m.position.size() && !m.normal.empty()) {
                c.normal = &m.normal[0].x;
            } else {
                nrmFill[i].assign(m.position.size() * 3, 0.f);
                c.normal = nrmFill[i].data();
            }
            c.texcrd = m.texCrd.size() == m.position.size() && !m.texCrd.empty()
                           ? &m.texCrd[0].x
                           : nullptr;
            c.nvert = uint32_t(m.position.size());
            c.index = m.index.data();
            c.nindex = uint32_t(m.index.size());
            c.material_id = m.materialID;
        }
        std::vector<crh_material> mats;
        for (const auto& m : data.materials) mats.push_back(m.toC());
        crh_scene_desc d{};
        d.meshes = meshes.data();
        d.nmesh = uint32_t(meshes.size());
        d.materials = mats.data();
        d.nmaterial = uint32_t(mats.size());
        d.emissive_index = data.emissiveIndex.data();
        d.nemissive = uint32_t(data.emissiveIndex.size());
        for (int k = 0; k < 3; ++k) d.env_radiance[k] = data.envRadiance[k];
        d.bvh = data.bvhParams;
        std::vector<float> cen, rad;
        for (const auto& sp : data.spheres) {
            cen.insert(cen.end(), {sp.center[0], sp.center[1], sp.center[2]});
            rad.push_back(sp.radius);
        }
        d.sphere_center = cen.data();
        d.sphere_radius = rad.data();
        d.nsphere = uint32_t(data.spheres.size());
        d.sphere_material = data.sphereMaterials.empty() ? nullptr : data.sphereMaterials.data();
        // Texture(const Image&) (Core/Texture.cu:46-62): channel c of pixel
        // (x, y) to texel y * width + x, interleaved.
        std::vector<crh_texture> tex;
        std::vector<std::vector<float>> texels(data.textureImages.size());
        for (size_t t = 0; t < data.textureImages.size(); ++t) {
            const Image& im = data.textureImages[t];
            const int ch = im.getChannels();
            texels[t].resize(size_t(im.getArea()) * ch);
            for (int c = 0; c < ch; ++c)
                for (int i = 0; i < im.getArea(); ++i) texels[t][size_t(i) * ch + c] = im.getPixel(i, c);
            tex.push_back(crh_texture{uint32_t(im.getWidth()), uint32_t(im.getHeight()),
                                      uint32_t(ch), texels[t].data()});
        }
        d.textures = tex.data();
        d.ntexture = uint32_t(tex.size());
        detail::check(crh_scene_create(detail::context(device), &d, &mScene));
    }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    ~Scene() {
        if (mScene) crh_scene_destroy(mScene);
    }
    Camera::Ref getCamera() { return mpCamera; }
    void setCamera(const Camera::Ref& c) { mpCamera = c; }
    crh_scene* handle() const { return mScene; }

    // SceneView::intersect / intersectOcclusion over a host ray batch.
    std::vector<crh_hit> intersect(const std::vector<crh_ray>& rays) {
        std::vector<crh_hit> hits(rays.size());
        detail::check(crh_trace_closest(mScene, rays.data(), rays.size(), hits.data()));
        return hits;
    }
    std::vector<uint8_t> intersectOcclusion(const std::vector<crh_ray>& rays) {
        std::vector<uint8_t> occ(rays.size());
        detail::check(crh_trace_any(mScene, rays.data(), rays.size(), occ.data()));
        return occ;
    }

   private:
    crh_scene* mScene = nullptr;
    Camera::Ref mpCamera;
};

class PathTraceIntegrator {
   public:
    using Ref = std::shared_ptr<PathTraceIntegrator>;
    // Core/Integrator.h:14-18, with RR start depth and mode added.  Unlike
    // the reference these values reach the device; Configs::reference()
    // gives the behaviour the reference kernel actually had.
    struct Configs {
        uint32_t maxDepth = 10u;
        Float rrThreshold = 0.05f;
        bool useMIS = true;
        uint32_t rrStartDepth = 6u;
        uint32_t mode = CRH_MODE_MEGAKERNEL;
        bool spectral = false;  // three of 95 wavelength bins per path; the film holds XYZ
        static Configs reference() { return Configs{}; }
    };

    explicit PathTraceIntegrator(Configs c) : mConfigs(c) {}

    // Core/Integrator.cu:233-262: renders sample indices [first, first+spp)
    // into the camera's sensor and reads the film back.
    crh_stats dispatch(Scene& scene, int spp, uint32_t firstSample = 0) const {
        auto cam = scene.getCamera();
        if (!cam || !cam->getSensor())
            throw RuntimeError(CRH_ERR_INVALID_ARG, "scene camera/sensor not set");
        auto sensor = cam->getSensor();
        sensor->setSPP(uint32_t(spp));
        crh_integrator_cfg c{};
        c.max_depth = mConfigs.maxDepth;
        c.rr_threshold = mConfigs.rrThreshold;
        c.rr_start_depth = mConfigs.rrStartDepth;
        c.use_mis = mConfigs.useMIS ? 1u : 0u;
        c.mode = mConfigs.mode;
        c.spectral = mConfigs.spectral ? 1u : 0u;
        crh_stats st{};
        detail::check(crh_render(scene.handle(), &cam->getData(), &c, firstSample,
                                 uint32_t(spp), sensor->handle(), &st));
        sensor->readbackDeviceData();
        return st;
    }

   private:
    Configs mConfigs;
};

}  // namespace CRay
