// Scene import and EXR output for C++ host programs, header-only, over
// CRaystal.h.  Mirrors the reference's Importer (Utils/Importer.cpp:23-236)
// and Image::writeEXR (Utils/Image.cpp:49-106) without assimp, pugixml or
// tinyexr:
//   * scene.xml: top-level <camera width height fovy> with <eye>, <lookat>,
//     <up> children (x, y, z attributes) and any number of
//     <light mtlname radiance="r, g, b"/>;
//   * scene.obj next to it (same stem) with its mtllib;
//   * every material is Phong; Kd, Ks, Ns, Ni and opacity d (transmittance
//     1 - d) are read; a material named by a <light> gets its radiance as
//     emission; primitive IDs follow mesh and face order and a triangle is
//     emissive iff its material is;
//   * a mesh starts at every o / g statement and material change; polygons
//     are fan-triangulated; faces without normals get their flat normal;
//     identical (position, uv, normal) corners are joined;
//   * a map_Kd diffuse texture (its last token, relative to the scene file)
//     is loaded by Image::load (PNG or EXR) into SceneData::textureImages,
//     and its material gets hasDiffuseTexture and the image's index
//     (Utils/Importer.cpp:113-120).
// Images are written with Image::writeEXR (Image.h).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "CRaystal.h"
#include "Image.h"

namespace CRay {

namespace detail {

inline std::string readText(const std::filesystem::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw RuntimeError(CRH_ERR_INVALID_ARG, "cannot open " + p.string());
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Elements of an XML text as (name, attributes), in document order; enough
// for the scene description (no text content, entities or CDATA).
struct XmlElement {
    std::string name;
    std::map<std::string, std::string> attr;
};
inline std::vector<XmlElement> xmlElements(const std::string& s) {
    std::vector<XmlElement> out;
    size_t i = 0;
    while ((i = s.find('<', i)) != std::string::npos) {
        if (s.compare(i, 4, "<!--") == 0) {
            size_t e = s.find("-->", i);
            if (e == std::string::npos) break;
            i = e + 3;
            continue;
        }
        size_t e = s.find('>', i);
        if (e == std::string::npos) throw RuntimeError(CRH_ERR_INVALID_ARG, "Invalid XML file.");
        char c = s[i + 1];
        if (c == '?' || c == '/' || c == '!') {
            i = e + 1;
            continue;
        }
        std::string tag = s.substr(i + 1, e - i - 1);
        if (!tag.empty() && tag.back() == '/') tag.pop_back();
        XmlElement el;
        size_t p = 0;
        while (p < tag.size() && !std::isspace(static_cast<unsigned char>(tag[p]))) ++p;
        el.name = tag.substr(0, p);
        while (p < tag.size()) {
            while (p < tag.size() && std::isspace(static_cast<unsigned char>(tag[p]))) ++p;
            size_t eq = tag.find('=', p);
            if (eq == std::string::npos) break;
            std::string key = tag.substr(p, eq - p);
            while (!key.empty() && std::isspace(static_cast<unsigned char>(key.back())))
                key.pop_back();
            size_t q0 = tag.find_first_of("\"'", eq);
            if (q0 == std::string::npos) break;
            size_t q1 = tag.find(tag[q0], q0 + 1);
            if (q1 == std::string::npos) break;
            el.attr[key] = tag.substr(q0 + 1, q1 - q0 - 1);
            p = q1 + 1;
        }
        out.push_back(std::move(el));
        i = e + 1;
    }
    return out;
}

inline float attrF(const XmlElement& e, const char* k, float def = 0.f) {
    auto it = e.attr.find(k);
    return it == e.attr.end() ? def : std::stof(it->second);
}

inline std::vector<std::string> splitWs(const std::string& line) {
    std::istringstream ss(line);
    std::vector<std::string> t;
    std::string w;
    while (ss >> w) t.push_back(w);
    return t;
}

inline std::string stripComment(const std::string& raw) {
    std::string l = raw.substr(0, raw.find('#'));
    while (!l.empty() && (l.back() == '\r' || std::isspace(static_cast<unsigned char>(l.back()))))
        l.pop_back();
    return l;
}

struct MtlEntry {
    Float3 kd{0, 0, 0}, ks{0, 0, 0};
    float ns = 0.f, ni = 1.f, d = 1.f;
    bool hasD = false;
    std::string diffuseMap;  // map_Kd file name, empty if none
};

inline std::map<std::string, MtlEntry> parseMtl(const std::string& text) {
    std::map<std::string, MtlEntry> mats;
    MtlEntry* cur = nullptr;
    std::istringstream in(text);
    std::string raw;
    while (std::getline(in, raw)) {
        auto t = splitWs(stripComment(raw));
        if (t.empty()) continue;
        auto rest = [&]() {
            std::string r;
            for (size_t k = 1; k < t.size(); ++k) r += (k > 1 ? " " : "") + t[k];
            return r;
        };
        if (t[0] == "newmtl") {
            cur = &mats[rest()];
        } else if (!cur) {
            continue;
        } else if ((t[0] == "Kd" || t[0] == "Ks") && t.size() >= 4) {
            Float3 v{std::stof(t[1]), std::stof(t[2]), std::stof(t[3])};
            (t[0] == "Kd" ? cur->kd : cur->ks) = v;
        } else if (t[0] == "Ns" && t.size() >= 2) {
            cur->ns = std::stof(t[1]);
        } else if (t[0] == "Ni" && t.size() >= 2) {
            cur->ni = std::stof(t[1]);
        } else if (t[0] == "d" && t.size() >= 2) {
            cur->d = std::stof(t[1]);
            cur->hasD = true;
        } else if (t[0] == "Tr" && t.size() >= 2) {
            cur->d = 1.f - std::stof(t[1]);
            cur->hasD = true;
        } else if (t[0] == "map_Kd" && t.size() >= 2) {
            cur->diffuseMap = t.back();
        }
    }
    return mats;
}

}  // namespace detail

struct CameraDesc {
    float fovY = 60.f;
    UInt2 size{0, 0};
    Float3 eye{0, 0, 0}, lookat{0, 0, -1}, up{0, 1, 0};
};

class Importer {
   public:
    // Importer::import (Utils/Importer.cpp:200-236).
    std::shared_ptr<Scene> import(const std::filesystem::path& path, int device = 0) {
        CameraDesc cd;
        SceneData data = loadData(path, cd);
        auto scene = std::make_shared<Scene>(std::move(data), device);
        auto cam = std::make_shared<Camera>();
        cam->setFovY(cd.fovY);
        cam->setWorldPosition(cd.eye);
        cam->setTarget(cd.lookat);
        cam->setUp(cd.up);
        cam->setSensor(std::make_shared<Sensor>(cd.size, 1u, device));
        cam->calculateCameraData();
        scene->setCamera(cam);
        return scene;
    }

    // The host part: SceneData and the camera description.
    SceneData loadData(const std::filesystem::path& path, CameraDesc& cam) {
        if (path.extension() != ".xml")
            throw RuntimeError(CRH_ERR_INVALID_ARG, "Imported file must be xml file.");
        mBasePath = path.parent_path();
        auto els = detail::xmlElements(detail::readText(path));
        std::map<std::string, Float3> lights;
        bool haveCam = false;
        std::string parent;
        for (const auto& e : els) {
            if (e.name == "camera") {
                haveCam = true;
                cam.fovY = detail::attrF(e, "fovy");
                cam.size = {uint32_t(detail::attrF(e, "width")),
                            uint32_t(detail::attrF(e, "height"))};
            } else if (e.name == "eye" || e.name == "lookat" || e.name == "up") {
                Float3 v{detail::attrF(e, "x"), detail::attrF(e, "y"), detail::attrF(e, "z")};
                (e.name == "eye" ? cam.eye : e.name == "lookat" ? cam.lookat : cam.up) = v;
            } else if (e.name == "light") {
                Float3 r{0, 0, 0};
                auto it = e.attr.find("radiance");
                if (it != e.attr.end()) {
                    std::istringstream ss(it->second);
                    std::string tok;
                    for (int k = 0; k < 3 && std::getline(ss, tok, ','); ++k) r[k] = std::stof(tok);
                }
                auto n = e.attr.find("mtlname");
                lights[n == e.attr.end() ? std::string() : n->second] = r;
            }
        }
        if (!haveCam) throw RuntimeError(CRH_ERR_INVALID_ARG, "No camera node found.");
        auto obj = path;
        obj.replace_extension(".obj");
        return buildSceneData(detail::readText(obj), lights);
    }

    SceneData buildSceneData(const std::string& objText, const std::map<std::string, Float3>& lights) {
        using Corner = std::tuple<int, int, int>;
        struct RawMesh {
            std::string material;
            bool hasMaterial;
            std::vector<std::array<Corner, 3>> tris;
        };
        std::vector<std::array<double, 3>> V, VN;
        std::vector<std::array<double, 2>> VT;
        std::vector<RawMesh> meshes;
        std::map<std::string, detail::MtlEntry> mtl;
        std::string curMat;
        bool haveMat = false, startNew = true;
        auto idx = [](const std::string& tok, size_t n) -> int {
            if (tok.empty()) return -1;
            int i = std::stoi(tok);
            return i > 0 ? i - 1 : int(n) + i;
        };
        std::istringstream in(objText);
        std::string raw;
        while (std::getline(in, raw)) {
            auto t = detail::splitWs(detail::stripComment(raw));
            if (t.empty()) continue;
            auto rest = [&]() {
                std::string r;
                for (size_t k = 1; k < t.size(); ++k) r += (k > 1 ? " " : "") + t[k];
                return r;
            };
            if (t[0] == "v" && t.size() >= 4) {
                V.push_back({std::stod(t[1]), std::stod(t[2]), std::stod(t[3])});
            } else if (t[0] == "vt" && t.size() >= 2) {
                VT.push_back({std::stod(t[1]), t.size() >= 3 ? std::stod(t[2]) : 0.0});
            } else if (t[0] == "vn" && t.size() >= 4) {
                VN.push_back({std::stod(t[1]), std::stod(t[2]), std::stod(t[3])});
            } else if (t[0] == "o" || t[0] == "g") {
                startNew = true;
            } else if (t[0] == "usemtl") {
                std::string n = rest();
                if (!haveMat || n != curMat) {
                    curMat = n;
                    haveMat = true;
                    startNew = true;
                }
            } else if (t[0] == "mtllib") {
                auto p = mBasePath / rest();
                if (std::filesystem::is_regular_file(p)) {
                    auto m = detail::parseMtl(detail::readText(p));
                    for (auto& kv : m) mtl[kv.first] = kv.second;
                }
            } else if (t[0] == "f" && t.size() >= 4) {
                std::vector<Corner> cs;
                for (size_t k = 1; k < t.size(); ++k) {
                    std::string parts[3];
                    std::istringstream ps(t[k]);
                    for (int q = 0; q < 3 && std::getline(ps, parts[q], '/'); ++q) {
                    }
                    cs.emplace_back(idx(parts[0], V.size()), idx(parts[1], VT.size()),
                                    idx(parts[2], VN.size()));
                }
                if (startNew || meshes.empty()) {
                    meshes.push_back({curMat, haveMat, {}});
                    startNew = false;
                }
                for (size_t k = 1; k + 1 < cs.size(); ++k)
                    meshes.back().tris.push_back({cs[0], cs[k], cs[k + 1]});
            }
        }
        SceneData data;
        std::vector<std::pair<bool, std::string>> names;
        auto nameId = [&](const RawMesh& m) {
            for (size_t k = 0; k < names.size(); ++k)
                if (names[k].first == m.hasMaterial && names[k].second == m.material)
                    return uint32_t(k);
            names.emplace_back(m.hasMaterial, m.material);
            return uint32_t(names.size() - 1);
        };
        std::vector<uint32_t> meshMat;
        for (const auto& m : meshes)
            if (!m.tris.empty()) meshMat.push_back(nameId(m));
        for (const auto& n : names) {
            detail::MtlEntry e;
            if (n.first) {
                auto it = mtl.find(n.second);
                if (it != mtl.end()) e = it->second;
            }
            MaterialData md;
            if (!e.diffuseMap.empty()) {
                md.flags |= uint32_t(MaterialFlags::hasDiffuseTexture);
                md.diffuseHandle = uint32_t(data.textureImages.size());
                data.textureImages.push_back(Image::load(resolveResourcePath(e.diffuseMap)));
            }
            md.diffuseRefl = e.kd;
            md.specularRefl = e.ks;
            if (e.hasD) md.transmittance = {1.f - e.d, 1.f - e.d, 1.f - e.d};
            md.shininess = e.ns;
            md.IoR = e.ni;
            md.type = MaterialType::Phong;
            auto lt = n.first ? lights.find(n.second) : lights.end();
            if (lt != lights.end()) md.emission = lt->second;
            md.finalize();
            data.materials.push_back(md);
        }
        uint32_t prim = 0, mi = 0;
        for (const auto& m : meshes) {
            if (m.tris.empty()) continue;
            MeshData out;
            out.materialID = meshMat[mi++];
            std::map<std::tuple<int, int, int, float, float, float>, uint32_t> ids;
            for (const auto& tri : m.tris) {
                for (const auto& c : tri)
                    if (std::get<0>(c) < 0 || std::get<0>(c) >= int(V.size()))
                        throw RuntimeError(CRH_ERR_INVALID_ARG, "face references a missing vertex");
                bool flat = false;
                for (const auto& c : tri) flat |= std::get<2>(c) < 0;
                std::array<double, 3> fn{0, 0, 1};
                if (flat) {
                    const auto& a = V[std::get<0>(tri[0])];
                    const auto& b = V[std::get<0>(tri[1])];
                    const auto& c = V[std::get<0>(tri[2])];
                    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                   e1[0] * e2[1] - e1[1] * e2[0]};
                    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                    if (l > 0) fn = {n[0] / l, n[1] / l, n[2] / l};
                }
                for (const auto& c : tri) {
                    int vi = std::get<0>(c), ti = std::get<1>(c), ni = std::get<2>(c);
                    Float3 nrm = ni >= 0 ? Float3{float(VN[ni][0]), float(VN[ni][1]), float(VN[ni][2])}
                                         : Float3{float(fn[0]), float(fn[1]), float(fn[2])};
                    auto key = ni >= 0 ? std::make_tuple(vi, ti, ni, 0.f, 0.f, 0.f)
                                       : std::make_tuple(vi, ti, -1, nrm[0], nrm[1], nrm[2]);
                    auto it = ids.find(key);
                    if (it == ids.end()) {
                        it = ids.emplace(key, uint32_t(out.position.size())).first;
                        out.position.push_back(
                            {float(V[vi][0]), float(V[vi][1]), float(V[vi][2])});
                        out.normal.push_back(nrm);
                        out.texCrd.push_back(ti >= 0 && ti < int(VT.size())
                                                 ? Float2{float(VT[ti][0]), float(VT[ti][1])}
                                                 : Float2{0.f, 0.f});
                    }
                    out.index.push_back(it->second);
                }
            }
            const uint32_t nt = uint32_t(out.index.size() / 3);
            if (data.materials[out.materialID].isEmissive())
                for (uint32_t k = 0; k < nt; ++k) data.emissiveIndex.push_back(prim + k);
            prim += nt;
            data.meshes.push_back(std::move(out));
        }
        return data;
    }

    std::filesystem::path resolveResourcePath(const std::filesystem::path& rel) const {
        return mBasePath / rel;
    }

   private:
    std::filesystem::path mBasePath;
};

}  // namespace CRay
