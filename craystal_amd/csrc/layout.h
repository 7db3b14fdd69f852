// GPU data layout of a scene in HBM (host + device visible).
//
// The reference stores 36-byte BVHNode variants in post-order and fetches
// both children (2 x 36 B, unaligned) per internal visit, then goes through
// two indirections (bvhToMesh, index buffer) per triangle test
// (Core/BVH.cu:13-84, Scene/Triangle.cu:40-75).  Here:
//   * GNode (64 B, one cache-line half): both child boxes + two child refs,
//     internal nodes only, in depth-first pre-order (left child follows its
//     parent) so a descent walks forward through memory;
//   * GTri (48 B, three 16 B loads): v0, e1 = v1 - v0, e2 = v2 - v0 and the
//     mesh primitive ID, stored in BVH leaf order so a leaf is one
//     contiguous run.  e1/e2 are the same IEEE single subtractions the
//     reference performs per test, so precomputing them is bit-exact.
// A child ref with bit 31 set is a leaf: the low 31 bits are the index of the
// leaf's first GTri; the last GTri of every leaf carries bit 31 in its
// primitive word.
#pragma once
#include <cstdint>

namespace crh {

constexpr uint32_t kLeafBit = 0x80000000u;
// Hit::prim of a sphere hit: the sphere index with this bit set.
constexpr uint32_t kSphereBit = 0x40000000u;

struct alignas(16) GNode {
    float lmin[3];
    float lmax[3];
    float rmin[3];
    float rmax[3];
    uint32_t left;
    uint32_t right;
    // bit 0: the left box has an axis with |max - min| < FLT_EPSILON (the
    // reference's flat-box case, Math/AABB.cu:10-21); bit 1: the right box.
    uint32_t flat;
    uint32_t pad1;
};
static_assert(sizeof(GNode) == 64, "GNode must be 64 B");

struct alignas(16) GTri {
    float v0[3];
    uint32_t prim;  // mesh primitive ID | (kLeafBit on the leaf's last tri)
    float e1[3];
    uint32_t pad0;
    float e2[3];
    uint32_t pad1;
};
static_assert(sizeof(GTri) == 48, "GTri must be 48 B");

// Material on the device: the reference MaterialData fields the hot path
// reads (Core/Material.h:34-66) plus values derived once per material.
struct alignas(16) GMaterial {
    float kd[3];
    uint32_t type;
    float ks[3];
    uint32_t flags;
    float emission[3];
    float shininess;
    uint32_t tex;  // diffuse texture (MaterialData::diffuseHandle)
    uint32_t pad[3];
};
static_assert(sizeof(GMaterial) == 64, "GMaterial must be 64 B");

// Everything a kernel needs about one scene, passed by value.
struct SceneView {
    const GNode* nodes;
    const GTri* tris;
    uint32_t rootRef;   // 0 = internal node 0, or a leaf ref
    uint32_t hasGeom;   // 0 for an empty scene
    uint32_t stackCap;  // tree depth in levels (stack entries needed)
    uint32_t nEmissive;
    float rootMin[3];
    float rootMax[3];
    // Shading data: the reference's global SOA vertex/index buffers.
    const float* pos;         // 3 floats per vertex
    const float* nrm;         // 3 floats per vertex
    const float* uv;          // 2 floats per vertex
    const uint32_t* idx;      // 3 global vertex indices per primitive
    const uint32_t* primMat;  // primitive -> material id
    const GMaterial* mats;
    const uint32_t* emissive;
    float env[3];
    uint32_t pad;
    // Spectral transport table (layout in spec_layout.h, kSpec*).
    const float* spec;
    // Spheres (SphereSOA, Scene/Sphere.h): 4 floats each (centre, radius);
    // sphereMat null = the reference's material rule (the material of the
    // triangle with the same index).
    const float* spheres;
    const uint32_t* sphereMat;
    // Textures: 4 floats per texel, all images back to back; per texture 4
    // words (first texel, width, height, 0).
    const float* texels;
    const uint32_t* texInfo;
    uint32_t nSphere;
    uint32_t nTexture;
};

struct CameraView {
    float pos[3];
    float hh;  // tan(radians(fovY * 0.5)) evaluated in double, as the reference
    float u[3];
    float aspect;
    float v[3];
    uint32_t width;
    float w[3];
    uint32_t height;
};

struct IntegratorView {
    uint32_t maxDepth;
    float rrThreshold;
    uint32_t rrStartDepth;
    uint32_t useMIS;
};

// One scene of a batched render: its view into the pooled scene data, its
// camera, the per-pass sample buffers and the film of its sensor.
struct BatchScene {
    SceneView view;
    CameraView cam;
    float* L;  // float4 per pixel
    float* P;  // float2 per pixel
    float* sum;
    float* wsum;
    uint32_t tilesX;
    uint32_t pad[3];
};

}  // namespace crh
