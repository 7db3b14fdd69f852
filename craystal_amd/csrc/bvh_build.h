// Host binned-SAH BVH builder and the conversion of its result to the GPU
// layout (layout.h).
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/crayhip.h"
#include "layout.h"

namespace crh {

struct BuildParams {
    uint32_t bins = 32;      // NUM_BINS (Core/BVH.h:104)
    uint32_t maxLeaf = 4;    // MIN_TRIS_PER_LEAF, used as count <= 4 (:105)
    uint32_t maxDepth = 20;  // MAX_DEPTH (:106)
};

struct BvhResult {
    std::vector<crh_bvh_node_ref> nodes;  // reference post-order, root last
    std::vector<uint32_t> primOrder;      // bvhToMeshPrimitiveIndex
    uint32_t root = 0;
    float rootMin[3], rootMax[3];
    uint32_t maxDepth = 0;  // deepest internal node (reference BuildStats)
    uint32_t maxLeaf = 0;
    uint32_t levels = 0;    // deepest node of any kind (root = 0)
    uint32_t internalCount = 0;
    double buildMs = 0;
};

BuildParams resolveParams(const crh_bvh_params* p);

// Bit-exact with BVH::build (Core/BVH.cu:88-338) for the same parameters.
// Subtrees above a size threshold are built on worker threads; the final
// node numbering is assigned afterwards in the reference's post-order.
BvhResult buildBvh(const float* pos, uint32_t nvert, const uint32_t* idx,
                   uint32_t ntri, const BuildParams& params);

struct GpuBvh {
    std::vector<GNode> nodes;  // internal nodes, depth-first pre-order
    std::vector<GTri> tris;    // triangle records in leaf order
    uint32_t rootRef = 0;
    uint32_t hasGeom = 0;
};

// pos: concatenated vertex positions; idx: global vertex indices.
GpuBvh toGpuLayout(const BvhResult& b, const float* pos, const uint32_t* idx);

}  // namespace crh
