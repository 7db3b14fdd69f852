// Host binned-SAH BVH builder (product path).
//
// Restates BVH::build (Core/BVH.cu:88-338) so that the node array and the
// primitive order are bit-identical to the reference's for the same
// parameters:
//  * centroid = (v0 + v1 + v2) / 3.0f, summed left to right (:138);
//  * bin = clamp(min(bins-1, int((c - min) * (bins / ext))), 0, bins-1), with
//    the float->int conversion of x86-64 (NaN / out of range -> INT_MIN) that
//    a host build of the reference performs (:226-235);
//  * surface area (xy + xz + yz) * 2.0 and cost 1 + (nL*SA_L + nR*SA_R)/SA,
//    first strict minimum over axis-major, bin-minor order (:258-301);
//  * split position min + (i+1) * ext / bins (:268-269);
//  * the bidirectional two-pointer partition that libstdc++'s std::partition
//    uses for vector iterators (:177-181), written out below;
//  * leaf when count <= maxLeaf, depth >= maxDepth or the split is degenerate;
//    left subtree before right; post-order numbering, root last.
// Differences that do not change the result: the three axes are binned in
// one pass over the triangles (per-bin min/max is order independent), the
// triangle data is in structure-of-arrays form, and large subtrees are built
// concurrently; node numbers are assigned in a final post-order walk.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <thread>

namespace crh {

namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();

inline float fminS(float a, float b) { return (b < a) ? b : a; }
inline float fmaxS(float a, float b) { return (a < b) ? b : a; }

// x86-64 cvttss2si semantics.
inline int truncX86(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return static_cast<int>(x);
}

struct Box {
    float mn[3] = {kInf, kInf, kInf};
    float mx[3] = {-kInf, -kInf, -kInf};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminS(mn[k], b.mn[k]);
            mx[k] = fmaxS(mx[k], b.mx[k]);
        }
    }
    float area() const {
        float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
        return float((d0 * d1 + d0 * d2 + d1 * d2) * 2.0);
    }
};

struct TmpNode {
    Box box;
    uint32_t leaf;  // 1 = leaf
    uint32_t a, b;  // leaf: offset,count; internal: TmpNode ids
};

struct Ctx {
    BuildParams p;
    std::vector<float> c[3];  // centroids
    std::vector<Box> tb;      // triangle bounds
    std::vector<uint32_t> ind;
    std::vector<TmpNode> pool;
    std::atomic<uint32_t> next{0};
    std::atomic<int> freeThreads{0};
    std::atomic<uint32_t> maxDepth{0}, maxLeaf{0};

    uint32_t alloc() { return next.fetch_add(1, std::memory_order_relaxed); }
    static void atomicMax(std::atomic<uint32_t>& a, uint32_t v) {
        uint32_t cur = a.load(std::memory_order_relaxed);
        while (cur < v && !a.compare_exchange_weak(cur, v)) {
        }
    }

    uint32_t makeLeaf(const Box& nb, uint32_t start, uint32_t count) {
        uint32_t id = alloc();
        pool[id] = TmpNode{nb, 1u, start, count};
        atomicMax(maxLeaf, count);
        return id;
    }

    void bestSplit(uint32_t start, uint32_t count, const Box& nb,
                   uint32_t& bestAxis, float& bestSplit) {
        const int NB = int(p.bins);
        std::vector<Box> bins(size_t(3) * NB);
        std::vector<int> cnt(size_t(3) * NB, 0);
        float ext[3], scale[3];
        for (int k = 0; k < 3; ++k) {
            ext[k] = nb.mx[k] - nb.mn[k];
            scale[k] = float(NB) / ext[k];
        }
        for (uint32_t i = start; i < start + count; ++i) {
            uint32_t t = ind[i];
            const Box& b = tb[t];
            for (int k = 0; k < 3; ++k) {
                int bi = std::min(NB - 1, truncX86((c[k][t] - nb.mn[k]) * scale[k]));
                bi = std::clamp(bi, 0, NB - 1);
                bins[size_t(k) * NB + bi].grow(b);
                cnt[size_t(k) * NB + bi]++;
            }
        }
        float bestCost = kInf;
        bestAxis = 0;
        bestSplit = 0.0f;
        std::vector<Box> lb(NB - 1), rb(NB - 1);
        std::vector<int> lc(NB - 1), rc(NB - 1);
        const float rootSA = nb.area();
        for (int k = 0; k < 3; ++k) {
            Box cur;
            int cc = 0;
            for (int i = 0; i < NB - 1; ++i) {
                cur.grow(bins[size_t(k) * NB + i]);
                cc += cnt[size_t(k) * NB + i];
                lb[i] = cur;
                lc[i] = cc;
            }
            Box curR;
            int cr = 0;
            for (int i = NB - 1; i > 0; --i) {
                curR.grow(bins[size_t(k) * NB + i]);
                cr += cnt[size_t(k) * NB + i];
                rb[i - 1] = curR;
                rc[i - 1] = cr;
            }
            for (int i = 0; i < NB - 1; ++i) {
                float splitPos = nb.mn[k] + float(i + 1) * ext[k] / float(NB);
                float lsa = lb[i].area(), rsa = rb[i].area();
                float cost = 1.0f + 1.0f * (lc[i] * lsa + rc[i] * rsa) / rootSA;
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = uint32_t(k);
                    bestSplit = splitPos;
                }
            }
        }
    }

    // Two-pointer partition, as libstdc++'s bidirectional std::__partition.
    uint32_t partition(uint32_t start, uint32_t count, uint32_t axis,
                       float split) {
        const float* cc = c[axis].data();
        uint32_t* first = ind.data() + start;
        uint32_t* last = first + count;
        uint32_t* begin = first;
        while (true) {
            while (true) {
                if (first == last) return uint32_t(first - begin);
                if (cc[*first] < split) ++first;
                else break;
            }
            --last;
            while (true) {
                if (first == last) return uint32_t(first - begin);
                if (!(cc[*last] < split)) --last;
                else break;
            }
            std::swap(*first, *last);
            ++first;
        }
    }

    uint32_t build(uint32_t start, uint32_t count, uint32_t depth) {
        Box nb;
        for (uint32_t i = start; i < start + count; ++i) nb.grow(tb[ind[i]]);
        if (count <= p.maxLeaf || depth >= p.maxDepth)
            return makeLeaf(nb, start, count);
        uint32_t axis;
        float split;
        bestSplit(start, count, nb, axis, split);
        uint32_t lc = partition(start, count, axis, split);
        if (lc == 0 || lc == count) return makeLeaf(nb, start, count);
        uint32_t l, r;
        bool spawn = count >= 32768 && freeThreads.fetch_sub(1) > 0;
        if (spawn) {
            auto fut = std::async(std::launch::async,
                                  [&] { return build(start, lc, depth + 1); });
            r = build(start + lc, count - lc, depth + 1);
            l = fut.get();
            freeThreads.fetch_add(1);
        } else {
            if (count >= 32768) freeThreads.fetch_add(1);
            l = build(start, lc, depth + 1);
            r = build(start + lc, count - lc, depth + 1);
        }
        uint32_t id = alloc();
        pool[id] = TmpNode{nb, 0u, l, r};
        atomicMax(maxDepth, depth);
        return id;
    }
};

// Post-order emission: left subtree, right subtree, then the node.
struct Emitter {
    const Ctx& cx;
    BvhResult& out;
    uint32_t emit(uint32_t id, uint32_t level) {
        const TmpNode& n = cx.pool[id];
        out.levels = std::max(out.levels, level);
        crh_bvh_node_ref r;
        if (n.leaf) {
            r.tag = 0;
            r.a = n.a;
            r.b = n.b;
        } else {
            uint32_t l = emit(n.a, level + 1);
            uint32_t rr = emit(n.b, level + 1);
            r.tag = 1;
            r.a = l;
            r.b = rr;
            out.internalCount++;
        }
        for (int k = 0; k < 3; ++k) {
            r.pmin[k] = n.box.mn[k];
            r.pmax[k] = n.box.mx[k];
        }
        out.nodes.push_back(r);
        return uint32_t(out.nodes.size() - 1);
    }
};

}  // namespace

BuildParams resolveParams(const crh_bvh_params* p) {
    BuildParams bp;
    if (p) {
        if (p->bins) bp.bins = p->bins;
        if (p->max_leaf) bp.maxLeaf = p->max_leaf;
        if (p->max_depth) bp.maxDepth = p->max_depth;
    }
    return bp;
}

BvhResult buildBvh(const float* pos, uint32_t nvert, const uint32_t* idx,
                   uint32_t ntri, const BuildParams& params) {
    (void)nvert;
    auto t0 = std::chrono::steady_clock::now();
    Ctx cx;
    cx.p = params;
    for (int k = 0; k < 3; ++k) cx.c[k].resize(ntri);
    cx.tb.resize(ntri);
    for (uint32_t t = 0; t < ntri; ++t) {
        const float* v0 = pos + 3 * size_t(idx[3 * size_t(t)]);
        const float* v1 = pos + 3 * size_t(idx[3 * size_t(t) + 1]);
        const float* v2 = pos + 3 * size_t(idx[3 * size_t(t) + 2]);
        Box b;
        for (int k = 0; k < 3; ++k) {
            cx.c[k][t] = (v0[k] + v1[k] + v2[k]) / 3.0f;
            b.mn[k] = fminS(fminS(fminS(b.mn[k], v0[k]), v1[k]), v2[k]);
            b.mx[k] = fmaxS(fmaxS(fmaxS(b.mx[k], v0[k]), v1[k]), v2[k]);
        }
        cx.tb[t] = b;
    }
    cx.ind.resize(ntri);
    for (uint32_t i = 0; i < ntri; ++i) cx.ind[i] = i;
    cx.pool.resize(size_t(2) * ntri + 1);
    unsigned hc = std::thread::hardware_concurrency();
    cx.freeThreads = int(hc > 1 ? std::min(hc, 32u) - 1 : 0);
    uint32_t rootTmp = cx.build(0, ntri, 0);

    BvhResult out;
    out.nodes.reserve(cx.next.load());
    Emitter em{cx, out};
    out.root = em.emit(rootTmp, 0);
    const TmpNode& rn = cx.pool[rootTmp];
    for (int k = 0; k < 3; ++k) {
        out.rootMin[k] = rn.box.mn[k];
        out.rootMax[k] = rn.box.mx[k];
    }
    out.primOrder = std::move(cx.ind);  // triIndex == build index
    out.maxDepth = cx.maxDepth.load();
    out.maxLeaf = cx.maxLeaf.load();
    out.buildMs = std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t0)
                      .count();
    return out;
}

GpuBvh toGpuLayout(const BvhResult& b, const float* pos, const uint32_t* idx) {
    GpuBvh g;
    const uint32_t ntri = uint32_t(b.primOrder.size());
    g.hasGeom = ntri > 0 ? 1u : 0u;
    g.tris.resize(ntri);
    for (uint32_t i = 0; i < ntri; ++i) {
        uint32_t prim = b.primOrder[i];
        const float* v0 = pos + 3 * size_t(idx[3 * size_t(prim)]);
        const float* v1 = pos + 3 * size_t(idx[3 * size_t(prim) + 1]);
        const float* v2 = pos + 3 * size_t(idx[3 * size_t(prim) + 2]);
        GTri& t = g.tris[i];
        for (int k = 0; k < 3; ++k) {
            t.v0[k] = v0[k];
            t.e1[k] = v1[k] - v0[k];
            t.e2[k] = v2[k] - v0[k];
        }
        t.prim = prim;
        t.pad0 = t.pad1 = 0;
    }
    if (!g.hasGeom) return g;
    // Mark the end of every leaf and assign pre-order ids to internal nodes.
    std::vector<uint32_t> gid(b.nodes.size(), 0xFFFFFFFFu);
    struct Item {
        uint32_t node;
    };
    std::vector<uint32_t> stack;
    stack.push_back(b.root);
    uint32_t nextId = 0;
    std::vector<uint32_t> order;
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        const crh_bvh_node_ref& r = b.nodes[n];
        if (r.tag == 0) {
            g.tris[r.a + r.b - 1].prim |= kLeafBit;
            continue;
        }
        gid[n] = nextId++;
        order.push_back(n);
        stack.push_back(r.b);  // right popped after the whole left subtree
        stack.push_back(r.a);
    }
    auto ref = [&](uint32_t n) -> uint32_t {
        const crh_bvh_node_ref& r = b.nodes[n];
        return r.tag == 0 ? (kLeafBit | r.a) : gid[n];
    };
    g.nodes.resize(order.size());
    for (uint32_t n : order) {
        const crh_bvh_node_ref& r = b.nodes[n];
        GNode& o = g.nodes[gid[n]];
        const crh_bvh_node_ref& L = b.nodes[r.a];
        const crh_bvh_node_ref& R = b.nodes[r.b];
        for (int k = 0; k < 3; ++k) {
            o.lmin[k] = L.pmin[k];
            o.lmax[k] = L.pmax[k];
            o.rmin[k] = R.pmin[k];
            o.rmax[k] = R.pmax[k];
        }
        o.left = ref(r.a);
        o.right = ref(r.b);
        // the same single-precision test the slab code makes per axis
        auto flat = [](const crh_bvh_node_ref& c) {
            for (int k = 0; k < 3; ++k) {
                const float ext = c.pmax[k] - c.pmin[k];
                if (std::fabs(ext) < FLT_EPSILON) return 1u;
            }
            return 0u;
        };
        o.flat = flat(L) | (flat(R) << 1);
        o.pad1 = 0;
    }
    g.rootRef = ref(b.root);
    return g;
}

}  // namespace crh
