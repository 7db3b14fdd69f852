// The wavefront's finishing launch (kWfFinish) for one traversal stack mode:
// the build compiles this file once per CRH_WF_SM (0 LDS, 1 private, 2 short
// stack), so its register-budget, MIS and spectral variants compile in
// parallel with the rest of the wavefront.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"
#include "wave.h"
#include "wf_device.h"

#ifndef CRH_WF_SM
#error "compile with -DCRH_WF_SM=<0|1|2>"
#endif

namespace crh {

using namespace wf;

namespace {

template <class Scenes>
CRH_D const SceneView& laneScene(const SceneView& sv, const Scenes& S) {
    if constexpr (Scenes::kBatch) return sv;
    else return S.view(0);
}

// Finishes every path still in flight after `d0` wavefront bounces in one
// persistent launch: a lane takes the next entry of the extension and MIS
// queues (a WaveDispenser: one atomic per chunk), traces and resolves the pending MIS ray of
// the path's last vertex (as kWfMisResolve), gives the record of its last
// guided vertex its snapshot (as kWfShadeHit), and, if the path continues,
// runs pathTraceKernel's loop from depth d0 one bounce per iteration (as
// the megakernel, without guiding); a lane whose path ends takes the next
// entry.  Each path performs the operations of the pure wavefront in the
// same order, so the film is the same; the launch replaces the tail of
// nearly empty bounces, each of which costs a full traversal launch.
template <bool MIS, bool SPEC, int SM, int W, class Scenes>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, 8))) void kWfFinish(
    Scenes S, WfState w, IntegratorView cfg, uint32_t cur, uint32_t d0) {
    extern __shared__ uint2 lds[];
    auto stk = makeStack<SM>(lds);
    // A batch's lanes hold paths of different scenes: sv is the scene of the
    // lane's current path (s refers to it); one scene's view is the argument.
    SceneView sv = S.view(0);
    const SceneView& s = laneScene(sv, S);
    PathCtx<false, typename StackOf<SM>::T> c{s, stk, 0, {0, 0}};
    const uint32_t nExt = w.counters[kCExt0 + cur];
    const uint32_t total = nExt + w.counters[kCMis0 + cur];
    bool live = false;
    WaveDispenser disp;
    const uint32_t chunk = dispenserChunk(total);
    uint32_t p = 0, depth = 0;
    Ray ray{};
    Spec rad = sp(0.f), beta = sp(1.f);
    Sampler smp(0u, 0u, 0u);
    Lanes lw = rgbLanes();
    while (true) {
        {
            const uint32_t i = disp.take(&w.counters[kCPtExt], total, chunk, !live);
            if (i < total) {
                const bool alive = i < nExt;
                p = alive ? w.qExt[cur][i] : w.qMis[cur][i - nExt];
                const uint32_t fl = w.flags[p];
                // the MIS entry of a path that continues is handled with its
                // extension entry (flags are not written here, so this test
                // holds whichever entry a lane takes first)
                if (alive || !(fl & F_ALIVE)) {
                    if constexpr (Scenes::kBatch) sv = S.view(sceneOf(fl));
                    rad = ldS(w.rad, p);
                    if (SPEC) lw = pathLanes(s, fl);
                    if (fl & F_MIS) {
                        Ray r = loadRay(w.misO, w.misD, p);
                        Hit h;
                        bool hit = closest(c, r, h);
                        float4 vs = w.misVal[p];
                        Spec value{vs.x, vs.y, vs.z};
                        float4 bw4 = w.misBw[p];
                        float4 pe = w.misPos[p];
                        misBsdfPart(s, mk3(pe.x, pe.y, pe.z), r, hit, h, __float_as_uint(pe.w),
                                    Spec{bw4.x, bw4.y, bw4.z}, bw4.w, value, lw);
                        rad = rad + (value / vs.w) * ldS(w.misBeta, p);
                    }
                    if (alive) {
                        if (w.recId) {
                            const uint32_t rid = w.recId[p];
                            if (rid != kNoRecord) {
                                w.rec[size_t(rid) * kRecF4 + 6] =
                                    make_float4(rad.r, rad.g, rad.b, 1.f);
                                w.recId[p] = kNoRecord;
                            }
                        }
                        ray = loadRay(w.rayO, w.rayD, p);
                        beta = ldS(w.beta, p);
                        smp = loadSampler(S, w, p, fl);
                        depth = d0;
                        live = true;
                    } else {
                        stS(w.rad, p, rad);
                    }
                }
            }
        }
        if (!__any(live)) {
            if (disp.exhausted) break;
            continue;
        }
        if (live) {
            bool done = !pathBounce<MIS, SPEC>(c, cfg, depth, ray, rad, beta, smp, lw);
            if (!done && ++depth >= cfg.maxDepth) done = true;
            if (done) {
                stS(w.rad, p, rad);
                live = false;
            }
        }
    }
    unsigned long long r64 = c.rays;
    for (int off = 32; off > 0; off >>= 1) r64 += __shfl_down(r64, off, 64);
    if (laneId() == 0 && r64) atomicAdd(w.work, r64);
}

}  // namespace

template <>
hipError_t launchWfFinishSM<CRH_WF_SM>(const WfFinishArgs& a) {
    constexpr int SM = CRH_WF_SM;
    withWaves(a.waves, [&](auto wv) {
        constexpr int W = decltype(wv)::value;
        auto go = [&](const auto& S) {
            using Scenes = std::decay_t<decltype(S)>;
#define CRH_WFF(MIS, SPEC)                                                                  \
    hipLaunchKernelGGL((kWfFinish<MIS, SPEC, SM, W, Scenes>), dim3(a.grid), dim3(256), a.lds, \
                       a.st, S, a.w, a.cfg, a.cur, a.d0)
            if (a.spectral && a.cfg.useMIS) CRH_WFF(true, true);
            else if (a.spectral) CRH_WFF(false, true);
            else if (a.cfg.useMIS) CRH_WFF(true, false);
            else CRH_WFF(false, false);
#undef CRH_WFF
        };
        if (a.batch) go(SceneTable{a.tab, a.pixPrefix, a.nActive});
        else go(OneScene{a.s});
    });
    return hipGetLastError();
}

}  // namespace crh
