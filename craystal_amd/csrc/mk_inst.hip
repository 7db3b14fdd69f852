// Launchers of one megakernel family for one traversal stack mode.  The
// build compiles this file once per (CRH_MK_FAMILY, CRH_MK_SM) pair:
// family 0 = kRenderPass, 1 = kRenderRegen, 2 = kBatchRender; SM = 0 LDS,
// 1 private, 2 short stack.  Each object then holds that family's register
// budget and MIS / spectral variants for one stack mode, so the variants
// compile in parallel.
#include "megakernel.h"

#if !defined(CRH_MK_FAMILY) || !defined(CRH_MK_SM)
#error "compile with -DCRH_MK_FAMILY=<0|1|2> -DCRH_MK_SM=<0|1|2>"
#endif

namespace crh {

#if CRH_MK_FAMILY == 0
template <>
hipError_t launchRenderSM<CRH_MK_SM>(const MkArgs& a) {
    constexpr int SM = CRH_MK_SM;
    const dim3 block(kRenderBlock);
    withWaves(a.waves, [&](auto wv) {
        constexpr int W = decltype(wv)::value;
#define CRH_RP(MIS, COUNT, SPEC)                                                                  \
    hipLaunchKernelGGL((kRenderPass<MIS, COUNT, SM, W, SPEC>), a.grid, block, a.lds, a.st, a.s, \
                       a.cam, a.cfg, a.si, a.tilesX, a.tiles, a.band, a.L, a.P, a.counters)
        if (a.spectral && a.mis) CRH_RP(true, false, true);
        else if (a.spectral) CRH_RP(false, false, true);
        else if (a.mis && a.count) CRH_RP(true, true, false);
        else if (a.mis) CRH_RP(true, false, false);
        else if (a.count) CRH_RP(false, true, false);
        else CRH_RP(false, false, false);
#undef CRH_RP
    });
    return hipGetLastError();
}
#elif CRH_MK_FAMILY == 1
template <>
hipError_t launchRegenSM<CRH_MK_SM>(const MkArgs& a) {
    constexpr int SM = CRH_MK_SM;
    const dim3 block(kRenderBlock);
    withWaves(a.waves, [&](auto wv) {
        constexpr int W = decltype(wv)::value;
#define CRH_RG(MIS, SPEC)                                                                      \
    hipLaunchKernelGGL((kRenderRegen<MIS, SPEC, SM, W>), a.grid, block, a.lds, a.st, a.s, a.cam, \
                       a.cfg, a.si, a.tilesX, a.tiles, a.total, a.next, a.L, a.P, a.counters)
        if (a.spectral && a.mis) CRH_RG(true, true);
        else if (a.spectral) CRH_RG(false, true);
        else if (a.mis) CRH_RG(true, false);
        else CRH_RG(false, false);
#undef CRH_RG
    });
    return hipGetLastError();
}
#elif CRH_MK_FAMILY == 2
template <>
hipError_t launchBatchSM<CRH_MK_SM>(const BatchArgs& a) {
    constexpr int SM = CRH_MK_SM;
    const dim3 block(kRenderBlock);
    withWaves(a.waves, [&](auto wv) {
        constexpr int W = decltype(wv)::value;
#define CRH_BR(MIS, SPEC)                                                                      \
    hipLaunchKernelGGL((kBatchRender<MIS, SPEC, SM, W>), a.grid, block, a.lds, a.st, a.tab,      \
                       a.tilePrefix, a.nActive, a.cfg, a.si, a.slots, a.tileW, a.chunks, a.next, \
                       a.counters)
        if (a.spectral && a.mis) CRH_BR(true, true);
        else if (a.spectral) CRH_BR(false, true);
        else if (a.mis) CRH_BR(true, false);
        else CRH_BR(false, false);
#undef CRH_BR
    });
    return hipGetLastError();
}
#else
#error "unknown CRH_MK_FAMILY"
#endif

}  // namespace crh
