// afd_fused.hip -- the stage update fused with the y faces (mixed RMHD WENO5, flux-divergence path).
#include "afd_launch.hpp"

namespace afd {

// Cells outside the flux range [r0, r1) per axis (the ghost shell of k_update_shell).
long shell_cells(const Grid &g) {
    long in = 1;
    for (int a = 0; a < 3; ++a) in *= (long)(g.r1[a] - g.r0[a]);
    return g.ncell - in;
}

// Workgroups of k_face_march_update: z chunks x runs along y x x lines.
long fused_blocks(const Grid &g) {
    const long nkc = (g.r1[2] - g.r0[2] + 63) / 64, nruns = (g.r1[1] - g.r0[1] + MARCH_CELLS - 1) / MARCH_CELLS;
    return nkc * nruns * (g.r1[0] - g.r0[0]);
}

template <int RIEM, bool LEARNED>
static void go(const FaceArgs<float> &fa, const UpdArgs<float> &ua, hipStream_t s) {
    hipLaunchKernelGGL((k_face_march_update<RIEM, LEARNED>), dim3((unsigned)fused_blocks(fa.g)), dim3(256), 0, s, fa,
                       ua);
    hipLaunchKernelGGL((k_update_shell<9, float, LEARNED>), grid1d(shell_cells(fa.g)), dim3(BLOCK), 0, s, ua);
}

hipError_t launch_stage_fused(const FaceLaunch &l, const UpdLaunch &u, hipStream_t s) {
    FaceArgs<float> fa = face_args<float>(l);
    const Grid &g = l.g;
    const unsigned nkc = (g.r1[2] - g.r0[2] + 63) / 64;
    const unsigned nruns = (g.r1[1] - g.r0[1] + MARCH_CELLS - 1) / MARCH_CELLS;
    fa.dk = make_udiv(nkc);
    fa.dn = make_udiv(nruns);
    UpdArgs<float> ua;
    static_assert(sizeof(UpdArgs<float>) == sizeof(UpdArgs<double>), "layout");
    memcpy((void *)&ua, (const void *)&u.a, sizeof(ua));
    const bool learned = ua.mlp != nullptr;
    if (ua.queue) (void)hipMemsetAsync(ua.qcount, 0, sizeof(unsigned) * QUEUE_SLOTS, s);
    const int key = 2 * (l.riemann < 0 ? 0 : (l.riemann > 2 ? 2 : l.riemann)) + (learned ? 1 : 0);
    switch (key) {
        case 0: go<0, false>(fa, ua, s); break;
        case 1: go<0, true>(fa, ua, s); break;
        case 2: go<1, false>(fa, ua, s); break;
        case 3: go<1, true>(fa, ua, s); break;
        case 4: go<2, false>(fa, ua, s); break;
        default: go<2, true>(fa, ua, s); break;
    }
    if (ua.queue) hipLaunchKernelGGL((k_c2p_queue<float>), dim3(16, QUEUE_SLOTS), dim3(BLOCK), 0, s, ua);
    return hipGetLastError();
}

// The FP64 repair of the cells k_update (or the fused kernels) queued with a non-finite FP32
// right-hand side; nothing runs past each slot's count, so with no such cell the launch only
// reads the counters.
hipError_t launch_repair(const FaceLaunch &l, const UpdLaunch &u, hipStream_t s) {
    UpdArgs<float> ua;
    static_assert(sizeof(UpdArgs<float>) == sizeof(UpdArgs<double>), "layout");
    memcpy((void *)&ua, (const void *)&u.a, sizeof(ua));
    if (!ua.rq) return hipSuccess;
    FaceArgs<double> fa = face_args<double>(l);
    const dim3 grid(16, QUEUE_SLOTS);
    switch (l.riemann < 0 ? 0 : (l.riemann > 2 ? 2 : l.riemann)) {
        case 0: hipLaunchKernelGGL((k_repair_weno<0>), grid, dim3(BLOCK), 0, s, fa, ua); break;
        case 1: hipLaunchKernelGGL((k_repair_weno<1>), grid, dim3(BLOCK), 0, s, fa, ua); break;
        default: hipLaunchKernelGGL((k_repair_weno<2>), grid, dim3(BLOCK), 0, s, fa, ua); break;
    }
    return hipGetLastError();
}

}  // namespace afd
