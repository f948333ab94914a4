// afd_rmhd.hip -- RMHD face-kernel instantiations (MUSCL/PPM/WENO x HLLE/hlld/hlld_full x x,y,z).
// The precision is selected at build time by AFD_RMHD_T (double or float) so the two
// halves compile as separate translation units in parallel.
#include "afd_launch.hpp"

#ifndef AFD_RMHD_T
#define AFD_RMHD_T double
#endif

namespace afd {

// FP32 WENO5 z faces: one lane per cell, the left neighbour's right edge from lane - 1.
template <int RIEM> static void go_weno_z_f32(const FaceLaunch &l, hipStream_t s) {
    FaceArgs<float> a = face_args<float>(l);
    const Grid &g = l.g;
    const unsigned nz = g.r1[2] - g.r0[2], ny = g.r1[1] - g.r0[1], nx = g.r1[0] - g.r0[0];
    const unsigned ncz = nz + 2, npos = ncz * ny * nx, waves = (npos - 1 + 62) / 63;
    a.dk = make_udiv(ncz);
    a.dn = make_udiv(ny);
    hipLaunchKernelGGL((k_face_rmhd_weno_lanes<RIEM>), grid1d(64L * waves), dim3(BLOCK), 0, s, a);
}

// FP32 WENO5 x and y faces: workgroups of ROW_WAVES cells along the normal sharing edges through LDS.
template <int RIEM> static void go_weno_xy_f32(const FaceLaunch &l, hipStream_t s) {
    FaceArgs<float> a = face_args<float>(l);
    const Grid &g = l.g;
    const unsigned nz = g.r1[2] - g.r0[2], nkc = (nz + 63) / 64;
    const unsigned nn = g.r1[l.dir] - g.r0[l.dir] + 1, ng = (nn + ROW_WAVES - 2) / (ROW_WAVES - 1);
    const unsigned no = l.dir == 0 ? g.r1[1] - g.r0[1] : g.r1[0] - g.r0[0];
    a.dk = make_udiv(nkc);
    a.dn = make_udiv(ng);
    const dim3 grid(nkc * ng * no), block(64 * ROW_WAVES);
    if (l.dir == 0) hipLaunchKernelGGL((k_face_rmhd_weno_rows<RIEM, 0>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_face_rmhd_weno_rows<RIEM, 1>), grid, block, 0, s, a);
}

// FP32 WENO5 flux divergence, z: waves of 62 cells overlapping by two positions.
template <int RIEM> static void go_weno_z_div(const FaceLaunch &l, hipStream_t s) {
    FaceArgs<float> a = face_args<float>(l);
    const Grid &g = l.g;
    const unsigned nz = g.r1[2] - g.r0[2], ny = g.r1[1] - g.r0[1], nx = g.r1[0] - g.r0[0];
    const unsigned ncz = nz + 2, npos = ncz * ny * nx, waves = (npos - 1 + 61) / 62;
    a.dk = make_udiv(ncz);
    a.dn = make_udiv(ny);
    if (l.dfirst) hipLaunchKernelGGL((k_face_rmhd_weno_lanes_div<RIEM, false>), grid1d(64L * waves), dim3(BLOCK), 0, s, a);
    else hipLaunchKernelGGL((k_face_rmhd_weno_lanes_div<RIEM, true>), grid1d(64L * waves), dim3(BLOCK), 0, s, a);
}

// FP32 WENO5 flux divergence, x and y: four-wave pencils sweeping runs of MARCH_CELLS cells.
template <int RIEM> static void go_weno_xy_march(const FaceLaunch &l, hipStream_t s) {
    FaceArgs<float> a = face_args<float>(l);
    const Grid &g = l.g;
    const unsigned nz = g.r1[2] - g.r0[2], nkc = (nz + 63) / 64;
    const unsigned nruns = (g.r1[l.dir] - g.r0[l.dir] + MARCH_CELLS - 1) / MARCH_CELLS;
    const unsigned no = l.dir == 0 ? g.r1[1] - g.r0[1] : g.r1[0] - g.r0[0];
    a.dk = make_udiv(nkc);
    a.dn = make_udiv(nruns);
    const dim3 grid(nkc * nruns * no), block(256);
    // x always adds (the z kernel runs before it: its carried fluxes need the z-back speeds); y
    // writes on the flux-divergence path, where it runs first
    if (l.dir == 0) hipLaunchKernelGGL((k_face_rmhd_weno_march<RIEM, 0, true>), grid, block, 0, s, a);
    else if (l.dfirst) hipLaunchKernelGGL((k_face_rmhd_weno_march<RIEM, 1, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_face_rmhd_weno_march<RIEM, 1, true>), grid, block, 0, s, a);
}

template <int RECON, int RIEM, typename T> static void go(const FaceLaunch &l, hipStream_t s) {
    if constexpr (RECON == 2 && sizeof(T) == 4) {
        if (l.D) {  // the context decided: shared edges, below 2^29 cells
            if (l.dir == 2) go_weno_z_div<RIEM>(l, s);
            else go_weno_xy_march<RIEM>(l, s);
            return;
        }
        if (l.shared_edges && l.g.ncell < (1L << 29)) {
            if (l.dir == 2) go_weno_z_f32<RIEM>(l, s);
            else go_weno_xy_f32<RIEM>(l, s);
            return;
        }
    }
    FaceArgs<T> a = face_args<T>(l);
    dim3 grid = grid1d(face_count(l.g, l.dir));
    if (l.dir == 0) hipLaunchKernelGGL((k_face_rmhd<RECON, RIEM, 0, T>), grid, dim3(BLOCK), 0, s, a);
    else if (l.dir == 1) hipLaunchKernelGGL((k_face_rmhd<RECON, RIEM, 1, T>), grid, dim3(BLOCK), 0, s, a);
    else hipLaunchKernelGGL((k_face_rmhd<RECON, RIEM, 2, T>), grid, dim3(BLOCK), 0, s, a);
}

template <typename T> void launch_face_rmhd_t(const FaceLaunch &l, hipStream_t s) {
    int key = l.recon * 3 + (l.riemann < 0 ? 0 : (l.riemann > 2 ? 2 : l.riemann));
    switch (key) {
        case 0: go<0, 0, T>(l, s); break;
        case 1: go<0, 1, T>(l, s); break;
        case 2: go<0, 2, T>(l, s); break;
        case 3: go<1, 0, T>(l, s); break;
        case 4: go<1, 1, T>(l, s); break;
        case 5: go<1, 2, T>(l, s); break;
        case 6: go<2, 0, T>(l, s); break;
        case 7: go<2, 1, T>(l, s); break;
        default: go<2, 2, T>(l, s); break;
    }
}

template void launch_face_rmhd_t<AFD_RMHD_T>(const FaceLaunch &l, hipStream_t s);

}  // namespace afd
