// afd_launch.hpp -- host-side launchers of the gfx950 kernels (one per kernel family).
#pragma once
#include <hip/hip_runtime.h>

#include "afd_kernels.hpp"

namespace afd {

// Arguments of a face-kernel launch in precision-neutral form.
struct FaceLaunch {
    const double *P;
    void *F, *Fcar, *zs;  // flux type: double (plain) or float (mixed)
    double *aux;
    Grid g;
    Phys<double> phd;
    Phys<float> phf;
    double p_max, v_max;
    int lim, recon, riemann, dir, diss;
    bool mixed;
    bool shared_edges;  // FP32 RMHD WENO5: the shared-edge kernels (afd_params.face_path == 0)
    float *D;           // FP32 RMHD WENO5 with shared edges: the flux-divergence accumulator (else null)
    bool dfirst;        // ... written by this launch (the stage's first face kernel), not added to
    float idn;          // 1 / spacing along dir
};

template <typename T> inline FaceArgs<T> face_args(const FaceLaunch &l) {
    FaceArgs<T> a;
    a.P = l.P;
    a.F = (T *)l.F;
    a.Fcar = (T *)l.Fcar;
    a.zs = (T *)l.zs;
    a.aux = l.aux;
    a.g = l.g;
    a.p_max = l.p_max;
    a.v_max = l.v_max;
    a.lim = l.lim;
    a.diss = l.diss;
    a.D = l.D;
    a.idn = l.idn;
    if constexpr (sizeof(T) == 8) a.ph = l.phd;
    else a.ph = l.phf;
    const int *r0 = l.g.r0, *r1 = l.g.r1;
    a.dk = make_udiv((unsigned)(r1[2] - r0[2] + (l.dir == 2 ? 1 : 0)));
    a.dn = make_udiv((unsigned)(l.dir == 0 ? r1[0] - r0[0] + 1 : r1[1] - r0[1] + (l.dir == 1 ? 1 : 0)));
    return a;
}

// Number of faces (threads) of the face box of direction dir.
inline long face_count(const Grid &g, int dir) {
    long n = 1;
    for (int a = 0; a < 3; ++a) n *= (long)(g.r1[a] - g.r0[a] + (a == dir ? 1 : 0));
    return n;
}

hipError_t launch_face_srhd(const FaceLaunch &l, hipStream_t s);
hipError_t launch_face_rmhd(const FaceLaunch &l, hipStream_t s);

// Stage update / rhs-only launch.  flux pointers typed by `mixed`.
struct UpdLaunch {
    UpdArgs<double> a;  // flux pointers reinterpreted as float when mixed
    bool mixed;
    int nb;
};
hipError_t launch_update(const UpdLaunch &l, bool rhs_only, hipStream_t s);

// Mixed RMHD WENO5 on the flux-divergence path: the y faces fused with the stage update
// (k_face_march_update), the ghost shell's update and the learned start's fallback queue.
hipError_t launch_stage_fused(const FaceLaunch &l, const UpdLaunch &u, hipStream_t s);
// ... and the FP64 repair of the cells whose FP32 right-hand side was not finite (k_repair_weno;
// l: the stage's faces, its primitives)
hipError_t launch_repair(const FaceLaunch &l, const UpdLaunch &u, hipStream_t s);
long shell_cells(const Grid &g);
long fused_blocks(const Grid &g);


}  // namespace afd
