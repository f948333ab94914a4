// afd_kernels.hpp -- kernel templates of the SRHD/RMHD finite-volume step (gfx950).
//
// Data layout in HBM: every field is the reference's C-order array
// (nv, nx+2ng, ny+2ng, nz+2ng) of float64, z fastest.  A stage of SSPRK runs as
//   face kernels, one per direction: one thread per face between cells c-1 and c,
//       lanes along z, writes the face flux at the index of cell c;
//   k_update: one thread per cell (ghosts included): flux differences, GLM and
//       resistive sources, the SSPRK combination in FP64 and FP64 con2prim.
// For RMHD PPM/WENO the x face of cell c is evaluated twice (Q3 of the
// reference): with its own HLL speeds (the right face of cell c-1) and with the
// speeds carried from the z-back face of the cell that precedes c in i->j->k order
// (the left face of cell c).  The z kernel runs first and stores those speeds.
#pragma once
#include "afd_device.hpp"
#include "afd_mlp_device.hpp"

#include <cstring>

namespace afd {

constexpr int BLOCK = 256;

// n / d for 0 <= n < 2^31 and a fixed 1 <= d < 2^31, as one multiply and one shift
// (Granlund & Montgomery, "Division by invariant integers using multiplication", PLDI 1994).
// With s = ceil(log2 d) and m = floor(2^(31+s) / d) + 1 (< 2^32), n / d = (n m) >> (31 + s):
// m d = 2^(31+s) + e with 0 < e <= d <= 2^s, so n m / 2^(31+s) = n / d + n e / (d 2^(31+s)),
// and n e < 2^(31+s) keeps the floor at n / d.
struct UDiv {
    unsigned m;
    int sh;
};
inline UDiv make_udiv(unsigned d) {
    int s = 0;
    while ((1ull << s) < d) ++s;
    return UDiv{(unsigned)((1ull << (31 + s)) / d + 1ull), 31 + s};
}
AFD_D unsigned udiv(const UDiv &q, unsigned n) { return (unsigned)(((unsigned long long)n * q.m) >> q.sh); }

struct Grid {
    int nv, nx, ny, nz;   // full dims (ghosts included)
    int r0[3], r1[3];     // cells that receive a flux divergence: [r0, r1) per axis
    long sx, sy, ncell;   // strides of x and y, number of cells (below 2^31)
    UDiv dnz, dny;        // division by nz and by ny
};

// Minimum waves per SIMD the face kernels are compiled for (the register budget follows:
// 3 waves <= 168 VGPRs).  Measured on the 256^3 RMHD WENO5 + HLLD step: FP64 faces gain ~3 %
// at 3 waves over the unconstrained 2; FP32 faces lose ~20 % at 4 (spills) against 3.
#ifndef AFD_FACE_WAVES_F64
#define AFD_FACE_WAVES_F64 3
#endif
#ifndef AFD_FACE_WAVES_F32
#define AFD_FACE_WAVES_F32 3
#endif
template <typename T> struct FaceOcc {
    static constexpr int waves = sizeof(T) == 8 ? AFD_FACE_WAVES_F64 : AFD_FACE_WAVES_F32;
};

inline dim3 grid1d(long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

AFD_D long cidx(const Grid &g, int i, int j, int k) { return ((long)i * g.ny + j) * g.nz + k; }
AFD_D void cell_ijk(const Grid &g, long c, int &i, int &j, int &k) {
    const unsigned r = udiv(g.dnz, (unsigned)c), q = udiv(g.dny, r);
    k = (int)((unsigned)c - r * (unsigned)g.nz);
    j = (int)(r - q * (unsigned)g.ny);
    i = (int)q;
}

template <typename T> struct FaceArgs {
    const double *P;   // primitives of the stage input
    T *F;              // face flux (own speeds), nb x ncell
    T *Fcar;           // RMHD PPM/WENO x only: flux with carried speeds
    T *zs;             // RMHD PPM/WENO: z-back face speeds (sL, sR) per cell, 2 x ncell
    double *aux;       // passives: (sL, sR, vL, vR) per face, 4 x ncell (or null)
    Grid g;
    Phys<T> ph;
    double p_max, v_max;
    int lim;
    int diss;          // SRHD dissipative hooks: bit 0 bulk pressure on the face pressures, bits 1-2 flux repeats
    UDiv dk, dn;       // face box: division by its z extent, then by its x (DIR 0) or y extent
    float *D;          // flux-divergence kernels: sum over directions of (F_right - F_left) / d, nb x ncell
    float idn;         //   1 / spacing along the face normal
};

// ------------------------------------------------------------------ reconstruction
// Face states of the face between cells c-1 and c along stride s, variable base pointer b.
// The mixed path reconstructs in FP32 each edge as a deviation from its own cell's value (the
// left state from q[c-1], the right state from q[c]) and adds it back in FP64, so that across a
// large jump neither state is the small difference of two large values.  Returns the smaller
// WENO5 weight sum of the FP32 path (1 otherwise); below WENO_SUM_MIN the caller evaluates the
// face again with RESCUE set (the rescaled FP32 WENO5).
template <int RECON, typename T, bool RESCUE = false>
AFD_D float recon_face(const double *__restrict__ b, long c, long s, int lim, double &qL, double &qR) {
    constexpr bool mixed = sizeof(T) == 4;
    const double rL = mixed ? b[c - s] : 0.0, rR = mixed ? b[c] : 0.0;
    float smin = 1.0f;
    T l, r;
    if (RECON == 0) {  // MUSCL: core/srhd_core.py:486-533
        const double bm2 = b[c - 2 * s], bm1 = b[c - s], b0 = b[c], bp1 = b[c + s];
        {  // left state: cell c-1 and its neighbours
            T qm2 = (T)(bm2 - rL), qm1 = (T)(bm1 - rL), q0 = (T)(b0 - rL);
            T slL = limiter<T>(lim, qm1 - qm2, q0 - qm1);
            l = qm1 + T(0.5) * slL;
        }
        {  // right state: cell c and its neighbours
            T qm1 = (T)(bm1 - rR), q0 = (T)(b0 - rR), qp1 = (T)(bp1 - rR);
            T slR = limiter<T>(lim, q0 - qm1, qp1 - q0);
            r = q0 - T(0.5) * slR;
        }
    } else {
        const double bm3 = b[c - 3 * s], bm2 = b[c - 2 * s], bm1 = b[c - s], b0 = b[c], bp1 = b[c + s],
                     bp2 = b[c + 2 * s];
        T lm3 = (T)(bm3 - rL), lm2 = (T)(bm2 - rL), lm1 = (T)(bm1 - rL), l0 = (T)(b0 - rL), lp1 = (T)(bp1 - rL);
        T um2 = (T)(bm2 - rR), um1 = (T)(bm1 - rR), u0 = (T)(b0 - rR), up1 = (T)(bp1 - rR), up2 = (T)(bp2 - rR);
        if (RECON == 1) {  // PPM: core/srhd_core.py:943-950
            T a, bb;
            ppm_reconstruct<T>(lm3, lm2, lm1, l0, lp1, a, l);
            ppm_reconstruct<T>(um2, um1, u0, up1, up2, r, bb);
        } else if constexpr (!mixed) {  // WENO5: core/srhd_core.py:1183-1184
            l = weno5_left<T>(lm3, lm2, lm1, l0, lp1);
            r = weno5_right<T>(up2, up1, u0, um1, um2);
        } else if constexpr (RESCUE) {
            l = weno5_left_f32_rescaled(lm3, lm2, lm1, l0, lp1);
            r = weno5_left_f32_rescaled(up2, up1, u0, um1, um2);
        } else {
            float s1, s2;
            l = weno5_left_f32_eps(lm3, lm2, lm1, l0, lp1, 1e-12f, s1);
            r = weno5_left_f32_eps(up2, up1, u0, um1, um2, 1e-12f, s2);
            smin = fminf(s1, s2);
        }
    }
    if (mixed) {
        qL = rL + (double)l;
        qR = rR + (double)r;
    } else {
        qL = (double)l;
        qR = (double)r;
    }
    return smin;
}

// The face states of NV variables (fields at b + v * n); the FP32 WENO5 lanes whose weights
// underflow evaluate again with the rescaled WENO5.
template <int RECON, typename T, int NV>
AFD_D void recon_face_all(const double *__restrict__ b, long n, long c, long s, int lim, double *qL, double *qR) {
    float smin = 1.0f;
#pragma unroll
    for (int v = 0; v < NV; ++v) smin = fminf(smin, recon_face<RECON, T>(b + v * n, c, s, lim, qL[v], qR[v]));
    if constexpr (RECON == 2 && sizeof(T) == 4) {
        if (__builtin_expect(!(smin >= WENO_SUM_MIN), 0)) {
            for (int v = 0; v < NV; ++v) recon_face<RECON, T, true>(b + v * n, c, s, lim, qL[v], qR[v]);
        }
    }
}

template <int DIR> AFD_D long dstride(const Grid &g) { return DIR == 0 ? g.sx : (DIR == 1 ? g.sy : 1L); }

// Decompose a linear thread index over the face box of direction DIR
// (faces [r0, r1] along DIR, cells [r0, r1) across it), z fastest.
// Workgroup b of a launch runs on XCD b % 8 (the round-robin dispatch of the 8-XCD MI355X).
// This maps b to a logical block so that each XCD walks one contiguous eighth of the grid in
// order: rows j and j +- 1 of a plane then share an XCD and its L2.  A bijection on [0, nb).
AFD_D long xcd_chunked_block(long b, long nb) {
    const long q = nb / 8, r = nb % 8, x = b % 8, idx = b / 8;
    return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Thread index of a face kernel.  The x and y faces' stencils span six rows (k lines) of the
// grid along the face normal.  Faces are ordered k fastest, then along the normal, so that the
// rows a stencil shares are consecutive blocks, and the blocks walk XCD-contiguous chunks so
// that those rows share an XCD and its L2.  (With the plain order the y stencil's rows landed on
// six XCDs and the y kernels fetched 3x their compulsory bytes from HBM at 256^3.)  The z
// stencil lies within a row.
template <int DIR> AFD_D unsigned face_thread() {
    const unsigned b = DIR != 2 ? (unsigned)xcd_chunked_block(blockIdx.x, gridDim.x) : blockIdx.x;
    return b * BLOCK + threadIdx.x;
}

// Face (i, j, k) of thread t; a face box has fewer faces than the grid has cells (< 2^31).
template <int DIR> AFD_D bool face_coords(const Grid &g, const UDiv &dk, const UDiv &dn, unsigned t, int &i, int &j,
                                          int &k) {
    int n[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) n[a] = g.r1[a] - g.r0[a] + (a == DIR ? 1 : 0);
    const unsigned tot = (unsigned)n[0] * n[1] * n[2];
    if (t >= tot) return false;
    const unsigned r = udiv(dk, t), q = udiv(dn, r);
    k = (int)(t - r * n[2]);
    if (DIR == 0) {  // k, then i (the normal), then j
        i = (int)(r - q * n[0]);
        j = (int)q;
    } else {         // k, then j, then i
        j = (int)(r - q * n[1]);
        i = (int)q;
    }
    i += g.r0[0];
    j += g.r0[1];
    k += g.r0[2];
    return true;
}

// ------------------------------------------------------------------ SRHD face
// apply_diss_flux_x/y/z (core/srhd_core.py:401-462): Pi, pi^ij and q^i of `cell` (fields 5..14)
// added to the physical flux F of a face state q (its velocities).
template <int DIR, typename T>
AFD_D void diss_flux(const double *__restrict__ P, long n, long cell, const T *q, T *F) {
    const T pi = (T)P[5 * n + cell], pxx = (T)P[6 * n + cell], pyy = (T)P[7 * n + cell], pzz = (T)P[8 * n + cell];
    const T pxy = (T)P[9 * n + cell], pxz = (T)P[10 * n + cell], pyz = (T)P[11 * n + cell];
    const T qx = (T)P[12 * n + cell], qy = (T)P[13 * n + cell], qz = (T)P[14 * n + cell];
    const T vx = q[1], vy = q[2], vz = q[3];
    if (DIR == 0) {
        F[1] += pi + pxx;
        F[2] += pxy;
        F[3] += pxz;
        F[4] += vx * (pi + pxx) + vy * pxy + vz * pxz + qx;
    } else if (DIR == 1) {
        F[1] += pxy;
        F[2] += pi + pyy;
        F[3] += pyz;
        F[4] += vx * pxy + vy * (pi + pyy) + vz * pyz + qy;
    } else {
        F[1] += pxz;
        F[2] += pyz;
        F[3] += pi + pzz;
        F[4] += vx * pxz + vy * pyz + vz * (pi + pzz) + qz;
    }
}

template <int RECON, int RIEM, int DIR, typename T>
__global__ __launch_bounds__(BLOCK) void k_face_srhd(FaceArgs<T> a) {
    const Grid &g = a.g;
    int i, j, k;
    if (!face_coords<DIR>(g, a.dk, a.dn, face_thread<DIR>(), i, j, k)) return;
    const long c = cidx(g, i, j, k), s = dstride<DIR>(g);
    double qLd[5], qRd[5];
    recon_face_all<RECON, T, 5>(a.P, g.ncell, c, s, a.lim, qLd, qRd);
    if (a.diss & 1) {  // add_pi_lr, core/srhd_core.py:394-399 (MUSCL faces, PASSIVE_OFFSET > 5)
        qLd[4] += a.P[5 * g.ncell + c - s];
        qRd[4] += a.P[5 * g.ncell + c];
    }
    srhd_floor(a.p_max, a.v_max, qLd);
    srhd_floor(a.p_max, a.v_max, qRd);
    T WL = (T)lorentz_guarded<sizeof(T) == 4>(qLd[1] * qLd[1] + qLd[2] * qLd[2] + qLd[3] * qLd[3]);
    T WR = (T)lorentz_guarded<sizeof(T) == 4>(qRd[1] * qRd[1] + qRd[2] * qRd[2] + qRd[3] * qRd[3]);
    T qL[5], qR[5], UL[5], UR[5], FL[5], FR[5], F[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        qL[v] = (T)qLd[v];
        qR[v] = (T)qRd[v];
    }
    srhd_p2c_W(a.ph, qL, WL, UL);
    srhd_p2c_W(a.ph, qR, WR, UR);
    srhd_flux<DIR>(qL, UL, FL);
    srhd_flux<DIR>(qR, UR, FR);
    for (int r = 0; r < (a.diss >> 1); ++r) {  // apply_diss_flux: once per MUSCL face, twice per PPM face
        diss_flux<DIR>(a.P, g.ncell, c - s, qL, FL);
        diss_flux<DIR>(a.P, g.ncell, c, qR, FR);
    }
    T csL = srhd_sound_speed(a.ph, qL[0], qL[4]), csR = srhd_sound_speed(a.ph, qR[0], qR[4]);
    T lmL, lpL, lmR, lpR;
    eig_speeds(qL[1 + DIR], csL, lmL, lpL);
    eig_speeds(qR[1 + DIR], csR, lmR, lpR);
    T sL = pymin(lmL, lmR), sR = pymax(lpL, lpR);
    if (RIEM == 1)
        hllc5<1 + DIR>(UL, UR, FL, FR, sL, sR, qL[4], qR[4], qL[1 + DIR], qR[1 + DIR], F);
    else
        hlle_vec<5>(UL, UR, FL, FR, sL, sR, F);
#pragma unroll
    for (int v = 0; v < 5; ++v) a.F[v * g.ncell + c] = F[v];
    if (a.aux) {
        a.aux[c] = sL;
        a.aux[g.ncell + c] = sR;
        a.aux[2 * g.ncell + c] = qL[1 + DIR];
        a.aux[3 * g.ncell + c] = qR[1 + DIR];
    }
}

// ------------------------------------------------------------------ RMHD face
template <int DIR, typename T> struct RmhdFace {
    T UL[9], UR[9], FL[9], FR[9];
    T pL, pR, vL, vR, sL, sR;
};

// The nine FP64 primitives of one face state.
struct State9 {
    double q[9];
};

// Face fluxes and speeds from the reconstructed FP64 face states.  The states come by value: with
// them passed as pointers into the caller's arrays, the FP64 per-face kernels spilled 76-120 bytes
// per lane at 3 waves per SIMD.
template <int DIR, typename T>
AFD_D void rmhd_face_states(const FaceArgs<T> &a, State9 L, State9 R, RmhdFace<DIR, T> &f) {
    double *qLd = L.q, *qRd = R.q;
    rmhd_floor(a.p_max, a.v_max, qLd);
    rmhd_floor(a.p_max, a.v_max, qRd);
    double vcL[3], vcR[3];
    constexpr bool FAST = sizeof(T) == 4;  // mixed path: FP64 Lorentz factors with the fast rsq
    double WLd = rmhd_cap_W<FAST>(a.v_max, qLd, vcL), WRd = rmhd_cap_W<FAST>(a.v_max, qRd, vcR);
    // rmhd_fast_speed's own Lorentz factor differs only when the floored v2 still reaches V_MAX^2
    double WfL = WLd, WfR = WRd;
    if (qLd[1] * qLd[1] + qLd[2] * qLd[2] + qLd[3] * qLd[3] >= a.v_max * a.v_max) WfL = rmhd_fast_W<FAST>(a.v_max, qLd);
    if (qRd[1] * qRd[1] + qRd[2] * qRd[2] + qRd[3] * qRd[3] >= a.v_max * a.v_max) WfR = rmhd_fast_W<FAST>(a.v_max, qRd);
    T qL[9], qR[9], vL3[3], vR3[3];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        qL[v] = (T)qLd[v];
        qR[v] = (T)qRd[v];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        vL3[m] = (T)vcL[m];
        vR3[m] = (T)vcR[m];
    }
    rmhd_p2c_W(a.ph, qL, vL3, (T)WLd, f.UL);
    rmhd_p2c_W(a.ph, qR, vR3, (T)WRd, f.UR);
    rmhd_flux<DIR>(a.ph, qL, f.UL, f.FL);
    rmhd_flux<DIR>(a.ph, qR, f.UR, f.FR);
    T cfL = rmhd_fast_speed_W(a.ph, qL, (T)WfL), cfR = rmhd_fast_speed_W(a.ph, qR, (T)WfR);
    T lmL, lpL, lmR, lpR;
    eig_speeds(qL[1 + DIR], cfL, lmL, lpL);
    eig_speeds(qR[1 + DIR], cfR, lmR, lpR);
    f.sL = pymin(lmL, lmR);
    f.sR = pymax(lpL, lpR);
    f.pL = qL[4];
    f.pR = qR[4];
    f.vL = qL[1 + DIR];
    f.vR = qR[1 + DIR];
}

template <int RECON, int DIR, typename T>
AFD_D void rmhd_face_setup(const FaceArgs<T> &a, long c, RmhdFace<DIR, T> &f) {
    const Grid &g = a.g;
    const long s = dstride<DIR>(g);
    State9 L, R;
    recon_face_all<RECON, T, 9>(a.P, g.ncell, c, s, a.lim, L.q, R.q);
    rmhd_face_states<DIR, T>(a, L, R, f);
}

// core/rmhd_core.py:569-578, 726-746 and the MUSCL-only hlld override (:811-831)
template <int RECON, int RIEM, int DIR, typename T>
AFD_D void rmhd_riemann(const Phys<T> &ph, const RmhdFace<DIR, T> &f, T sL, T sR, T *out) {
    if (RIEM == 2) {
        hlld_full<DIR>(ph, f.UL, f.UR, f.FL, f.FR, sL, sR, f.pL, f.pR, f.vL, f.vR, out);
    } else if (RIEM == 1) {
        hlle_vec<9>(f.UL, f.UR, f.FL, f.FR, sL, sR, out);
        hllc5<1 + DIR>(f.UL, f.UR, f.FL, f.FR, sL, sR, f.pL, f.pR, f.vL, f.vR, out);
        if (RECON == 0) {
            T sM = hllc_contact_speed<1 + DIR>(f.UL, f.UR, sL, sR, f.pL, f.pR, f.vL, f.vR);
            const T *Fm = (sM >= T(0)) ? f.FL : f.FR;
#pragma unroll
            for (int m = 5; m < 9; ++m) out[m] = Fm[m];
        }
    } else {
        hlle_vec<9>(f.UL, f.UR, f.FL, f.FR, sL, sR, out);
    }
}

// Riemann flux(es) of face (i, j, k) (between cells c - s and c) and the stored speeds.
template <int RECON, int RIEM, int DIR, typename T>
AFD_D void rmhd_face_store(const FaceArgs<T> &a, const RmhdFace<DIR, T> &f, int i, int j, int k, long c) {
    const Grid &g = a.g;
    T out[9];
    constexpr bool carried = (RECON != 0) && (DIR == 0);
    if (!carried || i > g.r0[0]) {
        rmhd_riemann<RECON, RIEM, DIR, T>(a.ph, f, f.sL, f.sR, out);
#pragma unroll
        for (int v = 0; v < 9; ++v) a.F[v * g.ncell + c] = out[v];
    }
    if (carried && i < g.r1[0]) {
        // predecessor of cell (i,j,k) in the reference's i->j->k loop (core/rmhd_core.py:1447-1449)
        T cL = T(-1), cR = T(1);
        long pc = -1;
        if (k > g.r0[2]) pc = c - 1;
        else if (j > g.r0[1]) pc = cidx(g, i, j - 1, g.r1[2] - 1);
        else if (i > g.r0[0]) pc = cidx(g, i - 1, g.r1[1] - 1, g.r1[2] - 1);
        if (pc >= 0) {
            cL = a.zs[pc];
            cR = a.zs[g.ncell + pc];
        }
        rmhd_riemann<RECON, RIEM, DIR, T>(a.ph, f, cL, cR, out);
#pragma unroll
        for (int v = 0; v < 9; ++v) a.Fcar[v * g.ncell + c] = out[v];
    }
    if (DIR == 2 && a.zs) {
        a.zs[c] = f.sL;
        a.zs[g.ncell + c] = f.sR;
    }
    if (a.aux) {
        a.aux[c] = f.sL;
        a.aux[g.ncell + c] = f.sR;
        a.aux[2 * g.ncell + c] = f.vL;
        a.aux[3 * g.ncell + c] = f.vR;
    }
}

template <int RECON, int RIEM, int DIR, typename T>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(FaceOcc<T>::waves)))
void k_face_rmhd(FaceArgs<T> a) {
    const Grid &g = a.g;
    int i, j, k;
    if (!face_coords<DIR>(g, a.dk, a.dn, face_thread<DIR>(), i, j, k)) return;
    const long c = cidx(g, i, j, k);
    RmhdFace<DIR, T> f;
    rmhd_face_setup<RECON, DIR, T>(a, c, f);
    rmhd_face_store<RECON, RIEM, DIR, T>(a, f, i, j, k, c);
}

// ------------------------------------------------------------------ RMHD WENO5 faces, FP32 path
// The mixed path's WENO5 z faces evaluate each cell's two edge values together
// (weno5_cell_pair), as deviations from the cell's own value.  Face (c-1 | c) takes the right
// edge of cell c-1 and the left edge of cell c.  Used for grids below 2^29 cells (32-bit byte
// offsets within a field); larger grids take the per-face kernel.  (A thread walking a run of
// faces along the normal, carrying the right edge, ran 10-15 % slower on the 256^3 RMHD step
// despite 20-28 % fewer VALU instructions per face.  The PPM and MUSCL faces keep the per-face
// kernels: with one cheap reconstruction per cell to share, the same LDS rows and lane shuffles
// ran 4-20 % slower per direction there.)

// Element e of a field whose byte offsets fit in 32 bits: the load takes the field's base from
// scalar registers and one 32-bit offset, shared by the nine fields.
AFD_D double ld32(const double *__restrict__ b, unsigned e) { return *(const double *)((const char *)b + e * 8u); }

// Edge values of cell x of one variable (base b) along stride s (grids below 2^29 cells), as
// deviations l, h from the cell's value v0; vm1 is the value of cell x - s.  Returns the smaller
// WENO5 weight sum (below WENO_SUM_MIN: weno5_cell_edges_rescaled gives the edges).
AFD_D float weno5_cell_edges(const double *__restrict__ b, unsigned x, unsigned s, double &vm1, double &v0, float &l,
                             float &h) {
    v0 = ld32(b, x);
    vm1 = ld32(b, x - s);
    return weno5_cell_pair((float)(ld32(b, x - 2 * s) - v0), (float)(vm1 - v0), (float)(ld32(b, x + s) - v0),
                           (float)(ld32(b, x + 2 * s) - v0), l, h);
}

AFD_D void weno5_cell_edges_rescaled(const double *__restrict__ b, unsigned x, unsigned s, float &l, float &h) {
    const double v0 = ld32(b, x);
    weno5_cell_pair_rescaled((float)(ld32(b, x - 2 * s) - v0), (float)(ld32(b, x - s) - v0),
                             (float)(ld32(b, x + s) - v0), (float)(ld32(b, x + 2 * s) - v0), l, h);
}

// Edge values of the nine variables of cell x: L.q = the values of cell x - s, R.q = those of
// cell x, lo / hi = the cell's left / right edge deviations.  Lanes whose WENO5 weights
// underflow evaluate again with the rescaled WENO5.
AFD_D void weno5_cell_edges9(const double *__restrict__ P, unsigned n, unsigned x, unsigned s, State9 &L, State9 &R,
                             float *lo, float *hi) {
    float smin = 1.0f;
#pragma unroll
    for (int v = 0; v < 9; ++v) smin = fminf(smin, weno5_cell_edges(P + v * n, x, s, L.q[v], R.q[v], lo[v], hi[v]));
    if (__builtin_expect(!(smin >= WENO_SUM_MIN), 0)) {
        for (int v = 0; v < 9; ++v) weno5_cell_edges_rescaled(P + v * n, x, s, lo[v], hi[v]);
    }
}

// x and y faces: a workgroup of ROW_WAVES waves covers 64 consecutive z positions of
// ROW_WAVES consecutive cells along the normal, n0 - 1 .. n0 + ROW_WAVES - 2.  Each lane evaluates
// its cell's two edges; the right edges go through LDS to the wave of the next cell, and waves
// 1 .. ROW_WAVES - 1 evaluate the faces n0 .. n0 + ROW_WAVES - 2.  Wave 0 only supplies cell
// n0 - 1.  Workgroups are ordered z chunk fastest, then group of faces along the normal, then the
// other transverse axis, and walk XCD-contiguous chunks.  a.dk divides by the number of z chunks,
// a.dn by the number of face groups per line.
#ifndef AFD_ROW_WAVES
#define AFD_ROW_WAVES 4
#endif
constexpr int ROW_WAVES = AFD_ROW_WAVES;

template <int RIEM, int DIR>
__global__ __launch_bounds__(64 * ROW_WAVES) __attribute__((amdgpu_waves_per_eu(FaceOcc<float>::waves)))
void k_face_rmhd_weno_rows(FaceArgs<float> a) {
    static_assert(DIR == 0 || DIR == 1, "rows of cells along x or y");
    constexpr int FG = ROW_WAVES - 1;  // faces per group
    __shared__ float edge[ROW_WAVES][9][64];
    const Grid &g = a.g;
    constexpr int O = DIR == 0 ? 1 : 0;  // the transverse axis other than z
    const int n2 = g.r1[2] - g.r0[2], nkc = (n2 + 63) / 64, nn = g.r1[DIR] - g.r0[DIR] + 1, ng = (nn + FG - 1) / FG;
    const unsigned b = (unsigned)xcd_chunked_block(blockIdx.x, gridDim.x);
    const unsigned r = udiv(a.dk, b), q = udiv(a.dn, r);
    const int kc = (int)(b - r * nkc), grp = (int)(r - q * ng), o = g.r0[O] + (int)q;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kraw = kc * 64 + lane, n = g.r0[DIR] + FG * grp + w - 1;
    const bool live = kraw < n2 && n <= g.r1[DIR];
    const int k = g.r0[2] + min(kraw, n2 - 1), nc = min(n, g.r1[DIR]);  // in range for the loads
    const unsigned s = (unsigned)dstride<DIR>(g);
    const unsigned c = (unsigned)(DIR == 0 ? cidx(g, nc, o, k) : cidx(g, o, nc, k));
    State9 L, R;
    float lo[9], hi[9];
    weno5_cell_edges9(a.P, (unsigned)g.ncell, c, s, L, R, lo, hi);
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        edge[w][v][lane] = hi[v];
        R.q[v] += (double)lo[v];
    }
    __syncthreads();
    if (w == 0 || !live) return;
#pragma unroll
    for (int v = 0; v < 9; ++v) L.q[v] += (double)edge[w - 1][v][lane];
    RmhdFace<DIR, float> fc;
    rmhd_face_states<DIR, float>(a, L, R, fc);
    rmhd_face_store<2, RIEM, DIR, float>(a, fc, DIR == 0 ? n : o, DIR == 0 ? o : n, k, (long)c);
}

// z faces: one lane per cell k = r0 - 1 .. r1 of a z line (ncz = r1 - r0 + 2 cells), the
// lines of the interior i, j range in order; lane l takes the right edge of cell k - 1 from
// lane l - 1.  Wave w covers line positions 63 w .. 63 w + 63 and writes the faces of lanes 1..63,
// so consecutive waves overlap by one cell.  a.dk divides by ncz, a.dn by the y extent.
template <int RIEM>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(FaceOcc<float>::waves)))
void k_face_rmhd_weno_lanes(FaceArgs<float> a) {
    const Grid &g = a.g;
    const int ncz = g.r1[2] - g.r0[2] + 2, ny = g.r1[1] - g.r0[1];
    const unsigned npos = (unsigned)ncz * ny * (g.r1[0] - g.r0[0]);
    const unsigned lane = threadIdx.x & 63u, w = (blockIdx.x * BLOCK + threadIdx.x) >> 6;
    if (63u * w + 1u >= npos) return;  // whole waves past the last face
    unsigned pos = 63u * w + lane;
    const bool valid = pos < npos;
    if (!valid) pos = npos - 1;  // keeps the lane's loads in range; it writes nothing
    const unsigned line = udiv(a.dk, pos), kk = pos - line * ncz, ii = udiv(a.dn, line);
    const int i = g.r0[0] + (int)ii, j = g.r0[1] + (int)(line - ii * ny), k = g.r0[2] - 1 + (int)kk;
    const unsigned c = (unsigned)cidx(g, i, j, k);
    State9 L, R;
    float lo[9], hi[9];
    weno5_cell_edges9(a.P, (unsigned)g.ncell, c, 1u, L, R, lo, hi);
#pragma unroll
    for (int v = 0; v < 9; ++v) {  // every lane takes part in the shuffles
        L.q[v] += (double)__shfl_up(hi[v], 1);
        R.q[v] += (double)lo[v];
    }
    if (!valid || lane == 0 || kk == 0) return;
    RmhdFace<2, float> fc;
    rmhd_face_states<2, float>(a, L, R, fc);
    rmhd_face_store<2, RIEM, 2, float>(a, fc, i, j, k, c);
}

// ------------------------------------------------------------------ RMHD WENO5 flux divergence, FP32 path
// The mixed path's shared-edge WENO5 kernels accumulate each cell's flux divergence
// D = sum_d (F(right face) - F(left face)) / d_d in FP32 (nb x ncell) instead of storing face
// fluxes: the first kernel of a stage writes D (ADD false), the others add their parts, and the
// stage update reads one FP32 array (36 B a cell) where it read four flux arrays (144 B).  The
// order is y, z, x (the y pencils write, so their rounds wait on no read of D; the z lanes,
// independent waves, hide the read's latency), or z, x when the stage update evaluates the y
// faces.  The z kernel also stores the z-back speeds the x kernel's carried fluxes need.

// The Riemann flux(es) of x or y face (i, j, k) (between cells c - s and c): Fown with the face's
// own speeds (the right face of cell c - s; wanted when own), Fcar with the speeds carried from
// the z-back face of the predecessor of cell c in the reference's i->j->k loop (x only; the left
// face of cell c; wanted when car), as rmhd_face_store computes them.
template <int RIEM, int DIR>
AFD_D void rmhd_face_pair(const FaceArgs<float> &a, const RmhdFace<DIR, float> &f, int i, int j, int k, long c,
                          bool own, bool car, float *Fown, float *Fcar) {
    const Grid &g = a.g;
    if (own) rmhd_riemann<2, RIEM, DIR, float>(a.ph, f, f.sL, f.sR, Fown);
    if (DIR == 0 && car) {  // reached for DIR 0 only
        float cL = -1.0f, cR = 1.0f;  // hlle_bounds_fast, core/rmhd_core.py:397-400
        long pc = -1;
        if (k > g.r0[2]) pc = c - 1;
        else if (j > g.r0[1]) pc = cidx(g, i, j - 1, g.r1[2] - 1);
        else if (i > g.r0[0]) pc = cidx(g, i - 1, g.r1[1] - 1, g.r1[2] - 1);
        if (pc >= 0) {
            cL = a.zs[pc];
            cR = a.zs[g.ncell + pc];
        }
        rmhd_riemann<2, RIEM, DIR, float>(a.ph, f, cL, cR, Fcar);
    }
}

// z faces: as k_face_rmhd_weno_lanes, with waves overlapping by two positions (wave w covers line
// positions 62 w .. 62 w + 63): lane l's cell takes its right face from lane l + 1, so lanes
// 1..62 write D for their cells (k = r0 .. r1 - 1 of a line) and the z-back speeds of their faces.
// Every lane evaluates a face (lanes without one use in-range duplicates and write nothing), so
// that every lane takes part in the shuffles.
#ifndef AFD_LANES_DIV_WAVES
#define AFD_LANES_DIV_WAVES AFD_FACE_WAVES_F32
#endif


template <int RIEM, bool ADD>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(AFD_LANES_DIV_WAVES)))
void k_face_rmhd_weno_lanes_div(FaceArgs<float> a) {
    const Grid &g = a.g;
    const int ncz = g.r1[2] - g.r0[2] + 2, ny = g.r1[1] - g.r0[1];
    const unsigned npos = (unsigned)ncz * ny * (g.r1[0] - g.r0[0]);
    const unsigned lane = threadIdx.x & 63u, w = (blockIdx.x * BLOCK + threadIdx.x) >> 6;
    if (62u * w + 1u >= npos) return;  // whole waves past the last cell
    unsigned pos = 62u * w + lane;
    const bool valid = pos < npos;
    if (!valid) pos = npos - 1;  // keeps the lane's loads in range; it writes nothing
    const unsigned line = udiv(a.dk, pos), kk = pos - line * ncz, ii = udiv(a.dn, line);
    const int i = g.r0[0] + (int)ii, j = g.r0[1] + (int)(line - ii * ny), k = g.r0[2] - 1 + (int)kk;
    const unsigned c = (unsigned)cidx(g, i, j, k), n = (unsigned)g.ncell;
    float dold[9];  // ADD: the cell's divergence so far, loaded with the stencil (c is in range)
#pragma unroll
    for (int v = 0; v < 9; ++v) dold[v] = ADD ? a.D[v * n + c] : 0.0f;
    State9 L, R;
    float lo[9], hi[9];
    weno5_cell_edges9(a.P, n, c, 1u, L, R, lo, hi);
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        L.q[v] += (double)__shfl_up(hi[v], 1);
        R.q[v] += (double)lo[v];
    }
    RmhdFace<2, float> fc;
    rmhd_face_states<2, float>(a, L, R, fc);
    float F[9];
    rmhd_riemann<2, RIEM, 2, float>(a.ph, fc, fc.sL, fc.sR, F);
    float Fr[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) Fr[v] = __shfl_down(F[v], 1);
    if (!valid || lane == 0u || lane == 63u || kk == 0u) return;
    a.zs[c] = fc.sL;
    a.zs[n + c] = fc.sR;
    if (kk + 2u > (unsigned)ncz) return;  // the face k = r1 closes a line: no cell of its own
#pragma unroll
    for (int v = 0; v < 9; ++v) a.D[v * n + c] = ADD ? dold[v] + (Fr[v] - F[v]) * a.idn : (Fr[v] - F[v]) * a.idn;
}

// x and y faces as pencils: a workgroup of four waves covers 64 consecutive z positions of one
// transverse line and sweeps a run of MARCH_CELLS cells m0 .. m1 - 1 along the normal.  Round r:
// wave w evaluates the edges of cell e = m0 - 1 + 4 r + w and the face (e - 1 | e), taking the
// left cell's right edge from the previous wave through LDS (wave 0 from wave 3 of the round
// before), then adds the face difference of cell e - 1 to D, taking the left face's flux from the
// previous wave the same way.  A run evaluates the edges of m1 - m0 + 2 cells and m1 - m0 + 1
// faces (its first face again at the next run's start); the LDS rows are double-buffered across
// rounds.  Workgroups are ordered z chunk fastest, then run, then the transverse line, and walk
// XCD-contiguous chunks.  a.dk divides by the number of z chunks, a.dn by the runs per line.
#ifndef AFD_MARCH_CELLS
#define AFD_MARCH_CELLS 30
#endif
constexpr int MARCH_CELLS = AFD_MARCH_CELLS;

#ifndef AFD_MARCH_WAVES
#define AFD_MARCH_WAVES AFD_FACE_WAVES_F32
#endif
#ifndef AFD_MARCH_PRIO
#define AFD_MARCH_PRIO 1
#endif

template <int RIEM, int DIR, bool ADD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AFD_MARCH_WAVES)))
void k_face_rmhd_weno_march(FaceArgs<float> a) {
    static_assert(DIR == 0 || DIR == 1, "pencils along x or y");
    constexpr int O = DIR == 0 ? 1 : 0;  // the transverse axis other than z
    __shared__ float edge[2][4][9][64];
    __shared__ float flx[2][4][9][64];
    const Grid &g = a.g;
    const int n2 = g.r1[2] - g.r0[2], nkc = (n2 + 63) / 64;
    const int nruns = (g.r1[DIR] - g.r0[DIR] + MARCH_CELLS - 1) / MARCH_CELLS;
    const unsigned b = (unsigned)xcd_chunked_block(blockIdx.x, gridDim.x);
    const unsigned r = udiv(a.dk, b), q = udiv(a.dn, r);
    const int kc = (int)(b - r * nkc), run = (int)(r - q * nruns), o = g.r0[O] + (int)q;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kraw = kc * 64 + lane;
    const bool lane_ok = kraw < n2;
    const int k = g.r0[2] + min(kraw, n2 - 1);  // in range for the loads
    const int m0 = g.r0[DIR] + run * MARCH_CELLS, m1 = min(m0 + MARCH_CELLS, g.r1[DIR]);
    const int count = m1 - m0 + 2;  // edge cells m0 - 1 .. m1
    const int rounds = (count + 3) / 4;
    const unsigned s = (unsigned)dstride<DIR>(g), n = (unsigned)g.ncell;
    for (int rd = 0; rd < rounds; ++rd) {
        const int buf = rd & 1, eidx = 4 * rd + w;
        const int e = m0 - 1 + min(eidx, count - 1);  // waves past the run repeat its last cell
        const unsigned c = (unsigned)(DIR == 0 ? cidx(g, e, o, k) : cidx(g, o, e, k));
        // the face e is the right face of cell e - 1 (eidx >= 2) and the left face of cell e (eidx <= count - 2)
        const bool right_of = eidx >= 2 && eidx < count, left_of = eidx >= 1 && eidx <= count - 2;
        const bool upd = lane_ok && right_of;
        const unsigned cl = c - s;
        // ADD, x: cell e - 1's divergence so far, loaded before the face arithmetic (1.42 -> 1.21
        // ms a launch at 256^3); loaded at the end, the loads' latency showed at every round.  In
        // y, loaded early it cost registers (1.42 ms), loaded late its latency (1.05 ms); y now
        // runs first and only writes
        float dold[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) dold[v] = (ADD && DIR == 0 && upd) ? a.D[v * n + cl] : 0.0f;
        State9 L, R;
        float lo[9], hi[9];
        // the wave's priority is raised while it loads its stencil, so that a round's loads go
        // out ahead of the other waves' arithmetic: y pencils 2.6 % faster, x 0.5 % (the z lanes,
        // independent waves, ran 15 % slower with it and do not take it)
        if (AFD_MARCH_PRIO) __builtin_amdgcn_s_setprio(1);
        weno5_cell_edges9(a.P, n, c, s, L, R, lo, hi);
        if (AFD_MARCH_PRIO) __builtin_amdgcn_s_setprio(0);
#pragma unroll
        for (int v = 0; v < 9; ++v) edge[buf][w][v][lane] = hi[v];
        __syncthreads();
        const bool first = rd == 0 && w == 0;  // the halo cell m0 - 1: no face
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            const float hp = first ? 0.0f : (w > 0 ? edge[buf][w - 1][v][lane] : edge[buf ^ 1][3][v][lane]);
            L.q[v] += (double)hp;
            R.q[v] += (double)lo[v];
        }
        RmhdFace<DIR, float> fc;
        rmhd_face_states<DIR, float>(a, L, R, fc);
        float Fown[9], Fex[9];
        if constexpr (DIR == 0) {
            rmhd_face_pair<RIEM, 0>(a, fc, e, o, k, (long)c, right_of, left_of, Fown, Fex);
        } else {
            rmhd_face_pair<RIEM, 1>(a, fc, o, e, k, (long)c, right_of || left_of, false, Fown, Fex);
#pragma unroll
            for (int v = 0; v < 9; ++v) Fex[v] = Fown[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) flx[buf][w][v][lane] = Fex[v];
        __syncthreads();
        if (upd) {  // cell e - 1: right face e (Fown), left face e - 1 (the previous wave's)
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                const float fl = w > 0 ? flx[buf][w - 1][v][lane] : flx[buf ^ 1][3][v][lane];
                const float d = (Fown[v] - fl) * a.idn;
                a.D[v * n + cl] = ADD ? (DIR == 0 ? dold[v] : a.D[v * n + cl]) + d : d;
            }
        }
    }
}

// ------------------------------------------------------------------ passive scalars (Q15)
// Passive fluxes of the reference's MUSCL (SRHD, RMHD) and PPM (SRHD) right-hand sides:
// the face states of the WENO, PPM and MUSCL variants are recomputed here and combined with
// the speeds/velocities the face kernels stored, in the reference's order of subtraction.
struct PassiveArgs {
    const double *P;
    const double *aux[3];   // per direction (sL, sR, vL, vR) x ncell
    double *rhs;            // n_passive x ncell, zero outside the range
    Grid g;
    int physics, recon, lim, np, po;
    double d[3];
};

AFD_D double pwrap(const Grid &g, const double *P, int v, int i, int j, int k) {
    // Python negative-index wrap of the reference's passive stencils (i-3 at i = 2);
    // indices past the end wrap as well (the reference reads out of range there).
    if (i < 0) i += g.nx;
    if (j < 0) j += g.ny;
    if (k < 0) k += g.nz;
    if (i >= g.nx) i -= g.nx;
    if (j >= g.ny) j -= g.ny;
    if (k >= g.nz) k -= g.nz;
    return P[v * g.ncell + cidx(g, i, j, k)];
}

// variant 0 WENO5, 1 PPM, 2 MUSCL (core/srhd_core.py:554-580) on the face (c-1 | c)
AFD_D double passive_flux(const PassiveArgs &a, int variant, int v, int dir, int i, int j, int k, double vL,
                          double vR, double sL, double sR) {
    double q[6];
    for (int o = -3; o <= 2; ++o)
        q[o + 3] = pwrap(a.g, a.P, v, i + (dir == 0 ? o : 0), j + (dir == 1 ? o : 0), k + (dir == 2 ? o : 0));
    double tL, tR;
    if (variant == 0) {
        tL = weno5_left<double>(q[0], q[1], q[2], q[3], q[4]);
        tR = weno5_right<double>(q[5], q[4], q[3], q[2], q[1]);
    } else if (variant == 1) {
        double a0, b0;
        ppm_reconstruct<double>(q[0], q[1], q[2], q[3], q[4], a0, tL);
        ppm_reconstruct<double>(q[1], q[2], q[3], q[4], q[5], tR, b0);
    } else {
        double slL = limiter<double>(a.lim, q[2] - q[1], q[3] - q[2]);
        double slR = limiter<double>(a.lim, q[3] - q[2], q[4] - q[3]);
        tL = q[2] + 0.5 * slL;
        tR = q[3] - 0.5 * slR;
    }
    return hlle_scalar<double>(tL, tR, tL * vL, tR * vR, sL, sR);
}

__global__ void k_passive_rhs(PassiveArgs a);

// ------------------------------------------------------------------ stage update + c2p
// Cells [lo, hi) per axis whose CFL speed a step's last stage takes: the flux range inside the
// interior [ng, n - ng), the part of max_char_speed's range [1, n-1)^3 that no fill writes.
struct SpeedBox {
    int lo[3], hi[3];
};

template <typename F> struct UpdArgs {
    const double *P0;     // step input primitives (U0 = p2c(P0) at every stage, passives t0)
    const double *Pin;    // stage input primitives (GLM, resistivity, passives t_{s-1})
    double *Pout;         // stage output primitives
    double *Uc;           // conserved: running stage value (FP64); the step start's is p2c(P0)
    const F *Fx, *FxL, *Fy, *Fz;  // face fluxes; FxL is the left-face x flux (carried speeds)
    const float *D;               // or (mixed RMHD WENO5 with shared edges) the FP32 flux divergence
    const double *rhs_pass;       // passive rhs (or null = 0)
    unsigned long long *stats;    // [hydro_iters, newton_evals, fallback, bisection, invalid, learned] or null
    unsigned char *flags;         // with stats: the last stage's con2prim branch bits per cell (C2PStats::flags)
    Grid g;
    int stage, nst, n_resist, np, po;
    double dt, dx, dy, dz, glm_cp2, resistivity;
    double gamma, gfac, p_max, v_max;
    const void *mlp;              // learned-guess weights (null: plain con2prim)
    int guess;                    // 0 plain, 1 learned, 2 previous-state Newton start
    int *queue;                   // learned guess: cells whose start failed the acceptance test, QUEUE_SLOTS regions
    unsigned *qcount;             //   of qcap entries each, filled per workgroup slot (or null)
    long qcap;
    int *rq;                      // mixed RMHD flux-divergence path: cells whose FP32 right-hand side is not
    unsigned *rqcount;            //   finite, queued as `queue` is, for k_repair_weno (or null)
    int ng, gr, gr_x0;            // GR sources: 0 none, 1 isotropic-Schwarzschild lapse, 2 Kerr-Schild
    double gr_mass, gr_gfac;      // GR_MASS; Gamma / (Gamma - 1) of the sources' enthalpy
    unsigned long long *amax;     // SRHD last stage: max characteristic speed over sbox of the new
                                  //   primitives, as the bits of a double >= 0 (atomicMax into
                                  //   AMAX_SLOTS slots), or null
    SpeedBox sbox;
};

// core/srhd_core.py:1268-1281 (max_char_speed): the largest |eigenvalue| of one cell's primitives
AFD_D double srhd_cell_speed(double gamma, double gfac, double rho, double vx, double vy, double vz, double p) {
    Phys<double> ph;
    ph.gamma = gamma;
    ph.gfac = gfac;
    const double cs = srhd_sound_speed(ph, rho, p);
    double l0, l1, l2, l3, l4, l5;
    eig_speeds(vx, cs, l0, l1);
    eig_speeds(vy, cs, l2, l3);
    eig_speeds(vz, cs, l4, l5);
    return pymax(pymax(pymax(pymax(pymax(fabs(l0), fabs(l1)), fabs(l2)), fabs(l3)), fabs(l4)), fabs(l5));
}

// The speed maximum is kept in AMAX_SLOTS slots AMAX_STRIDE words (256 B) apart, one per
// blockIdx % AMAX_SLOTS: one address taking every wave's atomic serialises them (a 256^3 update
// runs ~5x slower), spread slots run in parallel.
constexpr int AMAX_SLOTS = 256, AMAX_STRIDE = 32;
constexpr size_t AMAX_BOX_BYTES = sizeof(unsigned long long) * AMAX_SLOTS * AMAX_STRIDE;
constexpr size_t AMAX_BYTES = 2 * AMAX_BOX_BYTES;  // the speed box's slots, then the shell's

// The wave's largest v (every lane takes part; v >= 0, NaN counts as 0, as in the reduction of
// k_max_speed, which never takes a NaN) into the block's slot of dst by one atomicMax on the bits
// of a double.
AFD_D void wave_max_atomic(unsigned long long *dst, double v) {
    if (!(v > 0.0)) v = 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = pymax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0)
        atomicMax(dst + (blockIdx.x % AMAX_SLOTS) * AMAX_STRIDE, (unsigned long long)__double_as_longlong(v));
}

// Rejected cells of the learned guess are queued per workgroup slot (blockIdx % QUEUE_SLOTS),
// each slot a region of qcap = ceil(blocks / QUEUE_SLOTS) * BLOCK entries, so no slot overflows
// and few waves contend for one counter; k_c2p_queue then solves them densely.
constexpr int QUEUE_SLOTS = 64;
AFD_D void wave_enqueue(bool need, long c, int *queue, unsigned *qcount, long qcap) {
    const unsigned long long m = __ballot(need);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int slot = blockIdx.x % QUEUE_SLOTS;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(&qcount[slot], (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (need) queue[(long)slot * qcap + base + __popcll(m & ((1ull << lane) - 1ull))] = (int)c;
}

AFD_D double ssprk_combine(int stage, int nst, double u0, double uc, double r) {
#pragma clang fp contract(off)
    if (stage == 0) return u0 + r;
    if (nst == 2) return 0.5 * (u0 + uc + r);
    if (stage == 1) return 0.75 * u0 + 0.25 * (uc + r);
    return (1.0 / 3.0) * u0 + (2.0 / 3.0) * (uc + r);
}

// Solver counters: STATS_SLOTS copies of STATS_N counters; a wave adds its sums into the slot of
// its workgroup so that concurrent waves rarely hit the same address (the host sums the slots).
constexpr int STATS_N = 8, STATS_SLOTS = 64;
AFD_D void wave_add_stats(unsigned long long *stats, const C2PStats &st) {
    unsigned long long v[6] = {(unsigned long long)st.hydro_iters, (unsigned long long)st.newton_iters,
                               (unsigned long long)(st.flags & 1), (unsigned long long)((st.flags >> 1) & 1),
                               (unsigned long long)((st.flags >> 2) & 1), (unsigned long long)((st.flags >> 3) & 1)};
    unsigned long long *slot = stats + STATS_N * (blockIdx.x % STATS_SLOTS);
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        unsigned long long x = v[m];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(&slot[m], x);
    }
}

// Flux divergence and sources of one cell, in the reference's order of accumulation:
// x, y, z differences (core/srhd_core.py:662, 779, 922), then GLM damping and the resistive
// Laplacian of B (core/rmhd_core.py:1264-1275), n_resist times.
// core/gr_metric.py:6-16 and :18-38: d ln(alpha) / dx^i of the isotropic Schwarzschild lapse
// (metric 1) or the Kerr-Schild lapse (otherwise; the Kerr-Schild helper ignores the spin).
AFD_D void gr_lapse_grad(int metric, double x, double y, double z, double M, double *d) {
#pragma clang fp contract(off)
    const double r = sqrt(x * x + y * y + z * z) + 1e-12;
    if (metric == 1) {
        const double a = M / (2.0 * r), denom = 1.0 - a * a;
        const double dr = denom == 0.0 ? 0.0 : (2.0 * a) / (r * denom);
        d[0] = dr * (x / r);
        d[1] = dr * (y / r);
        d[2] = dr * (z / r);
    } else {
        const double H = M / r, dr = H / (r * (1.0 + 2.0 * H));
        d[0] = dr * (x / r);
        d[1] = dr * (y / r);
        d[2] = dr * (z / r);
    }
}

// core/gr_metric.py:18-22: the Kerr-Schild lapse (the flux-velocity rescale of every metric)
AFD_D double gr_ks_alpha(double x, double y, double z, double M) {
#pragma clang fp contract(off)
    const double r = sqrt(x * x + y * y + z * z) + 1e-12;
    return 1.0 / sqrt(1.0 + 2.0 * (M / r));
}

// Cell-centre coordinates of cell (i, j, k): core/grhd_core.py:26-30, 51-55
AFD_D void gr_xyz(int x0, int ng, int i, int j, int k, double dx, double dy, double dz, double *X) {
#pragma clang fp contract(off)
    X[0] = ((double)(x0 + (i - ng)) + 0.5) * dx;
    X[1] = ((double)(j - ng) + 0.5) * dy;
    X[2] = ((double)(k - ng) + 0.5) * dz;
}

// core/grhd_core.py:50-73 / core/grmhd_core.py:50-75: the lapse-gradient sources of an interior
// cell from the stage input's (unscaled) primitives.  The momentum source is (rho h W^2 [+ B^2] + p)
// d ln(alpha), the energy source S . d ln(alpha), with S from the flat p2c (the SRHD one keeps the
// definition-time 5/3; h takes the configured Gamma).
template <int NB, typename F> AFD_D void gr_source(const UpdArgs<F> &a, long c, double *rhs) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    if (i < a.ng || i >= g.nx - a.ng || j < a.ng || j >= g.ny - a.ng || k < a.ng || k >= g.nz - a.ng) return;
    double X[3], d[3], q[NB], U[NB];
    gr_xyz(a.gr_x0, a.ng, i, j, k, a.dx, a.dy, a.dz, X);
    gr_lapse_grad(a.gr, X[0], X[1], X[2], a.gr_mass, d);
#pragma unroll
    for (int v = 0; v < NB; ++v) q[v] = a.Pin[v * g.ncell + c];
    const double rho = q[0], p = q[4];
    double v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (v2 >= 1.0) v2 = 1.0 - 1e-14;
    const double W = 1.0 / sqrt(1.0 - v2);
    const double h = 1.0 + a.gr_gfac * p / pymax(rho, SMALL);
    double w;
    if constexpr (NB == 9) {
        const double B2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
        w = rho * h * W * W + B2 + p;
        rmhd_p2c_cell(a.gfac, a.p_max, a.v_max, q, U);
    } else {
        w = rho * h * W * W + p;
        srhd_p2c_cell(a.gfac, q, U);
    }
    rhs[1] += w * d[0];
    rhs[2] += w * d[1];
    rhs[3] += w * d[2];
    rhs[4] += U[1] * d[0] + U[2] * d[1] + U[3] * d[2];
}

// GLM damping and the resistive Laplacian of B (core/rmhd_core.py:1264-1275), n_resist times,
// after the flux divergence.
template <int NB, typename F> AFD_D void cell_rhs_glm_resist(const UpdArgs<F> &a, long c, double *rhs) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    if (NB == 9) {
        rhs[8 % NB] -= a.glm_cp2 * a.Pin[8 * g.ncell + c];
        double invdx2 = 1.0 / (a.dx * a.dx + SMALL), invdy2 = 1.0 / (a.dy * a.dy + SMALL),
               invdz2 = 1.0 / (a.dz * a.dz + SMALL);
        for (int rep = 0; rep < a.n_resist; ++rep) {
#pragma unroll
            for (int cc = 5; cc < 8; ++cc) {
                const double *q = a.Pin + cc * g.ncell;
                double lap = (q[c + g.sx] - 2.0 * q[c] + q[c - g.sx]) * invdx2 +
                             (q[c + g.sy] - 2.0 * q[c] + q[c - g.sy]) * invdy2 +
                             (q[c + 1] - 2.0 * q[c] + q[c - 1]) * invdz2;
                rhs[cc % NB] += a.resistivity * lap;
            }
        }
    }
}

// The flat RHS of a cell in the flux ranges: flux divergence, GLM and resistive sources.
template <int NB, typename F> AFD_D void cell_rhs_flat(const UpdArgs<F> &a, long c, double *rhs) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    if constexpr (sizeof(F) == 8) {  // FP64 path: the reference's divisions
#pragma unroll
        for (int v = 0; v < NB; ++v) {
            long o = v * g.ncell + c;
            rhs[v] -= ((double)a.Fx[o + g.sx] - (double)a.FxL[o]) / a.dx;
            rhs[v] -= ((double)a.Fy[o + g.sy] - (double)a.Fy[o]) / a.dy;
            rhs[v] -= ((double)a.Fz[o + 1] - (double)a.Fz[o]) / a.dz;
        }
    } else if (a.D) {  // the face kernels accumulated the divergence, already over the spacings
#pragma unroll
        for (int v = 0; v < NB; ++v) rhs[v] -= (double)a.D[v * g.ncell + c];
    } else {  // mixed path (FP32 fluxes): multiply by the inverse spacings
        const double ix = 1.0 / a.dx, iy = 1.0 / a.dy, iz = 1.0 / a.dz;
#pragma unroll
        for (int v = 0; v < NB; ++v) {
            long o = v * g.ncell + c;
            rhs[v] -= ((double)a.Fx[o + g.sx] - (double)a.FxL[o]) * ix;
            rhs[v] -= ((double)a.Fy[o + g.sy] - (double)a.Fy[o]) * iy;
            rhs[v] -= ((double)a.Fz[o + 1] - (double)a.Fz[o]) * iz;
        }
    }
    cell_rhs_glm_resist<NB, F>(a, c, rhs);
}

// The RHS of one cell: the flat RHS in the flux ranges (`in`), plus the GR sources on interior cells.
template <int NB, typename F> AFD_D void cell_rhs(const UpdArgs<F> &a, long c, bool in, double *rhs) {
#pragma clang fp contract(off)
#pragma unroll
    for (int v = 0; v < NB; ++v) rhs[v] = 0.0;
    if (in) cell_rhs_flat<NB, F>(a, c, rhs);
    if (a.gr) gr_source<NB, F>(a, c, rhs);
}

AFD_D bool in_range(const Grid &g, long c) {
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    return i >= g.r0[0] && i < g.r1[0] && j >= g.r0[1] && j < g.r1[1] && k >= g.r0[2] && k < g.r1[2];
}

// rhs of the stage input only (the reference's compute_rhs_*), written to Pout as (nb + np) x ncell
template <int NB, typename F>
__global__ __launch_bounds__(BLOCK) void k_rhs(UpdArgs<F> a) {
    const Grid &g = a.g;
    long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= g.ncell) return;
    double rhs[NB];
    cell_rhs<NB, F>(a, c, in_range(g, c), rhs);
#pragma unroll
    for (int v = 0; v < NB; ++v) a.Pout[v * g.ncell + c] = rhs[v];
    for (int t = 0; t < a.np; ++t) a.Pout[(NB + t) * g.ncell + c] = a.rhs_pass ? a.rhs_pass[t * g.ncell + c] : 0.0;
}

// Z = rho h W^2 of a primitive state (the previous-state Newton start)
AFD_D double prim_Z(double gfac, double rho, double vx, double vy, double vz, double p) {
    double v2 = vx * vx + vy * vy + vz * vz;
    if (!(v2 < 1.0)) v2 = 1.0 - 1e-14;
    return (rho + gfac * p) / (1.0 - v2);
}

// con2prim of one cell's conserved state U (FP64).  LEARNED: the RMHD Newton solve starts from
// the network's guess; guess == 2: from Z of the stage input's primitives (P0 at stage 0, else
// Pin).  Either start must pass the acceptance test and is backed by the plain solve: inline, or (queue set) by queueing
// the rejected cell for k_c2p_queue.  Reached by every lane of the wave (the network runs
// wave-wide); `live` marks lanes with a cell.  Returns whether q holds the cell's primitives.
template <int NB, bool LEARNED, typename F, bool LDSW = false>
AFD_D bool cell_c2p(const UpdArgs<F> &a, const double *U, bool live, long c, double *q, C2PStats &st,
                    MlpLds *lw, const MlpStaged<BLOCK> *ws) {
    // The mixed RMHD con2prim uses the fast FP64 rcp/rsq: its hydro stage only seeds a Newton solve
    // accepted by its residual.  The SRHD con2prim's output is the fixed point's own branch (converged
    // or the energy fallback), which ulp-level changes can flip, so it stays IEEE.
    constexpr bool FAST = sizeof(F) == 4 && NB == 9;
    if constexpr (NB == 9) {
        RmhdC2PParams cp = {a.gamma, a.gfac, a.p_max, a.v_max};
        bool done = false;
        if constexpr (LEARNED) {
            float x[5], y[2];
            bool fok = mlp_features(U, x) && live;
            if constexpr (LDSW) {  // the workgroup's LDS copy of the weights
                if (ws) mlp_stage_store<BLOCK>(*ws, *lw, threadIdx.x);  // loaded at the kernel's start
                mlp_infer_wave(MlpShared{lw}, x, y);
            } else {
                mlp_infer_wave(MlpGlobal{mlp_view(a.mlp)}, x, y);
            }
            if (fok) done = rmhd_c2p_from_guess<FAST>(cp, U, mlp_guess_Z(U[0], a.gfac, y), q, st);
        } else {
            if (a.guess == 2 && live) {
                const double *P = (a.stage == 0 ? a.P0 : a.Pin) + c;
                const long n = a.g.ncell;
                done = rmhd_c2p_from_guess<FAST>(cp, U, prim_Z(a.gfac, P[0], P[n], P[2 * n], P[3 * n], P[4 * n]), q,
                                                 st);
            }
        }
        if (a.queue) {
            wave_enqueue(live && !done, c, a.queue, a.qcount, a.qcap);
            return done;
        }
        if (live && !done) rmhd_c2p<FAST>(cp, U, q, st);
    } else {
        if (live) hydro_c2p<FAST>(a.gamma, a.gfac, a.p_max, a.v_max, U, q, st);
    }
    return live;
}

// The stage update of cell c (a lane with live set) from its right-hand side: the SSPRK
// combination into Uc (FP64), the passives on primitives and the con2prim of the new state into
// Pout.  Reached by every lane of the wave (the learned start runs wave-wide).  LDSW: the
// network's forward weights come from the workgroup's LDS copy lw.
template <int NB, typename F, bool LEARNED, bool LDSW = false>
AFD_D void cell_update(const UpdArgs<F> &a, long c, bool live, const double *rhs, MlpLds *lw = nullptr,
                       const MlpStaged<BLOCK> *ws = nullptr) {
    const Grid &g = a.g;
    if constexpr (NB == 9 && sizeof(F) == 4) {
        // a cell whose FP32 right-hand side overflowed (states beyond FP32's range, or a face
        // that no FP32 fallback saved) is left untouched here and queued for k_repair_weno,
        // which evaluates its faces in FP64 as the plain path does and then updates it
        if (a.rq) {
            bool bad = false;
            if (live) {
#pragma unroll
                for (int v = 0; v < NB; ++v) bad = bad || !isfinite(rhs[v]);
            }
            wave_enqueue(bad, c, a.rq, a.rqcount, a.qcap);
            live = live && !bad;
        }
    }
    C2PStats st = {0, 0, 0};
    double U[NB], q[NB];
#pragma unroll
    for (int v = 0; v < NB; ++v) U[v] = 0.0;
    const bool last = a.stage == a.nst - 1;
    if (live) {
        // U0 = p2c(P0) at every stage: the same arithmetic on the same input as at stage 0, so
        // recomputing it costs the read of P0 that storing and re-reading U0 would cost, and
        // saves the store
        double q0[NB];
#pragma unroll
        for (int v = 0; v < NB; ++v) q0[v] = a.P0[v * g.ncell + c];
        // every Uc load before the first Uc store: the compiler cannot tell the variables' planes
        // apart, and would otherwise wait on each load after the previous plane's store
        double uc[NB];
#pragma unroll
        for (int v = 0; v < NB; ++v) uc[v] = (a.stage == 0) ? 0.0 : a.Uc[v * g.ncell + c];
        // the mixed path takes the hardware reciprocals here as in its con2prim: U0 moves by a
        // few ulp, far below the FP32 flux rounding of the same stage
        if constexpr (NB == 9) rmhd_p2c_cell<sizeof(F) == 4>(a.gfac, a.p_max, a.v_max, q0, U);
        else srhd_p2c_cell(a.gfac, q0, U);
#pragma unroll
        for (int v = 0; v < NB; ++v) {
            double r;
            {
#pragma clang fp contract(off)
                r = a.dt * rhs[v];
            }
            U[v] = ssprk_combine(a.stage, a.nst, U[v], uc[v], r);
            if (!last) a.Uc[v * g.ncell + c] = U[v];  // the last stage's U is read only by k_c2p_queue
        }
    }
    const bool have = cell_c2p<NB, LEARNED, F, LDSW>(a, U, live, c, q, st, lw, ws);
    if (live) {
        if (have) {
#pragma unroll
            for (int v = 0; v < NB; ++v) a.Pout[v * g.ncell + c] = q[v];
        } else if (last) {  // queued for k_c2p_queue, which solves from Uc
#pragma unroll
            for (int v = 0; v < NB; ++v) a.Uc[v * g.ncell + c] = U[v];
        }
        if (last && a.flags) a.flags[c] = (unsigned char)st.flags;
        // SRHD fields between the base variables and the passives (the dissipative Pi, pi^ij,
        // q^i) pass through the stage unchanged: core/srhd_core.py:1308-1309, 1326-1327
        for (int v = NB; v < a.po; ++v) a.Pout[(long)v * g.ncell + c] = a.Pin[(long)v * g.ncell + c];
        for (int t = 0; t < a.np; ++t) {
            long o = (long)(a.po + t) * g.ncell + c;
            double rt = a.rhs_pass ? a.rhs_pass[t * g.ncell + c] : 0.0;
            double r;
            {
#pragma clang fp contract(off)
                r = a.dt * rt;
            }
            a.Pout[o] = ssprk_combine(a.stage, a.nst, a.P0[o], a.Pin[o], r);
        }
    }
    if (a.stats) wave_add_stats(a.stats, st);
    if constexpr (NB == 5) {
        // the next step's CFL speed over the speed box, from the primitives just stored (the
        // fills after the step write only outside it; k_max_speed_shell takes the rest of
        // max_char_speed's range after them), so that afd_max_char_speed needs no pass over the
        // whole state
        if (a.amax) {
            double v = 0.0;
            if (live) {
                int i, j, k;
                cell_ijk(g, c, i, j, k);
                const SpeedBox &b = a.sbox;
                if (i >= b.lo[0] && i < b.hi[0] && j >= b.lo[1] && j < b.hi[1] && k >= b.lo[2] && k < b.hi[2])
                    v = srhd_cell_speed(a.gamma, a.gfac, q[0], q[1], q[2], q[3], q[4]);
            }
            wave_max_atomic(a.amax, v);
        }
    }
}

// Stage update of one cell per thread: flux divergence + sources, then cell_update.  LEARNED
// selects the network's Newton start; a separate instantiation keeps its registers out of the
// plain kernel.  The learned instantiation is held to 4 waves per SIMD (<= 128 VGPRs; it takes
// 112, no spills, with the network's weights read from LDS, their loads held in registers until
// the network runs, and its two tiles evaluated one after the other): on the 256^3 RMHD WENO5 +
// HLLD step it ran 16 % slower at the compiler's own 2 waves, and 4 % slower at 3 waves with the
// tiles interleaved; 5 waves spilled 20 bytes and was 3 % slower.  The plain one is held to 5
// (<= 96 VGPRs, no spills; 100 registers and 4 waves otherwise): 6 % faster in FP32, 2.4 % in
// FP64 on that step (6 waves spills and gains nothing more).
#ifndef AFD_UPDATE_LEARNED_WAVES
#define AFD_UPDATE_LEARNED_WAVES 4
#endif
template <int NB, typename F, bool LEARNED>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(LEARNED ? AFD_UPDATE_LEARNED_WAVES : 5))) void k_update(
    UpdArgs<F> a) {
    const Grid &g = a.g;
    const long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    const bool live = c < g.ncell;
    if constexpr (LEARNED) {
        // the network's forward weights, read by every wave of the workgroup from LDS: loaded
        // first, written to LDS just before the network runs, so that their round trip overlaps
        // the cell's own loads
        __shared__ MlpLds mw;
        MlpStaged<BLOCK> ws;
        mlp_stage_load<BLOCK>(a.mlp, ws, threadIdx.x);
        double rhs[NB];
        if (live) cell_rhs<NB, F>(a, c, in_range(g, c), rhs);
        cell_update<NB, F, true, true>(a, c, live, rhs, &mw, &ws);
    } else {
        double rhs[NB];
        if (live) cell_rhs<NB, F>(a, c, in_range(g, c), rhs);
        cell_update<NB, F, false>(a, c, live, rhs);
    }
}

// y pencils fused with the stage update (mixed RMHD WENO5 on the flux-divergence path): the pencil
// of k_face_rmhd_weno_march<RIEM, 1> gives cell e - 1 its y difference, which is added to the z and
// x parts in D as that kernel would store it, and the same wave updates the cell at once
// (cell_update), so that the update's loads and stores overlap the other waves' face arithmetic.
// The result is the one of the separate y kernel and k_update.  Cells outside the flux range are
// updated by k_update_shell.
template <int RIEM, bool LEARNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FaceOcc<float>::waves)))
void k_face_march_update(FaceArgs<float> a, UpdArgs<float> u) {
    constexpr int DIR = 1, O = 0;
    __shared__ float edge[2][4][9][64];
    __shared__ float flx[2][4][9][64];
    const Grid &g = a.g;
    const int n2 = g.r1[2] - g.r0[2], nkc = (n2 + 63) / 64;
    const int nruns = (g.r1[DIR] - g.r0[DIR] + MARCH_CELLS - 1) / MARCH_CELLS;
    const unsigned b = (unsigned)xcd_chunked_block(blockIdx.x, gridDim.x);
    const unsigned r = udiv(a.dk, b), q = udiv(a.dn, r);
    const int kc = (int)(b - r * nkc), run = (int)(r - q * nruns), o = g.r0[O] + (int)q;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kraw = kc * 64 + lane;
    const bool lane_ok = kraw < n2;
    const int k = g.r0[2] + min(kraw, n2 - 1);  // in range for the loads
    const int m0 = g.r0[DIR] + run * MARCH_CELLS, m1 = min(m0 + MARCH_CELLS, g.r1[DIR]);
    const int count = m1 - m0 + 2;  // edge cells m0 - 1 .. m1
    const int rounds = (count + 3) / 4;
    const unsigned s = (unsigned)dstride<DIR>(g), n = (unsigned)g.ncell;
    for (int rd = 0; rd < rounds; ++rd) {
        const int buf = rd & 1, eidx = 4 * rd + w;
        const int e = m0 - 1 + min(eidx, count - 1);  // waves past the run repeat its last cell
        const unsigned c = (unsigned)cidx(g, o, e, k);
        const bool right_of = eidx >= 2 && eidx < count, left_of = eidx >= 1 && eidx <= count - 2;
        float Fown[9];
        {
            State9 L, R;
            float lo[9], hi[9];
            weno5_cell_edges9(a.P, n, c, s, L, R, lo, hi);
#pragma unroll
            for (int v = 0; v < 9; ++v) edge[buf][w][v][lane] = hi[v];
            __syncthreads();
            const bool first = rd == 0 && w == 0;  // the halo cell m0 - 1: no face
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                const float hp = first ? 0.0f : (w > 0 ? edge[buf][w - 1][v][lane] : edge[buf ^ 1][3][v][lane]);
                L.q[v] += (double)hp;
                R.q[v] += (double)lo[v];
            }
            RmhdFace<DIR, float> fc;
            rmhd_face_states<DIR, float>(a, L, R, fc);
            float unused[9];
            rmhd_face_pair<RIEM, 1>(a, fc, o, e, k, (long)c, right_of || left_of, false, Fown, unused);
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) flx[buf][w][v][lane] = Fown[v];
        __syncthreads();
        // keep the face arithmetic and the update apart: interleaved, their live ranges overlap
        __builtin_amdgcn_sched_barrier(0);
        if (right_of) {  // wave-uniform: cell e - 1, right face e (Fown), left face e - 1 (the previous wave's)
            const unsigned cl = c - s;
            double rhs[9];
            if (lane_ok) {
#pragma unroll
                for (int v = 0; v < 9; ++v) {
                    const float fl = w > 0 ? flx[buf][w - 1][v][lane] : flx[buf ^ 1][3][v][lane];
                    const float d = a.D[v * n + cl] + (Fown[v] - fl) * a.idn;
                    rhs[v] = 0.0;
                    rhs[v] -= (double)d;
                }
                cell_rhs_glm_resist<9, float>(u, (long)cl, rhs);
                if (u.gr) gr_source<9, float>(u, (long)cl, rhs);
            }
            cell_update<9, float, LEARNED>(u, (long)cl, lane_ok, rhs);
        }
    }
}

// The stage update of the cells outside the flux range [r0, r1) per axis (their flux divergence is
// zero; the reference's p2c -> c2p runs over the whole array), one thread per such cell: the x
// planes below r0 and from r1 (all j, k), then the rows of in-range x below r0 and from r1 in y (all
// k), then the in-range x, y lines' k below r0 and from r1.
template <int NB, typename F, bool LEARNED>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(LEARNED ? 4 : 5))) void k_update_shell(
    UpdArgs<F> a) {
    const Grid &g = a.g;
    const long ni = g.r1[0] - g.r0[0], nj = g.r1[1] - g.r0[1];
    const long ox = g.nx - ni, oy = g.ny - nj, oz = g.nz - (g.r1[2] - g.r0[2]);
    const long nA = ox * g.ny * g.nz, nB = ni * oy * g.nz, nC = ni * nj * oz;
    long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    const bool live = t < nA + nB + nC;
    long i = 0, j = 0, k = 0;
    if (live) {
        auto out = [&](long p, int ax) { return p < g.r0[ax] ? p : g.r1[ax] + (p - g.r0[ax]); };
        if (t < nA) {
            k = t % g.nz;
            j = (t / g.nz) % g.ny;
            i = out(t / ((long)g.nz * g.ny), 0);
        } else if (t < nA + nB) {
            t -= nA;
            k = t % g.nz;
            j = out((t / g.nz) % oy, 1);
            i = g.r0[0] + t / ((long)g.nz * oy);
        } else {
            t -= nA + nB;
            k = out(t % oz, 2);
            j = g.r0[1] + (t / oz) % nj;
            i = g.r0[0] + t / (oz * nj);
        }
    }
    const long c = live ? cidx(g, (int)i, (int)j, (int)k) : 0;
    double rhs[NB];
    if (live) cell_rhs<NB, F>(a, c, false, rhs);
    cell_update<NB, F, LEARNED>(a, c, live, rhs);
}

// FP64 right-hand side of cell c on the mixed RMHD WENO5 path, as the plain path computes it
// (k_face_rmhd<2, RIEM, dir, double> and cell_rhs_flat): the own-speed fluxes of its right faces,
// the y and z fluxes of its left faces, and the x left face's flux with the speeds of its
// predecessor's z-back face (Q3, rmhd_face_store), all from FP64 face states; then GLM damping,
// resistivity and GR sources.
template <int RIEM>
AFD_D void repair_rhs(const FaceArgs<double> &fa, const UpdArgs<float> &ua, long c, double *rhs) {
#pragma clang fp contract(off)
    const Grid &g = fa.g;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    double Fr[3][9], Fl[3][9];
    {
        RmhdFace<0, double> f;
        rmhd_face_setup<2, 0, double>(fa, c + g.sx, f);
        rmhd_riemann<2, RIEM, 0, double>(fa.ph, f, f.sL, f.sR, Fr[0]);
    }
    {
        double cL = -1.0, cR = 1.0;  // hlle_bounds_fast without a predecessor
        long pc = -1;
        if (k > g.r0[2]) pc = c - 1;
        else if (j > g.r0[1]) pc = cidx(g, i, j - 1, g.r1[2] - 1);
        else if (i > g.r0[0]) pc = cidx(g, i - 1, g.r1[1] - 1, g.r1[2] - 1);
        if (pc >= 0) {
            RmhdFace<2, double> f;
            rmhd_face_setup<2, 2, double>(fa, pc, f);
            cL = f.sL;
            cR = f.sR;
        }
        RmhdFace<0, double> f;
        rmhd_face_setup<2, 0, double>(fa, c, f);
        rmhd_riemann<2, RIEM, 0, double>(fa.ph, f, cL, cR, Fl[0]);
    }
    {
        RmhdFace<1, double> f;
        rmhd_face_setup<2, 1, double>(fa, c + g.sy, f);
        rmhd_riemann<2, RIEM, 1, double>(fa.ph, f, f.sL, f.sR, Fr[1]);
        rmhd_face_setup<2, 1, double>(fa, c, f);
        rmhd_riemann<2, RIEM, 1, double>(fa.ph, f, f.sL, f.sR, Fl[1]);
    }
    {
        RmhdFace<2, double> f;
        rmhd_face_setup<2, 2, double>(fa, c + 1, f);
        rmhd_riemann<2, RIEM, 2, double>(fa.ph, f, f.sL, f.sR, Fr[2]);
        rmhd_face_setup<2, 2, double>(fa, c, f);
        rmhd_riemann<2, RIEM, 2, double>(fa.ph, f, f.sL, f.sR, Fl[2]);
    }
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        rhs[v] = 0.0;
        rhs[v] -= (Fr[0][v] - Fl[0][v]) / ua.dx;
        rhs[v] -= (Fr[1][v] - Fl[1][v]) / ua.dy;
        rhs[v] -= (Fr[2][v] - Fl[2][v]) / ua.dz;
    }
    cell_rhs_glm_resist<9, float>(ua, c, rhs);
    if (ua.gr) gr_source<9, float>(ua, c, rhs);
}

// The cells k_update queued in ua.rq: FP64 right-hand side (repair_rhs), then the stage update
// (plain or previous-state start, con2prim inline).  blockIdx.y = queue slot; the block strides
// over the slot's entries with every lane in each round (cell_update is wave-wide).
template <int RIEM>
__global__ __launch_bounds__(BLOCK) void k_repair_weno(FaceArgs<double> fa, UpdArgs<float> ua) {
    const int slot = blockIdx.y;
    const long cnt = ua.rqcount[slot];
    UpdArgs<float> u = ua;
    u.queue = nullptr;
    u.rq = nullptr;
    u.mlp = nullptr;
    for (long base = (long)blockIdx.x * BLOCK; base < cnt; base += (long)gridDim.x * BLOCK) {
        const long e = base + threadIdx.x;
        const bool live = e < cnt;
        const long c = live ? (long)ua.rq[(long)slot * ua.qcap + e] : 0;
        double rhs[9];
        if (live) repair_rhs<RIEM>(fa, ua, c, rhs);
        cell_update<9, float, false>(u, c, live, rhs);
    }
}

// Plain RMHD con2prim of the queued cells (rejected learned starts), densely packed:
// blockIdx.y = queue slot, grid-stride over the slot's entries.
template <typename F> __global__ __launch_bounds__(BLOCK) void k_c2p_queue(UpdArgs<F> a) {
    const int slot = blockIdx.y;
    const long n = a.g.ncell;
    const unsigned cnt = a.qcount[slot];
    const int *q_slot = a.queue + (long)slot * a.qcap;
    RmhdC2PParams cp = {a.gamma, a.gfac, a.p_max, a.v_max};
    C2PStats st = {0, 0, 0};
    for (long k = (long)blockIdx.x * BLOCK + threadIdx.x; k < (long)cnt; k += (long)gridDim.x * BLOCK) {
        const long c = q_slot[k];
        double U[9], q[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) U[v] = a.Uc[v * n + c];
        C2PStats sc = {0, 0, 0};
        rmhd_c2p<sizeof(F) == 4>(cp, U, q, sc);
#pragma unroll
        for (int v = 0; v < 9; ++v) a.Pout[v * n + c] = q[v];
        if (a.flags && a.stage == a.nst - 1) a.flags[c] = (unsigned char)sc.flags;
        st.hydro_iters += sc.hydro_iters;
        st.newton_iters += sc.newton_iters;
        st.flags |= sc.flags;
    }
    if (a.stats) wave_add_stats(a.stats, st);
}

// ------------------------------------------------------------------ pointwise batches (test hooks)
struct BatchArgs {
    const double *in;
    double *out;
    int *iters;  // per state: hydro iterations + Newton evaluations (or null)
    unsigned long long *stats;  // solver counters (or null), as in the stage update
    long n;
    int physics;
    double gamma, gfac, p_max, v_max;
};

__global__ void k_c2p_batch(BatchArgs a);

__global__ void k_p2c_batch(BatchArgs a);

// ------------------------------------------------------------------ boundary fills
// core/boundary.py:4-8: y ghosts (all x, all z) from the opposite interior planes, then z
// ghosts (all x, all y) the same way.  bc = 0 periodic, 1 zero-gradient (additive option).
struct BcArgs {
    double *P;
    Grid g;
    int ng, bc;
};

__global__ void k_fill_y(BcArgs a);

__global__ void k_fill_z(BcArgs a);

// core/boundary.py:10-11: every +x ghost plane from the last interior plane
__global__ void k_outflow_x(BcArgs a);

// core/nozzle.py:27-92: the -x ghost planes over interior (j, k).  The host precomputes the
// profile of every variable the reference writes (nozzle_weight uses libm tanh/cos); the
// perturbation takes 3 uniform draws per cell with s > 0, in g -> j -> k order.
struct NozzleArgs {
    double *P;
    const double *tmpl;   // nv x ny_int x nz_int
    const int *idx;       // ny_int x nz_int: rank among cells with s > 0, or -1
    const double *draws;  // 3 * ng * nact uniforms in [0, 1)
    Grid g;
    int ng, ny_int, nz_int, nact, perturb;
    unsigned mask;        // bit v set: the nozzle writes variable v
    double vamp, pamp;
};

__global__ void k_nozzle(NozzleArgs a);

// The driver's post-step sources (solvers/srhd3d_mpi_muscl.py:324-326), pointwise over the
// interior, applied in that order to each cell in one pass: cooling/heating
// (core/source_terms.py:5-28), two-temperature relaxation (:30-55), H/He ion chemistry
// (core/chemistry.py:43-139).  One thread per cell of a k row; blockIdx.x = interior (i, j),
// blockIdx.y = chunk of k.
enum { SRC_COOL = 1, SRC_TWO_TEMP = 2, SRC_CHEM = 4 };
struct SourceArgs {
    double *P;
    Grid g;
    int ng, which;             // SRC_* bits
    double dt;
    double cool, heat, p_cap;  // COOLING_LAMBDA, HEATING_RATE, P_MAX
    int t_off;                 // Te, Ti
    double tau;                // TEI_TAU
    int c_off;                 // xHII, xHeII, xHeIII
    double X, Y, tunit, rate_scale, energy_scale, tmin, gamma;
};
__global__ void k_sources(SourceArgs a);

// core/dissipation.py:5-215: upwind advection of the ten dissipative fields (from the state as
// it was, :47-103) and their IMEX relaxation (:105-213), interior cells; the new fields go to S
// (10 planes of ncell) and k_diss_store copies them back, so that every cell reads old values.
struct DissArgs {
    const double *P;
    double *S;
    Grid g;
    int ng, advect;
    double dt, dx, dy, dz;
    double zeta, eta, kappa, tau_bulk, tau_shear, tau_heat, cap;
};
__global__ void k_diss(DissArgs a);

// core/grhd_core.py:22-38, core/grmhd_core.py:22-38 (ORTHONORMAL_FLUX): Q = P with the interior
// velocities divided by max(alpha_KS, 1e-12); the face kernels then read Q.  One thread per cell.
struct GrRescaleArgs {
    const double *P;
    double *Q;
    Grid g;
    int ng, x0;
    double mass, dx, dy, dz;
};
__global__ void k_gr_rescale(GrRescaleArgs a);
__global__ void k_diss_store(DissArgs a, double *P);

// ------------------------------------------------------------------ reductions
// core/srhd_core.py:1267-1281, the maximum over cells [1, n-1)^3, in two parts: the cells of the
// speed box (SpeedBox: the flux range inside the interior, which the fills never write), taken by
// the last stage of a step (cell_update) or by k_max_speed_box into the first AMAX_SLOTS slots,
// and the rest, taken by k_max_speed_shell after the fills into the next AMAX_SLOTS slots.  Both
// atomicMax the bits of a double >= 0.  k_amax_finish (one block of BLOCK threads) writes
// max(all slots, floor) as a double to out and zeroes the shell's slots for the next call.
__global__ void k_max_speed_box(const double *P, Grid g, SpeedBox b, double gamma, double gfac,
                                unsigned long long *amax);
__global__ void k_max_speed_shell(const double *P, Grid g, SpeedBox b, double gamma, double gfac,
                                  unsigned long long *amax);
__global__ void k_amax_finish(unsigned long long *amax, double floor_value, double *out);

// Diagnostics of utils/diagnostics.py over the interior [ng, ng + n) per axis: maximum Lorentz
// factor (:15-27), inlet flux sum of Sx dy dz on the plane i = ng (:31-43; the reference's
// prim_to_cons* are called there without gamma, so their definition-time 5/3 applies) and, for
// RMHD, the centred divergence of B (:143-153): max |divB|, sum, sum of squares.
constexpr int DIAG_N = 5;  // maxW, max|divB|, sum divB, sum divB^2, inlet flux
struct DiagArgs {
    const double *P;
    double *partial;  // DIAG_N x gridDim.x
    Grid g;
    int ng, rmhd;
    double dx, dy, dz, gfac53, p_max, v_max;
};
__global__ void k_diag(DiagArgs a);
__global__ void k_diag_reduce(const double *partial, int nblocks, double *out);
// the mid-plane centerline (utils/diagnostics.py:59-126): Gamma, rho, p, vx at j, k = mid, i interior
__global__ void k_centerline(const double *P, Grid g, int ng, double v_max, double *out);

// count of non-finite values over n doubles; one partial count per block
__global__ void k_count_nonfinite(const double *x, long n, unsigned long long *partial);

__global__ void k_reduce_sum_u64(const unsigned long long *partial, int n,
                                                           unsigned long long *out);

}  // namespace afd
