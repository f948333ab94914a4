// afd_device.hpp -- device-side physics of the SRHD/RMHD finite-volume step (gfx950).
//
// Pointwise building blocks of the face-flux and stage-update kernels.  Each one
// follows the reference function cited beside it; the face arithmetic is
// templated on the compute type T (double: plain FP64 path, float: mixed path).
// The Lorentz factor is always evaluated in FP64 from FP64 face velocities and
// handed in, because 1 - v^2 cancels catastrophically in FP32 at W ~ 7..50.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AFD_D __device__ __forceinline__

namespace afd {

constexpr double SMALL = 1e-12;

template <typename T> struct Lim;
template <> struct Lim<double> {
    static constexpr double small = 1e-12;
    static constexpr double one_minus = 1.0 - 1e-14;   // the reference's 1 - 1e-14 caps
    static constexpr double tiny = 1e-14;
};
template <> struct Lim<float> {
    static constexpr float small = 1e-12f;
    static constexpr float one_minus = 0.99999994f;      // largest float below 1
    static constexpr float tiny = 0.0f;                  // 1e-14 is below float resolution near 1
};

// Physical parameters in the compute type.  gfac = Gamma/(Gamma-1) is the first
// operation of the reference's enthalpy expression, so precomputing it keeps the order.
template <typename T> struct Phys {
    T gamma, gfac, gm1_over_g_inv;  // Gamma, Gamma/(Gamma-1), (Gamma-1) (used as (h-1) rho (G-1) / G)
    T p_max, v_max, vmax2, glm_ch2;
};

template <typename T> AFD_D T pymax(T a, T b) { return (b > a) ? b : a; }
// Division and square root in the compute type: IEEE in FP64 (the plain path rounds as the
// reference); in the FP32 path the hardware reciprocal and square root (about 1 ulp), whose
// error is far below the FP32 rounding the mixed path already accepts.
template <typename T> AFD_D T fdiv(T a, T b) { return a / b; }
template <> AFD_D float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename T> AFD_D T fsqrt(T x) { return sqrt(x); }
template <> AFD_D float fsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
// FP64 division / inverse square root.  FAST = false: IEEE (the plain path rounds as the
// reference).  FAST = true (mixed path: con2prim, Lorentz factors): the hardware reciprocal / reciprocal square
// root refined by one Newton step (measured on gfx950: <= 11 ulp for 1/x, <= 20 ulp for 1/sqrt x),
// at about a third of the instructions.  Newton-update divisions and the rare cap branches stay IEEE in both.
template <bool FAST> AFD_D double ddiv(double a, double b) {
    if constexpr (FAST) {
        double r = __builtin_amdgcn_rcp(b);
        r = fma(r, fma(-b, r, 1.0), r);
        return a * r;
    } else {
        return a / b;
    }
}
template <bool FAST> AFD_D double dinvsqrt(double x) {  // 1 / sqrt(x), x > 0
    if constexpr (FAST) {
        double y = __builtin_amdgcn_rsq(x);
        double h = 0.5 * x * y;
        return y * fma(-h, y, 1.5);
    } else {
        return 1.0 / sqrt(x);
    }
}

template <typename T> AFD_D T pymin(T a, T b) { return (b < a) ? b : a; }
template <typename T> AFD_D T sq(T x) { return x * x; }

// ---------------------------------------------------------------- limiters
// core/srhd_core.py:305-325, core/rmhd_core.py:52-72
template <typename T> AFD_D T minmod(T a, T b) {
    if (a * b <= T(0)) return T(0);
    return (fabs(a) < fabs(b)) ? a : b;
}
template <typename T> AFD_D T limiter(int lid, T dqL, T dqR) {
    if (lid == 1) return minmod(dqL, dqR);
    if (lid == 2) {
        if constexpr (sizeof(T) == 4) {
            // FP32 van Leer as 2 dqL (dqR / (dqL + dqR)): the quotient lies in (0, 1) when the
            // slopes share a sign, so the slope stays finite where the product of two slopes
            // above ~1.8e19 overflows (2 dqL dqR / (dqL + dqR) then gives inf or NaN).  The
            // product still decides the sign test: where it underflows (both slopes below
            // ~1e-22) the slope is 0, as before, and where it passes, |dqL + dqR| >= 3.7e-23, so
            // its reciprocal is finite (a denormal sum's is not)
            if (!(dqL * dqR > T(0))) return T(0);
            return T(2) * dqL * fdiv(dqR, dqL + dqR);
        } else {
            if (dqL * dqR <= T(0)) return T(0);
            T denom = dqL + dqR;
            if (denom == T(0)) return T(0);
            return fdiv(T(2) * dqL * dqR, denom);
        }
    }
    T mm = minmod(dqL, dqR);
    return minmod(T(0.5) * (dqL + dqR), T(2) * mm);
}

// core/srhd_core.py:327-341
template <typename T> AFD_D void ppm_reconstruct(T qm2, T qm1, T q0, T qp1, T qp2, T &oL, T &oR) {
    T qL = T(7.0 / 12.0) * (qm1 + q0) - T(1.0 / 12.0) * (qm2 + qp1);
    T qR = T(7.0 / 12.0) * (q0 + qp1) - T(1.0 / 12.0) * (qm1 + qp2);
    T qmax = pymax(pymax(qm1, q0), qp1);
    T qmin = pymin(pymin(qm1, q0), qp1);
    if (qL > qmax) qL = qmax;
    if (qL < qmin) qL = qmin;
    if (qR > qmax) qR = qmax;
    if (qR < qmin) qR = qmin;
    if ((qR - q0) * (q0 - qL) <= T(0)) {
        qL = q0;
        qR = q0;
    }
    oL = qL;
    oR = qR;
}

// core/srhd_core.py:343-359 (WENO5-JS, Jiang & Shu 1996, eps = 1e-12, weights 0.1/0.6/0.3)
template <typename T> AFD_D T weno5_left(T qm2, T qm1, T q0, T qp1, T qp2) {
    const T eps = T(1e-12);
    T IS0 = T(13.0 / 12.0) * sq(qm2 - T(2) * qm1 + q0) + T(0.25) * sq(qm2 - T(4) * qm1 + T(3) * q0);
    T IS1 = T(13.0 / 12.0) * sq(qm1 - T(2) * q0 + qp1) + T(0.25) * sq(qm1 - qp1);
    T IS2 = T(13.0 / 12.0) * sq(q0 - T(2) * qp1 + qp2) + T(0.25) * sq(T(3) * q0 - T(4) * qp1 + qp2);
    T a0 = fdiv(T(0.1), sq(eps + IS0));
    T a1 = fdiv(T(0.6), sq(eps + IS1));
    T a2 = fdiv(T(0.3), sq(eps + IS2));
    T s = a0 + a1 + a2;
    T w0 = fdiv(a0, s), w1 = fdiv(a1, s), w2 = fdiv(a2, s);
    T p0 = T(1.0 / 3.0) * qm2 - T(7.0 / 6.0) * qm1 + T(11.0 / 6.0) * q0;
    T p1 = T(-1.0 / 6.0) * qm1 + T(5.0 / 6.0) * q0 + T(1.0 / 3.0) * qp1;
    T p2 = T(1.0 / 3.0) * q0 + T(5.0 / 6.0) * qp1 - T(1.0 / 6.0) * qp2;
    return w0 * p0 + w1 * p1 + w2 * p2;
}
// core/srhd_core.py:361-377 (mirror image)
template <typename T> AFD_D T weno5_right(T qp2, T qp1, T q0, T qm1, T qm2) {
    return weno5_left(qp2, qp1, q0, qm1, qm2);
}

// FP32 WENO5 of large deviations.  The nonlinear weights a_k / sum(a_j) with a_k = c_k / (eps + IS_k)^2
// are invariant under q -> q / D when eps scales as eps / D^2 (IS_k scales as D^2), and every
// candidate polynomial is linear in q.  In FP32, (eps + IS_k)^2 overflows once the deviations
// exceed about 4e9 (IS ~ q^2, squared again), and the weights underflow well before that: every
// a_k becomes 0 and the weight sum 0 * (1 / 0) = NaN, where FP64 still returns the edge value.
// The kernels evaluate every variable of a face or cell as the reference does, track the smallest
// weight sum, and only where it falls below WENO_SUM_MIN evaluate that lane again on q / max|q|
// with the scaled eps (floored at 1e-18, so that 1 / eps^2 stays finite; on such stencils eps is
// far below every nonzero indicator either way), scaling the result back.  The floor only
// matters where the weights differ from FP64's by far less than FP32 rounding.
constexpr float WENO_SUM_MIN = 1e-30f;

// core/srhd_core.py:343-359 in FP32 with a given eps; s = the weight sum
AFD_D float weno5_left_f32_eps(float qm2, float qm1, float q0, float qp1, float qp2, float eps, float &s) {
    float IS0 = (13.0f / 12.0f) * sq(qm2 - 2.0f * qm1 + q0) + 0.25f * sq(qm2 - 4.0f * qm1 + 3.0f * q0);
    float IS1 = (13.0f / 12.0f) * sq(qm1 - 2.0f * q0 + qp1) + 0.25f * sq(qm1 - qp1);
    float IS2 = (13.0f / 12.0f) * sq(q0 - 2.0f * qp1 + qp2) + 0.25f * sq(3.0f * q0 - 4.0f * qp1 + qp2);
    float a0 = fdiv(0.1f, sq(eps + IS0));
    float a1 = fdiv(0.6f, sq(eps + IS1));
    float a2 = fdiv(0.3f, sq(eps + IS2));
    s = a0 + a1 + a2;
    float w0 = fdiv(a0, s), w1 = fdiv(a1, s), w2 = fdiv(a2, s);
    float p0 = (1.0f / 3.0f) * qm2 - (7.0f / 6.0f) * qm1 + (11.0f / 6.0f) * q0;
    float p1 = (-1.0f / 6.0f) * qm1 + (5.0f / 6.0f) * q0 + (1.0f / 3.0f) * qp1;
    float p2 = (1.0f / 3.0f) * q0 + (5.0f / 6.0f) * qp1 - (1.0f / 6.0f) * qp2;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

// The FP32 WENO5 edge with the rescaled evaluation where the weights underflow.
AFD_D float weno5_left_f32_rescaled(float qm2, float qm1, float q0, float qp1, float qp2) {
    float s;
    float r = weno5_left_f32_eps(qm2, qm1, q0, qp1, qp2, 1e-12f, s);
    if (!(s >= WENO_SUM_MIN)) {
        const float D = fmaxf(fmaxf(fmaxf(fabsf(qm2), fabsf(qm1)), fmaxf(fabsf(q0), fabsf(qp1))), fabsf(qp2));
        if (D > 0.0f) {
            const float iD = 1.0f / D;
            const float e = fmaxf(1e-12f * iD * iD, 1e-18f);
            r = D * weno5_left_f32_eps(qm2 * iD, qm1 * iD, q0 * iD, qp1 * iD, qp2 * iD, e, s);
        }
    }
    return r;
}

// Both WENO5 edge values of one cell, from the deviations of its four stencil neighbours from
// the cell's own value (the cell's deviation is 0): hi = weno5_left (the cell's right edge),
// lo = weno5_right (its left edge).  The mirror image has the same three smoothness indicators
// in reverse order, so one set of indicators and reciprocals serves both edges, and each edge
// is normalised once: sum(a_k p_k) / sum(a_k) equals sum(w_k p_k) with w_k = a_k / sum(a_k).
// Returns the smaller of the two weight sums.
AFD_D float weno5_cell_pair(float dm2, float dm1, float dp1, float dp2, float &lo, float &hi, float eps = 1e-12f) {
    const float IS0 = (13.0f / 12.0f) * sq(dm2 - 2.0f * dm1) + 0.25f * sq(dm2 - 4.0f * dm1);
    const float IS1 = (13.0f / 12.0f) * sq(dm1 + dp1) + 0.25f * sq(dm1 - dp1);
    const float IS2 = (13.0f / 12.0f) * sq(dp2 - 2.0f * dp1) + 0.25f * sq(dp2 - 4.0f * dp1);
    const float r0 = __builtin_amdgcn_rcpf(sq(eps + IS0));
    const float r1 = __builtin_amdgcn_rcpf(sq(eps + IS1));
    const float r2 = __builtin_amdgcn_rcpf(sq(eps + IS2));
    // right edge: weights 0.1, 0.6, 0.3 on stencils (-2..0), (-1..1), (0..2)
    const float h0 = (1.0f / 3.0f) * dm2 - (7.0f / 6.0f) * dm1;
    const float h1 = (-1.0f / 6.0f) * dm1 + (1.0f / 3.0f) * dp1;
    const float h2 = (5.0f / 6.0f) * dp1 - (1.0f / 6.0f) * dp2;
    const float ha0 = 0.1f * r0, ha1 = 0.6f * r1, ha2 = 0.3f * r2;
    const float sh = ha0 + ha1 + ha2;
    hi = (ha0 * h0 + ha1 * h1 + ha2 * h2) * __builtin_amdgcn_rcpf(sh);
    // left edge: the mirror image, weights 0.1, 0.6, 0.3 on stencils (0..2), (-1..1), (-2..0)
    const float l0 = (1.0f / 3.0f) * dp2 - (7.0f / 6.0f) * dp1;
    const float l1 = (-1.0f / 6.0f) * dp1 + (1.0f / 3.0f) * dm1;
    const float l2 = (5.0f / 6.0f) * dm1 - (1.0f / 6.0f) * dm2;
    const float la0 = 0.1f * r2, la2 = 0.3f * r0;
    const float sl = la0 + ha1 + la2;
    lo = (la0 * l0 + ha1 * l1 + la2 * l2) * __builtin_amdgcn_rcpf(sl);
    return fminf(sh, sl);
}

// weno5_cell_pair with the rescaled evaluation where the weights underflow.
AFD_D void weno5_cell_pair_rescaled(float dm2, float dm1, float dp1, float dp2, float &lo, float &hi) {
    const float s = weno5_cell_pair(dm2, dm1, dp1, dp2, lo, hi);
    if (!(s >= WENO_SUM_MIN)) {
        const float D = fmaxf(fmaxf(fabsf(dm2), fabsf(dm1)), fmaxf(fabsf(dp1), fabsf(dp2)));
        if (D > 0.0f) {
            const float iD = 1.0f / D;
            const float e = fmaxf(1e-12f * iD * iD, 1e-18f);
            weno5_cell_pair(dm2 * iD, dm1 * iD, dp1 * iD, dp2 * iD, lo, hi, e);
            lo *= D;
            hi *= D;
        }
    }
}

// core/srhd_core.py:171-186
template <typename T> AFD_D T hlle_scalar(T qL, T qR, T fL, T fR, T sL, T sR) {
    if (sL >= T(0)) return fL;
    if (sR <= T(0)) return fR;
    return fdiv(sR * fL - sL * fR + sL * sR * (qR - qL), sR - sL + Lim<T>::small);
}
template <int N, typename T>
AFD_D void hlle_vec(const T *UL, const T *UR, const T *FL, const T *FR, T sL, T sR, T *out) {
    if (sL >= T(0)) {
#pragma unroll
        for (int m = 0; m < N; ++m) out[m] = FL[m];
    } else if (sR <= T(0)) {
#pragma unroll
        for (int m = 0; m < N; ++m) out[m] = FR[m];
    } else {
        T inv = sR - sL + Lim<T>::small;
        T sLsR = sL * sR;
#pragma unroll
        for (int m = 0; m < N; ++m) out[m] = fdiv(sR * FL[m] - sL * FR[m] + sLsR * (UR[m] - UL[m]), inv);
    }
}

// core/srhd_core.py:142-154, core/rmhd_core.py:459-471
template <typename T> AFD_D void eig_speeds(T vn, T cs, T &lm, T &lp) {
    T dp = T(1) + vn * cs;
    T dm = T(1) - vn * cs;
    if (dp == T(0)) dp = Lim<T>::small;
    if (dm == T(0)) dm = Lim<T>::small;
    T p = fdiv(vn + cs, dp);
    T m = fdiv(vn - cs, dm);
    p = p > T(1) ? T(1) : (p < T(-1) ? T(-1) : p);
    m = m > T(1) ? T(1) : (m < T(-1) ? T(-1) : m);
    lm = m;
    lp = p;
}

// Newtonian-form HLLC on components [0:5] (Q13); NM = normal momentum index (1 + dir).
// core/srhd_core.py:188-282, core/rmhd_core.py:473-567
template <int NM, typename T>
AFD_D void hllc5(const T *UL, const T *UR, const T *FL, const T *FR, T sL, T sR, T pL, T pR, T vL, T vR, T *out) {
    if (sL >= T(0)) {
#pragma unroll
        for (int m = 0; m < 5; ++m) out[m] = FL[m];
        return;
    }
    if (sR <= T(0)) {
#pragma unroll
        for (int m = 0; m < 5; ++m) out[m] = FR[m];
        return;
    }
    T DL = UL[0], DR = UR[0];
    T denom = (DL * (sL - vL) - DR * (sR - vR)) + Lim<T>::small;
    T sM = fdiv(pR - pL + UL[NM] * (sL - vL) - UR[NM] * (sR - vR), denom);
    T pStar = T(0.5) * (pL + pR + DL * (sL - vL) * (sM - vL) + DR * (sR - vR) * (sM - vR));
    const bool left = sM >= T(0);
    const T *U = left ? UL : UR;
    const T *F = left ? FL : FR;
    T s = left ? sL : sR, p = left ? pL : pR, v = left ? vL : vR;
    T den = s - sM + Lim<T>::small;
    T fac = fdiv(s - v, den);
    T Ust[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) Ust[m] = U[m] * fac;
    Ust[NM] = U[NM] * fac + fdiv(pStar - p, den);
    Ust[4] = U[4] * fac + fdiv(pStar * sM - p * v, den);
#pragma unroll
    for (int m = 0; m < 5; ++m) out[m] = F[m] + s * (Ust[m] - U[m]);
    if constexpr (sizeof(T) == 4) {
        // FP32 only: at a face that opens a vacuum (a cold dense state, whose slower wave speed
        // rounds onto its velocity, beside a near-vacuum state) the contact-speed denominator
        // can cancel to 0, and the star pressure's (s - v)(sM - v) is then 0 x inf.  Where the
        // star state is not finite the face takes the HLLE flux (FP64, whose denominator does not
        // cancel, keeps the reference's HLLC)
        bool ok = true;
#pragma unroll
        for (int m = 0; m < 5; ++m) ok = ok && isfinite(out[m]);
        if (!ok) hlle_vec<5>(UL, UR, FL, FR, sL, sR, out);
    }
}

// ---------------------------------------------------------------- Lorentz factor (FP64)
// 1/sqrt(1 - v2) with the v2 >= 1 guard of core/srhd_core.py:49-51
template <bool FAST = false> AFD_D double lorentz_guarded(double v2) {
    if (v2 >= 1.0) v2 = 1.0 - 1e-14;
    return dinvsqrt<FAST>(1.0 - v2);
}

// ---------------------------------------------------------------- SRHD
// core/srhd_core.py:379-391 (in FP64 on the face states)
AFD_D void srhd_floor(double p_max, double v_max, double *q) {
    if (q[0] < SMALL) q[0] = SMALL;
    if (q[4] < SMALL) q[4] = SMALL;
    if (q[4] > p_max) q[4] = p_max;
    double v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (v2 >= v_max * v_max) {
        double fac = v_max / sqrt(v2 + 1e-32);
        q[1] *= fac;
        q[2] *= fac;
        q[3] *= fac;
    }
}

// core/srhd_core.py:46-59 with the Lorentz factor W supplied
template <typename T> AFD_D void srhd_p2c_W(const Phys<T> &ph, const T *q, T W, T *U) {
    T rho = q[0], p = q[4];
    T h = T(1) + fdiv(ph.gfac * p, pymax(rho, Lim<T>::small));
    T w = rho * h;
    T D = rho * W;
    T wWW = w * W * W;
    U[0] = D;
    U[1] = wWW * q[1];
    U[2] = wWW * q[2];
    U[3] = wWW * q[3];
    U[4] = wWW - p - D;
}

// core/srhd_core.py:156-169
template <int DIR, typename T> AFD_D void srhd_flux(const T *q, const T *U, T *F) {
    T vn = q[1 + DIR], p = q[4];
    F[0] = U[0] * vn;
    F[1] = U[1] * vn;
    F[2] = U[2] * vn;
    F[3] = U[3] * vn;
    F[1 + DIR] = U[1 + DIR] * vn + p;
    F[4] = (U[4] + p) * vn;
}

// core/srhd_core.py:133-140
template <typename T> AFD_D T srhd_sound_speed(const Phys<T> &ph, T rho, T p) {
    T h = T(1) + fdiv(ph.gfac * p, pymax(rho, Lim<T>::small));
    T w = rho * h;
    T cs2 = fdiv(ph.gamma * p, pymax(w, Lim<T>::small));
    if (cs2 < T(0)) cs2 = T(0);
    if (cs2 > Lim<T>::one_minus) cs2 = Lim<T>::one_minus;
    return fsqrt(cs2);
}

// ---------------------------------------------------------------- RMHD
// core/rmhd_core.py:125-135 (FP64, face states)
AFD_D void rmhd_floor(double p_max, double v_max, double *q) { srhd_floor(p_max, v_max, q); }

// Lorentz factor of the reference RMHD p2c after its velocity cap (core/rmhd_core.py:140-148):
// returns W and the capped velocity components in vc.
template <bool FAST = false> AFD_D double rmhd_cap_W(double v_max, const double *q, double *vc) {
    double vx = q[1], vy = q[2], vz = q[3];
    double v2 = vx * vx + vy * vy + vz * vz;
    if (v2 >= v_max * v_max) {
        double fac = v_max / sqrt(v2 + 1e-32);
        vx *= fac;
        vy *= fac;
        vz *= fac;
    }
    vc[0] = vx;
    vc[1] = vy;
    vc[2] = vz;
    return dinvsqrt<FAST>(1.0 - (vx * vx + vy * vy + vz * vz));
}

// core/rmhd_core.py:137-168; vc = capped velocity, W = its Lorentz factor, q = state
// FAST (FP64 only): the divisions by ddiv<true>, as the mixed path's con2prim
template <typename T, bool FAST = false> AFD_D void rmhd_p2c_W(const Phys<T> &ph, const T *q, const T *vc, T W, T *U) {
    auto dv = [](T a, T b) {
        if constexpr (sizeof(T) == 8) return ddiv<FAST>(a, b);
        else return fdiv(a, b);
    };
    T rho = q[0], p = q[4], Bx = q[5], By = q[6], Bz = q[7];
    T vx = vc[0], vy = vc[1], vz = vc[2];
    if (p < Lim<T>::small) p = Lim<T>::small;
    if (p > ph.p_max) p = ph.p_max;
    T B2 = Bx * Bx + By * By + Bz * Bz;
    T h = T(1) + dv(ph.gfac * p, pymax(rho, Lim<T>::small));
    T rhoh = rho * h;
    T vb = vx * Bx + vy * By + vz * Bz;
    T b0 = W * vb;
    T bx = dv(Bx + b0 * vx, W);
    T by = dv(By + b0 * vy, W);
    T bz = dv(Bz + b0 * vz, W);
    T b2 = (bx * bx + by * by + bz * bz) - b0 * b0;
    T rhohWW = rhoh * W * W;
    T wtot = rhohWW + B2;
    T D = rho * W;
    U[0] = D;
    U[1] = wtot * vx - (Bx * vb);
    U[2] = wtot * vy - (By * vb);
    U[3] = wtot * vz - (Bz * vb);
    U[4] = rhohWW - p + T(0.5) * B2 + T(0.5) * b2 - D;
    U[5] = Bx;
    U[6] = By;
    U[7] = Bz;
    U[8] = q[8];
}

// core/rmhd_core.py:343-395 (flux_rmhd_x/y/z) with U = p2c(q)
template <int DIR, typename T> AFD_D void rmhd_flux(const Phys<T> &ph, const T *q, const T *U, T *F) {
    T vx = q[1], vy = q[2], vz = q[3], p = q[4], Bx = q[5], By = q[6], Bz = q[7], psi = q[8];
    T B2 = Bx * Bx + By * By + Bz * Bz;
    T pt = p + T(0.5) * B2;
    T vB = Bx * vx + By * vy + Bz * vz;
    const T v[3] = {vx, vy, vz};
    const T B[3] = {Bx, By, Bz};
    const T vn = v[DIR], Bn = B[DIR];
    F[0] = U[0] * vn;
    F[1] = U[1] * vn - Bn * Bx;
    F[2] = U[2] * vn - Bn * By;
    F[3] = U[3] * vn - Bn * Bz;
    F[1 + DIR] = U[1 + DIR] * vn + pt - Bn * Bn;
    F[4] = (U[4] + pt) * vn - (Bn * vB);
    // induction: F[B_m] = v_m Bn - vn B_m for m != DIR, F[B_DIR] = psi
    F[5] = vx * Bn - vn * Bx;
    F[6] = vy * Bn - vn * By;
    F[7] = vz * Bn - vn * Bz;
    F[5 + DIR] = psi;
    F[8] = ph.glm_ch2 * Bn;
}

// core/rmhd_core.py:419-457 given W of the (uncapped-v, capped-v2) velocity
template <typename T> AFD_D T rmhd_fast_speed_W(const Phys<T> &ph, const T *q, T W) {
    T rho = q[0], p = q[4], vx = q[1], vy = q[2], vz = q[3], Bx = q[5], By = q[6], Bz = q[7];
    if (rho < Lim<T>::small) rho = Lim<T>::small;
    if (p < Lim<T>::small) p = Lim<T>::small;
    if (p > ph.p_max) p = ph.p_max;
    T h = T(1) + fdiv(ph.gfac * p, pymax(rho, Lim<T>::small));
    T rhoh = rho * h;
    T vb = vx * Bx + vy * By + vz * Bz;
    T b0 = W * vb;
    T bx = fdiv(Bx + b0 * vx, W);
    T by = fdiv(By + b0 * vy, W);
    T bz = fdiv(Bz + b0 * vz, W);
    T b2 = (bx * bx + by * by + bz * bz) - b0 * b0;
    T cs2 = fdiv(ph.gamma * p, pymax(rhoh, Lim<T>::small));
    if (cs2 < T(0)) cs2 = T(0);
    if (cs2 > Lim<T>::one_minus) cs2 = Lim<T>::one_minus;
    T ca2 = fdiv(b2, pymax(rhoh + b2, Lim<T>::small));
    if (ca2 < T(0)) ca2 = T(0);
    if (ca2 > Lim<T>::one_minus) ca2 = Lim<T>::one_minus;
    T a = cs2 + ca2 - cs2 * ca2;
    T disc = a * a - T(4) * cs2 * ca2;
    if (disc < T(0)) disc = T(0);
    T cf2 = T(0.5) * (a + fsqrt(disc));
    if (cf2 < T(0)) cf2 = T(0);
    if (cf2 > Lim<T>::one_minus) cf2 = Lim<T>::one_minus;
    return fsqrt(cf2);
}

// W used by rmhd_fast_speed: v2 capped at V_MAX^2 - 1e-14 (core/rmhd_core.py:422-425)
template <bool FAST = false> AFD_D double rmhd_fast_W(double v_max, const double *q) {
    double v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    if (v2 >= v_max * v_max) v2 = v_max * v_max - 1e-14;
    return dinvsqrt<FAST>(1.0 - v2);
}

// core/rmhd_core.py:580-599
template <int NM, typename T>
AFD_D T hllc_contact_speed(const T *UL, const T *UR, T sL, T sR, T pL, T pR, T vL, T vR) {
    T denom = (UL[0] * (sL - vL) - UR[0] * (sR - vR)) + Lim<T>::small;
    return fdiv(pR - pL + UL[NM] * (sL - vL) - UR[NM] * (sR - vR), denom);
}

// p2c of a star state built from T-valued primitives: the Lorentz factor in FP64
template <typename T> AFD_D void rmhd_p2c_state(const Phys<T> &ph, const T *q, T *U) {
    double qd[4] = {0.0, (double)q[1], (double)q[2], (double)q[3]};
    double vcd[3];
    double W = rmhd_cap_W<sizeof(T) == 4>((double)ph.v_max, qd, vcd);
    T vc[3] = {(T)vcd[0], (T)vcd[1], (T)vcd[2]};
    rmhd_p2c_W(ph, q, vc, (T)W, U);
}

// core/rmhd_core.py:601-724 (hll_d_flux_dir).  The star-state fluxes of the reference
// (:703-714) are never used and are not computed; only the star states the selected
// region needs are converted (ULs/Us on the left of the contact, URs/Us on the right).
template <int DIR, typename T>
AFD_D void hlld_full_regions(const Phys<T> &ph, const T *UL, const T *UR, const T *FL, const T *FR, T sL, T sR, T pL,
                     T pR, T vL, T vR, T *out) {
    if (sL >= T(0)) {
#pragma unroll
        for (int m = 0; m < 9; ++m) out[m] = FL[m];
        return;
    }
    if (sR <= T(0)) {
#pragma unroll
        for (int m = 0; m < 9; ++m) out[m] = FR[m];
        return;
    }
    constexpr int vt1 = DIR == 0 ? 2 : 1, vt2 = DIR == 2 ? 2 : 3;
    constexpr int bn = 5 + DIR, bt1 = DIR == 0 ? 6 : 5, bt2 = DIR == 2 ? 6 : 7;
    const T small = Lim<T>::small;
    bool fallback = (UL[0] <= small || UR[0] <= small);
    T Bn = T(0.5) * (UL[bn] + UR[bn]);
    fallback = fallback || (fabs(Bn) < T(1e-8));
    T sM = T(0), DstarL = T(0), DstarR = T(0);
    if (!fallback) {
        sM = hllc_contact_speed<1 + DIR>(UL, UR, sL, sR, pL, pR, vL, vR);
        DstarL = fdiv(UL[0] * (sL - vL), sL - sM + small);
        DstarR = fdiv(UR[0] * (sR - vR), sR - sM + small);
        fallback = (DstarL <= small || DstarR <= small);
    }
    T denomL = UL[0] * (sL - vL) - Bn * Bn;
    T denomR = UR[0] * (sR - vR) - Bn * Bn;
    fallback = fallback || fabs(denomL) < small || fabs(denomR) < small;
    if (fallback) {
        hlle_vec<9>(UL, UR, FL, FR, sL, sR, out);
        return;
    }
    T Bt1L_s = fdiv(UL[bt1] * (sL - vL), sL - sM + small);
    T Bt2L_s = fdiv(UL[bt2] * (sL - vL), sL - sM + small);
    T Bt1R_s = fdiv(UR[bt1] * (sR - vR), sR - sM + small);
    T Bt2R_s = fdiv(UR[bt2] * (sR - vR), sR - sM + small);
    T vt1L = fdiv(UL[vt1], pymax(UL[0], small));
    T vt2L = fdiv(UL[vt2], pymax(UL[0], small));
    T vt1R = fdiv(UR[vt1], pymax(UR[0], small));
    T vt2R = fdiv(UR[vt2], pymax(UR[0], small));
    T vt1L_s = vt1L - fdiv(Bn * (UL[bt1] - Bt1L_s), denomL);
    T vt2L_s = vt2L - fdiv(Bn * (UL[bt2] - Bt2L_s), denomL);
    T vt1R_s = vt1R - fdiv(Bn * (UR[bt1] - Bt1R_s), denomR);
    T vt2R_s = vt2R - fdiv(Bn * (UR[bt2] - Bt2R_s), denomR);
    T sL_star = sM - fdiv(fabs(Bn), fsqrt(DstarL + small));
    T sR_star = sM + fdiv(fabs(Bn), fsqrt(DstarR + small));
    // region of the fan containing x/t = 0 (core/rmhd_core.py:716-724)
    int region;
    if (sL <= T(0) && T(0) <= sL_star) region = 1;
    else if (sL_star <= T(0) && T(0) <= sM) region = 2;
    else if (sM <= T(0) && T(0) <= sR_star) region = 3;
    else if (sR_star <= T(0) && T(0) <= sR) region = 4;
    else region = 0;
    if (region == 0) {
        hlle_vec<9>(UL, UR, FL, FR, sL, sR, out);
        return;
    }
    const bool onL = region <= 2;
    T q[9], Us[9], Uo[9];
    // outer star state (ULs for regions 1-2, URs for 3-4): build_state (:684-693)
    T Dst = onL ? DstarL : DstarR;
    q[0] = pymax(Dst, small);
    q[1 + DIR] = sM;
    q[vt1] = onL ? vt1L_s : vt1R_s;
    q[vt2] = onL ? vt2L_s : vt2R_s;
    q[bn] = Bn;
    q[bt1] = onL ? Bt1L_s : Bt1R_s;
    q[bt2] = onL ? Bt2L_s : Bt2R_s;
    q[4] = (Dst == DstarL) ? pL : pR;
    q[8] = onL ? UL[8] : UR[8];
    rmhd_p2c_state(ph, q, Uo);
    const T *U = onL ? UL : UR;
    const T *F = onL ? FL : FR;
    T s = onL ? sL : sR;
    if (region == 1 || region == 4) {
#pragma unroll
        for (int m = 0; m < 9; ++m) out[m] = F[m] + s * (Uo[m] - U[m]);
        return;
    }
    // double-star state
    T sqrtDL = fsqrt(DstarL), sqrtDR = fsqrt(DstarR);
    T sgn = (Bn >= T(0)) ? T(1) : T(-1);
    T den = sqrtDL + sqrtDR + small;
    T vt1_ss = fdiv(sqrtDL * vt1L_s + sqrtDR * vt1R_s + sgn * (Bt1R_s - Bt1L_s), den);
    T vt2_ss = fdiv(sqrtDL * vt2L_s + sqrtDR * vt2R_s + sgn * (Bt2R_s - Bt2L_s), den);
    T Bt1_ss = fdiv(sqrtDL * Bt1R_s + sqrtDR * Bt1L_s + sgn * sqrtDL * sqrtDR * (vt1L_s - vt1R_s), den);
    T Bt2_ss = fdiv(sqrtDL * Bt2R_s + sqrtDR * Bt2L_s + sgn * sqrtDL * sqrtDR * (vt2L_s - vt2R_s), den);
    T Dss = T(0.5) * (DstarL + DstarR);
    q[0] = pymax(Dss, small);
    q[1 + DIR] = sM;
    q[vt1] = vt1_ss;
    q[vt2] = vt2_ss;
    q[bn] = Bn;
    q[bt1] = Bt1_ss;
    q[bt2] = Bt2_ss;
    q[4] = (Dss == DstarL) ? pL : pR;
    q[8] = T(0.5) * (UL[8] + UR[8]);
    rmhd_p2c_state(ph, q, Us);
    T ss = onL ? sL_star : sR_star;
#pragma unroll
    for (int m = 0; m < 9; ++m) out[m] = F[m] + s * (Uo[m] - U[m]) + ss * (Us[m] - Uo[m]);
}

// hll_d_flux_dir.  In FP32 a face whose star states are not finite (the contact-speed denominator
// cancelling to 0 at a face that opens a vacuum, as in hllc5) takes the HLLE flux; FP64 is the
// reference's solver unchanged.
template <int DIR, typename T>
AFD_D void hlld_full(const Phys<T> &ph, const T *UL, const T *UR, const T *FL, const T *FR, T sL, T sR, T pL,
                     T pR, T vL, T vR, T *out) {
    hlld_full_regions<DIR>(ph, UL, UR, FL, FR, sL, sR, pL, pR, vL, vR, out);
    if constexpr (sizeof(T) == 4) {
        bool ok = true;
#pragma unroll
        for (int m = 0; m < 9; ++m) ok = ok && isfinite(out[m]);
        if (!ok) hlle_vec<9>(UL, UR, FL, FR, sL, sR, out);
    }
}

// ---------------------------------------------------------------- con2prim (FP64)
struct C2PStats {
    int hydro_iters;  // fixed-point iterations of the hydro guess
    int newton_iters; // f(Z) evaluations in the Newton stage
    int flags;        // bit0: hydro fallback, bit1: bisection, bit2: invalid -> hydro guess, bit3: learned accepted
};

// core/srhd_core.py:61-131 and core/rmhd_core.py:170-226.  The loop stops early, with the
// same result, once p stops changing without meeting the tolerance: the map is a pure
// function of p, so the remaining iterations would repeat it and end in the fallback.
template <bool FAST = false>
AFD_D void hydro_c2p(double g, double gfac, double p_max, double v_max, const double *U, double *q, C2PStats &st) {
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4];
    double E = tau + D;
    double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
    double p = (g - 1.0) * (E - D);
    if (p < SMALL) p = SMALL;
    if (p > p_max) p = p_max;
    const double vmax2 = v_max * v_max;
    bool ok = false;
    int it = 0;
    for (; it < 60; ++it) {
        double Wm = E + p;
        double v2 = ddiv<FAST>(S2, Wm * Wm + SMALL);
        if (v2 >= vmax2) v2 = vmax2 - 1e-14;
        double W = dinvsqrt<FAST>(1.0 - v2);
        double rho = ddiv<FAST>(D, pymax(W, SMALL));
        double h = 1.0 + ddiv<FAST>(gfac * p, pymax(rho, SMALL));
        double w = rho * h;
        double f = Wm - w * W * W;
        double dp = -f;
        if (dp > 0.5 * p) dp = 0.5 * p;
        if (dp < -0.5 * p) dp = -0.5 * p;
        double p_new = p + dp;
        if (p_new < SMALL) p_new = SMALL;
        if (p_new > p_max) p_new = p_max;
        if (fabs(dp) < 1e-12 * pymax(1.0, p_new)) {
            p = p_new;
            ok = true;
            ++it;
            break;
        }
        if (p_new == p) {  // fixed point of the clamped map that fails the tolerance
            ++it;
            break;
        }
        p = p_new;
    }
    st.hydro_iters += it;
    if (!ok) {
        st.flags |= 1;
        p = (g - 1.0) * (E - D);
        if (p < SMALL) p = SMALL;
        if (p > p_max) p = p_max;
    }
    double Wm = E + p;
    double vx = Sx / pymax(Wm, SMALL);
    double vy = Sy / pymax(Wm, SMALL);
    double vz = Sz / pymax(Wm, SMALL);
    double v2 = vx * vx + vy * vy + vz * vz;
    if (v2 >= vmax2) {
        double fac = v_max / sqrt(v2 + 1e-32);
        vx *= fac;
        vy *= fac;
        vz *= fac;
    }
    v2 = vx * vx + vy * vy + vz * vz;
    if (v2 >= 1.0) v2 = 1.0 - 1e-14;
    double W = 1.0 / sqrt(1.0 - v2);
    double rho = D / pymax(W, SMALL);
    if (rho < SMALL) rho = SMALL;
    if (p < SMALL) p = SMALL;
    if (p > p_max) p = p_max;
    q[0] = rho;
    q[1] = vx;
    q[2] = vy;
    q[3] = vz;
    q[4] = p;
}

struct RmhdC2PParams {
    double gamma, gfac, p_max, v_max;
};

// core/rmhd_core.py:228-246
template <bool FAST = false>
AFD_D double rmhd_f_of_Z(const RmhdC2PParams &c, double Z, double D, double tau, double B2, double SB, double S2) {
    double denomZ = pymax(Z, SMALL);
    double vb = ddiv<FAST>(SB, denomZ);
    double ZpB = Z + B2;
    double v2 = ddiv<FAST>(S2 + ddiv<FAST>((SB * SB) * (2.0 * Z + B2), denomZ * denomZ), ZpB * ZpB + SMALL);
    double vmax2 = c.v_max * c.v_max;
    if (v2 >= vmax2) v2 = vmax2 - 1e-14;
    double W2 = ddiv<FAST>(1.0, 1.0 - v2);
    double W = FAST ? 1.0 * dinvsqrt<FAST>(1.0 - v2) : sqrt(W2);
    double rho = ddiv<FAST>(D, pymax(W, SMALL));
    double h = ddiv<FAST>(Z, pymax(rho * W2, SMALL));
    double p = (h - 1.0) * rho * (c.gamma - 1.0) / c.gamma;
    if (p < SMALL) p = SMALL;
    if (p > c.p_max) p = c.p_max;
    double b2 = ddiv<FAST>(B2, W2) + vb * vb;
    double tau_calc = Z - p + 0.5 * B2 + 0.5 * b2 - D;
    return tau_calc - tau;
}

// Newton on Z from Z_start (core/rmhd_core.py:273-290): returns true when |f| < tol
template <bool FAST = false>
AFD_D bool rmhd_newton(const RmhdC2PParams &c, double &Z, double D, double tau, double B2, double SB, double S2,
                       double tol, int maxit, C2PStats &st) {
    for (int it = 0; it < maxit; ++it) {
        double f = rmhd_f_of_Z<FAST>(c, Z, D, tau, B2, SB, S2);
        st.newton_iters += 1;
        if (fabs(f) < tol) return true;
        double eps = 1e-6 * pymax(1.0, Z);
        double fp = rmhd_f_of_Z<FAST>(c, Z + eps, D, tau, B2, SB, S2);
        double dfdZ = (fp - f) / eps;
        if (dfdZ == 0.0) return false;
        double dZ = -f / dfdZ;
        if (dZ > 0.5 * Z) dZ = 0.5 * Z;
        if (dZ < -0.5 * Z) dZ = -0.5 * Z;
        Z += dZ;
        if (Z < SMALL) Z = SMALL;
    }
    return false;
}

// Recover the primitives from Z (core/rmhd_core.py:320-341); the hydro guess q0 replaces invalid
// results.  With q0 == null an invalid recovery is reported (false) and q is left unset.
template <bool FAST = false>
AFD_D bool rmhd_recover(const RmhdC2PParams &c, double Z, const double *U, double B2, double SB, const double *q0,
                        double *q, C2PStats &st) {
    double Sx = U[1], Sy = U[2], Sz = U[3], D = U[0], Bx = U[5], By = U[6], Bz = U[7];
    double ZpB = Z + B2;
    double vb = ddiv<FAST>(SB, pymax(Z, SMALL));
    double rZ = ddiv<FAST>(1.0, pymax(ZpB, SMALL));
    double vx = FAST ? (Sx + vb * Bx) * rZ : (Sx + vb * Bx) / pymax(ZpB, SMALL);
    double vy = FAST ? (Sy + vb * By) * rZ : (Sy + vb * By) / pymax(ZpB, SMALL);
    double vz = FAST ? (Sz + vb * Bz) * rZ : (Sz + vb * Bz) / pymax(ZpB, SMALL);
    double v2 = vx * vx + vy * vy + vz * vz;
    double vmax2 = c.v_max * c.v_max;
    if (v2 >= vmax2) {
        double fac = c.v_max / sqrt(v2 + 1e-32);
        vx *= fac;
        vy *= fac;
        vz *= fac;
    }
    double W = dinvsqrt<FAST>(1.0 - (vx * vx + vy * vy + vz * vz));
    double rho = ddiv<FAST>(D, pymax(W, SMALL));
    double h = ddiv<FAST>(Z, pymax(rho * W * W, SMALL));
    double p = ddiv<FAST>((h - 1.0) * rho * (c.gamma - 1.0), c.gamma);
    if (!isfinite(rho) || !isfinite(p) || rho <= 0.0 || p <= 0.0) {
        if (!q0) return false;
        st.flags |= 4;
        rho = q0[0];
        vx = q0[1];
        vy = q0[2];
        vz = q0[3];
        p = q0[4];
    }
    q[0] = rho;
    q[1] = vx;
    q[2] = vy;
    q[3] = vz;
    q[4] = p;
    q[5] = Bx;
    q[6] = By;
    q[7] = Bz;
    q[8] = U[8];
    rmhd_floor(c.p_max, c.v_max, q);
    return true;
}

// core/rmhd_core.py:248-341 (hydro guess, FD-Newton from Z0 then Z1, bisection, recovery)
template <bool FAST = false>
AFD_D void rmhd_c2p(const RmhdC2PParams &c, const double *U, double *q, C2PStats &st) {
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4], Bx = U[5], By = U[6], Bz = U[7];
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double SB = Sx * Bx + Sy * By + Sz * Bz;
    double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
    double tau_h = tau - 0.5 * B2;
    if (tau_h < SMALL) tau_h = SMALL;
    double Uh[5] = {D, Sx, Sy, Sz, tau_h}, q0[5];
    hydro_c2p<FAST>(c.gamma, c.gfac, c.p_max, c.v_max, Uh, q0, st);
    double vmax2 = c.v_max * c.v_max;
    double v20 = q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3];
    if (v20 >= vmax2) v20 = vmax2 - 1e-14;
    double W0 = 1.0 / sqrt(1.0 - v20);
    double h0 = 1.0 + c.gfac * q0[4] / pymax(q0[0], SMALL);
    double Z0 = pymax(q0[0] * h0 * W0 * W0, SMALL);
    double Z1 = pymax(D + tau + B2, SMALL);
    double tol = 1e-10 * pymax(1.0, tau);
    double Z = Z0;
    bool ok = rmhd_newton<FAST>(c, Z, D, tau, B2, SB, S2, tol, 50, st);
    if (!ok) {
        Z = Z1;
        ok = rmhd_newton<FAST>(c, Z, D, tau, B2, SB, S2, tol, 50, st);
    }
    if (!ok) {
        st.flags |= 2;
        double Zmin = SMALL;
        double Zmax = pymax(pymax(Z, Z1), 1.0);
        double fmin = rmhd_f_of_Z<FAST>(c, Zmin, D, tau, B2, SB, S2);
        double fmax = rmhd_f_of_Z<FAST>(c, Zmax, D, tau, B2, SB, S2);
        for (int it = 0; it < 40; ++it) {
            if (fmin * fmax <= 0.0) break;
            Zmax *= 2.0;
            fmax = rmhd_f_of_Z<FAST>(c, Zmax, D, tau, B2, SB, S2);
        }
        if (fmin * fmax <= 0.0) {
            for (int it = 0; it < 60; ++it) {
                double Zmid = 0.5 * (Zmin + Zmax);
                double fmid = rmhd_f_of_Z<FAST>(c, Zmid, D, tau, B2, SB, S2);
                if (fabs(fmid) < tol) {
                    Z = Zmid;
                    break;
                }
                if (fmin * fmid <= 0.0) {
                    Zmax = Zmid;
                    fmax = fmid;
                } else {
                    Zmin = Zmid;
                    fmin = fmid;
                }
                Z = 0.5 * (Zmin + Zmax);
            }
        }
    }
    rmhd_recover<FAST>(c, Z, U, B2, SB, q0, q, st);
}

// Learned-guess con2prim: FD-Newton (the reference's iteration) from the network's Z0, accepted by a
// residual |f(Z)| < 1e-12 max(1, tau), a hundred times tighter than the reference's stopping test, and
// by a valid recovery.  Returns false when the acceptance test fails; the caller then runs rmhd_c2p.
template <bool FAST = false>
AFD_D bool rmhd_c2p_from_guess(const RmhdC2PParams &c, const double *U, double Z0, double *q, C2PStats &st) {
    if (!(Z0 > SMALL) || !isfinite(Z0)) return false;
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4], Bx = U[5], By = U[6], Bz = U[7];
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double SB = Sx * Bx + Sy * By + Sz * Bz;
    double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
    double Z = Z0;
    if (!rmhd_newton<FAST>(c, Z, D, tau, B2, SB, S2, 1e-12 * pymax(1.0, tau), 12, st)) return false;
    if (!rmhd_recover<FAST>(c, Z, U, B2, SB, nullptr, q, st)) return false;
    st.flags |= 8;
    return true;
}

// FP64 p2c of a cell state (U0 of the SSPRK stages): core/srhd_core.py:46-59, core/rmhd_core.py:137-168
AFD_D void srhd_p2c_cell(double gfac, const double *q, double *U) {
    double v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double W = lorentz_guarded(v2);
    Phys<double> ph;
    ph.gfac = gfac;
    srhd_p2c_W(ph, q, W, U);
}
// FAST: the Lorentz factor and the divisions by the hardware reciprocals (the mixed path's U0)
template <bool FAST = false>
AFD_D void rmhd_p2c_cell(double gfac, double p_max, double v_max, const double *q, double *U) {
    double vc[3];
    double W = rmhd_cap_W<FAST>(v_max, q, vc);
    Phys<double> ph;
    ph.gfac = gfac;
    ph.p_max = p_max;
    rmhd_p2c_W<double, FAST>(ph, q, vc, W, U);
}

}  // namespace afd
