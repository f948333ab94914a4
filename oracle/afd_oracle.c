/*
 * afd_oracle.c -- CPU FP64 restatement of AstroFD's SRHD/RMHD finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY (the results of the GPU library are compared with it): loaded
 * by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg; the product path never links or calls it.
 *
 * Every function follows the reference file:line it cites, in the reference's
 * evaluation order.  Build with -ffp-contract=off and without fast-math (see
 * oracle/Makefile) so each operation rounds exactly as numpy/CPython does:
 *   - "x**2" is libm pow(x, 2.0), as in numpy and CPython,
 *   - Python max(a, b) keeps a unless b > a; min(a, b) keeps a unless b < a,
 *   - negative indices in the reference's passive stencils wrap (Python indexing).
 *
 * Whole-array loops are parallel over the outermost index with OpenMP; the RMHD
 * PPM/WENO "carried" HLL speeds (Q3) are re-seeded at each x-plane from the
 * predecessor cell in i->j->k order, which reproduces the serial loop exactly.
 */
#include "afd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SMALL 1e-12

static inline double pymax(double a, double b) { return (b > a) ? b : a; }
static inline double pymin(double a, double b) { return (b < a) ? b : a; }
/* "x**2" in the reference is libm pow(x, 2.0) for both numpy float64 scalars and CPython floats;
 * glibc pow is not always the correctly rounded x*x, and gcc folds a literal pow(x, 2.0) into x*x,
 * so the call goes through a volatile pointer. */
static double (*volatile libm_pow)(double, double) = pow;
static inline double pow2(double x) { return libm_pow(x, 2.0); }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

#define NXF (P->nx)
#define NYF (P->ny)
#define NZF (P->nz)
#define IX(v, i, j, k) ((((size_t)(v) * NXF + (size_t)(i)) * NYF + (size_t)(j)) * NZF + (size_t)(k))
#define PR(v, i, j, k) pr[IX(v, i, j, k)]

/* =====================================================================
 * Shared scalar helpers (core/srhd_core.py and core/rmhd_core.py have
 * identical copies of these)
 * ===================================================================== */

/* core/srhd_core.py:305-311, core/rmhd_core.py:52-58 */
static double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    if (fabs(a) < fabs(b)) return a;
    return b;
}

/* core/srhd_core.py:313-325 (mc_lim), core/rmhd_core.py:60-72 (limiter) */
double orc_limiter(int lid, double dqL, double dqR) {
    if (lid == 1) return minmod(dqL, dqR);
    if (lid == 2) {
        if (dqL * dqR <= 0.0) return 0.0;
        double denom = dqL + dqR;
        if (denom == 0.0) return 0.0;
        return (2.0 * dqL * dqR) / denom;
    }
    double mm = minmod(dqL, dqR);
    return minmod(0.5 * (dqL + dqR), 2.0 * mm);
}

/* core/srhd_core.py:327-341, core/rmhd_core.py:74-87 */
static void ppm_reconstruct(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2,
                            double *outL, double *outR) {
    double qL = (7.0 / 12.0) * (q_im1 + q_i) - (1.0 / 12.0) * (q_im2 + q_ip1);
    double qR = (7.0 / 12.0) * (q_i + q_ip1) - (1.0 / 12.0) * (q_im1 + q_ip2);
    double qmax = pymax(pymax(q_im1, q_i), q_ip1);
    double qmin = pymin(pymin(q_im1, q_i), q_ip1);
    if (qL > qmax) qL = qmax;
    if (qL < qmin) qL = qmin;
    if (qR > qmax) qR = qmax;
    if (qR < qmin) qR = qmin;
    if ((qR - q_i) * (q_i - qL) <= 0.0) {
        qL = q_i;
        qR = q_i;
    }
    *outL = qL;
    *outR = qR;
}

void orc_ppm(const double *q, double *lr) { ppm_reconstruct(q[0], q[1], q[2], q[3], q[4], &lr[0], &lr[1]); }

/* core/srhd_core.py:343-359, core/rmhd_core.py:89-105 (WENO5-JS, Jiang & Shu 1996) */
static double weno5_left(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2) {
    const double eps = 1e-12;
    double IS0 = (13.0 / 12.0) * pow2(q_im2 - 2.0 * q_im1 + q_i) + 0.25 * pow2(q_im2 - 4.0 * q_im1 + 3.0 * q_i);
    double IS1 = (13.0 / 12.0) * pow2(q_im1 - 2.0 * q_i + q_ip1) + 0.25 * pow2(q_im1 - q_ip1);
    double IS2 = (13.0 / 12.0) * pow2(q_i - 2.0 * q_ip1 + q_ip2) + 0.25 * pow2(3.0 * q_i - 4.0 * q_ip1 + q_ip2);
    double a0 = 0.1 / pow2(eps + IS0);
    double a1 = 0.6 / pow2(eps + IS1);
    double a2 = 0.3 / pow2(eps + IS2);
    double s = a0 + a1 + a2;
    double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    double p0 = (1.0 / 3.0) * q_im2 - (7.0 / 6.0) * q_im1 + (11.0 / 6.0) * q_i;
    double p1 = (-1.0 / 6.0) * q_im1 + (5.0 / 6.0) * q_i + (1.0 / 3.0) * q_ip1;
    double p2 = (1.0 / 3.0) * q_i + (5.0 / 6.0) * q_ip1 - (1.0 / 6.0) * q_ip2;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

/* core/srhd_core.py:361-377, core/rmhd_core.py:107-123 (mirror image of weno5_left) */
static double weno5_right(double q_ip2, double q_ip1, double q_i, double q_im1, double q_im2) {
    const double eps = 1e-12;
    double IS0 = (13.0 / 12.0) * pow2(q_ip2 - 2.0 * q_ip1 + q_i) + 0.25 * pow2(q_ip2 - 4.0 * q_ip1 + 3.0 * q_i);
    double IS1 = (13.0 / 12.0) * pow2(q_ip1 - 2.0 * q_i + q_im1) + 0.25 * pow2(q_ip1 - q_im1);
    double IS2 = (13.0 / 12.0) * pow2(q_i - 2.0 * q_im1 + q_im2) + 0.25 * pow2(3.0 * q_i - 4.0 * q_im1 + q_im2);
    double a0 = 0.1 / pow2(eps + IS0);
    double a1 = 0.6 / pow2(eps + IS1);
    double a2 = 0.3 / pow2(eps + IS2);
    double s = a0 + a1 + a2;
    double w0 = a0 / s, w1 = a1 / s, w2 = a2 / s;
    double p0 = (1.0 / 3.0) * q_ip2 - (7.0 / 6.0) * q_ip1 + (11.0 / 6.0) * q_i;
    double p1 = (-1.0 / 6.0) * q_ip1 + (5.0 / 6.0) * q_i + (1.0 / 3.0) * q_im1;
    double p2 = (1.0 / 3.0) * q_i + (5.0 / 6.0) * q_im1 - (1.0 / 6.0) * q_im2;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

double orc_weno5_left(const double *q) { return weno5_left(q[0], q[1], q[2], q[3], q[4]); }
double orc_weno5_right(const double *q) { return weno5_right(q[0], q[1], q[2], q[3], q[4]); }

/* core/srhd_core.py:180-186, core/rmhd_core.py:411-417 */
static double hlle_scalar(double qL, double qR, double fL, double fR, double sL, double sR) {
    if (sL >= 0.0) return fL;
    if (sR <= 0.0) return fR;
    return (sR * fL - sL * fR + sL * sR * (qR - qL)) / (sR - sL + SMALL);
}

/* core/srhd_core.py:171-178, core/rmhd_core.py:402-409 (vector form over n components) */
static void hlle_vec(int n, const double *UL, const double *UR, const double *FL, const double *FR,
                     double sL, double sR, double *out) {
    if (sL >= 0.0) {
        for (int m = 0; m < n; ++m) out[m] = FL[m];
    } else if (sR <= 0.0) {
        for (int m = 0; m < n; ++m) out[m] = FR[m];
    } else {
        for (int m = 0; m < n; ++m)
            out[m] = (sR * FL[m] - sL * FR[m] + sL * sR * (UR[m] - UL[m])) / (sR - sL + SMALL);
    }
}

/* core/srhd_core.py:142-154, core/rmhd_core.py:459-471 */
static void eig_speeds(double vn, double cs, double *lm, double *lp) {
    double dp = 1.0 + vn * cs;
    double dm = 1.0 - vn * cs;
    if (dp == 0.0) dp = SMALL;
    if (dm == 0.0) dm = SMALL;
    double p = (vn + cs) / dp;
    double m = (vn - cs) / dm;
    if (p > 1.0) p = 1.0;
    if (p < -1.0) p = -1.0;
    if (m > 1.0) m = 1.0;
    if (m < -1.0) m = -1.0;
    *lm = m;
    *lp = p;
}

/* Newtonian-form HLLC (Q13) on the first five conserved components; normal momentum index
 * nm = 1 + dir.  core/srhd_core.py:188-282 (hllc_flux_x/y/z), core/rmhd_core.py:473-567. */
static void hllc5(int nm, const double *UL, const double *UR, const double *FL, const double *FR, double sL,
                  double sR, double pL, double pR, double vL, double vR, double *out) {
    double DL = UL[0], DR = UR[0];
    double SnL = UL[nm], SnR = UR[nm];
    double denom = (DL * (sL - vL) - DR * (sR - vR)) + SMALL;
    double sM = (pR - pL + SnL * (sL - vL) - SnR * (sR - vR)) / denom;
    double pStar = 0.5 * (pL + pR + DL * (sL - vL) * (sM - vL) + DR * (sR - vR) * (sM - vR));
    if (sL >= 0.0) {
        for (int m = 0; m < 5; ++m) out[m] = FL[m];
        return;
    }
    if (sR <= 0.0) {
        for (int m = 0; m < 5; ++m) out[m] = FR[m];
        return;
    }
    double Ust[5];
    if (sM >= 0.0) {
        double fac = (sL - vL) / (sL - sM + SMALL);
        for (int m = 0; m < 5; ++m) Ust[m] = UL[m] * fac;
        Ust[nm] = UL[nm] * fac + (pStar - pL) / (sL - sM + SMALL);
        Ust[4] = UL[4] * fac + (pStar * sM - pL * vL) / (sL - sM + SMALL);
        for (int m = 0; m < 5; ++m) out[m] = FL[m] + sL * (Ust[m] - UL[m]);
    } else {
        double fac = (sR - vR) / (sR - sM + SMALL);
        for (int m = 0; m < 5; ++m) Ust[m] = UR[m] * fac;
        Ust[nm] = UR[nm] * fac + (pStar - pR) / (sR - sM + SMALL);
        Ust[4] = UR[4] * fac + (pStar * sM - pR * vR) / (sR - sM + SMALL);
        for (int m = 0; m < 5; ++m) out[m] = FR[m] + sR * (Ust[m] - UR[m]);
    }
}

/* =====================================================================
 * SRHD (core/srhd_core.py)
 * ===================================================================== */

/* core/srhd_core.py:46-59 */
static void srhd_p2c(double g, const double *q, double *U) {
    double rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4];
    double v2 = vx * vx + vy * vy + vz * vz;
    if (v2 >= 1.0) v2 = 1.0 - 1e-14;
    double W = 1.0 / sqrt(1.0 - v2);
    double h = 1.0 + g / (g - 1.0) * p / fmax(rho, SMALL);
    double w = rho * h;
    double D = rho * W;
    U[0] = D;
    U[1] = w * W * W * vx;
    U[2] = w * W * W * vy;
    U[3] = w * W * W * vz;
    U[4] = w * W * W - p - D;
}

/* core/srhd_core.py:61-131 (also core/rmhd_core.py:170-226 with GAMMA explicit): the damped
 * fixed point with dfdp = 1 (Q4), 60 iterations, fallback, velocity cap, rho from final W. */
static void hydro_c2p(double g, double p_max, double v_max, const double *U, double *q) {
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4];
    double E = tau + D;
    double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
    double p = (g - 1.0) * (E - D);
    if (p < SMALL) p = SMALL;
    if (p > p_max) p = p_max;
    int ok = 0;
    for (int it = 0; it < 60; ++it) {
        double Wm = E + p;
        double v2 = S2 / (Wm * Wm + SMALL);
        if (v2 >= v_max * v_max) v2 = v_max * v_max - 1e-14;
        double W = 1.0 / sqrt(1.0 - v2);
        double rho = D / pymax(W, SMALL);
        double h = 1.0 + g / (g - 1.0) * p / pymax(rho, SMALL);
        double w = rho * h;
        double f = Wm - w * W * W;
        double dfdp = 1.0;
        double dp = -f / dfdp;
        if (dp > 0.5 * p) dp = 0.5 * p;
        if (dp < -0.5 * p) dp = -0.5 * p;
        double p_new = p + dp;
        if (p_new < SMALL) p_new = SMALL;
        if (p_new > p_max) p_new = p_max;
        if (fabs(dp) < 1e-12 * pymax(1.0, p_new)) {
            p = p_new;
            ok = 1;
            break;
        }
        p = p_new;
    }
    if (!ok) {
        p = (g - 1.0) * (E - D);
        if (p < SMALL) p = SMALL;
        if (p > p_max) p = p_max;
    }
    double Wm = E + p;
    double vx = Sx / pymax(Wm, SMALL);
    double vy = Sy / pymax(Wm, SMALL);
    double vz = Sz / pymax(Wm, SMALL);
    double v2 = vx * vx + vy * vy + vz * vz;
    double vmax2 = v_max * v_max;
    if (v2 >= vmax2) {
        double fac = v_max / sqrt(v2 + 1e-32);
        vx *= fac; vy *= fac; vz *= fac;
    }
    v2 = vx * vx + vy * vy + vz * vz;
    if (v2 >= 1.0) v2 = 1.0 - 1e-14;
    double W = 1.0 / sqrt(1.0 - v2);
    double rho = D / pymax(W, SMALL);
    if (rho < SMALL) rho = SMALL;
    if (p < SMALL) p = SMALL;
    if (p > p_max) p = p_max;
    q[0] = rho; q[1] = vx; q[2] = vy; q[3] = vz; q[4] = p;
}

/* core/srhd_core.py:133-140 */
static double srhd_sound_speed(double g, double rho, double p) {
    double h = 1.0 + g / (g - 1.0) * p / fmax(rho, SMALL);
    double w = rho * h;
    double cs2 = g * p / fmax(w, SMALL);
    if (cs2 < 0.0) cs2 = 0.0;
    if (cs2 > 1.0 - 1e-14) cs2 = 1.0 - 1e-14;
    return sqrt(cs2);
}

/* core/srhd_core.py:156-169 */
static void srhd_flux(double g, int dir, const double *q, double *F) {
    double U[5];
    srhd_p2c(g, q, U);
    double vn = q[1 + dir], p = q[4];
    F[0] = U[0] * vn;
    F[1] = U[1] * vn;
    F[2] = U[2] * vn;
    F[3] = U[3] * vn;
    F[1 + dir] = U[1 + dir] * vn + p;
    F[4] = (U[4] + p) * vn;
}

/* core/srhd_core.py:379-391 */
static void srhd_floor(double p_max, double v_max, double *q) {
    if (q[0] < SMALL) q[0] = SMALL;
    if (q[4] < SMALL) q[4] = SMALL;
    if (q[4] > p_max) q[4] = p_max;
    double v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double vmax2 = v_max * v_max;
    if (v2 >= vmax2) {
        double fac = v_max / sqrt(v2 + 1e-32);
        q[1] *= fac; q[2] *= fac; q[3] *= fac;
    }
}

/* apply_diss_flux_x/y/z (core/srhd_core.py:401-462): the bulk pressure Pi, the shear stress
 * pi^ij and the heat flux q^i of cell `cell` (fields 5..14) added to a physical flux, with the
 * face state's velocities. */
static void diss_flux(const orc_params *P, const double *pr, size_t cell, int dir, const double *q, double *F) {
    const size_t N = (size_t)NXF * NYF * NZF;
    const double pi = pr[5 * N + cell], pixx = pr[6 * N + cell], piyy = pr[7 * N + cell], pizz = pr[8 * N + cell];
    const double pixy = pr[9 * N + cell], pixz = pr[10 * N + cell], piyz = pr[11 * N + cell];
    const double qx = pr[12 * N + cell], qy = pr[13 * N + cell], qz = pr[14 * N + cell];
    const double vx = q[1], vy = q[2], vz = q[3];
    if (dir == 0) {
        F[1] += pi + pixx;
        F[2] += pixy;
        F[3] += pixz;
        F[4] += vx * (pi + pixx) + vy * pixy + vz * pixz + qx;
    } else if (dir == 1) {
        F[1] += pixy;
        F[2] += pi + piyy;
        F[3] += piyz;
        F[4] += vx * pixy + vy * (pi + piyy) + vz * piyz + qy;
    } else {
        F[1] += pixz;
        F[2] += piyz;
        F[3] += pi + pizz;
        F[4] += vx * pixz + vy * piyz + vz * (pi + pizz) + qz;
    }
}

/* One SRHD face: floors, p2c, physical fluxes, eigen-speed bounds and the Riemann solver
 * (core/srhd_core.py:536-549 and the identical blocks of every reconstruction).
 * qL/qR are floored in place; returns sL/sR.  With `pr` (the RHS input) the dissipative hooks
 * of the face's cells cL | cR (flat indices of variable 0) apply: add_pi_lr (:394-399, MUSCL
 * faces only, when PASSIVE_OFFSET > 5) before the floors, and apply_diss_flux (nv >= 15) once
 * per MUSCL face and twice per PPM face (:958-961); the WENO RHS has neither. */
static void srhd_face(const orc_params *P, int dir, double *qL, double *qR, double *F, double *sLo, double *sRo,
                      const double *pr, size_t cL, size_t cR) {
    double g = P->gamma;
    const size_t N = (size_t)NXF * NYF * NZF;
    if (pr && P->recon_id == 0 && P->passive_offset > 5) {
        qL[4] += pr[5 * N + cL];
        qR[4] += pr[5 * N + cR];
    }
    srhd_floor(P->p_max, P->v_max, qL);
    srhd_floor(P->p_max, P->v_max, qR);
    double UL[5], UR[5], FL[5], FR[5];
    srhd_p2c(g, qL, UL);
    srhd_p2c(g, qR, UR);
    srhd_flux(g, dir, qL, FL);
    srhd_flux(g, dir, qR, FR);
    if (pr && P->nv >= 15) {
        const int reps = P->recon_id == 0 ? 1 : (P->recon_id == 1 ? 2 : 0);
        for (int r = 0; r < reps; ++r) {
            diss_flux(P, pr, cL, dir, qL, FL);
            diss_flux(P, pr, cR, dir, qR, FR);
        }
    }
    double csL = srhd_sound_speed(g, qL[0], qL[4]), csR = srhd_sound_speed(g, qR[0], qR[4]);
    double lmL, lpL, lmR, lpR;
    eig_speeds(qL[1 + dir], csL, &lmL, &lpL);
    eig_speeds(qR[1 + dir], csR, &lmR, &lpR);
    double sL = pymin(lmL, lmR), sR = pymax(lpL, lpR);
    /* core/srhd_core.py:284-300 */
    if (P->riemann_id == 1)
        hllc5(1 + dir, UL, UR, FL, FR, sL, sR, qL[4], qR[4], qL[1 + dir], qR[1 + dir], F);
    else
        hlle_vec(5, UL, UR, FL, FR, sL, sR, F);
    *sLo = sL;
    *sRo = sR;
}

void orc_srhd_p2c_batch(const orc_params *P, int64_t n, const double *prim, double *cons) {
    for (int64_t m = 0; m < n; ++m) srhd_p2c(P->gamma, prim + 5 * m, cons + 5 * m);
}
void orc_srhd_c2p_batch(const orc_params *P, int64_t n, const double *cons, double *prim) {
    for (int64_t m = 0; m < n; ++m) hydro_c2p(P->gamma, P->p_max, P->v_max, cons + 5 * m, prim + 5 * m);
}
void orc_srhd_sound_speed_batch(const orc_params *P, int64_t n, const double *prim, double *cs) {
    for (int64_t m = 0; m < n; ++m) cs[m] = srhd_sound_speed(P->gamma, prim[5 * m], prim[5 * m + 4]);
}
void orc_srhd_flux_batch(const orc_params *P, int64_t n, int dir, const double *prim, double *flux) {
    for (int64_t m = 0; m < n; ++m) srhd_flux(P->gamma, dir, prim + 5 * m, flux + 5 * m);
}
void orc_srhd_riemann(const orc_params *P, int dir, const double *primL, const double *primR, double *F) {
    double qL[5], qR[5], sL, sR;
    memcpy(qL, primL, sizeof qL);
    memcpy(qR, primR, sizeof qR);
    srhd_face(P, dir, qL, qR, F, &sL, &sR, NULL, 0, 0);
}

/* ---------------------------------------------------------------------
 * Face states of the face between cells c-1 and c along `dir`, for a
 * variable v, with the three reconstructions.  (i,j,k) is cell c.
 * --------------------------------------------------------------------- */
static inline size_t dstride(const orc_params *P, int dir) {
    return dir == 0 ? (size_t)NYF * NZF : (dir == 1 ? (size_t)NZF : 1);
}

/* MUSCL: qL = q[c-1] + slope(c-1)/2, qR = q[c] - slope(c)/2 (core/srhd_core.py:486-533) */
static void muscl_face(const orc_params *P, const double *pr, size_t c, size_t s, double *qL, double *qR) {
    double qm2 = pr[c - 2 * s], qm1 = pr[c - s], q0 = pr[c], qp1 = pr[c + s];
    double slL = orc_limiter(P->limiter_id, qm1 - qm2, q0 - qm1);
    double slR = orc_limiter(P->limiter_id, q0 - qm1, qp1 - q0);
    *qL = qm1 + 0.5 * slL;
    *qR = q0 - 0.5 * slR;
}

/* PPM: qL = right value of the parabola in c-1, qR = left value in c (core/srhd_core.py:943-950) */
static void ppm_face(const double *pr, size_t c, size_t s, double *qL, double *qR) {
    double l_im1, r_im1, l_i, r_i;
    ppm_reconstruct(pr[c - 3 * s], pr[c - 2 * s], pr[c - s], pr[c], pr[c + s], &l_im1, &r_im1);
    ppm_reconstruct(pr[c - 2 * s], pr[c - s], pr[c], pr[c + s], pr[c + 2 * s], &l_i, &r_i);
    *qL = r_im1;
    *qR = l_i;
}

/* WENO5 (core/srhd_core.py:1183-1184) */
static void weno_face(const double *pr, size_t c, size_t s, double *qL, double *qR) {
    *qL = weno5_left(pr[c - 3 * s], pr[c - 2 * s], pr[c - s], pr[c], pr[c + s]);
    *qR = weno5_right(pr[c + 2 * s], pr[c + s], pr[c], pr[c - s], pr[c - 2 * s]);
}

static void face_states(const orc_params *P, const double *pr, int nvar, int dir, int i, int j, int k,
                        double *qL, double *qR) {
    size_t s = dstride(P, dir);
    for (int v = 0; v < nvar; ++v) {
        size_t c = IX(v, i, j, k);
        if (P->recon_id == 1)
            ppm_face(pr, c, s, &qL[v], &qR[v]);
        else if (P->recon_id == 2)
            weno_face(pr, c, s, &qL[v], &qR[v]);
        else
            muscl_face(P, pr, c, s, &qL[v], &qR[v]);
    }
}

/* Passive-scalar stencil reads with Python negative-index wrap (Q15).  Indices >= n wrap too:
 * the reference reads one plane past the end there (an IndexError in pure Python). */
static inline double pv(const orc_params *P, const double *pr, int v, int i, int j, int k) {
    if (i < 0) i += NXF;
    if (j < 0) j += NYF;
    if (k < 0) k += NZF;
    if (i >= NXF) i -= NXF;
    if (j >= NYF) j -= NYF;
    if (k >= NZF) k -= NZF;
    return PR(v, i, j, k);
}
/* value of passive v at offset o along dir from cell (i,j,k) */
static inline double pvd(const orc_params *P, const double *pr, int v, int dir, int i, int j, int k, int o) {
    return pv(P, pr, v, i + (dir == 0 ? o : 0), j + (dir == 1 ? o : 0), k + (dir == 2 ? o : 0));
}

/* The three passive-flux variants of the reference for the face between (c-1) and c along dir:
 * 0 = WENO5 (core/srhd_core.py:554-556), 1 = PPM (:561-567), 2 = MUSCL (:572-580).
 * (i,j,k) is cell c; vL/vR are the floored normal velocities of the face. */
static double passive_face_flux(const orc_params *P, const double *pr, int variant, int v, int dir, int i, int j,
                                int k, double vL, double vR, double sL, double sR) {
    double tL, tR;
#define Q(o) pvd(P, pr, v, dir, i, j, k, (o))
    if (variant == 0) {
        tL = weno5_left(Q(-3), Q(-2), Q(-1), Q(0), Q(1));
        tR = weno5_right(Q(2), Q(1), Q(0), Q(-1), Q(-2));
    } else if (variant == 1) {
        double l_im1, r_im1, l_i, r_i;
        ppm_reconstruct(Q(-3), Q(-2), Q(-1), Q(0), Q(1), &l_im1, &r_im1);
        ppm_reconstruct(Q(-2), Q(-1), Q(0), Q(1), Q(2), &l_i, &r_i);
        tL = r_im1;
        tR = l_i;
    } else {
        double slL = orc_limiter(P->limiter_id, Q(-1) - Q(-2), Q(0) - Q(-1));
        double slR = orc_limiter(P->limiter_id, Q(0) - Q(-1), Q(1) - Q(0));
        tL = Q(-1) + 0.5 * slL;
        tR = Q(0) - 0.5 * slR;
    }
#undef Q
    return hlle_scalar(tL, tR, tL * vL, tR * vR, sL, sR);
}

/* MUSCL passive right-face flux as written in the reference (slopes recomputed at c and c+1):
 * core/srhd_core.py:651-659.  Algebraically the same as the left face of cell c+1. */

/* core/srhd_core.py:468-924 (compute_rhs_muscl) */
static void srhd_rhs_muscl(const orc_params *P, const double *pr, double *rhs, double dx, double dy, double dz) {
    const int np = P->n_passive, po = P->passive_offset;
    const int i0 = 2, i1 = NXF - 3, j0 = 2, j1 = NYF - 3, k0 = 2, k1 = NZF - 3;
    const double dd[3] = {dx, dy, dz};
    for (int dir = 0; dir < 3; ++dir) {
#pragma omp parallel for schedule(static)
        for (int i = i0; i < i1; ++i)
            for (int j = j0; j < j1; ++j)
                for (int k = k0; k < k1; ++k) {
                    int ip = i + (dir == 0), jp = j + (dir == 1), kp = k + (dir == 2);
                    double qL[5], qR[5], FA[5], FB[5], sL, sR, sL2, sR2;
                    const size_t sd = dstride(P, dir), ca = IX(0, i, j, k), cb = IX(0, ip, jp, kp);
                    face_states(P, pr, 5, dir, i, j, k, qL, qR);
                    srhd_face(P, dir, qL, qR, FA, &sL, &sR, pr, ca - sd, ca);
                    double vLa = qL[1 + dir], vRa = qR[1 + dir];
                    double trL[8];
                    /* x: WENO, PPM and MUSCL blocks each overwrite FxL_tr (:550-580); z: MUSCL only (:847-859) */
                    if (np > 0 && dir != 1) {
                        for (int t = 0; t < np; ++t) {
                            if (dir == 0) {
                                trL[t] = passive_face_flux(P, pr, 0, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                                trL[t] = passive_face_flux(P, pr, 1, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                            }
                            trL[t] = passive_face_flux(P, pr, 2, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                        }
                    }
                    face_states(P, pr, 5, dir, ip, jp, kp, qL, qR);
                    srhd_face(P, dir, qL, qR, FB, &sL2, &sR2, pr, cb - sd, cb);
                    double vLb = qL[1 + dir], vRb = qR[1 + dir];
                    if (np > 0 && dir != 1) {
                        for (int t = 0; t < np; ++t) {
                            size_t o = IX(5 + t, i, j, k);
                            if (dir == 0) {
                                double a = passive_face_flux(P, pr, 0, po + t, dir, ip, jp, kp, vLb, vRb, sL2, sR2);
                                rhs[o] -= (a - trL[t]) / dd[dir];
                                double b = passive_face_flux(P, pr, 1, po + t, dir, ip, jp, kp, vLb, vRb, sL2, sR2);
                                rhs[o] -= (b - trL[t]) / dd[dir];
                            }
                            double c = passive_face_flux(P, pr, 2, po + t, dir, ip, jp, kp, vLb, vRb, sL2, sR2);
                            rhs[o] -= (c - trL[t]) / dd[dir];
                        }
                    }
                    for (int m = 0; m < 5; ++m) rhs[IX(m, i, j, k)] -= (FB[m] - FA[m]) / dd[dir];
                }
    }
}

/* core/srhd_core.py:926-1162 (compute_rhs_ppm) and :1166-1262 (compute_rhs_weno) */
static void srhd_rhs_ppm_weno(const orc_params *P, const double *pr, double *rhs, double dx, double dy,
                              double dz) {
    const int np = P->n_passive, po = P->passive_offset, ppm = (P->recon_id == 1);
    const int i0 = 3, i1 = NXF - 3, j0 = 3, j1 = NYF - 3, k0 = 3, k1 = NZF - 3;
    const double dd[3] = {dx, dy, dz};
#pragma omp parallel for schedule(static)
    for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j)
            for (int k = k0; k < k1; ++k) {
                for (int dir = 0; dir < 3; ++dir) {
                    int ip = i + (dir == 0), jp = j + (dir == 1), kp = k + (dir == 2);
                    double qL[5], qR[5], FA[5], FB[5], sL, sR, sL2, sR2, trL[8];
                    const size_t sd = dstride(P, dir), ca = IX(0, i, j, k), cb = IX(0, ip, jp, kp);
                    face_states(P, pr, 5, dir, i, j, k, qL, qR);
                    srhd_face(P, dir, qL, qR, FA, &sL, &sR, pr, ca - sd, ca);
                    double vLa = qL[1 + dir], vRa = qR[1 + dir];
                    /* PPM passives: y has WENO, PPM and MUSCL blocks (:1007-1037), z has WENO and PPM (:1110-1127) */
                    int pas = ppm && np > 0 && dir != 0;
                    if (pas) {
                        for (int t = 0; t < np; ++t) {
                            trL[t] = passive_face_flux(P, pr, 0, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                            trL[t] = passive_face_flux(P, pr, 1, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                            if (dir == 1) trL[t] = passive_face_flux(P, pr, 2, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                        }
                    }
                    face_states(P, pr, 5, dir, ip, jp, kp, qL, qR);
                    srhd_face(P, dir, qL, qR, FB, &sL2, &sR2, pr, cb - sd, cb);
                    double vLb = qL[1 + dir], vRb = qR[1 + dir];
                    if (pas) {
                        for (int t = 0; t < np; ++t) {
                            size_t o = IX(5 + t, i, j, k);
                            double a = passive_face_flux(P, pr, 0, po + t, dir, ip, jp, kp, vLb, vRb, sL2, sR2);
                            rhs[o] -= (a - trL[t]) / dd[dir];
                            double b = passive_face_flux(P, pr, 1, po + t, dir, ip, jp, kp, vLb, vRb, sL2, sR2);
                            rhs[o] -= (b - trL[t]) / dd[dir];
                            if (dir == 1) {
                                double c = passive_face_flux(P, pr, 2, po + t, dir, ip, jp, kp, vLb, vRb, sL2, sR2);
                                rhs[o] -= (c - trL[t]) / dd[dir];
                            }
                        }
                    }
                    for (int m = 0; m < 5; ++m) rhs[IX(m, i, j, k)] -= (FB[m] - FA[m]) / dd[dir];
                }
            }
}

/* core/srhd_core.py:1267-1281 (sound speed with the definition-time Gamma = 5/3, Q5) */
static double srhd_max_char_speed(const orc_params *P, const double *pr) {
    double amax = 0.0;
#pragma omp parallel for reduction(max : amax) schedule(static)
    for (int i = 1; i < NXF - 1; ++i)
        for (int j = 1; j < NYF - 1; ++j)
            for (int k = 1; k < NZF - 1; ++k) {
                double rho = PR(0, i, j, k), vx = PR(1, i, j, k), vy = PR(2, i, j, k), vz = PR(3, i, j, k),
                       p = PR(4, i, j, k);
                double cs = srhd_sound_speed(P->gamma, rho, p);
                double a[6];
                eig_speeds(vx, cs, &a[0], &a[1]);
                eig_speeds(vy, cs, &a[2], &a[3]);
                eig_speeds(vz, cs, &a[4], &a[5]);
                double loc = fabs(a[0]);
                for (int m = 1; m < 6; ++m) loc = pymax(loc, fabs(a[m]));
                if (loc > amax) amax = loc;
            }
    if (amax < SMALL) amax = SMALL;
    return amax;
}

/* =====================================================================
 * RMHD (core/rmhd_core.py)
 * ===================================================================== */

/* core/rmhd_core.py:125-135 */
static void rmhd_floor(double p_max, double v_max, double *q) {
    if (q[0] < SMALL) q[0] = SMALL;
    if (q[4] < SMALL) q[4] = SMALL;
    if (q[4] > p_max) q[4] = p_max;
    double v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double vmax2 = v_max * v_max;
    if (v2 >= vmax2) {
        double fac = v_max / sqrt(v2 + 1e-32);
        q[1] *= fac; q[2] *= fac; q[3] *= fac;
    }
}

/* core/rmhd_core.py:137-168 */
static void rmhd_p2c(double g, double p_max, double v_max, const double *q, double *U) {
    double rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4], Bx = q[5], By = q[6], Bz = q[7], psi = q[8];
    double v2 = vx * vx + vy * vy + vz * vz;
    double vmax2 = v_max * v_max;
    if (v2 >= vmax2) {
        double fac = v_max / sqrt(v2 + 1e-32);
        vx *= fac; vy *= fac; vz *= fac;
    }
    if (p < SMALL) p = SMALL;
    if (p > p_max) p = p_max;
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double W = 1.0 / sqrt(1.0 - (vx * vx + vy * vy + vz * vz));
    double h = 1.0 + g / (g - 1.0) * p / pymax(rho, SMALL);
    double rhoh = rho * h;
    double vb = vx * Bx + vy * By + vz * Bz;
    double b0 = W * vb;
    double bx = (Bx + b0 * vx) / W;
    double by = (By + b0 * vy) / W;
    double bz = (Bz + b0 * vz) / W;
    double b2 = (bx * bx + by * by + bz * bz) - b0 * b0;
    double wtot = rhoh * W * W + B2;
    double D = rho * W;
    U[0] = D;
    U[1] = wtot * vx - (Bx * vb);
    U[2] = wtot * vy - (By * vb);
    U[3] = wtot * vz - (Bz * vb);
    U[4] = rhoh * W * W - p + 0.5 * B2 + 0.5 * b2 - D;
    U[5] = Bx; U[6] = By; U[7] = Bz; U[8] = psi;
}

/* core/rmhd_core.py:343-395 (flux_rmhd_x/y/z): U is the p2c of q (shared with the caller) */
static void rmhd_flux_from(double glm_ch, int dir, const double *q, const double *U, double *F) {
    double vx = q[1], vy = q[2], vz = q[3], p = q[4], Bx = q[5], By = q[6], Bz = q[7], psi = q[8];
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double pt = p + 0.5 * B2;
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4];
    double vB = Bx * vx + By * vy + Bz * vz;
    if (dir == 0) {
        F[0] = D * vx;
        F[1] = Sx * vx + pt - Bx * Bx;
        F[2] = Sy * vx - Bx * By;
        F[3] = Sz * vx - Bx * Bz;
        F[4] = (tau + pt) * vx - (Bx * vB);
        F[5] = psi;
        F[6] = vy * Bx - vx * By;
        F[7] = vz * Bx - vx * Bz;
        F[8] = glm_ch * glm_ch * Bx;
    } else if (dir == 1) {
        F[0] = D * vy;
        F[1] = Sx * vy - By * Bx;
        F[2] = Sy * vy + pt - By * By;
        F[3] = Sz * vy - By * Bz;
        F[4] = (tau + pt) * vy - (By * vB);
        F[5] = vx * By - vy * Bx;
        F[6] = psi;
        F[7] = vz * By - vy * Bz;
        F[8] = glm_ch * glm_ch * By;
    } else {
        F[0] = D * vz;
        F[1] = Sx * vz - Bz * Bx;
        F[2] = Sy * vz - Bz * By;
        F[3] = Sz * vz + pt - Bz * Bz;
        F[4] = (tau + pt) * vz - (Bz * vB);
        F[5] = vx * Bz - vz * Bx;
        F[6] = vy * Bz - vz * By;
        F[7] = psi;
        F[8] = glm_ch * glm_ch * Bz;
    }
}

/* core/rmhd_core.py:228-246 */
static double rmhd_f_of_Z(const orc_params *P, double Z, double D, double tau, double B2, double SB, double S2) {
    double g = P->gamma;
    double denomZ = pymax(Z, SMALL);
    double vb = SB / denomZ;
    double ZpB = Z + B2;
    double v2 = (S2 + (SB * SB) * (2.0 * Z + B2) / (denomZ * denomZ)) / (ZpB * ZpB + SMALL);
    double vmax2 = P->v_max * P->v_max;
    if (v2 >= vmax2) v2 = vmax2 - 1e-14;
    double W2 = 1.0 / (1.0 - v2);
    double W = sqrt(W2);
    double rho = D / pymax(W, SMALL);
    double h = Z / pymax(rho * W2, SMALL);
    double p = (h - 1.0) * rho * (g - 1.0) / g;
    if (p < SMALL) p = SMALL;
    if (p > P->p_max) p = P->p_max;
    double b2 = B2 / W2 + vb * vb;
    double tau_calc = Z - p + 0.5 * B2 + 0.5 * b2 - D;
    return tau_calc - tau;
}

/* core/rmhd_core.py:248-341.  Returns flags: 2 = Newton failed from Z0 and Z1 (bisection),
 * 4 = invalid recovery (the hydro guess is returned). */
static int rmhd_c2p(const orc_params *P, const double *U, double *q) {
    int flags = 0;
    const double g = P->gamma;
    double D = U[0], Sx = U[1], Sy = U[2], Sz = U[3], tau = U[4], Bx = U[5], By = U[6], Bz = U[7], psi = U[8];
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double SB = Sx * Bx + Sy * By + Sz * Bz;
    double S2 = Sx * Sx + Sy * Sy + Sz * Sz;
    double tau_h = tau - 0.5 * B2;
    if (tau_h < SMALL) tau_h = SMALL;
    double Uh[5] = {D, Sx, Sy, Sz, tau_h}, q0[5];
    hydro_c2p(g, P->p_max, P->v_max, Uh, q0);
    double v20 = q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3];
    if (v20 >= P->v_max * P->v_max) v20 = P->v_max * P->v_max - 1e-14;
    double W0 = 1.0 / sqrt(1.0 - v20);
    double h0 = 1.0 + g / (g - 1.0) * q0[4] / pymax(q0[0], SMALL);
    double Z0 = pymax(q0[0] * h0 * W0 * W0, SMALL);
    double Z1 = pymax(D + tau + B2, SMALL);

    int ok = 0;
    double Z = Z0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (attempt == 1) Z = Z1;
        for (int it = 0; it < 50; ++it) {
            double f = rmhd_f_of_Z(P, Z, D, tau, B2, SB, S2);
            if (fabs(f) < 1e-10 * pymax(1.0, tau)) {
                ok = 1;
                break;
            }
            double eps = 1e-6 * pymax(1.0, Z);
            double fp = rmhd_f_of_Z(P, Z + eps, D, tau, B2, SB, S2);
            double dfdZ = (fp - f) / eps;
            if (dfdZ == 0.0) break;
            double dZ = -f / dfdZ;
            if (dZ > 0.5 * Z) dZ = 0.5 * Z;
            if (dZ < -0.5 * Z) dZ = -0.5 * Z;
            Z += dZ;
            if (Z < SMALL) Z = SMALL;
        }
        if (ok) break;
    }
    if (!ok) {
        flags |= 2;
        double Zmin = SMALL;
        double Zmax = pymax(pymax(Z, Z1), 1.0);
        double fmin = rmhd_f_of_Z(P, Zmin, D, tau, B2, SB, S2);
        double fmax = rmhd_f_of_Z(P, Zmax, D, tau, B2, SB, S2);
        for (int it = 0; it < 40; ++it) {
            if (fmin * fmax <= 0.0) break;
            Zmax *= 2.0;
            fmax = rmhd_f_of_Z(P, Zmax, D, tau, B2, SB, S2);
        }
        if (fmin * fmax <= 0.0) {
            for (int it = 0; it < 60; ++it) {
                double Zmid = 0.5 * (Zmin + Zmax);
                double fmid = rmhd_f_of_Z(P, Zmid, D, tau, B2, SB, S2);
                if (fabs(fmid) < 1e-10 * pymax(1.0, tau)) {
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
    double ZpB = Z + B2;
    double vb = SB / pymax(Z, SMALL);
    double vx = (Sx + vb * Bx) / pymax(ZpB, SMALL);
    double vy = (Sy + vb * By) / pymax(ZpB, SMALL);
    double vz = (Sz + vb * Bz) / pymax(ZpB, SMALL);
    double v2 = vx * vx + vy * vy + vz * vz;
    double vmax2 = P->v_max * P->v_max;
    if (v2 >= vmax2) {
        double fac = P->v_max / sqrt(v2 + 1e-32);
        vx *= fac; vy *= fac; vz *= fac;
    }
    double W = 1.0 / sqrt(1.0 - (vx * vx + vy * vy + vz * vz));
    double rho = D / pymax(W, SMALL);
    double h = Z / pymax(rho * W * W, SMALL);
    double p = (h - 1.0) * rho * (g - 1.0) / g;
    if (!isfinite(rho) || !isfinite(p) || rho <= 0.0 || p <= 0.0) {
        flags |= 4;
        rho = q0[0]; vx = q0[1]; vy = q0[2]; vz = q0[3]; p = q0[4];
    }
    q[0] = rho; q[1] = vx; q[2] = vy; q[3] = vz; q[4] = p;
    q[5] = Bx; q[6] = By; q[7] = Bz; q[8] = psi;
    rmhd_floor(P->p_max, P->v_max, q);
    return flags;
}

/* core/rmhd_core.py:419-457 */
static double rmhd_fast_speed(const orc_params *P, double rho, double p, double vx, double vy, double vz,
                              double Bx, double By, double Bz) {
    const double g = P->gamma;
    double v2 = vx * vx + vy * vy + vz * vz;
    if (v2 >= P->v_max * P->v_max) v2 = P->v_max * P->v_max - 1e-14;
    double W = 1.0 / sqrt(1.0 - v2);
    if (rho < SMALL) rho = SMALL;
    if (p < SMALL) p = SMALL;
    if (p > P->p_max) p = P->p_max;
    double h = 1.0 + g / (g - 1.0) * p / pymax(rho, SMALL);
    double rhoh = rho * h;
    double vb = vx * Bx + vy * By + vz * Bz;
    double b0 = W * vb;
    double bx = (Bx + b0 * vx) / W;
    double by = (By + b0 * vy) / W;
    double bz = (Bz + b0 * vz) / W;
    double b2 = (bx * bx + by * by + bz * bz) - b0 * b0;
    double cs2 = g * p / pymax(rhoh, SMALL);
    if (cs2 < 0.0) cs2 = 0.0;
    if (cs2 > 1.0 - 1e-14) cs2 = 1.0 - 1e-14;
    double ca2 = b2 / pymax(rhoh + b2, SMALL);
    if (ca2 < 0.0) ca2 = 0.0;
    if (ca2 > 1.0 - 1e-14) ca2 = 1.0 - 1e-14;
    double a = cs2 + ca2 - cs2 * ca2;
    double disc = a * a - 4.0 * cs2 * ca2;
    if (disc < 0.0) disc = 0.0;
    double cf2 = 0.5 * (a + sqrt(disc));
    if (cf2 < 0.0) cf2 = 0.0;
    if (cf2 > 1.0 - 1e-14) cf2 = 1.0 - 1e-14;
    return sqrt(cf2);
}

/* core/rmhd_core.py:580-599 */
static double hllc_contact_speed(int nm, const double *UL, const double *UR, double sL, double sR, double pL,
                                 double pR, double vL, double vR) {
    double DL = UL[0], DR = UR[0];
    double denom = (DL * (sL - vL) - DR * (sR - vR)) + SMALL;
    return (pR - pL + UL[nm] * (sL - vL) - UR[nm] * (sR - vR)) / denom;
}

/* core/rmhd_core.py:601-724 (hll_d_flux_dir).  The star-state fluxes FLs/FRs/Fs of the reference
 * (:703-714) are computed but never used, so they are omitted here (no side effects). */
static void hlld_full(const orc_params *P, int n, const double *UL, const double *UR, const double *FL,
                      const double *FR, double sL, double sR, double pL, double pR, double vL, double vR,
                      double *out) {
    int m;
    if (sL >= 0.0) { for (m = 0; m < 9; ++m) out[m] = FL[m]; return; }
    if (sR <= 0.0) { for (m = 0; m < 9; ++m) out[m] = FR[m]; return; }
    int vt1, vt2, bn, bt1, bt2;
    if (n == 0) { vt1 = 2; vt2 = 3; bn = 5; bt1 = 6; bt2 = 7; }
    else if (n == 1) { vt1 = 1; vt2 = 3; bn = 6; bt1 = 5; bt2 = 7; }
    else { vt1 = 1; vt2 = 2; bn = 7; bt1 = 5; bt2 = 6; }
    double vL_n = vL, vR_n = vR;
    if (UL[0] <= SMALL || UR[0] <= SMALL) { hlle_vec(9, UL, UR, FL, FR, sL, sR, out); return; }
    double BnL = UL[bn], BnR = UR[bn];
    double Bt1L = UL[bt1], Bt2L = UL[bt2], Bt1R = UR[bt1], Bt2R = UR[bt2];
    double Bn = 0.5 * (BnL + BnR);
    if (fabs(Bn) < 1e-8) { hlle_vec(9, UL, UR, FL, FR, sL, sR, out); return; }
    double sM = hllc_contact_speed(1 + n, UL, UR, sL, sR, pL, pR, vL_n, vR_n);
    double DL = UL[0], DR = UR[0];
    double DstarL = DL * (sL - vL_n) / (sL - sM + SMALL);
    double DstarR = DR * (sR - vR_n) / (sR - sM + SMALL);
    if (DstarL <= SMALL || DstarR <= SMALL) { hlle_vec(9, UL, UR, FL, FR, sL, sR, out); return; }
    double Bt1L_s = Bt1L * (sL - vL_n) / (sL - sM + SMALL);
    double Bt2L_s = Bt2L * (sL - vL_n) / (sL - sM + SMALL);
    double Bt1R_s = Bt1R * (sR - vR_n) / (sR - sM + SMALL);
    double Bt2R_s = Bt2R * (sR - vR_n) / (sR - sM + SMALL);
    double denomL = DL * (sL - vL_n) - Bn * Bn;
    double denomR = DR * (sR - vR_n) - Bn * Bn;
    if (fabs(denomL) < SMALL || fabs(denomR) < SMALL) { hlle_vec(9, UL, UR, FL, FR, sL, sR, out); return; }
    double vt1L = UL[vt1] / pymax(DL, SMALL);
    double vt2L = UL[vt2] / pymax(DL, SMALL);
    double vt1R = UR[vt1] / pymax(DR, SMALL);
    double vt2R = UR[vt2] / pymax(DR, SMALL);
    double vt1L_s = vt1L - Bn * (Bt1L - Bt1L_s) / denomL;
    double vt2L_s = vt2L - Bn * (Bt2L - Bt2L_s) / denomL;
    double vt1R_s = vt1R - Bn * (Bt1R - Bt1R_s) / denomR;
    double vt2R_s = vt2R - Bn * (Bt2R - Bt2R_s) / denomR;
    double sL_star = sM - fabs(Bn) / sqrt(DstarL + SMALL);
    double sR_star = sM + fabs(Bn) / sqrt(DstarR + SMALL);
    double sqrtDL = sqrt(DstarL), sqrtDR = sqrt(DstarR);
    double sgn = (Bn >= 0.0) ? 1.0 : -1.0;
    double vt1_ss = (sqrtDL * vt1L_s + sqrtDR * vt1R_s + sgn * (Bt1R_s - Bt1L_s)) / (sqrtDL + sqrtDR + SMALL);
    double vt2_ss = (sqrtDL * vt2L_s + sqrtDR * vt2R_s + sgn * (Bt2R_s - Bt2L_s)) / (sqrtDL + sqrtDR + SMALL);
    double Bt1_ss = (sqrtDL * Bt1R_s + sqrtDR * Bt1L_s + sgn * sqrtDL * sqrtDR * (vt1L_s - vt1R_s)) /
                    (sqrtDL + sqrtDR + SMALL);
    double Bt2_ss = (sqrtDL * Bt2R_s + sqrtDR * Bt2L_s + sgn * sqrtDL * sqrtDR * (vt2L_s - vt2R_s)) /
                    (sqrtDL + sqrtDR + SMALL);
    /* build_state (:684-693) */
    double st[3][9];
    const double Ds[3] = {DstarL, DstarR, 0.5 * (DstarL + DstarR)};
    const double vt1s[3] = {vt1L_s, vt1R_s, vt1_ss}, vt2s[3] = {vt2L_s, vt2R_s, vt2_ss};
    const double b1s[3] = {Bt1L_s, Bt1R_s, Bt1_ss}, b2s[3] = {Bt2L_s, Bt2R_s, Bt2_ss};
    const double psis[3] = {UL[8], UR[8], 0.5 * (UL[8] + UR[8])};
    for (int s = 0; s < 3; ++s) {
        double *q = st[s];
        q[0] = pymax(Ds[s], SMALL);
        q[1 + n] = sM;
        q[vt1] = vt1s[s];
        q[vt2] = vt2s[s];
        q[bn] = Bn;
        q[bt1] = b1s[s];
        q[bt2] = b2s[s];
        q[4] = (Ds[s] == DstarL) ? pL : pR;
        q[8] = psis[s];
    }
    double ULs[9], URs[9], Us[9];
    rmhd_p2c(P->gamma, P->p_max, P->v_max, st[0], ULs);
    rmhd_p2c(P->gamma, P->p_max, P->v_max, st[1], URs);
    rmhd_p2c(P->gamma, P->p_max, P->v_max, st[2], Us);
    if (sL <= 0.0 && 0.0 <= sL_star) {
        for (m = 0; m < 9; ++m) out[m] = FL[m] + sL * (ULs[m] - UL[m]);
    } else if (sL_star <= 0.0 && 0.0 <= sM) {
        for (m = 0; m < 9; ++m) out[m] = FL[m] + sL * (ULs[m] - UL[m]) + sL_star * (Us[m] - ULs[m]);
    } else if (sM <= 0.0 && 0.0 <= sR_star) {
        for (m = 0; m < 9; ++m) out[m] = FR[m] + sR * (URs[m] - UR[m]) + sR_star * (Us[m] - URs[m]);
    } else if (sR_star <= 0.0 && 0.0 <= sR) {
        for (m = 0; m < 9; ++m) out[m] = FR[m] + sR * (URs[m] - UR[m]);
    } else {
        hlle_vec(9, UL, UR, FL, FR, sL, sR, out);
    }
}

/* RMHD face: floors, p2c, fluxes, fast-speed bounds (unless speeds are given), the Riemann
 * solver (core/rmhd_core.py:569-578, 726-746) and, in MUSCL, the hlld override of [5:9] by the
 * physical flux of the upwind side of the contact (:811-831).  use_speeds: take sLo/sRo as input.
 * qL/qR are floored in place. */
static void rmhd_face(const orc_params *P, int dir, double *qL, double *qR, int use_speeds, double *F,
                      double *sLo, double *sRo) {
    rmhd_floor(P->p_max, P->v_max, qL);
    rmhd_floor(P->p_max, P->v_max, qR);
    double UL[9], UR[9], FL[9], FR[9];
    rmhd_p2c(P->gamma, P->p_max, P->v_max, qL, UL);
    rmhd_p2c(P->gamma, P->p_max, P->v_max, qR, UR);
    rmhd_flux_from(P->glm_ch, dir, qL, UL, FL);
    rmhd_flux_from(P->glm_ch, dir, qR, UR, FR);
    double sL, sR;
    if (use_speeds) {
        sL = *sLo;
        sR = *sRo;
    } else {
        double cfL = rmhd_fast_speed(P, qL[0], qL[4], qL[1], qL[2], qL[3], qL[5], qL[6], qL[7]);
        double cfR = rmhd_fast_speed(P, qR[0], qR[4], qR[1], qR[2], qR[3], qR[5], qR[6], qR[7]);
        double lmL, lpL, lmR, lpR;
        eig_speeds(qL[1 + dir], cfL, &lmL, &lpL);
        eig_speeds(qR[1 + dir], cfR, &lmR, &lpR);
        sL = pymin(lmL, lmR);
        sR = pymax(lpL, lpR);
    }
    double pL = qL[4], pR = qR[4], vL = qL[1 + dir], vR = qR[1 + dir];
    if (P->riemann_id == 2) {
        hlld_full(P, dir, UL, UR, FL, FR, sL, sR, pL, pR, vL, vR, F);
    } else if (P->riemann_id == 1) {
        hlle_vec(9, UL, UR, FL, FR, sL, sR, F);
        double h5[5];
        hllc5(1 + dir, UL, UR, FL, FR, sL, sR, pL, pR, vL, vR, h5);
        for (int m = 0; m < 5; ++m) F[m] = h5[m];
        if (P->recon_id == 0) {
            double sM = hllc_contact_speed(1 + dir, UL, UR, sL, sR, pL, pR, vL, vR);
            const double *Fm = (sM >= 0.0) ? FL : FR;
            for (int m = 5; m < 9; ++m) F[m] = Fm[m];
        }
    } else {
        hlle_vec(9, UL, UR, FL, FR, sL, sR, F);
    }
    *sLo = sL;
    *sRo = sR;
}

void orc_rmhd_p2c_batch(const orc_params *P, int64_t n, const double *prim, double *cons) {
    for (int64_t m = 0; m < n; ++m) rmhd_p2c(P->gamma, P->p_max, P->v_max, prim + 9 * m, cons + 9 * m);
}
void orc_rmhd_c2p_batch(const orc_params *P, int64_t n, const double *cons, double *prim) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t m = 0; m < n; ++m) rmhd_c2p(P, cons + 9 * m, prim + 9 * m);
}

void orc_rmhd_c2p_flags_batch(const orc_params *P, int64_t n, const double *cons, double *prim, int32_t *flags) {
    for (int64_t m = 0; m < n; ++m) flags[m] = rmhd_c2p(P, cons + 9 * m, prim + 9 * m);
}
void orc_rmhd_hydro_guess_batch(const orc_params *P, int64_t n, const double *cons5, double *prim5) {
    for (int64_t m = 0; m < n; ++m) hydro_c2p(P->gamma, P->p_max, P->v_max, cons5 + 5 * m, prim5 + 5 * m);
}
void orc_rmhd_f_of_Z_batch(const orc_params *P, int64_t n, const double *Z, const double *cons, double *f) {
    for (int64_t m = 0; m < n; ++m) {
        const double *U = cons + 9 * m;
        double B2 = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
        double SB = U[1] * U[5] + U[2] * U[6] + U[3] * U[7];
        double S2 = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
        f[m] = rmhd_f_of_Z(P, Z[m], U[0], U[4], B2, SB, S2);
    }
}
void orc_rmhd_fast_speed_batch(const orc_params *P, int64_t n, const double *prim, double *cf) {
    for (int64_t m = 0; m < n; ++m) {
        const double *q = prim + 9 * m;
        cf[m] = rmhd_fast_speed(P, q[0], q[4], q[1], q[2], q[3], q[5], q[6], q[7]);
    }
}
void orc_rmhd_riemann(const orc_params *P, int dir, const double *primL, const double *primR, int unit_speeds,
                      double *F) {
    double qL[9], qR[9], sL = -1.0, sR = 1.0;
    memcpy(qL, primL, sizeof qL);
    memcpy(qR, primR, sizeof qR);
    rmhd_face(P, dir, qL, qR, unit_speeds, F, &sL, &sR);
}

/* core/rmhd_core.py:1264 (GLM source) and :1265-1275 (resistive Laplacian of B) */
static void rmhd_sources(const orc_params *P, const double *pr, double *rhs, int i, int j, int k, int n_resist,
                         double dx, double dy, double dz) {
    rhs[IX(8, i, j, k)] -= P->glm_cp * P->glm_cp * PR(8, i, j, k);
    for (int rep = 0; rep < n_resist; ++rep) {
        if (P->resistive && P->resistivity > 0.0) {
            double invdx2 = 1.0 / (dx * dx + SMALL), invdy2 = 1.0 / (dy * dy + SMALL), invdz2 = 1.0 / (dz * dz + SMALL);
            for (int c = 5; c < 8; ++c) {
                double lap = (PR(c, i + 1, j, k) - 2.0 * PR(c, i, j, k) + PR(c, i - 1, j, k)) * invdx2 +
                             (PR(c, i, j + 1, k) - 2.0 * PR(c, i, j, k) + PR(c, i, j - 1, k)) * invdy2 +
                             (PR(c, i, j, k + 1) - 2.0 * PR(c, i, j, k) + PR(c, i, j, k - 1)) * invdz2;
                rhs[IX(c, i, j, k)] += P->resistivity * lap;
            }
        }
    }
}

/* core/rmhd_core.py:748-1277 (compute_rhs_rmhd, MUSCL): all faces use their own speeds; the
 * passive right faces reuse the left face's (sL, sR) (:915, :1075, :1235). */
static void rmhd_rhs_muscl(const orc_params *P, const double *pr, double *rhs, double dx, double dy, double dz) {
    const int np = P->n_passive, po = P->passive_offset;
    const int i0 = 2, i1 = NXF - 2, j0 = 2, j1 = NYF - 2, k0 = 2, k1 = NZF - 2;
    const double dd[3] = {dx, dy, dz};
#pragma omp parallel for schedule(static)
    for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j)
            for (int k = k0; k < k1; ++k) {
                for (int dir = 0; dir < 3; ++dir) {
                    int ip = i + (dir == 0), jp = j + (dir == 1), kp = k + (dir == 2);
                    double qL[9], qR[9], FA[9], FB[9], sL, sR, sL2, sR2, trL[8];
                    face_states(P, pr, 9, dir, i, j, k, qL, qR);
                    rmhd_face(P, dir, qL, qR, 0, FA, &sL, &sR);
                    double vLa = qL[1 + dir], vRa = qR[1 + dir];
                    for (int t = 0; t < np; ++t) {
                        trL[t] = passive_face_flux(P, pr, 0, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                        trL[t] = passive_face_flux(P, pr, 1, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                        trL[t] = passive_face_flux(P, pr, 2, po + t, dir, i, j, k, vLa, vRa, sL, sR);
                    }
                    face_states(P, pr, 9, dir, ip, jp, kp, qL, qR);
                    rmhd_face(P, dir, qL, qR, 0, FB, &sL2, &sR2);
                    double vLb = qL[1 + dir], vRb = qR[1 + dir];
                    for (int t = 0; t < np; ++t) {
                        size_t o = IX(9 + t, i, j, k);
                        double a = passive_face_flux(P, pr, 0, po + t, dir, ip, jp, kp, vLb, vRb, sL, sR);
                        rhs[o] -= (a - trL[t]) / dd[dir];
                        double b = passive_face_flux(P, pr, 1, po + t, dir, ip, jp, kp, vLb, vRb, sL, sR);
                        rhs[o] -= (b - trL[t]) / dd[dir];
                        double c = passive_face_flux(P, pr, 2, po + t, dir, ip, jp, kp, vLb, vRb, sL, sR);
                        rhs[o] -= (c - trL[t]) / dd[dir];
                    }
                    for (int m = 0; m < 9; ++m) rhs[IX(m, i, j, k)] -= (FB[m] - FA[m]) / dd[dir];
                }
                rmhd_sources(P, pr, rhs, i, j, k, 1, dx, dy, dz);
            }
}

/* z-back face speeds of cell (i,j,k) in PPM/WENO: the value the reference carries in (sL, sR)
 * out of that cell (core/rmhd_core.py:1395, :1537). */
static void zback_speeds(const orc_params *P, const double *pr, int i, int j, int k, double *sL, double *sR) {
    double qL[9], qR[9], F[9];
    face_states(P, pr, 9, 2, i, j, k, qL, qR);
    rmhd_face(P, 2, qL, qR, 0, F, sL, sR);
}

/* core/rmhd_core.py:1279-1437 (compute_rhs_ppm) and :1439-1557 (compute_rhs_weno):
 * the x-left face uses the speeds carried from the previous cell's z-back face (Q3), seeded
 * (-1, 1) by hlle_bounds_fast (:397-400); resistivity twice in PPM, never in WENO (Q9). */
static void rmhd_rhs_ppm_weno(const orc_params *P, const double *pr, double *rhs, double dx, double dy,
                              double dz) {
    const int i0 = 3, i1 = NXF - 3, j0 = 3, j1 = NYF - 3, k0 = 3, k1 = NZF - 3;
    const double dd[3] = {dx, dy, dz};
    const int n_resist = P->resist_once ? 1 : ((P->recon_id == 1) ? 2 : 0);
#pragma omp parallel for schedule(static)
    for (int i = i0; i < i1; ++i) {
        double sL = -1.0, sR = 1.0;
        if (i > i0) zback_speeds(P, pr, i - 1, j1 - 1, k1 - 1, &sL, &sR);
        for (int j = j0; j < j1; ++j)
            for (int k = k0; k < k1; ++k) {
                double qL[9], qR[9], FA[9], FB[9], s2L, s2R;
                /* x: left face with the carried speeds, right face with its own */
                face_states(P, pr, 9, 0, i, j, k, qL, qR);
                rmhd_face(P, 0, qL, qR, 1, FA, &sL, &sR);
                face_states(P, pr, 9, 0, i + 1, j, k, qL, qR);
                rmhd_face(P, 0, qL, qR, 0, FB, &s2L, &s2R);
                for (int m = 0; m < 9; ++m) rhs[IX(m, i, j, k)] -= (FB[m] - FA[m]) / dd[0];
                /* y: both faces with their own speeds; (sL, sR) overwritten */
                face_states(P, pr, 9, 1, i, j, k, qL, qR);
                rmhd_face(P, 1, qL, qR, 0, FA, &sL, &sR);
                face_states(P, pr, 9, 1, i, j + 1, k, qL, qR);
                rmhd_face(P, 1, qL, qR, 0, FB, &s2L, &s2R);
                for (int m = 0; m < 9; ++m) rhs[IX(m, i, j, k)] -= (FB[m] - FA[m]) / dd[1];
                /* z: (sL, sR) of the back face are carried into the next cell */
                face_states(P, pr, 9, 2, i, j, k, qL, qR);
                rmhd_face(P, 2, qL, qR, 0, FA, &sL, &sR);
                face_states(P, pr, 9, 2, i, j, k + 1, qL, qR);
                rmhd_face(P, 2, qL, qR, 0, FB, &s2L, &s2R);
                for (int m = 0; m < 9; ++m) rhs[IX(m, i, j, k)] -= (FB[m] - FA[m]) / dd[2];
                rmhd_sources(P, pr, rhs, i, j, k, n_resist, dx, dy, dz);
            }
    }
}

/* =====================================================================
 * Whole-array operators
 * ===================================================================== */
/* core/gr_metric.py:6-16 (schwarzschild_iso_lapse_and_grad): d ln(alpha) / dx^i */
static void gr_iso_grad(double x, double y, double z, double M, double *d) {
    double r = sqrt(x * x + y * y + z * z) + 1e-12;
    double a = M / (2.0 * r);
    double denom = (1.0 - a * a);
    double dlnadr = denom == 0.0 ? 0.0 : (2.0 * a) / (r * denom);
    d[0] = dlnadr * (x / r);
    d[1] = dlnadr * (y / r);
    d[2] = dlnadr * (z / r);
}

/* core/gr_metric.py:18-38 (schwarzschild_ks_lapse_shift_and_grad; the Kerr-Schild helper ignores
 * the spin): returns the lapse, and d ln(alpha) / dx^i into d when d != NULL */
static double gr_ks(double x, double y, double z, double M, double *d) {
    double r = sqrt(x * x + y * y + z * z) + 1e-12;
    double H = M / r;
    double alpha = 1.0 / sqrt(1.0 + 2.0 * H);
    if (d) {
        double lx = x / r, ly = y / r, lz = z / r;
        double dlnadr = H / (r * (1.0 + 2.0 * H));
        d[0] = dlnadr * lx;
        d[1] = dlnadr * ly;
        d[2] = dlnadr * lz;
    }
    return alpha;
}

static inline double gr_x(const orc_params *P, int i, double dx) { return ((double)(P->gr_x0 + (i - P->ng)) + 0.5) * dx; }

/* core/grhd_core.py:50-73, core/grmhd_core.py:50-75: the fixed-metric source terms from the
 * original primitives (not the rescaled ones), interior cells */
static void gr_sources(const orc_params *P, const double *pr, double *rhs, double dx, double dy, double dz) {
    const int ng = P->ng, rm = P->physics == 1;
#pragma omp parallel for schedule(static)
    for (int i = ng; i < NXF - ng; ++i)
        for (int j = ng; j < NYF - ng; ++j)
            for (int k = ng; k < NZF - ng; ++k) {
                double x = gr_x(P, i, dx), y = ((double)(j - ng) + 0.5) * dy, z = ((double)(k - ng) + 0.5) * dz;
                double d[3];
                if (P->gr_metric == 1) gr_iso_grad(x, y, z, P->gr_mass, d);
                else (void)gr_ks(x, y, z, P->gr_mass, d);
                double q[9], U[9];
                for (int v = 0; v < (rm ? 9 : 5); ++v) q[v] = PR(v, i, j, k);
                double rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4];
                double v2 = vx * vx + vy * vy + vz * vz;
                if (v2 >= 1.0) v2 = 1.0 - 1e-14;
                double W = 1.0 / sqrt(1.0 - v2);
                double G = P->gr_gamma;
                double h = 1.0 + G / (G - 1.0) * p / pymax(rho, SMALL);
                double w;
                if (rm) {
                    double B2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
                    w = rho * h * W * W + B2 + p;
                    rmhd_p2c(P->gamma, P->p_max, P->v_max, q, U);
                } else {
                    w = rho * h * W * W + p;
                    srhd_p2c(P->gamma, q, U);
                }
                const size_t N = (size_t)NXF * NYF * NZF, c = IX(0, i, j, k);
                rhs[1 * N + c] += w * d[0];
                rhs[2 * N + c] += w * d[1];
                rhs[3 * N + c] += w * d[2];
                rhs[4 * N + c] += U[1] * d[0] + U[2] * d[1] + U[3] * d[2];
            }
}

/* The RHS of the step: core/srhd_core.py and core/rmhd_core.py compute_rhs_*, wrapped for GR as
 * core/grhd_core.py:20-75 / core/grmhd_core.py:20-77: with ORTHONORMAL_FLUX and a Schwarzschild or
 * Kerr-Schild metric the fluxes see the interior velocities divided by the Kerr-Schild lapse
 * (max(alpha, 1e-12)); any metric but Minkowski adds the lapse-gradient sources. */
void orc_rhs(const orc_params *P, const double *pr, double *rhs, double dx, double dy, double dz) {
    int nrhs = (P->physics == 1 ? 9 : 5) + P->n_passive;
    const size_t N = (size_t)NXF * NYF * NZF;
    memset(rhs, 0, sizeof(double) * (size_t)nrhs * N);
    const double *pu = pr;
    double *tmp = NULL;
    if (P->gr_orthonormal && (P->gr_metric == 1 || P->gr_metric == 2)) {
        tmp = (double *)malloc(sizeof(double) * (size_t)P->nv * N);
        memcpy(tmp, pr, sizeof(double) * (size_t)P->nv * N);
        const int ng = P->ng;
        for (int i = ng; i < NXF - ng; ++i)
            for (int j = ng; j < NYF - ng; ++j)
                for (int k = ng; k < NZF - ng; ++k) {
                    double x = gr_x(P, i, dx), y = ((double)(j - ng) + 0.5) * dy, z = ((double)(k - ng) + 0.5) * dz;
                    double ax = pymax(gr_ks(x, y, z, P->gr_mass, NULL), 1e-12);
                    for (int v = 1; v <= 3; ++v) tmp[(size_t)v * N + IX(0, i, j, k)] = PR(v, i, j, k) / ax;
                }
        pu = tmp;
    }
    if (P->physics == 1) {
        if (P->recon_id == 0)
            rmhd_rhs_muscl(P, pu, rhs, dx, dy, dz);
        else
            rmhd_rhs_ppm_weno(P, pu, rhs, dx, dy, dz);
    } else {
        if (P->recon_id == 0)
            srhd_rhs_muscl(P, pu, rhs, dx, dy, dz);
        else
            srhd_rhs_ppm_weno(P, pu, rhs, dx, dy, dz);
    }
    if (P->gr_metric != 0) gr_sources(P, pr, rhs, dx, dy, dz);
    free(tmp);
}

double orc_max_char_speed(const orc_params *P, const double *pr) { return srhd_max_char_speed(P, pr); }

/* core/srhd_core.py:1286-1400 (step_ssprk2/3), core/rmhd_core.py:1559-1677: p2c and c2p run over
 * every cell incl. ghosts (Q6); conserved states carry across stages; passives are updated as
 * primitives (Q19).  SRHD copies fields [5, PASSIVE_OFFSET) through (none when PASSIVE_OFFSET = 5). */
void orc_step(const orc_params *P, const double *pr, double *out, double dx, double dy, double dz, double dt) {
    const int nb = (P->physics == 1) ? 9 : 5;
    const int np = P->n_passive, po = P->passive_offset;
    const size_t N = (size_t)NXF * NYF * NZF;
    const int nrhs = nb + np;
    double *U0 = (double *)malloc(sizeof(double) * nb * N);
    double *Uc = (double *)malloc(sizeof(double) * nb * N);
    double *rhs = (double *)malloc(sizeof(double) * nrhs * N);
    double *p1 = (double *)calloc((size_t)P->nv * N, sizeof(double));
    double *p2 = (double *)calloc((size_t)P->nv * N, sizeof(double));
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < N; ++c) {
        double q[9], U[9];
        for (int v = 0; v < nb; ++v) q[v] = pr[v * N + c];
        if (nb == 9) rmhd_p2c(P->gamma, P->p_max, P->v_max, q, U);
        else srhd_p2c(P->gamma, q, U);
        for (int v = 0; v < nb; ++v) U0[v * N + c] = U[v];
    }
    const double *src = pr;
    const int nst = (P->rk_order == 3) ? 3 : 2;
    for (int s = 0; s < nst; ++s) {
        double *dst = (s == nst - 1) ? out : ((s == 0) ? p1 : p2);
        orc_rhs(P, src, rhs, dx, dy, dz);
#pragma omp parallel for schedule(dynamic, 512)
        for (size_t c = 0; c < N; ++c) {
            double U[9], q[9];
            for (int v = 0; v < nb; ++v) {
                double u0 = U0[v * N + c], r = dt * rhs[v * N + c], uc;
                if (s == 0) uc = u0 + r;
                else if (nst == 2) uc = 0.5 * (u0 + Uc[v * N + c] + r);
                else if (s == 1) uc = 0.75 * u0 + 0.25 * (Uc[v * N + c] + r);
                else uc = (1.0 / 3.0) * u0 + (2.0 / 3.0) * (Uc[v * N + c] + r);
                Uc[v * N + c] = uc;
                U[v] = uc;
            }
            if (nb == 9) (void)rmhd_c2p(P, U, q);
            else hydro_c2p(P->gamma, P->p_max, P->v_max, U, q);
            for (int v = 0; v < nb; ++v) dst[v * N + c] = q[v];
            for (int v = nb; v < P->nv; ++v) dst[v * N + c] = 0.0;
            if (nb == 5 && P->nv > 5 && po > 5)
                for (int v = 5; v < po; ++v) dst[v * N + c] = src[v * N + c];
            for (int t = 0; t < np; ++t) {
                size_t o = (size_t)(po + t) * N + c;
                double t0 = pr[o], r = dt * rhs[(size_t)(nb + t) * N + c];
                if (s == 0) dst[o] = t0 + r;
                else if (nst == 2) dst[o] = 0.5 * (t0 + src[o] + r);
                else if (s == 1) dst[o] = 0.75 * t0 + 0.25 * (src[o] + r);
                else dst[o] = (1.0 / 3.0) * t0 + (2.0 / 3.0) * (src[o] + r);
            }
        }
        src = dst;
    }
    free(U0); free(Uc); free(rhs); free(p1); free(p2);
}

/* core/boundary.py:4-8 */
void orc_apply_periodic_yz(const orc_params *P, double *pr) {
    const int ng = P->ng;
    for (int v = 0; v < P->nv; ++v)
        for (int i = 0; i < NXF; ++i) {
            for (int j = 0; j < ng; ++j)
                for (int k = 0; k < NZF; ++k) PR(v, i, j, k) = PR(v, i, NYF - 2 * ng + j, k);
            for (int j = 0; j < ng; ++j)
                for (int k = 0; k < NZF; ++k) PR(v, i, NYF - ng + j, k) = PR(v, i, ng + j, k);
        }
    for (int v = 0; v < P->nv; ++v)
        for (int i = 0; i < NXF; ++i)
            for (int j = 0; j < NYF; ++j) {
                for (int k = 0; k < ng; ++k) PR(v, i, j, k) = PR(v, i, j, NZF - 2 * ng + k);
                for (int k = 0; k < ng; ++k) PR(v, i, j, NZF - ng + k) = PR(v, i, j, ng + k);
            }
}

/* core/boundary.py:10-11 */
void orc_apply_outflow_right_x(const orc_params *P, double *pr) {
    const int ng = P->ng;
    for (int v = 0; v < P->nv; ++v)
        for (int g = 0; g < ng; ++g)
            for (int j = 0; j < NYF; ++j)
                for (int k = 0; k < NZF; ++k) PR(v, NXF - ng + g, j, k) = PR(v, NXF - ng - 1, j, k);
}

/* Additive option: zero-gradient y and/or z ghosts, in the same y-then-z order as the periodic fill. */
void orc_apply_outflow_yz(const orc_params *P, double *pr, int y, int z) {
    const int ng = P->ng;
    for (int v = 0; v < P->nv; ++v)
        for (int i = 0; i < NXF; ++i)
            if (y)
                for (int g = 0; g < ng; ++g)
                    for (int k = 0; k < NZF; ++k) {
                        PR(v, i, g, k) = PR(v, i, ng, k);
                        PR(v, i, NYF - ng + g, k) = PR(v, i, NYF - ng - 1, k);
                    }
    for (int v = 0; v < P->nv; ++v)
        for (int i = 0; i < NXF; ++i)
            for (int j = 0; j < NYF; ++j)
                if (z)
                    for (int g = 0; g < ng; ++g) {
                        PR(v, i, j, g) = PR(v, i, j, ng);
                        PR(v, i, j, NZF - ng + g) = PR(v, i, j, NZF - ng - 1);
                    }
}

/* core/nozzle.py:5-25 */
double orc_nozzle_weight(double r, double radius, double shear_thick, int profile) {
    if (radius <= 0.0) return 0.0;
    if (profile == 1) {
        if (r >= radius) return 0.0;
        return 0.5 * (1.0 + cos(M_PI * r / radius));
    }
    if (profile == 2) {
        if (r >= radius) return 0.0;
        double x = r / radius;
        return pymax(0.0, 1.0 - x * x);
    }
    if (shear_thick > 0.0) return 0.5 * (1.0 - tanh((r - radius) / shear_thick));
    return (r <= radius) ? 1.0 : 0.0;
}

/* core/nozzle.py:27-92 (ghost planes g < ng, interior j and k only; p = P_EQ on the whole face) */
int64_t orc_apply_nozzle_left_x(const orc_params *P, const orc_nozzle *N, double *pr, const double *draws,
                                int64_t ndraws) {
    const int ng = P->ng;
    int64_t used = 0;
    for (int g = 0; g < ng; ++g)
        for (int j = ng; j < ng + N->ny_loc; ++j) {
            double y = (j - ng + 0.5) * N->dy;
            for (int k = ng; k < ng + N->nz_loc; ++k) {
                double z = (k - ng + 0.5) * N->dz;
                double rr = sqrt(pow2(y - N->y0) + pow2(z - N->z0));
                double s = orc_nozzle_weight(rr, N->jet_radius, N->shear_thick, N->profile);
                double rho = N->eta_rho * N->rho_amb * s + N->rho_amb * (1.0 - s);
                double p = N->p_eq;
                double beta = sqrt(1.0 - 1.0 / (N->gamma_jet * N->gamma_jet));
                double vx = beta * s + N->vx_amb * (1.0 - s);
                double vy = N->vy_amb, vz = N->vz_amb;
                if (N->perturb != 0 && s > 0.0) {
                    double r0 = used < ndraws ? draws[used] : 0.0; ++used;
                    double r1 = used < ndraws ? draws[used] : 0.0; ++used;
                    double r2 = used < ndraws ? draws[used] : 0.0; ++used;
                    vy += N->turb_vamp * (2.0 * r0 - 1.0);
                    vz += N->turb_vamp * (2.0 * r1 - 1.0);
                    p *= (1.0 + N->turb_pamp * (2.0 * r2 - 1.0));
                }
                PR(0, g, j, k) = rho;
                PR(1, g, j, k) = vx;
                PR(2, g, j, k) = vy;
                PR(3, g, j, k) = vz;
                PR(4, g, j, k) = p;
                if (P->nv >= 15)  /* core/nozzle.py:60-61 */
                    for (int v = 5; v < 15; ++v) PR(v, g, j, k) = 0.0;
                for (int t = 0; t < N->n_tracers; ++t) {
                    double tn = N->tracer_nozzle[t], ta = N->tracer_amb[t];
                    PR(N->tracer_offset + t, g, j, k) = ta + (tn - ta) * s;
                }
                if (N->two_temperature && P->nv > N->thermo_offset + 1) {
                    PR(N->thermo_offset, g, j, k) = N->te_nozzle;
                    PR(N->thermo_offset + 1, g, j, k) = N->ti_nozzle;
                }
                if (N->chemistry && P->nv > N->chem_offset + 2)  /* core/nozzle.py:77-82 */
                    for (int m = 0; m < 3; ++m) PR(N->chem_offset + m, g, j, k) = N->chem_nozzle[m];
                if (P->physics == 1) {
                    double Bx = 0.0, By = 0.0, Bz = 0.0;
                    if (N->b_init == 1) Bx = N->b0 * s;
                    else if (N->b_init == 2) By = N->b0 * s;
                    PR(5, g, j, k) = Bx + N->b_amb[0];
                    PR(6, g, j, k) = By + N->b_amb[1];
                    PR(7, g, j, k) = Bz + N->b_amb[2];
                    PR(8, g, j, k) = 0.0;
                }
            }
        }
    return used;
}

/* core/source_terms.py:30-55 */
void orc_apply_two_temperature(const orc_params *P, double *pr, double dt, double tau, int off) {
    if (P->nv < off + 2 || tau <= 0.0) return;
    const int ng = P->ng;
    for (int i = ng; i < NXF - ng; ++i)
        for (int j = ng; j < NYF - ng; ++j)
            for (int k = ng; k < NZF - ng; ++k) {
                double rho = PR(0, i, j, k), p = PR(4, i, j, k);
                double T = p / pymax(rho, 1e-12);
                double Te = PR(off, i, j, k), Ti = PR(off + 1, i, j, k);
                Te += dt * (T - Te) / tau;
                Ti += dt * (T - Ti) / tau;
                PR(off, i, j, k) = Te;
                PR(off + 1, i, j, k) = Ti;
            }
}

/* core/source_terms.py:5-28 (apply_cooling_heating), interior cells */
void orc_apply_cooling_heating(const orc_params *P, double *pr, double dt, double cool, double heat, double pmax) {
    if (cool == 0.0 && heat == 0.0) return;
    const int ng = P->ng;
    for (int i = ng; i < NXF - ng; ++i)
        for (int j = ng; j < NYF - ng; ++j)
            for (int k = ng; k < NZF - ng; ++k) {
                double rho = PR(0, i, j, k), p = PR(4, i, j, k);
                double dp = dt * (heat - cool * rho * p);
                double p_new = p + dp;
                if (p_new < 1e-12) p_new = 1e-12;
                if (p_new > pmax) p_new = pmax;
                PR(4, i, j, k) = p_new;
            }
}

/* core/chemistry.py:8-40 (rates; T in K); Python's "a ** b" is libm pow */
#define CHEM_SMALL 1e-20
static double chem_clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double rate_ci(double a, double e, double T) {
    return a * sqrt(T) * exp(-e / pymax(T, 1.0)) / (1.0 + sqrt(T / 1.0e5));
}

/* core/chemistry.py:43-139 (apply_ion_chemistry), interior cells */
void orc_apply_ion_chemistry(const orc_params *P, double *pr, double dt, const orc_chem *c) {
    const int ng = P->ng, off = c->offset;
    if (P->nv <= off + 2) return;
    const double e_h = 157809.1, e_he1 = 285335.4, e_he2 = 631515.0;
    for (int i = ng; i < NXF - ng; ++i)
        for (int j = ng; j < NYF - ng; ++j)
            for (int k = ng; k < NZF - ng; ++k) {
                double rho = PR(0, i, j, k), p = PR(4, i, j, k);
                if (rho <= 0.0 || p <= 0.0) continue;
                double T = (p / pymax(rho, CHEM_SMALL)) * c->tunit;
                if (T < c->tmin) T = c->tmin;
                double nH = c->X * rho, nHe = c->Y * rho / 4.0;
                double xHII = chem_clamp(PR(off, i, j, k), 0.0, 1.0);
                double xHeII = chem_clamp(PR(off + 1, i, j, k), 0.0, 1.0);
                double xHeIII = chem_clamp(PR(off + 2, i, j, k), 0.0, 1.0);
                xHeII = pymin(xHeII, 1.0 - xHeIII);
                double xHeI = pymax(0.0, 1.0 - xHeII - xHeIII);
                double xHI = 1.0 - xHII;
                double ne = nH * xHII + nHe * (xHeII + 2.0 * xHeIII);
                ne = pymax(ne, CHEM_SMALL);
                double C_H = c->rate_scale * rate_ci(5.85e-11, 157809.1, T);
                double C_HeI = c->rate_scale * rate_ci(2.38e-11, 285335.4, T);
                double C_HeII = c->rate_scale * rate_ci(5.68e-12, 631515.0, T);
                double A_H = c->rate_scale * (2.59e-13 * libm_pow(T / 1.0e4, -0.7));
                double A_HeII = c->rate_scale * (1.50e-10 * libm_pow(T / 1.0e4, -0.6353));
                double A_HeIII = c->rate_scale * (3.36e-10 * libm_pow(T / 1.0e4, -0.5));
                double xHII_new = (xHII + dt * ne * C_H * xHI) / (1.0 + dt * ne * A_H);
                double xHeIII_new = (xHeIII + dt * ne * C_HeII * xHeII) / (1.0 + dt * ne * A_HeIII);
                xHeIII_new = chem_clamp(xHeIII_new, 0.0, 1.0);
                double xHeII_new = (xHeII + dt * ne * (C_HeI * xHeI + A_HeIII * xHeIII_new)) /
                                   (1.0 + dt * ne * (A_HeII + C_HeII));
                xHII_new = chem_clamp(xHII_new, 0.0, 1.0);
                xHeII_new = chem_clamp(xHeII_new, 0.0, 1.0 - xHeIII_new);
                double dH = (xHII_new - xHII), dHeII = (xHeII_new - xHeII), dHeIII = (xHeIII_new - xHeIII);
                double dE = c->energy_scale * (nH * e_h * dH + nHe * e_he1 * dHeII + nHe * e_he2 * dHeIII);
                double dp = -(c->gamma - 1.0) * dE / pymax(c->tunit, CHEM_SMALL);
                double p_new = p + dp;
                if (p_new < CHEM_SMALL) p_new = CHEM_SMALL;
                PR(4, i, j, k) = p_new;
                PR(off, i, j, k) = xHII_new;
                PR(off + 1, i, j, k) = xHeII_new;
                PR(off + 2, i, j, k) = xHeIII_new;
            }
}

/* core/dissipation.py:5-215 (apply_causal_dissipation), hydro only, nv >= 15: first-order upwind
 * advection of the ten dissipative fields with the cell velocity (:47-103, from a copy of the
 * fields), then the IMEX relaxation of Pi, pi^ij and q^i towards their Navier-Stokes targets
 * from centred velocity and temperature (p / rho) gradients, each clamped to +-cap_frac P_MAX;
 * pi^zz is reset to -(pi^xx + pi^yy).  Interior cells only. */
void orc_apply_causal_dissipation(const orc_params *P, double *pr, double dt, double dx, double dy, double dz,
                                  const orc_diss *D) {
    if ((D->zeta == 0.0 && D->eta == 0.0 && D->kappa == 0.0) ||
        pymin(pymin(D->tau_bulk, D->tau_shear), D->tau_heat) <= 0.0)
        return;
    if (P->nv < 15) return;
    const int ng = P->ng;
    const size_t N = (size_t)NXF * NYF * NZF;
    const double inv2dx = 0.5 / pymax(dx, 1e-12), inv2dy = 0.5 / pymax(dy, 1e-12), inv2dz = 0.5 / pymax(dz, 1e-12);
    if (D->advect) {
        double *adv = (double *)malloc(sizeof(double) * 10 * N);
        memcpy(adv, pr + 5 * N, sizeof(double) * 10 * N);
#pragma omp parallel for schedule(static)
        for (int i = ng; i < NXF - ng; ++i)
            for (int j = ng; j < NYF - ng; ++j)
                for (int k = ng; k < NZF - ng; ++k) {
                    double vx = PR(1, i, j, k), vy = PR(2, i, j, k), vz = PR(3, i, j, k);
                    for (int f = 5; f < 15; ++f) {
                        double u0 = PR(f, i, j, k), dudx, dudy, dudz;
                        if (vx > 0.0) dudx = (u0 - PR(f, i - 1, j, k)) / pymax(dx, 1e-12);
                        else dudx = (PR(f, i + 1, j, k) - u0) / pymax(dx, 1e-12);
                        if (vy > 0.0) dudy = (u0 - PR(f, i, j - 1, k)) / pymax(dy, 1e-12);
                        else dudy = (PR(f, i, j + 1, k) - u0) / pymax(dy, 1e-12);
                        if (vz > 0.0) dudz = (u0 - PR(f, i, j, k - 1)) / pymax(dz, 1e-12);
                        else dudz = (PR(f, i, j, k + 1) - u0) / pymax(dz, 1e-12);
                        adv[(size_t)(f - 5) * N + IX(0, i, j, k)] = u0 - dt * (vx * dudx + vy * dudy + vz * dudz);
                    }
                }
        memcpy(pr + 5 * N, adv, sizeof(double) * 10 * N);
        free(adv);
    }
    const double cap = D->cap_frac * D->p_max;
#define CLAMP(x) do { if ((x) > cap) (x) = cap; if ((x) < -cap) (x) = -cap; } while (0)
#define TEMP(ii, jj, kk) (PR(4, ii, jj, kk) / pymax(PR(0, ii, jj, kk), 1e-12))
#pragma omp parallel for schedule(static)
    for (int i = ng; i < NXF - ng; ++i)
        for (int j = ng; j < NYF - ng; ++j)
            for (int k = ng; k < NZF - ng; ++k) {
                double dvx_dx = (PR(1, i + 1, j, k) - PR(1, i - 1, j, k)) * inv2dx;
                double dvx_dy = (PR(1, i, j + 1, k) - PR(1, i, j - 1, k)) * inv2dy;
                double dvx_dz = (PR(1, i, j, k + 1) - PR(1, i, j, k - 1)) * inv2dz;
                double dvy_dx = (PR(2, i + 1, j, k) - PR(2, i - 1, j, k)) * inv2dx;
                double dvy_dy = (PR(2, i, j + 1, k) - PR(2, i, j - 1, k)) * inv2dy;
                double dvy_dz = (PR(2, i, j, k + 1) - PR(2, i, j, k - 1)) * inv2dz;
                double dvz_dx = (PR(3, i + 1, j, k) - PR(3, i - 1, j, k)) * inv2dx;
                double dvz_dy = (PR(3, i, j + 1, k) - PR(3, i, j - 1, k)) * inv2dy;
                double dvz_dz = (PR(3, i, j, k + 1) - PR(3, i, j, k - 1)) * inv2dz;
                double divv = dvx_dx + dvy_dy + dvz_dz;
                if (D->zeta != 0.0) {
                    double pi_eq = -D->zeta * divv;
                    double pi_new = (PR(5, i, j, k) + dt * pi_eq / D->tau_bulk) / (1.0 + dt / D->tau_bulk);
                    CLAMP(pi_new);
                    PR(5, i, j, k) = pi_new;
                }
                if (D->eta != 0.0) {
                    const double third = 1.0 / 3.0, ts = D->tau_shear;
                    double sxx = dvx_dx - third * divv, syy = dvy_dy - third * divv, szz = dvz_dz - third * divv;
                    double sxy = 0.5 * (dvx_dy + dvy_dx), sxz = 0.5 * (dvx_dz + dvz_dx), syz = 0.5 * (dvy_dz + dvz_dy);
                    double m2e = -2.0 * D->eta;
                    double e[6] = {m2e * sxx, m2e * syy, m2e * szz, m2e * sxy, m2e * sxz, m2e * syz};
                    double v[6];
                    for (int m = 0; m < 6; ++m) v[m] = (PR(6 + m, i, j, k) + dt * e[m] / ts) / (1.0 + dt / ts);
                    v[2] = -(v[0] + v[1]);
                    for (int m = 0; m < 6; ++m) {
                        CLAMP(v[m]);
                        PR(6 + m, i, j, k) = v[m];
                    }
                }
                if (D->kappa != 0.0) {
                    double dTdx = (TEMP(i + 1, j, k) - TEMP(i - 1, j, k)) * inv2dx;
                    double dTdy = (TEMP(i, j + 1, k) - TEMP(i, j - 1, k)) * inv2dy;
                    double dTdz = (TEMP(i, j, k + 1) - TEMP(i, j, k - 1)) * inv2dz;
                    double g[3] = {dTdx, dTdy, dTdz};
                    for (int m = 0; m < 3; ++m) {
                        double q_eq = -D->kappa * g[m];
                        double qn = (PR(12 + m, i, j, k) + dt * q_eq / D->tau_heat) / (1.0 + dt / D->tau_heat);
                        CLAMP(qn);
                        PR(12 + m, i, j, k) = qn;
                    }
                }
            }
#undef CLAMP
#undef TEMP
}
