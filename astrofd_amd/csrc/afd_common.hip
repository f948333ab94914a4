// afd_common.hip -- stage update / rhs kernels and the RMHD face dispatcher.
#include "afd_launch.hpp"

namespace afd {

// ------------------------------------------------------------------ non-template kernels
__global__ __launch_bounds__(BLOCK) void k_passive_rhs(PassiveArgs a) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= g.ncell) return;
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    bool in = i >= g.r0[0] && i < g.r1[0] && j >= g.r0[1] && j < g.r1[1] && k >= g.r0[2] && k < g.r1[2];
    for (int t = 0; t < a.np; ++t) {
        double r = 0.0;
        if (in) {
            for (int dir = 0; dir < 3; ++dir) {
                // which variants the reference evaluates per direction, and which left flux survives
                int nvar;           // right-face variants subtracted, in order WENO, PPM, MUSCL
                int vars[3];
                int left;           // variant of the surviving left flux
                if (a.physics == 1) {            // RMHD MUSCL: all three, every direction
                    nvar = 3; vars[0] = 0; vars[1] = 1; vars[2] = 2; left = 2;
                } else if (a.recon == 0) {       // SRHD MUSCL: x all three, y none, z MUSCL
                    if (dir == 1) continue;
                    if (dir == 0) { nvar = 3; vars[0] = 0; vars[1] = 1; vars[2] = 2; left = 2; }
                    else { nvar = 1; vars[0] = 2; left = 2; }
                } else {                         // SRHD PPM: x none, y all three, z WENO and PPM
                    if (dir == 0) continue;
                    if (dir == 1) { nvar = 3; vars[0] = 0; vars[1] = 1; vars[2] = 2; left = 2; }
                    else { nvar = 2; vars[0] = 0; vars[1] = 1; left = 1; }
                }
                const double *ax = a.aux[dir];
                long cp = c + (dir == 0 ? g.sx : (dir == 1 ? g.sy : 1L));
                int ip = i + (dir == 0), jp = j + (dir == 1), kp = k + (dir == 2);
                double sLa = ax[c], sRa = ax[g.ncell + c], vLa = ax[2 * g.ncell + c], vRa = ax[3 * g.ncell + c];
                double sLb = ax[cp], sRb = ax[g.ncell + cp], vLb = ax[2 * g.ncell + cp], vRb = ax[3 * g.ncell + cp];
                if (a.physics == 1) {  // right faces keep the left face's speeds (core/rmhd_core.py:915)
                    sLb = sLa;
                    sRb = sRa;
                }
                double FLt = passive_flux(a, left, a.po + t, dir, i, j, k, vLa, vRa, sLa, sRa);
                for (int m = 0; m < nvar; ++m) {
                    double FRt = passive_flux(a, vars[m], a.po + t, dir, ip, jp, kp, vLb, vRb, sLb, sRb);
                    r -= (FRt - FLt) / a.d[dir];
                }
            }
        }
        a.rhs[t * g.ncell + c] = r;
    }
}

__global__ __launch_bounds__(BLOCK) void k_c2p_batch(BatchArgs a) {
    long m = (long)blockIdx.x * BLOCK + threadIdx.x;
    C2PStats st = {0, 0, 0};
    if (m < a.n) {
        if (a.physics == 1) {
            double U[9], q[9];
            for (int v = 0; v < 9; ++v) U[v] = a.in[9 * m + v];
            RmhdC2PParams cp = {a.gamma, a.gfac, a.p_max, a.v_max};
            rmhd_c2p(cp, U, q, st);
            for (int v = 0; v < 9; ++v) a.out[9 * m + v] = q[v];
        } else {
            double U[5], q[5];
            for (int v = 0; v < 5; ++v) U[v] = a.in[5 * m + v];
            hydro_c2p(a.gamma, a.gfac, a.p_max, a.v_max, U, q, st);
            for (int v = 0; v < 5; ++v) a.out[5 * m + v] = q[v];
        }
        if (a.iters) a.iters[m] = st.hydro_iters + st.newton_iters;
    }
    if (a.stats) wave_add_stats(a.stats, st);  // reached by every lane
}

__global__ __launch_bounds__(BLOCK) void k_p2c_batch(BatchArgs a) {
    long m = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (m >= a.n) return;
    if (a.physics == 1) {
        double q[9], U[9];
        for (int v = 0; v < 9; ++v) q[v] = a.in[9 * m + v];
        rmhd_p2c_cell(a.gfac, a.p_max, a.v_max, q, U);
        for (int v = 0; v < 9; ++v) a.out[9 * m + v] = U[v];
    } else {
        double q[5], U[5];
        for (int v = 0; v < 5; ++v) q[v] = a.in[5 * m + v];
        srhd_p2c_cell(a.gfac, q, U);
        for (int v = 0; v < 5; ++v) a.out[5 * m + v] = U[v];
    }
}

__global__ __launch_bounds__(BLOCK) void k_fill_y(BcArgs a) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    long tot = (long)g.nv * g.nx * 2 * a.ng * g.nz;
    if (t >= tot) return;
    int k = (int)(t % g.nz);
    long r = t / g.nz;
    int gg = (int)(r % (2 * a.ng));
    r /= 2 * a.ng;
    int i = (int)(r % g.nx);
    int v = (int)(r / g.nx);
    int jd, js;
    if (gg < a.ng) {
        jd = gg;
        js = a.bc ? a.ng : g.ny - 2 * a.ng + gg;
    } else {
        jd = g.ny - a.ng + (gg - a.ng);
        js = a.bc ? g.ny - a.ng - 1 : a.ng + (gg - a.ng);
    }
    double *b = a.P + (long)v * g.ncell;
    b[cidx(g, i, jd, k)] = b[cidx(g, i, js, k)];
}

__global__ __launch_bounds__(BLOCK) void k_fill_z(BcArgs a) {
    const Grid &g = a.g;
    long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    long tot = (long)g.nv * g.nx * g.ny * 2 * a.ng;
    if (t >= tot) return;
    int gg = (int)(t % (2 * a.ng));
    long r = t / (2 * a.ng);
    int j = (int)(r % g.ny);
    r /= g.ny;
    int i = (int)(r % g.nx);
    int v = (int)(r / g.nx);
    int kd, ks;
    if (gg < a.ng) {
        kd = gg;
        ks = a.bc ? a.ng : g.nz - 2 * a.ng + gg;
    } else {
        kd = g.nz - a.ng + (gg - a.ng);
        ks = a.bc ? g.nz - a.ng - 1 : a.ng + (gg - a.ng);
    }
    double *b = a.P + (long)v * g.ncell;
    b[cidx(g, i, j, kd)] = b[cidx(g, i, j, ks)];
}

__global__ __launch_bounds__(BLOCK) void k_outflow_x(BcArgs a) {
    const Grid &g = a.g;
    long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    long plane = (long)g.ny * g.nz;
    long tot = (long)g.nv * a.ng * plane;
    if (t >= tot) return;
    long jk = t % plane;
    long r = t / plane;
    int gi = (int)(r % a.ng);
    int v = (int)(r / a.ng);
    double *b = a.P + (long)v * g.ncell;
    b[(long)(g.nx - a.ng + gi) * plane + jk] = b[(long)(g.nx - a.ng - 1) * plane + jk];
}

__global__ __launch_bounds__(BLOCK) void k_nozzle(NozzleArgs a) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    long plane = (long)a.ny_int * a.nz_int;
    if (t >= (long)a.ng * plane) return;
    long jk = t % plane;
    int gi = (int)(t / plane);
    int jj = (int)(jk / a.nz_int), kk = (int)(jk % a.nz_int);
    long c = cidx(g, gi, a.ng + jj, a.ng + kk);
    for (int v = 0; v < g.nv && v < 32; ++v) {
        if (!((a.mask >> v) & 1u)) continue;
        double val = a.tmpl[(long)v * plane + jk];
        int rank = a.idx[jk];
        if (a.perturb && rank >= 0 && v >= 2 && v <= 4) {
            const double *d = a.draws + 3 * ((long)gi * a.nact + rank);
            if (v == 2) val = val + a.vamp * (2.0 * d[0] - 1.0);
            else if (v == 3) val = val + a.vamp * (2.0 * d[1] - 1.0);
            else val = val * (1.0 + a.pamp * (2.0 * d[2] - 1.0));
        }
        a.P[(long)v * g.ncell + c] = val;
    }
}

// core/chemistry.py:17-27: collisional ionization fit, T in K
__device__ static double rate_ci(double a, double e, double T) {
#pragma clang fp contract(off)
    return a * sqrt(T) * exp(-e / pymax(T, 1.0)) / (1.0 + sqrt(T / 1.0e5));
}

__device__ static double clamp01(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

__global__ __launch_bounds__(BLOCK) void k_sources(SourceArgs a) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    const int nyi = g.ny - 2 * a.ng, nzi = g.nz - 2 * a.ng;
    const int kk = blockIdx.y * BLOCK + threadIdx.x;
    if (kk >= nzi) return;
    const int i = a.ng + (int)blockIdx.x / nyi, j = a.ng + (int)blockIdx.x % nyi, k = a.ng + kk;
    const long c = cidx(g, i, j, k), n = g.ncell;
    double *P = a.P;
    const double rho = P[c];
    double p = P[4 * n + c];
    if (a.which & SRC_COOL) {  // core/source_terms.py:19-27
        double p_new = p + a.dt * (a.heat - a.cool * rho * p);
        if (p_new < 1e-12) p_new = 1e-12;
        if (p_new > a.p_cap) p_new = a.p_cap;
        p = p_new;
        P[4 * n + c] = p;
    }
    if (a.which & SRC_TWO_TEMP) {  // core/source_terms.py:45-54
        double T = p / pymax(rho, 1e-12);
        double *Te = P + (long)a.t_off * n + c, *Ti = Te + n;
        double te = *Te, ti = *Ti;
        te += a.dt * (T - te) / a.tau;
        ti += a.dt * (T - ti) / a.tau;
        *Te = te;
        *Ti = ti;
    }
    if ((a.which & SRC_CHEM) && rho > 0.0 && p > 0.0) {  // core/chemistry.py:78-137
        const double SM = 1e-20;
        double T = (p / pymax(rho, SM)) * a.tunit;
        if (T < a.tmin) T = a.tmin;
        double nH = a.X * rho, nHe = a.Y * rho / 4.0;
        double *x = P + (long)a.c_off * n + c;
        double xHII = clamp01(x[0], 0.0, 1.0), xHeII = clamp01(x[n], 0.0, 1.0), xHeIII = clamp01(x[2 * n], 0.0, 1.0);
        xHeII = pymin(xHeII, 1.0 - xHeIII);
        double xHeI = pymax(0.0, 1.0 - xHeII - xHeIII);
        double xHI = 1.0 - xHII;
        double ne = pymax(nH * xHII + nHe * (xHeII + 2.0 * xHeIII), SM);
        double C_H = a.rate_scale * rate_ci(5.85e-11, 157809.1, T);
        double C_HeI = a.rate_scale * rate_ci(2.38e-11, 285335.4, T);
        double C_HeII = a.rate_scale * rate_ci(5.68e-12, 631515.0, T);
        // case-B recombination, core/chemistry.py:30-40: (T / 1e4)^y as exp(y log(T / 1e4)), one
        // log for the three powers; |y log| <= ~20 over T in [1, 1e12] K, so the result is within
        // ~20 ulp of the correctly rounded power
        double lt4 = log(T / 1.0e4);
        double A_H = a.rate_scale * (2.59e-13 * exp(-0.7 * lt4));
        double A_HeII = a.rate_scale * (1.50e-10 * exp(-0.6353 * lt4));
        double A_HeIII = a.rate_scale * (3.36e-10 * exp(-0.5 * lt4));
        double dne = a.dt * ne;
        double xHII_new = (xHII + dne * C_H * xHI) / (1.0 + dne * A_H);
        double xHeIII_new = clamp01((xHeIII + dne * C_HeII * xHeII) / (1.0 + dne * A_HeIII), 0.0, 1.0);
        double xHeII_new = (xHeII + dne * (C_HeI * xHeI + A_HeIII * xHeIII_new)) / (1.0 + dne * (A_HeII + C_HeII));
        xHII_new = clamp01(xHII_new, 0.0, 1.0);
        xHeII_new = clamp01(xHeII_new, 0.0, 1.0 - xHeIII_new);
        double dE = a.energy_scale * (nH * 157809.1 * (xHII_new - xHII) + nHe * 285335.4 * (xHeII_new - xHeII) +
                                      nHe * 631515.0 * (xHeIII_new - xHeIII));
        double p_new = p + -(a.gamma - 1.0) * dE / pymax(a.tunit, SM);
        if (p_new < SM) p_new = SM;
        P[4 * n + c] = p_new;
        x[0] = xHII_new;
        x[n] = xHeII_new;
        x[2 * n] = xHeIII_new;
    }
}

__global__ __launch_bounds__(BLOCK) void k_diss(DissArgs a) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    const int nyi = g.ny - 2 * a.ng, nzi = g.nz - 2 * a.ng;
    const int kk = blockIdx.y * BLOCK + threadIdx.x;
    if (kk >= nzi) return;
    const int i = a.ng + (int)blockIdx.x / nyi, j = a.ng + (int)blockIdx.x % nyi, k = a.ng + kk;
    const long c = cidx(g, i, j, k), n = g.ncell, sx = g.sx, sy = g.sy;
    const double *P = a.P;
    double f[10];
#pragma unroll
    for (int m = 0; m < 10; ++m) f[m] = P[(5 + m) * n + c];
    const double vx = P[n + c], vy = P[2 * n + c], vz = P[3 * n + c];
    if (a.advect) {  // core/dissipation.py:60-92, first-order upwind with the cell's velocity
        const double hx = pymax(a.dx, 1e-12), hy = pymax(a.dy, 1e-12), hz = pymax(a.dz, 1e-12);
        const long ox = vx > 0.0 ? -sx : sx, oy = vy > 0.0 ? -sy : sy, oz = vz > 0.0 ? -1 : 1;
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            const double *q = P + (5 + m) * n + c;
            const double u0 = f[m];
            double dudx = vx > 0.0 ? (u0 - q[ox]) / hx : (q[ox] - u0) / hx;
            double dudy = vy > 0.0 ? (u0 - q[oy]) / hy : (q[oy] - u0) / hy;
            double dudz = vz > 0.0 ? (u0 - q[oz]) / hz : (q[oz] - u0) / hz;
            f[m] = u0 - a.dt * (vx * dudx + vy * dudy + vz * dudz);
        }
    }
    const double i2x = 0.5 / pymax(a.dx, 1e-12), i2y = 0.5 / pymax(a.dy, 1e-12), i2z = 0.5 / pymax(a.dz, 1e-12);
    const double *V1 = P + n + c, *V2 = P + 2 * n + c, *V3 = P + 3 * n + c;
    const double dvx_dx = (V1[sx] - V1[-sx]) * i2x, dvx_dy = (V1[sy] - V1[-sy]) * i2y, dvx_dz = (V1[1] - V1[-1]) * i2z;
    const double dvy_dx = (V2[sx] - V2[-sx]) * i2x, dvy_dy = (V2[sy] - V2[-sy]) * i2y, dvy_dz = (V2[1] - V2[-1]) * i2z;
    const double dvz_dx = (V3[sx] - V3[-sx]) * i2x, dvz_dy = (V3[sy] - V3[-sy]) * i2y, dvz_dz = (V3[1] - V3[-1]) * i2z;
    const double divv = dvx_dx + dvy_dy + dvz_dz;
    const double cap = a.cap;
    auto clamp = [cap](double x) { return x > cap ? cap : (x < -cap ? -cap : x); };
    if (a.zeta != 0.0) {  // bulk, :122-130
        double pi_eq = -a.zeta * divv;
        f[0] = clamp((f[0] + a.dt * pi_eq / a.tau_bulk) / (1.0 + a.dt / a.tau_bulk));
    }
    if (a.eta != 0.0) {  // shear, :132-177 (pi^zz reset to -(pi^xx + pi^yy) before the caps)
        const double third = 1.0 / 3.0, m2e = -2.0 * a.eta, ts = a.tau_shear;
        const double e[6] = {m2e * (dvx_dx - third * divv), m2e * (dvy_dy - third * divv),
                             m2e * (dvz_dz - third * divv), m2e * (0.5 * (dvx_dy + dvy_dx)),
                             m2e * (0.5 * (dvx_dz + dvz_dx)), m2e * (0.5 * (dvy_dz + dvz_dy))};
        double s6[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) s6[m] = (f[1 + m] + a.dt * e[m] / ts) / (1.0 + a.dt / ts);
        s6[2] = -(s6[0] + s6[1]);
#pragma unroll
        for (int m = 0; m < 6; ++m) f[1 + m] = clamp(s6[m]);
    }
    if (a.kappa != 0.0) {  // heat flux, :179-213, T = p / max(rho, 1e-12)
        const double *R = P + c, *Pp = P + 4 * n + c;
        auto T = [&](long o) { return Pp[o] / pymax(R[o], 1e-12); };
        const double g3[3] = {(T(sx) - T(-sx)) * i2x, (T(sy) - T(-sy)) * i2y, (T(1) - T(-1)) * i2z};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double q_eq = -a.kappa * g3[m];
            f[7 + m] = clamp((f[7 + m] + a.dt * q_eq / a.tau_heat) / (1.0 + a.dt / a.tau_heat));
        }
    }
#pragma unroll
    for (int m = 0; m < 10; ++m) a.S[m * n + c] = f[m];
}

__global__ __launch_bounds__(BLOCK) void k_diss_store(DissArgs a, double *P) {
    const Grid &g = a.g;
    const int nyi = g.ny - 2 * a.ng, nzi = g.nz - 2 * a.ng;
    const int kk = blockIdx.y * BLOCK + threadIdx.x;
    if (kk >= nzi) return;
    const int i = a.ng + (int)blockIdx.x / nyi, j = a.ng + (int)blockIdx.x % nyi, k = a.ng + kk;
    const long c = cidx(g, i, j, k), n = g.ncell;
#pragma unroll
    for (int m = 0; m < 10; ++m) P[(5 + m) * n + c] = a.S[m * n + c];
}

__global__ __launch_bounds__(BLOCK) void k_gr_rescale(GrRescaleArgs a) {
#pragma clang fp contract(off)
    const Grid &g = a.g;
    const long c = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= g.ncell) return;
    const long n = g.ncell;
    for (int v = 0; v < g.nv; ++v) a.Q[v * n + c] = a.P[v * n + c];
    int i, j, k;
    cell_ijk(g, c, i, j, k);
    if (i < a.ng || i >= g.nx - a.ng || j < a.ng || j >= g.ny - a.ng || k < a.ng || k >= g.nz - a.ng) return;
    double X[3];
    gr_xyz(a.x0, a.ng, i, j, k, a.dx, a.dy, a.dz, X);
    const double ax = pymax(gr_ks_alpha(X[0], X[1], X[2], a.mass), 1e-12);
    for (int v = 1; v <= 3; ++v) a.Q[v * n + c] = a.P[v * n + c] / ax;
}

__global__ __launch_bounds__(BLOCK) void k_max_speed_box(const double *P, Grid g, SpeedBox b, double gamma,
                                                          double gfac, unsigned long long *amax) {
    const long ni = b.hi[0] - b.lo[0], nj = b.hi[1] - b.lo[1], nk = b.hi[2] - b.lo[2], tot = ni * nj * nk;
    double m = 0.0;
    for (long t = (long)blockIdx.x * BLOCK + threadIdx.x; t < tot; t += (long)gridDim.x * BLOCK) {
        const int k = b.lo[2] + (int)(t % nk), j = b.lo[1] + (int)((t / nk) % nj), i = b.lo[0] + (int)(t / (nj * nk));
        const long c = cidx(g, i, j, k);
        const double loc = srhd_cell_speed(gamma, gfac, P[c], P[g.ncell + c], P[2 * g.ncell + c],
                                           P[3 * g.ncell + c], P[4 * g.ncell + c]);
        if (loc > m) m = loc;
    }
    wave_max_atomic(amax, m);
}

__global__ __launch_bounds__(BLOCK) void k_max_speed_shell(const double *P, Grid g, SpeedBox b, double gamma,
                                                            double gfac, unsigned long long *amax) {
    // range [1, n-1) per axis, box [lo, hi) inside it: slabs i outside (all j, k), then i inside
    // with j outside (all k), then i, j inside with k outside
    const long bi = g.nx - 2, bj = g.ny - 2, bk = g.nz - 2;
    const long ni = b.hi[0] - b.lo[0], nj = b.hi[1] - b.lo[1], nk = b.hi[2] - b.lo[2];
    const long oi = bi - ni, oj = bj - nj, ok = bk - nk;
    const long nA = oi * bj * bk, nB = ni * oj * bk, nC = ni * nj * ok;
    long t = (long)blockIdx.x * BLOCK + threadIdx.x;
    double v = 0.0;
    if (t < nA + nB + nC) {
        auto out = [&](long p, int ax) { return p < b.lo[ax] - 1 ? 1 + p : b.hi[ax] + (p - (b.lo[ax] - 1)); };
        long i, j, k;
        if (t < nA) {
            k = 1 + t % bk;
            j = 1 + (t / bk) % bj;
            i = out(t / (bk * bj), 0);
        } else if (t < nA + nB) {
            t -= nA;
            k = 1 + t % bk;
            j = out((t / bk) % oj, 1);
            i = b.lo[0] + t / (bk * oj);
        } else {
            t -= nA + nB;
            k = out(t % ok, 2);
            j = b.lo[1] + (t / ok) % nj;
            i = b.lo[0] + t / (ok * nj);
        }
        const long c = cidx(g, (int)i, (int)j, (int)k);
        v = srhd_cell_speed(gamma, gfac, P[c], P[g.ncell + c], P[2 * g.ncell + c], P[3 * g.ncell + c],
                            P[4 * g.ncell + c]);
    }
    wave_max_atomic(amax, v);
}

__global__ __launch_bounds__(BLOCK) void k_amax_finish(unsigned long long *amax, double floor_value, double *out) {
    __shared__ double red[BLOCK / 64];
    double m = 0.0;
    for (int t = threadIdx.x; t < 2 * AMAX_SLOTS; t += BLOCK) {
        unsigned long long *slot = amax + (long)t * AMAX_STRIDE;
        m = pymax(m, __longlong_as_double((long long)*slot));
        if (t >= AMAX_SLOTS) *slot = 0ull;  // each shell slot is read and cleared by this thread only
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = pymax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) m = pymax(m, red[w]);
        out[0] = m < floor_value ? floor_value : m;
    }
}

__global__ __launch_bounds__(BLOCK) void k_count_nonfinite(const double *x, long n, unsigned long long *partial) {
    __shared__ unsigned long long red[BLOCK];
    unsigned long long cnt = 0;
    for (long t = (long)blockIdx.x * BLOCK + threadIdx.x; t < n; t += (long)gridDim.x * BLOCK)
        cnt += isfinite(x[t]) ? 0ull : 1ull;
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(BLOCK) void k_reduce_sum_u64(const unsigned long long *partial, int n,
                                                           unsigned long long *out) {
    __shared__ unsigned long long red[BLOCK];
    unsigned long long m = 0;
    for (int t = threadIdx.x; t < n; t += BLOCK) m += partial[t];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <typename T> void launch_face_rmhd_t(const FaceLaunch &l, hipStream_t s);

hipError_t launch_face_rmhd(const FaceLaunch &l, hipStream_t s) {
    if (l.mixed) launch_face_rmhd_t<float>(l, s);
    else launch_face_rmhd_t<double>(l, s);
    return hipGetLastError();
}

// One stage update: k_update (flux divergence, SSPRK, con2prim).  With a Newton start other than
// the plain one, rejected cells are queued and solved densely by k_c2p_queue afterwards.
template <int NB, typename F, bool LEARNED> static void upd_stage(const UpdArgs<F> &a, dim3 grid, hipStream_t s) {
    if constexpr (NB == 9) {
        if (a.queue) {
            (void)hipMemsetAsync(a.qcount, 0, sizeof(unsigned) * QUEUE_SLOTS, s);
            hipLaunchKernelGGL((k_update<NB, F, LEARNED>), grid, dim3(BLOCK), 0, s, a);
            hipLaunchKernelGGL((k_c2p_queue<F>), dim3(16, QUEUE_SLOTS), dim3(BLOCK), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((k_update<NB, F, LEARNED>), grid, dim3(BLOCK), 0, s, a);
}

template <int NB, typename F> static void upd(const UpdLaunch &l, bool rhs_only, hipStream_t s) {
    UpdArgs<F> a;
    static_assert(sizeof(UpdArgs<F>) == sizeof(UpdArgs<double>), "layout");
    memcpy((void *)&a, (const void *)&l.a, sizeof(a));
    dim3 grid = grid1d(l.a.g.ncell);
    if (rhs_only) hipLaunchKernelGGL((k_rhs<NB, F>), grid, dim3(BLOCK), 0, s, a);
    else if constexpr (NB == 9) {
        if (l.a.mlp) upd_stage<NB, F, true>(a, grid, s);
        else upd_stage<NB, F, false>(a, grid, s);
    } else {
        upd_stage<NB, F, false>(a, grid, s);
    }
}

hipError_t launch_update(const UpdLaunch &l, bool rhs_only, hipStream_t s) {
    if (l.nb == 9) {
        if (l.mixed) upd<9, float>(l, rhs_only, s);
        else upd<9, double>(l, rhs_only, s);
    } else {
        if (l.mixed) upd<5, float>(l, rhs_only, s);
        else upd<5, double>(l, rhs_only, s);
    }
    return hipGetLastError();
}

}  // namespace afd

namespace afd {

__global__ __launch_bounds__(BLOCK) void k_diag(DiagArgs a) {
    const Grid &g = a.g;
    __shared__ double red[DIAG_N][BLOCK];
    const long ni = g.nx - 2 * a.ng, nj = g.ny - 2 * a.ng, nk = g.nz - 2 * a.ng, tot = ni * nj * nk;
    double v[DIAG_N] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const long n = g.ncell;
    for (long t = (long)blockIdx.x * BLOCK + threadIdx.x; t < tot; t += (long)gridDim.x * BLOCK) {
        const int k = a.ng + (int)(t % nk), j = a.ng + (int)((t / nk) % nj), i = a.ng + (int)(t / (nj * nk));
        const long c = cidx(g, i, j, k);
        double vx = a.P[n + c], vy = a.P[2 * n + c], vz = a.P[3 * n + c];
        double v2 = vx * vx + vy * vy + vz * vz;
        if (v2 >= 1.0) v2 = 1.0 - 1e-14;
        double W = 1.0 / sqrt(1.0 - v2);
        if (W > v[0]) v[0] = W;
        if (a.rmhd) {
            double dBx = (a.P[5 * n + c + g.sx] - a.P[5 * n + c - g.sx]) / (2.0 * a.dx);
            double dBy = (a.P[6 * n + c + g.sy] - a.P[6 * n + c - g.sy]) / (2.0 * a.dy);
            double dBz = (a.P[7 * n + c + 1] - a.P[7 * n + c - 1]) / (2.0 * a.dz);
            double d = dBx + dBy + dBz;
            if (fabs(d) > v[1]) v[1] = fabs(d);
            v[2] += d;
            v[3] += d * d;
        }
        if (i == a.ng) {
            double q[9], U[9];
            for (int m = 0; m < (a.rmhd ? 9 : 5); ++m) q[m] = a.P[m * n + c];
            if (a.rmhd) rmhd_p2c_cell(a.gfac53, a.p_max, a.v_max, q, U);
            else srhd_p2c_cell(a.gfac53, q, U);
            v[4] += U[1] * a.dy * a.dz;
        }
    }
    for (int m = 0; m < DIAG_N; ++m) red[m][threadIdx.x] = v[m];
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] = pymax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
            red[1][threadIdx.x] = pymax(red[1][threadIdx.x], red[1][threadIdx.x + o]);
            for (int m = 2; m < DIAG_N; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_N) a.partial[threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(BLOCK) void k_diag_reduce(const double *partial, int nblocks, double *out) {
    __shared__ double red[DIAG_N][BLOCK];
    double v[DIAG_N] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < nblocks; t += BLOCK) {
        v[0] = pymax(v[0], partial[t]);
        v[1] = pymax(v[1], partial[nblocks + t]);
        for (int m = 2; m < DIAG_N; ++m) v[m] += partial[m * nblocks + t];
    }
    for (int m = 0; m < DIAG_N; ++m) red[m][threadIdx.x] = v[m];
    __syncthreads();
    for (int o = BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] = pymax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
            red[1][threadIdx.x] = pymax(red[1][threadIdx.x], red[1][threadIdx.x + o]);
            for (int m = 2; m < DIAG_N; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_N) out[threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(BLOCK) void k_centerline(const double *P, Grid g, int ng, double v_max, double *out) {
    const int nxi = g.nx - 2 * ng;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= nxi) return;
    const int j = ng + (g.ny - 2 * ng) / 2, k = ng + (g.nz - 2 * ng) / 2;
    const long n = g.ncell, c = cidx(g, ng + t, j, k);
    double vx = P[n + c], vy = P[2 * n + c], vz = P[3 * n + c];
    double v2 = vx * vx + vy * vy + vz * vz;
    const double vmax2 = v_max * v_max;
    if (v2 >= vmax2) {
        double fac = v_max / sqrt(v2 + 1e-32);
        vx *= fac;
        v2 = vmax2 - 1e-14;
    }
    double v2c = v2 < 0.0 ? 0.0 : (v2 > 1.0 - 1e-14 ? 1.0 - 1e-14 : v2);
    out[t] = 1.0 / sqrt(1.0 - v2c);
    out[nxi + t] = P[c];
    out[2 * nxi + t] = P[4 * n + c];
    out[3 * nxi + t] = vx;
}

}  // namespace afd
