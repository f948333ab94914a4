// afd_capi.hip -- the extern "C" entry points of libafd.so (declared in include/afd.h).
//
// A context owns the device-resident state of one domain: three primitive buffers
// (step input and two stage outputs), the running conserved state U_s, the face-flux
// arrays of the three directions, the nozzle profile and the reduction scratch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "afd.h"
#include "afd_launch.hpp"
#include "afd_mlp.hpp"

using namespace afd;

static thread_local std::string g_err;
static int resolve_scratch(int device);

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(AFD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct afd_ctx {
    afd_params p{};
    Grid g{};
    int nb = 5;
    bool carried = false, passive_flux = false;
    bool shared_edges = false;  // mixed RMHD WENO5: the shared-edge face kernels
    bool dmode = false;  // ... accumulating the flux divergence in Fx
    bool fused = false;  // ... with the y faces evaluated by the stage update
    int diss = 0;  // FaceArgs::diss of the SRHD faces
    double *Puse = nullptr;  // GR ORTHONORMAL_FLUX: the stage input with rescaled velocities
    size_t fsize = 8;
    hipStream_t stream = nullptr;
    double *P[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    double *Uc = nullptr;
    void *Fx = nullptr, *Fxc = nullptr, *Fy = nullptr, *Fz = nullptr, *zs = nullptr;
    double *aux[3] = {nullptr, nullptr, nullptr};
    double *rhs_pass = nullptr, *rhs_out = nullptr;
    // nozzle
    bool noz_set = false;
    afd_nozzle noz{};
    double *noz_tmpl = nullptr, *noz_draws = nullptr;
    int *noz_idx = nullptr;
    int noz_nact = 0;
    unsigned noz_mask = 0;
    // reductions
    int red_blocks = 1024;
    double *red_out = nullptr;
    // the CFL speed's maximum over the speed box, as double bits in AMAX_SLOTS slots, from an SRHD
    // step's last stage or from k_max_speed_box, valid until something other than a step or a
    // fill changes the state; then AMAX_SLOTS slots of the shell's maximum, zero between calls
    unsigned long long *amax_dev = nullptr;
    SpeedBox sbox;
    bool speed_valid = false;
    unsigned long long *cnt_partial = nullptr, *cnt_out = nullptr;
    double *diag_partial = nullptr, *diag_out = nullptr;
    // instrumentation
    bool stats_on = false, timing_on = false;
    unsigned long long *stats_dev = nullptr;
    unsigned char *flags_dev = nullptr;  // per-cell con2prim branch bits of the last stage (stats on)
    afd_stats acc{};
    struct Mark {
        int cat;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> ev_pool;
    // learned guess
    MlpDevice mlp{};
    int *c2p_queue = nullptr;         // cells whose learned start failed the acceptance test (QUEUE_SLOTS regions)
    unsigned *c2p_qcount = nullptr;
    int *repair_queue = nullptr;      // flux-divergence path: cells whose FP32 right-hand side was not finite
    unsigned *repair_qcount = nullptr;
    long c2p_qcap = 0;
    std::vector<void *> allocs;
};

static int dalloc(afd_ctx *c, void **ptr, size_t bytes) {
    if (bytes == 0) {
        *ptr = nullptr;
        return AFD_OK;
    }
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(AFD_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    }
    c->allocs.push_back(*ptr);
    return AFD_OK;
}

#define ALLOC(ctx, ptr, bytes)                                         \
    do {                                                               \
        int rc_ = dalloc((ctx), (void **)&(ptr), (bytes));             \
        if (rc_ != AFD_OK) return rc_;                                 \
    } while (0)

static hipEvent_t ev_get(afd_ctx *c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// categories: 0,1,2 face x,y,z; 3 update; 4 other
struct Timed {
    afd_ctx *c;
    int cat;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(afd_ctx *ctx, int category) : c(ctx), cat(category) {
        if (c->timing_on) {
            a = ev_get(c);
            b = ev_get(c);
            if (a && b) (void)hipEventRecord(a, c->stream);
        }
    }
    ~Timed() {
        if (c->timing_on && a && b) {
            (void)hipEventRecord(b, c->stream);
            c->marks.push_back({cat, a, b});
        }
    }
};

static int resolve_marks(afd_ctx *c) {
    if (c->marks.empty()) return AFD_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &m : c->marks) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, m.a, m.b));
        if (m.cat < 3) {
            c->acc.ms_face[m.cat] += ms;
            c->acc.launches_face[m.cat] += 1;
        } else if (m.cat == 3) {
            c->acc.ms_update += ms;
            c->acc.launches_update += 1;
        } else {
            c->acc.ms_other += ms;
        }
        c->ev_pool.push_back(m.a);
        c->ev_pool.push_back(m.b);
    }
    c->marks.clear();
    return AFD_OK;
}

static Phys<double> make_phys(const afd_params &p) {
    Phys<double> ph;
    ph.gamma = p.gamma;
    ph.gfac = p.gamma / (p.gamma - 1.0);
    ph.gm1_over_g_inv = p.gamma - 1.0;
    ph.p_max = p.p_max;
    ph.v_max = p.v_max;
    ph.vmax2 = p.v_max * p.v_max;
    ph.glm_ch2 = p.glm_ch * p.glm_ch;
    return ph;
}

static Phys<float> to_float(const Phys<double> &d) {
    Phys<float> f;
    f.gamma = (float)d.gamma;
    f.gfac = (float)d.gfac;
    f.gm1_over_g_inv = (float)d.gm1_over_g_inv;
    f.p_max = (float)d.p_max;
    f.v_max = (float)d.v_max;
    f.vmax2 = (float)d.vmax2;
    f.glm_ch2 = (float)d.glm_ch2;
    return f;
}

extern "C" const char *afd_last_error(void) { return g_err.c_str(); }
extern "C" const char *afd_version(void) { return "astrofd_amd 0.1 (gfx950)"; }

extern "C" int afd_create(const afd_params *params, afd_ctx **out) {
    if (!params || !out) return fail(AFD_ERR_INVALID, "afd_create: null argument");
    afd_params p = *params;
    if (p.physics != AFD_PHYS_HYDRO && p.physics != AFD_PHYS_RMHD) return fail(AFD_ERR_INVALID, "physics must be 0 or 1");
    int nb = p.physics == AFD_PHYS_RMHD ? 9 : 5;
    if (p.ng < 2) return fail(AFD_ERR_INVALID, "NG must be >= 2");
    if (p.nx < 2 || p.ny < 2 || p.nz < 2) return fail(AFD_ERR_INVALID, "grid too small");
    // cells and faces are indexed in 32 bits inside the kernels
    if ((int64_t)(p.nx + 2 * p.ng) * (p.ny + 2 * p.ng) * (p.nz + 2 * p.ng) >= (int64_t(1) << 31))
        return fail(AFD_ERR_INVALID, "grid too large: 2^31 or more cells with ghosts");
    // passive_offset < nb means "right after the base variables"; SRHD may hold fields between
    // the base variables and the passives (the dissipative Pi, pi^ij, q^i at 5..14 with
    // PASSIVE_OFFSET = 15, utils/settings.py:250-252), RMHD may not
    if (p.passive_offset < nb) p.passive_offset = nb;
    if (nb == 9 && p.passive_offset != 9) return fail(AFD_ERR_INVALID, "RMHD passives follow the 9 base variables");
    if (p.n_passive < 0 || p.nv != p.passive_offset + p.n_passive)
        return fail(AFD_ERR_INVALID, "nv must equal passive_offset + n_passive");
    if (p.nv > 32) return fail(AFD_ERR_INVALID, "nv > 32");
    if (p.recon_id < 0 || p.recon_id > 2) return fail(AFD_ERR_INVALID, "recon_id");
    if (p.limiter_id < 0 || p.limiter_id > 2) return fail(AFD_ERR_INVALID, "limiter_id");
    if (p.riemann_id < 0 || p.riemann_id > (nb == 9 ? 2 : 1)) return fail(AFD_ERR_INVALID, "riemann_id");
    if (p.rk_order != 2 && p.rk_order != 3) return fail(AFD_ERR_INVALID, "rk_order must be 2 or 3");
    if (p.precision != AFD_PREC_FP64 && p.precision != AFD_PREC_MIXED) return fail(AFD_ERR_INVALID, "precision");
    if (p.c2p_guess != AFD_C2P_PLAIN && p.c2p_guess != AFD_C2P_LEARNED && p.c2p_guess != AFD_C2P_PREVIOUS)
        return fail(AFD_ERR_INVALID, "c2p_guess");
    if (p.c2p_guess != AFD_C2P_PLAIN && nb != 9) return fail(AFD_ERR_INVALID, "the con2prim guesses are RMHD only");
    if (p.two_temperature && (p.thermo_offset < nb || p.thermo_offset + 1 >= p.nv))
        return fail(AFD_ERR_INVALID, "thermo_offset outside the passive block");
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0)) return fail(AFD_ERR_INVALID, "dx, dy, dz must be > 0");
    if (p.n_resist < 0 || p.n_resist > 2) return fail(AFD_ERR_INVALID, "n_resist must be 0, 1 or 2");
    if (p.gr_metric < 0 || p.gr_metric > 3) return fail(AFD_ERR_INVALID, "gr_metric must be in [0, 3]");
    if (p.gr_metric != 0 && !(p.gr_gamma > 1.0)) return fail(AFD_ERR_INVALID, "gr_gamma must be > 1");
    if (p.face_path < 0 || p.face_path > 4) return fail(AFD_ERR_INVALID, "face_path must be in [0, 4]");

    HIPCHK(hipSetDevice(p.device));
    if (int rc0 = resolve_scratch(p.device)) return rc0;  // afd_c2p_batch_noalloc's per-device staging
    afd_ctx *c = new afd_ctx();
    c->p = p;
    c->nb = nb;
    Grid &g = c->g;
    g.nv = p.nv;
    g.nx = p.nx + 2 * p.ng;
    g.ny = p.ny + 2 * p.ng;
    g.nz = p.nz + 2 * p.ng;
    g.sy = g.nz;
    g.sx = (long)g.ny * g.nz;
    g.ncell = (long)g.nx * g.ny * g.nz;
    g.dnz = make_udiv((unsigned)g.nz);
    g.dny = make_udiv((unsigned)g.ny);
    const int full[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
        // core/srhd_core.py:477-479 [2, n-3); core/rmhd_core.py:751-753 [2, n-2); PPM/WENO [3, n-3)
        if (p.recon_id == AFD_RECON_MUSCL) {
            g.r0[a] = 2;
            g.r1[a] = nb == 9 ? full[a] - 2 : full[a] - 3;
        } else {
            g.r0[a] = 3;
            g.r1[a] = full[a] - 3;
        }
        if (g.r1[a] <= g.r0[a]) {
            delete c;
            return fail(AFD_ERR_INVALID, "grid too small for the reconstruction stencil");
        }
    }
    bool box_empty = false;
    for (int a = 0; a < 3; ++a) {
        c->sbox.lo[a] = std::max(g.r0[a], p.ng);
        c->sbox.hi[a] = std::min(g.r1[a], full[a] - p.ng);
        box_empty |= c->sbox.hi[a] <= c->sbox.lo[a];
    }
    if (box_empty)  // every cell of [1, n-1)^3 then goes to k_max_speed_shell
        for (int a = 0; a < 3; ++a) c->sbox.lo[a] = c->sbox.hi[a] = 1;
    c->carried = nb == 9 && p.recon_id != AFD_RECON_MUSCL;
    c->passive_flux = p.n_passive > 0 && (nb == 9 ? p.recon_id == AFD_RECON_MUSCL : p.recon_id != AFD_RECON_WENO);
    c->fsize = p.precision == AFD_PREC_MIXED ? 4 : 8;
    // the mixed RMHD WENO5 face variants (afd_params.face_path): shared edges below 2^29 cells, with the
    // face fluxes stored (2) or the flux divergence accumulated (3; 4: with the y faces in the update)
    const int fp = p.face_path == 0 ? AFD_FACE_PATH_DEFAULT : p.face_path;
    const bool shared = nb == 9 && p.precision == AFD_PREC_MIXED && p.recon_id == AFD_RECON_WENO && fp >= 2 &&
                        g.ncell < (1L << 29);
    c->shared_edges = shared;
    c->dmode = shared && fp >= 3;
    c->fused = shared && fp == 4;
    if (nb == 5) {  // core/srhd_core.py:393-462: bulk pressure (MUSCL), dissipative fluxes (MUSCL 1x, PPM 2x)
        int reps = p.nv >= 15 ? (p.recon_id == AFD_RECON_MUSCL ? 1 : (p.recon_id == AFD_RECON_PPM ? 2 : 0)) : 0;
        c->diss = (p.recon_id == AFD_RECON_MUSCL && p.passive_offset > 5 ? 1 : 0) | (reps << 1);
    }
    *out = nullptr;
    int rc;
#define TRY(x)                       \
    do {                             \
        if ((rc = (x)) != AFD_OK) {  \
            afd_destroy(c);          \
            return rc;               \
        }                            \
    } while (0)
    size_t N = (size_t)g.ncell;
    for (int b = 0; b < 3; ++b) TRY(dalloc(c, (void **)&c->P[b], N * p.nv * 8));
    if (p.gr_orthonormal && (p.gr_metric == 1 || p.gr_metric == 2)) TRY(dalloc(c, (void **)&c->Puse, N * p.nv * 8));
    TRY(dalloc(c, (void **)&c->Uc, N * nb * 8));
    TRY(dalloc(c, &c->Fx, N * nb * c->fsize));  // the flux-divergence accumulator in D-mode
    if (!c->dmode) {
        TRY(dalloc(c, &c->Fy, N * nb * c->fsize));
        TRY(dalloc(c, &c->Fz, N * nb * c->fsize));
    }
    if (c->carried) {
        if (!c->dmode) TRY(dalloc(c, &c->Fxc, N * nb * c->fsize));
        TRY(dalloc(c, &c->zs, N * 2 * c->fsize));
    }
    if (c->passive_flux) {
        for (int d = 0; d < 3; ++d) TRY(dalloc(c, (void **)&c->aux[d], N * 4 * 8));
        TRY(dalloc(c, (void **)&c->rhs_pass, N * p.n_passive * 8));
    }
    TRY(dalloc(c, (void **)&c->red_out, 8));
    TRY(dalloc(c, (void **)&c->amax_dev, AMAX_BYTES));
    if (hipMemset(c->amax_dev, 0, AMAX_BYTES) != hipSuccess) {  // the shell slots start at zero
        afd_destroy(c);
        return fail(AFD_ERR_HIP, "hipMemset of the speed slots failed");
    }
    TRY(dalloc(c, (void **)&c->cnt_partial, c->red_blocks * 8));
    TRY(dalloc(c, (void **)&c->cnt_out, 8));
    TRY(dalloc(c, (void **)&c->diag_partial, sizeof(double) * DIAG_N * c->red_blocks));
    TRY(dalloc(c, (void **)&c->diag_out, sizeof(double) * DIAG_N));
    TRY(dalloc(c, (void **)&c->stats_dev, sizeof(unsigned long long) * STATS_N * STATS_SLOTS));
    if (nb == 9) TRY(mlp_alloc(c->mlp, [c](void **ptr, size_t bytes) { return dalloc(c, ptr, bytes); }));
    if (nb == 9 && (p.c2p_guess != AFD_C2P_PLAIN || c->dmode)) {
        long blocks = (long)grid1d(c->g.ncell).x;
        c->c2p_qcap = (blocks + QUEUE_SLOTS - 1) / QUEUE_SLOTS * BLOCK;
        if (c->fused) {  // k_face_march_update (up to 64 x MARCH_CELLS cells a workgroup) + k_update_shell
            long fused = (fused_blocks(c->g) + QUEUE_SLOTS - 1) / QUEUE_SLOTS * 64L * MARCH_CELLS;
            long shell = ((long)grid1d(shell_cells(c->g)).x + QUEUE_SLOTS - 1) / QUEUE_SLOTS * BLOCK;
            c->c2p_qcap = std::max(c->c2p_qcap, fused + shell);
        }
        if (p.c2p_guess != AFD_C2P_PLAIN) {
            TRY(dalloc(c, (void **)&c->c2p_queue, sizeof(int) * (size_t)c->c2p_qcap * QUEUE_SLOTS));
            TRY(dalloc(c, (void **)&c->c2p_qcount, sizeof(unsigned) * QUEUE_SLOTS));
        }
        if (c->dmode) {
            TRY(dalloc(c, (void **)&c->repair_queue, sizeof(int) * (size_t)c->c2p_qcap * QUEUE_SLOTS));
            TRY(dalloc(c, (void **)&c->repair_qcount, sizeof(unsigned) * QUEUE_SLOTS));
        }
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        afd_destroy(c);
        return fail(AFD_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    e = hipMemsetAsync(c->stats_dev, 0, sizeof(unsigned long long) * STATS_N * STATS_SLOTS, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        afd_destroy(c);
        return fail(AFD_ERR_HIP, std::string("init: ") + hipGetErrorString(e));
    }
#undef TRY
    *out = c;
    return AFD_OK;
}

extern "C" int afd_destroy(afd_ctx *c) {
    if (!c) return AFD_OK;
    (void)hipSetDevice(c->p.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &m : c->marks) {
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (void *ptr : c->allocs) (void)hipFree(ptr);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return AFD_OK;
}

static int check_n(afd_ctx *c, int64_t n) {
    if (n != (int64_t)c->g.ncell * c->p.nv)
        return fail(AFD_ERR_INVALID, "array size " + std::to_string(n) + " != nv * cells " +
                                          std::to_string((int64_t)c->g.ncell * c->p.nv));
    return AFD_OK;
}

extern "C" int afd_set_prim(afd_ctx *c, const double *host, int64_t n) {
    if (!c || !host) return fail(AFD_ERR_INVALID, "afd_set_prim: null argument");
    if (int rc = check_n(c, n)) return rc;
    HIPCHK(hipSetDevice(c->p.device));
    c->speed_valid = false;
    HIPCHK(hipMemcpyAsync(c->P[c->cur], host, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

extern "C" int afd_get_prim(afd_ctx *c, double *host, int64_t n) {
    if (!c || !host) return fail(AFD_ERR_INVALID, "afd_get_prim: null argument");
    if (int rc = check_n(c, n)) return rc;
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipMemcpyAsync(host, c->P[c->cur], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

// ------------------------------------------------------------------ nozzle
static double (*volatile libm_pow)(double, double) = pow;

// core/nozzle.py:5-25
static double nozzle_weight(double r, double radius, double shear_thick, int profile) {
#pragma clang fp contract(off)
    if (radius <= 0.0) return 0.0;
    if (profile == 1) {
        if (r >= radius) return 0.0;
        return 0.5 * (1.0 + cos(M_PI * r / radius));
    }
    if (profile == 2) {
        if (r >= radius) return 0.0;
        double x = r / radius;
        double w = 1.0 - x * x;
        return (w > 0.0) ? w : 0.0;
    }
    if (shear_thick > 0.0) return 0.5 * (1.0 - tanh((r - radius) / shear_thick));
    return (r <= radius) ? 1.0 : 0.0;
}

extern "C" int afd_set_nozzle(afd_ctx *c, const afd_nozzle *nz) {
#pragma clang fp contract(off)
    if (!c || !nz) return fail(AFD_ERR_INVALID, "afd_set_nozzle: null argument");
    const afd_params &p = c->p;
    if (nz->n_tracers < 0 || nz->n_tracers > 8) return fail(AFD_ERR_INVALID, "n_tracers must be in [0, 8]");
    if (nz->n_tracers > 0 && (nz->tracer_offset < 0 || nz->tracer_offset + nz->n_tracers > p.nv))
        return fail(AFD_ERR_INVALID, "tracer block outside nv");
    const int NY = p.ny, NZ = p.nz, nv = p.nv;
    const long plane = (long)NY * NZ;
    std::vector<double> tmpl((size_t)nv * plane, 0.0);
    std::vector<int> idx((size_t)plane, -1);
    unsigned mask = 0x1f;  // rho, vx, vy, vz, p
    for (int t = 0; t < nz->n_tracers; ++t) mask |= 1u << (nz->tracer_offset + t);
    bool tt = nz->two_temperature && p.nv > nz->thermo_offset + 1;
    if (tt) mask |= (1u << nz->thermo_offset) | (1u << (nz->thermo_offset + 1));
    if (nz->chemistry && (nz->chem_offset < c->nb || nz->chem_offset > 29))
        return fail(AFD_ERR_INVALID, "chemistry offset outside [nb, 29]");
    bool chem = nz->chemistry && p.nv > nz->chem_offset + 2;  // core/nozzle.py:77-82
    if (chem) mask |= 7u << nz->chem_offset;
    if (nv >= 15) mask |= 0x7fe0u;  // core/nozzle.py:60-61: pr[5:15] = 0
    if (c->nb == 9) mask |= 0x1e0u;
    int nact = 0;
    double beta = sqrt(1.0 - 1.0 / (nz->gamma_jet * nz->gamma_jet));
    for (int jj = 0; jj < NY; ++jj) {
        double y = (jj + 0.5) * p.dy;
        for (int kk = 0; kk < NZ; ++kk) {
            double z = (kk + 0.5) * p.dz;
            double rr = sqrt(libm_pow(y - nz->y0, 2.0) + libm_pow(z - nz->z0, 2.0));
            double s = nozzle_weight(rr, nz->jet_radius, nz->shear_thick, nz->profile);
            long o = (long)jj * NZ + kk;
            tmpl[0 * plane + o] = nz->eta_rho * nz->rho_amb * s + nz->rho_amb * (1.0 - s);
            tmpl[1 * plane + o] = beta * s + nz->vx_amb * (1.0 - s);
            tmpl[2 * plane + o] = nz->vy_amb;
            tmpl[3 * plane + o] = nz->vz_amb;
            tmpl[4 * plane + o] = nz->p_eq;
            for (int t = 0; t < nz->n_tracers; ++t) {
                double tn = nz->tracer_nozzle[t], ta = nz->tracer_amb[t];
                tmpl[(size_t)(nz->tracer_offset + t) * plane + o] = ta + (tn - ta) * s;
            }
            if (tt) {
                tmpl[(size_t)nz->thermo_offset * plane + o] = nz->te_nozzle;
                tmpl[(size_t)(nz->thermo_offset + 1) * plane + o] = nz->ti_nozzle;
            }
            if (chem)
                for (int m = 0; m < 3; ++m) tmpl[(size_t)(nz->chem_offset + m) * plane + o] = nz->chem_nozzle[m];
            if (c->nb == 9) {
                double Bx = 0.0, By = 0.0, Bz = 0.0;
                if (nz->b_init == 1) Bx = nz->b0 * s;
                else if (nz->b_init == 2) By = nz->b0 * s;
                tmpl[5 * plane + o] = Bx + nz->b_amb[0];
                tmpl[6 * plane + o] = By + nz->b_amb[1];
                tmpl[7 * plane + o] = Bz + nz->b_amb[2];
                tmpl[8 * plane + o] = 0.0;
            }
            if (s > 0.0) idx[o] = nact++;
        }
    }
    HIPCHK(hipSetDevice(p.device));
    if (!c->noz_tmpl) {
        ALLOC(c, c->noz_tmpl, tmpl.size() * 8);
        ALLOC(c, c->noz_idx, idx.size() * sizeof(int));
        ALLOC(c, c->noz_draws, (size_t)3 * p.ng * plane * 8 + 8);
    }
    HIPCHK(hipMemcpyAsync(c->noz_tmpl, tmpl.data(), tmpl.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->noz_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->noz = *nz;
    c->noz_nact = nact;
    c->noz_mask = mask;
    c->noz_set = true;
    return AFD_OK;
}

extern "C" int afd_nozzle_draw_count(afd_ctx *c, int64_t *n) {
    if (!c || !n) return fail(AFD_ERR_INVALID, "afd_nozzle_draw_count: null argument");
    *n = (c->noz_set && c->noz.perturb) ? (int64_t)3 * c->p.ng * c->noz_nact : 0;
    return AFD_OK;
}

static int fill_yz(afd_ctx *c, int which) {
    BcArgs a{c->P[c->cur], c->g, c->p.ng, 0};
    Timed t(c, 4);
    if (which & 1) {
        a.bc = c->p.bc_y;
        hipLaunchKernelGGL(k_fill_y, grid1d((long)c->g.nv * c->g.nx * 2 * c->p.ng * c->g.nz), dim3(BLOCK), 0,
                           c->stream, a);
    }
    if (which & 2) {
        a.bc = c->p.bc_z;
        hipLaunchKernelGGL(k_fill_z, grid1d((long)c->g.nv * c->g.nx * c->g.ny * 2 * c->p.ng), dim3(BLOCK), 0,
                           c->stream, a);
    }
    HIPCHK(hipGetLastError());
    return AFD_OK;
}

static int outflow_x(afd_ctx *c) {
    BcArgs a{c->P[c->cur], c->g, c->p.ng, 1};
    Timed t(c, 4);
    hipLaunchKernelGGL(k_outflow_x, grid1d((long)c->g.nv * c->p.ng * c->g.ny * c->g.nz), dim3(BLOCK), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return AFD_OK;
}

static int nozzle_fill(afd_ctx *c, const double *draws, int64_t ndraws) {
    if (!c->noz_set) return fail(AFD_ERR_STATE, "nozzle not set (afd_set_nozzle)");
    int64_t need = 0;
    afd_nozzle_draw_count(c, &need);
    if (need > 0) {
        if (!draws || ndraws < need)
            return fail(AFD_ERR_INVALID, "perturbed nozzle needs " + std::to_string(need) + " draws");
        HIPCHK(hipMemcpyAsync(c->noz_draws, draws, (size_t)need * 8, hipMemcpyHostToDevice, c->stream));
    }
    NozzleArgs a;
    a.P = c->P[c->cur];
    a.tmpl = c->noz_tmpl;
    a.idx = c->noz_idx;
    a.draws = c->noz_draws;
    a.g = c->g;
    a.ng = c->p.ng;
    a.ny_int = c->p.ny;
    a.nz_int = c->p.nz;
    a.nact = c->noz_nact;
    a.perturb = need > 0;
    a.mask = c->noz_mask;
    a.vamp = c->noz.turb_vamp;
    a.pamp = c->noz.turb_pamp;
    Timed t(c, 4);
    hipLaunchKernelGGL(k_nozzle, grid1d((long)c->p.ng * c->p.ny * c->p.nz), dim3(BLOCK), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    if (need > 0) HIPCHK(hipStreamSynchronize(c->stream));  // the host draw buffer may be reused
    return AFD_OK;
}

extern "C" int afd_apply_periodic_yz(afd_ctx *c) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    int by = c->p.bc_y, bz = c->p.bc_z;
    c->p.bc_y = c->p.bc_z = AFD_BC_PERIODIC;
    int rc = fill_yz(c, 3);
    c->p.bc_y = by;
    c->p.bc_z = bz;
    return rc;
}

extern "C" int afd_apply_outflow_right_x(afd_ctx *c) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    return outflow_x(c);
}

extern "C" int afd_apply_nozzle_left_x(afd_ctx *c, const double *draws, int64_t ndraws) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    return nozzle_fill(c, draws, ndraws);
}

extern "C" int afd_apply_bcs(afd_ctx *c, const double *draws, int64_t ndraws) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    int rc;
    if ((rc = fill_yz(c, 3))) return rc;
    if ((rc = outflow_x(c))) return rc;
    if (c->noz_set && (rc = nozzle_fill(c, draws, ndraws))) return rc;
    return AFD_OK;
}

// ------------------------------------------------------------------ RHS / step
static FaceLaunch face_launch(afd_ctx *c, const double *P) {
    FaceLaunch l;
    l.P = P;
    l.F = l.Fcar = l.zs = nullptr;
    l.aux = nullptr;
    l.g = c->g;
    l.phd = make_phys(c->p);
    l.phf = to_float(l.phd);
    l.p_max = c->p.p_max;
    l.v_max = c->p.v_max;
    l.lim = c->p.limiter_id;
    l.recon = c->p.recon_id;
    l.riemann = c->p.riemann_id;
    l.dir = 0;
    l.diss = c->diss;
    l.mixed = c->p.precision == AFD_PREC_MIXED;
    l.shared_edges = c->shared_edges;
    l.D = c->dmode ? (float *)c->Fx : nullptr;
    l.dfirst = true;
    l.idn = 0.0f;
    return l;
}

static int run_faces(afd_ctx *c, const double *P, bool skip_y = false) {
    if (c->Puse) {  // GR: the fluxes see the rescaled velocities (core/grhd_core.py:22-45)
        GrRescaleArgs r{P, c->Puse, c->g, c->p.ng, c->p.gr_x0, c->p.gr_mass, c->p.dx, c->p.dy, c->p.dz};
        Timed t(c, 4);
        hipLaunchKernelGGL(k_gr_rescale, grid1d(c->g.ncell), dim3(BLOCK), 0, c->stream, r);
        HIPCHK(hipGetLastError());
        P = c->Puse;
    }
    FaceLaunch l = face_launch(c, P);
    void *Fd[3] = {c->Fx, c->Fy, c->Fz};
    // carried x speeds: z before x.  Flux divergence: the first launch writes D, so y (when
    // launched here) goes first
    const int order_carried[3] = {2, 0, 1}, order_plain[3] = {0, 1, 2}, order_div[3] = {1, 2, 0};
    const int *order = c->dmode && !skip_y ? order_div : (c->carried ? order_carried : order_plain);
    bool first = true;
    for (int n = 0; n < 3; ++n) {
        int d = order[n];
        if (skip_y && d == 1) continue;  // the stage update evaluates the y faces itself
        l.dfirst = first;
        first = false;
        l.dir = d;
        l.idn = (float)(1.0 / (d == 0 ? c->p.dx : (d == 1 ? c->p.dy : c->p.dz)));
        l.F = Fd[d];
        l.Fcar = (c->carried && d == 0) ? c->Fxc : nullptr;
        l.zs = c->carried ? c->zs : nullptr;
        l.aux = c->passive_flux ? c->aux[d] : nullptr;
        Timed t(c, d);
        hipError_t e = (c->nb == 9) ? launch_face_rmhd(l, c->stream) : launch_face_srhd(l, c->stream);
        if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("face kernel: ") + hipGetErrorString(e));
    }
    if (c->passive_flux) {
        PassiveArgs a;
        a.P = P;
        for (int d = 0; d < 3; ++d) a.aux[d] = c->aux[d];
        a.rhs = c->rhs_pass;
        a.g = c->g;
        a.physics = c->p.physics;
        a.recon = c->p.recon_id;
        a.lim = c->p.limiter_id;
        a.np = c->p.n_passive;
        a.po = c->p.passive_offset;
        a.d[0] = c->p.dx;
        a.d[1] = c->p.dy;
        a.d[2] = c->p.dz;
        Timed t(c, 4);
        hipLaunchKernelGGL(k_passive_rhs, grid1d(c->g.ncell), dim3(BLOCK), 0, c->stream, a);
        HIPCHK(hipGetLastError());
    }
    return AFD_OK;
}

static UpdLaunch upd_launch(afd_ctx *c, const double *Pin, const double *P0, double *Pout, int stage, int nst,
                            double dt) {
    UpdLaunch l;
    l.nb = c->nb;
    l.mixed = c->p.precision == AFD_PREC_MIXED;
    UpdArgs<double> &a = l.a;
    a.P0 = P0;
    a.Pin = Pin;
    a.Pout = Pout;
    a.Uc = c->Uc;
    a.Fx = (const double *)c->Fx;
    a.FxL = (const double *)(c->carried ? c->Fxc : c->Fx);
    a.Fy = (const double *)c->Fy;
    a.Fz = (const double *)c->Fz;
    a.D = c->dmode ? (const float *)c->Fx : nullptr;
    a.rhs_pass = c->passive_flux ? c->rhs_pass : nullptr;
    a.stats = c->stats_on ? c->stats_dev : nullptr;
    a.flags = c->stats_on ? c->flags_dev : nullptr;
    a.g = c->g;
    a.stage = stage;
    a.nst = nst;
    a.n_resist = c->nb == 9 ? c->p.n_resist : 0;
    a.np = c->p.n_passive;
    a.po = c->p.passive_offset;
    a.dt = dt;
    a.dx = c->p.dx;
    a.dy = c->p.dy;
    a.dz = c->p.dz;
    a.glm_cp2 = c->p.glm_cp * c->p.glm_cp;
    a.resistivity = c->p.resistivity;
    a.gamma = c->p.gamma;
    a.gfac = c->p.gamma / (c->p.gamma - 1.0);
    a.p_max = c->p.p_max;
    a.v_max = c->p.v_max;
    a.mlp = c->p.c2p_guess == AFD_C2P_LEARNED ? c->mlp.w : nullptr;
    a.guess = c->p.c2p_guess;
    a.queue = c->p.c2p_guess != AFD_C2P_PLAIN ? c->c2p_queue : nullptr;
    a.qcount = c->p.c2p_guess != AFD_C2P_PLAIN ? c->c2p_qcount : nullptr;
    a.rq = c->repair_queue;
    a.rqcount = c->repair_qcount;
    a.ng = c->p.ng;
    a.gr = c->p.gr_metric == 0 ? 0 : (c->p.gr_metric == 1 ? 1 : 2);
    a.gr_x0 = c->p.gr_x0;
    a.amax = nullptr;
    a.sbox = c->sbox;
    a.gr_mass = c->p.gr_mass;
    a.gr_gfac = c->p.gr_metric == 0 ? 0.0 : c->p.gr_gamma / (c->p.gr_gamma - 1.0);
    a.qcap = c->c2p_qcap;
    return l;
}

extern "C" int afd_step(afd_ctx *c, double dt) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    if (!(dt >= 0.0) || !std::isfinite(dt)) return fail(AFD_ERR_INVALID, "dt must be finite and >= 0");
    if (c->p.c2p_guess == AFD_C2P_LEARNED && !c->mlp.loaded)
        return fail(AFD_ERR_STATE, "learned guess selected but no weights loaded (afd_mlp_set_weights)");
    HIPCHK(hipSetDevice(c->p.device));
    const int nst = c->p.rk_order == 3 ? 3 : 2;
    const int in = c->cur, a = (in + 1) % 3, b = (in + 2) % 3;
    // SSPRK2: in -> a -> b ; SSPRK3: in -> a -> b -> a.  P[in] stays the step input.
    int src = in;
    c->speed_valid = false;
    for (int s = 0; s < nst; ++s) {
        int dst = (s % 2 == 0) ? a : b;
        int rc = run_faces(c, c->P[src], c->fused);
        if (rc) return rc;
        UpdLaunch l = upd_launch(c, c->P[src], c->P[in], c->P[dst], s, nst, dt);
        if (c->nb == 5 && s == nst - 1) {  // SRHD: the next step's CFL speed from the last stage
            HIPCHK(hipMemsetAsync(c->amax_dev, 0, AMAX_BOX_BYTES, c->stream));
            l.a.amax = c->amax_dev;
        }
        Timed t(c, 3);
        hipError_t e;
        if (c->repair_qcount)
            HIPCHK(hipMemsetAsync(c->repair_qcount, 0, sizeof(unsigned) * QUEUE_SLOTS, c->stream));
        if (c->fused) {  // the y faces with the stage update
            FaceLaunch fl = face_launch(c, c->Puse ? c->Puse : c->P[src]);
            fl.dir = 1;
            fl.idn = (float)(1.0 / c->p.dy);
            fl.zs = c->zs;
            e = launch_stage_fused(fl, l, c->stream);
        } else {
            e = launch_update(l, false, c->stream);
        }
        if (e == hipSuccess && c->repair_queue)  // cells whose FP32 right-hand side was not finite
            e = launch_repair(face_launch(c, c->Puse ? c->Puse : c->P[src]), l, c->stream);
        if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("update kernel: ") + hipGetErrorString(e));
        src = dst;
    }
    c->cur = src;
    c->speed_valid = c->nb == 5;
    c->acc.c2p_cells += (uint64_t)nst * c->g.ncell;
    return AFD_OK;
}

extern "C" int afd_compute_rhs(afd_ctx *c, double *host, int64_t n) {
    if (!c || !host) return fail(AFD_ERR_INVALID, "afd_compute_rhs: null argument");
    int nrhs = c->nb + c->p.n_passive;
    if (n != (int64_t)nrhs * c->g.ncell) return fail(AFD_ERR_INVALID, "rhs size must be (nb + n_passive) * cells");
    HIPCHK(hipSetDevice(c->p.device));
    if (!c->rhs_out) ALLOC(c, c->rhs_out, (size_t)nrhs * c->g.ncell * 8);
    int rc = run_faces(c, c->P[c->cur]);
    if (rc) return rc;
    UpdLaunch l = upd_launch(c, c->P[c->cur], c->P[c->cur], c->rhs_out, 0, 2, 0.0);
    hipError_t e = launch_update(l, true, c->stream);
    if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("rhs kernel: ") + hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync(host, c->rhs_out, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

// One pass of k_sources over the interior (the kernel applies the SRC_* bits in the driver's order).
static int run_sources(afd_ctx *c, SourceArgs &a) {
    if (!a.which) return AFD_OK;
    c->speed_valid = false;
    const int ng = c->p.ng;
    a.P = c->P[c->cur];
    a.g = c->g;
    a.ng = ng;
    HIPCHK(hipSetDevice(c->p.device));
    dim3 grid((unsigned)((long)c->p.nx * c->p.ny), (unsigned)((c->p.nz + BLOCK - 1) / BLOCK));
    Timed t(c, 4);
    hipLaunchKernelGGL(k_sources, grid, dim3(BLOCK), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return AFD_OK;
}

static bool two_temp_on(const afd_ctx *c) {  // core/source_terms.py:31-39
    return c->p.two_temperature && c->p.nv >= c->p.thermo_offset + 2 && c->p.tei_tau > 0.0;
}

static int check_chem(const afd_ctx *c, const afd_chem *ch, bool *on) {
    if (ch->offset < c->nb) return fail(AFD_ERR_INVALID, "chemistry offset overlaps the base variables");
    *on = c->p.nv > ch->offset + 2;  // core/chemistry.py:55-57
    return AFD_OK;
}

static void set_chem(SourceArgs &a, const afd_chem *ch) {
    a.c_off = ch->offset;
    a.X = ch->X;
    a.Y = ch->Y;
    a.tunit = ch->tunit_k;
    a.rate_scale = ch->rate_scale;
    a.energy_scale = ch->energy_scale;
    a.tmin = ch->tmin_k;
    a.gamma = ch->gamma;
}

extern "C" int afd_two_temperature(afd_ctx *c, double dt) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    SourceArgs a{};
    a.dt = dt;
    if (two_temp_on(c)) {
        a.which = SRC_TWO_TEMP;
        a.t_off = c->p.thermo_offset;
        a.tau = c->p.tei_tau;
    }
    return run_sources(c, a);
}

extern "C" int afd_cooling_heating(afd_ctx *c, double dt, double cool, double heat, double p_cap) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    SourceArgs a{};
    a.dt = dt;
    if (cool != 0.0 || heat != 0.0) {  // core/source_terms.py:10-11
        a.which = SRC_COOL;
        a.cool = cool;
        a.heat = heat;
        a.p_cap = p_cap;
    }
    return run_sources(c, a);
}

extern "C" int afd_ion_chemistry(afd_ctx *c, double dt, const afd_chem *chem) {
    if (!c || !chem) return fail(AFD_ERR_INVALID, "afd_ion_chemistry: null argument");
    bool on = false;
    int rc = check_chem(c, chem, &on);
    if (rc) return rc;
    SourceArgs a{};
    a.dt = dt;
    if (on) {
        a.which = SRC_CHEM;
        set_chem(a, chem);
    }
    return run_sources(c, a);
}

// core/dissipation.py:10-27: a no-op without terms, with a non-positive relaxation time, for
// RMHD or with fewer than 15 variables.
static bool diss_on(const afd_ctx *c, const afd_diss *d) {
    if (d->zeta == 0.0 && d->eta == 0.0 && d->kappa == 0.0) return false;
    if (std::min(std::min(d->tau_bulk, d->tau_shear), d->tau_heat) <= 0.0) return false;
    return c->nb == 5 && c->p.nv >= 15;
}

static int run_diss(afd_ctx *c, double dt, const afd_diss *d) {
    c->speed_valid = false;
    if (!diss_on(c, d)) return AFD_OK;
    HIPCHK(hipSetDevice(c->p.device));
    // the new fields are staged in variable planes 5..14 of a primitive buffer the step is not using
    double *S = c->P[(c->cur + 1) % 3] + 5 * c->g.ncell;
    DissArgs a{c->P[c->cur], S, c->g, c->p.ng, d->advect ? 1 : 0, dt, c->p.dx, c->p.dy, c->p.dz,
               d->zeta, d->eta, d->kappa, d->tau_bulk, d->tau_shear, d->tau_heat, d->cap_frac * d->p_max};
    dim3 grid((unsigned)((long)c->p.nx * c->p.ny), (unsigned)((c->p.nz + BLOCK - 1) / BLOCK));
    Timed t(c, 4);
    hipLaunchKernelGGL(k_diss, grid, dim3(BLOCK), 0, c->stream, a);
    hipLaunchKernelGGL(k_diss_store, grid, dim3(BLOCK), 0, c->stream, a, c->P[c->cur]);
    HIPCHK(hipGetLastError());
    return AFD_OK;
}

extern "C" int afd_causal_dissipation(afd_ctx *c, double dt, const afd_diss *diss) {
    if (!c || !diss) return fail(AFD_ERR_INVALID, "afd_causal_dissipation: null argument");
    return run_diss(c, dt, diss);
}

extern "C" int afd_post_step_sources(afd_ctx *c, double dt, const afd_sources *src) {
    if (!c || !src) return fail(AFD_ERR_INVALID, "afd_post_step_sources: null argument");
    if (src->dissipation) {  // solvers/srhd3d_mpi_muscl.py:323, before the pointwise sources
        int rc = run_diss(c, dt, &src->diss);
        if (rc) return rc;
    }
    SourceArgs a{};
    a.dt = dt;
    if (src->cooling && (src->cool != 0.0 || src->heat != 0.0)) {
        a.which |= SRC_COOL;
        a.cool = src->cool;
        a.heat = src->heat;
        a.p_cap = src->p_cap;
    }
    if (two_temp_on(c)) {
        a.which |= SRC_TWO_TEMP;
        a.t_off = c->p.thermo_offset;
        a.tau = c->p.tei_tau;
    }
    if (src->chemistry) {
        bool on = false;
        int rc = check_chem(c, &src->chem, &on);
        if (rc) return rc;
        if (on) {
            a.which |= SRC_CHEM;
            set_chem(a, &src->chem);
        }
    }
    return run_sources(c, a);
}

extern "C" int afd_max_char_speed(afd_ctx *c, double *amax) {
    if (!c || !amax) return fail(AFD_ERR_INVALID, "afd_max_char_speed: null argument");
    HIPCHK(hipSetDevice(c->p.device));
    const double gfac = c->p.gamma / (c->p.gamma - 1.0);
    const Grid &g = c->g;
    const SpeedBox &b = c->sbox;
    const long range = (long)(g.nx - 2) * (g.ny - 2) * (g.nz - 2);
    const long inner = (long)(b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]) * (b.hi[2] - b.lo[2]);
    {
        Timed t(c, 4);
        if (!c->speed_valid) {  // no step's last stage took the box's maximum of this state
            HIPCHK(hipMemsetAsync(c->amax_dev, 0, AMAX_BOX_BYTES, c->stream));
            if (inner > 0)
                hipLaunchKernelGGL(k_max_speed_box, dim3(c->red_blocks), dim3(BLOCK), 0, c->stream,
                                   (const double *)c->P[c->cur], g, b, c->p.gamma, gfac, c->amax_dev);
            HIPCHK(hipGetLastError());
            c->speed_valid = true;
        }
        // the rest of the range as the fills left it, then the floor of the reference's maximum
        if (range > inner)
            hipLaunchKernelGGL(k_max_speed_shell, grid1d(range - inner), dim3(BLOCK), 0, c->stream,
                               (const double *)c->P[c->cur], g, b, c->p.gamma, gfac,
                               c->amax_dev + AMAX_SLOTS * AMAX_STRIDE);
        hipLaunchKernelGGL(k_amax_finish, dim3(1), dim3(BLOCK), 0, c->stream, c->amax_dev, SMALL, c->red_out);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(amax, c->red_out, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

extern "C" int afd_count_nonfinite(afd_ctx *c, int64_t *count) {
    if (!c || !count) return fail(AFD_ERR_INVALID, "afd_count_nonfinite: null argument");
    HIPCHK(hipSetDevice(c->p.device));
    long n = c->g.ncell * c->p.nv;
    hipLaunchKernelGGL(k_count_nonfinite, dim3(c->red_blocks), dim3(BLOCK), 0, c->stream,
                       (const double *)c->P[c->cur], n, c->cnt_partial);
    hipLaunchKernelGGL(k_reduce_sum_u64, dim3(1), dim3(BLOCK), 0, c->stream,
                       (const unsigned long long *)c->cnt_partial, c->red_blocks, c->cnt_out);
    HIPCHK(hipGetLastError());
    unsigned long long v = 0;
    HIPCHK(hipMemcpyAsync(&v, c->cnt_out, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *count = (int64_t)v;
    return AFD_OK;
}

static int batch(afd_ctx *c, int64_t n, const double *in, double *out, int32_t *iters, bool c2p) {
    if (!c || (n > 0 && (!in || !out))) return fail(AFD_ERR_INVALID, "batch: null argument");
    if (n < 0) return fail(AFD_ERR_INVALID, "batch: n < 0");
    if (n == 0) return AFD_OK;
    HIPCHK(hipSetDevice(c->p.device));
    size_t bytes = (size_t)n * c->nb * 8;
    double *din = nullptr, *dout = nullptr;
    int *dit = nullptr;
    HIPCHK(hipMalloc(&din, bytes));
    hipError_t e = hipMalloc(&dout, bytes);
    if (e == hipSuccess && iters) e = hipMalloc(&dit, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(din, in, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        BatchArgs a{din,           dout,        dit,        c2p && c->stats_on ? c->stats_dev : nullptr,
                    (long)n,       c->p.physics, c->p.gamma, c->p.gamma / (c->p.gamma - 1.0),
                    c->p.p_max,    c->p.v_max};
        if (c2p && c->stats_on) c->acc.c2p_cells += (uint64_t)n;
        if (c2p) {
            C2PBatchExtra x{c->p.c2p_guess == AFD_C2P_LEARNED && c->mlp.loaded ? c->mlp.w : nullptr};
            launch_c2p_batch(a, x, c->stream);
        } else {
            hipLaunchKernelGGL(k_p2c_batch, grid1d(n), dim3(BLOCK), 0, c->stream, a);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && iters) e = hipMemcpyAsync(iters, dit, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(din);
    (void)hipFree(dout);
    if (dit) (void)hipFree(dit);
    if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
    return AFD_OK;
}

extern "C" int afd_p2c_batch(afd_ctx *c, int64_t n, const double *prim, double *cons) {
    return batch(c, n, prim, cons, nullptr, false);
}
extern "C" int afd_c2p_batch(afd_ctx *c, int64_t n, const double *cons, double *prim, int32_t *iters) {
    return batch(c, n, cons, prim, iters, true);
}

// ------------------------------------------------------------------ allocation-free con2prim batch
// Device staging of afd_c2p_batch_noalloc: the states in, the states out and the iteration
// counts, AFD_NOALLOC_SCRATCH_BYTES per device for the life of the process.  Host code never
// uses this name as an address: the per-device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[AFD_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device (set-up of afd_create, once per device).
static int resolve_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(AFD_ERR_INVALID, "device ordinal outside [0, 64) for the allocation-free entry point");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return AFD_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(AFD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return AFD_OK;
}

static void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

extern "C" int afd_c2p_batch_noalloc(afd_ctx *c, int64_t n, const double *cons, double *prim, int32_t *iters) {
    if (!c || (n > 0 && (!cons || !prim))) return fail(AFD_ERR_INVALID, "afd_c2p_batch_noalloc: null argument");
    if (n < 0) return fail(AFD_ERR_INVALID, "afd_c2p_batch_noalloc: n < 0");
    const int64_t nmax = AFD_C2P_NOALLOC_MAX(c->nb);
    if (n > nmax)
        return fail(AFD_ERR_INVALID, "afd_c2p_batch_noalloc: n = " + std::to_string(n) + " exceeds the " +
                                         std::to_string(nmax) + " states the " +
                                         std::to_string(AFD_NOALLOC_SCRATCH_BYTES) + "-byte device scratch holds");
    if (n == 0) return AFD_OK;
    HIPCHK(hipSetDevice(c->p.device));
    if (int rc = resolve_scratch(c->p.device)) return rc;  // set up by afd_create; kept for safety
    char *base = (char *)scratch_ptr(c->p.device);
    const size_t bytes = (size_t)n * c->nb * 8;
    double *din = (double *)base, *dout = (double *)(base + bytes);
    int *dit = iters ? (int *)(base + 2 * bytes) : nullptr;
    HIPCHK(hipMemcpyAsync(din, cons, bytes, hipMemcpyHostToDevice, c->stream));
    BatchArgs a{din,           dout,         dit,        c->stats_on ? c->stats_dev : nullptr,
                (long)n,       c->p.physics, c->p.gamma, c->p.gamma / (c->p.gamma - 1.0),
                c->p.p_max,    c->p.v_max};
    if (c->stats_on) c->acc.c2p_cells += (uint64_t)n;
    C2PBatchExtra x{c->p.c2p_guess == AFD_C2P_LEARNED && c->mlp.loaded ? c->mlp.w : nullptr};
    launch_c2p_batch(a, x, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(prim, dout, bytes, hipMemcpyDeviceToHost, c->stream));
    if (iters) HIPCHK(hipMemcpyAsync(iters, dit, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

// ------------------------------------------------------------------ diagnostics
extern "C" int afd_diagnostics(afd_ctx *c, afd_diag *out) {
    if (!c || !out) return fail(AFD_ERR_INVALID, "afd_diagnostics: null argument");
    HIPCHK(hipSetDevice(c->p.device));
    DiagArgs a{c->P[c->cur], c->diag_partial, c->g, c->p.ng, c->nb == 9 ? 1 : 0, c->p.dx, c->p.dy, c->p.dz,
               (5.0 / 3.0) / (5.0 / 3.0 - 1.0), c->p.p_max, c->p.v_max};
    {
        Timed t(c, 4);
        hipLaunchKernelGGL(k_diag, dim3(c->red_blocks), dim3(BLOCK), 0, c->stream, a);
        hipLaunchKernelGGL(k_diag_reduce, dim3(1), dim3(BLOCK), 0, c->stream, (const double *)c->diag_partial,
                           c->red_blocks, c->diag_out);
    }
    HIPCHK(hipGetLastError());
    double v[DIAG_N];
    HIPCHK(hipMemcpyAsync(v, c->diag_out, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double ncell = (double)c->p.nx * c->p.ny * c->p.nz;
    out->max_gamma = v[0];
    out->inlet_flux = v[4];
    out->divb_max = c->nb == 9 ? v[1] : 0.0;
    out->divb_mean = c->nb == 9 ? v[2] / ncell : 0.0;
    out->divb_rms = c->nb == 9 ? sqrt(v[3] / ncell) : 0.0;
    return AFD_OK;
}

extern "C" int afd_centerline(afd_ctx *c, double *out, int64_t n) {
    if (!c || !out) return fail(AFD_ERR_INVALID, "afd_centerline: null argument");
    if (n != 4 * (int64_t)c->p.nx) return fail(AFD_ERR_INVALID, "afd_centerline: n must be 4 * nx");
    HIPCHK(hipSetDevice(c->p.device));
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * n));
    hipLaunchKernelGGL(k_centerline, grid1d(c->p.nx), dim3(BLOCK), 0, c->stream, (const double *)c->P[c->cur], c->g,
                       c->p.ng, c->p.v_max, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("centerline: ") + hipGetErrorString(e));
    return AFD_OK;
}

// ------------------------------------------------------------------ instrumentation
extern "C" int afd_set_instrumentation(afd_ctx *c, int32_t stats, int32_t timing) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    if (stats && !c->flags_dev) {
        HIPCHK(hipSetDevice(c->p.device));
        ALLOC(c, c->flags_dev, (size_t)c->g.ncell);
        HIPCHK(hipMemsetAsync(c->flags_dev, 0, (size_t)c->g.ncell, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->stats_on = stats != 0;
    c->timing_on = timing != 0;
    return AFD_OK;
}

extern "C" int afd_get_stats(afd_ctx *c, afd_stats *out) {
    if (!c || !out) return fail(AFD_ERR_INVALID, "afd_get_stats: null argument");
    HIPCHK(hipSetDevice(c->p.device));
    if (int rc = resolve_marks(c)) return rc;
    unsigned long long slots[STATS_N * STATS_SLOTS], v[STATS_N] = {0};
    HIPCHK(hipMemcpyAsync(slots, c->stats_dev, sizeof(slots), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int s = 0; s < STATS_SLOTS; ++s)
        for (int m = 0; m < STATS_N; ++m) v[m] += slots[STATS_N * s + m];
    *out = c->acc;
    out->hydro_iters = v[0];
    out->newton_evals = v[1];
    out->hydro_fallbacks = v[2];
    out->bisections = v[3];
    out->invalid = v[4];
    out->learned_ok = v[5];
    return AFD_OK;
}

extern "C" int afd_get_c2p_flags(afd_ctx *c, uint8_t *out, int64_t n) {
    if (!c || !out) return fail(AFD_ERR_INVALID, "afd_get_c2p_flags: null argument");
    if (n != (int64_t)c->g.ncell) return fail(AFD_ERR_INVALID, "afd_get_c2p_flags: n must be the number of cells");
    if (!c->flags_dev) return fail(AFD_ERR_STATE, "afd_get_c2p_flags: no step has run with the counters on");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipMemcpyAsync(out, c->flags_dev, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

extern "C" int afd_reset_stats(afd_ctx *c) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    if (int rc = resolve_marks(c)) return rc;
    HIPCHK(hipMemsetAsync(c->stats_dev, 0, sizeof(unsigned long long) * STATS_N * STATS_SLOTS, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->acc = afd_stats{};
    return AFD_OK;
}

extern "C" int afd_synchronize(afd_ctx *c) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

// ------------------------------------------------------------------ learned guess weights
extern "C" int afd_mlp_param_count(int32_t *count) {
    if (!count) return fail(AFD_ERR_INVALID, "null argument");
    *count = MLP_NPARAM;
    return AFD_OK;
}

extern "C" int afd_mlp_set_weights(afd_ctx *c, const float *w, int32_t count) {
    if (!c || !w) return fail(AFD_ERR_INVALID, "afd_mlp_set_weights: null argument");
    if (c->nb != 9) return fail(AFD_ERR_STATE, "the learned guess is RMHD only");
    if (count != MLP_NPARAM) return fail(AFD_ERR_INVALID, "expected " + std::to_string(MLP_NPARAM) + " weights");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(mlp_upload(c->mlp, w, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mlp.loaded = true;  // the host copy may be released: the upload has completed
    return AFD_OK;
}

extern "C" int afd_mlp_get_weights(afd_ctx *c, float *w, int32_t count) {
    if (!c || !w) return fail(AFD_ERR_INVALID, "afd_mlp_get_weights: null argument");
    if (c->nb != 9 || !c->mlp.loaded) return fail(AFD_ERR_STATE, "no weights loaded");
    if (count != MLP_NPARAM) return fail(AFD_ERR_INVALID, "expected " + std::to_string(MLP_NPARAM) + " weights");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipMemcpyAsync(w, c->mlp.master, MLP_NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

// ------------------------------------------------------------------ learned guess: data, training, checks
static int need_mlp(afd_ctx *c) {
    if (!c) return fail(AFD_ERR_INVALID, "null context");
    if (c->nb != 9) return fail(AFD_ERR_STATE, "the learned guess is RMHD only");
    return AFD_OK;
}

template <typename T> static int tmp_alloc(T **p, size_t bytes) {
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(AFD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return AFD_OK;
}

extern "C" int afd_mlp_dataset_from_prims(afd_ctx *c, int64_t n, const double *prim, float *X, float *Y,
                                          int32_t *valid) {
    if (int rc = need_mlp(c)) return rc;
    if (n <= 0 || !prim || !X || !Y || !valid) return fail(AFD_ERR_INVALID, "afd_mlp_dataset_from_prims: bad argument");
    HIPCHK(hipSetDevice(c->p.device));
    const int64_t chunk = 1 << 22;
    double *dp = nullptr;
    float *dX = nullptr, *dY = nullptr;
    int *dv = nullptr;
    int64_t m = n < chunk ? n : chunk;
    int rc = tmp_alloc(&dp, (size_t)m * 72);
    if (!rc) rc = tmp_alloc(&dX, (size_t)m * 20);
    if (!rc) rc = tmp_alloc(&dY, (size_t)m * 8);
    if (!rc) rc = tmp_alloc(&dv, (size_t)m * 4);
    hipError_t e = hipSuccess;
    for (int64_t off = 0; !rc && e == hipSuccess && off < n; off += chunk) {
        int64_t k = (n - off) < chunk ? (n - off) : chunk;
        e = hipMemcpyAsync(dp, prim + 9 * off, (size_t)k * 72, hipMemcpyHostToDevice, c->stream);
        MlpDataArgs a{dp, dX, dY, dv, (long)k, c->p.gamma, c->p.gamma / (c->p.gamma - 1.0), c->p.p_max, c->p.v_max};
        if (e == hipSuccess) e = mlp_prepare_dataset(a, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(X + 5 * off, dX, (size_t)k * 20, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(Y + 2 * off, dY, (size_t)k * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(valid + off, dv, (size_t)k * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(dp);
    (void)hipFree(dX);
    (void)hipFree(dY);
    (void)hipFree(dv);
    if (rc) return rc;
    if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("dataset: ") + hipGetErrorString(e));
    return AFD_OK;
}

static void forget_alloc(afd_ctx *c, void *ptr) {
    for (auto it = c->allocs.begin(); it != c->allocs.end(); ++it)
        if (*it == ptr) {
            c->allocs.erase(it);
            break;
        }
    (void)hipFree(ptr);
}

extern "C" int afd_mlp_set_dataset(afd_ctx *c, int64_t n, const float *X, const float *Y) {
    if (int rc = need_mlp(c)) return rc;
    if (n <= 0 || !X || !Y) return fail(AFD_ERR_INVALID, "afd_mlp_set_dataset: bad argument");
    HIPCHK(hipSetDevice(c->p.device));
    MlpDevice &m = c->mlp;
    if (n > m.cap_data) {
        if (m.X) forget_alloc(c, m.X);
        if (m.Y) forget_alloc(c, m.Y);
        m.X = m.Y = nullptr;
        m.cap_data = 0;
        ALLOC(c, m.X, (size_t)n * 20);
        ALLOC(c, m.Y, (size_t)n * 8);
        m.cap_data = n;
    }
    HIPCHK(hipMemcpyAsync(m.X, X, (size_t)n * 20, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(m.Y, Y, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    m.n_data = n;
    return AFD_OK;
}

extern "C" int afd_mlp_train(afd_ctx *c, int64_t steps, int32_t batch, float lr, double *loss_history) {
    if (int rc = need_mlp(c)) return rc;
    MlpDevice &m = c->mlp;
    if (!m.loaded) return fail(AFD_ERR_STATE, "initial weights not set (afd_mlp_set_weights)");
    if (m.n_data <= 0) return fail(AFD_ERR_STATE, "no dataset (afd_mlp_set_dataset)");
    if (batch <= 0 || batch > m.n_data) return fail(AFD_ERR_INVALID, "batch must be in [1, dataset size]");
    if (steps < 0 || steps > 65536) return fail(AFD_ERR_INVALID, "steps must be in [0, 65536] per call");
    HIPCHK(hipSetDevice(c->p.device));
    const int64_t nbat = m.n_data / batch;
    for (int64_t st = 0; st < steps; ++st) {
        int64_t off = (st % nbat) * batch;
        HIPCHK(mlp_loss_grad(m.w, m.X + 5 * off, m.Y + 2 * off, batch, m.grad, m.loss + st, c->stream));
        HIPCHK(mlp_adam(m.master, m.adam_m, m.adam_v, m.grad, ++m.adam_t, lr, c->stream));
        HIPCHK(mlp_pack(m.master, m.w, c->stream));
    }
    if (loss_history && steps > 0)
        HIPCHK(hipMemcpyAsync(loss_history, m.loss, (size_t)steps * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AFD_OK;
}

extern "C" int afd_mlp_forward(afd_ctx *c, int64_t n, const float *X, float *Y) {
    if (int rc = need_mlp(c)) return rc;
    if (!c->mlp.loaded) return fail(AFD_ERR_STATE, "no weights loaded");
    if (n <= 0 || !X || !Y) return fail(AFD_ERR_INVALID, "afd_mlp_forward: bad argument");
    HIPCHK(hipSetDevice(c->p.device));
    float *dX = nullptr, *dY = nullptr;
    int rc = tmp_alloc(&dX, (size_t)n * 20);
    if (!rc) rc = tmp_alloc(&dY, (size_t)n * 8);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpyAsync(dX, X, (size_t)n * 20, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            Timed tm(c, 4);  // the network alone, in ms_other when timing is on
            e = mlp_forward(c->mlp.w, dX, n, dY, c->stream);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(Y, dY, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (rc) return rc;
    if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("mlp forward: ") + hipGetErrorString(e));
    return AFD_OK;
}

extern "C" int afd_mlp_loss_grad(afd_ctx *c, int64_t n, const float *X, const float *Y, float *grad, double *loss) {
    if (int rc = need_mlp(c)) return rc;
    if (!c->mlp.loaded) return fail(AFD_ERR_STATE, "no weights loaded");
    if (n <= 0 || !X || !Y || !grad || !loss) return fail(AFD_ERR_INVALID, "afd_mlp_loss_grad: bad argument");
    HIPCHK(hipSetDevice(c->p.device));
    float *dX = nullptr, *dY = nullptr;
    int rc = tmp_alloc(&dX, (size_t)n * 20);
    if (!rc) rc = tmp_alloc(&dY, (size_t)n * 8);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpyAsync(dX, X, (size_t)n * 20, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dY, Y, (size_t)n * 8, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = mlp_loss_grad(c->mlp.w, dX, dY, n, c->mlp.grad, c->mlp.loss, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(grad, c->mlp.grad, MLP_NPARAM * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(loss, c->mlp.loss, 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (rc) return rc;
    if (e != hipSuccess) return fail(AFD_ERR_HIP, std::string("mlp loss/grad: ") + hipGetErrorString(e));
    return AFD_OK;
}
