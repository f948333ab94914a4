/*
 * afd.h -- C-ABI of the MI355X (gfx950) SRHD/RMHD finite-volume step.
 *
 * The library libafd.so replaces the hot path of AstroFD: the SRHD and RMHD
 * cores (core/srhd_core.py, core/rmhd_core.py), the ghost-zone fills
 * (core/boundary.py), the nozzle inflow (core/nozzle.py) and the
 * two-temperature source (core/source_terms.py:30-55).  The reference is a
 * Python module-function API; a maintainer binds this ABI with ctypes (the
 * binding shipped in astrofd_amd/_lib.py).
 *
 * Conventions
 *   - Every entry point returns int: 0 on success, < 0 on error (AFD_ERR_*);
 *     afd_last_error() gives the message of the calling thread's last error.
 *     Numerical trouble is clamped as in the reference, never an error.
 *   - State arrays are float64 in the reference layout: C-order
 *     (nv, nx + 2 ng, ny + 2 ng, nz + 2 ng), z fastest, ghost zones included.
 *     The caller owns host arrays; device memory belongs to the context.
 *   - One context = one device + one HIP stream (default priority).  Calls on
 *     one context are not reentrant.
 */
#ifndef AFD_H
#define AFD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFD_OK 0
#define AFD_ERR_INVALID (-1) /* bad argument or size */
#define AFD_ERR_NOMEM (-2)   /* device allocation failed */
#define AFD_ERR_HIP (-3)     /* HIP runtime error */
#define AFD_ERR_STATE (-4)   /* call not valid in the context's state */

#define AFD_PHYS_HYDRO 0
#define AFD_PHYS_RMHD 1
#define AFD_RECON_MUSCL 0
#define AFD_RECON_PPM 1
#define AFD_RECON_WENO 2
#define AFD_PREC_FP64 0  /* everything in FP64 */
#define AFD_PREC_MIXED 1 /* reconstruction + Riemann in FP32; Lorentz factor, con2prim, SSPRK in FP64 */
#define AFD_C2P_PLAIN 0
#define AFD_C2P_LEARNED 1 /* MLP initial guess for the RMHD Newton solve; the plain solve checks it and falls back */
#define AFD_C2P_PREVIOUS 2 /* Newton start from the stage input's primitives (Z = rho h W^2); same acceptance test and fallback */
#define AFD_BC_PERIODIC 0
#define AFD_BC_OUTFLOW 1

/* Everything the reference's configure() reads (core/srhd_core.py:18-41,
 * core/rmhd_core.py:21-50) plus the grid.  Limiter/recon/Riemann IDs follow the
 * reference: LIMITER mc=0 minmod=1 vanleer=2; SRHD RIEMANN hlle=0 hllc=1;
 * RMHD RIEMANN hlle=0 hlld=1 (HLLC hydro + HLLE B) hlld_full=2. */
typedef struct afd_params {
    int32_t physics;          /* AFD_PHYS_* */
    int32_t nv;               /* variables incl. passives: 5 or 9, + n_passive */
    int32_t nx, ny, nz;       /* interior cells */
    int32_t ng;               /* ghost layers (>= 2) */
    double dx, dy, dz;
    double gamma;             /* SRHD computes with the 5/3 bound at definition time in the reference */
    double p_max, v_max;
    double glm_ch, glm_cp;
    int32_t limiter_id, recon_id, riemann_id, rk_order;
    int32_t n_passive, passive_offset; /* nv = passive_offset + n_passive; SRHD fields in [5, passive_offset)
                                          (the dissipative 5..14) pass through the step unchanged */
    int32_t n_resist;         /* applications of the resistive Laplacian per RHS (reference: 1 MUSCL, 2 PPM, 0 WENO) */
    double resistivity;
    int32_t precision;        /* AFD_PREC_* */
    int32_t c2p_guess;        /* AFD_C2P_* (RMHD only) */
    int32_t bc_y, bc_z;       /* AFD_BC_*; the reference is periodic in y and z */
    int32_t two_temperature, thermo_offset;
    double tei_tau;
    int32_t device;           /* HIP device ordinal */
    /* GRHD / GRMHD on a fixed metric (core/grhd_core.py, core/grmhd_core.py; PHYSICS "grhd" or
     * "grmhd" runs the SRHD or RMHD step with these set; zeros = flat space, the reference's
     * "hydro" / "rmhd"). */
    int32_t gr_metric;        /* 0 minkowski, 1 schwarzschild, 2 kerr-schild, 3 any other name */
    int32_t gr_orthonormal;   /* ORTHONORMAL_FLUX: fluxes see v / alpha_KS (metrics 1 and 2) */
    int32_t gr_x0;            /* global x index of the first interior cell (0 on one GPU) */
    double gr_mass;           /* GR_MASS */
    double gr_gamma;          /* Gamma of the sources' enthalpy (GRHD: the configured GAMMA) */
    int32_t face_path;        /* mixed RMHD WENO5 face kernels (no counterpart in the reference; all give the
                                 same results to FP32 rounding): 0 automatic; 1 one thread per face;
                                 2 cells share their edge values (LDS rows, lane shuffles), face fluxes
                                 stored; 3 as 2, the faces accumulate each cell's flux divergence instead;
                                 4 as 3, with the y faces evaluated by the stage update.  2-4 apply below
                                 2^29 cells, larger grids take 1 */
} afd_params;

/* The variant face_path 0 selects (the fastest measured on the 256^3 RMHD WENO5 step). */
#define AFD_FACE_PATH_DEFAULT 3

/* Nozzle inflow on the -x ghost planes (core/nozzle.py:27-92).  The library's host code
 * evaluates nozzle_weight (core/nozzle.py:5-25) per (j, k) and uploads the profile. */
typedef struct afd_nozzle {
    double y0, z0;            /* jet centre (JET_CENTER[1], JET_CENTER[2]) */
    double jet_radius, shear_thick;
    int32_t profile;          /* 0 top_hat, 1 taper/smooth, 2 parabolic/parabola */
    double eta_rho, rho_amb, p_eq, gamma_jet, vx_amb, vy_amb, vz_amb;
    int32_t perturb;          /* 0: NOZZLE_PERTURB == "none" */
    double turb_vamp, turb_pamp;
    int32_t n_tracers, tracer_offset;
    double tracer_nozzle[8], tracer_amb[8];
    int32_t two_temperature, thermo_offset;
    double te_nozzle, ti_nozzle;
    int32_t b_init;           /* 0 none, 1 poloidal (Bx = B0 s), 2 toroidal (By = B0 s) */
    double b0;
    double b_amb[3];          /* additive ambient field added to the nozzle B (zeros: reference) */
    int32_t chemistry;        /* CHEMISTRY_ENABLED: write the nozzle ion fractions */
    int32_t chem_offset;      /* CHEM_OFFSET */
    double chem_nozzle[3];    /* CHEM_X_HII_NOZZLE, CHEM_X_HEII_NOZZLE, CHEM_X_HEIII_NOZZLE */
} afd_nozzle;

typedef struct afd_stats {
    uint64_t hydro_iters;     /* fixed-point iterations of the hydro (guess) solve */
    uint64_t newton_evals;    /* f(Z) evaluations of the RMHD Newton stage */
    uint64_t hydro_fallbacks; /* cells where the fixed point took the energy fallback */
    uint64_t bisections;      /* RMHD cells that reached bisection */
    uint64_t invalid;         /* RMHD cells whose recovery was invalid (hydro guess kept) */
    uint64_t learned_ok;      /* RMHD cells solved from the learned guess (acceptance test passed) */
    uint64_t c2p_cells;       /* cells converted */
    double ms_face[3];        /* summed HIP-event time of the x, y, z face kernels (timing on) */
    double ms_update;         /* summed time of the stage-update (+ con2prim) kernel */
    double ms_other;          /* BCs, passives, reductions */
    uint64_t launches_face[3], launches_update;
} afd_stats;

/* Run diagnostics of utils/diagnostics.py on the current state (interior cells). */
typedef struct afd_diag {
    double max_gamma;         /* max Lorentz factor (:15-27) */
    double inlet_flux;        /* sum of Sx dy dz over the plane i = NG (:31-43; p2c with Gamma 5/3 as there) */
    double divb_max, divb_rms, divb_mean; /* centred div B, RMHD only (:128-183) */
} afd_diag;

typedef struct afd_ctx afd_ctx;

/* Context: replaces configure(params) (core/srhd_core.py:18, core/rmhd_core.py:21)
 * and the module globals it sets, plus init_block's allocation
 * (solvers/srhd3d_mpi_muscl.py:128-163).  AFD_ERR_INVALID for grids too small for the
 * reconstruction stencil and for grids of 2^31 or more cells, ghosts included (the kernels
 * index cells in 32 bits). */
int afd_create(const afd_params *params, afd_ctx **out);
int afd_destroy(afd_ctx *ctx);

/* Host <-> device copies of the whole primitive array `pr` (n = nv * full cells). */
int afd_set_prim(afd_ctx *ctx, const double *host, int64_t n);
int afd_get_prim(afd_ctx *ctx, double *host, int64_t n);

/* Nozzle profile; afd_nozzle_draw_count gives the uniform draws one fill consumes
 * (3 per ghost cell with s > 0 when perturbed, in g -> j -> k order: core/nozzle.py:50-53). */
int afd_set_nozzle(afd_ctx *ctx, const afd_nozzle *nozzle);
int afd_nozzle_draw_count(afd_ctx *ctx, int64_t *ndraws);

/* Ghost fills in the reference order: y/z (periodic, core/boundary.py:4-8, or outflow) ->
 * +x outflow (core/boundary.py:10-11) -> -x nozzle (core/nozzle.py:27-92).
 * draws: host uniforms in [0, 1), required when the nozzle is perturbed. */
int afd_apply_bcs(afd_ctx *ctx, const double *draws, int64_t ndraws);
int afd_apply_periodic_yz(afd_ctx *ctx);               /* core/boundary.py:4-8 */
int afd_apply_outflow_right_x(afd_ctx *ctx);           /* core/boundary.py:10-11 */
int afd_apply_nozzle_left_x(afd_ctx *ctx, const double *draws, int64_t ndraws); /* core/nozzle.py:27-92 */

/* CFL speed of the current state: core/srhd_core.py:1267-1281 (max_char_speed). */
int afd_max_char_speed(afd_ctx *ctx, double *amax);

/* One SSPRK2/3 step of the current state, device resident:
 * core/srhd_core.py:1286-1400, core/rmhd_core.py:1559-1677 (step_ssprk2/3); with gr_metric set,
 * core/grhd_core.py:77-166 and core/grmhd_core.py:78-183, whose RHS (:20-75 / :20-77) wraps the
 * flat one. */
int afd_step(afd_ctx *ctx, double dt);

/* RHS of the current state, (nb + n_passive) x full cells, copied to host:
 * core/srhd_core.py:468 (compute_rhs_muscl), :926 (ppm), :1166 (weno);
 * core/rmhd_core.py:748 (compute_rhs_rmhd), :1279 (ppm), :1439 (weno). */
int afd_compute_rhs(afd_ctx *ctx, double *rhs_host, int64_t n);

/* core/source_terms.py:30-55 (apply_two_temperature) on the current state's interior;
 * a no-op unless the context's two_temperature is set and tei_tau > 0. */
int afd_two_temperature(afd_ctx *ctx, double dt);

/* core/source_terms.py:5-28 (apply_cooling_heating) on the current state's interior:
 * p <- min(max(p + dt (heat - cool rho p), 1e-12), p_cap), with cool = COOLING_LAMBDA,
 * heat = HEATING_RATE and p_cap = the settings' P_MAX (the reference's default is 1.0).
 * A no-op when cool and heat are both zero. */
int afd_cooling_heating(afd_ctx *ctx, double dt, double cool, double heat, double p_cap);

/* Constants of core/chemistry.py:59-68, read from the settings by the caller. */
typedef struct afd_chem {
    int32_t offset;           /* CHEM_OFFSET: xHII, xHeII, xHeIII at offset, offset+1, offset+2 (>= nb) */
    double X, Y;              /* CHEM_X, CHEM_Y */
    double tunit_k;           /* CHEM_TUNIT_K */
    double rate_scale, energy_scale;
    double tmin_k;            /* CHEM_TMIN_K */
    double gamma;             /* GAMMA of the settings (default 5/3) */
} afd_chem;

/* core/chemistry.py:43-139 (apply_ion_chemistry) on the current state's interior: the IMEX
 * update of the H/He ion fractions and its pressure coupling.  A no-op when nv <= offset + 2;
 * the caller checks CHEMISTRY_ENABLED and N_CHEM >= 3 (core/chemistry.py:48-53). */
int afd_ion_chemistry(afd_ctx *ctx, double dt, const afd_chem *chem);

/* Constants of core/dissipation.py:16-45, read from the settings by the caller (a missing, None
 * or zero RELAX_TAU_* falls back to RELAX_TAU, :19-21). */
typedef struct afd_diss {
    double zeta, eta, kappa;                /* BULK_ZETA, SHEAR_ETA, HEAT_KAPPA */
    double tau_bulk, tau_shear, tau_heat;   /* relaxation times */
    double cap_frac, p_max;                 /* |Pi|, |pi^ij|, |q^i| <= DISSIPATION_CAP_FRAC * P_MAX */
    int32_t advect;                         /* DISSIPATION_ADVECT */
} afd_diss;

/* core/dissipation.py:5-215 (apply_causal_dissipation) on the current state's interior: upwind
 * advection of the ten dissipative fields 5..14 (if advect) and their IMEX relaxation towards the
 * Navier-Stokes targets.  A no-op for RMHD, nv < 15, no terms or a non-positive relaxation time;
 * the caller checks DISSIPATION_ENABLED.  The grid spacing is the context's. */
int afd_causal_dissipation(afd_ctx *ctx, double dt, const afd_diss *diss);

typedef struct afd_sources {
    int32_t dissipation;      /* DISSIPATION_ENABLED (hydro) */
    afd_diss diss;
    int32_t cooling;          /* COOLING_ENABLED */
    double cool, heat, p_cap; /* as afd_cooling_heating */
    int32_t chemistry;        /* CHEMISTRY_ENABLED and N_CHEM >= 3 */
    afd_chem chem;
} afd_sources;

/* The driver's post-step operators, solvers/srhd3d_mpi_muscl.py:323-326: causal dissipation (if
 * sources->dissipation; a stencil, so two passes of its own), then apply_cooling_heating,
 * apply_two_temperature and apply_ion_chemistry in one pass over the interior that applies them
 * to each cell in that order.  Two-temperature follows the context's parameters. */
int afd_post_step_sources(afd_ctx *ctx, double dt, const afd_sources *sources);

/* Non-finite values in the current state (the driver's NaN check,
 * solvers/srhd3d_mpi_muscl.py:337-343). */
int afd_count_nonfinite(afd_ctx *ctx, int64_t *count);

/* Pointwise conversions of n states (row-major n x nb) with the context's parameters:
 * prim_to_cons (core/srhd_core.py:46, core/rmhd_core.py:137) and cons_to_prim
 * (core/srhd_core.py:61, core/rmhd_core.py:248).  iters (optional, n): iterations used. */
int afd_p2c_batch(afd_ctx *ctx, int64_t n, const double *prim, double *cons);
int afd_c2p_batch(afd_ctx *ctx, int64_t n, const double *cons, double *prim, int32_t *iters);

/* afd_c2p_batch without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and results as afd_c2p_batch; the device
 * staging of the n states in, the n states out and the n iteration counts lives in a
 * library-owned static buffer of AFD_NOALLOC_SCRATCH_BYTES on each device the library runs on,
 * held until the process ends (its address is resolved by afd_create).  n above
 * AFD_C2P_NOALLOC_MAX(nb) (nb = 5 SRHD, 9 RMHD) returns AFD_ERR_INVALID.  Every call on a device
 * shares that buffer: calls that overlap on one GPU (several contexts or threads) race. */
#define AFD_NOALLOC_SCRATCH_BYTES (16u << 20)
#define AFD_C2P_NOALLOC_MAX(nb) ((int64_t)(AFD_NOALLOC_SCRATCH_BYTES / (2u * 8u * (unsigned)(nb) + 4u)))
int afd_c2p_batch_noalloc(afd_ctx *ctx, int64_t n, const double *cons, double *prim, int32_t *iters);

/* utils/diagnostics.py:5-57 (max Gamma, inlet flux) and :128-183 (div B), as device reductions. */
int afd_diagnostics(afd_ctx *ctx, afd_diag *out);
/* utils/diagnostics.py:59-126: the mid-plane centerline, out = 4 x nx (Gamma, rho, p, vx). */
int afd_centerline(afd_ctx *ctx, double *out, int64_t n);

/* Counters and per-kernel HIP-event timings (no counterpart in the reference). */
int afd_set_instrumentation(afd_ctx *ctx, int32_t stats, int32_t timing);
int afd_get_stats(afd_ctx *ctx, afd_stats *out);
int afd_reset_stats(afd_ctx *ctx);
/* Per-cell con2prim branch of the last stage of the latest step run with the counters on
 * (no counterpart in the reference): out = one byte per cell (ghosts included), bit 0 hydro
 * fixed point took the energy fallback, bit 1 RMHD bisection, bit 2 RMHD invalid recovery (hydro
 * guess kept), bit 3 learned or previous-state start accepted. */
int afd_get_c2p_flags(afd_ctx *ctx, uint8_t *out, int64_t n);
int afd_synchronize(afd_ctx *ctx);

/* Learned con2prim guess (no counterpart in the reference): an MLP 5 -> 64 -> 64 -> 2 (SiLU)
 * on scaled (D, |S|, tau, B^2, S.B) giving (log p, log W); BF16 weights, FP32 accumulation
 * on MFMA.  Weights are the packed FP32 master copy (4674 values, layout in afd_mlp_layout). */
int afd_mlp_param_count(int32_t *count);
int afd_mlp_set_weights(afd_ctx *ctx, const float *w, int32_t count);  /* also resets the Adam state */
int afd_mlp_get_weights(afd_ctx *ctx, float *w, int32_t count);
/* Training pairs from primitive states (n x 9): the context's p2c, then its own FP64 con2prim;
 * X = 5 scaled features of the conserved state, Y = (log10(p/D)/4, log10 W) of the con2prim
 * output; valid = 1 only where the con2prim converged by Newton to a valid state (no bisection,
 * no fallback to the hydro guess) and every feature and target is finite. */
int afd_mlp_dataset_from_prims(afd_ctx *ctx, int64_t n, const double *prim, float *X, float *Y, int32_t *valid);
int afd_mlp_set_dataset(afd_ctx *ctx, int64_t n, const float *X, const float *Y);
/* steps of minibatch Adam (lr, betas 0.9/0.999) on consecutive minibatches of the dataset;
 * loss_history (optional, steps): mean squared error of each minibatch before its update. */
int afd_mlp_train(afd_ctx *ctx, int64_t steps, int32_t batch, float lr, double *loss_history);
/* checks: network outputs, and loss + gradient w.r.t. the FP32 master layout, for host data */
int afd_mlp_forward(afd_ctx *ctx, int64_t n, const float *X, float *Y);
int afd_mlp_loss_grad(afd_ctx *ctx, int64_t n, const float *X, const float *Y, float *grad, double *loss);

const char *afd_last_error(void);
const char *afd_version(void);

#ifdef __cplusplus
}
#endif
#endif
