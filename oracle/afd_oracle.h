/*
 * afd_oracle.h -- CPU FP64 restatement of AstroFD's SRHD/RMHD finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * never as the thing measured or shipped.
 *
 * Each function restates the reference algorithm in plain C, in the reference's
 * own evaluation order (left-to-right sums, pow(x, 2.0) for "**2", Python
 * max()/min() tie semantics), so that it agrees with a pure-Python evaluation of
 * the reference to the last bit wherever the reference is deterministic.
 * The reference files are cited as path:line in afd_oracle.c.
 *
 * Layout: primitive arrays are C-order (nv, nx, ny, nz) float64 with ghosts,
 * exactly the reference's `pr` (nx, ny, nz here are the FULL dims incl. 2*ng).
 */
#ifndef AFD_ORACLE_H
#define AFD_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int physics;          /* 0 = hydro (SRHD), 1 = rmhd */
    int nv;               /* total variables (5 or 9, plus passives) */
    int nx, ny, nz;       /* full array dims incl. ghosts */
    int ng;
    double gamma;         /* SRHD: the 5/3 bound at definition time (Q5); RMHD: GAMMA */
    double p_max, v_max;
    double glm_ch, glm_cp;
    int limiter_id;       /* 0 mc, 1 minmod, 2 vanleer */
    int recon_id;         /* 0 muscl, 1 ppm, 2 weno */
    int riemann_id;       /* SRHD: 0 hlle, 1 hllc.  RMHD: 0 hlle, 1 hlld (hybrid), 2 hlld_full */
    int rk_order;         /* 2 or 3 */
    int n_passive, passive_offset;
    int resistive;
    double resistivity;
    /* GRHD/GRMHD wrappers (core/grhd_core.py, core/grmhd_core.py): 0 minkowski, 1 schwarzschild,
       2 kerr-schild, 3 any other GR_METRIC (sources from the Kerr-Schild lapse, no rescale) */
    int gr_metric, gr_orthonormal, gr_x0;
    double gr_mass, gr_gamma; /* GR_MASS; the Gamma of the source's enthalpy (GRHD: config GAMMA) */
    int resist_once;          /* additive RESISTIVE_MODE "once": the Laplacian once per RHS in every scheme
                                 (the reference: once in MUSCL, twice in PPM, never in WENO) */
} orc_params;

/* pointwise (batched over n states, row-major n x nvar) */
void orc_srhd_p2c_batch(const orc_params *P, int64_t n, const double *prim, double *cons);
void orc_srhd_c2p_batch(const orc_params *P, int64_t n, const double *cons, double *prim);
void orc_srhd_sound_speed_batch(const orc_params *P, int64_t n, const double *prim, double *cs);
void orc_srhd_flux_batch(const orc_params *P, int64_t n, int dir, const double *prim, double *flux);
double orc_limiter(int limiter_id, double a, double b);
void orc_ppm(const double *q5, double *lr);
double orc_weno5_left(const double *q5);
double orc_weno5_right(const double *q5);
/* Riemann solver on floored L/R primitive states along dir (0,1,2), speeds from the states */
void orc_srhd_riemann(const orc_params *P, int dir, const double *primL, const double *primR, double *F);

void orc_rmhd_p2c_batch(const orc_params *P, int64_t n, const double *prim, double *cons);
void orc_rmhd_c2p_batch(const orc_params *P, int64_t n, const double *cons, double *prim);
/* flags per state: 2 = bisection (Newton failed from both starts), 4 = invalid recovery */
void orc_rmhd_c2p_flags_batch(const orc_params *P, int64_t n, const double *cons, double *prim, int32_t *flags);
void orc_rmhd_hydro_guess_batch(const orc_params *P, int64_t n, const double *cons5, double *prim5);
void orc_rmhd_f_of_Z_batch(const orc_params *P, int64_t n, const double *Z, const double *cons, double *f);
void orc_rmhd_fast_speed_batch(const orc_params *P, int64_t n, const double *prim, double *cf);
/* unit_speeds != 0 replaces the estimated speeds by the (-1, 1) seed */
void orc_rmhd_riemann(const orc_params *P, int dir, const double *primL, const double *primR,
                      int unit_speeds, double *F);

/* whole-array operators */
void orc_rhs(const orc_params *P, const double *pr, double *rhs, double dx, double dy, double dz);
void orc_step(const orc_params *P, const double *pr, double *out, double dx, double dy, double dz, double dt);
double orc_max_char_speed(const orc_params *P, const double *pr);
void orc_apply_periodic_yz(const orc_params *P, double *pr);
void orc_apply_outflow_right_x(const orc_params *P, double *pr);
/* y/z zero-gradient copies (the additive BC_Y/BC_Z="outflow" option) */
void orc_apply_outflow_yz(const orc_params *P, double *pr, int y, int z);

typedef struct orc_nozzle {
    double dx, dy, dz, y0, z0;
    int ny_loc, nz_loc;
    double jet_radius, shear_thick;
    int profile;              /* 0 top_hat, 1 taper/smooth, 2 parabolic/parabola */
    double eta_rho, rho_amb, p_eq, gamma_jet, vx_amb, vy_amb, vz_amb;
    int perturb;              /* 0 none, else random */
    double turb_vamp, turb_pamp;
    int n_tracers, tracer_offset;
    const double *tracer_nozzle, *tracer_amb;
    int two_temperature, thermo_offset;
    double te_nozzle, ti_nozzle;
    int b_init;               /* 0 none, 1 poloidal, 2 toroidal */
    double b0;
    double b_amb[3];          /* additive B_AMB (zeros keep reference behaviour) */
    int chemistry, chem_offset;
    double chem_nozzle[3];    /* CHEM_X_HII/HEII/HEIII_NOZZLE */
} orc_nozzle;

double orc_nozzle_weight(double r, double radius, double shear_thick, int profile);
/* consumes 3 draws (uniform [0,1)) per nozzle cell with s > 0 when perturb != 0;
   returns the number of draws consumed */
int64_t orc_apply_nozzle_left_x(const orc_params *P, const orc_nozzle *N, double *pr,
                                const double *draws, int64_t ndraws);
void orc_apply_two_temperature(const orc_params *P, double *pr, double dt, double tei_tau, int thermo_offset);
void orc_apply_cooling_heating(const orc_params *P, double *pr, double dt, double cool, double heat, double pmax);
typedef struct orc_chem {
    int offset;
    double X, Y, tunit, rate_scale, energy_scale, tmin, gamma;
} orc_chem;
void orc_apply_ion_chemistry(const orc_params *P, double *pr, double dt, const orc_chem *chem);
typedef struct orc_diss {
    double zeta, eta, kappa;                /* BULK_ZETA, SHEAR_ETA, HEAT_KAPPA */
    double tau_bulk, tau_shear, tau_heat;   /* RELAX_TAU_* (resolved against RELAX_TAU) */
    double cap_frac, p_max;                 /* DISSIPATION_CAP_FRAC, P_MAX */
    int advect;                             /* DISSIPATION_ADVECT */
} orc_diss;
void orc_apply_causal_dissipation(const orc_params *P, double *pr, double dt, double dx, double dy, double dz,
                                  const orc_diss *diss);

/* number of OpenMP threads the whole-array operators use (1 when built without OpenMP) */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
