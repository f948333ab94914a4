"""GPU tests of the mixed-precision path (FP32 reconstruction and Riemann fluxes; FP64 face
states, Lorentz factor, con2prim and SSPRK accumulation) against the FP64 CPU oracle.

Bar: the metric's L1(rho) <= 1e-6 relative after a step, for rho, p and W, on smooth seeded
states, and max |error| within FP32 resolution of the flux update (dt / dx * |F| * 2^-24
times a modest growth factor, <= 1e-4 relative here)."""
import numpy as np
import pytest

import oracle as O
from tests.gpu_helpers import log, make_domain
from tests.golden_states import smooth_state

pytestmark = pytest.mark.gpu


def _lorentz(pr):
    v2 = pr[1] ** 2 + pr[2] ** 2 + pr[3] ** 2
    return 1.0 / np.sqrt(1.0 - np.minimum(v2, 1.0 - 1e-14))


CASES = [
    ("hydro", "muscl", "vanleer", "hllc", 3),
    ("hydro", "weno", "mc", "hlle", 2),
    ("hydro", "ppm", "mc", "hllc", 3),
    ("rmhd", "weno", "mc", "hlld_full", 3),
    ("rmhd", "ppm", "mc", "hlld", 3),
    ("rmhd", "muscl", "minmod", "hlle", 2),
]


@pytest.mark.parametrize("phys,recon,lim,riem,rk", CASES)
def test_mixed_step_vs_oracle(phys, recon, lim, riem, rk):
    log(f"test_mixed_step_vs_oracle {phys} {recon} {lim} {riem} rk{rk}")
    check_mixed_step(phys, recon, lim, riem, rk, (22, 20, 18))


@pytest.mark.parametrize("face_path", [2, 3, 4])
@pytest.mark.parametrize("interior", [(22, 20, 18), (19, 16, 70)])
def test_mixed_resistive_laplacian(interior, face_path):
    """The resistive Laplacian of B applied once per right-hand side (RESISTIVE_MODE "once") on
    the FP32 WENO5 face variants: face fluxes stored (2), the flux divergence accumulated (3) and
    the y faces in the update (4); the stage update adds the Laplacian in FP64 after the FP32
    divergence.  Resistivity 0.02 on unequal spacings: the term moves B by 0.2 to 4 % (L1) in the
    step, so a part dropped, counted twice, given the wrong sign or another axis's spacing fails
    the bars above (1e-5) by orders of magnitude."""
    log(f"test_mixed_resistive_laplacian {interior} face_path={face_path}")
    check_mixed_step("rmhd", "weno", "mc", "hlld_full", 3, interior, vamp=0.2, face_path=face_path,
                     resistivity=0.02, spacing=(1.0, 1.3, 0.8))


@pytest.mark.parametrize("recon,lim,riem", [("weno", "mc", "hlld_full"), ("ppm", "mc", "hlld"), ("muscl", "vanleer", "hlld")])
@pytest.mark.parametrize("interior", [(19, 16, 70), (21, 18, 130), (2, 3, 2)])
def test_mixed_shared_edges_shapes(interior, recon, lim, riem):
    """The FP32 RMHD faces on interior extents that leave partial 64-lane z chunks (70, 130), a last
    group along the normal of 2 or 1 faces, and partial last waves (the WENO5 faces share cell
    edges through 4-wave LDS rows for x/y and lane shuffles for z; PPM and MUSCL run one thread
    per face), against the FP64 oracle with the bars above.
    |v| <= 0.2 as in the FP64 parity tests: the modes span a fixed number of periods per axis,
    and at 0.5 on these extents hlld_full's branch thresholds make the per-face FP32 kernels miss
    the bars as well (L1(Bz) 3e-4)."""
    log(f"test_mixed_shared_edges_shapes {interior} {recon} {lim} {riem}")
    check_mixed_step("rmhd", recon, lim, riem, 3, interior, vamp=0.2)


@pytest.mark.parametrize("face_path", [2, 3, 4])
@pytest.mark.parametrize("interior", [(19, 16, 70), (21, 18, 130), (2, 3, 2), (22, 20, 18)])
def test_mixed_weno_face_variants(interior, face_path):
    """The shared-edge FP32 WENO5 RMHD face kernels (afd_params.face_path): face fluxes stored (2),
    the flux divergence accumulated by z lanes and x/y pencils (3), and with the y pencils fused into
    the stage update (4), on extents with partial z chunks, runs of pencils shorter than one round,
    and partial last waves, against the FP64 oracle with the bars above."""
    log(f"test_mixed_weno_face_variants {interior} face_path={face_path}")
    check_mixed_step("rmhd", "weno", "mc", "hlld_full", 3, interior, vamp=0.2, face_path=face_path)


@pytest.mark.parametrize("phys,recon,lim,riem", [("rmhd", "weno", "mc", "hlld_full"), ("rmhd", "muscl", "mc", "hlld"),
                                                 ("hydro", "muscl", "vanleer", "hllc")])
def test_mixed_step_passives_vs_oracle(phys, recon, lim, riem):
    """Two passive scalars against the FP64 oracle.  SRHD MUSCL and RMHD MUSCL evaluate the
    reference's passive fluxes (FP64) with the FP32 faces' HLL speeds and face velocities; RMHD WENO5
    has no passive flux in the reference (core/rmhd_core.py:1439-1557), so there the passives keep
    a zero right-hand side while the shared-edge WENO5 kernels run beside them."""
    log(f"test_mixed_step_passives_vs_oracle {phys} {recon}")
    check_mixed_step(phys, recon, lim, riem, 3, (22, 20, 18), vamp=0.2, npass=2)


@pytest.mark.parametrize("riem", ["hlle", "hllc"])
@pytest.mark.parametrize("interior", [(19, 16, 70), (21, 18, 130), (2, 3, 2)])
def test_mixed_srhd_weno_faces(interior, riem):
    """The FP32 SRHD WENO5 faces (one thread per face: sharing cell edges gained 12 % on the z faces
    only, 2 % of the 256^3 step) on odd extents, against the FP64 oracle with the bars above."""
    log(f"test_mixed_srhd_weno_faces {interior} {riem}")
    check_mixed_step("hydro", "weno", "mc", riem, 3, interior, vamp=0.2)


@pytest.mark.parametrize("riem", ["hlle", "hlld"])
def test_mixed_weno_other_riemann_solvers(riem):
    """The shared-edge FP32 WENO5 RMHD faces with HLLE (the reference's WENO RMHD config) and hlld,
    against the FP64 oracle with the bars above."""
    log(f"test_mixed_weno_other_riemann_solvers {riem}")
    check_mixed_step("rmhd", "weno", "mc", riem, 3, (22, 20, 18))


@pytest.mark.parametrize("riem", ["hlld_full", "hlld"])
def test_mixed_per_face_fp32_faces(riem):
    """afd_params.face_path = 1: the FP32 WENO5 RMHD faces with one thread per face, the path that
    grids of 2^29 cells or more take, against the FP64 oracle with the bars above."""
    log(f"test_mixed_per_face_fp32_faces {riem}")
    check_mixed_step("rmhd", "weno", "mc", riem, 3, (22, 20, 18), face_path=1)


@pytest.mark.parametrize("phys,riem,face_path", [("rmhd", "hlle", 1), ("rmhd", "hlle", 2), ("rmhd", "hlle", 3),
                                                ("rmhd", "hlle", 4), ("rmhd", "hlld_full", 0), ("hydro", "hllc", 0)])
def test_mixed_weno_large_jumps(phys, riem, face_path):
    """Density jumps of 1e10 across x, y and z faces.  FP32 WENO5 smoothness indicators of such
    deviations overflow when squared and the nonlinear weights underflow to 0 (0 / 0 = NaN before
    the rescaled evaluation); the FP32 faces must come out finite and the step within the bars
    above of the FP64 oracle for rho, B and psi.  The pressure and velocity of the dense box
    (p <= P_MAX against rho = 1e10) come out of conserved states 1e10 times those around them, so
    their bar is FP32 resolution of the conserved update: every conserved component within
    1e-5 max(tau + D).  The L1 of rho over a field whose 512 box cells are 1e10 times denser is the
    error of those cells: L1 <= 1e-5.  Covers every RMHD face-kernel variant (face_path 1-4) and the
    per-face SRHD kernel.  With hll_d_flux_dir the FP64 oracle's own step blows up at the contact
    (energies of 1e41 after one step), so that case checks only that the mixed step is finite."""
    log(f"test_mixed_weno_large_jumps {phys} {riem} face_path={face_path}")
    check_mixed_step(phys, "weno", "mc", riem, 3, (22, 20, 18), vamp=0.2, face_path=face_path, jump=1e10,
                     finite_only=riem == "hlld_full")


@pytest.mark.parametrize("riem,face_path", [("hlle", 3), ("hlle", 4), ("hlld_full", 3)])
def test_mixed_fp32_overflow_repaired_in_fp64(riem, face_path):
    """A box of density 1e40, beyond FP32's range (3.4e38): every FP32 face whose stencil touches
    it overflows, and the cells around it get a non-finite FP32 right-hand side.  On the
    flux-divergence path (face_path 3, and 4 with the y faces in the update) the stage update
    queues those cells and k_repair_weno evaluates their faces in FP64 as the plain path does, so
    the mixed step is finite wherever the FP64 step is.  Bars as for the 1e10 jumps; with
    hll_d_flux_dir, whose FP64 step blows up at the contact, finiteness only."""
    log(f"test_mixed_fp32_overflow_repaired_in_fp64 {riem} face_path={face_path}")
    check_mixed_step("rmhd", "weno", "mc", riem, 3, (22, 20, 18), vamp=0.2, face_path=face_path, jump=1e40,
                     finite_only=riem == "hlld_full")


@pytest.mark.parametrize("phys,riem", [("hydro", "hllc"), ("rmhd", "hlle")])
def test_mixed_vanleer_large_jumps(phys, riem):
    """MUSCL with the van Leer limiter across density jumps of 1e20: the FP32 slopes' product
    (above ~1.8e19 squared) overflows, and 2 a b / (a + b) gave inf / NaN slopes where FP64 is
    finite; the FP32 limiter forms 2 a (b / (a + b)) instead.  The SRHD 256^3 jet's own state
    reaches such contrasts (densities of 1e-12 next to 1e18) by step 90.  Bars as for the WENO5
    jumps: rho's L1, and every conserved component within 1e-5 max(tau + D)."""
    log(f"test_mixed_vanleer_large_jumps {phys} {riem}")
    check_mixed_step(phys, "muscl", "vanleer", riem, 3, (22, 20, 18), vamp=0.2, jump=1e20)


def test_mixed_vanleer_tiny_slopes():
    """MUSCL van Leer in FP32 where both slopes are tiny and share a sign: velocity deviations of a
    few 1e-39 (FP32 denormals), whose sum's reciprocal is not finite.  The slope must come out 0
    (the product of the slopes underflows), not inf or NaN.  Bars as in check_mixed_step."""
    log("test_mixed_vanleer_tiny_slopes")
    ng = 3
    pr = smooth_state((22 + 2 * ng, 20 + 2 * ng, 18 + 2 * ng), 97, "hydro", vamp=0.2)
    ramp = np.arange(pr.shape[1], dtype=np.float64)[:, None, None]
    pr[2] = 4e-39 * ramp * np.ones(pr.shape[1:])  # vy: same-sign deviations of 4e-39 along x
    pr[3] = 3e-39 * np.arange(pr.shape[3], dtype=np.float64)[None, None, :] * np.ones(pr.shape[1:])
    dx = 1.0 / 22
    dt = 0.4 * dx * 0.5
    d = make_domain("hydro", pr.shape, ng, dxyz=(dx, dx, dx), recon="muscl", limiter="vanleer", riemann="hllc",
                    rk_order=3, precision="mixed")
    d.set_prim(pr)
    d.step(dt)
    out = d.get_prim()
    d.close()
    P = O.make_params("hydro", pr.shape[1:], nv=5, ng=ng, p_max=2.0, v_max=0.9999, limiter="vanleer", recon="muscl",
                      riemann="hllc", rk_order=3)
    ref = O.step(P, pr, dx, dx, dx, dt)
    core = (slice(ng, -ng),) * 3
    assert np.all(np.isfinite(out))
    for name, a, b in (("rho", out[0], ref[0]), ("p", out[4], ref[4]), ("W", _lorentz(out), _lorentz(ref))):
        a, b = a[core], b[core]
        l1 = np.abs(a - b).sum() / np.abs(b).sum()
        log(f"  {name}: L1 rel {l1:.2e}")
        assert l1 <= (1e-6 if name == "rho" else 1e-5), name


# Seven cells along x of a state the SRHD 256^3 jet's mixed run reached (step 72): a cold dense
# cell (rho 5.4e5, p 1e-12) moving left beside a near-vacuum cell moving right.  (rho, vx, vy, vz, p)
VACUUM_ROW = np.array([
    [1.43245e+04, 9.69516e-01, 9.27624e-02, -1.43041e-01, 1.00000e-12],
    [5.40249e+05, -4.66499e-01, -5.39392e-01, 6.28299e-01, 1.00000e-12],
    [1.00000e-12, 5.00162e-01, 6.58097e-01, -5.62624e-01, 1.00000e-12],
    [2.14480e+03, -5.12765e-01, 7.96350e-01, 3.20468e-01, 2.00000e+00],
    [2.57216e+04, -2.87034e-01, -7.42998e-01, -6.04455e-01, 1.50000e-12],
    [3.15012e+04, -5.77672e-01, -1.61371e-02, 8.15987e-01, 1.00000e-12],
    [1.00000e-12, 6.13623e-01, -1.24587e-01, -7.79580e-01, 1.00000e-12],
])


@pytest.mark.parametrize("phys,riem", [("hydro", "hllc"), ("rmhd", "hlld")])
def test_mixed_hllc_vacuum_faces(phys, riem):
    """FP32 HLLC at faces that open a vacuum: the cold dense cell's slower wave speed rounds onto
    its velocity in FP32, the contact-speed denominator cancels to 0 and the star pressure became
    0 x inf = NaN where FP64 is finite.  Such faces take the HLLE flux in FP32 (FP64 keeps HLLC).
    The rows above, repeated through y and z in a smooth state, with MUSCL van Leer: every face of
    their vacuum planes is such a face, so the rows' cells differ from the FP64 HLLC step by what
    HLLE changes there (about 20 % of their density here; the jet has a few dozen such faces among
    10^8).  The step must be finite and non-negative in rho and p, and the cells outside the rows'
    reach (3 stages of MUSCL) within the usual bars."""
    log(f"test_mixed_hllc_vacuum_faces {phys} {riem}")
    ng = 2
    shape = (22 + 2 * ng, 12 + 2 * ng, 12 + 2 * ng)
    pr = smooth_state(shape, 98, phys, vamp=0.2)
    for r, row in enumerate(VACUUM_ROW):
        for v in range(5):
            pr[v][ng + 6 + r] = row[v]
    gamma = 4.0 / 3.0 if phys == "rmhd" else None
    dx = 1.0 / 256
    dt = 0.4 * dx
    d = make_domain(phys, pr.shape, ng, dxyz=(dx, dx, dx), gamma=gamma, recon="muscl", limiter="vanleer",
                    riemann=riem, rk_order=3, precision="mixed")
    d.set_prim(pr)
    d.step(dt)
    out = d.get_prim()
    d.close()
    P = O.make_params(phys, pr.shape[1:], nv=pr.shape[0], ng=ng, gamma=gamma, p_max=2.0, v_max=0.9999,
                      limiter="vanleer", recon="muscl", riemann=riem, rk_order=3)
    ref = O.step(P, pr, dx, dx, dx, dt)
    core = (slice(ng, -ng),) * 3
    assert np.all(np.isfinite(out))
    assert out[0][core].min() > 0.0 and out[4][core].min() > 0.0
    # the smooth part of the field, x planes away from the vacuum rows' stencils (3 stages of MUSCL)
    far = (slice(ng + 19, -ng),) + (slice(ng, -ng),) * 2
    for v in range(5):
        a, b = out[v][far], ref[v][far]
        l1 = np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)
        log(f"  var {v} away from the rows: L1 rel {l1:.2e}")
        assert l1 <= 1e-5, v


def check_mixed_step(phys, recon, lim, riem, rk, interior, vamp=0.5, npass=0, face_path=0, jump=None,
                     finite_only=False, resistivity=0.0, spacing=(1.0, 1.0, 1.0)):
    ng = 3
    shape = tuple(n + 2 * ng for n in interior)
    pr = smooth_state(shape, 91, phys, vamp=vamp, npass=npass)
    if jump is not None:  # a dense box with faces along every axis
        pr[0][ng + 6:ng + 14, ng + 5:ng + 13, ng + 4:ng + 12] *= jump
    gamma = 4.0 / 3.0 if phys == "rmhd" else None
    dx = 1.0 / 22
    dxyz = tuple(dx * f for f in spacing)
    dt = 0.4 * min(dxyz) * 0.5
    d = make_domain(phys, pr.shape, ng, dxyz=dxyz, gamma=gamma, recon=recon, limiter=lim, riemann=riem,
                    rk_order=rk, precision="mixed", n_passive=npass, face_path=face_path,
                    n_resist=1 if resistivity else 0, resistivity=resistivity)
    d.set_prim(pr)
    d.step(dt)
    out = d.get_prim()
    d.close()
    P = O.make_params(phys, pr.shape[1:], nv=pr.shape[0], ng=ng, gamma=gamma, p_max=2.0, v_max=0.9999,
                      limiter=lim, recon=recon, riemann=riem, rk_order=rk, n_passive=npass,
                      resistive=resistivity > 0.0, resistivity=resistivity, resist_once=resistivity > 0.0)
    ref = O.step(P, pr, *dxyz, dt)
    core = (slice(ng, -ng),) * 3
    # finite wherever the FP64 step is (everywhere, in every case here but the 1e40 box with
    # hll_d_flux_dir, whose FP64 step may itself overflow there)
    log(f"  non-finite: mixed {int(np.sum(~np.isfinite(out)))}, FP64 {int(np.sum(~np.isfinite(ref)))}")
    assert np.all(np.isfinite(out) | ~np.isfinite(ref))
    if finite_only:
        return
    assert np.all(np.isfinite(out))
    if jump is not None:
        # p and W of the cold dense gas come out of conserved states whose fluxes are 1e10 times
        # those around them: the bar is the FP32 resolution of the conserved update, every
        # conserved component within 1e-5 of the largest energy density tau + D
        nb = 9 if phys == "rmhd" else 5
        p2c = O.rmhd_p2c if phys == "rmhd" else O.srhd_p2c
        Ur, Uo = (p2c(P, np.moveaxis(x[:nb][(slice(None),) + core], 0, -1).reshape(-1, nb)) for x in (ref, out))
        E = np.abs(Ur[:, 4] + Ur[:, 0]).max()
        for c in range(5):
            dmax = np.abs(Uo[:, c] - Ur[:, c]).max()
            log(f"  conserved {c}: max |dU| / max(tau + D) {dmax / E:.2e}")
            assert dmax <= 1e-5 * E, c
    for name, a, b in (("rho", out[0], ref[0]), ("p", out[4], ref[4]), ("W", _lorentz(out), _lorentz(ref))):
        a, b = a[core], b[core]
        l1 = np.abs(a - b).sum() / np.abs(b).sum()
        mx = np.abs(a - b).max() / np.abs(b).max()
        log(f"  {name}: L1 rel {l1:.2e}  max rel {mx:.2e}")
        if jump is not None and name != "rho":  # bounded by the conserved update above
            continue
        # the metric's bar is L1(rho) <= 1e-6; p (and W) come out of tau - kinetic - magnetic energy,
        # which cancels: their relative error is larger by ~tau / p (~10 in these states)
        assert l1 <= (1e-6 if name == "rho" and jump is None else 1e-5), name
        # hll_d_flux_dir picks among five regions and HLLE fallbacks by thresholds (|Bn| < 1e-8,
        # D* <= SMALL, sign tests): an FP32 rounding next to one switches a face's branch, a
        # local jump of dt/dx times the difference of two consistent fluxes (up to ~1e-2 here), in
        # a few cells; every other cell stays within FP32 resolution of the update (1e-4)
        if riem == "hlld_full":
            frac = float(np.mean(np.abs(a - b) > 1e-4 * np.abs(b).max()))
            log(f"  {name}: cells beyond 1e-4 {frac:.2e}")
            assert mx <= 1e-2 and frac <= 1e-3, name
        else:
            assert mx <= 1e-4, name
    # B, psi and passives: the same bar as p (their fluxes come out of the FP32 Riemann solve, or
    # use its speeds and face velocities)
    for v in range(5, pr.shape[0]):
        a, b = out[v][core], ref[v][core]
        l1 = np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)
        log(f"  var {v}: L1 rel {l1:.2e}")
        assert l1 <= 1e-5, v


@pytest.mark.parametrize("case", ["dissipation_muscl", "dissipation_ppm", "grhd_schwarzschild", "grmhd_kerr_schild"])
def test_mixed_step_extensions_vs_oracle(case):
    """The mixed path with the dissipative SRHD face hooks (fields 5..14 in FP64, added to the FP32
    fluxes) and with the GR wrappers (the rescaled flux velocities in FP32 faces, the lapse
    sources in the FP64 update), against the FP64 oracle, with the bars above."""
    log(f"test_mixed_step_extensions_vs_oracle {case}")
    ng = 3
    dx = 1.0 / 22
    dt = 0.4 * dx * 0.5
    if case.startswith("dissipation"):
        recon = case.split("_")[1]
        base = smooth_state((22 + 2 * ng, 20 + 2 * ng, 18 + 2 * ng), 93, "hydro", vamp=0.5)
        diss = np.random.default_rng(94).uniform(-0.004, 0.004, (10,) + base.shape[1:])
        pr = np.concatenate([base[:5], diss])
        d = make_domain("hydro", pr.shape, ng, dxyz=(dx, dx, dx), recon=recon, limiter="mc", riemann="hllc",
                        rk_order=3, precision="mixed", passive_offset=15)
        P = O.make_params("hydro", pr.shape[1:], nv=15, ng=ng, p_max=2.0, v_max=0.9999, limiter="mc", recon=recon,
                          riemann="hllc", rk_order=3, passive_offset=15)
        phys = "hydro"
    else:
        phys = "rmhd" if case.startswith("grmhd") else "hydro"
        metric = "kerr-schild" if "kerr" in case else "schwarzschild"
        pr = smooth_state((22 + 2 * ng, 20 + 2 * ng, 18 + 2 * ng), 95, phys, vamp=0.3)
        gamma = 4.0 / 3.0 if phys == "rmhd" else None
        g_gr = 4.0 / 3.0 if phys == "rmhd" else 1.4
        d = make_domain(phys, pr.shape, ng, dxyz=(dx, dx, dx), gamma=gamma, recon="weno", limiter="mc",
                        riemann="hlld_full" if phys == "rmhd" else "hllc", rk_order=3, precision="mixed",
                        gr_metric=metric, gr_mass=0.05, gr_gamma=g_gr)
        P = O.make_params("grmhd" if phys == "rmhd" else "hydro", pr.shape[1:], nv=pr.shape[0], ng=ng, gamma=gamma,
                          p_max=2.0, v_max=0.9999, limiter="mc", recon="weno",
                          riemann="hlld_full" if phys == "rmhd" else "hllc", rk_order=3, gr_metric=metric,
                          gr_mass=0.05, gr_gamma=g_gr)
    d.set_prim(pr)
    d.step(dt)
    out = d.get_prim()
    d.close()
    ref = O.step(P, pr, dx, dx, dx, dt)
    core = (slice(ng, -ng),) * 3
    assert np.all(np.isfinite(out))
    for name, a, b in (("rho", out[0], ref[0]), ("p", out[4], ref[4]), ("W", _lorentz(out), _lorentz(ref))):
        a, b = a[core], b[core]
        l1 = np.abs(a - b).sum() / np.abs(b).sum()
        mx = np.abs(a - b).max() / np.abs(b).max()
        log(f"  {name}: L1 rel {l1:.2e}  max rel {mx:.2e}")
        assert l1 <= (1e-6 if name == "rho" else 1e-5), name
        assert mx <= (1e-3 if phys == "rmhd" else 1e-4), name
    assert np.array_equal(out[5:15], pr[5:15]) if case.startswith("dissipation") else True
