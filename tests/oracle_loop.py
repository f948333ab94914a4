"""Driver-loop restatement over the CPU oracle (test helper).

Follows the time loop of solvers/srhd3d_mpi_muscl.py:276-326 for one rank:
CFL from max_char_speed (SRHD) or amax = 1 (RMHD, :278-281), BCs before and
after the step in the order periodic y/z -> outflow +x -> nozzle -x
(:289-293, :318-322), then the post-step operators: Israel-Stewart dissipation (:323),
cooling/heating, two-temperature, ion chemistry (:324-326).
"""
import numpy as np

import oracle as O


def params_from_settings(s, shape):
    rm = s["PHYSICS"] in ("rmhd", "grmhd")
    gr = s["PHYSICS"] in ("grhd", "grmhd")
    return O.make_params("rmhd" if rm else "hydro", shape[1:], nv=shape[0], ng=s["NG"],
                         gr_metric=s.get("GR_METRIC", "minkowski") if gr else "minkowski",
                         gr_mass=s.get("GR_MASS", 1.0), orthonormal_flux=s.get("ORTHONORMAL_FLUX", True),
                         gr_gamma=s.get("GAMMA", 5.0 / 3.0),
                         gamma=s["GAMMA"] if rm else None, p_max=s["P_MAX"], v_max=s["V_MAX"],
                         glm_ch=s.get("GLM_CH", 1.0), glm_cp=s.get("GLM_CP", 0.1), limiter=s["LIMITER"],
                         recon=s["RECON"], riemann=s["RIEMANN"], rk_order=s["RK_ORDER"],
                         n_passive=s.get("N_PASSIVE", 0), passive_offset=s.get("PASSIVE_OFFSET"),
                         resistive=s.get("RESISTIVE_ENABLED", False), resistivity=s.get("RESISTIVITY", 0.0),
                         resist_once=str(s.get("RESISTIVE_MODE", "reference")).lower() == "once")


def apply_bcs(P, s, pr, rng, b_amb=(0.0, 0.0, 0.0), bc_y="periodic", bc_z="periodic"):
    dx, dy, dz = s["Lx"] / s["NX"], s["Ly"] / s["NY"], s["Lz"] / s["NZ"]
    if bc_y == "periodic" and bc_z == "periodic":
        O.apply_periodic_yz(P, pr)
    else:
        O.apply_outflow_yz(P, pr, bc_y == "outflow", bc_z == "outflow")
    O.apply_outflow_right_x(P, pr)
    ndraw = 0
    if s.get("NOZZLE_PERTURB", "none") != "none":
        ndraw = 3 * s["NG"] * s["NY"] * s["NZ"]
    draws = rng.random(ndraw) if ndraw else None
    used = O.apply_nozzle_left_x(P, pr, dx, dy, dz, s["NY"], s["NZ"], s["JET_CENTER"][1], s["JET_CENTER"][2],
                                 s, draws, b_amb)
    return draws, used


class DrawStream:
    """Wraps a numpy Generator so that the oracle consumes exactly the draws the reference
    would (3 per perturbed nozzle cell), in order."""

    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)
        self.buf = np.zeros(0)

    def random(self, n):
        if self.buf.size < n:
            self.buf = np.concatenate([self.buf, self.rng.random(n - self.buf.size)])
        return self.buf

    def consume(self, k):
        self.buf = self.buf[k:]


def run(s, pr, nsteps, seed=1234, b_amb=(0.0, 0.0, 0.0), bc_y="periodic", bc_z="periodic"):
    P = params_from_settings(s, pr.shape)
    dx, dy, dz = s["Lx"] / s["NX"], s["Ly"] / s["NY"], s["Lz"] / s["NZ"]
    rng = DrawStream(seed)

    def bcs(a):
        _, used = apply_bcs(P, s, a, rng, b_amb, bc_y, bc_z)
        rng.consume(used)

    bcs(pr)
    dts = []
    for _ in range(nsteps):
        amax = O.max_char_speed(P, pr) if s["PHYSICS"] in ("hydro", "grhd") else 1.0
        dt = s["CFL"] * min(dx, dy, dz) / max(amax, 1e-12)
        bcs(pr)
        pr = O.step(P, pr, dx, dy, dz, dt)
        bcs(pr)
        O.apply_causal_dissipation(P, pr, dt, dx, dy, dz, s)
        O.apply_cooling_heating(P, pr, dt, s)
        if s.get("TWO_TEMPERATURE", False):
            O.apply_two_temperature(P, pr, dt, s["TEI_TAU"], s["THERMO_OFFSET"])
        O.apply_ion_chemistry(P, pr, dt, s)
        dts.append(dt)
    return pr, np.array(dts)
