"""afd_c2p_batch_noalloc (the allocation-free variant of afd_c2p_batch) against afd_c2p_batch.

Both run the same kernels on the same states, so the primitives and the iteration counts must be
bitwise equal.  Two calls in a row on one context, the second with fewer states, show that the
shared device staging carries nothing from one call into the next; a size above the staging's
capacity must be refused with AFD_ERR_INVALID and a message.
"""
import ctypes

import numpy as np
import pytest

from astrofd_amd import _lib as L
from tests.gpu_helpers import log, make_domain

pytestmark = pytest.mark.gpu


def rmhd_states(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(-3, 1, n)
    W = 1.0 + 10 ** rng.uniform(-6, 1.3, n)
    v = np.sqrt(1.0 - 1.0 / W ** 2)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = rho * 10 ** rng.uniform(-3, 0, n)
    B = rng.normal(size=(n, 3)) * 0.3
    return np.column_stack([rho, v[:, None] * d, p, B, 1e-3 * rng.uniform(-1, 1, n)])


def srhd_states(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(-3, 1, n)
    v = rng.uniform(0, 0.99, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = 10 ** rng.uniform(-4, 0.2, n)
    return np.column_stack([rho, v[:, None] * d, p])


def noalloc(d, cons):
    cons = L.f64c(cons).reshape(-1, d.nb)
    out = np.full_like(cons, np.nan)
    it = np.full(cons.shape[0], -1, dtype=np.int32)
    rc = d.lib.afd_c2p_batch_noalloc(d.h, cons.shape[0], L.dptr(cons), L.dptr(out),
                                     it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return rc, out, it


@pytest.mark.parametrize("phys,guess", [("hydro", "plain"), ("rmhd", "plain"), ("rmhd", "learned")])
def test_noalloc_matches_c2p_batch(phys, guess):
    nb = 9 if phys == "rmhd" else 5
    make = rmhd_states if phys == "rmhd" else srhd_states
    kw = dict(gamma=4.0 / 3.0, p_max=1e4) if phys == "rmhd" else dict(p_max=2.0)
    with make_domain(phys, (nb, 10, 10, 10), 2, v_max=0.9999, c2p_guess=guess, **kw) as d:
        if guess == "learned":
            from astrofd_amd import learned
            d.set_mlp_weights(learned.load(learned.MODEL_PATH))
        nmax = L.c2p_noalloc_max(nb)
        for n, seed in ((50000, 1), (1234, 2)):
            cons = d.p2c(make(n, seed))
            log(f"[gpu] c2p_batch_noalloc {phys}/{guess} n={n} (max {nmax})")
            want, want_it = d.c2p(cons, with_iters=True)
            rc, got, got_it = noalloc(d, cons)
            assert rc == L.AFD_OK, L.load().afd_last_error()
            assert np.array_equal(got, want, equal_nan=True)
            assert np.array_equal(got_it, want_it)
            assert want_it.max() > 0
        # beyond the staging's capacity: refused, nothing written
        big = np.zeros((nmax + 1, nb))
        rc, out, _ = noalloc(d, big)
        assert rc == L.AFD_ERR_INVALID
        assert b"exceeds" in L.load().afd_last_error()
        assert np.isnan(out).all()
        # the largest accepted size still works
        cons = d.p2c(make(nmax, 3))
        rc, got, got_it = noalloc(d, cons)
        assert rc == L.AFD_OK
        want, want_it = d.c2p(cons, with_iters=True)
        assert np.array_equal(got, want, equal_nan=True) and np.array_equal(got_it, want_it)
