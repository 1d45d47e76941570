"""rp_pes_potforce_noalloc -- the entry point that allocates nothing -- against
rp_pes_potforce, through the C ABI.

Both run the same potential kernel, so the results must be bit for bit equal: two calls in a
row, the second with fewer beads (the staging array keeps the first call's contents), with
and without the gradient, for the 27-dof EVB and the 54-dof water cluster; then a call past
the staging limit, which must fail with RP_ERR_ARG before any work."""
import ctypes as C

import numpy as np
import pytest

from helpers import rand_chain
from rptunnel import Pes, _lib as L, specs

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20  # doubles of staging per GPU


def _call(name, pes, x, want_grad=True):
    D, nb = x.shape
    xs = np.ascontiguousarray(x)
    g = np.full((D, nb), np.nan) if want_grad else None
    e = np.full(nb, np.nan)
    rc = getattr(L.lib(), name)(pes.h, nb, L.ptr(xs), L.ptr(g), L.ptr(e))
    return rc, g, e


@pytest.mark.parametrize("name", ["evb", "water6"])
def test_noalloc_matches_potforce(name):
    sp = specs.evb() if name == "evb" else specs.water_cluster(nmol=6)
    pes = Pes(sp)
    pes.set_V0(0.0013)
    for nb, seed in ((4000, 1), (37, 2)):  # the second call is smaller
        x = rand_chain(sp, nb, seed, scale=0.05)
        rc0, g0, e0 = _call("rp_pes_potforce", pes, x)
        rc1, g1, e1 = _call("rp_pes_potforce_noalloc", pes, x)
        assert rc0 == rc1 == L.RP_OK, L.last_error()
        assert np.array_equal(g0, g1) and np.array_equal(e0, e1), nb
        rc2, _, e2 = _call("rp_pes_potforce_noalloc", pes, x, want_grad=False)
        assert rc2 == L.RP_OK and np.array_equal(e2, e0)


def test_noalloc_rejects_more_than_its_staging():
    sp = specs.evb()
    pes = Pes(sp)
    nb = LIMIT // (2 * sp.ndof + 1) + 1
    x = np.zeros((sp.ndof, nb))
    rc, _, _ = _call("rp_pes_potforce_noalloc", pes, x)
    assert rc == L.RP_ERR_ARG
    assert "limit" in L.last_error()
    # the largest call that fits still works, and agrees
    x = rand_chain(sp, nb - 1, 3, scale=0.05)
    rc0, g0, e0 = _call("rp_pes_potforce", pes, x)
    rc1, g1, e1 = _call("rp_pes_potforce_noalloc", pes, x)
    assert rc0 == rc1 == L.RP_OK
    assert np.array_equal(g0, g1) and np.array_equal(e0, e1)
