"""dmri_get_solution_noalloc (the allocation-free variant of dmri_get_solution, staging through a
file-scope device buffer) returns exactly what dmri_get_solution returns: two calls in a row through
the C ABI, the second on a smaller mesh, each compared bit for bit; then a third call on the first
context again (the buffer keeps the last call's contents, which must not leak into the result)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import oracle as O  # noqa: E402

DELTA, BIGDELTA, K = 10600.0, 43100.0, 800.0


def _solved(mesh, sig=3):
    import dmri_native as N

    fa, fb = O.gcloud_sequence(O.pgse_piecewise(DELTA, BIGDELTA), DELTA + BIGDELTA, K)
    c = N.Context(0)
    c.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    c.set_coeffs(D=(3e-3, 2e-3), kappa=1e-5)
    c.assemble()
    qs = [O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (500.0, 1000.0, 2000.0, 3000.0)]
    c.solve_batched(fa, fb, K, np.array(qs[:sig + 1]), np.array([[0.3, 1.0, 0.2]] * (sig + 1)), rtol=1e-12)
    return c


def _both(c, sig):
    import dmri_native as N

    n = c.info()["ndof"]
    u_ref = np.full(2 * n, np.nan)
    u_na = np.full(2 * n, np.nan)
    N._check(N.LIB.dmri_get_solution(c._h, sig, u_ref.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    N._check(N.LIB.dmri_get_solution_noalloc(c._h, sig, u_na.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return u_ref, u_na


def test_noalloc_solution_matches_allocating_getter():
    import dmri_native as N

    print("[no_alloc] large mesh", flush=True)
    big = _solved(dm.cylinder_mesh(8.0, 4.0, 6000, seed=21, interface_radii=(4.0,)))
    print("[no_alloc] small mesh", flush=True)
    small = _solved(dm.cylinder_mesh(5.0, 3.0, 1500, seed=22, interface_radii=(2.5,)), sig=1)
    assert small.info()["ndof"] < big.info()["ndof"]
    for c, sig in ((big, 3), (small, 1), (small, 0), (big, 2)):
        u_ref, u_na = _both(c, sig)
        assert np.all(np.isfinite(u_ref))
        np.testing.assert_array_equal(u_na, u_ref)
    # argument errors come back as error codes, as for dmri_get_solution
    u = np.zeros(2 * small.info()["ndof"])
    up = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert N.LIB.dmri_get_solution_noalloc(small._h, 2, up) == -1
    assert N.LIB.dmri_get_solution_noalloc(small._h, -1, up) == -1
    fresh = N.Context(0)
    assert N.LIB.dmri_get_solution_noalloc(fresh._h, 0, up) == -3
