"""Pins the CPU oracle (oracle/dmri_oracle.c) -- CPU only.

Each check is independent of the oracle's own formulas:
  * element matrices against tensor-product Gauss quadrature on collapsed coordinates
    (degree far above the P1 integrands) -- a dropped term, wrong factor or swapped index fails;
  * sparsity against a pure-Python set union;
  * assembled matrices against identities fixed by the mathematics (partition of unity,
    constant null space of K and of the interface coupling, exact P1 reproduction of x);
  * the theta scheme + BiCGStab against a dense direct solve of the same discrete scheme;
  * b = 0 and T2-only closed forms; the reference notebook's printed BoxMesh result.
"""
import itertools
import math

import numpy as np
import pytest

import dmri_mesh as dm
import oracle as O


# ----------------------------------------------------------------------------------------
# independent quadrature
# ----------------------------------------------------------------------------------------
def _gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def tet_quadrature(X, n=7):
    """Points (physical) and weights on tetrahedron X (4x3) via the Duffy map of the cube."""
    u, wu = _gl01(n)
    U, V, W = np.meshgrid(u, u, u, indexing="ij")
    WU, WV, WW = np.meshgrid(wu, wu, wu, indexing="ij")
    xi1 = U
    xi2 = V * (1 - U)
    xi3 = W * (1 - U) * (1 - V)
    jac = (1 - U) ** 2 * (1 - V)
    w = (WU * WV * WW * jac).ravel()
    xi = np.stack([xi1.ravel(), xi2.ravel(), xi3.ravel()], axis=1)
    E = (X[1:] - X[0]).T  # columns e1 e2 e3
    det = np.linalg.det(E)
    pts = X[0] + xi @ E.T
    lam = np.concatenate([1 - xi.sum(axis=1, keepdims=True), xi], axis=1)
    return pts, w * abs(det), lam, E


def tri_quadrature(X, n=7):
    u, wu = _gl01(n)
    U, V = np.meshgrid(u, u, indexing="ij")
    WU, WV = np.meshgrid(wu, wu, indexing="ij")
    xi1, xi2 = U.ravel(), (V * (1 - U)).ravel()
    w = (WU * WV * (1 - U)).ravel()
    area2 = np.linalg.norm(np.cross(X[1] - X[0], X[2] - X[0]))
    lam = np.stack([1 - xi1 - xi2, xi1, xi2], axis=1)
    return w * area2, lam


def _dense_block(S, which, dofs):
    A = S.dense(which)
    return A[np.ix_(dofs, dofs)]


def test_single_tet_element_matrices_match_quadrature():
    rng = np.random.default_rng(11)
    X = rng.uniform(-3, 3, size=(4, 3))
    if np.linalg.det((X[1:] - X[0]).T) < 0:
        X[[1, 2]] = X[[2, 1]]
    A = rng.normal(size=(3, 3))
    Dt = A @ A.T * 1e-3 + np.eye(3) * 1e-3
    S = O.System(X, np.array([[0, 1, 2, 3]]), None, D=0.0, invT2=0.37, Dcell=Dt.reshape(1, 9))
    pts, w, lam, E = tet_quadrature(X)
    grads = np.zeros((4, 3))
    Einv = np.linalg.inv(E)  # rows: grad of lambda_1..3
    grads[1:] = Einv
    grads[0] = -Einv.sum(axis=0)
    vol = w.sum()
    M = np.einsum("q,qa,qb->ab", w, lam, lam)
    K = vol * grads @ Dt @ grads.T
    J = [np.einsum("q,q,qa,qb->ab", w, pts[:, d], lam, lam) for d in range(3)]
    got = {k: S.dense(k) for k in ("m", "k", "r", "jx", "jy", "jz")}
    scale = lambda a: np.abs(a).max()
    assert np.abs(got["m"] - M).max() <= 1e-13 * scale(M)
    assert np.abs(got["r"] - 0.37 * M).max() <= 1e-13 * scale(M)
    assert np.abs(got["k"] - K).max() <= 1e-12 * scale(K)
    for d, k in enumerate(("jx", "jy", "jz")):
        assert np.abs(got[k] - J[d]).max() <= 1e-12 * scale(J[d]), k


def test_interface_facet_mass_matches_quadrature():
    # two tetrahedra sharing face (1,2,3), different phases
    X = np.array([[0.0, 0.0, 0.0], [1.3, 0.1, 0.2], [0.2, 1.1, -0.1], [0.1, 0.3, 0.9], [1.2, 1.0, 1.1]])
    tets = np.array([[0, 1, 2, 3], [4, 1, 2, 3]])
    S = O.System(X, tets, np.array([0, 1]), D=1e-3, kappa=2.5e-5)
    assert S.fv.shape == (1, 3) and tuple(S.fv[0]) == (1, 2, 3)
    w, lam = tri_quadrature(X[[1, 2, 3]])
    Fm = 2.5e-5 * np.einsum("q,qa,qb->ab", w, lam, lam)
    gi = S.dense("gi")
    d0 = [S.dof_of[v] for v in (1, 2, 3)]
    d1 = [S.dof_of[S.nv + v] for v in (1, 2, 3)]
    assert np.abs(gi[np.ix_(d0, d0)] - Fm).max() <= 1e-13 * Fm.max()
    assert np.abs(gi[np.ix_(d1, d1)] - Fm).max() <= 1e-13 * Fm.max()
    assert np.abs(gi[np.ix_(d0, d1)] + Fm).max() <= 1e-13 * Fm.max()
    assert np.abs(gi[np.ix_(d1, d0)] + Fm).max() <= 1e-13 * Fm.max()
    # vertex 0 only touches the phase-0 cell, vertex 4 only the phase-1 cell: dead DOFs dropped
    assert S.dof_of[S.nv + 0] == -1 and S.dof_of[4] == -1
    assert S.ndof == 8


def _tiny_two_comp():
    m = dm.box_mesh((0, 0, 0), (3, 3, 3), 3, 3, 3)
    c = m.xyz[m.tets].mean(axis=1)
    phase = ((np.abs(c - 1.5) < 0.6).all(axis=1)).astype(np.int32)
    return m, phase


def _python_pattern(tets, phase, dof_of, nv, fv):
    rows = {}
    for t, T in enumerate(tets):
        c = 0 if phase is None else phase[t]
        d = [dof_of[c * nv + v] for v in T]
        for a in d:
            rows.setdefault(a, set()).update(d)
    for F in fv:
        d = [dof_of[v] for v in F] + [dof_of[nv + v] for v in F]
        for a in d:
            rows.setdefault(a, set()).update(d)
    return rows


@pytest.mark.parametrize("two", [False, True])
def test_pattern_equals_python_set_union(two):
    if two:
        m, phase = _tiny_two_comp()
    else:
        m, phase = dm.cylinder_mesh(2.0, 2.0, 300, seed=3), None
    S = O.System(m.xyz, m.tets, phase)
    rows = _python_pattern(m.tets, phase, S.dof_of, S.nv, S.fv)
    assert len(rows) == S.ndof
    for r in range(S.ndof):
        assert list(S.col[S.rowptr[r]:S.rowptr[r + 1]]) == sorted(rows[r])
    if two:
        assert S.fv.shape[0] > 0


def test_assembled_identities_two_compartments():
    m, phase = _tiny_two_comp()
    S = O.System(m.xyz, m.tets, phase, D=(2e-3, 1e-3), kappa=1e-5)
    M = S.dense("m")
    assert abs(M.sum() - 27.0) <= 1e-12 * 27.0 and abs(S.volume - 27.0) <= 1e-12
    for k in ("m", "k", "r", "gi", "jx", "jy", "jz"):
        A = S.dense(k)
        assert np.abs(A - A.T).max() <= 1e-15 * max(np.abs(A).max(), 1e-300), k
    one = np.ones(S.ndof)
    assert np.abs(S.dense("k") @ one).max() <= 1e-15 * np.abs(S.dense("k")).max() * 30
    assert np.abs(S.dense("gi") @ one).max() <= 1e-18
    # x is reproduced exactly by P1: Jx 1 = M x_dof
    xyz_dof = np.zeros((S.ndof, 3))
    for c in range(2):
        for v in range(S.nv):
            d = S.dof_of[c * S.nv + v]
            if d >= 0:
                xyz_dof[d] = m.xyz[v]
    for d, k in enumerate(("jx", "jy", "jz")):
        assert np.abs(S.dense(k) @ one - M @ xyz_dof[:, d]).max() <= 1e-13 * np.abs(M).max() * 3
    # compartment volumes: cells of phase 1 are the central unit cube
    n0 = int((S.dof_of[:S.nv] >= 0).sum())
    assert abs(S.w[n0:].sum() - 1.0) <= 1e-13 and abs(S.w[:n0].sum() - 26.0) <= 1e-12


def _dense_theta(S, g, q, dt, fA, fB, theta, invT2=None):
    M, L = S.dense("m"), S.dense(S.l())
    Jg = S.dense(S.jg(g))
    u = np.ones(S.ndof, dtype=complex)
    for a, b in zip(fA, fB):
        A = M / dt + theta * L + 1j * theta * q * a * Jg
        B = M / dt - (1 - theta) * L - 1j * (1 - theta) * q * b * Jg
        u = np.linalg.solve(A, B @ u)
    return (S.w @ u.real) / S.volume, u


@pytest.mark.parametrize("two", [False, True])
def test_theta_bicgstab_matches_dense_direct_solve(two):
    if two:
        m, phase = _tiny_two_comp()
        S = O.System(m.xyz, m.tets, phase, D=(3e-3, 1.5e-3), kappa=2e-4, invT2=(1 / 5e4, 1 / 8e4))
    else:
        m = dm.cylinder_mesh(2.0, 3.0, 400, seed=5)
        S = O.System(m.xyz, m.tets, None, D=3e-3, invT2=1 / 6e4)
    delta, Delta, k = 3000.0, 9000.0, 300.0
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    q = O.q_from_b_pgse(3000.0, delta, Delta)
    g = (0.3, -1.0, 0.5)
    s_dense, u_dense = _dense_theta(S, g, q, k, fA, fB, 0.5)
    s_orc, u_orc, st = S.solve(g, q, k, fA, fB, rtol=1e-14)
    assert abs(s_orc - s_dense) <= 1e-11 * abs(s_dense)
    assert np.abs(u_orc - u_dense).max() <= 1e-10 * np.abs(u_dense).max()
    # theta = 1 (backward Euler) through the same path
    s_d1, _ = _dense_theta(S, g, q, k, fA, fB, 1.0)
    s_o1, _, _ = S.solve(g, q, k, fA, fB, theta=1.0, rtol=1e-14)
    assert abs(s_o1 - s_d1) <= 1e-11 * abs(s_d1)


def test_zero_gradient_and_T2_closed_form():
    m, phase = _tiny_two_comp()
    T2, k, theta = 8e4, 200.0, 0.5
    S = O.System(m.xyz, m.tets, phase, D=3e-3, kappa=1e-5, invT2=1.0 / T2)
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(10600.0, 43100.0), 53700.0, k)
    assert len(fA) == 270
    s, u, _ = S.solve((1, 0, 0), 0.0, k, fA, fB, rtol=1e-15)
    r = (1 / k - (1 - theta) / T2) / (1 / k + theta / T2)
    assert abs(s - r ** len(fA)) <= 1e-12
    # default T2 = 1e16 of MRI_parameters (DmriFemLib.py:807) still enters through 1/T2
    S0 = O.System(m.xyz, m.tets, phase, D=3e-3, kappa=1e-5)
    s0, _, _ = S0.solve((1, 0, 0), 0.0, k, fA, fB, rtol=1e-15)
    # (the 1e-16 relative decay per step sits at the FP64 rounding level, so only its size is checked)
    r0 = (1 / k - (1 - theta) / 1e16) / (1 / k + theta / 1e16)
    assert abs(s0 - r0 ** len(fA)) <= 1e-12 and abs(s0 - 1.0) < 1e-11


def test_sequences_and_q_match_reference_prints():
    # ExplicitImplementation.ipynb:472 prints "q: 1.500e-05" for b=1000, delta=10600, Delta=43100
    q = O.q_from_b_pgse(1000.0, 10600.0, 43100.0)
    assert f"{q:.3e}" == "1.500e-05"
    assert abs(q - 1.4997863706e-5) <= 1e-15
    # ConvergenceTest.ipynb:423 prints q 3.216e-04 and g 1.202 for delta=1000, Delta=10000, b=1000
    q2 = O.q_from_b_pgse(1000.0, 1000.0, 10000.0)
    assert f"{q2:.6e}" == "3.216338e-04" and f"{q2 / O.GAMMA_RATIO * 1e12:.3f}" == "1.202"
    # profile endpoints: sympy Piecewise exclusive at T, comri FT inclusive at Delta+delta
    f_py, f_c = O.pgse_piecewise(1000.0, 10000.0), O.pgse_comri(1000.0, 10000.0)
    assert f_py(11000.0) == 0.0 and f_c(11000.0) == -1.0
    assert f_py(1000.0) == 0.0 and f_c(1000.0) == 0.0 and f_py(0.0) == 1.0
    fA, fB = O.gcloud_sequence(f_py, 11000.0, 10.0)
    assert len(fA) == 1101 and fB[0] == fA[0] == 1.0 and np.all(fB[1:] == fA[:-1])
    fA2, fB2 = O.comri_sequence(f_c, 11000.0, 10.0)
    assert np.array_equal(fA2, fB2) and len(fA2) == 1101
    fA3, _ = O.comri_sequence(f_c, 11000.0, 10.0, until_T=True)
    assert len(fA3) == 1100


def test_free_diffusion_large_box_approaches_exp_minus_bD():
    m = dm.box_mesh((-30, -30, -30), (30, 30, 30), 12, 12, 12)  # h = 5 um, diffusion length ~5 um
    S = O.System(m.xyz, m.tets, None, D=2e-3)
    delta, Delta, k = 1000.0, 5000.0, 50.0
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    s, _, _ = S.solve((1, 0, 0), O.q_from_b_pgse(300.0, delta, Delta), k, fA, fB, rtol=1e-13)
    assert abs(s - math.exp(-300.0 * 2e-3)) < 1.5e-2  # discretisation error of the coarse mesh


@pytest.mark.slow
def test_boxmesh_golden_reference_notebook():
    """ConvergenceTest.ipynb:423: 'Normalized signal: 8.440078e-01, ... hmin: 5.413e-01, hmax:
    5.413e-01, whole_vol: 125.000' for BoxMesh [-2.5,2.5]^3 (n=16: hmin = sqrt(3)*5/16), D=2e-3,
    delta=1000, Delta=10000, b=1000, gdir z, k=10, loop while t < T with lagged f."""
    m = dm.box_mesh((-2.5,) * 3, (2.5,) * 3, 16, 16, 16)
    hmin, hmax = m.hminmax()
    assert f"{hmin:.3e}" == "5.413e-01" and f"{hmax:.3e}" == "5.413e-01"
    S = O.System(m.xyz, m.tets, None, D=2e-3)
    assert f"{S.volume:.3f}" == "125.000"
    fA, fB = O.lagged_sequence_until(O.pgse_piecewise(1000.0, 10000.0), 11000.0, 10.0)
    s, _, _ = S.solve((0, 0, 1), O.q_from_b_pgse(1000.0, 1000.0, 10000.0), 10.0, fA, fB, rtol=1e-14)
    assert f"{s:.6e}" == "8.440078e-01"


def test_oracle_is_thread_count_independent():
    m = dm.cylinder_mesh(3.0, 3.0, 1500, seed=9)
    S = O.System(m.xyz, m.tets, None)
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(3000.0, 9000.0), 12000.0, 400.0)
    q = O.q_from_b_pgse(2000.0, 3000.0, 9000.0)
    n0 = O.num_threads()
    try:
        O.set_threads(1)
        a, _, _ = S.solve((1, 1, 0), q, 400.0, fA, fB)
        O.set_threads(max(2, n0))
        b, _, _ = S.solve((1, 1, 0), q, 400.0, fA, fB)
    finally:
        O.set_threads(n0)
    assert a == b


def test_planted_sign_error_is_detected():
    """A J sign flip in the scheme (a plausible mistake) is caught by the dense check."""
    m = dm.cylinder_mesh(2.0, 3.0, 400, seed=5)
    S = O.System(m.xyz, m.tets, None, D=3e-3)
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(3000.0, 9000.0), 12000.0, 300.0)
    q = O.q_from_b_pgse(3000.0, 3000.0, 9000.0)
    s_dense, _ = _dense_theta(S, (1, 0, 0), q, 300.0, fA, fB, 0.5)
    # planted mistake: explicit side with the implicit sign
    s_bad, _ = _dense_theta(S, (1, 0, 0), q, 300.0, fA, -np.asarray(fB), 0.5)
    s_orc, _, _ = S.solve((1, 0, 0), q, 300.0, fA, fB, rtol=1e-14)
    assert abs(s_orc - s_dense) <= 1e-11 * abs(s_dense)
    assert abs(s_orc - s_bad) > 1e-6


# ----------------------------------------------------------------------------------------
# weak pseudo-periodic boundary condition (WeakPseudoPeridicMarker, DmriFemLib.py:601-610;
# ThetaMethodF/L_wBC1c/2c 58-145; WeakPseudoPeriodic_1c/2c 256-324, 386-450)
# ----------------------------------------------------------------------------------------
def _boundary_facets_py(tets):
    count = {}
    for t, T in enumerate(tets):
        for skip in range(4):
            f = tuple(sorted(int(T[a]) for a in range(4) if a != skip))
            count.setdefault(f, []).append(t)
    return [(f, cells[0]) for f, cells in count.items() if len(cells) == 1]


def _periodic_dense_py(S, xyz, tets, phase, pdir):
    """P (dense) by quadrature of kappa_e*u*v over every boundary facet, and the mirror operator
    Pi (dense, rows = DOFs taking mirrored values) and ell by 3-D point location over all cells."""
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    X = xyz[tets]
    ii, jj = np.triu_indices(4, 1)
    hmin = np.linalg.norm(X[:, ii] - X[:, jj], axis=2).max(axis=1).min()
    ks, eps = 3e-3 / hmin, 1e-2 * hmin
    nv, n = xyz.shape[0], S.ndof
    chi = np.zeros(nv)
    for d in range(3):
        if pdir[d]:
            chi = np.maximum(chi, ((xyz[:, d] < lo[d] + eps) | (xyz[:, d] > hi[d] - eps)).astype(float))
    P = np.zeros((n, n))
    for f, t in _boundary_facets_py(tets):
        c = 0 if phase is None else int(phase[t])
        Xf = xyz[list(f)]
        w, lam = tri_quadrature(Xf, n=4)
        kap = lam @ chi[list(f)]
        dofs = [S.dof_of[c * nv + v] for v in f]
        for a in range(3):
            for b in range(3):
                P[dofs[a], dofs[b]] += ks * np.sum(w * kap * lam[:, a] * lam[:, b])
    Pi = np.zeros((n, n))
    ell = np.zeros((n, 3))
    dof_vert = np.full(n, -1)
    dof_comp = np.full(n, -1)
    for c in range(S.ncomp):
        for v in range(nv):
            i = S.dof_of[c * nv + v]
            if i >= 0:
                dof_vert[i], dof_comp[i] = v, c
    Einv = [np.linalg.inv((x[1:] - x[0]).T) for x in X]
    for i in range(n):
        v, c = dof_vert[i], dof_comp[i]
        x = xyz[v].copy()
        mp, l, hit = x.copy(), np.zeros(3), False
        for d in range(3):
            if not pdir[d]:
                continue
            if abs(x[d] - lo[d]) <= 1e-7:
                mp, l, hit = x.copy(), np.zeros(3), True
                mp[d], l[d] = hi[d], hi[d] - x[d]
            if abs(x[d] - hi[d]) <= 1e-7:
                mp, l, hit = x.copy(), np.zeros(3), True
                mp[d], l[d] = lo[d], lo[d] - x[d]
        if not hit:
            continue
        best, bl, bt = -1e300, None, None
        for t in range(len(tets)):
            xi = Einv[t] @ (mp - X[t][0])
            lam = np.concatenate([[1 - xi.sum()], xi])
            if lam.min() > best:
                best, bl, bt = lam.min(), lam, t
        assert best > -1e-9
        for a in range(4):
            src = S.dof_of[c * nv + tets[bt][a]]
            if src >= 0:
                Pi[i, src] += bl[a]
        ell[i] = l
    return P, Pi, ell


def _periodic_cases():
    m = dm.box_mesh((0, 0, 0), (3, 2, 2), 3, 2, 2)
    c = m.xyz[m.tets].mean(axis=1)
    ph = (c[:, 1] < 1.0).astype(np.int32)
    return [(m, None, (1, 0, 0)), (m, None, (1, 1, 0)), (m, ph, (1, 0, 1)), (m, None, (1, 1, 1))]


@pytest.mark.parametrize("case", range(4))
def test_periodic_boundary_mass_and_mirror_map_match_brute_force(case):
    m, ph, pdir = _periodic_cases()[case]
    S = O.System(m.xyz, m.tets, ph, D=3e-3, periodic=pdir)
    Pd, Pi, ell = _periodic_dense_py(S, m.xyz, m.tets, ph, pdir)
    assert np.abs(S.dense("p") - Pd).max() <= 1e-14 * np.abs(Pd).max()
    per = S.periodic
    Po = np.zeros_like(Pi)
    for e, i in enumerate(per["dof"]):
        for a in range(3):
            if per["src"][e, a] >= 0:
                Po[i, per["src"][e, a]] += per["w"][e, a]
    rows = np.nonzero(np.abs(Pi).sum(axis=1) + np.abs(ell).sum(axis=1))[0]
    assert set(per["dof"].tolist()) == set(rows.tolist())
    rng = np.random.default_rng(case)
    u = rng.normal(size=S.ndof) + 1j * rng.normal(size=S.ndof)
    assert np.abs(Po @ u - Pi @ u).max() <= 1e-13 * np.abs(u).max()
    assert np.abs(per["ell"] - ell[per["dof"]]).max() == 0.0


def test_periodic_theta_scheme_matches_dense_direct_solve():
    m, ph, pdir = _periodic_cases()[2]
    S = O.System(m.xyz, m.tets, ph, D=(3e-3, 1e-3), kappa=1e-4, periodic=pdir)
    Pd, Pi, ell = _periodic_dense_py(S, m.xyz, m.tets, ph, pdir)
    delta, Delta, k = 3000.0, 9000.0, 300.0
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    _, FB = O.gcloud_sequence(O.pgse_integral(delta, Delta), delta + Delta, k)
    q = O.q_from_b_pgse(3000.0, delta, Delta)
    g = np.array([0.8, 0.3, -0.5])
    gn = g / np.linalg.norm(g)
    M, L, Jg = S.dense("m"), S.dense(S.l()), S.dense(S.jg(g))
    u = np.ones(S.ndof, dtype=complex)
    theta = 0.5
    for a, b, F in zip(fA, fB, FB):
        A = M / k + theta * (L + Pd) + 1j * theta * q * a * Jg
        ybc = np.exp(1j * q * F * (ell @ gn)) * (Pi @ u)
        rhs = (M / k - (1 - theta) * L - 1j * (1 - theta) * q * b * Jg) @ u + (1 - theta) * Pd @ ybc
        u = np.linalg.solve(A, rhs)
    s_dense = (S.w @ u.real) / S.volume
    s_orc, u_orc, _ = S.solve(g, q, k, fA, fB, FB=FB, rtol=1e-14)
    assert abs(s_orc - s_dense) <= 1e-11 * abs(s_dense)
    assert np.abs(u_orc - u).max() <= 1e-10 * np.abs(u).max()


def test_periodic_physics_known_answers():
    """All faces periodic: b = 0 keeps S = 1 (P 1 = P Pi 1); gradient along the periodic axis of a
    homogeneous box gives nearly free diffusion exp(-bD), far from the reflecting box's value.
    (With only some faces periodic the b = 0 signal is not 1: boundary facets next to a periodic
    face carry part of kappa_e while u_bc is 0 at their vertices off that face, as in the
    reference's forms.)"""
    delta, Delta, k = 10600.0, 43100.0, 400.0
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    _, FB = O.gcloud_sequence(O.pgse_integral(delta, Delta), delta + Delta, k)
    m = dm.box_mesh((0, 0, 0), (10, 10, 10), 6, 6, 6)
    S = O.System(m.xyz, m.tets, None, D=3e-3, periodic=(1, 1, 1))
    s0, u0, _ = S.solve((1, 0, 0), 0.0, k, fA, fB, FB=FB, rtol=1e-14)
    assert abs(s0 - 1.0) < 1e-10 and np.abs(u0 - 1).max() < 1e-9
    q = O.q_from_b_pgse(500.0, delta, Delta)
    # faces normal to x periodic only: where periodic faces meet, the mirror goes along the last
    # periodic axis (the reference eval's order) and drops the x phase, so (1,1,1) is less free
    S = O.System(m.xyz, m.tets, None, D=3e-3, periodic=(1, 0, 0))
    s_per = S.solve((1, 0, 0), q, k, fA, fB, FB=FB, rtol=1e-13)[0]
    s_neu = O.System(m.xyz, m.tets, None, D=3e-3).solve((1, 0, 0), q, k, fA, fB, rtol=1e-13)[0]
    free = math.exp(-500.0 * 3e-3)
    assert abs(s_per - free) < 0.1 * abs(s_neu - free)


def _notebook_cos_ogse(nperiod, delta, Delta):
    """Profile 3 of ArbitraryTimeSequence.ipynb:380-392 as the notebook builds it (sympy Piecewise,
    t0 = seq_ext = 100), with MRI_parameters' symbolic integrals (DmriFemLib.py:826-832)."""
    import sympy as sp

    s, u = sp.Symbol("s"), sp.Symbol("u")
    t0, seq_ext = 100, 100
    Dd = Delta + delta
    T = Dd + t0 + seq_ext
    omega = 2.0 * nperiod * math.pi / delta
    tau = Dd / 2.0
    fs = sp.Piecewise((0.0, s < t0), (sp.cos(omega * (s - t0)), s <= delta + t0), (0.0, s <= tau + t0),
                      (-sp.cos(omega * (s - t0 - tau)), s <= delta + tau + t0), (0.0, True))
    ifs = sp.integrate(fs.subs(s, u), (u, 0, s))
    i4 = float(sp.integrate(ifs * ifs, (s, 0, T)))
    return s, fs, ifs, i4, T


@pytest.mark.parametrize("nperiod,delta,Delta", [(4, 19900.0, 19900.0), (1, 10000.0, 30000.0)])
def test_cos_ogse_matches_sympy_profile_and_integrals(nperiod, delta, Delta):
    import dmri_sequences as DS

    s, fs, ifs, i4, T = _notebook_cos_ogse(nperiod, delta, Delta)
    f, T_o = O.cos_ogse(nperiod, delta, Delta)
    F = O.cos_ogse_integral(nperiod, delta, Delta)
    fl, T_l = DS.cos_ogse(nperiod, delta, Delta)
    Fl = DS.cos_ogse_integral(nperiod, delta, Delta)
    assert T_o == T and T_l == T
    # every sample of the GCloud loop at dt = 100 us, plus the branch ends (inclusive there)
    tau = 0.5 * (Delta + delta)
    ts = list(np.arange(0.0, T + 100.0, 100.0)) + [100.0, delta + 100.0, tau + 100.0, delta + tau + 100.0,
                                                   99.999, delta + 100.001]
    for t in ts:
        ref_f = float(fs.subs(s, t))
        ref_F = float(ifs.subs(s, t))
        assert abs(f(t) - ref_f) <= 1e-12 and abs(fl(t) - ref_f) <= 1e-12, t
        assert abs(F(t) - ref_F) <= 1e-9 * (1 + abs(ref_F)) and abs(Fl(t) - ref_F) <= 1e-9 * (1 + abs(ref_F)), t
    # q = sqrt(b / int_0^T F^2): the closed form against sympy's integral
    assert abs(O.int_F2_cos_ogse(nperiod, delta, Delta) - i4) <= 1e-10 * i4
    assert abs(DS.cos_ogse_int_F2(nperiod, delta, Delta) - i4) <= 1e-10 * i4
    # the 4-period workload: T = 40 ms and 401 theta steps at dt = 100 us (t < T + k)
    if nperiod == 4:
        fA, fB = O.gcloud_sequence(f, T, 100.0)
        assert T == 40000.0 and len(fA) == 401 and fA[0] == 0.0 and fA[1] == 1.0
        gA, gB, GB = DS.gcloud_loop(fl, T, 100.0, Fl)
        assert np.array_equal(gA, fA) and np.array_equal(gB, fB) and GB[0] == 0.0


def test_waveform_module_pgse_matches_oracle():
    import dmri_sequences as DS

    f_o, f_l = O.pgse_piecewise(10600.0, 43100.0), DS.pgse(10600.0, 43100.0)
    F_o, F_l = O.pgse_integral(10600.0, 43100.0), DS.pgse_integral(10600.0, 43100.0)
    for t in np.arange(0.0, 54000.0, 50.0):
        assert f_o(t) == f_l(t) and F_o(t) == F_l(t)
    assert abs(DS.q_from_b(1000.0, DS.pgse_int_F2(10600.0, 43100.0)) - O.q_from_b_pgse(1000.0, 10600.0, 43100.0)) \
        <= 1e-18


def _cell_gradients(X):
    E = (X[1:] - X[0]).T
    Einv = np.linalg.inv(E)
    G = np.zeros((4, 3))
    G[1:] = Einv
    G[0] = -Einv.sum(axis=0)
    return G


def test_cell_fields_nonsymmetric_tensor_and_facet_kappa_match_quadrature():
    """D_cell (non-symmetric tensors), invT2_cell and kappa_facet branches of orc_assemble against a
    per-cell quadrature assembly in Python: K_ab = int grad(phi_a) . D_t grad(phi_b) (row a the test
    function, inner(D*grad(u), grad(v)) of FuncF_wBC), R = sum_t invT2_t M_t, I per facet."""
    m, phase = _tiny_two_comp()
    rng = np.random.default_rng(5)
    nc = m.nc
    Dcell = rng.normal(size=(nc, 3, 3)) * 1e-3 + np.eye(3)[None] * 3e-3
    assert np.abs(Dcell - Dcell.transpose(0, 2, 1)).max() > 1e-4  # genuinely non-symmetric
    it2 = rng.uniform(1e-5, 1e-4, size=nc)
    S0 = O.System(m.xyz, m.tets, phase)
    kf = rng.uniform(1e-6, 1e-4, size=S0.fv.shape[0])
    S = O.System(m.xyz, m.tets, phase, Dcell=Dcell.reshape(nc, 9), invT2cell=it2, kappa_facet=kf)
    n = S.ndof
    K = np.zeros((n, n))
    R = np.zeros((n, n))
    for t in range(nc):
        X = m.xyz[m.tets[t]]
        pts, w, lam, _ = tet_quadrature(X, n=3)
        Mt = np.einsum("q,qa,qb->ab", w, lam, lam)
        G = _cell_gradients(X)
        Kt = w.sum() * G @ Dcell[t] @ G.T
        d = [S.dof_of[phase[t] * S.nv + v] for v in m.tets[t]]
        K[np.ix_(d, d)] += Kt
        R[np.ix_(d, d)] += it2[t] * Mt
    I = np.zeros((n, n))
    for f, F in enumerate(S.fv):
        wq, lamq = tri_quadrature(m.xyz[F], n=3)
        Fm = kf[f] * np.einsum("q,qa,qb->ab", wq, lamq, lamq)
        d0 = [S.dof_of[v] for v in F]
        d1 = [S.dof_of[S.nv + v] for v in F]
        I[np.ix_(d0, d0)] += Fm
        I[np.ix_(d1, d1)] += Fm
        I[np.ix_(d0, d1)] -= Fm
        I[np.ix_(d1, d0)] -= Fm
    assert np.abs(S.dense("k") - K).max() <= 1e-13 * np.abs(K).max()
    assert np.abs(S.dense("k") - S.dense("k").T).max() > 1e-6  # the transpose would fail the line above
    assert np.abs(S.dense("r") - R).max() <= 1e-13 * np.abs(R).max()
    assert np.abs(S.dense("gi") - I).max() <= 1e-13 * np.abs(I).max()


def _brute_interface_facets(tets, phase):
    """Every face of every cell keyed by its sorted vertex triple; a face shared by two cells of
    different phase is an interface facet (abs(jump(phase)) dS in icondition_wBC)."""
    faces = {}
    for t, T in enumerate(tets):
        for skip in range(4):
            key = tuple(sorted(int(T[i]) for i in range(4) if i != skip))
            faces.setdefault(key, []).append(t)
    out = [k for k, cells in faces.items() if len(cells) == 2 and phase[cells[0]] != phase[cells[1]]]
    assert all(len(c) <= 2 for c in faces.values())
    return np.array(sorted(out), dtype=np.int32).reshape(-1, 3)


@pytest.mark.parametrize("maker", ["box", "cyl"])
def test_interface_facets_match_brute_force_enumeration(maker):
    if maker == "box":
        m, phase = _tiny_two_comp()
    else:
        m = dm.cylinder_mesh(6.0, 3.0, 1500, seed=17, interface_radii=(3.0,))
        phase = m.phase
    S = O.System(m.xyz, m.tets, phase)
    ref = _brute_interface_facets(m.tets, phase)
    assert ref.shape[0] > 10
    assert np.array_equal(S.fv, ref)


def test_strong_periodic_direction_terms_match_quadrature_and_integration_by_parts():
    """orc_assemble_sbc: C_d = int 2 D (d phi_j/d x_d) phi_i against quadrature on one tet; over a
    two-compartment mesh, C_d + C_d^T = 2 B_d (int 2D g.grad(phi_i phi_j) = boundary integral of D g.n
    over each compartment's outer and interface facets), which pins the facet normals' signs."""
    X = np.array([[0.1, 0.0, 0.2], [1.4, 0.2, 0.1], [0.3, 1.2, -0.2], [0.2, 0.4, 1.1]])
    SP = O.StrongPeriodicSystem(X, np.array([[0, 1, 2, 3]]), None, D=1.7e-3, pdir=(0, 0, 0))
    pts, w, lam, E = tet_quadrature(X)
    G = _cell_gradients(X)
    S = SP.base
    for d, k in enumerate(("cx", "cy", "cz")):
        ref = np.einsum("q,qa->a", w, lam)[:, None] * (2 * 1.7e-3 * G[:, d])[None, :]  # [i, j]
        assert np.abs(S.dense(SP.ext[k]) - ref).max() <= 1e-13 * np.abs(ref).max()
    m = dm.cylinder_mesh(5.0, 3.0, 1200, seed=21, interface_radii=(2.5,))
    SP2 = O.StrongPeriodicSystem(m.xyz, m.tets, m.phase, D=(2e-3, 1e-3), pdir=(0, 0, 0))
    S2 = SP2.base
    for a in "xyz":
        C = S2.dense(SP2.ext["c" + a])
        B = S2.dense(SP2.ext["b" + a])
        assert np.abs(B).max() > 0
        assert np.abs(C + C.T - 2 * B).max() <= 1e-13 * np.abs(B).max()


def test_periodic_masters_brute_force_on_a_box():
    m = dm.box_mesh((0.0, -1.0, 0.0), (4.0, 2.0, 3.0), 4, 3, 3)
    master = O.periodic_masters(m.xyz, (1, 1, 0), 1e-3)
    for v, x in enumerate(m.xyz):
        t = x.copy()
        if abs(x[0] - 4.0) < 1e-9:
            t[0] = 0.0
        if abs(x[1] - 2.0) < 1e-9:
            t[1] = -1.0
        w_ = np.nonzero(np.all(np.abs(m.xyz - t) < 1e-9, axis=1))[0]
        assert w_.shape == (1,) and master[v] == w_[0]
    assert master[np.argmin(np.abs(m.xyz - [4.0, 2.0, 3.0]).sum(axis=1))] == \
        np.argmin(np.abs(m.xyz - [0.0, -1.0, 3.0]).sum(axis=1))


def test_strong_periodic_fully_periodic_box_closed_form():
    m = dm.box_mesh((0.0, 0.0, 0.0), (3.0, 2.0, 2.0), 6, 4, 4)
    SP = O.StrongPeriodicSystem(m.xyz, m.tets, None, D=2e-3, invT2=1 / 5e4, pdir=(1, 1, 1))
    k, theta, q, D, iT2 = 100.0, 0.5, 3e-4, 2e-3, 1 / 5e4
    F = O.pgse_integral(1000.0, 4000.0)
    FA, FB = O.gcloud_sequence(F, 5000.0, k)
    s, u = SP.solve((0.3, 1.0, 0.4), q, k, FA, FB)
    r = np.prod([(1 / k - theta * (q * q * D * fb * fb + iT2)) / (1 / k + theta * (q * q * D * fa * fa + iT2))
                 for fa, fb in zip(FA, FB)])
    assert abs(s - r) <= 1e-12 * r and np.abs(u - u[0]).max() <= 1e-12
