"""GPU parity of the weak pseudo-periodic boundary condition (WeakPseudoPeridicMarker,
DmriFemLib.py:601-610; ThetaMethodF/L_wBC1c/2c 58-145; WeakPseudoPeriodic_1c/2c 256-324,386-450)
against the CPU oracle, through the C ABI.

Bars: boundary mass P within 1e-13 of the largest entry; the mirror map as an operator within
1e-13 (the facet chosen for a point on a shared edge may differ, the interpolated value may not);
the mirror offsets exactly; signals within 1e-9 relative of the oracle at Krylov rtol 1e-13.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import oracle as O  # noqa: E402

DELTA, BIGDELTA = 10600.0, 43100.0


def _seq(k):
    f = O.pgse_piecewise(DELTA, BIGDELTA)
    fA, fB = O.gcloud_sequence(f, DELTA + BIGDELTA, k)
    _, FB = O.gcloud_sequence(O.pgse_integral(DELTA, BIGDELTA), DELTA + BIGDELTA, k)
    return fA, fB, FB


def _cases():
    box = dm.box_mesh((0, 0, 0), (10, 8, 6), 6, 5, 4)
    cen = box.xyz[box.tets].mean(axis=1)
    ph = (np.linalg.norm(cen[:, :2] - [5.0, 4.0], axis=1) < 2.5).astype(np.int32)
    cyl = dm.cylinder_mesh(4.0, 8.0, 1500, seed=17, interface_radii=(2.0,))
    return {
        "box_x": (box, None, (1, 0, 0), dict(D=3e-3, kappa=0.0)),
        "box_xy_2c": (box, ph, (1, 1, 0), dict(D=(3e-3, 1.5e-3), kappa=2e-4)),
        "box_xyz": (box, None, (1, 1, 1), dict(D=2e-3, kappa=0.0)),
        "cyl_z_2c": (cyl, cyl.phase, (0, 0, 1), dict(D=3e-3, kappa=1e-5)),
    }


CASES = _cases()


def _pair(name):
    import dmri_native as N

    m, ph, pdir, kw = CASES[name]
    ctx = N.Context(0)
    ctx.set_mesh(m.xyz, m.tets, ph)
    ctx.set_coeffs(D=kw["D"], invT2=1e-16, kappa=kw["kappa"])
    ctx.set_periodic(pdir)
    ctx.assemble()
    S = O.System(m.xyz, m.tets, ph, D=kw["D"], kappa=kw["kappa"], periodic=pdir)
    return ctx, S


@pytest.mark.parametrize("name", list(CASES))
def test_periodic_mass_and_mirror_map_match_oracle(name):
    ctx, S = _pair(name)
    info = ctx.info()
    assert info["ndof"] == S.ndof and info["nnz"] == S.nnz
    assert abs(info["kappa_e"] - S.periodic["kappa_e"]) <= 1e-15 * S.periodic["kappa_e"]
    P = ctx.export_values("p")
    ref = S.mats["p"]
    assert np.abs(P - ref).max() <= 1e-13 * np.abs(ref).max()
    g = ctx.export_periodic()
    o = S.periodic
    order = np.argsort(g["dof"])
    assert np.array_equal(g["dof"][order], o["dof"])
    assert np.array_equal(g["ell"][order], o["ell"])
    rng = np.random.default_rng(3)
    u = rng.normal(size=S.ndof) + 1j * rng.normal(size=S.ndof)

    def apply(e):
        y = np.zeros(len(e["dof"]), dtype=complex)
        for k in range(3):
            ok = e["src"][:, k] >= 0
            y[ok] += e["w"][ok, k] * u[e["src"][ok, k]]
        return y

    assert np.abs(apply(g)[order] - apply(o)).max() <= 1e-13 * np.abs(u).max()


@pytest.mark.parametrize("name,b,gdir", [("box_x", 1000.0, (1, 0, 0)), ("box_xy_2c", 2000.0, (0.6, 0.8, 0.0)),
                                         ("box_xyz", 500.0, (1, 1, 1)), ("cyl_z_2c", 1500.0, (0.3, 0.2, 1.0))])
def test_periodic_signal_matches_oracle(name, b, gdir):
    ctx, S = _pair(name)
    k = 400.0
    fA, fB, FB = _seq(k)
    q = O.q_from_b_pgse(b, DELTA, BIGDELTA)
    ref, _, _ = S.solve(gdir, q, k, fA, fB, FB=FB, rtol=1e-13)
    got, st = ctx.solve(fA, fB, k, q, gdir, rtol=1e-13, F_explicit=FB)
    assert abs(got.real - ref) <= 1e-9 * abs(ref), (got, ref)


def test_periodic_batched_equals_single_and_oracle():
    ctx, S = _pair("box_xy_2c")
    k = 400.0
    fA, fB, FB = _seq(k)
    qs = np.array([O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (0.0, 800.0, 2500.0)])
    gd = np.array([[1.0, 0.0, 0.0], [0.3, -0.9, 0.2], [0.0, 1.0, 1.0]])
    sig, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, F_explicit=FB)
    for j in range(3):
        one, _ = ctx.solve(fA, fB, k, qs[j], gd[j], rtol=1e-13, F_explicit=FB)
        assert abs(sig[j] - one) <= 1e-11 * abs(one)
        ref, _, _ = S.solve(gd[j], qs[j], k, fA, fB, FB=FB, rtol=1e-13)
        assert abs(sig[j].real - ref) <= 1e-9 * abs(ref)


def test_periodic_projection_guess_matches_oracle():
    ctx, S = _pair("box_xy_2c")
    k = 400.0
    fA, fB, FB = _seq(k)
    q = O.q_from_b_pgse(2000.0, DELTA, BIGDELTA)
    ref, _, _ = S.solve((0.6, 0.8, 0.0), q, k, fA, fB, FB=FB, rtol=1e-13)
    got, st = ctx.solve(fA, fB, k, q, (0.6, 0.8, 0.0), rtol=1e-13, F_explicit=FB, guess=0)
    quad, stq = ctx.solve(fA, fB, k, q, (0.6, 0.8, 0.0), rtol=1e-13, F_explicit=FB, guess=3)
    assert abs(got.real - ref) <= 1e-9 * abs(ref)
    assert abs(got - quad) <= 1e-11 * abs(quad)
    assert st["iters_total"] < stq["iters_total"]


def test_all_faces_periodic_keeps_b0_signal_one():
    ctx, _ = _pair("box_xyz")
    fA, fB, FB = _seq(400.0)
    s, _ = ctx.solve(fA, fB, 400.0, 0.0, (1, 0, 0), rtol=1e-14, F_explicit=FB)
    assert abs(s.real - 1.0) < 1e-10


def test_periodic_argument_errors_and_switch_off():
    import dmri_native as N

    ctx, _ = _pair("box_x")
    fA, fB, FB = _seq(400.0)
    q = O.q_from_b_pgse(1000.0, DELTA, BIGDELTA)
    with pytest.raises(N.DmriError):
        ctx.solve(fA, fB, 400.0, q, (1, 0, 0))  # no F_explicit
    with pytest.raises(N.DmriError):
        ctx.adjoint(fA, fB, 400.0, q, (1, 0, 0))
    ctx.set_periodic((0, 0, 0))
    ctx.assemble()
    assert ctx.info()["n_periodic"] == 0 and not ctx.export_values("p").any()
    m, _, _, kw = CASES["box_x"]
    ref, _, _ = O.System(m.xyz, m.tets, None, D=kw["D"]).solve((1, 0, 0), q, 400.0, fA, fB, rtol=1e-13)
    got, _ = ctx.solve(fA, fB, 400.0, q, (1, 0, 0), rtol=1e-13)
    assert abs(got.real - ref) <= 1e-9 * abs(ref)


def test_dmrifemlib_weak_periodic_domain_matches_oracle(capsys):
    import DmriFemLib as L

    mesh = L.BoxMesh(L.Point(0, 0, 0), L.Point(10, 8, 6), 6, 5, 4)
    p = L.MRI_parameters()
    p.bvalue = 1000
    p.delta, p.Delta = DELTA, BIGDELTA
    p.set_gradient_dir(mesh, 1, 0, 0)
    p.T = p.Delta + p.delta
    p.fs_sym = L.sp.Piecewise((1., p.s < p.delta), (0., p.s < p.Delta), (-1., p.s < p.T), (0., True))
    p.Apply()
    dom = L.MyDomain(mesh, p)
    dom.PeriodicDir = [1, 0, 0]
    dom.IsDomainPeriodic = False
    dom.IsDomainMultiple = False
    dom.D0 = 3e-3
    dom.Apply()
    simu = L.MRI_simulation()
    simu.k = 400
    solver = L.KrylovSolver("bicgstab", "jacobi")
    solver.parameters["relative_tolerance"] = 1e-13
    solver.parameters["absolute_tolerance"] = 0.0
    simu.solve(dom, p, solver)
    fA, fB, FB = _seq(400.0)
    m = CASES["box_x"][0]
    ref, _, _ = O.System(m.xyz, m.tets, None, D=3e-3, periodic=(1, 0, 0)).solve(
        (1, 0, 0), p.qvalue, 400.0, fA, fB, FB=FB, rtol=1e-13)
    assert abs(simu.signal_c.real - ref) <= 1e-9 * abs(ref)
