"""The restarted GMRES option (KrylovSolver("gmres"), comri/one-comp/fenics-cpp/main.cpp:180-183)
converges to the same discrete solution as BiCGStab: signals within 1e-9 of the oracle at rtol
1e-13, with and without the polynomial preconditioner, with restarts forced by a short cycle,
batched with per-signal directions, and through the DmriFemLib drop-in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import oracle as O  # noqa: E402

DELTA, BIGDELTA, K = 10600.0, 43100.0, 800.0
FA, FB = O.gcloud_sequence(O.pgse_piecewise(DELTA, BIGDELTA), DELTA + BIGDELTA, K)
MESHES = {
    "cyl1": (dm.cylinder_mesh(5.0, 8.0, 2500, seed=1000), dict(D=3e-3)),
    "cyl2": (dm.cylinder_mesh(8.0, 4.0, 3000, seed=1001, interface_radii=(4.0,)), dict(D=(3e-3, 2e-3), kappa=1e-5)),
}


def _ctx(name):
    import dmri_native as N

    m, kw = MESHES[name]
    c = N.Context(0)
    c.set_mesh(m.xyz, m.tets, m.phase)
    c.set_coeffs(**kw)
    c.assemble()
    return c


def _oracle(name, g, q):
    m, kw = MESHES[name]
    S = O.System(m.xyz, m.tets, m.phase, **kw)
    return S.solve(g, q, K, FA, FB, rtol=1e-13)[0]


@pytest.mark.parametrize("name,precond,restart", [("cyl1", 0, 0), ("cyl2", 0, 0), ("cyl2", 1, 0), ("cyl2", 0, 5),
                                                  ("cyl1", 1, 3)])
def test_gmres_signal_matches_oracle(name, precond, restart):
    print(f"[gmres] {name} precond={precond} restart={restart}", flush=True)
    q = O.q_from_b_pgse(2000.0, DELTA, BIGDELTA)
    g = (0.3, 1.0, 0.2)
    c = _ctx(name)
    s_g, st = c.solve(FA, FB, K, q, g, rtol=1e-13, krylov="gmres", restart=restart, precond=precond)
    s_b, _ = c.solve(FA, FB, K, q, g, rtol=1e-13, precond=precond)
    ref = _oracle(name, g, q)
    print(f"  gmres {s_g.real:.14e} bicgstab {s_b.real:.14e} oracle {ref:.14e} arnoldi/step "
          f"{st['iters_total'] / len(FA):.1f} relres {st['relres_max']:.2e}", flush=True)
    assert st["relres_max"] <= 1e-13
    assert abs(s_g.real - ref) <= 1e-9 * abs(ref)
    assert abs(s_g.real - s_b.real) <= 1e-10 * abs(ref)


def test_gmres_batched_per_direction_matches_single():
    print("[gmres] batched", flush=True)
    c = _ctx("cyl2")
    qs = np.array([O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (500.0, 1000.0, 2000.0, 4000.0, 3000.0)])
    gd = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0.2, 0.5, 1.0]], dtype=float)
    sig, _ = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13, krylov="gmres", restart=8)
    for j in (0, 3, 4):
        s1, _ = c.solve(FA, FB, K, qs[j], gd[j], rtol=1e-13)
        assert abs(sig[j] - s1) <= 1e-10 * abs(s1)


def test_dmrifemlib_gmres_solver():
    import DmriFemLib as L

    mesh = L.BoxMesh(L.Point(-2.5, -2.5, -2.5), L.Point(2.5, 2.5, 2.5), 5, 5, 5)
    p = L.MRI_parameters()
    p.bvalue = 1000
    p.delta, p.Delta = 1000.0, 10000.0
    p.set_gradient_dir(mesh, 0, 0, 1)
    p.T = p.Delta + p.delta
    p.fs_sym = L.sp.Piecewise((1., p.s < p.delta), (0., p.s < p.Delta), (-1., p.s < p.T), (0., True))
    p.Apply()
    dom = L.MyDomain(mesh, p)
    dom.D0 = 2e-3
    dom.Apply()
    simu = L.MRI_simulation()
    simu.k = 100
    solver = L.KrylovSolver("gmres")
    solver.parameters["relative_tolerance"] = 1e-13
    solver.parameters["absolute_tolerance"] = 0.0
    simu.solve(dom, p, solver)
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(1000.0, 10000.0), 11000.0, 100.0)
    ref, _, _ = O.System(mesh.xyz, mesh.tets, None, D=2e-3).solve((0, 0, 1), p.qvalue, 100.0, fA, fB, rtol=1e-13)
    assert abs(simu.signal_c.real - ref) <= 1e-9 * abs(ref)
