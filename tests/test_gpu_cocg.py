"""The COCG option (krylov = 2) converges to the same discrete solution as BiCGStab: signals within
1e-9 of the oracle at rtol 1e-13 on one- and two-compartment meshes, with the real-part, whole and
complex-diagonal polynomial, every initial guess, batched with per-signal directions (the folded
per-direction kernels), with the weak periodic term, in the resumable batch, under hipGraph replay
(bit for bit against plain launches), and in the adjoint (gradients against BiCGStab's); repeated
solves are bit for bit identical.  The strong periodic form, whose operator is not complex
symmetric, is refused."""
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
    "box": (dm.box_mesh((0.0, 0.0, 0.0), (12.0, 6.0, 6.0), 10, 5, 5), dict(D=2e-3)),
}


def _ctx(name, periodic=None):
    import dmri_native as N

    m, kw = MESHES[name]
    c = N.Context(0)
    c.set_mesh(m.xyz, m.tets, m.phase)
    c.set_coeffs(**kw)
    if periodic:
        c.set_periodic(periodic)
    c.assemble()
    return c


def _oracle(name, g, q):
    m, kw = MESHES[name]
    S = O.System(m.xyz, m.tets, m.phase, **kw)
    return S.solve(g, q, K, FA, FB, rtol=1e-13)[0]


@pytest.mark.parametrize("name,b,precond,guess", [("cyl1", 2000.0, 0, 0), ("cyl2", 2000.0, 0, 0),
                                                  ("cyl2", 4000.0, 2, 0), ("cyl2", 1000.0, 0, 3),
                                                  ("box", 3000.0, 0, 1), ("cyl1", 500.0, 0, 2)])
def test_cocg_signal_matches_oracle_and_bicgstab(name, b, precond, guess):
    print(f"[cocg] {name} b={b} precond={precond} guess={guess}", flush=True)
    q = O.q_from_b_pgse(b, DELTA, BIGDELTA)
    g = (0.3, 1.0, 0.2)
    c = _ctx(name)
    s_c, st_c = c.solve(FA, FB, K, q, g, rtol=1e-13, krylov="cocg", precond=precond, guess=guess)
    s_b, st_b = c.solve(FA, FB, K, q, g, rtol=1e-13, precond=precond, guess=guess)
    ref = _oracle(name, g, q)
    print(f"  cocg {s_c.real:.14e} bicgstab {s_b.real:.14e} oracle {ref:.14e}; iterations per step cocg "
          f"{st_c['iters_total'] / len(FA):.1f} bicgstab {st_b['iters_total'] / len(FA):.1f}; relres "
          f"{st_c['relres_max']:.2e}", flush=True)
    assert st_c["relres_max"] <= 1e-13 and st_c["restarts"] == 0
    assert abs(s_c.real - ref) <= 1e-9 * abs(ref)
    assert abs(s_c - s_b) <= 1e-10 * abs(ref)


def test_cocg_batched_per_direction_matches_single():
    print("[cocg] batched per-signal directions", flush=True)
    c = _ctx("cyl2")
    qs = np.array([O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (500.0, 1000.0, 2000.0, 4000.0, 3000.0, 1500.0)])
    gd = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0.2, 0.5, 1.0], [-0.4, 0.1, 0.7]], dtype=float)
    sig, st = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13, krylov="cocg")
    ref, _ = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13)
    assert st["relres_max"] <= 1e-13
    assert np.abs(sig - ref).max() <= 1e-10 * np.abs(ref).max()
    for lanes in (1, 2):  # both SpMV mappings that run COCG
        for mapping in (1, 2):
            s2, _ = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13, krylov="cocg", mapping=mapping, lanes=lanes)
            assert np.abs(s2 - ref).max() <= 1e-10 * np.abs(ref).max(), (mapping, lanes)
    # a mapping without COCG kernels is replaced by the staged one
    s4, _ = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13, krylov="cocg", mapping=4, lanes=1)
    assert np.abs(s4 - ref).max() <= 1e-10 * np.abs(ref).max()


def test_cocg_weak_periodic_matches_oracle():
    print("[cocg] weak periodic box", flush=True)
    m, kw = MESHES["box"]
    q = O.q_from_b_pgse(2000.0, DELTA, BIGDELTA)
    FBi = O.gcloud_sequence(O.pgse_integral(DELTA, BIGDELTA), DELTA + BIGDELTA, K)[1]
    c = _ctx("box", periodic=(1, 0, 0))
    s_c, _ = c.solve(FA, FB, K, q, (1.0, 0.0, 0.0), rtol=1e-13, F_explicit=FBi, krylov="cocg")
    s_b, _ = c.solve(FA, FB, K, q, (1.0, 0.0, 0.0), rtol=1e-13, F_explicit=FBi)
    S = O.System(m.xyz, m.tets, None, periodic=(1, 0, 0), **kw)
    ref = S.solve((1.0, 0.0, 0.0), q, K, FA, FB, rtol=1e-13, FB=FBi)[0]
    assert abs(s_c.real - ref) <= 1e-9 * abs(ref)
    assert abs(s_c - s_b) <= 1e-10 * abs(ref)


def test_cocg_resumable_batch_and_graphs_bit_for_bit():
    print("[cocg] resumable batch, graph replay", flush=True)
    c = _ctx("cyl2")
    qs = np.array([O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (1000.0, 2000.0, 4000.0)])
    s0, st0 = c.solve_batched(FA, FB, K, qs, (0, 0, 1), rtol=1e-13, krylov="cocg", snap_every=1)
    s1, st1 = c.solve_batched(FA, FB, K, qs, (0, 0, 1), rtol=1e-13, krylov="cocg", snap_every=1, graphs=1)
    assert st1["graph_launches"] > 0 and st1["iters_total"] == st0["iters_total"]
    assert np.array_equal(st1["snapshots"], st0["snapshots"]) and np.array_equal(s1, s0)
    with c.batch(FA, FB, K, qs, (0, 0, 1), rtol=1e-13, krylov="cocg") as b:
        b.advance(5)
        b.advance(len(FA) - 5)
        sb, done = b.signals()
    assert done == len(FA) and np.array_equal(sb, s0)


def test_cocg_per_direction_resumable_and_graphs_bit_for_bit():
    """Per-signal directions (the p-step folded into the q = A p SpMV): graph replay and the
    resumable batch give the one-call solve's states bit for bit."""
    print("[cocg] per-direction batch: graph replay, resumable batch", flush=True)
    c = _ctx("cyl2")
    qs = np.array([O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (1000.0, 2000.0, 4000.0, 3000.0)])
    gd = np.array([[1, 0, 0], [0, 1, 0], [0.2, 0.5, 1.0], [-0.4, 0.1, 0.7]], dtype=float)
    s0, st0 = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13, krylov="cocg", snap_every=3)
    s1, st1 = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13, krylov="cocg", snap_every=3, graphs=1)
    assert st1["graph_launches"] > 0 and st1["iters_total"] == st0["iters_total"]
    assert np.array_equal(st1["snapshots"], st0["snapshots"]) and np.array_equal(s1, s0)
    with c.batch(FA, FB, K, qs, gd, rtol=1e-13, krylov="cocg") as b:
        b.advance(4)
        b.advance(len(FA) - 4)
        sb, done = b.signals()
    assert done == len(FA) and np.array_equal(sb, s0)
    ref, _ = c.solve_batched(FA, FB, K, qs, gd, rtol=1e-13)  # BiCGStab
    assert np.abs(s0 - ref).max() <= 1e-10 * np.abs(ref).max()


def test_cocg_adjoint_gradients_match_bicgstab():
    print("[cocg] adjoint", flush=True)
    c = _ctx("cyl2")
    q = O.q_from_b_pgse(2000.0, DELTA, BIGDELTA)
    g = (0.0, 1.0, 0.3)
    s_c, g_c, _ = c.adjoint(FA, FB, K, q, g, rtol=1e-14, krylov="cocg", checkpoint_stride=4)
    s_b, g_b, _ = c.adjoint(FA, FB, K, q, g, rtol=1e-14)
    assert abs(s_c - s_b) <= 1e-11 * abs(s_b)
    for key in ("D", "invT2", "f_implicit", "f_explicit"):
        assert np.abs(g_c[key] - g_b[key]).max() <= 1e-8 * np.abs(g_b[key]).max(), key
    for key in ("kappa", "q"):
        assert abs(g_c[key] - g_b[key]) <= 1e-8 * abs(g_b[key]), key


@pytest.mark.parametrize("precond", [0, 2])
def test_cocg_repeated_solves_are_bit_identical(precond):
    """One signal on a small mesh (the update of r folded into the preconditioner's SpMV; with the
    complex diagonal, the update kernel): the p-kernel's blocks reduce the partial sums themselves,
    and every block must agree on the signals that stop (and apply their pending update of x)
    whatever the order in which the blocks run; three identical solves and a graph-replayed one
    give identical states."""
    print(f"[cocg] repeated solves, precond {precond}", flush=True)
    c = _ctx("cyl2")
    q = O.q_from_b_pgse(4000.0, DELTA, BIGDELTA)
    runs = [c.solve(FA, FB, K, q, (0.3, 1.0, 0.2), rtol=1e-13, krylov="cocg", precond=precond, snap_every=9,
                    graphs=graphs) for graphs in (0, 0, 0, 1)]
    assert runs[3][1]["graph_launches"] > 0  # the same iterations replayed from captured graphs
    for s_k, st_k in runs[1:]:
        assert s_k == runs[0][0] and st_k["iters_total"] == runs[0][1]["iters_total"]
        assert np.array_equal(st_k["snapshots"], runs[0][1]["snapshots"])


def test_cocg_refused_where_the_operator_is_not_complex_symmetric():
    import dmri_native as N

    c = _ctx("box")
    c.set_periodic_strong((1, 0, 0))
    c.assemble()
    q = O.q_from_b_pgse(1000.0, DELTA, BIGDELTA)
    FAi, FBi = O.gcloud_sequence(O.pgse_integral(DELTA, BIGDELTA), DELTA + BIGDELTA, K)
    c.solve(FA, FB, K, q, (1.0, 0.0, 0.0), F_explicit=FBi, F_implicit=FAi)  # BiCGStab runs it
    with pytest.raises(N.DmriError, match="complex-symmetric"):
        c.solve(FA, FB, K, q, (1.0, 0.0, 0.0), F_explicit=FBi, F_implicit=FAi, krylov="cocg")
    with pytest.raises(N.DmriError, match="polynomial preconditioner"):
        _ctx("cyl1").solve(FA, FB, K, q, (1.0, 0.0, 0.0), krylov="cocg", precond=1)
