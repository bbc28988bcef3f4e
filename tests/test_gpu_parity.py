"""GPU parity: libdmri (HIP, through the C ABI) against the CPU oracle.

Bars: CSR pattern and DOF map bit-exact; assembled values within 1e-13 relative to the
largest entry of their row; signals within 1e-9 relative of the oracle at Krylov rtol 1e-13
(the parity bar is 1e-8); the BoxMesh golden of the reference notebook to its 7 printed digits.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import oracle as O  # noqa: E402


def _native():
    import dmri_native

    return dmri_native


def _meshes():
    out = {
        "box4": (dm.box_mesh((-2.5,) * 3, (2.5,) * 3, 4, 4, 4), dict(D=2e-3, kappa=0.0)),
        "cyl1": (dm.cylinder_mesh(5.0, 10.0, 2500, seed=1000), dict(D=3e-3, kappa=0.0)),
        "cyl2": (dm.cylinder_mesh(8.0, 6.0, 4000, seed=1001, interface_radii=(4.0,)), dict(D=3e-3, kappa=1e-5)),
    }
    return out


MESHES = _meshes()


def _gpu_system(mesh, D, kappa, invT2=1e-16):
    N = _native()
    ctx = N.Context(0)
    ctx.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    ctx.set_coeffs(D=D, invT2=invT2, kappa=kappa)
    ctx.assemble()
    return ctx


@pytest.mark.parametrize("name", list(MESHES))
def test_pattern_bit_exact(name):
    mesh, kw = MESHES[name]
    print(f"[pattern] {name} nv={mesh.nv} nc={mesh.nc}", flush=True)
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    info = ctx.info()
    assert info["ndof"] == S.ndof and info["nnz"] == S.nnz
    assert info["n_iface"] == S.fv.shape[0]
    rowptr, col, dof_of = ctx.export_pattern()
    assert np.array_equal(rowptr, S.rowptr)
    assert np.array_equal(col, S.col)
    assert np.array_equal(dof_of, S.dof_of)
    fv = ctx.export_facets()
    assert np.array_equal(fv[np.lexsort(fv.T[::-1])], S.fv)


@pytest.mark.parametrize("name", list(MESHES))
def test_values_match_oracle(name):
    mesh, kw = MESHES[name]
    print(f"[values] {name}", flush=True)
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"], invT2=1.0 / 80000.0)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"], invT2=1.0 / 80000.0)
    rows = np.repeat(np.arange(S.ndof), np.diff(S.rowptr))
    for k in ("m", "k", "r", "gi", "jx", "jy", "jz"):
        ref = S.mats[k]
        got = ctx.export_values(k)
        rowmax = np.maximum.reduceat(np.abs(ref), S.rowptr[:-1])[rows] if ref.size else ref
        err = np.abs(got - ref)
        assert np.all(err <= 1e-13 * np.maximum(rowmax, 1e-300)), (k, float(err.max()))
    assert abs(ctx.info()["volume"] - S.volume) <= 1e-13 * S.volume


def _pgse_gcloud(delta, Delta, k):
    f = O.pgse_piecewise(delta, Delta)
    return O.gcloud_sequence(f, delta + Delta, k)


@pytest.mark.parametrize("name,b,g", [("cyl1", 1000.0, (1, 0, 0)), ("cyl2", 1000.0, (0, 1, 0)),
                                      ("cyl2", 4000.0, (1, 1, 0)), ("box4", 2000.0, (0, 0, 1))])
def test_signal_matches_oracle(name, b, g):
    mesh, kw = MESHES[name]
    print(f"[signal] {name} b={b} g={g}", flush=True)
    delta, Delta, k = 10600.0, 43100.0, 200.0
    q = O.q_from_b_pgse(b, delta, Delta)
    fA, fB = _pgse_gcloud(delta, Delta, k)
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    s_ref, _, _ = S.solve(g, q, k, fA, fB, rtol=1e-13)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    s_gpu, st = ctx.solve(fA, fB, k, q, g, rtol=1e-13)
    print(f"  oracle {s_ref:.12e} gpu {s_gpu.real:.12e} iters {st['iters_total']}", flush=True)
    assert abs(s_gpu.real - s_ref) <= 1e-9 * abs(s_ref)


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_lanes_per_row_do_not_change_signal(lanes):
    mesh, kw = MESHES["cyl2"]
    delta, Delta, k = 10600.0, 43100.0, 400.0
    fA, fB = _pgse_gcloud(delta, Delta, k)
    q = O.q_from_b_pgse(2000.0, delta, Delta)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    s_ref, _ = ctx.solve(fA, fB, k, q, (1, 0, 0), rtol=1e-13, lanes=8)
    s, _ = ctx.solve(fA, fB, k, q, (1, 0, 0), rtol=1e-13, lanes=lanes)
    assert abs(s - s_ref) <= 1e-11 * abs(s_ref)


@pytest.mark.parametrize("mode", ["shared_gdir", "per_gdir"])
def test_batched_equals_single_and_oracle(mode):
    mesh, kw = MESHES["cyl2"]
    print(f"[batched] {mode}", flush=True)
    delta, Delta, k = 10600.0, 43100.0, 400.0
    fA, fB = _pgse_gcloud(delta, Delta, k)
    bvals = np.array([0.0, 500.0, 1000.0, 2000.0, 4000.0, 3000.0, 1500.0])
    qs = np.array([O.q_from_b_pgse(b, delta, Delta) for b in bvals])
    if mode == "shared_gdir":
        gd = np.tile([0.0, 1.0, 0.0], (len(qs), 1))
    else:
        rng = np.random.default_rng(3)
        gd = rng.normal(size=(len(qs), 3))
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    sig, st = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13)
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    for j in range(len(qs)):
        s1, _ = ctx.solve(fA, fB, k, qs[j], gd[j], rtol=1e-13)
        assert abs(sig[j] - s1) <= 1e-11 * max(abs(s1), 1e-12), (j, sig[j], s1)
        if j in (0, 3, 4):
            s_ref, _, _ = S.solve(gd[j], qs[j], k, fA, fB, rtol=1e-13)
            assert abs(sig[j].real - s_ref) <= 1e-9 * abs(s_ref)
    assert abs(sig[0].real - 1.0) <= 1e-11


def test_batched_more_than_one_launch_group():
    mesh, kw = MESHES["box4"]
    delta, Delta, k = 1000.0, 10000.0, 200.0
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    rng = np.random.default_rng(5)
    nsig = 70
    gd = rng.normal(size=(nsig, 3))
    qs = np.sqrt(rng.uniform(0, 3000, nsig)) / np.sqrt(delta * delta * (Delta - delta / 3))
    ctx = _gpu_system(mesh, 2e-3, 0.0)
    sig, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13)
    for j in (0, 40, 63, 64, 69):
        s1, _ = ctx.solve(fA, fB, k, qs[j], gd[j], rtol=1e-13)
        assert abs(sig[j] - s1) <= 1e-11 * abs(s1)


def test_boxmesh_golden_gpu():
    """ConvergenceTest.ipynb:423 prints 'Normalized signal: 8.440078e-01' (hmin=hmax=5.413e-01,
    i.e. BoxMesh n=16 on [-2.5,2.5]^3; D=2e-3, delta=1000, Delta=10000, b=1000, gdir z, k=10,
    loop while t < T, lagged right-hand side)."""
    mesh = dm.box_mesh((-2.5,) * 3, (2.5,) * 3, 16, 16, 16)
    delta, Delta, k = 1000.0, 10000.0, 10.0
    q = O.q_from_b_pgse(1000.0, delta, Delta)
    fA, fB = O.lagged_sequence_until(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    ctx = _gpu_system(mesh, 2e-3, 0.0)
    s, st = ctx.solve(fA, fB, k, q, (0, 0, 1), rtol=1e-13)
    assert f"{s.real:.6e}" == "8.440078e-01"
    info = ctx.info()
    assert f"{info['hmin']:.3e}" == "5.413e-01" and f"{info['volume']:.3f}" == "125.000"


def test_zero_b_keeps_signal_one():
    mesh, kw = MESHES["cyl2"]
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 200.0)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    s, _ = ctx.solve(fA, fB, 200.0, 0.0, (1, 0, 0), rtol=1e-13)
    assert abs(s.real - 1.0) <= 1e-11 and abs(s.imag) <= 1e-12


def test_bad_arguments_raise():
    N = _native()
    mesh, kw = MESHES["box4"]
    ctx = _gpu_system(mesh, 2e-3, 0.0)
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 200.0)
    with pytest.raises(N.DmriError) as e:
        ctx.solve(fA, fB, 200.0, 1e-5, (0, 0, 0))
    assert e.value.code == -1
    with pytest.raises(N.DmriError) as e:
        ctx.solve(fA, fB, 200.0, 1e-5, (1, 0, 0), maxit=1, rtol=1e-15)
    assert e.value.code == -4
    bad = mesh.tets.copy()
    bad[0, 0] = mesh.nv + 5
    c2 = N.Context(0)
    with pytest.raises(N.DmriError) as e:
        c2.set_mesh(mesh.xyz, bad)
    assert e.value.code == -5


def test_autotuner_picks_valid_choice_and_results_agree():
    mesh, kw = MESHES["cyl2"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    for nsig, pg in ((1, False), (5, False), (16, True)):
        best, times = ctx.tune(nsig, pg)
        assert best in times and all(t > 0 for t in times.values())
        m, l = best
        assert ((m == 3 and nsig <= 8) or (m in (1, 2, 5) and l * nsig <= 64) or (m == 4 and l == 1)
                or (m == 6 and nsig >= 2 and l <= 4))
    assert {(6, 1), (6, 2), (6, 4)} <= set(times)  # 16 signals: two signals per lane with 1, 2 and 4 lanes
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 400.0)
    q = O.q_from_b_pgse(1000.0, 10600.0, 43100.0)
    s_tuned, _ = ctx.solve(fA, fB, 400.0, q, (0, 1, 0), rtol=1e-13)
    ctx.set_tuning(1, False, 4, 2)
    s_set, _ = ctx.solve(fA, fB, 400.0, q, (0, 1, 0), rtol=1e-13)
    assert abs(s_tuned - s_set) <= 1e-11 * abs(s_set)


@pytest.mark.parametrize("mapping,lanes", [(1, 1), (1, 4), (2, 1), (2, 4), (3, 1), (3, 2), (3, 8), (4, 1), (5, 1),
                                           (5, 2), (5, 8), (6, 1), (6, 2), (6, 4)])
@pytest.mark.parametrize("pg", [False, True])
def test_spmv_mappings_agree_batched(mapping, lanes, pg):
    mesh, kw = MESHES["cyl2"]
    delta, Delta, k = 10600.0, 43100.0, 400.0
    fA, fB = _pgse_gcloud(delta, Delta, k)
    qs = np.array([O.q_from_b_pgse(b, delta, Delta) for b in (0.0, 1000.0, 2500.0, 4000.0, 500.0)])
    gd = np.random.default_rng(7).normal(size=(5, 3)) if pg else np.tile([1.0, 0.0, 0.0], (5, 1))
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    ref, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, lanes=2, mapping=1)
    got, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, lanes=lanes, mapping=mapping)
    assert np.all(np.abs(got - ref) <= 1e-11 * np.abs(ref))


def test_initial_guess_modes_agree():
    mesh, kw = MESHES["cyl2"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 200.0)
    q = O.q_from_b_pgse(2000.0, 10600.0, 43100.0)
    s0, st0 = ctx.solve(fA, fB, 200.0, q, (0, 1, 0), rtol=1e-13, guess=0)
    s1, st1 = ctx.solve(fA, fB, 200.0, q, (0, 1, 0), rtol=1e-13, guess=1)
    s2, st2 = ctx.solve(fA, fB, 200.0, q, (0, 1, 0), rtol=1e-13, guess=2)
    s3, st3 = ctx.solve(fA, fB, 200.0, q, (0, 1, 0), rtol=1e-13, guess=3)
    print(f"  iterations: projection {st0['iters_total']} quadratic {st3['iters_total']} linear {st2['iters_total']} "
          f"previous {st1['iters_total']}", flush=True)
    assert abs(s0 - s1) <= 1e-10 * abs(s1) and abs(s2 - s1) <= 1e-10 * abs(s1) and abs(s3 - s1) <= 1e-10 * abs(s1)
    assert st0["iters_total"] < st3["iters_total"] < st2["iters_total"] < st1["iters_total"]
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    ref, _, _ = S.solve((0, 1, 0), q, 200.0, fA, fB, rtol=1e-13)
    assert abs(s0.real - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("pg", [False, True])
def test_projection_guess_batched_matches_quadratic(pg):
    """The residual-minimising guess over the last states changes only where Krylov starts:
    batched signals (shared or per-signal directions) agree with the quadratic extrapolation's
    and with single solves, in fewer iterations; a second launch group on the same context
    (history buffers reused) gives the same values."""
    mesh, kw = MESHES["cyl2"]
    delta, Delta, k = 10600.0, 43100.0, 400.0
    fA, fB = _pgse_gcloud(delta, Delta, k)
    qs = np.array([O.q_from_b_pgse(b, delta, Delta) for b in (0.0, 1000.0, 2500.0, 4000.0, 500.0)])
    gd = np.random.default_rng(11).normal(size=(5, 3)) if pg else np.tile([0.0, 1.0, 0.0], (5, 1))
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    ref, st_q = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, guess=3)
    got, st_p = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, guess=0)
    again, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, guess=0)
    print(f"  iterations: projection {st_p['iters_total']} quadratic {st_q['iters_total']}", flush=True)
    assert np.all(np.abs(got - ref) <= 1e-11 * np.abs(ref))
    assert np.array_equal(got, again)
    assert st_p["iters_total"] < st_q["iters_total"]
    one, _ = ctx.solve(fA, fB, k, qs[3], gd[3], rtol=1e-13, guess=0)
    assert abs(one - got[3]) <= 1e-11 * abs(got[3])


def _u0_vertex(mesh):
    """A discontinuous initial state: 1 on one half of the mesh, 0.25 on the other, 2 near the axis."""
    x = mesh.xyz
    u = np.where(x[:, 0] < np.median(x[:, 0]), 1.0, 0.25)
    u[np.hypot(x[:, 0], x[:, 1]) < 1.5] = 2.0
    return u


def _u0_dofs(S, mesh, u0v):
    nv = mesh.xyz.shape[0]
    u = np.zeros(S.ndof, dtype=np.complex128)
    for c in range(S.dof_of.shape[0] // nv):
        d = S.dof_of[c * nv:(c + 1) * nv]
        u[d[d >= 0]] = u0v[d >= 0]
    return u


@pytest.mark.parametrize("name,b", [("cyl1", 1000.0), ("cyl2", 2500.0)])
def test_custom_initial_state_matches_oracle(name, b):
    """u0 per vertex (the reference's InitialCondition with a P1 field): plain and batched solves
    against the oracle started from the same state."""
    mesh, kw = MESHES[name]
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 400.0)
    q = O.q_from_b_pgse(b, 10600.0, 43100.0)
    u0v = _u0_vertex(mesh)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    ref, _, _ = S.solve((0, 1, 0), q, 400.0, fA, fB, rtol=1e-13, u0=_u0_dofs(S, mesh, u0v))
    got, _ = ctx.solve(fA, fB, 400.0, q, (0, 1, 0), rtol=1e-13, u0=u0v)
    assert abs(got.real - ref) <= 1e-9 * abs(ref)
    sig, _ = ctx.solve_batched(fA, fB, 400.0, np.array([q, 0.0]), np.array([[0, 1, 0], [1, 0, 0]]), rtol=1e-13,
                               u0=u0v)
    assert abs(sig[0] - got) <= 1e-11 * abs(got)
    # b = 0: the state only relaxes, so the signal is int u0 / |Omega| up to the T2 decay
    w = np.add.reduceat(S.mats["m"], S.rowptr[:-1])
    assert abs(sig[1].real - (w * _u0_dofs(S, mesh, u0v).real).sum() / w.sum()) <= 1e-10
    with pytest.raises(ValueError):
        ctx.solve(fA, fB, 400.0, q, (0, 1, 0), u0=u0v[:-1])


def test_sampled_event_profile_and_spmv_replay():
    """profile=k times the SpMV pair of the first k Krylov iterations of each step only, leaves the
    result unchanged, and the back-to-back replay reports positive per-launch times."""
    mesh, kw = MESHES["cyl2"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 400.0)
    q = O.q_from_b_pgse(1000.0, 10600.0, 43100.0)
    s0, st0 = ctx.solve(fA, fB, 400.0, q, (0, 1, 0), rtol=1e-13)
    s1, st1 = ctx.solve(fA, fB, 400.0, q, (0, 1, 0), rtol=1e-13, profile=1)
    assert s0 == s1
    assert st0["spmv_launches"] == 0
    assert len(fA) <= st1["spmv_launches"] <= 2 * len(fA) and st1["spmv_ms"] > 0
    info = ctx.info()
    # every kernel of the timed iterations: the six of a preconditioned iteration, with their own bytes
    nl = st1["kern_launches"]
    assert len(set(nl)) == 1 and nl[0] == st1["spmv_launches"] // 2 > 0
    assert all(m > 0 for m in st1["kern_ms"])
    n, nnz = info["ndof"], info["nnz"]
    E2 = nnz * 20.0 + 4.0 * (n + 1)
    assert abs(st1["kern_bytes"][2] / nl[2] - (E2 + 48.0 * n)) <= 1e-9 * E2
    assert abs(st1["kern_bytes"][3] / nl[3] - 48.0 * n) <= 1e-9 * n
    assert abs(st1["kern_bytes"][6] / nl[6] - 128.0 * n) <= 1e-9 * n
    us_v, us_t, nb = ctx.time_spmv(1, False, 10)
    assert us_v > 0 and us_t > 0
    assert nb == info["nnz"] * 20.0 + 4.0 * (info["ndof"] + 1) + 48.0 * info["ndof"]


def test_wave_mapping_full_launch_group_matches_oracle():
    """64 signals in one launch group through the wave-per-row kernels (lane = signal)."""
    mesh, kw = MESHES["cyl2"]
    delta, Delta, k = 10600.0, 43100.0, 400.0
    fA, fB = _pgse_gcloud(delta, Delta, k)
    rng = np.random.default_rng(11)
    gd = rng.normal(size=(64, 3))
    qs = np.array([O.q_from_b_pgse(b, delta, Delta) for b in np.linspace(0.0, 4000.0, 64)])
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    got, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, mapping=4, lanes=1)
    ref, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, mapping=1, lanes=1)
    assert np.all(np.abs(got - ref) <= 1e-11 * np.abs(ref))
    til, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, mapping=5, lanes=1)
    assert np.all(np.abs(til - ref) <= 1e-11 * np.abs(ref))
    pr, _ = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, mapping=6, lanes=2)  # 32 signal pairs x 2 lanes
    assert np.all(np.abs(pr - ref) <= 1e-11 * np.abs(ref))
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    for j in (0, 17, 63):
        s, _, _ = S.solve(gd[j], qs[j], k, fA, fB, rtol=1e-13)
        assert abs(got[j].real - s) <= 1e-9 * abs(s)


def test_tile_mapping_single_signal_matches_oracle():
    """One signal through the tile-compressed kernels (8 lanes per row, 32-row tiles)."""
    mesh, kw = MESHES["cyl1"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 400.0)
    q = O.q_from_b_pgse(1500.0, 10600.0, 43100.0)
    got, _ = ctx.solve(fA, fB, 400.0, q, (0.2, 1.0, -0.3), rtol=1e-13, mapping=5, lanes=8)
    S = O.System(mesh.xyz, mesh.tets, None, D=kw["D"])
    ref, _, _ = S.solve((0.2, 1.0, -0.3), q, 400.0, fA, fB, rtol=1e-13)
    assert abs(got.real - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("mapping,lanes", [(2, 1), (1, 2), (4, 1), (6, 1), (6, 2)])
def test_polynomial_preconditioner_matches_jacobi_and_oracle(mapping, lanes):
    """Right preconditioning by 2I - A (scaled system) converges to the same signals as plain
    Jacobi-scaled BiCGStab, in fewer iterations."""
    mesh, kw = MESHES["cyl2"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 200.0)
    qs = np.array([O.q_from_b_pgse(b, 10600.0, 43100.0) for b in (0.0, 1000.0, 3000.0)])
    gd = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.2], [0.3, 0.3, 1.0]])
    poly, stp = ctx.solve_batched(fA, fB, 200.0, qs, gd, rtol=1e-13, mapping=mapping, lanes=lanes, precond=0)
    jac, stj = ctx.solve_batched(fA, fB, 200.0, qs, gd, rtol=1e-13, mapping=mapping, lanes=lanes, precond=1)
    print(f"  iterations: polynomial {stp['iters_total']} jacobi {stj['iters_total']}", flush=True)
    assert np.all(np.abs(poly - jac) <= 1e-10 * np.abs(jac))
    assert stp["iters_total"] < stj["iters_total"]
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    for j in (1, 2):
        ref, _, _ = S.solve(gd[j], qs[j], 200.0, fA, fB, rtol=1e-13)
        assert abs(poly[j].real - ref) <= 1e-9 * abs(ref)


def test_dmrifemlib_initial_condition_normalises_by_its_integral(tmp_path, monkeypatch):
    """MRI_simulation.solve(..., ic) with a discontinuous P1 initial state; PostProcessing divides
    by vol_of_interest = int ic dx as the reference does (DmriFemLib.py:865-876, 924)."""
    import re

    import DmriFemLib as L

    monkeypatch.chdir(tmp_path)
    mesh = L.BoxMesh(L.Point(0, 0, 0), L.Point(10, 8, 6), 6, 5, 4)
    p = L.MRI_parameters()
    p.bvalue = 1000
    p.delta, p.Delta = 10600.0, 43100.0
    p.set_gradient_dir(mesh, 1, 0, 0)
    p.T = p.Delta + p.delta
    p.fs_sym = L.sp.Piecewise((1., p.s < p.delta), (0., p.s < p.Delta), (-1., p.s < p.T), (0., True))
    p.Apply()
    dom = L.MyDomain(mesh, p)
    dom.PeriodicDir = [0, 0, 0]
    dom.IsDomainPeriodic = False
    dom.IsDomainMultiple = False
    dom.D0 = 3e-3
    dom.Apply()
    simu = L.MRI_simulation()
    simu.k = 400
    solver = L.KrylovSolver("bicgstab", "jacobi")
    solver.parameters["relative_tolerance"] = 1e-13
    solver.parameters["absolute_tolerance"] = 0.0

    def ic(x):
        return np.where(x[:, 0] < 5.0, 1.0, 0.5)

    simu.save_snapshots = True
    simu.nskip = 20
    simu.solve(dom, p, solver, ic=ic)
    out = L.PostProcessing(dom, p, simu, None)
    got = float(re.search(r"Normalized signal: ([0-9.e+-]+)", out).group(1))
    # u every nskip steps per compartment and vertex; the last one is u^N when N - 1 is a multiple
    assert simu.snapshots.shape == (len(simu.snapshot_steps), 1, mesh.nv)
    if simu.snapshot_steps[-1] == simu.nsteps - 1:
        _, _, dof_of = simu.ctx.export_pattern()
        assert np.array_equal(simu.snapshots[-1, 0], simu.ctx.solution(0)[dof_of])
    voi_got = float(re.search(r"vol_of_interest: ([0-9.e+-]+)", out).group(1))
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 400.0)
    S = O.System(mesh.xyz, mesh.tets, None, D=3e-3)
    u0 = _u0_dofs(S, mesh, ic(mesh.xyz))
    ref, _, _ = S.solve((1, 0, 0), p.qvalue, 400.0, fA, fB, rtol=1e-13, u0=u0)
    w = np.add.reduceat(S.mats["m"], S.rowptr[:-1])
    voi = float((w * u0.real).sum())
    assert abs(voi_got - voi) <= 1e-3 * voi
    assert abs(got - ref * w.sum() / voi) <= 2e-6 * abs(got)


def test_snapshots_match_oracle_states():
    """u after the steps whose 0-based index is a multiple of snap_every (comri's -s 1 -j nskip,
    comri/two-comp/hpc-fenics-cpp/main.cpp:886-890), in DOF order, against the oracle's state after
    the same number of steps; the last one equals dmri_get_solution when it is the final step."""
    mesh, kw = MESHES["cyl2"]
    delta, Delta, k = 10600.0, 43100.0, 200.0
    q = O.q_from_b_pgse(2000.0, delta, Delta)
    fA, fB = _pgse_gcloud(delta, Delta, k)
    fA, fB = fA[:61], fB[:61]  # steps through the first gradient pulse and into the gap
    every = 15
    print(f"[snapshots] cyl2 steps={len(fA)} every={every}", flush=True)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    s_gpu, st = ctx.solve(fA, fB, k, q, (0, 1, 0), rtol=1e-13, snap_every=every)
    snaps = st["snapshots"]
    assert snaps.shape == (5, ctx.info()["ndof"])
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    for j in range(snaps.shape[0]):
        m = j * every + 1
        _, u_ref, _ = S.solve((0, 1, 0), q, k, fA[:m], fB[:m], rtol=1e-13)
        err = np.linalg.norm(snaps[j] - u_ref) / np.linalg.norm(u_ref)
        print(f"  after {m} steps: rel err {err:.2e}", flush=True)
        assert err <= 1e-10
    assert np.array_equal(snaps[-1], ctx.solution(0))


def test_batched_snapshots_land_in_their_signal_slots():
    """70 signals (two launch groups): snapshot j of signal s sits at [j, s] and equals the plain
    solve's snapshot of that signal to the solver tolerance."""
    mesh, kw = MESHES["box4"]
    delta, Delta, k = 1000.0, 10000.0, 200.0
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(delta, Delta), delta + Delta, k)
    rng = np.random.default_rng(9)
    nsig = 70
    gd = rng.normal(size=(nsig, 3))
    qs = np.sqrt(rng.uniform(0, 3000, nsig)) / np.sqrt(delta * delta * (Delta - delta / 3))
    ctx = _gpu_system(mesh, 2e-3, 0.0)
    sig, st = ctx.solve_batched(fA, fB, k, qs, gd, rtol=1e-13, snap_every=20)
    snaps = st["snapshots"]
    assert snaps.shape == (3, nsig, ctx.info()["ndof"])
    for j in (0, 63, 64, 69):
        s1, st1 = ctx.solve(fA, fB, k, qs[j], gd[j], rtol=1e-13, snap_every=20)
        err = np.abs(snaps[:, j] - st1["snapshots"]).max() / np.abs(st1["snapshots"]).max()
        assert err <= 1e-11, (j, err)


def test_copy_bandwidth_reference_runs():
    """The copy kernel behind bench.py's measured-bandwidth reference (one element per thread over a
    covering grid) runs at sizes that are and are not a multiple of the block size."""
    mesh, kw = MESHES["box4"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    for nbytes in (1 << 26, (1 << 26) + 16 * 77):
        gbs = ctx.copy_bandwidth(nbytes, 3)
        assert 100.0 < gbs < 50000.0, gbs


@pytest.mark.parametrize("name,b,g", [("cyl2", 4000.0, (1, 1, 0)), ("box4", 2000.0, (0, 0, 1))])
def test_complex_diagonal_polynomial_matches_oracle(name, b, g):
    """The polynomial preconditioner in the variables scaled by the complex diagonal (forced with
    preconditioner 2) converges to the oracle's signal, plain and batched with per-signal directions."""
    mesh, kw = MESHES[name]
    delta, Delta, k = 10600.0, 43100.0, 200.0
    q = O.q_from_b_pgse(b, delta, Delta)
    fA, fB = _pgse_gcloud(delta, Delta, k)
    S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=kw["D"], kappa=kw["kappa"])
    s_ref, _, _ = S.solve(g, q, k, fA, fB, rtol=1e-13)
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    s_gpu, st = ctx.solve(fA, fB, k, q, g, rtol=1e-13, precond=2)
    print(f"[cdiag] {name} oracle {s_ref:.12e} gpu {s_gpu.real:.12e} iters {st['iters_total']}", flush=True)
    assert st["precond_used"] == 3
    assert abs(s_gpu.real - s_ref) <= 1e-9 * abs(s_ref)
    gd = np.array([g, (0.3, 0.5, 0.81), (1, 0, 0)], dtype=float)
    sig, stb = ctx.solve_batched(fA, fB, k, [q, q, 0.5 * q], gd, rtol=1e-13, precond=2)
    assert stb["precond_used"] == 3
    assert abs(sig[0].real - s_ref) <= 1e-9 * abs(s_ref)
    s2, _, _ = S.solve(gd[2], 0.5 * q, k, fA, fB, rtol=1e-13)
    assert abs(sig[2].real - s2) <= 1e-9 * abs(s2)
    # the same through the two-signals-per-lane kernels (their complex-diagonal variants)
    sp, stp = ctx.solve_batched(fA, fB, k, [q, q, 0.5 * q], gd, rtol=1e-13, precond=2, mapping=6, lanes=1)
    assert stp["precond_used"] == 3
    assert np.all(np.abs(sp - sig) <= 1e-10 * np.abs(sig))


def test_automatic_complex_diagonal_on_a_large_domain():
    """A strong gradient over a domain far from the origin (the gradient term outweighs the real part
    on the diagonal) selects the complex-diagonal polynomial by itself; the signal matches the oracle."""
    m0 = dm.box_mesh((0.0, 0.0, 0.0), (40.0, 8.0, 8.0), 20, 4, 4)
    mesh = dm.TetMesh(m0.xyz + np.array([200.0, 0.0, 0.0]), m0.tets)
    delta, Delta, k = 2000.0, 6000.0, 100.0
    q = O.q_from_b_pgse(4000.0, delta, Delta)
    fA, fB = _pgse_gcloud(delta, Delta, k)
    S = O.System(mesh.xyz, mesh.tets, None, D=3e-3)
    s_ref, _, _ = S.solve((1, 0, 0), q, k, fA, fB, rtol=1e-13)
    ctx = _gpu_system(mesh, 3e-3, 0.0)
    s_gpu, st = ctx.solve(fA, fB, k, q, (1, 0, 0), rtol=1e-13)
    s_pol, st1 = ctx.solve(fA, fB, k, q, (1, 0, 0), rtol=1e-13, precond=1)
    print(f"[cdiag auto] oracle {s_ref:.12e} gpu {s_gpu.real:.12e} precond {st['precond_used']} iters "
          f"{st['iters_total']} (jacobi {st1['iters_total']})", flush=True)
    assert st["precond_used"] == 3
    assert abs(s_gpu.real - s_ref) <= 1e-9 * abs(s_ref)


def test_snapshot_edge_cases():
    """No step: no snapshot and the initial signal 1; a negative interval is rejected."""
    mesh, kw = MESHES["box4"]
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    s0, st = ctx.solve(np.zeros(0), np.zeros(0), 200.0, 1e-5, (1, 0, 0), snap_every=3)
    assert st["snapshots"].shape == (0, ctx.info()["ndof"])
    assert abs(s0.real - 1.0) <= 1e-13
    with pytest.raises(ValueError):
        ctx.solve(np.zeros(4), np.zeros(4), 200.0, 1e-5, (1, 0, 0), snap_every=-1)
    # one snapshot per step with an interval of 1; the last is the final state
    fA, fB = _pgse_gcloud(1000.0, 3000.0, 200.0)
    _, st = ctx.solve(fA, fB, 200.0, O.q_from_b_pgse(1000.0, 1000.0, 3000.0), (1, 0, 0), rtol=1e-13,
                      snap_every=1)
    assert st["snapshots"].shape[0] == len(fA)
    assert np.array_equal(st["snapshots"][-1], ctx.solution(0))


def test_tuning_database_round_trip(tmp_path):
    """dmri_tune_db: a miss tunes and appends one line, a hit installs the stored choice without
    timing; another pattern misses; the stored choice gives the same signals."""
    mesh, kw = MESHES["cyl2"]
    db = tmp_path / "tune.db"
    ctx = _gpu_system(mesh, kw["D"], kw["kappa"])
    (m1, l1), hit1 = ctx.tune_db(db, 4, True)
    assert not hit1 and m1 in (1, 2, 4, 6)
    lines = db.read_text().split("\n")
    assert len([x for x in lines if x.strip()]) == 1 and lines[0].split()[5:7] == [str(m1), str(l1)]
    ctx2 = _gpu_system(mesh, kw["D"], kw["kappa"])
    (m2, l2), hit2 = ctx2.tune_db(db, 4, True)
    assert hit2 and (m2, l2) == (m1, l1)
    other, kwo = MESHES["cyl1"]
    ctx3 = _gpu_system(other, kwo["D"], kwo["kappa"])
    (_, _), hit3 = ctx3.tune_db(db, 4, True)
    assert not hit3 and len([x for x in db.read_text().split("\n") if x.strip()]) == 2
    fA, fB = _pgse_gcloud(10600.0, 43100.0, 800.0)
    qs = np.full(4, O.q_from_b_pgse(1500.0, 10600.0, 43100.0))
    gd = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=float)
    s_tuned, _ = ctx2.solve_batched(fA, fB, 800.0, qs, gd, rtol=1e-13)
    s_def, _ = _gpu_system(mesh, kw["D"], kw["kappa"]).solve_batched(fA, fB, 800.0, qs, gd, rtol=1e-13)
    assert np.abs(s_tuned - s_def).max() <= 1e-11
