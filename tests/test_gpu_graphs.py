"""hipGraph replay of the BiCGStab iteration runs (solver option graphs = 1) against plain launches.

Every theta step replays the captured graphs between the step's own eager kernels (right-hand side,
projection guess, residual) and a host convergence check, so a solve of N steps replays them about
N times.  The state after every step (snapshots) must equal the plain-launch run's bit for bit, as
must the signals, the Krylov iteration counts and the adjoint gradients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import oracle as O  # noqa: E402

MESH = dm.cylinder_mesh(8.0, 6.0, 4000, seed=1001, interface_radii=(4.0,))
DELTA, BIGDELTA = 10600.0, 43100.0
Q = O.q_from_b_pgse(3000.0, DELTA, BIGDELTA)


def _ctx():
    import dmri_native as N

    c = N.Context(0)
    c.set_mesh(MESH.xyz, MESH.tets, MESH.phase)
    c.set_coeffs(D=(3e-3, 2e-3), invT2=(1 / 8e4, 1 / 6e4), kappa=1e-5)
    c.assemble()
    return c


def _seq(k):
    return O.gcloud_sequence(O.pgse_piecewise(DELTA, BIGDELTA), DELTA + BIGDELTA, k)


@pytest.mark.parametrize("guess,precond", [(0, 0), (3, 0), (0, 1)])
def test_single_signal_graphs_replay_bit_for_bit(guess, precond):
    fA, fB = _seq(400.0)
    print(f"[graphs] single signal guess={guess} precond={precond} steps={len(fA)}", flush=True)
    c = _ctx()
    s0, st0 = c.solve(fA, fB, 400.0, Q, (0, 1, 0.3), guess=guess, precond=precond, snap_every=1, graphs=0)
    s1, st1 = c.solve(fA, fB, 400.0, Q, (0, 1, 0.3), guess=guess, precond=precond, snap_every=1, graphs=1)
    assert st0["graph_launches"] == 0 and st1["graph_launches"] >= len(fA)
    assert st1["iters_total"] == st0["iters_total"] and st1["restarts"] == st0["restarts"]
    assert np.array_equal(st1["snapshots"], st0["snapshots"])  # every replayed step
    assert s1 == s0


def test_batched_signals_graphs_replay_bit_for_bit():
    fA, fB = _seq(500.0)
    qs = Q * np.array([0.5, 1.0, 1.5, 2.0, 2.5])
    g = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0.3, 0.2, 1.0]], dtype=float)
    g /= np.linalg.norm(g, axis=1)[:, None]
    print(f"[graphs] batched 5 signals steps={len(fA)}", flush=True)
    c = _ctx()
    s0, st0 = c.solve_batched(fA, fB, 500.0, qs, g, snap_every=1, graphs=0)
    s1, st1 = c.solve_batched(fA, fB, 500.0, qs, g, snap_every=1, graphs=1)
    assert st1["graph_launches"] > 0 and st1["iters_total"] == st0["iters_total"]
    assert np.array_equal(st1["snapshots"], st0["snapshots"])
    assert np.array_equal(s1, s0)


def test_resumable_batch_keeps_its_graphs_across_calls():
    """dmri_batch_advance in pieces with other solves on the same context in between: the group's
    graphs outlive the calls and replay against the eager work of the other solves."""
    fA, fB = _seq(500.0)
    qs = Q * np.array([1.0, 2.0, 3.0])
    print(f"[graphs] resumable batch steps={len(fA)}", flush=True)
    c = _ctx()
    ref, _ = c.solve_batched(fA, fB, 500.0, qs, (0, 0, 1), graphs=0)
    other, _ = c.solve(fA, fB, 500.0, Q, (1, 0, 0), graphs=0)
    with c.batch(fA, fB, 500.0, qs, (0, 0, 1), graphs=1) as b:
        done = 0
        launches = 0
        for piece in (7, 1, 30, len(fA)):
            n = min(piece, len(fA) - done)
            if n == 0:
                break
            launches += b.advance(n)["graph_launches"]
            done += n
            again, _ = c.solve(fA, fB, 500.0, Q, (1, 0, 0), graphs=0)  # eager work in between
            assert again == other
        sig, steps = b.signals()
    assert steps == len(fA) and launches > 0
    assert np.array_equal(sig, ref)


def test_adjoint_graphs_replay_bit_for_bit():
    fA, fB = _seq(1000.0)
    print(f"[graphs] adjoint steps={len(fA)}", flush=True)
    c = _ctx()
    stride = int(np.ceil(np.sqrt(len(fA))))
    s0, g0, _ = c.adjoint(fA, fB, 1000.0, Q, (0, 1, 0.3), rtol=1e-13, checkpoint_stride=stride, graphs=0)
    s1, g1, st1 = c.adjoint(fA, fB, 1000.0, Q, (0, 1, 0.3), rtol=1e-13, checkpoint_stride=stride, graphs=1)
    assert st1["graph_launches"] > 0
    assert s1 == s0
    for k in ("D", "invT2", "f_implicit", "f_explicit"):
        assert np.array_equal(g1[k], g0[k]), k
    assert g1["kappa"] == g0["kappa"] and g1["q"] == g0["q"]
    assert (g1["iters_forward"], g1["iters_recompute"], g1["iters_backward"]) == (
        g0["iters_forward"], g0["iters_recompute"], g0["iters_backward"])
