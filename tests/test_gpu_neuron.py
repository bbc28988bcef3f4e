"""GPU parity on the neuron-like surrogate under the cos-OGSE waveform (the largest workload's
geometry and sequence, at a size the CPU oracle finishes in seconds): soma r = 8 um + 6 seeded
dendrite trees, voxelised to ~12k vertices; cos-OGSE with 4 periods (ArbitraryTimeSequence.ipynb:
380-392), delta = Delta = 19900 us, T = 40 ms, dt = 100 us: the full 401-step GCloud sequence.
Signals within 1e-9 of the oracle at rtol 1e-13; a 32-direction launch group equal to single
solves; the resumable batch equal to the one-call batched solve; adjoint dS/dD and dS/df(t_n)
against central differences of the GPU forward solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import dmri_sequences as DS  # noqa: E402
import oracle as O  # noqa: E402

MESH = dm.neuron_mesh(target_nv=12000, seed=1004)
NPER, DELTA, BIGDELTA, K = 4, 19900.0, 19900.0, 100.0
F_OGSE, T_OGSE = DS.cos_ogse(NPER, DELTA, BIGDELTA)
FA, FB, _ = DS.gcloud_loop(F_OGSE, T_OGSE, K)
I4 = DS.cos_ogse_int_F2(NPER, DELTA, BIGDELTA)
DIRS = DS.fibonacci_directions(32)


def _q(b):
    return DS.q_from_b(b, I4)


def _ctx(D=3e-3):
    import dmri_native as N

    c = N.Context(0)
    c.set_mesh(MESH.xyz, MESH.tets, None)
    c.set_coeffs(D=D, invT2=1e-16, kappa=0.0)
    c.assemble()
    return c


def test_sequence_is_the_full_cos_ogse():
    of, oT = O.cos_ogse(NPER, DELTA, BIGDELTA)
    oA, oB = O.gcloud_sequence(of, oT, K)
    assert len(FA) == 401 and np.array_equal(FA, oA) and np.array_equal(FB, oB)


@pytest.mark.parametrize("b,j", [(4000.0, 5), (1000.0, 20)])
def test_single_signal_matches_oracle(b, j):
    print(f"[neuron] nv={MESH.nv} b={b} dir={j}", flush=True)
    c = _ctx()
    s_gpu, st = c.solve(FA, FB, K, _q(b), DIRS[j], rtol=1e-13)
    S = O.System(MESH.xyz, MESH.tets, None, D=3e-3)
    s_cpu, _, _ = S.solve(DIRS[j], _q(b), K, FA, FB, rtol=1e-13)
    print(f"  gpu {s_gpu.real:.14e} oracle {s_cpu:.14e} iters/step {st['iters_total'] / len(FA):.1f}", flush=True)
    assert abs(s_gpu.real - s_cpu) <= 1e-9 * abs(s_cpu)


def test_32_direction_launch_group_matches_single_solves_and_oracle():
    print("[neuron] 32-direction batch", flush=True)
    c = _ctx()
    qs = np.full(32, _q(2000.0))
    sig, st = c.solve_batched(FA, FB, K, qs, DIRS, rtol=1e-13)
    for j in (0, 13, 31):
        s1, _ = c.solve(FA, FB, K, qs[j], DIRS[j], rtol=1e-13)
        print(f"  dir {j}: batched {sig[j].real:.14e} single {s1.real:.14e}", flush=True)
        assert abs(sig[j] - s1) <= 1e-10 * max(abs(s1), 1e-2)
    S = O.System(MESH.xyz, MESH.tets, None, D=3e-3)
    for j in (7, 26):
        s_cpu, _, _ = S.solve(DIRS[j], qs[j], K, FA, FB, rtol=1e-13)
        assert abs(sig[j].real - s_cpu) <= 1e-9 * max(abs(s_cpu), 1e-2), j


def test_resumable_batch_equals_one_call_batched_solve():
    print("[neuron] resumable batch", flush=True)
    c = _ctx()
    qs = np.repeat([_q(1000.0), _q(3000.0)], 16)
    sig, _ = c.solve_batched(FA, FB, K, qs, DIRS, rtol=1e-13)
    b = c.batch(FA, FB, K, qs, DIRS, rtol=1e-13, group=32)
    done_total = 0
    for chunk in (1, 99, 150, 1000):
        st = b.advance(chunk)
        done_total = min(len(FA), done_total + chunk)
        _, done = b.signals()
        assert done == done_total and st["iters_total"] >= 0
    got, done = b.signals()
    b.close()
    assert done == len(FA)
    assert np.abs(got - sig).max() <= 1e-13 * np.abs(sig).max()
    # groups of 8: another launch mapping, the same signals to the solver tolerance
    b8 = c.batch(FA, FB, K, qs, DIRS, rtol=1e-13, group=8)
    b8.advance(len(FA))
    g8, _ = b8.signals()
    b8.close()
    assert np.abs(g8 - sig).max() <= 1e-10 * np.abs(sig).max()


def test_adjoint_gradients_match_central_fd():
    print("[neuron] adjoint", flush=True)
    q, g = _q(2000.0), DIRS[5]
    s, grads, _ = _ctx().adjoint(FA, FB, K, q, g, rtol=1e-14, guess=3)

    def fwd(D=3e-3, fa=FA, fb=FB):
        return _ctx(D).solve(fa, fb, K, q, g, rtol=1e-14, guess=3)[0].real

    assert abs(s.real - fwd()) <= 1e-12
    h = 1e-6 * 3e-3
    fd = (fwd(D=3e-3 + h) - fwd(D=3e-3 - h)) / (2 * h)
    print(f"  dS/dD adjoint {grads['D'][0]:.10e} fd {fd:.10e}", flush=True)
    assert abs(grads["D"][0] - fd) <= 2e-5 * max(abs(fd), abs(grads["D"][0]))
    for n in (1, 120, 300):
        fa = FA.copy()
        d = 1e-6 * max(abs(fa[n]), 1.0)
        fa[n] += d
        up = fwd(fa=fa)
        fa[n] -= 2 * d
        dn = fwd(fa=fa)
        fd = (up - dn) / (2 * d)
        got = grads["f_implicit"][n]
        print(f"  dS/df_A[{n}] adjoint {got:.10e} fd {fd:.10e}", flush=True)
        assert abs(got - fd) <= 2e-5 * max(abs(got), abs(fd)) + 1e-12


def test_batch_argument_errors_and_stale_context():
    import dmri_native as N

    c = _ctx()
    qs = np.full(3, _q(1000.0))
    with pytest.raises(N.DmriError):
        c.batch(FA, FB, K, qs, np.zeros((3, 3)))          # |g| = 0
    with pytest.raises(N.DmriError):
        c.batch(FA, FB, K, qs, DIRS[:3], group=65)        # launch groups hold at most 64 signals
    b = c.batch(FA[:5], FB[:5], K, qs, DIRS[:3], group=2)
    b.advance(2)
    c.assemble()                                          # the batch's scaled system is now stale
    with pytest.raises(N.DmriError) as e:
        b.advance(1)
    assert e.value.code == -3
    b.close()
