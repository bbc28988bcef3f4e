"""GPU adjoint gradients (dmri_adjoint) against central finite differences of the GPU forward
solver (h = 1e-6 relative), the CPU oracle, and closed forms.  FD noise with rtol 1e-14 solves is
about 1e-14 / 1e-6 relative, so gradients are compared at 2e-5 relative (of the larger of the two)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import oracle as O  # noqa: E402

MESH = dm.cylinder_mesh(8.0, 4.0, 2500, seed=1001, interface_radii=(4.0,))
DELTA, BIGDELTA, K = 10600.0, 43100.0, 1000.0
FA, FB = O.gcloud_sequence(O.pgse_piecewise(DELTA, BIGDELTA), DELTA + BIGDELTA, K)
Q = O.q_from_b_pgse(2000.0, DELTA, BIGDELTA)
G = (0.0, 1.0, 0.3)
COEF = dict(D=(3e-3, 2e-3), invT2=(1 / 8e4, 1 / 6e4), kappa=1e-5)


def _ctx(**over):
    import dmri_native as N

    kw = dict(COEF)
    kw.update(over)
    c = N.Context(0)
    c.set_mesh(MESH.xyz, MESH.tets, MESH.phase)
    c.set_coeffs(**kw)
    c.assemble()
    return c


def _fwd(fA=FA, fB=FB, q=Q, **over):
    # the adjoint's forward sweep starts each Krylov solve from the quadratic extrapolation
    return _ctx(**over).solve(fA, fB, K, q, G, rtol=1e-14, guess=3)[0].real


def _close(a, b, tol=2e-5):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


@pytest.fixture(scope="module")
def adj():
    # quadratic extrapolation in both sweeps: the checkpointed and the stored-states runs then
    # take the same Krylov iterates, so their gradients agree bit for bit
    stride = int(np.ceil(np.sqrt(len(FA))))
    s, g, st = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, checkpoint_stride=stride, guess=3)
    s_all, g_all, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, checkpoint_stride=10 ** 6, guess=3)
    return s, g, s_all, g_all


@pytest.mark.parametrize("stride", [0, 17])
def test_projection_guess_gradients_match_quadratic(adj, stride):
    """Default guess: the projection over the last states in the backward sweep (and in the
    forward sweep when every state is kept); the gradients agree with the quadratic runs."""
    _, g_q, _, _ = adj
    s, g, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, checkpoint_stride=stride)
    assert abs(s.real - _fwd()) <= 1e-12
    for k in ("D", "invT2"):
        assert np.all(np.abs(g[k] - g_q[k]) <= 1e-9 * np.abs(g_q[k]).max())
    assert abs(g["kappa"] - g_q["kappa"]) <= 1e-9 * abs(g_q["kappa"])
    assert np.abs(g["f_implicit"] - g_q["f_implicit"]).max() <= 1e-9 * np.abs(g_q["f_implicit"]).max()


def test_automatic_checkpointing_keeps_every_state_when_it_fits():
    # a few MB of states against hundreds of GB of HBM: nothing is recomputed
    _, g, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14)
    assert g["checkpoints"] == 0 and g["recomputed_steps"] == 0


def test_adjoint_signal_equals_forward(adj):
    s, g, s_all, g_all = adj
    assert abs(s.real - _fwd()) <= 1e-13
    stride = int(np.ceil(np.sqrt(len(FA))))
    nck = -(-len(FA) // stride)
    # every segment but the last is recomputed; the first takes the remainder of the steps, so the
    # last, whose states the forward sweep keeps, is a full segment of `stride` steps
    assert g["checkpoints"] == nck > 1 and g["recomputed_steps"] == len(FA) - stride
    assert g_all["checkpoints"] == 0 and g_all["recomputed_steps"] == 0
    # recomputation from checkpoints reproduces the stored-states gradients exactly
    for k in ("D", "invT2"):
        assert np.array_equal(g[k], g_all[k])
    assert g["kappa"] == g_all["kappa"] and np.array_equal(g["f_implicit"], g_all["f_implicit"])


@pytest.mark.parametrize("which", ["D0", "D1", "kappa", "T2_0", "T2_1"])
def test_coefficient_gradients_match_central_fd(adj, which):
    s, g, _, _ = adj
    h = 1e-6
    if which.startswith("D"):
        c = int(which[1])
        base = COEF["D"][c]
        d = h * base
        up = list(COEF["D"]); up[c] += d
        dn = list(COEF["D"]); dn[c] -= d
        fd = (_fwd(D=tuple(up)) - _fwd(D=tuple(dn))) / (2 * d)
        got = g["D"][c]
    elif which == "kappa":
        d = h * COEF["kappa"]
        fd = (_fwd(kappa=COEF["kappa"] + d) - _fwd(kappa=COEF["kappa"] - d)) / (2 * d)
        got = g["kappa"]
    else:
        c = int(which[-1])
        base = COEF["invT2"][c]
        d = h * base
        up = list(COEF["invT2"]); up[c] += d
        dn = list(COEF["invT2"]); dn[c] -= d
        fd = (_fwd(invT2=tuple(up)) - _fwd(invT2=tuple(dn))) / (2 * d)
        got = g["invT2"][c]
    print(f"  {which}: adjoint {got:.10e} fd {fd:.10e}", flush=True)
    assert _close(got, fd)


def test_waveform_and_q_gradients_match_central_fd(adj):
    s, g, _, _ = adj
    h = 1e-6
    for n in (0, 5, 11, 50, len(FA) - 2):
        for side in ("f_implicit", "f_explicit"):
            fa, fb = FA.copy(), FB.copy()
            arr = fa if side == "f_implicit" else fb
            d = h * max(abs(arr[n]), 1.0)
            arr[n] += d
            up = _fwd(fA=fa, fB=fb)
            arr[n] -= 2 * d
            dn = _fwd(fA=fa, fB=fb)
            fd = (up - dn) / (2 * d)
            got = g[side][n]
            assert abs(got - fd) <= 2e-5 * max(abs(got), abs(fd)) + 1e-12, (side, n, got, fd)
    d = h * Q
    fd = (_fwd(q=Q + d) - _fwd(q=Q - d)) / (2 * d)
    assert _close(g["q"], fd)


def test_adjoint_matches_oracle_fd_for_D():
    """The same derivative from the CPU oracle (independent forward implementation)."""
    s, g, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14)
    h = 1e-6 * COEF["D"][1]
    f = []
    for dd in (+h, -h):
        S = O.System(MESH.xyz, MESH.tets, MESH.phase, D=(COEF["D"][0], COEF["D"][1] + dd), invT2=COEF["invT2"],
                     kappa=COEF["kappa"])
        f.append(S.solve(G, Q, K, FA, FB, rtol=1e-14)[0])
    fd = (f[0] - f[1]) / (2 * h)
    assert _close(g["D"][1], fd, 5e-5)


def test_closed_forms_at_zero_gradient():
    """b = 0: S = r^N exactly, so dS/dD = dS/dkappa = 0 and dS/d(1/T2) follows from r^N."""
    T2 = 8e4
    c = _ctx(invT2=1 / T2)
    s, g, _ = c.adjoint(FA, FB, K, 0.0, G, rtol=1e-15)
    th, N = 0.5, len(FA)
    r = (1 / K - (1 - th) / T2) / (1 / K + th / T2)
    drdi = (-(1 - th) * (1 / K + th / T2) - th * (1 / K - (1 - th) / T2)) / (1 / K + th / T2) ** 2
    dS = N * r ** (N - 1) * drdi  # derivative w.r.t. a common 1/T2
    assert abs(s.real - r ** N) <= 1e-12
    assert abs(g["invT2"][0] + g["invT2"][1] - dS) <= 1e-8 * abs(dS)
    assert abs(g["D"][0]) <= 1e-9 and abs(g["D"][1]) <= 1e-9 and abs(g["kappa"]) <= 1e-6


def test_torch_autograd_function_applies_adjoint_gradients():
    import torch

    import dmri_native as N
    import dmri_torch

    c = N.Context(0)
    c.set_mesh(MESH.xyz, MESH.tets, MESH.phase)
    D = torch.tensor(COEF["D"], dtype=torch.float64, requires_grad=True)
    T = torch.tensor(COEF["invT2"], dtype=torch.float64, requires_grad=True)
    kap = torch.tensor(COEF["kappa"], dtype=torch.float64, requires_grad=True)
    fa = torch.tensor(FA, requires_grad=True)
    S = dmri_torch.bloch_torrey_signal(c, D, T, kap, Q, fa, FB, K, G, rtol=1e-14)
    (3.0 * S).backward()
    s_ref, g, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14)
    assert abs(S.item() - s_ref.real) <= 1e-14
    assert torch.allclose(D.grad, torch.as_tensor(3.0 * g["D"]), rtol=1e-12, atol=0)
    assert torch.allclose(T.grad, torch.as_tensor(3.0 * g["invT2"]), rtol=1e-12, atol=0)
    assert abs(kap.grad.item() - 3.0 * g["kappa"]) <= 1e-12 * abs(g["kappa"])
    assert torch.allclose(fa.grad, torch.as_tensor(3.0 * g["f_implicit"]), rtol=1e-12, atol=1e-300)


def test_adjoint_with_custom_initial_state_matches_fd():
    """A custom initial state (u0 per vertex) changes x^0 of the sweep: the adjoint signal equals the
    forward one and dS/dD_c, dS/dkappa still match central differences."""
    x = MESH.xyz
    u0 = np.where(x[:, 0] < 0.0, 1.0, 0.4)
    s, g, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, guess=3, u0=u0)
    f0 = _ctx().solve(FA, FB, K, Q, G, rtol=1e-14, guess=3, u0=u0)[0].real
    assert abs(s.real - f0) <= 1e-13
    h = 1e-6
    for c in (0, 1):
        d = h * COEF["D"][c]
        up = list(COEF["D"]); up[c] += d
        dn = list(COEF["D"]); dn[c] -= d
        fd = (_ctx(D=tuple(up)).solve(FA, FB, K, Q, G, rtol=1e-14, guess=3, u0=u0)[0].real -
              _ctx(D=tuple(dn)).solve(FA, FB, K, Q, G, rtol=1e-14, guess=3, u0=u0)[0].real) / (2 * d)
        assert _close(g["D"][c], fd), (c, g["D"][c], fd)
    d = h * COEF["kappa"]
    fd = (_ctx(kappa=COEF["kappa"] + d).solve(FA, FB, K, Q, G, rtol=1e-14, guess=3, u0=u0)[0].real -
          _ctx(kappa=COEF["kappa"] - d).solve(FA, FB, K, Q, G, rtol=1e-14, guess=3, u0=u0)[0].real) / (2 * d)
    assert _close(g["kappa"], fd), (g["kappa"], fd)


def test_projection_checkpoints_recompute_bit_for_bit():
    """With the projection guess a checkpoint holds the solver's whole history, so the recomputed
    segments repeat the forward sweep: the gradients equal the stored-states run's bit for bit."""
    stride = int(np.ceil(np.sqrt(len(FA))))
    print(f"[adjoint] projection checkpoints, stride {stride}", flush=True)
    s1, g1, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, checkpoint_stride=stride)
    s0, g0, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, checkpoint_stride=10 ** 6)
    assert g1["recomputed_steps"] > 0 and g0["recomputed_steps"] == 0
    assert s1 == s0
    for k in ("D", "invT2", "f_implicit", "f_explicit"):
        np.testing.assert_array_equal(g1[k], g0[k])
    assert g1["kappa"] == g0["kappa"] and g1["q"] == g0["q"]
    assert g1["iters_forward"] == g0["iters_forward"] and g1["iters_backward"] == g0["iters_backward"]
    assert g1["iters_recompute"] <= g1["iters_forward"]
