"""COCG with the reduced-precision preconditioner input (pre_fp32 = 1: the preconditioner's SpMV reads
r and the polynomial's operator values rounded to FP32 and sums in FP64) converges to the same
discrete solution as FP64 COCG: signals within 1e-9 of the oracle at rtol 1e-13 and within 1e-10
of FP64 COCG, with the real-part polynomial (PGSE cylinders and box: its imaginary-part bound is
below 5 % of the real part's) and the whole-operator polynomial (the neuron surrogate under cos-OGSE
at low b), one direction and per-signal directions (the folded FP32 kernels), SpMV mappings 1 and 2,
the resumable batch and hipGraph replay bit for bit, and the adjoint's gradients.  The complex-
diagonal form keeps FP64 inputs (identical results), and the option is refused outside COCG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import dmri_sequences as DS  # noqa: E402
import oracle as O  # noqa: E402

DELTA, BIGDELTA, K = 10600.0, 43100.0, 800.0
FA, FB = O.gcloud_sequence(O.pgse_piecewise(DELTA, BIGDELTA), DELTA + BIGDELTA, K)
MESHES = {
    "cyl1": (dm.cylinder_mesh(5.0, 8.0, 2500, seed=1000), dict(D=3e-3)),
    "cyl2": (dm.cylinder_mesh(8.0, 4.0, 3000, seed=1001, interface_radii=(4.0,)), dict(D=(3e-3, 2e-3), kappa=1e-5)),
    "box": (dm.box_mesh((0.0, 0.0, 0.0), (12.0, 6.0, 6.0), 10, 5, 5), dict(D=2e-3)),
}
NEURON = dm.neuron_mesh(target_nv=12000, seed=1004)
F_OGSE, T_OGSE = DS.cos_ogse(4, 19900.0, 19900.0)
NFA, NFB, _ = DS.gcloud_loop(F_OGSE, T_OGSE, 100.0)
NI4 = DS.cos_ogse_int_F2(4, 19900.0, 19900.0)
DIRS = DS.fibonacci_directions(32)


def _ctx(name):
    import dmri_native as N

    c = N.Context(0)
    if name == "neuron":
        c.set_mesh(NEURON.xyz, NEURON.tets, None)
        c.set_coeffs(D=3e-3, invT2=1e-16, kappa=0.0)
    else:
        m, kw = MESHES[name]
        c.set_mesh(m.xyz, m.tets, m.phase)
        c.set_coeffs(**kw)
    c.assemble()
    return c


@pytest.mark.parametrize("name,b,form", [("cyl1", 2000.0, 2), ("cyl2", 4000.0, 2), ("box", 3000.0, 2),
                                         ("neuron", 50.0, 1)])
def test_fp32_preconditioner_signal_matches_oracle_and_fp64(name, b, form):
    print(f"[cocg fp32] {name} b={b}", flush=True)
    if name == "neuron":
        q, g, fa, fb, k = DS.q_from_b(b, NI4), DIRS[5], NFA, NFB, 100.0
    else:
        q, g, fa, fb, k = O.q_from_b_pgse(b, DELTA, BIGDELTA), (0.3, 1.0, 0.2), FA, FB, K
    c = _ctx(name)
    s32, st32 = c.solve(fa, fb, k, q, g, rtol=1e-13, krylov="cocg", pre_fp32=1)
    s64, st64 = c.solve(fa, fb, k, q, g, rtol=1e-13, krylov="cocg")
    if name == "neuron":
        ref = O.System(NEURON.xyz, NEURON.tets, None, D=3e-3).solve(g, q, k, fa, fb, rtol=1e-13)[0]
    else:
        m, kw = MESHES[name]
        ref = O.System(m.xyz, m.tets, m.phase, **kw).solve(g, q, k, fa, fb, rtol=1e-13)[0]
    print(f"  fp32-pre {s32.real:.14e} fp64 {s64.real:.14e} oracle {ref:.14e}; precond {st32['precond_used']}; "
          f"iterations {st32['iters_total']} against {st64['iters_total']}", flush=True)
    assert st32["precond_used"] == form and st64["precond_used"] == form
    assert st32["relres_max"] <= 1e-13 and st32["restarts"] == 0
    assert abs(s32.real - ref) <= 1e-9 * abs(ref)
    assert abs(s32 - s64) <= 1e-10 * abs(ref)
    assert st32["iters_total"] <= 1.3 * st64["iters_total"]


def test_fp32_preconditioner_per_direction_batch_both_mappings():
    print("[cocg fp32] neuron, 32 directions, mappings 1 and 2", flush=True)
    c = _ctx("neuron")
    qs = np.full(32, DS.q_from_b(10.0, NI4))
    ref, st64 = c.solve_batched(NFA, NFB, 100.0, qs, DIRS, rtol=1e-13, krylov="cocg")
    assert st64["precond_used"] == 1
    scale = np.abs(ref).max()
    for mapping, lanes in ((2, 1), (1, 1), (1, 2)):
        sig, st = c.solve_batched(NFA, NFB, 100.0, qs, DIRS, rtol=1e-13, krylov="cocg", pre_fp32=1,
                                  mapping=mapping, lanes=lanes)
        print(f"  mapping {mapping} lanes {lanes}: max diff {np.abs(sig - ref).max() / scale:.2e}, iterations "
              f"{st['iters_total']} against {st64['iters_total']}", flush=True)
        assert st["relres_max"] <= 1e-13 and st["precond_used"] == 1
        assert np.abs(sig - ref).max() <= 1e-10 * scale, (mapping, lanes)
    S = O.System(NEURON.xyz, NEURON.tets, None, D=3e-3)
    s_cpu = S.solve(DIRS[17], qs[17], 100.0, NFA, NFB, rtol=1e-13)[0]
    assert abs(sig[17].real - s_cpu) <= 1e-9 * abs(s_cpu)


def test_fp32_preconditioner_resumable_batch_and_graphs_bit_for_bit():
    print("[cocg fp32] resumable batch, graph replay", flush=True)
    c = _ctx("cyl2")
    qs = np.array([O.q_from_b_pgse(b, DELTA, BIGDELTA) for b in (1000.0, 2000.0, 4000.0)])
    s0, st0 = c.solve_batched(FA, FB, K, qs, (0, 0, 1), rtol=1e-13, krylov="cocg", pre_fp32=1, snap_every=1)
    s1, st1 = c.solve_batched(FA, FB, K, qs, (0, 0, 1), rtol=1e-13, krylov="cocg", pre_fp32=1, snap_every=1,
                              graphs=1)
    assert st1["graph_launches"] > 0 and st1["iters_total"] == st0["iters_total"]
    assert np.array_equal(st1["snapshots"], st0["snapshots"]) and np.array_equal(s1, s0)
    with c.batch(FA, FB, K, qs, (0, 0, 1), rtol=1e-13, krylov="cocg", pre_fp32=1) as b:
        b.advance(7)
        b.advance(len(FA) - 7)
        sb, done = b.signals()
    assert done == len(FA) and np.array_equal(sb, s0)


def test_fp32_preconditioner_adjoint_gradients_match_fp64():
    print("[cocg fp32] adjoint", flush=True)
    c = _ctx("cyl2")
    q = O.q_from_b_pgse(2000.0, DELTA, BIGDELTA)
    g = (0.0, 1.0, 0.3)
    s32, g32, _ = c.adjoint(FA, FB, K, q, g, rtol=1e-14, krylov="cocg", pre_fp32=1, checkpoint_stride=4)
    s64, g64, _ = c.adjoint(FA, FB, K, q, g, rtol=1e-14, krylov="cocg")
    assert abs(s32 - s64) <= 1e-11 * abs(s64)
    for key in ("D", "invT2", "f_implicit", "f_explicit"):
        assert np.abs(g32[key] - g64[key]).max() <= 1e-8 * np.abs(g64[key]).max(), key
    for key in ("kappa", "q"):
        assert abs(g32[key] - g64[key]) <= 1e-8 * abs(g64[key]), key


def test_fp32_preconditioner_complex_diagonal_keeps_fp64_and_refusals():
    import dmri_native as N

    print("[cocg fp32] complex diagonal; refusals", flush=True)
    c = _ctx("cyl2")
    q = O.q_from_b_pgse(4000.0, DELTA, BIGDELTA)
    s32, st32 = c.solve(FA, FB, K, q, (0.3, 1.0, 0.2), rtol=1e-13, krylov="cocg", precond=2, pre_fp32=1)
    s64, st64 = c.solve(FA, FB, K, q, (0.3, 1.0, 0.2), rtol=1e-13, krylov="cocg", precond=2)
    assert st32["precond_used"] == 3 and s32 == s64 and st32["iters_total"] == st64["iters_total"]
    # a repeated solve with the FP32 input is bit for bit the same
    q1 = O.q_from_b_pgse(1000.0, DELTA, BIGDELTA)
    r1 = c.solve(FA, FB, K, q1, (0.3, 1.0, 0.2), rtol=1e-13, krylov="cocg", pre_fp32=1)
    r2 = c.solve(FA, FB, K, q1, (0.3, 1.0, 0.2), rtol=1e-13, krylov="cocg", pre_fp32=1)
    assert r1[0] == r2[0] and r1[1]["iters_total"] == r2[1]["iters_total"]
    with pytest.raises(N.DmriError, match="COCG option"):
        c.solve(FA, FB, K, q, (0.3, 1.0, 0.2), pre_fp32=1)
    with pytest.raises(N.DmriError, match="pre_fp32 must be 0 or 1"):
        c.solve(FA, FB, K, q, (0.3, 1.0, 0.2), krylov="cocg", pre_fp32=2)
