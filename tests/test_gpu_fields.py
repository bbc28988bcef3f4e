"""GPU parity for the coefficient fields: a D tensor per cell (DG0, ImposeDiffusionTensor,
DmriFemLib.py:611-616), 1/T2 per cell (GCloudDmriSolver.py:157-171 reads T2 per cell from HDF5) and
kappa per interface facet (the CR-field permeability of MultilayeredDiskVariablePermeability.ipynb:
482-497).  Values within 1e-13 of the oracle's (row-relative), signals within 1e-9 at rtol 1e-13,
and the adjoint's field-multiplier gradients against central finite differences of the GPU forward
solver (h = 1e-6 relative, 2e-5 relative bar); the scalar gradients at kappa = 0 and D_c = 0 against
second-order one-sided differences."""
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


def _spd_tensors(nc, seed=3):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(nc, 3, 3))
    Qm, _ = np.linalg.qr(A)
    lam = rng.uniform(1e-3, 4e-3, size=(nc, 3))
    D = np.einsum("cij,cj,ckj->cik", Qm, lam, Qm)
    return 0.5 * (D + D.transpose(0, 2, 1))


DCELL = _spd_tensors(MESH.nc)
IT2CELL = np.random.default_rng(4).uniform(1e-5, 2e-5, size=MESH.nc)


def _kappa_of_facets(fv):
    """kappa per facet from its centroid (the same value in the GPU's and the oracle's facet order)."""
    c = MESH.xyz[fv].mean(axis=1)
    return 1e-5 * (1.0 + 0.5 * np.sin(c[:, 2]) * np.cos(0.7 * c[:, 0]))


def _ctx(Dcell=DCELL, it2=IT2CELL, kscale=1.0, D=(3e-3, 3e-3), kappa=1e-5):
    import dmri_native as N

    c = N.Context(0)
    c.set_mesh(MESH.xyz, MESH.tets, MESH.phase)
    c.set_coeffs(D=D, kappa=kappa, D_cell=Dcell, invT2_cell=it2)
    c.assemble()
    kf = None
    if kscale is not None:
        kf = kscale * _kappa_of_facets(c.export_facets())
        c.set_coeffs(D=D, kappa=kappa, D_cell=Dcell, invT2_cell=it2, kappa_facet=kf)
        c.assemble()
    return c


def _oracle(Dcell=DCELL, it2=IT2CELL):
    S0 = O.System(MESH.xyz, MESH.tets, MESH.phase)
    kf = _kappa_of_facets(S0.fv)
    return O.System(MESH.xyz, MESH.tets, MESH.phase, Dcell=None if Dcell is None else Dcell.reshape(-1, 9),
                    invT2cell=it2, kappa_facet=kf)


def _rowrel(S, ref, got):
    rows = np.repeat(np.arange(S.ndof), np.diff(S.rowptr))
    rowmax = np.maximum.reduceat(np.abs(ref), S.rowptr[:-1])[rows]
    return float((np.abs(got - ref) / np.maximum(rowmax, 1e-300)).max())


def test_field_values_and_signal_match_oracle():
    print("[fields] values + signal", flush=True)
    S = _oracle()
    c = _ctx()
    for k in ("m", "k", "r", "gi", "jx", "jy", "jz"):
        assert _rowrel(S, S.mats[k], c.export_values(k)) <= 1e-13, k
    s_gpu, _ = c.solve(FA, FB, K, Q, G, rtol=1e-13)
    s_cpu, _, _ = S.solve(G, Q, K, FA, FB, rtol=1e-13)
    print(f"  gpu {s_gpu.real:.14e} oracle {s_cpu:.14e}", flush=True)
    assert abs(s_gpu.real - s_cpu) <= 1e-9 * abs(s_cpu)


def test_nonsymmetric_tensor_assembled_with_the_test_function_as_row():
    print("[fields] non-symmetric D_cell", flush=True)
    rng = np.random.default_rng(9)
    Dn = DCELL + rng.normal(size=DCELL.shape) * 5e-4
    S = _oracle(Dcell=Dn)
    c = _ctx(Dcell=Dn)
    ref = S.mats["k"]
    assert _rowrel(S, ref, c.export_values("k")) <= 1e-13
    s_gpu, _ = c.solve(FA, FB, K, Q, G, rtol=1e-13)
    s_cpu, _, _ = S.solve(G, Q, K, FA, FB, rtol=1e-13)
    assert abs(s_gpu.real - s_cpu) <= 1e-9 * abs(s_cpu)
    import dmri_native as N

    with pytest.raises(N.DmriError) as e:
        c.adjoint(FA, FB, K, Q, G, rtol=1e-13)
    assert e.value.code == -1


def _fwd(**kw):
    return _ctx(**kw).solve(FA, FB, K, Q, G, rtol=1e-14, guess=3)[0].real


@pytest.fixture(scope="module")
def field_grads():
    s, g, _ = _ctx().adjoint(FA, FB, K, Q, G, rtol=1e-14, guess=3)
    return s, g


@pytest.mark.parametrize("which", ["D0", "D1", "T2_0", "T2_1", "kappa"])
def test_field_multiplier_gradients_match_central_fd(field_grads, which):
    _, g = field_grads
    h = 1e-6
    ph = MESH.phase
    if which.startswith("D"):
        c = int(which[1])
        up, dn = DCELL.copy(), DCELL.copy()
        up[ph == c] *= 1 + h
        dn[ph == c] *= 1 - h
        fd = (_fwd(Dcell=up) - _fwd(Dcell=dn)) / (2 * h)
        got = g["D"][c]
    elif which.startswith("T2"):
        c = int(which[-1])
        up, dn = IT2CELL.copy(), IT2CELL.copy()
        up[ph == c] *= 1 + h
        dn[ph == c] *= 1 - h
        fd = (_fwd(it2=up) - _fwd(it2=dn)) / (2 * h)
        got = g["invT2"][c]
    else:
        fd = (_fwd(kscale=1 + h) - _fwd(kscale=1 - h)) / (2 * h)
        got = g["kappa"]
    print(f"  {which}: adjoint {got:.10e} fd {fd:.10e}", flush=True)
    assert abs(got - fd) <= 2e-5 * max(abs(got), abs(fd))


@pytest.mark.parametrize("which", ["kappa", "D1"])
def test_scalar_gradients_at_zero_coefficient(which):
    """dS/dkappa at kappa = 0 and dS/dD_1 at D_1 = 0 are finite and match one-sided differences
    (f'(0) ~ (-3 f(0) + 4 f(h) - f(2h)) / (2h)); the coefficients cannot go negative."""
    print(f"[fields] gradient at zero {which}", flush=True)
    base = dict(Dcell=None, it2=None, kscale=None)
    if which == "kappa":
        _, g, _ = _ctx(kappa=0.0, **base).adjoint(FA, FB, K, Q, G, rtol=1e-14, guess=3)
        h = 1e-8
        f = [_fwd(kappa=k, **base) for k in (0.0, h, 2 * h)]
        got = g["kappa"]
    else:
        _, g, _ = _ctx(D=(3e-3, 0.0), **base).adjoint(FA, FB, K, Q, G, rtol=1e-14, guess=3)
        h = 3e-9
        f = [_fwd(D=(3e-3, d), **base) for d in (0.0, h, 2 * h)]
        got = g["D"][1]
    fd = (-3 * f[0] + 4 * f[1] - f[2]) / (2 * h)
    print(f"  adjoint {got:.10e} fd {fd:.10e}", flush=True)
    assert np.isfinite(got) and got != 0.0
    assert abs(got - fd) <= 1e-4 * max(abs(got), abs(fd))


def test_dmrifemlib_interface_kappa_by_marker():
    """mydomain.kappa = InterfaceKappa(kappa_tensor, partition_marker) (the CR-field permeability of
    MultilayeredDiskVariablePermeability.ipynb, cell 4) on three layers with markers 0, 1, 2: the
    drop-in's signal equals the oracle's with kappa_tensor[marker[a]][marker[b]] on every interface
    facet (cells found by brute force), and a uniform tensor gives the scalar-kappa signal."""
    import DmriFemLib as L

    print("[fields] DmriFemLib InterfaceKappa", flush=True)
    m = dm.cylinder_mesh(6.0, 2.0, 1500, seed=5, interface_radii=(2.5, 4.5))
    c = m.xyz[m.tets].mean(axis=1)
    r = np.hypot(c[:, 0], c[:, 1])
    marker = (r > 2.5).astype(np.int64) + (r > 4.5)
    phase = (marker % 2).astype(np.int32)
    K = np.array([[0.0, 1e-4, 0.0], [1e-4, 0.0, 1e-5], [0.0, 1e-5, 0.0]])

    def run(kappa):
        p = L.MRI_parameters()
        p.bvalue = 1000
        p.delta, p.Delta = 20000.0, 20000.0
        p.set_gradient_dir(m, 1, 0, 0)
        p.T = p.Delta + p.delta
        p.fs_sym = L.sp.Piecewise((1., p.s < p.delta), (0., p.s < p.Delta), (-1., p.s < p.T), (0., True))
        p.Apply()
        dom = L.MyDomain(m, p)
        dom.phase = phase
        dom.IsDomainMultiple = True
        dom.kappa = kappa
        dom.Apply()
        dom.D = 3e-3
        simu = L.MRI_simulation()
        simu.k = 1000
        solver = L.KrylovSolver("bicgstab", "jacobi")
        solver.parameters["relative_tolerance"] = 1e-13
        solver.parameters["absolute_tolerance"] = 0.0
        simu.solve(dom, p, solver)
        return simu, p

    simu, p = run(L.InterfaceKappa(K, marker))
    S0 = O.System(m.xyz, m.tets, phase)
    faces = {}
    for ci, t in enumerate(m.tets):
        for k in range(4):
            faces.setdefault(tuple(sorted(np.delete(t, k))), []).append(ci)
    kf = np.array([K[marker[faces[tuple(sorted(f))][0]], marker[faces[tuple(sorted(f))][1]]] for f in S0.fv])
    assert set(np.unique(kf)) == {1e-4, 1e-5}
    fA, fB = O.gcloud_sequence(O.pgse_piecewise(20000.0, 20000.0), 40000.0, 1000.0)
    S = O.System(m.xyz, m.tets, phase, D=3e-3, kappa_facet=kf)
    ref, _, _ = S.solve((1, 0, 0), p.qvalue, 1000.0, fA, fB, rtol=1e-13)
    print(f"  drop-in {simu.signal_c.real:.14e} oracle {ref:.14e}", flush=True)
    assert abs(simu.signal_c.real - ref) <= 1e-9 * abs(ref)
    uni, _ = run(L.InterfaceKappa(np.full((3, 3), 3e-5), marker))
    sca, _ = run(3e-5)
    assert abs(uni.signal_c.real - sca.signal_c.real) <= 1e-12 * abs(sca.signal_c.real)
