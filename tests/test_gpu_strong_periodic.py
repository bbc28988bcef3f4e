"""GPU parity of the strong pseudo-periodic formulation (IsDomainPeriodic = True: PeriodicBD,
DmriFemLib.py:327-375; ThetaMethodF/L_sBC1c/2c 167-238) against the oracle's constrained-space
restatement solved by sparse LU (the exact discrete scheme), and against the closed form of a fully
periodic box, where the uniform state is an exact discrete solution:
S_N = prod_n (1/k - theta q^2 D F_{n-1}^2) / (1/k + theta q^2 D F_n^2).
Bars: constrained pattern bit-exact, values 1e-13 row-relative, signals 1e-9 at rtol 1e-13."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dmri_mesh as dm  # noqa: E402
import dmri_sequences as DS  # noqa: E402
import oracle as O  # noqa: E402

DELTA, BIGDELTA, K = 1000.0, 6000.0, 100.0
T = DELTA + BIGDELTA
FA_f, FB_f, FB_F = DS.gcloud_loop(DS.pgse(DELTA, BIGDELTA), T, K, DS.pgse_integral(DELTA, BIGDELTA))
FA_F = DS.gcloud_loop(DS.pgse_integral(DELTA, BIGDELTA), T, K)[0]


def _q(b):
    return DS.q_from_b(b, DS.pgse_int_F2(DELTA, BIGDELTA))


def _slab_box():
    m = dm.box_mesh((0.0, 0.0, 0.0), (4.0, 3.0, 3.0), 8, 6, 6)
    cx = m.xyz[m.tets].mean(axis=1)[:, 0]
    phase = ((cx > 1.0) & (cx < 3.0)).astype(np.int32)
    return m, phase


CASES = {
    "box_xyz": (dm.box_mesh((0.0, 0.0, 0.0), (4.0, 3.0, 3.0), 8, 6, 6), None, (1, 1, 1), dict(D=2e-3)),
    "box_xy": (dm.box_mesh((-2.0, -1.5, 0.0), (2.0, 1.5, 3.0), 8, 6, 6), None, (1, 1, 0), dict(D=2e-3)),
    "slab_2c": (*_slab_box(), (0, 1, 1), dict(D=(2e-3, 1e-3), kappa=2e-5)),
}


def _pair(name):
    import dmri_native as N

    m, phase, pdir, kw = CASES[name]
    ctx = N.Context(0)
    ctx.set_mesh(m.xyz, m.tets, phase)
    ctx.set_coeffs(**kw)
    ctx.set_periodic_strong(pdir)
    ctx.assemble()
    SP = O.StrongPeriodicSystem(m.xyz, m.tets, phase, pdir=pdir, **kw)
    return ctx, SP


@pytest.mark.parametrize("name", list(CASES))
def test_constrained_pattern_and_values_match_oracle(name):
    print(f"[sbc] values {name}", flush=True)
    ctx, SP = _pair(name)
    info = ctx.info()
    assert info["ndof"] == SP.n and info["n_periodic"] == SP.base.ndof - SP.n > 0
    rowptr, col, dof_of = ctx.export_pattern()
    ref = {"m": SP.M, "gx": SP.G[0], "gy": SP.G[1], "gz": SP.G[2]}
    pat = SP.pattern.copy()
    pat.sort_indices()
    assert np.array_equal(rowptr, pat.indptr) and np.array_equal(col, pat.indices)
    # every vertex maps to the constrained DOF of its master
    for c in range(SP.base.ncomp):
        d = SP.base.dof_of[c * SP.base.nv:(c + 1) * SP.base.nv]
        live = d >= 0
        assert np.array_equal(dof_of[c * SP.base.nv:(c + 1) * SP.base.nv][live], SP.merged_of[d[live]])
    rows = np.repeat(np.arange(SP.n), np.diff(rowptr))
    for k, A in ref.items():
        A = A.tocsr()
        A.sort_indices()
        dense_ref = np.zeros(col.shape[0])
        for i in range(SP.n):
            seg = slice(A.indptr[i], A.indptr[i + 1])
            pos = np.searchsorted(col[rowptr[i]:rowptr[i + 1]], A.indices[seg]) + rowptr[i]
            dense_ref[pos] = A.data[seg]
        got = ctx.export_values(k)
        rowmax = np.maximum.reduceat(np.abs(dense_ref), rowptr[:-1])[rows]
        scale = np.maximum(rowmax, np.abs(dense_ref).max() * 1e-3)
        assert np.all(np.abs(got - dense_ref) <= 1e-13 * scale), (k, float(np.abs(got - dense_ref).max()))


@pytest.mark.parametrize("name,b,g", [("box_xyz", 1500.0, (1.0, 0.4, 0.2)), ("box_xy", 2000.0, (0.3, 1.0, 0.5)),
                                      ("slab_2c", 2000.0, (1.0, 0.6, 0.3))])
def test_signal_matches_oracle_direct_solve(name, b, g):
    print(f"[sbc] signal {name} b={b}", flush=True)
    ctx, SP = _pair(name)
    s_gpu, st = ctx.solve(FA_f, FB_f, K, _q(b), g, rtol=1e-13, F_explicit=FB_F, F_implicit=FA_F)
    s_cpu, _ = SP.solve(g, _q(b), K, FA_F, FB_F)
    print(f"  gpu {s_gpu.real:.14e} oracle {s_cpu:.14e}", flush=True)
    assert abs(s_gpu.real - s_cpu) <= 1e-9 * abs(s_cpu)


def test_fully_periodic_box_closed_form_and_batched_groups():
    print("[sbc] closed form", flush=True)
    ctx, _ = _pair("box_xyz")
    D, theta = 2e-3, 0.5
    bs = [500.0, 500.0, 3000.0, 3000.0, 3000.0]
    qs = np.array([_q(b) for b in bs])
    gd = np.array([[1, 0, 0], [0, 1, 0], [0.2, 0.3, 1.0], [1, 1, 1], [0, 0, 1]], dtype=float)
    sig, _ = ctx.solve_batched(FA_f, FB_f, K, qs, gd, rtol=1e-14, F_explicit=FB_F, F_implicit=FA_F)
    for q, s_ in zip(qs, sig):
        ref = np.prod([(1 / K - theta * q * q * D * fb * fb) / (1 / K + theta * q * q * D * fa * fa)
                       for fa, fb in zip(FA_F, FB_F)])
        assert abs(s_.real - ref) <= 1e-11 * ref, (q, s_.real, ref)
    # groups of equal q in one call equal one signal at a time
    s1, _ = ctx.solve(FA_f, FB_f, K, qs[3], gd[3], rtol=1e-14, F_explicit=FB_F, F_implicit=FA_F)
    assert abs(s1 - sig[3]) <= 1e-12


def test_strong_periodic_argument_errors():
    import dmri_native as N

    m = dm.box_mesh((0.0, 0.0, 0.0), (4.0, 3.0, 3.0), 8, 6, 6)
    xyz = m.xyz.copy()
    on = (np.abs(xyz[:, 0] - 4.0) < 1e-9) & (xyz[:, 1] > 0.1) & (xyz[:, 1] < 2.9)
    xyz[on, 1] += 0.1                              # the x faces no longer match
    c = N.Context(0)
    c.set_mesh(xyz, m.tets, None)
    c.set_coeffs(D=2e-3)
    c.set_periodic_strong((1, 0, 0))
    with pytest.raises(N.DmriError) as e:
        c.assemble()
    assert e.value.code == -5
    ctx, _ = _pair("box_xy")
    with pytest.raises(N.DmriError):               # F_implicit is required
        ctx.solve(FA_f, FB_f, K, _q(1000.0), (1, 0, 0), F_explicit=FB_F)
    with pytest.raises(N.DmriError):               # the adjoint needs a complex-symmetric operator
        ctx.adjoint(FA_f, FB_f, K, _q(1000.0), (1, 0, 0))
    c2 = N.Context(0)
    mb = CASES["box_xy"][0]
    c2.set_mesh(mb.xyz, mb.tets, None)
    c2.set_coeffs(D=2e-3, D_cell=np.tile(np.eye(3).ravel() * 2e-3, (mb.nc, 1)))
    c2.set_periodic_strong((1, 1, 0))
    with pytest.raises(N.DmriError) as e:          # needs a scalar D per compartment
        c2.assemble()
    assert e.value.code == -1


def test_dmrifemlib_strong_periodic_domain_matches_oracle():
    """IsDomainPeriodic = True with PeriodicDir selects the strong form in the drop-in, with F(t_n)
    and F(t_{n-1}) from MRI_parameters.itime_profile."""
    import DmriFemLib as L

    mesh = L.BoxMesh(L.Point(0, 0, 0), L.Point(10, 8, 6), 6, 5, 4)
    p = L.MRI_parameters()
    p.bvalue = 1000
    p.delta, p.Delta = DELTA, BIGDELTA
    p.set_gradient_dir(mesh, 1, 0.5, 0)
    p.T = p.Delta + p.delta
    p.fs_sym = L.sp.Piecewise((1., p.s < p.delta), (0., p.s < p.Delta), (-1., p.s < p.T), (0., True))
    p.Apply()
    dom = L.MyDomain(mesh, p)
    dom.PeriodicDir = [1, 1, 0]
    dom.IsDomainPeriodic = True
    dom.IsDomainMultiple = False
    dom.D0 = 3e-3
    dom.Apply()
    simu = L.MRI_simulation()
    simu.k = K
    solver = L.KrylovSolver("bicgstab", "jacobi")
    solver.parameters["relative_tolerance"] = 1e-13
    solver.parameters["absolute_tolerance"] = 0.0
    simu.solve(dom, p, solver)
    SP = O.StrongPeriodicSystem(mesh.xyz, mesh.tets, None, D=3e-3, pdir=(1, 1, 0))
    ref, _ = SP.solve((1, 0.5, 0), p.qvalue, K, FA_F, FB_F)
    assert abs(simu.signal_c.real - ref) <= 1e-9 * abs(ref)
