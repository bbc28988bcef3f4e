"""Host-side logic and the C ABI surface -- CPU only (no GPU calls)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import dmri_mesh as dm
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "dmri-fem-cloud_amd", "lib", "libdmri.so")
HEADER = os.path.join(ROOT, "include", "dmri.h")


def _declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dmri_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    if not os.path.exists(LIB):
        pytest.skip("libdmri.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dmri_[a-z_]+)\b", out))
    declared = _declared_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing


def test_ctypes_binding_loads_and_knows_every_declaration():
    if not os.path.exists(LIB):
        pytest.skip("libdmri.so not built")
    import dmri_native as N

    assert set(_declared_functions()) == set(N.EXPORTED)
    assert "gfx950" in N.version()


def _faces(tets):
    f = np.concatenate([np.sort(tets[:, [1, 2, 3]], 1), np.sort(tets[:, [0, 2, 3]], 1),
                        np.sort(tets[:, [0, 1, 3]], 1), np.sort(tets[:, [0, 1, 2]], 1)])
    _, counts = np.unique(f, axis=0, return_counts=True)
    return f, counts


@pytest.mark.parametrize("maker", ["box", "cyl", "cyl2", "cylbox"])
def test_generated_meshes_are_conforming_and_positive(maker):
    if maker == "box":
        m = dm.box_mesh((0, 0, 0), (2, 3, 4), 3, 4, 5)
        vol = 24.0
    elif maker == "cyl":
        m = dm.cylinder_mesh(5.0, 10.0, 4000, seed=1000)
        vol = None
    elif maker == "cyl2":
        m = dm.cylinder_mesh(8.0, 6.0, 5000, seed=1001, interface_radii=(4.0,))
        vol = None
    else:
        m = dm.cylinders_box_mesh(20.0, 20.0, 4.0, ncyl=6, rmin=1.5, rmax=3.0, target_nv=12000, seed=1002)
        vol = 20.0 * 20.0 * 4.0
    v = dm.cell_volumes(m.xyz, m.tets)
    assert v.min() > 1e-6 * v.max()
    _, counts = _faces(m.tets)
    assert counts.max() <= 2  # no face shared by three cells
    if vol is not None:
        assert abs(v.sum() - vol) <= 1e-9 * vol
    if m.phase is not None:
        assert set(np.unique(m.phase)) == {0, 1}


def test_cylinder_volume_and_interface_radius():
    m = dm.cylinder_mesh(8.0, 6.0, 20000, seed=1001, interface_radii=(4.0,))
    v = dm.cell_volumes(m.xyz, m.tets)
    inner = v[m.phase == 1].sum()
    assert abs(v.sum() / (np.pi * 64 * 6) - 1) < 5e-3
    assert abs(inner / (np.pi * 16 * 6) - 1) < 2e-2


def test_meshes_are_seed_deterministic():
    a = dm.cylinder_mesh(5.0, 10.0, 3000, seed=42)
    b = dm.cylinder_mesh(5.0, 10.0, 3000, seed=42)
    c = dm.cylinder_mesh(5.0, 10.0, 3000, seed=43)
    assert np.array_equal(a.xyz, b.xyz) and np.array_equal(a.tets, b.tets)
    assert not np.array_equal(a.xyz, c.xyz)


def test_dolfin_xml_round_trip(tmp_path):
    m = dm.box_mesh((0, 0, 0), (1, 1, 1), 2, 2, 2)
    p = tmp_path / "m.xml"
    dm.write_dolfin_xml(m, str(p))
    r = dm.read_dolfin_xml(str(p))
    assert np.array_equal(r.tets, m.tets) and np.allclose(r.xyz, m.xyz, rtol=0, atol=0)


def test_dmrifemlib_parameters_match_closed_forms():
    if not os.path.exists(LIB):
        pytest.skip("libdmri.so not built")
    import DmriFemLib as L

    mesh = L.BoxMesh(L.Point(-2.5, -2.5, -2.5), L.Point(2.5, 2.5, 2.5), 4, 4, 4)
    p = L.MRI_parameters()
    p.bvalue = 1000
    p.delta, p.Delta = 10600, 43100
    p.set_gradient_dir(mesh, 2, 0, 0)
    p.T = p.Delta + p.delta
    p.fs_sym = L.sp.Piecewise((1., p.s < p.delta), (0., p.s < p.Delta), (-1., p.s < p.T), (0., True))
    p.Apply()
    assert abs(p.qvalue - O.q_from_b_pgse(1000.0, 10600.0, 43100.0)) <= 1e-12 * p.qvalue
    assert abs(p.gdir.x() - 1.0) < 1e-15
    simu = L.MRI_simulation()
    simu.k = 200
    fA, fB, FB, FA = simu._sequence(p, integral=True)
    rA, rB = O.gcloud_sequence(O.pgse_piecewise(10600.0, 43100.0), 53700.0, 200.0)
    assert np.array_equal(fA, rA) and np.array_equal(fB, rB)
    rFA, rF = O.gcloud_sequence(O.pgse_integral(10600.0, 43100.0), 53700.0, 200.0)
    assert np.abs(FB - rF).max() <= 1e-9 and np.abs(FA - rFA).max() <= 1e-9
    dom = L.MyDomain(mesh, p)
    assert f"{dom.hmin:.3e}" == f"{np.sqrt(3) * 5 / 4:.3e}"


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stdout


def test_bench_sequence_matches_oracle():
    sys.path.insert(0, ROOT)
    import bench

    a = bench.pgse_sequence(10600.0, 43100.0, 200.0)
    b = O.gcloud_sequence(O.pgse_piecewise(10600.0, 43100.0), 53700.0, 200.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_profile_summary_samples_first_iterations_of_every_step(tmp_path):
    """prof_summary's sample: the first k dispatches of every Krylov iteration kernel after each
    step's residual kernel (the launches bench.py times with events), from a synthetic kernel trace;
    the names the traffic summary gives them."""
    import prof_summary as ps

    names = []
    for _ in range(3):  # three theta steps
        names += ["rhs_history_kernel<1, false>", "proj_gram_kernel<12>", "residual_kernel<1, false>"]
        for _ in range(4):  # four Krylov iterations
            names += ["bicg_p_kernel", "spmv_v_kernel<1, false, 1>", "spmv_v_kernel<1, false, 0>",
                      "spmv_t_kernel<1, false, 1>", "spmv_t_kernel<1, false, 2>", "bicg_update_kernel"]
    rows = [{"Kernel_Name": f"void dmri::(anonymous namespace)::{n}(dmri::SolveArgs, int)"} for n in names]
    picked = [ps.short(rows[i]["Kernel_Name"]) for i in sorted(ps.sampled(rows, 1))]
    assert len(picked) == 3 * 6
    assert picked.count("spmv_v_kernel<1, false, 0>") == 3 and picked.count("spmv_t_kernel<1, false, 2>") == 3
    assert picked.count("bicg_update_kernel") == 3
    assert len(ps.sampled(rows, 2)) == 3 * 6 * 2
    named = [ps.iteration_kernel(f"void dmri::{n}(dmri::SolveArgs, int)", 2) for n in names[3:9]]
    assert named == ["bicg_p", "spmv_v_pre", "spmv_v", "spmv_t_pre", "spmv_t", "bicg_update"]
    assert ps.iteration_kernel("void dmri::spmv_v_pair_kernel<1, true, 3>(dmri::SolveArgs, int)", 6) == "spmv_v_pre"
    assert ps.iteration_kernel("void dmri::spmv_v_pair_kernel<1, true, 0>(dmri::SolveArgs, int)", 2) is None
    # COCG's kernels: the p- and update kernels by name, the SpMVs by their operation
    assert [ps.iteration_kernel(f"void dmri::{k}(dmri::SolveArgs, int)", 2) for k in
            ("cocg_p_kernel", "spmv_v_kernel<1, true, 5>", "spmv_v_kernel<1, false, 6>", "spmv_v_kernel<1, true, 4>",
             "cocg_update_kernel")] == ["cocg_p", "spmv_pre", "spmv_pre", "spmv_q", "cocg_update"]
    # the FP32-input and folded-update preconditioner SpMVs, and the q-kernel with the p-step folded in
    assert [ps.iteration_kernel(f"void dmri::spmv_v_kernel<1, {pg}, {op}>(dmri::SolveArgs, int)", 2) for pg, op in
            (("true", 7), ("false", 8), ("true", 9))] == ["spmv_pre", "spmv_pre", "spmv_q"]


def test_bench_roofline_names_folded_kernels_and_matches_first_iteration_traffic():
    """bench.py's kernel table describes the folded COCG kernels (no p-kernel launches: the p-step is
    in the q-kernel; no update launches: the update of r is in the preconditioner's SpMV), and the
    roofline's traffic is the counters' first-iteration mean when the events time one iteration."""
    import bench

    def stats(launches):  # the q-kernel (slot 2) takes the most time
        return {"kern_ms": [(2.0 if j == 2 else 1.0) * (n > 0) for j, n in enumerate(launches)],
                "kern_launches": list(launches), "kern_bytes": [4e9 * n for n in launches]}

    traffic = {"spmv_q": {"bytes": 9.0e9, "bytes_first": 5.0e9, "source": "x"}}
    qf = stats([0, 4, 4, 0, 0, 0, 4])  # p-step folded into the q-kernel
    t = bench.kernel_table(qf, "cocg")
    assert "cocg_p" not in t and t["spmv_q"]["what"].startswith("x += alpha p")
    r1 = bench.roofline_block(qf, 6000.0, traffic, 32, 1, "cocg")
    r2 = bench.roofline_block(qf, 6000.0, traffic, 32, 2, "cocg")
    assert r1["traffic"] == 5.0e9 and r2["traffic"] == 9.0e9
    fr = bench.kernel_table(stats([4, 4, 4, 0, 0, 0, 0]), "cocg")  # update of r folded into z = M r
    assert "cocg_update" not in fr and fr["spmv_pre"]["what"].startswith("r -= alpha q")
    plain = bench.kernel_table(stats([4, 4, 4, 0, 0, 0, 4]), "cocg")
    assert plain["spmv_q"]["what"] == bench.KERNEL_DESC["spmv_q"]


def _layered_cylinder():
    """Three concentric layers (partition markers 0, 1, 2 by cell-centroid radius), phase = marker % 2."""
    m = dm.cylinder_mesh(6.0, 2.0, 700, seed=5, interface_radii=(2.5, 4.5))
    c = m.xyz[m.tets].mean(axis=1)
    r = np.hypot(c[:, 0], c[:, 1])
    marker = (r > 2.5).astype(np.int64) + (r > 4.5)
    return m, marker


def test_facet_cells_and_interface_kappa_against_brute_force():
    """DmriFemLib.facet_cells / interface_kappa_values: the two cells of every facet between
    compartments, found by testing every cell, and kappa_tensor[marker[a]][marker[b]] on them (the rule
    of the reference's variable-permeability example); the midpoint callable and the array forms."""
    import DmriFemLib as L

    m, marker = _layered_cylinder()
    phase = marker % 2
    tets = m.tets.astype(np.int64)
    faces = {}
    for ci, t in enumerate(tets):
        for k in range(4):
            f = tuple(sorted(np.delete(t, k)))
            faces.setdefault(f, []).append(ci)
    iface = [f for f, cs in faces.items() if len(cs) == 2 and phase[cs[0]] != phase[cs[1]]]
    assert len(iface) > 50
    rng = np.random.default_rng(0)
    fv = np.array([rng.permutation(f) for f in iface])  # vertex order inside a facet does not matter
    cells = L.facet_cells(m.tets, fv)
    for f, (a, b) in zip(iface, cells):
        assert [a, b] == sorted(faces[f])
    K = np.array([[0.0, 1e-4, 0.0], [1e-4, 0.0, 1e-5], [0.0, 1e-5, 0.0]])
    vals = L.interface_kappa_values(L.InterfaceKappa(K, marker), m.xyz, m.tets, fv)
    expect = np.array([K[marker[faces[f][0]], marker[faces[f][1]]] for f in iface])
    np.testing.assert_array_equal(vals, expect)
    assert set(np.unique(vals)) == {1e-4, 1e-5}
    mid = m.xyz[fv].mean(axis=1)
    np.testing.assert_array_equal(L.interface_kappa_values(lambda x: 2e-5 * x[:, 2], m.xyz, m.tets, fv), 2e-5 * mid[:, 2])
    np.testing.assert_array_equal(L.interface_kappa_values(expect, m.xyz, m.tets, fv), expect)
    with pytest.raises(ValueError):
        L.interface_kappa_values(expect[:-1], m.xyz, m.tets, fv)
    with pytest.raises(ValueError):
        L.facet_cells(m.tets, np.array([[0, 1, 2 + m.nv - 3]]))
    boundary = [f for f, cs in faces.items() if len(cs) == 1][:5]
    assert np.all(L.facet_cells(m.tets, np.array(boundary))[:, 1] == -1)


def test_profile_step_summary_idle_from_synthetic_trace(tmp_path):
    """prof_summary steps: span, kernel-covered time and idle of each theta step from a synthetic
    kernel trace (two steps of 3 kernels with known gaps, then a long pause that is left out)."""
    import prof_summary as ps

    rows = [("rhs_history_kernel<1, false>", 0, 10), ("residual_kernel<1, false>", 12, 20),
            ("bicg_p_kernel", 25, 40), ("rhs_history_kernel<1, false>", 50, 60),
            ("residual_kernel<1, false>", 61, 70), ("bicg_p_kernel", 70, 95),
            ("rhs_history_kernel<1, false>", 100, 110), ("rhs_history_kernel<1, false>", 900_000, 900_010)]
    p = tmp_path / "kt.csv"
    with open(p, "w") as fh:
        fh.write("Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for n, s, e in rows:
            fh.write(f"\"void dmri::{n}(dmri::SolveArgs, int)\",{s * 1000},{e * 1000}\n")
    out = ps.steps(str(p), "rhs_history_kernel")
    # steps: [0, 50): covered 10 + 8 + 15 = 33, idle 17; [50, 100): covered 10 + 34 = 44, idle 6
    assert out["steps"] == 2
    assert out["span_us"] == 50.0 and out["busy_us"] == 38.5 and out["idle_us"] == 11.5
