"""Mesh input on the reference's own gmsh files (tests/golden/*.msh.zip are copies of the data files
comri/meshes/cyl6_r_3E_6_vol.msh.zip, cyl12_r_3E_6_vol.msh.zip and undyAxon_Ax_1E_6_vol.msh.zip).

The reference turns them into DOLFIN XML with dolfin-convert (comri/meshes/neuron_download/
dolfin-convert.py:357-597, a Python 2 script that is read here, not run).  What pins the readers:
* the prism meshes fill a straight prism over a regular polygon, so their volume is the polygon's
  shoelace area times the length (a dropped, doubled or mis-numbered cell changes it);
* the axon mesh is a conforming mesh of a ball-like region: every interior triangle is shared by
  exactly two cells and V - E + F - T = 1 (a wrong renumbering breaks both);
* the converter prints every coordinate with "%g" (write_vertex, 889-892);
* the C++ reader of the dmri-solve CLI and the Python reader agree bit for bit.
"""
import os
import subprocess
import zipfile

import numpy as np
import pytest

import DmriFemLib as dfl
import dmri_mesh as dm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "dmri-fem-cloud_amd", "lib", "dmri-solve")


def _golden(name):
    return os.path.join(GOLDEN, name)


def _shoelace(pts):
    c = pts.mean(axis=0)
    p = pts[np.argsort(np.arctan2(pts[:, 1] - c[1], pts[:, 0] - c[0]))]
    x, y = p[:, 0], p[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


@pytest.mark.parametrize("name,sides,nv,nc", [("cyl6_r_3E_6_vol.msh.zip", 6, 54, 123),
                                              ("cyl12_r_3E_6_vol.msh.zip", 12, 178, 567)])
def test_prism_meshes_fill_their_prism(name, sides, nv, nc):
    m = dm.read_gmsh(_golden(name))
    assert (m.nv, m.nc) == (nv, nc)
    assert m.tets.min() == 0 and m.tets.max() == nv - 1
    vols = dm.cell_volumes(m.xyz, m.tets)
    assert vols.min() > 1e-3 * vols.mean()
    z0, z1 = m.xyz[:, 2].min(), m.xyz[:, 2].max()
    r = np.hypot(m.xyz[:, 0], m.xyz[:, 1])
    corners = m.xyz[(m.xyz[:, 2] == z0) & (r > 0.99 * r.max())][:, :2]
    assert corners.shape[0] == sides
    expect = _shoelace(corners) * (z1 - z0)
    assert abs(vols.sum() - expect) <= 1e-12 * expect


def _facets(tets):
    f = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]])
    return np.unique(np.sort(f, axis=1), axis=0, return_counts=True)


def test_axon_mesh_is_conforming_and_ball_like():
    m = dm.read_gmsh(_golden("undyAxon_Ax_1E_6_vol.msh.zip"))
    assert (m.nv, m.nc) == (4687, 22172)
    tri, count = _facets(m.tets)
    assert set(np.unique(count)) == {1, 2}
    e = np.concatenate([m.tets[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    edges = np.unique(np.sort(e, axis=1), axis=0)
    assert m.nv - edges.shape[0] + tri.shape[0] - m.nc == 1
    # the boundary is a closed surface: its Euler characteristic is 2
    bt = tri[count == 1]
    be = np.unique(np.sort(np.concatenate([bt[:, [0, 1]], bt[:, [1, 2]], bt[:, [0, 2]]]), axis=1), axis=0)
    assert np.unique(bt).size - be.shape[0] + bt.shape[0] == 2


def test_coordinates_pass_through_the_converters_format():
    m = dm.read_gmsh(_golden("cyl6_r_3E_6_vol.msh.zip"))
    full = dm.read_gmsh(_golden("cyl6_r_3E_6_vol.msh.zip"), dolfin_convert=False)
    # node 7 of the file: "7 9.29852441e-16 1.33226763e-15 0"
    assert m.xyz[6, 0] == 9.29852e-16 and full.xyz[6, 0] == 9.29852441e-16
    ax = dm.read_gmsh(_golden("undyAxon_Ax_1E_6_vol.msh.zip"))
    axf = dm.read_gmsh(_golden("undyAxon_Ax_1E_6_vol.msh.zip"), dolfin_convert=False)
    assert np.array_equal(ax.tets, axf.tets)
    assert all(float("%g" % v) == v for v in ax.xyz.ravel()[:3000])
    rel = np.abs(ax.xyz - axf.xyz) / np.maximum(np.abs(axf.xyz), 1e-300)
    assert rel.max() <= 5e-6 and (ax.xyz != axf.xyz).any()
    # vertex i is the i-th node line of the file (every node is used by a tetrahedron)
    with zipfile.ZipFile(_golden("undyAxon_Ax_1E_6_vol.msh.zip")) as zf:
        lines = zf.read(zf.namelist()[0]).decode().splitlines()
    i = lines.index("$Nodes")
    for k in (0, 1, 2345, 4686):
        vals = lines[i + 2 + k].split()
        assert int(vals[0]) == k + 1
        assert tuple(ax.xyz[k]) == tuple(float("%g" % float(v)) for v in vals[1:])


def test_dmrifemlib_mesh_reads_gmsh():
    a = dfl.Mesh(_golden("cyl12_r_3E_6_vol.msh.zip"))
    b = dm.read_gmsh(_golden("cyl12_r_3E_6_vol.msh.zip"))
    assert np.array_equal(a.xyz, b.xyz) and np.array_equal(a.tets, b.tets)


def test_gmsh_reader_rejects_binary_and_triangle_files(tmp_path):
    p = tmp_path / "bin.msh"
    p.write_text("$MeshFormat\n2.2 1 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError):
        dm.read_gmsh(str(p))
    p = tmp_path / "tri.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$EndNodes\n"
                 "$Elements\n1\n1 2 2 0 1 1 2 3\n$EndElements\n")
    with pytest.raises(ValueError):
        dm.read_gmsh(str(p))


@pytest.mark.skipif(not os.path.exists(CLI), reason="dmri-solve not built")
def test_cli_convert_matches_python_reader(tmp_path):
    with zipfile.ZipFile(_golden("undyAxon_Ax_1E_6_vol.msh.zip")) as zf:
        zf.extractall(tmp_path)
    msh = [f for f in os.listdir(tmp_path) if f.endswith(".msh")][0]
    r = subprocess.run([CLI, "-f", msh, "-convert", "axon.xml"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    a = dm.read_dolfin_xml(str(tmp_path / "axon.xml"))
    b = dm.read_gmsh(_golden("undyAxon_Ax_1E_6_vol.msh.zip"))
    assert np.array_equal(a.xyz, b.xyz) and np.array_equal(a.tets, b.tets)
