"""The dmri-solve CLI's host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only: the
cases below stop before any device call): argument parsing, the DOLFIN XML and gmsh readers on the
reference's gmsh files and on malformed inputs, the phase-file reader and the -convert writer.
Every malformed input must end in a clean error exit, never a sanitizer report or an uncaught
exception."""
import os
import shutil
import subprocess
import zipfile

import numpy as np
import pytest

import dmri_mesh as dm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=87")


@pytest.fixture(scope="module")
def cli(tmp_path_factory):
    import build_native

    if not os.path.exists(build_native.libdmri_path()):
        pytest.skip("libdmri.so is not built")
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    out = str(tmp_path_factory.mktemp("asan") / "dmri-solve-asan")
    try:
        build_native.build_cli_sanitized(out)
    except RuntimeError as e:
        if "asan" in str(e).lower() or "sanitize" in str(e).lower():
            pytest.skip("the compiler has no sanitizer runtime")
        raise
    return out


def _run(cli, args, cwd):
    r = subprocess.run([cli] + [str(a) for a in args], capture_output=True, text=True, cwd=cwd, env=ENV,
                       timeout=120)
    report = ("AddressSanitizer" in r.stderr or "runtime error:" in r.stderr or "LeakSanitizer" in r.stderr
              or "terminate called" in r.stderr)
    assert not report, r.stderr
    assert r.returncode not in (86, 87) and r.returncode >= 0, (r.returncode, r.stderr)
    return r


def test_argument_errors(cli, tmp_path):
    for args, code in ((["-h"], 0), ([], 2), (["-f"], 2), (["-gdir", "1", "0"], 2), (["-bogus"], 2),
                       (["-f", "m.xml", "-r", "1"], 2), (["-f", "m.xml", "-variant", "x"], 2),
                       (["-f", "m.xml", "-s", "1", "-j", "0"], 2), (["-f", "missing.xml"], 1),
                       (["-f", "m.xml", "-gdir", "0", "0", "0"], 1), (["-f", "m.xml", "-d", "9", "-D", "1"], 1)):
        r = _run(cli, args, tmp_path)
        assert r.returncode == code, (args, r.returncode, r.stderr)


def test_gmsh_to_xml_conversion_round_trip(cli, tmp_path):
    with zipfile.ZipFile(os.path.join(GOLDEN, "undyAxon_Ax_1E_6_vol.msh.zip")) as zf:
        zf.extractall(tmp_path)
    msh = [f for f in os.listdir(tmp_path) if f.endswith(".msh")][0]
    r = _run(cli, ["-f", msh, "-convert", "a.xml"], tmp_path)
    assert r.returncode == 0 and "4687 vertices, 22172 tetrahedra" in r.stdout
    r = _run(cli, ["-f", "a.xml", "-convert", "b.xml"], tmp_path)
    assert r.returncode == 0
    assert (tmp_path / "a.xml").read_text() == (tmp_path / "b.xml").read_text()
    ref = dm.read_gmsh(os.path.join(GOLDEN, "undyAxon_Ax_1E_6_vol.msh.zip"))
    got = dm.read_dolfin_xml(str(tmp_path / "a.xml"))
    np.testing.assert_array_equal(got.xyz, ref.xyz)
    np.testing.assert_array_equal(got.tets, ref.tets)


GOOD_MSH = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
5
1 0 0 0
2 1 0 0
3 0 1 0
4 0 0 1
5 1 1 1
$EndNodes
$Elements
3
1 2 2 0 1 1 2 3
2 4 2 0 1 1 2 3 4
3 4 2 0 1 2 3 4 5
$EndElements
"""


def _mutations():
    yield "truncated_nodes", GOOD_MSH.split("4 0 0 1")[0]
    yield "node_count_past_end", GOOD_MSH.replace("$Nodes\n5\n", "$Nodes\n900000000000\n")
    yield "element_count_past_end", GOOD_MSH.replace("$Elements\n3\n", "$Elements\n900000000000\n")
    yield "huge_tag_count", GOOD_MSH.replace("2 4 2 0 1 1 2 3 4", "2 4 900000000000000000 0 1 1 2 3 4")
    yield "negative_node_id", GOOD_MSH.replace("5 1 1 1", "-5 1 1 1")
    yield "unknown_node_in_cell", GOOD_MSH.replace("3 4 2 0 1 2 3 4 5", "3 4 2 0 1 2 3 4 77")
    yield "short_cell", GOOD_MSH.replace("3 4 2 0 1 2 3 4 5", "3 4 2 0 1 2 3")
    yield "garbage_coordinate", GOOD_MSH.replace("2 1 0 0", "2 1 zz 0")
    yield "binary_format", GOOD_MSH.replace("2.2 0 8", "2.2 1 8")


@pytest.mark.parametrize("name,text", list(_mutations()))
def test_malformed_gmsh_is_rejected(cli, tmp_path, name, text):
    (tmp_path / "m.msh").write_text(text)
    r = _run(cli, ["-f", "m.msh", "-convert", "o.xml"], tmp_path)
    assert r.returncode == 1 and "cannot read" in r.stderr, (name, r.stderr)


def _msh(ids):
    """GOOD_MSH's mesh with node k (1-based) given the id ids[k - 1]."""
    xyz = ["0 0 0", "1 0 0", "0 1 0", "0 0 1", "1 1 1"]
    nodes = "".join(f"{ids[k]} {xyz[k]}\n" for k in range(5))
    cell = lambda eid, typ, ks: f"{eid} {typ} 2 0 1 " + " ".join(str(ids[k - 1]) for k in ks) + "\n"
    return ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n5\n" + nodes + "$EndNodes\n$Elements\n3\n"
            + cell(1, 2, (1, 2, 3)) + cell(2, 4, (1, 2, 3, 4)) + cell(3, 4, (2, 3, 4, 5)) + "$EndElements\n")


def test_sparse_gmsh_node_ids(cli, tmp_path):
    """Node ids far above their count (read through the hash map) give the same mesh as 1..n."""
    assert _msh([1, 2, 3, 4, 5]) == GOOD_MSH
    (tmp_path / "a.msh").write_text(GOOD_MSH)
    (tmp_path / "b.msh").write_text(_msh([7 * 10**12, 3, 10**15, 9 * 10**11, 42]))
    for f in ("a", "b"):
        r = _run(cli, ["-f", f + ".msh", "-convert", f + ".xml"], tmp_path)
        assert r.returncode == 0, (f, r.stderr)
    assert (tmp_path / "a.xml").read_text() == (tmp_path / "b.xml").read_text()


@pytest.mark.parametrize("name,edit", [
    ("index_past_count", ('<vertex index="3"', '<vertex index="3000000000000"')),
    ("garbage_index", ('<tetrahedron index="0"', '<tetrahedron index="x"')),
    ("garbage_coordinate", ('x="1.0"', 'x="abc"')),
    ("out_of_range_vertex_id", ('v3="7"', 'v3="99999999999999999999"')),
    ("missing_attribute", (' v1="', ' w1="')),
])
def test_malformed_xml_is_rejected(cli, tmp_path, name, edit):
    m = dm.box_mesh((0,) * 3, (1,) * 3, 1, 1, 1)
    dm.write_dolfin_xml(m, str(tmp_path / "m.xml"))
    txt = (tmp_path / "m.xml").read_text()
    assert edit[0] in txt, name
    (tmp_path / "m.xml").write_text(txt.replace(edit[0], edit[1], 1))
    r = _run(cli, ["-f", "m.xml", "-convert", "o.xml"], tmp_path)
    assert r.returncode == 1 and "cannot read" in r.stderr, (name, r.stderr)


@pytest.mark.parametrize("content", ["0 1 x\\n", "0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 0 1 9\\n",
                                     '<value cell_index="zz" value="1" />\\n',
                                     '<value cell_index="999999" value="1" />\\n',
                                     '<value cell_index="0" value="99999999999999999999" />\\n'])
def test_malformed_phase_file_is_rejected(cli, tmp_path, content):
    m = dm.box_mesh((0,) * 3, (1,) * 3, 1, 1, 1)
    dm.write_dolfin_xml(m, str(tmp_path / "m.xml"))
    (tmp_path / "p.txt").write_text(content.replace("\\n", "\n"))
    r = _run(cli, ["-f", "m.xml", "-phase", "p.txt"], tmp_path)
    assert r.returncode == 1 and "-phase" in r.stderr, r.stderr
