"""Seeded tetrahedral mesh generators and mesh utilities.

The reference reads meshes produced by gmsh/mshr/Salome (GCloudDmriSolver.py:150-152,
PreprocessingOneCompt.py, PreprocessingMultiCompt.py) or builds a DOLFIN BoxMesh
(ConvergenceTest.ipynb cell 8).  None of those generators is available offline, so this
module builds the meshes the workloads need, deterministically from a seed:

* ``box_mesh``      -- the DOLFIN BoxMesh layout: vertices x-fastest, each cube split into
                       six tetrahedra around its main diagonal.
* ``cylinder_mesh`` -- a disk triangulated by 2-D Delaunay (scipy) of a seeded, jittered
                       hexagonal point set plus rings of points on the outer radius and on
                       every interface radius, extruded along z into prisms that are split
                       into three tetrahedra each.  Prisms are split by the global vertex
                       order, so neighbouring prisms always agree on a shared quad face.
* ``cylinders_box_mesh`` -- a rectangular box holding seeded non-overlapping cylinders along z,
                       built the same way (2-D Delaunay + extrusion).

Every generator returns a :class:`TetMesh` with FP64 coordinates in micrometres,
int32 vertex indices and, for two compartments, an int32 cell phase in {0, 1}
(phase 1 = inside a cylinder of the interface set, DmriFemLib.py:754-798 takes the
partition marker modulo 2 the same way).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "TetMesh",
    "box_mesh",
    "cylinder_mesh",
    "cylinders_box_mesh",
    "extrude_triangles",
    "cell_volumes",
    "read_dolfin_xml",
    "write_dolfin_xml",
]


@dataclass
class TetMesh:
    xyz: np.ndarray                     # (nv, 3) float64
    tets: np.ndarray                    # (nc, 4) int32
    phase: np.ndarray | None = None     # (nc,) int32 in {0,1}, None = one compartment
    info: dict = field(default_factory=dict)

    @property
    def nv(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nc(self) -> int:
        return int(self.tets.shape[0])

    def volume(self) -> float:
        return float(cell_volumes(self.xyz, self.tets).sum())

    def hminmax(self) -> tuple[float, float]:
        """Minimum and maximum cell size, the cell size being the greatest distance
        between two vertices of the cell (DOLFIN Cell::h, used by MyDomain through
        mymesh.hmin()/hmax(), DmriFemLib.py:588-589)."""
        x = self.xyz[self.tets]  # (nc,4,3)
        h = np.zeros(self.nc)
        for a in range(4):
            for b in range(a + 1, 4):
                h = np.maximum(h, np.linalg.norm(x[:, a] - x[:, b], axis=1))
        return float(h.min()), float(h.max())

    def bbox(self) -> np.ndarray:
        return np.concatenate([self.xyz.min(axis=0), self.xyz.max(axis=0)])


def cell_volumes(xyz: np.ndarray, tets: np.ndarray) -> np.ndarray:
    x = xyz[tets]
    e1, e2, e3 = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]
    return np.abs(np.einsum("ij,ij->i", e1, np.cross(e2, e3))) / 6.0


# --------------------------------------------------------------------------------------
# DOLFIN BoxMesh layout
# --------------------------------------------------------------------------------------
def box_mesh(p0, p1, nx: int, ny: int, nz: int) -> TetMesh:
    """Box [p0, p1] with nx*ny*nz cubes, 6 tetrahedra per cube sharing the diagonal v0-v7,
    vertex index iz*(nx+1)*(ny+1) + iy*(nx+1) + ix (the DOLFIN BoxMesh layout used by
    ConvergenceTest.ipynb cell 8, 216 vertices / 750 cells for n=5)."""
    a, b, c = (float(v) for v in p0)
    d, e, f = (float(v) for v in p1)
    ix = np.arange(nx + 1, dtype=np.float64)
    iy = np.arange(ny + 1, dtype=np.float64)
    iz = np.arange(nz + 1, dtype=np.float64)
    xs = a + ix * (d - a) / nx
    ys = b + iy * (e - b) / ny
    zs = c + iz * (f - c) / nz
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    cells = []
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx):
                v0 = kz * sz + ky * sy + kx
                v1 = v0 + 1
                v2 = v0 + sy
                v3 = v1 + sy
                v4 = v0 + sz
                v5 = v1 + sz
                v6 = v2 + sz
                v7 = v3 + sz
                cells += [
                    (v0, v1, v3, v7),
                    (v0, v1, v7, v5),
                    (v0, v5, v7, v4),
                    (v0, v3, v2, v7),
                    (v0, v6, v4, v7),
                    (v0, v2, v6, v7),
                ]
    return TetMesh(xyz, np.asarray(cells, dtype=np.int32), None, {"kind": "box", "n": (nx, ny, nz)})


# --------------------------------------------------------------------------------------
# 2-D point sets, Delaunay and extrusion
# --------------------------------------------------------------------------------------
def _hex_lattice(xmin, xmax, ymin, ymax, h, rng, jitter):
    dy = h * math.sqrt(3.0) / 2.0
    ny = int(math.ceil((ymax - ymin) / dy)) + 1
    nx = int(math.ceil((xmax - xmin) / h)) + 2
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = xmin + ii * h + (jj % 2) * 0.5 * h
    y = ymin + jj * dy
    pts = np.stack([x.ravel(), y.ravel()], axis=1)
    if jitter > 0:
        pts = pts + rng.uniform(-jitter * h, jitter * h, size=pts.shape)
    return pts


def _ring(cx, cy, r, h, phase0=0.0):
    n = max(8, int(math.ceil(2 * math.pi * r / h)))
    ang = phase0 + 2 * math.pi * np.arange(n) / n
    return np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], axis=1)


def _delaunay(pts):
    from scipy.spatial import Delaunay

    tri = Delaunay(pts, qhull_options="Qbb Qc Qz Q12").simplices.astype(np.int64)
    p = pts[tri]
    area = 0.5 * ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                  (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    keep = np.abs(area) > 1e-12 * max(1.0, float(np.abs(area).max()))
    tri = tri[keep]
    area = area[keep]
    flip = area < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return tri


def extrude_triangles(pts2d: np.ndarray, tri: np.ndarray, z0: float, z1: float, nz: int):
    """Extrude a 2-D triangulation into nz prism layers, each prism split into three
    tetrahedra.  With the bottom triangle sorted (v0 < v1 < v2) the tetrahedra are
    (v0,v1,v2,v2'), (v0,v1,v1',v2'), (v0,v0',v1',v2'): a quad face with bottom vertices
    p < q is always cut along p-q', so the split is conforming."""
    n2 = pts2d.shape[0]
    zs = z0 + (z1 - z0) * np.arange(nz + 1, dtype=np.float64) / nz
    xyz = np.empty(((nz + 1) * n2, 3))
    for k in range(nz + 1):
        xyz[k * n2:(k + 1) * n2, 0:2] = pts2d
        xyz[k * n2:(k + 1) * n2, 2] = zs[k]
    t = np.sort(tri, axis=1)
    tets = []
    for k in range(nz):
        b = t + k * n2
        u = b + n2
        v0, v1, v2 = b[:, 0], b[:, 1], b[:, 2]
        w0, w1, w2 = u[:, 0], u[:, 1], u[:, 2]
        tets.append(np.stack([v0, v1, v2, w2], axis=1))
        tets.append(np.stack([v0, v1, w1, w2], axis=1))
        tets.append(np.stack([v0, w0, w1, w2], axis=1))
    tets = np.concatenate(tets, axis=0)
    # interleave so that the three tets of a prism and prisms of a layer stay close in memory
    return xyz, tets.astype(np.int32)


def _order_points(pts):
    """Order 2-D points along rows of a coarse grid (locality of the vertex numbering)."""
    h = max(np.ptp(pts[:, 0]), np.ptp(pts[:, 1])) / max(1.0, math.sqrt(pts.shape[0]))
    row = np.floor((pts[:, 1] - pts[:, 1].min()) / (2 * h)).astype(np.int64)
    order = np.lexsort((pts[:, 0], row))
    return pts[order]


def _disk_points(radius, h, seed, rings, jitter):
    rng = np.random.default_rng(seed)
    lat = _hex_lattice(-radius, radius, -radius, radius, h, rng, jitter)
    rad = np.hypot(lat[:, 0], lat[:, 1])
    keep = rad < radius - 0.5 * h
    for r in rings:
        keep &= np.abs(rad - r) > 0.5 * h
    pts = [lat[keep], _ring(0.0, 0.0, radius, h)]
    for r in rings:
        pts.append(_ring(0.0, 0.0, r, h, phase0=0.5 * h / r))
    return _order_points(np.concatenate(pts, axis=0))


def cylinder_mesh(radius: float, length: float, target_nv: int = 20000, seed: int = 1000,
                  interface_radii=(), jitter: float = 0.15, h: float | None = None,
                  z0: float | None = None) -> TetMesh:
    """Cylinder of the given radius and length along z (z in [z0, z0+length], default
    centred at 0).  ``interface_radii`` adds point rings on those radii; with a non-empty
    list the mesh gets a phase (1 inside the smallest interface radius, 0 outside),
    i.e. concentric-cylinder two-compartment geometry."""
    rings = tuple(sorted(float(r) for r in interface_radii))

    def count(hh):
        p = _disk_points(radius, hh, seed, rings, jitter)
        nz = max(1, int(round(length / hh)))
        return p, nz, p.shape[0] * (nz + 1)

    if h is None:
        lo, hi = 1e-3 * radius, radius
        for _ in range(40):
            mid = math.sqrt(lo * hi)
            if count(mid)[2] > target_nv:
                lo = mid
            else:
                hi = mid
        h = hi
    pts, nz, _ = count(h)
    tri = _delaunay(pts)
    zb = -0.5 * length if z0 is None else z0
    xyz, tets = extrude_triangles(pts, tri, zb, zb + length, nz)
    phase = None
    if rings:
        c = xyz[tets].mean(axis=1)
        phase = (np.hypot(c[:, 0], c[:, 1]) < rings[0]).astype(np.int32)
    return TetMesh(xyz, tets, phase, {"kind": "cylinder", "radius": radius, "length": length, "h": h,
                                      "nz": nz, "rings": rings, "seed": seed})


def _place_circles(Lx, Ly, n, rmin, rmax, seed, gap, margin):
    rng = np.random.default_rng(seed)
    radii = np.sort(rng.uniform(rmin, rmax, size=n))[::-1]
    centers = []
    for r in radii:
        for _ in range(200000):
            c = rng.uniform([margin + r, margin + r], [Lx - margin - r, Ly - margin - r])
            if all(np.hypot(*(c - cc)) >= r + rr + gap for cc, rr in centers):
                centers.append((c, r))
                break
        else:
            raise RuntimeError("could not place non-overlapping cylinders; lower the radii")
    return [(float(c[0]), float(c[1]), float(r)) for c, r in centers]


def cylinders_box_mesh(Lx=30.0, Ly=30.0, Lz=10.0, ncyl=20, rmin=2.0, rmax=4.0, target_nv=500000,
                       seed=1002, jitter=0.15, gap=None, h=None, circles=None) -> TetMesh:
    """Box [0,Lx]x[0,Ly]x[0,Lz] holding ``ncyl`` seeded non-overlapping cylinders along z.
    Phase 1 inside a cylinder, 0 in the extra-cylindrical space.  The points on x=0/x=Lx and
    y=0/y=Ly are placed identically on opposite faces, so the mesh is periodic-conforming."""

    def build_points(hh):
        nonlocal circles
        g = 2.0 * hh if gap is None else gap
        if circles is None:
            circles = _place_circles(Lx, Ly, ncyl, rmin, rmax, seed, g, g)
        rng = np.random.default_rng(seed + 1)
        lat = _hex_lattice(0.0, Lx, 0.0, Ly, hh, rng, jitter)
        keep = (lat[:, 0] > 0.5 * hh) & (lat[:, 0] < Lx - 0.5 * hh) & (lat[:, 1] > 0.5 * hh) & (lat[:, 1] < Ly - 0.5 * hh)
        for cx, cy, r in circles:
            keep &= np.abs(np.hypot(lat[:, 0] - cx, lat[:, 1] - cy) - r) > 0.5 * hh
        nxb = max(2, int(math.ceil(Lx / hh)))
        nyb = max(2, int(math.ceil(Ly / hh)))
        sx = np.arange(nxb + 1) * (Lx / nxb)
        sy = np.arange(1, nyb) * (Ly / nyb)
        bnd = [np.stack([sx, np.zeros_like(sx)], 1), np.stack([sx, np.full_like(sx, Ly)], 1),
               np.stack([np.zeros_like(sy), sy], 1), np.stack([np.full_like(sy, Lx), sy], 1)]
        pts = [lat[keep]] + bnd + [_ring(cx, cy, r, hh, phase0=0.37) for cx, cy, r in circles]
        return _order_points(np.concatenate(pts, axis=0))

    def count(hh):
        p = build_points(hh)
        nz = max(1, int(round(Lz / hh)))
        return p, nz, p.shape[0] * (nz + 1)

    if h is None:
        lo, hi = 1e-3, 10.0
        for _ in range(30):
            mid = math.sqrt(lo * hi)
            if count(mid)[2] > target_nv:
                lo = mid
            else:
                hi = mid
        h = hi
    pts, nz, _ = count(h)
    tri = _delaunay(pts)
    xyz, tets = extrude_triangles(pts, tri, 0.0, Lz, nz)
    c = xyz[tets].mean(axis=1)
    phase = np.zeros(tets.shape[0], dtype=np.int32)
    for cx, cy, r in circles:
        phase[np.hypot(c[:, 0] - cx, c[:, 1] - cy) < r] = 1
    return TetMesh(xyz, tets, phase, {"kind": "cylinders_box", "L": (Lx, Ly, Lz), "circles": circles,
                                      "h": h, "nz": nz, "seed": seed})


# --------------------------------------------------------------------------------------
# DOLFIN XML mesh files (comri/meshes/*.xml, read by Mesh(fmesh), main.cpp:200)
# --------------------------------------------------------------------------------------
def read_dolfin_xml(path: str) -> TetMesh:
    import re

    with open(path, "r") as fh:
        txt = fh.read()
    vx = re.findall(r'<vertex\s+index="(\d+)"\s+x="([^"]+)"\s+y="([^"]+)"\s+z="([^"]+)"', txt)
    cx = re.findall(r'<tetrahedron\s+index="(\d+)"\s+v0="(\d+)"\s+v1="(\d+)"\s+v2="(\d+)"\s+v3="(\d+)"', txt)
    if not vx or not cx:
        raise ValueError(f"{path}: not a 3-D DOLFIN XML tetrahedral mesh")
    xyz = np.zeros((len(vx), 3))
    for i, x, y, z in vx:
        xyz[int(i)] = (float(x), float(y), float(z))
    tets = np.zeros((len(cx), 4), dtype=np.int32)
    for i, a, b, c, d in cx:
        tets[int(i)] = (int(a), int(b), int(c), int(d))
    return TetMesh(xyz, tets, None, {"kind": "dolfin_xml", "path": path})


def write_dolfin_xml(mesh: TetMesh, path: str) -> None:
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<dolfin xmlns:dolfin="http://fenicsproject.org">\n')
        fh.write('  <mesh celltype="tetrahedron" dim="3">\n')
        fh.write(f'    <vertices size="{mesh.nv}">\n')
        for i, (x, y, z) in enumerate(mesh.xyz.tolist()):
            fh.write(f'      <vertex index="{i}" x="{x!r}" y="{y!r}" z="{z!r}" />\n')
        fh.write('    </vertices>\n')
        fh.write(f'    <cells size="{mesh.nc}">\n')
        for i, (a, b, c, d) in enumerate(mesh.tets):
            fh.write(f'      <tetrahedron index="{i}" v0="{a}" v1="{b}" v2="{c}" v3="{d}" />\n')
        fh.write('    </cells>\n  </mesh>\n</dolfin>\n')
