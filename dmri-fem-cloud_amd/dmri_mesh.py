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
* ``neuron_mesh``   -- a neuron-like surrogate: a soma sphere and seeded branching cylindrical
                       dendrites, voxelised on a uniform grid (every voxel whose centre lies inside
                       the union is kept) and split like ``box_mesh``, so the boundary is
                       stair-stepped.  One compartment.

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
    "neuron_mesh",
    "neuron_segments",
    "extrude_triangles",
    "cell_volumes",
    "read_dolfin_xml",
    "read_gmsh",
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
    """n non-overlapping circles, radii uniform in [rmin, rmax] (seeded), at least ``gap`` apart and
    ``margin`` inside the rectangle.  Random sequential placement first; when that jams (dense
    packings), seeded random centres relaxed by pairwise overlap removal."""
    rng = np.random.default_rng(seed)
    radii = np.sort(rng.uniform(rmin, rmax, size=n))[::-1]
    centers = []
    for r in radii:
        for _ in range(20000):
            c = rng.uniform([margin + r, margin + r], [Lx - margin - r, Ly - margin - r])
            if all(np.hypot(*(c - cc)) >= r + rr + gap for cc, rr in centers):
                centers.append((c, r))
                break
        else:
            break
    if len(centers) == n:
        return [(float(c[0]), float(c[1]), float(r)) for c, r in centers]
    # seeded random centres, radii grown from half size in stages with overlap removal between
    pos = rng.uniform([margin + rmax, margin + rmax], [Lx - margin - rmax, Ly - margin - rmax], size=(n, 2))
    eye = np.eye(n, dtype=bool)
    for s in np.linspace(0.5, 1.0, 26):
        rr = s * radii
        lo = np.stack([margin + rr, margin + rr], axis=1)
        hi = np.stack([Lx - margin - rr, Ly - margin - rr], axis=1)
        need = rr[:, None] + rr[None, :] + gap
        for _ in range(4000):
            pos = np.clip(pos, lo, hi)
            d = pos[:, None, :] - pos[None, :, :]
            dist = np.hypot(d[..., 0], d[..., 1])
            ov = np.where(eye, -1.0, need - dist)
            if ov.max() <= 0.0:
                break
            # push overlapping pairs apart to a little beyond contact so the relaxation ends
            ov = np.where(ov > -1e-3, ov + 2e-3, 0.0)
            unit = d / np.where(dist > 1e-12, dist, 1.0)[..., None]
            pos = pos + 0.5 * (ov[..., None] * unit).sum(axis=1)
        pos = np.clip(pos, lo, hi)
    d = pos[:, None, :] - pos[None, :, :]
    dist = np.hypot(d[..., 0], d[..., 1])
    if np.where(eye, 0.0, radii[:, None] + radii[None, :] + gap - dist).max() > 0.0:
        raise RuntimeError("could not place non-overlapping cylinders; lower the radii")
    return [(float(pos[i, 0]), float(pos[i, 1]), float(radii[i])) for i in range(n)]


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
# --------------------------------------------------------------------------------------
# neuron-like surrogate (voxelised union of a sphere and capsules)
# --------------------------------------------------------------------------------------
def neuron_segments(seed=1004, soma_r=8.0, ndend=6, levels=3, rmin=0.5, rmax=2.0, lmin=20.0, lmax=60.0,
                    branch_angle=35.0):
    """Seeded dendrite tree: ``ndend`` roots leaving the soma along jittered Fibonacci-sphere
    directions, each bifurcating ``levels`` times (2**(levels+1)-1 segments per dendrite); root
    radius uniform in [0.75 rmax, rmax], radius x 0.8 per level (>= rmin), segment lengths uniform
    in [lmin, lmax].  Returns a list of (p0, p1, r) capsules."""
    rng = np.random.default_rng(seed)
    k = np.arange(ndend) + 0.5
    phi = np.arccos(1 - 2 * k / ndend)
    th = math.pi * (1 + 5 ** 0.5) * k
    dirs = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], axis=1)
    dirs += 0.15 * rng.normal(size=dirs.shape)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    segs = []

    def grow(p0, d, r, level):
        L = rng.uniform(lmin, lmax)
        p1 = p0 + L * d
        segs.append((p0, p1, r))
        if level == levels:
            return
        a = rng.normal(size=3)
        a -= a.dot(d) * d
        a /= np.linalg.norm(a)
        ang = math.radians(branch_angle)
        for sgn in (1.0, -1.0):
            c = math.cos(sgn * ang)
            s = math.sin(sgn * ang)
            nd = d * c + np.cross(a, d) * s
            grow(p1, nd / np.linalg.norm(nd), max(rmin, 0.8 * r), level + 1)

    for d in dirs:
        grow((soma_r - 0.5) * d, d, rng.uniform(0.75 * rmax, rmax), 0)
    return segs


def _capsule_voxels(p0, p1, r, h):
    """Integer voxel indices (n,3) whose centres (i+0.5)h lie within r of the segment p0-p1."""
    d = p1 - p0
    L = float(np.linalg.norm(d))
    ax = int(np.argmax(np.abs(d)))
    u, w = (ax + 1) % 3, (ax + 2) % 3
    lo = min(p0[ax], p1[ax]) - r
    hi = max(p0[ax], p1[ax]) + r
    ia = np.arange(int(math.floor(lo / h)), int(math.ceil(hi / h)) + 1)
    ca = (ia + 0.5) * h
    # axis point at each slice (clamped to the segment), the capsule's slice lies within r/|cos| of it
    t = np.clip((ca - p0[ax]) / d[ax], 0.0, 1.0) if abs(d[ax]) > 0 else np.zeros_like(ca)
    cen = p0[None, :] + t[:, None] * d[None, :]
    ext = r / max(abs(d[ax]) / max(L, 1e-300), 0.3) + h
    m = int(math.ceil(ext / h)) + 1
    off = np.arange(-m, m + 1)
    iu = np.floor(cen[:, u] / h).astype(np.int64)[:, None, None] + off[None, :, None]
    iw = np.floor(cen[:, w] / h).astype(np.int64)[:, None, None] + off[None, None, :]
    iu, iw = np.broadcast_arrays(iu, iw)
    ia3 = np.broadcast_to(ia[:, None, None], iu.shape)
    P = np.zeros(iu.shape + (3,))
    P[..., ax] = (ia3 + 0.5) * h
    P[..., u] = (iu + 0.5) * h
    P[..., w] = (iw + 0.5) * h
    q = P - p0
    tt = np.clip((q @ d) / max(L * L, 1e-300), 0.0, 1.0)
    dist2 = np.sum((q - tt[..., None] * d) ** 2, axis=-1)
    keep = dist2 <= r * r
    out = np.zeros((int(keep.sum()), 3), dtype=np.int64)
    out[:, ax] = ia3[keep]
    out[:, u] = iu[keep]
    out[:, w] = iw[keep]
    return out


def _neuron_voxels(h, soma_r, segs):
    m = int(math.ceil(soma_r / h)) + 1
    g = np.arange(-m, m + 1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    C = (np.stack([I, J, K], axis=-1) + 0.5) * h
    soma = np.stack([I, J, K], axis=-1)[np.sum(C * C, axis=-1) <= soma_r * soma_r]
    parts = [soma.reshape(-1, 3).astype(np.int64)]
    for p0, p1, r in segs:
        parts.append(_capsule_voxels(np.asarray(p0, float), np.asarray(p1, float), float(r), h))
    vox = np.concatenate(parts, axis=0)
    return np.unique(vox, axis=0)


def neuron_mesh(target_nv: int = 2_000_000, seed: int = 1004, soma_r: float = 8.0, ndend: int = 6, levels: int = 3,
                h: float | None = None, **tree) -> TetMesh:
    """Neuron-like surrogate of about ``target_nv`` vertices: soma sphere of radius ``soma_r``
    centred at the origin plus the dendrites of :func:`neuron_segments`, voxelised with edge h
    (chosen for the vertex target when not given) and split into six tetrahedra per voxel."""
    segs = neuron_segments(seed=seed, soma_r=soma_r, ndend=ndend, levels=levels, **tree)
    corner = np.array([[dx, dy, dz] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)], dtype=np.int64)

    def corners(vox):
        lo = vox.min(axis=0)
        ext = vox.max(axis=0) - lo + 2
        cid = (vox[:, None, :] - lo + corner[None, :, :])  # (nvox, 8, 3) corner index triples
        key = cid[..., 0] + ext[0] * (cid[..., 1] + ext[1] * cid[..., 2])
        return lo, ext, key

    if h is None:
        h = 0.6
        for _ in range(4):  # vertices scale like h^-3 (a little faster on thin branches)
            nvert = np.unique(corners(_neuron_voxels(h, soma_r, segs))[2].ravel()).shape[0]
            h *= (nvert / target_nv) ** (1.0 / 3.0)
    vox = _neuron_voxels(h, soma_r, segs)
    lo, ext, key = corners(vox)
    ukey, inv = np.unique(key.ravel(), return_inverse=True)
    inv = inv.reshape(-1, 8).astype(np.int64)
    iz, rem = np.divmod(ukey, ext[0] * ext[1])
    iy, ix = np.divmod(rem, ext[0])
    xyz = (np.stack([ix, iy, iz], axis=1) + lo).astype(np.float64) * h
    v = [inv[:, i] for i in range(8)]
    tets = np.stack([np.stack(t, axis=1) for t in (
        (v[0], v[1], v[3], v[7]), (v[0], v[1], v[7], v[5]), (v[0], v[5], v[7], v[4]),
        (v[0], v[3], v[2], v[7]), (v[0], v[6], v[4], v[7]), (v[0], v[2], v[6], v[7]))], axis=1).reshape(-1, 4)
    return TetMesh(xyz, tets.astype(np.int32), None,
                   {"kind": "neuron", "seed": seed, "h": h, "nvox": int(vox.shape[0]), "segments": len(segs)})


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


def _gmsh_text(path: str) -> str:
    if path.endswith(".zip"):
        import zipfile

        with zipfile.ZipFile(path) as zf:
            names = [n for n in zf.namelist() if n.endswith((".msh", ".gmsh")) and not n.startswith("__MACOSX")]
            if len(names) != 1:
                raise ValueError(f"{path}: expected one .msh file in the archive, found {names}")
            return zf.read(names[0]).decode("ascii")
    with open(path, "r") as fh:
        return fh.read()


def read_gmsh(path: str, dolfin_convert: bool = True) -> TetMesh:
    """Read an ASCII gmsh 2.x tetrahedral mesh (``.msh``, or a ``.zip`` holding one) as the
    reference's converter turns it into the DOLFIN XML mesh its solvers load
    (comri/meshes/neuron_download/dolfin-convert.py:357-597, gmsh2xml):

    * the cells are the tetrahedra (element type 4) in file order; other elements are skipped;
    * the vertices are the nodes that some tetrahedron uses, numbered in their ``$Nodes`` order
      (the converter's set-ordered ``vertex_dict`` is undone by its ``nodelist``, 449-453, 529, 560-565,
      so the final numbering does not depend on the set's iteration order);
    * each coordinate passes through the XML writer's ``"%g"`` (6 significant digits,
      write_vertex 889-892), so the mesh equals the one DOLFIN reads back from the converted file.
      ``dolfin_convert=False`` keeps the file's digits instead.

    The converter takes a tetrahedron's nodes as ``element[3 + ntags:9]``, i.e. it needs the two
    tags gmsh writes by default; here the four nodes after the tags are taken for any tag count.
    ``info["physical"]`` holds each cell's first tag (the gmsh physical group; the converter
    drops it).  Binary files and meshes without tetrahedra raise ValueError.
    """
    lines = _gmsh_text(path).splitlines()

    def block(name):
        try:
            i = lines.index("$" + name)
            j = lines.index("$End" + name, i)
        except ValueError:
            raise ValueError(f"{path}: no ${name} block") from None
        return i, j

    i, _ = block("MeshFormat")
    fmt = lines[i + 1].split()
    if len(fmt) != 3 or not fmt[0].startswith("2") or fmt[1] != "0":
        raise ValueError(f"{path}: not an ASCII gmsh 2.x file (format line {lines[i + 1]!r})")
    i, j = block("Nodes")
    nn = int(lines[i + 1])
    if j - i - 2 != nn:
        raise ValueError(f"{path}: $Nodes declares {nn} nodes but holds {j - i - 2} lines")
    tok = np.array(" ".join(lines[i + 2:j]).split(), dtype=object).reshape(nn, 4)
    node_id = tok[:, 0].astype(np.int64)
    if dolfin_convert:
        xyz_all = np.array([[float("%g" % float(v)) for v in row] for row in tok[:, 1:]], dtype=np.float64)
    else:
        xyz_all = tok[:, 1:].astype(np.float64)
    i, j = block("Elements")
    ne = int(lines[i + 1])
    if j - i - 2 != ne:
        raise ValueError(f"{path}: $Elements declares {ne} elements but holds {j - i - 2} lines")
    tets, phys = [], []
    for line in lines[i + 2:j]:
        e = line.split()
        if int(e[1]) != 4:
            continue
        nt = int(e[2])
        tets.append([int(v) for v in e[3 + nt:7 + nt]])
        phys.append(int(e[3]) if nt > 0 else 0)
    if not tets:
        raise ValueError(f"{path}: no tetrahedra (element type 4)")
    tets = np.asarray(tets, dtype=np.int64)
    # used nodes, renumbered in $Nodes order
    pos = np.full(int(node_id.max()) + 1, -1, dtype=np.int64)
    pos[node_id] = np.arange(nn)
    if tets.max() >= pos.size or (pos[tets] < 0).any():
        raise ValueError(f"{path}: a tetrahedron names a node that $Nodes does not define")
    used = np.zeros(nn, dtype=bool)
    used[pos[tets].ravel()] = True
    newidx = np.cumsum(used) - 1
    xyz = np.ascontiguousarray(xyz_all[used])
    tv = newidx[pos[tets]].astype(np.int32)
    return TetMesh(xyz, tv, None, {"kind": "gmsh", "path": path, "physical": np.asarray(phys, dtype=np.int32),
                                   "dolfin_convert": bool(dolfin_convert)})


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
