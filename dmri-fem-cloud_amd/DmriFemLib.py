"""DmriFemLib-compatible Python API over libdmri (the MI355X solver).

Mirrors the public surface of the reference library DmriFemLib.py so that driver code
written for it runs unchanged on the GPU solver:

    from DmriFemLib import *
    mymesh = BoxMesh(Point(-2.5, -2.5, -2.5), Point(2.5, 2.5, 2.5), 16, 16, 16)   # or Mesh("file.xml")
    mri_para = MRI_parameters(); mri_para.bvalue = 1000; mri_para.delta, mri_para.Delta = 10600, 43100
    mri_para.set_gradient_dir(mymesh, 1, 0, 0); mri_para.T = mri_para.Delta + mri_para.delta
    mri_para.fs_sym = sp.Piecewise((1., mri_para.s < mri_para.delta), (0., mri_para.s < mri_para.Delta),
                                   (-1., mri_para.s < mri_para.T), (0., True))
    mri_para.Apply()
    mri_simu = MRI_simulation(); mri_simu.k = 200
    mydomain = MyDomain(mymesh, mri_para); mydomain.phase = None; mydomain.PeriodicDir = [0, 0, 0]
    mydomain.IsDomainPeriodic = False; mydomain.IsDomainMultiple = False; mydomain.Apply()
    mydomain.D = 3e-3
    linsolver = KrylovSolver("bicgstab", "jacobi"); linsolver.parameters["relative_tolerance"] = 1e-12
    mri_simu.solve(mydomain, mri_para, linsolver)
    PostProcessing(mydomain, mri_para, mri_simu, None)

Reference behaviour kept (file:line in DmriFemLib.py of the reference):
  * MRI_parameters: defaults gdir (1,0,0), T2 1e16 (800-808); |g|=0 exits (809-824);
    q = sqrt(b / int_0^T F^2) from the sympy profile (826-842); bvalue or gvalue (847-858).
  * MRI_simulation: theta 0.5, nskip 5 (860-863); time loop while t < T + k with t += k and
    the right-hand side profile lagged one step (878-911); u0 = 1 on both compartments (865-876).
  * PostProcessing: signal = int (1-p) u0r + p u1r / int u(0) and the result lines and log.txt
    append of 917-988.
The PDE is solved by libdmri; this module only prepares inputs and formats outputs.
Not mirrored: dolfin plotting and .pvd output (MRI_simulation.save_snapshots keeps u every nskip
steps as arrays instead).
"""
from __future__ import annotations

import math
import sys
import time

import numpy as np
import sympy as sp

import dmri_mesh
import dmri_native

__all__ = ["sp", "Point", "Mesh", "BoxMesh", "MRI_parameters", "MRI_simulation", "MyDomain", "KrylovSolver",
           "PETScLUSolver", "PostProcessing", "convert_g2q", "convert_q2g", "CreatePhaseFunc", "MPI",
           "InterfaceKappa"]

G_RATIO = 2.675e8


class MPI:
    """Single-process stand-in for dolfin.MPI used by driver scripts (rank 0 only)."""

    class _Comm:
        def Get_rank(self):
            return 0

        def Get_size(self):
            return 1

    comm_world = _Comm()

    @staticmethod
    def min(comm, v):
        return v

    @staticmethod
    def max(comm, v):
        return v


class Point:
    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.v = np.array([float(x), float(y), float(z)])

    def x(self):
        return self.v[0]

    def y(self):
        return self.v[1]

    def z(self):
        return self.v[2]

    def norm(self):
        return float(np.linalg.norm(self.v))

    def __itruediv__(self, s):
        self.v = self.v / s
        return self


def Mesh(path_or_mesh):
    """Mesh from a DOLFIN XML file (Mesh(fmesh), comri main.cpp:200), a TetMesh, or an ASCII
    gmsh 2.x file (``.msh``, ``.msh.zip``) taken as dolfin-convert would convert it to XML
    (comri/meshes/neuron_download/dolfin-convert.py:357-597)."""
    if isinstance(path_or_mesh, dmri_mesh.TetMesh):
        return path_or_mesh
    path = str(path_or_mesh)
    if path.endswith((".msh", ".gmsh", ".msh.zip", ".gmsh.zip")):
        return dmri_mesh.read_gmsh(path)
    return dmri_mesh.read_dolfin_xml(path)


def BoxMesh(p0, p1, nx, ny, nz):
    return dmri_mesh.box_mesh(p0.v, p1.v, nx, ny, nz)


def convert_g2q(gvalue):
    return gvalue * G_RATIO * 1e-12


def convert_q2g(qvalue):
    return qvalue / G_RATIO * 1e12


class KrylovSolver:
    """Solver settings holder (KrylovSolver("bicgstab", "jacobi") in GCloudDmriSolver.py:219-222;
    KrylovSolver("gmres", ...) in the notebooks).  "gmres" runs libdmri's restarted GMRES, "cocg" its
    COCG (complex-symmetric operators only), every other method its BiCGStab; all with libdmri's
    Jacobi-scaled preconditioner and converged to the relative tolerance (the preconditioner name is
    kept for compatibility)."""

    def __init__(self, method="bicgstab", pc="jacobi"):
        self.method, self.pc = method, pc
        # "restart": Arnoldi steps per GMRES cycle (PETSc's default 30), unused by BiCGStab
        self.parameters = {"relative_tolerance": 1e-9, "absolute_tolerance": 1e-10, "maximum_iterations": 100000,
                           "nonzero_initial_guess": False, "restart": 30}


class PETScLUSolver(KrylovSolver):
    """A direct solve (PETScLUSolver("mumps")) is replaced by BiCGStab converged to rtol 1e-14,
    which reproduces the direct solution to far below the 1e-8 signal bar."""

    def __init__(self, method="mumps"):
        super().__init__("bicgstab", "jacobi")
        self.parameters["relative_tolerance"] = 1e-14
        self.parameters["absolute_tolerance"] = 0.0


def CreatePhaseFunc(mymesh, evengroup, oddgroup, partition_marker):
    """Phase from a partition marker (marker % 2, DmriFemLib.py:758-765)."""
    if partition_marker is None:
        raise NotImplementedError("phase from sub-meshes needs point location; pass a partition marker")
    pm = np.asarray(partition_marker)
    phase = (pm % 2).astype(np.int32)
    return phase, sorted(set(int(v) for v in pm))


class InterfaceKappa:
    """A permeability per interface facet, in place of the CR function of the reference's variable-
    permeability example (MultilayeredDiskVariablePermeability.ipynb, cell 4: facet_f[facet] =
    kappa_tensor[marker[cell 0]][marker[cell 1]] over the facets whose two cells carry different
    partition markers, copied into a CR Function's facet dofs and set as mydomain.kappa).

    InterfaceKappa(kappa_tensor, partition_marker): that rule, with the two cells of every interface
    facet found from the mesh.  mydomain.kappa may also be a callable of the (n, 3) facet midpoints
    (a CR interpolant's dofs are its values there) or an array with one value per interface facet in
    dmri_export_facets order.  Each facet's value is taken as constant over the facet, which is the
    CR function's restriction when no cell beside it has a second interface facet of another value.
    """

    def __init__(self, kappa_tensor, partition_marker):
        self.kappa_tensor = np.asarray(kappa_tensor, dtype=np.float64)
        self.partition_marker = np.asarray(partition_marker).ravel()


def facet_cells(tets, fv):
    """The two cells of each facet fv[i] (vertex triples) of the tetrahedral mesh tets: (n, 2) cell
    indices, lower index first; -1 for a facet on the boundary; ValueError for a facet not in the mesh."""
    tets = np.asarray(tets, dtype=np.int64)
    nc = tets.shape[0]
    faces = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]])
    owner = np.tile(np.arange(nc), 4)
    faces.sort(axis=1)
    nv = int(tets.max()) + 1
    key = (faces[:, 0] * nv + faces[:, 1]) * nv + faces[:, 2]
    order = np.lexsort((owner, key))
    key, owner = key[order], owner[order]
    q = np.sort(np.asarray(fv, dtype=np.int64), axis=1)
    qk = (q[:, 0] * nv + q[:, 1]) * nv + q[:, 2]
    lo = np.searchsorted(key, qk, side="left")
    hi = np.searchsorted(key, qk, side="right")
    if np.any(hi == lo):
        raise ValueError("a facet is not a face of the mesh")
    out = np.full((q.shape[0], 2), -1, dtype=np.int64)
    out[:, 0] = owner[lo]
    two = hi - lo >= 2
    out[two, 1] = owner[lo[two] + 1]
    return out


def interface_kappa_values(kappa, xyz, tets, fv):
    """mydomain.kappa (InterfaceKappa, callable or per-facet array) as one value per interface facet
    fv (dmri_export_facets order)."""
    fv = np.asarray(fv)
    if isinstance(kappa, InterfaceKappa):
        cells = facet_cells(tets, fv)
        if np.any(cells[:, 1] < 0):
            raise ValueError("an interface facet with a single cell")
        m = kappa.partition_marker
        if m.shape[0] != np.asarray(tets).shape[0]:
            raise ValueError("InterfaceKappa needs one partition marker per cell")
        return kappa.kappa_tensor[m[cells[:, 0]], m[cells[:, 1]]].astype(np.float64)
    if callable(kappa):
        vals = np.asarray(kappa(np.asarray(xyz)[fv].mean(axis=1)), dtype=np.float64).ravel()
    else:
        vals = np.asarray(kappa, dtype=np.float64).ravel()
    if vals.shape[0] != fv.shape[0]:
        raise ValueError("mydomain.kappa needs one value per interface facet (%d), got %d"
                         % (fv.shape[0], vals.shape[0]))
    return vals


class MRI_parameters:
    def __init__(self):
        self.bvalue = None
        self.gvalue = None
        self.gdir = [1, 0, 0]
        self.nperiod = 0
        self.T2 = 1e16
        self.s = sp.Symbol("s")

    def set_gradient_dir(self, mymesh, g0, g1, g2):
        g = np.array([float(g0), float(g1), float(g2)])
        n = float(np.linalg.norm(g))
        if n > 1e-10:
            g = g / n
        else:
            print("|g|=0! Please check again the gradient directions!")
            sys.exit()
        self.gdir = Point(*g)

    def itime_profile_sym(self):
        u = sp.Symbol("u")
        self.ifs_sym = sp.integrate(self.fs_sym.subs(self.s, u), (u, 0, self.s))

    def integral_term_for_gb(self):
        self.int4gb = float(sp.integrate(self.ifs_sym * self.ifs_sym, (self.s, 0, self.T)))

    def time_profile(self, t):
        return float(self.fs_sym.subs(self.s, t))

    def itime_profile(self, t):
        return float(self.ifs_sym.subs(self.s, t))

    def convert_b2q(self):
        self.qvalue = math.sqrt(self.bvalue) / math.sqrt(self.int4gb)
        return self.qvalue

    def convert_q2b(self):
        self.bvalue = self.qvalue * self.qvalue * self.int4gb
        return self.bvalue

    def Apply(self):
        self.itime_profile_sym()
        self.integral_term_for_gb()
        if self.bvalue is not None:
            self.qvalue = self.convert_b2q()
            self.gvalue = convert_q2g(self.qvalue)
        elif self.gvalue is not None:
            self.qvalue = convert_g2q(self.gvalue)
            self.bvalue = self.convert_q2b()
        else:
            print("bvalue or gvalue need to be specified.")
            sys.exit()


class MyDomain:
    def __init__(self, mymesh, mri_para):
        self.mymesh = mymesh
        self.porder = 1
        self.hmin, self.hmax = mymesh.hminmax()
        self.tol = 1e-2 * self.hmin
        self.gdim = self.tdim = 3
        bb = mymesh.bbox()
        self.xmin, self.ymin, self.zmin, self.xmax, self.ymax, self.zmax = (float(v) for v in bb)
        print("Domain size: xmin=%f, ymin=%f, zmin=%f, xmax=%f, ymax=%f, zmax=%f" % tuple(bb))
        self.gdir = mri_para.gdir
        self.qvalue = getattr(mri_para, "qvalue", None)
        self.kappa_e_scalar = 3e-3 / self.hmin
        self.phase = None
        self.PeriodicDir = [0, 0, 0]
        self.IsDomainPeriodic = False
        self.IsDomainMultiple = False
        self.kappa = 0.0
        self.D = None
        self.D0 = None

    def ImposeDiffusionTensor(self, k00, k01, k02, k10, k11, k12, k20, k21, k22):
        print("Impose Diffusion Tensor ...")
        nc = self.mymesh.nc
        comps = [np.broadcast_to(np.asarray(c, dtype=np.float64), (nc,)) for c in
                 (k00, k01, k02, k10, k11, k12, k20, k21, k22)]
        self.D = np.stack(comps, axis=1)

    def Apply(self):
        # PeriodicDir with IsDomainPeriodic = False: the weak pseudo-periodic condition; with
        # IsDomainPeriodic = True the strong one on the periodic (constrained) space (MyDomain.Apply,
        # DmriFemLib.py:618-637; PeriodicBD 327-375); both run in libdmri
        if sum(self.PeriodicDir) > 0:
            print("The pseudo-periodic BCS are strongly imposed." if self.IsDomainPeriodic else
                  "The pseudo-periodic BCS are weakly imposed.")
        if self.IsDomainMultiple:
            if self.phase is None:
                raise ValueError("IsDomainMultiple=True needs mydomain.phase (per-cell 0/1)")
            print("Function Space for Two-compartment Domains has 4 components")
        else:
            print("Function Space for Single Domains has 2 components")


class MRI_simulation:
    def __init__(self):
        self.nskip = 5
        self.theta = 0.5
        self.device = 0
        # True: keep u after every step whose counter is a multiple of nskip (the reference prints
        # its progress at those steps, DmriFemLib.py:898) in self.snapshots, shape
        # (nsnap, ncomp, nv) complex per compartment and vertex (0 where the compartment has no DOF),
        # with the 0-based step indices in self.snapshot_steps
        self.save_snapshots = False

    def _sequence(self, mri_para, integral=False):
        """f(t_n), f(t_{n-1}) and, for the periodic terms, F(t_n) = itime_profile(t) and F(t_{n-1}) =
        itime_profile(tp) (MRI_simulation.solve, DmriFemLib.py:897-910)."""
        fA, fB, FA, FB = [], [], [], []
        t, tp = 0.0, 0.0
        while t < mri_para.T + self.k:
            fA.append(mri_para.time_profile(t))
            fB.append(mri_para.time_profile(tp))
            if integral:
                FA.append(mri_para.itime_profile(t))
                FB.append(mri_para.itime_profile(tp))
            tp = t
            t += self.k
        if not integral:
            return np.asarray(fA), np.asarray(fB), None, None
        return np.asarray(fA), np.asarray(fB), np.asarray(FB), np.asarray(FA)

    def InitialCondition(self, mydomain, ic):
        """Initial state per mesh vertex (MRI_simulation.InitialCondition, DmriFemLib.py:865-876): the
        real part of every compartment takes it, the imaginary part is 0.  ic: None (1 everywhere),
        an array of one value per vertex, or a callable of the (nv, 3) vertex coordinates."""
        xyz = mydomain.mymesh.xyz
        if ic is None:
            return None
        if callable(ic):
            u0 = np.asarray(ic(xyz), dtype=np.float64).ravel()
        else:
            u0 = np.asarray(ic, dtype=np.float64).ravel()
        if u0.shape[0] != xyz.shape[0]:
            raise ValueError("the initial condition needs one value per mesh vertex (%d), got %d"
                             % (xyz.shape[0], u0.shape[0]))
        return u0

    def solve(self, mydomain, mri_para, linsolver, ic=None):
        self.u0_vertex = self.InitialCondition(mydomain, ic)
        mesh = mydomain.mymesh
        phase = None
        if mydomain.IsDomainMultiple:
            phase = np.asarray(mydomain.phase, dtype=np.int32).ravel()
        D = mydomain.D if mydomain.D is not None else mydomain.D0
        if D is None:
            raise ValueError("set mydomain.D (or call ImposeDiffusionTensor) before solve")
        invT2 = 1.0 / float(mri_para.T2) if np.isscalar(mri_para.T2) else None
        invT2_cell = None if invT2 is not None else 1.0 / np.asarray(mri_para.T2, dtype=np.float64)
        ctx = dmri_native.Context(self.device)
        ctx.set_mesh(mesh.xyz, mesh.tets, phase)
        kappa_field = not np.isscalar(mydomain.kappa)
        coeffs = dict(invT2=invT2 if invT2 is not None else 0.0, invT2_cell=invT2_cell,
                      kappa=0.0 if kappa_field else float(mydomain.kappa))
        if np.isscalar(D):
            coeffs.update(D=float(D))
        else:
            coeffs.update(D=0.0, D_cell=np.asarray(D))
        ctx.set_coeffs(**coeffs)
        periodic = sum(mydomain.PeriodicDir) > 0
        if periodic and mydomain.IsDomainPeriodic:
            ctx.set_periodic_strong(mydomain.PeriodicDir)
        else:
            ctx.set_periodic(mydomain.PeriodicDir if periodic else (0, 0, 0))
        ctx.assemble()
        if kappa_field:
            # the interface facets are known after the first assembly; the second takes their values
            kf = interface_kappa_values(mydomain.kappa, mesh.xyz, mesh.tets, ctx.export_facets())
            ctx.set_coeffs(kappa_facet=kf, **coeffs)
            ctx.assemble()
        fA, fB, FB, FA = self._sequence(mri_para, integral=periodic)
        p = linsolver.parameters
        start = time.time()
        g = mri_para.gdir.v if isinstance(mri_para.gdir, Point) else np.asarray(mri_para.gdir, dtype=np.float64)
        self.signal_c, self.stats = ctx.solve(fA, fB, self.k, mri_para.qvalue, g, theta=self.theta,
                                              rtol=p["relative_tolerance"], atol=p["absolute_tolerance"],
                                              maxit=min(int(p["maximum_iterations"]), 2 ** 31 - 1),
                                              F_explicit=FB, F_implicit=FA, u0=self.u0_vertex,
                                              krylov={"gmres": 1, "cocg": 2}.get(str(linsolver.method).lower(), 0),
                                              restart=int(p.get("restart", 30)),
                                              snap_every=self.nskip if self.save_snapshots else 0)
        self.elapsed_time = time.time() - start
        if self.save_snapshots:
            snaps = self.stats.pop("snapshots")
            _, _, dof_of = ctx.export_pattern()
            nv = mesh.nv
            ncomp = dof_of.shape[0] // nv
            out = np.zeros((snaps.shape[0], ncomp, nv), dtype=np.complex128)
            for c in range(ncomp):
                d = dof_of[c * nv:(c + 1) * nv]
                out[:, c, d >= 0] = snaps[:, d[d >= 0]]
            self.snapshots = out
            self.snapshot_steps = list(range(0, len(fA), self.nskip))
        self.t = len(fA) * self.k
        self.nsteps = len(fA)
        self.ctx = ctx
        print("Successfully Completed! Elapsed time: %f seconds" % self.elapsed_time)


def PostProcessing(mydomain, mri_para, mri_simu, plt, ms=""):
    ctx = mri_simu.ctx
    info = ctx.info()
    whole_vol = info["volume"]
    signal = mri_simu.signal_c.real * whole_vol
    rowptr, _, dof_of = ctx.export_pattern()
    w = np.add.reduceat(ctx.export_values("m"), rowptr[:-1])  # int phi_d over the DOF's compartment
    nv = info["nv"]
    u0 = getattr(mri_simu, "u0_vertex", None)
    u0d = np.ones(w.shape[0])  # the initial state per DOF
    if u0 is not None:
        for c in range(dof_of.shape[0] // nv):
            d = dof_of[c * nv:(c + 1) * nv]
            u0d[d[d >= 0]] = u0[d >= 0]
    voi = float((w * u0d).sum())  # int Dirac_Delta dx (DmriFemLib.py:924)
    if mydomain.IsDomainMultiple:
        u = ctx.solution()
        live0 = dof_of[:nv]
        n0 = int((live0 >= 0).sum())
        initial0, initial1 = float((w * u0d)[:n0].sum()), float((w * u0d)[n0:].sum())
        signal0, signal1 = float((w[:n0] * u[:n0].real).sum()), float((w[n0:] * u[n0:].real).sum())
        if np.isscalar(mydomain.kappa):
            out_text = ("b: %.3f, g: %.3f, q: %.3e, Signal: %.3e, Normalized signal: %.6e, kappa: %.3e, dt: %.3f, "
                        "hmin: %.3e, hmax: %.3e, whole_vol: %.3f, vol_of_interest: %.3f, elasped time %.3f (s)\n" %
                        (mri_para.bvalue, mri_para.gvalue, mri_para.qvalue, signal, signal / voi, mydomain.kappa,
                         mri_simu.k, mydomain.hmin, mydomain.hmax, whole_vol, voi, mri_simu.elapsed_time))
        else:
            out_text = ("b: %.3f, g: %.3f, q: %.3e, Signal: %.3e, Normalized signal: %.6e, dt: %.3f, hmin: %.3e, "
                        "hmax: %.3e, whole_vol: %.3f, vol_of_interest: %.3f, elasped time %.3f (s)\n" %
                        (mri_para.bvalue, mri_para.gvalue, mri_para.qvalue, signal, signal / voi, mri_simu.k,
                         mydomain.hmin, mydomain.hmax, whole_vol, voi, mri_simu.elapsed_time))
        print("Signal on each compartment")
        print("Sum initial0: %.3e, Signal0: %.3e" % (initial0, signal0))
        print("Sum initial1: %.3e, Signal1: %.3e" % (initial1, signal1))
        print(out_text)
    else:
        out_text = ("b: %.3f, g: %.3f, q: %.3e, Signal: %.3e, Normalized signal: %.6e, dt: %.3f, hmin: %.3e, "
                    "hmax: %.3e, whole_vol: %.3f, vol_of_interest: %.3f, elasped time %.3f (s)\n" %
                    (mri_para.bvalue, mri_para.gvalue, mri_para.qvalue, signal, signal / voi, mri_simu.k,
                     mydomain.hmin, mydomain.hmax, whole_vol, voi, mri_simu.elapsed_time))
        print(out_text)
    print("save to log.txt")
    with open("log.txt", "a") as fh:
        if ms != "":
            fh.write("%" + ms + "\n")
        fh.write(out_text)
    return out_text
