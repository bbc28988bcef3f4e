/*
 * dmri.h -- C ABI of libdmri, the MI355X (gfx950) Bloch-Torrey FEM solver.
 *
 * Plain C: pointers, sizes and status codes.  Every function returns DMRI_OK (0) or a
 * negative DMRI_ERR_* code; dmri_last_error() returns the message of the calling
 * thread's last failure.  No exception crosses this boundary.
 *
 * Ownership: the caller owns every host array and the library copies what it needs;
 * the context owns its device buffers.  Threading: one context per host thread at a
 * time (a context is not re-entrant); separate contexts may run concurrently.
 * All arithmetic is FP64.  Units follow the reference: um, us, D in um^2/us (== mm^2/s),
 * kappa in um/us (== m/s), b in us/um^2 (== s/mm^2), q in 1/(um us).
 *
 * The reference interfaces each entry point replaces are cited per function
 * (paths relative to the van-dang/DMRI-FEM-Cloud repository).
 */
#ifndef DMRI_H
#define DMRI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMRI_OK 0
#define DMRI_ERR_ARG (-1)     /* invalid argument */
#define DMRI_ERR_HIP (-2)     /* HIP runtime failure */
#define DMRI_ERR_STATE (-3)   /* call out of order (e.g. solve before assemble) */
#define DMRI_ERR_NOCONV (-4)  /* Krylov solver did not converge within maxit */
#define DMRI_ERR_MESH (-5)    /* mesh / phase / facet data inconsistent */
#define DMRI_ERR_ALLOC (-6)   /* device or host allocation failed */
#define DMRI_ERR_LIMIT (-7)   /* a size exceeds a documented limit */

typedef struct dmri_ctx dmri_ctx;

/* Value arrays produced by dmri_assemble (all on the DOF sparsity pattern). */
enum dmri_matrix {
    DMRI_MAT_M = 0,  /* mass  int phi_i phi_j                       (MassMatrix1c/2c, DmriFemLib.py:240-254) */
    DMRI_MAT_K = 1,  /* stiffness int grad phi_i . D grad phi_j     (FuncF_wBC, DmriFemLib.py:43-44) */
    DMRI_MAT_R = 2,  /* relaxation int (1/T2) phi_i phi_j           (FuncF_wBC, DmriFemLib.py:43-44) */
    DMRI_MAT_I = 3,  /* interface kappa int_Gamma (u0-u1)(v0-v1)    (icondition_wBC, DmriFemLib.py:47-50) */
    DMRI_MAT_JX = 4, /* int x phi_i phi_j  (GdotX, DmriFemLib.py:33-39; Jg = gx Jx + gy Jy + gz Jz) */
    DMRI_MAT_JY = 5,
    DMRI_MAT_JZ = 6,
    DMRI_MAT_P = 7,  /* weak pseudo-periodic boundary mass  int_{dOmega} kappa_e phi_i phi_j per compartment,
                        kappa_e = (3e-3/hmin) * P1 indicator of the periodic faces (ThetaMethodF_wBC1c/2c,
                        DmriFemLib.py:92,143; WeakPseudoPeridicMarker 601-610); zero without periodic dirs */
    DMRI_MAT_GX = 8, /* strong periodic form: (C - B)_x (dmri_periodic_strong_set); zero otherwise */
    DMRI_MAT_GY = 9,
    DMRI_MAT_GZ = 10
};

/* Create a context on HIP device `device`.
 * Replaces the process set-up of the drivers: MPI_Init + DOLFIN init
 * (comri/one-comp/hpc-fenics-cpp/main.cpp:117-127) and `from dolfin import *`
 * (GCloudDmriSolver.py:47-49). */
int dmri_create(dmri_ctx **out, int device);
int dmri_destroy(dmri_ctx *ctx);

/* Message of the calling thread's last failure ("" if none). */
const char *dmri_last_error(void);

/* Library version string. */
const char *dmri_version(void);

/* Mesh: nv vertices (xyz, nv*3, um), nc tetrahedra (tets, nc*4, 0-based), optional
 * cell phase in {0,1} (NULL = one compartment).
 * Replaces Mesh(fmesh) / HDF5File.read(mesh) and the phase function
 * (comri/two-comp/hpc-fenics-cpp/main.cpp:670, MarkPhase 242-337; GCloudDmriSolver.py:150-152,
 * 173-177; CreatePhaseFunc, DmriFemLib.py:754-798). */
int dmri_mesh_set(dmri_ctx *ctx, int64_t nv, const double *xyz, int64_t nc, const int32_t *tets,
                  const int32_t *cell_phase);

/* Coefficients.  Replaces MyDomain.D / D0 / ImposeDiffusionTensor, MyDomain.kappa and
 * MRI_parameters.T2 (DmriFemLib.py:583-616,807; GCloudDmriSolver.py:157-171,204-215) and the
 * comri kcoeff/kappa coefficients (comri/two-comp/hpc-fenics-cpp/main.cpp:741). */
typedef struct {
    double D[2];                /* scalar diffusivity per compartment */
    const double *D_cell;       /* optional nc*9 row-major tensor per cell (overrides D); K_ij =
                                   int grad(phi_i) . D grad(phi_j), row i the test function, as
                                   inner(D*grad(u), grad(v)) of FuncF_wBC (DmriFemLib.py:43-44) */
    double invT2[2];            /* 1/T2 per compartment (reference default T2 = 1e16) */
    const double *invT2_cell;   /* optional nc values (overrides invT2) */
    double kappa;               /* interface permeability */
    const double *kappa_facet;  /* optional, one value per interface facet in the order of
                                   dmri_export_facets (overrides kappa) */
} dmri_coeffs;
int dmri_coeffs_set(dmri_ctx *ctx, const dmri_coeffs *coeffs);

/* Build on the device: DOF map, interface facets, CSR pattern, element colouring and the
 * value arrays M, K, R, I, Jx, Jy, Jz.
 * Replaces DOLFIN's sparsity builder + UFC tabulate_dofs / tabulate_tensor
 * (comri/one-comp/hpc-fenics-cpp/ufc/Bloch_Torrey_NoTime3D.cpp:3869-3880,4110-4353;
 * comri/two-comp/hpc-fenics-cpp/ufc/Bloch_Torrey_NoTime3D.cpp:6699-7050,7421-8890) as driven by
 * assembler.assemble (main.cpp:261-281, two-comp 829-849) and MassMatrix / assemble(F)
 * (DmriFemLib.py:245,252,904). */
int dmri_assemble(dmri_ctx *ctx);

typedef struct {
    int64_t nv, nc, ncomp, ndof, nnz, n_iface, n_bfacet;
    int32_t ncolors_cells;  /* 0: the cell matrices are assembled by row gather, without colouring */
    int32_t ncolors_iface;  /* colours of the interface-facet assembly */
    double volume;        /* sum of M (|Omega|; whole_vol of PostProcessing, DmriFemLib.py:923) */
    double hmin, hmax;    /* min/max greatest vertex distance of a cell (mymesh.hmin/hmax) */
    double setup_ms;      /* device time of dmri_assemble */
    int64_t n_periodic;   /* weak periodic condition: DOFs on a periodic face that take a mirrored value
                             (dmri_export_periodic); strong: DOFs identified with their master's */
    double kappa_e;       /* 3e-3/hmin when periodic directions are set (MyDomain.kappa_e_scalar), else 0 */
} dmri_info;
int dmri_get_info(dmri_ctx *ctx, dmri_info *info);

/* Parity hooks: copy the device results to host buffers sized from dmri_get_info.
 * rowptr: ndof+1, col: nnz, dof_of: ncomp*nv (-1 = dead DOF), val: nnz, fv: 3*n_iface. */
int dmri_export_pattern(dmri_ctx *ctx, int64_t *rowptr, int32_t *col, int32_t *dof_of);
int dmri_export_values(dmri_ctx *ctx, int which, double *val);
int dmri_export_facets(dmri_ctx *ctx, int32_t *fv);

/* Weak pseudo-periodic boundary condition (MyDomain.PeriodicDir with IsDomainPeriodic = False:
 * WeakPseudoPeridicMarker, DmriFemLib.py:601-610; ThetaMethodF/L_wBC1c/2c 58-145;
 * WeakPseudoPeriodic_1c/2c 256-324,386-450).  dir[d] != 0 makes the two faces of the mesh's
 * bounding box normal to axis d periodic.  Takes effect at the next dmri_assemble, which then adds
 *   implicit side:  theta * P,              P_ij = int_{dOmega} kappa_e phi_i phi_j  (per compartment)
 *   explicit side:  (1-theta) * P * u_bc,   u_bc(x) = u^n(x + l) exp(i q (g.l) F(t_{n-1}))
 * where kappa_e = (3e-3/hmin) * the P1 interpolant of the indicator "within 1e-2 hmin of a periodic
 * face", and for a vertex x within 1e-7 of a periodic face, x + l is its mirror on the opposite face
 * (l along that axis; where several periodic faces meet the last axis in x, y, z order wins, as in
 * the reference's eval); u_bc = 0 at every other vertex.  u^n at the mirror point is the P1
 * interpolant on the boundary facet of the opposite face that contains it.  Each periodic
 * direction needs flat opposite faces.  The time profile's integral F must be given
 * (dmri_seq.F_explicit).  dir = {0,0,0} switches the condition off.  The adjoint does not support
 * it (DMRI_ERR_ARG). */
int dmri_periodic_set(dmri_ctx *ctx, const int32_t dir[3]);

/* Strong pseudo-periodic condition (MyDomain.PeriodicDir with IsDomainPeriodic = True: PeriodicBD,
 * DmriFemLib.py:327-375, on the function spaces of MyFunctionSpaces 453-486; the forms
 * ThetaMethodF/L_sBC1c/2c 167-238 with FuncF_sBC, outer_interface and inner_interface 148-165).
 * dir[d] != 0 identifies the two faces of the bounding box normal to axis d: every vertex within
 * 1e-2 hmin of the max face shares the DOF of the vertex at its image on the min face (found within
 * the same distance; the mesh must be periodic there, DMRI_ERR_MESH otherwise; a vertex on several
 * max faces takes the min corner).  The unknown is the magnetisation times exp(i q F(t) g.x), and
 * each step solves (theta on both sides, as the reference's forms)
 *   A_n = M/k + theta (K + R + I) + theta q^2 F_n^2 D_c M + i theta q F_n (C - B).g
 *   B_n = M/k - theta (K + R + I) - theta q^2 F_{n-1}^2 D_c M - i theta q F_{n-1} (C - B).g
 * on the constrained space, with C_d = int 2 D_c (d phi_j/d x_d) phi_i and B_d the integral of D_c n_d
 * phi_i phi_j over each compartment's boundary (outer facets and the interface, n outward of that
 * compartment).  Each compartment's outer term is taken over its own facets; the reference takes each
 * compartment's over every exterior facet, which differs only where the other compartment's cells
 * reach a non-periodic face with g.n != 0.  The reference's 1e-16 added to q F in outer_interface is
 * left out (below FP64 resolution of the signal).  Needs a scalar D per compartment (no D_cell),
 * dmri_seq.F_implicit and F_explicit, and one q per launch group (batched signals are grouped by q).
 * The initial guess is the quadratic extrapolation when the projection is asked for.  The exports
 * then describe the constrained space (dof_of maps every vertex of a max face to its master's DOF;
 * DMRI_MAT_GX..GZ hold (C - B)_d).  dir = {0,0,0}, or dmri_periodic_set, switches it off.  The
 * adjoint does not support it (DMRI_ERR_ARG).  Takes effect at the next dmri_assemble. */
int dmri_periodic_strong_set(dmri_ctx *ctx, const int32_t dir[3]);

/* Parity hook: the n_periodic mirrored DOFs of the last dmri_assemble.  dof[e] is the DOF that takes
 * the mirrored value; src[3e..3e+2] the DOFs (same compartment, -1 = dead) of the facet vertices
 * whose P1 interpolant gives u^n at the mirror point, with weights w[3e..3e+2]; ell[3e..3e+2] = l. */
int dmri_export_periodic(dmri_ctx *ctx, int32_t *dof, int32_t *src, double *w, double *ell);

/* Time sequence for one signal.
 * f_implicit[n] multiplies J on the implicit side of step n, f_explicit[n] on the explicit
 * side.  GCloud: f(t_n) and f(t_{n-1}) with t_{-1}=0 (MRI_simulation.solve, DmriFemLib.py:897-910);
 * comri: both f(t_n) (main.cpp:300-318). */
typedef struct {
    int32_t nsteps;
    const double *f_implicit;
    const double *f_explicit;
    double dt;        /* k (MRI_simulation.k, -k) */
    double theta;     /* 0.5 (MRI_simulation.theta, DmriFemLib.py:863) */
    double q;         /* gamma*G (MRI_parameters.qvalue; comri gnorm) */
    double gdir[3];   /* gradient direction (normalised inside) */
    const double *F_explicit; /* nsteps values of F(t) = int_0^t f on the explicit side's time
                                 (GCloud: t_{n-1}, MRI_parameters.itime_profile, DmriFemLib.py:834-835,
                                 901-902): the phase of the pseudo-periodic mirror values; required
                                 when periodic directions are set, ignored otherwise (may be NULL) */
    const double *u0_vertex;  /* initial state: nv real values per mesh vertex, taken by the real
                                 part of every compartment's DOF at that vertex (the imaginary part
                                 is 0), as MRI_simulation.InitialCondition assigns a P1 field to
                                 u_0.sub(0) and u_0.sub(2) (DmriFemLib.py:865-876); NULL = 1
                                 everywhere.  The signal stays sum_i w_i Re u_i / |Omega|: a caller
                                 with a custom state divides by int u0 / |Omega| (PostProcessing's
                                 vol_of_interest, DmriFemLib.py:924) */
    int32_t snap_every;       /* > 0: copy the state to snap_out after every step whose 0-based index
                                 is a multiple of snap_every, as comri's time loop writes u after the
                                 solve of every step with step_counter % nskip == 0 (-j nskip, -s 1;
                                 comri/two-comp/hpc-fenics-cpp/main.cpp:886-890); 0 = no snapshots.
                                 dmri_solve and dmri_solve_batched only (dmri_adjoint ignores it). */
    double *snap_out;         /* with snap_every > 0: ceil(nsteps / snap_every) * nsig * 2 * ndof
                                 doubles; snapshot j of signal s starts at (j * nsig + s) * 2 * ndof
                                 and holds u in dmri_get_solution's layout (interleaved re/im in DOF
                                 order; nsig = 1 for dmri_solve).  A device-to-host copy (and a wait
                                 for it) per snapshot, so only as often as needed */
    const double *F_implicit; /* nsteps values of F(t) = int_0^t f on the implicit side's time (t_n;
                                 MRI_parameters.itime_profile at the step's t, DmriFemLib.py:901): required
                                 by the strong periodic condition, ignored otherwise (may be NULL) */
} dmri_seq;

typedef struct {
    double rtol;           /* ||r|| <= max(rtol ||b||, atol); reference 1e-9 (GCloudDmriSolver.py:220) */
    double atol;
    int32_t maxit;         /* per time step */
    int32_t profile;       /* k > 0: time the SpMV launches of the first k Krylov iterations of every
                              time step with HIP events (dmri_stats.spmv_ms, spmv_launches); the
                              events cost GPU time of their own, so a small k perturbs least */
    int32_t lanes_per_row; /* SpMV lanes per row: 0 = tuned value (dmri_tune / dmri_set_tuning) or the
                              library default; else a power of two in 1..32 */
    int32_t initial_guess; /* Krylov start of step n: 0 = the combination of the last 12 states
                              (fewer in the first steps) that minimises the residual of step n
                              (the default; the history takes 36 state-sized buffers; 10, 8, 6
                              or 4 states are kept when those do not fit in half of the free
                              device memory, and the quadratic extrapolation below when even 4
                              do not; dmri_adjoint restarts the history at its checkpoints),
                              1 = u^n, 2 = 2u^n - u^(n-1) (linear extrapolation), 3 = 3u^n -
                              3u^(n-1) + u^(n-2) (quadratic extrapolation; linear at n = 1) */
    int32_t mapping;       /* SpMV mapping: 0 = tuned/default, 1 = signals spread over lanes
                              (lanes * signals <= 64), 2 = the same with rows staged in LDS,
                              3 = signals kept in registers (at most 8 signals per launch group),
                              4 = one wave per row, lane = signal, row data through the scalar
                              cache (lanes_per_row 1; for large launch groups),
                              5 = mapping 1 over tiles whose distinct columns are staged once
                              in LDS (lanes * signals <= 64), 6 = mapping 2 with two signals per
                              lane (at least 2 signals; lanes_per_row 1, 2 or 4) */
    int32_t preconditioner; /* 0 = Jacobi (the symmetric diagonal scaling) with the degree-1
                               polynomial c (2I - c A) of the scaled system, applied from the right
                               (the residual the tolerance tests stays the true one; the default):
                               - of its real part alone when the gradient term's bound is below 5 %
                                 of the real part's;
                               - in the variables scaled by the complex diagonal D = diag(A)
                                 (c D^-1 (2I - c A D^-1)) when that bound reaches the real part's
                                 (strong gradients over a large domain);
                               - of the whole operator otherwise.
                               c = 1.8 / b where b bounds the largest eigenvalue of the real part
                               (1.05 x a 30-step power iteration estimate, capped by the Gershgorin
                               bound), or, with the complex diagonal, the larger of that and a
                               Gershgorin-type bound of D^-1 A over every step and signal.
                               1 = Jacobi only, as the reference's KrylovSolver("bicgstab",
                               "jacobi").  2 = the complex-diagonal polynomial whatever the bounds.
                               Every choice converges to the same solution; the polynomial needs
                               SpMV mapping 1, 2 or 4 (others fall back to mapping 2).
                               dmri_stats.precond_used reports the choice. */
    int32_t krylov;         /* 0 = BiCGStab (KrylovSolver("bicgstab", ...), GCloudDmriSolver.py:219;
                               the default), 1 = restarted GMRES (KrylovSolver("gmres", ...),
                               comri/one-comp/fenics-cpp/main.cpp:180-183) with the same right
                               preconditioner: classical Gram-Schmidt applied twice, Givens rotations
                               on the device, the true residual checked at the end of every cycle.
                               2 = COCG (conjugate orthogonal CG) for the complex-symmetric operator
                               (every form but the strong periodic one; needs the polynomial
                               preconditioner, preconditioner 0 or 2): one preconditioned product
                               per iteration where BiCGStab takes two, with unconjugated products
                               r^T z, p^T A p.  All stop at ||r|| <= max(rtol ||b||, atol).
                               dmri_stats.iters_* count Arnoldi steps for GMRES; for COCG the timed
                               kernels are [0] p = z + beta p, [1] z = M r, [2] q = A p, [6] the
                               update of x and r ([6] stays 0 when a single-signal solve whose matrix
                               and vectors fit in the 256 MB MALL folds the update of r into [1]). */
    int32_t restart;        /* GMRES restart length, 0 = 30, at most 64 (m + 1 basis vectors per signal
                               are kept in HBM) */
    int32_t graphs;         /* 1: BiCGStab and COCG launch each run of iterations between two convergence
                               checks as hipGraphs of 2^k iterations, each captured once per (kernel
                               arguments, first iteration, length) and replayed (the step's
                               coefficients go to a fixed device slot and the projection guess iterates
                               in a fixed buffer so that the arguments repeat); results are identical
                               to 0 (plain launches); timed iterations (profile) go out one by one */
    int32_t pre_fp32;       /* 1 (COCG only): the preconditioner's SpMV reads r and the polynomial's
                               operator values rounded to FP32 and sums in FP64; z, p, q, x and r stay
                               FP64 and the stopping test is on the FP64 residual, so the solution meets
                               the same tolerance (a few % more iterations, half the gathered bytes in
                               the preconditioner's product).  The complex-diagonal form keeps FP64
                               inputs.  0 = FP64 throughout (the default). */
} dmri_solver_opts;

typedef struct {
    int64_t iters_total;    /* Krylov iterations over all steps */
    int32_t iters_max;      /* max iterations of one step */
    int32_t restarts;       /* BiCGStab breakdown restarts */
    double relres_max;      /* max final ||r||/||b|| over the steps */
    double solve_ms;        /* wall time of the time loop (device synchronised) */
    double spmv_ms;         /* summed device time of the timed SpMV launches (profile > 0) */
    int64_t spmv_launches;  /* number of timed SpMV launches */
    double spmv_bytes;      /* algorithmic bytes of one SpMV launch */
    int32_t precond_used;   /* preconditioner of the last launch group: 0 = Jacobi only, 1 = the
                               polynomial of the scaled operator, 2 = the polynomial of its real part,
                               3 = the polynomial in the variables scaled by the complex diagonal */
    int32_t graph_launches; /* hipGraph launches of captured iteration runs (graphs = 1; 0 when the
                               runtime refused the capture and the iterations went out one by one) */
    /* Every kernel of the timed Krylov iterations (profile > 0), by its own dispatch timestamps:
     * [0] bicg_p (p = r + beta (p - omega v)), [1] preconditioner SpMV on p (ph), [2] v = A ph (or A p),
     * [3] the s-pass (alpha, s = r - alpha v), [4] preconditioner SpMV on s (sh), [5] t = A sh (or A s),
     * [6] the update of x and r.  [1], [3] and [4] stay 0 without the polynomial preconditioner.
     * kern_ms: summed device time; kern_bytes: summed algorithmic HBM bytes of those launches (matrix
     * entries, column indices and row offsets once, every vector read or written once; per launch:
     * [0] 64 n S (32 n S in a solve's first iteration), [1] E(npre) + 32 n S, [2] E(nval) + 48 n S,
     * [3] 48 n S, [4] E(npre) + 32 n S, [5] E(nval) + 48 n S, [6] 128 n S, with E(k) = nnz (4 + 8k) +
     * 4 (n+1), nval = 2 (4 with per-signal directions) and npre = 1 for the real-part polynomial, else
     * nval; with COCG: [0] 80 n S (32 n S in a solve's first iteration), [1] E(npre) + 32 n S
     * (E(npre) + 64 n S with the update of r folded in), [2] E(nval) + 32 n S, [6] 48 n S). */
    double kern_ms[7];
    int64_t kern_launches[7];
    double kern_bytes[7];
} dmri_stats;

/* Run the theta scheme from u0 = 1 (real part, every live DOF) and return the signal
 * sig_out[0] = sum_i w_i Re u_i / |Omega|, sig_out[1] = the same for Im u (w = M 1).
 * Replaces MRI_simulation.solve + PostProcessing (DmriFemLib.py:878-988) and the comri
 * time loop + Comp_Sig3D (main.cpp:296-333, two-comp 854-900). */
int dmri_solve(dmri_ctx *ctx, const dmri_seq *seq, const dmri_solver_opts *opts, double sig_out[2],
               dmri_stats *stats);

/* Batched solve: nsig signals with their own q[s] and gdir[3*s], sharing the mesh, the
 * coefficients and the time profile of seq (seq->q and seq->gdir are ignored); signals go
 * to the device in launch groups of up to 64 that advance together through every theta step
 * and Krylov iteration on the shared sparsity pattern.  sig_out has 2*nsig doubles.
 * No reference counterpart for the batching itself: the reference runs one signal per
 * process (GCloudDmriSolver.py:232) and loops over b in its notebooks; each batched
 * signal equals the one dmri_solve returns for that (q, gdir). */
int dmri_solve_batched(dmri_ctx *ctx, int32_t nsig, const double *q, const double *gdir, const dmri_seq *seq,
                       const dmri_solver_opts *opts, double *sig_out, dmri_stats *stats);

/* Resumable batched solve: the signals of dmri_solve_batched advanced a given number of theta
 * steps per call, every launch group (up to `group` signals, 0 = 64) keeping its state, its
 * Krylov workspace and its initial-guess history in HBM between calls.  dmri_batch_begin checks
 * the inputs, scales the system and sets u0 (the time profile is copied to the device, so the
 * caller's arrays may be freed after it returns; seq->snap_every must be 0); dmri_batch_advance
 * runs up to nsteps more theta steps of every signal (fewer at the end of the sequence; stats as
 * in dmri_solve, solve_ms = this call); dmri_batch_signals gives the signals of the current states
 * (2*nsig doubles, as sig_out of dmri_solve_batched; after the last step they equal its result)
 * and the number of steps done; dmri_batch_end frees the batch.  A batch stays bound to its
 * context, which must not be assembled again (or given new inputs) while the batch lives
 * (DMRI_ERR_STATE), and every batch must be ended before its context is destroyed.  Several
 * batches of one context run one after another on its stream.
 * No reference counterpart: the reference runs each signal's time loop to its end in one call
 * (MRI_simulation.solve, DmriFemLib.py:878-915). */
typedef struct dmri_batch dmri_batch;
int dmri_batch_begin(dmri_ctx *ctx, int32_t nsig, const double *q, const double *gdir, const dmri_seq *seq,
                     const dmri_solver_opts *opts, int32_t group, dmri_batch **out);
int dmri_batch_advance(dmri_batch *batch, int32_t nsteps, dmri_stats *stats);
int dmri_batch_signals(dmri_batch *batch, double *sig_out, int32_t *steps_done);
int dmri_batch_end(dmri_batch *batch);

/* Autotuner: time the four SpMV kernels of a polynomial-preconditioned BiCGStab iteration (the default
 * solve: the preconditioner SpMV on p, v = A ph, the preconditioner SpMV on s, t = A sh) for every
 * (mapping, lanes per row 1..32) that runs them (mappings 1, 2, 4 and 6) on the assembled pattern
 * (us[36] = [mapping-1][log2 lanes] microseconds per iteration's four SpMVs, -1 = not valid) and keep
 * the fastest for later solves of this context whose options leave mapping and lanes_per_row at 0.
 * dmri_set_tuning installs a stored choice without timing.  No reference counterpart (the reference
 * has no GPU kernels); results do not depend on the choice beyond rounding. */
int dmri_tune(dmri_ctx *ctx, int32_t nsig, int32_t per_gdir, int32_t *best_lanes, int32_t *best_mapping,
              double *us);
int dmri_set_tuning(dmri_ctx *ctx, int32_t nsig, int32_t per_gdir, int32_t lanes, int32_t mapping);

/* Tuning database: a text file with one "device ndof nnz nsig per_gdir mapping lanes us" line per
 * tuned (device, pattern size, launch group) key.  dmri_tune_db installs the stored choice for this
 * context's pattern when the file holds one (*hit = 1, nothing is timed; the last matching line
 * wins); otherwise it runs dmri_tune and appends the choice (*hit = 0).  path NULL or "" = tune
 * without a database.  No reference counterpart. */
int dmri_tune_db(dmri_ctx *ctx, const char *path, int32_t nsig, int32_t per_gdir, int32_t *best_lanes,
                 int32_t *best_mapping, int32_t *hit);

/* Replay the two operator SpMV kernels of a polynomial-preconditioned BiCGStab iteration (v = A ph and
 * t = A sh with s = r - alpha v read at the own row) `reps` times each, back to back, on this context's
 * solver buffers for a launch group of nsig signals, with the mapping a solve of that group would use.
 * us_v, us_t = microseconds per launch (kernel-bound HIP events, no marker packets between launches);
 * bytes_per_launch (may be NULL) = algorithmic HBM bytes of one launch: nnz*(4 + 8*nv) column indices
 * and values (nv = 2, or 4 with per-signal gradient directions) + 4*(n+1) row pointers + 48*n*nsig for
 * the complex vectors (one gathered, one read at the own row, one written).  No reference counterpart. */
int dmri_time_spmv(dmri_ctx *ctx, int32_t nsig, int32_t per_gdir, int32_t reps, double *us_v, double *us_t,
                   double *bytes_per_launch);

/* Bandwidth of a plain device copy kernel over `bytes` (read + write counted), averaged over
 * `reps` launches: the measured reference beside the nominal HBM peak, and the calibration
 * kernel for the HBM byte counters.  No reference counterpart. */
int dmri_copy_bandwidth(dmri_ctx *ctx, int64_t bytes, int32_t reps, double *gbps);

/* Adjoint (reverse-mode) gradients of the normalised signal Re S of one dmri_solve.
 * With A_n, B_n complex symmetric the discrete adjoint of the theta scheme is
 *   lambda^N = w/|Omega|,  A_n mu_n = lambda^(n+1),  lambda^n = B_n mu_n,
 *   dS/dp = Re sum_n mu_n^T (dB_n/dp u^n - dA_n/dp u^(n+1)),
 * so the backward sweep runs the forward kernels.  The N steps form segments of checkpoint_stride
 * steps, the first taking the remainder (so the last is a full one).  A checkpoint at each segment's
 * first step holds the solver's state history (the projection guess's whole history, the default,
 * or (u^k, u^(k-1), u^(k-2)) with the extrapolated guesses), so a recomputed segment repeats the
 * forward sweep bit for bit; when the whole histories would take more than a quarter of the free
 * device memory, a checkpoint keeps three states and the recomputed states agree with the forward
 * ones to the solver tolerance.  Every segment but the last is recomputed; the last one's states
 * are kept from the forward sweep itself.  checkpoint_stride >= N keeps every state instead.
 * 0 = automatic:
 * every state when the N+1 states fit in a quarter of the free device memory (then no step is
 * recomputed), else checkpoints every ceil(sqrt(N)) steps.
 * A D_cell field must hold symmetric tensors (DMRI_ERR_ARG otherwise: the operator is then not
 * complex symmetric).
 * No reference counterpart: the reference computes signals only. */
typedef struct {
    int32_t checkpoint_stride;
    int32_t reserved;
} dmri_adjoint_opts;

typedef struct {
    double dS_dD[2];        /* per compartment: d/dD_c for scalar D (also at D_c = 0), or d/dm_c where the
                               D_cell tensors of compartment c's cells are scaled by m_c (at m_c = 1) */
    double dS_dinvT2[2];    /* per compartment: d/d(1/T2_c) for scalar 1/T2, or d/dm_c of the invT2_cell
                               values of compartment c's cells scaled by m_c (at m_c = 1) */
    double dS_dkappa;       /* d/dkappa for scalar kappa (also at kappa = 0), or d/dm of kappa_facet scaled
                               by m (at m = 1); 0 for one compartment */
    double dS_dq;           /* at fixed time profile */
    double *dS_df_implicit; /* optional out, nsteps: d/d f_implicit[n] at fixed q */
    double *dS_df_explicit; /* optional out, nsteps: d/d f_explicit[n] at fixed q */
    double ms_forward, ms_adjoint;
    int32_t checkpoints, recomputed_steps;
    /* Krylov iterations of the forward sweep, of the steps recomputed from checkpoints, and of the
       backward (adjoint) solves */
    int64_t iters_forward, iters_recompute, iters_backward;
} dmri_grads;
int dmri_adjoint(dmri_ctx *ctx, const dmri_seq *seq, const dmri_solver_opts *opts, const dmri_adjoint_opts *aopts,
                 double sig_out[2], dmri_grads *grads, dmri_stats *stats);

/* Copy u^N of signal `sig` of the last launch group to host (2*ndof doubles, interleaved
 * re/im, DOF order; the DOF of (compartment c, vertex v) is dof_of[c*nv+v]). */
int dmri_get_solution(dmri_ctx *ctx, int32_t sig, double *u);

/* Largest ndof that dmri_get_solution_noalloc takes: the size of its device staging buffer (2^21
 * complex doubles, 32 MiB on every GPU the library runs on, held until the process ends). */
#define DMRI_NOALLOC_MAX_DOF 2097152

/* dmri_get_solution without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same result as dmri_get_solution, bit for bit; the rescaled state in DOF order
 * is staged in a file-scope device buffer instead of a temporary allocation, then copied to u.
 * ndof > DMRI_NOALLOC_MAX_DOF: DMRI_ERR_LIMIT.  Every call on one GPU shares that buffer, which
 * keeps the last call's contents; overlapping calls on one GPU would race on it, so the library
 * holds a per-device lock for the length of a call and they run one after another. */
int dmri_get_solution_noalloc(dmri_ctx *ctx, int32_t sig, double *u);

#ifdef __cplusplus
}
#endif

#endif /* DMRI_H */
