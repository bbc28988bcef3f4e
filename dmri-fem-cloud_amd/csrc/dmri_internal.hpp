// Internal declarations shared by the libdmri host code and its gfx950 kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmri {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define DMRI_HIP(call)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw ::dmri::Error(-2, std::string(#call) + ": " + hipGetErrorString(e_) + " at " __FILE__ \
                                                                                        ":" +            \
                                        std::to_string(__LINE__));                                       \
    } while (0)

// Device buffer owned by the host object (freed on destruction).
template <class T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { release(); }
    void alloc(size_t count) {
        release();
        if (count) {
            hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
            if (e != hipSuccess) {
                p = nullptr;
                throw Error(-6 /* DMRI_ERR_ALLOC */, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                                hipGetErrorString(e));
            }
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) DMRI_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) {
        if (n) DMRI_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
};

// Per-signal BiCGStab state.  Scalars of iteration i live in slot i&1: a kernel only
// writes the slot of its own iteration and reads slots written by earlier kernels, so no
// block ever reads a value another block of the same launch writes.
struct KrylovState {
    double2 rho[2], alpha[2], omega[2];
    double tol2, bnorm2, rnorm2, pad;
    int done, conv, breakdown, iters;
    int it0, restarts;
    int stop;  // COCG: the p-kernel stopped the signal (converged, breakdown or maxit); the update
               // kernel of the same iteration (with the folded update: the next iteration's first
               // kernel) then marks it done (see cocg_p_kernel)
    int pad2;
};

constexpr int kThreads = 256;   // target threads per block of every solver kernel
constexpr int kMaxGrid = 512;      // blocks of the vector kernels (grid-stride)
// upper bound of the SpMV grid; fit_spmv_grid lowers it to the blocks that stay resident at once
constexpr int kMaxGridSpmv = 2048;
constexpr int kMaxSignals = 64; // signals per batched launch
constexpr int kNV = 4;          // doubles per (block, signal) partial slot
constexpr int kStageNnzPerRow = 18; // LDS budget of a staged SpMV tile, entries per row
constexpr int kStageBatch = 4;      // independent loads per thread per batch when staging a tile

// SpMV mappings: signals spread over lanes (entries from global memory, or staged in LDS per
// tile of rows), or all signals of a lane's entries kept in registers (S <= 8).
constexpr int kMapSL = 1;
constexpr int kMapStaged = 2;
constexpr int kMapSR = 3;
constexpr int kMapWave = 4;  // one wave per row, lane = signal; row data through the scalar cache
constexpr int kMapTile = 5;  // tiles of rpb rows: the tile's distinct columns staged once in LDS
constexpr int kMapPair = 6;  // mapping 2 with two signals per lane (G in 1, 2, 4; at least 2 signals)
// SpMV kernel operations: v = A p / t = A s (plain BiCGStab), the degree-1 polynomial
// preconditioner z -> 2z - A z applied to p or s, and t = A sh (preconditioned BiCGStab)
constexpr int kOpV = 0, kOpT = 0, kOpPre = 1, kOpTP = 2;
// kernel-side variant of kOpPre with the complex diagonal (SolveArgs::pdiag), compiled apart so
// that the plain preconditioner kernels keep their registers
constexpr int kOpPreCD = 3;
// COCG's SpMVs (mappings 1 and 2): q = A p with the partials of p^T q, and z = M p (the polynomial,
// plain or with the complex diagonal) with the partials of p^T z; unconjugated products, since
// COCG works with the complex-symmetric bilinear form
constexpr int kOpVDot = 4, kOpPreDot = 5, kOpPreCDDot = 6;
// kOpPreDot from FP32 inputs (SolveArgs::r32, SolveArgs::pvals32), accumulated in FP64: COCG's
// reduced-precision preconditioner (the real-part and whole-operator polynomials)
constexpr int kOpPreDot32 = 7;
// kOpPreDot with COCG's update of r folded in (one gradient direction, mappings 1 and 2; the solver
// takes it for single-signal solves): alpha from
// the previous q = A p, r_k = r_(k-1) - alpha q formed per gathered column and written for the own
// rows to the other of the two residual buffers (a.r, a.s: r_k lives in a.s for odd k), then
// z = M r_k with partials [r^T z, |r|^2, 0]
constexpr int kOpPreDotR = 8;
// COCG's q = A p with the p-step folded in (per-signal directions, mappings 1 and 2): the head is
// cocg_p_kernel's (convergence test, beta, the pending x update), the body gathers z and forms
// w = A z, then for its own rows x += alpha_prev p_prev, p = z + beta p_prev and q = w + beta q_prev
// (= A p), with partials of p^T q
constexpr int kOpQFold = 9;
// MI355X's last-level cache (MALL, AMD Infinity Cache) in bytes
constexpr double kMallBytes = 256.0 * 1024 * 1024;
// the preconditioner polynomial takes the real part alone when the imaginary part's bound is at
// most this fraction of the real part's (measured: 0.002 -> same iterations, 3.3 -> twice as many)
constexpr double kRealPreconditioner = 0.05;
// the polynomial acts on the operator scaled by its complex diagonal when that bound of the gradient
// term reaches this fraction of the real part's (below it the iterations it saves do not pay for the
// extra work of its SpMVs)
constexpr double kComplexDiagonal = 1.0;
constexpr int kTileMaxEntries = 4096;  // entries of a tile the compression build sorts in LDS
constexpr int kTileLdsBytes = 32768;   // LDS cap of the staged x values of one tile (distinct columns x S)
constexpr int kMaxSRSignals = 8;

// Launch geometry of the solver kernels for a (mapping, G lanes per row, S signals) choice.
struct SolveGeom {
    int G = 8, S = 1, mapping = kMapSL;
    int rpb = 32;            // rows per block pass in the spread-signal SpMV kernels
    int spmv_threads = 256;  // rpb * G * S
    int G_rhs = 8;           // lanes per row of the residual kernel (always spread-signal)
    int rhs_rpb = 32, rhs_threads = 256;
    int sr_rpb = 32, sr_threads = 256;  // signals-in-registers kernels: rows per pass, threads
    int vec_rpb = 256;       // rows per block in the vector kernels
    int vec_threads = 256;   // vec_rpb * S
    int grid = 8;            // blocks of every SpMV-type kernel (a multiple of 8)
    int tile_rows = 0;       // kMapTile: rows of a compressed tile (= rpb)
    int tile_ucap = 0;       // kMapTile: distinct columns staged per tile (LDS), set from the build
    int pair_sp = 0;         // kMapPair: signal pairs per row, ceil(S / 2) rounded up to a power of two
    int vgrid = 1;           // blocks of the vector kernels
    int pgrid = 1;           // blocks of the projection Gram kernel
};
SolveGeom make_geom(int64_t n, int G, int S, int mapping = kMapSL);
// Lower g.grid to the number of SpMV blocks the device keeps resident at once (occupancy of the
// mapping's preconditioner t-kernel, the most constrained one, times the CU count, a multiple of
// 8), so every block starts in the first round and the launch has no partial last round.
void fit_spmv_grid(SolveGeom &g, bool per_gdir);
int stage_rows(const SolveGeom &g, bool per_gdir);
size_t stage_lds_bytes(int TR, bool per_gdir);

// ---------------------------------------------------------------- setup launchers
void exclusive_scan(const int *d_in, int *d_out, int64_t n, hipStream_t s);  // d_out has n+1 entries

void launch_count_incidence(const int *tets, int64_t nc, int *cnt, hipStream_t s);
void launch_fill_incidence(const int *tets, int64_t nc, const int *ptr, int *cursor, int *list, hipStream_t s);
void launch_count_incidence3(const int *fv, int64_t nf, int *cnt, hipStream_t s);
void launch_fill_incidence3(const int *fv, int64_t nf, const int *ptr, int *cursor, int *list, hipStream_t s);
void launch_sort_lists(const int *ptr, int *list, int64_t n, hipStream_t s);
void launch_mark_live(const int *tets, const int *phase, int64_t nc, int64_t nv, int *live, hipStream_t s);
void launch_dof_from_scan(const int *live, const int *scan, int64_t n, int *dof_of, int64_t nv, int *dof_vert,
                          int *dof_comp, hipStream_t s);
void launch_face_classify(const int *tets, const int *phase, int64_t nc, const int *vptr, const int *vlist,
                          int *iface_flag, int *bnd_flag, int *nbr, hipStream_t s);
void launch_face_compact(const int *tets, int64_t nc, const int *nbr, const int *iface_flag, const int *iface_scan,
                         const int *bnd_flag, const int *bnd_scan, int *iface_v, int *iface_c, int *bnd_v,
                         int *bnd_c, hipStream_t s);
void launch_pattern(int pass, int64_t ndof, int64_t nv, const int *dof_vert, const int *dof_comp, const int *tets,
                    const int *phase, const int *vptr, const int *vlist, const int *fptr, const int *flist,
                    const int *iface_v, const int *dof_of, int *row_count, const int *rowptr, int *col,
                    int *overflow, hipStream_t s);
void launch_jp_round(int k, int64_t ne, int nvert, const int *ev, const int *vptr, const int *vlist, int *color,
                     int *remaining, hipStream_t s);
void launch_color_hist(const int *color, int64_t ne, int *hist, hipStream_t s);
void launch_color_scatter(const int *color, int64_t ne, const int *offs, int *cursor, int *perm, hipStream_t s);
// M, K, R, Jx, Jy, Jz of every pattern entry by row gather over the incident cells (writes every
// entry; phase may be null for one compartment)
void launch_assemble_rows(int64_t n, const int *dof_vert, const int *dof_comp, const int *rowptr, const int *col,
                          const int *tets, const int *phase, const double *xyz, const int *vptr, const int *vlist,
                          const double *Dc, const double *Dcell, const double *iT2c, const double *iT2cell, double *m,
                          double *k, double *r, double *jx, double *jy, double *jz, hipStream_t s,
                          double *cx = nullptr, double *cy = nullptr, double *cz = nullptr);
// Strong pseudo-periodic constraint: g = (c - c^T)/2; per DOF the DOF of its master vertex (key),
// the constrained ids (mid, from the scan of is_master); the constrained pattern P^T X P of the CSR
// (rows, cols of the merged runs of equal (mid row, mid col) keys, sorted stably) and the run sums.
void launch_antisym(int64_t n, const int *rowptr, const int *col, const double *c, double *g, int *miss,
                    hipStream_t s);
void launch_master_dofs(int64_t n, int64_t nv, const int *dof_vert, const int *dof_comp, const int *master,
                        const int *dof_of, int *key, int *is_master, int *bad, hipStream_t s);
void launch_merged_ids(int64_t n, const int *key, const int *scan, int *mid, hipStream_t s);
// returns the constrained entry count; rstart[u], sidx: run starts and the sorted entry positions
int64_t build_merge(int64_t n, int64_t nnz, const int *rowptr, const int *col, const int *mid, int64_t nm,
                    DevArray<int> &mrowptr, DevArray<int> &mcol, DevArray<int> &rstart, DevArray<int> &sidx,
                    hipStream_t s);
void launch_run_sum(int64_t nu, int64_t nnz, const int *rstart, const int *sidx, const double *val, double *merged,
                    hipStream_t s);
void launch_assemble_facets(int64_t start, int64_t count, const int *perm, const int *iface_v, const double *xyz,
                            int64_t nv, const int *dof_of, const int *rowptr, const int *col, double kappa,
                            const double *kappa_f, double *gi, int *miss, hipStream_t s);
// out[0] += sum of cnt[0..n) in 64 bits (out zeroed by the caller)
void launch_count_sum64(const int *cnt, int64_t n, unsigned long long *out, hipStream_t s);
void launch_diag_and_weights(int64_t n, const int *rowptr, const int *col, const double *m, int *diag, double *w,
                             int *missing_diag, hipStream_t s);
// deterministic sum of x[0..n) into out[0] (two kernels, no atomics); part >= kMaxGrid doubles
void launch_sum(const double *x, int64_t n, double *part, double *out, hipStream_t s);

// ---------------------------------------------------------------- weak pseudo-periodic set-up
struct PeriodicBox {
    int dir[3];            // periodic axes
    double lo[3], hi[3];   // bounding box of the mesh
    double hmin;           // min over cells of the longest edge
};
struct PlaneGrid {         // centroid bins of the boundary facets lying on one face
    int u, w;              // in-plane axes
    int nb_u, nb_w, off;   // bins per axis, first bin id
    double lo_u, lo_w, h_u, h_w;
};
struct PlaneGrids {
    PlaneGrid p[6];        // face 2d = lo[d], 2d+1 = hi[d]
};
void launch_periodic_mass(int64_t n, int64_t nv, const int *dof_vert, const int *dof_comp, const int *bptr,
                          const int *blist, const int *bnd_v, const int *bnd_c, const int *phase, const double *xyz,
                          const int *dof_of, const int *rowptr, const int *col, const PeriodicBox &B, double *pv,
                          int *rowflag, int *miss, hipStream_t s);
void launch_periodic_flag(int64_t n, const int *dof_vert, const double *xyz, const PeriodicBox &B, int *flag,
                          hipStream_t s);
void launch_bfacet_bin(int64_t nbf, const int *bnd_v, const double *xyz, const PeriodicBox &B, const PlaneGrids &G,
                       int *fbin, int *cnt, hipStream_t s);
void launch_bfacet_fill(int64_t nbf, const int *fbin, const int *binptr, int *cursor, int *binlist, hipStream_t s);
void launch_periodic_entries(int64_t n, int64_t nv, const int *flag, const int *scan, const int *dof_vert,
                             const int *dof_comp, const int *dof_of, const double *xyz, const int *bnd_v,
                             const int *binptr, const int *binlist, const PeriodicBox &B, const PlaneGrids &G,
                             int *ent_dof, int *ent_src, double *ent_w, double *ent_ell, int *bad, hipStream_t s);
void launch_compact_rows(int64_t n, const int *flag, const int *scan, int *rows, hipStream_t s);
// per step: ybc from x through the mirror entries (phase h = q F g per signal), then
// bper = (1-theta) S P ybc on the P rows
void launch_periodic_step(int64_t ne, int64_t nrows, int S, const int *ent_dof, const int *ent_src,
                          const double *ent_w, const double *ent_ell, const int *rows, const int *rowptr,
                          const int *col, const double *pv, const double *sc, const double2 *x, const double *h,
                          double omt, double2 *ybc, double2 *bper, hipStream_t s);

// ---------------------------------------------------------------- solve launchers
// pv: weak pseudo-periodic boundary mass (implicit side only), or null
void launch_prep_scale(int64_t n, const int *diag, const double *m, const double *k, const double *r,
                       const double *gi, const double *pv, double dt, double theta, double *sc, hipStream_t s);
// per-gdir == false: val2[e] = {a, g.J} ; per-gdir == true: val4[e] = {a, Jx, Jy, Jz}; all scaled by sc_i sc_j.
// rowptr/col in the solver numbering; m..jz are the assembled arrays, read at emap[e].
void launch_prep_values(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc,
                        const double *m, const double *k, const double *r, const double *gi, const double *pv,
                        const double *jx, const double *jy, const double *jz, bool per_gdir, double gx, double gy,
                        double gz, double dt, double theta, double *vals, double *bre, hipStream_t s,
                        double theta_b = -1.0);  // explicit-side weight of L (default 1 - theta)
// strong periodic form: mgg = S (D_c M) S in the solver numbering (scomp null: one compartment), and
// per step vals.a = afix + s2a mgg (and areal), bre = bfix + s2b mgg
void launch_sbc_mgg(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc, const double *m,
                    const int *scomp, double D0, double D1, double *mgg, hipStream_t s);
void launch_sbc_step(int64_t nnz, int vw, const double *afix, const double *bfix, const double *mgg, double s2a,
                     double s2b, double *vals, double *bre, double *areal, hipStream_t s);
// the DOF map on the constrained space (is_master, mid from launch_master_dofs / launch_merged_ids)
void launch_constrained_dofs(int64_t n, int64_t nl, const int *is_master, const int *mid, const int *dof_vert,
                             const int *dof_comp, int *dof_of, int *nvert, int *ncomp, hipStream_t s);
// Solver numbering along a Morton curve of the DOF vertices: perm (new -> old), iperm, the
// renumbered CSR (srow, scol) with emap[e'] = old entry index, and per new row the old diagonal
// position sdiag, the signal weight sw and (two compartments) the compartment scomp.
void launch_solver_order(int64_t n, int64_t nnz, const int *dof_vert, const int *dof_comp, const double *xyz,
                         const double lo[3], double ext, const int *rowptr, const int *col, const int *diag,
                         const double *w, bool two_comp, int *perm, int *iperm, int *srow, int *scol, int *emap,
                         int *sdiag, double *sw, int *scomp, hipStream_t s);
void launch_gather_double(int64_t n, const int *map, const double *src, double *dst, hipStream_t s);
// Column compression of the solver CSR in tiles of TR rows: tinfo (ntiles), tptr (ntiles+1),
// ucol (tptr[ntiles]) and lcol (nnz) as used by the kMapTile SpMV kernels.  The arrays are
// allocated by the callee through the allocator callback.
struct TileCompression {
    DevArray<int> tinfo, tptr, ucol;
    DevArray<unsigned short> lcol;
    int TR = 0;
    int64_t nunique = 0;
    int maxnu = 0;  // most distinct columns of a compressed tile
};
void build_tile_compression(int64_t n, int64_t nnz, const int *srow, const int *scol, int TR, TileCompression &tc,
                            hipStream_t s);
// dst[i] = iperm[src[i]] (-1 stays -1)
void launch_renumber(int64_t n, const int *iperm, const int *src, int *dst, hipStream_t s);
// u[perm[i]] = sc_i x[i*S+sig]: a solver-numbered state back in the assembly numbering
void launch_unscale_perm(int64_t n, int S, int sig, const int *perm, const double *sc, const double2 *x, double2 *u,
                         hipStream_t s);
void launch_init_x(int64_t n, int S, const double *sc, double2 *x, hipStream_t s);
// x[i][s] = u0[dof_vert[perm[i]]] / sc[i] (u0 per mesh vertex)
void launch_init_x_vertex(int64_t n, int S, const double *sc, const int *perm, const int *dof_vert, const double *u0,
                          double2 *x, hipStream_t s);
// Gershgorin bound max_i sum_j |a_ij| of the real implicit part of the scaled operator (the
// first value of every vals entry, vw values per entry) into out[0]; part >= kMaxGrid doubles
// Diagonal values of the scaled operator per row into pdiag[vw*i ..] and, into out, a bound of the
// spectrum of the operator scaled by its complex diagonal over every signal and step:
// max over rows i and signals s of max over 0 <= c <= cmax_s of (min(R_i, breal) + c Q_is) / |a_ii + i c j_is|,
// with R_i = sum_j |a_ij|, Q_is = sum_k |g_sk| sum_j |jk_ij|, j_is = |g_s . j_ii| (one direction: the
// stored j, g = 1).  gc = [S][4]: g_s and cmax_s.
void launch_cdiag_prep(int64_t n, const int *rowptr, const int *col, const double *vals, int vw, int S,
                       const double *gc, double breal, double *pdiag, double *part, double *out, hipStream_t s);
void launch_gershgorin(int64_t n, const int *rowptr, const double *vals, int vw, double *part, double *out,
                       hipStream_t s);
// y = Re(A) x for the real implicit part (vw values per entry), and xy = x.*y, yy = y.*y
void launch_real_spmv(int64_t n, const int *rowptr, const int *col, const double *vals, int vw, const double *x,
                      double *y, double *xy, double *yy, hipStream_t s);
// x = y * scale, and the deterministic start vector of the power iteration
void launch_scale_vec(int64_t n, const double *y, double scale, double *x, hipStream_t s);
void launch_power_start(int64_t n, double *x, hipStream_t s);

struct SolveArgs {
    int64_t n;
    const int *rowptr, *col;
    const double *vals;    // 2 or 4 doubles per nnz
    const double *bre;     // explicit-side real part per nnz
    bool per_gdir;
    const double *coefA;   // [S*3] implicit-side imaginary coefficients of this step
    const double *coefB;   // [S*3] explicit-side coefficients of this step
    double2 *x, *r, *rh, *p, *v, *t, *b, *xprev;
    double2 *xpp;                   // mode 2 writes the initial guess gc0 x + gc1 xprev + gc2 xpp here
    // kMapTile: per tile of tile_rows rows, tinfo = number of distinct columns (-1: not compressed),
    // tptr = offset of its sorted distinct columns ucol; lcol = per entry, index into that list
    const int *tinfo, *tptr, *ucol;
    const unsigned short *lcol;
    double gc0, gc1, gc2;
    const double2 *bper;            // weak pseudo-periodic right-hand side of this step, or null
    double2 *ph, *sh;               // preconditioned p and s (polynomial preconditioner), or null
    double2 *s;                     // s = r - alpha v, written by the s-pass before the preconditioner's
                                    // t-kernel, which gathers it; read by the next t-kernel and the
                                    // update (null: the consumers form it from r and v)
    double pc;                      // polynomial scale c: z -> c (2 z - c Are z)
    const double *areal;            // [nnz] the scaled real part alone (preconditioner SpMVs)
    const double *pdiag;            // [n*vw] diagonal values of the scaled operator per row (a, j or
                                    // a, jx, jy, jz): the polynomial preconditioner in the variables
                                    // scaled by the complex diagonal a_ii + i (c_s . j_ii); null = off
    double2 *hp, *hq;               // mode 4 writes Are x and Jg x here (projection history)
    const double *gsig;             // [S*3] unit gradient directions (per-direction values only)
    double *partU, *partV, *partT;  // [grid][S][kNV] each
    double *totU, *totV, *totT;     // [S][kNV] finalize outputs, or null: consumers reduce the partials
    // reduced-precision preconditioner input (COCG, kOpPreDot32), or null: r rounded to FP32 pairs
    // (written beside r by the update kernel; launch_to_f32 converts it at a solve's start) and
    // the preconditioner's operator values in FP32 (one per entry with areal set, else 2 or 4)
    float2 *r32;
    const float *pvals32;
    int nparts_upd;                 // blocks of the vector kernels (partials U of the update kernel)
    KrylovState *st;
    int *all_done;
    double rtol, atol;
    int maxit;
};
// mode 0: b and r from x; 1: r from the stored b (masked signals, nullptr = all); 2: b only and
// xpp <- gc0 x + gc1 xprev + gc2 xpp (the extrapolated initial guess; xpp is not read when
// gc2 = 0); 3: b only; 4: b, hp = Are x and hq = Jg x (Jg = g_s.J with per-signal directions)
void launch_rhs_residual(const SolveGeom &g, const SolveArgs &a, int mode, const int *mask, hipStream_t s);

// Residual-minimising initial guess over the last nh <= kMaxHist states y_0 (newest) .. y_nh-1 of
// the scaled system A_n = Are + i c_s Jg.  With P_j = Are y_j and Q_j = Jg y_j kept from the step
// that produced y_j, the columns W_j = P_j + i c_s Q_j span A_n times the history; the guess
// y0 = sum_j gamma_j y_j minimises |b - A_n y0| per signal.  The normal equations are formed on
// the backward differences of the history (the raw states are nearly parallel, their Gram matrix
// is singular to working precision), Jacobi-scaled and solved by a Cholesky factorisation that
// stops at the first pivot below kProjPivot (the remaining differences are dropped).
constexpr int kMaxHist = 12;
constexpr int kMaxProjGrid = 1024;
constexpr double kProjPivot = 1e-12;
struct ProjArgs {
    int64_t n;
    int S, nh;
    bool per_gdir;
    const double2 *Y[kMaxHist], *P[kMaxHist], *Q[kMaxHist];  // newest first
    const double2 *b;
    const double *coefA;  // [S*3] this step's implicit imaginary coefficients (c_s g_s per direction)
    const double *gsig;   // [S*3] unit directions (per-direction values only)
    double *part;         // [pgrid][S][proj_nv(mh)] block partials
    double *tot;          // [S][proj_nv(mh)]
    double *gam;          // [S][kMaxHist] complex coefficients on the raw states
    double2 *y0;          // the guess
};
// doubles per signal of the normal equations: upper triangle of the Gram matrix and W^H b, complex
constexpr int proj_nv(int mh) { return 2 * (mh * (mh + 1) / 2 + mh); }
// mh: history capacity (4, 6, 8, 10 or 12) >= p.nh; part >= proj_grid() * S * proj_nv(mh) doubles
int proj_grid(const SolveGeom &g);
void launch_proj_guess(const SolveGeom &g, const ProjArgs &p, int mh, hipStream_t s);
void launch_bicg_p(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr);
// op: kOpV / kOpPre for the v-kernel (kOpPre writes ph from p), kOpT / kOpPre / kOpTP for the
// t-kernel (kOpPre writes sh from s = r - alpha v; kOpTP gathers sh).  Mappings 1, 2 and 4 take
// every op, mappings 3 and 5 only kOpV / kOpT.
void launch_spmv_v(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr, int op = kOpV);
void launch_spmv_t(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr, int op = kOpT);
void launch_bicg_update(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
// COCG (conjugate orthogonal conjugate gradients) for the complex-symmetric operator, right-
// preconditioned by the same polynomial M (also complex symmetric).  One iteration: z = M r with
// rho = r^T z (launch_cocg_spmv, pre), the convergence test, the previous iteration's
// x += alpha p and p = z + (rho / rho_prev) p (launch_cocg_p), q = A p with p^T q
// (launch_cocg_spmv), then alpha = rho / p^T q and r -= alpha q with |r|^2 (launch_cocg_update).
// a.ph holds z, a.v holds q.
void launch_cocg_spmv(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                      bool pre);
// rfold: the iteration's |r|^2 comes with rho from launch_cocg_pre_update (partials V) instead of
// from the update kernel
void launch_cocg_p(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr, bool rfold = false);
// The folded iteration's first kernel (kOpPreDotR): update of r from the previous iteration's q,
// z = M r, rho = r^T z and |r|^2; it also turns the previous p-kernel's stop into done and raises
// the all-done flag.  The iteration is then launch_cocg_p(rfold) and launch_cocg_spmv(q): no update
// kernel.
void launch_cocg_pre_update(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                            hipEvent_t e1 = nullptr);
// The iteration's second kernel with the p-step folded in (kOpQFold; per-signal directions): the
// iteration is then launch_cocg_spmv(pre), launch_cocg_qfold and launch_cocg_update.
void launch_cocg_qfold(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr);
void launch_cocg_update(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
// dst[i] = (float)src[i], i < n (round to nearest)
void launch_to_f32(const double *src, int64_t n, float *dst, hipStream_t s);
// alpha = rho / <r^, v> per signal and s = r - alpha v into a.s (the polynomial preconditioner's t-kernel
// then gathers one vector instead of forming s from two)
void launch_spass(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr);
// areal[e] = vals[vw e]
void launch_extract_real(int64_t nnz, const double *vals, int vw, double *areal, hipStream_t s);
// zero st[0..S) and *flag
void launch_reset_state(KrylovState *st, int S, int *flag, hipStream_t s);
// st[0..S) and *flag into host-mapped pinned memory (device addresses of hipHostMalloc buffers)
void launch_publish_state(const KrylovState *st, int S, const int *flag, KrylovState *hst, int *hflag, hipStream_t s);
// per-signal sum_i w_i sc_i x[i*S+s] into out[2*s..2*s+1]
void launch_signal(int64_t n, int S, const double *w, const double *sc, const double2 *x, double *part, double *out,
                   hipStream_t s);
void launch_unscale(int64_t n, int S, int sig, const double *sc, const double2 *x, double2 *u, hipStream_t s);
void launch_copy(const double2 *src, double2 *dst, int64_t n, hipStream_t s);
// ---------------------------------------------------------------- restarted GMRES
constexpr int kGmChunk = 4;        // basis vectors per dot-product launch
constexpr int kGmMaxRestart = 64;  // largest restart length
struct GmArgs {
    int64_t n;
    int S, m;                // signals, restart length
    double2 *V;              // m+1 basis vectors of n*S, vector i at V + i*n*S
    double2 *w;              // the new Arnoldi vector
    double *part;            // block partials
    double *h;               // [S][2(m+1)] this pass's Gram-Schmidt coefficients (complex)
    double *hsum;            // [S][2(m+1)] the sum of both passes
    double2 *H;              // [S][m][m+1] Hessenberg columns (upper triangular after the rotations)
    double *cc;              // [S][m] Givens cosines
    double2 *cs;             // [S][m] Givens sines
    double2 *g;              // [S][m+1] rotated right-hand side
    double2 *y;              // [S][m] the cycle's coefficients
    double *inv;             // [S] 1/|w| (1/|r| at a cycle start)
    int *nk;                 // [S] Arnoldi steps of the cycle
    KrylovState *st;         // tol2, rnorm2, done (for this cycle), conv, iters
    int *all_done;
    double rtol, atol;
    int maxit;
};
void launch_gm_dots(const SolveGeom &g, const GmArgs &a, int i0, int cnt, hipStream_t s);
void launch_gm_fin(const double *parts, int nparts, int S, int pstride, int nv, double *out, int ostride, int o,
                   hipStream_t s);
void launch_gm_orth(const SolveGeom &g, const GmArgs &a, int j, bool norm, hipStream_t s);
void launch_gm_hess(const GmArgs &a, int j, const double *wn2, hipStream_t s);
void launch_gm_scale(const SolveGeom &g, const GmArgs &a, int j, hipStream_t s);
void launch_gm_start(const GmArgs &a, const double *tot, bool first, hipStream_t s);
void launch_gm_v0(const SolveGeom &g, const GmArgs &a, const double2 *r, hipStream_t s);
void launch_gm_combine(const SolveGeom &g, const GmArgs &a, double2 *out, hipStream_t s);
void launch_gm_axpy(int64_t total, const double2 *z, double2 *x, hipStream_t s);
// st[s].done = 0 for every signal and *flag = 0, the rest of the states kept
void launch_reset_state_flags(KrylovState *st, int S, int *flag, hipStream_t s);

// the contraction's values in solver entry order (s_i s_j K, s_i s_j M, s_i s_j I, Jg), once per sweep,
// and the per-step contraction over them (G lanes per row); returns the number of block partials
void launch_adj_values(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc,
                       const double *mm, const double *kk, const double *gi, const double *vals, double4 *adjv,
                       hipStream_t s);
int launch_adj_contract_rows(int64_t n, const int *rowptr, const int *col, const int *dof_comp, const double4 *adjv,
                             const double2 *mu, const double2 *xn, const double2 *xn1, double theta, double *parts,
                             hipStream_t s);
void launch_adj_init_lambda(int64_t n, const double *w, const double *sc, double inv_vol, double2 *lam,
                            hipStream_t s);

}  // namespace dmri
