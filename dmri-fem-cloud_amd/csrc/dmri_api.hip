// libdmri host code: the C ABI of include/dmri.h, the device set-up pipeline and the
// theta-scheme time loop driving the gfx950 kernels of dmri_kernels.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "dmri.h"
#include "dmri_internal.hpp"

using namespace dmri;

namespace {

thread_local std::string g_last_error;

template <class T>
T read_scalar(const T *d, hipStream_t s) {
    T h;
    DMRI_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    return h;
}

constexpr int kEventPool = 512;

struct SolverWork {
    DevArray<double2> x, r, rh, p, v, t, b, xprev, xprev2, ybc, bper, ph, sh, sv;
    DevArray<double2> yh[kMaxHist], hp[kMaxHist], hq[kMaxHist];  // projection history
    DevArray<double> sc, vals, bre, areal, pdiag, part, tot, coef, coefP, sig, spart, ppart, ptot, pgam, gsig, u0v;
    DevArray<double> afix, bfix, mgg;  // strong periodic form: fixed real parts and the q^2 F^2 mass term
    DevArray<double> coefcur;          // graph mode: the current step's coefficients (A then B)
    DevArray<double2> xk;              // graph mode: the projection guess's Krylov iterate
    DevArray<float> r32, pvals32;      // COCG's FP32 preconditioner input: r and the polynomial's values
    // restarted GMRES: basis, new vector, partials, coefficients, Hessenberg, rotations, totals
    DevArray<double2> gV, gw, gH, gcs, gg, gy;
    DevArray<double> gpart, gh, ghsum, gcc, ginv, gtot, gwn;
    DevArray<int> gnk;
    DevArray<KrylovState> st;
    DevArray<int> all_done, mask;
    KrylovState *h_st = nullptr, *d_h_st = nullptr;  // pinned host copy and its device address
    int *h_flag = nullptr, *d_h_flag = nullptr;
    SolverWork() = default;
    SolverWork(const SolverWork &) = delete;
    SolverWork &operator=(const SolverWork &) = delete;
    ~SolverWork() {
        if (h_st) (void)hipHostFree(h_st);
        if (h_flag) (void)hipHostFree(h_flag);
    }
};

}  // namespace

struct dmri_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh
    int64_t nv = 0, nc = 0;
    int ncomp = 1;
    bool has_mesh = false;
    double hmin = 0, hmax = 0;
    double bb_lo[3] = {0, 0, 0}, bb_hi[3] = {0, 0, 0};
    DevArray<double> xyz;
    DevArray<int> tets, phase;
    // coefficients
    double D[2] = {3e-3, 3e-3};
    double invT2[2] = {1e-16, 1e-16};
    double kappa = 0.0;
    DevArray<double> Dc, iT2c, Dcell, iT2cell, kappa_f;
    bool has_Dcell = false, has_iT2cell = false, has_kappa_f = false;
    bool Dcell_symmetric = true;  // every D_cell tensor symmetric (the adjoint needs it)
    // assembled
    bool assembled = false;
    int64_t ndof = 0, nnz = 0, n_iface = 0, n_bfacet = 0;
    int ncol_cells = 0, ncol_iface = 0;
    double volume = 0, setup_ms = 0;
    DevArray<int> dof_of, dof_vert, dof_comp, rowptr, col, diag;
    DevArray<int> vptr, vlist, fptr, flist, iface_v, iface_c, bnd_v, bnd_c;
    DevArray<double> val[7];
    DevArray<double> w;
    // weak pseudo-periodic boundary condition
    int pdir[3] = {0, 0, 0};
    bool strong = false;                   // the periodic directions are imposed by the strong form
    bool strong_active = false;            // assembled with the strong periodic constraint
    std::vector<double> h_xyz;             // host copy of the coordinates (periodic vertex matching)
    DevArray<double> gval[3];              // strong form: (C - B)_d = (C_d - C_d^T)/2 on the constrained space
    bool periodic = false;                 // assembled with periodic directions
    int64_t n_per = 0, n_prow = 0;
    DevArray<double> pval, per_w, per_ell;
    DevArray<int> per_dof, per_src, prow, bptr, blist;
    // solver numbering (Morton order of the DOF vertices); s* arrays are in that numbering
    DevArray<int> perm, iperm, srow, scol, emap, sdiag, scomp, sper_dof, sper_src, sprow;
    DevArray<double> sw, spval;
    // column compression of the solver CSR per tile height (built on first use by a solve)
    std::map<int, std::unique_ptr<TileCompression>> tiles;
    // solver workspace
    SolverWork work;
    DevArray<double> part, dsig;
    int last_S = 1;
    std::vector<std::pair<long long, int>> tuned;  // (S*2 + per_gdir) -> lanes per row
    bool solved = false;
    uint64_t generation = 0;  // bumped by every dmri_assemble: batches made before it are stale
    hipEvent_t ev0[kEventPool], ev1[kEventPool];
    bool events_made = false;

    ~dmri_ctx() {
        if (events_made)
            for (int i = 0; i < kEventPool; ++i) {
                (void)hipEventDestroy(ev0[i]);
                (void)hipEventDestroy(ev1[i]);
            }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class F>
int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return DMRI_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return DMRI_ERR_ALLOC;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return DMRI_ERR_ARG;
    }
}

void require(bool cond, int code, const std::string &msg) {
    if (!cond) throw Error(code, msg);
}

void set_device(dmri_ctx *c) { DMRI_HIP(hipSetDevice(c->device)); }

// ---- staging buffer of dmri_get_solution_noalloc: DMRI_NOALLOC_MAX_DOF complex doubles (32 MiB)
// in every GPU's copy of this code object, held until the process ends.  On the host, `scratch`
// names a stand-in rather than device memory, so host code never hands it to a kernel or a copy:
// each device's address is looked up once, with that device current, and kept in g_scratch.
__device__ __align__(256) double2 scratch[DMRI_NOALLOC_MAX_DOF];

struct ScratchSlot {
    std::mutex mu;  // held for the whole of a call that uses this device's buffer
    double2 *ptr = nullptr;
};
std::mutex g_scratch_mu;
std::map<int, std::unique_ptr<ScratchSlot>> g_scratch;  // device -> its buffer

// The slot of the current device `device`, registered on first use (dmri_create does that).
ScratchSlot &scratch_slot(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return *it->second;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        throw Error(DMRI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    auto slot = std::make_unique<ScratchSlot>();
    slot->ptr = static_cast<double2 *>(p);
    return *(g_scratch[device] = std::move(slot));
}

// vertex -> element incidence lists (sorted)
void build_incidence(int nvert, const int *ev, int64_t ne, int64_t nv, DevArray<int> &ptr, DevArray<int> &list,
                     hipStream_t s) {
    DevArray<int> cnt;
    cnt.alloc(nv + 1);
    cnt.zero(s);
    ptr.alloc(nv + 1);
    if (nvert == 4) launch_count_incidence(ev, ne, cnt.p, s);
    else launch_count_incidence3(ev, ne, cnt.p, s);
    exclusive_scan(cnt.p, ptr.p, nv, s);
    int64_t total = read_scalar(ptr.p + nv, s);
    require(total == (int64_t)nvert * ne, DMRI_ERR_MESH, "incidence count mismatch");
    list.alloc(std::max<int64_t>(total, 1));
    cnt.zero(s);
    if (nvert == 4) launch_fill_incidence(ev, ne, ptr.p, cnt.p, list.p, s);
    else launch_fill_incidence3(ev, ne, ptr.p, cnt.p, list.p, s);
    launch_sort_lists(ptr.p, list.p, nv, s);
}

// Jones-Plassmann colouring; returns colour offsets (size ncolors+1) and the element
// permutation grouped by colour.
std::vector<int64_t> colour_elements(int nvert, const int *ev, int64_t ne, const int *vptr, const int *vlist,
                                     DevArray<int> &perm, hipStream_t s) {
    std::vector<int64_t> offs(1, 0);
    perm.alloc(std::max<int64_t>(ne, 1));
    if (ne == 0) return offs;
    DevArray<int> color, remaining;
    color.alloc(ne);
    DMRI_HIP(hipMemsetAsync(color.p, 0xFF, ne * sizeof(int), s));
    remaining.alloc(1);
    int ncolors = 0;
    for (int k = 0;; ++k) {
        require(k < 4096, DMRI_ERR_MESH, "element colouring did not terminate");
        remaining.zero(s);
        launch_jp_round(k, ne, nvert, ev, vptr, vlist, color.p, remaining.p, s);
        int left = read_scalar(remaining.p, s);
        if (left == 0) {
            ncolors = k + 1;
            break;
        }
    }
    DevArray<int> hist, doffs, cursor;
    hist.alloc(ncolors);
    hist.zero(s);
    launch_color_hist(color.p, ne, hist.p, s);
    std::vector<int> h(ncolors);
    DMRI_HIP(hipMemcpyAsync(h.data(), hist.p, sizeof(int) * ncolors, hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    std::vector<int> ho(ncolors + 1, 0);
    offs.assign(ncolors + 1, 0);
    for (int c = 0; c < ncolors; ++c) {
        ho[c + 1] = ho[c] + h[c];
        offs[c + 1] = ho[c + 1];
    }
    doffs.upload(ho.data(), ncolors + 1, s);
    cursor.alloc(ncolors);
    cursor.zero(s);
    launch_color_scatter(color.p, ne, doffs.p, cursor.p, perm.p, s);
    DMRI_HIP(hipStreamSynchronize(s));
    return offs;
}

// Weak pseudo-periodic set-up: P on the pattern (rows gather their boundary facets), the rows
// holding P, and the mirror entries located on facet bins of the opposite faces.
void assemble_periodic(dmri_ctx *c) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof, nv = c->nv, nbf = c->n_bfacet;
    PeriodicBox B{};
    for (int d = 0; d < 3; ++d) {
        B.dir[d] = c->pdir[d] ? 1 : 0;
        B.lo[d] = c->bb_lo[d];
        B.hi[d] = c->bb_hi[d];
    }
    B.hmin = c->hmin;
    const int *phase = c->ncomp == 2 ? c->phase.p : nullptr;
    build_incidence(3, c->bnd_v.p, nbf, nv, c->bptr, c->blist, s);
    c->pval.alloc(c->nnz);
    c->pval.zero(s);
    DevArray<int> flag, scan, miss;
    flag.alloc(n);
    scan.alloc(n + 1);
    miss.alloc(1);
    miss.zero(s);
    launch_periodic_mass(n, nv, c->dof_vert.p, c->dof_comp.p, c->bptr.p, c->blist.p, c->bnd_v.p, c->bnd_c.p, phase,
                         c->xyz.p, c->dof_of.p, c->rowptr.p, c->col.p, B, c->pval.p, flag.p, miss.p, s);
    require(read_scalar(miss.p, s) == 0, DMRI_ERR_MESH, "periodic boundary entry outside the CSR pattern");
    exclusive_scan(flag.p, scan.p, n, s);
    c->n_prow = read_scalar(scan.p + n, s);
    c->prow.alloc(std::max<int64_t>(c->n_prow, 1));
    launch_compact_rows(n, flag.p, scan.p, c->prow.p, s);
    // facet bins per face: bin edge >= hmax >= any facet's diameter, so the facet that holds a
    // point has its centroid within one bin of the point's bin
    PlaneGrids G{};
    int nbins = 0;
    for (int p = 0; p < 6; ++p) {
        const int d = p / 2;
        PlaneGrid &g = G.p[p];
        g.u = (d + 1) % 3;
        g.w = (d + 2) % 3;
        g.lo_u = c->bb_lo[g.u];
        g.lo_w = c->bb_lo[g.w];
        const double eu = c->bb_hi[g.u] - c->bb_lo[g.u], ew = c->bb_hi[g.w] - c->bb_lo[g.w];
        g.h_u = std::max(c->hmax, eu / 1023.0);
        g.h_w = std::max(c->hmax, ew / 1023.0);
        g.nb_u = (int)std::floor(eu / g.h_u) + 1;
        g.nb_w = (int)std::floor(ew / g.h_w) + 1;
        g.off = nbins;
        nbins += c->pdir[d] ? g.nb_u * g.nb_w : 0;
        if (!c->pdir[d]) g.nb_u = g.nb_w = 0;
    }
    DevArray<int> fbin, cnt, binptr, cursor, binlist, bad;
    fbin.alloc(std::max<int64_t>(nbf, 1));
    cnt.alloc(std::max(nbins, 1));
    cnt.zero(s);
    binptr.alloc(nbins + 1);
    cursor.alloc(std::max(nbins, 1));
    cursor.zero(s);
    launch_bfacet_bin(nbf, c->bnd_v.p, c->xyz.p, B, G, fbin.p, cnt.p, s);
    exclusive_scan(cnt.p, binptr.p, nbins, s);
    const int64_t nplane = read_scalar(binptr.p + nbins, s);
    binlist.alloc(std::max<int64_t>(nplane, 1));
    launch_bfacet_fill(nbf, fbin.p, binptr.p, cursor.p, binlist.p, s);
    // mirror entries
    launch_periodic_flag(n, c->dof_vert.p, c->xyz.p, B, flag.p, s);
    exclusive_scan(flag.p, scan.p, n, s);
    c->n_per = read_scalar(scan.p + n, s);
    const size_t ne = (size_t)std::max<int64_t>(c->n_per, 1);
    c->per_dof.alloc(ne);
    c->per_src.alloc(3 * ne);
    c->per_w.alloc(3 * ne);
    c->per_ell.alloc(3 * ne);
    bad.alloc(1);
    bad.zero(s);
    if (c->n_per > 0) {
        require(nplane > 0, DMRI_ERR_MESH, "periodic direction without flat boundary faces on the bounding box");
        launch_periodic_entries(n, nv, flag.p, scan.p, c->dof_vert.p, c->dof_comp.p, c->dof_of.p, c->xyz.p,
                                c->bnd_v.p, binptr.p, binlist.p, B, G, c->per_dof.p, c->per_src.p, c->per_w.p,
                                c->per_ell.p, bad.p, s);
        require(read_scalar(bad.p, s) == 0, DMRI_ERR_MESH,
                "a mirror point of a periodic face has no boundary facet on the opposite face");
    }
}

// Master vertex of every vertex under the strong periodic constraint (PeriodicBD, DmriFemLib.py:
// 327-375): a vertex within tol of the max face of a periodic axis stands for the vertex at the same
// point moved to the min face (map() moves x by -L along that axis); the chain of maps is followed to
// its end, so a vertex on several max faces takes the vertex at the corner of the min faces.  The
// image is matched within tol (max norm) among the vertices near a min face; it must be unique.
std::vector<int> periodic_masters(const dmri_ctx *c, double tol, int64_t &moved) {
    const int64_t nv = c->nv;
    const double *x = c->h_xyz.data();
    std::vector<int> master(nv);
    auto cell = [&](const double *p, int d) { return (int64_t)std::floor((p[d] - c->bb_lo[d]) / tol); };
    auto hkey = [](int64_t i, int64_t j, int64_t k) {
        return ((uint64_t)(i & 0x1FFFFF) << 42) | ((uint64_t)(j & 0x1FFFFF) << 21) | (uint64_t)(k & 0x1FFFFF);
    };
    std::unordered_multimap<uint64_t, int> grid;
    for (int64_t v = 0; v < nv; ++v) {
        master[v] = (int)v;
        const double *p = x + 3 * v;
        bool near_min = false;
        for (int d = 0; d < 3; ++d) near_min = near_min || (c->pdir[d] && std::fabs(p[d] - c->bb_lo[d]) < tol);
        if (near_min) grid.emplace(hkey(cell(p, 0), cell(p, 1), cell(p, 2)), (int)v);
    }
    moved = 0;
    for (int64_t v = 0; v < nv; ++v) {
        const double *p = x + 3 * v;
        double q[3] = {p[0], p[1], p[2]};
        bool mv = false;
        for (int d = 0; d < 3; ++d)
            if (c->pdir[d] && std::fabs(p[d] - c->bb_hi[d]) < tol) {
                q[d] = c->bb_lo[d];
                mv = true;
            }
        if (!mv) continue;
        int found = -1, count = 0;
        const int64_t ci = cell(q, 0), cj = cell(q, 1), ck = cell(q, 2);
        for (int64_t di = -1; di <= 1; ++di)
            for (int64_t dj = -1; dj <= 1; ++dj)
                for (int64_t dk = -1; dk <= 1; ++dk) {
                    auto rg = grid.equal_range(hkey(ci + di, cj + dj, ck + dk));
                    for (auto it = rg.first; it != rg.second; ++it) {
                        const double *o = x + 3 * (int64_t)it->second;
                        if (std::fabs(o[0] - q[0]) < tol && std::fabs(o[1] - q[1]) < tol && std::fabs(o[2] - q[2]) < tol) {
                            found = it->second;
                            ++count;
                        }
                    }
                }
        require(count == 1, DMRI_ERR_MESH,
                "strong periodic condition: vertex " + std::to_string(v) + " has " + std::to_string(count) +
                    " vertices at its periodic image (the mesh is not periodic within 1e-2 hmin)");
        master[v] = found;
        ++moved;
    }
    return master;
}

template <class T>
void take(DevArray<T> &dst, DevArray<T> &src) {
    std::swap(dst.p, src.p);
    std::swap(dst.n, src.n);
}

// The strong periodic constraint on the assembled unconstrained system: the direction terms
// (C - B)_d = (C_d - C_d^T)/2 (C + C^T = 2B, the boundary integral of every compartment's outer
// normal, for a scalar D per compartment), then every value array becomes P^T X P on the space
// where a DOF on a max face is its master's DOF (DOLFIN's constrained function space), with the DOF
// map, pattern and counts replaced by the constrained ones.
void apply_periodic_constraint(dmri_ctx *c, DevArray<double> *cvals) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof, nv = c->nv, nnz = c->nnz;
    int64_t moved = 0;
    const std::vector<int> master = periodic_masters(c, 1e-2 * c->hmin, moved);
    DevArray<int> dmaster, key, ism, scan, mid, bad;
    dmaster.upload(master.data(), nv, s);
    key.alloc(n);
    ism.alloc(n);
    scan.alloc(n + 1);
    mid.alloc(n);
    bad.alloc(1);
    bad.zero(s);
    launch_master_dofs(n, nv, c->dof_vert.p, c->dof_comp.p, dmaster.p, c->dof_of.p, key.p, ism.p, bad.p, s);
    require(read_scalar(bad.p, s) == 0, DMRI_ERR_MESH,
            "strong periodic condition: a periodic image has no DOF in its master's compartment");
    exclusive_scan(ism.p, scan.p, n, s);
    const int64_t nm = read_scalar(scan.p + n, s);
    launch_merged_ids(n, key.p, scan.p, mid.p, s);
    DevArray<double> g[3];
    {
        DevArray<int> miss;
        miss.alloc(1);
        miss.zero(s);
        for (int d = 0; d < 3; ++d) {
            g[d].alloc(nnz);
            launch_antisym(n, c->rowptr.p, c->col.p, cvals[d].p, g[d].p, miss.p, s);
        }
        require(read_scalar(miss.p, s) == 0, DMRI_ERR_MESH, "the sparsity pattern is not symmetric");
    }
    DevArray<int> mrow, mcol, rstart, sidx;
    const int64_t nu = build_merge(n, nnz, c->rowptr.p, c->col.p, mid.p, nm, mrow, mcol, rstart, sidx, s);
    auto merge = [&](DevArray<double> &a) {
        DevArray<double> t;
        t.alloc(nu);
        launch_run_sum(nu, nnz, rstart.p, sidx.p, a.p, t.p, s);
        take(a, t);
    };
    for (auto &a : c->val) merge(a);
    for (int d = 0; d < 3; ++d) {
        merge(g[d]);
        take(c->gval[d], g[d]);
    }
    DevArray<int> nvert, ncomp;
    nvert.alloc(nm);
    ncomp.alloc(nm);
    launch_constrained_dofs(n, (int64_t)c->ncomp * nv, ism.p, mid.p, c->dof_vert.p, c->dof_comp.p, c->dof_of.p,
                            nvert.p, ncomp.p, s);
    take(c->dof_vert, nvert);
    take(c->dof_comp, ncomp);
    take(c->rowptr, mrow);
    take(c->col, mcol);
    DMRI_HIP(hipStreamSynchronize(s));
    c->ndof = nm;
    c->nnz = nu;
    c->n_per = n - nm;
    (void)moved;
}

// Interface matrix kappa int_Gamma (u0-u1)(v0-v1) into gi (zeroed first), one Jones-Plassmann
// colour of facets per launch; returns the number of colours.
int assemble_interface(dmri_ctx *c, double kappa, const double *kappa_f, double *gi, int *miss) {
    hipStream_t s = c->stream;
    DMRI_HIP(hipMemsetAsync(gi, 0, sizeof(double) * c->nnz, s));
    DevArray<int> perm;
    std::vector<int64_t> offs = colour_elements(3, c->iface_v.p, c->n_iface, c->fptr.p, c->flist.p, perm, s);
    const int ncol = (int)offs.size() - 1;
    for (int k = 0; k < ncol; ++k)
        launch_assemble_facets(offs[k], offs[k + 1] - offs[k], perm.p, c->iface_v.p, c->xyz.p, c->nv, c->dof_of.p,
                               c->rowptr.p, c->col.p, kappa, kappa_f, gi, miss, s);
    DMRI_HIP(hipStreamSynchronize(s));
    return ncol;
}

void assemble_impl(dmri_ctx *c) {
    require(c->has_mesh, DMRI_ERR_STATE, "dmri_assemble: no mesh (call dmri_mesh_set first)");
    ++c->generation;
    c->tiles.clear();
    set_device(c);
    hipStream_t s = c->stream;
    hipEvent_t e0, e1;
    DMRI_HIP(hipEventCreate(&e0));
    DMRI_HIP(hipEventCreate(&e1));
    DMRI_HIP(hipEventRecord(e0, s));
    const int64_t nv = c->nv, nc = c->nc;
    const int *phase = c->ncomp == 2 ? c->phase.p : nullptr;

    // 1. vertex -> cell incidence
    build_incidence(4, c->tets.p, nc, nv, c->vptr, c->vlist, s);

    // 2. DOF map
    {
        const int64_t nl = (int64_t)c->ncomp * nv;
        DevArray<int> live, scan;
        live.alloc(nl);
        live.zero(s);
        scan.alloc(nl + 1);
        launch_mark_live(c->tets.p, phase, nc, nv, live.p, s);
        exclusive_scan(live.p, scan.p, nl, s);
        c->ndof = read_scalar(scan.p + nl, s);
        require(c->ndof > 0, DMRI_ERR_MESH, "mesh has no live DOF");
        c->dof_of.alloc(nl);
        c->dof_vert.alloc(c->ndof);
        c->dof_comp.alloc(c->ndof);
        launch_dof_from_scan(live.p, scan.p, nl, c->dof_of.p, nv, c->dof_vert.p, c->dof_comp.p, s);
    }

    // 3. faces: interface (phase jump) and boundary facets
    {
        const int64_t nfc = 4 * nc;
        DevArray<int> iflag, bflag, nbr, iscan, bscan;
        iflag.alloc(nfc);
        bflag.alloc(nfc);
        nbr.alloc(nfc);
        iscan.alloc(nfc + 1);
        bscan.alloc(nfc + 1);
        launch_face_classify(c->tets.p, phase, nc, c->vptr.p, c->vlist.p, iflag.p, bflag.p, nbr.p, s);
        exclusive_scan(iflag.p, iscan.p, nfc, s);
        exclusive_scan(bflag.p, bscan.p, nfc, s);
        c->n_iface = read_scalar(iscan.p + nfc, s);
        c->n_bfacet = read_scalar(bscan.p + nfc, s);
        c->iface_v.alloc(std::max<int64_t>(3 * c->n_iface, 1));
        c->iface_c.alloc(std::max<int64_t>(2 * c->n_iface, 1));
        c->bnd_v.alloc(std::max<int64_t>(3 * c->n_bfacet, 1));
        c->bnd_c.alloc(std::max<int64_t>(c->n_bfacet, 1));
        launch_face_compact(c->tets.p, nc, nbr.p, iflag.p, iscan.p, bflag.p, bscan.p, c->iface_v.p, c->iface_c.p,
                            c->bnd_v.p, c->bnd_c.p, s);
    }
    if (c->has_kappa_f)
        require((int64_t)c->kappa_f.n == c->n_iface, DMRI_ERR_ARG,
                "kappa_facet length does not match the interface facet count");
    const bool have_iface = c->ncomp == 2 && c->n_iface > 0;
    if (have_iface) build_incidence(3, c->iface_v.p, c->n_iface, nv, c->fptr, c->flist, s);

    // 4. CSR pattern
    {
        DevArray<int> cnt, overflow;
        cnt.alloc(c->ndof);
        overflow.alloc(1);
        overflow.zero(s);
        c->rowptr.alloc(c->ndof + 1);
        const int *fptr = have_iface ? c->fptr.p : nullptr;
        const int *flist = have_iface ? c->flist.p : nullptr;
        launch_pattern(0, c->ndof, nv, c->dof_vert.p, c->dof_comp.p, c->tets.p, phase, c->vptr.p, c->vlist.p, fptr,
                       flist, c->iface_v.p, c->dof_of.p, cnt.p, nullptr, nullptr, overflow.p, s);
        require(read_scalar(overflow.p, s) == 0, DMRI_ERR_LIMIT,
                "a vertex touches more than 512 cells; pattern build limit exceeded");
        // the CSR offsets are 32-bit: check the entry count in 64 bits before the 32-bit scan
        DevArray<unsigned long long> total;
        total.alloc(1);
        total.zero(s);
        launch_count_sum64(cnt.p, c->ndof, total.p, s);
        const unsigned long long nnz64 = read_scalar(total.p, s);
        require(nnz64 < (1ULL << 31), DMRI_ERR_LIMIT,
                "the sparsity pattern has " + std::to_string(nnz64) + " entries; the limit is 2^31 - 1");
        exclusive_scan(cnt.p, c->rowptr.p, c->ndof, s);
        c->nnz = read_scalar(c->rowptr.p + c->ndof, s);
        require(c->nnz > 0, DMRI_ERR_MESH, "empty pattern");
        c->col.alloc(c->nnz);
        launch_pattern(1, c->ndof, nv, c->dof_vert.p, c->dof_comp.p, c->tets.p, phase, c->vptr.p, c->vlist.p, fptr,
                       flist, c->iface_v.p, c->dof_of.p, nullptr, c->rowptr.p, c->col.p, overflow.p, s);
    }

    // 5. cell matrices by row gather (every entry written once); interface facets by colour
    const bool strong = c->strong && (c->pdir[0] || c->pdir[1] || c->pdir[2]);
    require(!strong || !c->has_Dcell, DMRI_ERR_ARG,
            "the strong periodic condition needs a scalar D per compartment (no D_cell field)");
    for (auto &a : c->val) a.alloc(c->nnz);
    c->val[DMRI_MAT_I].zero(s);
    DevArray<double> cvals[3];  // strong form: int 2 D (d phi_j / d x_d) phi_i per axis
    if (strong)
        for (auto &a : cvals) a.alloc(c->nnz);
    DevArray<int> miss;
    miss.alloc(1);
    miss.zero(s);
    c->ncol_cells = 0;
    launch_assemble_rows(c->ndof, c->dof_vert.p, c->dof_comp.p, c->rowptr.p, c->col.p, c->tets.p, phase, c->xyz.p,
                         c->vptr.p, c->vlist.p, c->Dc.p, c->has_Dcell ? c->Dcell.p : nullptr, c->iT2c.p,
                         c->has_iT2cell ? c->iT2cell.p : nullptr, c->val[DMRI_MAT_M].p, c->val[DMRI_MAT_K].p,
                         c->val[DMRI_MAT_R].p, c->val[DMRI_MAT_JX].p, c->val[DMRI_MAT_JY].p, c->val[DMRI_MAT_JZ].p, s,
                         strong ? cvals[0].p : nullptr, strong ? cvals[1].p : nullptr, strong ? cvals[2].p : nullptr);
    c->ncol_iface = 0;
    if (have_iface)
        c->ncol_iface = assemble_interface(c, c->kappa, c->has_kappa_f ? c->kappa_f.p : nullptr,
                                           c->val[DMRI_MAT_I].p, miss.p);
    require(read_scalar(miss.p, s) == 0, DMRI_ERR_MESH, "element entry outside the CSR pattern");

    // 5b. weak pseudo-periodic boundary mass and mirror entries, or the strong periodic constraint
    c->periodic = !strong && (c->pdir[0] || c->pdir[1] || c->pdir[2]) && c->n_bfacet > 0;
    c->n_per = c->n_prow = 0;
    c->strong_active = strong;
    for (auto &a : c->gval) a.release();
    if (c->periodic) assemble_periodic(c);
    if (strong) apply_periodic_constraint(c, cvals);

    // 6. diagonal positions, signal weights w = M 1, volume
    {
        DevArray<int> missing;
        missing.alloc(1);
        missing.zero(s);
        c->diag.alloc(c->ndof);
        c->w.alloc(c->ndof);
        launch_diag_and_weights(c->ndof, c->rowptr.p, c->col.p, c->val[DMRI_MAT_M].p, c->diag.p, c->w.p, missing.p,
                                s);
        require(read_scalar(missing.p, s) == 0, DMRI_ERR_MESH, "missing diagonal entry");
        c->part.alloc((size_t)kMaxGrid * kNV);
        c->dsig.alloc(kNV);
        launch_sum(c->w.p, c->ndof, c->part.p, c->dsig.p, s);
        c->volume = read_scalar(c->dsig.p, s);
    }

    // 7. solver numbering
    {
        const int64_t n = c->ndof;
        c->perm.alloc(n);
        c->iperm.alloc(n);
        c->srow.alloc(n + 1);
        c->scol.alloc(c->nnz);
        c->emap.alloc(c->nnz);
        c->sdiag.alloc(n);
        c->sw.alloc(n);
        if (c->ncomp == 2) c->scomp.alloc(n);
        else c->scomp.release();
        double ext = 0.0;
        for (int d = 0; d < 3; ++d) ext = std::max(ext, c->bb_hi[d] - c->bb_lo[d]);
        launch_solver_order(n, c->nnz, c->dof_vert.p, c->dof_comp.p, c->xyz.p, c->bb_lo, ext, c->rowptr.p, c->col.p,
                            c->diag.p, c->w.p, c->ncomp == 2, c->perm.p, c->iperm.p, c->srow.p, c->scol.p, c->emap.p,
                            c->sdiag.p, c->sw.p, c->ncomp == 2 ? c->scomp.p : nullptr, s);
        if (c->periodic) {
            c->spval.alloc(c->nnz);
            launch_gather_double(c->nnz, c->emap.p, c->pval.p, c->spval.p, s);
            const size_t ne = (size_t)std::max<int64_t>(c->n_per, 1);
            c->sper_dof.alloc(ne);
            c->sper_src.alloc(3 * ne);
            c->sprow.alloc(std::max<int64_t>(c->n_prow, 1));
            launch_renumber(c->n_per, c->iperm.p, c->per_dof.p, c->sper_dof.p, s);
            launch_renumber(3 * c->n_per, c->iperm.p, c->per_src.p, c->sper_src.p, s);
            launch_renumber(c->n_prow, c->iperm.p, c->prow.p, c->sprow.p, s);
        }
    }
    DMRI_HIP(hipEventRecord(e1, s));
    DMRI_HIP(hipEventSynchronize(e1));
    float ms = 0;
    DMRI_HIP(hipEventElapsedTime(&ms, e0, e1));
    c->setup_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->assembled = true;
    c->solved = false;
}

// SpMV mapping and lanes per row when neither the caller nor the tuner chose: vector CSR for
// one signal, LDS-staged rows with one lane per (row, signal) up to 8 signals, one wave per row
// (lane = signal) beyond.
void default_mapping(int S, int &mapping, int &lanes) {
    if (S == 1) {
        mapping = kMapSL;
        lanes = 8;
    } else if (S <= kMaxSRSignals) {
        mapping = kMapStaged;
        lanes = 1;
    } else {
        mapping = kMapWave;
        lanes = 1;
    }
}
int auto_lanes(int S) {
    int m, l;
    default_mapping(S, m, l);
    return l;
}

void ensure_group_workspace(dmri_ctx *c, SolverWork &w, const SolveGeom &g, int nsteps, bool per_gdir) {
    hipStream_t s = c->stream;
    const size_t nS = (size_t)c->ndof * g.S;
    if (w.x.n != nS) {
        w.x.alloc(nS);
        w.r.alloc(nS);
        w.rh.alloc(nS);
        w.p.alloc(nS);
        w.v.alloc(nS);
        w.t.alloc(nS);
        w.b.alloc(nS);
        w.xprev.alloc(nS);
        w.xprev2.alloc(nS);
    }
    if (w.sc.n != (size_t)c->ndof) w.sc.alloc(c->ndof);
    const size_t nval = (size_t)c->nnz * (per_gdir ? 4 : 2);
    if (w.vals.n != nval) w.vals.alloc(nval);
    if (w.bre.n != (size_t)c->nnz) w.bre.alloc(c->nnz);
    const size_t np = (size_t)std::max(g.grid, g.vgrid) * g.S * kNV;
    if (w.part.n < 3 * np) w.part.alloc(3 * np);
    if (w.tot.n < (size_t)3 * g.S * kNV) w.tot.alloc(3 * (size_t)g.S * kNV);
    // at least one step's worth: the host always uploads max(N, 1) steps of coefficients
    const size_t ncoef = (size_t)2 * std::max(nsteps, 1) * g.S * 3;
    if (w.coef.n < ncoef) w.coef.alloc(ncoef);
    if (w.st.n < (size_t)g.S) w.st.alloc(kMaxSignals);
    if (!w.all_done.p) w.all_done.alloc(1);
    if (!w.mask.p) w.mask.alloc(kMaxSignals);
    if (w.sig.n < (size_t)kMaxSignals * kNV) w.sig.alloc(kMaxSignals * kNV);
    if (w.spart.n < (size_t)kMaxGrid * kMaxSignals * kNV) w.spart.alloc((size_t)kMaxGrid * kMaxSignals * kNV);
    if (!w.h_st) {
        DMRI_HIP(hipHostMalloc((void **)&w.h_st, sizeof(KrylovState) * kMaxSignals,
                               hipHostMallocMapped | hipHostMallocCoherent));
        DMRI_HIP(hipHostGetDevicePointer((void **)&w.d_h_st, w.h_st, 0));
    }
    if (!w.h_flag) {
        DMRI_HIP(hipHostMalloc((void **)&w.h_flag, sizeof(int) * 4, hipHostMallocMapped | hipHostMallocCoherent));
        DMRI_HIP(hipHostGetDevicePointer((void **)&w.d_h_flag, w.h_flag, 0));
    }
    if (!c->events_made) {
        for (int i = 0; i < kEventPool; ++i) {
            DMRI_HIP(hipEventCreate(&c->ev0[i]));
            DMRI_HIP(hipEventCreate(&c->ev1[i]));
        }
        c->events_made = true;
    }
    (void)s;
}

// One launch group of S <= 64 signals sharing the time profile.  q[s], g[3*s] (unit).
// Theta-scheme stepping engine for one launch group of S <= 64 signals sharing the time profile.
// Scale c of the polynomial preconditioner p(A) = c (2I - c A): c = 1.8 / b with b an upper
// estimate of lambda_max of the real, symmetric part of the scaled implicit operator (diagonal 1):
// 1.05 x the Rayleigh quotient after 30 power iterations, capped by the Gershgorin bound.  The
// preconditioned real spectrum 1 - (1 - c lambda)^2 then stays in (0, 1] while the smallest
// eigenvalues grow by about 2c.
// largest row sum of |vals[vw e]| over the solver pattern
double row_abs_sum_max(dmri_ctx *c, const double *vals, int vw) {
    DevArray<double> part, out;
    part.alloc((size_t)kMaxGrid * kNV);
    out.alloc(kNV);
    launch_gershgorin(c->ndof, c->srow.p, vals, vw, part.p, out.p, c->stream);
    return read_scalar(out.p, c->stream);
}

double poly_scale(dmri_ctx *c, const double *vals, int vw) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    DevArray<double> x, y, xy, yy, part, out;
    x.alloc(n);
    y.alloc(n);
    xy.alloc(n);
    yy.alloc(n);
    part.alloc((size_t)kMaxGrid * kNV);
    out.alloc(kNV);
    launch_gershgorin(n, c->srow.p, vals, vw, part.p, out.p, s);
    const double gersh = read_scalar(out.p, s);
    launch_power_start(n, x.p, s);
    launch_real_spmv(n, c->srow.p, c->scol.p, vals, vw, x.p, y.p, xy.p, yy.p, s);
    double rq = 0.0;
    for (int k = 0; k < 30; ++k) {
        // x <- y / |y|, y <- A x, rq = x.y
        launch_sum(yy.p, n, part.p, out.p, s);
        const double ny = std::sqrt(read_scalar(out.p, s));
        if (!(ny > 0.0)) break;
        launch_scale_vec(n, y.p, 1.0 / ny, x.p, s);
        launch_real_spmv(n, c->srow.p, c->scol.p, vals, vw, x.p, y.p, xy.p, yy.p, s);
        launch_sum(xy.p, n, part.p, out.p, s);
        rq = read_scalar(out.p, s);
    }
    double b = rq > 0.0 ? 1.05 * rq : gersh;
    if (gersh > 0.0) b = std::min(b, gersh);
    return b > 0.0 ? 1.8 / b : 1.0;
}

// Tile compression for a kMapTile geometry (built once per tile height and kept), with the
// staged-column capacity of the launch set from its largest tile; null for other mappings.
const TileCompression *tiles_for(dmri_ctx *c, SolveGeom &geo) {
    if (geo.mapping != kMapTile) return nullptr;
    auto it = c->tiles.find(geo.tile_rows);
    if (it == c->tiles.end()) {
        auto tc = std::make_unique<TileCompression>();
        build_tile_compression(c->ndof, c->nnz, c->srow.p, c->scol.p, geo.tile_rows, *tc, c->stream);
        it = c->tiles.emplace(geo.tile_rows, std::move(tc)).first;
    }
    geo.tile_ucap = std::max(1, std::min(kTileLdsBytes / (16 * geo.S), it->second->maxnu));
    return it->second.get();
}
void attach_tiles(SolveArgs &a, const TileCompression *tc) {
    if (!tc) return;
    a.tinfo = tc->tinfo.p;
    a.tptr = tc->tptr.p;
    a.ucol = tc->ucol.p;
    a.lcol = tc->lcol.p;
}

// History of the projection guess: up to m states newest first, the Are / Jg products of each
// (same order), and the free buffers.  m + 1 state buffers circulate: the m states and the one
// the next guess is written to.
struct ProjHistory {
    int m = 0;  // capacity, 0 = off
    std::deque<double2 *> ys, yfree;
    std::deque<std::pair<double2 *, double2 *>> pq, pqfree;
    void reset(int m_, const DevArray<double2> *yb, const DevArray<double2> *pb, const DevArray<double2> *qb) {
        m = m_;
        ys.clear();
        yfree.clear();
        pq.clear();
        pqfree.clear();
        for (int j = 0; j < m; ++j) {
            yfree.push_back(yb[j].p);
            pqfree.emplace_back(pb[j].p, qb[j].p);
        }
    }
    int nh() const { return (int)std::min(ys.size(), pq.size()); }
    // buffers for the products of the state that becomes the newest (the oldest pair when full)
    std::pair<double2 *, double2 *> take_slot() {
        std::pair<double2 *, double2 *> sl;
        if (!pqfree.empty()) {
            sl = pqfree.front();
            pqfree.pop_front();
        } else {
            sl = pq.back();
            pq.pop_back();
        }
        return sl;
    }
    double2 *take_free() {
        double2 *y = yfree.front();
        yfree.pop_front();
        return y;
    }
    void push(double2 *y, std::pair<double2 *, double2 *> sl) {
        pq.push_front(sl);
        ys.push_front(y);
        if ((int)ys.size() > m) {
            yfree.push_back(ys.back());
            ys.pop_back();
        }
    }
};

// capacity of a projection history of S-signal states: 12, 10, 8, 6 or 4 states when 3 x capacity
// state-sized buffers take at most half of the free device memory (plus what the caller already
// holds for it), else 0
int proj_capacity(size_t nS, size_t held_bytes) {
    size_t fr = 0, tot = 0;
    DMRI_HIP(hipMemGetInfo(&fr, &tot));
    for (int m : {12, 10, 8, 6, 4})
        if ((double)m * 3.0 * (double)nS * sizeof(double2) <= 0.5 * (double)(fr + held_bytes)) return m;
    return 0;
}

struct Stepper {
    dmri_ctx *c = nullptr;
    SolverWork *wk = nullptr;  // the launch group's buffers (the context's own, or a batch group's)
    hipStream_t s = nullptr;
    int S = 1, N = 0;
    int64_t n = 0;
    bool per_gdir = false, extrap = true, prof = false;
    SolveGeom geo;
    const TileCompression *tiles = nullptr;
    SolveArgs a{};
    const dmri_solver_opts *opts = nullptr;
    dmri_stats *st = nullptr;
    double2 *xc = nullptr, *xp = nullptr, *xq = nullptr;  // u^n, u^(n-1), u^(n-2)
    int guess_order = 2;
    ProjHistory fh;  // projection guess of the forward steps (fh.m = 0: off); fh.ys[0] = xc
    bool poly = true;       // degree-1 polynomial preconditioner
    bool real_pre = false;  // the polynomial of the real part alone
    double kbytes[7] = {0, 0, 0, 0, 0, 0, 0};  // algorithmic bytes per launch of the iteration kernels
    int precond = 0;                         // dmri_stats.precond_used of this launch group
    bool strong = false;                     // the strong periodic form (per-step real parts)
    bool gmres_on = false;                   // restarted GMRES instead of BiCGStab
    bool cocg_on = false;                    // COCG instead of BiCGStab (complex-symmetric operators)
    bool pre32 = false;                      // COCG's preconditioner SpMV reads FP32 r and values
    bool cocg_fold = false;                  // COCG's update of r folded into the preconditioner SpMV
    bool cocg_qfold = false;                 // COCG's p-step folded into the q = A p SpMV
    GmArgs gm{};
    std::vector<double> s2A, s2B;            // strong form: theta q^2 F^2 per step, implicit and explicit
    bool timed_first[kEventPool / 7] = {};   // timed iteration k of a chunk is the solve's first
    double omt = 0.5;
    int chunk = 8;
    // hipGraph replay of BiCGStab iterations (opts->graphs): the captured runs, keyed by the kernel
    // arguments, the first iteration and the count
    struct ChainGraph {
        SolveArgs args;
        int it0, count;
        hipGraph_t g;
        hipGraphExec_t e;
    };
    bool graphs_on = false;
    std::vector<ChainGraph> graphs;
    Stepper() = default;
    Stepper(const Stepper &) = delete;
    Stepper &operator=(const Stepper &) = delete;
    ~Stepper() { drop_graphs(); }
    void drop_graphs() {
        if (!graphs.empty() && s) (void)hipStreamSynchronize(s);  // no replay still running
        for (auto &cg : graphs) {
            (void)hipGraphExecDestroy(cg.e);
            (void)hipGraphDestroy(cg.g);
        }
        graphs.clear();
    }

    // q[s], g[3*s] (unit).  Scales the system, uploads the per-step coefficients, sets u0 = 1.
    void setup(dmri_ctx *ctx, int S_, const double *q, const double *g, bool pg, const dmri_seq *seq,
               const dmri_solver_opts *o, dmri_stats *stats, SolverWork *work = nullptr) {
        c = ctx;
        wk = work ? work : &c->work;
        s = c->stream;
        S = S_;
        n = c->ndof;
        per_gdir = pg;
        opts = o;
        st = stats;
        N = seq->nsteps;
        int mapping, lanes;
        default_mapping(S, mapping, lanes);
        if (opts->lanes_per_row <= 0 && opts->mapping <= 0) {
            for (const auto &t : c->tuned)
                if (t.first == (long long)S * 2 + (per_gdir ? 1 : 0)) {
                    mapping = t.second / 64;
                    lanes = t.second % 64;
                }
        } else {
            if (opts->mapping > 0) mapping = opts->mapping;
            if (opts->lanes_per_row > 0) lanes = opts->lanes_per_row;
        }
        poly = opts->preconditioner == 0 || opts->preconditioner == 2;
        if ((poly && (mapping == kMapSR || mapping == kMapTile)) ||  // kernels without the polynomial ops
            (opts->krylov == 2 && mapping != kMapSL && mapping != kMapStaged)) {  // COCG's SpMVs: mappings 1, 2
            mapping = kMapStaged;
            while (lanes > 1 && lanes * S > 64) lanes >>= 1;
        }
        geo = make_geom(n, lanes, S, mapping);
        fit_spmv_grid(geo, per_gdir);
        tiles = tiles_for(c, geo);
        ensure_group_workspace(c, *wk, geo, N, per_gdir);
        SolverWork &w = *wk;
        auto &V = c->val;
        const double theta = seq->theta, dt = seq->dt;
        const double *pv = c->periodic ? c->pval.p : nullptr;
        launch_prep_scale(n, c->sdiag.p, V[DMRI_MAT_M].p, V[DMRI_MAT_K].p, V[DMRI_MAT_R].p, V[DMRI_MAT_I].p, pv, dt,
                          theta, w.sc.p, s);
        // the strong periodic form (ThetaMethodF/L_sBC1c/2c, DmriFemLib.py:167-238) takes theta on both
        // sides, the direction terms (C - B).g in place of the gradient moments and F = int f in place of f
        strong = c->strong_active;
        const double *j3[3] = {V[DMRI_MAT_JX].p, V[DMRI_MAT_JY].p, V[DMRI_MAT_JZ].p};
        if (strong)
            for (int d = 0; d < 3; ++d) j3[d] = c->gval[d].p;
        launch_prep_values(n, c->srow.p, c->scol.p, c->emap.p, w.sc.p, V[DMRI_MAT_M].p, V[DMRI_MAT_K].p,
                           V[DMRI_MAT_R].p, V[DMRI_MAT_I].p, pv, j3[0], j3[1], j3[2], per_gdir, g[0], g[1], g[2], dt,
                           theta, w.vals.p, w.bre.p, s, strong ? theta : 1.0 - theta);
        omt = 1.0 - theta;
        const int vw0 = per_gdir ? 4 : 2;
        const double *fimp = strong ? seq->F_implicit : seq->f_implicit;
        const double *fexp = strong ? seq->F_explicit : seq->f_explicit;
        double s2max = 0.0;
        if (strong) {
            for (int j = 1; j < S; ++j)
                require(q[j] == q[0], DMRI_ERR_ARG, "strong periodic form: the signals of a launch group share q");
            s2A.assign(std::max(N, 1), 0.0);
            s2B.assign(std::max(N, 1), 0.0);
            for (int k = 0; k < N; ++k) {
                s2A[k] = theta * q[0] * q[0] * fimp[k] * fimp[k];
                s2B[k] = -theta * q[0] * q[0] * fexp[k] * fexp[k];
                s2max = std::max(s2max, s2A[k]);
            }
            if (w.afix.n != (size_t)c->nnz) w.afix.alloc(c->nnz);
            if (w.bfix.n != (size_t)c->nnz) w.bfix.alloc(c->nnz);
            if (w.mgg.n != (size_t)c->nnz) w.mgg.alloc(c->nnz);
            launch_extract_real(c->nnz, w.vals.p, vw0, w.afix.p, s);
            DMRI_HIP(hipMemcpyAsync(w.bfix.p, w.bre.p, sizeof(double) * c->nnz, hipMemcpyDeviceToDevice, s));
            launch_sbc_mgg(n, c->srow.p, c->scol.p, c->emap.p, w.sc.p, V[DMRI_MAT_M].p,
                           c->ncomp == 2 ? c->scomp.p : nullptr, c->D[0], c->D[1], w.mgg.p, s);
            // the preconditioner's bounds are taken of the largest real part of the sequence
            launch_sbc_step(c->nnz, vw0, w.afix.p, w.bfix.p, w.mgg.p, s2max, 0.0, w.vals.p, w.bre.p, nullptr, s);
        }
        if (c->periodic) {
            // mirror phase per step and signal: q_s F(t_{n-1}) g_s
            std::vector<double> hP((size_t)std::max(N, 1) * S * 3, 0.0);
            for (int k = 0; k < N; ++k)
                for (int j = 0; j < S; ++j)
                    for (int d = 0; d < 3; ++d) {
                        const double gd = g[3 * j + d];
                        hP[((size_t)k * S + j) * 3 + d] = q[j] * seq->F_explicit[k] * gd;
                    }
            if (!hP.empty()) {
                if (w.coefP.n < hP.size()) w.coefP.alloc(hP.size());
                DMRI_HIP(hipMemcpyAsync(w.coefP.p, hP.data(), sizeof(double) * hP.size(), hipMemcpyHostToDevice, s));
            }
            const size_t nS = (size_t)n * S;
            if (w.ybc.n != nS) w.ybc.alloc(nS);
            if (w.bper.n != nS) w.bper.alloc(nS);
            w.ybc.zero(s);
            w.bper.zero(s);
        }
        // per-step imaginary coefficients: implicit side theta q f_A(n) (g), explicit -(1-theta) q f_B(n) (g)
        std::vector<double> coef((size_t)2 * std::max(N, 1) * S * 3, 0.0);
        double *cA = coef.data(), *cB = coef.data() + (size_t)std::max(N, 1) * S * 3;
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < S; ++j) {
                const double av = theta * q[j] * fimp[k];
                const double bv = -(strong ? theta : 1.0 - theta) * q[j] * fexp[k];
                for (int d = 0; d < 3; ++d) {
                    const double gd = per_gdir ? g[3 * j + d] : (d == 0 ? 1.0 : 0.0);
                    cA[((size_t)k * S + j) * 3 + d] = av * gd;
                    cB[((size_t)k * S + j) * 3 + d] = bv * gd;
                }
            }
        DMRI_HIP(hipMemcpyAsync(w.coef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, s));
        if (seq->u0_vertex) {
            if (w.u0v.n != (size_t)c->nv) w.u0v.alloc(c->nv);
            DMRI_HIP(hipMemcpyAsync(w.u0v.p, seq->u0_vertex, sizeof(double) * c->nv, hipMemcpyHostToDevice, s));
            launch_init_x_vertex(n, S, w.sc.p, c->perm.p, c->dof_vert.p, w.u0v.p, w.x.p, s);
        } else {
            launch_init_x(n, S, w.sc.p, w.x.p, s);
        }
        // Finalize kernels reduce the partials unless one signal's few partials can be summed by
        // every consumer block directly (mixed first/later iterations need per-signal counts).
        const bool use_fin = S > 1 || std::max(geo.grid, geo.vgrid) > 512;
        const size_t np = (size_t)std::max(geo.grid, geo.vgrid) * S * kNV;
        a = SolveArgs{};
        a.n = n;
        a.rowptr = c->srow.p;
        a.col = c->scol.p;
        a.vals = w.vals.p;
        a.bre = w.bre.p;
        a.per_gdir = per_gdir;
        xc = w.x.p;
        xp = w.xprev.p;
        xq = w.xprev2.p;
        // initial guess of step n: 0 = residual-minimising projection on the last states (the
        // quadratic extrapolation when the history does not fit), 1 = u^n, 2 = linear
        // 2u^n - u^(n-1), 3 = quadratic extrapolation 3u^n - 3u^(n-1) + u^(n-2)
        guess_order = opts->initial_guess == 1 ? 0 : (opts->initial_guess == 2 ? 1 : 2);
        extrap = guess_order > 0;
        // the projection keeps products with the real part, which the strong form changes every step:
        // it takes the quadratic extrapolation instead
        if (opts->initial_guess == 0 && !strong) setup_projection(g);
        a.x = xc;
        a.xprev = xp;
        a.r = w.r.p;
        a.rh = w.rh.p;
        a.p = w.p.p;
        a.v = w.v.p;
        a.t = w.t.p;
        a.b = w.b.p;
        a.partU = w.part.p;
        a.partV = w.part.p + np;
        a.partT = w.part.p + 2 * np;
        a.totU = use_fin ? w.tot.p : nullptr;
        a.totV = use_fin ? w.tot.p + S * kNV : nullptr;
        a.totT = use_fin ? w.tot.p + 2 * S * kNV : nullptr;
        a.st = w.st.p;
        a.all_done = w.all_done.p;
        a.bper = c->periodic ? w.bper.p : nullptr;
        attach_tiles(a, tiles);
        if (poly) {
            const size_t nS = (size_t)n * S;
            if (w.ph.n != nS) w.ph.alloc(nS);
            if (w.sh.n != nS) w.sh.alloc(nS);
            if (w.sv.n != nS) w.sv.alloc(nS);
            a.ph = w.ph.p;
            a.sh = w.sh.p;
            a.s = w.sv.p;
            const int vw = per_gdir ? 4 : 2;
            a.pc = poly_scale(c, w.vals.p, vw);
            // The polynomial takes the real part alone when the imaginary part of the scaled
            // operator is small beside it: the iterations do not change, and the preconditioner
            // SpMVs read one value per entry instead of vw.  Where the gradient term is large
            // (high-q OGSE) the real part alone would double the iterations.
            double jb = 0.0;
            for (int d = 0; d < vw - 1; ++d) jb = std::max(jb, row_abs_sum_max(c, w.vals.p + 1 + d, vw));
            double qf = 0.0, gs = per_gdir ? 0.0 : 1.0;
            for (int j = 0; j < S; ++j) {
                qf = std::max(qf, std::fabs(q[j]));
                if (per_gdir) gs = std::max(gs, std::fabs(g[3 * j]) + std::fabs(g[3 * j + 1]) + std::fabs(g[3 * j + 2]));
            }
            double fmax = 0.0;
            for (int k = 0; k < N; ++k) fmax = std::max(fmax, std::fabs(fimp[k]));
            const double imag = theta * qf * fmax * gs * jb;
            const double breal = 1.8 / a.pc;
            real_pre = opts->preconditioner == 0 && imag <= kRealPreconditioner * breal;
            st->precond_used = 1;
            if (real_pre) {
                if (w.areal.n != (size_t)c->nnz) w.areal.alloc(c->nnz);
                launch_extract_real(c->nnz, w.vals.p, vw, w.areal.p, s);
                a.areal = w.areal.p;
                st->precond_used = 2;
            } else if ((geo.mapping == kMapSL || geo.mapping == kMapStaged || geo.mapping == kMapPair) && !strong &&
                       (opts->preconditioner == 2 || imag >= kComplexDiagonal * breal)) {
                // The gradient term dominates the diagonal of rows far along g: the polynomial then
                // acts on the operator scaled by its complex diagonal, whose spectrum stays near
                // the real part's; c = 1.8 / (a bound of that spectrum over every step and signal).
                std::vector<double> gc((size_t)4 * S);
                for (int j = 0; j < S; ++j) {
                    for (int k = 0; k < 3; ++k) gc[4 * (size_t)j + k] = per_gdir ? g[3 * j + k] : (k == 0 ? 1.0 : 0.0);
                    gc[4 * (size_t)j + 3] = theta * std::fabs(q[j]) * fmax;
                }
                DevArray<double> gcd, part, out;
                gcd.alloc(gc.size());
                part.alloc(kMaxGrid);
                out.alloc(1);
                DMRI_HIP(hipMemcpyAsync(gcd.p, gc.data(), sizeof(double) * gc.size(), hipMemcpyHostToDevice, s));
                if (w.pdiag.n != (size_t)n * vw) w.pdiag.alloc((size_t)n * vw);
                launch_cdiag_prep(n, c->srow.p, c->scol.p, w.vals.p, vw, S, gcd.p, breal, w.pdiag.p, part.p, out.p, s);
                const double hb = read_scalar(out.p, s);
                a.pc = 1.8 / std::max(breal, hb);
                a.pdiag = w.pdiag.p;
                st->precond_used = 3;
            }
        }
        a.rtol = opts->rtol;
        a.atol = opts->atol;
        a.maxit = opts->maxit;
        gmres_on = opts->krylov == 1;
        cocg_on = opts->krylov == 2;
        // COCG needs A^T = A: every form but the strong periodic one (its direction terms are
        // antisymmetric)
        require(!(cocg_on && strong), DMRI_ERR_ARG,
                "COCG needs a complex-symmetric operator: not available with the strong periodic form");
        if (gmres_on) setup_gmres();
        // COCG's reduced-precision preconditioner input: FP32 copies of the polynomial's values (the
        // real part alone or the whole operator; the complex-diagonal form keeps FP64 inputs) and of
        // r, which the update kernel writes beside r
        pre32 = cocg_on && poly && opts->pre_fp32 != 0 && !a.pdiag;
        a.r32 = nullptr;
        a.pvals32 = nullptr;
        if (pre32) {
            const size_t nv32 = (size_t)c->nnz * (real_pre ? 1 : (per_gdir ? 4 : 2));
            if (w.pvals32.n != nv32) w.pvals32.alloc(nv32);
            launch_to_f32(real_pre ? w.areal.p : w.vals.p, (int64_t)nv32, w.pvals32.p, s);
            if (w.r32.n != (size_t)2 * n * S) w.r32.alloc((size_t)2 * n * S);
            a.r32 = reinterpret_cast<float2 *>(w.r32.p);
            a.pvals32 = w.pvals32.p;
        }
        // A single signal's COCG folds its update of r into the next iteration's preconditioner
        // SpMV (three kernels and two reductions per iteration instead of four and three) when the
        // iteration's working set (the matrix and six vectors) fits in the 256 MB MALL: its
        // launches are then short and the second gathered vector costs little (C1 one signal at a
        // time: 3.70 -> 4.49 signals/s; C3 forward 414 -> 356 ms).  Beyond the MALL the second
        // gather costs more than the update kernel (C4, one signal: forward 22.7 -> 24.4 s), and
        // with several signals about as much (C1, 5 signals: 10.33 -> 10.19 signals/s): the update
        // kernel stays.  r alternates between a.r and a.s (the BiCGStab s buffer, unused here).
        {
            const int vw = per_gdir ? 4 : 2;
            const double ws = (double)c->nnz * (4.0 + 8.0 * vw) + 4.0 * (double)(n + 1) + 96.0 * (double)n * S;
            cocg_fold = cocg_on && S == 1 && !per_gdir && !pre32 && !a.pdiag && a.s != nullptr &&
                        ws <= kMallBytes;
        }
        // With per-signal directions the p-step (x += alpha p, p = z + beta p) runs in the q = A p
        // SpMV's own rows, with q = A z + beta q_prev: three kernels per iteration instead of four.
        cocg_qfold = cocg_on && per_gdir && !pre32 && !a.pdiag;
        a.nparts_upd = geo.vgrid;
        // algorithmic bytes per launch of each iteration kernel (dmri_stats.kern_bytes): the operator
        // SpMVs v = A ph and t = A sh read their matrix entries and 48 B per row and signal (one vector
        // gathered, one read at the own row, one written)
        {
            const int vw = per_gdir ? 4 : 2, npre = real_pre ? 1 : vw;
            const double nS = (double)n * S;
            auto E = [&](int k) { return (double)c->nnz * (4.0 + 8.0 * k) + 4.0 * (double)(n + 1); };
            const double cdx = a.pdiag ? 8.0 * vw * (double)n : 0.0;  // the complex diagonal per row
            kbytes[0] = 64.0 * nS;
            kbytes[1] = poly ? E(npre) + 32.0 * nS + cdx : 0.0;
            kbytes[2] = E(vw) + 48.0 * nS;
            kbytes[3] = poly ? 48.0 * nS : 0.0;
            kbytes[4] = poly ? E(npre) + 32.0 * nS + cdx : 0.0;
            kbytes[5] = E(vw) + 48.0 * nS;
            kbytes[6] = 128.0 * nS;
            if (cocg_on) {
                // [0] x += alpha_prev p, p = z + beta p (x, p, z read, x, p written), [1] z = M r,
                // [2] q = A p (one vector gathered, one written, the own row of the gathered vector
                // counted once), [6] r -= alpha q (r, q read, r written)
                kbytes[0] = 80.0 * nS;
                kbytes[2] = E(vw) + 32.0 * nS;
                kbytes[3] = kbytes[4] = kbytes[5] = 0.0;
                kbytes[6] = 48.0 * nS;
                if (pre32) {
                    // FP32 values and r (8 B per row and signal, gathered once); z written in FP64;
                    // the update also writes r in FP32
                    kbytes[1] = (double)c->nnz * (4.0 + 4.0 * npre) + 4.0 * (double)(n + 1) + 24.0 * nS;
                    kbytes[6] = 56.0 * nS;
                }
                if (cocg_fold) {
                    // r_(k-1) and q gathered, r_k and z written; no update kernel
                    kbytes[1] = E(npre) + 64.0 * nS;
                    kbytes[6] = 0.0;
                }
                if (cocg_qfold) {
                    // z gathered; p, q, x read and written at the own rows; no p-kernel
                    kbytes[0] = 0.0;
                    kbytes[2] = E(vw) + 112.0 * nS;
                }
            }
        }
        st->spmv_bytes = cocg_on ? kbytes[2] : 0.5 * (kbytes[2] + kbytes[5]);
        precond = poly ? st->precond_used : 0;
        prof = opts->profile != 0;
        chunk = 8;
        drop_graphs();
        graphs_on = opts->graphs != 0 && !gmres_on;
        if (graphs_on) {
            if (w.coefcur.n < (size_t)6 * S) w.coefcur.alloc((size_t)6 * S);
            if (fh.m > 0 && w.xk.n != (size_t)n * S) w.xk.alloc((size_t)n * S);
        }
    }

    void setup_gmres() {
        SolverWork &w = *wk;
        const int m = opts->restart > 0 ? opts->restart : 30;
        require(m >= 1 && m <= kGmMaxRestart, DMRI_ERR_ARG, "GMRES restart length must lie in [1, 64]");
        const size_t nS = (size_t)n * S;
        auto need = [](auto &arr, size_t cnt) {
            if (arr.n != cnt) arr.alloc(cnt);
        };
        need(w.gV, (size_t)(m + 1) * nS);
        need(w.gw, nS);
        need(w.gH, (size_t)S * m * (m + 1));
        need(w.gcs, (size_t)S * m);
        need(w.gg, (size_t)S * (m + 1));
        need(w.gy, (size_t)S * m);
        need(w.gpart, (size_t)std::max(geo.vgrid, 1) * S * 2 * kGmChunk);
        need(w.gh, (size_t)S * 2 * (m + 1));
        need(w.ghsum, (size_t)S * 2 * (m + 1));
        need(w.gcc, (size_t)S * m);
        need(w.ginv, (size_t)S);
        need(w.gtot, (size_t)S * kNV);
        need(w.gwn, (size_t)S * kNV);
        need(w.gnk, (size_t)S);
        gm = GmArgs{};
        gm.n = n;
        gm.S = S;
        gm.m = m;
        gm.V = w.gV.p;
        gm.w = w.gw.p;
        gm.part = w.gpart.p;
        gm.h = w.gh.p;
        gm.hsum = w.ghsum.p;
        gm.H = w.gH.p;
        gm.cc = w.gcc.p;
        gm.cs = w.gcs.p;
        gm.g = w.gg.p;
        gm.y = w.gy.p;
        gm.inv = w.ginv.p;
        gm.nk = w.gnk.p;
        gm.st = w.st.p;
        gm.all_done = w.all_done.p;
        gm.rtol = opts->rtol;
        gm.atol = opts->atol;
        gm.maxit = opts->maxit;
    }

    // Restarted GMRES on a.x with the residual already formed (r and the partials of |r|^2, |b|^2):
    // cycles of up to m Arnoldi steps queued at once (steps past a signal's end exit at once), then
    // x += M (V y), the true residual and the next cycle's start; one host check per cycle.
    void gmres(int step) {
        SolverWork &w = *wk;
        const int m = gm.m;
        const size_t nS = (size_t)n * S;
        launch_gm_fin(a.partU, geo.grid, S, kNV, kNV, w.gtot.p, kNV, 0, s);
        launch_gm_start(gm, w.gtot.p, true, s);
        launch_gm_v0(geo, gm, a.r, s);
        for (int cycle = 0;; ++cycle) {
            for (int j = 0; j < m; ++j) {
                SolveArgs av = a;
                double2 *vj = w.gV.p + (size_t)j * nS;
                if (poly) {
                    av.p = vj;
                    launch_spmv_v(geo, av, 0, s, nullptr, nullptr, kOpPre);  // ph = M v_j
                    av.p = a.ph;
                } else {
                    av.p = vj;
                }
                av.v = w.gw.p;
                av.rh = w.gV.p;
                launch_spmv_v(geo, av, 0, s);  // w = A M v_j
                for (int pass = 0; pass < 2; ++pass) {
                    for (int i0 = 0; i0 <= j; i0 += kGmChunk) {
                        const int cnt = std::min(kGmChunk, j + 1 - i0);
                        launch_gm_dots(geo, gm, i0, cnt, s);
                        launch_gm_fin(w.gpart.p, geo.vgrid, S, 2 * kGmChunk, 2 * cnt, w.gh.p, 2 * (m + 1), 2 * i0, s);
                    }
                    launch_gm_orth(geo, gm, j, pass == 1, s);
                }
                launch_gm_fin(w.gpart.p, geo.vgrid, S, kNV, kNV, w.gwn.p, kNV, 0, s);
                launch_gm_hess(gm, j, w.gwn.p, s);
                launch_gm_scale(geo, gm, j, s);
            }
            // x += M (V y) for every signal that took steps this cycle
            launch_reset_state_flags(w.st.p, S, w.all_done.p, s);
            launch_gm_combine(geo, gm, a.p, s);
            if (poly) {
                SolveArgs av = a;
                av.p = a.p;
                av.ph = a.sh;
                launch_spmv_v(geo, av, 0, s, nullptr, nullptr, kOpPre);
                launch_gm_axpy((int64_t)nS, a.sh, a.x, s);
            } else {
                launch_gm_axpy((int64_t)nS, a.p, a.x, s);
            }
            launch_rhs_residual(geo, a, 1, nullptr, s);
            launch_gm_fin(a.partU, geo.grid, S, kNV, kNV, w.gtot.p, kNV, 0, s);
            launch_gm_start(gm, w.gtot.p, false, s);
            launch_gm_v0(geo, gm, a.r, s);
            launch_publish_state(w.st.p, S, w.all_done.p, w.d_h_st, w.d_h_flag, s);
            DMRI_HIP(hipStreamSynchronize(s));
            if (w.h_flag[0] == 0) continue;  // signals left that have not converged: another cycle
            for (int j = 0; j < S; ++j)
                if (!w.h_st[j].conv)
                    throw Error(DMRI_ERR_NOCONV, "GMRES did not converge in " + std::to_string(opts->maxit) +
                                                     " iterations at step " + std::to_string(step) + " (signal " +
                                                     std::to_string(j) + ")");
            break;
        }
        int step_max = 0;
        for (int j = 0; j < S; ++j) {
            const KrylovState &h = w.h_st[j];
            st->iters_total += h.iters;
            step_max = std::max(step_max, h.iters);
            const double rel = h.bnorm2 > 0 ? std::sqrt(h.rnorm2 / h.bnorm2) : std::sqrt(h.rnorm2);
            st->relres_max = std::max(st->relres_max, rel);
        }
        st->iters_max = std::max(st->iters_max, step_max);
    }

    void solve_krylov(int step) {
        if (gmres_on) gmres(step);
        else krylov(step);
    }

    // Residual-minimising guess over the last fh.m states (see proj_capacity); the quadratic
    // extrapolation stays in use when the history does not fit.
    void setup_projection(const double *g) {
        SolverWork &w = *wk;
        const size_t nS = (size_t)n * S;
        size_t held = 0;  // history buffers this context already holds count as free
        for (int j = 0; j < kMaxHist; ++j) held += (w.yh[j].n + w.hp[j].n + w.hq[j].n) * sizeof(double2);
        const int m = proj_capacity(nS, held);
        fh.reset(0, nullptr, nullptr, nullptr);
        if (m == 0) return;
        for (int j = 0; j < kMaxHist; ++j) {
            if (j < m) {
                if (w.yh[j].n != nS) w.yh[j].alloc(nS);
                if (w.hp[j].n != nS) w.hp[j].alloc(nS);
                if (w.hq[j].n != nS) w.hq[j].alloc(nS);
            } else {
                w.yh[j].release();
                w.hp[j].release();
                w.hq[j].release();
            }
        }
        ensure_proj_scratch(m, g);
        fh.reset(m, w.yh, w.hp, w.hq);
        fh.ys.push_back(xc);  // u^0: its products are formed at step 0
        extrap = false;
        guess_order = 0;
    }

    void ensure_proj_scratch(int m, const double *g) {
        SolverWork &w = *wk;
        const size_t npart = (size_t)proj_grid(geo) * S * proj_nv(m);
        if (w.ppart.n < npart) w.ppart.alloc(npart);
        if (w.ptot.n < (size_t)S * proj_nv(m)) w.ptot.alloc((size_t)S * proj_nv(m));
        if (w.pgam.n < (size_t)S * 2 * kMaxHist) w.pgam.alloc((size_t)S * 2 * kMaxHist);
        if (per_gdir) {
            if (w.gsig.n < (size_t)S * 3) w.gsig.alloc((size_t)S * 3);
            DMRI_HIP(hipMemcpyAsync(w.gsig.p, g, sizeof(double) * S * 3, hipMemcpyHostToDevice, s));
            a.gsig = w.gsig.p;
        }
    }

    // Guess of the current step (coefficients set by begin_step, b in work.b) from history h
    // into a free buffer of h, which it returns.
    double2 *projected_guess(ProjHistory &h, double2 *into = nullptr) {
        SolverWork &w = *wk;
        double2 *y0 = into ? into : h.take_free();
        ProjArgs pa{};
        pa.n = n;
        pa.S = S;
        pa.nh = h.nh();
        pa.per_gdir = per_gdir;
        for (int j = 0; j < pa.nh; ++j) {
            pa.Y[j] = h.ys[j];
            pa.P[j] = h.pq[j].first;
            pa.Q[j] = h.pq[j].second;
        }
        pa.b = a.b;
        pa.coefA = a.coefA;
        pa.gsig = per_gdir ? w.gsig.p : nullptr;
        pa.part = w.ppart.p;
        pa.tot = w.ptot.p;
        pa.gam = w.pgam.p;
        pa.y0 = y0;
        launch_proj_guess(geo, pa, h.m, s);
        return y0;
    }

    // Step k with the projection guess: b and the history products of u^n (= fh.ys[0]) in one
    // pass, the guess into a free state buffer, r = b - A x0, Krylov; the solution becomes ys[0].
    void projection_step(int k) {
        const auto slot = fh.take_slot();
        a.x = xc;
        a.hp = slot.first;
        a.hq = slot.second;
        launch_rhs_residual(geo, a, 4, nullptr, s);
        a.hp = a.hq = nullptr;
        fh.pq.push_front(slot);
        double2 *y0;
        if (graphs_on) {
            // the Krylov iterate stays in one buffer (the captured kernels' argument), then moves to
            // the history's free buffer
            y0 = fh.take_free();
            a.x = projected_guess(fh, wk->xk.p);
            launch_rhs_residual(geo, a, 1, nullptr, s);
            solve_krylov(k);
            DMRI_HIP(hipMemcpyAsync(y0, wk->xk.p, sizeof(double2) * (size_t)n * S, hipMemcpyDeviceToDevice, s));
        } else {
            y0 = projected_guess(fh);
            a.x = y0;
            launch_rhs_residual(geo, a, 1, nullptr, s);
            solve_krylov(k);
        }
        xc = y0;
        a.x = xc;
        fh.ys.push_front(y0);
        if ((int)fh.ys.size() > fh.m) {
            fh.yfree.push_back(fh.ys.back());
            fh.ys.pop_back();
        }
    }

    // State b steps back from the newest (0 = xc), for checkpoints; the newest when there is none
    const double2 *state_back(int b) const {
        if (fh.m > 0) return b < (int)fh.ys.size() ? fh.ys[b] : xc;
        return b == 0 ? xc : (b == 1 ? xp : xq);
    }

    // Restart the projection history from count stored states (newest first): the older ones get
    // their history products, the newest becomes xc (its products come with the next step).
    void reseed_projection(const double2 *const *states, int count) {
        SolverWork &w = *wk;
        fh.reset(fh.m, w.yh, w.hp, w.hq);
        fh.yfree.push_back(w.x.p);
        const size_t vb = sizeof(double2) * (size_t)n * S;
        for (int i = count - 1; i >= 0; --i) {
            double2 *y = fh.take_free();
            DMRI_HIP(hipMemcpyAsync(y, states[i], vb, hipMemcpyDeviceToDevice, s));
            if (i > 0) {
                const auto slot = fh.take_slot();
                a.x = y;
                a.hp = slot.first;
                a.hq = slot.second;
                launch_rhs_residual(geo, a, 4, nullptr, s);  // products; b is formed again by the step
                fh.push(y, slot);
            } else {
                fh.ys.push_front(y);
            }
        }
        xc = fh.ys[0];
        a.x = xc;
        a.hp = a.hq = nullptr;
    }

    // Adjoint solve of step k: A_k y = rhs, from the projection over h (from *start when h is
    // empty); returns the buffer holding the solution.
    double2 *solve_projected(int k, const double2 *rhs, ProjHistory &h, double2 *start) {
        begin_step(k);
        a.b = const_cast<double2 *>(rhs);  // read in place (the residual passes only read b)
        double2 *y = h.nh() > 0 ? projected_guess(h) : start;
        a.x = y;
        launch_rhs_residual(geo, a, 1, nullptr, s);
        solve_krylov(k);
        a.x = xc;
        a.b = wk->b.p;
        return y;
    }

    // out = B_k y, with the history products of y (y becomes the newest state of h)
    void apply_B_record(int k, double2 *y, double2 *out, ProjHistory &h) {
        begin_step(k);
        const auto slot = h.take_slot();
        a.x = y;
        a.hp = slot.first;
        a.hq = slot.second;
        a.b = out;  // the right-hand-side pass writes B_k y straight into out
        launch_rhs_residual(geo, a, 4, nullptr, s);
        h.push(y, slot);
        a.x = xc;
        a.b = wk->b.p;
        a.hp = a.hq = nullptr;
    }

    void begin_step(int step) {
        SolverWork &w = *wk;
        if (strong)
            launch_sbc_step(c->nnz, per_gdir ? 4 : 2, w.afix.p, w.bfix.p, w.mgg.p, s2A[step], s2B[step], w.vals.p,
                            w.bre.p, real_pre ? w.areal.p : nullptr, s);
        const double *cA = w.coef.p + (size_t)step * S * 3;
        const double *cB = w.coef.p + (size_t)std::max(N, 1) * S * 3 + (size_t)step * S * 3;
        if (graphs_on) {  // a fixed slot, so that the captured kernels' arguments repeat from step to step
            DMRI_HIP(hipMemcpyAsync(w.coefcur.p, cA, sizeof(double) * 3 * S, hipMemcpyDeviceToDevice, s));
            DMRI_HIP(hipMemcpyAsync(w.coefcur.p + 3 * S, cB, sizeof(double) * 3 * S, hipMemcpyDeviceToDevice, s));
            a.coefA = w.coefcur.p;
            a.coefB = w.coefcur.p + 3 * S;
        } else {
            a.coefA = cA;
            a.coefB = cB;
        }
        launch_reset_state(w.st.p, S, w.all_done.p, s);
    }

    // The seven kernels of BiCGStab iteration it (e0(j), e1(j): the events of kernel j, or null)
    template <class E0, class E1>
    void launch_iteration(int it, E0 e0, E1 e1) {
        if (cocg_fold) {
            launch_cocg_pre_update(geo, a, it, s, e0(1), e1(1));   // r -= alpha q, z = M r, r^T z, |r|^2
            launch_cocg_p(geo, a, it, s, e0(0), e1(0), true);     // p = z + beta p
            launch_cocg_spmv(geo, a, it, s, e0(2), e1(2), false);  // q = A p (into a.v) with p^T q
            return;
        }
        if (cocg_qfold) {
            SolveArgs ar = a;  // z = M r (into a.ph) with r^T z
            ar.p = a.r;
            launch_cocg_spmv(geo, ar, it, s, e0(1), e1(1), true);
            launch_cocg_qfold(geo, a, it, s, e0(2), e1(2));  // p-step and q = A p, p^T q
            launch_cocg_update(geo, a, it, s, e0(6), e1(6));
            return;
        }
        if (cocg_on) {
            SolveArgs ar = a;  // z = M r (into a.ph) with r^T z
            ar.p = a.r;
            launch_cocg_spmv(geo, ar, it, s, e0(1), e1(1), true);
            launch_cocg_p(geo, a, it, s, e0(0), e1(0));      // p = z + beta p
            launch_cocg_spmv(geo, a, it, s, e0(2), e1(2), false);  // q = A p (into a.v) with p^T q
            launch_cocg_update(geo, a, it, s, e0(6), e1(6));
            return;
        }
        launch_bicg_p(geo, a, it, s, e0(0), e1(0));
        if (poly) {
            // ph = 2p - Ap, v = A ph; sh = 2s - As, t = A sh (right preconditioning)
            launch_spmv_v(geo, a, it, s, e0(1), e1(1), kOpPre);
            SolveArgs av = a;
            av.p = a.ph;
            launch_spmv_v(geo, av, it, s, e0(2), e1(2));
            launch_spass(geo, a, it, s, e0(3), e1(3));
            launch_spmv_t(geo, a, it, s, e0(4), e1(4), kOpPre);
            launch_spmv_t(geo, a, it, s, e0(5), e1(5), kOpTP);
        } else {
            launch_spmv_v(geo, a, it, s, e0(2), e1(2));
            launch_spmv_t(geo, a, it, s, e0(5), e1(5));
        }
        launch_bicg_update(geo, a, it, s, e0(6), e1(6));
    }

    // Iterations it0 .. it0+count-1 from captured graphs of 2^k iterations each (at most 64 at the
    // largest), so that the few run lengths a solve needs recur from step to step.  A graph is keyed
    // by the bytes of the kernel arguments, its first iteration and its length: a replay launches
    // exactly what the plain launches would.  A capture the runtime refuses turns graphs off and
    // launches the iterations one by one.
    void launch_chain(int it0, int count) {
        auto none = [](int) -> hipEvent_t { return nullptr; };
        while (count > 0) {
            int len = 64;
            while (len > count) len >>= 1;
            if (!graphs_on) {
                for (int i = 0; i < len; ++i) launch_iteration(it0 + i, none, none);
            } else {
                int hit = -1;
                for (size_t q = 0; q < graphs.size() && hit < 0; ++q)
                    if (graphs[q].it0 == it0 && graphs[q].count == len &&
                        std::memcmp(&graphs[q].args, &a, sizeof(SolveArgs)) == 0)
                        hit = (int)q;
                if (hit < 0) hit = capture_chain(it0, len);
                if (hit >= 0) {
                    DMRI_HIP(hipGraphLaunch(graphs[hit].e, s));
                    if (st) st->graph_launches += 1;
                } else {
                    for (int i = 0; i < len; ++i) launch_iteration(it0 + i, none, none);
                }
            }
            it0 += len;
            count -= len;
        }
    }
    // Captures iterations it0 .. it0+len-1 into the cache; returns its index, or -1 (graphs off)
    int capture_chain(int it0, int len) {
        auto none = [](int) -> hipEvent_t { return nullptr; };
        if (graphs.size() >= 256) drop_graphs();
        ChainGraph cg{};
        cg.args = a;
        cg.it0 = it0;
        cg.count = len;
        bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            try {
                for (int i = 0; i < len; ++i) launch_iteration(it0 + i, none, none);
            } catch (...) {
                ok = false;
            }
            hipGraph_t g = nullptr;
            const bool ended = hipStreamEndCapture(s, &g) == hipSuccess;
            ok = ok && ended && g != nullptr && hipGraphInstantiate(&cg.e, g, nullptr, nullptr, 0) == hipSuccess;
            cg.g = g;
            if (!ok) {
                if (cg.e) (void)hipGraphExecDestroy(cg.e);
                if (g) (void)hipGraphDestroy(g);
            }
        }
        (void)hipGetLastError();
        if (!ok) {
            graphs_on = false;
            return -1;
        }
        graphs.push_back(cg);
        return (int)graphs.size() - 1;
    }

    // BiCGStab iterations on a.x with the residual already formed; restarts on breakdown.
    // r in FP32 for the reduced-precision preconditioner (before a solve's first iteration; the
    // update kernel keeps it current after that)
    void round_r32() {
        if (a.r32) launch_to_f32(reinterpret_cast<const double *>(a.r), 2 * (int64_t)n * S,
                                 reinterpret_cast<float *>(a.r32), s);
    }

    void krylov(int step) {
        SolverWork &w = *wk;
        int it = 0, it_before = 0, restarts = 0;
        round_r32();
        constexpr int kK = 7;  // timed kernels per iteration (dmri_stats.kern_*)
        for (;;) {
            const int nl = chunk;
            int ntimed = 0;
            int k = 0;
            for (; k < nl; ++k, ++it) {
                // every kernel of the iterations it < profile carries start/end events that take the
                // timestamps of its own dispatch (no marker packets between kernels)
                const bool timed = prof && it < opts->profile && ntimed < kEventPool / kK;
                if (!timed && graphs_on) break;
                auto e0 = [&](int j) { return timed ? c->ev0[kK * ntimed + j] : nullptr; };
                auto e1 = [&](int j) { return timed ? c->ev1[kK * ntimed + j] : nullptr; };
                launch_iteration(it, e0, e1);
                if (timed) timed_first[ntimed++] = it == 0;
            }
            if (k < nl) {  // the untimed rest of the run as one graph
                launch_chain(it, nl - k);
                it += nl - k;
            }
            launch_publish_state(w.st.p, S, w.all_done.p, w.d_h_st, w.d_h_flag, s);
            DMRI_HIP(hipStreamSynchronize(s));
            const bool all = w.h_flag[0] != 0;
            int it_last = it;
            if (all) {
                it_last = 0;
                for (int j = 0; j < S; ++j) it_last = std::max(it_last, w.h_st[j].iters);
            }
            if (prof) {
                const int real = std::max(0, std::min(ntimed, it_last - it_before));
                for (int k = 0; k < real; ++k)
                    for (int j = 0; j < kK; ++j) {
                        if (!poly && (j == 1 || j == 3 || j == 4)) continue;
                        if (cocg_on && (j == 3 || j == 4 || j == 5)) continue;
                        if (cocg_fold && j == 6) continue;
                        if (cocg_qfold && j == 0) continue;
                        float ms = 0;
                        DMRI_HIP(hipEventElapsedTime(&ms, c->ev0[kK * k + j], c->ev1[kK * k + j]));
                        st->kern_ms[j] += ms;
                        st->kern_launches[j] += 1;
                        // the first iteration's p = r (COCG: p = z) reads one vector and writes one;
                        // with the folded p-step, the q-kernel's first iteration reads no p, q or x
                        const double first_less = cocg_qfold ? 64.0 : (cocg_on ? 48.0 : 32.0);
                        const int first_j = cocg_qfold ? 2 : 0;
                        st->kern_bytes[j] +=
                            kbytes[j] - (j == first_j && timed_first[k] ? first_less * (double)n * S : 0.0);
                        if (j == 2 || j == 5) {
                            st->spmv_ms += ms;
                            st->spmv_launches += 1;
                        }
                    }
            }
            it_before = it;
            if (!all) {
                chunk = std::max(2, chunk / 4);
                continue;
            }
            // every signal stopped: converged, broke down, or hit maxit
            bool any_restart = false;
            std::vector<int> mask(S, 0);
            for (int j = 0; j < S; ++j) {
                KrylovState &h = w.h_st[j];
                if (h.conv) continue;
                if (h.breakdown) {
                    mask[j] = 1;
                    any_restart = true;
                    h.done = 0;
                    h.stop = 0;
                    h.breakdown = 0;
                    h.it0 = it;
                    continue;
                }
                throw Error(DMRI_ERR_NOCONV, "BiCGStab did not converge in " + std::to_string(opts->maxit) +
                                                 " iterations at step " + std::to_string(step) + " (signal " +
                                                 std::to_string(j) + ", relres " +
                                                 std::to_string(std::sqrt(h.rnorm2 / std::max(h.bnorm2, 1e-300))) +
                                                 ")");
            }
            if (!any_restart) break;
            ++restarts;
            require(restarts <= 20, DMRI_ERR_NOCONV, "BiCGStab: repeated breakdown at step " + std::to_string(step));
            DMRI_HIP(hipMemcpyAsync(w.st.p, w.h_st, sizeof(KrylovState) * S, hipMemcpyHostToDevice, s));
            DMRI_HIP(hipMemcpyAsync(w.mask.p, mask.data(), sizeof(int) * S, hipMemcpyHostToDevice, s));
            DMRI_HIP(hipMemsetAsync(w.all_done.p, 0, sizeof(int), s));
            launch_rhs_residual(geo, a, 1, w.mask.p, s);
            round_r32();
            DMRI_HIP(hipStreamSynchronize(s));
        }
        int step_max = 0;
        for (int j = 0; j < S; ++j) {
            const KrylovState &h = w.h_st[j];
            st->iters_total += h.iters;
            step_max = std::max(step_max, h.iters);
            const double rel = h.bnorm2 > 0 ? std::sqrt(h.rnorm2 / h.bnorm2) : std::sqrt(h.rnorm2);
            st->relres_max = std::max(st->relres_max, rel);
        }
        st->iters_max = std::max(st->iters_max, step_max);
        st->restarts += restarts;
        // the next step converges in about as many iterations; its convergence shows in the
        // head of the iteration after the last one, so that iteration is queued too
        chunk = std::max(1, step_max) + 1;
    }

    // One theta step of the forward problem on xc (xp, xq hold the two previous states).
    void step(int k) {
        begin_step(k);
        a.x = xc;
        a.xprev = xp;
        a.hp = a.hq = nullptr;
        if (c->periodic) {
            SolverWork &w = *wk;
            launch_periodic_step(c->n_per, c->n_prow, S, c->sper_dof.p, c->sper_src.p, c->per_w.p, c->per_ell.p,
                                 c->sprow.p, c->srow.p, c->scol.p, c->spval.p, w.sc.p, xc,
                                 w.coefP.p + (size_t)k * S * 3, omt, w.ybc.p, w.bper.p, s);
        }
        if (fh.m > 0) {
            projection_step(k);
            return;
        }
        const int order = std::min(guess_order, k);
        if (order > 0) {
            // b from u^n; xq <- the extrapolated guess; iterate on it, then rotate the history
            a.xpp = xq;
            a.gc0 = order == 1 ? 2.0 : 3.0;
            a.gc1 = order == 1 ? -1.0 : -3.0;
            a.gc2 = order == 1 ? 0.0 : 1.0;
            launch_rhs_residual(geo, a, 2, nullptr, s);
            double2 *u_n = xc, *u_n1 = xp;
            xc = xq;
            xp = u_n;
            xq = u_n1;
            a.x = xc;
            a.xprev = xp;
            launch_rhs_residual(geo, a, 1, nullptr, s);
        } else {
            launch_rhs_residual(geo, a, 0, nullptr, s);
            if (extrap)
                DMRI_HIP(hipMemcpyAsync(xp, xc, sizeof(double2) * (size_t)n * S, hipMemcpyDeviceToDevice, s));
        }
        solve_krylov(k);
    }

    // Solve A_k y = rhs (complex-symmetric A_k) starting from y; one signal.
    void solve_given(int k, const double2 *rhs, double2 *y) {
        begin_step(k);
        a.b = const_cast<double2 *>(rhs);
        a.x = y;
        launch_rhs_residual(geo, a, 1, nullptr, s);
        solve_krylov(k);
        a.x = xc;
        a.b = wk->b.p;
    }

    // out = B_k y
    void apply_B(int k, double2 *y, double2 *out) {
        begin_step(k);
        a.x = y;
        a.b = out;
        launch_rhs_residual(geo, a, 3, nullptr, s);
        a.x = xc;
        a.b = wk->b.p;
    }

    // Signals of the current state u^n (the state buffers are left as they are, so stepping may go on)
    void signal(double *sig) {
        SolverWork &w = *wk;
        launch_signal(n, S, c->sw.p, w.sc.p, xc, w.spart.p, w.sig.p, s);
        std::vector<double> hs((size_t)S * kNV);
        DMRI_HIP(hipMemcpyAsync(hs.data(), w.sig.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s));
        DMRI_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < S; ++j) {
            sig[2 * j] = hs[(size_t)j * kNV] / c->volume;
            sig[2 * j + 1] = hs[(size_t)j * kNV + 1] / c->volume;
        }
    }

    // Move the final state into the workspace's x buffer (dmri_get_solution reads it there).
    void park() {
        SolverWork &w = *wk;
        if (xc != w.x.p)
            DMRI_HIP(hipMemcpyAsync(w.x.p, xc, sizeof(double2) * (size_t)n * S, hipMemcpyDeviceToDevice, s));
        xc = w.x.p;
    }
};

// Signals j0 .. j0+S-1 of nsig; snapshots (seq->snap_every > 0) go to their slots of seq->snap_out.
void solve_group(dmri_ctx *c, int S, int j0, int nsig, const double *q, const double *g, bool per_gdir,
                 const dmri_seq *seq, const dmri_solver_opts *opts, double *sig, dmri_stats &st) {
    Stepper run;
    run.setup(c, S, q, g, per_gdir, seq, opts, &st);
    const int every = seq->snap_every;
    const int64_t n = c->ndof;
    DevArray<double2> snap;
    if (every > 0) snap.alloc((size_t)S * n);
    for (int k = 0; k < run.N; ++k) {
        run.step(k);
        if (every > 0 && k % every == 0) {
            for (int s = 0; s < S; ++s)
                launch_unscale_perm(n, S, s, c->perm.p, c->work.sc.p, run.xc, snap.p + (size_t)s * n, c->stream);
            double *dst = seq->snap_out + ((size_t)(k / every) * nsig + j0) * 2 * (size_t)n;
            DMRI_HIP(hipMemcpyAsync(dst, snap.p, sizeof(double2) * (size_t)S * n, hipMemcpyDeviceToHost, c->stream));
            DMRI_HIP(hipStreamSynchronize(c->stream));
        }
    }
    run.signal(sig);
    run.park();
    c->last_S = S;
}

// Discrete adjoint of the theta scheme for one signal (see the C ABI comment of dmri_adjoint).
void adjoint_impl(dmri_ctx *c, const dmri_seq *seq, const dmri_solver_opts *opts, const dmri_adjoint_opts *aopts,
                  double *sig, dmri_grads *grads, dmri_stats *stats) {
    hipStream_t s = c->stream;
    const double gn = std::sqrt(seq->gdir[0] * seq->gdir[0] + seq->gdir[1] * seq->gdir[1] + seq->gdir[2] * seq->gdir[2]);
    require(gn > 1e-10 && std::isfinite(gn), DMRI_ERR_ARG, "|g|=0: gradient direction must be non-zero");
    require(std::isfinite(seq->q), DMRI_ERR_ARG, "q is not finite");
    const double g[3] = {seq->gdir[0] / gn, seq->gdir[1] / gn, seq->gdir[2] / gn};
    dmri_stats st{};
    Stepper run;
    const auto t0 = std::chrono::steady_clock::now();
    const int N = seq->nsteps;
    const int64_t n = c->ndof;
    // stride 0: keep every state when the N+1 states fit in a quarter of the free HBM (no step is
    // recomputed), else checkpoints every ceil(sqrt(N)) steps
    int K = aopts ? aopts->checkpoint_stride : 0;
    if (K <= 0) {
        size_t fr = 0, tot = 0;
        DMRI_HIP(hipMemGetInfo(&fr, &tot));
        const double all = (double)(N + 1) * (double)n * sizeof(double2);
        K = all <= 0.25 * (double)fr ? std::max(N, 1) : (int)std::ceil(std::sqrt((double)std::max(N, 1)));
    }
    K = std::max(1, std::min(K, std::max(N, 1)));
    const bool keep_all = K >= N;
    // A checkpoint holds the state the next step restarts from: with the extrapolated guesses
    // (initial_guess 1-3) three states, with the projection guess (the default) the whole history
    // of up to 12 states (their Are / Jg products are formed again when a segment is recomputed).
    // Either way a recomputed segment repeats the forward sweep bit for bit, so it costs what the
    // forward sweep cost.  When the full histories of every checkpoint would take more than a quarter
    // of the free device memory, a checkpoint keeps the newest three states of the history and the
    // recomputed states agree with the first sweep's to the solver tolerance.  The backward sweep
    // keeps its own projection history of adjoint states.
    const bool proj = opts->initial_guess == 0;
    run.setup(c, 1, &seq->q, g, false, seq, opts, &st);
    const int nck = (N + K - 1) / K;
    // Segment j covers steps kbeg(j) .. kend(j) - 1.  The first segment takes the remainder
    // (1 .. K steps) and every later one K steps, so the last segment, whose states the forward
    // sweep keeps and the backward sweep does not recompute, is a full one.
    const int k_first = N - (nck - 1) * K;
    auto kbeg = [&](int j) { return j == 0 ? 0 : k_first + (j - 1) * K; };
    auto kend = [&](int j) { return j == nck - 1 ? N : kbeg(j + 1); };
    const int klast = keep_all ? 0 : kbeg(nck - 1);  // first step of the last segment
    const size_t vb = sizeof(double2) * (size_t)n;
    int ckm = 3;  // states per checkpoint
    if (!keep_all && run.fh.m > 0) {
        size_t fr = 0, tot = 0;
        DMRI_HIP(hipMemGetInfo(&fr, &tot));
        if ((double)nck * run.fh.m * (double)vb <= 0.25 * (double)fr) ckm = run.fh.m;
    }
    // checkpoints at the segments' first steps: the history newest first (x^k, x^(k-1), ...); a
    // segment holds x^k0 .. x^k1
    DevArray<double2> ck, seg, mu, lam, lam2;
    std::vector<int> ck_count(std::max(nck, 1), 0);
    ck.alloc((size_t)std::max(nck, 1) * ckm * n);
    seg.alloc((size_t)(std::min(K, std::max(N, 1)) + 1) * n);
    mu.alloc(n);
    lam.alloc(n);
    lam2.alloc(n);
    auto segp = [&](int j) { return seg.p + (size_t)j * n; };
    auto ckp = [&](int j, int b) { return ck.p + ((size_t)j * ckm + b) * n; };
    // forward sweep
    for (int k = 0, jn = 0; k < N; ++k) {
        if (!keep_all && jn < nck && k == kbeg(jn)) {
            const int j = jn++;
            if (run.fh.m > 0) {
                ck_count[j] = std::min(ckm, (int)run.fh.ys.size());
                for (int b = 0; b < ck_count[j]; ++b)
                    DMRI_HIP(hipMemcpyAsync(ckp(j, b), run.state_back(b), vb, hipMemcpyDeviceToDevice, s));
            } else {
                ck_count[j] = 3;
                DMRI_HIP(hipMemcpyAsync(ckp(j, 0), run.xc, vb, hipMemcpyDeviceToDevice, s));
                DMRI_HIP(hipMemcpyAsync(ckp(j, 2), run.state_back(k > 1 ? 2 : 0), vb, hipMemcpyDeviceToDevice, s));
                DMRI_HIP(hipMemcpyAsync(ckp(j, 1), run.state_back(k > 0 ? 1 : 0), vb, hipMemcpyDeviceToDevice, s));
            }
        }
        // the last segment's states stay in the segment buffer, so the backward sweep starts
        // without recomputing it
        if (k == klast) DMRI_HIP(hipMemcpyAsync(segp(0), run.xc, vb, hipMemcpyDeviceToDevice, s));
        run.step(k);
        if (k >= klast) DMRI_HIP(hipMemcpyAsync(segp(k - klast + 1), run.xc, vb, hipMemcpyDeviceToDevice, s));
    }
    run.signal(sig);
    const auto t1 = std::chrono::steady_clock::now();
    const int64_t it_fwd = st.iters_total;
    int64_t it_rec = 0;
    // backward sweep
    SolverWork &w = c->work;
    const double theta = seq->theta;
    // What each parameter multiplies in the operator.  A scalar coefficient multiplies a matrix of
    // unit coefficient, assembled here so that the gradient does not depend on the coefficient's
    // value (D = 0 and kappa = 0 included); a field's multiplier multiplies the assembled matrix.
    DevArray<double> unitK, unitI, junk, ones;
    const double *kdiff = c->val[DMRI_MAT_K].p, *idiff = c->val[DMRI_MAT_I].p;
    const double *mdiff = c->has_iT2cell ? c->val[DMRI_MAT_R].p : c->val[DMRI_MAT_M].p;
    if (!c->has_Dcell) {
        const double one[2] = {1.0, 1.0};
        ones.upload(one, 2, s);
        unitK.alloc(c->nnz);
        junk.alloc(c->nnz);
        launch_assemble_rows(c->ndof, c->dof_vert.p, c->dof_comp.p, c->rowptr.p, c->col.p, c->tets.p,
                             c->ncomp == 2 ? c->phase.p : nullptr, c->xyz.p, c->vptr.p, c->vlist.p, ones.p, nullptr,
                             c->iT2c.p, nullptr, junk.p, unitK.p, junk.p, junk.p, junk.p, junk.p, s);
        kdiff = unitK.p;
    }
    if (c->ncomp == 2 && c->n_iface > 0 && !c->has_kappa_f) {
        DevArray<int> miss;
        miss.alloc(1);
        miss.zero(s);
        unitI.alloc(c->nnz);
        assemble_interface(c, 1.0, nullptr, unitI.p, miss.p);
        idiff = unitI.p;
    }
    // the contraction's four values per entry, scaled and in solver order, read contiguously each step
    DevArray<double4> adjv;
    adjv.alloc(c->nnz);
    launch_adj_values(n, c->srow.p, c->scol.p, c->emap.p, w.sc.p, mdiff, kdiff, idiff, w.vals.p, adjv.p, s);
    launch_adj_init_lambda(n, c->sw.p, w.sc.p, 1.0 / c->volume, lam.p, s);
    DMRI_HIP(hipMemsetAsync(mu.p, 0, vb, s));
    // per-step contraction partials stay on the device and are summed on the host at the end
    DevArray<double> parts;
    parts.alloc((size_t)std::max(N, 1) * kMaxGrid * 14);
    std::vector<int> nparts_of(std::max(N, 1), 0);
    double gK[2] = {0, 0}, gM[2] = {0, 0}, gI = 0, gq = 0;
    std::vector<double> dfa(std::max(N, 1), 0.0), dfb(std::max(N, 1), 0.0);
    int recomputed = 0;
    ProjHistory bh;
    DevArray<double2> byh[kMaxHist], bhp[kMaxHist], bhq[kMaxHist];
    if (proj) {
        const int m = proj_capacity((size_t)n, 0);
        if (m > 0) {
            for (int j = 0; j < m; ++j) {
                byh[j].alloc(n);
                bhp[j].alloc(n);
                bhq[j].alloc(n);
            }
            run.ensure_proj_scratch(m, g);
            bh.reset(m, byh, bhp, bhq);
        }
    }
    double2 *mu_cur = mu.p;
    // Segment j's states x^k0 .. x^k1 into dst, recomputed by stepper r from checkpoint j; returns
    // the Krylov iterations taken.
    auto recompute = [&](Stepper &r, int j, double2 *dst) -> int64_t {
        const int k0 = kbeg(j), k1 = kend(j);
        const hipStream_t rs = r.s;
        if (r.fh.m > 0) {
            const double2 *states[kMaxHist];
            for (int b = 0; b < ck_count[j]; ++b) states[b] = ckp(j, b);
            r.reseed_projection(states, ck_count[j]);
        } else {
            DMRI_HIP(hipMemcpyAsync(r.xc, ckp(j, 0), vb, hipMemcpyDeviceToDevice, rs));
            DMRI_HIP(hipMemcpyAsync(r.xp, ckp(j, 1), vb, hipMemcpyDeviceToDevice, rs));
            DMRI_HIP(hipMemcpyAsync(r.xq, ckp(j, 2), vb, hipMemcpyDeviceToDevice, rs));
        }
        DMRI_HIP(hipMemcpyAsync(dst, r.xc, vb, hipMemcpyDeviceToDevice, rs));
        const int64_t before = r.st->iters_total;
        for (int k = k0; k < k1; ++k) {
            r.step(k);
            DMRI_HIP(hipMemcpyAsync(dst + (size_t)(k - k0 + 1) * n, r.xc, vb, hipMemcpyDeviceToDevice, rs));
        }
        DMRI_HIP(hipStreamSynchronize(rs));
        return r.st->iters_total - before;
    };
    for (int j = nck - 1; j >= 0; --j) {
        const int k0 = kbeg(j), k1 = kend(j);
        if (!keep_all && j != nck - 1) {
            it_rec += recompute(run, j, seg.p);
            recomputed += k1 - k0;
        }
        for (int k = k1 - 1; k >= k0; --k) {
            const int base = keep_all ? 0 : k0;
            // A_k mu = lambda^(k+1): from the projection over the later adjoint states, or warm
            // started from mu_(k+1)
            if (bh.m > 0)
                mu_cur = run.solve_projected(k, lam.p, bh, mu_cur);
            else
                run.solve_given(k, lam.p, mu_cur);
            nparts_of[k] = launch_adj_contract_rows(n, c->srow.p, c->scol.p, c->ncomp == 2 ? c->scomp.p : nullptr,
                                                    adjv.p, mu_cur, segp(k - base), segp(k - base + 1), theta,
                                                    parts.p + (size_t)k * kMaxGrid * 14, s);
            if (bh.m > 0)
                run.apply_B_record(k, mu_cur, lam2.p, bh);  // lambda^k = B_k mu_k, mu_k into the history
            else
                run.apply_B(k, mu_cur, lam2.p);  // lambda^k = B_k mu_k
            std::swap(lam.p, lam2.p);
        }
    }
    std::vector<double> hp((size_t)std::max(N, 1) * kMaxGrid * 14);
    DMRI_HIP(hipMemcpyAsync(hp.data(), parts.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    for (int j = nck - 1; j >= 0; --j) {
        const int k0 = kbeg(j), k1 = kend(j);
        for (int k = k1 - 1; k >= k0; --k) {
            const double *pk = hp.data() + (size_t)k * kMaxGrid * 14;
            double t[14] = {0};
            for (int b = 0; b < nparts_of[k]; ++b)
                for (int q = 0; q < 14; ++q) t[q] += pk[(size_t)b * 14 + q];
            // dS/dp = Re sum_k mu_k^T (dB_k x^k - dA_k x^(k+1))
            gK[0] -= t[0];
            gK[1] -= t[2];
            gM[0] -= t[4];
            gM[1] -= t[6];
            gI -= t[8];
            const double ia = theta * t[13];          // Im(mu^T Jg x^(k+1)) * theta
            const double ib = (1.0 - theta) * t[11];  // Im(mu^T Jg x^k) * (1-theta)
            dfa[k] = seq->q * ia;
            dfb[k] = seq->q * ib;
            gq += seq->f_implicit[k] * ia + seq->f_explicit[k] * ib;
        }
    }
    const auto t2 = std::chrono::steady_clock::now();
    for (int cc = 0; cc < 2; ++cc) {
        const bool live = cc < c->ncomp;
        grads->dS_dD[cc] = live ? gK[cc] : 0.0;
        grads->dS_dinvT2[cc] = live ? gM[cc] : 0.0;
    }
    grads->dS_dkappa = c->ncomp == 2 ? gI : 0.0;
    grads->dS_dq = gq;
    if (grads->dS_df_implicit)
        for (int k = 0; k < N; ++k) grads->dS_df_implicit[k] = dfa[k];
    if (grads->dS_df_explicit)
        for (int k = 0; k < N; ++k) grads->dS_df_explicit[k] = dfb[k];
    grads->ms_forward = std::chrono::duration<double, std::milli>(t1 - t0).count();
    grads->ms_adjoint = std::chrono::duration<double, std::milli>(t2 - t1).count();
    grads->checkpoints = keep_all ? 0 : nck;
    grads->recomputed_steps = recomputed;
    grads->iters_forward = it_fwd;
    grads->iters_recompute = it_rec;
    grads->iters_backward = st.iters_total - it_fwd - it_rec;
    if (stats) *stats = st;
    c->solved = false;  // work.x no longer holds u^N of a plain solve
}

// Replay the SpMV kernels of one preconditioned BiCGStab iteration as a solve launches them
// (preconditioner SpMV on p, v = A ph, preconditioner SpMV on s = r - alpha v writing s, t = A sh),
// `reps` times each, back to back, on the context's solver buffers; us[k] = mean duration of kernel
// k from the start and end timestamps of its own dispatch (kernel-bound events, no marker packets).
// poly = false replays the plain pair v = A p, t = A s instead (us[1] = us[3] = 0).
void replay_spmv(dmri_ctx *c, SolveGeom geo, int S, bool per_gdir, const std::vector<KrylovState> &st0, int reps,
                 bool poly, double us[4]) {
    hipStream_t s = c->stream;
    const TileCompression *tiles = tiles_for(c, geo);
    SolverWork &w = c->work;
    const int64_t n = c->ndof;
    const size_t np = (size_t)std::max(geo.grid, geo.vgrid) * S * kNV;
    if (w.part.n < 3 * np) w.part.alloc(3 * np);
    const size_t nS = (size_t)n * S;
    if (poly) {
        if (w.ph.n != nS) w.ph.alloc(nS);
        if (w.sh.n != nS) w.sh.alloc(nS);
        if (w.sv.n != nS) w.sv.alloc(nS);
        launch_init_x(n, S, w.sc.p, w.ph.p, s);
        launch_init_x(n, S, w.sc.p, w.sh.p, s);
        launch_init_x(n, S, w.sc.p, w.sv.p, s);
    }
    SolveArgs a{};
    a.n = n;
    a.rowptr = c->srow.p;
    a.col = c->scol.p;
    a.vals = w.vals.p;
    a.bre = w.bre.p;
    a.per_gdir = per_gdir;
    a.coefA = w.coef.p;
    a.coefB = w.coef.p;
    a.x = w.x.p;
    a.r = w.r.p;
    a.rh = w.rh.p;
    a.p = w.p.p;
    a.v = w.v.p;
    a.t = w.t.p;
    a.b = w.b.p;
    a.xprev = w.xprev.p;
    a.partU = w.part.p;
    a.partV = w.part.p + np;
    a.partT = w.part.p + 2 * np;
    if (poly) {
        a.ph = w.ph.p;
        a.sh = w.sh.p;
        a.s = w.sv.p;
        a.pc = 0.9;
    }
    const bool use_fin = S > 1 || std::max(geo.grid, geo.vgrid) > 512;  // as in a solve
    if (w.tot.n < (size_t)3 * S * kNV) w.tot.alloc(3 * (size_t)S * kNV);
    a.totU = use_fin ? w.tot.p : nullptr;
    a.totV = use_fin ? w.tot.p + S * kNV : nullptr;
    a.totT = use_fin ? w.tot.p + 2 * S * kNV : nullptr;
    a.st = w.st.p;
    a.all_done = w.all_done.p;
    a.rtol = 1e-12;
    a.maxit = 1;
    attach_tiles(a, tiles);
    DMRI_HIP(hipMemcpyAsync(w.st.p, st0.data(), sizeof(KrylovState) * S, hipMemcpyHostToDevice, s));
    DMRI_HIP(hipMemsetAsync(w.all_done.p, 0, sizeof(int), s));
    SolveArgs av = a;
    if (poly) av.p = a.ph;
    auto run = [&](int k, hipEvent_t e0, hipEvent_t e1) {
        switch (k) {
            case 0: launch_spmv_v(geo, a, 0, s, e0, e1, kOpPre); break;
            case 1: launch_spmv_v(geo, av, 0, s, e0, e1); break;
            case 2: launch_spmv_t(geo, a, 0, s, e0, e1, kOpPre); break;
            default: launch_spmv_t(geo, a, 0, s, e0, e1, poly ? kOpTP : kOpT); break;
        }
    };
    const int k0 = poly ? 0 : 1;
    for (int k = k0; k < 4; k += poly ? 1 : 2) run(k, nullptr, nullptr);  // warm-up
    reps = std::min(reps, kEventPool / 4);
    for (int r = 0; r < reps; ++r)
        for (int k = k0; k < 4; k += poly ? 1 : 2) run(k, c->ev0[4 * r + k], c->ev1[4 * r + k]);
    DMRI_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 4; ++k) {
        us[k] = 0.0;
        if (!poly && (k == 0 || k == 2)) continue;
        double sum = 0;
        for (int r = 0; r < reps; ++r) {
            float ms = 0;
            DMRI_HIP(hipEventElapsedTime(&ms, c->ev0[4 * r + k], c->ev1[4 * r + k]));
            sum += ms;
        }
        us[k] = 1e3 * sum / reps;
    }
}

// Scaled values and vectors of a tuning or replay run on the context's buffers (dt = 1, theta = 1/2,
// the x direction, a small imaginary coefficient)
void prepare_replay(dmri_ctx *c, const SolveGeom &g0, int S, bool per_gdir) {
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    ensure_group_workspace(c, c->work, g0, 1, per_gdir);
    SolverWork &w = c->work;
    auto &V = c->val;
    launch_prep_scale(n, c->sdiag.p, V[DMRI_MAT_M].p, V[DMRI_MAT_K].p, V[DMRI_MAT_R].p, V[DMRI_MAT_I].p, nullptr, 1.0,
                      0.5, w.sc.p, s);
    launch_prep_values(n, c->srow.p, c->scol.p, c->emap.p, w.sc.p, V[DMRI_MAT_M].p, V[DMRI_MAT_K].p, V[DMRI_MAT_R].p,
                       V[DMRI_MAT_I].p, nullptr, V[DMRI_MAT_JX].p, V[DMRI_MAT_JY].p, V[DMRI_MAT_JZ].p, per_gdir, 1.0, 0.0,
                       0.0, 1.0, 0.5, w.vals.p, w.bre.p, s);
    launch_init_x(n, S, w.sc.p, w.p.p, s);
    launch_init_x(n, S, w.sc.p, w.r.p, s);
    launch_init_x(n, S, w.sc.p, w.v.p, s);
    launch_init_x(n, S, w.sc.p, w.rh.p, s);
    std::vector<double> coef((size_t)S * 3, 1e-4);
    DMRI_HIP(hipMemcpyAsync(w.coef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, s));
}

std::vector<KrylovState> replay_states(int S) {
    std::vector<KrylovState> st0(S);
    std::memset(st0.data(), 0, sizeof(KrylovState) * S);
    for (auto &x : st0) {
        x.rho[0] = make_double2(1.0, 0.0);
        x.alpha[0] = make_double2(0.5, 0.0);
    }
    return st0;
}

void install_tuning(dmri_ctx *c, int S, bool per_gdir, int mapping, int lanes) {
    const long long key = (long long)S * 2 + (per_gdir ? 1 : 0);
    const int val = mapping * 64 + lanes;
    for (auto &t : c->tuned)
        if (t.first == key) {
            t.second = val;
            return;
        }
    c->tuned.emplace_back(key, val);
}

// Time the SpMV kernels of a polynomial-preconditioned BiCGStab iteration (the library's default
// solve) for every (mapping, lanes per row) candidate that runs it (mappings 1, 2, 4 and 6; the
// signals-in-registers and tile-compressed mappings have no preconditioner kernels) for S signals on
// the assembled pattern, and remember the fastest for later solves of this context.
// us[6*6]: [mapping-1][log2 lanes] microseconds per iteration's four SpMVs, -1 where not valid.
// Tuner timings: us[(mapping - 1) * 6 + log2(lanes)] for mappings 1..kMapPair and lanes 1..32
constexpr int kTuneSlots = kMapPair * 6;

void tune_impl(dmri_ctx *c, int S, bool per_gdir, int *best_lanes, int *best_mapping, double *us) {
    require(c->assembled, DMRI_ERR_STATE, "dmri_tune: call dmri_assemble first");
    require(S >= 1 && S <= kMaxSignals, DMRI_ERR_ARG, "dmri_tune: nsig must lie in [1, 64]");
    set_device(c);
    const int64_t n = c->ndof;
    const int cand[6] = {1, 2, 4, 8, 16, 32};
    prepare_replay(c, make_geom(n, 1, S, kMapSL), S, per_gdir);
    SolverWork &w = c->work;
    const std::vector<KrylovState> st0 = replay_states(S);
    double best_us = 1e300;
    int def_mapping, def_lanes;
    default_mapping(S, def_mapping, def_lanes);
    if (def_mapping == kMapSR || def_mapping == kMapTile) def_mapping = kMapStaged;
    *best_mapping = def_mapping;
    *best_lanes = def_lanes;
    double def_us = -1.0;
    // 20 replays per candidate, then the default is kept unless a candidate beats it by 3 %
    // (timing noise must not flip the choice between runs)
    int pair_sp = 1;
    while (2 * pair_sp < S) pair_sp *= 2;
    for (int m = kMapSL; m <= kMapPair; ++m)
        for (int i = 0; i < 6; ++i) {
            const int G = cand[i];
            double &slot = us[(m - 1) * 6 + i];
            slot = -1.0;
            if (m == kMapSR || m == kMapTile) continue;
            if (m == kMapWave && G != 1) continue;
            const bool pairs = m == kMapPair;
            if (pairs && (S < 2 || G > 4 || G * pair_sp > 64)) continue;
            if (m == kMapPair && (uint64_t)n * S * 16u >= (1ull << 32)) continue;
            if (m != kMapWave && !pairs && G * S > 64) continue;
            SolveGeom geo = make_geom(n, G, S, m);
            fit_spmv_grid(geo, per_gdir);
            const size_t np = (size_t)std::max(geo.grid, geo.vgrid) * S * kNV;
            if (w.part.n < 3 * np) w.part.alloc(3 * np);
            double u4[4];
            replay_spmv(c, geo, S, per_gdir, st0, 20, true, u4);
            slot = u4[0] + u4[1] + u4[2] + u4[3];
            if (m == def_mapping && G == def_lanes) def_us = slot;
            if (slot < best_us) {
                best_us = slot;
                *best_lanes = G;
                *best_mapping = m;
            }
        }
    if (def_us > 0.0 && best_us > 0.97 * def_us) {
        *best_mapping = def_mapping;
        *best_lanes = def_lanes;
    }
    install_tuning(c, S, per_gdir, *best_mapping, *best_lanes);
}

// Tuning database: a text file of "device ndof nnz nsig per_gdir mapping lanes us" lines
std::string device_name(const dmri_ctx *c) {
    hipDeviceProp_t prop{};
    DMRI_HIP(hipGetDeviceProperties(&prop, c->device));
    std::string name = prop.gcnArchName;
    for (auto &ch : name)
        if (ch == ' ' || ch == ':') ch = '_';
    return name.empty() ? "unknown" : name;
}

void tune_db_impl(dmri_ctx *c, const char *path, int S, bool per_gdir, int *best_lanes, int *best_mapping, int *hit) {
    require(c->assembled, DMRI_ERR_STATE, "dmri_tune_db: call dmri_assemble first");
    require(S >= 1 && S <= kMaxSignals, DMRI_ERR_ARG, "dmri_tune_db: nsig must lie in [1, 64]");
    const std::string dev = device_name(c);
    *hit = 0;
    if (path && *path) {
        if (FILE *f = std::fopen(path, "r")) {
            char d[128];
            long long nd = 0, nz = 0;
            int ns = 0, pg = 0, mp = 0, ln = 0;
            double u = 0;
            int found_m = 0, found_l = 0;
            while (std::fscanf(f, "%127s %lld %lld %d %d %d %d %lf", d, &nd, &nz, &ns, &pg, &mp, &ln, &u) == 8)
                if (dev == d && nd == c->ndof && nz == c->nnz && ns == S && pg == (per_gdir ? 1 : 0)) {
                    found_m = mp;  // the last matching line wins
                    found_l = ln;
                }
            std::fclose(f);
            if (found_m > 0) {
                (void)make_geom(c->ndof, found_l, S, found_m);  // validates the stored choice
                install_tuning(c, S, per_gdir, found_m, found_l);
                *best_mapping = found_m;
                *best_lanes = found_l;
                *hit = 1;
                return;
            }
        }
    }
    std::vector<double> us(kTuneSlots);
    tune_impl(c, S, per_gdir, best_lanes, best_mapping, us.data());
    if (path && *path) {
        const double u = us[(*best_mapping - 1) * 6 + (int)std::log2((double)*best_lanes)];
        FILE *f = std::fopen(path, "a");
        require(f != nullptr, DMRI_ERR_ARG, std::string("dmri_tune_db: cannot append to ") + path);
        std::fprintf(f, "%s %lld %lld %d %d %d %d %.3f\n", dev.c_str(), (long long)c->ndof, (long long)c->nnz, S,
                     per_gdir ? 1 : 0, *best_mapping, *best_lanes, u);
        std::fclose(f);
    }
}

void validate_seq(const dmri_seq *seq, const dmri_solver_opts *opts) {
    require(seq != nullptr && opts != nullptr, DMRI_ERR_ARG, "null argument");
    require(seq->nsteps >= 0, DMRI_ERR_ARG, "nsteps < 0");
    require(seq->nsteps == 0 || (seq->f_implicit && seq->f_explicit), DMRI_ERR_ARG, "null time profile");
    require(seq->dt > 0 && std::isfinite(seq->dt), DMRI_ERR_ARG, "dt must be > 0");
    require(seq->theta >= 0.0 && seq->theta <= 1.0, DMRI_ERR_ARG, "theta must lie in [0, 1]");
    require(opts->maxit > 0, DMRI_ERR_ARG, "maxit must be > 0");
    require(opts->rtol >= 0 && opts->atol >= 0 && (opts->rtol > 0 || opts->atol > 0), DMRI_ERR_ARG,
            "need rtol > 0 or atol > 0");
    require(opts->lanes_per_row >= 0, DMRI_ERR_ARG, "lanes_per_row must be >= 0");
    require(opts->initial_guess >= 0 && opts->initial_guess <= 3, DMRI_ERR_ARG, "initial_guess must lie in [0, 3]");
    require(opts->preconditioner >= 0 && opts->preconditioner <= 2, DMRI_ERR_ARG, "preconditioner must lie in [0, 2]");
    require(opts->krylov >= 0 && opts->krylov <= 2, DMRI_ERR_ARG, "krylov must be 0 (BiCGStab), 1 (GMRES) or 2 (COCG)");
    require(opts->krylov != 2 || opts->preconditioner != 1, DMRI_ERR_ARG,
            "COCG runs with the polynomial preconditioner (preconditioner 0 or 2)");
    require(opts->pre_fp32 == 0 || opts->pre_fp32 == 1, DMRI_ERR_ARG, "pre_fp32 must be 0 or 1");
    require(opts->pre_fp32 == 0 || opts->krylov == 2, DMRI_ERR_ARG,
            "pre_fp32 (FP32 preconditioner input) is a COCG option (krylov 2)");
    for (int i = 0; i < seq->nsteps; ++i)
        require(std::isfinite(seq->f_implicit[i]) && std::isfinite(seq->f_explicit[i]), DMRI_ERR_ARG,
                "time profile has a non-finite value");
}
void validate_u0(const dmri_ctx *c, const dmri_seq *seq) {
    if (!seq->u0_vertex) return;
    for (int64_t v = 0; v < c->nv; ++v)
        require(std::isfinite(seq->u0_vertex[v]), DMRI_ERR_ARG, "u0_vertex has a non-finite value");
}

// Checks of a batched solve's inputs; returns the unit gradient directions (3 per signal).
std::vector<double> checked_signals(dmri_ctx *c, int nsig, const double *q, const double *gdir, const dmri_seq *seq,
                                    const dmri_solver_opts *opts) {
    require(c->assembled, DMRI_ERR_STATE, "solve: call dmri_assemble first");
    require(nsig >= 1 && q && gdir, DMRI_ERR_ARG, "solve: need nsig >= 1 and non-null q, gdir");
    validate_seq(seq, opts);
    validate_u0(c, seq);
    require(seq->snap_every >= 0, DMRI_ERR_ARG, "snap_every must be >= 0");
    require(seq->snap_every == 0 || seq->nsteps == 0 || seq->snap_out != nullptr, DMRI_ERR_ARG,
            "snap_every > 0 needs snap_out");
    if ((c->periodic || c->strong_active) && seq->nsteps > 0) {
        require(seq->F_explicit != nullptr, DMRI_ERR_ARG,
                "periodic directions are set: dmri_seq.F_explicit (int_0^t f per step) is required");
        for (int i = 0; i < seq->nsteps; ++i)
            require(std::isfinite(seq->F_explicit[i]), DMRI_ERR_ARG, "F_explicit has a non-finite value");
    }
    if (c->strong_active && seq->nsteps > 0) {
        require(seq->F_implicit != nullptr, DMRI_ERR_ARG,
                "the strong periodic condition needs dmri_seq.F_implicit (int_0^t f at each step's time)");
        for (int i = 0; i < seq->nsteps; ++i)
            require(std::isfinite(seq->F_implicit[i]), DMRI_ERR_ARG, "F_implicit has a non-finite value");
    }
    std::vector<double> g(3 * (size_t)nsig);
    for (int j = 0; j < nsig; ++j) {
        require(std::isfinite(q[j]), DMRI_ERR_ARG, "q is not finite");
        const double *d = gdir + 3 * (size_t)j;
        const double gn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        // |g| = 0 is rejected by the reference (set_gradient_dir, DmriFemLib.py:813-823)
        require(gn > 1e-10 && std::isfinite(gn), DMRI_ERR_ARG, "|g|=0: gradient direction must be non-zero");
        for (int k = 0; k < 3; ++k) g[3 * (size_t)j + k] = d[k] / gn;
    }
    return g;
}

// Launch groups of up to G consecutive signals; with the strong periodic form a group also ends where
// q changes (its real part depends on q).  (first signal, count) pairs.
std::vector<std::pair<int, int>> launch_groups(const dmri_ctx *c, int nsig, const double *q, int G) {
    std::vector<std::pair<int, int>> out;
    for (int j0 = 0; j0 < nsig;) {
        int S = 1;
        while (j0 + S < nsig && S < G && (!c->strong_active || q[j0 + S] == q[j0])) ++S;
        out.emplace_back(j0, S);
        j0 += S;
    }
    return out;
}

// true when signals j0 .. j0+S-1 do not all share one gradient direction
bool group_per_gdir(const std::vector<double> &g, int j0, int S) {
    bool same = true;
    for (int j = 1; j < S && same; ++j)
        for (int k = 0; k < 3; ++k) same = same && g[3 * (size_t)(j0 + j) + k] == g[3 * (size_t)j0 + k];
    return !same;
}

// nsig signals with per-signal q and gradient direction, in launch groups of <= 64.
void solve_batched_impl(dmri_ctx *c, int nsig, const double *q, const double *gdir, const dmri_seq *seq,
                        const dmri_solver_opts *opts, double *sig, dmri_stats *stats) {
    require(sig != nullptr, DMRI_ERR_ARG, "solve: sig_out is null");
    const std::vector<double> g = checked_signals(c, nsig, q, gdir, seq, opts);
    set_device(c);
    dmri_stats st{};
    const auto t0 = std::chrono::steady_clock::now();
    for (const auto &gr : launch_groups(c, nsig, q, kMaxSignals)) {
        const int j0 = gr.first, S = gr.second;
        solve_group(c, S, j0, nsig, q + j0, g.data() + 3 * (size_t)j0, group_per_gdir(g, j0, S), seq, opts,
                    sig + 2 * (size_t)j0, st);
    }
    const auto t1 = std::chrono::steady_clock::now();
    st.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (stats) *stats = st;
    c->solved = true;
}

}  // namespace

// A batched solve that advances on request: every launch group keeps its state, its Krylov
// workspace and its projection history in HBM between calls.
struct dmri_batch {
    struct Group {
        SolverWork work;
        Stepper run;
        dmri_stats st{};
        int j0 = 0, S = 0;
    };
    dmri_ctx *c = nullptr;
    uint64_t generation = 0;
    int nsig = 0, N = 0, done = 0;
    dmri_solver_opts opts{};
    std::vector<std::unique_ptr<Group>> groups;
};

namespace {

void batch_begin_impl(dmri_ctx *c, int nsig, const double *q, const double *gdir, const dmri_seq *seq,
                      const dmri_solver_opts *opts, int group, dmri_batch **out) {
    require(out != nullptr, DMRI_ERR_ARG, "dmri_batch_begin: out is null");
    *out = nullptr;
    require(group >= 0 && group <= kMaxSignals, DMRI_ERR_ARG, "dmri_batch_begin: group must lie in [0, 64]");
    require(seq != nullptr && seq->snap_every == 0, DMRI_ERR_ARG,
            "dmri_batch_begin: snapshots are not taken by a resumable batch (snap_every must be 0)");
    const std::vector<double> g = checked_signals(c, nsig, q, gdir, seq, opts);
    set_device(c);
    auto b = std::make_unique<dmri_batch>();
    b->c = c;
    b->generation = c->generation;
    b->nsig = nsig;
    b->N = seq->nsteps;
    b->opts = *opts;
    const int G = group > 0 ? group : kMaxSignals;
    for (const auto &rg : launch_groups(c, nsig, q, G)) {
        const int j0 = rg.first, S = rg.second;
        auto gr = std::make_unique<dmri_batch::Group>();
        gr->j0 = j0;
        gr->S = S;
        gr->run.setup(c, S, q + j0, g.data() + 3 * (size_t)j0, group_per_gdir(g, j0, S), seq, &b->opts, &gr->st,
                      &gr->work);
        b->groups.push_back(std::move(gr));
    }
    DMRI_HIP(hipStreamSynchronize(c->stream));
    *out = b.release();
}

void batch_check(dmri_batch *b) {
    require(b != nullptr, DMRI_ERR_ARG, "null batch");
    require(b->c->assembled && b->generation == b->c->generation, DMRI_ERR_STATE,
            "the context was assembled again (or its inputs changed) after dmri_batch_begin");
}

void batch_advance_impl(dmri_batch *b, int nsteps, dmri_stats *stats) {
    batch_check(b);
    require(nsteps >= 0, DMRI_ERR_ARG, "nsteps < 0");
    set_device(b->c);
    for (auto &gr : b->groups) gr->st = dmri_stats{};
    const auto t0 = std::chrono::steady_clock::now();
    const int end = std::min(b->N, b->done + nsteps);
    for (; b->done < end; ++b->done)
        for (auto &gr : b->groups) gr->run.step(b->done);
    DMRI_HIP(hipStreamSynchronize(b->c->stream));
    const auto t1 = std::chrono::steady_clock::now();
    if (!stats) return;
    dmri_stats st{};
    for (auto &gr : b->groups) {
        const dmri_stats &x = gr->st;
        st.iters_total += x.iters_total;
        st.iters_max = std::max(st.iters_max, x.iters_max);
        st.restarts += x.restarts;
        st.graph_launches += x.graph_launches;
        st.relres_max = std::max(st.relres_max, x.relres_max);
        st.spmv_ms += x.spmv_ms;
        st.spmv_launches += x.spmv_launches;
        st.spmv_bytes = gr->run.cocg_on ? gr->run.kbytes[2] : 0.5 * (gr->run.kbytes[2] + gr->run.kbytes[5]);
        st.precond_used = gr->run.precond;
        for (int j = 0; j < 7; ++j) {
            st.kern_ms[j] += x.kern_ms[j];
            st.kern_launches[j] += x.kern_launches[j];
            st.kern_bytes[j] += x.kern_bytes[j];
        }
    }
    st.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    *stats = st;
}

}  // namespace

extern "C" {

const char *dmri_last_error(void) { return g_last_error.c_str(); }

const char *dmri_version(void) { return "libdmri 0.1 (gfx950, FP64)"; }

int dmri_create(dmri_ctx **out, int device) {
    return guarded([&] {
        require(out != nullptr, DMRI_ERR_ARG, "dmri_create: out is null");
        *out = nullptr;
        int ndev = 0;
        DMRI_HIP(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, DMRI_ERR_ARG, "dmri_create: no HIP device " + std::to_string(device));
        DMRI_HIP(hipSetDevice(device));
        (void)scratch_slot(device);
        auto *c = new dmri_ctx();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error(DMRI_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        *out = c;
    });
}

int dmri_destroy(dmri_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int dmri_mesh_set(dmri_ctx *c, int64_t nv, const double *xyz, int64_t nc, const int32_t *tets,
                  const int32_t *cell_phase) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        require(nv > 0 && nc > 0 && xyz && tets, DMRI_ERR_ARG, "dmri_mesh_set: empty mesh");
        require(nv < (1LL << 30) && nc < (1LL << 29), DMRI_ERR_LIMIT, "mesh too large for 32-bit indices");
        for (int64_t i = 0; i < 4 * nc; ++i)
            require(tets[i] >= 0 && tets[i] < nv, DMRI_ERR_MESH,
                    "cell " + std::to_string(i / 4) + " has a vertex index out of range");
        for (int64_t i = 0; i < 3 * nv; ++i) require(std::isfinite(xyz[i]), DMRI_ERR_MESH, "non-finite coordinate");
        if (cell_phase)
            for (int64_t i = 0; i < nc; ++i)
                require(cell_phase[i] == 0 || cell_phase[i] == 1, DMRI_ERR_MESH, "cell phase must be 0 or 1");
        double hmin = 1e300, hmax = 0;
        for (int64_t t = 0; t < nc; ++t) {
            double h = 0;
            for (int a = 0; a < 4; ++a)
                for (int b = a + 1; b < 4; ++b) {
                    const double *p = xyz + 3 * tets[4 * t + a], *q = xyz + 3 * tets[4 * t + b];
                    double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
                    h = std::max(h, std::sqrt(dx * dx + dy * dy + dz * dz));
                }
            require(h > 0, DMRI_ERR_MESH, "degenerate cell " + std::to_string(t));
            hmin = std::min(hmin, h);
            hmax = std::max(hmax, h);
        }
        set_device(c);
        c->xyz.upload(xyz, 3 * nv, c->stream);
        c->h_xyz.assign(xyz, xyz + 3 * nv);
        c->tets.upload(tets, 4 * nc, c->stream);
        if (cell_phase) c->phase.upload(cell_phase, nc, c->stream);
        else c->phase.release();
        DMRI_HIP(hipStreamSynchronize(c->stream));
        c->nv = nv;
        c->nc = nc;
        c->ncomp = cell_phase ? 2 : 1;
        c->hmin = hmin;
        c->hmax = hmax;
        for (int d = 0; d < 3; ++d) {
            c->bb_lo[d] = 1e300;
            c->bb_hi[d] = -1e300;
        }
        for (int64_t v = 0; v < nv; ++v)
            for (int d = 0; d < 3; ++d) {
                c->bb_lo[d] = std::min(c->bb_lo[d], xyz[3 * v + d]);
                c->bb_hi[d] = std::max(c->bb_hi[d], xyz[3 * v + d]);
            }
        c->has_mesh = true;
        c->assembled = false;
        c->solved = false;
        c->has_Dcell = c->has_iT2cell = c->has_kappa_f = false;
        double d2[2] = {c->D[0], c->D[1]}, t2[2] = {c->invT2[0], c->invT2[1]};
        c->Dc.upload(d2, 2, c->stream);
        c->iT2c.upload(t2, 2, c->stream);
        DMRI_HIP(hipStreamSynchronize(c->stream));
    });
}

int dmri_coeffs_set(dmri_ctx *c, const dmri_coeffs *k) {
    return guarded([&] {
        require(c != nullptr && k != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->has_mesh, DMRI_ERR_STATE, "dmri_coeffs_set: set the mesh first");
        for (int i = 0; i < 2; ++i) {
            require(std::isfinite(k->D[i]) && k->D[i] >= 0, DMRI_ERR_ARG, "D must be finite and >= 0");
            require(std::isfinite(k->invT2[i]) && k->invT2[i] >= 0, DMRI_ERR_ARG, "1/T2 must be finite and >= 0");
        }
        require(std::isfinite(k->kappa) && k->kappa >= 0, DMRI_ERR_ARG, "kappa must be finite and >= 0");
        set_device(c);
        c->D[0] = k->D[0];
        c->D[1] = k->D[1];
        c->invT2[0] = k->invT2[0];
        c->invT2[1] = k->invT2[1];
        c->kappa = k->kappa;
        c->Dc.upload(c->D, 2, c->stream);
        c->iT2c.upload(c->invT2, 2, c->stream);
        c->has_Dcell = k->D_cell != nullptr;
        c->Dcell_symmetric = true;
        if (c->has_Dcell) {
            for (int64_t t = 0; t < c->nc; ++t) {
                const double *d = k->D_cell + 9 * t;
                double mx = 0.0;
                for (int i = 0; i < 9; ++i) {
                    require(std::isfinite(d[i]), DMRI_ERR_ARG, "D_cell has a non-finite value");
                    mx = std::max(mx, std::fabs(d[i]));
                }
                const double tol = 1e-12 * mx;
                if (std::fabs(d[1] - d[3]) > tol || std::fabs(d[2] - d[6]) > tol || std::fabs(d[5] - d[7]) > tol)
                    c->Dcell_symmetric = false;
            }
            c->Dcell.upload(k->D_cell, 9 * c->nc, c->stream);
        }
        c->has_iT2cell = k->invT2_cell != nullptr;
        if (c->has_iT2cell) c->iT2cell.upload(k->invT2_cell, c->nc, c->stream);
        c->has_kappa_f = k->kappa_facet != nullptr;
        if (c->has_kappa_f) {
            require(c->assembled, DMRI_ERR_STATE,
                    "kappa_facet needs the interface facet count: assemble once, then set it and assemble again");
            c->kappa_f.upload(k->kappa_facet, c->n_iface, c->stream);
        }
        DMRI_HIP(hipStreamSynchronize(c->stream));
        c->assembled = false;
        c->solved = false;
    });
}

int dmri_assemble(dmri_ctx *c) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        assemble_impl(c);
    });
}

int dmri_get_info(dmri_ctx *c, dmri_info *info) {
    return guarded([&] {
        require(c != nullptr && info != nullptr, DMRI_ERR_ARG, "null argument");
        std::memset(info, 0, sizeof(*info));
        info->nv = c->nv;
        info->nc = c->nc;
        info->ncomp = c->ncomp;
        info->hmin = c->hmin;
        info->hmax = c->hmax;
        if (c->assembled) {
            info->ndof = c->ndof;
            info->nnz = c->nnz;
            info->n_iface = c->n_iface;
            info->n_bfacet = c->n_bfacet;
            info->ncolors_cells = c->ncol_cells;
            info->ncolors_iface = c->ncol_iface;
            info->volume = c->volume;
            info->setup_ms = c->setup_ms;
            info->n_periodic = (c->periodic || c->strong_active) ? c->n_per : 0;
            info->kappa_e = c->periodic ? 3e-3 / c->hmin : 0.0;
        }
    });
}

int dmri_export_pattern(dmri_ctx *c, int64_t *rowptr, int32_t *col, int32_t *dof_of) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        require(c->assembled, DMRI_ERR_STATE, "not assembled");
        set_device(c);
        hipStream_t s = c->stream;
        if (rowptr) {
            std::vector<int> rp(c->ndof + 1);
            DMRI_HIP(hipMemcpyAsync(rp.data(), c->rowptr.p, sizeof(int) * rp.size(), hipMemcpyDeviceToHost, s));
            DMRI_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < rp.size(); ++i) rowptr[i] = rp[i];
        }
        if (col) DMRI_HIP(hipMemcpyAsync(col, c->col.p, sizeof(int) * c->nnz, hipMemcpyDeviceToHost, s));
        if (dof_of)
            DMRI_HIP(hipMemcpyAsync(dof_of, c->dof_of.p, sizeof(int) * c->ncomp * c->nv, hipMemcpyDeviceToHost, s));
        DMRI_HIP(hipStreamSynchronize(s));
    });
}

int dmri_export_values(dmri_ctx *c, int which, double *val) {
    return guarded([&] {
        require(c != nullptr && val != nullptr, DMRI_ERR_ARG, "null argument");
        require(which >= 0 && which <= DMRI_MAT_GZ, DMRI_ERR_ARG, "unknown matrix id");
        require(c->assembled, DMRI_ERR_STATE, "not assembled");
        set_device(c);
        const bool gdir_term = which >= DMRI_MAT_GX;
        if ((which == DMRI_MAT_P && !c->periodic) || (gdir_term && !c->strong_active)) {
            std::memset(val, 0, sizeof(double) * c->nnz);
            return;
        }
        const double *src = which == DMRI_MAT_P ? c->pval.p : (gdir_term ? c->gval[which - DMRI_MAT_GX].p
                                                                           : c->val[which].p);
        DMRI_HIP(hipMemcpyAsync(val, src, sizeof(double) * c->nnz, hipMemcpyDeviceToHost, c->stream));
        DMRI_HIP(hipStreamSynchronize(c->stream));
    });
}

int dmri_export_facets(dmri_ctx *c, int32_t *fv) {
    return guarded([&] {
        require(c != nullptr && fv != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->assembled, DMRI_ERR_STATE, "not assembled");
        set_device(c);
        if (c->n_iface)
            DMRI_HIP(hipMemcpyAsync(fv, c->iface_v.p, sizeof(int) * 3 * c->n_iface, hipMemcpyDeviceToHost, c->stream));
        DMRI_HIP(hipStreamSynchronize(c->stream));
    });
}

int dmri_periodic_set(dmri_ctx *c, const int32_t dir[3]) {
    return guarded([&] {
        require(c != nullptr && dir != nullptr, DMRI_ERR_ARG, "null argument");
        for (int d = 0; d < 3; ++d) c->pdir[d] = dir[d] != 0 ? 1 : 0;
        c->strong = false;
        c->assembled = false;
        c->solved = false;
    });
}

int dmri_periodic_strong_set(dmri_ctx *c, const int32_t dir[3]) {
    return guarded([&] {
        require(c != nullptr && dir != nullptr, DMRI_ERR_ARG, "null argument");
        for (int d = 0; d < 3; ++d) c->pdir[d] = dir[d] != 0 ? 1 : 0;
        c->strong = true;
        c->assembled = false;
        c->solved = false;
    });
}

int dmri_export_periodic(dmri_ctx *c, int32_t *dof, int32_t *src, double *w, double *ell) {
    return guarded([&] {
        require(c != nullptr && dof && src && w && ell, DMRI_ERR_ARG, "null argument");
        require(c->assembled, DMRI_ERR_STATE, "not assembled");
        if (!c->periodic || c->n_per == 0) return;
        set_device(c);
        hipStream_t s = c->stream;
        const size_t ne = (size_t)c->n_per;
        DMRI_HIP(hipMemcpyAsync(dof, c->per_dof.p, sizeof(int) * ne, hipMemcpyDeviceToHost, s));
        DMRI_HIP(hipMemcpyAsync(src, c->per_src.p, sizeof(int) * 3 * ne, hipMemcpyDeviceToHost, s));
        DMRI_HIP(hipMemcpyAsync(w, c->per_w.p, sizeof(double) * 3 * ne, hipMemcpyDeviceToHost, s));
        DMRI_HIP(hipMemcpyAsync(ell, c->per_ell.p, sizeof(double) * 3 * ne, hipMemcpyDeviceToHost, s));
        DMRI_HIP(hipStreamSynchronize(s));
    });
}

int dmri_solve(dmri_ctx *c, const dmri_seq *seq, const dmri_solver_opts *opts, double sig_out[2],
               dmri_stats *stats) {
    return guarded([&] {
        require(c != nullptr && seq != nullptr, DMRI_ERR_ARG, "null argument");
        solve_batched_impl(c, 1, &seq->q, seq->gdir, seq, opts, sig_out, stats);
    });
}

int dmri_solve_batched(dmri_ctx *c, int32_t nsig, const double *q, const double *gdir, const dmri_seq *seq,
                       const dmri_solver_opts *opts, double *sig_out, dmri_stats *stats) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        solve_batched_impl(c, nsig, q, gdir, seq, opts, sig_out, stats);
    });
}

int dmri_tune(dmri_ctx *c, int32_t nsig, int32_t per_gdir, int32_t *best_lanes, int32_t *best_mapping,
              double *us) {
    return guarded([&] {
        require(c != nullptr && best_lanes != nullptr && best_mapping != nullptr && us != nullptr, DMRI_ERR_ARG,
                "null argument");
        int bl = 0, bm = 0;
        tune_impl(c, nsig, per_gdir != 0, &bl, &bm, us);
        *best_lanes = bl;
        *best_mapping = bm;
    });
}

int dmri_time_spmv(dmri_ctx *c, int32_t nsig, int32_t per_gdir, int32_t reps, double *us_v, double *us_t,
                   double *bytes_per_launch) {
    return guarded([&] {
        require(c != nullptr && us_v != nullptr && us_t != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->assembled, DMRI_ERR_STATE, "dmri_time_spmv: call dmri_assemble first");
        require(nsig >= 1 && nsig <= kMaxSignals && reps >= 1, DMRI_ERR_ARG, "need 1 <= nsig <= 64 and reps >= 1");
        set_device(c);
        const bool pg = per_gdir != 0;
        int mapping, lanes;
        default_mapping(nsig, mapping, lanes);
        for (const auto &t : c->tuned)
            if (t.first == (long long)nsig * 2 + (pg ? 1 : 0)) {
                mapping = t.second / 64;
                lanes = t.second % 64;
            }
        // the preconditioned solve runs mappings 1, 2, 4 and 6 (others fall back to 2)
        if (mapping == kMapSR || mapping == kMapTile) {
            mapping = kMapStaged;
            while (lanes > 1 && lanes * nsig > 64) lanes >>= 1;
        }
        SolveGeom geo = make_geom(c->ndof, lanes, nsig, mapping);
        fit_spmv_grid(geo, pg);
        // the buffers keep the last solve's scaled values and vectors when it had this shape
        if (c->work.x.n != (size_t)c->ndof * nsig || c->work.vals.n != (size_t)c->nnz * (pg ? 4 : 2))
            prepare_replay(c, geo, nsig, pg);
        double u4[4];
        replay_spmv(c, geo, nsig, pg, replay_states(nsig), reps, true, u4);
        *us_v = u4[1];
        *us_t = u4[3];
        if (bytes_per_launch)
            *bytes_per_launch = (double)c->nnz * (4.0 + 8.0 * (pg ? 4 : 2)) + 4.0 * (double)(c->ndof + 1) +
                                48.0 * (double)c->ndof * nsig;
    });
}

int dmri_tune_db(dmri_ctx *c, const char *path, int32_t nsig, int32_t per_gdir, int32_t *best_lanes,
                 int32_t *best_mapping, int32_t *hit) {
    return guarded([&] {
        require(c != nullptr && best_lanes != nullptr && best_mapping != nullptr && hit != nullptr, DMRI_ERR_ARG,
                "null argument");
        set_device(c);
        int bl = 0, bm = 0, h = 0;
        tune_db_impl(c, path, nsig, per_gdir != 0, &bl, &bm, &h);
        *best_lanes = bl;
        *best_mapping = bm;
        *hit = h;
    });
}

int dmri_set_tuning(dmri_ctx *c, int32_t nsig, int32_t per_gdir, int32_t lanes, int32_t mapping) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        require(nsig >= 1 && nsig <= kMaxSignals, DMRI_ERR_ARG, "nsig must lie in [1, 64]");
        (void)make_geom(c->ndof > 0 ? c->ndof : 1, lanes, nsig, mapping);  // validates the combination
        const long long key = (long long)nsig * 2 + (per_gdir ? 1 : 0);
        for (auto &t : c->tuned)
            if (t.first == key) {
                t.second = mapping * 64 + lanes;
                return;
            }
        c->tuned.emplace_back(key, mapping * 64 + lanes);
    });
}

int dmri_copy_bandwidth(dmri_ctx *c, int64_t bytes, int32_t reps, double *gbps) {
    return guarded([&] {
        require(c != nullptr && gbps != nullptr, DMRI_ERR_ARG, "null argument");
        require(bytes >= 16 && reps >= 1, DMRI_ERR_ARG, "need bytes >= 16 and reps >= 1");
        set_device(c);
        const int64_t n = bytes / 16;
        DevArray<double2> src, dst;
        src.alloc(n);
        dst.alloc(n);
        src.zero(c->stream);
        hipEvent_t e0, e1;
        DMRI_HIP(hipEventCreate(&e0));
        DMRI_HIP(hipEventCreate(&e1));
        launch_copy(src.p, dst.p, n, c->stream);
        DMRI_HIP(hipEventRecord(e0, c->stream));
        for (int i = 0; i < reps; ++i) launch_copy(src.p, dst.p, n, c->stream);
        DMRI_HIP(hipEventRecord(e1, c->stream));
        DMRI_HIP(hipEventSynchronize(e1));
        float ms = 0;
        DMRI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = 2.0 * 16.0 * (double)n * reps / (ms * 1e-3) / 1e9;
    });
}

int dmri_adjoint(dmri_ctx *c, const dmri_seq *seq, const dmri_solver_opts *opts, const dmri_adjoint_opts *aopts,
                 double sig_out[2], dmri_grads *grads, dmri_stats *stats) {
    return guarded([&] {
        require(c != nullptr && grads != nullptr && sig_out != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->assembled, DMRI_ERR_STATE, "dmri_adjoint: call dmri_assemble first");
        require(!c->periodic && !c->strong_active, DMRI_ERR_ARG,
                "dmri_adjoint: the pseudo-periodic conditions are not supported");
        require(!c->has_Dcell || c->Dcell_symmetric, DMRI_ERR_ARG,
                "dmri_adjoint: a non-symmetric D_cell tensor makes the operator non-symmetric; the adjoint "
                "sweep needs a complex-symmetric operator");
        validate_seq(seq, opts);
        validate_u0(c, seq);
        set_device(c);
        adjoint_impl(c, seq, opts, aopts, sig_out, grads, stats);
    });
}

int dmri_get_solution(dmri_ctx *c, int32_t sig, double *u) {
    return guarded([&] {
        require(c != nullptr && u != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->solved, DMRI_ERR_STATE, "no solution (call dmri_solve first)");
        require(sig >= 0 && sig < c->last_S, DMRI_ERR_ARG, "signal index outside the last launch group");
        set_device(c);
        DevArray<double2> tmp;
        tmp.alloc(c->ndof);
        launch_unscale_perm(c->ndof, c->last_S, sig, c->perm.p, c->work.sc.p, c->work.x.p, tmp.p, c->stream);
        DMRI_HIP(hipMemcpyAsync(u, tmp.p, sizeof(double2) * c->ndof, hipMemcpyDeviceToHost, c->stream));
        DMRI_HIP(hipStreamSynchronize(c->stream));
    });
}

int dmri_get_solution_noalloc(dmri_ctx *c, int32_t sig, double *u) {
    return guarded([&] {
        require(c != nullptr && u != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->solved, DMRI_ERR_STATE, "no solution (call dmri_solve first)");
        require(sig >= 0 && sig < c->last_S, DMRI_ERR_ARG, "signal index outside the last launch group");
        require(c->ndof <= DMRI_NOALLOC_MAX_DOF, DMRI_ERR_LIMIT,
                "dmri_get_solution_noalloc: " + std::to_string(c->ndof) + " DOFs exceed its buffer of " +
                    std::to_string(DMRI_NOALLOC_MAX_DOF) + " (use dmri_get_solution)");
        set_device(c);
        ScratchSlot &slot = scratch_slot(c->device);
        std::lock_guard<std::mutex> lk(slot.mu);
        launch_unscale_perm(c->ndof, c->last_S, sig, c->perm.p, c->work.sc.p, c->work.x.p, slot.ptr, c->stream);
        DMRI_HIP(hipMemcpyAsync(u, slot.ptr, sizeof(double2) * c->ndof, hipMemcpyDeviceToHost, c->stream));
        DMRI_HIP(hipStreamSynchronize(c->stream));
    });
}

int dmri_batch_begin(dmri_ctx *c, int32_t nsig, const double *q, const double *gdir, const dmri_seq *seq,
                     const dmri_solver_opts *opts, int32_t group, dmri_batch **out) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        batch_begin_impl(c, nsig, q, gdir, seq, opts, group, out);
    });
}

int dmri_batch_advance(dmri_batch *b, int32_t nsteps, dmri_stats *stats) {
    return guarded([&] { batch_advance_impl(b, nsteps, stats); });
}

int dmri_batch_signals(dmri_batch *b, double *sig_out, int32_t *steps_done) {
    return guarded([&] {
        batch_check(b);
        require(sig_out != nullptr, DMRI_ERR_ARG, "null argument");
        set_device(b->c);
        for (auto &gr : b->groups) gr->run.signal(sig_out + 2 * (size_t)gr->j0);
        if (steps_done) *steps_done = b->done;
    });
}

int dmri_batch_end(dmri_batch *b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->c->device);
        (void)hipStreamSynchronize(b->c->stream);
        delete b;
    });
}

}  // extern "C"
