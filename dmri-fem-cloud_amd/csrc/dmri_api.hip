// libdmri host code: the C ABI of include/dmri.h, the device set-up pipeline and the
// theta-scheme time loop driving the gfx950 kernels of dmri_kernels.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dmri.h"
#include "dmri_internal.hpp"

using namespace dmri;

namespace {

thread_local std::string g_last_error;

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
                throw Error(DMRI_ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
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

template <class T>
T read_scalar(const T *d, hipStream_t s) {
    T h;
    DMRI_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    return h;
}

constexpr int kEventPool = 512;

}  // namespace

struct dmri_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh
    int64_t nv = 0, nc = 0;
    int ncomp = 1;
    bool has_mesh = false;
    double hmin = 0, hmax = 0;
    DevArray<double> xyz;
    DevArray<int> tets, phase;
    // coefficients
    double D[2] = {3e-3, 3e-3};
    double invT2[2] = {1e-16, 1e-16};
    double kappa = 0.0;
    DevArray<double> Dc, iT2c, Dcell, iT2cell, kappa_f;
    bool has_Dcell = false, has_iT2cell = false, has_kappa_f = false;
    // assembled
    bool assembled = false;
    int64_t ndof = 0, nnz = 0, n_iface = 0, n_bfacet = 0;
    int ncol_cells = 0, ncol_iface = 0;
    double volume = 0, setup_ms = 0;
    DevArray<int> dof_of, dof_vert, dof_comp, rowptr, col, diag;
    DevArray<int> vptr, vlist, fptr, flist, iface_v, iface_c, bnd_v, bnd_c;
    DevArray<double> val[7];
    DevArray<double> w;
    // solver workspace
    DevArray<double2> aj, x, r, rh, p, v, t;
    DevArray<double> bre, sc, part, dsig;
    DevArray<KrylovState> st;
    DevArray<unsigned> counters;
    KrylovState *h_st = nullptr;
    bool solved = false;
    hipEvent_t ev0[kEventPool], ev1[kEventPool];
    bool events_made = false;

    ~dmri_ctx() {
        if (h_st) (void)hipHostFree(h_st);
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

void assemble_impl(dmri_ctx *c) {
    require(c->has_mesh, DMRI_ERR_STATE, "dmri_assemble: no mesh (call dmri_mesh_set first)");
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
        exclusive_scan(cnt.p, c->rowptr.p, c->ndof, s);
        c->nnz = read_scalar(c->rowptr.p + c->ndof, s);
        require(c->nnz > 0, DMRI_ERR_MESH, "empty pattern");
        c->col.alloc(c->nnz);
        launch_pattern(1, c->ndof, nv, c->dof_vert.p, c->dof_comp.p, c->tets.p, phase, c->vptr.p, c->vlist.p, fptr,
                       flist, c->iface_v.p, c->dof_of.p, nullptr, c->rowptr.p, c->col.p, overflow.p, s);
    }

    // 5. colouring + assembly
    for (auto &a : c->val) {
        a.alloc(c->nnz);
        a.zero(s);
    }
    DevArray<int> miss;
    miss.alloc(1);
    miss.zero(s);
    {
        DevArray<int> perm, ct, cph;
        std::vector<int64_t> offs = colour_elements(4, c->tets.p, nc, c->vptr.p, c->vlist.p, perm, s);
        c->ncol_cells = (int)offs.size() - 1;
        ct.alloc(4 * nc);
        cph.alloc(nc);
        launch_gather_cells(perm.p, nc, c->tets.p, phase, ct.p, cph.p, s);
        for (int k = 0; k < c->ncol_cells; ++k)
            launch_assemble_cells(offs[k], offs[k + 1] - offs[k], nc, perm.p, ct.p, cph.p, c->xyz.p, nv, c->dof_of.p,
                                  c->rowptr.p, c->col.p, c->Dc.p, c->has_Dcell ? c->Dcell.p : nullptr, c->iT2c.p,
                                  c->has_iT2cell ? c->iT2cell.p : nullptr, c->val[DMRI_MAT_M].p,
                                  c->val[DMRI_MAT_K].p, c->val[DMRI_MAT_R].p, c->val[DMRI_MAT_JX].p,
                                  c->val[DMRI_MAT_JY].p, c->val[DMRI_MAT_JZ].p, miss.p, s);
        DMRI_HIP(hipStreamSynchronize(s));
    }
    c->ncol_iface = 0;
    if (have_iface) {
        DevArray<int> perm;
        std::vector<int64_t> offs = colour_elements(3, c->iface_v.p, c->n_iface, c->fptr.p, c->flist.p, perm, s);
        c->ncol_iface = (int)offs.size() - 1;
        for (int k = 0; k < c->ncol_iface; ++k)
            launch_assemble_facets(offs[k], offs[k + 1] - offs[k], perm.p, c->iface_v.p, c->xyz.p, nv, c->dof_of.p,
                                   c->rowptr.p, c->col.p, c->kappa, c->has_kappa_f ? c->kappa_f.p : nullptr,
                                   c->val[DMRI_MAT_I].p, miss.p, s);
        DMRI_HIP(hipStreamSynchronize(s));
    }
    require(read_scalar(miss.p, s) == 0, DMRI_ERR_MESH, "element entry outside the CSR pattern");

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
        c->part.alloc(std::max<size_t>(c->part.n, 4 * (size_t)kVecBlocks + 16));
        c->counters.alloc(8);
        c->counters.zero(s);
        c->dsig.alloc(4);
        launch_sum(c->w.p, c->ndof, c->part.p, c->counters.p, c->dsig.p, s);
        c->volume = read_scalar(c->dsig.p, s);
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

void ensure_solver_workspace(dmri_ctx *c) {
    hipStream_t s = c->stream;
    const size_t n = c->ndof;
    if (c->x.n != n) {
        c->x.alloc(n);
        c->r.alloc(n);
        c->rh.alloc(n);
        c->p.alloc(n);
        c->v.alloc(n);
        c->t.alloc(n);
        c->sc.alloc(n);
    }
    if (c->aj.n != (size_t)c->nnz) {
        c->aj.alloc(c->nnz);
        c->bre.alloc(c->nnz);
    }
    const size_t nblk = (n + kRowsPerBlock - 1) / kRowsPerBlock;
    const size_t need = 4 * std::max<size_t>(nblk, kVecBlocks) + 16;
    if (c->part.n < need) c->part.alloc(need);
    if (!c->st.p) {
        c->st.alloc(1);
        c->st.zero(s);
    }
    if (!c->h_st) DMRI_HIP(hipHostMalloc((void **)&c->h_st, sizeof(KrylovState), hipHostMallocDefault));
    if (!c->events_made) {
        for (int i = 0; i < kEventPool; ++i) {
            DMRI_HIP(hipEventCreate(&c->ev0[i]));
            DMRI_HIP(hipEventCreate(&c->ev1[i]));
        }
        c->events_made = true;
    }
}

void solve_impl(dmri_ctx *c, const dmri_seq *seq, const dmri_solver_opts *opts, double *sig, dmri_stats *stats) {
    require(c->assembled, DMRI_ERR_STATE, "dmri_solve: call dmri_assemble first");
    require(seq != nullptr && opts != nullptr && sig != nullptr, DMRI_ERR_ARG, "dmri_solve: null argument");
    require(seq->nsteps >= 0, DMRI_ERR_ARG, "nsteps < 0");
    require(seq->nsteps == 0 || (seq->f_implicit && seq->f_explicit), DMRI_ERR_ARG, "null time profile");
    require(seq->dt > 0 && std::isfinite(seq->dt), DMRI_ERR_ARG, "dt must be > 0");
    require(seq->theta >= 0.0 && seq->theta <= 1.0, DMRI_ERR_ARG, "theta must lie in [0, 1]");
    require(std::isfinite(seq->q), DMRI_ERR_ARG, "q is not finite");
    const double gn = std::sqrt(seq->gdir[0] * seq->gdir[0] + seq->gdir[1] * seq->gdir[1] + seq->gdir[2] * seq->gdir[2]);
    // |g| = 0 is rejected by the reference (set_gradient_dir, DmriFemLib.py:813-823)
    require(gn > 1e-10, DMRI_ERR_ARG, "|g|=0: gradient direction must be non-zero");
    require(opts->maxit > 0, DMRI_ERR_ARG, "maxit must be > 0");
    require(opts->rtol >= 0 && opts->atol >= 0 && (opts->rtol > 0 || opts->atol > 0), DMRI_ERR_ARG,
            "need rtol > 0 or atol > 0");
    for (int i = 0; i < seq->nsteps; ++i)
        require(std::isfinite(seq->f_implicit[i]) && std::isfinite(seq->f_explicit[i]), DMRI_ERR_ARG,
                "time profile has a non-finite value");
    set_device(c);
    ensure_solver_workspace(c);
    hipStream_t s = c->stream;
    const int64_t n = c->ndof;
    const double g[3] = {seq->gdir[0] / gn, seq->gdir[1] / gn, seq->gdir[2] / gn};
    const double theta = seq->theta, dt = seq->dt, q = seq->q;
    auto &V = c->val;

    launch_prep_scale(n, c->diag.p, V[DMRI_MAT_M].p, V[DMRI_MAT_K].p, V[DMRI_MAT_R].p, V[DMRI_MAT_I].p, dt, theta,
                      c->sc.p, s);
    launch_prep_values(n, c->rowptr.p, c->col.p, c->sc.p, V[DMRI_MAT_M].p, V[DMRI_MAT_K].p, V[DMRI_MAT_R].p,
                       V[DMRI_MAT_I].p, V[DMRI_MAT_JX].p, V[DMRI_MAT_JY].p, V[DMRI_MAT_JZ].p, g[0], g[1], g[2], dt,
                       theta, c->aj.p, c->bre.p, s);
    launch_init_x(n, c->sc.p, c->x.p, s);
    DMRI_HIP(hipStreamSynchronize(s));

    dmri_stats st{};
    st.spmv_bytes = (double)c->nnz * 20.0 + 4.0 * (double)(n + 1) + 48.0 * (double)n;
    const auto t0 = std::chrono::steady_clock::now();
    int chunk = 8;
    const bool prof = opts->profile != 0;
    for (int step = 0; step < seq->nsteps; ++step) {
        const double sA = theta * q * seq->f_implicit[step];
        const double sB = -(1.0 - theta) * q * seq->f_explicit[step];
        int step_iters = 0, restarts = 0;
        launch_rhs_residual(n, c->rowptr.p, c->col.p, c->aj.p, c->bre.p, sA, sB, c->x.p, c->r.p, c->rh.p, opts->rtol,
                            opts->atol, opts->maxit, c->part.p, c->st.p, s);
        int iter_before = 0;
        for (;;) {
            const int nl = std::min(chunk, kEventPool / 2);
            for (int k = 0; k < nl; ++k) {
                launch_bicg_p(n, c->r.p, c->p.p, c->v.p, c->st.p, s);
                if (prof) DMRI_HIP(hipEventRecord(c->ev0[2 * k], s));
                launch_spmv_v(n, c->rowptr.p, c->col.p, c->aj.p, sA, c->p.p, c->v.p, c->rh.p, c->part.p, c->st.p, s);
                if (prof) {
                    DMRI_HIP(hipEventRecord(c->ev1[2 * k], s));
                    DMRI_HIP(hipEventRecord(c->ev0[2 * k + 1], s));
                }
                launch_spmv_t(n, c->rowptr.p, c->col.p, c->aj.p, sA, c->r.p, c->v.p, c->t.p, c->part.p, c->st.p, s);
                if (prof) DMRI_HIP(hipEventRecord(c->ev1[2 * k + 1], s));
                launch_bicg_update(n, c->x.p, c->p.p, c->r.p, c->v.p, c->t.p, c->rh.p, c->part.p, c->st.p, s);
            }
            DMRI_HIP(hipMemcpyAsync(c->h_st, c->st.p, sizeof(KrylovState), hipMemcpyDeviceToHost, s));
            DMRI_HIP(hipStreamSynchronize(s));
            const KrylovState &h = *c->h_st;
            if (prof) {
                const int real = std::min(nl, h.iter - iter_before);
                for (int k = 0; k < real; ++k)
                    for (int j = 0; j < 2; ++j) {
                        float ms = 0;
                        DMRI_HIP(hipEventElapsedTime(&ms, c->ev0[2 * k + j], c->ev1[2 * k + j]));
                        st.spmv_ms += ms;
                        st.spmv_launches += 1;
                    }
            }
            iter_before = h.iter;
            if (h.done) {
                if (h.breakdown && !(h.rnorm2 <= h.tol2)) {
                    ++restarts;
                    require(restarts <= 20, DMRI_ERR_NOCONV, "BiCGStab: repeated breakdown at step " +
                                                                 std::to_string(step));
                    step_iters += h.iter;
                    launch_rhs_residual(n, c->rowptr.p, c->col.p, c->aj.p, c->bre.p, sA, sB, c->x.p, c->r.p, c->rh.p,
                                        opts->rtol, opts->atol, opts->maxit, c->part.p, c->st.p, s);
                    iter_before = 0;
                    continue;
                }
                if (!(h.rnorm2 <= h.tol2))
                    throw Error(DMRI_ERR_NOCONV, "BiCGStab did not converge in " + std::to_string(opts->maxit) +
                                                     " iterations at step " + std::to_string(step) +
                                                     " (relres " + std::to_string(std::sqrt(h.rnorm2 / h.bnorm2)) +
                                                     ")");
                break;
            }
            if (h.iter >= opts->maxit)
                throw Error(DMRI_ERR_NOCONV, "BiCGStab did not converge at step " + std::to_string(step));
            chunk = std::max(2, chunk / 4);
        }
        const KrylovState &h = *c->h_st;
        step_iters += h.iter;
        st.iters_total += step_iters;
        st.iters_max = std::max(st.iters_max, step_iters);
        st.restarts += restarts;
        const double rel = h.bnorm2 > 0 ? std::sqrt(h.rnorm2 / h.bnorm2) : std::sqrt(h.rnorm2);
        st.relres_max = std::max(st.relres_max, rel);
        chunk = std::max(1, h.iter);
    }
    launch_signal(n, c->w.p, c->sc.p, c->x.p, c->part.p, c->counters.p + 1, c->dsig.p, s);
    double hs[2];
    DMRI_HIP(hipMemcpyAsync(hs, c->dsig.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    st.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    sig[0] = hs[0] / c->volume;
    sig[1] = hs[1] / c->volume;
    if (stats) *stats = st;
    c->solved = true;
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
        c->tets.upload(tets, 4 * nc, c->stream);
        if (cell_phase) c->phase.upload(cell_phase, nc, c->stream);
        else c->phase.release();
        DMRI_HIP(hipStreamSynchronize(c->stream));
        c->nv = nv;
        c->nc = nc;
        c->ncomp = cell_phase ? 2 : 1;
        c->hmin = hmin;
        c->hmax = hmax;
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
        if (c->has_Dcell) c->Dcell.upload(k->D_cell, 9 * c->nc, c->stream);
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
        require(which >= 0 && which < 7, DMRI_ERR_ARG, "unknown matrix id");
        require(c->assembled, DMRI_ERR_STATE, "not assembled");
        set_device(c);
        DMRI_HIP(hipMemcpyAsync(val, c->val[which].p, sizeof(double) * c->nnz, hipMemcpyDeviceToHost, c->stream));
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

int dmri_solve(dmri_ctx *c, const dmri_seq *seq, const dmri_solver_opts *opts, double sig_out[2],
               dmri_stats *stats) {
    return guarded([&] {
        require(c != nullptr, DMRI_ERR_ARG, "null context");
        solve_impl(c, seq, opts, sig_out, stats);
    });
}

int dmri_get_solution(dmri_ctx *c, double *u) {
    return guarded([&] {
        require(c != nullptr && u != nullptr, DMRI_ERR_ARG, "null argument");
        require(c->solved, DMRI_ERR_STATE, "no solution (call dmri_solve first)");
        set_device(c);
        DevArray<double2> tmp;
        tmp.alloc(c->ndof);
        launch_unscale(c->ndof, c->sc.p, c->x.p, tmp.p, c->stream);
        DMRI_HIP(hipMemcpyAsync(u, tmp.p, sizeof(double2) * c->ndof, hipMemcpyDeviceToHost, c->stream));
        DMRI_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
