/*
 * dmri_oracle.c -- CPU FP64 restatement of the reference's Bloch-Torrey FEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py may load it.  The product path
 * (libdmri.so, HIP kernels) never calls, links or falls back to it.
 *
 * What it restates (reference files, van-dang/DMRI-FEM-Cloud):
 *   - DOF layout: mixed P1 space, 2 (1-compartment) or 4 (2-compartment) real
 *     components per vertex; DmriFemLib.py:453-486, comri/two-comp/hpc-fenics-cpp/
 *     ufc/Bloch_Torrey_NoTime3D.cpp:6050-6397 (dof = comp*Nv + vertex).  Here the
 *     real/imag pair is one complex unknown and dead compartment DOFs (vertices no
 *     cell of that compartment touches; M.ident_zeros() at DmriFemLib.py:246) are
 *     dropped, which leaves the live solution unchanged.
 *   - Sparsity: union of element cliques (DOLFIN's sparsity builder; the reference
 *     relies on third-party code here, so the definition is the brute-force set union).
 *   - Element matrices (closed forms of the P1 integrals in the forms of
 *     DmriFemLib.py:41-50,79-93,125-145,240-254 and comri/one-comp/hpc-fenics-cpp/ufc/
 *     Bloch_Torrey_NoTime3D.ufl): mass, stiffness with scalar or DG0 tensor D, T2
 *     relaxation, gradient moments Jx/Jy/Jz, interface facet mass on phase jumps.
 *   - Weak pseudo-periodic boundary condition (WeakPseudoPeridicMarker, DmriFemLib.py:601-610;
 *     ThetaMethodF/L_wBC1c/2c 58-145; WeakPseudoPeriodic_1c/2c eval 256-324,386-450): the
 *     boundary mass P with the P1-interpolated face indicator, and the mirror map of u^n with its
 *     phase exp(i q (g.l) F(t_{n-1})), located by brute force over the opposite face's facets.
 *   - Theta time stepping (DmriFemLib.py:878-915): A_n u^{n+1} = B_n u^n with
 *     A_n = M/k + theta*L + i*theta*q*fA[n]*Jg, B_n = M/k - (1-theta)*L - i*(1-theta)*q*fB[n]*Jg,
 *     L = K + R + I, solved by BiCGStab with Jacobi (GCloudDmriSolver.py:219-222,
 *     comri/one-comp/hpc-fenics-cpp/main.cpp:232).
 *
 * Every parallel loop uses a static partition and every reduction sums fixed
 * chunks in a fixed order, so results do not depend on the thread count.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define ORC_CHUNK 4096

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* DOF map.  dof_of[c*nv + v] = DOF index of (compartment c, vertex v) or -1.   */
/* Numbering is compartment-major, increasing vertex index inside a compartment */
/* (the DOLFIN-HPC serial layout, Bloch_Torrey_NoTime3D.cpp:6050-6397).        */
/* ------------------------------------------------------------------------- */
int64_t orc_dof_map(int64_t nv, int64_t nc, const int32_t *tets, const int32_t *phase,
                    int ncomp, int32_t *dof_of) {
    memset(dof_of, 0, sizeof(int32_t) * (size_t)(ncomp * nv));
    for (int64_t t = 0; t < nc; ++t) {
        int c = phase ? phase[t] : 0;
        if (c < 0 || c >= ncomp) return -1;
        for (int a = 0; a < 4; ++a) {
            int32_t v = tets[4 * t + a];
            if (v < 0 || v >= nv) return -2;
            dof_of[c * nv + v] = 1;
        }
    }
    int64_t d = 0;
    for (int c = 0; c < ncomp; ++c)
        for (int64_t v = 0; v < nv; ++v)
            dof_of[c * nv + v] = dof_of[c * nv + v] ? (int32_t)(d++) : -1;
    return d;
}

/* ------------------------------------------------------------------------- */
/* Faces.  Interface facets Gamma are the faces shared by two cells of         */
/* different phase (abs(jump(phase))*dS, DmriFemLib.py:112,139).  Boundary     */
/* facets are faces owned by exactly one cell.                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t v[3];
    int32_t cell;
} face_rec;

static int face_cmp(const void *pa, const void *pb) {
    const face_rec *a = (const face_rec *)pa, *b = (const face_rec *)pb;
    for (int i = 0; i < 3; ++i) {
        if (a->v[i] != b->v[i]) return a->v[i] < b->v[i] ? -1 : 1;
    }
    return a->cell < b->cell ? -1 : (a->cell > b->cell);
}

static void sort3(int32_t *v) {
    int32_t t;
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

static face_rec *collect_faces(int64_t nc, const int32_t *tets) {
    face_rec *f = (face_rec *)malloc(sizeof(face_rec) * (size_t)(4 * nc));
    if (!f) return NULL;
    for (int64_t t = 0; t < nc; ++t) {
        const int32_t *T = tets + 4 * t;
        for (int o = 0; o < 4; ++o) { /* face opposite local vertex o */
            face_rec *r = f + 4 * t + o;
            int k = 0;
            for (int a = 0; a < 4; ++a)
                if (a != o) r->v[k++] = T[a];
            sort3(r->v);
            r->cell = (int32_t)t;
        }
    }
    qsort(f, (size_t)(4 * nc), sizeof(face_rec), face_cmp);
    return f;
}

/* Interface facets sorted by vertex triple.  fv[3*i..] vertices (ascending),
 * fc[2*i] = cell on phase 0, fc[2*i+1] = cell on phase 1.  Returns the count
 * (writes at most cap records; call with cap=0 to size).                     */
int64_t orc_interface_facets(int64_t nc, const int32_t *tets, const int32_t *phase,
                             int32_t *fv, int32_t *fc, int64_t cap) {
    if (!phase) return 0;
    face_rec *f = collect_faces(nc, tets);
    if (!f) return -1;
    int64_t n = 0;
    for (int64_t i = 0; i + 1 < 4 * nc; ++i) {
        if (f[i].v[0] == f[i + 1].v[0] && f[i].v[1] == f[i + 1].v[1] && f[i].v[2] == f[i + 1].v[2]) {
            int32_t c0 = f[i].cell, c1 = f[i + 1].cell;
            if (phase[c0] != phase[c1]) {
                if (n < cap) {
                    memcpy(fv + 3 * n, f[i].v, sizeof(int32_t) * 3);
                    fc[2 * n] = phase[c0] == 0 ? c0 : c1;
                    fc[2 * n + 1] = phase[c0] == 0 ? c1 : c0;
                }
                ++n;
            }
            ++i;
        }
    }
    free(f);
    return n;
}

/* Boundary (exterior) facets sorted by vertex triple; fc[i] = owning cell.    */
int64_t orc_boundary_facets(int64_t nc, const int32_t *tets, int32_t *fv, int32_t *fc, int64_t cap) {
    face_rec *f = collect_faces(nc, tets);
    if (!f) return -1;
    int64_t n = 0, m = 4 * nc;
    for (int64_t i = 0; i < m;) {
        int64_t j = i + 1;
        while (j < m && f[j].v[0] == f[i].v[0] && f[j].v[1] == f[i].v[1] && f[j].v[2] == f[i].v[2]) ++j;
        if (j - i == 1) {
            if (n < cap) {
                memcpy(fv + 3 * n, f[i].v, sizeof(int32_t) * 3);
                fc[n] = f[i].cell;
            }
            ++n;
        }
        i = j;
    }
    free(f);
    return n;
}

/* ------------------------------------------------------------------------- */
/* Weak pseudo-periodic boundary condition.                                     */
/*   P_ij += kappa_s * sum_k chi_k int_F phi_k phi_i phi_j over every boundary   */
/*   facet F, in the compartment of F's cell (the (1-phase)*ds and phase*ds of   */
/*   ThetaMethodF_wBC2c, DmriFemLib.py:143), kappa_s = 3e-3/hmin (MyDomain 599), */
/*   chi_k = vertex k lies within eps = 1e-2*hmin of a periodic face (the degree-1 */
/*   Expression of WeakPseudoPeridicMarker 601-610, interpolated at vertices),   */
/*   int_F l_k l_i l_j = |F|/60 (1+d_ij)(1+d_ik+d_jk).                            */
/*   Mirror entries: every live DOF whose vertex lies within 1e-7 of a periodic   */
/*   face (eval's tol, 256-324) takes u^n at the point with that coordinate moved */
/*   to the opposite face; later axes override earlier ones (the eval's order);   */
/*   l = mirror - x.  u^n there is the P1 interpolant on the facet of the         */
/*   opposite face with the largest smallest barycentric coordinate.              */
/* Returns the number of entries (<= cap written), or -1 (allocation), -2 (a     */
/* mirror point with no facet on its face).  pv (nnz) is accumulated into.       */
/* ------------------------------------------------------------------------- */
static double tri_area(const double *A, const double *B, const double *C) {
    double e1[3], e2[3];
    for (int q = 0; q < 3; ++q) { e1[q] = B[q] - A[q]; e2[q] = C[q] - A[q]; }
    double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    return 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

static int64_t csr_find64(const int64_t *rowptr, const int32_t *col, int64_t r, int32_t c);

int64_t orc_periodic(int64_t nv, const double *xyz, int64_t nc, const int32_t *tets, const int32_t *phase,
                     const int32_t *dof_of, const int32_t *pdir, double hmin, const int64_t *rowptr,
                     const int32_t *col, double *pv, int32_t *ent_dof, int32_t *ent_src, double *ent_w,
                     double *ent_ell, int64_t cap) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t v = 0; v < nv; ++v)
        for (int d = 0; d < 3; ++d) {
            if (xyz[3 * v + d] < lo[d]) lo[d] = xyz[3 * v + d];
            if (xyz[3 * v + d] > hi[d]) hi[d] = xyz[3 * v + d];
        }
    const double ks = 3e-3 / hmin, eps = 1e-2 * hmin, tol = 1e-7;
    const int ncomp = phase ? 2 : 1;
    int32_t *fv = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(4 * nc));
    int32_t *fc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(4 * nc));
    if (!fv || !fc) { free(fv); free(fc); return -1; }
    int64_t nbf = orc_boundary_facets(nc, tets, fv, fc, 4 * nc);
    /* boundary mass P */
    for (int64_t f = 0; f < nbf; ++f) {
        int c = phase ? phase[fc[f]] : 0;
        int32_t dd[3];
        double chi[3], ssum = 0;
        for (int a = 0; a < 3; ++a) {
            const double *X = xyz + 3 * (int64_t)fv[3 * f + a];
            int on = 0;
            for (int d = 0; d < 3; ++d)
                if (pdir[d] && (X[d] < lo[d] + eps || X[d] > hi[d] - eps)) on = 1;
            chi[a] = on;
            ssum += on;
            dd[a] = dof_of[(int64_t)c * nv + fv[3 * f + a]];
        }
        if (ssum == 0) continue;
        double area = tri_area(xyz + 3 * (int64_t)fv[3 * f], xyz + 3 * (int64_t)fv[3 * f + 1],
                               xyz + 3 * (int64_t)fv[3 * f + 2]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += chi[k] * (1.0 + (a == b)) * (1.0 + (a == k) + (b == k));
                pv[csr_find64(rowptr, col, dd[a], dd[b])] += ks * area / 60.0 * s;
            }
    }
    /* mirror entries, in DOF order (compartment-major, increasing vertex) */
    int64_t ne = 0;
    int rc = 0;
    for (int c = 0; c < ncomp && !rc; ++c)
        for (int64_t v = 0; v < nv; ++v) {
            int32_t i = dof_of[(int64_t)c * nv + v];
            if (i < 0) continue;
            const double *X = xyz + 3 * v;
            int td = -1;
            double tgt = 0;
            for (int d = 0; d < 3; ++d) {
                if (!pdir[d]) continue;
                if (fabs(X[d] - lo[d]) <= tol) { td = d; tgt = hi[d]; }
                if (fabs(X[d] - hi[d]) <= tol) { td = d; tgt = lo[d]; }
            }
            if (td < 0) continue;
            const int u = (td + 1) % 3, w = (td + 2) % 3;
            const double pu = X[u], pw = X[w];
            double best = -1e300, bl[3] = {0, 0, 0};
            int64_t bf = -1;
            for (int64_t f = 0; f < nbf; ++f) {
                const double *A = xyz + 3 * (int64_t)fv[3 * f], *B = xyz + 3 * (int64_t)fv[3 * f + 1],
                             *C = xyz + 3 * (int64_t)fv[3 * f + 2];
                if (fabs(A[td] - tgt) > eps || fabs(B[td] - tgt) > eps || fabs(C[td] - tgt) > eps) continue;
                double det = (B[u] - A[u]) * (C[w] - A[w]) - (C[u] - A[u]) * (B[w] - A[w]);
                if (det == 0) continue;
                double l1 = ((pu - A[u]) * (C[w] - A[w]) - (C[u] - A[u]) * (pw - A[w])) / det;
                double l2 = ((B[u] - A[u]) * (pw - A[w]) - (pu - A[u]) * (B[w] - A[w])) / det;
                double l0 = 1.0 - l1 - l2;
                double m = l0 < l1 ? l0 : l1;
                m = m < l2 ? m : l2;
                if (m > best) { best = m; bf = f; bl[0] = l0; bl[1] = l1; bl[2] = l2; }
            }
            if (bf < 0) { rc = -2; break; }
            if (ne < cap) {
                ent_dof[ne] = i;
                for (int a = 0; a < 3; ++a) {
                    ent_src[3 * ne + a] = dof_of[(int64_t)c * nv + fv[3 * bf + a]];
                    ent_w[3 * ne + a] = bl[a];
                    ent_ell[3 * ne + a] = a == td ? tgt - X[td] : 0.0;
                }
            }
            ++ne;
        }
    free(fv);
    free(fc);
    return rc ? rc : ne;
}

/* ------------------------------------------------------------------------- */
/* Sparsity pattern: set union of the DOF cliques of every cell (4 DOFs in its  */
/* own compartment) and every interface facet (3 vertices x 2 compartments).   */
/* Keys row*ndof+col are radix-sorted and made unique.                          */
/* ------------------------------------------------------------------------- */
static void radix_sort_u64(uint64_t *a, uint64_t *tmp, size_t n, int nbits) {
    size_t cnt[65536];
    for (int shift = 0; shift < nbits; shift += 16) {
        memset(cnt, 0, sizeof(cnt));
        for (size_t i = 0; i < n; ++i) cnt[(a[i] >> shift) & 0xFFFF]++;
        size_t s = 0;
        for (int d = 0; d < 65536; ++d) { size_t c = cnt[d]; cnt[d] = s; s += c; }
        for (size_t i = 0; i < n; ++i) tmp[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
        memcpy(a, tmp, n * sizeof(uint64_t));
    }
}

int64_t orc_pattern(int64_t ndof, int64_t nv, int64_t nc, const int32_t *tets, const int32_t *phase,
                    const int32_t *dof_of, int64_t nf, const int32_t *fv,
                    int64_t *rowptr, int32_t *col) {
    size_t npairs = (size_t)(16 * nc + 36 * nf);
    uint64_t *keys = (uint64_t *)malloc(npairs * sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)malloc(npairs * sizeof(uint64_t));
    if (!keys || !tmp) { free(keys); free(tmp); return -1; }
    size_t p = 0;
    for (int64_t t = 0; t < nc; ++t) {
        int c = phase ? phase[t] : 0;
        int32_t d[4];
        for (int a = 0; a < 4; ++a) d[a] = dof_of[c * nv + tets[4 * t + a]];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) keys[p++] = (uint64_t)d[a] * (uint64_t)ndof + (uint64_t)d[b];
    }
    for (int64_t f = 0; f < nf; ++f) {
        int32_t d[6];
        for (int a = 0; a < 3; ++a) {
            d[a] = dof_of[0 * nv + fv[3 * f + a]];
            d[3 + a] = dof_of[1 * nv + fv[3 * f + a]];
        }
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) keys[p++] = (uint64_t)d[a] * (uint64_t)ndof + (uint64_t)d[b];
    }
    int nbits = 0;
    uint64_t maxkey = (uint64_t)ndof * (uint64_t)ndof;
    while (nbits < 64 && (maxkey >> nbits)) nbits += 16;
    radix_sort_u64(keys, tmp, p, nbits);
    free(tmp);
    for (int64_t r = 0; r <= ndof; ++r) rowptr[r] = 0;
    int64_t nnz = 0;
    for (size_t i = 0; i < p; ++i) {
        if (i > 0 && keys[i] == keys[i - 1]) continue;
        int64_t r = (int64_t)(keys[i] / (uint64_t)ndof);
        if (col) col[nnz] = (int32_t)(keys[i] % (uint64_t)ndof);
        rowptr[r + 1]++;
        ++nnz;
    }
    for (int64_t r = 0; r < ndof; ++r) rowptr[r + 1] += rowptr[r];
    free(keys);
    return nnz;
}

/* ------------------------------------------------------------------------- */
/* Closed-form P1 element matrices.                                            */
/*   int_T l_a l_b       = |T|(1+d_ab)/20                                      */
/*   int_T l_a l_b l_c   = |T|/20, /60, /120 for 3, 2, 0 equal indices          */
/*   => Jx_ab = |T|/120 (1+d_ab) (sum_c x_c + x_a + x_b)                       */
/*   K_ab = |T| grad(l_a) . D grad(l_b)                                        */
/*   int_F l_a l_b       = |F|(1+d_ab)/12                                      */
/* ------------------------------------------------------------------------- */
static double tet_geom(const double X[4][3], double G[4][3]) {
    double e[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) e[i][k] = X[i + 1][k] - X[0][k];
    double c23[3] = {e[1][1] * e[2][2] - e[1][2] * e[2][1], e[1][2] * e[2][0] - e[1][0] * e[2][2],
                     e[1][0] * e[2][1] - e[1][1] * e[2][0]};
    double c31[3] = {e[2][1] * e[0][2] - e[2][2] * e[0][1], e[2][2] * e[0][0] - e[2][0] * e[0][2],
                     e[2][0] * e[0][1] - e[2][1] * e[0][0]};
    double c12[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                     e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    double det = e[0][0] * c23[0] + e[0][1] * c23[1] + e[0][2] * c23[2];
    for (int k = 0; k < 3; ++k) {
        G[1][k] = c23[k] / det;
        G[2][k] = c31[k] / det;
        G[3][k] = c12[k] / det;
        G[0][k] = -(G[1][k] + G[2][k] + G[3][k]);
    }
    return fabs(det) / 6.0;
}

static int64_t csr_find(const int64_t *rowptr, const int32_t *col, int64_t r, int32_t c);
static int64_t csr_find64(const int64_t *rowptr, const int32_t *col, int64_t r, int32_t c) {
    return csr_find(rowptr, col, r, c);
}
static int64_t csr_find(const int64_t *rowptr, const int32_t *col, int64_t r, int32_t c) {
    int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        if (col[mid] == c) return mid;
        if (col[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* Assemble the time-independent matrices on the DOF pattern.
 *   m  : mass                      (MassMatrix1c/2c, DmriFemLib.py:240-254)
 *   k  : stiffness, D per cell tensor (Dcell, row-major 3x3) or per-compartment scalar Dc
 *   r  : T2 relaxation, invT2 per cell or per compartment (FuncF_wBC, DmriFemLib.py:43-44)
 *   gi : permeable interface kappa*int_F (u0-u1)(v0-v1) (icondition_wBC, DmriFemLib.py:47-50)
 *   jx,jy,jz : int_T x_d phi_i phi_j  (GdotX, DmriFemLib.py:33-39; Jg = g.(Jx,Jy,Jz))
 * kappa_f (per interface facet) overrides kappa when non-NULL.
 * Returns 0, or a negative code when an entry falls outside the pattern.     */
int orc_assemble(int64_t nv, const double *xyz, int64_t nc, const int32_t *tets, const int32_t *phase,
                 const int32_t *dof_of, const double *Dc, const double *Dcell, const double *invT2c,
                 const double *invT2cell, double kappa, const double *kappa_f, int64_t nf, const int32_t *fv,
                 const int64_t *rowptr, const int32_t *col, int64_t nnz, double *m, double *k, double *r,
                 double *gi, double *jx, double *jy, double *jz) {
    double *outs[7] = {m, k, r, gi, jx, jy, jz};
    for (int o = 0; o < 7; ++o) memset(outs[o], 0, sizeof(double) * (size_t)nnz);
    for (int64_t t = 0; t < nc; ++t) {
        int c = phase ? phase[t] : 0;
        double X[4][3], G[4][3];
        int32_t d[4];
        for (int a = 0; a < 4; ++a) {
            int32_t v = tets[4 * t + a];
            d[a] = dof_of[c * nv + v];
            for (int q = 0; q < 3; ++q) X[a][q] = xyz[3 * v + q];
        }
        double vol = tet_geom(X, G);
        double D[9];
        if (Dcell) memcpy(D, Dcell + 9 * t, sizeof(D));
        else { memset(D, 0, sizeof(D)); D[0] = D[4] = D[8] = Dc[c]; }
        double it2 = invT2cell ? invT2cell[t] : invT2c[c];
        double S[3] = {0, 0, 0};
        for (int a = 0; a < 4; ++a)
            for (int q = 0; q < 3; ++q) S[q] += X[a][q];
        for (int a = 0; a < 4; ++a) {
            for (int b = 0; b < 4; ++b) {
                /* row = test function phi_a, column = trial phi_b: inner(D grad(u), grad(v)) gives
                   K_ab = grad(phi_a) . D grad(phi_b) (FuncF_wBC, DmriFemLib.py:43-44) */
                double DG[3];
                for (int q = 0; q < 3; ++q) DG[q] = D[3 * q] * G[b][0] + D[3 * q + 1] * G[b][1] + D[3 * q + 2] * G[b][2];
                int64_t pos = csr_find(rowptr, col, d[a], d[b]);
                if (pos < 0) return -3;
                double dab = (a == b) ? 2.0 : 1.0;
                double mab = vol * dab / 20.0;
                m[pos] += mab;
                r[pos] += it2 * mab;
                k[pos] += vol * (G[a][0] * DG[0] + G[a][1] * DG[1] + G[a][2] * DG[2]);
                double w = vol * dab / 120.0;
                jx[pos] += w * (S[0] + X[a][0] + X[b][0]);
                jy[pos] += w * (S[1] + X[a][1] + X[b][1]);
                jz[pos] += w * (S[2] + X[a][2] + X[b][2]);
            }
        }
    }
    for (int64_t f = 0; f < nf; ++f) {
        double X[3][3];
        int32_t d0[3], d1[3];
        for (int a = 0; a < 3; ++a) {
            int32_t v = fv[3 * f + a];
            d0[a] = dof_of[v];
            d1[a] = dof_of[nv + v];
            for (int q = 0; q < 3; ++q) X[a][q] = xyz[3 * v + q];
        }
        double e1[3], e2[3];
        for (int q = 0; q < 3; ++q) { e1[q] = X[1][q] - X[0][q]; e2[q] = X[2][q] - X[0][q]; }
        double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
        double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        double kap = kappa_f ? kappa_f[f] : kappa;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double val = kap * area * ((a == b) ? 2.0 : 1.0) / 12.0;
                int64_t p00 = csr_find(rowptr, col, d0[a], d0[b]);
                int64_t p11 = csr_find(rowptr, col, d1[a], d1[b]);
                int64_t p01 = csr_find(rowptr, col, d0[a], d1[b]);
                int64_t p10 = csr_find(rowptr, col, d1[a], d0[b]);
                if (p00 < 0 || p11 < 0 || p01 < 0 || p10 < 0) return -4;
                gi[p00] += val;
                gi[p11] += val;
                gi[p01] -= val;
                gi[p10] -= val;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Strong pseudo-periodic formulation (IsDomainPeriodic = True): the terms of   */
/* FuncF_sBC, outer_interface and inner_interface (DmriFemLib.py:148-165) that   */
/* carry the gradient direction, per axis d on the unconstrained DOF pattern,   */
/* for a scalar D per compartment:                                              */
/*   c_d[ij] = int 2 D_c (d phi_j / d x_d) phi_i   (inner(g, D grad u) +         */
/*             inner(grad u, D g) tested with v, FuncF_sBC 162-165);             */
/*   b_d[ij] = int D_c n_d phi_i phi_j over every exterior facet of a cell of     */
/*             compartment c (outer_interface 157-159, n the outward normal) and, */
/*             for two compartments, every interface facet with n the outward     */
/*             normal of compartment c's cell: inner_interface (149-155) reduces  */
/*             to -i q F D_0 (g.n0) u0 v0 + i q F D_1 (g.n0) u1 v1 with n0 the    */
/*             outward normal of the phase-0 cell, the outer term of each side.   */
/* The reference integrates each compartment's outer term over every exterior   */
/* facet, also those of the other compartment's cells, where that field has     */
/* only dead DOFs but for the interface vertices; here it is taken over the      */
/* compartment's own facets.                                                    */
/* In the scheme the operator adds i theta q F (c - b) . g.                     */
/* Returns 0, or a negative code when an entry falls outside the pattern.       */
/* ------------------------------------------------------------------------- */
static void outward_normal(const double X[4][3], int o, double n[3], double *area) {
    /* facet opposite local vertex o of a cell with vertices X */
    int k = 0, idx[3];
    for (int a = 0; a < 4; ++a)
        if (a != o) idx[k++] = a;
    double e1[3], e2[3];
    for (int q = 0; q < 3; ++q) {
        e1[q] = X[idx[1]][q] - X[idx[0]][q];
        e2[q] = X[idx[2]][q] - X[idx[0]][q];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    *area = 0.5 * len;
    double dir = 0;
    for (int q = 0; q < 3; ++q) dir += n[q] * (X[o][q] - X[idx[0]][q]);
    double sg = dir > 0 ? -1.0 : 1.0; /* away from the opposite vertex */
    for (int q = 0; q < 3; ++q) n[q] *= sg / len;
}

static int add_facet_normal_mass(const int64_t *rowptr, const int32_t *col, const int32_t dd[3], double Dc,
                                 const double n[3], double area, double *bx, double *by, double *bz) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            int64_t pos = csr_find(rowptr, col, dd[a], dd[b]);
            if (pos < 0) return -1;
            double w = Dc * area * ((a == b) ? 2.0 : 1.0) / 12.0;
            bx[pos] += w * n[0];
            by[pos] += w * n[1];
            bz[pos] += w * n[2];
        }
    return 0;
}

int orc_assemble_sbc(int64_t nv, const double *xyz, int64_t nc, const int32_t *tets, const int32_t *phase,
                     const int32_t *dof_of, const double *Dc, const int64_t *rowptr, const int32_t *col, int64_t nnz,
                     double *cx, double *cy, double *cz, double *bx, double *by, double *bz) {
    double *outs[6] = {cx, cy, cz, bx, by, bz};
    for (int o = 0; o < 6; ++o) memset(outs[o], 0, sizeof(double) * (size_t)nnz);
    for (int64_t t = 0; t < nc; ++t) {
        int c = phase ? phase[t] : 0;
        double X[4][3], G[4][3];
        int32_t d[4];
        for (int a = 0; a < 4; ++a) {
            int32_t v = tets[4 * t + a];
            d[a] = dof_of[c * nv + v];
            for (int q = 0; q < 3; ++q) X[a][q] = xyz[3 * v + q];
        }
        double vol = tet_geom(X, G);
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                int64_t pos = csr_find(rowptr, col, d[a], d[b]);
                if (pos < 0) return -3;
                /* int phi_a = vol / 4, grad phi_b constant on the cell */
                cx[pos] += 2.0 * Dc[c] * G[b][0] * vol / 4.0;
                cy[pos] += 2.0 * Dc[c] * G[b][1] * vol / 4.0;
                cz[pos] += 2.0 * Dc[c] * G[b][2] * vol / 4.0;
            }
    }
    /* exterior facets (each belongs to one cell) and interface facets (both sides) */
    face_rec *f = collect_faces(nc, tets);
    if (!f) return -1;
    int64_t m = 4 * nc;
    int rc = 0;
    for (int64_t i = 0; i < m && !rc;) {
        int64_t j = i + 1;
        while (j < m && f[j].v[0] == f[i].v[0] && f[j].v[1] == f[i].v[1] && f[j].v[2] == f[i].v[2]) ++j;
        int take = (j - i == 1) || (j - i == 2 && phase && phase[f[i].cell] != phase[f[i + 1].cell]);
        for (int64_t k = i; take && k < j && !rc; ++k) {
            int64_t t = f[k].cell;
            int c = phase ? phase[t] : 0;
            double X[4][3];
            int o = -1;
            for (int a = 0; a < 4; ++a) {
                int32_t v = tets[4 * t + a];
                for (int q = 0; q < 3; ++q) X[a][q] = xyz[3 * v + q];
                if (v != f[k].v[0] && v != f[k].v[1] && v != f[k].v[2]) o = a;
            }
            double n[3], area;
            outward_normal(X, o, n, &area);
            int32_t dd[3];
            for (int a = 0; a < 3; ++a) dd[a] = dof_of[c * nv + f[k].v[a]];
            if (add_facet_normal_mass(rowptr, col, dd, Dc[c], n, area, bx, by, bz)) rc = -4;
        }
        i = j;
    }
    free(f);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Linear algebra helpers (complex vectors, real CSR value arrays).            */
/* ------------------------------------------------------------------------- */

/* y = (va + i*s*vj) x */
static void spmv_c(int64_t n, const int64_t *rowptr, const int32_t *col, const double *va, const double *vj,
                   double s, const cplx *x, cplx *y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        cplx accA = 0, accJ = 0;
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            cplx xv = x[col[p]];
            accA += va[p] * xv;
            accJ += vj[p] * xv;
        }
        y[i] = accA + I * s * accJ;
    }
}

/* deterministic x^H y */
static cplx dotc(int64_t n, const cplx *x, const cplx *y) {
    int64_t nch = (n + ORC_CHUNK - 1) / ORC_CHUNK;
    cplx *part = (cplx *)malloc(sizeof(cplx) * (size_t)(nch > 0 ? nch : 1));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nch; ++c) {
        cplx s = 0;
        int64_t e = (c + 1) * ORC_CHUNK < n ? (c + 1) * ORC_CHUNK : n;
        for (int64_t i = c * ORC_CHUNK; i < e; ++i) s += conj(x[i]) * y[i];
        part[c] = s;
    }
    cplx s = 0;
    for (int64_t c = 0; c < nch; ++c) s += part[c];
    free(part);
    return s;
}

static double nrm2(int64_t n, const cplx *x) { return sqrt(creal(dotc(n, x, x))); }

/* ------------------------------------------------------------------------- */
/* Theta-scheme time stepping (MRI_simulation.solve, DmriFemLib.py:878-915;    */
/* comri/one-comp/hpc-fenics-cpp/main.cpp:296-328).                            */
/*   a  = m/dt + theta*l                      (time independent)               */
/*   A_n = a + i*theta*q*fA[n]*jg                                              */
/*   b_n = (m/dt - (1-theta)*l) u^n - i*(1-theta)*q*fB[n]*jg u^n               */
/* Krylov: BiCGStab, Jacobi preconditioner diag(a) -- the diagonal of the real  */
/* 2x2-block operator PETSc sees (the J coupling is off the block diagonal).    */
/* Stop when ||r|| <= max(rtol*||b||, atol).  u (2n doubles, interleaved      */
/* re/im) is the initial condition on entry and u^N on exit.                   */
/* stats[0] = total iterations, stats[1] = max iterations in one step,         */
/* stats[2] = max final relative residual, stats[3] = breakdown restarts.      */
/* Returns 0, or -1 when a step did not converge within maxit.                 */
/* ------------------------------------------------------------------------- */
/* Theta time stepping with the weak pseudo-periodic terms: a = M/k + theta (L + P), and per step
 * b += (1-theta) P ybc with ybc[ent_dof[e]] = exp(i q (g.ell_e) FB[n]) sum_a w u^n[src] (ThetaMethodL
 * _wBC1c/2c, DmriFemLib.py:70-75,114-120).  pv == NULL: no periodic terms. */
int orc_theta_solve_p(int64_t n, const int64_t *rowptr, const int32_t *col, const double *m, const double *l,
                      const double *jg, double dt, double theta, double q, int nsteps, const double *fA,
                      const double *fB, double rtol, double atol, int maxit, double *u_io, double *stats,
                      const double *pv, int64_t ne, const int32_t *ent_dof, const int32_t *ent_src,
                      const double *ent_w, const double *ent_ell, const double *g, const double *FB) {
    cplx *u = (cplx *)u_io;
    int64_t nnz = rowptr[n];
    double *a = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *bre = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *dinv = (double *)malloc(sizeof(double) * (size_t)n);
    cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)n), *r = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *rh = (cplx *)malloc(sizeof(cplx) * (size_t)n), *p = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *v = (cplx *)malloc(sizeof(cplx) * (size_t)n), *s = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *t = (cplx *)malloc(sizeof(cplx) * (size_t)n), *ph = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *sh = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    int rc = 0;
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nnz; ++e) {
        a[e] = m[e] / dt + theta * (l[e] + (pv ? pv[e] : 0.0));
        bre[e] = m[e] / dt - (1.0 - theta) * l[e];
    }
    cplx *ybc = pv ? (cplx *)calloc((size_t)n, sizeof(cplx)) : NULL;
    for (int64_t i = 0; i < n; ++i) {
        double d = 0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
            if (col[e] == i) d = a[e];
        dinv[i] = d != 0 ? 1.0 / d : 1.0;
    }
    for (int step = 0; step < nsteps; ++step) {
        double sA = theta * q * fA[step];
        double sB = -(1.0 - theta) * q * fB[step];
        spmv_c(n, rowptr, col, bre, jg, sB, u, b);
        if (pv) {
            for (int64_t e = 0; e < ne; ++e) {
                cplx y = 0;
                for (int k = 0; k < 3; ++k)
                    if (ent_src[3 * e + k] >= 0) y += ent_w[3 * e + k] * u[ent_src[3 * e + k]];
                double ph = q * FB[step] * (g[0] * ent_ell[3 * e] + g[1] * ent_ell[3 * e + 1] + g[2] * ent_ell[3 * e + 2]);
                ybc[ent_dof[e]] = y * (cos(ph) + I * sin(ph));
            }
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                cplx acc = 0;
                for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += pv[e] * ybc[col[e]];
                b[i] += (1.0 - theta) * acc;
            }
        }
        double bn = nrm2(n, b);
        double tol = rtol * bn > atol ? rtol * bn : atol;
        /* warm start x0 = u^n (the previous solution, DOLFIN reuses the vector) */
        spmv_c(n, rowptr, col, a, jg, sA, u, t);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) { r[i] = b[i] - t[i]; rh[i] = r[i]; p[i] = 0; v[i] = 0; }
        double rn = nrm2(n, r);
        cplx rho_old = 1, alpha = 1, omega = 1;
        int it = 0;
        while (rn > tol) {
            if (it >= maxit) { rc = -1; break; }
            cplx rho = dotc(n, rh, r);
            if (cabs(rho) == 0.0 || cabs(omega) == 0.0) { /* breakdown: restart with r^ = r */
                stats[3] += 1;
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; ++i) { rh[i] = r[i]; p[i] = 0; v[i] = 0; }
                rho = dotc(n, rh, r);
                rho_old = 1; alpha = 1; omega = 1;
            }
            cplx beta = (rho / rho_old) * (alpha / omega);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                p[i] = r[i] + beta * (p[i] - omega * v[i]);
                ph[i] = dinv[i] * p[i];
            }
            spmv_c(n, rowptr, col, a, jg, sA, ph, v);
            cplx rv = dotc(n, rh, v);
            alpha = rho / rv;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                s[i] = r[i] - alpha * v[i];
                sh[i] = dinv[i] * s[i];
            }
            spmv_c(n, rowptr, col, a, jg, sA, sh, t);
            cplx ts = dotc(n, t, s);
            double tt = creal(dotc(n, t, t));
            omega = tt > 0 ? ts / tt : 0;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                u[i] += alpha * ph[i] + omega * sh[i];
                r[i] = s[i] - omega * t[i];
            }
            rho_old = rho;
            rn = nrm2(n, r);
            ++it;
        }
        stats[0] += it;
        if (it > stats[1]) stats[1] = it;
        double rel = bn > 0 ? rn / bn : rn;
        if (rel > stats[2]) stats[2] = rel;
        if (rc) break;
    }
    free(a); free(bre); free(dinv); free(b); free(r); free(rh); free(p); free(v); free(s); free(t);
    free(ph); free(sh); free(ybc);
    return rc;
}

int orc_theta_solve(int64_t n, const int64_t *rowptr, const int32_t *col, const double *m, const double *l,
                    const double *jg, double dt, double theta, double q, int nsteps, const double *fA,
                    const double *fB, double rtol, double atol, int maxit, double *u_io, double *stats) {
    return orc_theta_solve_p(n, rowptr, col, m, l, jg, dt, theta, q, nsteps, fA, fB, rtol, atol, maxit, u_io, stats,
                             NULL, 0, NULL, NULL, NULL, NULL, NULL, NULL);
}

/* y = (va + i*s*vj) x, exported for tests (x, y interleaved complex). */
void orc_spmv(int64_t n, const int64_t *rowptr, const int32_t *col, const double *va, const double *vj, double s,
              const double *x, double *y) {
    spmv_c(n, rowptr, col, va, vj, s, (const cplx *)x, (cplx *)y);
}

/* Signal: sum_i w_i Re(u_i) with w = M 1 (PostProcessing, DmriFemLib.py:921-931,971;
 * Comp_Sig3D.ufl).  Fixed-chunk summation.  out[0] = sum w Re u, out[1] = sum w Im u. */
void orc_signal(int64_t n, const double *w, const double *u_io, double *out) {
    const cplx *u = (const cplx *)u_io;
    int64_t nch = (n + ORC_CHUNK - 1) / ORC_CHUNK;
    double re = 0, im = 0;
    for (int64_t c = 0; c < nch; ++c) {
        double sr = 0, si = 0;
        int64_t e = (c + 1) * ORC_CHUNK < n ? (c + 1) * ORC_CHUNK : n;
        for (int64_t i = c * ORC_CHUNK; i < e; ++i) { sr += w[i] * creal(u[i]); si += w[i] * cimag(u[i]); }
        re += sr;
        im += si;
    }
    out[0] = re;
    out[1] = im;
}
