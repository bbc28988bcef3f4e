// gfx950 kernels of libdmri: device set-up (incidence, DOF map, facets, CSR pattern,
// element colouring, closed-form P1 assembly) and the FP64 theta-scheme / BiCGStab solve.
//
// Wave64 throughout.  Every reduction is a fixed-shape tree (wave butterfly, then the
// waves of a block in index order, then the blocks' partials in index order, summed by the
// consuming kernel or a small finalize kernel), so results are bitwise reproducible run to run.
#include <algorithm>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "dmri_internal.hpp"

namespace dmri {

namespace {

// --------------------------------------------------------------------------------------
// small helpers
// --------------------------------------------------------------------------------------
__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_div(double2 a, double2 b) {
    double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// y = a + i*s*j  for complex accumulators a = sum(va x), j = sum(vj x)
__device__ __forceinline__ double2 c_axpj(double2 a, double s, double2 j) {
    return make_double2(a.x - s * j.y, a.y + s * j.x);
}

// Loads and stores with the non-temporal hint, for the streaming vector kernels (s-pass and the
// BiCGStab update), where each element is touched once.  In the SpMV kernels the same hint on the
// matrix entries, own-row values or results measured slower on the 2M-vertex neuron workload (up to
// +15 % per launch for the stores), so they keep plain accesses.
typedef double dmri_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_once(const double2 *p) {
    const dmri_d2v v = __builtin_nontemporal_load(reinterpret_cast<const dmri_d2v *>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_once(double2 *p, double2 v) {
    dmri_d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<dmri_d2v *>(p));
}


// --------------------------------------------------------------------------------------
// deterministic reductions without atomics or device-scope fences
//
// A producing kernel writes one partial per (block, signal) with plain stores; the next
// kernel (a kernel boundary orders the accesses) sums the partials in block order, either
// in every block's prologue (few partials) or through a finalize kernel with one block per
// signal (many partials).  Every sum has a fixed shape, so results are bitwise
// reproducible.
// --------------------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void wave_reduce(double (&v)[NV]) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, 64);
    }
}

// Block reduction; the result is valid in thread 0.  Fixed order.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
    __shared__ double sh[NV][kThreads / 64];
    wave_reduce<NV>(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int i = 0; i < nw; ++i) s += sh[k][i];
            v[k] = s;
        }
    }
    __syncthreads();
}

// One LDS scratch of kNV x kThreads doubles per kernel, shared by the reductions below (the
// kernel's prologue reduction and its epilogue partials use it at different times), so a kernel
// that does both holds 8 KB of it instead of 16 KB and fits more blocks per CU.
__device__ __forceinline__ double (&reduce_scratch())[kNV][kThreads] {
    __shared__ double buf[kNV][kThreads];
    return buf;
}

// Per-signal block partials.  slot = (row in block)*S + signal; every slot in [0, nslots)
// is written by exactly one thread (writer == true).
__device__ __forceinline__ void block_signal_partials(const double (&v)[kNV], bool writer, int slot, int nslots,
                                                      int S, double *parts) {
    double(&buf)[kNV][kThreads] = reduce_scratch();
    if (writer) {
#pragma unroll
        for (int k = 0; k < kNV; ++k) buf[k][slot] = v[k];
    }
    __syncthreads();
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        double acc[kNV] = {0.0, 0.0, 0.0, 0.0};
        for (int j = s; j < nslots; j += S) {
#pragma unroll
            for (int k = 0; k < kNV; ++k) acc[k] += buf[k][j];
        }
#pragma unroll
        for (int k = 0; k < kNV; ++k) parts[((size_t)blockIdx.x * S + s) * kNV + k] = acc[k];
    }
}

// Totals of every signal into tot[s][k]: from the finalize output when given, else by
// summing the nparts block partials in block order.
__device__ __forceinline__ void signal_totals(const double *parts, const double *ftot, int nparts, int S,
                                              double (*tot)[kNV]) {
    if (ftot) {
        for (int i = threadIdx.x; i < S * kNV; i += blockDim.x) tot[i / kNV][i % kNV] = ftot[i];
        __syncthreads();
        return;
    }
    double(&buf)[kNV][kThreads] = reduce_scratch();
    const int J = blockDim.x / S;
    const int j = threadIdx.x / S, s = threadIdx.x % S;
    if (j < J) {
        double acc[kNV] = {0.0, 0.0, 0.0, 0.0};
        for (int b = j; b < nparts; b += J) {
#pragma unroll
            for (int k = 0; k < kNV; ++k) acc[k] += parts[((size_t)b * S + s) * kNV + k];
        }
#pragma unroll
        for (int k = 0; k < kNV; ++k) buf[k][j * S + s] = acc[k];
    }
    __syncthreads();
    if ((int)threadIdx.x < S) {
        double acc[kNV] = {0.0, 0.0, 0.0, 0.0};
        for (int jj = 0; jj < J; ++jj) {
#pragma unroll
            for (int k = 0; k < kNV; ++k) acc[k] += buf[k][jj * S + threadIdx.x];
        }
#pragma unroll
        for (int k = 0; k < kNV; ++k) tot[threadIdx.x][k] = acc[k];
    }
    __syncthreads();
}

// One block per signal: tot[s][k] = sum over blocks of parts[b][s][k], in block order.
__global__ void __launch_bounds__(kThreads) finalize_kernel(const double *parts, int nparts, int S, double *tot) {
    const int s = blockIdx.x;
    double acc[kNV] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < kNV; ++k) acc[k] += parts[((size_t)b * S + s) * kNV + k];
    }
    block_reduce<kNV>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kNV; ++k) tot[s * kNV + k] = acc[k];
    }
}

__global__ void __launch_bounds__(kThreads) sum_part_kernel(const double *x, int64_t n, double *parts) {
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[0] += x[i];
    block_reduce<1>(v);
    if (threadIdx.x == 0) {
        parts[(size_t)blockIdx.x * kNV + 0] = v[0];
        parts[(size_t)blockIdx.x * kNV + 1] = 0.0;
        parts[(size_t)blockIdx.x * kNV + 2] = 0.0;
        parts[(size_t)blockIdx.x * kNV + 3] = 0.0;
    }
}

inline unsigned grid_for(int64_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }

// Largest absolute row sum of the real implicit part (block maxima, then one block over them).
__global__ void __launch_bounds__(kThreads) row_absmax_kernel(int64_t n, const int *rowptr, const double *vals,
                                                              int vw, double *parts) {
    __shared__ double sm[kThreads];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) s += fabs(vals[(int64_t)vw * e]);
        m = fmax(m, s);
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sm[0];
}
// Diagonal values per row and the bound of the complex-diagonal-scaled operator (launch_cdiag_prep).
// h(c) = (R + c Q) / sqrt(a^2 + c^2 j^2) rises up to c* = Q a^2 / (R j^2) and falls after it, so its
// largest value on [0, cmax] is at min(c*, cmax).
__global__ void __launch_bounds__(kThreads) cdiag_prep_kernel(int64_t n, const int *rowptr, const int *col,
                                                              const double *vals, int vw, int S, const double *gc,
                                                              double breal, double *pdiag, double *parts) {
    __shared__ double sm[kThreads];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double R = 0.0, Q[3] = {0.0, 0.0, 0.0}, dg[4] = {0.0, 0.0, 0.0, 0.0};
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const double *v = vals + (int64_t)vw * e;
            R += fabs(v[0]);
            for (int k = 1; k < vw; ++k) Q[k - 1] += fabs(v[k]);
            if (col[e] == (int)i)
                for (int k = 0; k < vw; ++k) dg[k] = v[k];
        }
        for (int k = 0; k < vw; ++k) pdiag[(int64_t)vw * i + k] = dg[k];
        const double a = dg[0], Rb = fmin(R, breal);
        for (int s = 0; s < S; ++s) {
            const double *g = gc + 4 * s;
            double q, j;
            if (vw == 4) {
                q = fabs(g[0]) * Q[0] + fabs(g[1]) * Q[1] + fabs(g[2]) * Q[2];
                j = fabs(g[0] * dg[1] + g[1] * dg[2] + g[2] * dg[3]);
            } else {
                q = Q[0];
                j = fabs(dg[1]);
            }
            const double cmax = g[3];
            double c = cmax;
            if (Rb > 0.0 && j > 0.0) c = fmin(cmax, q * a * a / (Rb * j * j));
            const double h = (Rb + c * q) / sqrt(a * a + c * c * j * j);
            const double h1 = (Rb + cmax * q) / sqrt(a * a + cmax * cmax * j * j);
            m = fmax(m, fmax(fmax(h, h1), Rb / fabs(a)));
        }
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sm[0];
}
__global__ void __launch_bounds__(kThreads) max_final_kernel(const double *parts, int np, double *out) {
    __shared__ double sm[kThreads];
    double m = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax(m, parts[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}

// --------------------------------------------------------------------------------------
// exclusive scan (int32, non-negative counts)
// --------------------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void __launch_bounds__(kScanBlock) scan_tile_kernel(const int *in, int *out, int *tile_sums, int64_t n) {
    __shared__ int s[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int vals[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        vals[k] = i < n ? in[i] : 0;
        sum += vals[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;
    if (threadIdx.x == kScanBlock - 1) tile_sums[blockIdx.x] = s[threadIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
}

__global__ void __launch_bounds__(kScanBlock) scan_add_kernel(int *out, const int *tile_offs, int64_t n) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    const int off = tile_offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] += off;
    }
}

__global__ void set_total_kernel(int *out, int64_t n, const int *in, const int *tile_offs, int ntiles) {
    // out[n] = total = exclusive[n-1] + in[n-1]
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = n > 0 ? out[n - 1] + in[n - 1] : 0;
}

// --------------------------------------------------------------------------------------
// incidence lists (vertex -> cells, vertex -> interface facets)
// --------------------------------------------------------------------------------------
template <int NVE>
__global__ void count_incidence_kernel(const int *ev, int64_t ne, int *cnt) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
#pragma unroll
    for (int a = 0; a < NVE; ++a) atomicAdd(&cnt[ev[NVE * e + a]], 1);
}

template <int NVE>
__global__ void fill_incidence_kernel(const int *ev, int64_t ne, const int *ptr, int *cursor, int *list) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
#pragma unroll
    for (int a = 0; a < NVE; ++a) {
        int v = ev[NVE * e + a];
        int pos = atomicAdd(&cursor[v], 1);
        list[ptr[v] + pos] = (int)e;
    }
}

// Sort each list (insertion sort; lists are short) so that the lists, and everything
// derived from their order, are deterministic.
__global__ void sort_lists_kernel(const int *ptr, int *list, int64_t n) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int b = ptr[v], e = ptr[v + 1];
    for (int i = b + 1; i < e; ++i) {
        int x = list[i];
        int j = i - 1;
        while (j >= b && list[j] > x) {
            list[j + 1] = list[j];
            --j;
        }
        list[j + 1] = x;
    }
}

// --------------------------------------------------------------------------------------
// DOF map: live (compartment, vertex) pairs numbered compartment-major
// --------------------------------------------------------------------------------------
__global__ void mark_live_kernel(const int *tets, const int *phase, int64_t nc, int64_t nv, int *live) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc) return;
    int c = phase ? phase[t] : 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) live[(int64_t)c * nv + tets[4 * t + a]] = 1;
}

__global__ void dof_from_scan_kernel(const int *live, const int *scan, int64_t n, int *dof_of, int64_t nv,
                                     int *dof_vert, int *dof_comp) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (live[i]) {
        int d = scan[i];
        dof_of[i] = d;
        dof_vert[d] = (int)(i % nv);
        dof_comp[d] = (int)(i / nv);
    } else {
        dof_of[i] = -1;
    }
}

// --------------------------------------------------------------------------------------
// faces: neighbour across each face, interface (phase 0 | phase 1) and boundary facets
// --------------------------------------------------------------------------------------
__device__ __forceinline__ void face_verts(const int *T, int o, int &f0, int &f1, int &f2) {
    int v[3], k = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
        if (a != o) v[k++] = T[a];
    // ascending
    int a0 = min(v[0], min(v[1], v[2]));
    int a2 = max(v[0], max(v[1], v[2]));
    int a1 = v[0] + v[1] + v[2] - a0 - a2;
    f0 = a0;
    f1 = a1;
    f2 = a2;
}

__global__ void face_classify_kernel(const int *tets, const int *phase, int64_t nc, const int *vptr,
                                     const int *vlist, int *iface_flag, int *bnd_flag, int *nbr) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * nc) return;
    int64_t t = idx >> 2;
    int o = (int)(idx & 3);
    int T[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) T[a] = tets[4 * t + a];
    int f0, f1, f2;
    face_verts(T, o, f0, f1, f2);
    // scan the shortest of the three incidence lists
    int best = f0, bl = vptr[f0 + 1] - vptr[f0];
    int l1 = vptr[f1 + 1] - vptr[f1], l2 = vptr[f2 + 1] - vptr[f2];
    if (l1 < bl) { best = f1; bl = l1; }
    if (l2 < bl) { best = f2; bl = l2; }
    int found = -1;
    for (int p = vptr[best]; p < vptr[best + 1]; ++p) {
        int u = vlist[p];
        if (u == (int)t) continue;
        int hit = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            int w = tets[4 * (int64_t)u + a];
            hit += (w == f0) + (w == f1) + (w == f2);
        }
        if (hit == 3) { found = u; break; }
    }
    nbr[idx] = found;
    int pt = phase ? phase[t] : 0;
    iface_flag[idx] = (found >= 0 && phase && pt == 0 && phase[found] == 1) ? 1 : 0;
    bnd_flag[idx] = found < 0 ? 1 : 0;
}

__global__ void face_compact_kernel(const int *tets, int64_t nc, const int *nbr, const int *iface_flag,
                                    const int *iface_scan, const int *bnd_flag, const int *bnd_scan, int *iface_v,
                                    int *iface_c, int *bnd_v, int *bnd_c) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * nc) return;
    int64_t t = idx >> 2;
    int o = (int)(idx & 3);
    if (!iface_flag[idx] && !bnd_flag[idx]) return;
    int T[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) T[a] = tets[4 * t + a];
    int f0, f1, f2;
    face_verts(T, o, f0, f1, f2);
    if (iface_flag[idx]) {
        int j = iface_scan[idx];
        iface_v[3 * j] = f0;
        iface_v[3 * j + 1] = f1;
        iface_v[3 * j + 2] = f2;
        iface_c[2 * j] = (int)t;
        iface_c[2 * j + 1] = nbr[idx];
    }
    if (bnd_flag[idx]) {
        int j = bnd_scan[idx];
        bnd_v[3 * j] = f0;
        bnd_v[3 * j + 1] = f1;
        bnd_v[3 * j + 2] = f2;
        bnd_c[j] = (int)t;
    }
}

// --------------------------------------------------------------------------------------
// CSR pattern: one wave per DOF row.  The candidate columns of row (c, v) are the DOFs
// of every cell of phase c touching v and of every interface facet touching v (both
// compartments).  Sorted unique set, computed by ranking in LDS (no atomics), so the
// result equals the brute-force set union exactly.
// --------------------------------------------------------------------------------------
constexpr int kPatCap = 2048;

__global__ void __launch_bounds__(64) pattern_kernel(int pass, const int *dof_vert, const int *dof_comp,
                                                     const int *tets, const int *phase, int64_t nv,
                                                     const int *vptr, const int *vlist, const int *fptr,
                                                     const int *flist, const int *iface_v, const int *dof_of,
                                                     int *row_count, const int *rowptr, int *col, int *overflow) {
    __shared__ int cand[kPatCap];
    __shared__ unsigned char first[kPatCap];
    const int d = blockIdx.x;
    const int lane = threadIdx.x;
    const int v = dof_vert[d], c = dof_comp[d];
    const int c0 = vptr[v], ncell = vptr[v + 1] - c0;
    const int f0 = fptr ? fptr[v] : 0, nface = fptr ? fptr[v + 1] - f0 : 0;
    const int m = 4 * ncell + 6 * nface;
    if (m > kPatCap) {
        if (lane == 0) atomicOr(overflow, 1);
        return;
    }
    for (int i = lane; i < ncell; i += 64) {
        int t = vlist[c0 + i];
        bool ok = (phase ? phase[t] : 0) == c;
#pragma unroll
        for (int a = 0; a < 4; ++a) cand[4 * i + a] = ok ? dof_of[(int64_t)c * nv + tets[4 * (int64_t)t + a]] : -1;
    }
    for (int i = lane; i < nface; i += 64) {
        int f = flist[f0 + i];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int w = iface_v[3 * (int64_t)f + a];
            cand[4 * ncell + 6 * i + a] = dof_of[w];
            cand[4 * ncell + 6 * i + 3 + a] = dof_of[nv + w];
        }
    }
    __syncthreads();
    for (int i = lane; i < m; i += 64) {
        int ci = cand[i];
        unsigned char fl = ci >= 0;
        for (int j = 0; j < i && fl; ++j)
            if (cand[j] == ci) fl = 0;
        first[i] = fl;
    }
    __syncthreads();
    if (pass == 0) {
        int cnt = 0;
        for (int i = lane; i < m; i += 64) cnt += first[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
        if (lane == 0) row_count[d] = cnt;
    } else {
        const int base = rowptr[d];
        for (int i = lane; i < m; i += 64) {
            if (!first[i]) continue;
            int ci = cand[i], rank = 0;
            for (int j = 0; j < m; ++j) rank += (first[j] && cand[j] < ci);
            col[base + rank] = ci;
        }
    }
}

// --------------------------------------------------------------------------------------
// Jones-Plassmann colouring of elements (cells or facets) that share a vertex.
// An uncoloured element takes colour k when it beats every neighbour that is uncoloured
// or coloured k in this round.  Priorities are a fixed hash, so the colouring is
// deterministic.
// --------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ bool beats(int a, int b) {
    unsigned ha = hash32((unsigned)a), hb = hash32((unsigned)b);
    return ha > hb || (ha == hb && a > b);
}

__global__ void jp_round_kernel(int k, int64_t ne, int nvert, const int *ev, const int *vptr, const int *vlist,
                                int *color, int *remaining) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool left = false;
    if (e < ne && color[e] < 0) {
        bool win = true;
        for (int a = 0; a < nvert && win; ++a) {
            int v = ev[(int64_t)nvert * e + a];
            for (int p = vptr[v]; p < vptr[v + 1]; ++p) {
                int u = vlist[p];
                if (u == (int)e) continue;
                int cu = color[u];
                if ((cu < 0 || cu == k) && beats(u, (int)e)) { win = false; break; }
            }
        }
        if (win) color[e] = k;
        else left = true;
    }
    unsigned long long b = __ballot(left);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(remaining, __popcll(b));
}

__global__ void color_hist_kernel(const int *color, int64_t ne, int *hist) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < ne) atomicAdd(&hist[color[e]], 1);
}

__global__ void color_scatter_kernel(const int *color, int64_t ne, const int *offs, int *cursor, int *perm) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int c = color[e];
    int pos = atomicAdd(&cursor[c], 1);
    perm[offs[c] + pos] = (int)e;
}

// --------------------------------------------------------------------------------------
// closed-form P1 element assembly by row gather: a group of kRowLanes lanes per DOF row, lane j
// owning the row's entries j, j + kRowLanes, ...  Each lane walks the incident cells of the row's
// vertex (sorted) and adds, for every cell of the row's compartment that holds the entry's
// column vertex, the element entry.  Every entry is written once by one lane, so there is no
// colouring and no atomic; the summation order per entry is the sorted cell order.
// --------------------------------------------------------------------------------------
__device__ __forceinline__ int csr_find(const int *rowptr, const int *col, int r, int c) {
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        int cm = col[mid];
        if (cm == c) return mid;
        if (cm < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

constexpr int kRowLanes = 16;

__global__ void __launch_bounds__(256) assemble_rows_kernel(int64_t n, const int *dof_vert, const int *dof_comp,
                                                            const int *rowptr, const int *col, const int *tets,
                                                            const int *phase, const double *xyz, const int *vptr,
                                                            const int *vlist, const double *Dc, const double *Dcell,
                                                            const double *iT2c, const double *iT2cell, double *mv,
                                                            double *kv, double *rv, double *jxv, double *jyv,
                                                            double *jzv, double *cxv, double *cyv, double *czv) {
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes;
    const int lane = threadIdx.x % kRowLanes;
    if (r >= n) return;
    const int v = dof_vert[r], cr = dof_comp[r];
    const int e0 = rowptr[r], e1 = rowptr[r + 1];
    const int q0 = vptr[v], q1 = vptr[v + 1];
    for (int e = e0 + lane; e < e1; e += kRowLanes) {
        const int cl = col[e];
        const int vj = dof_vert[cl];
        double m = 0.0, kk = 0.0, rr = 0.0, jx = 0.0, jy = 0.0, jz = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        if (dof_comp[cl] == cr) {
            for (int q = q0; q < q1; ++q) {
                const int t = vlist[q];
                if ((phase ? phase[t] : 0) != cr) continue;
                const int4 tv = *reinterpret_cast<const int4 *>(tets + 4 * (int64_t)t);
                const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
                int a = 0, b = -1;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (vid[k] == v) a = k;
                    if (vid[k] == vj) b = k;
                }
                if (b < 0) continue;
                double X[4][3];
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int d = 0; d < 3; ++d) X[k][d] = xyz[3 * (int64_t)vid[k] + d];
                double E[3][3];
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int d = 0; d < 3; ++d) E[k][d] = X[k + 1][d] - X[0][d];
                // rows of the inverse Jacobian (gradients of the barycentric functions 1..3)
                const double c23[3] = {E[1][1] * E[2][2] - E[1][2] * E[2][1], E[1][2] * E[2][0] - E[1][0] * E[2][2],
                                       E[1][0] * E[2][1] - E[1][1] * E[2][0]};
                const double c31[3] = {E[2][1] * E[0][2] - E[2][2] * E[0][1], E[2][2] * E[0][0] - E[2][0] * E[0][2],
                                       E[2][0] * E[0][1] - E[2][1] * E[0][0]};
                const double c12[3] = {E[0][1] * E[1][2] - E[0][2] * E[1][1], E[0][2] * E[1][0] - E[0][0] * E[1][2],
                                       E[0][0] * E[1][1] - E[0][1] * E[1][0]};
                const double det = E[0][0] * c23[0] + E[0][1] * c23[1] + E[0][2] * c23[2];
                double G[4][3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    G[1][d] = c23[d] / det;
                    G[2][d] = c31[d] / det;
                    G[3][d] = c12[d] / det;
                    G[0][d] = -(G[1][d] + G[2][d] + G[3][d]);
                }
                const double vol = fabs(det) / 6.0;
                double Ga[3], Gb[3], Xa[3], Xb[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    Ga[d] = G[0][d];
                    Gb[d] = G[0][d];
                    Xa[d] = X[0][d];
                    Xb[d] = X[0][d];
                }
#pragma unroll
                for (int k = 1; k < 4; ++k)
                    if (a == k || b == k)
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            if (a == k) { Ga[d] = G[k][d]; Xa[d] = X[k][d]; }
                            if (b == k) { Gb[d] = G[k][d]; Xb[d] = X[k][d]; }
                        }
                double Dm[9];
                if (Dcell) {
#pragma unroll
                    for (int d = 0; d < 9; ++d) Dm[d] = Dcell[9 * (int64_t)t + d];
                } else {
#pragma unroll
                    for (int d = 0; d < 9; ++d) Dm[d] = 0.0;
                    Dm[0] = Dm[4] = Dm[8] = Dc[cr];
                }
                const double it2 = iT2cell ? iT2cell[t] : iT2c[cr];
                const double dab = (a == b) ? 2.0 : 1.0;
                const double mab = vol * dab / 20.0;
                m += mab;
                rr += it2 * mab;
                // row i = test function (vertex a), column j = trial function (vertex b):
                // inner(D grad(u), grad(v)) gives K_ij = grad(phi_i) . D grad(phi_j)
                double kab = 0.0;
#pragma unroll
                for (int d = 0; d < 3; ++d) kab += Ga[d] * (Dm[3 * d] * Gb[0] + Dm[3 * d + 1] * Gb[1] + Dm[3 * d + 2] * Gb[2]);
                kk += vol * kab;
                const double w = vol * dab / 120.0;
                const double S0 = X[0][0] + X[1][0] + X[2][0] + X[3][0];
                const double S1 = X[0][1] + X[1][1] + X[2][1] + X[3][1];
                const double S2 = X[0][2] + X[1][2] + X[2][2] + X[3][2];
                jx += w * (S0 + Xa[0] + Xb[0]);
                jy += w * (S1 + Xa[1] + Xb[1]);
                jz += w * (S2 + Xa[2] + Xb[2]);
                // strong periodic form: int 2 D (d phi_j / d x_d) phi_i = 2 D grad(phi_j)_d |T| / 4
                const double cd = 0.5 * vol * Dc[cr];
                cx += cd * Gb[0];
                cy += cd * Gb[1];
                cz += cd * Gb[2];
            }
        }
        if (cxv) {
            cxv[e] = cx;
            cyv[e] = cy;
            czv[e] = cz;
        }
        mv[e] = m;
        kv[e] = kk;
        rv[e] = rr;
        jxv[e] = jx;
        jyv[e] = jy;
        jzv[e] = jz;
    }
}

__global__ void assemble_facets_kernel(int64_t start, int64_t count, const int *perm, const int *iface_v,
                                       const double *xyz, int64_t nv, const int *dof_of, const int *rowptr,
                                       const int *col, double kappa, const double *kappa_f, double *gi, int *miss) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    const int f = perm[start + g];
    double X[3][3];
    int d0[3], d1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int v = iface_v[3 * (int64_t)f + a];
        d0[a] = dof_of[v];
        d1[a] = dof_of[nv + v];
#pragma unroll
        for (int q = 0; q < 3; ++q) X[a][q] = xyz[3 * (int64_t)v + q];
    }
    double e1[3], e2[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        e1[q] = X[1][q] - X[0][q];
        e2[q] = X[2][q] - X[0][q];
    }
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                 cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    const double kap = kappa_f ? kappa_f[f] : kappa;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double val = kap * area * ((a == b) ? 2.0 : 1.0) / 12.0;
            int p00 = csr_find(rowptr, col, d0[a], d0[b]);
            int p11 = csr_find(rowptr, col, d1[a], d1[b]);
            int p01 = csr_find(rowptr, col, d0[a], d1[b]);
            int p10 = csr_find(rowptr, col, d1[a], d0[b]);
            if (p00 < 0 || p11 < 0 || p01 < 0 || p10 < 0) {
                atomicOr(miss, 1);
                continue;
            }
            gi[p00] += val;
            gi[p11] += val;
            gi[p01] -= val;
            gi[p10] -= val;
        }
}

// sum of the row counts in 64 bits (the CSR offsets are 32-bit: the caller checks the total fits)
__global__ void count_sum64_kernel(const int *cnt, int64_t n, unsigned long long *out) {
    unsigned long long v = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += (unsigned long long)(unsigned)cnt[i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, v);
}

__global__ void diag_weights_kernel(int64_t n, const int *rowptr, const int *col, const double *m, int *diag,
                                    double *w, int *missing) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int p = csr_find(rowptr, col, (int)i, (int)i);
    diag[i] = p;
    if (p < 0) atomicOr(missing, 1);
    double s = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) s += m[e];
    w[i] = s;
}

// --------------------------------------------------------------------------------------
// Weak pseudo-periodic boundary condition (set-up).
// --------------------------------------------------------------------------------------
__device__ inline bool near_periodic_face(const double *X, const PeriodicBox &B, double eps) {
    bool on = false;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (B.dir[d] && (X[d] < B.lo[d] + eps || X[d] > B.hi[d] - eps)) on = true;
    return on;
}

// One thread per DOF row: P_ij += kappa_s |F|/60 sum_k chi_k (1+d_ij)(1+d_ik+d_jk) over the boundary
// facets F of the row's vertex whose cell lies in the row's compartment (the row owns its entries,
// so no colouring is needed; facets are visited in the sorted incidence order).  rowflag[i] = 1
// when the row received a contribution.
__global__ void periodic_mass_kernel(int64_t n, int64_t nv, const int *dof_vert, const int *dof_comp,
                                     const int *bptr, const int *blist, const int *bnd_v, const int *bnd_c,
                                     const int *phase, const double *xyz, const int *dof_of, const int *rowptr,
                                     const int *col, PeriodicBox B, double *pv, int *rowflag, int *miss) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = dof_vert[i], c = dof_comp[i];
    const double eps = 1e-2 * B.hmin, ks = 3e-3 / B.hmin;
    int touched = 0;
    for (int q = bptr[v]; q < bptr[v + 1]; ++q) {
        const int f = blist[q];
        if ((phase ? phase[bnd_c[f]] : 0) != c) continue;
        int vv[3], a = -1;
        double X[3][3], chi[3], cs = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vv[k] = bnd_v[3 * (int64_t)f + k];
            if (vv[k] == v) a = k;
#pragma unroll
            for (int d = 0; d < 3; ++d) X[k][d] = xyz[3 * (int64_t)vv[k] + d];
            chi[k] = near_periodic_face(X[k], B, eps) ? 1.0 : 0.0;
            cs += chi[k];
        }
        if (cs == 0.0 || a < 0) continue;
        double e1[3], e2[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            e1[d] = X[1][d] - X[0][d];
            e2[d] = X[2][d] - X[0][d];
        }
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        const double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
        for (int b = 0; b < 3; ++b) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                sum += chi[k] * (a == b ? 2.0 : 1.0) * (1.0 + (a == k ? 1.0 : 0.0) + (b == k ? 1.0 : 0.0));
            const int j = dof_of[(int64_t)c * nv + vv[b]];
            const int p = j < 0 ? -1 : csr_find(rowptr, col, (int)i, j);
            if (p < 0) {
                atomicOr(miss, 1);
                continue;
            }
            pv[p] += ks * area / 60.0 * sum;
            touched = 1;
        }
    }
    rowflag[i] = touched;
}

// DOFs whose vertex lies within 1e-7 of a periodic face take a mirrored value (the tolerance of
// WeakPseudoPeriodic eval): flag[i] = 1.
__global__ void periodic_flag_kernel(int64_t n, const int *dof_vert, const double *xyz, PeriodicBox B, int *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *X = xyz + 3 * (int64_t)dof_vert[i];
    int hit = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (B.dir[d] && (fabs(X[d] - B.lo[d]) <= 1e-7 || fabs(X[d] - B.hi[d]) <= 1e-7)) hit = 1;
    flag[i] = hit;
}

// Plane of a boundary facet: 2d (all vertices within eps of lo[d]) or 2d+1 (of hi[d]) for a
// periodic axis d, else -1; facets bin by centroid on their plane's grid.
__global__ void bfacet_bin_kernel(int64_t nbf, const int *bnd_v, const double *xyz, PeriodicBox B, PlaneGrids G,
                                  int *fbin, int *cnt) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nbf) return;
    double X[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) X[k][d] = xyz[3 * (int64_t)bnd_v[3 * f + k] + d];
    const double eps = 1e-2 * B.hmin;
    int plane = -1;
    for (int d = 0; d < 3 && plane < 0; ++d) {
        if (!B.dir[d]) continue;
        for (int side = 0; side < 2 && plane < 0; ++side) {
            const double bound = side ? B.hi[d] : B.lo[d];
            if (fabs(X[0][d] - bound) <= eps && fabs(X[1][d] - bound) <= eps && fabs(X[2][d] - bound) <= eps)
                plane = 2 * d + side;
        }
    }
    if (plane < 0) {
        fbin[f] = -1;
        return;
    }
    const PlaneGrid &g = G.p[plane];
    const double cu = (X[0][g.u] + X[1][g.u] + X[2][g.u]) / 3.0, cw = (X[0][g.w] + X[1][g.w] + X[2][g.w]) / 3.0;
    const int iu = min(max((int)floor((cu - g.lo_u) / g.h_u), 0), g.nb_u - 1);
    const int iw = min(max((int)floor((cw - g.lo_w) / g.h_w), 0), g.nb_w - 1);
    const int bin = g.off + iu * g.nb_w + iw;
    fbin[f] = bin;
    atomicAdd(cnt + bin, 1);
}

__global__ void bfacet_fill_kernel(int64_t nbf, const int *fbin, const int *binptr, int *cursor, int *binlist) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nbf || fbin[f] < 0) return;
    const int b = fbin[f];
    binlist[binptr[b] + atomicAdd(cursor + b, 1)] = (int)f;
}

// Mirror entries, one thread per flagged DOF: target face by the eval's order (a later periodic
// axis overrides an earlier one), l = mirror - x, and the facet of the target face whose 2-D
// barycentric coordinates of the mirror point have the largest minimum (ties: smaller facet
// index) among the facets binned within one bin of it.  bad[0] = 1 when none is found.
__global__ void periodic_entry_kernel(int64_t n, int64_t nv, const int *flag, const int *scan, const int *dof_vert,
                                      const int *dof_comp, const int *dof_of, const double *xyz, const int *bnd_v,
                                      const int *binptr, const int *binlist, PeriodicBox B, PlaneGrids G,
                                      int *ent_dof, int *ent_src, double *ent_w, double *ent_ell, int *bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const int e = scan[i];
    const int v = dof_vert[i], c = dof_comp[i];
    double X[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) X[d] = xyz[3 * (int64_t)v + d];
    int plane = -1;
    double ell[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) {
        if (!B.dir[d]) continue;
        if (fabs(X[d] - B.lo[d]) <= 1e-7) {
            plane = 2 * d + 1;
            ell[0] = ell[1] = ell[2] = 0.0;
            ell[d] = B.hi[d] - X[d];
        }
        if (fabs(X[d] - B.hi[d]) <= 1e-7) {
            plane = 2 * d;
            ell[0] = ell[1] = ell[2] = 0.0;
            ell[d] = B.lo[d] - X[d];
        }
    }
    ent_dof[e] = (int)i;
#pragma unroll
    for (int d = 0; d < 3; ++d) ent_ell[3 * (int64_t)e + d] = ell[d];
    const PlaneGrid &g = G.p[plane];
    const double pu = X[g.u], pw = X[g.w];
    const int iu = min(max((int)floor((pu - g.lo_u) / g.h_u), 0), g.nb_u - 1);
    const int iw = min(max((int)floor((pw - g.lo_w) / g.h_w), 0), g.nb_w - 1);
    double best = -1e300, bl[3] = {0.0, 0.0, 0.0};
    int bf = -1;
    for (int du = -1; du <= 1; ++du)
        for (int dw = -1; dw <= 1; ++dw) {
            const int ju = iu + du, jw = iw + dw;
            if (ju < 0 || ju >= g.nb_u || jw < 0 || jw >= g.nb_w) continue;
            const int b = g.off + ju * g.nb_w + jw;
            for (int q = binptr[b]; q < binptr[b + 1]; ++q) {
                const int f = binlist[q];
                double A[2], Bv[2], C[2];
                const double *xa = xyz + 3 * (int64_t)bnd_v[3 * (int64_t)f];
                const double *xb = xyz + 3 * (int64_t)bnd_v[3 * (int64_t)f + 1];
                const double *xc = xyz + 3 * (int64_t)bnd_v[3 * (int64_t)f + 2];
                A[0] = xa[g.u]; A[1] = xa[g.w];
                Bv[0] = xb[g.u]; Bv[1] = xb[g.w];
                C[0] = xc[g.u]; C[1] = xc[g.w];
                const double det = (Bv[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (Bv[1] - A[1]);
                if (det == 0.0) continue;
                const double l1 = ((pu - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (pw - A[1])) / det;
                const double l2 = ((Bv[0] - A[0]) * (pw - A[1]) - (pu - A[0]) * (Bv[1] - A[1])) / det;
                const double l0 = 1.0 - l1 - l2;
                const double m = fmin(l0, fmin(l1, l2));
                if (m > best || (m == best && f < bf)) {
                    best = m;
                    bf = f;
                    bl[0] = l0;
                    bl[1] = l1;
                    bl[2] = l2;
                }
            }
        }
    if (bf < 0) {
        atomicOr(bad, 1);
        bf = 0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ent_src[3 * (int64_t)e + k] = dof_of[(int64_t)c * nv + bnd_v[3 * (int64_t)bf + k]];
        ent_w[3 * (int64_t)e + k] = bl[k];
    }
}

__global__ void compact_rows_kernel(int64_t n, const int *flag, const int *scan, int *rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) rows[scan[i]] = (int)i;
}

// --------------------------------------------------------------------------------------
// Solver ordering.  The solver numbers DOFs along a Morton (Z-order) curve of their vertex
// coordinates (the two compartments of a vertex next to each other), so that the rows a block
// or an XCD works on at once form a compact region and the columns they gather hit in L2.
// Assembly and every export keep the compartment-major numbering.
// --------------------------------------------------------------------------------------
__device__ inline uint64_t spread_bits_20(uint64_t v) {
    // bit k of v -> bit 3k
    v &= 0xFFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ void morton_key_kernel(int64_t n, const int *dof_vert, const int *dof_comp, const double *xyz, double lo0,
                                  double lo1, double lo2, double inv_ext, uint64_t *key, int *idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *X = xyz + 3 * (int64_t)dof_vert[i];
    const double lo[3] = {lo0, lo1, lo2};
    uint64_t code = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double t = (X[d] - lo[d]) * inv_ext;
        t = fmin(fmax(t, 0.0), 1.0);
        const uint64_t q = (uint64_t)(t * 1048575.0);
        code |= spread_bits_20(q) << d;
    }
    key[i] = (code << 1) | (uint64_t)(dof_comp[i] & 1);
    idx[i] = (int)i;
}

__global__ void invert_perm_kernel(int64_t n, const int *perm, int *iperm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = (int)i;
}

__global__ void perm_row_len_kernel(int64_t n, const int *perm, const int *rowptr, int *len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) len[i] = rowptr[perm[i] + 1] - rowptr[perm[i]];
}

// New row i = old row perm[i] with columns renumbered and sorted; emap[e'] = old entry index.
// Also the per-row arrays the solver reads in the new numbering.
__global__ void perm_fill_kernel(int64_t n, const int *perm, const int *iperm, const int *rowptr, const int *col,
                                 const int *srow, const int *diag, const double *w, const int *dof_comp, int *scol,
                                 int *emap, int *sdiag, double *sw, int *scomp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = perm[i];
    const int e0 = rowptr[r], len = rowptr[r + 1] - e0, o = srow[i];
    for (int k = 0; k < len; ++k) {
        const int c = iperm[col[e0 + k]];
        int p = k;
        while (p > 0 && scol[o + p - 1] > c) {
            scol[o + p] = scol[o + p - 1];
            emap[o + p] = emap[o + p - 1];
            --p;
        }
        scol[o + p] = c;
        emap[o + p] = e0 + k;
    }
    sdiag[i] = diag[r];
    sw[i] = w[r];
    if (scomp) scomp[i] = dof_comp[r];
}

__global__ void gather_double_kernel(int64_t n, const int *map, const double *src, double *dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[map[i]];
}

// Column compression of a tile of TR rows (one block per tile): sort the tile's column indices
// in LDS (bitonic), keep the distinct ones.  pass 0 writes their count (or -1 when the tile has
// more than kTileMaxEntries entries); pass 1 writes them to ucol at tptr[tile] and the index of
// every entry's column among them to lcol.
__global__ void __launch_bounds__(256) tile_compress_kernel(int pass, int64_t n, int TR, const int *srow,
                                                            const int *scol, int *tinfo, const int *tptr, int *ucol,
                                                            unsigned short *lcol) {
    __shared__ int sk[kTileMaxEntries];
    __shared__ int sflag[kTileMaxEntries];
    const int64_t t = blockIdx.x;
    const int64_t r0 = t * TR;
    const int64_t r1 = r0 + TR < n ? r0 + TR : n;
    const int E0 = srow[r0], ne = srow[r1] - E0;
    if (ne > kTileMaxEntries) {
        if (pass == 0 && threadIdx.x == 0) tinfo[t] = -1;
        return;
    }
    int m = 1;
    while (m < ne) m <<= 1;
    for (int i = threadIdx.x; i < m; i += blockDim.x) sk[i] = i < ne ? scol[E0 + i] : 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const int x = sk[i], y = sk[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        sk[i] = y;
                        sk[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    // first occurrences of each column, their rank by a block scan over the flags
    for (int i = threadIdx.x; i < m; i += blockDim.x) sflag[i] = (i < ne && (i == 0 || sk[i] != sk[i - 1])) ? 1 : 0;
    __syncthreads();
    {  // inclusive scan: serial within each thread's chunk, then the chunk offsets
        __shared__ int part[256];
        const int per = (m + (int)blockDim.x - 1) / (int)blockDim.x;
        const int b0 = (int)threadIdx.x * per, b1 = b0 + per < m ? b0 + per : m;
        int run = 0;
        for (int i = b0; i < b1; ++i) {
            run += sflag[i];
            sflag[i] = run;
        }
        part[threadIdx.x] = run;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int j = 0; j < (int)blockDim.x; ++j) {
                const int v = part[j];
                part[j] = acc;
                acc += v;
            }
        }
        __syncthreads();
        for (int i = b0; i < b1; ++i) sflag[i] += part[threadIdx.x];
        __syncthreads();
    }
    const int nu = ne > 0 ? sflag[ne - 1] : 0;
    if (pass == 0) {
        if (threadIdx.x == 0) tinfo[t] = nu;
        return;
    }
    const int u0 = tptr[t];
    for (int i = threadIdx.x; i < ne; i += blockDim.x)
        if (i == 0 || sk[i] != sk[i - 1]) ucol[u0 + sflag[i] - 1] = sk[i];
    __syncthreads();
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const int c = scol[E0 + i];
        int lo = 0, hi = ne - 1;  // first position of c in the sorted keys
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sk[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        lcol[E0 + i] = (unsigned short)(sflag[lo] - 1);
    }
}

__global__ void clamp_count_kernel(int64_t n, const int *in, int *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] > 0 ? in[i] : 0;
}

__global__ void renumber_kernel(int64_t n, const int *iperm, const int *src, int *dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] < 0 ? -1 : iperm[src[i]];
}

__global__ void unscale_perm_kernel(int64_t n, int S, int sig, const int *perm, const double *sc, const double2 *x,
                                    double2 *u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 xi = x[i * S + sig];
        u[perm[i]] = make_double2(sc[i] * xi.x, sc[i] * xi.y);
    }
}

// --------------------------------------------------------------------------------------
// Solve-time preparation.  Jacobi preconditioning is applied as the symmetric scaling
// S = diag(a)^(-1/2) of the whole system (a = M/dt + theta*L): the solver works on
// S A S, whose real diagonal is 1, with x = S^-1 u.  (PETSc's Jacobi on the reference's
// real 2x2-block operator uses the same diagonal: the J coupling is off the block diagonal.)
// --------------------------------------------------------------------------------------
// pv (weak pseudo-periodic boundary mass, may be null) enters the implicit side only.
__global__ void prep_scale_kernel(int64_t n, const int *diag, const double *m, const double *k, const double *r,
                                  const double *gi, const double *pv, double dt, double theta, double *sc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int p = diag[i];
    double a = m[p] / dt + theta * (k[p] + r[p] + gi[p] + (pv ? pv[p] : 0.0));
    sc[i] = 1.0 / sqrt(a);
}

// kPrepLanes lanes per row walk the row's entries, so the value arrays are read in runs.
constexpr int kPrepLanes = 8;
// Solver rows/columns (rowptr, col) in the solver numbering; the assembled arrays are read at the
// old entry index emap[e].
__global__ void prep_values_kernel(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc,
                                   const double *m, const double *k, const double *r, const double *gi,
                                   const double *pv, const double *jx, const double *jy, const double *jz,
                                   int per_gdir, double gx, double gy, double gz, double dt, double theta,
                                   double theta_b, double *vals, double *bre) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / kPrepLanes;
    if (i >= n) return;
    const double si = sc[i];
    for (int e = rowptr[i] + (int)(t % kPrepLanes); e < rowptr[i + 1]; e += kPrepLanes) {
        const double s = si * sc[col[e]];
        const int o = emap[e];
        const double l = k[o] + r[o] + gi[o];
        const double mk = m[o] / dt;
        bre[e] = s * (mk - theta_b * l);
        const double impl = s * (mk + theta * (l + (pv ? pv[o] : 0.0)));
        if (per_gdir) {
            vals[4 * (int64_t)e + 0] = impl;
            vals[4 * (int64_t)e + 1] = s * jx[o];
            vals[4 * (int64_t)e + 2] = s * jy[o];
            vals[4 * (int64_t)e + 3] = s * jz[o];
        } else {
            vals[2 * (int64_t)e + 0] = impl;
            vals[2 * (int64_t)e + 1] = s * (gx * jx[o] + gy * jy[o] + gz * jz[o]);
        }
    }
}

__global__ void init_x_kernel(int64_t n, int S, const double *sc, double2 *x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * S) x[i] = make_double2(1.0 / sc[i / S], 0.0);
}

// Plain streaming copy, one 16 B element per thread over a grid that covers the array (the
// hardware dispatcher keeps every CU fed; beyond the 256 MB MALL this reaches ~6.2 TB/s, where a
// 4096-block grid-stride loop reaches ~4.6): the bandwidth reference for the roofline and the
// calibration kernel for the HBM byte counters.
__global__ void __launch_bounds__(kThreads) copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                        int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void unscale_kernel(int64_t n, int S, int sig, const double *sc, const double2 *x, double2 *u) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 xi = x[i * S + sig];
        u[i] = make_double2(sc[i] * xi.x, sc[i] * xi.y);
    }
}

// --------------------------------------------------------------------------------------
// Solver kernels for S signals sharing one pattern.  SpMV mapping: thread =
// (row in block, signal s, lane g of G); the G lanes of a (row, signal) split the row's
// entries and are reduced with a butterfly.  Vectors are stored x[row*S + s] so that the
// signals of one gathered column are contiguous.  Vector kernels: thread = (row, s).
//
// y_s = (Are + i (c_s . J)) x_s with Are = S(M/dt + theta L)S and J = S(g.J)S (one
// direction, PG = false) or J = S(Jx, Jy, Jz)S with per-signal c_s = theta q_s f_n g_s
// (PG = true).
// --------------------------------------------------------------------------------------
template <bool PG>
struct Acc {
    double2 a;
    double2 j[PG ? 3 : 1];
};

template <int G, bool PG>
__device__ __forceinline__ double2 imag_coef_apply(const Acc<PG> &acc, const double *coef, int s) {
    double2 cj;
    if (PG) {
        const double cx = coef[3 * s], cy = coef[3 * s + 1], cz = coef[3 * s + 2];
        cj = make_double2(cx * acc.j[0].x + cy * acc.j[1].x + cz * acc.j[2].x,
                          cx * acc.j[0].y + cy * acc.j[1].y + cz * acc.j[2].y);
    } else {
        const double c = coef[3 * s];
        cj = make_double2(c * acc.j[0].x, c * acc.j[0].y);
    }
    // a + i*cj
    return make_double2(acc.a.x - cj.y, acc.a.y + cj.x);
}

// A row's (a + i c_s.J) x with the signal's coefficients folded into every entry (per-signal
// directions): one complex sum instead of Acc<true>'s four, which frees registers in the SpMV
// kernels that apply one coefficient set (the right-hand-side passes, which apply two, keep Acc).
struct FoldAcc {
    double2 a;
    double cx, cy, cz;
};
__device__ __forceinline__ FoldAcc fold_acc(const double *coef, int s) {
    return FoldAcc{make_double2(0.0, 0.0), coef[3 * s], coef[3 * s + 1], coef[3 * s + 2]};
}
template <int G, bool PG>
__device__ __forceinline__ double2 imag_coef_apply(const FoldAcc &acc, const double *, int) {
    return acc.a;
}
template <bool PG>
__device__ __forceinline__ void acc_zero(Acc<PG> &acc) {
    acc.a = make_double2(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < (PG ? 3 : 1); ++d) acc.j[d] = make_double2(0.0, 0.0);
}
__device__ __forceinline__ void acc_zero(FoldAcc &acc) { acc.a = make_double2(0.0, 0.0); }

// Inverse of the complex diagonal d = a_ii + i (c_s . j_ii) of the scaled operator at row i for
// signal s: conj(d) times the hardware reciprocal of |d|^2 (an approximate inverse, but the same
// one at every application, so the preconditioner stays one fixed linear operator).
template <bool PG>
__device__ __forceinline__ double2 cdiag_inv(const double *pdiag, const double *coef, int64_t i, int s) {
    double re, im;
    if (PG) {
        const double4 d = *reinterpret_cast<const double4 *>(pdiag + 4 * i);
        re = d.x;
        im = coef[3 * s] * d.y + coef[3 * s + 1] * d.z + coef[3 * s + 2] * d.w;
    } else {
        const double2 d = *reinterpret_cast<const double2 *>(pdiag + 2 * i);
        re = d.x;
        im = coef[3 * s] * d.y;
    }
    const double r = __builtin_amdgcn_rcp(re * re + im * im);
    return make_double2(re * r, -im * r);
}

template <int G>
__device__ __forceinline__ void group_reduce(double2 &v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
}

template <int G, bool PG>
__device__ __forceinline__ void group_reduce_acc(Acc<PG> &acc) {
    group_reduce<G>(acc.a);
#pragma unroll
    for (int d = 0; d < (PG ? 3 : 1); ++d) group_reduce<G>(acc.j[d]);
}
template <int G, bool PG>
__device__ __forceinline__ void group_reduce_acc(FoldAcc &acc) {
    group_reduce<G>(acc.a);
}

// Accumulate row entries g, g+G, ... of x = X(col) (X(c) returns the gathered value).
template <bool PG>
__device__ __forceinline__ void acc_entry(const double *vals, int e, double2 xv, Acc<PG> &acc) {
    if (PG) {
        const double4 v = *reinterpret_cast<const double4 *>(vals + 4 * (int64_t)e);
        acc.a.x += v.x * xv.x;
        acc.a.y += v.x * xv.y;
        acc.j[0].x += v.y * xv.x;
        acc.j[0].y += v.y * xv.y;
        acc.j[1].x += v.z * xv.x;
        acc.j[1].y += v.z * xv.y;
        acc.j[2].x += v.w * xv.x;
        acc.j[2].y += v.w * xv.y;
    } else {
        const double2 v = *reinterpret_cast<const double2 *>(vals + 2 * (int64_t)e);
        acc.a.x += v.x * xv.x;
        acc.a.y += v.x * xv.y;
        acc.j[0].x += v.y * xv.x;
        acc.j[0].y += v.y * xv.y;
    }
}

// FoldAcc (per-signal directions): w = c_s.J of the entry, then (a + i w) x, six FMAs
template <bool PG>
__device__ __forceinline__ void acc_entry(const double *vals, int e, double2 xv, FoldAcc &acc) {
    const double4 v = *reinterpret_cast<const double4 *>(vals + 4 * (int64_t)e);
    const double w = fma(acc.cz, v.w, fma(acc.cy, v.z, acc.cx * v.y));
    acc.a.x = fma(v.x, xv.x, fma(-w, xv.y, acc.a.x));
    acc.a.y = fma(v.x, xv.y, fma(w, xv.x, acc.a.y));
}

// FP32 values (the reduced-precision preconditioner input), converted and accumulated in FP64
template <bool PG>
__device__ __forceinline__ void acc_entry(const float *vals, int e, double2 xv, Acc<PG> &acc) {
    if (PG) {
        const float4 v = *reinterpret_cast<const float4 *>(vals + 4 * (int64_t)e);
        const double a = v.x, j0 = v.y, j1 = v.z, j2 = v.w;
        acc.a.x += a * xv.x;
        acc.a.y += a * xv.y;
        acc.j[0].x += j0 * xv.x;
        acc.j[0].y += j0 * xv.y;
        acc.j[1].x += j1 * xv.x;
        acc.j[1].y += j1 * xv.y;
        acc.j[2].x += j2 * xv.x;
        acc.j[2].y += j2 * xv.y;
    } else {
        const float2 v = *reinterpret_cast<const float2 *>(vals + 2 * (int64_t)e);
        const double a = v.x, j0 = v.y;
        acc.a.x += a * xv.x;
        acc.a.y += a * xv.y;
        acc.j[0].x += j0 * xv.x;
        acc.j[0].y += j0 * xv.y;
    }
}
template <bool PG>
__device__ __forceinline__ void acc_entry(const float *vals, int e, double2 xv, FoldAcc &acc) {
    const float4 v = *reinterpret_cast<const float4 *>(vals + 4 * (int64_t)e);
    const double a = v.x;
    const double w = fma(acc.cz, (double)v.w, fma(acc.cy, (double)v.z, acc.cx * (double)v.y));
    acc.a.x = fma(a, xv.x, fma(-w, xv.y, acc.a.x));
    acc.a.y = fma(a, xv.y, fma(w, xv.x, acc.a.y));
}

// Accumulate row entries g, g+G, ... of x = X(col) (X(c) returns the gathered value).
// Two entries per step, so each lane keeps two independent gathers in flight.
template <int G, bool PG, class Gather, class AccT, class V>
__device__ __forceinline__ void row_accumulate(const int *rowptr, const int *col, const V *vals, int64_t row,
                                               int g, Gather X, AccT &acc) {
    acc_zero(acc);
    const int e1 = rowptr[row + 1];
    int e = rowptr[row] + g;
    for (; e + G < e1; e += 2 * G) {
        const int c0 = col[e], c1 = col[e + G];
        const double2 x0 = X(c0), x1 = X(c1);
        acc_entry<PG>(vals, e, x0, acc);
        acc_entry<PG>(vals, e + G, x1, acc);
    }
    if (e < e1) acc_entry<PG>(vals, e, X(col[e]), acc);
}

// Same accumulation with the row's column indices and values already staged in LDS;
// kRowStep entries per step keep that many independent gathers in flight per lane, and the entries
// left go in one masked step (their gathers in flight together, not one after another).
constexpr int kRowStep = 4;
template <int G, bool PG, class Gather, class AccT, class V>
__device__ __forceinline__ void lds_row_accumulate(const int *srp, const int *scol, const V *svals, int lr,
                                                   int g, Gather X, AccT &acc) {
    acc_zero(acc);
    const int e1 = srp[lr + 1];
    int e = srp[lr] + g;
    constexpr int B = kRowStep;
    for (; e + (B - 1) * G < e1; e += B * G) {
        int c[B];
        double2 x[B];
#pragma unroll
        for (int k = 0; k < B; ++k) c[k] = scol[e + k * G];
#pragma unroll
        for (int k = 0; k < B; ++k) x[k] = X(c[k]);
#pragma unroll
        for (int k = 0; k < B; ++k) acc_entry<PG>(svals, e + k * G, x[k], acc);
    }
    if (e < e1) {
        int c[B - 1];
        double2 x[B - 1];
#pragma unroll
        for (int k = 0; k < B - 1; ++k) c[k] = e + k * G < e1 ? scol[e + k * G] : -1;
#pragma unroll
        for (int k = 0; k < B - 1; ++k)
            if (c[k] >= 0) x[k] = X(c[k]);
#pragma unroll
        for (int k = 0; k < B - 1; ++k)
            if (c[k] >= 0) acc_entry<PG>(svals, e + k * G, x[k], acc);
    }
}

// Rows of this block.  Workgroups are dealt to the 8 XCDs round-robin, so block b runs on XCD
// b % 8: every XCD owns one contiguous eighth of the rows (with the solver's Morton numbering,
// a compact region whose gathered columns stay in that XCD's L2), and its blocks sweep that
// eighth together in tiles of TW rows: block k of the XCD takes tiles k, k + grid/8, ...  The
// XCD's rows in flight at any time are one window of (grid/8)*TW consecutive rows, whose
// neighbours' entries are reused from L2 instead of re-fetched.  The grid is a multiple of 8.
constexpr int kXcds = 8;
struct TileWalk {
    int64_t first, stride, end;  // tile starts first, first + stride, ... < end; a tile is TW rows, cut at end
};
__device__ __forceinline__ TileWalk xcd_tiles(int64_t n, int TW) {
    const int per = (int)gridDim.x / kXcds;
    const int x = (int)(blockIdx.x % kXcds), k = (int)(blockIdx.x / kXcds);
    int64_t q = (n + kXcds - 1) / kXcds;
    q = (q + TW - 1) / TW * TW;
    const int64_t xb = (int64_t)x * q < n ? (int64_t)x * q : n;
    const int64_t xe = xb + q < n ? xb + q : n;
    return TileWalk{xb + (int64_t)k * TW, (int64_t)per * TW, xe};
}

struct Lane {
    int lrow, s, g;
};
template <int G>
__device__ __forceinline__ Lane spmv_lane(int S) {
    const int gs = G * S;
    const int rem = threadIdx.x % gs;
    return Lane{(int)threadIdx.x / gs, rem / G, rem % G};
}

// LDS-staged SpMV over the block's rows.  A tile of TR consecutive rows is staged with
// coalesced loads (row offsets, column indices, values), then every (row, signal) lane group
// reads its entries from LDS and gathers x from global memory.  A tile whose entries exceed the
// LDS budget (TR * kStageNnzPerRow) is processed straight from global memory instead.
// emit(row, acc) receives the group-reduced accumulators (acted on by lane g == 0).
// Real part only (RO): y = sum a_e x_col(e) from the separate array a.areal, one value per entry;
// acc.a receives it and acc.j stays zero.
template <int G, bool PG, class Gather, class AccT, class V>
__device__ __forceinline__ void real_row_accumulate(const int *col, const V *av, int e0, int e1, int g, Gather X,
                                                    AccT &acc) {
    acc_zero(acc);
    int e = e0 + g;
    for (; e + 3 * G < e1; e += 4 * G) {
        const int c0 = col[e], c1 = col[e + G], c2 = col[e + 2 * G], c3 = col[e + 3 * G];
        const double2 x0 = X(c0), x1 = X(c1), x2 = X(c2), x3 = X(c3);
        const double v0 = av[e], v1 = av[e + G], v2 = av[e + 2 * G], v3 = av[e + 3 * G];
        acc.a.x += v0 * x0.x + v1 * x1.x + v2 * x2.x + v3 * x3.x;
        acc.a.y += v0 * x0.y + v1 * x1.y + v2 * x2.y + v3 * x3.y;
    }
    if (e < e1) {
        int c[3];
        double2 x[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = e + k * G < e1 ? col[e + k * G] : -1;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (c[k] >= 0) x[k] = X(c[k]);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (c[k] >= 0) {
                const double v = av[e + k * G];
                acc.a.x += v * x[k].x;
                acc.a.y += v * x[k].y;
            }
    }
}

// Copy rows t0 .. t0+nr-1 of the pattern into LDS (the caller holds a barrier before and after):
// their ne entries' column indices (scol) and values (svals: VW per entry from vals, which is
// a.vals, a.areal (RO: one per entry) or their FP32 copy a.pvals32), and the row offsets relative
// to the tile's first entry (srp).  Loads go through registers in batches of kStageBatch
// independent loads per thread, so a thread waits on memory once per batch rather than once per
// element.
template <bool RO, int VW, class V>
__device__ __forceinline__ void stage_tile(const SolveArgs &a, const V *vals, int64_t t0, int nr, int E0, int ne,
                                           V *svals, int *scol, int *srp) {
    const int nt = blockDim.x;
    const int rp = threadIdx.x <= nr ? a.rowptr[t0 + threadIdx.x] : 0;
    if (RO) {
        const V *gv = vals + E0;
        for (int base = 0; base < ne; base += kStageBatch * nt) {
            int cv[kStageBatch];
            V vv[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = base + u * nt + threadIdx.x;
                const bool ok = i < ne;
                cv[u] = ok ? a.col[E0 + i] : 0;
                vv[u] = ok ? gv[i] : V(0);
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = base + u * nt + threadIdx.x;
                if (i < ne) {
                    scol[i] = cv[u];
                    svals[i] = vv[u];
                }
            }
        }
    } else {
        // an entry's VW values move as NC chunks of 16 bytes (8 for two FP32 values)
        using Chunk = std::conditional_t<(VW * sizeof(V)) % 16 == 0, uint4, uint2>;
        constexpr int NC = (int)(VW * sizeof(V) / sizeof(Chunk));
        const Chunk *gv = reinterpret_cast<const Chunk *>(vals + (int64_t)E0 * VW);
        Chunk *lv = reinterpret_cast<Chunk *>(svals);
        for (int base = 0; base < ne; base += kStageBatch * nt) {
            int cv[kStageBatch];
            Chunk vv[kStageBatch][NC];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = base + u * nt + threadIdx.x;
                const bool ok = i < ne;
                cv[u] = ok ? a.col[E0 + i] : 0;
#pragma unroll
                for (int h = 0; h < NC; ++h) vv[u][h] = ok ? gv[(int64_t)i * NC + h] : Chunk{};
            }
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) {
                const int i = base + u * nt + threadIdx.x;
                if (i < ne) {
                    scol[i] = cv[u];
#pragma unroll
                    for (int h = 0; h < NC; ++h) lv[i * NC + h] = vv[u][h];
                }
            }
        }
    }
    if ((int)threadIdx.x <= nr) srp[threadIdx.x] = rp - E0;
    for (int i = threadIdx.x + nt; i <= nr; i += nt) srp[i] = a.rowptr[t0 + i] - E0;
}

// The operator values a staged SpMV reads: FP64 (a.vals, or a.areal for the real part alone) or
// their FP32 copy (a.pvals32, laid out the same way)
template <class V, bool RO>
__device__ __forceinline__ const V *spmv_values(const SolveArgs &a) {
    if constexpr (std::is_same_v<V, float>) return a.pvals32;
    else return RO ? a.areal : a.vals;
}

template <int G, bool PG, bool RO = false, class V = double, class Gather, class Emit, class AccT = Acc<PG>>
__device__ __forceinline__ void staged_spmv(const SolveArgs &a, int rpb, int TR, const Lane &L, bool act, Gather X,
                                            Emit emit, AccT proto = AccT{}) {
    constexpr int VW = RO ? 1 : (PG ? 4 : 2);
    const V *vals = spmv_values<V, RO>(a);
    if (TR <= 0) {  // direct: entries straight from global memory
        const TileWalk W = xcd_tiles(a.n, rpb);
        if (!act) return;
        for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
            const int64_t row = t0 + L.lrow;
            if (row >= W.end) continue;
            AccT acc = proto;
            if (RO)
                real_row_accumulate<G, PG>(a.col, vals, a.rowptr[row], a.rowptr[row + 1], L.g, X, acc);
            else
                row_accumulate<G, PG>(a.rowptr, a.col, vals, row, L.g, X, acc);
            group_reduce_acc<G, PG>(acc);
            emit(row, acc);
        }
        return;
    }
    const int ECAP = TR * kStageNnzPerRow;
    extern __shared__ double4 dmri_stage_lds[];
    V *svals = reinterpret_cast<V *>(dmri_stage_lds);
    int *scol = reinterpret_cast<int *>(svals + (size_t)ECAP * VW);
    int *srp = scol + ECAP;
    const TileWalk W = xcd_tiles(a.n, TR);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int nr = (int)((W.end - t0) < TR ? (W.end - t0) : TR);
        const int E0 = a.rowptr[t0];
        const int ne = a.rowptr[t0 + nr] - E0;
        const bool fits = ne <= ECAP;
        __syncthreads();
        if (fits) stage_tile<RO, VW>(a, vals, t0, nr, E0, ne, svals, scol, srp);
        __syncthreads();
        if (!act) continue;
        for (int lr = L.lrow; lr < nr; lr += rpb) {
            AccT acc = proto;
            if (RO) {
                if (fits) real_row_accumulate<G, PG>(scol, svals, srp[lr], srp[lr + 1], L.g, X, acc);
                else real_row_accumulate<G, PG>(a.col, vals, a.rowptr[t0 + lr], a.rowptr[t0 + lr + 1], L.g, X, acc);
            } else {
                if (fits) lds_row_accumulate<G, PG>(srp, scol, svals, lr, L.g, X, acc);
                else row_accumulate<G, PG>(a.rowptr, a.col, vals, t0 + lr, L.g, X, acc);
            }
            group_reduce_acc<G, PG>(acc);
            emit(t0 + lr, acc);
        }
    }
}

// Accumulation of entries e0 + g, e0 + g + G, ... < e1 of a row: the implicit values (acc) and the
// explicit real part sb = sum bre_e x_col(e); the pointers are the staged LDS copies or the
// global arrays.
template <int G, bool PG, class Gather>
__device__ __forceinline__ void accumulate_with_bre(const int *col, const double *vals, const double *bre, int e0,
                                                    int e1, int g, Gather X, Acc<PG> &acc, double2 &sb) {
    acc.a = make_double2(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < (PG ? 3 : 1); ++d) acc.j[d] = make_double2(0.0, 0.0);
    sb = make_double2(0.0, 0.0);
    int e = e0 + g;
    for (; e + G < e1; e += 2 * G) {
        const int c0 = col[e], c1 = col[e + G];
        const double2 x0 = X(c0), x1 = X(c1);
        acc_entry<PG>(vals, e, x0, acc);
        acc_entry<PG>(vals, e + G, x1, acc);
        const double b0 = bre[e], b1 = bre[e + G];
        sb.x += b0 * x0.x + b1 * x1.x;
        sb.y += b0 * x0.y + b1 * x1.y;
    }
    if (e < e1) {
        const double2 x0 = X(col[e]);
        acc_entry<PG>(vals, e, x0, acc);
        sb.x += bre[e] * x0.x;
        sb.y += bre[e] * x0.y;
    }
}

// staged_spmv with the explicit real part: a tile's column indices, values and bre are staged in
// LDS (stage_lds_bytes_bre); emit(row, acc, sb) receives the group-reduced sums.
template <int G, bool PG, class Gather, class Emit>
__device__ __forceinline__ void staged_spmv_bre(const SolveArgs &a, int rpb, int TR, const Lane &L, Gather X,
                                                Emit emit) {
    constexpr int VW = PG ? 4 : 2;
    if (TR <= 0) {
        const TileWalk W = xcd_tiles(a.n, rpb);
        for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
            const int64_t row = t0 + L.lrow;
            if (row >= W.end) continue;
            Acc<PG> acc;
            double2 sb;
            accumulate_with_bre<G, PG>(a.col, a.vals, a.bre, a.rowptr[row], a.rowptr[row + 1], L.g, X, acc, sb);
            group_reduce_acc<G, PG>(acc);
            group_reduce<G>(sb);
            emit(row, acc, sb);
        }
        return;
    }
    const int ECAP = TR * kStageNnzPerRow;
    extern __shared__ double4 dmri_stage_lds[];
    double *svals = reinterpret_cast<double *>(dmri_stage_lds);
    double *sbre = svals + (size_t)ECAP * VW;
    int *scol = reinterpret_cast<int *>(sbre + ECAP);
    int *srp = scol + ECAP;
    const TileWalk W = xcd_tiles(a.n, TR);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int nr = (int)((W.end - t0) < TR ? (W.end - t0) : TR);
        const int E0 = a.rowptr[t0];
        const int ne = a.rowptr[t0 + nr] - E0;
        const bool fits = ne <= ECAP;
        __syncthreads();
        if (fits) {
            const double2 *gv = reinterpret_cast<const double2 *>(a.vals + (int64_t)E0 * VW);
            double2 *lv = reinterpret_cast<double2 *>(svals);
            const int nt = blockDim.x;
            const int rp = threadIdx.x <= nr ? a.rowptr[t0 + threadIdx.x] : 0;
            for (int base = 0; base < ne; base += kStageBatch * nt) {
                int cv[kStageBatch];
                double bv[kStageBatch];
                double2 vv[kStageBatch][VW / 2];
#pragma unroll
                for (int u = 0; u < kStageBatch; ++u) {
                    const int i = base + u * nt + threadIdx.x;
                    const bool ok = i < ne;
                    cv[u] = ok ? a.col[E0 + i] : 0;
                    bv[u] = ok ? a.bre[E0 + i] : 0.0;
#pragma unroll
                    for (int h = 0; h < VW / 2; ++h)
                        vv[u][h] = ok ? gv[(int64_t)i * (VW / 2) + h] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < kStageBatch; ++u) {
                    const int i = base + u * nt + threadIdx.x;
                    if (i < ne) {
                        scol[i] = cv[u];
                        sbre[i] = bv[u];
#pragma unroll
                        for (int h = 0; h < VW / 2; ++h) lv[i * (VW / 2) + h] = vv[u][h];
                    }
                }
            }
            if ((int)threadIdx.x <= nr) srp[threadIdx.x] = rp - E0;
            for (int i = threadIdx.x + nt; i <= nr; i += nt) srp[i] = a.rowptr[t0 + i] - E0;
        }
        __syncthreads();
        for (int lr = L.lrow; lr < nr; lr += rpb) {
            Acc<PG> acc;
            double2 sb;
            if (fits)
                accumulate_with_bre<G, PG>(scol, svals, sbre, srp[lr], srp[lr + 1], L.g, X, acc, sb);
            else
                accumulate_with_bre<G, PG>(a.col, a.vals, a.bre, a.rowptr[t0 + lr], a.rowptr[t0 + lr + 1], L.g, X,
                                           acc, sb);
            group_reduce_acc<G, PG>(acc);
            group_reduce<G>(sb);
            emit(t0 + lr, acc, sb);
        }
    }
}

// Mode 4 of the residual kernel in the SpMV kernels' (staged) row mapping: b = (Bre + i cB.J) x
// (+ the periodic right-hand side), hp = Are x, hq = Jg x.
template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) rhs_history_kernel(SolveArgs a, int S, int rpb, int TR) {
    const Lane L = spmv_lane<G>(S);
    const int S_ = S;
    const int s = L.s;
    const double2 *x = a.x;
    staged_spmv_bre<G, PG>(
        a, rpb, TR, L, [&](int c) { return x[(int64_t)c * S_ + s]; },
        [&](int64_t row, const Acc<PG> &acc, double2 sb) {
            if (L.g != 0) return;
            const int64_t idx = row * S_ + s;
            Acc<PG> bacc = acc;
            bacc.a = sb;
            double2 b = imag_coef_apply<G, PG>(bacc, a.coefB, s);
            if (a.bper) {
                const double2 bp = a.bper[idx];
                b.x += bp.x;
                b.y += bp.y;
            }
            a.b[idx] = b;
            a.hp[idx] = acc.a;
            double2 q = acc.j[0];
            if (PG) {
                const double g0 = a.gsig[3 * s], g1 = a.gsig[3 * s + 1], g2 = a.gsig[3 * s + 2];
                q.x = g0 * acc.j[0].x + g1 * acc.j[1].x + g2 * acc.j[2].x;
                q.y = g0 * acc.j[0].y + g1 * acc.j[1].y + g2 * acc.j[2].y;
            }
            a.hq[idx] = q;
        });
}

// mode 0: b = (Bre + i cB.J) x, r = b - (Are + i cA.J) x, r^ = r          (start of a step)
// mode 1: r = b - (Are + i cA.J) x with the stored b, for the masked signals (restart, or
//         after an extrapolated initial guess; mask == nullptr selects every signal)
// mode 2: b = (Bre + i cB.J) x only, and xpp <- gc0 x + gc1 xprev + gc2 xpp (the extrapolated guess)
// mode 3: b = (Bre + i cB.J) x only (the adjoint's lambda^n = B_n mu_n)
// mode 4: b = (Bre + i cB.J) x, hp = Are x, hq = Jg x (the projection history of state x)
// partials U (modes 0, 1): [|r|^2, 0, |r|^2, |b|^2]
template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) rhs_residual_kernel(SolveArgs a, int S, int rpb, int mode,
                                                                const int *mask) {
    const Lane L = spmv_lane<G>(S);
    const bool sel = mode != 1 || mask == nullptr || mask[L.s];
    const int S_ = S;
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const TileWalk W = xcd_tiles(a.n, rpb);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int64_t row = t0 + L.lrow;
        if (!sel || row >= W.end) continue;
        const double2 *x = a.x;
        const int s = L.s;
        Acc<PG> acc;
        if (mode != 2) {
            row_accumulate<G, PG>(a.rowptr, a.col, a.vals, row, L.g,
                                  [&](int c) { return x[(int64_t)c * S_ + s]; }, acc);
            group_reduce_acc<G, PG>(acc);
        }
        double2 sb = make_double2(0.0, 0.0);
        Acc<PG> jacc;
        const bool jsep = mode == 2 || mode == 3;  // J part of x not in acc: accumulate it here
        if (mode != 1) {
            jacc.a = make_double2(0.0, 0.0);
#pragma unroll
            for (int d = 0; d < (PG ? 3 : 1); ++d) jacc.j[d] = make_double2(0.0, 0.0);
            const int e1 = a.rowptr[row + 1];
            for (int e = a.rowptr[row] + L.g; e < e1; e += G) {
                const double bv = a.bre[e];
                const double2 xv = x[(int64_t)a.col[e] * S_ + s];
                sb.x += bv * xv.x;
                sb.y += bv * xv.y;
                if (jsep) {
                    // J part of the explicit side (modes 0 and 4 reuse the implicit accumulation)
                    if (PG) {
                        const double4 v = *reinterpret_cast<const double4 *>(a.vals + 4 * (int64_t)e);
                        jacc.j[0].x += v.y * xv.x; jacc.j[0].y += v.y * xv.y;
                        jacc.j[1].x += v.z * xv.x; jacc.j[1].y += v.z * xv.y;
                        jacc.j[2].x += v.w * xv.x; jacc.j[2].y += v.w * xv.y;
                    } else {
                        const double jv = a.vals[2 * (int64_t)e + 1];
                        jacc.j[0].x += jv * xv.x; jacc.j[0].y += jv * xv.y;
                    }
                }
            }
            group_reduce<G>(sb);
            if (jsep) group_reduce_acc<G, PG>(jacc);
        }
        if (L.g == 0) {
            const int64_t idx = row * S_ + s;
            double2 b;
            if (mode == 1) {
                b = a.b[idx];
            } else {
                Acc<PG> bacc = jsep ? jacc : acc;
                bacc.a = sb;
                b = imag_coef_apply<G, PG>(bacc, a.coefB, s);
                if (a.bper) {
                    const double2 bp = a.bper[idx];
                    b.x += bp.x;
                    b.y += bp.y;
                }
                a.b[idx] = b;
            }
            if (mode == 2) {
                const double2 xi = x[idx], xp = a.xprev[idx];
                double2 g = make_double2(a.gc0 * xi.x + a.gc1 * xp.x, a.gc0 * xi.y + a.gc1 * xp.y);
                if (a.gc2 != 0.0) {
                    const double2 xq = a.xpp[idx];
                    g.x += a.gc2 * xq.x;
                    g.y += a.gc2 * xq.y;
                }
                a.xpp[idx] = g;
            } else if (mode == 3) {
            } else if (mode == 4) {
                a.hp[idx] = acc.a;
                double2 q = acc.j[0];
                if (PG) {
                    const double g0 = a.gsig[3 * s], g1 = a.gsig[3 * s + 1], g2 = a.gsig[3 * s + 2];
                    q.x = g0 * acc.j[0].x + g1 * acc.j[1].x + g2 * acc.j[2].x;
                    q.y = g0 * acc.j[0].y + g1 * acc.j[1].y + g2 * acc.j[2].y;
                }
                a.hq[idx] = q;
            } else {
                const double2 ax = imag_coef_apply<G, PG>(acc, a.coefA, s);
                const double2 r = make_double2(b.x - ax.x, b.y - ax.y);
                a.r[idx] = r;
                a.rh[idx] = r;
                const double rr = r.x * r.x + r.y * r.y;
                red[0] += rr;
                red[2] += rr;
                red[3] += b.x * b.x + b.y * b.y;
            }
        }
    }
    if (mode < 2) block_signal_partials(red, L.g == 0, L.lrow * S + L.s, rpb * S, S, a.partU);
}

// Iteration head: totals of U give rho_i = <r^, r_i> and |r_i|^2; convergence, breakdown
// and iteration-limit decisions; p = r (first iteration) or p = r + beta (p - omega v).
// The partials of U come from the residual kernel (SpMV grid) on a signal's first iteration
// and from the update kernel (vector grid) afterwards; without a finalize pass S is 1.
__global__ void __launch_bounds__(kThreads) bicg_p_kernel(SolveArgs a, int S, int vrpb, int nparts_rhs,
                                                          int nparts_upd, int it) {
    if (*a.all_done) return;
    __shared__ double tot[kMaxSignals][kNV];
    __shared__ double2 sbeta[kMaxSignals], somega[kMaxSignals];
    __shared__ int sact[kMaxSignals], sfirst[kMaxSignals];
    const int nparts = (it - a.st[0].it0) == 0 ? nparts_rhs : nparts_upd;
    signal_totals(a.partU, a.totU, nparts, S, tot);
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        KrylovState &st = a.st[s];
        int act = !st.done;
        const int local = it - st.it0;
        sfirst[s] = local == 0;
        if (act) {
            const double2 rho = make_double2(tot[s][0], tot[s][1]);
            const double rn2 = tot[s][2];
            double tol2;
            if (local == 0) {
                const double t1 = a.rtol * a.rtol * tot[s][3], t2 = a.atol * a.atol;
                tol2 = t1 > t2 ? t1 : t2;
            } else {
                tol2 = st.tol2;
            }
            const bool conv = rn2 <= tol2;
            const double2 om_prev = local > 0 ? st.omega[(local - 1) & 1] : make_double2(1.0, 0.0);
            const bool brk = !conv && local > 0 &&
                             ((rho.x == 0.0 && rho.y == 0.0) || (om_prev.x == 0.0 && om_prev.y == 0.0));
            const bool lim = !conv && local >= a.maxit;
            const bool newly = conv || brk || lim;
            if (blockIdx.x == 0) {
                if (local == 0) {
                    st.tol2 = tol2;
                    st.bnorm2 = tot[s][3];
                }
                st.rnorm2 = rn2;
                st.iters = it;
                if (newly) {
                    st.done = 1;
                    st.conv = conv ? 1 : 0;
                    st.breakdown = brk ? 1 : 0;
                } else {
                    st.rho[local & 1] = rho;
                }
            }
            if (newly) {
                act = 0;
            } else if (local > 0) {
                const double2 rho_prev = st.rho[(local - 1) & 1];
                const double2 al_prev = st.alpha[(local - 1) & 1];
                sbeta[s] = c_mul(c_div(rho, rho_prev), c_div(al_prev, om_prev));
                somega[s] = om_prev;
            }
        }
        sact[s] = act;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int all = 1;
        for (int s = 0; s < S; ++s)
            if (sact[s]) all = 0;
        if (all) *a.all_done = 1;
    }
    const int64_t total = a.n * S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S || !sact[s]) return;
    const bool first = sfirst[s];
    const double2 beta = sbeta[s], omega = somega[s];
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
        const double2 ri = a.r[i];
        if (first) {
            a.p[i] = ri;
        } else {
            const double2 w = c_sub(a.p[i], c_mul(omega, a.v[i]));
            a.p[i] = c_add(ri, c_mul(beta, w));
        }
    }
}

__device__ __forceinline__ void cocg_head(const SolveArgs &a, int S, int nparts_rhs, int nparts_upd, int nparts_spmv,
                                          int it, int rfold, double2 *sbeta, double2 *salpha, int *sact, int *sfirst,
                                          int *sx);

// OP kOpV:  v = A p ; partials V: [<r^, v>.re, <r^, v>.im, 0, 0]
// OP kOpPre: ph = c (2 p - c Are p) (a degree-1 polynomial preconditioner of the scaled system,
//            whose diagonal is 1; c from a bound of its spectrum), no partials
// With per-signal directions (PG) the staged kernels are held to at most 80 VGPRs, so 6 waves
// per SIMD fit (with their 16-row tiles' LDS, 6 blocks per CU) instead of 5: more gathers in flight
// per CU, which is what bounds these kernels (C4, 32 signals: 3.8 % less time per theta step; 7
// waves spilled too much).
template <int G, bool PG, int OP>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(PG ? 6 : 1)))
spmv_v_kernel(SolveArgs a, int S, int rpb, int TR, int it) {
    if (*a.all_done) return;
    const Lane L = spmv_lane<G>(S);
    // kOpPreDotR: COCG's update of r folded in.  The head turns the previous p-kernel's stop into
    // done, forms alpha = rho_prev / p^T q from the previous q = A p (partials T, one per block of
    // that launch, which had this grid) and raises the all-done flag when no signal is left; every
    // block reads the same values, so all agree.  State bits per signal: 1 active, 2 r update
    // pending (every iteration but a solve's first), 4 r_(k-1) in a.s (k - 1 odd).
    constexpr bool fold_r = OP == kOpPreDotR;
    __shared__ double2 fr_alpha[fold_r ? kMaxSignals : 1];
    __shared__ int fr_state[fold_r ? kMaxSignals : 1];
    if constexpr (fold_r) {
        __shared__ double ftot[kMaxSignals][kNV];
        signal_totals(a.partT, a.totT, (int)gridDim.x, S, ftot);
        if ((int)threadIdx.x < S) {
            const int sg = threadIdx.x;
            KrylovState &st = a.st[sg];
            if (st.stop && blockIdx.x == 0) st.done = 1;  // stopped by the previous p-kernel
            int act = !st.done && !st.stop;
            const int local = it - st.it0;
            double2 al = make_double2(0.0, 0.0);
            if (act && local > 0) {
                const double2 pq = make_double2(ftot[sg][0], ftot[sg][1]);
                if (pq.x == 0.0 && pq.y == 0.0) {
                    if (blockIdx.x == 0) {
                        st.done = 1;
                        st.breakdown = 1;
                        st.conv = 0;
                    }
                    act = 0;
                } else {
                    al = c_div(st.rho[(local - 1) & 1], pq);
                    if (blockIdx.x == 0) st.alpha[(local - 1) & 1] = al;
                }
            }
            fr_alpha[sg] = al;
            fr_state[sg] = (act ? 1 : 0) | (act && local > 0 ? 2 : 0) | (local > 0 && ((local - 1) & 1) ? 4 : 0);
        }
        __syncthreads();
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            int any = 0;
            for (int sg = 0; sg < S; ++sg) any |= fr_state[sg] & 1;
            if (!any) *a.all_done = 1;
        }
    }
    // kOpQFold: cocg_p_kernel's head; a lane also runs for a signal that stops here, to apply its
    // pending x update
    constexpr bool qfold = OP == kOpQFold;
    __shared__ double2 qf_beta[qfold ? kMaxSignals : 1], qf_alpha[qfold ? kMaxSignals : 1];
    __shared__ int qf_act[qfold ? kMaxSignals : 1], qf_first[qfold ? kMaxSignals : 1], qf_x[qfold ? kMaxSignals : 1];
    if constexpr (qfold)
        cocg_head(a, S, (int)gridDim.x, a.nparts_upd, (int)gridDim.x, it, 0, qf_beta, qf_alpha, qf_act, qf_first,
                  qf_x);
    const int fst = fold_r ? fr_state[L.s] : 0;
    const bool act = fold_r ? (fst & 1) != 0
                   : qfold  ? (qf_act[L.s] | qf_x[L.s]) != 0
                            : !a.st[L.s].done;
    const bool rupd = (fst & 2) != 0;
    const double2 ralpha = fold_r ? fr_alpha[L.s] : make_double2(0.0, 0.0);
    const double2 *rin = (fst & 4) ? a.s : a.r;
    double2 *rout = (fst & 4) ? a.r : a.s;
    const int S_ = S;
    const int s = L.s;
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *p = qfold ? a.ph : a.p;  // the gathered vector (z for the folded p-step)
    // kOpPreCD, with the complex diagonal d: ph = c D^-1 (2 p - c A D^-1 p), A D^-1 p gathering p / d
    constexpr bool cd = OP == kOpPreCD || OP == kOpPreCDDot;
    // kOpPreDot32: the input as FP32 pairs (a.r32) and FP32 operator values, summed in FP64
    constexpr bool f32 = OP == kOpPreDot32;
    constexpr bool pre = OP == kOpPre || OP == kOpPreCD || OP == kOpPreDot || OP == kOpPreCDDot || f32 || fold_r;
    // COCG's operations: the unconjugated product of the input's own row with the output
    constexpr bool pdot = OP == kOpVDot || OP == kOpPreDot || OP == kOpPreCDDot || f32 || fold_r;
    auto input = [&](int64_t i) {
        if constexpr (f32) {
            const float2 z = a.r32[i];
            return make_double2(z.x, z.y);
        } else if constexpr (fold_r) {
            const double2 r0 = rin[i];
            return rupd ? c_sub(r0, c_mul(ralpha, a.v[i])) : r0;
        } else {
            return p[i];
        }
    };
    auto gather = [&](int c) {
        const double2 z = input((int64_t)c * S_ + s);
        if constexpr (cd) return c_mul(z, cdiag_inv<PG>(a.pdiag, a.coefA, c, s));
        return z;
    };
    auto emit = [&](int64_t row, const auto &acc) {
        if (L.g != 0) return;
        const int64_t idx = row * S_ + s;
        const double2 y = imag_coef_apply<G, PG>(acc, a.coefA, s);
        if (pre) {
            const double2 z = input(idx);
            const double2 w = make_double2(a.pc * (2.0 * z.x - a.pc * y.x), a.pc * (2.0 * z.y - a.pc * y.y));
            double2 out = w;
            if constexpr (cd) out = c_mul(w, cdiag_inv<PG>(a.pdiag, a.coefA, row, s));
            a.ph[idx] = out;
            if constexpr (pdot) {
                const double2 zo = c_mul(z, out);
                red[0] += zo.x;
                red[1] += zo.y;
            }
            if constexpr (fold_r) {
                if (rupd) rout[idx] = z;
                red[2] += z.x * z.x + z.y * z.y;
            }
            return;
        }
        if constexpr (qfold) {
            const int sg = L.s;
            const double2 z = a.ph[idx];
            if (qf_first[sg]) {  // a solve's first iteration: p = z, q = A z
                if (qf_act[sg]) {
                    a.p[idx] = z;
                    a.v[idx] = y;
                    const double2 pq = c_mul(z, y);
                    red[0] += pq.x;
                    red[1] += pq.y;
                }
                return;
            }
            const double2 pp = a.p[idx];
            if (qf_x[sg]) a.x[idx] = c_add(a.x[idx], c_mul(qf_alpha[sg], pp));
            if (qf_act[sg]) {
                const double2 b = qf_beta[sg];
                const double2 pn = c_add(z, c_mul(b, pp));
                const double2 qn = c_add(y, c_mul(b, a.v[idx]));
                a.p[idx] = pn;
                a.v[idx] = qn;
                const double2 pq = c_mul(pn, qn);
                red[0] += pq.x;
                red[1] += pq.y;
            }
            return;
        }
        a.v[idx] = y;
        if constexpr (pdot) {
            const double2 py = c_mul(p[idx], y);
            red[0] += py.x;
            red[1] += py.y;
        } else {
            const double2 h = a.rh[idx];
            red[0] += h.x * y.x + h.y * y.y;
            red[1] += h.x * y.y - h.y * y.x;
        }
    };
    // the preconditioner takes the real part alone when a.areal is set; with per-signal directions
    // the signal's coefficients are folded into every entry
    using V = std::conditional_t<f32, float, double>;
    if ((OP == kOpPre || OP == kOpPreDot || f32 || fold_r) && a.areal)
        staged_spmv<G, PG, true, V>(a, rpb, TR, L, act, gather, emit);
    else if constexpr (PG)
        staged_spmv<G, PG, false, V>(a, rpb, TR, L, act, gather, emit, fold_acc(a.coefA, s));
    else
        staged_spmv<G, PG, false, V>(a, rpb, TR, L, act, gather, emit);
    if (OP == kOpV || OP == kOpPreDot || OP == kOpPreCDDot || f32 || fold_r)
        block_signal_partials(red, L.g == 0, L.lrow * S + L.s, rpb * S, S, a.partV);
    if (OP == kOpVDot || qfold) block_signal_partials(red, L.g == 0, L.lrow * S + L.s, rpb * S, S, a.partT);
}

// r = b - (Are + i cA.J) x with the stored b, r^ = r, in the SpMV kernels' staged row mapping
// (the unmasked mode 1 of the residual kernel); partials U: [|r|^2, 0, |r|^2, |b|^2]
template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) residual_kernel(SolveArgs a, int S, int rpb, int TR) {
    const Lane L = spmv_lane<G>(S);
    const int S_ = S;
    const int s = L.s;
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *x = a.x;
    staged_spmv<G, PG>(
        a, rpb, TR, L, true, [&](int c) { return x[(int64_t)c * S_ + s]; },
        [&](int64_t row, const Acc<PG> &acc) {
            if (L.g != 0) return;
            const int64_t idx = row * S_ + s;
            const double2 ax = imag_coef_apply<G, PG>(acc, a.coefA, s);
            const double2 b = a.b[idx];
            const double2 r = make_double2(b.x - ax.x, b.y - ax.y);
            a.r[idx] = r;
            a.rh[idx] = r;
            const double rr = r.x * r.x + r.y * r.y;
            red[0] += rr;
            red[2] += rr;
            red[3] += b.x * b.x + b.y * b.y;
        });
    block_signal_partials(red, L.g == 0, L.lrow * S + L.s, rpb * S, S, a.partU);
}

// Head of every t = A s kernel: alpha_s = rho_s / <r^, v>_s per signal into salpha, the
// active flags into sact (breakdown when <r^, v> = 0).
__device__ __forceinline__ void spmv_t_head(const SolveArgs &a, int S, int nparts, int it, double2 *salpha,
                                            int *sact) {
    __shared__ double tot[kMaxSignals][kNV];
    signal_totals(a.partV, a.totV, nparts, S, tot);
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        KrylovState &st = a.st[s];
        int act = !st.done;
        const int local = it - st.it0;
        if (act) {
            const double2 rv = make_double2(tot[s][0], tot[s][1]);
            if (rv.x == 0.0 && rv.y == 0.0) {
                if (blockIdx.x == 0) {
                    st.done = 1;
                    st.breakdown = 1;
                    st.conv = 0;
                    st.iters = it;
                }
                act = 0;
            } else {
                const double2 al = c_div(st.rho[local & 1], rv);
                salpha[s] = al;
                if (blockIdx.x == 0) st.alpha[local & 1] = al;
            }
        }
        sact[s] = act;
    }
    __syncthreads();
}

// alpha = rho / <r^, v> and s = r - alpha v (formed for the gathered columns), then
// OP kOpT:   t = A s ; partials T: [<t, s>.re, <t, s>.im, |t|^2, |s|^2]
// OP kOpPre: sh = c (2 s - c Are s) (the polynomial preconditioner applied to s), no partials
// OP kOpTP:  t = A sh with sh gathered from a.sh; partials T as for kOpT (s at the own row)
template <int G, bool PG, int OP>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(PG ? 6 : 1)))
spmv_t_kernel(SolveArgs a, int S, int rpb, int TR, int nparts, int it) {
    if (*a.all_done) return;
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    spmv_t_head(a, S, nparts, it, salpha, sact);
    const Lane L = spmv_lane<G>(S);
    const bool act = sact[L.s];
    // The preconditioned path (every OP but kOpT) reads s as the s-pass materialised it (the
    // launcher checks that it is there); plain BiCGStab forms s = r - alpha v where it reads it.
    constexpr bool given_s = OP != kOpT;
    const double2 alpha = given_s ? make_double2(0.0, 0.0) : salpha[L.s];
    const int S_ = S;
    const int s = L.s;
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *r = a.r, *v = a.v, *sh = a.sh;
    constexpr bool cd = OP == kOpPreCD;  // complex diagonal, as in spmv_v_kernel
    constexpr bool pre = OP == kOpPre || OP == kOpPreCD;
    const double2 *sm = a.s;
    auto gather = [&](int c) {
        const int64_t j = (int64_t)c * S_ + s;
        if (OP == kOpTP) return sh[j];
        const double2 z = given_s ? sm[j] : c_sub(r[j], c_mul(alpha, v[j]));
        if constexpr (cd) return c_mul(z, cdiag_inv<PG>(a.pdiag, a.coefA, c, s));
        return z;
    };
    auto emit = [&](int64_t row, const auto &acc) {
        if (L.g != 0) return;
        const int64_t idx = row * S_ + s;
        const double2 y = imag_coef_apply<G, PG>(acc, a.coefA, s);
        const double2 sv = given_s ? sm[idx] : c_sub(r[idx], c_mul(alpha, v[idx]));
        if (pre) {
            const double2 w = make_double2(a.pc * (2.0 * sv.x - a.pc * y.x), a.pc * (2.0 * sv.y - a.pc * y.y));
            if constexpr (cd)
                a.sh[idx] = c_mul(w, cdiag_inv<PG>(a.pdiag, a.coefA, row, s));
            else
                a.sh[idx] = w;
            return;
        }
        a.t[idx] = y;
        red[0] += y.x * sv.x + y.y * sv.y;
        red[1] += y.x * sv.y - y.y * sv.x;
        red[2] += y.x * y.x + y.y * y.y;
        red[3] += sv.x * sv.x + sv.y * sv.y;
    };
    if (OP == kOpPre && a.areal)
        staged_spmv<G, PG, true>(a, rpb, TR, L, act, gather, emit);
    else if constexpr (PG)
        staged_spmv<G, PG, false>(a, rpb, TR, L, act, gather, emit, fold_acc(a.coefA, s));
    else
        staged_spmv<G, PG, false>(a, rpb, TR, L, act, gather, emit);
    if (!pre) block_signal_partials(red, L.g == 0, L.lrow * S + L.s, rpb * S, S, a.partT);
}

// ---- two signals per lane (kMapPair): LDS-staged tiles as in staged_spmv, but a lane serves
// signals 2k and 2k+1 of its row, so every staged column index and value it reads, and every
// address it forms, serves two signals, and it keeps two gathers in flight per entry.  A row takes
// LP = G x SP lanes (SP: signal pairs rounded up to a power of two; G lanes split its entries), a
// wave holds 64 / LP whole rows.  Each entry's imaginary coefficient is folded in before the
// multiply (a + i c_s.j), so a signal needs one complex accumulator.
struct PairLane {
    int lrow, sp, g;
};
template <int G>
__device__ __forceinline__ PairLane pair_lane(int SP) {
    const int LP = G * SP;
    const int rem = threadIdx.x % LP;
    return PairLane{(int)threadIdx.x / LP, rem / G, rem % G};
}

// One entry's contribution (a + i j) x to both signals' sums, four FMAs per signal.
template <bool PG, bool RO>
__device__ __forceinline__ void pair_entry(const double *vals, int64_t e, const double (&cf)[2][3], const double2 &x0,
                                           const double2 &x1, double2 (&y)[2]) {
    constexpr int VW = RO ? 1 : (PG ? 4 : 2);
    const int64_t ek = e * VW;
    double av, j0 = 0.0, j1 = 0.0;
    if (RO) {
        av = vals[ek];
    } else if (PG) {
        const double4 v = *reinterpret_cast<const double4 *>(vals + ek);
        av = v.x;
        j0 = fma(cf[0][2], v.w, fma(cf[0][1], v.z, cf[0][0] * v.y));
        j1 = fma(cf[1][2], v.w, fma(cf[1][1], v.z, cf[1][0] * v.y));
    } else {
        const double2 v = *reinterpret_cast<const double2 *>(vals + ek);
        av = v.x;
        j0 = cf[0][0] * v.y;
        j1 = cf[1][0] * v.y;
    }
    y[0].x = fma(av, x0.x, y[0].x);
    y[0].y = fma(av, x0.y, y[0].y);
    y[1].x = fma(av, x1.x, y[1].x);
    y[1].y = fma(av, x1.y, y[1].y);
    if (!RO) {
        y[0].x = fma(-j0, x0.y, y[0].x);
        y[0].y = fma(j0, x0.x, y[0].y);
        y[1].x = fma(-j1, x1.y, y[1].x);
        y[1].y = fma(j1, x1.x, y[1].y);
    }
}

// Row entries e0 + g, e0 + g + G, ... < e1 (column indices col[], values vals[] with VW per entry,
// or RO: one real value per entry) against the two gathered signals X(c, x0, x1): full steps of
// four entries without branches (every step's eight gathers are in flight together), then the up
// to three entries left in one masked step.  cf[k]: the imaginary coefficients of the lane's
// signal k (per direction when PG).
template <int G, bool PG, bool RO, class Gather>
__device__ __forceinline__ void pair_row(const int *col, const double *vals, int e0, int e1, int g,
                                         const double (&cf)[2][3], Gather X, double2 (&y)[2]) {
    y[0] = make_double2(0.0, 0.0);
    y[1] = make_double2(0.0, 0.0);
    int e = e0 + g;
    for (; e + 3 * G < e1; e += 4 * G) {
        int c[4];
        double2 x0[4], x1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = col[e + k * G];
#pragma unroll
        for (int k = 0; k < 4; ++k) X(c[k], x0[k], x1[k]);
#pragma unroll
        for (int k = 0; k < 4; ++k) pair_entry<PG, RO>(vals, e + k * G, cf, x0[k], x1[k], y);
    }
    if (e < e1) {
        int c[3];
        double2 x0[3], x1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = e + k * G < e1 ? col[e + k * G] : -1;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (c[k] >= 0) X(c[k], x0[k], x1[k]);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (c[k] >= 0) pair_entry<PG, RO>(vals, e + k * G, cf, x0[k], x1[k], y);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        y[0].x += __shfl_xor(y[0].x, off, 64);
        y[0].y += __shfl_xor(y[0].y, off, 64);
        y[1].x += __shfl_xor(y[1].x, off, 64);
        y[1].y += __shfl_xor(y[1].y, off, 64);
    }
}

// Both signals of a lane at column c of a signal-interleaved vector: element c S + s0 and the
// next, through a 32-bit byte offset from the kernel argument's base (the load then takes the
// base from scalar registers and a single 32-bit offset from the lane; make_geom keeps n S 16 B of
// every vector of this mapping below 4 GiB).  Signal s0 + 1 past the last signal reads s0 again.
__device__ __forceinline__ void pair_load(const double2 *base, uint32_t row_bytes, uint32_t lane_bytes,
                                          uint32_t next_bytes, int c, double2 &x0, double2 &x1) {
    const char *p = reinterpret_cast<const char *>(base);
    const uint32_t o0 = (uint32_t)c * row_bytes + lane_bytes, o1 = o0 + next_bytes;
    x0 = *reinterpret_cast<const double2 *>(p + o0);
    x1 = *reinterpret_cast<const double2 *>(p + o1);
}

// The tiles of the block (XCD-aware walk of xcd_tiles) staged in LDS, then every row of a tile for
// the lane's two signals; emit(row, y) acts on lanes g == 0 and receives both signals' sums.  A tile
// whose entries exceed the LDS budget is read straight from global memory.
template <int G, bool PG, bool RO, class Gather, class Emit>
__device__ __forceinline__ void pair_spmv(const SolveArgs &a, int SP, int TR, const PairLane &L, bool act,
                                          const double (&cf)[2][3], Gather X, Emit emit) {
    constexpr int VW = RO ? 1 : (PG ? 4 : 2);
    const int rpb = (int)blockDim.x / (G * SP);
    const int ECAP = TR * kStageNnzPerRow;
    extern __shared__ double4 dmri_stage_lds[];
    double *svals = reinterpret_cast<double *>(dmri_stage_lds);
    int *scol = reinterpret_cast<int *>(svals + (size_t)ECAP * VW);
    int *srp = scol + ECAP;
    const double *gvals = RO ? a.areal : a.vals;
    const TileWalk W = xcd_tiles(a.n, TR);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int nr = (int)((W.end - t0) < TR ? (W.end - t0) : TR);
        const int E0 = a.rowptr[t0];
        const int ne = a.rowptr[t0 + nr] - E0;
        const bool fits = ne <= ECAP;
        __syncthreads();
        if (fits) stage_tile<RO, VW>(a, gvals, t0, nr, E0, ne, svals, scol, srp);
        __syncthreads();
        if (!act) continue;
        for (int lr = L.lrow; lr < nr; lr += rpb) {
            double2 y[2];
            if (fits)
                pair_row<G, PG, RO>(scol, svals, srp[lr], srp[lr + 1], L.g, cf, X, y);
            else
                pair_row<G, PG, RO>(a.col, gvals, a.rowptr[t0 + lr], a.rowptr[t0 + lr + 1], L.g, cf, X, y);
            if (L.g == 0) emit(t0 + lr, y);
        }
    }
}

// Per-signal block partials of a pair kernel (rk[k]: lane's sums for its signal 2 sp + k, zero on
// lanes g != 0): the rows of a wave sit LP lanes apart and are added with shuffles, then lane sp
// of every wave writes LDS and thread s < S adds the waves in order.
__device__ __forceinline__ void pair_partials(double (&r0)[kNV], double (&r1)[kNV], int LP, int sp, int g, int S,
                                              double *parts) {
#pragma unroll
    for (int k = 0; k < kNV; ++k)
        for (int off = LP; off < 64; off <<= 1) {
            r0[k] += __shfl_xor(r0[k], off, 64);
            r1[k] += __shfl_xor(r1[k], off, 64);
        }
    double(&buf)[kNV][kThreads] = reduce_scratch();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = (int)blockDim.x >> 6;
    if (lane < LP && g == 0) {
        const int s0 = 2 * sp;
#pragma unroll
        for (int k = 0; k < kNV; ++k) {
            if (s0 < S) buf[k][wave * S + s0] = r0[k];
            if (s0 + 1 < S) buf[k][wave * S + s0 + 1] = r1[k];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < S) {
        double acc[kNV] = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < nw; ++w) {
#pragma unroll
            for (int k = 0; k < kNV; ++k) acc[k] += buf[k][w * S + threadIdx.x];
        }
#pragma unroll
        for (int k = 0; k < kNV; ++k) parts[((size_t)blockIdx.x * S + threadIdx.x) * kNV + k] = acc[k];
    }
}

// The imaginary coefficients of the lane's two signals (zero past S)
template <bool PG>
__device__ __forceinline__ void pair_coefs(const double *coef, int s0, int S, double (&cf)[2][3]) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) cf[k][d] = s0 + k < S && (PG || d == 0) ? coef[3 * (s0 + k) + d] : 0.0;
}

// spmv_v_kernel's operations (kOpV, kOpPre, kOpPreCD) with two signals per lane
template <int G, bool PG, int OP>
__global__ void __launch_bounds__(kThreads) spmv_v_pair_kernel(SolveArgs a, int S, int SP, int TR, int it) {
    if (*a.all_done) return;
    const PairLane L = pair_lane<G>(SP);
    const int s0 = 2 * L.sp, s1 = s0 + 1;
    const bool on0 = s0 < S && !a.st[s0].done;
    const bool on1 = s1 < S && !a.st[s1].done;
    constexpr bool cd = OP == kOpPreCD;
    constexpr bool pre = OP == kOpPre || OP == kOpPreCD;
    double cf[2][3];
    pair_coefs<PG>(a.coefA, s0, S, cf);
    double r0[kNV] = {0.0, 0.0, 0.0, 0.0}, r1[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *p = a.p;
    const int64_t S_ = S;
    // signals of finished or absent lanes are gathered and multiplied too, but never written
    const uint32_t rb = 16u * (uint32_t)S, lb = 16u * (uint32_t)s0, nb = s1 < S ? 16u : 0u;
    const int s1c = s1 < S ? s1 : s0;
    auto gather = [&](int c, double2 &x0, double2 &x1) {
        pair_load(p, rb, lb, nb, c, x0, x1);
        if constexpr (cd) {
            x0 = c_mul(x0, cdiag_inv<PG>(a.pdiag, a.coefA, c, s0));
            x1 = c_mul(x1, cdiag_inv<PG>(a.pdiag, a.coefA, c, s1c));
        }
    };
    auto emit = [&](int64_t row, const double2 (&y)[2]) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (!(k == 0 ? on0 : on1)) continue;
            const int s = s0 + k;
            const int64_t idx = row * S_ + s;
            if (pre) {
                const double2 z = p[idx];
                const double2 w = make_double2(a.pc * (2.0 * z.x - a.pc * y[k].x), a.pc * (2.0 * z.y - a.pc * y[k].y));
                if constexpr (cd)
                    a.ph[idx] = c_mul(w, cdiag_inv<PG>(a.pdiag, a.coefA, row, s));
                else
                    a.ph[idx] = w;
                continue;
            }
            a.v[idx] = y[k];
            const double2 h = a.rh[idx];
            double(&r)[kNV] = k == 0 ? r0 : r1;
            r[0] += h.x * y[k].x + h.y * y[k].y;
            r[1] += h.x * y[k].y - h.y * y[k].x;
        }
    };
    const bool act = on0 || on1;
    if (OP == kOpPre && a.areal)
        pair_spmv<G, PG, true>(a, SP, TR, L, act, cf, gather, emit);
    else
        pair_spmv<G, PG, false>(a, SP, TR, L, act, cf, gather, emit);
    if (OP == kOpV) pair_partials(r0, r1, G * SP, L.sp, L.g, S, a.partV);
}

// spmv_t_kernel's operations (kOpT, kOpTP, kOpPre, kOpPreCD) with two signals per lane
template <int G, bool PG, int OP>
__global__ void __launch_bounds__(kThreads) spmv_t_pair_kernel(SolveArgs a, int S, int SP, int TR, int nparts,
                                                               int it) {
    if (*a.all_done) return;
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    spmv_t_head(a, S, nparts, it, salpha, sact);
    const PairLane L = pair_lane<G>(SP);
    const int s0 = 2 * L.sp, s1 = s0 + 1;
    const bool on0 = s0 < S && sact[s0];
    const bool on1 = s1 < S && sact[s1];
    const double2 al0 = on0 ? salpha[s0] : make_double2(0.0, 0.0);
    const double2 al1 = on1 ? salpha[s1] : make_double2(0.0, 0.0);
    constexpr bool cd = OP == kOpPreCD;
    constexpr bool pre = OP == kOpPre || OP == kOpPreCD;
    double cf[2][3];
    pair_coefs<PG>(a.coefA, s0, S, cf);
    double r0[kNV] = {0.0, 0.0, 0.0, 0.0}, r1[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *r = a.r, *v = a.v, *sh = a.sh, *sm = a.s;
    const int64_t S_ = S;
    // s at entry j of signal k: materialised by the s-pass, or formed from r and v
    auto sval = [&](int64_t j, const double2 &al) { return sm ? sm[j] : c_sub(r[j], c_mul(al, v[j])); };
    // signals of finished or absent lanes are gathered and multiplied too, but never written
    const uint32_t rb = 16u * (uint32_t)S, lb = 16u * (uint32_t)s0, nb = s1 < S ? 16u : 0u;
    const int s1c = s1 < S ? s1 : s0;
    auto gather = [&](int c, double2 &x0, double2 &x1) {
        if (OP == kOpTP || sm) {
            pair_load(OP == kOpTP ? sh : sm, rb, lb, nb, c, x0, x1);
        } else {
            const int64_t j = (int64_t)c * S_ + s0;
            x0 = c_sub(r[j], c_mul(al0, v[j]));
            x1 = c_sub(r[j + nb / 16], c_mul(al1, v[j + nb / 16]));
        }
        if constexpr (cd) {
            x0 = c_mul(x0, cdiag_inv<PG>(a.pdiag, a.coefA, c, s0));
            x1 = c_mul(x1, cdiag_inv<PG>(a.pdiag, a.coefA, c, s1c));
        }
    };
    auto emit = [&](int64_t row, const double2 (&y)[2]) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (!(k == 0 ? on0 : on1)) continue;
            const int s = s0 + k;
            const int64_t idx = row * S_ + s;
            const double2 sv = sval(idx, k == 0 ? al0 : al1);
            if (pre) {
                const double2 w =
                    make_double2(a.pc * (2.0 * sv.x - a.pc * y[k].x), a.pc * (2.0 * sv.y - a.pc * y[k].y));
                if constexpr (cd)
                    a.sh[idx] = c_mul(w, cdiag_inv<PG>(a.pdiag, a.coefA, row, s));
                else
                    a.sh[idx] = w;
                continue;
            }
            a.t[idx] = y[k];
            double(&rr)[kNV] = k == 0 ? r0 : r1;
            rr[0] += y[k].x * sv.x + y[k].y * sv.y;
            rr[1] += y[k].x * sv.y - y[k].y * sv.x;
            rr[2] += y[k].x * y[k].x + y[k].y * y[k].y;
            rr[3] += sv.x * sv.x + sv.y * sv.y;
        }
    };
    const bool act = on0 || on1;
    if (OP == kOpPre && a.areal)
        pair_spmv<G, PG, true>(a, SP, TR, L, act, cf, gather, emit);
    else
        pair_spmv<G, PG, false>(a, SP, TR, L, act, cf, gather, emit);
    if (!pre) pair_partials(r0, r1, G * SP, L.sp, L.g, S, a.partT);
}

// ---- tile-compressed rows: a tile is the TR rows of one pass of the (row, signal) lane groups.
// Per tile, the values at the tile's distinct columns (sorted, listed at build time) are staged
// in LDS once for every active signal: neighbouring rows share most columns, so each x value
// (or s = r - alpha v) is read from memory once per tile instead of once per entry, and the
// entries index the staged values through 16-bit local column numbers.
template <int G, bool PG, class Gather>
__device__ __forceinline__ void tile_row_accumulate(const int *rowptr, const unsigned short *lcol,
                                                    const double *vals, int64_t row, int g, Gather X,
                                                    Acc<PG> &acc) {
    acc.a = make_double2(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < (PG ? 3 : 1); ++d) acc.j[d] = make_double2(0.0, 0.0);
    const int e1 = rowptr[row + 1];
    int e = rowptr[row] + g;
    for (; e + 3 * G < e1; e += 4 * G) {
        const int l0 = lcol[e], l1 = lcol[e + G], l2 = lcol[e + 2 * G], l3 = lcol[e + 3 * G];
        const double2 x0 = X(l0), x1 = X(l1), x2 = X(l2), x3 = X(l3);
        acc_entry<PG>(vals, e, x0, acc);
        acc_entry<PG>(vals, e + G, x1, acc);
        acc_entry<PG>(vals, e + 2 * G, x2, acc);
        acc_entry<PG>(vals, e + 3 * G, x3, acc);
    }
    for (; e < e1; e += G) acc_entry<PG>(vals, e, X(lcol[e]), acc);
}

// X(c, s): value of signal s at column c.  ucap: distinct columns the dynamic LDS holds.
template <int G, bool PG, class Stage, class Emit>
__device__ __forceinline__ void tile_spmv(const SolveArgs &a, int S, int TR, int ucap, const Lane &L, bool act,
                                          const int *sact, Stage X, Emit emit) {
    extern __shared__ double2 tile_x[];
    const TileWalk W = xcd_tiles(a.n, TR);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int64_t tile = t0 / TR;
        const int nu = a.tinfo[tile];
        const bool comp = nu >= 0 && nu <= ucap;
        __syncthreads();  // the previous tile's readers are done
        if (comp) {
            const int u0 = a.tptr[tile];
            for (int i = threadIdx.x; i < nu * S; i += blockDim.x) {
                const int u = i / S, s = i - u * S;
                tile_x[i] = sact[s] ? X(a.ucol[u0 + u], s) : make_double2(0.0, 0.0);
            }
        }
        __syncthreads();
        const int64_t row = t0 + L.lrow;
        if (!act || row >= W.end) continue;
        Acc<PG> acc;
        if (comp)
            tile_row_accumulate<G, PG>(a.rowptr, a.lcol, a.vals, row, L.g,
                                       [&](int lc) { return tile_x[lc * S + L.s]; }, acc);
        else
            row_accumulate<G, PG>(a.rowptr, a.col, a.vals, row, L.g, [&](int c) { return X(c, L.s); }, acc);
        group_reduce_acc<G, PG>(acc);
        emit(row, acc);
    }
}

template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) spmv_v_tile_kernel(SolveArgs a, int S, int TR, int ucap, int it) {
    if (*a.all_done) return;
    __shared__ int sact[kMaxSignals];
    if ((int)threadIdx.x < S) sact[threadIdx.x] = !a.st[threadIdx.x].done;
    __syncthreads();
    const Lane L = spmv_lane<G>(S);
    const bool act = sact[L.s];
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *p = a.p;
    tile_spmv<G, PG>(
        a, S, TR, ucap, L, act, sact, [&](int c, int s) { return p[(int64_t)c * S + s]; },
        [&](int64_t row, const Acc<PG> &acc) {
            if (L.g != 0) return;
            const int64_t idx = row * S + L.s;
            const double2 y = imag_coef_apply<G, PG>(acc, a.coefA, L.s);
            a.v[idx] = y;
            const double2 h = a.rh[idx];
            red[0] += h.x * y.x + h.y * y.y;
            red[1] += h.x * y.y - h.y * y.x;
        });
    block_signal_partials(red, L.g == 0, L.lrow * S + L.s, TR * S, S, a.partV);
}

template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) spmv_t_tile_kernel(SolveArgs a, int S, int TR, int ucap, int nparts,
                                                               int it) {
    if (*a.all_done) return;
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    spmv_t_head(a, S, nparts, it, salpha, sact);
    const Lane L = spmv_lane<G>(S);
    const bool act = sact[L.s];
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *r = a.r, *v = a.v;
    tile_spmv<G, PG>(
        a, S, TR, ucap, L, act, sact,
        [&](int c, int s) {
            const int64_t j = (int64_t)c * S + s;
            return c_sub(r[j], c_mul(salpha[s], v[j]));
        },
        [&](int64_t row, const Acc<PG> &acc) {
            if (L.g != 0) return;
            const int64_t idx = row * S + L.s;
            const double2 y = imag_coef_apply<G, PG>(acc, a.coefA, L.s);
            a.t[idx] = y;
            const double2 sv = c_sub(r[idx], c_mul(salpha[L.s], v[idx]));
            red[0] += y.x * sv.x + y.y * sv.y;
            red[1] += y.x * sv.y - y.y * sv.x;
            red[2] += y.x * y.x + y.y * y.y;
            red[3] += sv.x * sv.x + sv.y * sv.y;
        });
    block_signal_partials(red, L.g == 0, L.lrow * S + L.s, TR * S, S, a.partT);
}

// ---- one wave per row, lane = signal (any S <= 64).  The row's column indices and values are
// the same for every lane: with the row index made wave-uniform they load through the scalar
// cache once per wave (s_load), the vector registers hold only the gathered x and the
// accumulators, and each lane gathers its own signal's 16 B of x (64 signals: one 1 KB run).
template <bool PG, class Gather, class Emit, class AccT = Acc<PG>>
__device__ __forceinline__ void wave_rows(int64_t n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                          const double *__restrict__ vals, bool act, Gather X, Emit emit,
                                          AccT proto = AccT{}) {
    const int w = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    const TileWalk W = xcd_tiles(n, nw);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int64_t row = t0 + w;
        if (row >= W.end) continue;
        const int r = __builtin_amdgcn_readfirstlane((int)row);
        const int e0 = rowptr[r], e1 = rowptr[r + 1];
        AccT acc = proto;
        acc_zero(acc);
        int e = e0;
        for (; e + 4 <= e1; e += 4) {
            int c[4];
            double2 xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = col[e + u];
#pragma unroll
            for (int u = 0; u < 4; ++u) xv[u] = act ? X(c[u]) : make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < 4; ++u) acc_entry<PG>(vals, e + u, xv[u], acc);
        }
        for (; e < e1; ++e) {
            const double2 xv = act ? X(col[e]) : make_double2(0.0, 0.0);
            acc_entry<PG>(vals, e, xv, acc);
        }
        emit(row, acc);
    }
}

template <bool PG, int OP>
__global__ void __launch_bounds__(kThreads) spmv_v_wave_kernel(SolveArgs a, const int *__restrict__ rowptr,
                                                               const int *__restrict__ col,
                                                               const double *__restrict__ vals, int S, int it) {
    if (*a.all_done) return;
    const int lane = (int)(threadIdx.x & 63);
    const bool act = lane < S && !a.st[lane].done;
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *p = a.p;
    auto emit = [&](int64_t row, const auto &acc) {
        if (!act) return;
        const int64_t idx = row * S + lane;
        const double2 y = imag_coef_apply<1, PG>(acc, a.coefA, lane);
        if (OP == kOpPre) {
            const double2 z = p[idx];
            a.ph[idx] = make_double2(a.pc * (2.0 * z.x - a.pc * y.x), a.pc * (2.0 * z.y - a.pc * y.y));
            return;
        }
        a.v[idx] = y;
        const double2 h = a.rh[idx];
        red[0] += h.x * y.x + h.y * y.y;
        red[1] += h.x * y.y - h.y * y.x;
    };
    auto gather = [&](int c) { return p[(int64_t)c * S + lane]; };
    if constexpr (PG)  // lanes beyond the signals read no coefficients
        wave_rows<PG>(a.n, rowptr, col, vals, act, gather, emit, lane < S ? fold_acc(a.coefA, lane) : FoldAcc{});
    else
        wave_rows<PG>(a.n, rowptr, col, vals, act, gather, emit);
    if (OP == kOpV)
        block_signal_partials(red, lane < S, (int)(threadIdx.x >> 6) * S + lane, (int)(blockDim.x >> 6) * S, S,
                              a.partV);
}

template <bool PG, int OP>
__global__ void __launch_bounds__(kThreads) spmv_t_wave_kernel(SolveArgs a, const int *__restrict__ rowptr,
                                                               const int *__restrict__ col,
                                                               const double *__restrict__ vals, int S, int nparts,
                                                               int it) {
    if (*a.all_done) return;
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    spmv_t_head(a, S, nparts, it, salpha, sact);
    const int lane = (int)(threadIdx.x & 63);
    const bool act = lane < S && sact[lane];
    const double2 alpha = act ? salpha[lane] : make_double2(0.0, 0.0);
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    const double2 *r = a.r, *v = a.v, *sh = a.sh, *sm = a.s;
    auto gather = [&](int c) {
        const int64_t j = (int64_t)c * S + lane;
        if (OP == kOpTP) return sh[j];
        return sm ? sm[j] : c_sub(r[j], c_mul(alpha, v[j]));
    };
    auto emit = [&](int64_t row, const auto &acc) {
        if (!act) return;
        const int64_t idx = row * S + lane;
        const double2 y = imag_coef_apply<1, PG>(acc, a.coefA, lane);
        const double2 sv = sm ? sm[idx] : c_sub(r[idx], c_mul(alpha, v[idx]));
        if (OP == kOpPre) {
            a.sh[idx] = make_double2(a.pc * (2.0 * sv.x - a.pc * y.x), a.pc * (2.0 * sv.y - a.pc * y.y));
            return;
        }
        a.t[idx] = y;
        red[0] += y.x * sv.x + y.y * sv.y;
        red[1] += y.x * sv.y - y.y * sv.x;
        red[2] += y.x * y.x + y.y * y.y;
        red[3] += sv.x * sv.x + sv.y * sv.y;
    };
    if constexpr (PG)
        wave_rows<PG>(a.n, rowptr, col, vals, act, gather, emit, lane < S ? fold_acc(a.coefA, lane) : FoldAcc{});
    else
        wave_rows<PG>(a.n, rowptr, col, vals, act, gather, emit);
    if (OP != kOpPre)
        block_signal_partials(red, lane < S, (int)(threadIdx.x >> 6) * S + lane, (int)(blockDim.x >> 6) * S, S,
                              a.partT);
}

// --------------------------------------------------------------------------------------
// Restarted GMRES for S signals in lockstep (right-preconditioned like BiCGStab): Arnoldi with
// classical Gram-Schmidt applied twice, complex Givens rotations per signal on the device, and the
// update x += M (V y) at the end of a cycle.  The per-signal KrylovState marks a signal done for the
// rest of the cycle (converged, the cycle full, or a lucky breakdown), which the SpMV kernels honour.
// --------------------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_signal_partials_n(const double (&v)[NV], bool writer, int slot, int nslots,
                                                        int S, double *parts) {
    __shared__ double buf[NV][kThreads];
    if (writer) {
#pragma unroll
        for (int k = 0; k < NV; ++k) buf[k][slot] = v[k];
    }
    __syncthreads();
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        double acc[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] = 0.0;
        for (int j = s; j < nslots; j += S) {
#pragma unroll
            for (int k = 0; k < NV; ++k) acc[k] += buf[k][j];
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) parts[((size_t)blockIdx.x * S + s) * NV + k] = acc[k];
    }
}

// conj(v_i) . w for basis vectors i0 .. i0+cnt-1 (cnt <= kGmChunk): block partials [grid][S][2 kGmChunk]
__global__ void __launch_bounds__(kThreads) gm_dots_kernel(GmArgs a, int i0, int cnt, int vrpb) {
    if (*a.all_done) return;
    constexpr int NV = 2 * kGmChunk;
    const int S = a.S;
    const int64_t total = a.n * S, nS = a.n * S;
    const int s = threadIdx.x % S;
    const bool valid = (int)threadIdx.x < vrpb * S;
    const bool act = valid && !a.st[s].done;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    if (act)
        for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
            const double2 wv = a.w[i];
#pragma unroll
            for (int k = 0; k < kGmChunk; ++k)
                if (k < cnt) {
                    const double2 v = a.V[(int64_t)(i0 + k) * nS + i];
                    acc[2 * k] += v.x * wv.x + v.y * wv.y;
                    acc[2 * k + 1] += v.x * wv.y - v.y * wv.x;
                }
        }
    block_signal_partials_n<NV>(acc, valid, threadIdx.x, vrpb * S, S, a.part);
}

// One block per signal: out[s][o + k] = sum over blocks of parts[b][s][k] (pstride doubles per block and
// signal), k < nv, in a fixed order
__global__ void __launch_bounds__(kThreads) gm_fin_kernel(const double *parts, int nparts, int S, int pstride, int nv,
                                                          double *out, int ostride, int o) {
    const int s = blockIdx.x;
    for (int k = 0; k < nv; ++k) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) v += parts[((size_t)b * S + s) * pstride + k];
        double r[1] = {v};
        block_reduce<1>(r);
        if (threadIdx.x == 0) out[(size_t)s * ostride + o + k] = r[0];
    }
}

// w -= sum_{i <= j} h_i v_i (h from a.h); norm (the second pass): partials [|w|^2, 0, 0, 0] (kNV
// layout).  Column j of the Hessenberg matrix is the sum of both passes' coefficients: the first pass
// stores its h in a.hsum, the second adds to it.
__global__ void __launch_bounds__(kThreads) gm_orth_kernel(GmArgs a, int j, int norm, int vrpb) {
    if (*a.all_done) return;
    const int S = a.S;
    const int64_t total = a.n * S, nS = a.n * S;
    const int s = threadIdx.x % S;
    const bool valid = (int)threadIdx.x < vrpb * S;
    const bool act = valid && !a.st[s].done;
    const int hst = 2 * (a.m + 1);
    if (blockIdx.x == 0 && (int)threadIdx.x < S && !a.st[threadIdx.x].done)
        for (int i = 0; i <= j; ++i) {
            double *hs = a.hsum + (size_t)threadIdx.x * hst + 2 * i;
            const double *h = a.h + (size_t)threadIdx.x * hst + 2 * i;
            hs[0] = norm ? hs[0] + h[0] : h[0];
            hs[1] = norm ? hs[1] + h[1] : h[1];
        }
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    if (act)
        for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
            double2 wv = a.w[i];
            for (int q = 0; q <= j; ++q) {
                const double hr = a.h[(size_t)s * hst + 2 * q], hi = a.h[(size_t)s * hst + 2 * q + 1];
                const double2 v = a.V[(int64_t)q * nS + i];
                wv.x -= hr * v.x - hi * v.y;
                wv.y -= hr * v.y + hi * v.x;
            }
            a.w[i] = wv;
            red[0] += wv.x * wv.x + wv.y * wv.y;
        }
    if (norm) block_signal_partials(red, valid, threadIdx.x, vrpb * S, S, a.part);
}

// One block, thread = signal: column j of the Hessenberg matrix from hsum and |w|, the earlier Givens
// rotations applied to it, a new rotation zeroing its subdiagonal, the residual estimate |g_{j+1}|;
// a signal ends its cycle when that estimate meets its tolerance, at a lucky breakdown (|w| = 0) or
// when the cycle is full.  wn2 = [S][kNV] totals of the norm pass.
__global__ void gm_hess_kernel(GmArgs a, int j, const double *wn2) {
    if (*a.all_done) return;
    const int S = a.S, m = a.m, hst = 2 * (m + 1);
    const int s = threadIdx.x;
    __shared__ int sact[kMaxSignals];
    if (s < S) {
        KrylovState &st = a.st[s];
        int act = !st.done;
        if (act) {
            double2 *H = a.H + (size_t)s * m * (m + 1) + (size_t)j * (m + 1);
            for (int i = 0; i <= j; ++i)
                H[i] = make_double2(a.hsum[(size_t)s * hst + 2 * i], a.hsum[(size_t)s * hst + 2 * i + 1]);
            const double hn = sqrt(wn2[(size_t)s * kNV]);
            double2 *cs = a.cs + (size_t)s * m;
            double *cc = a.cc + (size_t)s * m;
            for (int i = 0; i < j; ++i) {
                const double2 x0 = H[i], x1 = H[i + 1];
                H[i] = c_add(make_double2(cc[i] * x0.x, cc[i] * x0.y), c_mul(cs[i], x1));
                H[i + 1] = c_sub(make_double2(cc[i] * x1.x, cc[i] * x1.y), c_mul(make_double2(cs[i].x, -cs[i].y), x0));
            }
            const double2 hj = H[j];
            const double ah = sqrt(hj.x * hj.x + hj.y * hj.y);
            const double t = sqrt(ah * ah + hn * hn);
            double c;
            double2 sn, r;
            if (t == 0.0) {
                c = 1.0;
                sn = make_double2(0.0, 0.0);
                r = make_double2(0.0, 0.0);
            } else if (ah == 0.0) {
                c = 0.0;
                sn = make_double2(1.0, 0.0);
                r = make_double2(hn, 0.0);
            } else {
                const double2 ph = make_double2(hj.x / ah, hj.y / ah);
                c = ah / t;
                sn = make_double2(ph.x * hn / t, ph.y * hn / t);
                r = make_double2(ph.x * t, ph.y * t);
            }
            H[j] = r;
            cc[j] = c;
            cs[j] = sn;
            double2 *g = a.g + (size_t)s * (m + 1);
            const double2 gj = g[j];
            g[j] = make_double2(c * gj.x, c * gj.y);
            g[j + 1] = c_mul(make_double2(-sn.x, sn.y), gj);
            const double res2 = g[j + 1].x * g[j + 1].x + g[j + 1].y * g[j + 1].y;
            st.rnorm2 = res2;
            st.iters += 1;
            a.nk[s] = j + 1;
            a.inv[s] = hn > 0.0 ? 1.0 / hn : 0.0;
            if (res2 <= st.tol2 || hn == 0.0 || j + 1 >= m || st.iters >= a.maxit) {
                st.done = 1;
                act = 0;
            }
        }
        sact[s] = act;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 1;
        for (int q = 0; q < S; ++q)
            if (sact[q]) all = 0;
        if (all) *a.all_done = 1;
    }
}

// v_{j+1} = w / |w| for the signals still in the cycle
__global__ void __launch_bounds__(kThreads) gm_scale_kernel(GmArgs a, int j, int vrpb) {
    if (*a.all_done) return;
    const int S = a.S;
    const int64_t total = a.n * S, nS = a.n * S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S || a.st[s].done) return;
    const double iv = a.inv[s];
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
        const double2 wv = a.w[i];
        a.V[(int64_t)(j + 1) * nS + i] = make_double2(wv.x * iv, wv.y * iv);
    }
}

// Start of a cycle from the residual totals tot[s] = [|r|^2, 0, |r|^2, |b|^2] (first: also the
// tolerance): g = (|r|, 0, ...), hsum = 0, signals within tolerance converge; the others restart
// their Arnoldi process (st.done = 0).  One block, thread = signal.
__global__ void gm_start_kernel(GmArgs a, const double *tot, int first) {
    const int S = a.S, m = a.m;
    const int s = threadIdx.x;
    __shared__ int sact[kMaxSignals];
    if (s < S) {
        KrylovState &st = a.st[s];
        const double rn2 = tot[(size_t)s * kNV], bn2 = tot[(size_t)s * kNV + 3];
        if (first) {
            const double t1 = a.rtol * a.rtol * bn2, t2 = a.atol * a.atol;
            st.tol2 = t1 > t2 ? t1 : t2;
            st.bnorm2 = bn2;
            st.iters = 0;
            st.conv = 0;
            st.breakdown = 0;
        }
        st.rnorm2 = rn2;
        const bool conv = rn2 <= st.tol2;
        const bool lim = !conv && st.iters >= a.maxit;
        st.conv = conv ? 1 : 0;
        st.done = (conv || lim) ? 1 : 0;
        a.nk[s] = 0;
        const double beta = sqrt(rn2);
        a.inv[s] = beta > 0.0 ? 1.0 / beta : 0.0;
        double2 *g = a.g + (size_t)s * (m + 1);
        g[0] = make_double2(beta, 0.0);
        for (int i = 1; i <= m; ++i) g[i] = make_double2(0.0, 0.0);
        for (int i = 0; i < 2 * (m + 1); ++i) a.hsum[(size_t)s * 2 * (m + 1) + i] = 0.0;
        sact[s] = !st.done;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int any = 0;
        for (int q = 0; q < S; ++q) any |= sact[q];
        *a.all_done = any ? 0 : 1;
    }
}

// v_0 = r / |r| for the signals starting a cycle
__global__ void __launch_bounds__(kThreads) gm_v0_kernel(GmArgs a, const double2 *r, int vrpb) {
    const int S = a.S;
    const int64_t total = a.n * S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S || a.st[s].done) return;
    const double iv = a.inv[s];
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
        const double2 rv = r[i];
        a.V[i] = make_double2(rv.x * iv, rv.y * iv);
    }
}

// End of a cycle: y from R y = g by back substitution over the nk steps of each signal (one block,
// thread = signal) ...
__global__ void gm_solve_kernel(GmArgs a) {
    const int S = a.S, m = a.m;
    const int s = threadIdx.x;
    if (s >= S) return;
    const int k = a.nk[s];
    const double2 *H = a.H + (size_t)s * m * (m + 1);
    const double2 *g = a.g + (size_t)s * (m + 1);
    double2 *y = a.y + (size_t)s * m;
    for (int i = k - 1; i >= 0; --i) {
        double2 acc = g[i];
        for (int q = i + 1; q < k; ++q) acc = c_sub(acc, c_mul(H[(size_t)q * (m + 1) + i], y[q]));
        y[i] = c_div(acc, H[(size_t)i * (m + 1) + i]);
    }
}

// ... then out = sum_i y_i v_i (0 for a signal that took no step)
__global__ void __launch_bounds__(kThreads) gm_combine_kernel(GmArgs a, double2 *out, int vrpb) {
    const int S = a.S, m = a.m;
    const int64_t total = a.n * S, nS = a.n * S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S) return;
    const int k = a.nk[s];
    const double2 *y = a.y + (size_t)s * m;
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
        double2 acc = make_double2(0.0, 0.0);
        for (int q = 0; q < k; ++q) acc = c_add(acc, c_mul(y[q], a.V[(int64_t)q * nS + i]));
        out[i] = acc;
    }
}

// x += z
__global__ void __launch_bounds__(kThreads) gm_axpy_kernel(int64_t total, const double2 *z, double2 *x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = c_add(x[i], z[i]);
}

// every signal active again and the all-done flag cleared (the other state is kept)
__global__ void reopen_kernel(KrylovState *st, int S, int *flag) {
    const int s = threadIdx.x;
    if (s < S) st[s].done = 0;
    if (s == 0) *flag = 0;
}
// ---- signals in registers: each lane takes entries g, g+G, ... of a row for all S <= 8
// signals, so column indices and values are loaded once per entry instead of once per signal.
constexpr int kMaxSR = 8;

template <int NVS>
__device__ __forceinline__ void block_partials_sr(double (&red)[kMaxSR][NVS], int S, double *parts) {
    for (int s = 0; s < S; ++s) {
        double v[NVS];
#pragma unroll
        for (int k = 0; k < NVS; ++k) v[k] = red[s][k];
        block_reduce<NVS>(v);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < kNV; ++k) parts[((size_t)blockIdx.x * S + s) * kNV + k] = k < NVS ? v[k] : 0.0;
        }
    }
}

template <int G, bool PG, class Gather>
__device__ __forceinline__ void sr_row(const SolveArgs &a, int64_t row, int g, int S, Gather X,
                                       Acc<PG> (&acc)[kMaxSR]) {
#pragma unroll
    for (int s = 0; s < kMaxSR; ++s) {
        acc[s].a = make_double2(0.0, 0.0);
#pragma unroll
        for (int d = 0; d < (PG ? 3 : 1); ++d) acc[s].j[d] = make_double2(0.0, 0.0);
    }
    const int e1 = a.rowptr[row + 1];
    for (int e = a.rowptr[row] + g; e < e1; e += G) {
        const int c = a.col[e];
#pragma unroll
        for (int s = 0; s < kMaxSR; ++s)
            if (s < S) acc_entry<PG>(a.vals, e, X(c, s), acc[s]);
    }
#pragma unroll
    for (int s = 0; s < kMaxSR; ++s)
        if (s < S) group_reduce_acc<G, PG>(acc[s]);
}

template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) spmv_v_sr_kernel(SolveArgs a, int S, int rpb, int it) {
    if (*a.all_done) return;
    const int lrow = threadIdx.x / G, g = threadIdx.x % G;
    bool act[kMaxSR];
    double red[kMaxSR][2];
#pragma unroll
    for (int s = 0; s < kMaxSR; ++s) {
        act[s] = s < S && !a.st[s].done;
        red[s][0] = red[s][1] = 0.0;
    }
    const double2 *p = a.p;
    const TileWalk W = xcd_tiles(a.n, rpb);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int64_t row = t0 + lrow;
        if (row >= W.end) continue;
        Acc<PG> acc[kMaxSR];
        sr_row<G, PG>(a, row, g, S, [&](int c, int s) { return p[(int64_t)c * S + s]; }, acc);
        if (g != 0) continue;
#pragma unroll
        for (int s = 0; s < kMaxSR; ++s) {
            if (!act[s]) continue;
            const int64_t idx = row * S + s;
            const double2 y = imag_coef_apply<G, PG>(acc[s], a.coefA, s);
            a.v[idx] = y;
            const double2 h = a.rh[idx];
            red[s][0] += h.x * y.x + h.y * y.y;
            red[s][1] += h.x * y.y - h.y * y.x;
        }
    }
    block_partials_sr<2>(red, S, a.partV);
}

template <int G, bool PG>
__global__ void __launch_bounds__(kThreads) spmv_t_sr_kernel(SolveArgs a, int S, int rpb, int nparts, int it) {
    if (*a.all_done) return;
    __shared__ double tot[kMaxSignals][kNV];
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    signal_totals(a.partV, a.totV, nparts, S, tot);
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        KrylovState &st = a.st[s];
        int actv = !st.done;
        const int local = it - st.it0;
        if (actv) {
            const double2 rv = make_double2(tot[s][0], tot[s][1]);
            if (rv.x == 0.0 && rv.y == 0.0) {
                if (blockIdx.x == 0) {
                    st.done = 1;
                    st.breakdown = 1;
                    st.conv = 0;
                    st.iters = it;
                }
                actv = 0;
            } else {
                const double2 al = c_div(st.rho[local & 1], rv);
                salpha[s] = al;
                if (blockIdx.x == 0) st.alpha[local & 1] = al;
            }
        }
        sact[s] = actv;
    }
    __syncthreads();
    const int lrow = threadIdx.x / G, g = threadIdx.x % G;
    bool act[kMaxSR];
    double2 alpha[kMaxSR];
    double red[kMaxSR][4];
#pragma unroll
    for (int s = 0; s < kMaxSR; ++s) {
        act[s] = s < S && sact[s];
        alpha[s] = s < S ? salpha[s] : make_double2(0.0, 0.0);
        red[s][0] = red[s][1] = red[s][2] = red[s][3] = 0.0;
    }
    const double2 *r = a.r, *v = a.v;
    const TileWalk W = xcd_tiles(a.n, rpb);
    for (int64_t t0 = W.first; t0 < W.end; t0 += W.stride) {
        const int64_t row = t0 + lrow;
        if (row >= W.end) continue;
        Acc<PG> acc[kMaxSR];
        sr_row<G, PG>(
            a, row, g, S,
            [&](int c, int s) {
                const int64_t j = (int64_t)c * S + s;
                return c_sub(r[j], c_mul(alpha[s], v[j]));
            },
            acc);
        if (g != 0) continue;
#pragma unroll
        for (int s = 0; s < kMaxSR; ++s) {
            if (!act[s]) continue;
            const int64_t idx = row * S + s;
            const double2 y = imag_coef_apply<G, PG>(acc[s], a.coefA, s);
            a.t[idx] = y;
            const double2 sv = c_sub(r[idx], c_mul(alpha[s], v[idx]));
            red[s][0] += y.x * sv.x + y.y * sv.y;
            red[s][1] += y.x * sv.y - y.y * sv.x;
            red[s][2] += y.x * y.x + y.y * y.y;
            red[s][3] += sv.x * sv.x + sv.y * sv.y;
        }
    }
    block_partials_sr<4>(red, S, a.partT);
}

// omega = <t, s> / <t, t> (0 when |s| already meets the tolerance);
// x += alpha p + omega s ; r = s - omega t ; partials U: [<r^, r>.re, .im, |r|^2, 0]
// s-pass: alpha = rho / <r^, v> (the head every t-kernel runs too) and s = r - alpha v for the active
// signals, so the preconditioner's t-kernel gathers one vector instead of two
__global__ void __launch_bounds__(kThreads) spass_kernel(SolveArgs a, int S, int vrpb, int nparts, int it) {
    if (*a.all_done) return;
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    spmv_t_head(a, S, nparts, it, salpha, sact);
    const int64_t total = a.n * S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S || !sact[s]) return;
    const double2 alpha = salpha[s];
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S)
        st_once(a.s + i, c_sub(ld_once(a.r + i), c_mul(alpha, ld_once(a.v + i))));
}

__global__ void __launch_bounds__(kThreads) bicg_update_kernel(SolveArgs a, int S, int vrpb, int nparts, int it) {
    if (*a.all_done) return;
    __shared__ double tot[kMaxSignals][kNV];
    __shared__ double2 salpha[kMaxSignals], somega[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    signal_totals(a.partT, a.totT, nparts, S, tot);
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        KrylovState &st = a.st[s];
        const int act = !st.done;
        const int local = it - st.it0;
        if (act) {
            const double2 ts = make_double2(tot[s][0], tot[s][1]);
            const double tt = tot[s][2], s2 = tot[s][3];
            double2 om = make_double2(0.0, 0.0);
            if (!(s2 <= st.tol2) && tt != 0.0) om = make_double2(ts.x / tt, ts.y / tt);
            somega[s] = om;
            salpha[s] = st.alpha[local & 1];
            if (blockIdx.x == 0) st.omega[local & 1] = om;
        }
        sact[s] = act;
    }
    __syncthreads();
    const int64_t total = a.n * S;
    const int s = threadIdx.x % S;
    const bool valid = (int)threadIdx.x < vrpb * S;
    const bool act = valid && sact[s];
    const double2 alpha = salpha[s], omega = somega[s];
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    if (act) {
        for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total;
             i += (int64_t)gridDim.x * vrpb * S) {
            const double2 sv = a.s ? ld_once(a.s + i) : c_sub(ld_once(a.r + i), c_mul(alpha, ld_once(a.v + i)));
            // right preconditioning: the iterate moves along the preconditioned directions
            const double2 xi = ld_once(a.x + i);
            if (a.ph)
                st_once(a.x + i, c_add(xi, c_add(c_mul(alpha, ld_once(a.ph + i)), c_mul(omega, ld_once(a.sh + i)))));
            else
                st_once(a.x + i, c_add(xi, c_add(c_mul(alpha, ld_once(a.p + i)), c_mul(omega, sv))));
            const double2 rn = c_sub(sv, c_mul(omega, ld_once(a.t + i)));
            st_once(a.r + i, rn);
            const double2 h = ld_once(a.rh + i);
            red[0] += h.x * rn.x + h.y * rn.y;
            red[1] += h.x * rn.y - h.y * rn.x;
            red[2] += rn.x * rn.x + rn.y * rn.y;
        }
    }
    block_signal_partials(red, valid, threadIdx.x, vrpb * S, S, a.partU);
}

// ---- COCG (see launch_cocg_spmv in dmri_internal.hpp).  Per-signal scalars in KrylovState: rho
// (r^T z) and alpha in the slot of the iteration's parity, as for BiCGStab.  The iterate's update
// x += alpha p of an iteration is applied by the next iteration's p-kernel, which reads that p
// anyway before it overwrites it, so the update kernel streams only r and q.
// Head: the convergence test on |r|^2 (partials U, from the residual kernel at a solve's first
// iteration, else from the update), rho = r^T z (partials V, from the preconditioner SpMV of this
// iteration), beta = rho / rho_prev; body: x += alpha_prev p (every iteration but a solve's first,
// also for a signal that stops here), then p = z (first iteration) or z + beta p.
// Head of COCG's p-step (cocg_p_kernel, or the q-kernel that folds it in): the convergence test on
// |r|^2, beta = rho / rho_prev, and which signals still owe the previous iteration's x += alpha p;
// per-signal results in the shared arrays, followed by a barrier.
__device__ __forceinline__ void cocg_head(const SolveArgs &a, int S, int nparts_rhs, int nparts_upd, int nparts_spmv,
                                          int it, int rfold, double2 *sbeta, double2 *salpha, int *sact, int *sfirst,
                                          int *sx) {
    __shared__ double totU[kMaxSignals][kNV], totV[kMaxSignals][kNV];
    signal_totals(a.partU, a.totU, (it - a.st[0].it0) == 0 ? nparts_rhs : nparts_upd, S, totU);
    signal_totals(a.partV, a.totV, nparts_spmv, S, totV);
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        KrylovState &st = a.st[s];
        // The decision rests on done (changed only by the update kernel) and on the totals, which
        // every block reads alike: all blocks agree on which signals stop here and still owe their
        // pending update of x.  stop is written below by block 0 and must not be read here (a
        // block that saw it would skip that update on its rows); it is set only by this kernel
        // and turned into done by the update kernel of the same iteration, so at this kernel's
        // start it is 0 for every signal not done.
        int act = !st.done;
        const int local = it - st.it0;
        sfirst[s] = local == 0;
        sx[s] = 0;
        if (act) {
            // |r|^2 of this iteration's r: from the update kernel (partials U; the residual kernel's
            // at a solve's first iteration), or with the folded update from the preconditioner SpMV
            // (partials V); |b|^2 from the residual kernel
            const double rn2 = rfold ? totV[s][2] : totU[s][2];
            double tol2;
            if (local == 0) {
                const double t1 = a.rtol * a.rtol * totU[s][3], t2 = a.atol * a.atol;
                tol2 = t1 > t2 ? t1 : t2;
            } else {
                tol2 = st.tol2;
            }
            const double2 rho = make_double2(totV[s][0], totV[s][1]);
            const bool conv = rn2 <= tol2;
            const bool brk = !conv && rho.x == 0.0 && rho.y == 0.0;
            const bool lim = !conv && local >= a.maxit;
            const bool newly = conv || brk || lim;
            // the previous iteration's x += alpha p is still to be applied
            sx[s] = local > 0;
            salpha[s] = local > 0 ? st.alpha[(local - 1) & 1] : make_double2(0.0, 0.0);
            if (blockIdx.x == 0) {
                if (local == 0) {
                    st.tol2 = tol2;
                    st.bnorm2 = totU[s][3];
                }
                st.rnorm2 = rn2;
                st.iters = it;
                if (newly) {  // done is set by the update kernel (see stop)
                    st.stop = 1;
                    st.conv = conv ? 1 : 0;
                    st.breakdown = brk ? 1 : 0;
                } else {
                    st.rho[local & 1] = rho;
                }
            }
            if (newly) act = 0;
            else if (local > 0) sbeta[s] = c_div(rho, st.rho[(local - 1) & 1]);
        }
        sact[s] = act;
    }
    // (the all-done flag is raised by the update kernel: raised here, it would stop this kernel's
    // later blocks before they apply the pending updates of the signals that stop now)
    __syncthreads();
}

__global__ void __launch_bounds__(kThreads) cocg_p_kernel(SolveArgs a, int S, int vrpb, int nparts_rhs,
                                                          int nparts_upd, int nparts_spmv, int it, int rfold) {
    if (*a.all_done) return;
    __shared__ double2 sbeta[kMaxSignals], salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals], sfirst[kMaxSignals], sx[kMaxSignals];
    cocg_head(a, S, nparts_rhs, nparts_upd, nparts_spmv, it, rfold, sbeta, salpha, sact, sfirst, sx);
    const int64_t total = a.n * S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S) return;
    const bool act = sact[s], upd_x = sx[s];
    if (!act && !upd_x) return;
    const bool first = sfirst[s];
    const double2 beta = sbeta[s], alpha = salpha[s];
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
        if (first) {  // a solve's first iteration: no update pending, p = z
            a.p[i] = a.ph[i];
            continue;
        }
        const double2 pi = a.p[i];
        st_once(a.x + i, c_add(ld_once(a.x + i), c_mul(alpha, pi)));
        if (act) a.p[i] = c_add(a.ph[i], c_mul(beta, pi));
    }
}

// alpha = rho / p^T q (partials T, from q = A p) and r -= alpha q; partials U: [0, 0, |r|^2, 0].
// A zero p^T q before convergence is a breakdown (the host restarts the signal from its x).
__global__ void __launch_bounds__(kThreads) cocg_update_kernel(SolveArgs a, int S, int vrpb, int nparts, int it) {
    if (*a.all_done) return;
    __shared__ double tot[kMaxSignals][kNV];
    __shared__ double2 salpha[kMaxSignals];
    __shared__ int sact[kMaxSignals];
    signal_totals(a.partT, a.totT, nparts, S, tot);
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        KrylovState &st = a.st[s];
        int act = !st.done && !st.stop;
        const int local = it - st.it0;
        if (st.stop && blockIdx.x == 0) st.done = 1;  // stopped by this iteration's p-kernel
        if (act) {
            const double2 pq = make_double2(tot[s][0], tot[s][1]);
            if (pq.x == 0.0 && pq.y == 0.0) {
                if (blockIdx.x == 0) {
                    st.done = 1;
                    st.breakdown = 1;
                    st.conv = 0;
                }
                act = 0;
            } else {
                const double2 al = c_div(st.rho[local & 1], pq);
                salpha[s] = al;
                if (blockIdx.x == 0) st.alpha[local & 1] = al;
            }
        }
        sact[s] = act;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // every signal stopped: later launches exit at once
        int all = 1;
        for (int s = 0; s < S; ++s)
            if (sact[s]) all = 0;
        if (all) *a.all_done = 1;
    }
    const int64_t total = a.n * S;
    const int s = threadIdx.x % S;
    const bool valid = (int)threadIdx.x < vrpb * S;
    const bool act = valid && sact[s];
    const double2 alpha = salpha[s];
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    if (act) {
        for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total;
             i += (int64_t)gridDim.x * vrpb * S) {
            const double2 rn = c_sub(ld_once(a.r + i), c_mul(alpha, ld_once(a.v + i)));
            st_once(a.r + i, rn);
            if (a.r32) a.r32[i] = make_float2((float)rn.x, (float)rn.y);
            red[2] += rn.x * rn.x + rn.y * rn.y;
        }
    }
    block_signal_partials(red, valid, threadIdx.x, vrpb * S, S, a.partU);
}

// ---- residual-minimising initial guess over the last states (see ProjArgs)
__device__ __forceinline__ double2 c_conj_mul(double2 a, double2 b) {  // conj(a) b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// d[j] <- j-th backward difference of d[0] (newest) .. d[nh-1]: sum_i (-1)^i C(j,i) d[i]
template <int MH>
__device__ __forceinline__ void backward_differences(double2 (&d)[MH], int nh) {
#pragma unroll
    for (int lvl = 1; lvl < MH; ++lvl)
#pragma unroll
        for (int j = MH - 1; j >= lvl; --j)
            if (j < nh) d[j] = c_sub(d[j - 1], d[j]);
}

__device__ __forceinline__ double proj_coef(const ProjArgs &p, int s) {
    if (!p.per_gdir) return p.coefA[3 * s];
    return p.coefA[3 * s] * p.gsig[3 * s] + p.coefA[3 * s + 1] * p.gsig[3 * s + 1] +
           p.coefA[3 * s + 2] * p.gsig[3 * s + 2];
}

// Blocks (x, y): y selects a window of Sw signals [y Sw, y Sw + Sl); x walks tiles of
// rpt = 256 / Sl rows (thread t <-> element t of the tile: row t / Sl, window signal t % Sl).
// Load phase: each thread forms W_j = P_j + i c_s Q_j of its element for the nh states, takes the
// backward differences in registers and stores them, with b, in LDS.  Pair phase: thread t owns
// the (pair, signal) combinations q = t, t + 256 (pairs: the upper triangle (j, l), j <= l,
// row-major, then (j, b)) and sums conj(W_j) W_l over the tile's rows from LDS.  Sw keeps the
// combinations of a window within 2 per thread.  Every combination belongs to one thread, so the
// block partials need no further reduction: part[x][s][2 pair + re/im].
template <int MH>
__global__ void __launch_bounds__(kThreads) proj_gram_kernel(ProjArgs p, int Sw) {
    constexpr int NG = MH * (MH + 1) / 2;
    constexpr int NP = NG + MH;
    constexpr int NC = 2;
    extern __shared__ double2 sw[];  // [MH + 1][kThreads]
    const int S = p.S;
    const int s0 = blockIdx.y * Sw;
    const int Sl = S - s0 < Sw ? S - s0 : Sw;
    const int rpt = kThreads / Sl;
    const int nel = rpt * Sl;
    const int t = threadIdx.x;
    const int nh = p.nh;
    double2 acc[NC];
    int cx[NC], cy[NC];  // LDS offsets of the two columns of each combination (signal included)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        acc[k] = make_double2(0.0, 0.0);
        const int q = t + k * kThreads;
        int pp = q / Sl, j = 0, l;
        const int sl = q % Sl;
        if (pp < NG) {
            while (pp >= MH - j) {
                pp -= MH - j;
                ++j;
            }
            l = j + pp;
        } else {
            j = pp - NG;
            l = MH;  // the b column
        }
        cx[k] = j * kThreads + sl;
        cy[k] = l * kThreads + sl;
    }
    const int64_t ntiles = (p.n + rpt - 1) / rpt;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // the previous tile's pair phase is done with the LDS
        if (t < nel) {
            const int64_t row = tile * rpt + t / Sl;
            const int s = s0 + t % Sl;
            double2 w[MH];
#pragma unroll
            for (int j = 0; j < MH; ++j) w[j] = make_double2(0.0, 0.0);
            double2 bv = make_double2(0.0, 0.0);
            if (row < p.n) {
                const int64_t e = row * S + s;
                const double c = proj_coef(p, s);
#pragma unroll
                for (int j = 0; j < MH; ++j)
                    if (j < nh) {
                        const double2 P = p.P[j][e], Q = p.Q[j][e];
                        w[j] = make_double2(P.x - c * Q.y, P.y + c * Q.x);  // P + i c Q
                    }
                backward_differences<MH>(w, nh);
                bv = p.b[e];
            }
#pragma unroll
            for (int j = 0; j < MH; ++j) sw[j * kThreads + t] = w[j];
            sw[MH * kThreads + t] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (t + k * kThreads < NP * Sl) {
                const double2 *x = sw + cx[k];
                const double2 *y = sw + cy[k];
                double2 v = acc[k];
#pragma unroll 2
                for (int r = 0; r < rpt; ++r) v = c_add(v, c_conj_mul(x[r * Sl], y[r * Sl]));
                acc[k] = v;
            }
        }
    }
    double *part = p.part + (size_t)blockIdx.x * S * (2 * NP);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const int q = t + k * kThreads;
        if (q < NP * Sl) {
            const int pp = q / Sl, s = s0 + q % Sl;
            part[(size_t)s * (2 * NP) + 2 * pp] = acc[k].x;
            part[(size_t)s * (2 * NP) + 2 * pp + 1] = acc[k].y;
        }
    }
}

// tot[s][k] = sum over the nparts blocks of part[b][s][k], in block order per thread then a
// fixed block reduction; one block per (signal, k).
__global__ void __launch_bounds__(kThreads) proj_finalize_kernel(const double *part, int nparts, int S, int nv,
                                                                 double *tot) {
    const int s = blockIdx.x % S, k = blockIdx.x / S;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v[0] += part[((size_t)b * S + s) * nv + k];
    block_reduce<1>(v);
    if (threadIdx.x == 0) tot[s * nv + k] = v[0];
}

// Per signal (one block of one wave): Jacobi-scaled Cholesky of the Gram matrix in LDS with
// truncation at the first pivot below kProjPivot (column by column, the rows of a column in
// parallel), the two triangular solves, and the coefficients on the raw states.
template <int MH>
__global__ void __launch_bounds__(64) proj_solve_kernel(ProjArgs p) {
    constexpr int NG = MH * (MH + 1) / 2;
    constexpr int NV = proj_nv(MH);
    __shared__ double sm[NV];
    __shared__ double2 L[MH][MH];  // lower triangle: the scaled matrix, then its factor
    __shared__ double2 z[MH];
    __shared__ double D[MH];
    __shared__ int sm_m, sm_ok;
    const int s = blockIdx.x;
    const int t = threadIdx.x;
    const int nh = p.nh;
    for (int i = t; i < NV; i += blockDim.x) sm[i] = p.tot[(size_t)s * NV + i];
    __syncthreads();
    if (t < MH) {
        const int k = t * MH - t * (t - 1) / 2;  // packed index of (t, t)
        const double g = sm[2 * k];
        D[t] = (t < nh && g > 0.0) ? sqrt(g) : 0.0;
        z[t] = make_double2(sm[2 * NG + 2 * t], sm[2 * NG + 2 * t + 1]);
    }
    if (t == 0) {
        sm_m = 0;
        sm_ok = 1;
    }
    __syncthreads();
    // L(i, j) = conj(G(j, i)) / (D_i D_j), j <= i (G holds the upper triangle, row-major)
    for (int q = t; q < MH * MH; q += blockDim.x) {
        const int i = q / MH, j = q % MH;
        if (j <= i) {
            const int k = j * MH - j * (j - 1) / 2 + (i - j);
            const double sc = (D[i] > 0.0 && D[j] > 0.0) ? 1.0 / (D[i] * D[j]) : 0.0;
            L[i][j] = make_double2(sm[2 * k] * sc, -sm[2 * k + 1] * sc);
        }
    }
    __syncthreads();
    for (int j = 0; j < MH; ++j) {
        if (t == 0) {
            if (sm_ok && j < nh && D[j] > 0.0) {
                double d = L[j][j].x;
                for (int k = 0; k < j; ++k) d -= L[j][k].x * L[j][k].x + L[j][k].y * L[j][k].y;
                if (d > kProjPivot) {
                    L[j][j] = make_double2(sqrt(d), 0.0);
                    sm_m = j + 1;
                } else {
                    sm_ok = 0;
                }
            } else {
                sm_ok = 0;
            }
        }
        __syncthreads();
        if (!sm_ok) break;
        const int i = j + 1 + t;
        if (i < MH && i < nh) {
            double2 v = L[i][j];
            for (int k = 0; k < j; ++k) v = c_sub(v, c_mul(L[i][k], make_double2(L[j][k].x, -L[j][k].y)));
            L[i][j] = make_double2(v.x / L[j][j].x, v.y / L[j][j].x);
        }
        __syncthreads();
    }
    const int m = sm_m;
    if (t == 0) {
        // forward: L y = W^H b / D ; back: L^H x = y (in place in z)
        for (int j = 0; j < m; ++j) {
            double2 v = make_double2(z[j].x / D[j], z[j].y / D[j]);
            for (int k = 0; k < j; ++k) v = c_sub(v, c_mul(L[j][k], z[k]));
            z[j] = make_double2(v.x / L[j][j].x, v.y / L[j][j].x);
        }
        for (int j = m - 1; j >= 0; --j) {
            double2 v = z[j];
            for (int k = j + 1; k < m; ++k) v = c_sub(v, c_mul(make_double2(L[k][j].x, -L[k][j].y), z[k]));
            z[j] = make_double2(v.x / L[j][j].x, v.y / L[j][j].x);
        }
    }
    __syncthreads();
    // y0 = sum_j c_j Delta^j = sum_i gamma_i y_i, gamma_i = sum_{j >= i} c_j (-1)^i C(j, i), c_j = x_j / D_j
    if (t < kMaxHist) {
        double2 g = make_double2(0.0, 0.0);
        if (t < MH) {
            for (int j = t; j < m; ++j) {
                double binom = 1.0;  // C(j, t)
                for (int r = 0; r < t; ++r) binom = binom * (double)(j - r) / (double)(r + 1);
                const double f = (t & 1) ? -binom : binom;
                g.x += f * z[j].x / D[j];
                g.y += f * z[j].y / D[j];
            }
            if (m == 0 && t == 0) g = make_double2(1.0, 0.0);  // no usable direction: keep the newest state
        }
        p.gam[(size_t)s * 2 * kMaxHist + 2 * t] = g.x;
        p.gam[(size_t)s * 2 * kMaxHist + 2 * t + 1] = g.y;
    }
}

// y0 = sum_j gamma_{s,j} y_j (vector lanes: element i of the [n][S] arrays, signal i mod S)
template <int MH>
__global__ void __launch_bounds__(kThreads) proj_combine_kernel(ProjArgs p, int vrpb) {
    const int S = p.S;
    const int s = threadIdx.x % S;
    if ((int)threadIdx.x >= vrpb * S) return;
    const int nh = p.nh;
    double2 gm[MH];
#pragma unroll
    for (int j = 0; j < MH; ++j)
        gm[j] = make_double2(p.gam[(size_t)s * 2 * kMaxHist + 2 * j], p.gam[(size_t)s * 2 * kMaxHist + 2 * j + 1]);
    const int64_t total = p.n * S;
    for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total; i += (int64_t)gridDim.x * vrpb * S) {
        double2 y = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < MH; ++j)
            if (j < nh) y = c_add(y, c_mul(gm[j], p.Y[j][i]));
        p.y0[i] = y;
    }
}

// The contraction's value arrays in the solver's entry order, each entry's four values together and
// scaled like the solver's operator: adjv[e] = (s_i s_j K, s_i s_j M, s_i s_j I, Jg) (Jg already
// scaled in vals).  Built once per adjoint sweep; one thread per row.
__global__ void __launch_bounds__(kThreads) adj_values_kernel(int64_t n, const int *rowptr, const int *col,
                                                              const int *emap, const double *sc, const double *mm,
                                                              const double *kk, const double *gi,
                                                              const double *vals, double4 *adjv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double si = sc[i];
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        const double sij = si * sc[col[e]];
        const int o = emap[e];
        adjv[e] = make_double4(sij * kk[o], sij * mm[o], sij * gi[o], vals[2 * (int64_t)e + 1]);
    }
}

// Adjoint contraction of one theta step (one signal).  With the scaled states x_n, x_n1, the
// scaled adjoint mu and y = (1-theta) x_n + theta x_n1, every row i adds (complex products,
// no conjugation; s_ij = sc_i sc_j, the values from adjv):
//   k[0..3]  mu_i sum_j s_ij K_ij y_j  into compartment 0 / 1 of row i
//   k[4..7]  mu_i sum_j s_ij M_ij y_j  into compartment 0 / 1
//   k[8,9]   mu_i sum_j s_ij I_ij y_j
//   k[10,11] mu_i sum_j Jg_ij x_n,j    (Jg: the scaled g.J)
//   k[12,13] mu_i sum_j Jg_ij x_n1,j
// G lanes per row take every G-th entry (coalesced entry reads) and add their sums with shuffles.
// Block partials go to parts[block * 14 + k] and are summed on the host in block order.
template <int G>
__global__ void __launch_bounds__(kThreads) adj_contract_rows_kernel(int64_t n, const int *rowptr, const int *col,
                                                                     const int *dof_comp, const double4 *adjv,
                                                                     const double2 *mu, const double2 *xn,
                                                                     const double2 *xn1, double theta, double *parts) {
    double red[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) red[k] = 0.0;
    const int g = threadIdx.x % G;
    const int rpb = blockDim.x / G;
    for (int64_t i = (int64_t)blockIdx.x * rpb + threadIdx.x / G; i < n; i += (int64_t)gridDim.x * rpb) {
        double2 aK = make_double2(0, 0), aM = make_double2(0, 0), aI = make_double2(0, 0);
        double2 aJ0 = make_double2(0, 0), aJ1 = make_double2(0, 0);
        const int e1 = rowptr[i + 1];
        for (int e = rowptr[i] + g; e < e1; e += G) {
            const int j = col[e];
            const double4 v = adjv[e];
            const double2 x0 = xn[j], x1 = xn1[j];
            const double2 y = make_double2((1.0 - theta) * x0.x + theta * x1.x, (1.0 - theta) * x0.y + theta * x1.y);
            aK.x += v.x * y.x; aK.y += v.x * y.y;
            aM.x += v.y * y.x; aM.y += v.y * y.y;
            aI.x += v.z * y.x; aI.y += v.z * y.y;
            aJ0.x += v.w * x0.x; aJ0.y += v.w * x0.y;
            aJ1.x += v.w * x1.x; aJ1.y += v.w * x1.y;
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            aK.x += __shfl_xor(aK.x, off, 64); aK.y += __shfl_xor(aK.y, off, 64);
            aM.x += __shfl_xor(aM.x, off, 64); aM.y += __shfl_xor(aM.y, off, 64);
            aI.x += __shfl_xor(aI.x, off, 64); aI.y += __shfl_xor(aI.y, off, 64);
            aJ0.x += __shfl_xor(aJ0.x, off, 64); aJ0.y += __shfl_xor(aJ0.y, off, 64);
            aJ1.x += __shfl_xor(aJ1.x, off, 64); aJ1.y += __shfl_xor(aJ1.y, off, 64);
        }
        if (g != 0) continue;
        const double2 m = mu[i];
        const int c = dof_comp ? dof_comp[i] : 0;
        const double2 pK = c_mul(m, aK), pM = c_mul(m, aM), pI = c_mul(m, aI);
        const double2 pJ0 = c_mul(m, aJ0), pJ1 = c_mul(m, aJ1);
        red[2 * c] += pK.x;
        red[2 * c + 1] += pK.y;
        red[4 + 2 * c] += pM.x;
        red[4 + 2 * c + 1] += pM.y;
        red[8] += pI.x;
        red[9] += pI.y;
        red[10] += pJ0.x;
        red[11] += pJ0.y;
        red[12] += pJ1.x;
        red[13] += pJ1.y;
    }
    block_reduce<14>(red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 14; ++k) parts[(size_t)blockIdx.x * 14 + k] = red[k];
}

// lambda^N = c^ (scaled signal weights sc_i w_i / |Omega|)
__global__ void adj_init_lambda_kernel(int64_t n, const double *w, const double *sc, double inv_vol, double2 *lam) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lam[i] = make_double2(w[i] * sc[i] * inv_vol, 0.0);
}

// per-signal partials of sum_i w_i sc_i x[i*S+s]
// Weak pseudo-periodic terms of one step (ThetaMethodL_wBC1c/2c, DmriFemLib.py:70-75,114-120):
//   ybc[dof_e] = exp(i q_s (g_s.l_e) F) sum_k w_ek u^n[src_ek]   (u = sc x, unscaled)
// h[3s..3s+2] = q_s F g_s of this step.  One thread per (entry, signal).
__global__ void periodic_mirror_kernel(int64_t ne, int S, const int *ent_dof, const int *ent_src, const double *ent_w,
                                       const double *ent_ell, const double *sc, const double2 *x, const double *h,
                                       double2 *ybc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * S) return;
    const int64_t e = t / S;
    const int s = (int)(t % S);
    double yr = 0.0, yi = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int j = ent_src[3 * e + k];
        if (j < 0) continue;
        const double wk = ent_w[3 * e + k] * sc[j];
        const double2 xv = x[(int64_t)j * S + s];
        yr += wk * xv.x;
        yi += wk * xv.y;
    }
    const double ph = h[3 * s] * ent_ell[3 * e] + h[3 * s + 1] * ent_ell[3 * e + 1] + h[3 * s + 2] * ent_ell[3 * e + 2];
    double sn, cs;
    sincos(ph, &sn, &cs);
    ybc[(int64_t)ent_dof[e] * S + s] = make_double2(yr * cs - yi * sn, yr * sn + yi * cs);
}

// bper[i] = (1-theta) sc_i sum_j P_ij ybc_j on the rows that hold P entries (scaled right-hand side).
__global__ void periodic_rhs_kernel(int64_t nrows, int S, const int *rows, const int *rowptr, const int *col,
                                    const double *pv, const double *sc, const double2 *ybc, double omt, double2 *bper) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * S) return;
    const int i = rows[t / S];
    const int s = (int)(t % S);
    double ar = 0.0, ai = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        const double pvv = pv[e];
        if (pvv == 0.0) continue;
        const double2 y = ybc[(int64_t)col[e] * S + s];
        ar += pvv * y.x;
        ai += pvv * y.y;
    }
    const double f = omt * sc[i];
    bper[(int64_t)i * S + s] = make_double2(f * ar, f * ai);
}

__global__ void __launch_bounds__(kThreads) signal_part_kernel(int64_t n, int S, int vrpb, const double *w,
                                                               const double *sc, const double2 *x, double *parts) {
    const int64_t total = n * S;
    const bool valid = (int)threadIdx.x < vrpb * S;
    double red[kNV] = {0.0, 0.0, 0.0, 0.0};
    if (valid) {
        for (int64_t i = (int64_t)blockIdx.x * vrpb * S + threadIdx.x; i < total;
             i += (int64_t)gridDim.x * vrpb * S) {
            const int64_t row = i / S;
            const double ws = w[row] * sc[row];
            red[0] += ws * x[i].x;
            red[1] += ws * x[i].y;
        }
    }
    block_signal_partials(red, valid, threadIdx.x, vrpb * S, S, parts);
}

}  // namespace

// ======================================================================================
// launchers
// ======================================================================================
static inline void check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(-2, std::string("kernel launch: ") + hipGetErrorString(e));
}

void exclusive_scan(const int *d_in, int *d_out, int64_t n, hipStream_t s) {
    if (n <= 0) {
        DMRI_HIP(hipMemsetAsync(d_out, 0, sizeof(int), s));
        return;
    }
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    int *tile_sums = nullptr, *tile_offs = nullptr;
    DMRI_HIP(hipMallocAsync((void **)&tile_sums, sizeof(int) * ntiles, s));
    DMRI_HIP(hipMallocAsync((void **)&tile_offs, sizeof(int) * (ntiles + 1), s));
    scan_tile_kernel<<<(unsigned)ntiles, kScanBlock, 0, s>>>(d_in, d_out, tile_sums, n);
    check_launch();
    if (ntiles > 1) {
        exclusive_scan(tile_sums, tile_offs, ntiles, s);
        scan_add_kernel<<<(unsigned)ntiles, kScanBlock, 0, s>>>(d_out, tile_offs, n);
        check_launch();
    }
    set_total_kernel<<<1, 1, 0, s>>>(d_out, n, d_in, tile_offs, (int)ntiles);
    check_launch();
    DMRI_HIP(hipFreeAsync(tile_sums, s));
    DMRI_HIP(hipFreeAsync(tile_offs, s));
}

void launch_count_incidence(const int *tets, int64_t nc, int *cnt, hipStream_t s) {
    if (nc <= 0) return;
    count_incidence_kernel<4><<<grid_for(nc, 256), 256, 0, s>>>(tets, nc, cnt);
    check_launch();
}
void launch_fill_incidence(const int *tets, int64_t nc, const int *ptr, int *cursor, int *list, hipStream_t s) {
    if (nc <= 0) return;
    fill_incidence_kernel<4><<<grid_for(nc, 256), 256, 0, s>>>(tets, nc, ptr, cursor, list);
    check_launch();
}
void launch_count_incidence3(const int *fv, int64_t nf, int *cnt, hipStream_t s) {
    if (nf <= 0) return;
    count_incidence_kernel<3><<<grid_for(nf, 256), 256, 0, s>>>(fv, nf, cnt);
    check_launch();
}
void launch_fill_incidence3(const int *fv, int64_t nf, const int *ptr, int *cursor, int *list, hipStream_t s) {
    if (nf <= 0) return;
    fill_incidence_kernel<3><<<grid_for(nf, 256), 256, 0, s>>>(fv, nf, ptr, cursor, list);
    check_launch();
}
void launch_sort_lists(const int *ptr, int *list, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    sort_lists_kernel<<<grid_for(n, 256), 256, 0, s>>>(ptr, list, n);
    check_launch();
}
void launch_mark_live(const int *tets, const int *phase, int64_t nc, int64_t nv, int *live, hipStream_t s) {
    mark_live_kernel<<<grid_for(nc, 256), 256, 0, s>>>(tets, phase, nc, nv, live);
    check_launch();
}
void launch_dof_from_scan(const int *live, const int *scan, int64_t n, int *dof_of, int64_t nv, int *dof_vert,
                          int *dof_comp, hipStream_t s) {
    dof_from_scan_kernel<<<grid_for(n, 256), 256, 0, s>>>(live, scan, n, dof_of, nv, dof_vert, dof_comp);
    check_launch();
}
void launch_face_classify(const int *tets, const int *phase, int64_t nc, const int *vptr, const int *vlist,
                          int *iface_flag, int *bnd_flag, int *nbr, hipStream_t s) {
    face_classify_kernel<<<grid_for(4 * nc, 256), 256, 0, s>>>(tets, phase, nc, vptr, vlist, iface_flag, bnd_flag,
                                                                nbr);
    check_launch();
}
void launch_face_compact(const int *tets, int64_t nc, const int *nbr, const int *iface_flag, const int *iface_scan,
                         const int *bnd_flag, const int *bnd_scan, int *iface_v, int *iface_c, int *bnd_v,
                         int *bnd_c, hipStream_t s) {
    face_compact_kernel<<<grid_for(4 * nc, 256), 256, 0, s>>>(tets, nc, nbr, iface_flag, iface_scan, bnd_flag,
                                                               bnd_scan, iface_v, iface_c, bnd_v, bnd_c);
    check_launch();
}
void launch_pattern(int pass, int64_t ndof, int64_t nv, const int *dof_vert, const int *dof_comp, const int *tets,
                    const int *phase, const int *vptr, const int *vlist, const int *fptr, const int *flist,
                    const int *iface_v, const int *dof_of, int *row_count, const int *rowptr, int *col,
                    int *overflow, hipStream_t s) {
    if (ndof <= 0) return;
    pattern_kernel<<<(unsigned)ndof, 64, 0, s>>>(pass, dof_vert, dof_comp, tets, phase, nv, vptr, vlist, fptr, flist,
                                                 iface_v, dof_of, row_count, rowptr, col, overflow);
    check_launch();
}
void launch_jp_round(int k, int64_t ne, int nvert, const int *ev, const int *vptr, const int *vlist, int *color,
                     int *remaining, hipStream_t s) {
    jp_round_kernel<<<grid_for(ne, 256), 256, 0, s>>>(k, ne, nvert, ev, vptr, vlist, color, remaining);
    check_launch();
}
void launch_color_hist(const int *color, int64_t ne, int *hist, hipStream_t s) {
    color_hist_kernel<<<grid_for(ne, 256), 256, 0, s>>>(color, ne, hist);
    check_launch();
}
void launch_color_scatter(const int *color, int64_t ne, const int *offs, int *cursor, int *perm, hipStream_t s) {
    color_scatter_kernel<<<grid_for(ne, 256), 256, 0, s>>>(color, ne, offs, cursor, perm);
    check_launch();
}
void launch_assemble_rows(int64_t n, const int *dof_vert, const int *dof_comp, const int *rowptr, const int *col,
                          const int *tets, const int *phase, const double *xyz, const int *vptr, const int *vlist,
                          const double *Dc, const double *Dcell, const double *iT2c, const double *iT2cell, double *m,
                          double *k, double *r, double *jx, double *jy, double *jz, hipStream_t s, double *cx,
                          double *cy, double *cz) {
    if (n <= 0) return;
    assemble_rows_kernel<<<grid_for(n * kRowLanes, 256), 256, 0, s>>>(n, dof_vert, dof_comp, rowptr, col, tets, phase,
                                                                       xyz, vptr, vlist, Dc, Dcell, iT2c, iT2cell, m,
                                                                       k, r, jx, jy, jz, cx, cy, cz);
    check_launch();
}

// --------------------------------------------------------------------------------------
// Strong pseudo-periodic constraint (set-up).
// --------------------------------------------------------------------------------------
// g[e] = (c[e] - c[e^T]) / 2 on a structurally symmetric pattern (one thread per row)
__global__ void antisym_kernel(int64_t n, const int *rowptr, const int *col, const double *c, double *g, int *miss) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        const int pt = csr_find(rowptr, col, col[e], (int)i);
        if (pt < 0) {
            atomicOr(miss, 1);
            continue;
        }
        g[e] = 0.5 * (c[e] - c[pt]);
    }
}

// key[d] = DOF of (compartment of d, master vertex of d's vertex); bad when that DOF is dead
__global__ void master_dof_kernel(int64_t n, int64_t nv, const int *dof_vert, const int *dof_comp, const int *master,
                                  const int *dof_of, int *key, int *is_master, int *bad) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const int v = dof_vert[d], c = dof_comp[d];
    const int k = dof_of[(int64_t)c * nv + master[v]];
    if (k < 0) atomicOr(bad, 1);
    key[d] = k < 0 ? (int)d : k;
    is_master[d] = key[d] == (int)d;
}

// mid[d] = constrained id of d (the scan position of its master)
__global__ void merged_id_kernel(int64_t n, const int *key, const int *scan, int *mid) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    mid[d] = scan[key[d]];
}

// entry keys (constrained row * nm + constrained column) and their original positions
__global__ void merge_key_kernel(int64_t n, const int *rowptr, const int *col, const int *mid, int64_t nm,
                                 uint64_t *ekey, int *eidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = (uint64_t)mid[i];
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        ekey[e] = r * (uint64_t)nm + (uint64_t)mid[col[e]];
        eidx[e] = e;
    }
}

// head[p] = 1 where a sorted key starts a run
__global__ void run_head_kernel(int64_t nnz, const uint64_t *skey, int *head) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    head[p] = (p == 0 || skey[p] != skey[p - 1]) ? 1 : 0;
}

// run starts rstart[u] (u = run index = hscan[p] at a head) and the constrained CSR column and row counts
__global__ void run_fill_kernel(int64_t nnz, int64_t nm, const uint64_t *skey, const int *head, const int *hscan,
                                int *rstart, int *mcol, int *rowcnt) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz || !head[p]) return;
    const int u = hscan[p];
    rstart[u] = (int)p;
    mcol[u] = (int)(skey[p] % (uint64_t)nm);
    atomicAdd(&rowcnt[skey[p] / (uint64_t)nm], 1);
}

// merged[u] = sum of val over the run's entries in sorted (= original, the sort is stable) order
__global__ void run_sum_kernel(int64_t nu, int64_t nnz, const int *rstart, const int *sidx, const double *val,
                               double *merged) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int p1 = u + 1 < nu ? rstart[u + 1] : (int)nnz;
    double acc = 0.0;
    for (int p = rstart[u]; p < p1; ++p) acc += val[sidx[p]];
    merged[u] = acc;
}

void launch_antisym(int64_t n, const int *rowptr, const int *col, const double *c, double *g, int *miss,
                    hipStream_t s) {
    antisym_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, rowptr, col, c, g, miss);
    check_launch();
}

void launch_master_dofs(int64_t n, int64_t nv, const int *dof_vert, const int *dof_comp, const int *master,
                        const int *dof_of, int *key, int *is_master, int *bad, hipStream_t s) {
    master_dof_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, nv, dof_vert, dof_comp, master, dof_of, key, is_master, bad);
    check_launch();
}

void launch_merged_ids(int64_t n, const int *key, const int *scan, int *mid, hipStream_t s) {
    merged_id_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, key, scan, mid);
    check_launch();
}

int64_t build_merge(int64_t n, int64_t nnz, const int *rowptr, const int *col, const int *mid, int64_t nm,
                    DevArray<int> &mrowptr, DevArray<int> &mcol, DevArray<int> &rstart, DevArray<int> &sidx,
                    hipStream_t s) {
    DevArray<uint64_t> key, skey;
    DevArray<int> idx, head, hscan, cnt;
    key.alloc(nnz);
    skey.alloc(nnz);
    idx.alloc(nnz);
    sidx.alloc(nnz);
    merge_key_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, rowptr, col, mid, nm, key.p, idx.p);
    check_launch();
    int bits = 1;
    while (bits < 64 && ((uint64_t)1 << bits) < (uint64_t)nm * (uint64_t)nm) ++bits;
    size_t tmp_bytes = 0;
    DMRI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, skey.p, idx.p, sidx.p, (int)nnz, 0, bits, s));
    DevArray<char> tmp;
    tmp.alloc(std::max<size_t>(tmp_bytes, 1));
    DMRI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, skey.p, idx.p, sidx.p, (int)nnz, 0, bits, s));
    head.alloc(nnz);
    hscan.alloc(nnz + 1);
    run_head_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, skey.p, head.p);
    check_launch();
    exclusive_scan(head.p, hscan.p, nnz, s);
    int nu = 0;
    DMRI_HIP(hipMemcpyAsync(&nu, hscan.p + nnz, sizeof(int), hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    rstart.alloc(std::max(nu, 1));
    mcol.alloc(std::max(nu, 1));
    cnt.alloc(nm);
    cnt.zero(s);
    run_fill_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, nm, skey.p, head.p, hscan.p, rstart.p, mcol.p, cnt.p);
    check_launch();
    mrowptr.alloc(nm + 1);
    exclusive_scan(cnt.p, mrowptr.p, nm, s);
    DMRI_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
    return nu;
}

void launch_run_sum(int64_t nu, int64_t nnz, const int *rstart, const int *sidx, const double *val, double *merged,
                    hipStream_t s) {
    run_sum_kernel<<<grid_for(nu, 256), 256, 0, s>>>(nu, nnz, rstart, sidx, val, merged);
    check_launch();
}
void launch_assemble_facets(int64_t start, int64_t count, const int *perm, const int *iface_v, const double *xyz,
                            int64_t nv, const int *dof_of, const int *rowptr, const int *col, double kappa,
                            const double *kappa_f, double *gi, int *miss, hipStream_t s) {
    if (count <= 0) return;
    assemble_facets_kernel<<<grid_for(count, 128), 128, 0, s>>>(start, count, perm, iface_v, xyz, nv, dof_of, rowptr,
                                                                col, kappa, kappa_f, gi, miss);
    check_launch();
}
void launch_count_sum64(const int *cnt, int64_t n, unsigned long long *out, hipStream_t s) {
    unsigned g = grid_for(n, 256);
    if (g > 1024) g = 1024;
    if (g == 0) g = 1;
    count_sum64_kernel<<<g, 256, 0, s>>>(cnt, n, out);
    check_launch();
}

void launch_diag_and_weights(int64_t n, const int *rowptr, const int *col, const double *m, int *diag, double *w,
                             int *missing_diag, hipStream_t s) {
    diag_weights_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, rowptr, col, m, diag, w, missing_diag);
    check_launch();
}

SolveGeom make_geom(int64_t n, int G, int S, int mapping) {
    if (S < 1 || S > kMaxSignals) throw Error(-1, "signals per launch must lie in [1, 64]");
    if (G < 1 || G > 32 || (G & (G - 1))) throw Error(-1, "lanes per row must be a power of two in [1, 32]");
    if (mapping != kMapSL && mapping != kMapStaged && mapping != kMapSR && mapping != kMapWave && mapping != kMapTile &&
        mapping != kMapPair)
        throw Error(-1, "unknown SpMV mapping");
    int pair_sp = 1;
    while (2 * pair_sp < S) pair_sp *= 2;
    const bool pairs = mapping == kMapPair;
    if (pairs && (S < 2 || G > 4 || G * pair_sp > 64))
        throw Error(-1, "the two-signals-per-lane mapping needs at least 2 signals, lanes 1, 2 or 4 and "
                        "lanes x signal pairs <= 64");
    if (mapping == kMapPair && (uint64_t)n * (uint64_t)S * 16u >= (1ull << 32))
        throw Error(-1, "the two-signals-per-lane mapping addresses vectors of at most 4 GiB (n x signals x 16 B)");
    if (mapping == kMapWave && G != 1) throw Error(-1, "the wave-per-row mapping takes one lane per signal (lanes 1)");
    if (mapping == kMapSR && S > kMaxSRSignals)
        throw Error(-1, "signals-in-registers mapping needs at most 8 signals per launch group");
    if (mapping != kMapSR && !pairs && G * S > 64) throw Error(-1, "lanes per row times signals must not exceed 64");
    SolveGeom g;
    g.G = G;
    g.S = S;
    g.mapping = mapping;
    // the residual kernel always spreads signals over lanes
    g.G_rhs = mapping == kMapSR || pairs ? std::max(1, std::min(G, 64 / S)) : G;
    while (g.G_rhs & (g.G_rhs - 1)) g.G_rhs &= g.G_rhs - 1;
    g.rhs_rpb = std::max(1, kThreads / (g.G_rhs * S));
    g.rhs_threads = g.rhs_rpb * g.G_rhs * S;
    g.rpb = mapping == kMapSR ? g.rhs_rpb : std::max(1, kThreads / (G * S));
    g.spmv_threads = mapping == kMapSR ? g.rhs_threads : g.rpb * G * S;
    if (mapping == kMapWave) {  // one wave per row
        g.rpb = kThreads / 64;
        g.spmv_threads = kThreads;
    }
    if (mapping == kMapTile) g.tile_rows = g.rpb;
    if (pairs) {  // G x pair_sp lanes per row
        g.pair_sp = pair_sp;
        g.rpb = kThreads / (G * pair_sp);
        g.spmv_threads = kThreads;
    }
    g.sr_rpb = std::max(1, kThreads / G);
    g.sr_threads = g.sr_rpb * G;
    g.vec_rpb = std::max(1, kThreads / S);
    g.vec_threads = g.vec_rpb * S;
    int64_t need = (n + g.rhs_rpb - 1) / g.rhs_rpb;
    need = (need + 7) / 8 * 8;  // XCD-contiguous row chunks need a multiple of 8 blocks
    g.grid = (int)std::max<int64_t>(8, std::min<int64_t>(need, kMaxGridSpmv));
    int64_t vneed = (n + g.vec_rpb - 1) / g.vec_rpb;
    g.vgrid = (int)std::max<int64_t>(1, std::min<int64_t>(vneed, kMaxGrid));
    g.pgrid = (int)std::max<int64_t>(1, std::min<int64_t>(vneed, kMaxProjGrid));
    return g;
}

// rows staged per tile: a multiple of the rows one pass of the block covers, about 64 rows
// (32 with per-signal directions, whose entries carry four values)
int stage_rows(const SolveGeom &g, bool per_gdir) {
    // per-signal directions: 16 rows (36 B per staged entry), so that the tile's LDS lets the 6
    // blocks per CU that the kernels' register budget allows be resident
    const int target = per_gdir ? 16 : 64;
    return g.rpb * std::max(1, target / g.rpb);
}
size_t stage_lds_bytes(int TR, bool per_gdir) {
    const size_t ecap = (size_t)TR * kStageNnzPerRow;
    return ecap * (8 * (per_gdir ? 4 : 2) + 4) + 4 * (size_t)(TR + 1) + 32;
}
// staged tile of the real-part-only (preconditioner) kernels: one value per entry
static size_t stage_lds_bytes_ro(int TR) {
    const size_t ecap = (size_t)TR * kStageNnzPerRow;
    return ecap * (8 + 4) + 4 * (size_t)(TR + 1) + 32;
}
// staged tile of FP32 values, vw per entry
static size_t stage_lds_bytes_f32(int TR, int vw) {
    const size_t ecap = (size_t)TR * kStageNnzPerRow;
    return ecap * (4 * (size_t)vw + 4) + 4 * (size_t)(TR + 1) + 32;
}
static size_t stage_lds_bytes_bre(int TR, bool per_gdir) {
    return stage_lds_bytes(TR, per_gdir) + (size_t)TR * kStageNnzPerRow * 8;
}

void launch_sum(const double *x, int64_t n, double *part, double *out, hipStream_t s) {
    unsigned g = grid_for(n, kThreads);
    if (g > (unsigned)kMaxGrid) g = kMaxGrid;
    if (g == 0) g = 1;
    sum_part_kernel<<<g, kThreads, 0, s>>>(x, n, part);
    check_launch();
    finalize_kernel<<<1, kThreads, 0, s>>>(part, (int)g, 1, out);
    check_launch();
}

void launch_prep_scale(int64_t n, const int *diag, const double *m, const double *k, const double *r,
                       const double *gi, const double *pv, double dt, double theta, double *sc, hipStream_t s) {
    prep_scale_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, diag, m, k, r, gi, pv, dt, theta, sc);
    check_launch();
}
void launch_prep_values(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc,
                        const double *m, const double *k, const double *r, const double *gi, const double *pv,
                        const double *jx, const double *jy, const double *jz, bool per_gdir, double gx, double gy,
                        double gz, double dt, double theta, double *vals, double *bre, hipStream_t s,
                        double theta_b) {
    if (theta_b < 0.0) theta_b = 1.0 - theta;
    prep_values_kernel<<<grid_for(n * kPrepLanes, 256), 256, 0, s>>>(n, rowptr, col, emap, sc, m, k, r, gi, pv, jx,
                                                                      jy, jz, per_gdir ? 1 : 0, gx, gy, gz, dt, theta,
                                                                      theta_b, vals, bre);
    check_launch();
}

// strong periodic form: mgg[e] = sc_i sc_j D_c M (the (g.D g) q^2 F^2 mass term per unit q^2 F^2)
__global__ void sbc_mgg_kernel(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc,
                               const double *m, const int *scomp, double D0, double D1, double *mgg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Dc = (scomp && scomp[i] == 1) ? D1 : D0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) mgg[e] = sc[i] * sc[col[e]] * Dc * m[emap[e]];
}

// per step: the real implicit part afix + s2a mgg into vals (vw values per entry; and into areal when
// set), the explicit real part bfix + s2b mgg into bre
__global__ void sbc_step_kernel(int64_t nnz, int vw, const double *afix, const double *bfix, const double *mgg,
                                double s2a, double s2b, double *vals, double *bre, double *areal) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const double mg = mgg[e];
    const double a = afix[e] + s2a * mg;
    vals[(int64_t)vw * e] = a;
    if (areal) areal[e] = a;
    bre[e] = bfix[e] + s2b * mg;
}

void launch_sbc_mgg(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc, const double *m,
                    const int *scomp, double D0, double D1, double *mgg, hipStream_t s) {
    sbc_mgg_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, rowptr, col, emap, sc, m, scomp, D0, D1, mgg);
    check_launch();
}

void launch_sbc_step(int64_t nnz, int vw, const double *afix, const double *bfix, const double *mgg, double s2a,
                     double s2b, double *vals, double *bre, double *areal, hipStream_t s) {
    sbc_step_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(nnz, vw, afix, bfix, mgg, s2a, s2b, vals, bre, areal);
    check_launch();
}

// constrained DOF map: dof_of[c nv + v] <- mid[dof_of] (-1 stays), and per kept (master) DOF its
// vertex and compartment at its constrained id
__global__ void remap_dof_of_kernel(int64_t nl, const int *mid, int *dof_of) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const int d = dof_of[i];
    if (d >= 0) dof_of[i] = mid[d];
}
__global__ void compact_dofs_kernel(int64_t n, const int *is_master, const int *mid, const int *dof_vert,
                                    const int *dof_comp, int *nvert, int *ncomp) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n || !is_master[d]) return;
    nvert[mid[d]] = dof_vert[d];
    ncomp[mid[d]] = dof_comp[d];
}
void launch_constrained_dofs(int64_t n, int64_t nl, const int *is_master, const int *mid, const int *dof_vert,
                             const int *dof_comp, int *dof_of, int *nvert, int *ncomp, hipStream_t s) {
    remap_dof_of_kernel<<<grid_for(nl, 256), 256, 0, s>>>(nl, mid, dof_of);
    check_launch();
    compact_dofs_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, is_master, mid, dof_vert, dof_comp, nvert, ncomp);
    check_launch();
}

void launch_solver_order(int64_t n, int64_t nnz, const int *dof_vert, const int *dof_comp, const double *xyz,
                         const double lo[3], double ext, const int *rowptr, const int *col, const int *diag,
                         const double *w, bool two_comp, int *perm, int *iperm, int *srow, int *scol, int *emap,
                         int *sdiag, double *sw, int *scomp, hipStream_t s) {
    DevArray<uint64_t> key, key_sorted;
    DevArray<int> idx, len;
    key.alloc(n);
    key_sorted.alloc(n);
    idx.alloc(n);
    len.alloc(n);
    morton_key_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, dof_vert, dof_comp, xyz, lo[0], lo[1], lo[2],
                                                        ext > 0 ? 1.0 / ext : 0.0, key.p, idx.p);
    check_launch();
    size_t tmp_bytes = 0;
    DMRI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key_sorted.p, idx.p, perm, (int)n, 0, 61,
                                                s));
    DevArray<char> tmp;
    tmp.alloc(std::max<size_t>(tmp_bytes, 1));
    DMRI_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, key.p, key_sorted.p, idx.p, perm, (int)n, 0, 61, s));
    invert_perm_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, perm, iperm);
    check_launch();
    perm_row_len_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, perm, rowptr, len.p);
    check_launch();
    exclusive_scan(len.p, srow, n, s);
    perm_fill_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, perm, iperm, rowptr, col, srow, diag, w, dof_comp, scol, emap,
                                                       sdiag, sw, two_comp ? scomp : nullptr);
    check_launch();
    (void)nnz;
    DMRI_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
}
void build_tile_compression(int64_t n, int64_t nnz, const int *srow, const int *scol, int TR, TileCompression &tc,
                            hipStream_t s) {
    const int64_t nt = (n + TR - 1) / TR;
    tc.TR = TR;
    tc.tinfo.alloc(nt);
    tc.tptr.alloc(nt + 1);
    tc.lcol.alloc(nnz);
    tile_compress_kernel<<<(unsigned)nt, 256, 0, s>>>(0, n, TR, srow, scol, tc.tinfo.p, nullptr, nullptr, nullptr);
    check_launch();
    DevArray<int> cnt;
    cnt.alloc(nt);
    clamp_count_kernel<<<grid_for(nt, 256), 256, 0, s>>>(nt, tc.tinfo.p, cnt.p);
    check_launch();
    exclusive_scan(cnt.p, tc.tptr.p, nt, s);
    int total = 0;
    DMRI_HIP(hipMemcpyAsync(&total, tc.tptr.p + nt, sizeof(int), hipMemcpyDeviceToHost, s));
    DMRI_HIP(hipStreamSynchronize(s));
    tc.nunique = total;
    {
        std::vector<int> ti(nt);
        DMRI_HIP(hipMemcpy(ti.data(), tc.tinfo.p, sizeof(int) * nt, hipMemcpyDeviceToHost));
        tc.maxnu = 0;
        for (int v : ti) tc.maxnu = std::max(tc.maxnu, v);
    }
    tc.ucol.alloc(std::max(total, 1));
    tile_compress_kernel<<<(unsigned)nt, 256, 0, s>>>(1, n, TR, srow, scol, tc.tinfo.p, tc.tptr.p, tc.ucol.p,
                                                       tc.lcol.p);
    check_launch();
    DMRI_HIP(hipStreamSynchronize(s));
}
void launch_gather_double(int64_t n, const int *map, const double *src, double *dst, hipStream_t s) {
    if (n <= 0) return;
    gather_double_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, map, src, dst);
    check_launch();
}
void launch_renumber(int64_t n, const int *iperm, const int *src, int *dst, hipStream_t s) {
    if (n <= 0) return;
    renumber_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, iperm, src, dst);
    check_launch();
}
void launch_unscale_perm(int64_t n, int S, int sig, const int *perm, const double *sc, const double2 *x, double2 *u,
                         hipStream_t s) {
    unscale_perm_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, S, sig, perm, sc, x, u);
    check_launch();
}
void launch_periodic_mass(int64_t n, int64_t nv, const int *dof_vert, const int *dof_comp, const int *bptr,
                          const int *blist, const int *bnd_v, const int *bnd_c, const int *phase, const double *xyz,
                          const int *dof_of, const int *rowptr, const int *col, const PeriodicBox &B, double *pv,
                          int *rowflag, int *miss, hipStream_t s) {
    periodic_mass_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, nv, dof_vert, dof_comp, bptr, blist, bnd_v, bnd_c, phase,
                                                           xyz, dof_of, rowptr, col, B, pv, rowflag, miss);
    check_launch();
}
void launch_periodic_flag(int64_t n, const int *dof_vert, const double *xyz, const PeriodicBox &B, int *flag,
                          hipStream_t s) {
    periodic_flag_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, dof_vert, xyz, B, flag);
    check_launch();
}
void launch_bfacet_bin(int64_t nbf, const int *bnd_v, const double *xyz, const PeriodicBox &B, const PlaneGrids &G,
                       int *fbin, int *cnt, hipStream_t s) {
    if (nbf <= 0) return;
    bfacet_bin_kernel<<<grid_for(nbf, 256), 256, 0, s>>>(nbf, bnd_v, xyz, B, G, fbin, cnt);
    check_launch();
}
void launch_bfacet_fill(int64_t nbf, const int *fbin, const int *binptr, int *cursor, int *binlist, hipStream_t s) {
    if (nbf <= 0) return;
    bfacet_fill_kernel<<<grid_for(nbf, 256), 256, 0, s>>>(nbf, fbin, binptr, cursor, binlist);
    check_launch();
}
void launch_periodic_entries(int64_t n, int64_t nv, const int *flag, const int *scan, const int *dof_vert,
                             const int *dof_comp, const int *dof_of, const double *xyz, const int *bnd_v,
                             const int *binptr, const int *binlist, const PeriodicBox &B, const PlaneGrids &G,
                             int *ent_dof, int *ent_src, double *ent_w, double *ent_ell, int *bad, hipStream_t s) {
    periodic_entry_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, nv, flag, scan, dof_vert, dof_comp, dof_of, xyz, bnd_v,
                                                            binptr, binlist, B, G, ent_dof, ent_src, ent_w, ent_ell,
                                                            bad);
    check_launch();
}
void launch_compact_rows(int64_t n, const int *flag, const int *scan, int *rows, hipStream_t s) {
    compact_rows_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, flag, scan, rows);
    check_launch();
}
void launch_periodic_step(int64_t ne, int64_t nrows, int S, const int *ent_dof, const int *ent_src,
                          const double *ent_w, const double *ent_ell, const int *rows, const int *rowptr,
                          const int *col, const double *pv, const double *sc, const double2 *x, const double *h,
                          double omt, double2 *ybc, double2 *bper, hipStream_t s) {
    if (ne > 0) {
        periodic_mirror_kernel<<<grid_for(ne * S, 256), 256, 0, s>>>(ne, S, ent_dof, ent_src, ent_w, ent_ell, sc, x,
                                                                      h, ybc);
        check_launch();
    }
    if (nrows > 0) {
        periodic_rhs_kernel<<<grid_for(nrows * S, 256), 256, 0, s>>>(nrows, S, rows, rowptr, col, pv, sc, ybc, omt,
                                                                      bper);
        check_launch();
    }
}
__global__ void real_spmv_kernel(int64_t n, const int *rowptr, const int *col, const double *vals, int vw,
                                 const double *x, double *y, double *xy, double *yy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += vals[(int64_t)vw * e] * x[col[e]];
    y[i] = acc;
    xy[i] = x[i] * acc;
    yy[i] = acc * acc;
}
__global__ void scale_vec_kernel(int64_t n, const double *y, double scale, double *x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = y[i] * scale;
}
__global__ void power_start_kernel(int64_t n, double *x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 1.0 + 0.5 * sin(0.7 * (double)i);  // no symmetry the top eigenvector could miss
}
void launch_real_spmv(int64_t n, const int *rowptr, const int *col, const double *vals, int vw, const double *x,
                      double *y, double *xy, double *yy, hipStream_t s) {
    real_spmv_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, rowptr, col, vals, vw, x, y, xy, yy);
    check_launch();
}
void launch_scale_vec(int64_t n, const double *y, double scale, double *x, hipStream_t s) {
    scale_vec_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, y, scale, x);
    check_launch();
}
void launch_power_start(int64_t n, double *x, hipStream_t s) {
    power_start_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, x);
    check_launch();
}
void launch_gershgorin(int64_t n, const int *rowptr, const double *vals, int vw, double *part, double *out,
                       hipStream_t s) {
    unsigned g = grid_for(n, kThreads);
    if (g > (unsigned)kMaxGrid) g = kMaxGrid;
    if (g == 0) g = 1;
    row_absmax_kernel<<<g, kThreads, 0, s>>>(n, rowptr, vals, vw, part);
    check_launch();
    max_final_kernel<<<1, kThreads, 0, s>>>(part, (int)g, out);
    check_launch();
}
void launch_cdiag_prep(int64_t n, const int *rowptr, const int *col, const double *vals, int vw, int S,
                       const double *gc, double breal, double *pdiag, double *part, double *out, hipStream_t s) {
    if (vw != 2 && vw != 4) throw Error(-1, "cdiag_prep: 2 or 4 values per entry");
    unsigned g = grid_for(n, kThreads);
    if (g > (unsigned)kMaxGrid) g = kMaxGrid;
    if (g == 0) g = 1;
    cdiag_prep_kernel<<<g, kThreads, 0, s>>>(n, rowptr, col, vals, vw, S, gc, breal, pdiag, part);
    check_launch();
    max_final_kernel<<<1, kThreads, 0, s>>>(part, (int)g, out);
    check_launch();
}
// x[i][s] = u0(vertex of solver DOF i) / sc_i: a custom initial state in the scaled variables
__global__ void init_x_vertex_kernel(int64_t n, int S, const double *sc, const int *perm, const int *dof_vert,
                                     const double *u0, double2 *x) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n * S) return;
    const int64_t i = k / S;
    x[k] = make_double2(u0[dof_vert[perm[i]]] / sc[i], 0.0);
}
void launch_init_x_vertex(int64_t n, int S, const double *sc, const int *perm, const int *dof_vert, const double *u0,
                          double2 *x, hipStream_t s) {
    init_x_vertex_kernel<<<grid_for(n * S, 256), 256, 0, s>>>(n, S, sc, perm, dof_vert, u0, x);
    check_launch();
}
void launch_init_x(int64_t n, int S, const double *sc, double2 *x, hipStream_t s) {
    init_x_kernel<<<grid_for(n * S, 256), 256, 0, s>>>(n, S, sc, x);
    check_launch();
}
void launch_adj_values(int64_t n, const int *rowptr, const int *col, const int *emap, const double *sc,
                       const double *mm, const double *kk, const double *gi, const double *vals, double4 *adjv,
                       hipStream_t s) {
    adj_values_kernel<<<grid_for(n, kThreads), kThreads, 0, s>>>(n, rowptr, col, emap, sc, mm, kk, gi, vals, adjv);
    check_launch();
}
int launch_adj_contract_rows(int64_t n, const int *rowptr, const int *col, const int *dof_comp, const double4 *adjv,
                             const double2 *mu, const double2 *xn, const double2 *xn1, double theta, double *parts,
                             hipStream_t s) {
    constexpr int G = 8;  // lanes per row
    unsigned g = (unsigned)std::min<int64_t>((n + kThreads / G - 1) / (kThreads / G), kMaxGrid);
    if (g == 0) g = 1;
    adj_contract_rows_kernel<G><<<g, kThreads, 0, s>>>(n, rowptr, col, dof_comp, adjv, mu, xn, xn1, theta, parts);
    check_launch();
    return (int)g;
}
void launch_adj_init_lambda(int64_t n, const double *w, const double *sc, double inv_vol, double2 *lam,
                            hipStream_t s) {
    adj_init_lambda_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, w, sc, inv_vol, lam);
    check_launch();
}

void launch_copy(const double2 *src, double2 *dst, int64_t n, hipStream_t s) {
    // a launch holds fewer than 2^32 work-items
    if (n > (int64_t(1) << 32) - kThreads) throw Error(-1, "copy: more than 64 GiB in one launch");
    const int64_t blocks = (n + kThreads - 1) / kThreads;
    copy_kernel<<<(unsigned)std::max<int64_t>(1, blocks), kThreads, 0, s>>>(src, dst, n);
    check_launch();
}

void launch_unscale(int64_t n, int S, int sig, const double *sc, const double2 *x, double2 *u, hipStream_t s) {
    unscale_kernel<<<grid_for(n, 256), 256, 0, s>>>(n, S, sig, sc, x, u);
    check_launch();
}

// dispatch on the lanes-per-row template parameter and the gradient mode
#define DMRI_DISPATCH_G(G_, PG_, CALL)                                                      \
    do {                                                                                    \
        switch (G_) {                                                                       \
            case 1: { constexpr int GG = 1; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            case 2: { constexpr int GG = 2; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            case 4: { constexpr int GG = 4; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            case 8: { constexpr int GG = 8; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            case 16: { constexpr int GG = 16; if (PG_) { constexpr bool PP = true; CALL; } \
                       else { constexpr bool PP = false; CALL; } } break;                  \
            case 32: { constexpr int GG = 32; if (PG_) { constexpr bool PP = true; CALL; } \
                       else { constexpr bool PP = false; CALL; } } break;                  \
            default: throw Error(-1, "unsupported lanes per row");                         \
        }                                                                                   \
    } while (0)

#define DMRI_DISPATCH_GP(G_, PG_, CALL)                                                     \
    do {                                                                                    \
        switch (G_) {                                                                       \
            case 1: { constexpr int GG = 1; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            case 2: { constexpr int GG = 2; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            case 4: { constexpr int GG = 4; if (PG_) { constexpr bool PP = true; CALL; }   \
                      else { constexpr bool PP = false; CALL; } } break;                   \
            default: throw Error(-1, "unsupported lanes per row for two signals per lane"); \
        }                                                                                   \
    } while (0)

void fit_spmv_grid(SolveGeom &g, bool per_gdir) {
    const void *fn = nullptr;
    size_t lds = 0;
    if (g.mapping == kMapPair) {
        // the least occupancy of the two t-kernels with a full tile, as for mapping 2
        lds = stage_lds_bytes(stage_rows(g, per_gdir), per_gdir);
        const void *fpre = nullptr;
        DMRI_DISPATCH_GP(g.G, per_gdir, fn = reinterpret_cast<const void *>(&spmv_t_pair_kernel<GG, PP, kOpTP>));
        DMRI_DISPATCH_GP(g.G, per_gdir, fpre = reinterpret_cast<const void *>(&spmv_t_pair_kernel<GG, PP, kOpPre>));
        int dev = 0, cus = 0, nb = 0, nb_pre = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, g.spmv_threads, lds) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_pre, fpre, g.spmv_threads, lds) != hipSuccess)
            throw Error(-2, "occupancy query of the SpMV kernels failed");
        const int64_t cap = (int64_t)std::min(nb, nb_pre) * cus / 8 * 8;
        if (cap >= 8 && cap < g.grid) g.grid = (int)cap;
        return;
    }
    if (g.mapping == kMapStaged || g.mapping == kMapSL) {
        // the least occupancy of the t-kernels (operator and preconditioner) with a full tile
        const int TR = g.mapping == kMapStaged ? stage_rows(g, per_gdir) : 0;
        lds = TR > 0 ? stage_lds_bytes(TR, per_gdir) : 0;
        DMRI_DISPATCH_G(g.G, per_gdir, fn = reinterpret_cast<const void *>(&spmv_t_kernel<GG, PP, kOpTP>));
        const void *fpre = nullptr;
        DMRI_DISPATCH_G(g.G, per_gdir, fpre = reinterpret_cast<const void *>(&spmv_t_kernel<GG, PP, kOpPre>));
        int nb_pre = 0;  // with the full tile (the preconditioner may take the imaginary part too)
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_pre, fpre, g.spmv_threads, lds) != hipSuccess)
            throw Error(-2, "occupancy query of the SpMV kernels failed");
        int dev = 0, cus = 0, nb = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, g.spmv_threads, lds) != hipSuccess)
            throw Error(-2, "occupancy query of the SpMV kernels failed");
        const int64_t cap = (int64_t)std::min(nb, nb_pre) * cus / 8 * 8;
        if (cap >= 8 && cap < g.grid) g.grid = (int)cap;
        return;
    } else if (g.mapping == kMapWave) {
        fn = per_gdir ? reinterpret_cast<const void *>(&spmv_t_wave_kernel<true, kOpPre>)
                      : reinterpret_cast<const void *>(&spmv_t_wave_kernel<false, kOpPre>);
    }
    if (!fn) return;
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, g.spmv_threads, lds) != hipSuccess)
        throw Error(-2, "occupancy query of the SpMV kernels failed");
    const int64_t cap = (int64_t)nb * cus / 8 * 8;
    if (cap >= 8 && cap < g.grid) g.grid = (int)cap;
}

void launch_rhs_residual(const SolveGeom &g, const SolveArgs &a, int mode, const int *mask, hipStream_t s) {
    if (mode == 4 && (g.mapping == kMapStaged || g.mapping == kMapSL)) {
        const int TR = g.mapping == kMapStaged ? stage_rows(g, a.per_gdir) : 0;
        const size_t lds = TR > 0 ? stage_lds_bytes_bre(TR, a.per_gdir) : 0;
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        (rhs_history_kernel<GG, PP><<<g.grid, g.spmv_threads, lds, s>>>(a, g.S, g.rpb, TR)));
    } else if (mode == 1 && mask == nullptr && (g.mapping == kMapStaged || g.mapping == kMapSL)) {
        // the plain residual takes the SpMV kernels' (staged) rows; same grid, same partials
        const int TR = g.mapping == kMapStaged ? stage_rows(g, a.per_gdir) : 0;
        const size_t lds = TR > 0 ? stage_lds_bytes(TR, a.per_gdir) : 0;
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        (residual_kernel<GG, PP><<<g.grid, g.spmv_threads, lds, s>>>(a, g.S, g.rpb, TR)));
    } else {
        DMRI_DISPATCH_G(g.G_rhs, a.per_gdir,
                        (rhs_residual_kernel<GG, PP><<<g.grid, g.rhs_threads, 0, s>>>(a, g.S, g.rhs_rpb, mode, mask)));
    }
    check_launch();
    if (a.totU && mode < 2) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partU, g.grid, g.S, a.totU);
        check_launch();
    }
}
void launch_bicg_p(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    hipExtLaunchKernelGGL(bicg_p_kernel, dim3(g.vgrid), dim3(g.vec_threads), 0, s, e0, e1, 0, a, g.S, g.vec_rpb,
                          g.grid, g.vgrid, it);
    check_launch();
}
// The SpMV kernels go through hipExtLaunchKernelGGL so that e0/e1 (when not null) take the start
// and end timestamps of the kernel's own dispatch: no marker packet, hence no extra gap.
// dynamic LDS of a tile kernel: the staged values of up to ucap distinct columns
static void tile_lds(const SolveGeom &g, int &ucap, uint32_t &bytes) {
    ucap = kTileLdsBytes / (16 * g.S);
    if (g.tile_ucap > 0 && g.tile_ucap < ucap) ucap = g.tile_ucap;
    bytes = (uint32_t)(ucap * 16 * g.S);
}
template <int OP>
static void spmv_v_dispatch(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1) {
    if constexpr (OP == kOpVDot || OP == kOpPreDot || OP == kOpPreCDDot || OP == kOpPreDot32 || OP == kOpPreDotR ||
                  OP == kOpQFold) {
        if (g.mapping != kMapSL && g.mapping != kMapStaged) throw Error(-1, "COCG needs SpMV mapping 1 or 2");
        const int TR = g.mapping == kMapStaged ? stage_rows(g, a.per_gdir) : 0;
        uint32_t lds = 0;
        if (TR > 0 && OP == kOpPreDot32)
            lds = (uint32_t)stage_lds_bytes_f32(TR, a.areal ? 1 : (a.per_gdir ? 4 : 2));
        else if (TR > 0)
            lds = (uint32_t)((OP == kOpPreDot || OP == kOpPreDotR) && a.areal ? stage_lds_bytes_ro(TR)
                                                                                : stage_lds_bytes(TR, a.per_gdir));
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_v_kernel<GG, PP, OP>), dim3(g.grid), dim3(g.spmv_threads), lds, s,
                                              e0, e1, 0, a, g.S, g.rpb, TR, it));
    } else if (g.mapping == kMapPair) {
        const int TR = stage_rows(g, a.per_gdir);
        const uint32_t lds =
            (uint32_t)(OP == kOpPre && a.areal ? stage_lds_bytes_ro(TR) : stage_lds_bytes(TR, a.per_gdir));
        DMRI_DISPATCH_GP(g.G, a.per_gdir,
                         hipExtLaunchKernelGGL((spmv_v_pair_kernel<GG, PP, OP>), dim3(g.grid), dim3(g.spmv_threads),
                                               lds, s, e0, e1, 0, a, g.S, g.pair_sp, TR, it));
    } else if (g.mapping == kMapWave) {
        if constexpr (OP == kOpPreCD) throw Error(-1, "the complex-diagonal preconditioner needs SpMV mapping 1 or 2");
        else if (a.per_gdir)
            hipExtLaunchKernelGGL((spmv_v_wave_kernel<true, OP>), dim3(g.grid), dim3(g.spmv_threads), 0, s, e0, e1,
                                  0, a, a.rowptr, a.col, a.vals, g.S, it);
        else
            hipExtLaunchKernelGGL((spmv_v_wave_kernel<false, OP>), dim3(g.grid), dim3(g.spmv_threads), 0, s, e0, e1,
                                  0, a, a.rowptr, a.col, a.vals, g.S, it);
    } else {
        const int TR = g.mapping == kMapStaged ? stage_rows(g, a.per_gdir) : 0;
        const uint32_t lds =
            TR > 0 ? (uint32_t)(OP == kOpPre && a.areal ? stage_lds_bytes_ro(TR) : stage_lds_bytes(TR, a.per_gdir))
                   : 0u;
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_v_kernel<GG, PP, OP>), dim3(g.grid), dim3(g.spmv_threads), lds, s,
                                              e0, e1, 0, a, g.S, g.rpb, TR, it));
    }
}
template <int OP>
static void spmv_t_dispatch(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1) {
    if (g.mapping == kMapPair) {
        const int TR = stage_rows(g, a.per_gdir);
        const uint32_t lds =
            (uint32_t)(OP == kOpPre && a.areal ? stage_lds_bytes_ro(TR) : stage_lds_bytes(TR, a.per_gdir));
        DMRI_DISPATCH_GP(g.G, a.per_gdir,
                         hipExtLaunchKernelGGL((spmv_t_pair_kernel<GG, PP, OP>), dim3(g.grid), dim3(g.spmv_threads),
                                               lds, s, e0, e1, 0, a, g.S, g.pair_sp, TR, (int)g.grid, it));
    } else if (g.mapping == kMapWave) {
        if constexpr (OP == kOpPreCD) throw Error(-1, "the complex-diagonal preconditioner needs SpMV mapping 1 or 2");
        else if (a.per_gdir)
            hipExtLaunchKernelGGL((spmv_t_wave_kernel<true, OP>), dim3(g.grid), dim3(g.spmv_threads), 0, s, e0, e1,
                                  0, a, a.rowptr, a.col, a.vals, g.S, (int)g.grid, it);
        else
            hipExtLaunchKernelGGL((spmv_t_wave_kernel<false, OP>), dim3(g.grid), dim3(g.spmv_threads), 0, s, e0, e1,
                                  0, a, a.rowptr, a.col, a.vals, g.S, (int)g.grid, it);
    } else {
        const int TR = g.mapping == kMapStaged ? stage_rows(g, a.per_gdir) : 0;
        const uint32_t lds =
            TR > 0 ? (uint32_t)(OP == kOpPre && a.areal ? stage_lds_bytes_ro(TR) : stage_lds_bytes(TR, a.per_gdir))
                   : 0u;
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_t_kernel<GG, PP, OP>), dim3(g.grid), dim3(g.spmv_threads), lds, s,
                                              e0, e1, 0, a, g.S, g.rpb, TR, (int)g.grid, it));
    }
}

void launch_spmv_v(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                   int op) {
    if (op != kOpV && (g.mapping == kMapSR || g.mapping == kMapTile))
        throw Error(-1, "the polynomial preconditioner needs SpMV mapping 1, 2 or 4");
    if (g.mapping == kMapTile) {
        int ucap;
        uint32_t lds;
        tile_lds(g, ucap, lds);
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_v_tile_kernel<GG, PP>), dim3(g.grid), dim3(g.spmv_threads), lds,
                                              s, e0, e1, 0, a, g.S, g.tile_rows, ucap, it));
    } else if (g.mapping == kMapSR) {
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_v_sr_kernel<GG, PP>), dim3(g.grid), dim3(g.sr_threads), 0, s, e0,
                                              e1, 0, a, g.S, g.sr_rpb, it));
    } else if (op == kOpPre) {
        if (a.pdiag && g.mapping != kMapWave)
            spmv_v_dispatch<kOpPreCD>(g, a, it, s, e0, e1);
        else
            spmv_v_dispatch<kOpPre>(g, a, it, s, e0, e1);
    } else {
        spmv_v_dispatch<kOpV>(g, a, it, s, e0, e1);
    }
    check_launch();
    if (a.totV && op == kOpV) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partV, g.grid, g.S, a.totV);
        check_launch();
    }
}
void launch_spmv_t(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                   int op) {
    if (op != kOpT && (g.mapping == kMapSR || g.mapping == kMapTile))
        throw Error(-1, "the polynomial preconditioner needs SpMV mapping 1, 2 or 4");
    if (op != kOpT && !a.s) throw Error(-1, "the preconditioned t-kernels read s from the s-pass (a.s is null)");
    if (g.mapping == kMapTile) {
        int ucap;
        uint32_t lds;
        tile_lds(g, ucap, lds);
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_t_tile_kernel<GG, PP>), dim3(g.grid), dim3(g.spmv_threads), lds,
                                              s, e0, e1, 0, a, g.S, g.tile_rows, ucap, (int)g.grid, it));
    } else if (g.mapping == kMapSR) {
        DMRI_DISPATCH_G(g.G, a.per_gdir,
                        hipExtLaunchKernelGGL((spmv_t_sr_kernel<GG, PP>), dim3(g.grid), dim3(g.sr_threads), 0, s, e0,
                                              e1, 0, a, g.S, g.sr_rpb, (int)g.grid, it));
    } else if (op == kOpPre) {
        if (a.pdiag && g.mapping != kMapWave)
            spmv_t_dispatch<kOpPreCD>(g, a, it, s, e0, e1);
        else
            spmv_t_dispatch<kOpPre>(g, a, it, s, e0, e1);
    } else if (op == kOpTP) {
        spmv_t_dispatch<kOpTP>(g, a, it, s, e0, e1);
    } else {
        spmv_t_dispatch<kOpT>(g, a, it, s, e0, e1);
    }
    check_launch();
    if (a.totT && op != kOpPre) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partT, g.grid, g.S, a.totT);
        check_launch();
    }
}
void launch_spass(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    hipExtLaunchKernelGGL(spass_kernel, dim3(g.vgrid), dim3(g.vec_threads), 0, s, e0, e1, 0, a, g.S, g.vec_rpb,
                          g.grid, it);
    check_launch();
}

void launch_bicg_update(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
    hipExtLaunchKernelGGL(bicg_update_kernel, dim3(g.vgrid), dim3(g.vec_threads), 0, s, e0, e1, 0, a, g.S,
                          g.vec_rpb, g.grid, it);
    check_launch();
    if (a.totU) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partU, g.vgrid, g.S, a.totU);
        check_launch();
    }
}
void launch_cocg_spmv(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                      bool pre) {
    if (pre) {
        if (a.pdiag) spmv_v_dispatch<kOpPreCDDot>(g, a, it, s, e0, e1);
        else if (a.r32) spmv_v_dispatch<kOpPreDot32>(g, a, it, s, e0, e1);
        else spmv_v_dispatch<kOpPreDot>(g, a, it, s, e0, e1);
    } else {
        spmv_v_dispatch<kOpVDot>(g, a, it, s, e0, e1);
    }
    check_launch();
    double *part = pre ? a.partV : a.partT, *tot = pre ? a.totV : a.totT;
    if (tot) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(part, g.grid, g.S, tot);
        check_launch();
    }
}
void launch_cocg_p(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                   bool rfold) {
    hipExtLaunchKernelGGL(cocg_p_kernel, dim3(g.vgrid), dim3(g.vec_threads), 0, s, e0, e1, 0, a, g.S, g.vec_rpb,
                          g.grid, g.vgrid, g.grid, it, rfold ? 1 : 0);
    check_launch();
}
void launch_cocg_qfold(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (!a.per_gdir || a.pdiag) throw Error(-1, "the folded COCG p-step takes per-signal directions");
    spmv_v_dispatch<kOpQFold>(g, a, it, s, e0, e1);
    check_launch();
    if (a.totT) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partT, g.grid, g.S, a.totT);
        check_launch();
    }
}
void launch_cocg_pre_update(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0,
                            hipEvent_t e1) {
    if (a.per_gdir || a.pdiag || a.r32) throw Error(-1, "the folded COCG update takes one gradient direction");
    spmv_v_dispatch<kOpPreDotR>(g, a, it, s, e0, e1);
    check_launch();
    if (a.totV) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partV, g.grid, g.S, a.totV);
        check_launch();
    }
}
void launch_cocg_update(const SolveGeom &g, const SolveArgs &a, int it, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
    hipExtLaunchKernelGGL(cocg_update_kernel, dim3(g.vgrid), dim3(g.vec_threads), 0, s, e0, e1, 0, a, g.S,
                          g.vec_rpb, g.grid, it);
    check_launch();
    if (a.totU) {
        finalize_kernel<<<g.S, kThreads, 0, s>>>(a.partU, g.vgrid, g.S, a.totU);
        check_launch();
    }
}
// areal[e] = vals[vw e]: the scaled real part alone, for the preconditioner SpMVs
__global__ void __launch_bounds__(256) extract_real_kernel(int64_t nnz, const double *vals, int vw, double *areal) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        areal[e] = vals[vw * e];
}
__global__ void __launch_bounds__(256) to_f32_kernel(int64_t n, const double *src, float *dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}
void launch_to_f32(const double *src, int64_t n, float *dst, hipStream_t s) {
    if (n <= 0) return;
    unsigned g = grid_for(n, 256);
    if (g > 4096u) g = 4096u;
    to_f32_kernel<<<g, 256, 0, s>>>(n, src, dst);
    check_launch();
}
void launch_extract_real(int64_t nnz, const double *vals, int vw, double *areal, hipStream_t s) {
    unsigned g = grid_for(nnz, 256);
    if (g > 4096u) g = 4096u;
    extract_real_kernel<<<g, 256, 0, s>>>(nnz, vals, vw, areal);
    check_launch();
}

// Zero the Krylov states of a step and the all-done flag (one dispatch instead of two fills)
__global__ void __launch_bounds__(256) reset_state_kernel(int *st, int nwords, int *flag) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) st[i] = 0;
    if (threadIdx.x == 0) flag[0] = 0;
}
void launch_reset_state(KrylovState *st, int S, int *flag, hipStream_t s) {
    const int nwords = (int)(S * sizeof(KrylovState) / sizeof(int));
    reset_state_kernel<<<1, 256, 0, s>>>(reinterpret_cast<int *>(st), nwords, flag);
    check_launch();
}

// Krylov states and the all-done flag into pinned host memory: one dispatch at every convergence
// check instead of two device-to-host copies
__global__ void __launch_bounds__(256) publish_state_kernel(const int *st, int nwords, const int *flag, int *hst,
                                                            int *hflag) {
    for (int i = threadIdx.x; i < nwords; i += blockDim.x) hst[i] = st[i];
    if (threadIdx.x == 0) hflag[0] = flag[0];
}
void launch_publish_state(const KrylovState *st, int S, const int *flag, KrylovState *hst, int *hflag, hipStream_t s) {
    const int nwords = (int)(S * sizeof(KrylovState) / sizeof(int));
    publish_state_kernel<<<1, 256, 0, s>>>(reinterpret_cast<const int *>(st), nwords, flag,
                                           reinterpret_cast<int *>(hst), hflag);
    check_launch();
}

int proj_grid(const SolveGeom &g) { return g.pgrid; }

void launch_proj_guess(const SolveGeom &g, const ProjArgs &p, int mh, hipStream_t s) {
    if (p.nh < 1 || p.nh > mh || !(mh == 4 || mh == 6 || mh == 8 || mh == 10 || mh == 12))
        throw Error(-1, "projection guess: history size must satisfy 1 <= nh <= mh, mh in {4, 6, 8, 10, 12}");
    if (p.S < 1 || p.S > kMaxSignals) throw Error(-1, "projection guess: signals per launch must lie in [1, 64]");
    const int nv = proj_nv(mh);
    const size_t lds = (size_t)(mh + 1) * kThreads * sizeof(double2);
    // Gram blocks: at most what stays resident at once (its LDS tile bounds it), so no block
    // waits for a second round
    int grid = proj_grid(g);
    {
        const void *fn = mh == 4    ? reinterpret_cast<const void *>(&proj_gram_kernel<4>)
                         : mh == 6  ? reinterpret_cast<const void *>(&proj_gram_kernel<6>)
                         : mh == 8  ? reinterpret_cast<const void *>(&proj_gram_kernel<8>)
                         : mh == 10 ? reinterpret_cast<const void *>(&proj_gram_kernel<10>)
                                    : reinterpret_cast<const void *>(&proj_gram_kernel<12>);
        int dev = 0, cus = 0, nb = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kThreads, lds) != hipSuccess)
            throw Error(-2, "occupancy query of the projection Gram kernel failed");
        const int nwin = (p.S + std::max(1, std::min(p.S, 2 * kThreads / (nv / 2))) - 1) /
                         std::max(1, std::min(p.S, 2 * kThreads / (nv / 2)));
        const int cap = nb * cus / nwin;
        if (cap >= 1 && cap < grid) grid = cap;
    }
    // signal window of the Gram kernel: at most 2 (pair, signal) combinations per thread
    const int Sw = std::max(1, std::min(p.S, 2 * kThreads / (nv / 2)));
    const dim3 ggrid(grid, (p.S + Sw - 1) / Sw);
    // the combination streams the history once: one element per thread over a covering grid
    const int64_t cblocks = (p.n * p.S + (int64_t)g.vec_rpb * p.S - 1) / ((int64_t)g.vec_rpb * p.S);
    const unsigned cgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(cblocks, 1 << 30));
#define DMRI_PROJ(M)                                                                   \
    proj_gram_kernel<M><<<ggrid, kThreads, lds, s>>>(p, Sw);                            \
    check_launch();                                                                    \
    proj_finalize_kernel<<<p.S * nv, kThreads, 0, s>>>(p.part, grid, p.S, nv, p.tot); \
    check_launch();                                                                    \
    proj_solve_kernel<M><<<p.S, 64, 0, s>>>(p);                                        \
    check_launch();                                                                    \
    proj_combine_kernel<M><<<cgrid, g.vec_threads, 0, s>>>(p, g.vec_rpb);              \
    check_launch();
    if (mh == 4) {
        DMRI_PROJ(4)
    } else if (mh == 6) {
        DMRI_PROJ(6)
    } else if (mh == 8) {
        DMRI_PROJ(8)
    } else if (mh == 10) {
        DMRI_PROJ(10)
    } else {
        DMRI_PROJ(12)
    }
#undef DMRI_PROJ
}

void launch_signal(int64_t n, int S, const double *w, const double *sc, const double2 *x, double *part, double *out,
                   hipStream_t s) {
    const int vrpb = std::max(1, kThreads / S);
    int64_t need = (n + vrpb - 1) / vrpb;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(need, kMaxGrid));
    signal_part_kernel<<<grid, vrpb * S, 0, s>>>(n, S, vrpb, w, sc, x, part);
    check_launch();
    finalize_kernel<<<S, kThreads, 0, s>>>(part, grid, S, out);
    check_launch();
}

void launch_gm_dots(const SolveGeom &g, const GmArgs &a, int i0, int cnt, hipStream_t s) {
    gm_dots_kernel<<<g.vgrid, g.vec_threads, 0, s>>>(a, i0, cnt, g.vec_rpb);
    check_launch();
}
void launch_gm_fin(const double *parts, int nparts, int S, int pstride, int nv, double *out, int ostride, int o,
                   hipStream_t s) {
    gm_fin_kernel<<<S, kThreads, 0, s>>>(parts, nparts, S, pstride, nv, out, ostride, o);
    check_launch();
}
void launch_gm_orth(const SolveGeom &g, const GmArgs &a, int j, bool norm, hipStream_t s) {
    gm_orth_kernel<<<g.vgrid, g.vec_threads, 0, s>>>(a, j, norm ? 1 : 0, g.vec_rpb);
    check_launch();
}
void launch_gm_hess(const GmArgs &a, int j, const double *wn2, hipStream_t s) {
    gm_hess_kernel<<<1, 64, 0, s>>>(a, j, wn2);
    check_launch();
}
void launch_gm_scale(const SolveGeom &g, const GmArgs &a, int j, hipStream_t s) {
    gm_scale_kernel<<<g.vgrid, g.vec_threads, 0, s>>>(a, j, g.vec_rpb);
    check_launch();
}
void launch_gm_start(const GmArgs &a, const double *tot, bool first, hipStream_t s) {
    gm_start_kernel<<<1, 64, 0, s>>>(a, tot, first ? 1 : 0);
    check_launch();
}
void launch_gm_v0(const SolveGeom &g, const GmArgs &a, const double2 *r, hipStream_t s) {
    gm_v0_kernel<<<g.vgrid, g.vec_threads, 0, s>>>(a, r, g.vec_rpb);
    check_launch();
}
void launch_gm_combine(const SolveGeom &g, const GmArgs &a, double2 *out, hipStream_t s) {
    gm_solve_kernel<<<1, 64, 0, s>>>(a);
    check_launch();
    gm_combine_kernel<<<g.vgrid, g.vec_threads, 0, s>>>(a, out, g.vec_rpb);
    check_launch();
}
void launch_reset_state_flags(KrylovState *st, int S, int *flag, hipStream_t s) {
    reopen_kernel<<<1, 64, 0, s>>>(st, S, flag);
    check_launch();
}

void launch_gm_axpy(int64_t total, const double2 *z, double2 *x, hipStream_t s) {
    unsigned gr = grid_for(total, kThreads);
    if (gr > (unsigned)kMaxGrid) gr = kMaxGrid;
    if (gr == 0) gr = 1;
    gm_axpy_kernel<<<gr, kThreads, 0, s>>>(total, z, x);
    check_launch();
}


}  // namespace dmri
