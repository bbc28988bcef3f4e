// Internal declarations shared by the libdmri host code and its gfx950 kernels.
#pragma once

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

// Device state of one Krylov solve, written by the last block of each reducing kernel
// and read by the kernels that follow (kernel boundaries order the accesses).
struct KrylovState {
    double2 rho, rho_old, alpha, omega, beta;
    double bnorm2, rnorm2, snorm2, tol2;
    int iter, done, breakdown, maxit;
    double ts_re, ts_im, tt, pad0;
    unsigned int counter[8];
};

constexpr int kBlock = 256;          // threads per block of every solver kernel
constexpr int kRowGroup = 8;         // lanes per matrix row in the CSR SpMV kernels
constexpr int kRowsPerBlock = kBlock / kRowGroup;
constexpr int kVecBlocks = 1024;     // grid of the vector-update kernels (grid-stride)
constexpr int kMaxPartials = 1 << 20;

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
void launch_gather_cells(const int *perm, int64_t ne, const int *tets, const int *phase, int *ct, int *cph,
                         hipStream_t s);
void launch_assemble_cells(int64_t start, int64_t count, int64_t ne, const int *perm, const int *ct, const int *cph,
                           const double *xyz, int64_t nv, const int *dof_of, const int *rowptr, const int *col,
                           const double *Dc, const double *Dcell, const double *iT2c, const double *iT2cell,
                           double *m, double *k, double *r, double *jx, double *jy, double *jz, int *miss,
                           hipStream_t s);
void launch_assemble_facets(int64_t start, int64_t count, const int *perm, const int *iface_v, const double *xyz,
                            int64_t nv, const int *dof_of, const int *rowptr, const int *col, double kappa,
                            const double *kappa_f, double *gi, int *miss, hipStream_t s);
void launch_diag_and_weights(int64_t n, const int *rowptr, const int *col, const double *m, int *diag, double *w,
                             int *missing_diag, hipStream_t s);
void launch_sum(const double *x, int64_t n, double *part, unsigned *counter, double *out, hipStream_t s);

// ---------------------------------------------------------------- solve launchers
void launch_prep_scale(int64_t n, const int *diag, const double *m, const double *k, const double *r,
                       const double *gi, double dt, double theta, double *sc, hipStream_t s);
void launch_prep_values(int64_t n, const int *rowptr, const int *col, const double *sc, const double *m,
                        const double *k, const double *r, const double *gi, const double *jx, const double *jy,
                        const double *jz, double gx, double gy, double gz, double dt, double theta, double2 *aj,
                        double *bre, hipStream_t s);
void launch_init_x(int64_t n, const double *sc, double2 *x, hipStream_t s);
void launch_rhs_residual(int64_t n, const int *rowptr, const int *col, const double2 *aj, const double *bre,
                         double sA, double sB, const double2 *x, double2 *r, double2 *rh, double rtol, double atol,
                         int maxit, double *part, KrylovState *st, hipStream_t s);
void launch_bicg_p(int64_t n, const double2 *r, double2 *p, const double2 *v, const KrylovState *st,
                   hipStream_t s);
void launch_spmv_v(int64_t n, const int *rowptr, const int *col, const double2 *aj, double sA, const double2 *p,
                   double2 *v, const double2 *rh, double *part, KrylovState *st, hipStream_t s);
void launch_spmv_t(int64_t n, const int *rowptr, const int *col, const double2 *aj, double sA, const double2 *r,
                   const double2 *v, double2 *t, double *part, KrylovState *st, hipStream_t s);
void launch_bicg_update(int64_t n, double2 *x, const double2 *p, double2 *r, const double2 *v, const double2 *t,
                        const double2 *rh, double *part, KrylovState *st, hipStream_t s);
void launch_signal(int64_t n, const double *w, const double *sc, const double2 *x, double *part,
                   unsigned *counter, double *out, hipStream_t s);
void launch_unscale(int64_t n, const double *sc, const double2 *x, double2 *u, hipStream_t s);

}  // namespace dmri
