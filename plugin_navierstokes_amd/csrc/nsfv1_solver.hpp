// nsfv1_solver.hpp -- launchers of the device linear algebra used by the
// Newton loop (Dirichlet rows, SpMV, point-block Jacobi, GMRES BLAS-1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

constexpr int GMRES_MAX_RESTART = 60;
constexpr int RED_PARTIALS = 1024 * 64;
constexpr int MG_MAX_LEVELS = 12;
constexpr int MG_DENSE_MAX = 1200;  // largest coarsest-level system inverted densely

hipError_t launch_dirichlet(int jprec, const int64_t* rowptr, const int32_t* colidx, const int64_t* rows,
                            int64_t nrows, void* J, double* d, const double* u, const double* g, hipStream_t s);
// adj: node-block neighbour lists (nullptr: plain CSR column indices); nc = dim+1;
// n rows, nnz entries
hipError_t launch_spmv(int jprec, const int64_t* rowptr, const int32_t* colidx, const int32_t* adj, int nc,
                       const void* A, const double* x, double* y, int64_t n, int64_t nnz, hipStream_t s);
hipError_t launch_block_jacobi_setup(int jprec, int nc, const int64_t* rowptr, const int32_t* colidx, const void* A,
                                     int64_t nv, double* Binv, int* err, hipStream_t s);
hipError_t launch_block_jacobi_apply(int nc, const double* Binv, const double* x, double* y, int64_t nv, hipStream_t s);
hipError_t launch_multi_dot(const double* V, int64_t ld, int k, const double* w, int64_t n, double* partial,
                            double* out, hipStream_t s);
hipError_t launch_multi_axpy(const double* V, int64_t ld, int k, const double* h, double* w, int64_t n, hipStream_t s);
hipError_t launch_multi_add(const double* V, int64_t ld, int k, const double* c, double* x, int64_t n, hipStream_t s);
hipError_t launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);
// device-side pieces of the fixed-iteration GMRES of the capturable Newton
// step (no host synchronisation): x = v; beta = sqrt(*n2) (and *res2 = *n2);
// Hessenberg column j (m <= hs <= ...); y = x / *h (0 if *h = 0); Givens
// least squares of the (m+1) x m Hessenberg H with right-hand side beta e1
hipError_t launch_fill(double* x, double v, int64_t n, hipStream_t s);
hipError_t launch_zero_words(void* x, int64_t nwords, hipStream_t s);
hipError_t launch_gm_beta(const double* n2, double* beta, double* res2, hipStream_t s);
hipError_t launch_gm_hcol(const double* hd, int hs, double* H, int m, int j, hipStream_t s);
hipError_t launch_gm_scale(const double* x, const double* h, double* y, int64_t n, hipStream_t s);
hipError_t launch_gm_lsq(double* H, const double* beta, int m, double* y, hipStream_t s);
hipError_t launch_residual(int jprec, const int64_t* rowptr, const int32_t* colidx, const int32_t* adj, int nc,
                           const void* A, const double* x, const double* b, double* r, int64_t n, int64_t nnz,
                           hipStream_t s);
// one damped point-block Jacobi sweep y = x + omega Binv (b - A x) (y must not
// alias x; r: scratch of n doubles, used only without node blocks)
hipError_t launch_smooth(int jprec, const int64_t* rowptr, const int32_t* colidx, const int32_t* adj, int nc,
                         const void* A, const double* x, const double* b, const double* Binv, double omega, double* r,
                         double* y, int64_t n, int64_t nnz, hipStream_t s);
hipError_t launch_bj_update(int nc, const double* Binv, const double* r, double* x, double omega, int first,
                            int64_t nv, hipStream_t s);
hipError_t launch_node_transfer(const int64_t* rowptr, const int32_t* colidx, const double* w, const double* x,
                                double* y, int64_t nrows, int nc, int accumulate, hipStream_t s);
hipError_t launch_zero_rows(const int64_t* rows, int64_t nrows, double* y, hipStream_t s);
hipError_t launch_inject(const int64_t* inj, const double* uf, double* uc, int64_t nvc, int nc, hipStream_t s);
hipError_t launch_dense_matvec(const double* A, const double* x, double* y, int n, hipStream_t s);
// inverse of the n x n CSR matrix (n <= MG_DENSE_MAX) into the dense row-major
// inv; *err (device) = 1 when a pivot is zero (singular matrix)
size_t dense_inverse_work_bytes(int n);
hipError_t launch_dense_inverse(int jprec, const int64_t* rowptr, const int32_t* colidx, const void* J, int n,
                                void* work, double* inv, int* err, hipStream_t s);

}  // namespace nsfv1
