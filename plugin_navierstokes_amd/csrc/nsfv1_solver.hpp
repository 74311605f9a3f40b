// nsfv1_solver.hpp -- launchers of the device linear algebra used by the
// Newton loop (Dirichlet rows, SpMV, point-block Jacobi, GMRES BLAS-1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

constexpr int GMRES_MAX_RESTART = 60;
constexpr int RED_PARTIALS = 1024 * 64;

hipError_t launch_dirichlet(int jprec, const int64_t* rowptr, const int32_t* colidx, const int64_t* rows,
                            int64_t nrows, void* J, double* d, const double* u, const double* g, hipStream_t s);
hipError_t launch_spmv(int jprec, const int64_t* rowptr, const int32_t* colidx, const void* A, const double* x,
                       double* y, int64_t n, hipStream_t s);
hipError_t launch_block_jacobi_setup(int jprec, int nc, const int64_t* rowptr, const int32_t* colidx, const void* A,
                                     int64_t nv, double* Binv, int* err, hipStream_t s);
hipError_t launch_block_jacobi_apply(int nc, const double* Binv, const double* x, double* y, int64_t nv, hipStream_t s);
hipError_t launch_multi_dot(const double* V, int64_t ld, int k, const double* w, int64_t n, double* partial,
                            double* out, hipStream_t s);
hipError_t launch_multi_axpy(const double* V, int64_t ld, int k, const double* h, double* w, int64_t n, hipStream_t s);
hipError_t launch_multi_add(const double* V, int64_t ld, int k, const double* c, double* x, int64_t n, hipStream_t s);
hipError_t launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);

}  // namespace nsfv1
