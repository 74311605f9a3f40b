// nsfv1_export.hip -- the disc's data exports at points of every element
// (navier_stokes_fv1.cpp:894-1491): velocity, velocity gradient, pressure,
// pressure gradient, and the ip velocity away from SCVF ips.
//
// One lane per (element, point).  The points are the element's SCVF ips
// (FV1 local ip = average of the SCVF's local corners), its SCV ips (the
// corners) or caller-given local points.  Values are P1/Q1 interpolations of
// the corner unknowns; gradients use the element map's inverse Jacobian at
// the point, except the velocity gradient away from SCVF ips, which the
// reference returns in reference-element coordinates (:1072-1080).  The
// stabilised ip velocity at SCVF ips is computed by the assembly's SCVF code
// (launch_export_stab_vel).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_device.hpp"
#include "nsfv1_export.hpp"
#include "nsfv1_topo.hpp"

namespace nsfv1 {

template <int ET>
__global__ void __launch_bounds__(256)
export_points(ExportArgs A, const double* __restrict__ u, double* __restrict__ out, int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.ne * A.npt) return;
    const int64_t e = t / A.npt;
    const int q = (int)(t - e * A.npt);
    double co[N][D], ul[N][NC];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int64_t w = A.conn[e * N + k];
#pragma unroll
        for (int i = 0; i < D; ++i) co[k][i] = A.xyz[w * D + i];
#pragma unroll
        for (int a = 0; a < NC; ++a) ul[k][a] = u[w * NC + a];
    }
    double x[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < D; ++i)
        x[i] = A.at == EX_AT_SCVF ? local_ip<ET>(q, i) : (A.at == EX_AT_SCV ? T::lc[q][i] : A.lips[q][i]);
    double phi[N], gl[N][D];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        phi[k] = shape_at<ET>(k, -1, x[0], x[1], x[2]);
#pragma unroll
        for (int j = 0; j < D; ++j) gl[k][j] = shape_at<ET>(k, j, x[0], x[1], x[2]);
    }
    double* o = out + t * A.ncomp;
    const int w = A.what;
    if (w == EX_PRESSURE && A.at == EX_AT_SCV) {  // u(P, co): :1342-1364
        o[0] = ul[q][D];
        return;
    }
    if (w == EX_VELOCITY || w == EX_VELOCITY_IP || w == EX_PRESSURE) {
        const int c0 = w == EX_PRESSURE ? D : 0, nc = w == EX_PRESSURE ? 1 : D;
        for (int c = 0; c < nc; ++c) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) s += ul[k][c0 + c] * phi[k];
            o[c] = s;
        }
        return;
    }
    if (w == EX_VELOCITY_GRAD && A.at != EX_AT_SCVF) {
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
            for (int d2 = 0; d2 < D; ++d2) {
                double s = 0;
#pragma unroll
                for (int k = 0; k < N; ++k) s += ul[k][d1] * gl[k][d2];
                o[d1 * D + d2] = s;
            }
        return;
    }
    // global gradients: g_k = J^-T gl_k with J_ij = dx_i / dxi_j
    double J[D][D], Ji[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) s += co[k][i] * gl[k][j];
            J[i][j] = s;
        }
    if constexpr (D == 2) {
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        if (det == 0.0) { kernel_error(err, KERR_DEGENERATE); return; }
        Ji[0][0] = J[1][1] / det;  Ji[0][1] = -J[0][1] / det;
        Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
    } else {
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (det == 0.0) { kernel_error(err, KERR_DEGENERATE); return; }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int a0 = (j + 1) % 3, a1 = (j + 2) % 3, b0 = (i + 1) % 3, b1 = (i + 2) % 3;
                Ji[i][j] = (J[a0][b0] * J[a1][b1] - J[a0][b1] * J[a1][b0]) / det;
            }
    }
    double g[N][D];
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < D; ++j) s += Ji[j][i] * gl[k][j];
            g[k][i] = s;
        }
    if (w == EX_VELOCITY_GRAD) {
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
            for (int d2 = 0; d2 < D; ++d2) {
                double s = 0;
#pragma unroll
                for (int k = 0; k < N; ++k) s += ul[k][d1] * g[k][d2];
                o[d1 * D + d2] = s;
            }
    } else {  // EX_PRESSURE_GRAD
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) s += ul[k][D] * g[k][d];
            o[d] = s;
        }
    }
}

hipError_t launch_export_points(int et, const ExportArgs& a, const double* u, double* out, int* err, hipStream_t s)
{
    const int64_t lanes = a.ne * a.npt;
    if (lanes == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((lanes + 255) / 256);
    switch (et) {
    case 0: export_points<0><<<blocks, 256, 0, s>>>(a, u, out, err); break;
    case 1: export_points<1><<<blocks, 256, 0, s>>>(a, u, out, err); break;
    case 2: export_points<2><<<blocks, 256, 0, s>>>(a, u, out, err); break;
    case 3: export_points<3><<<blocks, 256, 0, s>>>(a, u, out, err); break;
    case 4: export_points<4><<<blocks, 256, 0, s>>>(a, u, out, err); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace nsfv1
