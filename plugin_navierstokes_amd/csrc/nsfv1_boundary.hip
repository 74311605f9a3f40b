// nsfv1_boundary.hip -- FV1 boundary element discs added to the rows of the
// nodes on the listed boundary sides: the no-normal-stress outflow condition
// (incompressible/fv1/bnd/no_normal_stress_outflow_fv1.cpp:188-426) and the
// symmetry condition (incompressible/fv1/bnd/symmetric_boundary_fv1.cpp:
// 224-386).
//
// Boundary faces (ugcore FV1Geometry BF, restated): on side s of an element,
// corner c owns the part of the side inside its sub-control volume --
// 2D: [x_c, side midpoint]; 3D: [x_c, mid(c, next), side centre,
// mid(prev, c)].  ip = average of the part's corners, normal = outward area
// vector (2D: the rotated segment; 3D: half the cross product of the
// diagonals), shapes and global gradients at the local ip.
//
// Per boundary face (node c, normal n, ip) the condition adds
//   momentum:   the tangential part of -nu rho (grad u + grad u^T) n, i.e.
//               F - (F.n) n with the area vector n (as :261 does), and the
//               outgoing convective flux max(rho u.n, 0) u;
//   continuity: rho u.n;
// and the matching fixpoint Jacobian (:291-311, 360-368).
//
// The sliding wall (incompressible/fv1/bnd/wall_sliding_fv1.cpp:229-425)
// adds per boundary face, with the unit normal nh: the convective term
// (t.n) rho t of the tangential velocity t (zero up to rounding, kept as
// the reference writes it), -k t_w and -lim t_w/|t_w| where t_w is the
// tangential part of the interpolation over the first dim shapes only
// (:391-397), and the reference's approximate Jacobian (:265-365).
//
// The symmetry condition adds per boundary face
//   momentum d: F[d] n[d] with F = -nu rho (grad u + grad u^T) n (the
//               component-wise product of :373-375) plus p(ip) n[d];
//   continuity: nothing (u.n = 0 there);
// and its Jacobian (:290-317).  It is linear in u.
//
// One lane per boundary node walks that node's boundary faces in a fixed
// order and adds into its own rows of J and d (read-modify-write of rows no
// other lane touches), so the result is deterministic.  Runs after the
// element assembly, on the same stream.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_boundary.hpp"
#include "nsfv1_device.hpp"
#include "nsfv1_topo.hpp"

namespace nsfv1 {

// P1/Q1 (prism: P1 x P1) shapes and reference gradients at a run-time local point
template <int ET>
__device__ inline void shapes_at_point(const double* x, double* phi, double (*gl)[3])
{
    if constexpr (ET == 0) {
        phi[0] = 1.0 - x[0] - x[1]; phi[1] = x[0]; phi[2] = x[1];
        gl[0][0] = -1; gl[0][1] = -1; gl[1][0] = 1; gl[1][1] = 0; gl[2][0] = 0; gl[2][1] = 1;
    } else if constexpr (ET == 1) {
        const double a = x[0], b = x[1];
        phi[0] = (1 - a) * (1 - b); phi[1] = a * (1 - b); phi[2] = a * b; phi[3] = (1 - a) * b;
        gl[0][0] = -(1 - b); gl[0][1] = -(1 - a);
        gl[1][0] = (1 - b);  gl[1][1] = -a;
        gl[2][0] = b;        gl[2][1] = a;
        gl[3][0] = -b;       gl[3][1] = (1 - a);
    } else if constexpr (ET == 2) {
        phi[0] = 1.0 - x[0] - x[1] - x[2]; phi[1] = x[0]; phi[2] = x[1]; phi[3] = x[2];
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < 3; ++j) gl[k][j] = 0.0;
        gl[0][0] = gl[0][1] = gl[0][2] = -1;
        gl[1][0] = 1; gl[2][1] = 1; gl[3][2] = 1;
    } else if constexpr (ET == 4) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            phi[k] = shape_at<4>(k, -1, x[0], x[1], x[2]);
#pragma unroll
            for (int j = 0; j < 3; ++j) gl[k][j] = shape_at<4>(k, j, x[0], x[1], x[2]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double f[3], df[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const bool hi = Topo<3>::lc[k][j] > 0.5;
                f[j] = hi ? x[j] : 1.0 - x[j];
                df[j] = hi ? 1.0 : -1.0;
            }
            phi[k] = f[0] * f[1] * f[2];
            gl[k][0] = df[0] * f[1] * f[2];
            gl[k][1] = f[0] * df[1] * f[2];
            gl[k][2] = f[0] * f[1] * df[2];
        }
    }
}

template <int ET, typename JT>
__global__ void __launch_bounds__(128)
outflow_rows(OutflowArgs A, const double* __restrict__ u, JT* __restrict__ J, double* __restrict__ dout,
             int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nbn) return;
    const int64_t v = A.bn_node[t];
    const double rho = A.rho;
    int64_t rows[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) rows[a] = A.rowptr[v * NC + a];
    double dsum[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) dsum[a] = 0.0;

    for (int it = A.bn_ptr[t]; it < A.bn_ptr[t + 1]; ++it) {
        const double nu = A.nu_item ? A.nu_item[it] : A.nu;  // at this face's ip
        const int q = A.item_inc[it];
        const int s = A.item_side[it];
        const int e = A.inc_elem[q];
        const int kv = A.inc_info[q] & 0xff;
        const uint64_t pw = A.inc_pos[q];
        double co[N][D], ul[N][NC], c[D];
#pragma unroll
        for (int i = 0; i < D; ++i) c[i] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int64_t w = A.conn[(int64_t)e * N + k];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                co[k][i] = A.xyz[w * D + i];
                c[i] += co[k][i] / N;
            }
#pragma unroll
            for (int a = 0; a < NC; ++a) ul[k][a] = u[w * NC + a];
        }
        // the part of side s owned by corner kv
        const int ns = T::side_n[s];
        int j = 0;
        for (int m = 0; m < ns; ++m)
            if (T::side[s][m] == kv) j = m;
        double p[4][3], pl[4][3];
        int np;
        if constexpr (D == 2) {
            const int other = (j == 0) ? T::side[s][1] : T::side[s][0];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                p[0][i] = co[kv][i];
                p[1][i] = 0.5 * (co[kv][i] + co[other][i]);
                pl[0][i] = T::lc[kv][i];
                pl[1][i] = 0.5 * (T::lc[kv][i] + T::lc[other][i]);
            }
            np = 2;
        } else {
            const int nx = T::side[s][(j + 1) % ns], pv = T::side[s][(j + ns - 1) % ns];
            double fc[3] = {0, 0, 0}, fcl[3] = {0, 0, 0};
            for (int m = 0; m < ns; ++m) {
                const int cm = T::side[s][m];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    fc[i] += co[cm][i] / ns;
                    fcl[i] += T::lc[cm][i] / ns;
                }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                p[0][i] = co[kv][i];
                p[1][i] = 0.5 * (co[kv][i] + co[nx][i]);
                p[2][i] = fc[i];
                p[3][i] = 0.5 * (co[pv][i] + co[kv][i]);
                pl[0][i] = T::lc[kv][i];
                pl[1][i] = 0.5 * (T::lc[kv][i] + T::lc[nx][i]);
                pl[2][i] = fcl[i];
                pl[3][i] = 0.5 * (T::lc[pv][i] + T::lc[kv][i]);
            }
            np = 4;
        }
        double ipg[3] = {0, 0, 0}, ipl[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double a = 0, b = 0;
            for (int m = 0; m < np; ++m) { a += p[m][i]; b += pl[m][i]; }
            ipg[i] = a / np;
            ipl[i] = b / np;
        }
        double n[3] = {0, 0, 0};
        if constexpr (D == 2) {
            n[0] = p[1][1] - p[0][1];
            n[1] = -(p[1][0] - p[0][0]);
        } else {
            double d1[3], d2[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) { d1[i] = p[2][i] - p[0][i]; d2[i] = p[3][i] - p[1][i]; }
            n[0] = 0.5 * (d1[1] * d2[2] - d1[2] * d2[1]);
            n[1] = 0.5 * (d1[2] * d2[0] - d1[0] * d2[2]);
            n[2] = 0.5 * (d1[0] * d2[1] - d1[1] * d2[0]);
        }
        {
            double o = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) o += n[i] * (ipg[i] - c[i]);
            if (o < 0)
#pragma unroll
                for (int i = 0; i < D; ++i) n[i] = -n[i];
        }
        // shapes and global gradients at the local ip
        double phi[8], gl[8][3];
        shapes_at_point<ET>(ipl, phi, gl);
        double Jm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int jj = 0; jj < D; ++jj) Jm[i][jj] += co[k][i] * gl[k][jj];
        double Ji[3][3];
        if constexpr (D == 2) {
            const double det = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
            if (det == 0.0) { kernel_error(err, KERR_DEGENERATE); continue; }
            Ji[0][0] = Jm[1][1] / det;  Ji[0][1] = -Jm[0][1] / det;
            Ji[1][0] = -Jm[1][0] / det; Ji[1][1] = Jm[0][0] / det;
        } else {
            const double det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1]) -
                               Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[1][2] * Jm[2][0]) +
                               Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0]);
            if (det == 0.0) { kernel_error(err, KERR_DEGENERATE); continue; }
            Ji[0][0] = (Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1]) / det;
            Ji[0][1] = (Jm[0][2] * Jm[2][1] - Jm[0][1] * Jm[2][2]) / det;
            Ji[0][2] = (Jm[0][1] * Jm[1][2] - Jm[0][2] * Jm[1][1]) / det;
            Ji[1][0] = (Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2]) / det;
            Ji[1][1] = (Jm[0][0] * Jm[2][2] - Jm[0][2] * Jm[2][0]) / det;
            Ji[1][2] = (Jm[0][2] * Jm[1][0] - Jm[0][0] * Jm[1][2]) / det;
            Ji[2][0] = (Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0]) / det;
            Ji[2][1] = (Jm[0][1] * Jm[2][0] - Jm[0][0] * Jm[2][1]) / det;
            Ji[2][2] = (Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0]) / det;
        }
        double g[N][D];
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = 0;
#pragma unroll
                for (int jj = 0; jj < D; ++jj) acc += Ji[jj][i] * gl[k][jj];
                g[k][i] = acc;
            }
        double vel[3] = {0, 0, 0};
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
            for (int k = 0; k < N; ++k) vel[d1] += ul[k][d1] * phi[k];
        double vn = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) vn += vel[i] * n[i];
        double flux = vn * rho;
        if (flux < 0) flux = 0;
        const double sa = A.s_a;
        if (A.kind == BND_SYMMETRY) {
            if (dout) {
                double gv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, F[3] = {0, 0, 0}, pr = 0;
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2)
#pragma unroll
                        for (int k = 0; k < N; ++k) gv[d1][d2] += g[k][d2] * ul[k][d1];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) F[d1] += gv[d1][d2] * n[d2];
                if (!A.laplace)
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) F[d1] += gv[d2][d1] * n[d2];
#pragma unroll
                for (int k = 0; k < N; ++k) pr += phi[k] * ul[k][D];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) dsum[d1] += sa * (F[d1] * (-nu * rho) * n[d1] + pr * n[d1]);
            }
            if (J) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const int col = (int)((pw >> (8 * k)) & 0xffu) * NC;
                    double gn = 0;
#pragma unroll
                    for (int i = 0; i < D; ++i) gn += g[k][i] * n[i];
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) {
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) {
                            double val = (d1 == d2) ? gn * n[d1] : 0.0;
                            if (!A.laplace) val += g[k][d1] * n[d2] * n[d1];
                            J[rows[d1] + col + d2] += JT(sa * (-nu * rho) * val);
                        }
                        J[rows[d1] + col + D] += JT(sa * (phi[k] * n[d1]));
                    }
                }
            }
            continue;
        }
        if (A.kind == BND_SLIDING) {
            double len = 0, nh[3] = {0, 0, 0};
#pragma unroll
            for (int i = 0; i < D; ++i) len += n[i] * n[i];
            len = sqrt(len);
#pragma unroll
            for (int i = 0; i < D; ++i) nh[i] = n[i] / len;
            if (!A.stokes) {
                double t[3] = {0, 0, 0}, tn = 0, f = 0;
#pragma unroll
                for (int i = 0; i < D; ++i) tn += vel[i] * nh[i];
#pragma unroll
                for (int i = 0; i < D; ++i) t[i] = vel[i] - tn * nh[i];
#pragma unroll
                for (int i = 0; i < D; ++i) f += t[i] * n[i];
                f *= rho;
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) dsum[d1] += sa * (f * t[d1]);
                if (J)
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        const int col = (int)((pw >> (8 * k)) & 0xffu) * NC;
#pragma unroll
                        for (int d1 = 0; d1 < D; ++d1) J[rows[d1] + col + d1] += JT(sa * (f * phi[k]));
                    }
            }
            double sum_all = 0, sum_dim = 0, w[3] = {0, 0, 0}, wn = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) sum_all += phi[k];
#pragma unroll
            for (int k = 0; k < D; ++k) sum_dim += phi[k];
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                for (int k = 0; k < D; ++k) w[d1] += ul[k][d1] * phi[k];
#pragma unroll
            for (int i = 0; i < D; ++i) wn += w[i] * nh[i];
#pragma unroll
            for (int i = 0; i < D; ++i) w[i] -= wn * nh[i];
            double wl = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) wl += w[i] * w[i];
            wl = sqrt(wl);
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1) {
                double r = -1.0 * w[d1] * A.slide_k;
                if (wl != 0.0) r += -1.0 * (w[d1] / wl) * A.slide_lim;
                dsum[d1] += sa * r;
            }
            if (J) {
                const double sl = sqrt((double)D * sum_all * sum_all);
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const int col = (int)((pw >> (8 * k)) & 0xffu) * NC;
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) {
                            double val = -1.0 * ((d1 == d2 ? sum_all : 0.0) - sum_dim * nh[d1] * nh[d2]) * A.slide_k;
                            if (d1 == d2 && sl != 0.0) val += -1.0 * (sum_all / sl) * A.slide_lim;
                            J[rows[d1] + col + d2] += JT(sa * val);
                        }
                }
            }
            continue;
        }
        // defect
        if (dout) {
            double gv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, F[3] = {0, 0, 0};
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2)
#pragma unroll
                    for (int k = 0; k < N; ++k) gv[d1][d2] += g[k][d2] * ul[k][d1];
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) F[d1] += gv[d1][d2] * n[d2];
            if (!A.laplace)
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) F[d1] += gv[d2][d1] * n[d2];
            double fn = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) fn += F[i] * n[i];
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1) {
                F[d1] -= fn * n[d1];
                double r = F[d1] * (-nu * rho);
                if (!A.stokes) r += flux * vel[d1];
                dsum[d1] += sa * r;
            }
            dsum[D] += sa * (vn * rho);
        }
        // Jacobian
        if (J) {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const int col = (int)((pw >> (8 * k)) & 0xffu) * NC;
                double df[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, nst[3] = {0, 0, 0};
                double gn = 0;
#pragma unroll
                for (int i = 0; i < D; ++i) gn += g[k][i] * n[i];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) df[d1][d1] = gn;
                if (!A.laplace)
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) df[d1][d2] += g[k][d1] * n[d2];
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2)
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) nst[d2] += df[d1][d2] * n[d1];
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2)
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) df[d1][d2] -= n[d1] * nst[d2];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) {
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) {
                        double val = df[d1][d2] * (-nu * rho);
                        if (!A.stokes && d1 == d2) val += flux * phi[k];
                        J[rows[d1] + col + d2] += JT(sa * val);
                    }
                }
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) J[rows[D] + col + d2] += JT(sa * (phi[k] * n[d2] * rho));
            }
        }
    }
    if (dout)
#pragma unroll
        for (int a = 0; a < NC; ++a) dout[v * NC + a] += dsum[a];
}

hipError_t launch_outflow(int et, int jprec, const OutflowArgs& a, const double* u, void* J, double* d, int* err,
                          hipStream_t s)
{
    if (a.nbn <= 0) return hipSuccess;
    const unsigned bl = (unsigned)((a.nbn + 127) / 128);
#define NSFV1_OUT(ETV)                                                                                  \
    if (jprec == 64) outflow_rows<ETV, double><<<bl, 128, 0, s>>>(a, u, static_cast<double*>(J), d, err); \
    else outflow_rows<ETV, float><<<bl, 128, 0, s>>>(a, u, static_cast<float*>(J), d, err);
    switch (et) {
    case 0: NSFV1_OUT(0) break;
    case 1: NSFV1_OUT(1) break;
    case 2: NSFV1_OUT(2) break;
    case 3: NSFV1_OUT(3) break;
    case 4: NSFV1_OUT(4) break;
    default: return hipErrorInvalidValue;
    }
#undef NSFV1_OUT
    return hipGetLastError();
}

}  // namespace nsfv1
