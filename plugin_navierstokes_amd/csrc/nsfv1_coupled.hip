// nsfv1_coupled.hip -- assembly of the configurations whose SCVFs within an
// element are coupled:
//   * the FLOW stabilisation (stabilization.cpp:436-772): velocity
//     components connected, plus the downwind of its upwind;
//   * upwinds with ip shapes -- the positive upwind (upwind.cpp:643-786):
//     the upwind velocity of an SCVF takes the velocities of the SCVFs that
//     flow into the same SCV, so the stabilised ip velocities solve a
//     numIp x numIp system per element (stabilization.cpp:244-403,
//     604-772) and the convective Jacobian folds the ip shapes into corner
//     shapes (navier_stokes_fv1.cpp:430-457).
//
// Kernel coupled_scvf: a workgroup is one wavefront holding EPW = 64 / nip
// elements, one lane per (element, SCVF).  The lanes of an element exchange
// their SCVF data through LDS (ip velocities, shapes, gradients, mass
// fluxes, upwind shapes) and invert the ip system together: Gauss-Jordan
// with partial pivoting, one row per lane, the pivot search by the
// element's first lane.  Each lane then writes its SCVF's contribution to
// the row block of its `from` corner -- (dim+1) x nsh x (dim+1) Jacobian
// values and dim+1 defect values; the `to` corner receives the negative --
// or, for the data export, the stabilised velocity.
//
// Kernel coupled_rows: the row-owner gather shared with the other paths:
// one lane per (node, element) incidence adds the records of the SCVFs at
// its corner with their sign and the mass and source parts, and the lanes
// of a workgroup add their rows into LDS in rounds of the incidence slot,
// so the sums are reproducible without atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_internal.hpp"
#include "nsfv1_ray.hpp"
#include "nsfv1_topo.hpp"

namespace nsfv1 {

// per-element data the lanes of an element share
template <int ET>
struct CplShared {
    static constexpr int D = Topo<ET>::dim, N = Topo<ET>::nsh, NIP = Topo<ET>::nip;
    double sv[NIP][D], ipg[NIP][D];
    double phi[NIP][N], g[NIP][N][D];
    double mflux[NIP];
    int has[NIP];
    double vv[NIP], vc[NIP], vd[NIP];
    double Us[NIP][N], Ud[NIP][N];
    double aug[NIP][2 * NIP];
    int piv;
};

// corner c (runtime) of a per-lane coordinate array, by selects over the
// compile-time corners, so the array stays in registers
template <int N, int D>
__device__ inline void pick_corner(const double (&co)[N][D], int c, double (&x)[D])
{
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) v = k == c ? co[k][i] : v;
        x[i] = v;
    }
}

template <int ET>
struct CplCfg {
    static constexpr int NIP = Topo<ET>::nip;
    static constexpr int EPW = 64 / NIP;  // elements per workgroup
};

// The ray o - t v (t > 0) against the segment [a, b] (2D): true with the
// segment parameter lambda (from a, clamped to [0, 1], tolerance 1e-10) when
// it crosses; every product rounded on its own, as in the oracle.
__device__ inline bool ray_segment_2d(const double* o, const double* v, const double* a, const double* b,
                                      double& lam)
{
#pragma clang fp contract(off)
    const double d0 = b[0] - a[0], d1 = b[1] - a[1];
    const double r0 = o[0] - a[0], r1 = o[1] - a[1];
    const double det = d0 * v[1] - d1 * v[0];
    if (det == 0.0) return false;
    const double l = (r0 * v[1] - r1 * v[0]) / det;
    const double t = (d0 * r1 - d1 * r0) / det;
    if (!(t > 0.0) || l < -1e-10 || l > 1.0 + 1e-10) return false;
    lam = l < 0.0 ? 0.0 : (l > 1.0 ? 1.0 : l);
    return true;
}

// Regular upwind of one SCVF (upwind.cpp:792-964, 2D).  ugcore's
// SCVFofSCVRayIntersection is taken with this convention: the SCV of corner
// k is the quad [x_k, m_k, c, m_{k-1}] (m_j the midpoint of edge j, c the
// element centre); the ray ip - t v leaves it through the other SCVF (than
// the lane's), side 0 = [x_k, m_k] or side 3 = [m_{k-1}, x_k], tried in that
// order; lambda runs from x_k to m_k on side 0, from m_{k-1} to x_k on side
// 3, and from the SCVF's edge midpoint to c on the SCVF.
template <int ET>
__device__ inline bool regular_lane(int ip, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* n,
                                    const double* ipg, const double* v, double (&U)[Topo<ET>::nsh],
                                    double (&W)[Topo<ET>::nip])
{
    using T = Topo<ET>;
    constexpr int N = T::nsh;
    const double eps = 2.220446049250313e-16 * 10;
    if (fabs(dot_nc<2>(v, v)) <= eps) return true;  // no convection: no upwind shapes
    double flux = dot_nc<2>(v, n);
    if (fabs(flux) <= 100 * eps) {
        // flow along the SCVF: the neighbouring ips (toward the element
        // boundary) or the SCVF's corners (toward the centre)
        {
#pragma clang fp contract(off)
            flux = v[0] * n[1] - v[1] * n[0];
        }
        if (flux > 0) W[(ip + N - 1) % N] = W[(ip + 1) % N] = 0.5;
        else U[ip] = U[(ip + 1) % N] = 0.5;
        return true;
    }
    const int k = flux > 0.0 ? T::edge[ip][0] : T::edge[ip][1];
    const int oth = flux > 0.0 ? (ip + N - 1) % N : (ip + 1) % N;
    double c[2], xk[2], mo[2], mk[2], mkm[2];
    elem_centre<ET>(co, c);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        double a = 0.0, b = 0.0, x = 0.0, bb = 0.0, cc = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) {
            x = q == k ? co[q][i] : x;
            a = q == T::edge[oth][0] ? co[q][i] : a;
            b = q == T::edge[oth][1] ? co[q][i] : b;
            bb = q == T::edge[k][1] ? co[q][i] : bb;
            cc = q == T::edge[(k + N - 1) % N][0] ? co[q][i] : cc;
        }
        xk[i] = x;
        mo[i] = 0.5 * (a + b);
        mk[i] = 0.5 * (x + bb);
        mkm[i] = 0.5 * (cc + x);
    }
    double lam = 0.0;
    if (ray_segment_2d(ipg, v, mo, c, lam)) {
        int ipI = (ip + 1) % N, ipO = (ip + N - 1) % N;
        if (flux > 0.0) {
            ipI = (ip + N - 1) % N;
            ipO = (ip + 1) % N;
        }
        if (lam <= 0.5) {
            W[ipO] = lam - 0.5;
            W[ipI] = 1.0 - (lam - 0.5);
        } else {
            U[T::edge[ip][0]] = 0.5 * 2 * (lam - 0.5);
            U[T::edge[ip][1]] = 0.5 * 2 * (lam - 0.5);
            W[ipI] = 1.0 - 2 * (lam - 0.5);
        }
        return true;
    }
    int side = -1;
    if (ray_segment_2d(ipg, v, xk, mk, lam)) side = 0;
    else if (ray_segment_2d(ipg, v, mkm, xk, lam)) side = 3;
    if (side < 0) return false;
    if (flux > 0.0) {
        if (side == 0) {
            U[ip] = 1.0 - 0.5 * lam;
            U[(ip + 1) % N] = 0.5 * lam;
        } else {
            U[(ip + N - 1) % N] = 1.0 - 0.5 * (lam + 1.0);
            U[ip] = 0.5 * (lam + 1.0);
        }
    } else {
        if (side == 0) {
            U[(ip + 1) % N] = 1.0 - 0.5 * lam;
            U[(ip + 2) % N] = 0.5 * lam;
        } else {
            U[ip] = 1.0 - 0.5 * (lam + 1.0);
            U[(ip + 1) % N] = 0.5 * (lam + 1.0);
        }
    }
    return true;
}

// upwind shapes of the lane's SCVF for velocity sign * v (sign = -1: the
// downwind of upwind_interface.h:157-165): corner shapes U, ip shapes W,
// convection length.  Returns false when a ray finds no side.
template <int ET>
__device__ inline bool upwind_lane(int type, double sign, int ip, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                   const double* n, const CplShared<ET>& S, double (&U)[Topo<ET>::nsh],
                                   double (&W)[Topo<ET>::nip], double& len)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NIP = T::nip;
    const int fr = T::edge[ip][0], to = T::edge[ip][1];
    const double* ipg = S.ipg[ip];
    double v[D];
#pragma unroll
    for (int i = 0; i < D; ++i) v[i] = sign * S.sv[ip][i];
#pragma unroll
    for (int k = 0; k < N; ++k) U[k] = 0.0;
#pragma unroll
    for (int j = 0; j < NIP; ++j) W[j] = 0.0;
    len = 1.0;
    if (type == UPW_NO) {  // upwind.cpp:52-80
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = S.phi[ip][k];
        return true;
    }
    if (type == UPW_FULL) {  // upwind.cpp:133-172
        const int c = dot_nc<D>(n, v) > 0.0 ? fr : to;
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = k == c ? 1.0 : 0.0;
        double xc[D];
        pick_corner(co, c, xc);
        len = sqrt(dist2_nc<D>(ipg, xc));
        return true;
    }
    if (type == UPW_REGULAR) {
        if constexpr (D == 2) {
            if (!regular_lane<ET>(ip, co, n, ipg, v, U, W)) return false;
        } else {
            return false;  // refused on the host for 3D meshes
        }
        // convection length to the upwind point (upwind.cpp:945-962)
        double up[D];
#pragma unroll
        for (int i = 0; i < D; ++i) up[i] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) up[i] += U[k] * co[k][i];
#pragma unroll
        for (int j = 0; j < NIP; ++j)
#pragma unroll
            for (int i = 0; i < D; ++i) up[i] += W[j] * S.ipg[j][i];
        len = sqrt(dist2_nc<D>(ipg, up));
        return true;
    }
    if (type == UPW_POSITIVE) {  // upwind.cpp:643-786
        if (!S.has[ip]) {
#pragma unroll
            for (int k = 0; k < N; ++k) U[k] = (k == fr || k == to) ? 0.5 : 0.0;
        } else {
            const int sc = sign * S.mflux[ip] > 0 ? fr : to;  // the SCV the SCVF leaves
            double m_in = 0.0, m_out = 0.0;
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                if (!S.has[j]) continue;
                const int a = T::edge[j][0], b = T::edge[j][1];
                if (a != sc && b != sc) continue;
                const double f = a == sc ? sign * S.mflux[j] : -1.0 * (sign * S.mflux[j]);
                m_in += -1.0 * fmin(f, 0.0);
                m_out += fmax(f, 0.0);
            }
            const double F = fmax(m_in, m_out);
            double sum = 0.0;
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                if (!S.has[j]) continue;
                const int a = T::edge[j][0], b = T::edge[j][1];
                if (a != sc && b != sc) continue;
                const double f = a == sc ? sign * S.mflux[j] : -1.0 * (sign * S.mflux[j]);
                if (f < 0) {
                    W[j] = -1.0 * f / F;
                    sum += W[j];
                }
            }
#pragma unroll
            for (int k = 0; k < N; ++k) U[k] = k == sc ? 1.0 - sum : 0.0;
        }
        double up[D];
#pragma unroll
        for (int i = 0; i < D; ++i) up[i] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) up[i] += U[k] * co[k][i];
#pragma unroll
        for (int j = 0; j < NIP; ++j)
#pragma unroll
            for (int i = 0; i < D; ++i) up[i] += W[j] * S.ipg[j][i];
        len = sqrt(dist2_nc<D>(ipg, up));
        return true;
    }
    // skewed and linear-profile skewed: upwind.cpp:337-430, 505-575
    if (sqrt(dot_nc<D>(v, v)) < 1e-14) return true;
    int side = -1;
    double gi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
    side_loop<ET, 0>(co, ipg, v, side, gi, li);
    if (side < 0) return false;
    if (type == UPW_SKEWED) {
        double mn = NSFV1_DBL_MAX;
        int best = 0;
        for (int j = 0; j < T::side_n[side]; ++j) {
            const int c = T::side[side][j];
            double xc[D];
            pick_corner(co, c, xc);
            const double dd = dist2_nc<D>(gi, xc);
            if (dd < mn) { mn = dd; best = c; }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = k == best ? 1.0 : 0.0;
        double xb[D];
        pick_corner(co, best, xb);
        len = sqrt(dist2_nc<D>(ipg, xb));
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = side_has<ET>(side, k) ? shape_at<ET>(k, -1, li[0], li[1], li[2]) : 0.0;
        len = sqrt(dist2_nc<D>(ipg, gi));
    }
    return true;
}

// RHOIP: the density is imported at the ips (M.rho_ip); a separate
// instance, so that the constant-density kernel keeps its registers
template <int ET, typename JT, bool RHOIP>
__global__ void __launch_bounds__(64)
coupled_scvf(KMeshPtrs M, KParams P, const double* __restrict__ u, const double* __restrict__ uold, int64_t ne,
             JT* __restrict__ recJ, double* __restrict__ recD, double* __restrict__ vel_out, int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1, NIP = T::nip, EPW = CplCfg<ET>::EPW;
    __shared__ CplShared<ET> shm[EPW];
    const int lane = threadIdx.x;
    const int es = lane / NIP, ip = lane - (lane / NIP) * NIP;
    const int64_t e = (int64_t)blockIdx.x * EPW + es;
    const bool act = es < EPW && e < ne;
    CplShared<ET>& S = shm[es < EPW ? es : EPW - 1];
    const int fr = T::edge[ip][0], to = T::edge[ip][1];
    const bool flow = P.stab_type == STAB_FLOW;
    const bool stokes = P.stokes != 0;
    const bool matrix = !stokes && P.stab_type != STAB_WITHOUT &&
                        (P.stab_upwind == UPW_POSITIVE || P.stab_upwind == UPW_REGULAR);
    bool ok_geo = true, ok_cut = true;

    // ---- geometry of the lane's SCVF (restated ugcore FV1Geometry)
    double co[N][D], ul[N][NC], n[D], l2 = 0.0, nu = P.nu;
    if (act) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int64_t w = M.conn[e * N + k];
#pragma unroll
            for (int i = 0; i < D; ++i) co[k][i] = M.xyz[w * D + i];
#pragma unroll
            for (int a = 0; a < NC; ++a) ul[k][a] = u[w * NC + a];
        }
        if (M.nu_elem) nu = M.nu_elem[e];
        if (M.nu_ip) nu = M.nu_ip[e * NIP + ip];  // function-valued viscosity at the lane's ip
        double c[D], me[D], ipg[D], nn[3] = {0, 0, 0};
        elem_centre<ET>(co, c);
        double cfr[D], cto[D];
        pick_corner(co, fr, cfr);
        pick_corner(co, to, cto);
#pragma unroll
        for (int i = 0; i < D; ++i) me[i] = 0.5 * (cfr[i] + cto[i]);
        if constexpr (D == 2) {
#pragma unroll
            for (int i = 0; i < 2; ++i) ipg[i] = 0.5 * (me[i] + c[i]);
            nn[0] = c[1] - me[1];
            nn[1] = -(c[0] - me[0]);
        } else {
            const int A = edge_face<ET>(ip, 0), B = edge_face<ET>(ip, 1);
            double fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0}, d1[3], d2[3];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double xa[3], xb[3];
                pick_corner(co, T::side[A][j], xa);
                pick_corner(co, T::side[B][j], xb);
                if (j < T::side_n[A])
#pragma unroll
                    for (int i = 0; i < 3; ++i) fa[i] += xa[i];
                if (j < T::side_n[B])
#pragma unroll
                    for (int i = 0; i < 3; ++i) fb[i] += xb[i];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                fa[i] /= T::side_n[A];
                fb[i] /= T::side_n[B];
                ipg[i] = 0.25 * (me[i] + fa[i] + c[i] + fb[i]);
                d1[i] = c[i] - me[i];
                d2[i] = fb[i] - fa[i];
            }
            cross_nc(d1, d2, nn);
#pragma unroll
            for (int i = 0; i < 3; ++i) nn[i] *= 0.5;
        }
        {
            double ed[D];
#pragma unroll
            for (int i = 0; i < D; ++i) ed[i] = cto[i] - cfr[i];
            const double s = dot_nc<D>(nn, ed);
#pragma unroll
            for (int i = 0; i < D; ++i) n[i] = s < 0 ? -nn[i] : nn[i];
        }
        // shapes and global gradients at the local ip
        double x[3] = {0, 0, 0}, gl[N][D], J[D][D], Ji[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i) x[i] = local_ip<ET>(ip, i);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            S.phi[ip][k] = shape_at<ET>(k, -1, x[0], x[1], x[2]);
#pragma unroll
            for (int j = 0; j < D; ++j) gl[k][j] = shape_at<ET>(k, j, x[0], x[1], x[2]);
        }
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
            ok_geo = det != 0.0;
            const double id = ok_geo ? 1.0 / det : 0.0;
            Ji[0][0] = J[1][1] * id;  Ji[0][1] = -J[0][1] * id;
            Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
        } else {
            const double a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            const double a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
            const double a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
            const double det = J[0][0] * a00 + J[1][0] * a01 + J[2][0] * a02;
            ok_geo = det != 0.0;
            const double id = ok_geo ? 1.0 / det : 0.0;
            Ji[0][0] = a00 * id; Ji[0][1] = a01 * id; Ji[0][2] = a02 * id;
            Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        }
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double s = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) s += Ji[j][i] * gl[k][j];
                S.g[ip][k][i] = s;
            }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) s += ul[k][d] * S.phi[ip][k];
            S.sv[ip][d] = s;
            S.ipg[ip][d] = ipg[d];
        }
        // mass flux of the positive upwind (upwind.cpp:660-690)
        {
            const double eps = 2.220446049250313e-16 * 10;
            const double nsq = dot_nc<D>(S.sv[ip], S.sv[ip]);
            S.has[ip] = 1;
            S.mflux[ip] = 0.0;
            if (fabs(nsq) <= eps) {
                S.has[ip] = 0;
            } else {
                S.mflux[ip] = dot_nc<D>(S.sv[ip], n);
                const double vl = sqrt(nsq), ln = sqrt(dot_nc<D>(n, n));
                if (fabs(S.mflux[ip] / sqrt(vl * ln)) <= eps) S.has[ip] = 0;
            }
        }
        l2 = M.l2inv[e * NIP + ip];
    }
    __syncthreads();

    // ---- upwind and downwind of the stabilisation; ip system row
    double Us[N], Ws[NIP], Ls = 1.0, Ud[N], Wd[NIP], Ld = 1.0, vv = 0.0, vc = 0.0, vd = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) Us[k] = Ud[k] = 0.0;
#pragma unroll
    for (int j = 0; j < NIP; ++j) Ws[j] = Wd[j] = 0.0;
    if (act && P.stab_type != STAB_WITHOUT) {
        if (!stokes) {
            ok_cut &= upwind_lane<ET>(P.stab_upwind, 1.0, ip, co, n, S, Us, Ws, Ls);
            if (flow) ok_cut &= upwind_lane<ET>(P.stab_upwind, -1.0, ip, co, n, S, Ud, Wd, Ld);
        }
        vv = nu * l2;
        if (!stokes) {
            const double nrm = sqrt(dot_nc<D>(S.sv[ip], S.sv[ip]));
            vc = nrm / Ls;
            if (flow) vd = nrm / (Ld + Ls);
        }
        if (matrix) {
            S.vv[ip] = vv;
            S.vc[ip] = vc;
            S.vd[ip] = vd;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                S.Us[ip][k] = Us[k];
                S.Ud[ip][k] = Ud[k];
            }
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                double m = 0.0;
                if (j == ip) {
                    if (P.has_old) m += P.inv_dt;
                    m += vv;
                    m += vc;
                }
                m -= Ws[j] * vc;
                if (flow) m += vd * (Ws[j] - Wd[j]);
                S.aug[ip][j] = m;
                S.aug[ip][NIP + j] = j == ip ? 1.0 : 0.0;
            }
        }
    }
    __syncthreads();

    // ---- Gauss-Jordan inversion of the ip system, one row per lane
    if (matrix) {
        for (int c = 0; c < NIP; ++c) {
            if (act && ip == 0) {
                int pv = c;
                for (int r = c + 1; r < NIP; ++r)
                    if (fabs(S.aug[r][c]) > fabs(S.aug[pv][c])) pv = r;
                if (S.aug[pv][c] == 0.0) ok_geo = false;
                if (pv != c)
                    for (int j = 0; j < 2 * NIP; ++j) {
                        const double t = S.aug[c][j];
                        S.aug[c][j] = S.aug[pv][j];
                        S.aug[pv][j] = t;
                    }
            }
            __syncthreads();
            if (act && ip == c && S.aug[c][c] != 0.0) {
                const double iv = 1.0 / S.aug[c][c];
                for (int j = 0; j < 2 * NIP; ++j) S.aug[c][j] *= iv;
            }
            __syncthreads();
            if (act && ip != c) {
                const double f = S.aug[ip][c];
                if (f != 0.0)
                    for (int j = 0; j < 2 * NIP; ++j) S.aug[ip][j] -= f * S.aug[c][j];
            }
            __syncthreads();
        }
    }
    if (!act) return;

    // ---- stabilised velocity of the lane's ip.  Its shapes (the derivatives
    //      by each corner's unknowns) are formed corner by corner in the
    //      Jacobian loop below, from the same expressions.
    const double rho = RHOIP ? M.rho_ip[e * NIP + ip] : P.rho;  // density at the lane's ip
    double svel[D], diag = 1.0;
    if (P.stab_type == STAB_WITHOUT) {  // stabilization.cpp:805-850
#pragma unroll
        for (int d = 0; d < D; ++d) svel[d] = S.sv[ip][d];
    } else if (!matrix) {  // diagonal case: stabilization.cpp:163-242, 486-602
        diag = vv;
        if (P.has_old) diag += P.inv_dt;
        if (!stokes) diag += vc;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double rhs = P.has_source ? (M.f_scvf ? M.f_scvf[(e * NIP + ip) * D + d] : P.f[d]) : 0.0;
            if (P.has_old) {
                double o = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) o += S.phi[ip][k] * uold[(int64_t)M.conn[e * N + k] * NC + d];
                rhs += o * P.inv_dt;
            }
#pragma unroll
            for (int k = 0; k < N; ++k) {
                double sumVel = vv * S.phi[ip][k];
                if (!stokes) {
                    sumVel += vc * Us[k];
                    if (flow) sumVel += vd * (Ud[k] - Us[k]);
                }
                if (flow)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2)
                        if (d2 != d) sumVel -= S.sv[ip][d2] * S.g[ip][k][d2];
                rhs += sumVel * ul[k][d];
                if (flow)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) {
                        if (d2 == d) continue;
                        const double sumVel2 = S.sv[ip][d] * S.g[ip][k][d2];
                        rhs += sumVel2 * ul[k][d2];
                    }
                const double sumP = -1.0 * S.g[ip][k][d] / rho;
                rhs += sumP * ul[k][D];
            }
            svel[d] = rhs / diag;
        }
    } else {  // matrix case: stabilization.cpp:244-403, 604-772
        // corner contributions of every velocity component and the pressure
        // to each ip's right-hand side f[d][j], then mapped through row ip of
        // the inverse; each f[d][j] takes its corners in ascending order, and
        // the ip data of (j, k) is read once for all components
        const double* iv = &S.aug[ip][NIP];
        double f[D][NIP];
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                f[d][j] = 0.0;
                if (!flow) {
                    if (P.has_source) f[d][j] = M.f_scvf ? M.f_scvf[(e * NIP + j) * D + d] : P.f[d];
                    if (P.has_old) {
                        double o = 0.0;
#pragma unroll
                        for (int k = 0; k < N; ++k) o += S.phi[j][k] * uold[(int64_t)M.conn[e * N + k] * NC + d];
                        f[d][j] += o * P.inv_dt;
                    }
                }
            }
#pragma unroll
        for (int k = 0; k < N; ++k) {
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                double gj[D], svj[D];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    gj[d] = S.g[j][k][d];
                    svj[d] = S.sv[j][d];
                }
                double cd0 = S.vv[j] * S.phi[j][k];
                cd0 += S.vc[j] * S.Us[j][k];
                if (flow) cd0 += S.vd[j] * (S.Ud[j][k] - S.Us[j][k]);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double cp = -1.0 * gj[d] / (RHOIP ? M.rho_ip[e * NIP + j] : rho);
                    double cd = cd0;
                    if (flow) {
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) {
                            if (d2 == d) continue;
                            cd -= svj[d2] * gj[d2];
                        }
                        // f: corner contributions of every component, then pressure
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2)
                            f[d][j] = f[d][j] + ul[k][d2] * (d2 == d ? cd : svj[d] * gj[d2]);
                        f[d][j] = f[d][j] + ul[k][D] * cp;
                    } else {
                        f[d][j] = f[d][j] + ul[k][d] * cd;
                        f[d][j] = f[d][j] + ul[k][D] * cp;
                    }
                }
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (flow)
#pragma unroll
                for (int j = 0; j < NIP; ++j) {
                    if (P.has_source) f[d][j] += M.f_scvf ? M.f_scvf[(e * NIP + j) * D + d] : P.f[d];
                    if (P.has_old) {
                        double o = 0.0;
#pragma unroll
                        for (int k = 0; k < N; ++k) o += S.phi[j][k] * uold[(int64_t)M.conn[e * N + k] * NC + d];
                        f[d][j] += o * P.inv_dt;
                    }
                }
            double x = 0.0;
#pragma unroll
            for (int j = 0; j < NIP; ++j) x += iv[j] * f[d][j];
            svel[d] = x;
        }
    }
    if (!ok_cut) kernel_error(err, KERR_CUT_SIDE);
    if (!ok_geo) kernel_error(err, KERR_DEGENERATE);
    if (vel_out) {
#pragma unroll
        for (int d = 0; d < D; ++d) vel_out[(e * NIP + ip) * D + d] = svel[d];
        return;
    }

    // ---- convective upwind of the momentum equation
    double Uc[N], Wc[NIP], Lc = 1.0, upv[D];
#pragma unroll
    for (int k = 0; k < N; ++k) Uc[k] = Us[k];
#pragma unroll
    for (int j = 0; j < NIP; ++j) Wc[j] = Ws[j];
    const bool pac = P.pac != 0;
    bool conv_ip = false;
    if (!stokes && !pac) {
        if (P.conv_upwind != P.stab_upwind || P.stab_type == STAB_WITHOUT)
            ok_cut &= upwind_lane<ET>(P.conv_upwind, 1.0, ip, co, n, S, Uc, Wc, Lc);
        conv_ip = P.conv_upwind == UPW_POSITIVE || P.conv_upwind == UPW_REGULAR;
        if (!ok_cut) kernel_error(err, KERR_CUT_SIDE);
    }
    (void)Lc;
    // upwind velocity: upwind_interface.h:335-358, or the stabilised one
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double a = 0.0;
        if (pac) {
            a = svel[d];
        } else {
#pragma unroll
            for (int k = 0; k < N; ++k) a += Uc[k] * ul[k][d];
            if (conv_ip)
#pragma unroll
                for (int j = 0; j < NIP; ++j) a += Wc[j] * S.sv[j][d];
        }
        upv[d] = a;
    }
    const double* sv = S.sv[ip];
    double w = 1.0;
    if (!stokes && P.peclet) {  // navier_stokes_fv1.cpp:871-892
        double cfr[D], cto[D];
        pick_corner(co, fr, cfr);
        pick_corner(co, to, cto);
        const double Pe = dot_nc<D>(sv, n) / dot_nc<D>(n, n) * sqrt(dist2_nc<D>(cto, cfr)) / nu;
        w = Pe * Pe / (5.0 + Pe * Pe);
#pragma unroll
        for (int d = 0; d < D; ++d) upv[d] = w * upv[d] + (1.0 - w) * sv[d];
    }
    const double prod = dot_nc<D>(sv, n) * rho;
    const bool conn = flow;
    const double sa = P.s_a;

    // ---- defect of the `from` row: navier_stokes_fv1.cpp:667-777
    {
        double gv[D][D];
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
            for (int d2 = 0; d2 < D; ++d2) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) s += S.g[ip][k][d2] * ul[k][d1];
                gv[d1][d2] = s;
            }
        double pr = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) pr += S.phi[ip][k] * ul[k][D];
        double* rd = recD + (e * NIP + ip) * NC;
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1) {
            double df = 0.0;
#pragma unroll
            for (int d2 = 0; d2 < D; ++d2) df += gv[d1][d2] * n[d2];
            if (!P.laplace)
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) df += gv[d2][d1] * n[d2];
            double r = df * (-1.0 * nu * rho);
            if (!stokes) r += upv[d1] * prod;
            r += pr * n[d1];
            rd[d1] = sa * r;
        }
        rd[D] = sa * (dot_nc<D>(svel, n) * rho);
    }

    // ---- Jacobian block of the `from` row, one column block per corner:
    //      navier_stokes_fv1.cpp:317-593
#pragma unroll
    for (int sh = 0; sh < N; ++sh) {
        // stabilised-velocity shapes of corner sh: Sv[d][d2] = d svel_d /
        // d u_{sh,d2}, Sp[d] = d svel_d / d p_sh
        double Sv[D][D], Sp[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            Sp[d] = 0.0;
#pragma unroll
            for (int d2 = 0; d2 < D; ++d2) Sv[d][d2] = 0.0;
        }
        if (P.stab_type == STAB_WITHOUT) {
#pragma unroll
            for (int d = 0; d < D; ++d) Sv[d][d] = S.phi[ip][sh];
        } else if (!matrix) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double sumVel = vv * S.phi[ip][sh];
                if (!stokes) {
                    sumVel += vc * Us[sh];
                    if (flow) sumVel += vd * (Ud[sh] - Us[sh]);
                }
                if (flow)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2)
                        if (d2 != d) sumVel -= S.sv[ip][d2] * S.g[ip][sh][d2];
                Sv[d][d] = sumVel / diag;
                if (flow)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) {
                        if (d2 == d) continue;
                        const double sumVel2 = S.sv[ip][d] * S.g[ip][sh][d2];
                        Sv[d][d2] = sumVel2 / diag;
                    }
                const double sumP = -1.0 * S.g[ip][sh][d] / rho;
                Sp[d] = sumP / diag;
            }
        } else {
            // every accumulator takes its terms in ascending j, as in the
            // stabilised velocity above; the ip data of j is read once for
            // all components
            const double* iv = &S.aug[ip][NIP];
            double xp[D], xv[D][D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                xp[d] = 0.0;
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) xv[d][d2] = 0.0;
            }
#pragma unroll
            for (int j = 0; j < NIP; ++j) {
                const double ivj = iv[j];
                double gj[D], svj[D];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    gj[d] = S.g[j][sh][d];
                    svj[d] = S.sv[j][d];
                }
                double cd0 = S.vv[j] * S.phi[j][sh];
                cd0 += S.vc[j] * S.Us[j][sh];
                if (flow) cd0 += S.vd[j] * (S.Ud[j][sh] - S.Us[j][sh]);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double cp = -1.0 * gj[d] / (RHOIP ? M.rho_ip[e * NIP + j] : rho);
                    double cd = cd0;
                    double cvo[D];
                    if (flow) {
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) {
                            if (d2 == d) continue;
                            cd -= svj[d2] * gj[d2];
                            cvo[d2] = svj[d] * gj[d2];
                        }
                    }
                    xp[d] += ivj * cp;
                    xv[d][d] += ivj * cd;
                    if (flow)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2)
                            if (d2 != d) xv[d][d2] += ivj * cvo[d2];
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                Sp[d] = xp[d];
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2)
                    if (flow || d2 == d) Sv[d][d2] = xv[d][d2];
            }
        }
        double Jb[NC][NC];
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = 0; b < NC; ++b) Jb[a][b] = 0.0;
        const double* gsh = S.g[ip][sh];
        const double phsh = S.phi[ip][sh];
        const double fsh = -1.0 * nu * rho * dot_nc<D>(gsh, n);
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1) {
            Jb[d1][d1] += fsh;
            if (!P.laplace)
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) Jb[d1][d2] += -1.0 * nu * rho * gsh[d1] * n[d2];
            Jb[d1][D] += phsh * n[d1];
        }
        if (!stokes) {
            if (pac) {
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) {
                    if (conn) {
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) Jb[d1][d2] += prod * w * Sv[d1][d2];
                    } else {
                        Jb[d1][d1] += prod * w * Sv[d1][d1];
                    }
                    Jb[d1][D] += prod * w * Sp[d1];
                }
            } else {
                double cv = Uc[sh];
                if (conv_ip)
#pragma unroll
                    for (int j = 0; j < NIP; ++j) cv += S.phi[j][sh] * Wc[j];
                cv *= prod * w;
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) Jb[d1][d1] += cv;
            }
            if (P.peclet) {
                const double cpe = prod * (1.0 - w) * phsh;
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) Jb[d1][d1] += cpe;
            }
            if (P.exact != 0.0) {
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) {
                    if (pac) {
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) {
                            double pv = 0.0;
                            if (conn) {
#pragma unroll
                                for (int k = 0; k < D; ++k) pv += w * Sv[k][d2] * n[k];
                            } else {
                                pv = Sv[d1][d1] * n[d1];
                            }
                            pv *= P.exact * rho;
                            Jb[d1][d2] += pv * upv[d1];
                        }
                        double pp = 0.0;
#pragma unroll
                        for (int k = 0; k < D; ++k) pp += Sp[k] * n[k];
                        pp *= P.exact * rho;
                        Jb[d1][D] += pp * upv[d1];
                    } else {
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) Jb[d1][d2] += (w * Uc[sh] * n[d2] * rho) * upv[d1];
                    }
                    if (P.peclet)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) Jb[d1][d2] += upv[d1] * (1.0 - w) * phsh * n[d2] * rho;
                }
            }
        }
        // continuity: navier_stokes_fv1.cpp:561-592
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1) {
            double cv = 0.0;
            if (conn) {
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) cv += Sv[d2][d1] * n[d2] * rho;
            } else {
                cv = Sv[d1][d1] * n[d1] * rho;
            }
            Jb[D][d1] += cv;
        }
        {
            double cpp = 0.0;
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1) cpp += Sp[d1] * n[d1] * rho;
            Jb[D][D] += cpp;
        }
        if (recJ) {
            JT* rj = recJ + ((e * NIP + ip) * N + sh) * (NC * NC);
#pragma unroll
            for (int a = 0; a < NC; ++a)
#pragma unroll
                for (int b = 0; b < NC; ++b) rj[a * NC + b] = JT(sa * Jb[a][b]);
        }
    }
}

// ---------------------------------------------------------------------
// row gather of the coupled records (one lane per (node, element))
// ---------------------------------------------------------------------
template <int ET, typename JT>
__global__ void __launch_bounds__(256)
coupled_rows(KMeshPtrs M, KParams P, const double* __restrict__ u, const JT* __restrict__ recJ,
             const double* __restrict__ recD, JT* __restrict__ J, double* __restrict__ dout)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1, NIP = T::nip;
    extern __shared__ __align__(16) unsigned char smem[];
    // consecutive workgroups go to the same XCD (block i runs on XCD i % 8)
    const int w = [] {
        const int nwg = (int)gridDim.x, i = (int)blockIdx.x;
        const int per = nwg >> 3, rem = nwg & 7, x = i & 7, k = i >> 3;
        return x * per + (x < rem ? x : rem) + k;
    }();
    const int v0 = M.wg_node[w], v1 = M.wg_node[w + 1];
    const int64_t jbase = M.rowptr[(int64_t)v0 * NC];
    const int jlen = (int)(M.rowptr[(int64_t)v1 * NC] - jbase);
    const int dlen = (v1 - v0) * NC;
    JT* jbuf = reinterpret_cast<JT*>(smem);
    // 16 pad bytes after each node's rows: consecutive nodes then start in
    // different LDS banks, so the lanes of one round do not conflict, and
    // every (row, column block) stays 16-byte aligned for vector access
    constexpr int PADE = NC == 4 ? 16 / (int)sizeof(JT) : 0;  // 2D rows already have odd strides
    double* dbuf = reinterpret_cast<double*>(smem + (((jlen + (v1 - v0) * PADE) * (int)sizeof(JT) + 15) & ~15));
    const bool wantJ = J != nullptr;
    const int i0 = M.node_inc[v0], i1 = M.node_inc[v1];
    // element groups of a mixed mesh (J and d zeroed before the first): a
    // workgroup none of whose nodes touch the group leaves its rows as they
    // are; the first group writes its rows (accum 2), each further group
    // adds to the rows the groups before it wrote (accum 1, same stream)
    if (P.accum && i0 == i1) return;
    if (P.accum == 1) {
        if (wantJ)
            for (int j = 0; j < v1 - v0; ++j) {
                const int off = (int)(M.rowptr[(int64_t)(v0 + j) * NC] - jbase);
                const int len = (int)(M.rowptr[(int64_t)(v0 + j + 1) * NC] - jbase) - off;
                for (int i = threadIdx.x; i < len; i += blockDim.x) jbuf[off + j * PADE + i] = J[jbase + off + i];
            }
        for (int i = threadIdx.x; i < dlen; i += blockDim.x) dbuf[i] = dout ? dout[(int64_t)v0 * NC + i] : 0.0;
    } else {
        if (wantJ)
            for (int i = threadIdx.x; i < jlen + (v1 - v0) * PADE; i += blockDim.x) jbuf[i] = JT(0);
        for (int i = threadIdx.x; i < dlen; i += blockDim.x) dbuf[i] = 0.0;
    }

    const int lane_inc = i0 + (int)threadIdx.x;
    const bool active = lane_inc < i1;
    JT jacc[NC][N][NC];
    double dacc[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        dacc[a] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int b = 0; b < NC; ++b) jacc[a][k][b] = JT(0);
    }
    int v = v0, slot = -1;
    int pos[N];
#pragma unroll
    for (int k = 0; k < N; ++k) pos[k] = 0;
    if (active) {
        const int inc = M.perm[lane_inc];
        const int64_t e = M.inc_elem[inc];
        v = M.inc_node[inc];
        const int info = M.inc_info[inc];
        const int kv = info & 0xff;
        slot = info >> 8;
        if (P.parts & 1) {
#pragma unroll
            for (int ip = 0; ip < NIP; ++ip) {
                const bool isf = T::edge[ip][0] == kv, ist = T::edge[ip][1] == kv;
                if (!isf && !ist) continue;
                const int64_t r = e * NIP + ip;
                if (wantJ) {
                    const JT* rj = recJ + r * N * (NC * NC);
#pragma unroll
                    for (int k = 0; k < N; ++k)
#pragma unroll
                        for (int a = 0; a < NC; ++a)
#pragma unroll
                            for (int b = 0; b < NC; ++b) {
                                const JT x = rj[k * (NC * NC) + a * NC + b];
                                jacc[a][k][b] += isf ? x : -x;
                            }
                }
#pragma unroll
                for (int a = 0; a < NC; ++a) dacc[a] += isf ? recD[r * NC + a] : -recD[r * NC + a];
            }
        }
        // mass part and source: navier_stokes_fv1.cpp:781-869
        if (P.parts & 6) {
            const double volk = M.vol[e * N + kv];
            const double rk = M.rho_scv ? M.rho_scv[e * N + kv] : P.rho;  // density at the SCV ip
            if (P.parts & 2) {
                const double m = volk * rk;
                const double* uv = u + (int64_t)v * NC;
#pragma unroll
                for (int d = 0; d < D; ++d) {
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        if (k == kv && wantJ) jacc[d][k][d] += JT(P.s_m * m);
                    dacc[d] += P.s_m * (uv[d] * m);
                }
            }
            if ((P.parts & 4) && P.has_source)
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double fk = M.f_scv ? M.f_scv[((int64_t)e * N + kv) * D + d] : P.f[d];
                    dacc[d] += P.s_f * (fk * volk * rk);
                }
        }
        const uint64_t pw = M.inc_pos[inc];
#pragma unroll
        for (int k = 0; k < N; ++k) pos[k] = (int)((pw >> (8 * k)) & 0xffu) * NC;
    }
    int rbase = 0, rlen = 0;
    if (active) {
        const int64_t rs = M.rowptr[(int64_t)v * NC];
        rbase = (int)(rs - jbase) + (v - v0) * PADE;
        rlen = (int)(M.rowptr[(int64_t)v * NC + 1] - rs);
    }
    __syncthreads();
    const int rounds = M.wg_rounds[w];
    for (int r = 0; r < rounds; ++r) {
        if (active && slot == r) {
            if (wantJ) {
#pragma unroll
                for (int a = 0; a < NC; ++a) {
                    const int ra = rbase + a * rlen;
#pragma unroll
                    for (int k = 0; k < N; ++k) lds_add_block<NC>(jbuf + ra + pos[k], jacc[a][k]);
                }
            }
#pragma unroll
            for (int a = 0; a < NC; ++a) dbuf[(v - v0) * NC + a] += dacc[a];
        }
        __syncthreads();
    }
    if (wantJ) {
        if constexpr (PADE == 0) {
            for (int i = threadIdx.x; i < jlen; i += blockDim.x) J[jbase + i] = jbuf[i];
        } else {
            for (int j = 0; j < v1 - v0; ++j) {
                const int off = (int)(M.rowptr[(int64_t)(v0 + j) * NC] - jbase);
                const int len = (int)(M.rowptr[(int64_t)(v0 + j + 1) * NC] - jbase) - off;
                for (int i = threadIdx.x; i < len; i += blockDim.x) J[jbase + off + i] = jbuf[off + j * PADE + i];
            }
        }
    }
    if (dout)
        for (int i = threadIdx.x; i < dlen; i += blockDim.x) dout[(int64_t)v0 * NC + i] = dbuf[i];
}

int64_t coupled_record_bytes(int et, int jprec, int64_t ne)
{
    const int nsh = et == ET_HEX ? 8 : (et == ET_PRISM ? 6 : (et == ET_TRI ? 3 : 4));
    const int nip = et == ET_TRI ? 3 : (et == ET_QUAD ? 4 : (et == ET_TET ? 6 : (et == ET_PRISM ? 9 : 12)));
    const int nc = (et == ET_TRI || et == ET_QUAD) ? 3 : 4;
    const int64_t jb = ne * nip * nsh * nc * nc * (jprec == 64 ? 8 : 4);
    return ((jb + 255) / 256) * 256 + ne * nip * nc * 8;
}

template <int ET, typename JT>
static hipError_t launch_coupled_et(const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                                    JT* J, double* d, int* err, int nwg, size_t lds, int64_t ne, void* recs,
                                    double* vel_out, hipStream_t s)
{
    constexpr int NIP = Topo<ET>::nip, N = Topo<ET>::nsh, NC = Topo<ET>::dim + 1;
    static_assert(sizeof(CplShared<ET>) * CplCfg<ET>::EPW <= 64 * 1024, "coupled LDS exceeds 64 KiB");
    const unsigned blocks = (unsigned)((ne + CplCfg<ET>::EPW - 1) / CplCfg<ET>::EPW);
    JT* recJ = nullptr;
    double* recD = nullptr;
    if (!vel_out) {
        recJ = static_cast<JT*>(recs);
        const int64_t jb = ne * NIP * N * NC * NC * (int64_t)sizeof(JT);
        recD = reinterpret_cast<double*>(static_cast<char*>(recs) + ((jb + 255) / 256) * 256);
    }
    if (blocks > 0) {
        if (m.rho_ip)
            coupled_scvf<ET, JT, true><<<blocks, 64, 0, s>>>(m, p, u, uold, ne, recJ, recD, vel_out, err);
        else
            coupled_scvf<ET, JT, false><<<blocks, 64, 0, s>>>(m, p, u, uold, ne, recJ, recD, vel_out, err);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || vel_out) return e;
    coupled_rows<ET, JT><<<nwg, m.block, lds, s>>>(m, p, u, recJ, recD, J, d);
    return hipGetLastError();
}

hipError_t launch_assemble_coupled(int et, int jprec, const KMeshPtrs& m, const KParams& p, const double* u,
                                   const double* uold, void* J, double* d, int* err, int nwg, size_t lds, int64_t ne,
                                   void* recs, double* vel_out, hipStream_t s)
{
#define NSFV1_CPL(ETV)                                                                                          \
    return jprec == 64 ? launch_coupled_et<ETV, double>(m, p, u, uold, static_cast<double*>(J), d, err, nwg, lds, ne, \
                                                        recs, vel_out, s)                                        \
                       : launch_coupled_et<ETV, float>(m, p, u, uold, static_cast<float*>(J), d, err, nwg, lds, ne,   \
                                                       recs, vel_out, s);
    switch (et) {
    case ET_TRI: NSFV1_CPL(ET_TRI)
    case ET_QUAD: NSFV1_CPL(ET_QUAD)
    case ET_TET: NSFV1_CPL(ET_TET)
    case ET_HEX: NSFV1_CPL(ET_HEX)
    case ET_PRISM: NSFV1_CPL(ET_PRISM)
    default: return hipErrorInvalidValue;
    }
#undef NSFV1_CPL
}

}  // namespace nsfv1
