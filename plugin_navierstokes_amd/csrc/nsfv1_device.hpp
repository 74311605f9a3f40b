// nsfv1_device.hpp -- device-side restatement of the FV1 incompressible
// Navier-Stokes element arithmetic for gfx950.
//
// The reference computes, per element, the FV1 dual-mesh geometry
// (ugcore FV1Geometry, used at incompressible/fv1/navier_stokes_fv1.cpp:214),
// the upwind shapes (upwind.cpp:52-575), the Schneider-Raw FIELDS
// stabilisation (incompressible/fv1/stabilization.cpp:122-242) and the SCVF
// flux loops (navier_stokes_fv1.cpp:250-778).  Here the same arithmetic is
// evaluated per (row node, incident element) pair: only the sub-control-
// volume faces that touch the row node are formed, because every SCVF adds
// to exactly two rows (from, to) and the row owner gathers its own row.
//
// Decision-carrying quantities (geometry, interpolated velocity, the upwind
// flux sign and the side/ray intersection) are evaluated without FMA
// contraction, in the operation order of the CPU oracle, so that upwind
// choices agree bit for bit; value arithmetic is free to contract.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

enum { ET_TRI = 0, ET_QUAD = 1, ET_TET = 2, ET_HEX = 3 };
enum { UPW_NO = 0, UPW_FULL = 1, UPW_SKEWED = 2, UPW_LPS = 3 };
enum { STAB_FIELDS = 0, STAB_WITHOUT = 1 };
enum { DL_RAW = 0, DL_FIVEPOINT = 1, DL_COR = 2 };

// error bits reported by kernels
enum { KERR_CUT_SIDE = 1, KERR_DEGENERATE = 2 };

template <int ET> struct Shape;
template <> struct Shape<ET_TRI>  { static constexpr int dim = 2, nsh = 3, nip = 3, nside = 3, cip = 2; };
template <> struct Shape<ET_QUAD> { static constexpr int dim = 2, nsh = 4, nip = 4, nside = 4, cip = 2; };
template <> struct Shape<ET_TET>  { static constexpr int dim = 3, nsh = 4, nip = 6, nside = 4, cip = 3; };
template <> struct Shape<ET_HEX>  { static constexpr int dim = 3, nsh = 8, nip = 12, nside = 6, cip = 3; };

// Reference-element topology (UG4 numbering), filled by the host.
struct RefTopo {
    int8_t edge[12][2];   // SCVF ip <-> edge, from/to corners
    int8_t side_n[6];
    int8_t side[6][4];    // sides (2D: edges, 3D: faces) in reference order
    int8_t ef[12][2];     // 3D: the two faces containing edge i, in side order
    int8_t cip[8][3];     // edges (= SCVF ids) at each corner, ascending
    double lc[8][3];      // local corner coordinates
};

struct KParams {
    int conv_upwind, stab_upwind, stab_type, pac, difflen;
    int stokes, laplace, peclet, has_source, has_old;
    int parts;
    double exact, nu, rho, dt;
    double f[3];
    double s_a, s_m, s_f;
};

// ---------------------------------------------------------------------
// P1/Q1 Lagrange shapes on the reference element (same formulas and
// operation order as the oracle so that shape values agree bitwise)
// ---------------------------------------------------------------------
template <int ET>
__device__ inline void ref_shapes(const double* x, double* phi, double (*gl)[3])
{
#pragma clang fp contract(off)
    if constexpr (ET == ET_TRI) {
        phi[0] = 1.0 - x[0] - x[1]; phi[1] = x[0]; phi[2] = x[1];
        gl[0][0] = -1; gl[0][1] = -1; gl[1][0] = 1; gl[1][1] = 0; gl[2][0] = 0; gl[2][1] = 1;
    } else if constexpr (ET == ET_QUAD) {
        const double a = x[0], b = x[1];
        phi[0] = (1 - a) * (1 - b); phi[1] = a * (1 - b); phi[2] = a * b; phi[3] = (1 - a) * b;
        gl[0][0] = -(1 - b); gl[0][1] = -(1 - a);
        gl[1][0] = (1 - b);  gl[1][1] = -a;
        gl[2][0] = b;        gl[2][1] = a;
        gl[3][0] = -b;       gl[3][1] = (1 - a);
    } else if constexpr (ET == ET_TET) {
        phi[0] = 1.0 - x[0] - x[1] - x[2]; phi[1] = x[0]; phi[2] = x[1]; phi[3] = x[2];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < 3; ++j) gl[k][j] = 0;
        gl[0][0] = gl[0][1] = gl[0][2] = -1;
        gl[1][0] = 1; gl[2][1] = 1; gl[3][2] = 1;
    } else {
        // corner k of the hexahedron sits at (k&1 ^ k>>1&1, k>>1&1, k>>2)
        const double m0 = 1.0 - x[0], m1 = 1.0 - x[1], m2 = 1.0 - x[2];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int cy = (k >> 1) & 1, cz = (k >> 2) & 1;
            const int cx = (k & 1) ^ cy;
            const double f0 = cx ? x[0] : m0, f1 = cy ? x[1] : m1, f2 = cz ? x[2] : m2;
            const double d0 = cx ? 1.0 : -1.0, d1 = cy ? 1.0 : -1.0, d2 = cz ? 1.0 : -1.0;
            phi[k] = f0 * f1 * f2;
            gl[k][0] = d0 * f1 * f2;
            gl[k][1] = f0 * d1 * f2;
            gl[k][2] = f0 * f1 * d2;
        }
    }
}

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <int D>
__device__ inline double dotd(const double* a, const double* b)
{
#pragma clang fp contract(off)
    double s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += a[i] * b[i];
    return s;
}

template <int D>
__device__ inline double dist2d(const double* a, const double* b)
{
#pragma clang fp contract(off)
    double s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
    return s;
}

__device__ inline void cross3(const double* a, const double* b, double* c)
{
#pragma clang fp contract(off)
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---------------------------------------------------------------------
// FV1 geometry of one SCVF (restated ugcore FV1Geometry::update)
// ---------------------------------------------------------------------
template <int ET>
struct IpGeo {
    static constexpr int D = Shape<ET>::dim, N = Shape<ET>::nsh;
    double ipg[D];      // global ip
    double n[D];        // normal, |n| = face measure, oriented from -> to
    double phi[N];      // shapes at the ip
    double g[N][D];     // global gradients at the ip
    double c0[D], c2[D];  // SCVF corner 0 (edge midpoint) and the element centre
};

template <int ET>
__device__ inline bool ip_geometry(const RefTopo& R, const double (*co)[Shape<ET>::dim], int ip, IpGeo<ET>& G)
{
    constexpr int D = Shape<ET>::dim, N = Shape<ET>::nsh;
    const int f = R.edge[ip][0], t = R.edge[ip][1];
    double c[D], cl[D], me[D], mel[D];
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int i = 0; i < D; ++i) { c[i] = 0; cl[i] = 0; }
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) { c[i] += co[k][i]; cl[i] += R.lc[k][i]; }
#pragma unroll
        for (int i = 0; i < D; ++i) { c[i] /= N; cl[i] /= N; }
#pragma unroll
        for (int i = 0; i < D; ++i) {
            me[i] = 0.5 * (co[f][i] + co[t][i]);
            mel[i] = 0.5 * (R.lc[f][i] + R.lc[t][i]);
        }
    }
    double ipl[3] = {0, 0, 0};
    double n[3] = {0, 0, 0};
    if constexpr (D == 2) {
#pragma clang fp contract(off)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            G.ipg[i] = 0.5 * (me[i] + c[i]);
            ipl[i] = 0.5 * (mel[i] + cl[i]);
            G.c0[i] = me[i];
            G.c2[i] = c[i];
        }
        const double d0 = c[0] - me[0], d1 = c[1] - me[1];
        n[0] = d1; n[1] = -d0;
    } else {
#pragma clang fp contract(off)
        const int A = R.ef[ip][0], B = R.ef[ip][1];
        double fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0}, fal[3] = {0, 0, 0}, fbl[3] = {0, 0, 0};
        const int na = R.side_n[A], nb = R.side_n[B];
        for (int j = 0; j < na; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) { fa[i] += co[R.side[A][j]][i]; fal[i] += R.lc[R.side[A][j]][i]; }
        for (int j = 0; j < nb; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) { fb[i] += co[R.side[B][j]][i]; fbl[i] += R.lc[R.side[B][j]][i]; }
#pragma unroll
        for (int i = 0; i < 3; ++i) { fa[i] /= na; fal[i] /= na; fb[i] /= nb; fbl[i] /= nb; }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            G.ipg[i] = 0.25 * (me[i] + fa[i] + c[i] + fb[i]);
            ipl[i] = 0.25 * (mel[i] + fal[i] + cl[i] + fbl[i]);
            G.c0[i] = me[i];
            G.c2[i] = c[i];
        }
        double d1[3], d2[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) { d1[i] = c[i] - me[i]; d2[i] = fb[i] - fa[i]; }
        cross3(d1, d2, n);
#pragma unroll
        for (int i = 0; i < 3; ++i) n[i] *= 0.5;
    }
    {
#pragma clang fp contract(off)
        double e[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < D; ++i) e[i] = co[t][i] - co[f][i];
        const double s = dotd<D>(n, e);
#pragma unroll
        for (int i = 0; i < D; ++i) G.n[i] = s < 0 ? -n[i] : n[i];
    }
    double gl[N][3];
    ref_shapes<ET>(ipl, G.phi, gl);
    // Jacobian of the element map at the local ip, its inverse, J^-T gl_k
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) s += co[k][i] * gl[k][j];
            J[i][j] = s;
        }
    double Ji[D][D];
    if constexpr (D == 2) {
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        if (det == 0.0) return false;
        const double id = 1.0 / det;
        Ji[0][0] = J[1][1] * id;  Ji[0][1] = -J[0][1] * id;
        Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
    } else {
        const double a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
        const double a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        const double det = J[0][0] * a00 + J[1][0] * a01 + J[2][0] * a02;
        if (det == 0.0) return false;
        const double id = 1.0 / det;
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
            G.g[k][i] = s;
        }
    return true;
}

// ---------------------------------------------------------------------
// SCV volumes (FV1Geometry SCV: corner, edge midpoints, face centres,
// element centre).  Simplices: |element| / nsh (affine).  Quadrilateral:
// planar quad area through the diagonals.  Hexahedron: exact integral of
// det J of the trilinear element map over the corner's octant of the
// reference cube (det J is tri-quadratic, so its Q2 interpolant on the
// 3x3x3 grid is exact; the octant weights are the integrals of the 1D
// quadratic Lagrange basis over [0,1/2] and [1/2,1]).
// ---------------------------------------------------------------------
template <int ET>
__device__ inline void scv_volumes(const double (*co)[Shape<ET>::dim], double* vol)
{
    constexpr int N = Shape<ET>::nsh;
    if constexpr (ET == ET_TRI) {
        const double a = ((co[1][0] - co[0][0]) * (co[2][1] - co[0][1]) - (co[1][1] - co[0][1]) * (co[2][0] - co[0][0]));
        const double v = fabs(a) * (0.5 / 3.0);
        vol[0] = vol[1] = vol[2] = v;
    } else if constexpr (ET == ET_QUAD) {
        double c[2] = {0, 0};
#pragma unroll
        for (int k = 0; k < 4; ++k) { c[0] += co[k][0]; c[1] += co[k][1]; }
        c[0] /= 4; c[1] /= 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int kn = (k + 1) & 3, kp = (k + 3) & 3;
            // quad [x_k, m(k,kn), c, m(kp,k)]: diagonals c-x_k and m(k,kn)-m(kp,k)
            const double d1x = c[0] - co[k][0], d1y = c[1] - co[k][1];
            const double d2x = 0.5 * (co[kn][0] - co[kp][0]), d2y = 0.5 * (co[kn][1] - co[kp][1]);
            vol[k] = 0.5 * fabs(d1x * d2y - d1y * d2x);
        }
    } else if constexpr (ET == ET_TET) {
        double a[3], b[3], c[3], x[3];
        for (int i = 0; i < 3; ++i) { a[i] = co[1][i] - co[0][i]; b[i] = co[2][i] - co[0][i]; c[i] = co[3][i] - co[0][i]; }
        cross3(b, c, x);
        const double v = fabs(dot3(a, x)) * (1.0 / 24.0);
        vol[0] = vol[1] = vol[2] = vol[3] = v;
    } else {
        // edge vectors of the three directions at the four (other-axis) corners
        // X-columns: dx/dxi at (eta,zeta) in {0,1}^2: x1-x0, x2-x3, x5-x4, x6-x7
        double ex[4][3], ey[4][3], ez[4][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ex[0][i] = co[1][i] - co[0][i]; ex[1][i] = co[2][i] - co[3][i];
            ex[2][i] = co[5][i] - co[4][i]; ex[3][i] = co[6][i] - co[7][i];
            // dx/deta at (xi,zeta): x3-x0, x2-x1, x7-x4, x6-x5
            ey[0][i] = co[3][i] - co[0][i]; ey[1][i] = co[2][i] - co[1][i];
            ey[2][i] = co[7][i] - co[4][i]; ey[3][i] = co[6][i] - co[5][i];
            // dx/dzeta at (xi,eta): x4-x0, x5-x1, x7-x3, x6-x2
            ez[0][i] = co[4][i] - co[0][i]; ez[1][i] = co[5][i] - co[1][i];
            ez[2][i] = co[7][i] - co[3][i]; ez[3][i] = co[6][i] - co[2][i];
        }
        // bilinear interpolation on the 3x3 grid {0,1/2,1}^2: index (p,q)
        auto bil = [](const double (*e)[3], int p, int q, double* out) {
            const double wp[3][2] = {{1, 0}, {0.5, 0.5}, {0, 1}};
#pragma unroll
            for (int i = 0; i < 3; ++i)
                out[i] = wp[p][0] * wp[q][0] * e[0][i] + wp[p][1] * wp[q][0] * e[1][i]
                       + wp[p][0] * wp[q][1] * e[2][i] + wp[p][1] * wp[q][1] * e[3][i];
        };
        double X[3][3][3], Y[3][3][3], Z[3][3][3];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                bil(ex, p, q, X[p][q]);   // X[eta][zeta]
                bil(ey, p, q, Y[p][q]);   // Y[xi][zeta]
                bil(ez, p, q, Z[p][q]);   // Z[xi][eta]
            }
        // 1D weights: integral over [0,1/2] (w[0]) and [1/2,1] (w[1]) of the
        // quadratic Lagrange basis on nodes {0, 1/2, 1}
        const double w[2][3] = {{5.0 / 24, 8.0 / 24, -1.0 / 24}, {-1.0 / 24, 8.0 / 24, 5.0 / 24}};
        double acc[2][2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double yz[3];
                    cross3(Y[a][c], Z[a][b], yz);
                    const double det = dot3(X[b][c], yz);
#pragma unroll
                    for (int oa = 0; oa < 2; ++oa)
#pragma unroll
                        for (int ob = 0; ob < 2; ++ob)
#pragma unroll
                            for (int oc = 0; oc < 2; ++oc) acc[oa][ob][oc] += w[oa][a] * w[ob][b] * w[oc][c] * det;
                }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int cy = (k >> 1) & 1, cz = (k >> 2) & 1;
            const int cx = (k & 1) ^ cy;
            vol[k] = fabs(acc[cx][cy][cz]);
        }
    }
    (void)N;
}

// ---------------------------------------------------------------------
// ElementSideRayIntersection (ugcore, restated) searching upwind: first
// side in reference order hit at ray parameter t <= 0, side parameters
// inside with tolerance 1e-10 (upwind.cpp:351-353, 547-549)
// ---------------------------------------------------------------------
#define NSFV1_RAY_EPS 1e-10

__device__ inline bool ray_tri(const double* p0, const double* p1, const double* p2,
                               const double* from, const double* dir, double& b1o, double& b2o)
{
#pragma clang fp contract(off)
    double e1[3], e2[3], pv[3], tv[3], qv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { e1[i] = p1[i] - p0[i]; e2[i] = p2[i] - p0[i]; }
    cross3(dir, e2, pv);
    const double det = dotd<3>(e1, pv);
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i) tv[i] = from[i] - p0[i];
    const double b1 = dotd<3>(tv, pv) * inv;
    cross3(tv, e1, qv);
    const double b2 = dotd<3>(dir, qv) * inv;
    const double t = dotd<3>(e2, qv) * inv;
    if (b1 < -NSFV1_RAY_EPS || b2 < -NSFV1_RAY_EPS || b1 + b2 > 1.0 + NSFV1_RAY_EPS) return false;
    if (t > 0.0) return false;
    b1o = b1; b2o = b2;
    return true;
}

template <int ET>
__device__ inline bool side_ray(const RefTopo& R, const double (*co)[Shape<ET>::dim], const double* from,
                                const double* dir, int& side, double* gi, double* li)
{
#pragma clang fp contract(off)
    constexpr int D = Shape<ET>::dim, NS = Shape<ET>::nside;
    if constexpr (D == 2) {
        for (int s = 0; s < NS; ++s) {
            const double* p0 = co[R.side[s][0]];
            const double* p1 = co[R.side[s][1]];
            const double ex = p1[0] - p0[0], ey = p1[1] - p0[1];
            const double rx = p0[0] - from[0], ry = p0[1] - from[1];
            const double det = -dir[0] * ey + ex * dir[1];
            if (det == 0.0) continue;
            const double t = (-rx * ey + ex * ry) / det;
            const double lam = (dir[0] * ry - dir[1] * rx) / det;
            if (lam < -NSFV1_RAY_EPS || lam > 1.0 + NSFV1_RAY_EPS) continue;
            if (t > 0.0) continue;
            side = s;
            gi[0] = p0[0] + lam * ex;
            gi[1] = p0[1] + lam * ey;
            const double* l0 = R.lc[R.side[s][0]];
            const double* l1 = R.lc[R.side[s][1]];
#pragma unroll
            for (int i = 0; i < 2; ++i) li[i] = (1.0 - lam) * l0[i] + lam * l1[i];
            return true;
        }
    } else {
        for (int s = 0; s < NS; ++s) {
            const int ntri = R.side_n[s] == 3 ? 1 : 2;
            for (int tr = 0; tr < ntri; ++tr) {
                const int a = R.side[s][0];
                const int b = R.side[s][tr == 0 ? 1 : 2];
                const int c = R.side[s][tr == 0 ? 2 : 3];
                double b1, b2;
                if (!ray_tri(co[a], co[b], co[c], from, dir, b1, b2)) continue;
                side = s;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    gi[i] = co[a][i] + b1 * (co[b][i] - co[a][i]) + b2 * (co[c][i] - co[a][i]);
                    li[i] = (1.0 - b1 - b2) * R.lc[a][i] + b1 * R.lc[b][i] + b2 * R.lc[c][i];
                }
                return true;
            }
        }
    }
    return false;
}

// Upwind shapes U[nsh] and convection length of one ip:
// upwind.cpp:52-80 (no), 133-172 (full), 337-430 (skewed), 505-575 (LPS)
template <int ET>
__device__ inline bool upwind_ip(int type, const RefTopo& R, const double (*co)[Shape<ET>::dim], int ip,
                                 const IpGeo<ET>& G, const double* vel, double* U, double& convlen)
{
#pragma clang fp contract(off)
    constexpr int D = Shape<ET>::dim, N = Shape<ET>::nsh;
#pragma unroll
    for (int k = 0; k < N; ++k) U[k] = 0.0;
    if (type == UPW_NO) {
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = G.phi[k];
        convlen = 1.0;
        return true;
    }
    if (type == UPW_FULL) {
        const double flux = dotd<D>(G.n, vel);
        const int c = flux > 0.0 ? R.edge[ip][0] : R.edge[ip][1];
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = (k == c) ? 1.0 : 0.0;
        convlen = sqrt(dist2d<D>(G.ipg, co[c]));
        return true;
    }
    if (sqrt(dotd<D>(vel, vel)) < 1e-14) { convlen = 1.0; return true; }
    int side = 0;
    double gi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
    if (!side_ray<ET>(R, co, G.ipg, vel, side, gi, li)) return false;
    if (type == UPW_SKEWED) {
        double mn = 1.7976931348623157e308;
        int coOut = 0;
        for (int i = 0; i < R.side_n[side]; ++i) {
            const int c = R.side[side][i];
            const double dd = dist2d<D>(gi, co[c]);
            if (dd < mn) { mn = dd; coOut = c; }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = (k == coOut) ? 1.0 : 0.0;
        convlen = sqrt(dist2d<D>(G.ipg, co[coOut]));
    } else {
        double phi[N], gl[N][3];
        ref_shapes<ET>(li, phi, gl);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            bool on = false;
            for (int j = 0; j < R.side_n[side]; ++j) on |= (R.side[side][j] == k);
            U[k] = on ? phi[k] : 0.0;
        }
        convlen = sqrt(dist2d<D>(G.ipg, gi));
    }
    return true;
}

}  // namespace nsfv1
