// nsfv1_device.hpp -- shared device definitions: enums, kernel parameters
// and the SCV volume formulas of the FV1 dual mesh.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

enum { ET_TRI = 0, ET_QUAD = 1, ET_TET = 2, ET_HEX = 3, ET_PRISM = 4 };
enum { UPW_NO = 0, UPW_FULL = 1, UPW_SKEWED = 2, UPW_LPS = 3, UPW_POSITIVE = 4, UPW_REGULAR = 5 };
enum { STAB_FIELDS = 0, STAB_WITHOUT = 1, STAB_FLOW = 2 };
enum { DL_RAW = 0, DL_FIVEPOINT = 1, DL_COR = 2 };

// error flags reported by kernels: one int word per kind, set by a plain
// store of 1 (any lane may set it; no lane ever clears it), so concurrent
// writers of different kinds cannot overwrite each other
enum { KERR_CUT_SIDE = 0, KERR_DEGENERATE = 1, KERR_SINGULAR_BLOCK = 2, KERR_WORDS = 4 };

__device__ inline void kernel_error(int* err, int kind) { err[kind] = 1; }

struct KParams {
    int conv_upwind, stab_upwind, stab_type, pac, difflen;
    int stokes, laplace, peclet, has_source, has_old;
    int parts;
    int upw_mode;  // skewed-upwind decisions: 0 fresh, 1 fresh + record, 2 replay the record
    int accum;     // row gathers: 0 every row of the plan from zero; element groups of a mixed mesh
                   // (empty node ranges skipped): 2 written from zero, 1 added to what J and d hold
    double exact, nu, rho, dt, inv_rho, inv_dt;
    double f[3];
    double s_a, s_m, s_f;
};

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ inline void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
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
__device__ inline void scv_volumes(const double (&co)[ET <= ET_QUAD ? (ET == ET_TRI ? 3 : 4) : (ET == ET_TET ? 4 : 8)][ET <= ET_QUAD ? 2 : 3], double* vol)
{
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
}

}  // namespace nsfv1
