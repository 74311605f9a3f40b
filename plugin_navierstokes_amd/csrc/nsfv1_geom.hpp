// nsfv1_geom.hpp -- FV1 geometry shared by the assembly kernels: compile-time
// shape tables at the SCVF integration points, the SCVF frame (ip, normal,
// edge midpoint; restated ugcore FV1Geometry), the prism's SCV volumes,
// the inverse element Jacobian at an ip, and the reference-coordinate ray
// cast of affine elements used by the skewed upwinds.
#pragma once

#include "nsfv1_device.hpp"
#include "nsfv1_ray.hpp"
#include "nsfv1_topo.hpp"

namespace nsfv1 {

// ---------------------------------------------------------------------
// compile-time tables per element type
// ---------------------------------------------------------------------
template <int ET>
struct Tab {
    double phi[12][8];
    double gl[12][8][3];
    int cip[8][3];
    int efA[12], efB[12];
};

template <int ET>
constexpr Tab<ET> make_tab()
{
    Tab<ET> t{};
    for (int ip = 0; ip < Topo<ET>::nip; ++ip) {
        for (int k = 0; k < Topo<ET>::nsh; ++k) {
            t.phi[ip][k] = ip_shape<ET>(ip, k);
            for (int j = 0; j < Topo<ET>::dim; ++j) t.gl[ip][k][j] = ip_grad<ET>(ip, k, j);
        }
        t.efA[ip] = edge_face<ET>(ip, 0);
        t.efB[ip] = edge_face<ET>(ip, 1);
    }
    for (int k = 0; k < Topo<ET>::nsh; ++k)
        for (int q = 0; q < Topo<ET>::cip; ++q) t.cip[k][q] = corner_ip<ET>(k, q);
    return t;
}

template <int ET>
struct TabH {
    static constexpr Tab<ET> t = make_tab<ET>();
};

// ip, normal and the edge midpoint of SCVF IP -- restated ugcore FV1Geometry
template <int ET, int IP>
__device__ inline void scvf_frame(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* c,
                                  double* ipg, double* n, double* me)
{
    using T = Topo<ET>;
    constexpr int D = T::dim;
    constexpr int f = T::edge[IP][0], t = T::edge[IP][1];
#pragma unroll
    for (int i = 0; i < D; ++i) me[i] = 0.5 * (co[f][i] + co[t][i]);
    double nn[3] = {0, 0, 0};
    if constexpr (D == 2) {
#pragma unroll
        for (int i = 0; i < 2; ++i) ipg[i] = 0.5 * (me[i] + c[i]);
        const double d0 = c[0] - me[0], d1 = c[1] - me[1];
        nn[0] = d1;
        nn[1] = -d0;
    } else {
        constexpr int A = TabH<ET>::t.efA[IP], B = TabH<ET>::t.efB[IP];
        constexpr int na = T::side_n[A], nb = T::side_n[B];
        double fa[3], fb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double sa = 0, sb = 0;
#pragma unroll
            for (int j = 0; j < na; ++j) sa += co[T::side[A][j]][i];
#pragma unroll
            for (int j = 0; j < nb; ++j) sb += co[T::side[B][j]][i];
            fa[i] = sa / na;
            fb[i] = sb / nb;
        }
        double d1[3], d2[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ipg[i] = 0.25 * (me[i] + fa[i] + c[i] + fb[i]);
            d1[i] = c[i] - me[i];
            d2[i] = fb[i] - fa[i];
        }
        cross_nc(d1, d2, nn);
#pragma unroll
        for (int i = 0; i < 3; ++i) nn[i] *= 0.5;
    }
    double e[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < D; ++i) e[i] = co[t][i] - co[f][i];
    const double s = dot_nc<D>(nn, e);
#pragma unroll
    for (int i = 0; i < D; ++i) n[i] = s < 0 ? -nn[i] : nn[i];
}

// SCV volumes of a prism.  Every corner has three edges, so its SCV is the
// hexahedron [corner, midpoints of edges e0 < e1 < e2, the face centres of
// the edge pairs, the element centre]; its volume is that of the trilinear
// hexahedron through the eight points (2x2x2 Gauss on det J, exact for a
// trilinear map), as FV1Geometry's hexahedral SCVs.
__device__ inline double trilinear_hex_volume(const double (&p)[8][3])
{
    const double g[2] = {0.5 - 0.5 / sqrt(3.0), 0.5 + 0.5 / sqrt(3.0)};
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const double x[3] = {g[a], g[b], g[c]};
                double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    double f[3], df[3];
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const bool hi = Topo<3>::lc[k][j] > 0.5;
                        f[j] = hi ? x[j] : 1.0 - x[j];
                        df[j] = hi ? 1.0 : -1.0;
                    }
                    const double gl[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) J[i][j] += p[k][i] * gl[j];
                }
                const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                   J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                v += 0.125 * det;
            }
    return fabs(v);
}

__device__ inline void scv_volumes_prism(const double (&co)[6][3], double* vol)
{
    using T = Topo<ET_PRISM>;
    double c[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) c[i] += co[k][i];
#pragma unroll
    for (int i = 0; i < 3; ++i) c[i] /= 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double p[8][3];
        const int e[3] = {corner_ip<ET_PRISM>(k, 0), corner_ip<ET_PRISM>(k, 1), corner_ip<ET_PRISM>(k, 2)};
        const int f[3] = {pair_face<ET_PRISM>(k, 0), pair_face<ET_PRISM>(k, 1), pair_face<ET_PRISM>(k, 2)};
        double me[3][3], fc[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                me[q][i] = 0.5 * (co[T::edge[e[q]][0]][i] + co[T::edge[e[q]][1]][i]);
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j < T::side_n[f[q]]) s += co[T::side[f[q]][j]][i];
                fc[q][i] = s / T::side_n[f[q]];
            }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            p[0][i] = co[k][i];
            p[1][i] = me[0][i];
            p[2][i] = fc[0][i];
            p[3][i] = me[1][i];
            p[4][i] = me[2][i];
            p[5][i] = fc[1][i];
            p[6][i] = c[i];
            p[7][i] = fc[2][i];
        }
        vol[k] = trilinear_hex_volume(p);
    }
}

// J^-1 of the element map at SCVF IP (J_ij = dx_i/dxi_j); false if singular
template <int ET, int IP>
__device__ inline bool inv_jacobian(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                    double (&Ji)[Topo<ET>::dim][Topo<ET>::dim])
{
    constexpr int D = Topo<ET>::dim, N = Topo<ET>::nsh;
    double J[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) s += co[k][i] * TabH<ET>::t.gl[IP][k][j];
            J[i][j] = s;
        }
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
    return true;
}

// reference-element facts used by ray_affine (constant-folded after unrolling)
template <int ET>
constexpr int side_axis(int S)  // box elements: the reference axis normal to side S
{
    using T = Topo<ET>;
    int r = 0;
    for (int a = 0; a < T::dim; ++a) {
        bool same = true;
        for (int j = 1; j < T::side_n[S]; ++j)
            if (T::lc[T::side[S][j]][a] != T::lc[T::side[S][0]][a]) same = false;
        if (same) r = a;
    }
    return r;
}
template <int ET>
constexpr bool side_high(int S)
{
    return Topo<ET>::lc[Topo<ET>::side[S][0]][side_axis<ET>(S)] > 0.5;
}
template <int ET>
constexpr int side_opposite(int S)  // simplices: the corner not on side S
{
    for (int c = 0; c < Topo<ET>::nsh; ++c)
        if (!side_has<ET>(S, c)) return c;
    return 0;
}

// Ray cast of an affine element (every simplex; parallelogram quads and
// parallelepiped hexahedra per the geometry cache) in reference
// coordinates: xi(s) = xi_ip + s * J^-1 (-v), s >= 0 (upwind).  The side
// whose plane the ray leaves through is found by one division per
// candidate side instead of a ray/triangle test per side triangle; among
// sides that pass the inside test (tolerance NSFV1_RAY_EPS) the first in
// reference order wins, as in the general test.  Leaves side < 0 when no
// side passes (the caller then runs the general test).
template <int ET>
__device__ inline void ray_affine(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* ipg,
                                  const double* sv, int& side, double* gi, double* li)
{
    using T = Topo<ET>;
    constexpr int D = T::dim;
    constexpr bool box = (ET == 1 || ET == 3);
    constexpr int a1 = 1, a2 = box ? 3 : 2, a3 = box ? 4 : 3;
    double Jm[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        Jm[i][0] = co[a1][i] - co[0][i];
        Jm[i][1] = co[a2][i] - co[0][i];
        if constexpr (D == 3) Jm[i][2] = co[a3][i] - co[0][i];
    }
    double K[D][D];
    if constexpr (D == 2) {
        const double det = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
        if (det == 0.0) return;
        const double id = 1.0 / det;
        K[0][0] = Jm[1][1] * id; K[0][1] = -Jm[0][1] * id;
        K[1][0] = -Jm[1][0] * id; K[1][1] = Jm[0][0] * id;
    } else {
        const double c00 = Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1];
        const double c01 = Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2];
        const double c02 = Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0];
        const double det = Jm[0][0] * c00 + Jm[0][1] * c01 + Jm[0][2] * c02;
        if (det == 0.0) return;
        const double id = 1.0 / det;
        K[0][0] = c00 * id;
        K[1][0] = c01 * id;
        K[2][0] = c02 * id;
        K[0][1] = (Jm[0][2] * Jm[2][1] - Jm[0][1] * Jm[2][2]) * id;
        K[1][1] = (Jm[0][0] * Jm[2][2] - Jm[0][2] * Jm[2][0]) * id;
        K[2][1] = (Jm[0][1] * Jm[2][0] - Jm[0][0] * Jm[2][1]) * id;
        K[0][2] = (Jm[0][1] * Jm[1][2] - Jm[0][2] * Jm[1][1]) * id;
        K[1][2] = (Jm[0][2] * Jm[1][0] - Jm[0][0] * Jm[1][2]) * id;
        K[2][2] = (Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0]) * id;
    }
    double xi[D], dh[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double x = 0.0, d = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            x += K[a][i] * (ipg[i] - co[0][i]);
            d -= K[a][i] * sv[i];
        }
        xi[a] = x;
        dh[a] = d;
    }
    constexpr double eps = NSFV1_RAY_EPS;
    double xs[D];
    bool found = false;
#pragma unroll
    for (int S = 0; S < T::nside; ++S) {
        if (found) continue;
        if constexpr (box) {
            // side S lies on the plane xi_ax = hi (reference corner coordinates)
            const int ax = side_axis<ET>(S);
            const bool hi = side_high<ET>(S);
            const bool toward = hi ? dh[ax] > 0.0 : dh[ax] < 0.0;
            if (!toward) continue;
            const double sp = (hi ? 1.0 - xi[ax] : -xi[ax]) / dh[ax];
            bool in = true;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const double p = a == ax ? (hi ? 1.0 : 0.0) : xi[a] + sp * dh[a];
                xs[a] = p;
                if (a != ax) in &= (p >= -eps) && (p <= 1.0 + eps);
            }
            if (in) { side = S; found = true; }
        } else {
            // side S is opposite the corner it does not contain
            const int opp = side_opposite<ET>(S);
            double lam[D + 1], dl[D + 1];
            lam[0] = 1.0;
            dl[0] = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                lam[0] -= xi[a];
                dl[0] -= dh[a];
                lam[a + 1] = xi[a];
                dl[a + 1] = dh[a];
            }
            if (!(dl[opp] < 0.0)) continue;
            const double sp = -lam[opp] / dl[opp];
            bool in = true;
#pragma unroll
            for (int c = 0; c <= D; ++c)
                if (c != opp) in &= (lam[c] + sp * dl[c] >= -eps);
#pragma unroll
            for (int a = 0; a < D; ++a) xs[a] = (a + 1 == opp) ? 0.0 : xi[a] + sp * dh[a];
            if (in) { side = S; found = true; }
        }
    }
    if (!found) return;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double x = co[0][i];
#pragma unroll
        for (int a = 0; a < D; ++a) x += Jm[i][a] * xs[a];
        gi[i] = x;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) li[a] = a < D ? xs[a] : 0.0;
}

}  // namespace nsfv1
