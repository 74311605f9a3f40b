// nsfv1_topo.hpp -- compile-time reference-element topology and FV1 tables.
//
// UG4 corner/edge/side numbering of the triangle, quadrilateral,
// tetrahedron, hexahedron and prism; the FV1 sub-control-volume-face (SCVF) of
// edge i has from/to = the edge's corners and, in 3D, spans the edge
// midpoint, the centres of the two faces containing the edge and the
// element centre.  Local integration points, shape values and reference
// gradients at those points are evaluated here by constexpr code with the
// same operation order as the run-time construction, so every table value
// is the correctly rounded double the run-time code would produce.
#pragma once

namespace nsfv1 {

template <int ET> struct Topo;

template <> struct Topo<0> {  // triangle
    static constexpr int dim = 2, nsh = 3, nip = 3, nside = 3, cip = 2;
    static constexpr int edge[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    static constexpr int side_n[3] = {2, 2, 2};
    static constexpr int side[3][4] = {{0, 1, 0, 0}, {1, 2, 0, 0}, {2, 0, 0, 0}};
    static constexpr double lc[3][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
};
template <> struct Topo<1> {  // quadrilateral
    static constexpr int dim = 2, nsh = 4, nip = 4, nside = 4, cip = 2;
    static constexpr int edge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static constexpr int side_n[4] = {2, 2, 2, 2};
    static constexpr int side[4][4] = {{0, 1, 0, 0}, {1, 2, 0, 0}, {2, 3, 0, 0}, {3, 0, 0, 0}};
    static constexpr double lc[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
};
template <> struct Topo<2> {  // tetrahedron
    static constexpr int dim = 3, nsh = 4, nip = 6, nside = 4, cip = 3;
    static constexpr int edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    static constexpr int side_n[4] = {3, 3, 3, 3};
    static constexpr int side[4][4] = {{0, 2, 1, 0}, {1, 2, 3, 0}, {0, 3, 2, 0}, {0, 1, 3, 0}};
    static constexpr double lc[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};
template <> struct Topo<3> {  // hexahedron
    static constexpr int dim = 3, nsh = 8, nip = 12, nside = 6, cip = 3;
    static constexpr int edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                        {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    static constexpr int side_n[6] = {4, 4, 4, 4, 4, 4};
    static constexpr int side[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                       {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
    static constexpr double lc[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                        {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
};

template <> struct Topo<4> {  // prism: triangle 0-1-2 at zeta = 0, 3-4-5 above it
    static constexpr int dim = 3, nsh = 6, nip = 9, nside = 5, cip = 3;
    static constexpr int edge[9][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 4}, {2, 5}, {3, 4}, {4, 5}, {5, 3}};
    static constexpr int side_n[5] = {3, 4, 4, 4, 3};
    static constexpr int side[5][4] = {{0, 2, 1, 0}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}, {3, 4, 5, 0}};
    static constexpr double lc[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}};
};

template <int ET>
constexpr bool side_has(int s, int c)
{
    for (int j = 0; j < Topo<ET>::side_n[s]; ++j)
        if (Topo<ET>::side[s][j] == c) return true;
    return false;
}

// face A (q = 0) or B (q = 1) containing edge ip, in side order (3D)
template <int ET>
constexpr int edge_face(int ip, int q)
{
    int m = 0;
    for (int s = 0; s < Topo<ET>::nside; ++s)
        if (side_has<ET>(s, Topo<ET>::edge[ip][0]) && side_has<ET>(s, Topo<ET>::edge[ip][1])) {
            if (m == q) return s;
            ++m;
        }
    return -1;
}

// q-th edge (= SCVF) at corner k, ascending
template <int ET>
constexpr int corner_ip(int k, int q)
{
    int m = 0;
    for (int e = 0; e < Topo<ET>::nip; ++e)
        if (Topo<ET>::edge[e][0] == k || Topo<ET>::edge[e][1] == k) {
            if (m == q) return e;
            ++m;
        }
    return -1;
}

// local ip coordinate i of SCVF ip (average of the local SCVF corners)
template <int ET>
constexpr double local_ip(int ip, int i)
{
    using T = Topo<ET>;
    double cl = 0;
    for (int k = 0; k < T::nsh; ++k) cl += T::lc[k][i];
    cl /= T::nsh;
    const double mel = 0.5 * (T::lc[T::edge[ip][0]][i] + T::lc[T::edge[ip][1]][i]);
    if (T::dim == 2) return 0.5 * (mel + cl);
    const int A = edge_face<ET>(ip, 0), B = edge_face<ET>(ip, 1);
    double fa = 0, fb = 0;
    for (int j = 0; j < T::side_n[A]; ++j) fa += T::lc[T::side[A][j]][i];
    for (int j = 0; j < T::side_n[B]; ++j) fb += T::lc[T::side[B][j]][i];
    fa /= T::side_n[A];
    fb /= T::side_n[B];
    return 0.25 * (mel + fa + cl + fb);
}

// P1/Q1 shape value (j < 0) or reference derivative d/dx_j of shape k at x
template <int ET>
constexpr double shape_at(int k, int j, double x0, double x1, double x2)
{
    if (ET == 0) {
        const double v[3] = {1.0 - x0 - x1, x0, x1};
        const double g[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
        return j < 0 ? v[k] : g[k][j];
    }
    if (ET == 1) {
        const double a = x0, b = x1;
        const double v[4] = {(1 - a) * (1 - b), a * (1 - b), a * b, (1 - a) * b};
        const double g[4][2] = {{-(1 - b), -(1 - a)}, {(1 - b), -a}, {b, a}, {-b, (1 - a)}};
        return j < 0 ? v[k] : g[k][j];
    }
    if (ET == 2) {
        const double v[4] = {1.0 - x0 - x1 - x2, x0, x1, x2};
        const double g[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        return j < 0 ? v[k] : g[k][j];
    }
    if (ET == 4) {
        // P1 triangle (barycentric lambda) times P1 segment in zeta
        const double lam[3] = {1.0 - x0 - x1, x0, x1};
        const double dl[3][2] = {{-1, -1}, {1, 0}, {0, 1}};
        const int t = k % 3;
        const double w = k < 3 ? 1.0 - x2 : x2, dw = k < 3 ? -1.0 : 1.0;
        if (j < 0) return lam[t] * w;
        return j < 2 ? dl[t][j] * w : lam[t] * dw;
    }
    const double* c = Topo<3>::lc[k];
    const double x[3] = {x0, x1, x2};
    double f[3] = {0, 0, 0}, df[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        f[i] = c[i] > 0.5 ? x[i] : 1.0 - x[i];
        df[i] = c[i] > 0.5 ? 1.0 : -1.0;
    }
    if (j < 0) return f[0] * f[1] * f[2];
    if (j == 0) return df[0] * f[1] * f[2];
    if (j == 1) return f[0] * df[1] * f[2];
    return f[0] * f[1] * df[2];
}

template <int ET>
constexpr double ip_shape(int ip, int k)
{
    return shape_at<ET>(k, -1, local_ip<ET>(ip, 0), local_ip<ET>(ip, 1), Topo<ET>::dim == 3 ? local_ip<ET>(ip, 2) : 0.0);
}
template <int ET>
constexpr double ip_grad(int ip, int k, int j)
{
    return shape_at<ET>(k, j, local_ip<ET>(ip, 0), local_ip<ET>(ip, 1), Topo<ET>::dim == 3 ? local_ip<ET>(ip, 2) : 0.0);
}

// sub-control volumes of a corner: its three edges (3D) and the faces of
// each edge pair (hexahedral SCV through corner, edge midpoints, face
// centres, element centre)
template <int ET>
constexpr int pair_face(int k, int p)  // p: 0 -> (e0,e1), 1 -> (e0,e2), 2 -> (e1,e2)
{
    const int a = p == 2 ? 1 : 0, b = p == 0 ? 1 : 2;
    const int ea = corner_ip<ET>(k, a), eb = corner_ip<ET>(k, b);
    for (int s = 0; s < Topo<ET>::nside; ++s)
        if (side_has<ET>(s, Topo<ET>::edge[ea][0]) && side_has<ET>(s, Topo<ET>::edge[ea][1]) &&
            side_has<ET>(s, Topo<ET>::edge[eb][0]) && side_has<ET>(s, Topo<ET>::edge[eb][1]))
            return s;
    return -1;
}

// ---------------------------------------------------------------------
// canonical frame of a corner: a symmetry of the reference element (as a
// corner permutation) that maps reference corner 0 to corner kv.
// canonical corner k' is original corner perm(kv, k').  Hexahedra and
// quadrilaterals use the reflections x -> x xor c(kv); tetrahedra (any
// permutation of the corners is affine) and triangles list kv first;
// prisms rotate their triangles and swap levels.
// ---------------------------------------------------------------------
template <int ET>
constexpr int corner_of_coords(int x, int y, int z)
{
    for (int k = 0; k < Topo<ET>::nsh; ++k)
        if ((int)Topo<ET>::lc[k][0] == x && (int)Topo<ET>::lc[k][1] == y && (int)Topo<ET>::lc[k][2] == z) return k;
    return -1;
}

template <int ET>
constexpr int canon_perm(int kv, int kp)
{
    using T = Topo<ET>;
    if (ET == 1 || ET == 3) {
        const int x = (int)T::lc[kp][0] ^ (int)T::lc[kv][0];
        const int y = (int)T::lc[kp][1] ^ (int)T::lc[kv][1];
        const int z = ET == 3 ? ((int)T::lc[kp][2] ^ (int)T::lc[kv][2]) : 0;
        return corner_of_coords<ET>(x, y, z);
    }
    if (ET == 0) return (kv + kp) % 3;
    // prism: rotate the triangles (any corner permutation of a triangle is
    // affine) and swap the two levels
    if (ET == 4) return (kp % 3 + kv % 3) % 3 + 3 * ((kp / 3) ^ (kv / 3));
    // tetrahedron: kv first, the others ascending
    if (kp == 0) return kv;
    int m = 0;
    for (int k = 0; k < 4; ++k)
        if (k != kv) {
            if (++m == kp) return k;
        }
    return -1;
}

// original edge index of the q-th canonical SCVF at canonical corner 0,
// and whether kv is that edge's `from` corner
template <int ET>
constexpr int canon_orig_ip(int kv, int q)
{
    using T = Topo<ET>;
    const int cip = corner_ip<ET>(0, q);
    const int a = canon_perm<ET>(kv, T::edge[cip][0]), b = canon_perm<ET>(kv, T::edge[cip][1]);
    for (int e = 0; e < T::nip; ++e)
        if ((T::edge[e][0] == a && T::edge[e][1] == b) || (T::edge[e][0] == b && T::edge[e][1] == a)) return e;
    return -1;
}

template <int ET>
constexpr bool canon_is_from(int kv, int q)
{
    return Topo<ET>::edge[canon_orig_ip<ET>(kv, q)][0] == kv;
}

}  // namespace nsfv1
