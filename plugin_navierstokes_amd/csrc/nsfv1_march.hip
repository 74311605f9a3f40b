// nsfv1_march.hip -- plane-marching assembly of structured hexahedral grids
// (FP32 Jacobian + FP64 defect, gfx950).
//
// For a lexicographic nx x ny x nz grid of parallelepiped hexahedra the
// SCVF records of the split path never need to leave the chip.  A workgroup
// owns a TX x TY tile of node columns and a range of node planes and marches
// upward through the element layers.  Per layer:
//   phase 1  the FP64 SCVF work of the layer (geometry, upwind, FIELDS
//            stabilisation, flux) for every SCVF that touches a node of the
//            tile, once.  One wave takes one edge index of the hexahedron
//            for the tile's element columns, in the canonical frame of the
//            edge's `from` corner, so every shape value and reference
//            gradient is a compile-time constant of the frame (the frame's
//            axis Q is uniform over the workgroup in each pass).  The flux
//            and a compact FP32 Jacobian record (flux weight, stabilisation
//            factors, upwind corner) go to LDS, with each element's inverse
//            Jacobian, SCVF normals and SCV volume.
//   phase 2  the node plane below the layer is complete.  Its defect rows
//            add the fluxes of both adjacent layers (the lower layer's
//            part is carried in registers from the step before).  Its
//            Jacobian rows are formed by lanes that each own one family of
//            4x4 column blocks of one node (the centre block, the face, edge
//            or corner neighbours along one family of offsets: eight element
//            contributions per lane), summed in a fixed order and stored as
//            16-byte row segments.
// Every value is written once, with no atomics and fixed summation orders,
// so the assembly is bitwise reproducible.  LDS holds the records of two
// element layers (68.5 KB per workgroup, two workgroups per CU).
//
// The arithmetic follows navier_stokes_fv1.cpp:250-595 (Jacobian) and
// :597-778 (defect), 781-869 (mass, source), stabilization.cpp:122-242
// (FIELDS, diagonal case), upwind.cpp:52-80, 133-172, 337-430 (no, full and
// skewed upwind) and diffusion_length.h:91-133 (RAW), in the canonical-frame
// form of the split path's record kernel: the same expressions in the same
// order, so the defect and the upwind decisions are those of that kernel.
#include "nsfv1_geom.hpp"
#include "nsfv1_internal.hpp"
#include "nsfv1_march.hpp"

namespace nsfv1 {
namespace march {

constexpr int TX = MARCH_TX, TY = MARCH_TY;  // node columns of a tile (x, y)
constexpr int CX = TX + 1, CY = TY + 1;      // element columns touching them
constexpr int NCOL = CX * CY;                // 72
constexpr int NNX = TX + 2, NNY = TY + 2;    // nodes of those columns (with the halo)
constexpr int NN = NNX * NNY;                // 90
constexpr int NIP = 12;
constexpr int NT = 256;                      // threads per workgroup
constexpr int NSTAGE = (7 * NN + NT - 1) / NT;  // node values staged per thread

struct Smem {
    float4 rec[2][NIP][NCOL];   // per SCVF: flux weight (from -> to), vv*idg, vc*idg, -idg/rho
    double flux[NIP][NCOL][4];  // per SCVF of the current layer: flux added to `from`, taken from `to`
    float ejn[2][18][NCOL];     // per element: J^-1 (row-major, d xi_r / d x_c), then N_x, N_y, N_z
    double evol[2][NCOL];       // per element: SCV volume |det J| / 8
    double ncx[2][3][NN];       // coordinates of the nodes of two node planes (plane p in slot p & 1)
    double ncu[2][4][NN];       // states of those nodes
    uint8_t idx[2][NIP][NCOL];  // per SCVF: upwind corner (original numbering), 255 = none
};
static_assert(sizeof(Smem) <= 80 * 1024, "two workgroups per CU must fit the 160 KB of LDS");

using T3 = Topo<ET_HEX>;

constexpr int lca(int k, int a) { return (int)T3::lc[k][a]; }
constexpr int corner_at(int x, int y, int z) { return corner_of_coords<ET_HEX>(x, y, z); }

// edge (= SCVF) of the hexahedron between corner k and its neighbour along axis q
constexpr int edge_at(int k, int q)
{
    const int x = lca(k, 0) ^ (q == 0 ? 1 : 0), y = lca(k, 1) ^ (q == 1 ? 1 : 0), z = lca(k, 2) ^ (q == 2 ? 1 : 0);
    const int k2 = corner_at(x, y, z);
    for (int e = 0; e < 12; ++e)
        if ((T3::edge[e][0] == k && T3::edge[e][1] == k2) || (T3::edge[e][0] == k2 && T3::edge[e][1] == k)) return e;
    return -1;
}

struct Tables {
    int axis_edges[3][4];  // the four edges along each axis, ascending
    int perm[8][8];        // canonical corner k' of the frame of corner kf -> original corner
    int sgn[8][3];         // 1 - 2 lc[kf][a]: the frame's reflection of axis a
    int efrom[12];         // `from` corner of each edge
    int lc[8][3];
};

constexpr Tables make_tables()
{
    Tables t{};
    for (int q = 0; q < 3; ++q) {
        int m = 0;
        for (int e = 0; e < 12; ++e)
            if (lca(T3::edge[e][0], q) != lca(T3::edge[e][1], q)) t.axis_edges[q][m++] = e;
    }
    for (int kf = 0; kf < 8; ++kf) {
        for (int k = 0; k < 8; ++k) t.perm[kf][k] = canon_perm<ET_HEX>(kf, k);
        for (int a = 0; a < 3; ++a) {
            t.sgn[kf][a] = 1 - 2 * lca(kf, a);
            t.lc[kf][a] = lca(kf, a);
        }
    }
    for (int e = 0; e < 12; ++e) t.efrom[e] = T3::edge[e][0];
    return t;
}

__constant__ Tables kTab = make_tables();

// node offset of original corner k from corner 0 (runtime k)
__device__ inline int64_t corner_off(const MarchGrid& G, int k)
{
    return (int64_t)kTab.lc[k][0] + G.NX * kTab.lc[k][1] + G.NXY * kTab.lc[k][2];
}

// Affine frame of the element from its original corners 0, 1, 3, 4 (node
// ln0 of plane slot s0, its x and y neighbours, and node ln0 of slot s1) with
// the expressions of ray_affine: columns of J, J^-1 by cofactors, det J.
__device__ inline double elem_frame(const Smem& S, int s0, int s1, int ln0, double (&x0)[3], double (&Jm)[3][3],
                                    double (&K)[3][3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        x0[i] = S.ncx[s0][i][ln0];
        Jm[i][0] = S.ncx[s0][i][ln0 + 1] - x0[i];
        Jm[i][1] = S.ncx[s0][i][ln0 + NNX] - x0[i];
        Jm[i][2] = S.ncx[s1][i][ln0] - x0[i];
    }
    const double c00 = Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1];
    const double c01 = Jm[1][2] * Jm[2][0] - Jm[1][0] * Jm[2][2];
    const double c02 = Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0];
    const double det = Jm[0][0] * c00 + Jm[0][1] * c01 + Jm[0][2] * c02;
    if (det == 0.0) return 0.0;
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
    return det;
}

// Upwind side of the ray from ipg against the velocity sv, in reference
// coordinates of the parallelepiped (ray_affine's arithmetic): the first
// side in reference order whose plane the ray leaves through inside the
// side (tolerance NSFV1_RAY_EPS); side < 0 if none.
__device__ inline void ray_ref_hex(const double (&x0)[3], const double (&Jm)[3][3], const double (&K)[3][3],
                                   const double* ipg, const double* sv, int& side, double* gi)
{
    double xi[3], dh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double x = 0.0, d = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x += K[a][i] * (ipg[i] - x0[i]);
            d -= K[a][i] * sv[i];
        }
        xi[a] = x;
        dh[a] = d;
    }
    constexpr double eps = NSFV1_RAY_EPS;
    double xs[3] = {0.0, 0.0, 0.0};
    bool found = false;
#pragma unroll
    for (int S = 0; S < T3::nside; ++S) {
        if (found) continue;
        const int ax = side_axis<ET_HEX>(S);
        const bool hi = side_high<ET_HEX>(S);
        const bool toward = hi ? dh[ax] > 0.0 : dh[ax] < 0.0;
        if (!toward) continue;
        const double sp = (hi ? 1.0 - xi[ax] : -xi[ax]) / dh[ax];
        bool in = true;
        double p3[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double p = a == ax ? (hi ? 1.0 : 0.0) : xi[a] + sp * dh[a];
            p3[a] = p;
            if (a != ax) in &= (p >= -eps) && (p <= 1.0 + eps);
        }
        if (in) {
            side = S;
            found = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) xs[a] = p3[a];
        }
    }
    if (!found) return;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double x = x0[i];
#pragma unroll
        for (int a = 0; a < 3; ++a) x += Jm[i][a] * xs[a];
        gi[i] = x;
    }
}

// ---------------------------------------------------------------------
// phase 1: one SCVF of the layer per lane, canonical frame of its `from`
// corner (axis Q of the edge is uniform over the pass, the edge w of that
// axis uniform over the wave)
// ---------------------------------------------------------------------
template <int Q, int FAST>
__device__ inline void scvf_lane(Smem& S, const MarchGrid& G, const double* __restrict__ xyz, const KParams& P,
                                 const double* __restrict__ u, const double* __restrict__ uold,
                                 uint8_t* __restrict__ dec, int* __restrict__ err, int sl, int ck, int i0, int j0,
                                 int w, int lane)
{
    constexpr int IPc = corner_ip<ET_HEX>(0, Q);  // the frame's SCVF at corner 0 along axis Q
    constexpr int OT = T3::edge[IPc][0] == 0 ? T3::edge[IPc][1] : T3::edge[IPc][0];
    const int ip = kTab.axis_edges[Q][w];
    const int kf = kTab.efrom[ip];
    const int fx = kTab.lc[kf][0], fy = kTab.lc[kf][1];
    // the tile's element columns whose edge `ip` touches a node of the tile
    int ci, cj;
    bool act;
    if constexpr (Q == 0) {
        act = lane < CX * TY;
        ci = i0 - 1 + lane % CX;
        cj = j0 - fy + lane / CX;
    } else if constexpr (Q == 1) {
        act = lane < TX * CY;
        ci = i0 - fx + lane % TX;
        cj = j0 - 1 + lane / TX;
    } else {
        act = lane < TX * TY;
        ci = i0 - fx + lane % TX;
        cj = j0 - fy + lane / TX;
    }
    act = act && ci >= 0 && ci < G.nx && cj >= 0 && cj < G.ny;
    if (!act) return;
    const int col = (ci - i0 + 1) + CX * (cj - j0 + 1);
    const int64_t e = ci + (int64_t)G.nx * (cj + (int64_t)G.ny * ck);
    const int64_t v0 = ci + G.NX * cj + G.NXY * ck;
    // the element's nodes in the staged planes: local node of original
    // corner k is ln0 + lc_x + NNX lc_y in plane slot (ck + lc_z) & 1; the
    // frame's canonical corner k' is the original corner with coordinates
    // lc(k') xor lc(kf)
    const int ln0 = (ci - i0 + 1) + NNX * (cj - j0 + 1);
    const int sz0 = ck & 1, sz1 = (ck + 1) & 1;
    const int fz = kTab.lc[kf][2];

    // corners in the canonical order of the frame: first the coordinates
    // (the SCVF frame: centre, ip, normal, edge midpoint), then the state,
    // streamed into the ip value, ip gradient and old ip velocity
    double ipg[3], n[3], me[3], c[3], cf[3], ct[3];
    {
        double co[8][3];
#pragma unroll
        for (int kp = 0; kp < 8; ++kp) {
            const int ln = ln0 + ((int)T3::lc[kp][0] ^ fx) + NNX * ((int)T3::lc[kp][1] ^ fy);
            const int sz = ((int)T3::lc[kp][2] ^ fz) ? sz1 : sz0;
#pragma unroll
            for (int i = 0; i < 3; ++i) co[kp][i] = S.ncx[sz][i][ln];
        }
        elem_centre<ET_HEX>(co, c);
        double nn_[3], ev[3];
        scvf_frame<ET_HEX, IPc>(co, c, ipg, nn_, me);
#pragma unroll
        for (int i = 0; i < 3; ++i) ev[i] = co[OT][i] - co[0][i];
        const double s = dot_nc<3>(nn_, ev);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            n[i] = s < 0 ? -nn_[i] : nn_[i];
            cf[i] = co[0][i];
            ct[i] = co[OT][i];
        }
    }
    double gh[4][3], sv[3] = {0.0, 0.0, 0.0}, pr = 0.0, o[3] = {0.0, 0.0, 0.0};
    double uf[3] = {0.0, 0.0, 0.0}, ut[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) gh[a][j] = 0.0;
#pragma unroll
    for (int kp = 0; kp < 8; ++kp) {
        const int ox = (int)T3::lc[kp][0] ^ fx, oy = (int)T3::lc[kp][1] ^ fy, oz = (int)T3::lc[kp][2] ^ fz;
        const int ln = ln0 + ox + NNX * oy;
        const int sz = oz ? sz1 : sz0;
        double uk[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) uk[a] = S.ncu[sz][a][ln];
        const double ph = TabH<ET_HEX>::t.phi[IPc][kp];
#pragma unroll
        for (int d = 0; d < 3; ++d) sv[d] += uk[d] * ph;
        pr += ph * uk[3];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int j = 0; j < 3; ++j) gh[a][j] += uk[a] * TabH<ET_HEX>::t.gl[IPc][kp][j];
        if (P.has_old) {
            const int64_t node = v0 + ox + G.NX * oy + G.NXY * oz;
#pragma unroll
            for (int d = 0; d < 3; ++d) o[d] += ph * uold[node * 4 + d];
        }
        if constexpr (FAST == UPW_FULL) {
            if (kp == 0)
#pragma unroll
                for (int d = 0; d < 3; ++d) uf[d] = uk[d];
            if (kp == OT)
#pragma unroll
                for (int d = 0; d < 3; ++d) ut[d] = uk[d];
        }
    }
    // the element's affine frame (original corners 0, 1, 3, 4)
    double x0[3], Jm[3][3], K[3][3];
    const double det = elem_frame(S, sz0, sz1, ln0, x0, Jm, K);
    if (det == 0.0) {
        kernel_error(err, KERR_DEGENERATE);
        S.rec[sl][ip][col] = make_float4(0.f, 0.f, 0.f, 0.f);
        S.idx[sl][ip][col] = 255;
#pragma unroll
        for (int a = 0; a < 4; ++a) S.flux[ip][col][a] = 0.0;
        return;
    }
    // physical gradients: the frame's inverse Jacobian is the original one
    // with reflected rows
    double gv[4][3];
    {
        double Kc[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) Kc[j][i] = kTab.sgn[kf][j] < 0 ? -K[j][i] : K[j][i];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double acc = 0;
#pragma unroll
                for (int j = 0; j < 3; ++j) acc += gh[a][j] * Kc[j][i];
                gv[a][i] = acc;
            }
    }
    // RAW diffusion length (diffusion_length.h:91-133): SCV volumes of a
    // parallelepiped are |det J| / 8
    const double vol = fabs(det) * 0.125;
    double l2;
    {
        const double nnq = dot_nc<3>(n, n);
        const double ds = dist2_nc<3>(me, c);
        double a2 = 0.5 * (vol + vol);
        a2 *= a2;
        l2 = 1. / (0.5 * a2 / nnq + 3.0 * ds / 8.);
    }
    const double nu = P.nu, rho = P.rho;

    // upwind of the stabilisation and of the convective term (one object)
    int best = 255;
    double Ls = 1.0, upS[3] = {0.0, 0.0, 0.0};
    if constexpr (FAST == UPW_NO) {
#pragma unroll
        for (int d = 0; d < 3; ++d) upS[d] = sv[d];
    } else if constexpr (FAST == UPW_FULL) {
        // upwind.cpp:150-166: `from` iff n(from->to) . v > 0
        const bool pick0 = dot_nc<3>(n, sv) > 0.0;
        best = pick0 ? kf : kTab.perm[kf][OT];
        double cc[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) cc[i] = pick0 ? cf[i] : ct[i];
        Ls = sqrt(dist2_nc<3>(ipg, cc));
#pragma unroll
        for (int d = 0; d < 3; ++d) upS[d] = pick0 ? uf[d] : ut[d];
    } else {
        // upwind.cpp:337-430 (GetNodeNextToCut + skewed upwind)
        uint8_t* dp = dec ? dec + e * NIP + ip : nullptr;
        if (P.upw_mode == 2 && dp) {
            best = *dp;
        } else if (sqrt(dot_nc<3>(sv, sv)) < 1e-14) {
            best = 255;
            if (P.upw_mode == 1 && dp) *dp = 255;
        } else {
            int side = -1;
            double gi[3] = {0.0, 0.0, 0.0};
            ray_ref_hex(x0, Jm, K, ipg, sv, side, gi);
            if (side < 0) {
                kernel_error(err, KERR_CUT_SIDE);
            } else {
                double mn = NSFV1_DBL_MAX;
#pragma unroll
                for (int s = 0; s < T3::nside; ++s) {
                    if (s != side) continue;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int ck2 = T3::side[s][j];
                        const int ln = ln0 + (int)T3::lc[ck2][0] + NNX * (int)T3::lc[ck2][1];
                        const int sz = T3::lc[ck2][2] > 0.5 ? sz1 : sz0;
                        double xc[3];
#pragma unroll
                        for (int i = 0; i < 3; ++i) xc[i] = S.ncx[sz][i][ln];
                        const double dd = dist2_nc<3>(gi, xc);
                        if (dd < mn) {
                            mn = dd;
                            best = ck2;
                        }
                    }
                }
            }
            if (P.upw_mode == 1 && dp) *dp = (uint8_t)best;
        }
        if (best != 255) {
            // best is per lane (replayed or cast): its coordinates by bits
            const int bx = ((best + 1) >> 1) & 1, by = (best >> 1) & 1, bz = (best >> 2) & 1;
            const int ln = ln0 + bx + NNX * by;
            const int sz = bz ? sz1 : sz0;
            double cc[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) cc[i] = S.ncx[sz][i][ln];
            Ls = sqrt(dist2_nc<3>(ipg, cc));
#pragma unroll
            for (int d = 0; d < 3; ++d) upS[d] = S.ncu[sz][d][ln];
        }
    }

    // FIELDS stabilisation, diagonal case (stabilization.cpp:163-242)
    const double vv = nu * l2;
    const double vc = sqrt(dot_nc<3>(sv, sv)) / Ls;
    double diag = vv;
    if (P.has_old) diag += P.inv_dt;
    diag += vc;
    const double idg = 1.0 / diag;
    double svel[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double rhs = P.has_source ? P.f[d] : 0.0;
        if (P.has_old) rhs += o[d] * P.inv_dt;
        rhs += vv * sv[d] + vc * upS[d] - gv[3][d] * P.inv_rho;
        svel[d] = rhs * idg;
    }
    // flux of the defect (navier_stokes_fv1.cpp:667-777), from-row sign
    const double sgn = P.s_a;
    double fl[4];
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1) {
        double acc = 0;
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) acc += gv[d1][d2] * n[d2];
        if (!P.laplace)
#pragma unroll
            for (int d2 = 0; d2 < 3; ++d2) acc += gv[d2][d1] * n[d2];
        fl[d1] = sgn * (acc * (-1.0 * nu * rho) + pr * n[d1]);
    }
    const double prod = dot_nc<3>(sv, n) * rho;
#pragma unroll
    for (int d1 = 0; d1 < 3; ++d1) fl[d1] += sgn * (upS[d1] * prod);
    fl[3] = sgn * (dot_nc<3>(svel, n) * rho);

    S.rec[sl][ip][col] = make_float4(float(prod), float(vv * idg), float(vc * idg), float(-P.inv_rho * idg));
    S.idx[sl][ip][col] = (uint8_t)best;
    *reinterpret_cast<double2*>(&S.flux[ip][col][0]) = make_double2(fl[0], fl[1]);
    *reinterpret_cast<double2*>(&S.flux[ip][col][2]) = make_double2(fl[2], fl[3]);
    // element data: one designated writer per element column (edge 0 for
    // the rows j0.., edge 2 for the row below the tile)
    if (ip == 0 || (ip == 2 && cj == j0 - 1)) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) S.ejn[sl][3 * r + cc][col] = float(K[r][cc]);
        const double sd = det > 0.0 ? 0.25 : -0.25;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int q1 = (q + 1) % 3, q2 = (q + 2) % 3;
            double b[3], cv[3], x[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                b[i] = Jm[i][q1];
                cv[i] = Jm[i][q2];
            }
            cross_nc(b, cv, x);
#pragma unroll
            for (int i = 0; i < 3; ++i) S.ejn[sl][9 + 3 * q + i][col] = float(sd * x[i]);
        }
        S.evol[sl][col] = vol;
    }
}

template <int FAST>
__device__ inline void layer_scvfs(Smem& S, const MarchGrid& G, const double* __restrict__ xyz, const KParams& P,
                                   const double* __restrict__ u, const double* __restrict__ uold,
                                   uint8_t* __restrict__ dec, int* __restrict__ err, int ck, int i0, int j0, int wave,
                                   int lane)
{
    const int sl = ck & 1;
    scvf_lane<0, FAST>(S, G, xyz, P, u, uold, dec, err, sl, ck, i0, j0, wave, lane);
    scvf_lane<1, FAST>(S, G, xyz, P, u, uold, dec, err, sl, ck, i0, j0, wave, lane);
    scvf_lane<2, FAST>(S, G, xyz, P, u, uold, dec, err, sl, ck, i0, j0, wave, lane);
}

// ---------------------------------------------------------------------
// phase 2, defect: the four elements of layer ck around node column (i, j),
// the node at height PZ (0 bottom, 1 top) of them
// ---------------------------------------------------------------------
// flux of SCVF edge(KV, Q) into the row of corner KV: added at `from`,
// subtracted at `to`
template <int KV, int Q>
__device__ inline void node_scvf(const Smem& S, int col, double (&acc)[4])
{
    constexpr int ip = edge_at(KV, Q);
    constexpr bool from = T3::edge[ip][0] == KV;
    const double2 f01 = *reinterpret_cast<const double2*>(&S.flux[ip][col][0]);
    const double2 f23 = *reinterpret_cast<const double2*>(&S.flux[ip][col][2]);
    acc[0] += from ? f01.x : -f01.x;
    acc[1] += from ? f01.y : -f01.y;
    acc[2] += from ? f23.x : -f23.x;
    acc[3] += from ? f23.y : -f23.y;
}

template <int SX, int SY, int PZ>
__device__ inline void node_elem(const Smem& S, const MarchGrid& G, const KParams& P, int sl, int i, int j, int li,
                                 int lj, const double (&uv)[3], double (&acc)[4])
{
    const int ci = i + (SX < 0 ? -1 : 0), cj = j + (SY < 0 ? -1 : 0);
    if (ci < 0 || ci >= G.nx || cj < 0 || cj >= G.ny) return;
    const int col = li + (SX < 0 ? 0 : 1) + CX * (lj + (SY < 0 ? 0 : 1));
    constexpr int kv = corner_at(SX < 0 ? 1 : 0, SY < 0 ? 1 : 0, PZ);
    node_scvf<kv, 0>(S, col, acc);
    node_scvf<kv, 1>(S, col, acc);
    node_scvf<kv, 2>(S, col, acc);
    // mass part and source (navier_stokes_fv1.cpp:781-869)
    if (P.parts & 6) {
        const double volk = S.evol[sl][col];
        if (P.parts & 2) {
            const double m = volk * P.rho;
#pragma unroll
            for (int d = 0; d < 3; ++d) acc[d] += P.s_m * (uv[d] * m);
        }
        if ((P.parts & 4) && P.has_source)
#pragma unroll
            for (int d = 0; d < 3; ++d) acc[d] += P.s_f * (P.f[d] * volk * P.rho);
    }
}

template <int PZ>
__device__ inline void node_layer(const Smem& S, const MarchGrid& G, const KParams& P, int ck, int i, int j, int li,
                                  int lj, const double (&uv)[3], double (&acc)[4])
{
    const int sl = ck & 1;
    node_elem<-1, -1, PZ>(S, G, P, sl, i, j, li, lj, uv, acc);
    node_elem<1, -1, PZ>(S, G, P, sl, i, j, li, lj, uv, acc);
    node_elem<-1, 1, PZ>(S, G, P, sl, i, j, li, lj, uv, acc);
    node_elem<1, 1, PZ>(S, G, P, sl, i, j, li, lj, uv, acc);
}

// ---------------------------------------------------------------------
// phase 2, Jacobian.  A node row has 27 column blocks (fewer on the grid's
// faces).  Lane task t (uniform over a wave) covers one family of blocks:
// the centre (t = 0), the face neighbours along x, y, z (t = 1..3), the
// edge neighbours in the xy, xz, yz planes (t = 4..6) or the corner
// neighbours (t = 7).  Each family is eight (block, element) pairs; a pair
// adds the element's three SCVFs at the node to the block.  The pairs come
// from a uniform table in constant memory (element position, the column
// corner, the three SCVFs with their signs, shape values and reference
// gradients), so every constant is a scalar operand.
// ---------------------------------------------------------------------
struct Pair {
    int dx, dy, dz;       // block offset (column node - row node)
    int sx, sy, sz;       // element position: the element spans [v, v + s] on each axis
    int kw;               // corner of the column node in the element (original numbering)
    int first, last;      // the pair opens / closes its block
    int ip[3];            // SCVF of the element's edge at the row node along x, y, z
    float sg[3];          // +1 if the row node is the edge's `from` corner, -1 if `to`
    float sn[3];          // the normal N_q points away from the row node (+1) or toward it (-1)
    float gl[3][3];       // reference gradient of shape kw at each SCVF's ip
    float phi[3];         // shape kw at each SCVF's ip
    float phs[3];         // phi sn
    float phg[3];         // phi sg
};
struct PairTabs {
    Pair p[8][8];
};

constexpr PairTabs make_pairs()
{
    PairTabs t{};
    const int fam[8][8][3] = {
        {{0, 0, 0}},
        {{-1, 0, 0}, {1, 0, 0}},
        {{0, -1, 0}, {0, 1, 0}},
        {{0, 0, -1}, {0, 0, 1}},
        {{-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}},
        {{-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1}},
        {{0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}},
        {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}}};
    const int nblk[8] = {1, 2, 2, 2, 4, 4, 4, 8};
    for (int f = 0; f < 8; ++f) {
        int m = 0;
        for (int b = 0; b < nblk[f]; ++b) {
            const int dx = fam[f][b][0], dy = fam[f][b][1], dz = fam[f][b][2];
            const int m0 = m;
            for (int sz = -1; sz <= 1; sz += 2)
                for (int sy = -1; sy <= 1; sy += 2)
                    for (int sx = -1; sx <= 1; sx += 2) {
                        if ((dx != 0 && dx != sx) || (dy != 0 && dy != sy) || (dz != 0 && dz != sz)) continue;
                        Pair& P = t.p[f][m++];
                        P.dx = dx; P.dy = dy; P.dz = dz;
                        P.sx = sx; P.sy = sy; P.sz = sz;
                        const int px = sx < 0 ? 1 : 0, py = sy < 0 ? 1 : 0, pz = sz < 0 ? 1 : 0;
                        const int kv = corner_at(px, py, pz);
                        P.kw = corner_at(px + dx, py + dy, pz + dz);
                        const int pq[3] = {px, py, pz};
                        for (int q = 0; q < 3; ++q) {
                            const int ip = edge_at(kv, q);
                            P.ip[q] = ip;
                            P.sg[q] = T3::edge[ip][0] == kv ? 1.f : -1.f;
                            P.sn[q] = pq[q] ? -1.f : 1.f;
                            for (int r = 0; r < 3; ++r) P.gl[q][r] = float(TabH<ET_HEX>::t.gl[ip][P.kw][r]);
                            P.phi[q] = float(TabH<ET_HEX>::t.phi[ip][P.kw]);
                            P.phs[q] = P.phi[q] * P.sn[q];
                            P.phg[q] = P.phi[q] * P.sg[q];
                        }
                        P.first = 0;
                        P.last = 0;
                    }
            t.p[f][m0].first = 1;
            t.p[f][m - 1].last = 1;
        }
    }
    return t;
}

__constant__ PairTabs kPairs = make_pairs();

struct RowCtx {
    int i, j, k, li, lj;
    int64_t rb;  // first value of the node's rows
    int rl;      // row length
    int ax, ay, x0, y0, z0;
};

// LDS data of one (block, element) pair of a node row: the element's
// inverse Jacobian and SCVF normals, the records and upwind corners of its
// three SCVFs at the node.  Every index stays inside the LDS arrays also for
// elements outside the grid (their values are then unused).
struct PairData {
    float Ji[3][3];
    float N[3][3];
    float4 R[3];
    int id[3];
};

__device__ inline void load_pair(const Smem& S, const Pair& Q, const RowCtx& X, PairData& D)
{
    const int layer = X.k + (Q.sz < 0 ? -1 : 0);
    const int col = X.li + (Q.sx < 0 ? 0 : 1) + CX * (X.lj + (Q.sy < 0 ? 0 : 1));
    const int sl = layer & 1;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            D.Ji[r][c] = S.ejn[sl][3 * r + c][col];
            D.N[r][c] = S.ejn[sl][9 + 3 * r + c][col];
        }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        D.R[q] = S.rec[sl][Q.ip[q]][col];
        D.id[q] = S.idx[sl][Q.ip[q]][col];
    }
}

// Pair contributions to the open block B (navier_stokes_fv1.cpp:317-593 for
// FIELDS with one upwind object): per SCVF at the node, diffusion,
// pressure, convection through the upwind corner and continuity through
// the stabilisation's shapes.  The normal of the SCVF is sn N_q (sn = +-1
// orients it away from the node); sn and the edge sign sg are folded into
// the scalar factors, and row 3 is accumulated without the density, which
// is applied once per block when it is stored.
template <int FAST, bool LAP>
__device__ inline void add_pair(const PairData& D, const Pair& Q, float nr, float (&B)[4][4])
{
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        float g[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] = Q.gl[q][0] * D.Ji[0][c] + Q.gl[q][1] * D.Ji[1][c] + Q.gl[q][2] * D.Ji[2][c];
        const float gN = g[0] * D.N[q][0] + g[1] * D.N[q][1] + g[2] * D.N[q][2];
        float Usg, Usn;
        if constexpr (FAST == UPW_NO) {
            Usg = Q.phg[q];
            Usn = Q.phs[q];
        } else {
            const bool up = D.id[q] == Q.kw;
            Usg = up ? Q.sg[q] : 0.f;
            Usn = up ? Q.sn[q] : 0.f;
        }
        const float nrs = nr * Q.sn[q];
        const float dg = nrs * gN + Usg * D.R[q].x;
#pragma unroll
        for (int d1 = 0; d1 < 3; ++d1) {
            B[d1][d1] += dg;
            if constexpr (!LAP) {
                const float gd = nrs * g[d1];
#pragma unroll
                for (int d2 = 0; d2 < 3; ++d2) B[d1][d2] += gd * D.N[q][d2];
            }
            B[d1][3] += Q.phs[q] * D.N[q][d1];
        }
        const float Svs = D.R[q].y * Q.phs[q] + D.R[q].z * Usn;
#pragma unroll
        for (int d = 0; d < 3; ++d) B[3][d] += Svs * D.N[q][d];
        B[3][3] += (Q.sn[q] * D.R[q].w) * gN;
    }
}

// The family t of blocks of the node's rows: its eight (block, element)
// pairs in table order, each pair's LDS data loaded one pair ahead of its
// arithmetic; a block is stored (with s_a, the density of row 3 and the
// mass part on the centre block's diagonal) after its last pair.
template <int FAST, bool LAP>
__device__ inline void row_family(int t, const Smem& S, const MarchGrid& G, const RowCtx& X, const KParams& P,
                                  float nr, float* __restrict__ J)
{
    const float sa = float(P.s_a), sr = float(P.s_a * P.rho);
    const float msc = (P.parts & 2) ? float(P.s_m * P.rho) : 0.f;
    float B[4][4];
    float mvol = 0.f;
    PairData D[2];
    load_pair(S, kPairs.p[t][0], X, D[0]);
#pragma unroll 2
    for (int m = 0; m < 8; ++m) {
        const Pair& Q = kPairs.p[t][m];
        if (m < 7) load_pair(S, kPairs.p[t][m + 1], X, D[(m + 1) & 1]);
        if (Q.first) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) B[r][c] = 0.f;
            mvol = 0.f;
        }
        const int layer = X.k + (Q.sz < 0 ? -1 : 0);
        const int ci = X.i + (Q.sx < 0 ? -1 : 0), cj = X.j + (Q.sy < 0 ? -1 : 0);
        if (layer >= 0 && layer < G.nz && ci >= 0 && ci < G.nx && cj >= 0 && cj < G.ny) {
            if (t == 0) {
                const int col = X.li + (Q.sx < 0 ? 0 : 1) + CX * (X.lj + (Q.sy < 0 ? 0 : 1));
                mvol += float(S.evol[layer & 1][col]);
            }
            add_pair<FAST, LAP>(D[m & 1], Q, nr, B);
        }
        if (Q.last) {
            const int dx = Q.dx, dy = Q.dy, dz = Q.dz;
            if (X.i + dx >= 0 && X.i + dx <= G.nx && X.j + dy >= 0 && X.j + dy <= G.ny && X.k + dz >= 0 &&
                X.k + dz <= G.nz) {
                const float mass = msc * mvol;
                const int slot = (dz - X.z0) * X.ay * X.ax + (dy - X.y0) * X.ax + (dx - X.x0);
                float* base = J + X.rb + (int64_t)slot * 4;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float f = r == 3 ? sr : sa;
                    float4 v = make_float4(f * B[r][0], f * B[r][1], f * B[r][2], f * B[r][3]);
                    if (r == 0) v.x += mass;
                    if (r == 1) v.y += mass;
                    if (r == 2) v.z += mass;
                    *reinterpret_cast<float4*>(base + (int64_t)r * X.rl) = v;
                }
            }
        }
    }
}

// CSR position of a node's rows: node (i, j, k) of the lexicographic grid has
// ax ay az neighbours (ax = 2 on the x faces, 3 inside), and its 4 rows of
// 4 ax ay az values follow those of every node before it
__device__ inline RowCtx row_ctx(const MarchGrid& G, int i, int j, int k, int li, int lj)
{
    RowCtx X;
    X.i = i;
    X.j = j;
    X.k = k;
    X.li = li;
    X.lj = lj;
    X.ax = 1 + (i > 0) + (i < G.nx);
    X.ay = 1 + (j > 0) + (j < G.ny);
    const int az = 1 + (k > 0) + (k < G.nz);
    X.x0 = i > 0 ? -1 : 0;
    X.y0 = j > 0 ? -1 : 0;
    X.z0 = k > 0 ? -1 : 0;
    const int64_t sx = i > 0 ? 3 * (int64_t)i - 1 : 0;
    const int64_t sy = j > 0 ? 3 * (int64_t)j - 1 : 0;
    const int64_t sz = k > 0 ? 3 * (int64_t)k - 1 : 0;
    const int64_t SX = 3 * (int64_t)G.nx + 1, SY = 3 * (int64_t)G.ny + 1;
    X.rb = 16 * (SX * SY * sz + SX * sy * az + sx * X.ay * az);
    X.rl = 4 * X.ax * X.ay * az;
    return X;
}

// Node plane p of the tile (its (TX + 2) x (TY + 2) nodes: coordinates and
// state) staged through registers: fetch issues the global loads, store
// writes them to slot p & 1 after the plane that occupied it is done.
struct Stage {
    double v[NSTAGE];
};

__device__ inline void fetch_plane(Stage& r, const MarchGrid& G, const double* __restrict__ xyz,
                                   const double* __restrict__ u, int i0, int j0, int p)
{
#pragma unroll
    for (int q = 0; q < NSTAGE; ++q) {
        const int it = (int)threadIdx.x + q * NT;
        double x = 0.0;
        if (it < 7 * NN && p <= G.nz) {
            const int comp = it / NN, ln = it % NN;
            const int i = i0 - 1 + ln % NNX, j = j0 - 1 + ln / NNX;
            if (i >= 0 && i <= G.nx && j >= 0 && j <= G.ny) {
                const int64_t node = i + G.NX * j + G.NXY * p;
                x = comp < 3 ? xyz[node * 3 + comp] : u[node * 4 + comp - 3];
            }
        }
        r.v[q] = x;
    }
}

__device__ inline void store_plane(Smem& S, const Stage& r, int p)
{
    const int sl = p & 1;
#pragma unroll
    for (int q = 0; q < NSTAGE; ++q) {
        const int it = (int)threadIdx.x + q * NT;
        if (it < 7 * NN) {
            const int comp = it / NN, ln = it % NN;
            if (comp < 3) S.ncx[sl][comp][ln] = r.v[q];
            else S.ncu[sl][comp - 3][ln] = r.v[q];
        }
    }
}

template <int FAST, bool LAP>
__global__ void __launch_bounds__(NT, 2)
march_hex(MarchGrid G, const double* __restrict__ xyz, KParams P, const double* __restrict__ u,
          const double* __restrict__ uold, uint8_t* __restrict__ dec, float* __restrict__ J, double* __restrict__ dout,
          int* __restrict__ err)
{
    __shared__ Smem S;
    // the wave index is uniform: in an SGPR, the per-wave tables (edge of
    // phase 1, block family of phase 2) are scalar loads and offsets
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = (int)(threadIdx.x & 63);
    const float nr = float(-1.0 * P.nu * P.rho);
    const int64_t per = (int64_t)G.tiles_x * G.tiles_y;
    const int li = lane % TX, lj = lane / TX;
    for (int64_t unit = blockIdx.x; unit < G.units; unit += gridDim.x) {
        const int tz = (int)(unit / per);
        const int rem = (int)(unit % per);
        const int i0 = (rem % G.tiles_x) * TX, j0 = (rem / G.tiles_x) * TY;
        const int kz0 = tz * G.zc, kz1 = min(kz0 + G.zc, G.nz + 1);
        const int ni = i0 + li, nj = j0 + lj;
        const bool node = lane < TX * TY && ni <= G.nx && nj <= G.ny;
        // defect: wave 0 owns the tile's node columns and carries the part of
        // the next plane's rows that the current layer contributes
        const bool dlane = dout != nullptr && wave == 0 && node;
        double held[4] = {0.0, 0.0, 0.0, 0.0};
        // the first step of a unit above the bottom only computes the layer
        // below its first plane (for that plane's rows)
        const int kstart = kz0 > 0 ? kz0 - 1 : kz0;
        Stage st;
        fetch_plane(st, G, xyz, u, i0, j0, kstart);
        store_plane(S, st, kstart);
        fetch_plane(st, G, xyz, u, i0, j0, kstart + 1);
        const int lnv = (li + 1) + NNX * (lj + 1);  // the lane's node in the staged planes
        for (int k = kstart; k < kz1; ++k) {
            const bool emit = k >= kz0;
            // plane k + 1 goes to the slot of plane k - 1; plane k + 2 is
            // fetched while this layer is processed
            store_plane(S, st, k + 1);
            fetch_plane(st, G, xyz, u, i0, j0, k + 2);
            __syncthreads();
            if (k < G.nz) layer_scvfs<FAST>(S, G, xyz, P, u, uold, dec, err, k, i0, j0, wave, lane);
            __syncthreads();
            if (dlane) {
                const int64_t v = ni + G.NX * nj + G.NXY * k;
                double uv[3];
                if (emit) {
                    double acc[4] = {held[0], held[1], held[2], held[3]};
#pragma unroll
                    for (int d = 0; d < 3; ++d) uv[d] = S.ncu[k & 1][d][lnv];
                    if (k < G.nz) node_layer<0>(S, G, P, k, ni, nj, li, lj, uv, acc);
                    *reinterpret_cast<double2*>(dout + v * 4) = make_double2(acc[0], acc[1]);
                    *reinterpret_cast<double2*>(dout + v * 4 + 2) = make_double2(acc[2], acc[3]);
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) held[c] = 0.0;
                if (k < G.nz) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) uv[d] = S.ncu[(k + 1) & 1][d][lnv];
                    node_layer<1>(S, G, P, k, ni, nj, li, lj, uv, held);
                }
            }
            if (emit && J != nullptr && node) {
                const RowCtx X = row_ctx(G, ni, nj, k, li, lj);
                for (int pass = 0; pass < 2; ++pass) row_family<FAST, LAP>(wave + 4 * pass, S, G, X, P, nr, J);
            }
            __syncthreads();
        }
    }
}

}  // namespace march

hipError_t launch_assemble_march(int fast, const MarchGrid& g, const double* xyz, const KParams& p, const double* u,
                                 const double* uold, uint8_t* upw_dec, float* J, double* d, int* err,
                                 hipStream_t stream)
{
    if (g.units <= 0 || g.units > 0x7fffffff) return hipErrorInvalidValue;
    const dim3 grid((unsigned)g.units), block(march::NT);
    const bool lap = p.laplace != 0;
    switch (fast * 2 + (lap ? 1 : 0)) {
    case UPW_NO * 2:
        march::march_hex<UPW_NO, false><<<grid, block, 0, stream>>>(g, xyz, p, u, uold, upw_dec, J, d, err);
        break;
    case UPW_NO * 2 + 1:
        march::march_hex<UPW_NO, true><<<grid, block, 0, stream>>>(g, xyz, p, u, uold, upw_dec, J, d, err);
        break;
    case UPW_FULL * 2:
        march::march_hex<UPW_FULL, false><<<grid, block, 0, stream>>>(g, xyz, p, u, uold, upw_dec, J, d, err);
        break;
    case UPW_FULL * 2 + 1:
        march::march_hex<UPW_FULL, true><<<grid, block, 0, stream>>>(g, xyz, p, u, uold, upw_dec, J, d, err);
        break;
    case UPW_SKEWED * 2:
        march::march_hex<UPW_SKEWED, false><<<grid, block, 0, stream>>>(g, xyz, p, u, uold, upw_dec, J, d, err);
        break;
    case UPW_SKEWED * 2 + 1:
        march::march_hex<UPW_SKEWED, true><<<grid, block, 0, stream>>>(g, xyz, p, u, uold, upw_dec, J, d, err);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace nsfv1
