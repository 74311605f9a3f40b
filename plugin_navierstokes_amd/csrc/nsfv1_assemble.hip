// nsfv1_assemble.hip -- row-owner gather assembly of the FV1 incompressible
// Navier-Stokes defect and Jacobian (gfx950).
//
// Work decomposition.  Every SCVF of an element adds its flux to exactly two
// rows (from, to: navier_stokes_fv1.cpp:342-343, 713-714).  The row of node
// v therefore only needs, for each incident element e, the SCVFs of e that
// touch v.  One lane handles one (v, e) pair: it forms those SCVFs (3 for
// hexahedra and tetrahedra, 2 in 2D), evaluates the upwind, the FIELDS
// stabilisation and the fixpoint linearisation for them, and holds e's
// contribution to v's block row ((dim+1) x nsh x (dim+1) Jacobian values
// and dim+1 defect values) in registers.
//
// Canonical frame.  The lane loads e's corners permuted by a symmetry of the
// reference element that puts v at reference corner 0, so a single code
// path, with every SCVF index, shape value and reference gradient a
// compile-time constant, serves every lane, and no register array is
// indexed dynamically.  Normals point away from v, so all row contributions
// carry the same sign.  The side/ray intersection of the skewed upwinds runs
// on the original corner order and side order of the reference element.
//
// A workgroup owns consecutive nodes (a contiguous range of CSR values that
// fits 64 KiB of LDS).  Lanes add their rows into LDS in rounds ordered by
// the incidence slot (no two lanes of one round share a row), so the sums
// are bitwise reproducible without atomics; finished rows leave with
// coalesced stores.  Static geometry (diffusion lengths per SCVF, SCV
// volumes) comes from a per-mesh cache built by geometry_cache<>.
//
// Decisions (flux sign, zero-velocity guard, cut side, nearest corner)
// agree with the CPU oracle except at exact ties, whose outcome depends on
// the rounding of the canonical-order sums.
#include "nsfv1_geom.hpp"
#include "nsfv1_internal.hpp"
#include "nsfv1_ray.hpp"
#include "nsfv1_topo.hpp"

namespace nsfv1 {

// ---------------------------------------------------------------------
// per-mesh static geometry: SCV volumes and the diffusion-length factor
// 1/L^2 of every SCVF (diffusion_length.h:48-199 -- geometry only)
// ---------------------------------------------------------------------
template <int ET, int IP>
__device__ inline void cache_ips(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* c, double* nn,
                                 double* ds)
{
    if constexpr (IP < Topo<ET>::nip) {
        constexpr int D = Topo<ET>::dim;
        double ipg[3], n[3], me[3];
        scvf_frame<ET, IP>(co, c, ipg, n, me);
        nn[IP] = dot_nc<D>(n, n);
        ds[IP] = D == 3 ? dist2_nc<D>(me, c) : 0.0;
        cache_ips<ET, IP + 1>(co, c, nn, ds);
    }
}

template <int ET>
__global__ void __launch_bounds__(256)
geometry_cache(const double* __restrict__ xyz, const int32_t* __restrict__ conn, int64_t ne, int difflen,
               double* __restrict__ l2inv, double* __restrict__ volc, uint8_t* __restrict__ affine,
               int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NIP = T::nip;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double co[N][D];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int64_t v = conn[e * N + k];
#pragma unroll
        for (int i = 0; i < D; ++i) co[k][i] = xyz[v * D + i];
    }
    double vol[N];
    if constexpr (ET == ET_PRISM) scv_volumes_prism(co, vol);
    else scv_volumes<ET>(co, vol);
#pragma unroll
    for (int k = 0; k < N; ++k) volc[e * N + k] = vol[k];
    {
        // parallelogram / parallelepiped test (simplices are always affine;
        // prisms take the general ray test)
        bool aff = ET != ET_PRISM;
        if constexpr (ET == 1 || ET == 3) {
            double scale = 0.0, dev = 0.0;
#pragma unroll
            for (int k = 1; k < N; ++k)
#pragma unroll
                for (int i = 0; i < D; ++i) scale = fmax(scale, fabs(co[k][i] - co[0][i]));
#pragma unroll
            for (int i = 0; i < D; ++i) {
                dev = fmax(dev, fabs(co[1][i] + co[3][i] - co[0][i] - co[2][i]));
                if constexpr (ET == 3) {
                    dev = fmax(dev, fabs(co[1][i] + co[4][i] - co[0][i] - co[5][i]));
                    dev = fmax(dev, fabs(co[3][i] + co[4][i] - co[0][i] - co[7][i]));
                    dev = fmax(dev, fabs(co[1][i] + co[3][i] + co[4][i] - 2.0 * co[0][i] - co[6][i]));
                }
            }
            aff = dev <= 1e-12 * scale;
        }
        affine[e] = aff ? 1 : 0;
    }
    double c[D];
    elem_centre<ET>(co, c);
    double nn[NIP], ds[NIP];
    cache_ips<ET, 0>(co, c, nn, ds);
    double minN = NSFV1_DBL_MAX, minD = NSFV1_DBL_MAX, avgN = 0.0;
#pragma unroll
    for (int ip = 0; ip < NIP; ++ip) {
        minN = nn[ip] < minN ? nn[ip] : minN;
        avgN += nn[ip];
        minD = ds[ip] < minD ? ds[ip] : minD;
    }
    avgN /= NIP;
    bool ok = true;
#pragma unroll
    for (int ip = 0; ip < NIP; ++ip) {
        double a2 = 0.5 * (vol[T::edge[ip][0]] + vol[T::edge[ip][1]]);
        a2 *= a2;
        double L;
        if (difflen == DL_RAW)
            L = (D == 2) ? 1. / (0.5 * a2 / nn[ip] + 3.0 * nn[ip] / 8.) : 1. / (0.5 * a2 / nn[ip] + 3.0 * ds[ip] / 8.);
        else if (difflen == DL_FIVEPOINT)
            L = (D == 2) ? 2.0 * nn[ip] / a2 + 8.0 / nn[ip] : 2.0 * nn[ip] / a2 + 8.0 * ds[ip] / nn[ip];
        else
            L = (D == 2) ? 2.0 * minN / a2 + 8.0 / (3.0 * avgN) : 2.0 * minN / a2 + 8.0 * minD / (3.0 * avgN);
        ok &= (nn[ip] > 0.0) && (a2 > 0.0);
        l2inv[e * NIP + ip] = L;
    }
    if (!ok) kernel_error(err, KERR_DEGENERATE);
}

// ---------------------------------------------------------------------
// canonical frame tables (lane-varying lookups -> constant memory)
// ---------------------------------------------------------------------
template <int ET>
struct Canon {
    int perm[8][8];   // canonical corner k' -> original corner
    int oip[8][3];    // original SCVF of the q-th canonical SCVF at corner 0
    int from[8][3];   // 1 if kv is the `from` corner of that SCVF
    int A[8][3][3];   // original reference coordinates = A * canonical ones + corner offset
    int Ainv[8][3][3];
    int kq[12];       // canonical index of SCVF ip in the frame of its `from` corner
};

template <int ET>
constexpr Canon<ET> make_canon()
{
    Canon<ET> c{};
    for (int kv = 0; kv < Topo<ET>::nsh; ++kv) {
        for (int k = 0; k < Topo<ET>::nsh; ++k) c.perm[kv][k] = canon_perm<ET>(kv, k);
        for (int q = 0; q < Topo<ET>::cip; ++q) {
            c.oip[kv][q] = canon_orig_ip<ET>(kv, q);
            c.from[kv][q] = canon_is_from<ET>(kv, q) ? 1 : 0;
        }
        // reference-coordinate map of the frame: columns are the images of
        // the canonical unit corners
        constexpr bool box = (ET == 1 || ET == 3);
        constexpr int D = Topo<ET>::dim;
        const int o = canon_perm<ET>(kv, 0);
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) c.A[kv][i][j] = 0;
        for (int j = 0; j < D; ++j) {
            const int kj = box ? (j == 0 ? 1 : (j == 1 ? 3 : 4)) : j + 1;
            const int t = canon_perm<ET>(kv, kj);
            for (int i = 0; i < D; ++i) c.A[kv][i][j] = (int)Topo<ET>::lc[t][i] - (int)Topo<ET>::lc[o][i];
        }
        int (&a)[3][3] = c.A[kv];
        int (&b)[3][3] = c.Ainv[kv];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) b[i][j] = 0;
        if (D == 2) {
            const int det = a[0][0] * a[1][1] - a[0][1] * a[1][0];
            b[0][0] = a[1][1] * det; b[0][1] = -a[0][1] * det;
            b[1][0] = -a[1][0] * det; b[1][1] = a[0][0] * det;
        } else {
            const int det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                            a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                            a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);  // +-1
            b[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * det;
            b[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * det;
            b[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * det;
            b[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * det;
            b[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * det;
            b[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * det;
            b[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * det;
            b[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * det;
            b[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * det;
        }
    }
    for (int ip = 0; ip < 12; ++ip) c.kq[ip] = 0;
    for (int ip = 0; ip < Topo<ET>::nip; ++ip) {
        const int kv = Topo<ET>::edge[ip][0];
        for (int q = 0; q < Topo<ET>::cip; ++q)
            if (c.oip[kv][q] == ip) c.kq[ip] = q;
    }
    return c;
}

template <int ET>
struct CanonH {
    static constexpr Canon<ET> c = make_canon<ET>();
};

// configuration: FAST >= 0 is the common case (FIELDS stabilisation, one
// upwind object of type FAST for stabilisation and convection, no Stokes,
// PAC, Peclet blend or exact Jacobian) compiled with those switches fixed;
// FAST < 0 reads every switch at run time.
template <int FAST>
struct Cfg {
    const KParams& P;
    double nu;  // kinematic viscosity of the lane's element (batched instances carry their own)
    // the element's rows of the imports at the integration points
    // (navier_stokes_fv1.cpp:188-247), or null: [nip] viscosity, [nip][dim]
    // source at the SCVF ips, indexed by the original SCVF (= edge) number
    const double* nu_ip = nullptr;
    const double* f_scvf = nullptr;
    const double* rho_ip = nullptr;  // [nip] density at the SCVF ips, or null
    __device__ bool stokes() const { return FAST >= 0 ? false : P.stokes != 0; }
    __device__ bool pac() const { return FAST >= 0 ? false : P.pac != 0; }
    __device__ bool peclet() const { return FAST >= 0 ? false : P.peclet != 0; }
    __device__ bool exact() const { return FAST >= 0 ? false : P.exact != 0.0; }
    __device__ bool fields() const { return FAST >= 0 ? true : P.stab_type == STAB_FIELDS; }
    __device__ int stab_up() const { return FAST >= 0 ? FAST : P.stab_upwind; }
    __device__ int conv_up() const { return FAST >= 0 ? FAST : P.conv_upwind; }
    __device__ bool two_upwinds() const { return FAST >= 0 ? false : (!P.pac && P.conv_upwind != P.stab_upwind); }
};

// compact per-SCVF data that the Jacobian expansion needs (precision JT)
template <int ET, typename JT>
struct Rec {
    static constexpr int D = Topo<ET>::dim, N = Topo<ET>::nsh;
    JT n[D], Ji[D][D], pw, pw1, vv, vc, sp, w, up[D];
    JT nr;  // -nu rho at the SCVF
    JT Us[N], Uc[N];
};

template <int ET, typename JT>
struct Lane {
    static constexpr int D = Topo<ET>::dim, N = Topo<ET>::nsh, NC = D + 1;
    JT jacc[NC][N][NC];
    double dacc[NC];
};

// upwind shapes in canonical indexing.  `pk` maps canonical -> original
// corners; the ray cast runs on the original corners (`coo`) in original
// side order (upwind.cpp:337-430, 505-575).
template <int ET, int Q>
__device__ inline bool upwind_canon(int type, bool kvfrom, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                    const double* __restrict__ xyz, const int32_t* __restrict__ econn, const int* pk,
                                    const double* ipg, const double* n, const double* sv,
                                    double (&U)[Topo<ET>::nsh], double& convlen, bool& have_cut, int& side,
                                    double* gi, double* li, uint8_t* decp, int mode, bool affine)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh;
    constexpr int CIP = corner_ip<ET>(0, Q);
    constexpr int OT = T::edge[CIP][0] == 0 ? T::edge[CIP][1] : T::edge[CIP][0];  // other end of the edge
    if (type == UPW_NO) {
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = TabH<ET>::t.phi[CIP][k];
        convlen = 1.0;
        return true;
    }
    if (type == UPW_FULL) {
        // n points away from canonical corner 0 (= kv): the reference picks
        // `from` iff n_(from->to) . v > 0 (upwind.cpp:150-166)
        const double fl = dot_nc<D>(n, sv);
        const bool pick0 = kvfrom ? (fl > 0.0) : !(fl < 0.0);
#pragma unroll
        for (int k = 0; k < N; ++k) U[k] = (k == 0) ? (pick0 ? 1.0 : 0.0) : (k == OT ? (pick0 ? 0.0 : 1.0) : 0.0);
        double cc[D];
#pragma unroll
        for (int i = 0; i < D; ++i) cc[i] = pick0 ? co[0][i] : co[OT][i];
        convlen = sqrt(dist2_nc<D>(ipg, cc));
        return true;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) U[k] = 0.0;
    if (type == UPW_SKEWED && mode == 2 && decp) {
        // replay a recorded decision (no ray cast)
        const int best = *decp;
        if (best == 255) {
            convlen = 1.0;
            return true;
        }
        double cc[D];
#pragma unroll
        for (int i = 0; i < D; ++i) cc[i] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const bool on = (pk[k] == best);
            U[k] = on ? 1.0 : 0.0;
#pragma unroll
            for (int i = 0; i < D; ++i) cc[i] = on ? co[k][i] : cc[i];
        }
        convlen = sqrt(dist2_nc<D>(ipg, cc));
        return true;
    }
    if (sqrt(dot_nc<D>(sv, sv)) < 1e-14) {
        if (type == UPW_SKEWED && mode == 1 && decp && kvfrom) *decp = 255;
        convlen = 1.0;
        return true;
    }
    // the element's corners in original order, for the ray cast
    double coo[N][D];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int64_t on = econn[k];
#pragma unroll
        for (int i = 0; i < D; ++i) coo[k][i] = xyz[on * D + i];
    }
    if (!have_cut) {
        side = -1;
        // the skewed upwind only takes a decision from the cut, so the
        // cheaper reference-coordinate cast serves affine elements; the LPS
        // shapes are continuous functions of the cut point, which is
        // ill-conditioned for rays nearly parallel to a side, so LPS uses the
        // reference's own intersection formulas (same rounding as the oracle)
        if (affine && type == UPW_SKEWED) ray_affine<ET>(coo, ipg, sv, side, gi, li);
        if (side < 0) side_loop<ET, 0>(coo, ipg, sv, side, gi, li);
        have_cut = true;
    }
    if (side < 0) return false;
    if (type == UPW_SKEWED) {
        double mn = NSFV1_DBL_MAX;
        int best = 0;
#pragma unroll
        for (int s = 0; s < T::nside; ++s) {
            if (s != side) continue;
#pragma unroll
            for (int j = 0; j < T::side_n[s]; ++j) {
                const double dd = dist2_nc<D>(gi, coo[T::side[s][j]]);
                if (dd < mn) { mn = dd; best = T::side[s][j]; }
            }
        }
        if (mode == 1 && decp && kvfrom) *decp = (uint8_t)best;
        double cc[D];
#pragma unroll
        for (int i = 0; i < D; ++i) cc[i] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const bool on = (pk[k] == best);
            U[k] = on ? 1.0 : 0.0;
#pragma unroll
            for (int i = 0; i < D; ++i) cc[i] = on ? co[k][i] : cc[i];
        }
        convlen = sqrt(dist2_nc<D>(ipg, cc));
    } else {
#pragma clang fp contract(off)
        double val[N];
#pragma unroll
        for (int c = 0; c < N; ++c) {
            bool on = false;
#pragma unroll
            for (int s = 0; s < T::nside; ++s) on |= (s == side) && side_has<ET>(s, c);
            double v;
            if constexpr (ET == 0) v = c == 0 ? 1.0 - li[0] - li[1] : (c == 1 ? li[0] : li[1]);
            else if constexpr (ET == 1) {
                const double a = li[0], b = li[1];
                v = c == 0 ? (1 - a) * (1 - b) : (c == 1 ? a * (1 - b) : (c == 2 ? a * b : (1 - a) * b));
            } else if constexpr (ET == 2) v = c == 0 ? 1.0 - li[0] - li[1] - li[2] : li[c - 1];
            else if constexpr (ET == ET_PRISM) {
                const int t = c % 3;
                const double lam = t == 0 ? 1.0 - li[0] - li[1] : (t == 1 ? li[0] : li[1]);
                v = lam * (c < 3 ? 1.0 - li[2] : li[2]);
            } else {
                const double f0 = T::lc[c][0] > 0.5 ? li[0] : 1.0 - li[0];
                const double f1 = T::lc[c][1] > 0.5 ? li[1] : 1.0 - li[1];
                const double f2 = T::lc[c][2] > 0.5 ? li[2] : 1.0 - li[2];
                v = f0 * f1 * f2;
            }
            val[c] = on ? v : 0.0;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double x = 0.0;
#pragma unroll
            for (int c = 0; c < N; ++c) x = (pk[k] == c) ? val[c] : x;
            U[k] = x;
        }
        convlen = sqrt(dist2_nc<D>(ipg, gi));
    }
    return true;
}

// original number of the edge between original corners a and b
template <int ET>
__device__ inline int edge_between(int a, int b)
{
    int r = 0;
#pragma unroll
    for (int e = 0; e < Topo<ET>::nip; ++e)
        if ((Topo<ET>::edge[e][0] == a && Topo<ET>::edge[e][1] == b) ||
            (Topo<ET>::edge[e][0] == b && Topo<ET>::edge[e][1] == a))
            r = e;
    return r;
}

// scvf_frame of a runtime SCVF number (uniform over a wave where it is used)
template <int ET, int E = 0>
__device__ inline void scvf_frame_rt(int ip, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* c,
                                     double* ipg, double* nn, double* me)
{
    if constexpr (E < Topo<ET>::nip) {
        if (ip == E) scvf_frame<ET, E>(co, c, ipg, nn, me);
        else scvf_frame_rt<ET, E + 1>(ip, co, c, ipg, nn, me);
    }
}

// FP64 part of canonical SCVF Q at corner 0 (= the row node): geometry,
// upwind, stabilisation, defect; emits the record for the Jacobian.
// ORIGF: the SCVF's ip, normal and edge midpoint are formed from the corners
// in the element's own order (the reference's sums: element centre, face
// centres), so they do not depend on the frame; the records of the split
// path, which serve both rows of an SCVF, take this form for tetrahedra.
template <int ET, int FAST, int Q, typename JT, bool ORIGF = false>
__device__ inline void scvf_canon(const Cfg<FAST>& C, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                  const double* __restrict__ xyz, const int32_t* __restrict__ econn,
                                  const double (&ul)[Topo<ET>::nsh][Topo<ET>::dim + 1], const int* pk,
                                  const int32_t* ccn, const double* c, const double* __restrict__ uold,
                                  double l2, bool kvfrom, bool wantJ, Lane<ET, JT>& L, Rec<ET, JT>& R,
                                  bool& okg, bool& oku, bool& have_cut, int& side, double* gi, double* li,
                                  uint8_t* decp, bool affine, double* sv_out = nullptr)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh;
    constexpr int IP = corner_ip<ET>(0, Q);
    constexpr int OT = T::edge[IP][0] == 0 ? T::edge[IP][1] : T::edge[IP][0];
    const KParams& P = C.P;
    double nu = C.nu;
    double rho = P.rho, inv_rho = P.inv_rho;
    const double* fip = nullptr;  // source at this SCVF's ip (import), or the constant
    if (C.nu_ip || C.f_scvf || C.rho_ip) {
        const int oip = edge_between<ET>(pk[0], pk[OT]);
        if (C.nu_ip) nu = C.nu_ip[oip];
        if (C.f_scvf) fip = C.f_scvf + oip * D;
        if (C.rho_ip) {
            rho = C.rho_ip[oip];
            inv_rho = 1.0 / rho;
        }
    }

    double ipg[D], n[D], me[D];
    {
        // frame of canonical SCVF IP with the normal pointing away from
        // corner 0 (the row node): row contributions are then all `+`
        double nn_[D];
        if constexpr (ORIGF) {
            double cor[N][D], cc[D];
#pragma unroll
            for (int k = 0; k < N; ++k)
#pragma unroll
                for (int i = 0; i < D; ++i) cor[k][i] = xyz[(int64_t)econn[k] * D + i];
            elem_centre<ET>(cor, cc);
            scvf_frame_rt<ET>(edge_between<ET>(pk[0], pk[OT]), cor, cc, ipg, nn_, me);
        } else {
            scvf_frame<ET, IP>(co, c, ipg, nn_, me);
        }
        double e[D];
#pragma unroll
        for (int i = 0; i < D; ++i) e[i] = co[OT][i] - co[0][i];
        const double s = dot_nc<D>(nn_, e);
#pragma unroll
        for (int i = 0; i < D; ++i) n[i] = s < 0 ? -nn_[i] : nn_[i];
    }
    double Ji[D][D];
    okg &= inv_jacobian<ET, IP>(co, Ji);
    const double sgn = P.s_a;

    double sv[D];
    {
#pragma unroll
        for (int d = 0; d < D; ++d) sv[d] = 0.0;
#pragma unroll
        for (int sh = 0; sh < N; ++sh)
#pragma unroll
            for (int d = 0; d < D; ++d) sv[d] += ul[sh][d] * TabH<ET>::t.phi[IP][sh];
    }
    double gv[D + 1][D];
    {
        double gh[D + 1][D];
#pragma unroll
        for (int a = 0; a <= D; ++a)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double acc = 0;
#pragma unroll
                for (int k = 0; k < N; ++k) acc += ul[k][a] * TabH<ET>::t.gl[IP][k][j];
                gh[a][j] = acc;
            }
#pragma unroll
        for (int a = 0; a <= D; ++a)
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) acc += gh[a][j] * Ji[j][i];
                gv[a][i] = acc;
            }
    }
    double pr = 0;
#pragma unroll
    for (int sh = 0; sh < N; ++sh) pr += TabH<ET>::t.phi[IP][sh] * ul[sh][D];

    // upwinds (the stabilisation's and the convective term's)
    double Ls = 1.0, upS[D], upC[D];
    {
        double Us[N];
#pragma unroll
        for (int k = 0; k < N; ++k) Us[k] = 0.0;
        if (!C.stokes()) oku &= upwind_canon<ET, Q>(C.stab_up(), kvfrom, co, xyz, econn, pk, ipg, n, sv, Us, Ls, have_cut, side, gi, li, decp, P.upw_mode, affine);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < N; ++k) acc += Us[k] * ul[k][d];
            upS[d] = acc;
            upC[d] = acc;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) { R.Us[k] = JT(Us[k]); R.Uc[k] = JT(Us[k]); }
        if (!C.stokes() && C.two_upwinds()) {
            double Uc[N], Lc;
            oku &= upwind_canon<ET, Q>(C.conv_up(), kvfrom, co, xyz, econn, pk, ipg, n, sv, Uc, Lc, have_cut, side, gi, li, decp, P.upw_mode, affine);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double acc = 0;
#pragma unroll
                for (int k = 0; k < N; ++k) acc += Uc[k] * ul[k][d];
                upC[d] = acc;
            }
#pragma unroll
            for (int k = 0; k < N; ++k) R.Uc[k] = JT(Uc[k]);
        }
    }

    // stabilisation: stabilization.cpp:122-242 (FIELDS), 805-850 (without)
    double svel[D], vv = 0.0, vc = 0.0, idg = 0.0;
    if (C.fields()) {
        vv = nu * l2;
        vc = C.stokes() ? 0.0 : sqrt(dot_nc<D>(sv, sv)) / Ls;
        double diag = vv;
        if (P.has_old) diag += P.inv_dt;
        if (!C.stokes()) diag += vc;
        idg = 1.0 / diag;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double rhs = P.has_source ? (fip ? fip[d] : P.f[d]) : 0.0;
            if (P.has_old) {
                double o = 0.0;
#pragma unroll
                for (int sh = 0; sh < N; ++sh) o += TabH<ET>::t.phi[IP][sh] * uold[(int64_t)ccn[sh] * (D + 1) + d];
                rhs += o * P.inv_dt;
            }
            rhs += vv * sv[d] + vc * upS[d] - gv[D][d] * inv_rho;
            svel[d] = rhs * idg;
        }
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) svel[d] = sv[d];
    }
    if (sv_out)
#pragma unroll
        for (int d = 0; d < D; ++d) sv_out[d] = svel[d];

    // defect: navier_stokes_fv1.cpp:667-777
#pragma unroll
    for (int d1 = 0; d1 < D; ++d1) {
        double acc = 0;
#pragma unroll
        for (int d2 = 0; d2 < D; ++d2) acc += gv[d1][d2] * n[d2];
        if (!P.laplace)
#pragma unroll
            for (int d2 = 0; d2 < D; ++d2) acc += gv[d2][d1] * n[d2];
        L.dacc[d1] += sgn * (acc * (-1.0 * nu * rho) + pr * n[d1]);
    }
    double prod = 0.0, wpe = 1.0, upv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) upv[d] = 0.0;
    if (!C.stokes()) {
#pragma unroll
        for (int d = 0; d < D; ++d) upv[d] = C.pac() ? svel[d] : upC[d];
        if (C.peclet()) {
            // peclet_blend: navier_stokes_fv1.cpp:871-892
            double nn = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) nn += n[i] * n[i];
            const double Pe = dot_nc<D>(sv, n) / nn * sqrt(dist2_nc<D>(co[OT], co[0])) / nu;
            wpe = Pe * Pe / (5.0 + Pe * Pe);
#pragma unroll
            for (int d = 0; d < D; ++d) upv[d] = wpe * upv[d] + (1.0 - wpe) * sv[d];
        }
        prod = dot_nc<D>(sv, n) * rho;
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1) L.dacc[d1] += sgn * (upv[d1] * prod);
    }
    L.dacc[D] += sgn * (dot_nc<D>(svel, n) * rho);

    if (!wantJ) return;
    R.nr = JT(-1.0 * nu * rho);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        R.n[i] = JT(n[i]);
        R.up[i] = JT(upv[i]);
#pragma unroll
        for (int j = 0; j < D; ++j) R.Ji[i][j] = JT(Ji[i][j]);
    }
    R.pw = JT(prod * wpe);
    R.pw1 = JT(prod * (1.0 - wpe));
    R.vv = JT(vv * idg);
    R.vc = JT(vc * idg);
    R.sp = JT(-inv_rho * idg);
    R.w = JT(wpe);
}

// Jacobian row contributions of canonical SCVF Q from its record:
// navier_stokes_fv1.cpp:317-593
template <int ET, int FAST, int Q, typename JT>
__device__ inline void scvf_jac(const Cfg<FAST>& C, const Rec<ET, JT>& R, Lane<ET, JT>& L,
                                const JT (*G)[Topo<ET>::dim] = nullptr, int kv = 0)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh;
    constexpr int IP = corner_ip<ET>(0, Q);
    const KParams& P = C.P;
    const bool fields = C.fields();
    // density at the SCVF ip: the import, read again here so that the
    // records carry no extra register, or the constant
    const double rd = C.rho_ip ? C.rho_ip[CanonH<ET>::c.oip[kv][Q]] : P.rho;
    const JT s_ = JT(P.s_a), rh = JT(rd), nr = R.nr, ex = JT(P.exact * rd);
#pragma unroll
    for (int sh = 0; sh < N; ++sh) {
        JT g[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (G) {
                g[i] = G[sh][i];
            } else {
                JT acc = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) acc += R.Ji[j][i] * JT(TabH<ET>::t.gl[IP][sh][j]);
                g[i] = acc;
            }
        }
        const JT ph = JT(TabH<ET>::t.phi[IP][sh]);
        const JT Sv = fields ? R.vv * ph + R.vc * R.Us[sh] : ph;
        JT Sp[D];
#pragma unroll
        for (int d = 0; d < D; ++d) Sp[d] = fields ? g[d] * R.sp : JT(0);
        JT gn = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) gn += g[i] * R.n[i];
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1) {
            L.jacc[d1][sh][d1] += s_ * (nr * gn);
            if (!P.laplace)
#pragma unroll
                for (int d2 = 0; d2 < D; ++d2) L.jacc[d1][sh][d2] += s_ * (nr * g[d1] * R.n[d2]);
            L.jacc[d1][sh][D] += s_ * (ph * R.n[d1]);
        }
        if (!C.stokes()) {
            if (C.pac()) {
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) {
                    L.jacc[d1][sh][d1] += s_ * (R.pw * Sv);
                    L.jacc[d1][sh][D] += s_ * (R.pw * Sp[d1]);
                }
            } else {
                const JT cv = R.Uc[sh] * R.pw;
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) L.jacc[d1][sh][d1] += s_ * cv;
            }
            if (C.peclet())
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) L.jacc[d1][sh][d1] += s_ * (R.pw1 * ph);
            if (C.exact()) {
                if (C.pac()) {
                    JT pp = 0;
#pragma unroll
                    for (int k = 0; k < D; ++k) pp += Sp[k] * R.n[k];
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) {
                        const JT pv = Sv * R.n[d1] * ex;
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) L.jacc[d1][sh][d2] += s_ * (pv * R.up[d1]);
                        L.jacc[d1][sh][D] += s_ * (pp * ex * R.up[d1]);
                    }
                } else {
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) L.jacc[d1][sh][d2] += s_ * (R.w * R.Uc[sh] * R.n[d2] * rh * R.up[d1]);
                }
                if (C.peclet())
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2)
                            L.jacc[d1][sh][d2] += s_ * (R.up[d1] * (JT(1) - R.w) * ph * R.n[d2] * rh);
            }
        }
        JT cpp = 0;
#pragma unroll
        for (int d1 = 0; d1 < D; ++d1) {
            L.jacc[D][sh][d1] += s_ * (Sv * R.n[d1] * rh);
            cpp += Sp[d1] * R.n[d1] * rh;
        }
        L.jacc[D][sh][D] += s_ * cpp;
    }
}

template <int ET, int FAST, int Q, typename JT>
__device__ inline void scvf_all(const Cfg<FAST>& C, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                const double* __restrict__ xyz, const int32_t* __restrict__ econn,
                                const double (&ul)[Topo<ET>::nsh][Topo<ET>::dim + 1], const int* pk,
                                const int32_t* ccn, const double* c, const double* uold, const double* l2row,
                                uint8_t* decrow, int kv, bool wantJ, Lane<ET, JT>& L, Rec<ET, JT>* R, bool& okg,
                                bool& oku, bool affine)
{
    if constexpr (Q < Topo<ET>::cip) {
        bool have_cut = false;
        int side = -1;
        double gi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
        const int oip = CanonH<ET>::c.oip[kv][Q];
        const double l2 = l2row[oip];
        const bool kvfrom = CanonH<ET>::c.from[kv][Q] != 0;
        scvf_canon<ET, FAST, Q, JT>(C, co, xyz, econn, ul, pk, ccn, c, uold, l2, kvfrom, wantJ, L, R[Q], okg, oku, have_cut,
                                    side, gi, li, decrow ? decrow + oip : nullptr, affine);
        scvf_all<ET, FAST, Q + 1, JT>(C, co, xyz, econn, ul, pk, ccn, c, uold, l2row, decrow, kv, wantJ, L, R, okg, oku,
                                      affine);
    }
}

template <int ET, int FAST, int Q, typename JT>
__device__ inline void jac_all(const Cfg<FAST>& C, const Rec<ET, JT>* R, Lane<ET, JT>& L, int kv)
{
    if constexpr (Q < Topo<ET>::cip) {
        scvf_jac<ET, FAST, Q, JT>(C, R[Q], L, nullptr, kv);
        jac_all<ET, FAST, Q + 1, JT>(C, R, L, kv);
    }
}

// ---------------------------------------------------------------------
// assembly kernel: one lane per (row node, incident element), node-major;
// a workgroup owns the consecutive nodes [v0, v1) and their contiguous CSR
// values; rows are accumulated in LDS in rounds of the incidence slot
// ---------------------------------------------------------------------
template <int ET, typename JT, int FAST>
__global__ void __launch_bounds__(256)
assemble_rows(KMeshPtrs M, KParams P, const double* __restrict__ u, const double* __restrict__ uold,
              JT* __restrict__ J, double* __restrict__ dout, int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1, NIP = T::nip;
    extern __shared__ __align__(16) unsigned char smem[];

    // consecutive workgroups (neighbouring node ranges, shared element data)
    // go to the same XCD: hardware hands block i to XCD i % 8
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

    Lane<ET, JT> L;
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        L.dacc[a] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int b = 0; b < NC; ++b) L.jacc[a][k][b] = JT(0);
    }
    int v = v0, slot = -1;
    int pos[N];
#pragma unroll
    for (int k = 0; k < N; ++k) pos[k] = 0;
    if (active) {
        const int inc = M.perm[lane_inc];
        const int e = M.inc_elem[inc];
        v = M.inc_node[inc];
        const int info = M.inc_info[inc];
        const int kv = info & 0xff;
        slot = info >> 8;
        int pk[N];
        int32_t ccn[N];
        double co[N][D], ul[N][NC];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            pk[k] = CanonH<ET>::c.perm[kv][k];
            ccn[k] = M.conn[(int64_t)e * N + pk[k]];
#pragma unroll
            for (int i = 0; i < D; ++i) co[k][i] = M.xyz[(int64_t)ccn[k] * D + i];
#pragma unroll
            for (int a = 0; a < NC; ++a) ul[k][a] = u[(int64_t)ccn[k] * NC + a];
        }
        bool okg = true, oku = true;
        if (P.parts & 1) {
            double c[D];
            elem_centre<ET>(co, c);
            Rec<ET, JT> R[T::cip];
            const Cfg<FAST> C{P, M.nu_elem ? M.nu_elem[e] : P.nu, M.nu_ip ? M.nu_ip + (int64_t)e * NIP : nullptr,
                              M.f_scvf ? M.f_scvf + (int64_t)e * NIP * D : nullptr,
                              M.rho_ip ? M.rho_ip + (int64_t)e * NIP : nullptr};
            scvf_all<ET, FAST, 0, JT>(C, co, M.xyz, M.conn + (int64_t)e * N, ul, pk, ccn, c, uold, M.l2inv + (int64_t)e * NIP,
                                      (P.upw_mode != 0 && M.upw_dec) ? M.upw_dec + (int64_t)e * NIP : nullptr, kv, wantJ, L, R,
                                      okg, oku, (ET == 0 || ET == 2) || (M.affine != nullptr && M.affine[e] != 0));
            if (wantJ) jac_all<ET, FAST, 0, JT>(C, R, L, kv);
        }
        // mass part and source: navier_stokes_fv1.cpp:781-869
        if (P.parts & 6) {
            const double volk = M.vol[(int64_t)e * N + kv];
            const double rk = M.rho_scv ? M.rho_scv[(int64_t)e * N + kv] : P.rho;  // density at the SCV ip
            if (P.parts & 2) {
                const double m = volk * rk;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    if (wantJ) L.jacc[d][0][d] += JT(P.s_m * m);
                    L.dacc[d] += P.s_m * (ul[0][d] * m);
                }
            }
            if ((P.parts & 4) && P.has_source)
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double fk = M.f_scv ? M.f_scv[((int64_t)e * N + kv) * D + d] : P.f[d];
                    L.dacc[d] += P.s_f * (fk * volk * rk);
                }
        }
        if (!oku) kernel_error(err, KERR_CUT_SIDE);
    if (!okg) kernel_error(err, KERR_DEGENERATE);
        // column slots of the (canonically ordered) corners in v's block row
        const uint64_t pw = M.inc_pos[inc];
#pragma unroll
        for (int k = 0; k < N; ++k) pos[k] = (int)((pw >> (8 * pk[k])) & 0xffu) * NC;
    }
    int rbase = 0, rlen = 0;
    if (active) {
        const int64_t rs = M.rowptr[(int64_t)v * NC];
        rbase = (int)(rs - jbase) + (v - v0) * PADE;
        rlen = (int)(M.rowptr[(int64_t)v * NC + 1] - rs);
    }
    __syncthreads();

    // ordered accumulation: round r adds the incidences of slot r
    const int rounds = M.wg_rounds[w];
    for (int r = 0; r < rounds; ++r) {
        if (active && slot == r) {
            if (wantJ) {
#pragma unroll
                for (int a = 0; a < NC; ++a) {
                    const int ra = rbase + a * rlen;
#pragma unroll
                    for (int k = 0; k < N; ++k) lds_add_block<NC>(jbuf + ra + pos[k], L.jacc[a][k]);
                }
            }
#pragma unroll
            for (int a = 0; a < NC; ++a) dbuf[(v - v0) * NC + a] += L.dacc[a];
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


// ---------------------------------------------------------------------
// Split path (FP32 Jacobian, common configuration): every SCVF is
// evaluated once.
//   scvf_records: one lane per (SCVF, element), in the frame of the SCVF's
//     `from` corner; writes the FP64 flux (added to `from`, subtracted
//     from `to`) and the FP32 record of its Jacobian block, expressed in
//     the original corner order and reference frame (128 B per SCVF).
//   assemble_rows_split: the row-owner gather of assemble_rows, reading the
//     three records of each (node, element) pair instead of recomputing
//     them; the `to` row takes the negated block (every term of the block is
//     linear in the normal or in the flux weight).
// ---------------------------------------------------------------------
constexpr int REC_FLOATS = 32;
// A record is REC_CHUNKS 16-byte chunks.  Chunk c of SCVF ip of element e
// lives at float4 index (ip * REC_CHUNKS + c) * ne_pad + e: the lanes of a
// wavefront, which hold consecutive elements, read and write each chunk as
// one contiguous run (whole cache lines per instruction).
constexpr int REC_CHUNKS = REC_FLOATS / 4;
// record layout: box elements keep the inverse reference Jacobian (their
// frame changes are axis reflections, exact in FP32); simplices keep the
// physical shape gradients (a frame change would round)
template <int ET>
struct RecLayout {
    static constexpr bool box = (ET == 1 || ET == 3);
    static constexpr int G = 0;               // Ji[D][D] (box) or grad[N][D] (simplex), original frame/corners
    static constexpr int Nrm = box ? 9 : 12;  // normal, from -> to
    static constexpr int S = Nrm + 3;         // pw, vv, vc, sp
    static constexpr int U = S + 4;           // upwind shapes, original corners
    static constexpr int F = 24;              // flux (4 x FP64)
    static_assert(U + Topo<ET>::nsh <= F, "record layout");
};

template <int ET, int FAST, int Q>
__device__ inline void record_one(const Cfg<FAST>& C, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                  const double* __restrict__ xyz, const int32_t* __restrict__ econn,
                                  const double (&ul)[Topo<ET>::nsh][Topo<ET>::dim + 1], const int* pk,
                                  const int32_t* ccn, const double* c, const double* __restrict__ uold, double l2,
                                  int kv, float4* __restrict__ rec, int64_t cs, bool& okg, bool& oku, uint8_t* decp,
                                  bool affine)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh;
    Lane<ET, float> L;
#pragma unroll
    for (int a = 0; a <= D; ++a) L.dacc[a] = 0.0;
    Rec<ET, float> R;
    bool have_cut = false;
    int side = -1;
    double gi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
    scvf_canon<ET, FAST, Q, float, ET == ET_TET>(C, co, xyz, econn, ul, pk, ccn, c, uold, l2, true, true, L, R, okg,
                                                 oku, have_cut, side, gi, li, decp, affine);
    using RL = RecLayout<ET>;
    float rv[RL::F];
#pragma unroll
    for (int i = 0; i < RL::F; ++i) rv[i] = 0.f;
    if constexpr (RL::box) {
        // canonical -> original reference frame: Ji_o = A Ji_c (A = diag(+-1))
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                float acc = 0.f;
#pragma unroll
                for (int m = 0; m < D; ++m) acc += float(CanonH<ET>::c.A[kv][i][m]) * R.Ji[m][j];
                rv[RL::G + i * D + j] = acc;
            }
    } else {
        // physical gradients of every corner, computed once in the original
        // frame in FP64 (constant on a simplex): J = sum_k x_k (x) grad_ref
        // phi_k, cofactor inverse, grad phi_c = J^-T grad_ref phi_c; every
        // product rounded on its own (the oracle's arithmetic), so a
        // gradient component that vanishes there is exactly zero here too
#pragma clang fp contract(off)
        double J[D][D], Ki[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double glk = k == 0 ? -1.0 : (k == j + 1 ? 1.0 : 0.0);
                    acc += xyz[(int64_t)econn[k] * D + i] * glk;
                }
                J[i][j] = acc;
            }
        if constexpr (D == 2) {
            const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            Ki[0][0] = J[1][1] / det;  Ki[0][1] = -J[0][1] / det;
            Ki[1][0] = -J[1][0] / det; Ki[1][1] = J[0][0] / det;
        } else {
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            Ki[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
            Ki[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            Ki[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            Ki[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
            Ki[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            Ki[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            Ki[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
            Ki[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            Ki[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
#pragma unroll
        for (int c = 0; c < N; ++c)
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) acc += Ki[j][i] * (c == 0 ? -1.0 : (c == j + 1 ? 1.0 : 0.0));
                rv[RL::G + c * D + i] = float(acc);
            }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) rv[RL::Nrm + i] = R.n[i];
    rv[RL::S + 0] = R.pw;
    rv[RL::S + 1] = R.vv;
    rv[RL::S + 2] = R.vc;
    rv[RL::S + 3] = R.sp;
    if constexpr (FAST == UPW_FULL || FAST == UPW_SKEWED) {
        // one-hot shapes: the original number of the upwind corner (-1: none)
        int idx = -1;
#pragma unroll
        for (int k = 0; k < N; ++k) idx = R.Us[k] != 0.f ? pk[k] : idx;
        rv[RL::U] = __int_as_float(idx);
    } else if constexpr (FAST == UPW_LPS) {
#pragma unroll
        for (int c = 0; c < N; ++c) {
            float x = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) x = pk[k] == c ? R.Us[k] : x;
            rv[RL::U + c] = x;
        }
    }
#pragma unroll
    for (int q = 0; q < RL::F / 4; ++q) rec[q * cs] = make_float4(rv[4 * q], rv[4 * q + 1], rv[4 * q + 2], rv[4 * q + 3]);
    *reinterpret_cast<double2*>(rec + (RL::F / 4) * cs) = make_double2(L.dacc[0], L.dacc[1]);
    *reinterpret_cast<double2*>(rec + (RL::F / 4 + 1) * cs) = make_double2(L.dacc[2], D == 3 ? L.dacc[3] : 0.0);
}

template <int ET, int Q>
struct IpsOfQ {
    static constexpr int count = [] {
        int n = 0;
        for (int ip = 0; ip < Topo<ET>::nip; ++ip)
            if (CanonH<ET>::c.kq[ip] == Q) ++n;
        return n;
    }();
    static constexpr auto list = [] {
        struct L { int ip[12]; } l{};
        int n = 0;
        for (int ip = 0; ip < Topo<ET>::nip; ++ip)
            if (CanonH<ET>::c.kq[ip] == Q) l.ip[n++] = ip;
        return l;
    }();
};

template <int ET, int FAST, int Q>
__global__ void __launch_bounds__(256)
scvf_records(KMeshPtrs M, KParams P, const double* __restrict__ u, const double* __restrict__ uold, int64_t ne,
             int64_t ne_pad, float* __restrict__ recs, int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1, NIP = T::nip;
    // wavefront w takes SCVF w % count of the 64 elements block w / count:
    // consecutive wavefronts share their elements' data in L2
    constexpr int CNT = IpsOfQ<ET, Q>::count;
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t wave = L >> 6;
    const int idx = (int)(wave % CNT);
    const int64_t e = (wave / CNT) * 64 + (L & 63);
    if (e >= ne || e >= ne_pad) return;
    const int ip = IpsOfQ<ET, Q>::list.ip[idx];
    const int kv = T::edge[ip][0];
    int pk[N];
    int32_t ccn[N];
    double co[N][D], ul[N][NC];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        pk[k] = CanonH<ET>::c.perm[kv][k];
        ccn[k] = M.conn[e * N + pk[k]];
#pragma unroll
        for (int i = 0; i < D; ++i) co[k][i] = M.xyz[(int64_t)ccn[k] * D + i];
#pragma unroll
        for (int a = 0; a < NC; ++a) ul[k][a] = u[(int64_t)ccn[k] * NC + a];
    }
    double c[D];
    elem_centre<ET>(co, c);
    const Cfg<FAST> C{P, M.nu_elem ? M.nu_elem[e] : P.nu};
    const double l2 = M.l2inv[e * NIP + ip];
    uint8_t* decp = (P.upw_mode != 0 && M.upw_dec) ? M.upw_dec + e * NIP + ip : nullptr;
    const bool affine = (ET == 0 || ET == 2) || (M.affine != nullptr && M.affine[e] != 0);
    float4* rec = reinterpret_cast<float4*>(recs) + (int64_t)ip * REC_CHUNKS * ne_pad + e;
    bool okg = true, oku = true;
    record_one<ET, FAST, Q>(C, co, M.xyz, M.conn + e * N, ul, pk, ccn, c, uold, l2, kv, rec, ne_pad, okg, oku, decp,
                            affine);
    if (!oku) kernel_error(err, KERR_CUT_SIDE);
    if (!okg) kernel_error(err, KERR_DEGENERATE);
}

template <int ET, int FAST, int Q>
__device__ inline void gather_jac(const Cfg<FAST>& C, const float* __restrict__ recs, int64_t ne_pad, int64_t e,
                                  int kv, const int* pk, Lane<ET, float>& L, bool wantJ)
{
    using T = Topo<ET>;
    if constexpr (Q < T::cip) {
        constexpr int D = T::dim, N = T::nsh;
        const int ip = CanonH<ET>::c.oip[kv][Q];
        const float sg = CanonH<ET>::c.from[kv][Q] ? 1.f : -1.f;
        const float4* rec = reinterpret_cast<const float4*>(recs) + (int64_t)ip * REC_CHUNKS * ne_pad + e;
        using RL = RecLayout<ET>;
        const double2 fa = *reinterpret_cast<const double2*>(rec + (RL::F / 4) * ne_pad);
        const double2 fb = *reinterpret_cast<const double2*>(rec + (RL::F / 4 + 1) * ne_pad);
        L.dacc[0] += double(sg) * fa.x;
        L.dacc[1] += double(sg) * fa.y;
        L.dacc[2] += double(sg) * fb.x;
        if constexpr (D == 3) L.dacc[3] += double(sg) * fb.y;
        if (wantJ) {
            float rv[RL::F];
#pragma unroll
            for (int q = 0; q < RL::F / 4; ++q) {
                const float4 t = rec[q * ne_pad];
                rv[4 * q] = t.x;
                rv[4 * q + 1] = t.y;
                rv[4 * q + 2] = t.z;
                rv[4 * q + 3] = t.w;
            }
            Rec<ET, float> R;
            float G[N][D];
            if constexpr (RL::box) {
                // original -> this lane's canonical frame: Ji_c = A^-1 Ji_o
#pragma unroll
                for (int i = 0; i < D; ++i)
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        float acc = 0.f;
#pragma unroll
                        for (int m = 0; m < D; ++m) acc += float(CanonH<ET>::c.Ainv[kv][i][m]) * rv[RL::G + m * D + j];
                        R.Ji[i][j] = acc;
                    }
            } else {
#pragma unroll
                for (int k = 0; k < N; ++k)
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        float x = 0.f;
#pragma unroll
                        for (int c = 0; c < N; ++c) x = pk[k] == c ? rv[RL::G + c * D + i] : x;
                        G[k][i] = x;
                    }
            }
#pragma unroll
            for (int i = 0; i < D; ++i) {
                R.n[i] = sg * rv[RL::Nrm + i];
                R.up[i] = 0.f;
            }
            R.pw = sg * rv[RL::S + 0];
            R.pw1 = 0.f;
            R.nr = float(-1.0 * C.nu * C.P.rho);
            R.vv = rv[RL::S + 1];
            R.vc = rv[RL::S + 2];
            R.sp = rv[RL::S + 3];
            R.w = 1.f;
            if constexpr (FAST == UPW_FULL || FAST == UPW_SKEWED) {
                const int idx = __float_as_int(rv[RL::U]);
#pragma unroll
                for (int k = 0; k < N; ++k) R.Us[k] = pk[k] == idx ? 1.f : 0.f;
            } else if constexpr (FAST == UPW_LPS) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    float x = 0.f;
#pragma unroll
                    for (int c = 0; c < N; ++c) x = pk[k] == c ? rv[RL::U + c] : x;
                    R.Us[k] = x;
                }
            } else {
                constexpr int IP = corner_ip<ET>(0, Q);
#pragma unroll
                for (int k = 0; k < N; ++k) R.Us[k] = float(TabH<ET>::t.phi[IP][k]);
            }
#pragma unroll
            for (int k = 0; k < N; ++k) R.Uc[k] = R.Us[k];
            if constexpr (RL::box) scvf_jac<ET, FAST, Q, float>(C, R, L);
            else scvf_jac<ET, FAST, Q, float>(C, R, L, G);
        }
        gather_jac<ET, FAST, Q + 1>(C, recs, ne_pad, e, kv, pk, L, wantJ);
    }
}

template <int ET, int FAST>
__global__ void __launch_bounds__(256)
assemble_rows_split(KMeshPtrs M, KParams P, const double* __restrict__ u, const float* __restrict__ recs,
                    int64_t ne_pad, float* __restrict__ J, double* __restrict__ dout)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1;
    using JT = float;
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = [] {
        const int nwg = (int)gridDim.x, i = (int)blockIdx.x;
        const int per = nwg >> 3, rem = nwg & 7, x = i & 7, k = i >> 3;
        return x * per + (x < rem ? x : rem) + k;
    }();
    // the workgroup's and the lane's plan data come packed, one load each
    const int4 wa = M.wg_a[2 * w], wb = M.wg_a[2 * w + 1];
    const int v0 = wa.x, v1 = wa.y;
    const int64_t jbase = (int64_t)(uint32_t)wb.x | ((int64_t)wb.y << 32);
    const int jlen = wb.z;
    const int dlen = (v1 - v0) * NC;
    JT* jbuf = reinterpret_cast<JT*>(smem);
    // 16 pad bytes after each node's rows: consecutive nodes then start in
    // different LDS banks, so the lanes of one round do not conflict, and
    // every (row, column block) stays 16-byte aligned for vector access
    constexpr int PADE = NC == 4 ? 16 / (int)sizeof(JT) : 0;  // 2D rows already have odd strides
    double* dbuf = reinterpret_cast<double*>(smem + (((jlen + (v1 - v0) * PADE) * (int)sizeof(JT) + 15) & ~15));
    const bool wantJ = J != nullptr;
    if (wantJ)
        for (int i = threadIdx.x; i < jlen + (v1 - v0) * PADE; i += blockDim.x) jbuf[i] = JT(0);
    for (int i = threadIdx.x; i < dlen; i += blockDim.x) dbuf[i] = 0.0;

    const int i0 = wa.z, i1 = wa.w;
    const int lane_inc = i0 + (int)threadIdx.x;
    const bool active = lane_inc < i1;

    Lane<ET, JT> L;
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        L.dacc[a] = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int b = 0; b < NC; ++b) L.jacc[a][k][b] = JT(0);
    }
    int v = v0, slot = -1, rbase = 0, rlen = 0;
    int pos[N];
#pragma unroll
    for (int k = 0; k < N; ++k) pos[k] = 0;
    if (active) {
        const int4 la = M.lane_a[lane_inc];
        const uint64_t pw = M.lane_pos[lane_inc];
        const int e = la.x;
        v = la.y;
        const int kv = la.z & 0xff;
        slot = (la.z >> 8) & 0xff;
        rlen = la.z >> 16;
        rbase = la.w;
        int pk[N];
#pragma unroll
        for (int k = 0; k < N; ++k) pk[k] = CanonH<ET>::c.perm[kv][k];
        if (P.parts & 1) {
            const Cfg<FAST> C{P, M.nu_elem ? M.nu_elem[e] : P.nu};
            gather_jac<ET, FAST, 0>(C, recs, ne_pad, e, kv, pk, L, wantJ);
        }
        if (P.parts & 6) {
            const double volk = M.vol[(int64_t)e * N + kv];
            if (P.parts & 2) {
                const double m = volk * P.rho;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    if (wantJ) L.jacc[d][0][d] += JT(P.s_m * m);
                    L.dacc[d] += P.s_m * (u[(int64_t)v * NC + d] * m);
                }
            }
            if ((P.parts & 4) && P.has_source)
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double fk = M.f_scv ? M.f_scv[((int64_t)e * N + kv) * D + d] : P.f[d];
                    L.dacc[d] += P.s_f * (fk * volk * P.rho);
                }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) pos[k] = (int)((pw >> (8 * pk[k])) & 0xffu) * NC;
    }
    __syncthreads();
    const int rounds = wb.w;
    for (int r = 0; r < rounds; ++r) {
        if (active && slot == r) {
            if (wantJ) {
#pragma unroll
                for (int a = 0; a < NC; ++a) {
                    const int ra = rbase + a * rlen;
#pragma unroll
                    for (int k = 0; k < N; ++k) lds_add_block<NC>(jbuf + ra + pos[k], L.jacc[a][k]);
                }
            }
#pragma unroll
            for (int a = 0; a < NC; ++a) dbuf[(v - v0) * NC + a] += L.dacc[a];
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

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
// ---------------------------------------------------------------------
// Row-lane variant of the split row gather: four lanes (dim+1 in general)
// per (node, element) incidence, lane a forms row a of the element's block.
// A lane holds its three records (loaded together, one memory round trip)
// and one row of the block instead of all dim+1 rows, so more wavefronts
// fit and each keeps all its loads in flight.  The terms and their order
// of addition per entry are those of scvf_jac for the split configurations
// (FIELDS, one upwind, no Stokes/PAC/Peclet/exact), so the result is
// bitwise that of assemble_rows_split.
// ---------------------------------------------------------------------
template <int ET, int FAST, int Q>
__device__ inline void row_terms(const Cfg<FAST>& C, const float4* __restrict__ rec, int64_t cs, int kv, const int* pk,
                                 int a, float (&jrow)[Topo<ET>::nsh][Topo<ET>::dim + 1])
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1;
    constexpr int IP = corner_ip<ET>(0, Q);
    using RL = RecLayout<ET>;
    using JT = float;
    const KParams& P = C.P;
    const float sg = CanonH<ET>::c.from[kv][Q] ? 1.f : -1.f;
    float rv[RL::F];
#pragma unroll
    for (int q = 0; q < RL::F / 4; ++q) {
        const float4 t = rec[q * cs];
        rv[4 * q] = t.x;
        rv[4 * q + 1] = t.y;
        rv[4 * q + 2] = t.z;
        rv[4 * q + 3] = t.w;
    }
    JT Ji[D][D], G[N][D], n[D], Us[N];
    if constexpr (RL::box) {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                float acc = 0.f;
#pragma unroll
                for (int m = 0; m < D; ++m) acc += float(CanonH<ET>::c.Ainv[kv][i][m]) * rv[RL::G + m * D + j];
                Ji[i][j] = acc;
            }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) {
                float x = 0.f;
#pragma unroll
                for (int c = 0; c < N; ++c) x = pk[k] == c ? rv[RL::G + c * D + i] : x;
                G[k][i] = x;
            }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) n[i] = sg * rv[RL::Nrm + i];
    const JT pw = sg * rv[RL::S + 0], vv = rv[RL::S + 1], vc = rv[RL::S + 2], sp = rv[RL::S + 3];
    if constexpr (FAST == UPW_FULL || FAST == UPW_SKEWED) {
        const int idx = __float_as_int(rv[RL::U]);
#pragma unroll
        for (int k = 0; k < N; ++k) Us[k] = pk[k] == idx ? 1.f : 0.f;
    } else if constexpr (FAST == UPW_LPS) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            float x = 0.f;
#pragma unroll
            for (int c = 0; c < N; ++c) x = pk[k] == c ? rv[RL::U + c] : x;
            Us[k] = x;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) Us[k] = float(TabH<ET>::t.phi[IP][k]);
    }
    const JT s_ = JT(P.s_a), rh = JT(P.rho), nr = JT(-1.0 * C.nu * P.rho);
    const bool mom = a < D;
#pragma unroll
    for (int sh = 0; sh < N; ++sh) {
        JT g[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if constexpr (RL::box) {
                JT acc = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) acc += Ji[j][i] * JT(TabH<ET>::t.gl[IP][sh][j]);
                g[i] = acc;
            } else {
                g[i] = G[sh][i];
            }
        }
        const JT ph = JT(TabH<ET>::t.phi[IP][sh]);
        JT ga = 0, na = 0;  // g[a], n[a] of the lane's momentum row
#pragma unroll
        for (int i = 0; i < D; ++i) {
            ga = a == i ? g[i] : ga;
            na = a == i ? n[i] : na;
        }
        if (mom) {
            JT gn = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) gn += g[i] * n[i];
            const JT cv = Us[sh] * pw;
#pragma unroll
            for (int b = 0; b < NC; ++b) {
                JT x = jrow[sh][b];
                if (b == a) x += s_ * (nr * gn);
                if (b < D && !P.laplace) x += s_ * (nr * ga * n[b < D ? b : 0]);
                if (b == D) x += s_ * (ph * na);
                if (b == a) x += s_ * cv;
                jrow[sh][b] = x;
            }
        } else {
            const JT Sv = vv * ph + vc * Us[sh];
            JT cpp = 0;
#pragma unroll
            for (int d1 = 0; d1 < D; ++d1) {
                jrow[sh][d1] += s_ * (Sv * n[d1] * rh);
                cpp += (g[d1] * sp) * n[d1] * rh;
            }
            jrow[sh][D] += s_ * cpp;
        }
    }
}

template <int ET, int FAST, int Q>
__device__ inline void row_all(const Cfg<FAST>& C, const float4* const* recq, int64_t cs, int kv, const int* pk, int a,
                               float (&jrow)[Topo<ET>::nsh][Topo<ET>::dim + 1])
{
    if constexpr (Q < Topo<ET>::cip) {
        row_terms<ET, FAST, Q>(C, recq[Q], cs, kv, pk, a, jrow);
        row_all<ET, FAST, Q + 1>(C, recq, cs, kv, pk, a, jrow);
    }
}

template <int ET, int FAST>
__global__ void __launch_bounds__(64)
assemble_rows_rowlane(KMeshPtrs M, KParams P, const double* __restrict__ u, const float* __restrict__ recs,
                      int64_t ne_pad, float* __restrict__ J, double* __restrict__ dout)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1, CIP = T::cip;
    using JT = float;
    using RL = RecLayout<ET>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int w = [] {
        const int nwg = (int)gridDim.x, i = (int)blockIdx.x;
        const int per = nwg >> 3, rem = nwg & 7, x = i & 7, k = i >> 3;
        return x * per + (x < rem ? x : rem) + k;
    }();
    const int4 wa = M.wg_a[2 * w], wb = M.wg_a[2 * w + 1];
    const int v0 = wa.x, v1 = wa.y;
    const int64_t jbase = (int64_t)(uint32_t)wb.x | ((int64_t)wb.y << 32);
    const int jlen = wb.z;
    const int dlen = (v1 - v0) * NC;
    JT* jbuf = reinterpret_cast<JT*>(smem);
    constexpr int PADE = NC == 4 ? 16 / (int)sizeof(JT) : 0;  // 2D rows already have odd strides
    double* dbuf = reinterpret_cast<double*>(smem + (((jlen + (v1 - v0) * PADE) * (int)sizeof(JT) + 15) & ~15));
    const bool wantJ = J != nullptr;
    if (wantJ)
        for (int i = threadIdx.x; i < jlen + (v1 - v0) * PADE; i += blockDim.x) jbuf[i] = JT(0);
    for (int i = threadIdx.x; i < dlen; i += blockDim.x) dbuf[i] = 0.0;

    const int li = (int)threadIdx.x / NC, a = (int)threadIdx.x - li * NC;
    const int i0 = wa.z, i1 = wa.w;
    const int lane_inc = i0 + li;
    const bool active = li < 64 / NC && lane_inc < i1;
    float jrow[N][NC];
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
        for (int b = 0; b < NC; ++b) jrow[k][b] = 0.f;
    double dsum = 0.0;
    int v = v0, slot = -1, rbase = 0, rlen = 0;
    int pos[N];
#pragma unroll
    for (int k = 0; k < N; ++k) pos[k] = 0;
    if (active) {
        const int4 la = M.lane_a[lane_inc];
        const uint64_t pw = M.lane_pos[lane_inc];
        const int64_t e = la.x;
        v = la.y;
        const int kv = la.z & 0xff;
        slot = (la.z >> 8) & 0xff;
        rlen = la.z >> 16;
        rbase = la.w;
        int pk[N];
#pragma unroll
        for (int k = 0; k < N; ++k) pk[k] = CanonH<ET>::c.perm[kv][k];
        if (P.parts & 1) {
            const float4* recq[CIP];
            double fl[CIP];
#pragma unroll
            for (int q = 0; q < CIP; ++q) {
                const int ip = CanonH<ET>::c.oip[kv][q];
                recq[q] = reinterpret_cast<const float4*>(recs) + (int64_t)ip * REC_CHUNKS * ne_pad + e;
                const double2 f = *reinterpret_cast<const double2*>(recq[q] + (RL::F / 4 + (a >> 1)) * ne_pad);
                fl[q] = (a & 1) ? f.y : f.x;
            }
#pragma unroll
            for (int q = 0; q < CIP; ++q) dsum += double(CanonH<ET>::c.from[kv][q] ? 1.f : -1.f) * fl[q];
            const Cfg<FAST> C{P, M.nu_elem ? M.nu_elem[e] : P.nu};
            if (wantJ) row_all<ET, FAST, 0>(C, recq, ne_pad, kv, pk, a, jrow);
        }
        if (P.parts & 6) {
            const double volk = M.vol[e * N + kv];
            if ((P.parts & 2) && a < D) {
                const double m = volk * P.rho;
#pragma unroll
                for (int b = 0; b < NC; ++b)
                    if (b == a && wantJ) jrow[0][b] += JT(P.s_m * m);
                dsum += P.s_m * (u[(int64_t)v * NC + a] * m);
            }
            if ((P.parts & 4) && P.has_source && a < D) {
                double f = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) f = d == a ? P.f[d] : f;
                dsum += P.s_f * (f * volk * P.rho);
            }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) pos[k] = (int)((pw >> (8 * pk[k])) & 0xffu) * NC;
    }
    __syncthreads();
    const int rounds = wb.w;
    for (int r = 0; r < rounds; ++r) {
        if (active && slot == r) {
            if (wantJ) {
                const int ra = rbase + a * rlen;
#pragma unroll
                for (int k = 0; k < N; ++k) lds_add_block<NC>(jbuf + ra + pos[k], jrow[k]);
            }
            dbuf[(v - v0) * NC + a] += dsum;
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

template <int ET, typename JT>
static hipError_t launch_cfg(int fast, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                             JT* J, double* d, int* err, int nwg, size_t lds, hipStream_t s)
{
    switch (fast) {
    case UPW_NO: assemble_rows<ET, JT, UPW_NO><<<nwg, m.block, lds, s>>>(m, p, u, uold, J, d, err); break;
    case UPW_FULL: assemble_rows<ET, JT, UPW_FULL><<<nwg, m.block, lds, s>>>(m, p, u, uold, J, d, err); break;
    case UPW_SKEWED: assemble_rows<ET, JT, UPW_SKEWED><<<nwg, m.block, lds, s>>>(m, p, u, uold, J, d, err); break;
    case UPW_LPS: assemble_rows<ET, JT, UPW_LPS><<<nwg, m.block, lds, s>>>(m, p, u, uold, J, d, err); break;
    default: assemble_rows<ET, JT, -1><<<nwg, m.block, lds, s>>>(m, p, u, uold, J, d, err); break;
    }
    return hipGetLastError();
}

// the common configuration runs the specialised kernel
static int fast_path(const KParams& p)
{
    const bool common = p.stab_type == STAB_FIELDS && !p.stokes && !p.pac && !p.peclet && p.exact == 0.0 &&
                        p.conv_upwind == p.stab_upwind && p.conv_upwind >= 0 && p.conv_upwind <= UPW_LPS;
    return common ? p.conv_upwind : -1;
}

template <int ET>
static hipError_t launch_et(int jprec, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                            void* J, double* d, int* err, int nwg, size_t lds, hipStream_t s)
{
    const int fast = fast_path(p);
    if (jprec == 64) return launch_cfg<ET, double>(fast, m, p, u, uold, static_cast<double*>(J), d, err, nwg, lds, s);
    return launch_cfg<ET, float>(fast, m, p, u, uold, static_cast<float*>(J), d, err, nwg, lds, s);
}

hipError_t launch_assemble(int et, int jprec, const KMeshPtrs& m, const KParams& p, const double* u,
                           const double* uold, void* J, double* d, int* err, int nwg, size_t lds, hipStream_t s)
{
    if (lds > 65536 || m.block < 64 || m.block > 256 || m.block % 64) return hipErrorInvalidValue;
    switch (et) {
    case ET_TRI:  return launch_et<ET_TRI>(jprec, m, p, u, uold, J, d, err, nwg, lds, s);
    case ET_QUAD: return launch_et<ET_QUAD>(jprec, m, p, u, uold, J, d, err, nwg, lds, s);
    case ET_TET:  return launch_et<ET_TET>(jprec, m, p, u, uold, J, d, err, nwg, lds, s);
    case ET_HEX:  return launch_et<ET_HEX>(jprec, m, p, u, uold, J, d, err, nwg, lds, s);
    case ET_PRISM: return launch_et<ET_PRISM>(jprec, m, p, u, uold, J, d, err, nwg, lds, s);
    }
    return hipErrorInvalidValue;
}


// split path: records then row gather (FP32 Jacobian, common configuration)
template <int ET, int FAST>
static hipError_t launch_split_cfg(const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                                   float* J, double* d, int* err, int nwg, size_t lds, int64_t ne, float* recs,
                                   bool rowlane, hipStream_t s)
{
    const int64_t ne_pad = (ne + 63) / 64 * 64;
    if (p.parts & 1) {
        // one launch per frame index Q that has SCVFs (triangles use Q = 0 only)
        if constexpr (IpsOfQ<ET, 0>::count > 0)
            scvf_records<ET, FAST, 0><<<(unsigned)((IpsOfQ<ET, 0>::count * ne_pad + 255) / 256), 256, 0, s>>>(
                m, p, u, uold, ne, ne_pad, recs, err);
        if constexpr (IpsOfQ<ET, 1>::count > 0)
            scvf_records<ET, FAST, 1><<<(unsigned)((IpsOfQ<ET, 1>::count * ne_pad + 255) / 256), 256, 0, s>>>(
                m, p, u, uold, ne, ne_pad, recs, err);
        if constexpr (Topo<ET>::cip > 2 && IpsOfQ<ET, 2>::count > 0)
            scvf_records<ET, FAST, 2><<<(unsigned)((IpsOfQ<ET, 2>::count * ne_pad + 255) / 256), 256, 0, s>>>(
                m, p, u, uold, ne, ne_pad, recs, err);
    }
    if (rowlane) assemble_rows_rowlane<ET, FAST><<<nwg, 64, lds, s>>>(m, p, u, recs, ne_pad, J, d);
    else assemble_rows_split<ET, FAST><<<nwg, m.block, lds, s>>>(m, p, u, recs, ne_pad, J, d);
    return hipGetLastError();
}

template <int ET>
static hipError_t launch_split_et(const KMeshPtrs& m, const KParams& p, const double* u, const double* uold, float* J,
                                  double* d, int* err, int nwg, size_t lds, int64_t ne, float* recs, bool rl,
                                  hipStream_t s)
{
    switch (fast_path(p)) {
    case UPW_NO: return launch_split_cfg<ET, UPW_NO>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rl, s);
    case UPW_FULL: return launch_split_cfg<ET, UPW_FULL>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rl, s);
    case UPW_SKEWED: return launch_split_cfg<ET, UPW_SKEWED>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rl, s);
    case UPW_LPS: return launch_split_cfg<ET, UPW_LPS>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rl, s);
    }
    return hipErrorInvalidValue;
}

// Tetrahedra stay on the fused kernel: a record built in the `from` corner's
// frame rounds some exactly-zero normal/gradient products of axis-aligned
// tetrahedra differently from a row's own frame (O(1e-19) where the
// reference arithmetic gives an exact 0).
// prisms keep the fused path: the record layout and the canonical record
// frames cover the four element types of the reference's 2D/3D workloads
bool split_supported(int et, const KParams& p) { return et != ET_PRISM && fast_path(p) >= 0; }

int64_t split_record_floats(int et, int64_t ne)
{
    const int nip = et == ET_HEX ? 12 : et == ET_TET ? 6 : et == ET_QUAD ? 4 : 3;
    return (int64_t)nip * ((ne + 63) / 64 * 64) * REC_FLOATS;
}

hipError_t launch_assemble_split(int et, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                                 float* J, double* d, int* err, int nwg, size_t lds, int64_t ne, float* recs,
                                 bool rowlane, hipStream_t s)
{
    if (lds > 65536 || m.block < 64 || m.block > 256 || m.block % 64 || !split_supported(et, p)) return hipErrorInvalidValue;
    if (rowlane && m.block != 64) return hipErrorInvalidValue;
    switch (et) {
    case ET_TRI:  return launch_split_et<ET_TRI>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rowlane, s);
    case ET_QUAD: return launch_split_et<ET_QUAD>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rowlane, s);
    case ET_TET:  return launch_split_et<ET_TET>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rowlane, s);
    case ET_HEX:  return launch_split_et<ET_HEX>(m, p, u, uold, J, d, err, nwg, lds, ne, recs, rowlane, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_geometry_cache(int et, const double* xyz, const int32_t* conn, int64_t ne, int difflen,
                                 double* l2inv, double* vol, uint8_t* affine, int* err, hipStream_t s)
{
    const int blocks = (int)((ne + 255) / 256);
    switch (et) {
    case ET_TRI:  geometry_cache<ET_TRI><<<blocks, 256, 0, s>>>(xyz, conn, ne, difflen, l2inv, vol, affine, err); break;
    case ET_QUAD: geometry_cache<ET_QUAD><<<blocks, 256, 0, s>>>(xyz, conn, ne, difflen, l2inv, vol, affine, err); break;
    case ET_TET:  geometry_cache<ET_TET><<<blocks, 256, 0, s>>>(xyz, conn, ne, difflen, l2inv, vol, affine, err); break;
    case ET_HEX:  geometry_cache<ET_HEX><<<blocks, 256, 0, s>>>(xyz, conn, ne, difflen, l2inv, vol, affine, err); break;
    case ET_PRISM: geometry_cache<ET_PRISM><<<blocks, 256, 0, s>>>(xyz, conn, ne, difflen, l2inv, vol, affine, err); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------
// stabilised SCVF velocities (ex_velocity_ip, navier_stokes_fv1.cpp:1105-
// 1204): one lane per (element, corner); the lane evaluates the SCVFs whose
// `from` corner it is, in that corner's canonical frame, and writes
// out[e][ip][d].  Upwind decisions are always fresh.
// ---------------------------------------------------------------------
template <int ET, int Q>
__device__ inline void stab_vel_all(const Cfg<-1>& C, const double (&co)[Topo<ET>::nsh][Topo<ET>::dim],
                                    const double* __restrict__ xyz, const int32_t* __restrict__ econn,
                                    const double (&ul)[Topo<ET>::nsh][Topo<ET>::dim + 1], const int* pk,
                                    const int32_t* ccn, const double* c, const double* uold, const double* l2row, int kv,
                                    double* oute, bool& okg, bool& oku, bool affine)
{
    if constexpr (Q < Topo<ET>::cip) {
        if (CanonH<ET>::c.from[kv][Q]) {
            bool have_cut = false;
            int side = -1;
            double gi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
            const int oip = CanonH<ET>::c.oip[kv][Q];
            Lane<ET, double> L;
#pragma unroll
            for (int a = 0; a <= Topo<ET>::dim; ++a) L.dacc[a] = 0.0;
            Rec<ET, double> R;
            scvf_canon<ET, -1, Q, double>(C, co, xyz, econn, ul, pk, ccn, c, uold, l2row[oip], true, false, L, R, okg,
                                          oku, have_cut, side, gi, li, nullptr, affine, oute + oip * Topo<ET>::dim);
        }
        stab_vel_all<ET, Q + 1>(C, co, xyz, econn, ul, pk, ccn, c, uold, l2row, kv, oute, okg, oku, affine);
    }
}

template <int ET>
__global__ void __launch_bounds__(256)
export_stab_vel(KMeshPtrs M, KParams P, const double* __restrict__ u, const double* __restrict__ uold, int64_t ne,
                double* __restrict__ out, int* __restrict__ err)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh, NC = D + 1, NIP = T::nip;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * N) return;
    const int64_t e = t / N;
    const int kv = (int)(t - e * N);
    int pk[N];
    int32_t ccn[N];
    double co[N][D], ul[N][NC];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        pk[k] = CanonH<ET>::c.perm[kv][k];
        ccn[k] = M.conn[e * N + pk[k]];
#pragma unroll
        for (int i = 0; i < D; ++i) co[k][i] = M.xyz[(int64_t)ccn[k] * D + i];
#pragma unroll
        for (int a = 0; a < NC; ++a) ul[k][a] = u[(int64_t)ccn[k] * NC + a];
    }
    double c[D];
    elem_centre<ET>(co, c);
    bool okg = true, oku = true;
    const Cfg<-1> C{P, M.nu_elem ? M.nu_elem[e] : P.nu, M.nu_ip ? M.nu_ip + e * NIP : nullptr,
                    M.f_scvf ? M.f_scvf + e * NIP * D : nullptr,
                        M.rho_ip ? M.rho_ip + e * NIP : nullptr};
    stab_vel_all<ET, 0>(C, co, M.xyz, M.conn + e * N, ul, pk, ccn, c, uold, M.l2inv + e * NIP, kv,
                        out + e * NIP * D, okg, oku, (ET == 0 || ET == 2) || (M.affine != nullptr && M.affine[e] != 0));
    if (!oku) kernel_error(err, KERR_CUT_SIDE);
    if (!okg) kernel_error(err, KERR_DEGENERATE);
}

hipError_t launch_export_stab_vel(int et, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                                  int64_t ne, double* out, int* err, hipStream_t s)
{
    const int nsh = et == ET_HEX ? 8 : (et == ET_PRISM ? 6 : (et == ET_TRI ? 3 : 4));
    const int64_t lanes = ne * nsh;
    if (lanes == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((lanes + 255) / 256);
    switch (et) {
    case ET_TRI:  export_stab_vel<ET_TRI><<<blocks, 256, 0, s>>>(m, p, u, uold, ne, out, err); break;
    case ET_QUAD: export_stab_vel<ET_QUAD><<<blocks, 256, 0, s>>>(m, p, u, uold, ne, out, err); break;
    case ET_TET:  export_stab_vel<ET_TET><<<blocks, 256, 0, s>>>(m, p, u, uold, ne, out, err); break;
    case ET_HEX:  export_stab_vel<ET_HEX><<<blocks, 256, 0, s>>>(m, p, u, uold, ne, out, err); break;
    case ET_PRISM: export_stab_vel<ET_PRISM><<<blocks, 256, 0, s>>>(m, p, u, uold, ne, out, err); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace nsfv1
