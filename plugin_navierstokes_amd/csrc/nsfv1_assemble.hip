// nsfv1_assemble.hip -- row-owner gather assembly of the FV1 incompressible
// Navier-Stokes defect and Jacobian (gfx950).
//
// One thread per (row node v, incident element e) incidence.  The thread
// forms the sub-control-volume faces of e that touch v (3 for hexahedra and
// tetrahedra, 2 for 2D elements), evaluates the upwind, the FIELDS
// stabilisation and the flux linearisation of navier_stokes_fv1.cpp:250-778
// for those faces only, and holds the element's contribution to v's block
// row: (dim+1) x nsh x (dim+1) Jacobian values and dim+1 defect values.
// A workgroup owns a contiguous range of nodes, hence a contiguous range of
// CSR values; it accumulates them in LDS in rounds ordered by the
// incidence slot (no two threads of one round share a row, so the sum
// order is fixed and the result is bitwise reproducible, without atomics)
// and streams the finished rows to HBM with coalesced stores.
#include "nsfv1_internal.hpp"

namespace nsfv1 {

__constant__ RefTopo c_topo[4];

// Mesh arrays (device):  xyz [nv][dim], conn [ne][nsh], CSR rowptr/colidx,
// incidences sorted by node (inc_elem, inc_node, inc_info = corner | slot<<8,
// node_inc [nv+1]), workgroup plan wg_node [nwg+1] / wg_rounds [nwg].
using KMesh = KMeshPtrs;

template <int ET, typename JT>
__global__ void __launch_bounds__(256)
assemble_rows(KMesh M, KParams P, const double* __restrict__ u, const double* __restrict__ uold,
              JT* __restrict__ J, double* __restrict__ dout, int* __restrict__ err)
{
    constexpr int D = Shape<ET>::dim, N = Shape<ET>::nsh, NC = D + 1, CIP = Shape<ET>::cip;
    extern __shared__ __align__(16) unsigned char smem[];
    const RefTopo& R = c_topo[ET];

    const int w = blockIdx.x;
    const int v0 = M.wg_node[w], v1 = M.wg_node[w + 1];
    const int64_t jbase = M.rowptr[(int64_t)v0 * NC];
    const int64_t jlen = M.rowptr[(int64_t)v1 * NC] - jbase;
    const int dlen = (v1 - v0) * NC;
    JT* jbuf = reinterpret_cast<JT*>(smem);
    double* dbuf = reinterpret_cast<double*>(smem + ((jlen * sizeof(JT) + 15) & ~(int64_t)15));
    for (int64_t i = threadIdx.x; i < jlen; i += blockDim.x) jbuf[i] = JT(0);
    for (int i = threadIdx.x; i < dlen; i += blockDim.x) dbuf[i] = 0.0;

    const int64_t i0 = M.node_inc[v0], i1 = M.node_inc[v1];
    const int64_t inc = i0 + threadIdx.x;
    const bool active = inc < i1;

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
    int v = 0, kv = 0, slot = -1;
    int32_t cn[N];
    int pos[N];
    if (active) {
        const int e = M.inc_elem[inc];
        v = M.inc_node[inc];
        const int info = M.inc_info[inc];
        kv = info & 0xff;
        slot = info >> 8;
        double co[N][D], ul[N][NC], uo[N][NC];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            cn[k] = M.conn[(int64_t)e * N + k];
#pragma unroll
            for (int i = 0; i < D; ++i) co[k][i] = M.xyz[(int64_t)cn[k] * D + i];
#pragma unroll
            for (int a = 0; a < NC; ++a) {
                ul[k][a] = u[(int64_t)cn[k] * NC + a];
                uo[k][a] = P.has_old ? uold[(int64_t)cn[k] * NC + a] : 0.0;
            }
        }
        double vol[N];
        scv_volumes<ET>(co, vol);

        // COR diffusion length needs element-wide extrema of |n|^2 (and of
        // the edge-midpoint/centre distance in 3D): diffusion_length.h:135-199
        double corMinN = 0, corAvgN = 0, corMinD = 0;
        bool ok = true, okg = true;
        if (P.stab_type == STAB_FIELDS && P.difflen == DL_COR) {
            corMinN = 1.7976931348623157e308; corMinD = 1.7976931348623157e308;
            for (int ip = 0; ip < Shape<ET>::nip; ++ip) {
                IpGeo<ET> Gq;
                okg &= ip_geometry<ET>(R, co, ip, Gq);
                const double nn = dotd<D>(Gq.n, Gq.n);
                corMinN = nn < corMinN ? nn : corMinN;
                corAvgN += nn;
                if (D == 3) {
                    const double ds = dist2d<D>(Gq.c0, Gq.c2);
                    corMinD = ds < corMinD ? ds : corMinD;
                }
            }
            corAvgN /= Shape<ET>::nip;
        }

        const double nu = P.nu, rho = P.rho;
        const int ncip = (P.parts & 1) ? CIP : 0;
        for (int q = 0; q < ncip; ++q) {
            const int ip = R.cip[kv][q];
            const int fr = R.edge[ip][0], to = R.edge[ip][1];
            const double sgn = (fr == kv) ? P.s_a : -P.s_a;
            IpGeo<ET> G;
            okg &= ip_geometry<ET>(R, co, ip, G);

            // StdVel: navier_stokes_fv1.cpp:631-642
            double sv[D];
            {
#pragma clang fp contract(off)
#pragma unroll
                for (int d = 0; d < D; ++d) sv[d] = 0.0;
#pragma unroll
                for (int sh = 0; sh < N; ++sh)
#pragma unroll
                    for (int d = 0; d < D; ++d) sv[d] += ul[sh][d] * G.phi[sh];
            }
            // upwind of the stabilisation and of the convective term
            double Us[N], Ls = 1.0;
#pragma unroll
            for (int k = 0; k < N; ++k) Us[k] = 0.0;
            if (!P.stokes) ok &= upwind_ip<ET>(P.stab_upwind, R, co, ip, G, sv, Us, Ls);
            double Uc[N];
#pragma unroll
            for (int k = 0; k < N; ++k) Uc[k] = Us[k];
            if (!P.stokes && !P.pac && P.conv_upwind != P.stab_upwind) {
                double Lc;
                ok &= upwind_ip<ET>(P.conv_upwind, R, co, ip, G, sv, Uc, Lc);
            }

            // stabilisation: stabilization.cpp:122-242 (FIELDS, diagonal),
            // 805-850 (without)
            double nn = 0;
#pragma unroll
            for (int i = 0; i < D; ++i) nn += G.n[i] * G.n[i];
            double svel[D], Sv[N], Sp[D][N];
            if (P.stab_type == STAB_FIELDS) {
                double a2 = 0.5 * (vol[fr] + vol[to]);
                a2 *= a2;
                double dist = 0;
                if (D == 3) dist = dist2d<D>(G.c0, G.c2);
                double L2inv;
                if (P.difflen == DL_RAW) {
                    L2inv = (D == 2) ? 1. / (0.5 * a2 / nn + 3.0 * nn / 8.) : 1. / (0.5 * a2 / nn + 3.0 * dist / 8.);
                } else if (P.difflen == DL_FIVEPOINT) {
                    L2inv = (D == 2) ? 2.0 * nn / a2 + 8.0 / nn : 2.0 * nn / a2 + 8.0 * dist / nn;
                } else {
                    L2inv = (D == 2) ? 2.0 * corMinN / a2 + 8.0 / (3.0 * corAvgN)
                                     : 2.0 * corMinN / a2 + 8.0 * corMinD / (3.0 * corAvgN);
                }
                const double vv = nu * L2inv;
                const double vc = P.stokes ? 0.0 : sqrt(dotd<D>(sv, sv)) / Ls;
                double diag = vv;
                if (P.has_old) diag += 1. / P.dt;
                if (!P.stokes) diag += vc;
#pragma unroll
                for (int k = 0; k < N; ++k) Sv[k] = (vv * G.phi[k] + vc * Us[k]);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    double rhs = P.has_source ? P.f[d] : 0.0;
                    if (P.has_old) {
                        double o = 0.0;
#pragma unroll
                        for (int sh = 0; sh < N; ++sh) o += G.phi[sh] * uo[sh][d];
                        rhs += o / P.dt;
                    }
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        rhs += Sv[k] * ul[k][d];
                        const double sumP = -1.0 * G.g[k][d] / rho;
                        rhs += sumP * ul[k][D];
                        Sp[d][k] = sumP / diag;
                    }
                    svel[d] = rhs / diag;
                }
#pragma unroll
                for (int k = 0; k < N; ++k) Sv[k] = Sv[k] / diag;
            } else {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    Sv[k] = G.phi[k];
#pragma unroll
                    for (int d = 0; d < D; ++d) Sp[d][k] = 0.0;
                }
#pragma unroll
                for (int d = 0; d < D; ++d) svel[d] = sv[d];
            }

            // ---- defect: navier_stokes_fv1.cpp:667-777 ----
            {
                double gv[D][D];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) {
                        double s = 0;
#pragma unroll
                        for (int sh = 0; sh < N; ++sh) s += G.g[sh][d2] * ul[sh][d1];
                        gv[d1][d2] = s;
                    }
                double df[D];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) {
                    double s = 0;
#pragma unroll
                    for (int d2 = 0; d2 < D; ++d2) s += gv[d1][d2] * G.n[d2];
                    if (!P.laplace)
#pragma unroll
                        for (int d2 = 0; d2 < D; ++d2) s += gv[d2][d1] * G.n[d2];
                    df[d1] = s * (-1.0 * nu * rho);
                }
                double pr = 0;
#pragma unroll
                for (int sh = 0; sh < N; ++sh) pr += G.phi[sh] * ul[sh][D];
#pragma unroll
                for (int d1 = 0; d1 < D; ++d1) dacc[d1] += sgn * (df[d1] + pr * G.n[d1]);
                if (!P.stokes) {
                    double upv[D];
                    if (!P.pac) {
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            double s = 0;
#pragma unroll
                            for (int k = 0; k < N; ++k) s += Uc[k] * ul[k][d];
                            upv[d] = s;
                        }
                    } else {
#pragma unroll
                        for (int d = 0; d < D; ++d) upv[d] = svel[d];
                    }
                    double wpe = 1.0;
                    if (P.peclet) {
                        // peclet_blend: navier_stokes_fv1.cpp:871-892
                        const double Pe = dotd<D>(sv, G.n) / nn * sqrt(dist2d<D>(co[to], co[fr])) / nu;
                        wpe = Pe * Pe / (5.0 + Pe * Pe);
#pragma unroll
                        for (int d = 0; d < D; ++d) upv[d] = wpe * upv[d] + (1.0 - wpe) * sv[d];
                    }
                    const double prod = dotd<D>(sv, G.n) * rho;
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) dacc[d1] += sgn * (upv[d1] * prod);

                    // ---- Jacobian convection: navier_stokes_fv1.cpp:375-553 ----
                    if (J) {
#pragma unroll
                        for (int sh = 0; sh < N; ++sh) {
                            if (P.pac) {
#pragma unroll
                                for (int d1 = 0; d1 < D; ++d1) {
                                    jacc[d1][sh][d1] += JT(sgn * (prod * wpe * Sv[sh]));
                                    jacc[d1][sh][D] += JT(sgn * (prod * wpe * Sp[d1][sh]));
                                }
                            } else {
                                const double cv = Uc[sh] * prod * wpe;
#pragma unroll
                                for (int d1 = 0; d1 < D; ++d1) jacc[d1][sh][d1] += JT(sgn * cv);
                            }
                            if (P.peclet) {
                                const double cpe = prod * (1.0 - wpe) * G.phi[sh];
#pragma unroll
                                for (int d1 = 0; d1 < D; ++d1) jacc[d1][sh][d1] += JT(sgn * cpe);
                            }
                            if (P.exact != 0.0) {
                                if (P.pac) {
#pragma unroll
                                    for (int d1 = 0; d1 < D; ++d1) {
                                        const double pv = Sv[sh] * G.n[d1] * (P.exact * rho);
#pragma unroll
                                        for (int d2 = 0; d2 < D; ++d2) jacc[d1][sh][d2] += JT(sgn * (pv * upv[d1]));
                                        double pp = 0;
#pragma unroll
                                        for (int k = 0; k < D; ++k) pp += Sp[k][sh] * G.n[k];
                                        pp *= P.exact * rho;
                                        jacc[d1][sh][D] += JT(sgn * (pp * upv[d1]));
                                    }
                                } else {
#pragma unroll
                                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                                        for (int d2 = 0; d2 < D; ++d2)
                                            jacc[d1][sh][d2] += JT(sgn * (wpe * Uc[sh] * G.n[d2] * rho * upv[d1]));
                                }
                                if (P.peclet)
#pragma unroll
                                    for (int d1 = 0; d1 < D; ++d1)
#pragma unroll
                                        for (int d2 = 0; d2 < D; ++d2)
                                            jacc[d1][sh][d2] += JT(sgn * (upv[d1] * (1.0 - wpe) * G.phi[sh] * G.n[d2] * rho));
                            }
                        }
                    }
                }
                dacc[D] += sgn * (dotd<D>(svel, G.n) * rho);
            }

            // ---- Jacobian: diffusion, pressure, continuity ----
            if (J) {
                const double nr = -1.0 * nu * rho;
#pragma unroll
                for (int sh = 0; sh < N; ++sh) {
                    double gn = 0;
#pragma unroll
                    for (int i = 0; i < D; ++i) gn += G.g[sh][i] * G.n[i];
                    const double flux_sh = nr * gn;
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) {
                        jacc[d1][sh][d1] += JT(sgn * flux_sh);
                        if (!P.laplace)
#pragma unroll
                            for (int d2 = 0; d2 < D; ++d2) jacc[d1][sh][d2] += JT(sgn * (nr * G.g[sh][d1] * G.n[d2]));
                        jacc[d1][sh][D] += JT(sgn * (G.phi[sh] * G.n[d1]));
                        jacc[D][sh][d1] += JT(sgn * (Sv[sh] * G.n[d1] * rho));
                    }
                    double cpp = 0;
#pragma unroll
                    for (int d1 = 0; d1 < D; ++d1) cpp += Sp[d1][sh] * G.n[d1] * rho;
                    jacc[D][sh][D] += JT(sgn * cpp);
                }
            }
        }

        // ---- mass part and source: navier_stokes_fv1.cpp:781-869 ----
        if (P.parts & 2) {
            const double m = vol[kv] * rho;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (J) jacc[d][kv][d] += JT(P.s_m * m);
                dacc[d] += P.s_m * (ul[kv][d] * m);
            }
        }
        if ((P.parts & 4) && P.has_source) {
#pragma unroll
            for (int d = 0; d < D; ++d) dacc[d] += P.s_f * (P.f[d] * vol[kv] * rho);
        }
        if (!ok || !okg) err[0] = (ok ? 0 : KERR_CUT_SIDE) | (okg ? 0 : KERR_DEGENERATE);

        // column slots of the element's corners in v's block row
        const int64_t rs = M.rowptr[(int64_t)v * NC];
        const int deg = (int)((M.rowptr[(int64_t)v * NC + 1] - rs) / NC);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int32_t key = cn[k] * NC;
            int lo = 0, hi = deg - 1, found = 0;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int32_t c = M.colidx[rs + (int64_t)mid * NC];
                if (c == key) { found = mid; break; }
                if (c < key) lo = mid + 1; else hi = mid - 1;
            }
            pos[k] = found * NC;
        }
    }
    __syncthreads();

    // ---- ordered accumulation of the rows in LDS ----
    const int rounds = M.wg_rounds[w];
    for (int r = 0; r < rounds; ++r) {
        if (active && slot == r) {
            const int64_t rs = M.rowptr[(int64_t)v * NC] - jbase;
            const int64_t rlen = M.rowptr[(int64_t)v * NC + 1] - M.rowptr[(int64_t)v * NC];
            if (J) {
#pragma unroll
                for (int a = 0; a < NC; ++a)
#pragma unroll
                    for (int k = 0; k < N; ++k)
#pragma unroll
                        for (int b = 0; b < NC; ++b) jbuf[rs + a * rlen + pos[k] + b] += jacc[a][k][b];
            }
#pragma unroll
            for (int a = 0; a < NC; ++a) dbuf[(v - v0) * NC + a] += dacc[a];
        }
        __syncthreads();
    }

    // ---- coalesced write-back ----
    if (J)
        for (int64_t i = threadIdx.x; i < jlen; i += blockDim.x) J[jbase + i] = jbuf[i];
    if (dout)
        for (int i = threadIdx.x; i < dlen; i += blockDim.x) dout[(int64_t)v0 * NC + i] = dbuf[i];
}

}  // namespace nsfv1

namespace nsfv1 {

hipError_t upload_topology(const RefTopo* topo4)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(c_topo), topo4, sizeof(RefTopo) * 4);
}

template <int ET>
static hipError_t launch_et(int jprec, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                            void* J, double* d, int* err, int nwg, int threads, size_t lds, hipStream_t s)
{
    if (jprec == 64)
        assemble_rows<ET, double><<<nwg, threads, lds, s>>>(m, p, u, uold, static_cast<double*>(J), d, err);
    else
        assemble_rows<ET, float><<<nwg, threads, lds, s>>>(m, p, u, uold, static_cast<float*>(J), d, err);
    return hipGetLastError();
}

hipError_t launch_assemble(int et, int jprec, const KMeshPtrs& m, const KParams& p,
                           const double* u, const double* uold, void* J, double* d, int* err,
                           int nwg, int threads, size_t lds, hipStream_t s)
{
    switch (et) {
    case ET_TRI:  return launch_et<ET_TRI>(jprec, m, p, u, uold, J, d, err, nwg, threads, lds, s);
    case ET_QUAD: return launch_et<ET_QUAD>(jprec, m, p, u, uold, J, d, err, nwg, threads, lds, s);
    case ET_TET:  return launch_et<ET_TET>(jprec, m, p, u, uold, J, d, err, nwg, threads, lds, s);
    case ET_HEX:  return launch_et<ET_HEX>(jprec, m, p, u, uold, J, d, err, nwg, threads, lds, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace nsfv1
