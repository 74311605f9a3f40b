/*
 * fv1_oracle.c -- CPU restatement of the FV1 incompressible Navier-Stokes
 * element assembly of UG4/plugin_NavierStokes, in plain C, FP64.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Each function cites the reference file:line it restates.  The geometry
 * (FV1Geometry), the ray/side intersection and the CSR scatter are ugcore
 * code that is not present; they are restated here from the box-method
 * construction:
 *   - SCVF i <-> reference edge i, from/to = the edge's corners;
 *   - 2D SCVF = segment [edge midpoint, element centre];
 *     3D SCVF = quad [edge midpoint, centre of face A, element centre,
 *     centre of face B], A and B the two faces containing the edge;
 *   - ip = average of the SCVF corners (local and global);
 *   - normal: 2D the rotated segment, 3D half the cross product of the
 *     diagonals; its sign makes n.(x_to - x_from) > 0;
 *   - shapes and gradients of the P1/Q1 Lagrange space at the local ip,
 *     gradients mapped with J^-T;
 *   - SCV of corner k: bounded by the corner, the midpoints of its edges, the
 *     centres of its faces and the element centre; 2D area by the diagonal
 *     cross product, 3D volume = exact volume of the trilinear hexahedron
 *     through those eight points (2x2x2 Gauss on det J).
 */
#include "fv1_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[256];
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char* or_last_error(void) { return g_err; }

/* the imports at the integration points: the kinematic viscosity and the
   source at the SCVF ips, the source at the SCV ips
   (navier_stokes_fv1.cpp:188-247; used at :296, 336, 351, 390, 855-868) */
static double nu_at(const or_params* p, int ip) { return p->nu_ip ? p->nu_ip[ip] : p->nu; }
static double rho_at(const or_params* p, int ip) { return p->rho_ip ? p->rho_ip[ip] : p->rho; }
static double rho_sh(const or_params* p, int sh) { return p->rho_scv ? p->rho_scv[sh] : p->rho; }
static double src_scvf(const or_params* p, int ip, int d, int dim) { return p->f_scvf ? p->f_scvf[ip * dim + d] : p->f[d]; }
static double src_scv(const or_params* p, int sh, int d, int dim) { return p->f_scv ? p->f_scv[sh * dim + d] : p->f[d]; }

/* ------------------------------------------------------------------ */
/* reference elements (UG4 corner / edge / side numbering)            */
/* ------------------------------------------------------------------ */
typedef struct {
    int dim, nsh, nip, nside;
    double lc[8][3];
    int edge[12][2];
    int side_n[6];
    int side[6][4];
} refelem;

static const refelem REF[5] = {
    /* triangle */
    {2, 3, 3, 3,
     {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}},
     {{0, 1}, {1, 2}, {2, 0}},
     {2, 2, 2},
     {{0, 1}, {1, 2}, {2, 0}}},
    /* quadrilateral */
    {2, 4, 4, 4,
     {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}},
     {{0, 1}, {1, 2}, {2, 3}, {3, 0}},
     {2, 2, 2, 2},
     {{0, 1}, {1, 2}, {2, 3}, {3, 0}}},
    /* tetrahedron */
    {3, 4, 6, 4,
     {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
     {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}},
     {3, 3, 3, 3},
     {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}}},
    /* hexahedron */
    {3, 8, 12, 6,
     {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},
     {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}},
     {4, 4, 4, 4, 4, 4},
     {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}}},
    /* prism: triangle 0-1-2 at zeta = 0, 3-4-5 above it (UG4 numbering) */
    {3, 6, 9, 5,
     {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
     {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 4}, {2, 5}, {3, 4}, {4, 5}, {5, 3}},
     {3, 4, 4, 4, 3},
     {{0, 2, 1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}, {3, 4, 5}}},
};

int or_nsh(int et) { return (et >= 0 && et < 5) ? REF[et].nsh : -1; }
int or_nip(int et) { return (et >= 0 && et < 5) ? REF[et].nip : -1; }
int or_dim(int et) { return (et >= 0 && et < 5) ? REF[et].dim : -1; }

static int side_has(const refelem* R, int s, int c)
{
    for (int i = 0; i < R->side_n[s]; ++i)
        if (R->side[s][i] == c) return 1;
    return 0;
}

/* the two sides containing edge e, in side order (3D) */
static void edge_faces(const refelem* R, int e, int* A, int* B)
{
    int k = 0;
    *A = *B = -1;
    for (int s = 0; s < R->nside && k < 2; ++s)
        if (side_has(R, s, R->edge[e][0]) && side_has(R, s, R->edge[e][1])) {
            if (k == 0) *A = s; else *B = s;
            ++k;
        }
}

/* ------------------------------------------------------------------ */
/* P1 / Q1 Lagrange shapes on the reference elements                  */
/* ------------------------------------------------------------------ */
static void shapes(int et, const double* x, double* phi, double (*gl)[3])
{
    switch (et) {
    case OR_TRI:
        phi[0] = 1.0 - x[0] - x[1]; phi[1] = x[0]; phi[2] = x[1];
        gl[0][0] = -1; gl[0][1] = -1; gl[1][0] = 1; gl[1][1] = 0; gl[2][0] = 0; gl[2][1] = 1;
        break;
    case OR_QUAD: {
        const double a = x[0], b = x[1];
        phi[0] = (1 - a) * (1 - b); phi[1] = a * (1 - b); phi[2] = a * b; phi[3] = (1 - a) * b;
        gl[0][0] = -(1 - b); gl[0][1] = -(1 - a);
        gl[1][0] = (1 - b);  gl[1][1] = -a;
        gl[2][0] = b;        gl[2][1] = a;
        gl[3][0] = -b;       gl[3][1] = (1 - a);
        break;
    }
    case OR_TET:
        phi[0] = 1.0 - x[0] - x[1] - x[2]; phi[1] = x[0]; phi[2] = x[1]; phi[3] = x[2];
        for (int k = 0; k < 4; ++k) for (int j = 0; j < 3; ++j) gl[k][j] = 0;
        gl[0][0] = gl[0][1] = gl[0][2] = -1;
        gl[1][0] = 1; gl[2][1] = 1; gl[3][2] = 1;
        break;
    case OR_PRISM:
        /* P1 triangle (barycentric lambda) times P1 segment in zeta */
        for (int k = 0; k < 6; ++k) {
            const int t = k % 3;
            const double lam = t == 0 ? 1.0 - x[0] - x[1] : (t == 1 ? x[0] : x[1]);
            const double dl0 = t == 0 ? -1.0 : (t == 1 ? 1.0 : 0.0), dl1 = t == 0 ? -1.0 : (t == 2 ? 1.0 : 0.0);
            const double w = k < 3 ? 1.0 - x[2] : x[2], dw = k < 3 ? -1.0 : 1.0;
            phi[k] = lam * w;
            gl[k][0] = dl0 * w;
            gl[k][1] = dl1 * w;
            gl[k][2] = lam * dw;
        }
        break;
    case OR_HEX:
        for (int k = 0; k < 8; ++k) {
            const double* c = REF[OR_HEX].lc[k];
            double f[3], df[3];
            for (int j = 0; j < 3; ++j) {
                f[j] = c[j] > 0.5 ? x[j] : 1.0 - x[j];
                df[j] = c[j] > 0.5 ? 1.0 : -1.0;
            }
            phi[k] = f[0] * f[1] * f[2];
            gl[k][0] = df[0] * f[1] * f[2];
            gl[k][1] = f[0] * df[1] * f[2];
            gl[k][2] = f[0] * f[1] * df[2];
        }
        break;
    }
}

static double dot(int dim, const double* a, const double* b)
{
    double s = 0;
    for (int i = 0; i < dim; ++i) s += a[i] * b[i];
    return s;
}
static double dist2(int dim, const double* a, const double* b)
{
    double s = 0;
    for (int i = 0; i < dim; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
    return s;
}
static void cross(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* volume of the trilinear hexahedron through p[8] (hex corner order) */
static double trilinear_hex_volume(const double p[8][3])
{
    const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
    const double gp[2] = {g0, g1};
    double v = 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                double x[3] = {gp[a], gp[b], gp[c]}, phi[8], gl[8][3], J[3][3] = {{0}};
                shapes(OR_HEX, x, phi, gl);
                for (int k = 0; k < 8; ++k)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) J[i][j] += p[k][i] * gl[k][j];
                double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                           - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                           + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                v += 0.125 * det;
            }
    return fabs(v);
}

/* ------------------------------------------------------------------ */
/* FV1Geometry::update (ugcore, restated); used at                      */
/* navier_stokes_fv1.cpp:214-216                                       */
/* ------------------------------------------------------------------ */
int or_geometry(int et, const double* corners, or_geo* G)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const refelem* R = &REF[et];
    const int dim = R->dim, nsh = R->nsh, nip = R->nip;
    memset(G, 0, sizeof *G);
    G->et = et; G->dim = dim; G->nsh = nsh; G->nip = nip;
    for (int k = 0; k < nsh; ++k)
        for (int i = 0; i < dim; ++i) G->co[k][i] = corners[k * dim + i];

    double c[3] = {0, 0, 0}, cl[3] = {0, 0, 0};
    for (int k = 0; k < nsh; ++k)
        for (int i = 0; i < dim; ++i) { c[i] += G->co[k][i]; cl[i] += R->lc[k][i]; }
    for (int i = 0; i < dim; ++i) { c[i] /= nsh; cl[i] /= nsh; }

    double me[12][3], mel[12][3];
    for (int e = 0; e < nip; ++e)
        for (int i = 0; i < dim; ++i) {
            me[e][i] = 0.5 * (G->co[R->edge[e][0]][i] + G->co[R->edge[e][1]][i]);
            mel[e][i] = 0.5 * (R->lc[R->edge[e][0]][i] + R->lc[R->edge[e][1]][i]);
        }
    double fc[6][3] = {{0}}, fcl[6][3] = {{0}};
    if (dim == 3)
        for (int s = 0; s < R->nside; ++s) {
            for (int j = 0; j < R->side_n[s]; ++j)
                for (int i = 0; i < 3; ++i) {
                    fc[s][i] += G->co[R->side[s][j]][i];
                    fcl[s][i] += R->lc[R->side[s][j]][i];
                }
            for (int i = 0; i < 3; ++i) { fc[s][i] /= R->side_n[s]; fcl[s][i] /= R->side_n[s]; }
        }

    for (int ip = 0; ip < nip; ++ip) {
        const int f = R->edge[ip][0], t = R->edge[ip][1];
        G->from[ip] = f; G->to[ip] = t;
        double n[3] = {0, 0, 0};
        if (dim == 2) {
            for (int i = 0; i < 2; ++i) {
                G->ipg[ip][i] = 0.5 * (me[ip][i] + c[i]);
                G->ipl[ip][i] = 0.5 * (mel[ip][i] + cl[i]);
                G->c0[ip][i] = me[ip][i];
                G->c2[ip][i] = c[i];
            }
            const double d0 = c[0] - me[ip][0], d1 = c[1] - me[ip][1];
            n[0] = d1; n[1] = -d0;
        } else {
            int A, B;
            edge_faces(R, ip, &A, &B);
            for (int i = 0; i < 3; ++i) {
                G->ipg[ip][i] = 0.25 * (me[ip][i] + fc[A][i] + c[i] + fc[B][i]);
                G->ipl[ip][i] = 0.25 * (mel[ip][i] + fcl[A][i] + cl[i] + fcl[B][i]);
                G->c0[ip][i] = me[ip][i];
                G->c2[ip][i] = c[i];
            }
            double d1[3], d2[3];
            for (int i = 0; i < 3; ++i) { d1[i] = c[i] - me[ip][i]; d2[i] = fc[B][i] - fc[A][i]; }
            cross(d1, d2, n);
            for (int i = 0; i < 3; ++i) n[i] *= 0.5;
        }
        double e[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) e[i] = G->co[t][i] - G->co[f][i];
        if (dot(dim, n, e) < 0)
            for (int i = 0; i < dim; ++i) n[i] = -n[i];
        for (int i = 0; i < dim; ++i) G->n[ip][i] = n[i];

        /* shapes and global gradients at the local ip */
        double gl[8][3] = {{0}};
        shapes(et, G->ipl[ip], G->phi[ip], gl);
        double J[3][3] = {{0}};
        for (int k = 0; k < nsh; ++k)
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) J[i][j] += G->co[k][i] * gl[k][j];
        double Ji[3][3] = {{0}};
        if (dim == 2) {
            const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            if (det == 0.0) { set_err("degenerate element"); return -2; }
            Ji[0][0] = J[1][1] / det;  Ji[0][1] = -J[0][1] / det;
            Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
        } else {
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                             - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                             + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            if (det == 0.0) { set_err("degenerate element"); return -2; }
            Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
        /* grad_k = J^-T gl_k */
        for (int k = 0; k < nsh; ++k)
            for (int i = 0; i < dim; ++i) {
                double s = 0;
                for (int j = 0; j < dim; ++j) s += Ji[j][i] * gl[k][j];
                G->g[ip][k][i] = s;
            }
    }

    /* SCV volumes */
    for (int k = 0; k < nsh; ++k) {
        int es[3], ne = 0;
        for (int e = 0; e < nip; ++e)
            if (R->edge[e][0] == k || R->edge[e][1] == k) es[ne++] = e;
        if (dim == 2) {
            /* quad [x_k, m_ea, c, m_eb] */
            const double d1x = c[0] - G->co[k][0], d1y = c[1] - G->co[k][1];
            const double d2x = me[es[1]][0] - me[es[0]][0], d2y = me[es[1]][1] - me[es[0]][1];
            G->vol[k] = 0.5 * fabs(d1x * d2y - d1y * d2x);
        } else {
            int fs[3]; /* face of (e0,e1), (e0,e2), (e1,e2) */
            const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
            for (int q = 0; q < 3; ++q) {
                fs[q] = -1;
                const int ea = es[pairs[q][0]], eb = es[pairs[q][1]];
                for (int s = 0; s < R->nside; ++s)
                    if (side_has(R, s, R->edge[ea][0]) && side_has(R, s, R->edge[ea][1]) &&
                        side_has(R, s, R->edge[eb][0]) && side_has(R, s, R->edge[eb][1])) { fs[q] = s; break; }
            }
            double p[8][3];
            for (int i = 0; i < 3; ++i) {
                p[0][i] = G->co[k][i];
                p[1][i] = me[es[0]][i];
                p[2][i] = fc[fs[0]][i];
                p[3][i] = me[es[1]][i];
                p[4][i] = me[es[2]][i];
                p[5][i] = fc[fs[1]][i];
                p[6][i] = c[i];
                p[7][i] = fc[fs[2]][i];
            }
            G->vol[k] = trilinear_hex_volume(p);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* ElementSideRayIntersection (ugcore, restated), searching upwind:     */
/* the first side in reference order whose intersection parameters lie  */
/* in the side (tolerance 1e-10) at ray parameter t <= 0.               */
/* called from upwind.cpp:351-353 and 547-549                          */
/* ------------------------------------------------------------------ */
#define OR_RAY_EPS 1e-10

static int ray_tri(const double* p0, const double* p1, const double* p2,
                   const double* from, const double* dir, double* b1o, double* b2o, double* to)
{
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    for (int i = 0; i < 3; ++i) { e1[i] = p1[i] - p0[i]; e2[i] = p2[i] - p0[i]; }
    cross(dir, e2, pv);
    const double det = dot(3, e1, pv);
    if (det == 0.0) return 0;
    const double inv = 1.0 / det;
    for (int i = 0; i < 3; ++i) tv[i] = from[i] - p0[i];
    const double b1 = dot(3, tv, pv) * inv;
    cross(tv, e1, qv);
    const double b2 = dot(3, dir, qv) * inv;
    const double t = dot(3, e2, qv) * inv;
    if (b1 < -OR_RAY_EPS || b2 < -OR_RAY_EPS || b1 + b2 > 1.0 + OR_RAY_EPS) return 0;
    if (t > 0.0) return 0;
    *b1o = b1; *b2o = b2; *to = t;
    return 1;
}

int or_side_ray(int et, const double* corners, const double* from, const double* dir,
                int* side, double* glob, double* loc)
{
    const refelem* R = &REF[et];
    const int dim = R->dim;
#define CO(k) (corners + (k) * dim)
    if (dim == 2) {
        for (int s = 0; s < R->nside; ++s) {
            const double* p0 = CO(R->side[s][0]);
            const double* p1 = CO(R->side[s][1]);
            const double ex = p1[0] - p0[0], ey = p1[1] - p0[1];
            const double rx = p0[0] - from[0], ry = p0[1] - from[1];
            const double det = -dir[0] * ey + ex * dir[1];
            if (det == 0.0) continue;
            const double t = (-rx * ey + ex * ry) / det;
            const double lam = (dir[0] * ry - dir[1] * rx) / det;
            if (lam < -OR_RAY_EPS || lam > 1.0 + OR_RAY_EPS) continue;
            if (t > 0.0) continue;
            *side = s;
            glob[0] = p0[0] + lam * ex;
            glob[1] = p0[1] + lam * ey;
            const double* l0 = R->lc[R->side[s][0]];
            const double* l1 = R->lc[R->side[s][1]];
            for (int i = 0; i < 2; ++i) loc[i] = (1.0 - lam) * l0[i] + lam * l1[i];
            return 0;
        }
    } else {
        for (int s = 0; s < R->nside; ++s) {
            const int ntri = R->side_n[s] == 3 ? 1 : 2;
            for (int tr = 0; tr < ntri; ++tr) {
                const int a = R->side[s][0];
                const int b = R->side[s][tr == 0 ? 1 : 2];
                const int cc = R->side[s][tr == 0 ? 2 : 3];
                double b1, b2, t;
                if (!ray_tri(CO(a), CO(b), CO(cc), from, dir, &b1, &b2, &t)) continue;
                *side = s;
                for (int i = 0; i < 3; ++i) {
                    glob[i] = CO(a)[i] + b1 * (CO(b)[i] - CO(a)[i]) + b2 * (CO(cc)[i] - CO(a)[i]);
                    loc[i] = (1.0 - b1 - b2) * R->lc[a][i] + b1 * R->lc[b][i] + b2 * R->lc[cc][i];
                }
                return 0;
            }
        }
    }
#undef CO
    set_err("Cannot find cut side.");
    return -10;
}

/* ------------------------------------------------------------------ */
/* positive upwind: upwind.cpp:643-786.  Corner shapes U and ip shapes  */
/* W: the velocity at an ip is taken from the SCV it flows out of, as a */
/* convex combination of that SCV's corner value and the inflowing ip   */
/* velocities weighted by their share of the SCV's throughput F =       */
/* max(inflow, outflow).  Faces with (numerically) no flux take the     */
/* mean of their two corners.  W is zero for every other upwind.        */
/* ------------------------------------------------------------------ */
static void positive_upwind(const or_geo* G, const double (*vel)[3], double (*U)[8], double (*W)[12],
                            double* convlen)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh;
    const double eps = DBL_EPSILON * 10;
    double mflux[12];
    int has[12], nno = 0;
    for (int ip = 0; ip < nip; ++ip) {
        for (int sh = 0; sh < nsh; ++sh) U[ip][sh] = 0.0;
        for (int j = 0; j < nip; ++j) W[ip][j] = 0.0;
        mflux[ip] = 0.0;
        has[ip] = 1;
        const double nsq = dot(dim, vel[ip], vel[ip]);
        if (fabs(nsq) <= eps) {
            U[ip][G->from[ip]] = 0.5; U[ip][G->to[ip]] = 0.5;
            has[ip] = 0; ++nno;
            continue;
        }
        mflux[ip] = dot(dim, vel[ip], G->n[ip]);
        const double v = sqrt(nsq), len = sqrt(dot(dim, G->n[ip], G->n[ip]));
        if (fabs(mflux[ip] / sqrt(v * len)) <= eps) {
            U[ip][G->from[ip]] = 0.5; U[ip][G->to[ip]] = 0.5;
            has[ip] = 0; ++nno;
        }
    }
    if (nno != nip)
        for (int sh = 0; sh < nsh; ++sh) {
            double m_in = 0, m_out = 0, fl[12];
            int ips[12], cnt = 0;
            for (int ip = 0; ip < nip; ++ip) {
                if (!has[ip]) continue;
                double f;
                if (G->from[ip] == sh) f = mflux[ip];
                else if (G->to[ip] == sh) f = -1.0 * mflux[ip];
                else continue;
                ips[cnt] = ip; fl[cnt] = f; ++cnt;
                m_in += -1.0 * fmin(f, 0.0);
                m_out += fmax(f, 0.0);
            }
            const double F = fmax(m_in, m_out);
            for (int i = 0; i < cnt; ++i) {
                if (!(fl[i] > 0)) continue;
                double sum = 0.0;
                for (int j = 0; j < cnt; ++j)
                    if (fl[j] < 0) {
                        W[ips[i]][ips[j]] = -1.0 * fl[j] / F;
                        sum += W[ips[i]][ips[j]];
                    }
                U[ips[i]][sh] = 1.0 - sum;
            }
        }
    for (int ip = 0; ip < nip; ++ip) {
        double up[3] = {0, 0, 0};
        for (int sh = 0; sh < nsh; ++sh)
            for (int i = 0; i < dim; ++i) up[i] += U[ip][sh] * G->co[sh][i];
        for (int j = 0; j < nip; ++j)
            for (int i = 0; i < dim; ++i) up[i] += W[ip][j] * G->ipg[j][i];
        convlen[ip] = sqrt(dist2(dim, G->ipg[ip], up));
    }
}

/* ------------------------------------------------------------------ */
/* regular upwind (2D): upwind.cpp:792-964.  Its ray/SCV intersection   */
/* is ugcore's SCVFofSCVRayIntersection, which the reference tree does  */
/* not hold; it is restated with this convention:                       */
/*   - the SCV of corner k is the quad [x_k, m_k, c, m_{k-1}] (m_j the  */
/*     midpoint of edge j = (j, j+1), c the element centre); side 0 =   */
/*     [x_k, m_k] and side 3 = [m_{k-1}, x_k] lie on the element's      */
/*     edges, the other two sides are the SCVFs of edges k and k-1;     */
/*   - the ray x_ip - t v (t > 0) leaves the upwind SCV through the     */
/*     SCVF other than ip, side 0 or side 3, tried in that order; the   */
/*     first hit with 0 <= lambda <= 1 (tolerance 1e-10) wins;          */
/*   - lambda runs on side 0 from x_k to m_k, on side 3 from m_{k-1}   */
/*     to x_k (corner order), and on the SCVF from its edge midpoint    */
/*     (0) to c (1): the reading under which the reference's branch     */
/*     a.1 reproduces the intersection point exactly for either sign    */
/*     of the flux.                                                     */
/* ------------------------------------------------------------------ */
static int ray_segment_2d(const double* o, const double* v, const double* a, const double* b, double* lam)
{
    const double d0 = b[0] - a[0], d1 = b[1] - a[1];
    const double r0 = o[0] - a[0], r1 = o[1] - a[1];
    const double det = d0 * v[1] - d1 * v[0];
    if (det == 0.0) return 0;
    const double l = (r0 * v[1] - r1 * v[0]) / det;
    const double t = (d0 * r1 - d1 * r0) / det;
    if (!(t > 0.0) || l < -1e-10 || l > 1.0 + 1e-10) return 0;
    *lam = l < 0.0 ? 0.0 : (l > 1.0 ? 1.0 : l);
    return 1;
}

static int regular_upwind(const or_geo* G, const double (*vel)[3], double (*U)[8], double (*W)[12],
                          double* convlen)
{
    if (G->dim != 2) { set_err("RegularUpwind only implemented for 2d."); return -11; }
    const int nc = G->nsh, nip = G->nip;
    const double eps = DBL_EPSILON * 10;
    for (int ip = 0; ip < nip; ++ip) {
        for (int sh = 0; sh < nc; ++sh) U[ip][sh] = 0.0;
        for (int j = 0; j < nip; ++j) W[ip][j] = 0.0;
        const double* v = vel[ip];
        const double* n = G->n[ip];
        if (fabs(dot(2, v, v)) <= eps) continue;
        double flux = dot(2, v, n);
        if (fabs(flux) <= 100 * eps) {
            flux = v[0] * n[1] - v[1] * n[0];
            if (flux > 0) W[ip][(ip + nc - 1) % nc] = W[ip][(ip + 1) % nc] = 0.5;
            else U[ip][ip] = U[ip][(ip + 1) % nc] = 0.5;
            continue;
        }
        const int k = flux > 0.0 ? G->from[ip] : G->to[ip];
        const int oth = flux > 0.0 ? (ip + nc - 1) % nc : (ip + 1) % nc;
        double lam = 0.0;
        if (ray_segment_2d(G->ipg[ip], v, G->c0[oth], G->c2[oth], &lam)) {
            int ipI = (ip + 1) % nc, ipO = (ip + nc - 1) % nc;
            if (flux > 0.0) { ipI = (ip + nc - 1) % nc; ipO = (ip + 1) % nc; }
            if (lam <= 0.5) {
                W[ip][ipO] = lam - 0.5;
                W[ip][ipI] = 1.0 - (lam - 0.5);
            } else {
                U[ip][G->from[ip]] = 0.5 * 2 * (lam - 0.5);
                U[ip][G->to[ip]] = 0.5 * 2 * (lam - 0.5);
                W[ip][ipI] = 1.0 - 2 * (lam - 0.5);
            }
            continue;
        }
        int side = -1;
        if (ray_segment_2d(G->ipg[ip], v, G->co[k], G->c0[k], &lam)) side = 0;
        else if (ray_segment_2d(G->ipg[ip], v, G->c0[(k + nc - 1) % nc], G->co[k], &lam)) side = 3;
        if (side < 0) { set_err("RegularUpwind: the ray leaves the SCV through no side."); return -10; }
        if (flux > 0.0) {
            if (side == 0) { U[ip][ip] = 1.0 - 0.5 * lam; U[ip][(ip + 1) % nc] = 0.5 * lam; }
            else { U[ip][(ip + nc - 1) % nc] = 1.0 - 0.5 * (lam + 1.0); U[ip][ip] = 0.5 * (lam + 1.0); }
        } else {
            if (side == 0) { U[ip][(ip + 1) % nc] = 1.0 - 0.5 * lam; U[ip][(ip + 2) % nc] = 0.5 * lam; }
            else { U[ip][ip] = 1.0 - 0.5 * (lam + 1.0); U[ip][(ip + 1) % nc] = 0.5 * (lam + 1.0); }
        }
    }
    /* convection length to the upwind point (upwind.cpp:945-962) */
    for (int ip = 0; ip < nip; ++ip) {
        double up[3] = {0, 0, 0};
        for (int sh = 0; sh < nc; ++sh)
            for (int i = 0; i < 2; ++i) up[i] += U[ip][sh] * G->co[sh][i];
        for (int j = 0; j < nip; ++j)
            for (int i = 0; i < 2; ++i) up[i] += W[ip][j] * G->ipg[j][i];
        convlen[ip] = sqrt(dist2(2, G->ipg[ip], up));
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* upwind shapes: upwind.cpp:52-80 (no), 133-172 (full),                */
/* 337-430 (skewed + GetNodeNextToCut), 505-575 (LPS), 643-786          */
/* (positive).  W (ip shapes) may be NULL for the corner-only upwinds.  */
/* ------------------------------------------------------------------ */
static int upwind_geo(int type, const or_geo* G, const double (*vel)[3], double (*U)[8], double (*W)[12],
                      double* convlen)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh;
    if (type == OR_UPW_POSITIVE) {
        if (!W) { set_err("positive upwind needs ip shapes"); return -4; }
        positive_upwind(G, vel, U, W, convlen);
        return 0;
    }
    if (type == OR_UPW_REGULAR) {
        if (!W) { set_err("regular upwind needs ip shapes"); return -4; }
        return regular_upwind(G, vel, U, W, convlen);
    }
    if (W)
        for (int ip = 0; ip < nip; ++ip)
            for (int j = 0; j < 12; ++j) W[ip][j] = 0.0;
    double cflat[24];
    for (int k = 0; k < nsh; ++k)
        for (int i = 0; i < dim; ++i) cflat[k * dim + i] = G->co[k][i];
    for (int ip = 0; ip < nip; ++ip) {
        switch (type) {
        case OR_UPW_NO:
            for (int sh = 0; sh < nsh; ++sh) U[ip][sh] = G->phi[ip][sh];
            convlen[ip] = 1.0;
            break;
        case OR_UPW_FULL: {
            for (int sh = 0; sh < nsh; ++sh) U[ip][sh] = 0.0;
            const double flux = dot(dim, G->n[ip], vel[ip]);
            if (flux > 0.0) {
                U[ip][G->from[ip]] = 1.0;
                convlen[ip] = sqrt(dist2(dim, G->ipg[ip], G->co[G->from[ip]]));
            } else {
                U[ip][G->to[ip]] = 1.0;
                convlen[ip] = sqrt(dist2(dim, G->ipg[ip], G->co[G->to[ip]]));
            }
            break;
        }
        case OR_UPW_SKEWED:
        case OR_UPW_LPS: {
            for (int sh = 0; sh < nsh; ++sh) U[ip][sh] = 0.0;
            if (sqrt(dot(dim, vel[ip], vel[ip])) < 1e-14) { convlen[ip] = 1.0; break; }
            int side;
            double gi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
            if (or_side_ray(G->et, cflat, G->ipg[ip], vel[ip], &side, gi, li)) return -10;
            const refelem* R = &REF[G->et];
            if (type == OR_UPW_SKEWED) {
                double mn = DBL_MAX;
                int coOut = 0;
                for (int i = 0; i < R->side_n[side]; ++i) {
                    const int co = R->side[side][i];
                    const double dd = dist2(dim, gi, G->co[co]);
                    if (dd < mn) { mn = dd; coOut = co; }
                }
                U[ip][coOut] = 1.0;
                convlen[ip] = sqrt(dist2(dim, G->ipg[ip], G->co[coOut]));
            } else {
                double phi[8], gl[8][3];
                shapes(G->et, li, phi, gl);
                for (int j = 0; j < R->side_n[side]; ++j) {
                    const int co = R->side[side][j];
                    U[ip][co] = phi[co];
                }
                convlen[ip] = sqrt(dist2(dim, G->ipg[ip], gi));
            }
            break;
        }
        default:
            set_err("upwind not set");
            return -4;
        }
    }
    return 0;
}

int or_upwind(int type, int et, const double* corners, const double* ipvel, double* U, double* W, double* convlen)
{
    or_geo G;
    int rc = or_geometry(et, corners, &G);
    if (rc) return rc;
    double vel[12][3] = {{0}};
    for (int ip = 0; ip < G.nip; ++ip)
        for (int i = 0; i < G.dim; ++i) vel[ip][i] = ipvel[ip * G.dim + i];
    double Ul[12][8], Wl[12][12];
    rc = upwind_geo(type, &G, (const double (*)[3])vel, Ul, Wl, convlen);
    if (rc) return rc;
    for (int ip = 0; ip < G.nip; ++ip)
        for (int k = 0; k < 8; ++k) U[ip * 8 + k] = k < G.nsh ? Ul[ip][k] : 0.0;
    if (W)
        for (int ip = 0; ip < G.nip; ++ip)
            for (int j = 0; j < 12; ++j) W[ip * 12 + j] = j < G.nip ? Wl[ip][j] : 0.0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* diffusion length: diffusion_length.h:48-89 (FIVEPOINT),               */
/* 91-133 (RAW), 135-199 (COR); selection stabilization.cpp:102-116     */
/* ------------------------------------------------------------------ */
static void diff_length(int method, const or_geo* G, double* L2inv)
{
    const int dim = G->dim, nip = G->nip;
    if (method == OR_DL_COR) {
        double minN = DBL_MAX, minD = DBL_MAX, avgN = 0.0;
        for (int i = 0; i < nip; ++i) {
            const double nn = dot(dim, G->n[i], G->n[i]);
            if (nn < minN) minN = nn;
            avgN += nn;
            if (dim == 3) {
                const double ds = dist2(3, G->c0[i], G->c2[i]);
                if (ds < minD) minD = ds;
            }
        }
        avgN /= nip;
        for (int i = 0; i < nip; ++i) {
            double a = 0.5 * (G->vol[G->from[i]] + G->vol[G->to[i]]);
            a *= a;
            if (dim == 2) L2inv[i] = 2.0 * minN / a + 8.0 / (3.0 * avgN);
            else L2inv[i] = 2.0 * minN / a + 8.0 * minD / (3.0 * avgN);
        }
        return;
    }
    for (int i = 0; i < nip; ++i) {
        const double nn = dot(dim, G->n[i], G->n[i]);
        double a = 0.5 * (G->vol[G->from[i]] + G->vol[G->to[i]]);
        a *= a;
        if (method == OR_DL_RAW) {
            if (dim == 2) L2inv[i] = 1. / (0.5 * a / nn + 3.0 * nn / 8.);
            else L2inv[i] = 1. / (0.5 * a / nn + 3.0 * dist2(3, G->c0[i], G->c2[i]) / 8.);
        } else {
            if (dim == 2) L2inv[i] = 2.0 * nn / a + 8.0 / nn;
            else L2inv[i] = 2.0 * nn / a + 8.0 * dist2(3, G->c0[i], G->c2[i]) / nn;
        }
    }
}

/* ------------------------------------------------------------------ */
/* stabilisation update: FIELDS diagonal case stabilization.cpp:122-242, */
/* without stabilisation stabilization.cpp:805-850                       */
/* ------------------------------------------------------------------ */
typedef struct {
    double vel[12][3];
    double Svel[12][3][3][8]; /* stab_shape_vel(ip, d, d2, sh); d != d2 only when connected */
    double Sp[12][3][8];      /* stab_shape_p(ip, d, sh) */
    double diag[12];          /* divisor of the diagonal ip system */
    double cst[12][3];        /* the part of vel from the source and the old time step */
    int connected;            /* vel_comp_connected(): FLOW */
    int nondiag;              /* the ip system was solved as a matrix */
} stabdata;

/* upwind shapes of one upwind evaluation */
typedef struct {
    double U[12][8], W[12][12], len[12];
} updata;

static void u_at(const double* u, int nc, int d, int sh, double* out) { *out = u[sh * nc + d]; }
#define UL(d, sh) (uloc[(sh) * nc + (d)])
#define UO(d, sh) (uold[(sh) * nc + (d)])

/* inverse of the n x n ip system (Gauss-Jordan with partial pivoting);
   returns 0 if singular (GetInverse failing, stabilization.cpp:289) */
static int invert(int n, const double (*A)[12], double (*X)[12])
{
    double M[12][24];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j) M[i][j] = j < n ? A[i][j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0) return 0;
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) { const double t = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = t; }
        const double ip = 1.0 / M[c][c];
        for (int j = 0; j < 2 * n; ++j) M[c][j] *= ip;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 2 * n; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) X[i][j] = M[i][n + j];
    return 1;
}

/* x = A y */
static void matvec(int n, const double (*A)[12], const double* y, double* x)
{
    for (int i = 0; i < n; ++i) {
        double a = 0.0;
        for (int j = 0; j < n; ++j) a += A[i][j] * y[j];
        x[i] = a;
    }
}

/* Stabilisation update: FIELDS stabilization.cpp:122-403 (diagonal case
   :163-242, matrix case :244-403), FLOW :436-772 (diagonal :486-602, matrix
   :604-772), without stabilisation :805-850.  `up` holds the upwind of
   the stabilisation, `dn` its downwind (FLOW); the ip system is a matrix
   when the upwind has ip shapes (positive upwind) and the equations are
   not Stokes. */
static int stab_update(const or_params* p, const or_geo* G, const double* uloc, const double* uold, double dt,
                       const double (*stdvel)[3], const updata* up, const updata* dn, int ipshapes, stabdata* S)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh, nc = dim + 1;
    (void)u_at;
    memset(S->Svel, 0, sizeof S->Svel);
    memset(S->cst, 0, sizeof S->cst);
    S->connected = p->stab_type == OR_STAB_FLOW;
    S->nondiag = 0;
    if (p->stab_type == OR_STAB_WITHOUT) {
        for (int ip = 0; ip < nip; ++ip) {
            for (int i = 0; i < dim; ++i)
                for (int sh = 0; sh < nsh; ++sh) { S->Sp[ip][i][sh] = 0; S->Svel[ip][i][i][sh] = G->phi[ip][sh]; }
            for (int i = 0; i < dim; ++i) S->vel[ip][i] = stdvel[ip][i];
            S->diag[ip] = 1.0;
        }
        return 0;
    }
    if (p->stab_type != OR_STAB_FIELDS && p->stab_type != OR_STAB_FLOW) {
        set_err("Stabilization has not been set.");
        return -3;
    }
    const int flow = p->stab_type == OR_STAB_FLOW;
    double L2inv[12];
    diff_length(p->difflen, G, L2inv);
    double vvisc[12], vconv[12] = {0}, vdown[12] = {0};
    for (int ip = 0; ip < nip; ++ip) vvisc[ip] = nu_at(p, ip) * L2inv[ip];
    if (!p->stokes)
        for (int ip = 0; ip < nip; ++ip) {
            const double nrm = sqrt(dot(dim, stdvel[ip], stdvel[ip]));
            vconv[ip] = nrm / up->len[ip];
            if (flow) vdown[ip] = nrm / (dn->len[ip] + up->len[ip]);
        }
    double src[12][3] = {{0}};
    if (p->has_source)
        for (int ip = 0; ip < nip; ++ip)
            for (int d = 0; d < dim; ++d) src[ip][d] = src_scvf(p, ip, d, dim);
    double old[12][3] = {{0}};
    if (uold)
        for (int ip = 0; ip < nip; ++ip)
            for (int d = 0; d < dim; ++d) {
                double o = 0.0;
                for (int sh = 0; sh < nsh; ++sh) o += G->phi[ip][sh] * UO(d, sh);
                old[ip][d] = o;
            }

    if (p->stokes || !ipshapes) {
        /* diagonal case: ip by ip */
        for (int ip = 0; ip < nip; ++ip) {
            double diag = vvisc[ip];
            if (uold) diag += 1. / dt;
            if (!p->stokes) diag += vconv[ip];
            S->diag[ip] = diag;
            const double* sv = stdvel[ip];
            for (int d = 0; d < dim; ++d) {
                double rhs = src[ip][d];
                if (uold) rhs += old[ip][d] / dt;
                S->cst[ip][d] = rhs / diag;
                for (int k = 0; k < nsh; ++k) {
                    double sumVel = vvisc[ip] * G->phi[ip][k];
                    if (!p->stokes) {
                        sumVel += vconv[ip] * up->U[ip][k];
                        if (flow) sumVel += vdown[ip] * (dn->U[ip][k] - up->U[ip][k]);
                    }
                    if (flow)
                        for (int d2 = 0; d2 < dim; ++d2)
                            if (d2 != d) sumVel -= sv[d2] * G->g[ip][k][d2];
                    rhs += sumVel * UL(d, k);
                    S->Svel[ip][d][d][k] = sumVel / diag;
                    if (flow)
                        for (int d2 = 0; d2 < dim; ++d2) {
                            if (d2 == d) continue;
                            const double sumVel2 = sv[d] * G->g[ip][k][d2];
                            rhs += sumVel2 * UL(d2, k);
                            S->Svel[ip][d][d2][k] = sumVel2 / diag;
                        }
                    const double sumP = -1.0 * G->g[ip][k][d] / rho_at(p, ip);
                    rhs += sumP * UL(dim, k);
                    S->Sp[ip][d][k] = sumP / diag;
                }
                S->vel[ip][d] = rhs / diag;
            }
        }
        return 0;
    }

    /* matrix case */
    S->nondiag = 1;
    double mat[12][12] = {{0}}, inv[12][12];
    for (int ip = 0; ip < nip; ++ip) {
        if (uold) mat[ip][ip] += 1. / dt;
        mat[ip][ip] += vvisc[ip];
        mat[ip][ip] += vconv[ip];
        for (int ip2 = 0; ip2 < nip; ++ip2) {
            mat[ip][ip2] -= up->W[ip][ip2] * vconv[ip];
            if (flow) mat[ip][ip2] += vdown[ip] * (up->W[ip][ip2] - dn->W[ip][ip2]);
        }
        S->diag[ip] = mat[ip][ip];
    }
    if (!invert(nip, (const double (*)[12])mat, inv)) { set_err("Could not compute inverse."); return -11; }
    const double (*iv)[12] = (const double (*)[12])inv;
    for (int d = 0; d < dim; ++d) {
        double contVel[3][8][12], contP[8][12], x[12], f[12], f0[12], c0[12];
        for (int ip = 0; ip < nip; ++ip) {
            const double* sv = stdvel[ip];
            for (int k = 0; k < nsh; ++k) {
                contP[k][ip] = -1.0 * G->g[ip][k][d] / rho_at(p, ip);
                contVel[d][k][ip] = vvisc[ip] * G->phi[ip][k];
                contVel[d][k][ip] += vconv[ip] * up->U[ip][k];
                if (flow) {
                    contVel[d][k][ip] += vdown[ip] * (dn->U[ip][k] - up->U[ip][k]);
                    for (int d2 = 0; d2 < dim; ++d2) {
                        if (d2 == d) continue;
                        contVel[d][k][ip] -= sv[d2] * G->g[ip][k][d2];
                        contVel[d2][k][ip] = sv[d] * G->g[ip][k][d2];
                    }
                }
            }
        }
        for (int k = 0; k < nsh; ++k) {
            matvec(nip, iv, contP[k], x);
            for (int ip = 0; ip < nip; ++ip) S->Sp[ip][d][k] = x[ip];
            for (int d2 = 0; d2 < dim; ++d2) {
                if (!flow && d2 != d) continue;
                matvec(nip, iv, contVel[d2][k], x);
                for (int ip = 0; ip < nip; ++ip) S->Svel[ip][d][d2][k] = x[ip];
            }
        }
        for (int ip = 0; ip < nip; ++ip) {
            f0[ip] = src[ip][d];
            if (uold) f0[ip] += old[ip][d] / dt;
        }
        if (!flow) {
            /* FIELDS: source and old time first, then corner contributions */
            for (int ip = 0; ip < nip; ++ip) f[ip] = f0[ip];
            for (int k = 0; k < nsh; ++k)
                for (int ip = 0; ip < nip; ++ip) {
                    f[ip] = f[ip] + UL(d, k) * contVel[d][k][ip];
                    f[ip] = f[ip] + UL(dim, k) * contP[k][ip];
                }
        } else {
            /* FLOW: corner contributions first, then source and old time */
            for (int ip = 0; ip < nip; ++ip) f[ip] = 0.0;
            for (int k = 0; k < nsh; ++k)
                for (int ip = 0; ip < nip; ++ip) {
                    for (int d2 = 0; d2 < dim; ++d2) f[ip] = f[ip] + UL(d2, k) * contVel[d2][k][ip];
                    f[ip] = f[ip] + UL(dim, k) * contP[k][ip];
                }
            for (int ip = 0; ip < nip; ++ip) {
                if (p->has_source) f[ip] += src[ip][d];
                if (uold) f[ip] += old[ip][d] / dt;
            }
        }
        matvec(nip, iv, f, x);
        matvec(nip, iv, f0, c0);
        for (int ip = 0; ip < nip; ++ip) {
            S->vel[ip][d] = x[ip];
            S->cst[ip][d] = c0[ip];
        }
    }
    return 0;
}

/* peclet_blend: navier_stokes_fv1.cpp:871-892 */
static double peclet_blend(const or_params* p, const or_geo* G, int ip, const double* stdvel, double* upvel)
{
    const int dim = G->dim;
    const double Pe = dot(dim, stdvel, G->n[ip]) / dot(dim, G->n[ip], G->n[ip])
                    * sqrt(dist2(dim, G->co[G->to[ip]], G->co[G->from[ip]])) / nu_at(p, ip);
    const double Pe2 = Pe * Pe;
    const double w = Pe2 / (5.0 + Pe2);
    for (int i = 0; i < dim; ++i) upvel[i] = w * upvel[i] + (1.0 - w) * stdvel[i];
    return w;
}

/* common prologue of add_jac_A_elem / add_def_A_elem:
   navier_stokes_fv1.cpp:282-308 and 631-658 */
typedef struct {
    double stdvel[12][3];
    double Uc[12][8], Wc[12][12], lc[12]; /* convective upwind (if used) */
    int conv_ip;                          /* the convective upwind has ip shapes */
    stabdata S;
} prolog;

/* the stabilisation's update: its upwind (and for FLOW its downwind, the
   same upwind of the negated ip velocities, upwind_interface.h:157-165),
   then the ip system */
static int stab_prologue(const or_params* p, const or_geo* G, const double* uloc, const double* uold, double dt,
                         const double (*stdvel)[3], updata* st, stabdata* S)
{
    const int nip = G->nip, dim = G->dim;
    updata dn;
    memset(st, 0, sizeof *st);
    memset(&dn, 0, sizeof dn);
    for (int ip = 0; ip < 12; ++ip) st->len[ip] = dn.len[ip] = 1.0;
    if (!p->stokes) {
        if (p->stab_upwind < 0) { set_err("No upwind object"); return -4; }
        int rc = upwind_geo(p->stab_upwind, G, stdvel, st->U, st->W, st->len);
        if (rc) return rc;
        if (p->stab_type == OR_STAB_FLOW) {
            double neg[12][3] = {{0}};
            for (int ip = 0; ip < nip; ++ip)
                for (int d = 0; d < dim; ++d) neg[ip][d] = -1.0 * stdvel[ip][d];
            rc = upwind_geo(p->stab_upwind, G, (const double (*)[3])neg, dn.U, dn.W, dn.len);
            if (rc) return rc;
        }
    }
    return stab_update(p, G, uloc, uold, dt, stdvel, st, &dn, p->stab_upwind == OR_UPW_POSITIVE || p->stab_upwind == OR_UPW_REGULAR || p->force_matrix, S);
}

static int prologue(const or_params* p, const or_geo* G, const double* uloc, const double* uold, double dt, prolog* W)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh, nc = dim + 1;
    if (p->stab_type < 0) { set_err("Stabilization has not been set."); return -3; }
    if (!p->stokes && p->conv_upwind < 0 && !p->pac) {
        set_err("Upwinding for convective Term in Momentum eq. not set.");
        return -4;
    }
    for (int ip = 0; ip < nip; ++ip) {
        for (int d = 0; d < dim; ++d) W->stdvel[ip][d] = 0.0;
        for (int sh = 0; sh < nsh; ++sh)
            for (int d = 0; d < dim; ++d) W->stdvel[ip][d] += UL(d, sh) * G->phi[ip][sh];
    }
    updata st;
    int rc = stab_prologue(p, G, uloc, uold, dt, (const double (*)[3])W->stdvel, &st, &W->S);
    if (rc) return rc;
    memset(W->Wc, 0, sizeof W->Wc);
    W->conv_ip = 0;
    if (!p->stokes && !p->pac) {
        if (p->conv_upwind == p->stab_upwind) {
            memcpy(W->Uc, st.U, sizeof st.U);
            memcpy(W->Wc, st.W, sizeof st.W);
            memcpy(W->lc, st.len, sizeof st.len);
        } else {
            rc = upwind_geo(p->conv_upwind, G, (const double (*)[3])W->stdvel, W->Uc, W->Wc, W->lc);
            if (rc) return rc;
        }
        W->conv_ip = p->conv_upwind == OR_UPW_POSITIVE || p->conv_upwind == OR_UPW_REGULAR;
    }
    return 0;
}

/* upwind velocity of the convective term: upwind_interface.h:335-358 */
static void conv_upwind_vel(const or_geo* G, const double* uloc, const prolog* W, int ip, double* upv)
{
    const int dim = G->dim, nsh = G->nsh, nip = G->nip, nc = dim + 1;
    for (int d = 0; d < dim; ++d) upv[d] = 0.0;
    for (int k = 0; k < nsh; ++k)
        for (int d = 0; d < dim; ++d) upv[d] += W->Uc[ip][k] * UL(d, k);
    if (W->conv_ip)
        for (int j = 0; j < nip; ++j)
            for (int d = 0; d < dim; ++d) upv[d] += W->Wc[ip][j] * W->stdvel[j][d];
}

/* add_jac_A_elem: navier_stokes_fv1.cpp:250-595 */
static int jac_A(const or_params* p, const or_geo* G, const double* uloc, const prolog* W, double* Jl)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh, nc = dim + 1, P = dim, nl = nsh * nc;
#define JL(a, i, b, j) Jl[((i) * nc + (a)) * nl + (j) * nc + (b)]
    for (int ip = 0; ip < nip; ++ip) {
        const double nu = nu_at(p, ip), rho = rho_at(p, ip);
        const int fr = G->from[ip], to = G->to[ip];
        const double* n = G->n[ip];
        for (int sh = 0; sh < nsh; ++sh) {
            const double* gsh = G->g[ip][sh];
            /* diffusion :336-356 */
            const double flux_sh = -1.0 * nu * rho * dot(dim, gsh, n);
            for (int d1 = 0; d1 < dim; ++d1) { JL(d1, fr, d1, sh) += flux_sh; JL(d1, to, d1, sh) -= flux_sh; }
            if (!p->laplace)
                for (int d1 = 0; d1 < dim; ++d1)
                    for (int d2 = 0; d2 < dim; ++d2) {
                        const double f2 = -1.0 * nu * rho * gsh[d1] * n[d2];
                        JL(d1, fr, d2, sh) += f2; JL(d1, to, d2, sh) -= f2;
                    }
            /* pressure :363-368 */
            for (int d1 = 0; d1 < dim; ++d1) {
                const double fp = G->phi[ip][sh] * n[d1];
                JL(d1, fr, P, sh) += fp; JL(d1, to, P, sh) -= fp;
            }
            /* convection :375-553 */
            if (!p->stokes) {
                double upv[3] = {0, 0, 0};
                if (!p->pac) {
                    conv_upwind_vel(G, uloc, W, ip, upv);
                } else {
                    for (int d = 0; d < dim; ++d) upv[d] = W->S.vel[ip][d];
                }
                double w = 1.0;
                if (p->peclet) w = peclet_blend(p, G, ip, W->stdvel[ip], upv);
                const double prod = dot(dim, W->stdvel[ip], n) * rho;
                if (p->pac) {
                    if (W->S.connected)
                        for (int d1 = 0; d1 < dim; ++d1)
                            for (int d2 = 0; d2 < dim; ++d2) {
                                const double cv = prod * w * W->S.Svel[ip][d1][d2][sh];
                                JL(d1, fr, d2, sh) += cv; JL(d1, to, d2, sh) -= cv;
                            }
                    else
                        for (int d1 = 0; d1 < dim; ++d1) {
                            const double cv = prod * w * W->S.Svel[ip][d1][d1][sh];
                            JL(d1, fr, d1, sh) += cv; JL(d1, to, d1, sh) -= cv;
                        }
                    for (int d1 = 0; d1 < dim; ++d1) {
                        const double cp = prod * w * W->S.Sp[ip][d1][sh];
                        JL(d1, fr, P, sh) += cp; JL(d1, to, P, sh) -= cp;
                    }
                } else {
                    /* ip shapes folded into corner shapes (:437-447) */
                    double cv = W->Uc[ip][sh];
                    if (W->conv_ip)
                        for (int j = 0; j < nip; ++j) cv += G->phi[j][sh] * W->Wc[ip][j];
                    cv *= prod * w;
                    for (int d1 = 0; d1 < dim; ++d1) { JL(d1, fr, d1, sh) += cv; JL(d1, to, d1, sh) -= cv; }
                }
                if (p->peclet) {
                    const double cpe = prod * (1.0 - w) * G->phi[ip][sh];
                    for (int d1 = 0; d1 < dim; ++d1) { JL(d1, fr, d1, sh) += cpe; JL(d1, to, d1, sh) -= cpe; }
                }
                if (p->exact != 0.0) {
                    if (p->pac) {
                        for (int d1 = 0; d1 < dim; ++d1) {
                            for (int d2 = 0; d2 < dim; ++d2) {
                                double pv = 0.0;
                                if (W->S.connected)
                                    for (int k = 0; k < dim; ++k) pv += w * W->S.Svel[ip][k][d2][sh] * n[k];
                                else
                                    pv = W->S.Svel[ip][d1][d1][sh] * n[d1];
                                pv *= p->exact * rho;
                                JL(d1, fr, d2, sh) += pv * upv[d1]; JL(d1, to, d2, sh) -= pv * upv[d1];
                            }
                            double pp = 0.0;
                            for (int k = 0; k < dim; ++k) pp += W->S.Sp[ip][k][sh] * n[k];
                            pp *= p->exact * rho;
                            JL(d1, fr, P, sh) += pp * upv[d1]; JL(d1, to, P, sh) -= pp * upv[d1];
                        }
                    } else {
                        for (int d1 = 0; d1 < dim; ++d1)
                            for (int d2 = 0; d2 < dim; ++d2) {
                                const double pv = w * W->Uc[ip][sh] * n[d2] * rho;
                                JL(d1, fr, d2, sh) += pv * upv[d1]; JL(d1, to, d2, sh) -= pv * upv[d1];
                            }
                    }
                    if (p->peclet)
                        for (int d1 = 0; d1 < dim; ++d1)
                            for (int d2 = 0; d2 < dim; ++d2) {
                                const double cpe = upv[d1] * (1.0 - w) * G->phi[ip][sh] * n[d2] * rho;
                                JL(d1, fr, d2, sh) += cpe; JL(d1, to, d2, sh) -= cpe;
                            }
                }
            }
            /* continuity :561-592 */
            for (int d1 = 0; d1 < dim; ++d1) {
                double cv = 0.0;
                if (W->S.connected)
                    for (int d2 = 0; d2 < dim; ++d2) cv += W->S.Svel[ip][d2][d1][sh] * n[d2] * rho;
                else
                    cv = W->S.Svel[ip][d1][d1][sh] * n[d1] * rho;
                JL(P, fr, d1, sh) += cv; JL(P, to, d1, sh) -= cv;
            }
            double cpp = 0.0;
            for (int d1 = 0; d1 < dim; ++d1) cpp += W->S.Sp[ip][d1][sh] * n[d1] * rho;
            JL(P, fr, P, sh) += cpp; JL(P, to, P, sh) -= cpp;
        }
    }
#undef JL
    return 0;
}

/* add_def_A_elem: navier_stokes_fv1.cpp:597-778 */
static int def_A(const or_params* p, const or_geo* G, const double* uloc, const prolog* W, double* dl)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh, nc = dim + 1, P = dim;
#define DL(a, i) dl[(i) * nc + (a)]
    for (int ip = 0; ip < nip; ++ip) {
        const double nu = nu_at(p, ip), rho = rho_at(p, ip);
        const int fr = G->from[ip], to = G->to[ip];
        const double* n = G->n[ip];
        double gv[3][3];
        for (int d1 = 0; d1 < dim; ++d1)
            for (int d2 = 0; d2 < dim; ++d2) {
                gv[d1][d2] = 0.0;
                for (int sh = 0; sh < nsh; ++sh) gv[d1][d2] += G->g[ip][sh][d2] * UL(d1, sh);
            }
        double df[3];
        for (int d1 = 0; d1 < dim; ++d1) {
            df[d1] = 0.0;
            for (int d2 = 0; d2 < dim; ++d2) df[d1] += gv[d1][d2] * n[d2];
        }
        if (!p->laplace)
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2) df[d1] += gv[d2][d1] * n[d2];
        for (int d1 = 0; d1 < dim; ++d1) df[d1] *= (-1.0) * nu * rho;
        for (int d1 = 0; d1 < dim; ++d1) { DL(d1, fr) += df[d1]; DL(d1, to) -= df[d1]; }
        if (!p->stokes) {
            double upv[3] = {0, 0, 0};
            if (!p->pac) {
                conv_upwind_vel(G, uloc, W, ip, upv);
            } else {
                for (int d = 0; d < dim; ++d) upv[d] = W->S.vel[ip][d];
            }
            if (p->peclet) peclet_blend(p, G, ip, W->stdvel[ip], upv);
            const double prod = dot(dim, W->stdvel[ip], n) * rho;
            for (int d1 = 0; d1 < dim; ++d1) { DL(d1, fr) += upv[d1] * prod; DL(d1, to) -= upv[d1] * prod; }
        }
        double pr = 0.0;
        for (int sh = 0; sh < nsh; ++sh) pr += G->phi[ip][sh] * UL(P, sh);
        for (int d1 = 0; d1 < dim; ++d1) { DL(d1, fr) += pr * n[d1]; DL(d1, to) -= pr * n[d1]; }
        const double cf = dot(dim, W->S.vel[ip], n) * rho;
        DL(P, fr) += cf; DL(P, to) -= cf;
    }
#undef DL
    return 0;
}

/* ------------------------------------------------------------------ */
/* Term-magnitude scale: for every entry the sum of |terms| that the     */
/* restated loops add to it, with every interpolated factor (StdVel,    */
/* upwind velocity, velocity gradient, ip pressure, stabilised velocity,*/
/* ip flux) replaced by the sum of the magnitudes it is formed from.    */
/* This bounds the rounding of any evaluation order; the parity tests   */
/* state their tolerances against it.                                   */
/* ------------------------------------------------------------------ */
static void elem_A_abs(const or_params* p, const or_geo* G, const double* uloc, const double* uold, double dt,
                       const prolog* W, double* Ja, double* da)
{
    const int dim = G->dim, nip = G->nip, nsh = G->nsh, nc = dim + 1, P = dim, nl = nsh * nc;
#define JA(a, i, b, j) Ja[((i) * nc + (a)) * nl + (j) * nc + (b)]
#define DA(a, i) da[(i) * nc + (a)]
#define ADDJ(a, b, j, X) do { const double _x = fabs(X); if (Ja) { JA(a, fr, b, j) += _x; JA(a, to, b, j) += _x; } } while (0)
#define ADDD(a, X) do { const double _x = fabs(X); if (da) { DA(a, fr) += _x; DA(a, to) += _x; } } while (0)
    for (int ip = 0; ip < nip; ++ip) {
        const double nu = fabs(nu_at(p, ip)), rho = fabs(rho_at(p, ip));
        const int fr = G->from[ip], to = G->to[ip];
        double n[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) n[i] = fabs(G->n[ip][i]);
        double sva[3] = {0, 0, 0}, pra = 0, upa[3] = {0, 0, 0}, gva[3][3] = {{0}}, sta[3] = {0, 0, 0};
        for (int sh = 0; sh < nsh; ++sh) {
            for (int d = 0; d < dim; ++d) sva[d] += fabs(UL(d, sh) * G->phi[ip][sh]);
            pra += fabs(G->phi[ip][sh] * UL(P, sh));
            for (int d = 0; d < dim; ++d) upa[d] += fabs(W->Uc[ip][sh] * UL(d, sh));
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2) gva[d1][d2] += fabs(G->g[ip][sh][d2] * UL(d1, sh));
        }
        if (W->conv_ip)
            for (int j = 0; j < nip; ++j)
                for (int d = 0; d < dim; ++d) {
                    double a = 0;
                    for (int k = 0; k < nsh; ++k) a += fabs(G->phi[j][k] * UL(d, k));
                    upa[d] += fabs(W->Wc[ip][j]) * a;
                }
        /* stabilised velocity magnitude bound */
        if (p->stab_type == OR_STAB_WITHOUT) {
            for (int d = 0; d < dim; ++d) sta[d] = sva[d];
        } else if (W->S.connected || W->S.nondiag) {
            for (int d = 0; d < dim; ++d) {
                double a = fabs(W->S.cst[ip][d]);
                for (int k = 0; k < nsh; ++k) {
                    for (int d2 = 0; d2 < dim; ++d2) a += fabs(W->S.Svel[ip][d][d2][k] * UL(d2, k));
                    a += fabs(W->S.Sp[ip][d][k] * UL(P, k));
                }
                sta[d] = a;
            }
        } else {
            const double diag = W->S.diag[ip];
            for (int d = 0; d < dim; ++d) {
                double a = p->has_source ? fabs(src_scvf(p, ip, d, dim)) : 0.0;
                if (uold) {
                    double o = 0;
                    for (int sh = 0; sh < nsh; ++sh) o += fabs(G->phi[ip][sh] * UO(d, sh));
                    a += o / dt;
                }
                for (int k = 0; k < nsh; ++k) {
                    a += fabs(W->S.Svel[ip][d][d][k] * diag * UL(d, k));
                    a += fabs(G->g[ip][k][d] / rho_at(p, ip) * UL(P, k));
                }
                sta[d] = a / diag;
            }
        }
        if (p->pac)
            for (int d = 0; d < dim; ++d) upa[d] = sta[d];
        double proda = 0;
        for (int d = 0; d < dim; ++d) proda += sva[d] * n[d];
        proda *= rho;
        double w = 1.0, upv_w[3];
        for (int d = 0; d < dim; ++d) upv_w[d] = upa[d];
        if (p->peclet && !p->stokes) {
            double sv[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) sv[d] = W->stdvel[ip][d];
            double tmp[3] = {0, 0, 0};
            w = peclet_blend(p, G, ip, sv, tmp);
            for (int d = 0; d < dim; ++d) upv_w[d] = fabs(w) * upa[d] + fabs(1.0 - w) * sva[d];
        }
        /* defect */
        for (int d1 = 0; d1 < dim; ++d1) {
            double df = 0;
            for (int d2 = 0; d2 < dim; ++d2) df += gva[d1][d2] * n[d2];
            if (!p->laplace)
                for (int d2 = 0; d2 < dim; ++d2) df += gva[d2][d1] * n[d2];
            ADDD(d1, nu * rho * df);
            if (!p->stokes) ADDD(d1, upv_w[d1] * proda);
            ADDD(d1, pra * n[d1]);
        }
        {
            double cf = 0;
            for (int d = 0; d < dim; ++d) cf += sta[d] * n[d];
            ADDD(P, cf * rho);
        }
        /* Jacobian */
        for (int sh = 0; sh < nsh; ++sh) {
            const double* g = G->g[ip][sh];
            double gn = 0;
            for (int i = 0; i < dim; ++i) gn += fabs(g[i]) * n[i];
            for (int d1 = 0; d1 < dim; ++d1) {
                ADDJ(d1, d1, sh, nu * rho * gn);
                if (!p->laplace)
                    for (int d2 = 0; d2 < dim; ++d2) ADDJ(d1, d2, sh, nu * rho * g[d1] * n[d2]);
                ADDJ(d1, P, sh, G->phi[ip][sh] * n[d1]);
            }
            if (!p->stokes) {
                for (int d1 = 0; d1 < dim; ++d1) {
                    if (p->pac) {
                        if (W->S.connected)
                            for (int d2 = 0; d2 < dim; ++d2) ADDJ(d1, d2, sh, proda * w * W->S.Svel[ip][d1][d2][sh]);
                        else
                            ADDJ(d1, d1, sh, proda * w * W->S.Svel[ip][d1][d1][sh]);
                        ADDJ(d1, P, sh, proda * w * W->S.Sp[ip][d1][sh]);
                    } else {
                        double ue = fabs(W->Uc[ip][sh]);
                        if (W->conv_ip)
                            for (int j = 0; j < nip; ++j) ue += fabs(G->phi[j][sh] * W->Wc[ip][j]);
                        ADDJ(d1, d1, sh, ue * proda * w);
                    }
                    if (p->peclet) ADDJ(d1, d1, sh, proda * (1.0 - w) * G->phi[ip][sh]);
                    if (p->exact != 0.0) {
                        for (int d2 = 0; d2 < dim; ++d2) {
                            double pv;
                            if (p->pac && W->S.connected) {
                                pv = 0.0;
                                for (int k = 0; k < dim; ++k) pv += fabs(w * W->S.Svel[ip][k][d2][sh]) * n[k];
                                pv *= p->exact * rho;
                            } else {
                                pv = p->pac ? W->S.Svel[ip][d1][d1][sh] * n[d1] * p->exact * rho
                                            : w * W->Uc[ip][sh] * n[d2] * rho;
                            }
                            ADDJ(d1, d2, sh, pv * upv_w[d1]);
                            if (p->peclet) ADDJ(d1, d2, sh, upv_w[d1] * (1.0 - w) * G->phi[ip][sh] * n[d2] * rho);
                        }
                        if (p->pac) {
                            double pp = 0;
                            for (int k = 0; k < dim; ++k) pp += fabs(W->S.Sp[ip][k][sh]) * n[k];
                            ADDJ(d1, P, sh, pp * p->exact * rho * upv_w[d1]);
                        }
                    }
                }
            }
            double cpp = 0;
            for (int d1 = 0; d1 < dim; ++d1) {
                if (W->S.connected) {
                    double a = 0;
                    for (int d2 = 0; d2 < dim; ++d2) a += fabs(W->S.Svel[ip][d2][d1][sh]) * n[d2];
                    ADDJ(P, d1, sh, a * rho);
                } else {
                    ADDJ(P, d1, sh, W->S.Svel[ip][d1][d1][sh] * n[d1] * rho);
                }
                cpp += fabs(W->S.Sp[ip][d1][sh]) * n[d1] * rho;
            }
            ADDJ(P, P, sh, cpp);
        }
    }
#undef JA
#undef DA
#undef ADDJ
#undef ADDD
}

int or_elem_A_abs(const or_params* p, int et, const double* corners, const double* uloc,
                  const double* uold, double dt, double* Ja, double* da)
{
    or_geo G;
    int rc = or_geometry(et, corners, &G);
    if (rc) return rc;
    prolog W;
    rc = prologue(p, &G, uloc, uold, dt, &W);
    if (rc) return rc;
    elem_A_abs(p, &G, uloc, uold, dt, &W, Ja, da);
    return 0;
}

int or_elem_A(const or_params* p, int et, const double* corners, const double* uloc,
              const double* uold, double dt, double* Jl, double* dl)
{
    or_geo G;
    int rc = or_geometry(et, corners, &G);
    if (rc) return rc;
    prolog W;
    rc = prologue(p, &G, uloc, uold, dt, &W);
    if (rc) return rc;
    if (Jl) jac_A(p, &G, uloc, &W, Jl);
    if (dl) def_A(p, &G, uloc, &W, dl);
    return 0;
}

/* add_jac_M_elem / add_def_M_elem: navier_stokes_fv1.cpp:781-838 */
int or_elem_M(const or_params* p, int et, const double* corners, const double* uloc, double* Jl, double* dl)
{
    or_geo G;
    int rc = or_geometry(et, corners, &G);
    if (rc) return rc;
    const int dim = G.dim, nsh = G.nsh, nc = dim + 1, nl = nsh * nc;
    for (int sh = 0; sh < nsh; ++sh)
        for (int d1 = 0; d1 < dim; ++d1) {
            if (Jl) Jl[(sh * nc + d1) * nl + sh * nc + d1] += G.vol[sh] * rho_sh(p, sh);
            if (dl) dl[sh * nc + d1] += UL(d1, sh) * G.vol[sh] * rho_sh(p, sh);
        }
    return 0;
}

/* add_rhs_elem: navier_stokes_fv1.cpp:841-869 */
int or_elem_rhs(const or_params* p, int et, const double* corners, double* fl)
{
    if (!p->has_source) return 0;
    or_geo G;
    int rc = or_geometry(et, corners, &G);
    if (rc) return rc;
    const int dim = G.dim, nsh = G.nsh, nc = dim + 1;
    for (int sh = 0; sh < nsh; ++sh)
        for (int d1 = 0; d1 < dim; ++d1) fl[sh * nc + d1] += src_scv(p, sh, d1, dim) * G.vol[sh] * rho_sh(p, sh);
    return 0;
}

/* ------------------------------------------------------------------ */
/* global scatter (ugcore's local->global add, restated)               */
/* ------------------------------------------------------------------ */
static int64_t find_col(const int64_t* rowptr, const int32_t* colidx, int64_t row, int32_t col)
{
    int64_t lo = rowptr[row], hi = rowptr[row + 1] - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) / 2;
        if (colidx[mid] == col) return mid;
        if (colidx[mid] < col) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* one element's local defect/Jacobian (or their magnitudes, absmode) added
   into the global vector and CSR values */
static int assemble_elem(const or_params* p, int et, const double* xyz, const int32_t* conn, int64_t e,
                         const double* u, const double* uold, double dt,
                         double s_a, double s_m, double s_f, int parts,
                         const int64_t* rowptr, const int32_t* colidx,
                         double* J, double* d, int absmode)
{
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1, nl = nsh * nc, nip = REF[et].nip;
    double co[24], ul[32], uo[32], Jl[1024], Ja[1024], dl[32], da[32], fl[32];
    const int32_t* c = conn + e * nsh;
    /* this element's rows of the integration-point imports */
    or_params q = *p;
    if (q.nu_ip) q.nu_ip += e * nip;
    if (q.rho_ip) q.rho_ip += e * nip;
    if (q.rho_scv) q.rho_scv += e * nsh;
    if (q.f_scv) q.f_scv += e * nsh * dim;
    if (q.f_scvf) q.f_scvf += e * nip * dim;
    p = &q;
    for (int k = 0; k < nsh; ++k) {
        for (int i = 0; i < dim; ++i) co[k * dim + i] = xyz[(int64_t)c[k] * dim + i];
        for (int a = 0; a < nc; ++a) {
            ul[k * nc + a] = u[(int64_t)c[k] * nc + a];
            if (uold) uo[k * nc + a] = uold[(int64_t)c[k] * nc + a];
        }
    }
    memset(Jl, 0, sizeof(double) * nl * nl);
    memset(dl, 0, sizeof(double) * nl);
    memset(Ja, 0, sizeof(double) * nl * nl);
    memset(da, 0, sizeof(double) * nl);
    int rc;
    if (parts & 1) {
        double Jt[1024], dt_[32];
        memset(Jt, 0, sizeof(double) * nl * nl);
        memset(dt_, 0, sizeof(double) * nl);
        rc = or_elem_A(p, et, co, ul, uold ? uo : NULL, dt, J ? Jt : NULL, d ? dt_ : NULL);
        if (rc) return rc;
        if (absmode) {
            double Jb[1024], db[32];
            memset(Jb, 0, sizeof(double) * nl * nl);
            memset(db, 0, sizeof(double) * nl);
            rc = or_elem_A_abs(p, et, co, ul, uold ? uo : NULL, dt, J ? Jb : NULL, d ? db : NULL);
            if (rc) return rc;
            for (int i = 0; i < nl * nl; ++i) Ja[i] += fabs(s_a) * Jb[i];
            for (int i = 0; i < nl; ++i) da[i] += fabs(s_a) * db[i];
        }
        for (int i = 0; i < nl * nl; ++i) Jl[i] += s_a * Jt[i];
        for (int i = 0; i < nl; ++i) dl[i] += s_a * dt_[i];
    }
    if (parts & 2) {
        double Jt[1024], dt_[32];
        memset(Jt, 0, sizeof(double) * nl * nl);
        memset(dt_, 0, sizeof(double) * nl);
        rc = or_elem_M(p, et, co, ul, J ? Jt : NULL, d ? dt_ : NULL);
        if (rc) return rc;
        for (int i = 0; i < nl * nl; ++i) { Jl[i] += s_m * Jt[i]; Ja[i] += fabs(s_m * Jt[i]); }
        for (int i = 0; i < nl; ++i) { dl[i] += s_m * dt_[i]; da[i] += fabs(s_m * dt_[i]); }
    }
    if ((parts & 4) && d) {
        memset(fl, 0, sizeof(double) * nl);
        rc = or_elem_rhs(p, et, co, fl);
        if (rc) return rc;
        for (int i = 0; i < nl; ++i) { dl[i] += s_f * fl[i]; da[i] += fabs(s_f * fl[i]); }
    }
    const double* Jsrc = absmode ? Ja : Jl;
    const double* dsrc = absmode ? da : dl;
    for (int i = 0; i < nsh; ++i)
        for (int a = 0; a < nc; ++a) {
            const int64_t row = (int64_t)c[i] * nc + a;
            if (d) d[row] += dsrc[i * nc + a];
            if (J)
                for (int j = 0; j < nsh; ++j)
                    for (int b = 0; b < nc; ++b) {
                        const int64_t pos = find_col(rowptr, colidx, row, (int32_t)((int64_t)c[j] * nc + b));
                        if (pos < 0) { set_err("pattern misses an entry"); return -6; }
                        J[pos] += Jsrc[(i * nc + a) * nl + j * nc + b];
                    }
        }
    return 0;
}

static int assemble_impl(const or_params* p, int et, const double* xyz, int64_t nv,
                         const int32_t* conn, int64_t e0, int64_t e1,
                         const double* u, const double* uold, double dt,
                         double s_a, double s_m, double s_f, int parts,
                         const int64_t* rowptr, const int32_t* colidx,
                         double* J, double* d, int absmode)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const int dim = REF[et].dim, nc = dim + 1;
    if (p->dim != dim) { set_err("Wrong number of functions"); return -5; }
    const int64_t ndof = nv * nc;
    if (J) memset(J, 0, sizeof(double) * rowptr[ndof]);
    if (d) memset(d, 0, sizeof(double) * ndof);
    for (int64_t e = e0; e < e1; ++e) {
        const int rc = assemble_elem(p, et, xyz, conn, e, u, uold, dt, s_a, s_m, s_f, parts, rowptr, colidx, J, d,
                                     absmode);
        if (rc) return rc;
    }
    return 0;
}

/* The same assembly with the elements taken colour by colour
   (order[color_ptr[c] .. color_ptr[c+1]) holds colour c) and the elements of
   one colour spread over OpenMP threads: no two elements of a colour may
   share a node, so no two threads add into the same row.  The CPU baseline's
   all-cores restatement of UG4's one-rank-per-core element loop; the sums
   per entry then follow the colour order, not the element order. */
int or_assemble_colored(const or_params* p, int et, const double* xyz, int64_t nv,
                        const int32_t* conn, const int64_t* order, const int64_t* color_ptr, int ncolors,
                        const double* u, const double* uold, double dt,
                        double s_a, double s_m, double s_f, int parts,
                        const int64_t* rowptr, const int32_t* colidx,
                        double* J, double* d, int nthreads)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const int dim = REF[et].dim, nc = dim + 1;
    if (p->dim != dim) { set_err("Wrong number of functions"); return -5; }
    const int64_t ndof = nv * nc;
    if (J) memset(J, 0, sizeof(double) * rowptr[ndof]);
    if (d) memset(d, 0, sizeof(double) * ndof);
    int err = 0;
    for (int c = 0; c < ncolors && !err; ++c) {
        const int64_t a = color_ptr[c], b = color_ptr[c + 1];
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(min : err)
        for (int64_t i = a; i < b; ++i) {
            const int rc = assemble_elem(p, et, xyz, conn, order[i], u, uold, dt, s_a, s_m, s_f, parts, rowptr,
                                         colidx, J, d, 0);
            if (rc < err) err = rc;
        }
    }
    return err;
}

int or_assemble(const or_params* p, int et, const double* xyz, int64_t nv,
                const int32_t* conn, int64_t e0, int64_t e1,
                const double* u, const double* uold, double dt,
                double s_a, double s_m, double s_f, int parts,
                const int64_t* rowptr, const int32_t* colidx,
                double* J, double* d)
{
    return assemble_impl(p, et, xyz, nv, conn, e0, e1, u, uold, dt, s_a, s_m, s_f, parts,
                         rowptr, colidx, J, d, 0);
}

int or_assemble_abs(const or_params* p, int et, const double* xyz, int64_t nv,
                    const int32_t* conn, int64_t ne,
                    const double* u, const double* uold, double dt,
                    double s_a, double s_m, int parts,
                    const int64_t* rowptr, const int32_t* colidx,
                    double* Jabs, double* dabs)
{
    return assemble_impl(p, et, xyz, nv, conn, 0, ne, u, uold, dt, s_a, s_m, 1.0, parts,
                         rowptr, colidx, Jabs, dabs, 1);
}

/* ------------------------------------------------------------------ */
/* CSR pattern: node adjacency closure (x) (dim+1)^2, sorted columns    */
/* ------------------------------------------------------------------ */
static int cmp_i32(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

int or_pattern(int et, int64_t nv, const int32_t* conn, int64_t ne,
               int64_t* rowptr, int32_t* colidx, int64_t* nnz_out)
{
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1;
    /* node -> element incidence */
    int64_t* cnt = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
    for (int64_t e = 0; e < ne; ++e)
        for (int k = 0; k < nsh; ++k) cnt[conn[e * nsh + k] + 1]++;
    for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nv);
    int64_t* inc = (int64_t*)malloc(sizeof(int64_t) * (size_t)cnt[nv]);
    for (int64_t v = 0; v < nv; ++v) fill[v] = cnt[v];
    for (int64_t e = 0; e < ne; ++e)
        for (int k = 0; k < nsh; ++k) inc[fill[conn[e * nsh + k]]++] = e;
    int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * 512);
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int64_t v = 0; v < nv; ++v) {
        int m = 0;
        for (int64_t q = cnt[v]; q < cnt[v + 1]; ++q)
            for (int k = 0; k < nsh; ++k) {
                const int32_t w = conn[inc[q] * nsh + k];
                int dup = 0;
                for (int t = 0; t < m; ++t) if (nb[t] == w) { dup = 1; break; }
                if (!dup && m < 512) nb[m++] = w;
            }
        if (m == 0) nb[m++] = (int32_t)v; /* isolated node: keep the diagonal */
        qsort(nb, (size_t)m, sizeof(int32_t), cmp_i32);
        for (int a = 0; a < nc; ++a) {
            const int64_t row = v * nc + a;
            if (colidx)
                for (int t = 0; t < m; ++t)
                    for (int b = 0; b < nc; ++b) colidx[rowptr[row] + t * nc + b] = nb[t] * nc + b;
            rowptr[row + 1] = rowptr[row] + (int64_t)m * nc;
        }
        nnz += (int64_t)m * nc * nc;
    }
    *nnz_out = nnz;
    free(cnt); free(fill); free(inc); free(nb);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FV1 boundary faces (ugcore FV1Geometry BF, restated): on side s of  */
/* the element, corner c owns the part of the side inside its SCV.     */
/* 2D: [x_c, edge midpoint]; 3D: [x_c, mid(c,next), side centre,       */
/* mid(prev,c)].  ip = average of the part's corners, normal = outward */
/* area vector (2D: rotated segment; 3D: half the cross product of the */
/* diagonals), shapes and global gradients at the local ip.            */
/* ------------------------------------------------------------------ */
int or_bnd_faces(int et, const double* corners, int side, or_bfgeo* B)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const refelem* R = &REF[et];
    if (side < 0 || side >= R->nside) { set_err("side index out of range"); return -1; }
    const int dim = R->dim, nsh = R->nsh, k = R->side_n[side];
    memset(B, 0, sizeof *B);
    double co[8][3] = {{0}}, c[3] = {0, 0, 0};
    for (int q = 0; q < nsh; ++q)
        for (int i = 0; i < dim; ++i) { co[q][i] = corners[q * dim + i]; c[i] += co[q][i] / nsh; }
    double fc[3] = {0, 0, 0}, fcl[3] = {0, 0, 0};
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < dim; ++i) {
            fc[i] += co[R->side[side][j]][i] / k;
            fcl[i] += R->lc[R->side[side][j]][i] / k;
        }
    B->nbf = k;
    for (int j = 0; j < k; ++j) {
        const int cc = R->side[side][j];
        const int nx = R->side[side][(j + 1) % k], pv = R->side[side][(j + k - 1) % k];
        double p[4][3] = {{0}}, pl[4][3] = {{0}};
        int np;
        if (dim == 2) {
            /* the side is the edge (side[0], side[1]); the part of corner cc ends at the midpoint */
            const int other = (j == 0) ? R->side[side][1] : R->side[side][0];
            for (int i = 0; i < 2; ++i) {
                p[0][i] = co[cc][i];
                p[1][i] = 0.5 * (co[cc][i] + co[other][i]);
                pl[0][i] = R->lc[cc][i];
                pl[1][i] = 0.5 * (R->lc[cc][i] + R->lc[other][i]);
            }
            np = 2;
        } else {
            for (int i = 0; i < 3; ++i) {
                p[0][i] = co[cc][i];
                p[1][i] = 0.5 * (co[cc][i] + co[nx][i]);
                p[2][i] = fc[i];
                p[3][i] = 0.5 * (co[pv][i] + co[cc][i]);
                pl[0][i] = R->lc[cc][i];
                pl[1][i] = 0.5 * (R->lc[cc][i] + R->lc[nx][i]);
                pl[2][i] = fcl[i];
                pl[3][i] = 0.5 * (R->lc[pv][i] + R->lc[cc][i]);
            }
            np = 4;
        }
        B->node[j] = cc;
        for (int i = 0; i < dim; ++i) {
            double a = 0, b = 0;
            for (int q = 0; q < np; ++q) { a += p[q][i]; b += pl[q][i]; }
            B->ipg[j][i] = a / np;
            B->ipl[j][i] = b / np;
        }
        double n[3] = {0, 0, 0};
        if (dim == 2) {
            n[0] = p[1][1] - p[0][1];
            n[1] = -(p[1][0] - p[0][0]);
        } else {
            double d1[3], d2[3];
            for (int i = 0; i < 3; ++i) { d1[i] = p[2][i] - p[0][i]; d2[i] = p[3][i] - p[1][i]; }
            cross(d1, d2, n);
            for (int i = 0; i < 3; ++i) n[i] *= 0.5;
        }
        double out[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) out[i] = B->ipg[j][i] - c[i];
        if (dot(dim, n, out) < 0)
            for (int i = 0; i < dim; ++i) n[i] = -n[i];
        for (int i = 0; i < dim; ++i) B->n[j][i] = n[i];
        /* shapes and global gradients at the local ip */
        double gl[8][3] = {{0}};
        shapes(et, B->ipl[j], B->phi[j], gl);
        double Jm[3][3] = {{0}};
        for (int q = 0; q < nsh; ++q)
            for (int i = 0; i < dim; ++i)
                for (int jj = 0; jj < dim; ++jj) Jm[i][jj] += co[q][i] * gl[q][jj];
        double Ji[3][3] = {{0}};
        if (dim == 2) {
            const double det = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
            if (det == 0.0) { set_err("degenerate element"); return -2; }
            Ji[0][0] = Jm[1][1] / det;  Ji[0][1] = -Jm[0][1] / det;
            Ji[1][0] = -Jm[1][0] / det; Ji[1][1] = Jm[0][0] / det;
        } else {
            const double det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1])
                             - Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[1][2] * Jm[2][0])
                             + Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0]);
            if (det == 0.0) { set_err("degenerate element"); return -2; }
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
        for (int q = 0; q < nsh; ++q)
            for (int i = 0; i < dim; ++i) {
                double s = 0;
                for (int jj = 0; jj < dim; ++jj) s += Ji[jj][i] * gl[q][jj];
                B->g[j][q][i] = s;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* NavierStokesNoNormalStressOutflowFV1 (restated):                    */
/* incompressible/fv1/bnd/no_normal_stress_outflow_fv1.cpp:188-426.   */
/* Per boundary face (node c, normal n, ip):                           */
/*   momentum: tangential part of -nu rho (grad u + grad u^T) n, with  */
/*     the normal part removed as F - (F.n) n (n is the area vector,   */
/*     as in :261); convection max(rho u.n, 0) u (:305-309, 329-335);  */
/*   continuity: rho u.n (:424).                                       */
/* Jacobian: the same terms with the convective factor frozen          */
/* (:291-311, 360-368).  Element-local J (nl x nl) and d (nl), ADDED.  */
/* ------------------------------------------------------------------ */
int or_outflow_elem(const or_params* p, int et, const double* corners, int side, const double* uloc,
                    double* Jl, double* dl)
{
    or_bfgeo B;
    int rc = or_bnd_faces(et, corners, side, &B);
    if (rc) return rc;
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1, nl = nsh * nc;
    const double rho = p->rho;
    for (int f = 0; f < B.nbf; ++f) {
        const double nu = p->nu_bf ? p->nu_bf[f] : p->nu;  /* the import at the face's ip */
        const double* n = B.n[f];
        const int row = B.node[f];
        double vel[3] = {0, 0, 0};
        for (int d1 = 0; d1 < dim; ++d1)
            for (int sh = 0; sh < nsh; ++sh) vel[d1] += uloc[sh * nc + d1] * B.phi[f][sh];
        double flux = dot(dim, vel, n) * rho;
        if (flux < 0) flux = 0;
        if (dl) {
            double gv[3][3] = {{0}}, F[3] = {0, 0, 0};
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2)
                    for (int sh = 0; sh < nsh; ++sh) gv[d1][d2] += B.g[f][sh][d2] * uloc[sh * nc + d1];
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2) F[d1] += gv[d1][d2] * n[d2];
            if (!p->laplace)
                for (int d1 = 0; d1 < dim; ++d1)
                    for (int d2 = 0; d2 < dim; ++d2) F[d1] += gv[d2][d1] * n[d2];
            const double fn = dot(dim, F, n);
            for (int d1 = 0; d1 < dim; ++d1) F[d1] -= fn * n[d1];
            for (int d1 = 0; d1 < dim; ++d1) {
                dl[row * nc + d1] += F[d1] * (-nu * rho);
                if (!p->stokes) dl[row * nc + d1] += flux * vel[d1];
            }
            dl[row * nc + dim] += dot(dim, vel, n) * rho;
        }
        if (Jl) {
            for (int sh = 0; sh < nsh; ++sh) {
                const double* g = B.g[f][sh];
                double df[3][3] = {{0}}, ns[3] = {0, 0, 0};
                const double gn = dot(dim, g, n);
                for (int d1 = 0; d1 < dim; ++d1) df[d1][d1] = gn;
                if (!p->laplace)
                    for (int d1 = 0; d1 < dim; ++d1)
                        for (int d2 = 0; d2 < dim; ++d2) df[d1][d2] += g[d1] * n[d2];
                for (int d2 = 0; d2 < dim; ++d2)
                    for (int d1 = 0; d1 < dim; ++d1) ns[d2] += df[d1][d2] * n[d1];
                for (int d2 = 0; d2 < dim; ++d2)
                    for (int d1 = 0; d1 < dim; ++d1) df[d1][d2] -= n[d1] * ns[d2];
                for (int d1 = 0; d1 < dim; ++d1)
                    for (int d2 = 0; d2 < dim; ++d2)
                        Jl[(row * nc + d1) * nl + sh * nc + d2] += df[d1][d2] * (-nu * rho);
                if (!p->stokes)
                    for (int d1 = 0; d1 < dim; ++d1) Jl[(row * nc + d1) * nl + sh * nc + d1] += flux * B.phi[f][sh];
                for (int d2 = 0; d2 < dim; ++d2)
                    Jl[(row * nc + dim) * nl + sh * nc + d2] += B.phi[f][sh] * n[d2] * rho;
            }
        }
    }
    return 0;
}

/* adds s_a * (outflow contributions of the listed boundary sides) to J and d */
/* element-local boundary contributions of one kind (0 outflow, 1 symmetry,
   2 sliding wall with factor k and limit lim) */
static int bnd_elem(int kind, const or_params* p, int et, const double* co, int side, const double* ul, double k,
                    double lim, double* Jl, double* dl);

/* adds s_a * (contributions of the listed boundary sides) to J and d */
static int bnd_scatter(int kind, const or_params* p, int et, const double* xyz, const int32_t* conn,
                       const int32_t* belem, const int32_t* bside, int64_t nb, const double* u, double s_a,
                       double k, double lim, const int64_t* rowptr, const int32_t* colidx, double* J, double* d)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1, nl = nsh * nc;
    double co[24], ul[32], Jl[1024], dl[32];
    for (int64_t b = 0; b < nb; ++b) {
        const int32_t* c = conn + (int64_t)belem[b] * nsh;
        for (int q = 0; q < nsh; ++q) {
            for (int i = 0; i < dim; ++i) co[q * dim + i] = xyz[(int64_t)c[q] * dim + i];
            for (int a = 0; a < nc; ++a) ul[q * nc + a] = u[(int64_t)c[q] * nc + a];
        }
        memset(Jl, 0, sizeof(double) * nl * nl);
        memset(dl, 0, sizeof(double) * nl);
        or_params q = *p;
        if (p->nu_bf) q.nu_bf = p->nu_bf + b * 4;  /* the side's own faces */
        int rc = bnd_elem(kind, &q, et, co, bside[b], ul, k, lim, J ? Jl : NULL, d ? dl : NULL);
        if (rc) return rc;
        for (int i = 0; i < nsh; ++i)
            for (int a = 0; a < nc; ++a) {
                const int64_t row = (int64_t)c[i] * nc + a;
                if (d) d[row] += s_a * dl[i * nc + a];
                if (J)
                    for (int j = 0; j < nsh; ++j)
                        for (int bb = 0; bb < nc; ++bb) {
                            const int64_t pos = find_col(rowptr, colidx, row, (int32_t)((int64_t)c[j] * nc + bb));
                            if (pos < 0) { set_err("pattern misses an entry"); return -6; }
                            J[pos] += s_a * Jl[(i * nc + a) * nl + j * nc + bb];
                        }
            }
    }
    return 0;
}

int or_outflow(const or_params* p, int et, const double* xyz, int64_t nv, const int32_t* conn,
               const int32_t* belem, const int32_t* bside, int64_t nb, const double* u, double s_a,
               const int64_t* rowptr, const int32_t* colidx, double* J, double* d)
{
    (void)nv;
    return bnd_scatter(0, p, et, xyz, conn, belem, bside, nb, u, s_a, 0.0, 0.0, rowptr, colidx, J, d);
}

/* ------------------------------------------------------------------ */
/* NavierStokesSymBCFV1 (restated):                                    */
/* incompressible/fv1/bnd/symmetric_boundary_fv1.cpp:224-386.          */
/* Per boundary face (node c, area vector n, ip):                      */
/*   momentum d1: F[d1] n[d1] with F = -nu rho (grad u + grad u^T) n   */
/*     (component-wise product, as :373-375), plus p(ip) n[d1];        */
/*   continuity: nothing (u.n = 0 on the boundary).                    */
/* Jacobian (:290-317) matches; element-local, ADDED.                  */
/* ------------------------------------------------------------------ */
int or_symmetry_elem(const or_params* p, int et, const double* corners, int side, const double* uloc,
                     double* Jl, double* dl)
{
    or_bfgeo B;
    int rc = or_bnd_faces(et, corners, side, &B);
    if (rc) return rc;
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1, nl = nsh * nc;
    for (int f = 0; f < B.nbf; ++f) {
        const double nr = -1.0 * (p->nu_bf ? p->nu_bf[f] : p->nu) * p->rho;  /* the import at the face's ip */
        const double* n = B.n[f];
        const int row = B.node[f];
        if (dl) {
            double gv[3][3] = {{0}}, F[3] = {0, 0, 0};
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2)
                    for (int sh = 0; sh < nsh; ++sh) gv[d1][d2] += B.g[f][sh][d2] * uloc[sh * nc + d1];
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2) F[d1] += gv[d1][d2] * n[d2];
            if (!p->laplace)
                for (int d1 = 0; d1 < dim; ++d1)
                    for (int d2 = 0; d2 < dim; ++d2) F[d1] += gv[d2][d1] * n[d2];
            double pr = 0;
            for (int sh = 0; sh < nsh; ++sh) pr += B.phi[f][sh] * uloc[sh * nc + dim];
            for (int d1 = 0; d1 < dim; ++d1) dl[row * nc + d1] += F[d1] * nr * n[d1] + pr * n[d1];
        }
        if (Jl)
            for (int sh = 0; sh < nsh; ++sh) {
                const double* g = B.g[f][sh];
                const double fl = nr * dot(dim, g, n);
                for (int d1 = 0; d1 < dim; ++d1) Jl[(row * nc + d1) * nl + sh * nc + d1] += fl * n[d1];
                if (!p->laplace)
                    for (int d1 = 0; d1 < dim; ++d1)
                        for (int d2 = 0; d2 < dim; ++d2)
                            Jl[(row * nc + d1) * nl + sh * nc + d2] += nr * g[d1] * n[d2] * n[d1];
                for (int d1 = 0; d1 < dim; ++d1) Jl[(row * nc + d1) * nl + sh * nc + dim] += B.phi[f][sh] * n[d1];
            }
    }
    return 0;
}

int or_symmetry(const or_params* p, int et, const double* xyz, int64_t nv, const int32_t* conn,
                const int32_t* belem, const int32_t* bside, int64_t nb, const double* u, double s_a,
                const int64_t* rowptr, const int32_t* colidx, double* J, double* d)
{
    (void)nv;
    return bnd_scatter(1, p, et, xyz, conn, belem, bside, nb, u, s_a, 0.0, 0.0, rowptr, colidx, J, d);
}

/* ------------------------------------------------------------------ */
/* NavierStokesWSBCFV1, the sliding wall (restated):                   */
/* incompressible/fv1/bnd/wall_sliding_fv1.cpp:229-425.  Per boundary  */
/* face (node c, area vector n, unit normal nh = n/|n|, ip):           */
/*  convection (not Stokes): v = sum_sh u_sh phi_sh, t = v - (v.nh)nh, */
/*    f = (t.n) rho; d(d1,c) += f t[d1]; J(d1,c,d1,sh) += f phi_sh     */
/*    (t is tangential, so f is a rounding residue, kept as written);  */
/*  sliding factor k: w = sum over the FIRST dim shapes of u_sh phi_sh */
/*    (the loop bound of :391-397), t = w - (w.nh)nh,                  */
/*    d(d1,c) -= k t[d1]; J(d1,c,d2,sh) -= k P(d1,d2) for every sh,    */
/*    with P = (sum_sh phi_sh) I - (sum_{sh<dim} phi_sh) nh nh^T       */
/*    (:328-344);                                                      */
/*  sliding limit lim: if |t| != 0, d(d1,c) -= lim t[d1]/|t|;          */
/*    J(d1,c,d1,sh) -= lim s[d1] for every sh, s = (1,..,1)(sum phi)   */
/*    normalised (:350-365).                                           */
/* The Jacobian is the reference's approximation, not d'(u).           */
/* ------------------------------------------------------------------ */
int or_sliding_elem(const or_params* p, int et, const double* corners, int side, const double* uloc, double k,
                    double lim, double* Jl, double* dl)
{
    or_bfgeo B;
    int rc = or_bnd_faces(et, corners, side, &B);
    if (rc) return rc;
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1, nl = nsh * nc;
    const double rho = p->rho;
    for (int f = 0; f < B.nbf; ++f) {
        const double* n = B.n[f];
        const int row = B.node[f];
        const double len = sqrt(dot(dim, n, n));
        double nh[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) nh[i] = n[i] / len;
        if (!p->stokes) {
            double v[3] = {0, 0, 0};
            for (int d1 = 0; d1 < dim; ++d1)
                for (int sh = 0; sh < nsh; ++sh) v[d1] += uloc[sh * nc + d1] * B.phi[f][sh];
            const double vn = dot(dim, v, nh);
            for (int d1 = 0; d1 < dim; ++d1) v[d1] -= vn * nh[d1];
            const double flux = dot(dim, v, n) * rho;
            if (dl)
                for (int d1 = 0; d1 < dim; ++d1) dl[row * nc + d1] += flux * v[d1];
            if (Jl)
                for (int sh = 0; sh < nsh; ++sh)
                    for (int d1 = 0; d1 < dim; ++d1) Jl[(row * nc + d1) * nl + sh * nc + d1] += flux * B.phi[f][sh];
        }
        double sum_all = 0.0, sum_dim = 0.0;
        for (int sh = 0; sh < nsh; ++sh) sum_all += B.phi[f][sh];
        for (int sh = 0; sh < dim; ++sh) sum_dim += B.phi[f][sh];
        if (dl) {
            double w[3] = {0, 0, 0};
            for (int d1 = 0; d1 < dim; ++d1)
                for (int sh = 0; sh < dim; ++sh) w[d1] += uloc[sh * nc + d1] * B.phi[f][sh];
            const double wn = dot(dim, w, nh);
            for (int d1 = 0; d1 < dim; ++d1) w[d1] -= wn * nh[d1];
            for (int d1 = 0; d1 < dim; ++d1) dl[row * nc + d1] += -1.0 * w[d1] * k;
            const double wl = sqrt(dot(dim, w, w));
            if (wl != 0.0)
                for (int d1 = 0; d1 < dim; ++d1) dl[row * nc + d1] += -1.0 * (w[d1] / wl) * lim;
        }
        if (Jl) {
            double P[3][3];
            for (int d1 = 0; d1 < dim; ++d1)
                for (int d2 = 0; d2 < dim; ++d2) P[d1][d2] = (d1 == d2 ? sum_all : 0.0) - sum_dim * nh[d1] * nh[d2];
            double s[3] = {0, 0, 0};
            for (int d1 = 0; d1 < dim; ++d1) s[d1] = sum_all;
            const double sl = sqrt(dot(dim, s, s));
            for (int sh = 0; sh < nsh; ++sh) {
                for (int d1 = 0; d1 < dim; ++d1)
                    for (int d2 = 0; d2 < dim; ++d2) Jl[(row * nc + d1) * nl + sh * nc + d2] += -1.0 * P[d1][d2] * k;
                if (sl != 0.0)
                    for (int d1 = 0; d1 < dim; ++d1) Jl[(row * nc + d1) * nl + sh * nc + d1] += -1.0 * (s[d1] / sl) * lim;
            }
        }
    }
    return 0;
}

int or_sliding(const or_params* p, int et, const double* xyz, const int32_t* conn, const int32_t* belem,
               const int32_t* bside, int64_t nb, const double* u, double s_a, double k, double lim,
               const int64_t* rowptr, const int32_t* colidx, double* J, double* d)
{
    return bnd_scatter(2, p, et, xyz, conn, belem, bside, nb, u, s_a, k, lim, rowptr, colidx, J, d);
}

static int bnd_elem(int kind, const or_params* p, int et, const double* co, int side, const double* ul, double k,
                    double lim, double* Jl, double* dl)
{
    if (kind == 0) return or_outflow_elem(p, et, co, side, ul, Jl, dl);
    if (kind == 1) return or_symmetry_elem(p, et, co, side, ul, Jl, dl);
    return or_sliding_elem(p, et, co, side, ul, k, lim, Jl, dl);
}

/* ------------------------------------------------------------------ */
/* Data exports (restated): navier_stokes_fv1.cpp:894-1491.            */
/* what: 0 velocity (ex_nodal_velocity :898-990), 1 velocity gradient  */
/* (ex_velocity_grad :997-1098), 2 stabilised ip velocity              */
/* (ex_velocity_ip :1105-1278), 3 pressure (ex_nodal_pressure          */
/* :1285-1392), 4 pressure gradient (ex_pressure_grad :1399-1491).     */
/* at: 0 the element's SCVF ips, 1 its SCV ips (the corners), 2 the    */
/* nip local points lips[nip][dim] ("general case").                   */
/* out: [nip][ncomp], ncomp = dim, dim*dim (row d1, column d2), dim, 1,*/
/* dim.                                                                */
/* ------------------------------------------------------------------ */
int or_export_ncomp(int what, int dim)
{
    switch (what) {
    case 0: case 2: case 4: return dim;
    case 1: return dim * dim;
    case 3: return 1;
    default: return -1;
    }
}

/* shapes, reference and global gradients at local point x */
static int point_grads(int et, const double (*co)[3], const double* x, double* phi, double (*gl)[3], double (*g)[3])
{
    const int dim = REF[et].dim, nsh = REF[et].nsh;
    shapes(et, x, phi, gl);
    double J[3][3] = {{0}}, Ji[3][3] = {{0}};
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j)
            for (int k = 0; k < nsh; ++k) J[i][j] += co[k][i] * gl[k][j];
    double det;
    if (dim == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        if (det == 0.0) { set_err("degenerate element"); return -2; }
        Ji[0][0] = J[1][1] / det;  Ji[0][1] = -J[0][1] / det;
        Ji[1][0] = -J[1][0] / det; Ji[1][1] = J[0][0] / det;
    } else {
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (det == 0.0) { set_err("degenerate element"); return -2; }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                /* cofactor of J[j][i] / det */
                const int a0 = (j + 1) % 3, a1 = (j + 2) % 3, b0 = (i + 1) % 3, b1 = (i + 2) % 3;
                Ji[i][j] = (J[a0][b0] * J[a1][b1] - J[a0][b1] * J[a1][b0]) / det;
            }
    }
    for (int k = 0; k < nsh; ++k)
        for (int i = 0; i < dim; ++i) {
            double s = 0;
            for (int j = 0; j < dim; ++j) s += Ji[j][i] * gl[k][j];
            g[k][i] = s;
        }
    return 0;
}

int or_export_elem(const or_params* p, int et, const double* corners, const double* uloc, const double* uold,
                   double dt, int what, int at, const double* lips, int nip, double* out)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1, P = dim;
    const int ncomp = or_export_ncomp(what, dim);
    if (ncomp < 0 || at < 0 || at > 2) { set_err("unknown export"); return -1; }
    or_geo G;
    int rc = or_geometry(et, corners, &G);
    if (rc) return rc;
    if (at == 0) {
        if (what == 2) {
            /* the stabilisation's update only (ex_velocity_ip :1146-1189) */
            double stdvel[12][3];
            stabdata S;
            updata st;
            for (int ip = 0; ip < G.nip; ++ip)
                for (int d = 0; d < dim; ++d) {
                    stdvel[ip][d] = 0.0;
                    for (int sh = 0; sh < nsh; ++sh) stdvel[ip][d] += UL(d, sh) * G.phi[ip][sh];
                }
            if (p->stab_type < 0) { set_err("Stabilization has not been set."); return -3; }
            rc = stab_prologue(p, &G, uloc, uold, dt, (const double (*)[3])stdvel, &st, &S);
            if (rc) return rc;
            for (int ip = 0; ip < G.nip; ++ip)
                for (int d = 0; d < dim; ++d) out[ip * ncomp + d] = S.vel[ip][d];
            return 0;
        }
        for (int ip = 0; ip < G.nip; ++ip) {
            double* o = out + ip * ncomp;
            for (int c = 0; c < ncomp; ++c) o[c] = 0.0;
            for (int sh = 0; sh < nsh; ++sh) {
                if (what == 0)
                    for (int d = 0; d < dim; ++d) o[d] += UL(d, sh) * G.phi[ip][sh];
                else if (what == 1)
                    for (int d1 = 0; d1 < dim; ++d1)
                        for (int d2 = 0; d2 < dim; ++d2) o[d1 * dim + d2] += UL(d1, sh) * G.g[ip][sh][d2];
                else if (what == 3)
                    o[0] += UL(P, sh) * G.phi[ip][sh];
                else
                    for (int d = 0; d < dim; ++d) o[d] += UL(P, sh) * G.g[ip][sh][d];
            }
        }
        return 0;
    }
    const int npt = at == 1 ? nsh : nip;
    for (int q = 0; q < npt; ++q) {
        double x[3] = {0, 0, 0};
        for (int i = 0; i < dim; ++i) x[i] = at == 1 ? REF[et].lc[q][i] : lips[q * dim + i];
        double phi[8], gl[8][3], g[8][3];
        rc = point_grads(et, (const double (*)[3])G.co, x, phi, gl, g);
        if (rc) return rc;
        double* o = out + q * ncomp;
        for (int c = 0; c < ncomp; ++c) o[c] = 0.0;
        if (what == 3 && at == 1) { o[0] = UL(P, q); continue; }  /* SCV branch :1342-1364 */
        for (int sh = 0; sh < nsh; ++sh) {
            if (what == 0 || what == 2)
                for (int d = 0; d < dim; ++d) o[d] += UL(d, sh) * phi[sh];
            else if (what == 1)  /* reference-element gradient, as :1072-1080 */
                for (int d1 = 0; d1 < dim; ++d1)
                    for (int d2 = 0; d2 < dim; ++d2) o[d1 * dim + d2] += UL(d1, sh) * gl[sh][d2];
            else if (what == 3)
                o[0] += UL(P, sh) * phi[sh];
            else
                for (int d = 0; d < dim; ++d) o[d] += UL(P, sh) * g[sh][d];
        }
    }
    return 0;
}

int or_export(const or_params* p, int et, const double* xyz, const int32_t* conn, int64_t ne, const double* u,
              const double* uold, double dt, int what, int at, const double* lips, int nip, double* out)
{
    if (et < 0 || et > 4) { set_err("unknown element type"); return -1; }
    const int dim = REF[et].dim, nsh = REF[et].nsh, nc = dim + 1;
    const int npt = at == 0 ? REF[et].nip : (at == 1 ? nsh : nip);
    const int ncomp = or_export_ncomp(what, dim);
    if (ncomp < 0) { set_err("unknown export"); return -1; }
    double co[24], ul[32], uo[32];
    for (int64_t e = 0; e < ne; ++e) {
        const int32_t* c = conn + e * nsh;
        for (int k = 0; k < nsh; ++k) {
            for (int i = 0; i < dim; ++i) co[k * dim + i] = xyz[(int64_t)c[k] * dim + i];
            for (int a = 0; a < nc; ++a) {
                ul[k * nc + a] = u[(int64_t)c[k] * nc + a];
                if (uold) uo[k * nc + a] = uold[(int64_t)c[k] * nc + a];
            }
        }
        /* this element's rows of the integration-point imports */
        or_params q = *p;
        const int64_t sip = (int64_t)REF[et].nip;
        if (q.nu_ip) q.nu_ip += e * sip;
        if (q.rho_ip) q.rho_ip += e * sip;
        if (q.rho_scv) q.rho_scv += e * nsh;
        if (q.f_scv) q.f_scv += e * nsh * dim;
        if (q.f_scvf) q.f_scvf += e * sip * dim;
        int rc = or_export_elem(&q, et, co, ul, uold ? uo : NULL, dt, what, at, lips, nip, out + e * npt * ncomp);
        if (rc) return rc;
    }
    return 0;
}
