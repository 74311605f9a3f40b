// nsfv1_api.cpp -- C++ host layer and C ABI of the MI355X FV1 Navier-Stokes
// assembly.  The handle is the host-side counterpart of the reference's
// NavierStokesFV1<TDomain> object (incompressible/fv1/navier_stokes_fv1.h:
// 145-629): it carries the same configuration state (viscosity, density,
// source, exact-Jacobian factor, Stokes/Laplace/Peclet flags, the
// stabilisation object and the upwind objects with their sharing rules) and
// replaces ugcore's element loop by one device launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nsfv1.h"
#include "nsfv1_march.hpp"
#include "nsfv1_internal.hpp"
#include "nsfv1_solver.hpp"
#include "nsfv1_mlp.hpp"
#include "nsfv1_boundary.hpp"
#include "nsfv1_export.hpp"
#include "nsfv1_topo.hpp"

using namespace nsfv1;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

std::string trim_lower(const char* s)
{
    std::string n = s ? s : "";
    size_t a = n.find_first_not_of(" \t\n\r"), b = n.find_last_not_of(" \t\n\r");
    n = (a == std::string::npos) ? std::string() : n.substr(a, b - a + 1);
    std::transform(n.begin(), n.end(), n.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return n;
}

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t bytes)
    {
        release();
        n = bytes;
        return hipMalloc(&p, bytes ? bytes : 16);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Plan {
    bool valid = false;
    int nwg = 0;
    int block = 256;
    size_t lds = 0;
    DevBuf wg_node, wg_rounds, perm;
    // the same plan packed for one dependent load per lane and per
    // workgroup (split row kernel): lane_a[lane] = {element, node,
    // corner | slot << 8 | row length << 16, LDS row base}, lane_pos[lane] =
    // column blocks of the corners; wg_a[2w] = {v0, v1, i0, i1},
    // wg_a[2w+1] = {jbase low, jbase high, jlen, rounds}
    DevBuf lane_a, lane_pos, wg_a;
    bool rowlane = false;  // built for the row-lane split gather
};

}  // namespace

// one coarse level of the geometric multigrid preconditioner
struct MGLevel {
    nsfv1_handle* h = nullptr;      // coarse discretisation (owned by the caller)
    int64_t nv_fine = 0;             // nodes of the next finer level
    DevBuf P_rp, P_ci, P_w;          // prolongation: rows = finer nodes, cols = this level's nodes
    DevBuf R_rp, R_ci, R_w;          // restriction = transpose of P
    DevBuf inj;                      // this level's node -> finer node (state injection)
    DevBuf J, Binv, u, uold, b, x, r, t;  // t: second iterate of the smoothing sweeps
    int jprec = 0;
};

// boundary discs (no-normal-stress outflow, symmetry, sliding wall) of one
// element group: per boundary node, its boundary faces as (incidence, side)
// items
struct BndList {
    DevBuf ptr, node, inc, side;
    int64_t nbn = 0;
    std::vector<int32_t> src;  // per item: its face in the input list (pair b, side corner j) as 4 b + j
    DevBuf nu;                 // per item: the viscosity at the face's ip (function-valued viscosity), or empty
};

// A further element type of a mixed mesh (the first type lives in the
// handle itself): its elements, their node incidences with column slots in
// the mesh's common pattern, per-element geometry, launch plans and
// boundary lists.
struct ElemGroup {
    int et = -1;
    int64_t ne = 0, ninc = 0;
    DevBuf conn, node_inc, inc_elem, inc_node, inc_info, inc_pos, l2inv, vol, affine, upw_dec;
    std::vector<int64_t> h_node_inc;
    std::vector<int32_t> h_inc_info, h_inc_elem;
    Plan plan32, plan64;
    BndList bnd[BND_KINDS];
};

// enum values of the configuration state
enum { UPW_UNSET = -1 };
enum { STAB_UNSET = -1 };

struct nsfv1_handle {
    int dim = 2;
    int device = 0;
    std::vector<std::string> functions, subsets;
    // NavierStokesBase / IncompressibleNavierStokesBase state
    double nu = 0.0;
    bool nu_set = false;         // a viscosity (global or per instance) is available
    bool nu_global_set = false;  // set_kinematic_viscosity was called
    double rho = 1.0;  // navier_stokes_fv1.cpp:82
    bool has_source = false;
    double f[3] = {0, 0, 0};
    double exact = 0.0;  // navier_stokes_base.cpp:57
    bool stokes = false, laplace = false, peclet = false;  // incompressible_navier_stokes_base.cpp:57-59
    double grad_div = 0.0;
    // stabilisation / upwind objects (types + sharing)
    int stab_type = STAB_UNSET;
    int stab_upwind = UPW_UNSET;
    int difflen = DL_RAW;
    int conv_upwind = UPW_UNSET;
    bool pac = false;
    // mesh
    int et = -1;
    int64_t nv = 0, ne = 0, ndof = 0, nnz = 0, ninc = 0;
    DevBuf adj;  // node-block neighbour lists: adj[rowptr[v nc] / nc^2 + s] = s-th neighbour of v
    DevBuf xyz, conn, rowptr, colidx, errflag, l2inv, vol, affine, node_inc, inc_elem, inc_node, inc_info, inc_pos;
    std::vector<int32_t> h_inc_info, h_inc_elem;
    // element types after the first of a mixed mesh (empty: one type)
    std::vector<std::unique_ptr<ElemGroup>> groups;
    bool mixed() const { return !groups.empty(); }
    // boundary discs of the handle's own element group
    BndList bnd[BND_KINDS];
    std::vector<int32_t> bnd_elem_in[BND_KINDS], bnd_side_in[BND_KINDS];  // the input (element, side) lists
    double slide_k = 0.0, slide_lim = 0.0;
    std::vector<int64_t> h_rowptr, h_node_inc;
    int max_threads = 256;
    size_t lds_cap = 64 * 1024;
    Plan plan32, plan64;
    // Newton-loop linear algebra
    DevBuf cpl_recs;  // coupled path: per-SCVF row blocks
    DevBuf dir_rows, gm_V, gm_Binv, gm_part, gm_small, gm_w;
    int64_t ndir = 0;
    int gm_restart = 0;
    // geometric multigrid preconditioner
    std::vector<std::unique_ptr<MGLevel>> mg;
    int mg_pre = 2, mg_post = 2;
    double mg_omega = 0.7;
    bool mg_ready = false, use_mg = false;
    int mg_jprec = 0;
    int mg_ncoarse = 0;
    DevBuf mg_r0, mg_t0, mg_dense, mg_gj;  // mg_gj: Gauss-Jordan tableaux of the coarsest level
    bool geo_valid = false;
    int geo_difflen = -1;
    // split assembly path (per-SCVF records)
    DevBuf recs;
    // 0: split where supported, 1: fused kernel only, 2: split with row lanes,
    // 3: plane march on structured grids of parallelepipeds, otherwise 0
    int path = 3;
    // structured grid of parallelepipeds (plane-marching assembly path)
    bool march_ok = false;
    int march_n[3] = {0, 0, 0};
    int march_zc = 0;  // node planes per work unit (0: chosen from the grid)
    int last_path = -1;  // kernel family of the last element assembly (nsfv1_assembly_path_used)
    // batched instances: per-element viscosity
    DevBuf nu_elem;
    int64_t ninst = 0;
    // recorded skewed-upwind decisions
    DevBuf upw_dec;
    int upw_mode = 0;
    bool upw_recorded = false;
    // capturable Newton step: its own Krylov workspace (V: (m+1) ndof, w: 2
    // ndof, dev: Gram-Schmidt projections | ||r||^2 | beta | y | Hessenberg)
    // and the HIP graph of the step
    DevBuf nt_V, nt_w, nt_Binv, nt_part, nt_dev;
    DevBuf pred_part;  // learned predictor: partial sums of the per-step scale
    // imports at the integration points (function-valued viscosity and
    // source, evaluated by the caller): [ne][nip], [ne][nsh][dim], [ne][nip][dim]
    DevBuf nu_ip, f_scv, f_scvf, rho_ip, rho_scv;
    bool ip_imports() const { return nu_ip.p != nullptr || f_scv.p != nullptr || rho_ip.p != nullptr; }
    int nt_m = 0;
    hipGraph_t nt_graph = nullptr;
    hipGraphExec_t nt_exec = nullptr;
    ~nsfv1_handle()
    {
        if (nt_exec) (void)hipGraphExecDestroy(nt_exec);
        if (nt_graph) (void)hipGraphDestroy(nt_graph);
    }
};

// ---- imports at the integration points (helpers of the entry points below) ----
// corners and SCVFs (= edges) per element type: triangle, quadrilateral,
// tetrahedron, hexahedron, prism
static int nsh_count(int et) { static const int n[5] = {3, 4, 4, 8, 6}; return n[et]; }
static int nip_of(int et) { static const int n[5] = {3, 4, 6, 12, 9}; return n[et]; }

// ---- mesh set-up (helpers of nsfv1_set_mesh / nsfv1_set_mesh_mixed) ----
struct GroupIn {
    int et;
    const int32_t* conn;
    int64_t ne;
};

// node -> (element, corner) incidences of one element group, sorted by node
// then element: cnt = node offsets, ie/in = element and node, ii = corner |
// slot << 8 (slot = rank among the node's incidences of the group)
struct IncList {
    std::vector<int64_t> cnt;
    std::vector<int32_t> ie, in, ii;
};

static int build_incidences(const GroupIn& g, int64_t nv, int max_threads, IncList& L)
{
    const int nsh = nsh_count(g.et);
    L.cnt.assign((size_t)nv + 1, 0);
    for (int64_t i = 0; i < g.ne * nsh; ++i) L.cnt[g.conn[i] + 1]++;
    for (int64_t v = 0; v < nv; ++v) L.cnt[v + 1] += L.cnt[v];
    const int64_t ninc = L.cnt[nv];
    if (ninc >= (int64_t(1) << 31)) return fail(NSFV1_ERR_ARGUMENT, "too many incidences");
    L.ie.resize(ninc);
    L.in.resize(ninc);
    L.ii.resize(ninc);
    std::vector<int64_t> fill(L.cnt.begin(), L.cnt.end() - 1);
    for (int64_t e = 0; e < g.ne; ++e)
        for (int k = 0; k < nsh; ++k) {
            const int32_t v = g.conn[e * nsh + k];
            const int64_t q = fill[v]++;
            L.ie[q] = (int32_t)e;
            L.in[q] = v;
            L.ii[q] = k | (int32_t)((q - L.cnt[v]) << 8);
        }
    int maxinc = 0;
    for (int64_t v = 0; v < nv; ++v) maxinc = std::max<int>(maxinc, (int)(L.cnt[v + 1] - L.cnt[v]));
    if (maxinc > 255 || maxinc > max_threads)
        return fail(NSFV1_ERR_ARGUMENT, "a node has more incident elements than one workgroup can hold");
    return NSFV1_OK;
}

// the incidences a launch plan is built from (the handle's own element
// group or a further group of a mixed mesh)
struct IncView {
    const std::vector<int64_t>& node_inc;
    const std::vector<int32_t>& info;
    const std::vector<int32_t>& elem;
    int64_t ninc;
    const DevBuf& pos;
};

static void invalidate_plans(nsfv1_handle* h)
{
    h->plan32.valid = h->plan64.valid = false;
    for (auto& g : h->groups) g->plan32.valid = g->plan64.valid = false;
}

// integration points over all element groups (kind 0: SCVF ips, 1: SCV
// ips = corners); imports and positions run through the groups in order
static int64_t group_points(const nsfv1_handle* h, int kind)
{
    int64_t n = h->ne * (kind == 0 ? nip_of(h->et) : nsh_count(h->et));
    for (auto& g : h->groups) n += g->ne * (kind == 0 ? nip_of(g->et) : nsh_count(g->et));
    return n;
}

static int upload(DevBuf& b, const double* x, int64_t n)
{
    if (b.alloc(sizeof(double) * (size_t)n) != hipSuccess ||
        hipMemcpy(b.p, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "import upload");
    return NSFV1_OK;
}

template <int ET>
static void ip_positions_et(const std::vector<double>& xyz, const std::vector<int32_t>& conn, int64_t ne, int kind,
                            double* out)
{
    using T = Topo<ET>;
    constexpr int D = T::dim, N = T::nsh;
    for (int64_t e = 0; e < ne; ++e) {
        double co[N][3] = {}, c[3] = {0, 0, 0};
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < D; ++i) {
                co[k][i] = xyz[(size_t)conn[(size_t)e * N + k] * D + i];
                c[i] += co[k][i];
            }
        for (int i = 0; i < D; ++i) c[i] /= N;
        if (kind == 1) {
            for (int k = 0; k < N; ++k)
                for (int i = 0; i < D; ++i) out[((size_t)e * N + k) * D + i] = co[k][i];
            continue;
        }
        for (int ip = 0; ip < T::nip; ++ip) {
            double me[3];
            for (int i = 0; i < D; ++i) me[i] = 0.5 * (co[T::edge[ip][0]][i] + co[T::edge[ip][1]][i]);
            double* o = out + ((size_t)e * T::nip + ip) * D;
            if constexpr (D == 2) {
                for (int i = 0; i < 2; ++i) o[i] = 0.5 * (me[i] + c[i]);
            } else {
                double fc[2][3] = {};
                for (int q = 0; q < 2; ++q) {
                    const int s = edge_face<ET>(ip, q);
                    for (int j = 0; j < T::side_n[s]; ++j)
                        for (int i = 0; i < 3; ++i) fc[q][i] += co[T::side[s][j]][i];
                    for (int i = 0; i < 3; ++i) fc[q][i] /= T::side_n[s];
                }
                for (int i = 0; i < 3; ++i) o[i] = 0.25 * (me[i] + fc[0][i] + c[i] + fc[1][i]);
            }
        }
    }
}

extern "C" {

const char* nsfv1_last_error(void) { return g_last_error.c_str(); }

int nsfv1_create(int dim, const char* functions, const char* subsets, nsfv1_handle** out)
{
    if (!out) return fail(NSFV1_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(NSFV1_ERR_ARGUMENT, "NavierStokesFV1: dim must be 2 or 3");
    std::vector<std::string> fct;
    {
        std::string s = functions ? functions : "";
        size_t p = 0;
        while (p <= s.size()) {
            size_t q = s.find(',', p);
            if (q == std::string::npos) q = s.size();
            std::string t = s.substr(p, q - p);
            t.erase(0, t.find_first_not_of(" \t"));
            t.erase(t.find_last_not_of(" \t") + 1);
            if (!t.empty()) fct.push_back(t);
            p = q + 1;
        }
    }
    // navier_stokes_fv1.cpp:65-68
    if ((int)fct.size() != dim + 1)
        return fail(NSFV1_ERR_FUNCTION_COUNT, "Wrong number of functions: The ElemDisc 'NavierStokes' needs exactly " +
                                                  std::to_string(dim + 1) + " symbolic function.");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "no HIP device");
    auto h = std::make_unique<nsfv1_handle>();
    h->dim = dim;
    h->device = dev;
    h->functions = fct;
    h->subsets.push_back(subsets ? subsets : "");
    if (h->errflag.alloc(sizeof(int) * KERR_WORDS) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "hipMalloc");
    if (hipMemset(h->errflag.p, 0, sizeof(int) * KERR_WORDS) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "hipMemset");
    // device address of the no-allocation entry's scratch, once per device;
    // a failure here is reported by that entry, which retries the lookup
    (void)noalloc_scratch_setup(dev);
    *out = h.release();
    return NSFV1_OK;
}

int nsfv1_destroy(nsfv1_handle* h)
{
    delete h;
    return NSFV1_OK;
}

int nsfv1_prepare_setting(nsfv1_handle* h, const int* lfe_type, const int* order, int n, int non_regular)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (non_regular) return fail(NSFV1_ERR_HANGING_NODES, "NavierStokes: only regular grid implemented.");
    if (n != h->dim + 1) return fail(NSFV1_ERR_TRIAL_SPACE, "NavierStokes: Need exactly " + std::to_string(h->dim + 1) + " functions");
    for (int d = 0; d < n; ++d)
        if (!lfe_type || !order || lfe_type[d] != 0 || order[d] != 1)
            return fail(NSFV1_ERR_TRIAL_SPACE, "NavierStokes: 'fv1' expects Lagrange P1 trial space for velocity and pressure.");
    return NSFV1_OK;
}

int nsfv1_disc_type(nsfv1_handle* h, char* buf, size_t n)
{
    if (!h || !buf || n < 4) return fail(NSFV1_ERR_ARGUMENT, "bad argument");
    std::snprintf(buf, n, "fv1");
    return NSFV1_OK;
}

int nsfv1_set_kinematic_viscosity(nsfv1_handle* h, double nu)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->nu_ip.release();
    // the constant replaces a function's values at the boundary faces too
    for (int kind = 0; kind < BND_KINDS; ++kind) {
        h->bnd[kind].nu.release();
        for (auto& g : h->groups) g->bnd[kind].nu.release();
    }
    h->nu = nu;
    h->nu_set = true;
    h->nu_global_set = true;
    return NSFV1_OK;
}

int nsfv1_set_source(nsfv1_handle* h, const double* f, int n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->f_scv.release();
    h->f_scvf.release();
    if (!f || n <= 0) {
        h->has_source = false;
        return NSFV1_OK;
    }
    for (int i = 0; i < 3; ++i) h->f[i] = i < n && i < h->dim ? f[i] : 0.0;
    h->has_source = true;
    return NSFV1_OK;
}

int nsfv1_set_kinematic_viscosity_ip(nsfv1_handle* h, const double* nu, int64_t n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (!nu) {
        // the function's values at the boundary faces go with it
        h->nu_ip.release();
        for (int kind = 0; kind < BND_KINDS; ++kind) {
            h->bnd[kind].nu.release();
            for (auto& g : h->groups) g->bnd[kind].nu.release();
        }
        return NSFV1_OK;
    }
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (n != group_points(h, 0))
        return fail(NSFV1_ERR_ARGUMENT, "viscosity at the SCVF ips: expected ne * nip = " +
                                            std::to_string(group_points(h, 0)) + " values");
    int rc = upload(h->nu_ip, nu, n);
    if (rc) return rc;
    h->nu_set = true;
    h->nu_global_set = true;
    return NSFV1_OK;
}

int nsfv1_set_density_ip(nsfv1_handle* h, const double* rho_scvf, int64_t n_scvf, const double* rho_scv, int64_t n_scv)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (!rho_scvf && !rho_scv) {
        h->rho_ip.release();
        h->rho_scv.release();
        return NSFV1_OK;
    }
    if (!rho_scvf || !rho_scv)
        return fail(NSFV1_ERR_ARGUMENT, "the density import needs its values at both the SCVF and the SCV ips");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    const int64_t a = group_points(h, 0), b = group_points(h, 1);
    if (n_scvf != a || n_scv != b)
        return fail(NSFV1_ERR_ARGUMENT, "density at the ips: expected ne * nip = " + std::to_string(a) +
                                            " and ne * nsh = " + std::to_string(b) + " values");
    for (int64_t i = 0; i < a; ++i)
        if (!(rho_scvf[i] != 0.0)) return fail(NSFV1_ERR_ARGUMENT, "density at an SCVF ip is zero or NaN");
    int rc = upload(h->rho_ip, rho_scvf, n_scvf);
    if (!rc) rc = upload(h->rho_scv, rho_scv, n_scv);
    return rc;
}

int nsfv1_set_source_ip(nsfv1_handle* h, const double* f_scv, int64_t n_scv, const double* f_scvf, int64_t n_scvf)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (!f_scv && !f_scvf) {
        h->f_scv.release();
        h->f_scvf.release();
        h->has_source = false;
        return NSFV1_OK;
    }
    if (!f_scv || !f_scvf) return fail(NSFV1_ERR_ARGUMENT, "the source import needs its values at both the SCV and the SCVF ips");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    const int64_t a = group_points(h, 1) * h->dim, b = group_points(h, 0) * h->dim;
    if (n_scv != a || n_scvf != b)
        return fail(NSFV1_ERR_ARGUMENT, "source at the ips: expected ne * nsh * dim = " + std::to_string(a) +
                                            " and ne * nip * dim = " + std::to_string(b) + " values");
    int rc = upload(h->f_scv, f_scv, n_scv);
    if (!rc) rc = upload(h->f_scvf, f_scvf, n_scvf);
    if (rc) return rc;
    for (int i = 0; i < 3; ++i) h->f[i] = 0.0;
    h->has_source = true;
    return NSFV1_OK;
}

int nsfv1_ip_positions(nsfv1_handle* h, int kind, double* out, int64_t n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (kind != 0 && kind != 1) return fail(NSFV1_ERR_ARGUMENT, "kind: 0 = SCVF ips, 1 = SCV ips");
    const int64_t want = group_points(h, kind) * h->dim;
    if (!out || n != want) return fail(NSFV1_ERR_ARGUMENT, "expected " + std::to_string(want) + " values");
    std::vector<double> xyz((size_t)h->nv * h->dim);
    if (hipMemcpy(xyz.data(), h->xyz.p, sizeof(double) * xyz.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "mesh download");
    // group after group (the handle's own elements first)
    auto group = [&](int et, int64_t ne, const DevBuf& cb, double* o) -> bool {
        std::vector<int32_t> conn((size_t)ne * nsh_count(et));
        if (hipMemcpy(conn.data(), cb.p, sizeof(int32_t) * conn.size(), hipMemcpyDeviceToHost) != hipSuccess) return false;
        switch (et) {
        case 0: ip_positions_et<0>(xyz, conn, ne, kind, o); break;
        case 1: ip_positions_et<1>(xyz, conn, ne, kind, o); break;
        case 2: ip_positions_et<2>(xyz, conn, ne, kind, o); break;
        case 3: ip_positions_et<3>(xyz, conn, ne, kind, o); break;
        default: ip_positions_et<4>(xyz, conn, ne, kind, o); break;
        }
        return true;
    };
    if (!group(h->et, h->ne, h->conn, out)) return fail(NSFV1_ERR_DEVICE, "mesh download");
    int64_t off = h->ne * (kind == 0 ? nip_of(h->et) : nsh_count(h->et)) * h->dim;
    for (auto& g : h->groups) {
        if (!group(g->et, g->ne, g->conn, out + off)) return fail(NSFV1_ERR_DEVICE, "mesh download");
        off += g->ne * (kind == 0 ? nip_of(g->et) : nsh_count(g->et)) * h->dim;
    }
    return NSFV1_OK;
}

int nsfv1_set_exact_jacobian(nsfv1_handle* h, double factor)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->exact = factor;
    return NSFV1_OK;
}

int nsfv1_set_density(nsfv1_handle* h, double rho)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->rho_ip.release();  // the constant replaces a function's values
    h->rho_scv.release();
    h->rho = rho;
    return NSFV1_OK;
}

int nsfv1_set_stokes(nsfv1_handle* h, int on)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->stokes = on != 0;
    return NSFV1_OK;
}

int nsfv1_set_laplace(nsfv1_handle* h, int on)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->laplace = on != 0;
    return NSFV1_OK;
}

int nsfv1_set_peclet_blend(nsfv1_handle* h, int on)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->peclet = on != 0;
    return NSFV1_OK;
}

int nsfv1_set_grad_div(nsfv1_handle* h, double factor)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    h->grad_div = factor;  // used by the FVCR discretisation only
    return NSFV1_OK;
}

// upwind factory: upwind_interface.cpp:43-62
static int parse_upwind(const char* name, int* type)
{
    const std::string n = trim_lower(name);
    if (n == "no") { *type = UPW_NO; return NSFV1_OK; }
    if (n == "full") { *type = UPW_FULL; return NSFV1_OK; }
    if (n == "skewed") { *type = UPW_SKEWED; return NSFV1_OK; }
    if (n == "linearprofileskewed" || n == "lps") { *type = UPW_LPS; return NSFV1_OK; }
    if (n == "positive" || n == "pos") { *type = UPW_POSITIVE; return NSFV1_OK; }
    if (n == "regular" || n == "reg") { *type = UPW_REGULAR; return NSFV1_OK; }
    return fail(NSFV1_ERR_UNKNOWN_NAME, "NavierStokes: upwind type '" + std::string(name ? name : "") +
                                            "' not found. Options are: no, full, weighted, skewed, "
                                            "LinearProfileSkewed (lps), positive (pos), regular (reg)");
}

int nsfv1_set_upwind(nsfv1_handle* h, const char* name)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    int t;
    int rc = parse_upwind(name, &t);
    if (rc) return rc;
    // navier_stokes_fv1.h:213-215
    h->pac = false;
    h->conv_upwind = t;
    if (h->stab_type != STAB_UNSET && h->stab_upwind == UPW_UNSET) h->stab_upwind = t;
    return NSFV1_OK;
}

int nsfv1_set_stabilization(nsfv1_handle* h, const char* name, const char* diff_length)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    const std::string n = trim_lower(name);
    if (n != "fields" && n != "flow")
        return fail(NSFV1_ERR_UNKNOWN_NAME, "NavierStokes: stabilization type '" + std::string(name ? name : "") +
                                                "' not a valid name of a Schneider-Raw stabilization. Options are: fields, flow");
    int dl = DL_RAW;
    if (diff_length) {
        const std::string d = trim_lower(diff_length);
        if (d == "raw") dl = DL_RAW;
        else if (d == "fivepoint") dl = DL_FIVEPOINT;
        else if (d == "cor") dl = DL_COR;
        else return fail(NSFV1_ERR_UNKNOWN_NAME, "Diffusion Length calculation method not found. Use one of [Raw, Fivepoint, Cor].");
    }
    // navier_stokes_fv1.h:190-200: a new stabilisation object, handed the
    // convective upwind if one is set
    h->stab_type = n == "flow" ? STAB_FLOW : STAB_FIELDS;
    h->difflen = dl;
    h->stab_upwind = (h->conv_upwind != UPW_UNSET) ? h->conv_upwind : UPW_UNSET;
    return NSFV1_OK;
}

int nsfv1_set_stabilization_without(nsfv1_handle* h)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    // object form of set_stabilization: no upwind is handed over
    h->stab_type = STAB_WITHOUT;
    h->stab_upwind = UPW_UNSET;
    return NSFV1_OK;
}

int nsfv1_set_stabilization_upwind(nsfv1_handle* h, const char* name)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->stab_type == STAB_UNSET) return fail(NSFV1_ERR_NO_STABILIZATION, "Stabilization has not been set.");
    int t;
    int rc = parse_upwind(name, &t);
    if (rc) return rc;
    h->stab_upwind = t;
    return NSFV1_OK;
}

int nsfv1_set_pac_upwind(nsfv1_handle* h, int on)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (!on) return NSFV1_OK;
    // navier_stokes_fv1.h:217-226
    if (h->conv_upwind == UPW_UNSET) return fail(NSFV1_ERR_NO_UPWIND, "Upwind must be specified previously.\n");
    if (h->stab_type == STAB_UNSET) return fail(NSFV1_ERR_NO_STABILIZATION, "Stabilization must be specified previously.\n");
    h->stab_upwind = h->conv_upwind;
    h->pac = true;
    h->conv_upwind = UPW_UNSET;
    return NSFV1_OK;
}

static int nsh_of(int et) { return nsh_count(et); }
static int dim_of(int et) { return et <= ET_QUAD ? 2 : 3; }

// A hexahedral mesh that is a lexicographic nx x ny x nz grid (node (i,j,k)
// = i + (nx+1)(j + (ny+1)k), element (i,j,k) = i + nx(j + ny k) with the UG4
// corner order) of parallelepipeds (the geometry cache's test, tolerance
// 1e-12 of the element size), whose CSR rows are the grid's closed form.
// Fills n[3] and returns true when all of it holds.
static bool detect_structured_hex(const double* xyz, int64_t nv, const int32_t* conn, int64_t ne,
                                  const std::vector<int64_t>& rowptr, int* n)
{
    if (ne < 1 || conn[0] != 0) return false;
    const int64_t NX = conn[3], NXY = conn[4];
    if (NX < 2 || NXY < 2 * NX || NXY % NX) return false;
    const int64_t nx = NX - 1, ny = NXY / NX - 1;
    if (ne % (nx * ny)) return false;
    const int64_t nz = ne / (nx * ny);
    if (nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20) || nv != NXY * (nz + 1)) return false;
    const int64_t off[8] = {0, 1, NX + 1, NX, NXY, NXY + 1, NXY + NX + 1, NXY + NX};
    for (int64_t k = 0, e = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i, ++e) {
                const int64_t v0 = i + NX * j + NXY * k;
                const int32_t* c = conn + e * 8;
                for (int q = 0; q < 8; ++q)
                    if (c[q] != v0 + off[q]) return false;
                double scale = 0.0, dev = 0.0;
                const double* x[8];
                for (int q = 0; q < 8; ++q) x[q] = xyz + 3 * (v0 + off[q]);
                for (int q = 1; q < 8; ++q)
                    for (int a = 0; a < 3; ++a) scale = std::max(scale, std::fabs(x[q][a] - x[0][a]));
                for (int a = 0; a < 3; ++a) {
                    dev = std::max(dev, std::fabs(x[1][a] + x[3][a] - x[0][a] - x[2][a]));
                    dev = std::max(dev, std::fabs(x[1][a] + x[4][a] - x[0][a] - x[5][a]));
                    dev = std::max(dev, std::fabs(x[3][a] + x[4][a] - x[0][a] - x[7][a]));
                    dev = std::max(dev, std::fabs(x[1][a] + x[3][a] + x[4][a] - 2.0 * x[0][a] - x[6][a]));
                }
                if (!(dev <= 1e-12 * scale)) return false;
            }
    // rows: 4 per node, 4 ax ay az values each, nodes in order
    const int64_t SX = 3 * nx + 1, SY = 3 * ny + 1;
    for (int64_t k = 0, v = 0; k <= nz; ++k)
        for (int64_t j = 0; j <= ny; ++j)
            for (int64_t i = 0; i <= nx; ++i, ++v) {
                const int64_t ax = 1 + (i > 0) + (i < nx), ay = 1 + (j > 0) + (j < ny), az = 1 + (k > 0) + (k < nz);
                const int64_t sx = i > 0 ? 3 * i - 1 : 0, sy = j > 0 ? 3 * j - 1 : 0, sz = k > 0 ? 3 * k - 1 : 0;
                const int64_t rb = 16 * (SX * SY * sz + SX * sy * az + sx * ay * az);
                for (int r = 0; r <= 4; ++r)
                    if (rowptr[v * 4 + r] != rb + r * 4 * ax * ay * az) return false;
            }
    n[0] = (int)nx;
    n[1] = (int)ny;
    n[2] = (int)nz;
    return true;
}

// The mesh as groups of elements of one type each (a mixed mesh numbers its
// elements group after group): one common CSR pattern over all groups, the
// node incidences of each group with their column slots in it.
static int set_mesh_groups(nsfv1_handle* h, const double* xyz, int64_t nv, const std::vector<GroupIn>& G)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (G.empty()) return fail(NSFV1_ERR_ARGUMENT, "no element groups");
    for (const GroupIn& g : G) {
        if (g.et == NSFV1_PYRAMID)
            return fail(NSFV1_ERR_NOT_IMPLEMENTED, "pyramids are not provided: the apex corner has four edges, so its "
                                                   "FV1 sub-control volume is not the hexahedron the other elements use");
        if (g.et < 0 || g.et > 4) return fail(NSFV1_ERR_ARGUMENT, "unknown element type");
        if (dim_of(g.et) != h->dim) return fail(NSFV1_ERR_ARGUMENT, "element type does not match the domain dimension");
        if (g.ne <= 0 || !g.conn) return fail(NSFV1_ERR_ARGUMENT, "empty mesh");
    }
    if (nv <= 0 || !xyz) return fail(NSFV1_ERR_ARGUMENT, "empty mesh");
    if (nv >= (int64_t(1) << 31) / (h->dim + 1)) return fail(NSFV1_ERR_ARGUMENT, "too many nodes for int32 column indices");
    const int nc = h->dim + 1;
    const size_t ng = G.size();
    for (const GroupIn& g : G)
        for (int64_t i = 0; i < g.ne * nsh_of(g.et); ++i)
            if (g.conn[i] < 0 || g.conn[i] >= nv) return fail(NSFV1_ERR_ARGUMENT, "connectivity index out of range");

    std::vector<IncList> inc(ng);
    for (size_t g = 0; g < ng; ++g) {
        int rc = build_incidences(G[g], nv, h->max_threads, inc[g]);
        if (rc) return rc;
    }
    // the neighbours of node v: every corner of its incident elements (of
    // all groups), sorted, without repeats; an isolated node keeps itself
    auto neighbours = [&](int64_t v, std::vector<int32_t>& nb) {
        nb.clear();
        for (size_t g = 0; g < ng; ++g) {
            const int nsh = nsh_of(G[g].et);
            for (int64_t q = inc[g].cnt[v]; q < inc[g].cnt[v + 1]; ++q)
                for (int k = 0; k < nsh; ++k) nb.push_back(G[g].conn[(int64_t)inc[g].ie[q] * nsh + k]);
        }
        if (nb.empty()) nb.push_back((int32_t)v);
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    };

    // CSR pattern: node adjacency closure (x) (dim+1)^2, sorted columns
    std::vector<int64_t> rowptr(nv * nc + 1);
    {
        std::vector<int32_t> nb;
        nb.reserve(256);
        rowptr[0] = 0;
        for (int64_t v = 0; v < nv; ++v) {
            neighbours(v, nb);
            if (nb.size() > 255) return fail(NSFV1_ERR_ARGUMENT, "a node has more than 255 neighbours");
            for (int a = 0; a < nc; ++a) rowptr[v * nc + a + 1] = rowptr[v * nc + a] + (int64_t)nb.size() * nc;
        }
    }
    const int64_t nnz = rowptr[nv * nc];
    int mn[3] = {0, 0, 0};
    const bool structured = ng == 1 && G[0].et == ET_HEX && detect_structured_hex(xyz, nv, G[0].conn, G[0].ne, rowptr, mn);
    std::vector<int32_t> colidx, adjv;
    std::vector<std::vector<uint64_t>> ipos(ng);
    try {
        colidx.resize(nnz);
        adjv.resize(nnz / (nc * nc));
        for (size_t g = 0; g < ng; ++g) ipos[g].resize(inc[g].ie.size());
    } catch (...) {
        return fail(NSFV1_ERR_ARGUMENT, "pattern too large for host memory");
    }
    {
        std::vector<int32_t> nb;
        nb.reserve(256);
        for (int64_t v = 0; v < nv; ++v) {
            neighbours(v, nb);
            // column slots of each incident element's corners in v's row
            for (size_t g = 0; g < ng; ++g) {
                const int nsh = nsh_of(G[g].et);
                for (int64_t q = inc[g].cnt[v]; q < inc[g].cnt[v + 1]; ++q) {
                    uint64_t pk = 0;
                    for (int k = 0; k < nsh; ++k) {
                        const int32_t w = G[g].conn[(int64_t)inc[g].ie[q] * nsh + k];
                        const uint64_t slot = (uint64_t)(std::lower_bound(nb.begin(), nb.end(), w) - nb.begin());
                        pk |= slot << (8 * k);
                    }
                    ipos[g][q] = pk;
                }
            }
            for (size_t t = 0; t < nb.size(); ++t) adjv[rowptr[v * nc] / (nc * nc) + t] = nb[t];
            for (int a = 0; a < nc; ++a) {
                int32_t* out = colidx.data() + rowptr[v * nc + a];
                for (size_t t = 0; t < nb.size(); ++t)
                    for (int b = 0; b < nc; ++b) out[t * nc + b] = nb[t] * nc + b;
            }
        }
    }

    auto up = [](DevBuf& b, const void* src, size_t bytes) -> bool {
        if (b.alloc(bytes) != hipSuccess) return false;
        return hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    // the device arrays of group g: its connectivity, incidences and
    // per-element geometry buffers
    auto up_group = [&](size_t g, DevBuf& conn, DevBuf& node_inc, DevBuf& inc_elem, DevBuf& inc_node, DevBuf& inc_info,
                        DevBuf& inc_pos, DevBuf& l2inv, DevBuf& vol, DevBuf& affine) -> bool {
        const GroupIn& gi = G[g];
        const IncList& L = inc[g];
        const size_t ninc = L.ie.size();
        const std::vector<int32_t> ni32(L.cnt.begin(), L.cnt.end());
        return up(conn, gi.conn, sizeof(int32_t) * gi.ne * nsh_of(gi.et)) &&
               up(inc_elem, L.ie.data(), sizeof(int32_t) * std::max<size_t>(ninc, 1)) &&
               up(inc_node, L.in.data(), sizeof(int32_t) * std::max<size_t>(ninc, 1)) &&
               up(inc_info, L.ii.data(), sizeof(int32_t) * std::max<size_t>(ninc, 1)) &&
               up(inc_pos, ipos[g].data(), sizeof(uint64_t) * std::max<size_t>(ninc, 1)) &&
               up(node_inc, ni32.data(), sizeof(int32_t) * ni32.size()) &&
               l2inv.alloc(sizeof(double) * gi.ne * nip_of(gi.et)) == hipSuccess &&
               vol.alloc(sizeof(double) * gi.ne * nsh_of(gi.et)) == hipSuccess && affine.alloc((size_t)gi.ne) == hipSuccess;
    };
    std::vector<std::unique_ptr<ElemGroup>> more;
    bool ok = up(h->xyz, xyz, sizeof(double) * nv * h->dim) &&
              up(h->rowptr, rowptr.data(), sizeof(int64_t) * rowptr.size()) &&
              up(h->colidx, colidx.data(), sizeof(int32_t) * nnz) &&
              up(h->adj, adjv.data(), sizeof(int32_t) * std::max<size_t>(adjv.size(), 1)) &&
              up_group(0, h->conn, h->node_inc, h->inc_elem, h->inc_node, h->inc_info, h->inc_pos, h->l2inv, h->vol,
                       h->affine);
    for (size_t g = 1; ok && g < ng; ++g) {
        auto eg = std::make_unique<ElemGroup>();
        ok = up_group(g, eg->conn, eg->node_inc, eg->inc_elem, eg->inc_node, eg->inc_info, eg->inc_pos, eg->l2inv,
                      eg->vol, eg->affine);
        eg->et = G[g].et;
        eg->ne = G[g].ne;
        eg->ninc = (int64_t)inc[g].ie.size();
        eg->h_node_inc.swap(inc[g].cnt);
        eg->h_inc_info.swap(inc[g].ii);
        eg->h_inc_elem.swap(inc[g].ie);
        more.push_back(std::move(eg));
    }
    if (!ok) return fail(NSFV1_ERR_DEVICE, std::string("mesh upload: ") + hipGetErrorString(hipGetLastError()));
    h->et = G[0].et;
    h->nv = nv;
    h->ne = G[0].ne;
    h->ndof = nv * nc;
    h->march_ok = structured;
    for (int a = 0; a < 3; ++a) h->march_n[a] = mn[a];
    h->nnz = nnz;
    h->ninc = (int64_t)inc[0].ie.size();
    h->h_rowptr.swap(rowptr);
    h->h_node_inc.swap(inc[0].cnt);
    h->h_inc_info.swap(inc[0].ii);
    h->h_inc_elem.swap(inc[0].ie);
    h->groups.swap(more);
    for (auto& b : h->bnd) b.nbn = 0;
    invalidate_plans(h);
    h->geo_valid = false;
    // solver state sized for the previous mesh
    h->gm_restart = 0;
    h->ndir = 0;
    h->mg.clear();
    h->mg_ready = false;
    h->upw_dec.release();
    h->upw_recorded = false;
    h->nu_elem.release();
    h->nu_ip.release();
    h->rho_ip.release();
    h->rho_scv.release();
    if (h->f_scv.p) {
        h->f_scv.release();
        h->f_scvf.release();
        h->has_source = false;
    }
    h->ninst = 0;
    h->recs.release();
    h->cpl_recs.release();
    return NSFV1_OK;
}

int nsfv1_set_mesh(nsfv1_handle* h, int et, const double* xyz, int64_t nv, const int32_t* conn, int64_t ne)
{
    return set_mesh_groups(h, xyz, nv, {GroupIn{et, conn, ne}});
}

int nsfv1_set_mesh_mixed(nsfv1_handle* h, const double* xyz, int64_t nv, int ngroups, const int* et,
                         const int32_t* const* conn, const int64_t* ne)
{
    if (ngroups <= 0 || ngroups > 8 || !et || !conn || !ne)
        return fail(NSFV1_ERR_ARGUMENT, "mixed mesh: 1..8 element groups with their types, connectivities and counts");
    std::vector<GroupIn> G;
    for (int g = 0; g < ngroups; ++g) G.push_back(GroupIn{et[g], conn[g], ne[g]});
    return set_mesh_groups(h, xyz, nv, G);
}

int nsfv1_num_dofs(nsfv1_handle* h, int64_t* ndof)
{
    if (!h || !ndof) return fail(NSFV1_ERR_ARGUMENT, "bad argument");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    *ndof = h->ndof;
    return NSFV1_OK;
}

int nsfv1_nnz(nsfv1_handle* h, int64_t* nnz)
{
    if (!h || !nnz) return fail(NSFV1_ERR_ARGUMENT, "bad argument");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    *nnz = h->nnz;
    return NSFV1_OK;
}

int nsfv1_get_pattern(nsfv1_handle* h, int64_t* rowptr, int32_t* colidx)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (rowptr) std::memcpy(rowptr, h->h_rowptr.data(), sizeof(int64_t) * (h->ndof + 1));
    if (colidx && hipMemcpy(colidx, h->colidx.p, sizeof(int32_t) * h->nnz, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "pattern copy");
    return NSFV1_OK;
}

int nsfv1_pattern_device(nsfv1_handle* h, const int64_t** rowptr, const int32_t** colidx)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (rowptr) *rowptr = h->rowptr.as<int64_t>();
    if (colidx) *colidx = h->colidx.as<int32_t>();
    return NSFV1_OK;
}

int nsfv1_set_launch(nsfv1_handle* h, int threads, int lds_bytes)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (threads) {
        if (threads < 64 || threads > 256 || threads % 64) return fail(NSFV1_ERR_ARGUMENT, "threads must be 64..256, a multiple of 64");
        h->max_threads = threads;
    }
    if (lds_bytes && h->path == 3) {
        // the plane march reads the budget in KiB as node planes per work unit
        if (lds_bytes < 2048 || lds_bytes > 65536) return fail(NSFV1_ERR_ARGUMENT, "plane march: 2..64 node planes per work unit");
        h->march_zc = lds_bytes / 1024;
    } else if (lds_bytes) {
        if (lds_bytes < 4096 || lds_bytes > 65536) return fail(NSFV1_ERR_ARGUMENT, "LDS budget must be 4..64 KiB");
        h->lds_cap = (size_t)lds_bytes;
    }
    invalidate_plans(h);
    return NSFV1_OK;
}

// Workgroup plan: consecutive nodes while their incidences fit the thread
// budget and their rows (plus defect) fit the LDS budget (<= 64 KiB).
static int make_plan(nsfv1_handle* h, const IncView& g, int jbytes, Plan& P)
{
    const int nc = h->dim + 1;
    std::vector<int32_t> wn, wr;
    wn.push_back(0);
    int64_t v = 0;
    size_t maxlds = 0;
    // the row-lane split gather puts dim+1 lanes on each incidence of a
    // 64-lane workgroup
    const bool rowlane = h->path == 2 && jbytes == 4 && h->max_threads == 64;
    const int64_t inc_budget = rowlane ? 64 / nc : h->max_threads;
    // element groups of a mixed mesh: a run of nodes without incidences of
    // the group is one empty range (its workgroup returns at once), so the
    // other ranges hold only the group's nodes
    const bool skip_empty = h->mixed();
    auto empty = [&](int64_t u) { return g.node_inc[u + 1] == g.node_inc[u]; };
    while (v < h->nv) {
        const int64_t v0 = v;
        int rounds = 0;
        if (skip_empty && empty(v)) {
            while (v < h->nv && empty(v)) ++v;
            wn.push_back((int32_t)v);
            wr.push_back(0);
            continue;
        }
        while (v < h->nv && !(skip_empty && empty(v))) {
            const int64_t incs = g.node_inc[v + 1] - g.node_inc[v0];
            const int64_t vals = h->h_rowptr[(v + 1) * nc] - h->h_rowptr[v0 * nc];
            const size_t lds = ((size_t)vals * jbytes + (size_t)(v + 1 - v0) * 16 + 15) / 16 * 16 + (size_t)(v + 1 - v0) * nc * 8;
            if (incs > inc_budget || lds > h->lds_cap) {
                if (v == v0) return fail(NSFV1_ERR_ARGUMENT, "one node row exceeds the workgroup budget");
                break;
            }
            rounds = std::max<int>(rounds, (int)(g.node_inc[v + 1] - g.node_inc[v]));
            maxlds = std::max(maxlds, lds);
            ++v;
        }
        wn.push_back((int32_t)v);
        wr.push_back(rounds);
    }
    P.nwg = (int)wr.size();
    P.lds = maxlds;
    P.block = h->max_threads;
    P.rowlane = rowlane;
    // lane order inside each workgroup: by incidence slot, then node, so a
    // wavefront takes part in as few accumulation rounds as possible
    std::vector<int32_t> perm((size_t)g.ninc);
    for (int w = 0; w < P.nwg; ++w) {
        const int64_t i0 = g.node_inc[wn[w]], i1 = g.node_inc[wn[w + 1]];
        std::vector<std::pair<int32_t, int32_t>> key;
        key.reserve((size_t)(i1 - i0));
        for (int64_t q = i0; q < i1; ++q) key.push_back({g.info[q] >> 8, (int32_t)q});
        std::stable_sort(key.begin(), key.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (int64_t t = 0; t < i1 - i0; ++t) perm[i0 + t] = key[t].second;
    }
    auto up = [](DevBuf& b, const std::vector<int32_t>& x) {
        return b.alloc(sizeof(int32_t) * std::max<size_t>(x.size(), 1)) == hipSuccess &&
               (x.empty() || hipMemcpy(b.p, x.data(), sizeof(int32_t) * x.size(), hipMemcpyHostToDevice) == hipSuccess);
    };
    if (!up(P.wg_node, wn) || !up(P.wg_rounds, wr) || !up(P.perm, perm)) return fail(NSFV1_ERR_DEVICE, "plan upload");
    {
        const int pade = nc == 4 ? 16 / jbytes : 0;  // the row kernels pad 3D node rows only
        std::vector<int32_t> la((size_t)g.ninc * 4), wa((size_t)P.nwg * 8);
        std::vector<uint64_t> lp((size_t)g.ninc);
        std::vector<uint64_t> hpos((size_t)g.ninc);
        if (g.ninc > 0 &&
            hipMemcpy(hpos.data(), g.pos.p, sizeof(uint64_t) * hpos.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(NSFV1_ERR_DEVICE, "plan download");
        for (int w = 0; w < P.nwg; ++w) {
            const int64_t v0 = wn[w], v1 = wn[w + 1];
            const int64_t i0 = g.node_inc[v0], i1 = g.node_inc[v1];
            const int64_t jbase = h->h_rowptr[v0 * nc];
            for (int64_t t = i0; t < i1; ++t) {
                const int32_t q = perm[t];
                int64_t v = v0;
                while (g.node_inc[v + 1] <= q) ++v;
                const int32_t info = g.info[q];
                const int64_t rlen = h->h_rowptr[v * nc + 1] - h->h_rowptr[v * nc];
                if (rlen >= (1 << 15)) return fail(NSFV1_ERR_ARGUMENT, "node row too long for the plan");
                la[t * 4 + 0] = g.elem[q];
                la[t * 4 + 1] = (int32_t)v;
                la[t * 4 + 2] = (info & 0xffff) | (int32_t)(rlen << 16);
                la[t * 4 + 3] = (int32_t)(h->h_rowptr[v * nc] - jbase + (v - v0) * pade);
                lp[t] = hpos[q];
            }
            const int64_t jlen = h->h_rowptr[v1 * nc] - jbase;
            wa[w * 8 + 0] = (int32_t)v0;
            wa[w * 8 + 1] = (int32_t)v1;
            wa[w * 8 + 2] = (int32_t)i0;
            wa[w * 8 + 3] = (int32_t)i1;
            wa[w * 8 + 4] = (int32_t)(uint32_t)(jbase & 0xffffffffu);
            wa[w * 8 + 5] = (int32_t)(jbase >> 32);
            wa[w * 8 + 6] = (int32_t)jlen;
            wa[w * 8 + 7] = wr[w];
        }
        auto upb = [](DevBuf& b, const void* x, size_t bytes) {
            return b.alloc(std::max<size_t>(bytes, 16)) == hipSuccess &&
                   (bytes == 0 || hipMemcpy(b.p, x, bytes, hipMemcpyHostToDevice) == hipSuccess);
        };
        if (!upb(P.lane_a, la.data(), sizeof(int32_t) * la.size()) ||
            !upb(P.lane_pos, lp.data(), sizeof(uint64_t) * lp.size()) ||
            !upb(P.wg_a, wa.data(), sizeof(int32_t) * wa.size()))
            return fail(NSFV1_ERR_DEVICE, "plan upload");
    }
    P.valid = true;
    return NSFV1_OK;
}

// configurations whose SCVFs are coupled within an element: the FLOW
// stabilisation (connected velocity components) and upwinds with ip shapes
static bool coupled_config(const nsfv1_handle* h)
{
    if (h->stab_type == STAB_FLOW) return true;
    if (h->stokes) return false;
    if (h->stab_type != STAB_WITHOUT && (h->stab_upwind == UPW_POSITIVE || h->stab_upwind == UPW_REGULAR)) return true;
    return !h->pac && (h->conv_upwind == UPW_POSITIVE || h->conv_upwind == UPW_REGULAR);
}

// per-element SCV volumes and diffusion-length factors, built on first use
// and again when the diffusion-length method changes
static int ensure_geometry(nsfv1_handle* h, void* stream)
{
    if (h->geo_valid && h->geo_difflen == h->difflen) return NSFV1_OK;
    hipError_t ge = launch_geometry_cache(h->et, h->xyz.as<double>(), h->conn.as<int32_t>(), h->ne, h->difflen,
                                          h->l2inv.as<double>(), h->vol.as<double>(), h->affine.as<uint8_t>(),
                                          h->errflag.as<int>(), (hipStream_t)stream);
    for (size_t g = 0; ge == hipSuccess && g < h->groups.size(); ++g) {
        ElemGroup& G = *h->groups[g];
        ge = launch_geometry_cache(G.et, h->xyz.as<double>(), G.conn.as<int32_t>(), G.ne, h->difflen, G.l2inv.as<double>(),
                                   G.vol.as<double>(), G.affine.as<uint8_t>(), h->errflag.as<int>(), (hipStream_t)stream);
    }
    if (ge != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("geometry cache: ") + hipGetErrorString(ge));
    h->geo_valid = true;
    h->geo_difflen = h->difflen;
    return NSFV1_OK;
}

// checks of prep_elem_loop (navier_stokes_fv1.cpp:136-182)
static int check_setup(nsfv1_handle* h)
{
    if (h->stab_type == STAB_UNSET) return fail(NSFV1_ERR_NO_STABILIZATION, "Stabilization has not been set.");
    if (!h->stokes) {
        if (!h->pac && h->conv_upwind == UPW_UNSET)
            return fail(NSFV1_ERR_NO_UPWIND, "Upwinding for convective Term in Momentum eq. not set.");
        if (h->stab_upwind == UPW_UNSET)
            return fail(NSFV1_ERR_NO_UPWIND, "The stabilization has no upwind object (INavierStokesFV1Stabilization::set_upwind).");
    }
    if (!h->nu_set)
        return fail(NSFV1_ERR_NO_VISCOSITY, "NavierStokes::prep_elem_loop: Kinematic Viscosity has not been set, but is required.");
    // upwind.cpp:806 (thrown when the upwind is computed)
    const bool reg = (!h->stokes && h->stab_type != STAB_WITHOUT && h->stab_upwind == UPW_REGULAR) ||
                     (!h->stokes && !h->pac && h->conv_upwind == UPW_REGULAR);
    if (reg && h->dim != 2) return fail(NSFV1_ERR_NOT_IMPLEMENTED, "RegularUpwind only implemented for 2d.");
    return NSFV1_OK;
}

// Body of nsfv1_assemble and nsfv1_assemble_noalloc.  With `scratch` set
// (the no-allocation entry) the record buffers of the split and coupled
// paths come from that fixed device array of NOALLOC_SCRATCH_BYTES, and
// every state the call would otherwise allocate (launch plan, upwind
// record buffer) must exist already.
static int assemble_impl(nsfv1_handle* h, const double* u, const double* u_old, double dt, double s_a, double s_m,
                         double s_f, int parts, void* jvals, int jprec, double* defect, void* stream, void* scratch)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (!u) return fail(NSFV1_ERR_ARGUMENT, "u is NULL");
    if (jprec != NSFV1_F32 && jprec != NSFV1_F64) return fail(NSFV1_ERR_ARGUMENT, "jprec must be 32 or 64");
    if (parts & ~7) return fail(NSFV1_ERR_ARGUMENT, "unknown part bits");
    if (u_old && !(dt > 0.0)) return fail(NSFV1_ERR_TIME_SERIES, "NavierStokes: time step must be positive for the time-dependent stabilization");
    if (parts & NSFV1_PART_A) {
        int rc = check_setup(h);
        if (rc) return rc;
        bool bnd = false;
        for (int kind = 0; kind < BND_KINDS; ++kind) {
            bnd |= h->bnd[kind].nbn > 0;
            for (auto& g : h->groups) bnd |= g->bnd[kind].nbn > 0;
        }
        if (bnd && h->nu_elem.p)
            return fail(NSFV1_ERR_NOT_IMPLEMENTED,
                        "boundary discs with per-instance viscosity (batched instances) are not provided");
        if (bnd && h->rho_ip.p)
            return fail(NSFV1_ERR_NOT_IMPLEMENTED,
                        "boundary discs with a function-valued density are not provided (their density is a constant)");
        // a function-valued viscosity reaches the outflow and symmetry discs
        // through its values at their boundary faces
        if (h->nu_ip.p)
            for (int kind = 0; kind < BND_KINDS; ++kind) {
                if (kind == BND_SLIDING) continue;  // the sliding wall has no viscous term
                bool miss = h->bnd[kind].nbn > 0 && !h->bnd[kind].nu.p;
                for (auto& g : h->groups) miss |= g->bnd[kind].nbn > 0 && !g->bnd[kind].nu.p;
                if (miss)
                    return fail(NSFV1_ERR_ARGUMENT, "a function-valued viscosity needs its values at the boundary faces "
                                                    "of the outflow and symmetry discs (nsfv1_set_boundary_viscosity)");
            }
    }
    Plan& P = jprec == NSFV1_F64 ? h->plan64 : h->plan32;
    if (!P.valid) {
        if (scratch)
            return fail(NSFV1_ERR_ARGUMENT, "no launch plan for this Jacobian precision yet: the no-allocation entry "
                                            "needs one nsfv1_assemble call with the same jprec first (it builds the plan)");
        int rc = make_plan(h, IncView{h->h_node_inc, h->h_inc_info, h->h_inc_elem, h->ninc, h->inc_pos},
                           jprec == NSFV1_F64 ? 8 : 4, P);
        if (rc) return rc;
    }
    for (auto& gp : h->groups) {
        ElemGroup& G = *gp;
        Plan& PG = jprec == NSFV1_F64 ? G.plan64 : G.plan32;
        if (PG.valid) continue;
        if (scratch)
            return fail(NSFV1_ERR_ARGUMENT, "no launch plan for this Jacobian precision yet: the no-allocation entry "
                                            "needs one nsfv1_assemble call with the same jprec first (it builds the plan)");
        int rc = make_plan(h, IncView{G.h_node_inc, G.h_inc_info, G.h_inc_elem, G.ninc, G.inc_pos},
                           jprec == NSFV1_F64 ? 8 : 4, PG);
        if (rc) return rc;
    }
    KParams k{};
    k.conv_upwind = h->conv_upwind;
    k.stab_upwind = h->stab_upwind;
    k.stab_type = h->stab_type;
    k.pac = h->pac;
    k.difflen = h->difflen;
    k.stokes = h->stokes;
    k.laplace = h->laplace;
    k.peclet = h->peclet;
    k.has_source = h->has_source;
    k.has_old = u_old != nullptr;
    k.parts = parts;
    k.exact = h->exact;
    k.nu = h->nu;
    k.rho = h->rho;
    k.dt = dt;
    k.inv_dt = dt > 0.0 ? 1.0 / dt : 0.0;
    k.inv_rho = 1.0 / h->rho;
    for (int i = 0; i < 3; ++i) k.f[i] = h->f[i];
    k.s_a = s_a;
    k.s_m = s_m;
    k.s_f = s_f;
    {
        int rc = ensure_geometry(h, stream);
        if (rc) return rc;
    }
    k.upw_mode = h->upw_mode;
    if (h->upw_mode == NSFV1_UPWIND_REPLAY && !h->upw_recorded)
        return fail(NSFV1_ERR_ARGUMENT, "no recorded upwind decisions to replay");
    if (h->upw_mode == NSFV1_UPWIND_RECORD && h->upw_dec.p == nullptr) {
        if (scratch)
            return fail(NSFV1_ERR_ARGUMENT, "no upwind record buffer yet: the no-allocation entry needs one "
                                            "nsfv1_assemble call in NSFV1_UPWIND_RECORD mode first (it allocates it)");
        const int nip = nip_of(h->et);
        if (h->upw_dec.alloc((size_t)h->ne * nip) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "upwind record buffer");
        for (auto& gp : h->groups)
            if (gp->upw_dec.alloc((size_t)gp->ne * nip_of(gp->et)) != hipSuccess)
                return fail(NSFV1_ERR_DEVICE, "upwind record buffer");
    }
    KMeshPtrs m{h->xyz.as<double>(), h->conn.as<int32_t>(), h->rowptr.as<int64_t>(), h->colidx.as<int32_t>(),
                h->node_inc.as<int32_t>(), h->inc_elem.as<int32_t>(), h->inc_node.as<int32_t>(),
                h->inc_info.as<int32_t>(), P.wg_node.as<int32_t>(), P.wg_rounds.as<int32_t>(), h->l2inv.as<double>(),
                h->vol.as<double>(), h->upw_dec.as<uint8_t>(), h->affine.as<uint8_t>(), h->inc_pos.as<uint64_t>(),
                P.perm.as<int32_t>(), h->nu_elem.as<double>(), h->nu_ip.as<double>(), h->f_scv.as<double>(),
                h->f_scvf.as<double>(), h->rho_ip.as<double>(), h->rho_scv.as<double>(), P.block, P.lane_a.as<int4>(),
                P.lane_pos.as<uint64_t>(), P.wg_a.as<int4>()};
    hipError_t e;
    int mfast = -1;
    if (h->path == 3 && h->march_ok && jprec == NSFV1_F32 && (parts & NSFV1_PART_A) && h->difflen == DL_RAW &&
        !h->nu_elem.p && !h->ip_imports() && k.stab_type == STAB_FIELDS && !k.stokes && !k.pac && !k.peclet &&
        k.exact == 0.0 &&
        k.conv_upwind == k.stab_upwind && (k.conv_upwind == UPW_NO || k.conv_upwind == UPW_FULL || k.conv_upwind == UPW_SKEWED) &&
        ((uintptr_t)jvals & 15) == 0 && ((uintptr_t)defect & 15) == 0)
        mfast = k.conv_upwind;
    if (h->mixed()) {
        // a mixed mesh: every element group on the fused kernel (or the
        // coupled path), in group order on the one stream; the first group
        // assembles every row from zero, each further group adds its
        // elements to the rows it touches.  The coupled path's record
        // buffer is reused group after group (stream order).
        const bool cpl = coupled_config(h);
        void* crecs = nullptr;
        if (cpl) {
            if (h->upw_mode != NSFV1_UPWIND_FRESH)
                return fail(NSFV1_ERR_NOT_IMPLEMENTED, "recorded upwind decisions are not available with the FLOW "
                                                       "stabilization or the positive upwind");
            size_t need = (size_t)coupled_record_bytes(h->et, jprec, h->ne);
            for (auto& gp : h->groups) need = std::max(need, (size_t)coupled_record_bytes(gp->et, jprec, gp->ne));
            if (scratch) {
                if (need > NOALLOC_SCRATCH_BYTES)
                    return fail(NSFV1_ERR_ARGUMENT, "the coupled path needs " + std::to_string(need) +
                                                        " bytes of SCVF records for this mesh; the no-allocation entry holds at most " +
                                                        std::to_string(NOALLOC_SCRATCH_BYTES));
                crecs = scratch;
            } else {
                if (h->cpl_recs.n < need && h->cpl_recs.alloc(need) != hipSuccess)
                    return fail(NSFV1_ERR_DEVICE, "coupled record buffer");
                crecs = h->cpl_recs.p;
            }
        }
        auto launch = [&](int et, const KMeshPtrs& mm, const KParams& kk, const Plan& PP, int64_t ne) {
            return cpl ? launch_assemble_coupled(et, jprec, mm, kk, u, u_old, jvals, defect, h->errflag.as<int>(), PP.nwg,
                                                 PP.lds, ne, crecs, nullptr, (hipStream_t)stream)
                       : launch_assemble(et, jprec, mm, kk, u, u_old, jvals, defect, h->errflag.as<int>(), PP.nwg, PP.lds,
                                         (hipStream_t)stream);
        };
        // J and d start from zero; the first group writes the rows of its
        // nodes, the others add to them
        e = hipSuccess;
        if (jvals) e = launch_zero_words(jvals, h->nnz * (jprec == NSFV1_F64 ? 2 : 1), (hipStream_t)stream);
        if (e == hipSuccess && defect) e = launch_zero_words(defect, h->ndof * 2, (hipStream_t)stream);
        KParams k0 = k;
        k0.accum = 2;
        if (e == hipSuccess) e = launch(h->et, m, k0, P, h->ne);
        // imports at the integration points run through the groups in order
        int64_t off_ip = h->ne * nip_of(h->et), off_sh = h->ne * nsh_count(h->et);
        for (size_t g = 0; e == hipSuccess && g < h->groups.size(); ++g) {
            ElemGroup& G = *h->groups[g];
            const Plan& PG = jprec == NSFV1_F64 ? G.plan64 : G.plan32;
            KMeshPtrs mg = m;
            mg.conn = G.conn.as<int32_t>();
            mg.node_inc = G.node_inc.as<int32_t>();
            mg.inc_elem = G.inc_elem.as<int32_t>();
            mg.inc_node = G.inc_node.as<int32_t>();
            mg.inc_info = G.inc_info.as<int32_t>();
            mg.wg_node = PG.wg_node.as<int32_t>();
            mg.wg_rounds = PG.wg_rounds.as<int32_t>();
            mg.l2inv = G.l2inv.as<double>();
            mg.vol = G.vol.as<double>();
            mg.upw_dec = G.upw_dec.as<uint8_t>();
            mg.affine = G.affine.as<uint8_t>();
            mg.inc_pos = G.inc_pos.as<uint64_t>();
            mg.perm = PG.perm.as<int32_t>();
            mg.block = PG.block;
            mg.lane_a = PG.lane_a.as<int4>();
            mg.lane_pos = PG.lane_pos.as<uint64_t>();
            mg.wg_a = PG.wg_a.as<int4>();
            mg.nu_ip = h->nu_ip.p ? h->nu_ip.as<double>() + off_ip : nullptr;
            mg.f_scv = h->f_scv.p ? h->f_scv.as<double>() + off_sh * h->dim : nullptr;
            mg.f_scvf = h->f_scvf.p ? h->f_scvf.as<double>() + off_ip * h->dim : nullptr;
            mg.rho_ip = h->rho_ip.p ? h->rho_ip.as<double>() + off_ip : nullptr;
            mg.rho_scv = h->rho_scv.p ? h->rho_scv.as<double>() + off_sh : nullptr;
            off_ip += G.ne * nip_of(G.et);
            off_sh += G.ne * nsh_count(G.et);
            KParams kg = k;
            kg.accum = 1;
            e = launch(G.et, mg, kg, PG, G.ne);
        }
        h->last_path = cpl ? 4 : 1;
    } else if (mfast >= 0) {
        MarchGrid g{};
        g.nx = h->march_n[0];
        g.ny = h->march_n[1];
        g.nz = h->march_n[2];
        g.NX = g.nx + 1;
        g.NXY = g.NX * (g.ny + 1);
        g.tiles_x = (g.nx + 1 + MARCH_TX - 1) / MARCH_TX;
        g.tiles_y = (g.ny + 1 + MARCH_TY - 1) / MARCH_TY;
        g.zc = h->march_zc > 0 ? std::min(std::max(h->march_zc, 2), g.nz + 1) : std::min(16, g.nz + 1);
        g.nzc = (g.nz + 1 + g.zc - 1) / g.zc;
        g.units = (int64_t)g.tiles_x * g.tiles_y * g.nzc;
        e = launch_assemble_march(mfast, g, h->xyz.as<double>(), k, u, u_old, h->upw_dec.as<uint8_t>(),
                                  static_cast<float*>(jvals), defect, h->errflag.as<int>(), (hipStream_t)stream);
        h->last_path = 3;
    } else if (coupled_config(h)) {
        if (h->upw_mode != NSFV1_UPWIND_FRESH)
            return fail(NSFV1_ERR_NOT_IMPLEMENTED, "recorded upwind decisions are not available with the FLOW "
                                                   "stabilization or the positive upwind");
        const size_t need = (size_t)coupled_record_bytes(h->et, jprec, h->ne);
        void* recs = scratch;
        if (scratch) {
            if (need > NOALLOC_SCRATCH_BYTES)
                return fail(NSFV1_ERR_ARGUMENT, "the coupled path needs " + std::to_string(need) +
                                                    " bytes of SCVF records for this mesh; the no-allocation entry holds at most " +
                                                    std::to_string(NOALLOC_SCRATCH_BYTES));
        } else {
            if (h->cpl_recs.n < need && h->cpl_recs.alloc(need) != hipSuccess)
                return fail(NSFV1_ERR_DEVICE, "coupled record buffer");
            recs = h->cpl_recs.p;
        }
        e = launch_assemble_coupled(h->et, jprec, m, k, u, u_old, jvals, defect, h->errflag.as<int>(), P.nwg, P.lds,
                                    h->ne, recs, nullptr, (hipStream_t)stream);
        h->last_path = 4;
    } else if ((h->path == 0 || h->path == 2 || h->path == 3) && jprec == NSFV1_F32 && (parts & NSFV1_PART_A) &&
               !h->ip_imports() && split_supported(h->et, k)) {
        // also without a Jacobian (defect only): the gather then reads only
        // the records' fluxes, and the defect is bitwise the one assembled
        // together with the Jacobian
        const size_t need = sizeof(float) * (size_t)split_record_floats(h->et, h->ne);
        float* recs = static_cast<float*>(scratch);
        if (scratch) {
            if (need > NOALLOC_SCRATCH_BYTES)
                return fail(NSFV1_ERR_ARGUMENT, "the split path needs " + std::to_string(need) +
                                                    " bytes of SCVF records for this mesh; the no-allocation entry holds at most " +
                                                    std::to_string(NOALLOC_SCRATCH_BYTES));
        } else {
            if (h->recs.n < need && h->recs.alloc(need) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "record buffer");
            recs = h->recs.as<float>();
        }
        e = launch_assemble_split(h->et, m, k, u, u_old, static_cast<float*>(jvals), defect, h->errflag.as<int>(), P.nwg,
                                  P.lds, h->ne, recs, h->path == 2 && P.rowlane, (hipStream_t)stream);
        h->last_path = (h->path == 2 && P.rowlane) ? 2 : 0;
    } else {
        e = launch_assemble(h->et, jprec, m, k, u, u_old, jvals, defect, h->errflag.as<int>(), P.nwg, P.lds,
                            (hipStream_t)stream);
        h->last_path = 1;
    }
    if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("assemble launch: ") + hipGetErrorString(e));
    // boundary discs, group by group in order (each launch adds into the
    // rows of its boundary nodes after the launches before it)
    for (int kind = 0; kind < BND_KINDS && (parts & NSFV1_PART_A); ++kind)
        for (size_t g = 0; g <= h->groups.size(); ++g) {
            const bool own = g == 0;
            const ElemGroup* G = own ? nullptr : h->groups[g - 1].get();
            const BndList& B = own ? h->bnd[kind] : G->bnd[kind];
            if (B.nbn == 0) continue;
            OutflowArgs oa{h->xyz.as<double>(), (own ? h->conn : G->conn).as<int32_t>(), h->rowptr.as<int64_t>(),
                           (own ? h->inc_elem : G->inc_elem).as<int32_t>(), (own ? h->inc_info : G->inc_info).as<int32_t>(),
                           (own ? h->inc_pos : G->inc_pos).as<uint64_t>(), B.ptr.as<int32_t>(), B.node.as<int64_t>(),
                           B.inc.as<int32_t>(), B.side.as<int32_t>(), B.nbn, h->nu, h->rho, s_a, h->laplace ? 1 : 0,
                           h->stokes ? 1 : 0, kind, h->slide_k, h->slide_lim, B.nu.as<double>()};
            e = launch_outflow(own ? h->et : G->et, jprec, oa, u, jvals, defect, h->errflag.as<int>(), (hipStream_t)stream);
            if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("boundary disc launch: ") + hipGetErrorString(e));
        }
    if (h->upw_mode == NSFV1_UPWIND_RECORD && (parts & NSFV1_PART_A)) h->upw_recorded = true;
    return NSFV1_OK;
}

int nsfv1_assemble(nsfv1_handle* h, const double* u, const double* u_old, double dt, double s_a, double s_m,
                   double s_f, int parts, void* jvals, int jprec, double* defect, void* stream)
{
    return assemble_impl(h, u, u_old, dt, s_a, s_m, s_f, parts, jvals, jprec, defect, stream, nullptr);
}

int nsfv1_assemble_noalloc(nsfv1_handle* h, const double* u, const double* u_old, double dt, double s_a, double s_m,
                           double s_f, int parts, void* jvals, int jprec, double* defect, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    void* scratch = noalloc_scratch(h->device);
    if (!scratch) {
        // the handle was created before the lookup could run on this device
        hipError_t e = noalloc_scratch_setup(h->device);
        if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("no-allocation scratch: ") + hipGetErrorString(e));
        scratch = noalloc_scratch(h->device);
    }
    return assemble_impl(h, u, u_old, dt, s_a, s_m, s_f, parts, jvals, jprec, defect, stream, scratch);
}

int nsfv1_batch_set_instances(nsfv1_handle* h, int64_t ninst, const int64_t* elem_offset, const double* nu)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (h->mixed() && ninst != 0) return fail(NSFV1_ERR_NOT_IMPLEMENTED, "batched instances of mixed-element meshes are not provided");
    if (ninst == 0) {
        h->nu_elem.release();
        h->ninst = 0;
        h->nu_set = h->nu_global_set;  // back to the global viscosity, if one was set
        return NSFV1_OK;
    }
    if (ninst < 0 || !elem_offset || !nu) return fail(NSFV1_ERR_ARGUMENT, "bad instance arrays");
    if (elem_offset[0] != 0 || elem_offset[ninst] != h->ne)
        return fail(NSFV1_ERR_ARGUMENT, "instance element ranges must cover the mesh: offset[0] = 0, offset[n] = ne");
    // validate every range and viscosity before anything is written
    for (int64_t i = 0; i < ninst; ++i) {
        if (elem_offset[i + 1] < elem_offset[i] || elem_offset[i + 1] > h->ne)
            return fail(NSFV1_ERR_ARGUMENT, "instance element ranges not monotone within [0, ne]");
        if (!(nu[i] > 0.0)) return fail(NSFV1_ERR_ARGUMENT, "instance viscosity must be positive");
    }
    std::vector<double> ne_nu((size_t)h->ne);
    for (int64_t i = 0; i < ninst; ++i)
        for (int64_t e = elem_offset[i]; e < elem_offset[i + 1]; ++e) ne_nu[e] = nu[i];
    if (h->nu_elem.alloc(sizeof(double) * h->ne) != hipSuccess ||
        hipMemcpy(h->nu_elem.p, ne_nu.data(), sizeof(double) * h->ne, hipMemcpyHostToDevice) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "instance viscosity upload");
    h->ninst = ninst;
    h->nu_set = true;
    return NSFV1_OK;
}

}  // extern "C"

template <int ET>
static void side_corners(int s, int* n, int* c)
{
    *n = Topo<ET>::side_n[s];
    for (int j = 0; j < *n; ++j) c[j] = Topo<ET>::side[s][j];
}

// Groups the corners of the listed (element, side) pairs by node, keeping the
// input order within a node, and uploads them as boundary list `kind` of
// each element group (a mixed mesh numbers its elements group after group).
static int set_boundary_list(nsfv1_handle* h, int kind, const int32_t* elem, const int32_t* side, int64_t n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    h->bnd[kind].nbn = 0;
    h->bnd[kind].nu.release();
    for (auto& g : h->groups) {
        g->bnd[kind].nbn = 0;
        g->bnd[kind].nu.release();
    }
    h->bnd_elem_in[kind].clear();
    h->bnd_side_in[kind].clear();
    if (n == 0) return NSFV1_OK;
    if (n < 0 || !elem || !side) return fail(NSFV1_ERR_ARGUMENT, "bad boundary side arrays");
    // the element groups: type, element range, connectivity, incidences, list
    struct GV {
        int et;
        int64_t e0, ne;
        const DevBuf* conn;
        const std::vector<int64_t>* node_inc;
        const std::vector<int32_t>* inc_elem;
        BndList* B;
        std::vector<int32_t> cn;  // connectivity, downloaded when first needed
        std::vector<std::pair<int64_t, std::pair<int32_t, int32_t>>> items;  // (node, (incidence, side))
        std::vector<int32_t> src;                                           // 4 b + j of each item
    };
    std::vector<GV> gv;
    gv.push_back({h->et, 0, h->ne, &h->conn, &h->h_node_inc, &h->h_inc_elem, &h->bnd[kind], {}, {}, {}});
    int64_t e0 = h->ne;
    for (auto& g : h->groups) {
        gv.push_back({g->et, e0, g->ne, &g->conn, &g->h_node_inc, &g->h_inc_elem, &g->bnd[kind], {}, {}, {}});
        e0 += g->ne;
    }
    for (int64_t b = 0; b < n; ++b) {
        const int64_t eg = elem[b];
        const int s = side[b];
        if (eg < 0 || eg >= e0) return fail(NSFV1_ERR_ARGUMENT, "boundary element out of range");
        size_t g = 0;
        while (eg >= gv[g].e0 + gv[g].ne) ++g;
        GV& G = gv[g];
        const int64_t e = eg - G.e0;
        const int nsh = nsh_of(G.et);
        const int nside = G.et == ET_HEX ? 6 : (G.et == ET_PRISM ? 5 : (G.et == ET_QUAD ? 4 : (G.et == ET_TET ? 4 : 3)));
        if (s < 0 || s >= nside) return fail(NSFV1_ERR_ARGUMENT, "boundary side out of range");
        if (G.cn.empty()) {
            G.cn.resize((size_t)G.ne * nsh);
            if (hipMemcpy(G.cn.data(), G.conn->p, sizeof(int32_t) * G.cn.size(), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(NSFV1_ERR_DEVICE, "connectivity download");
        }
        int k = 0, c[4];
        switch (G.et) {
        case 0: side_corners<0>(s, &k, c); break;
        case 1: side_corners<1>(s, &k, c); break;
        case 2: side_corners<2>(s, &k, c); break;
        case 3: side_corners<3>(s, &k, c); break;
        default: side_corners<4>(s, &k, c); break;
        }
        for (int j = 0; j < k; ++j) {
            const int64_t v = G.cn[e * nsh + c[j]];
            int64_t q = -1;
            for (int64_t t = (*G.node_inc)[v]; t < (*G.node_inc)[v + 1]; ++t)
                if ((*G.inc_elem)[t] == e) { q = t; break; }
            if (q < 0) return fail(NSFV1_ERR_ARGUMENT, "boundary side corner without incidence");
            G.items.push_back({v, {(int32_t)q, (int32_t)s}});
            G.src.push_back((int32_t)(4 * b + j));
        }
    }
    auto up = [](DevBuf& d, const void* src, size_t bytes) {
        return d.alloc(bytes) == hipSuccess && hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    for (GV& G : gv) {
        if (G.items.empty()) continue;
        // stable order by node (input order within a node), the source of
        // each item carried along
        std::vector<size_t> ord(G.items.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return G.items[x].first < G.items[y].first; });
        std::vector<int64_t> nodes;
        std::vector<int32_t> ptr{0}, iq, is;
        G.B->src.clear();
        for (size_t o : ord) {
            if (nodes.empty() || nodes.back() != G.items[o].first) {
                if (!nodes.empty()) ptr.push_back((int32_t)iq.size());
                nodes.push_back(G.items[o].first);
            }
            iq.push_back(G.items[o].second.first);
            is.push_back(G.items[o].second.second);
            G.B->src.push_back(G.src[o]);
        }
        ptr.push_back((int32_t)iq.size());
        BndList& B = *G.B;
        if (!up(B.ptr, ptr.data(), sizeof(int32_t) * ptr.size()) ||
            !up(B.node, nodes.data(), sizeof(int64_t) * nodes.size()) ||
            !up(B.inc, iq.data(), sizeof(int32_t) * iq.size()) || !up(B.side, is.data(), sizeof(int32_t) * is.size()))
            return fail(NSFV1_ERR_DEVICE, "boundary list upload");
        B.nbn = (int64_t)nodes.size();
    }
    h->bnd_elem_in[kind].assign(elem, elem + n);
    h->bnd_side_in[kind].assign(side, side + n);
    return NSFV1_OK;
}

extern "C" {

int nsfv1_set_outflow(nsfv1_handle* h, const int32_t* elem, const int32_t* side, int64_t n)
{
    return set_boundary_list(h, BND_OUTFLOW, elem, side, n);
}

int nsfv1_set_symmetry(nsfv1_handle* h, const int32_t* elem, const int32_t* side, int64_t n)
{
    return set_boundary_list(h, BND_SYMMETRY, elem, side, n);
}

int nsfv1_set_sliding_wall(nsfv1_handle* h, const int32_t* elem, const int32_t* side, int64_t n,
                           double sliding_factor, double sliding_limit)
{
    if (h && n > 0 && !(std::isfinite(sliding_factor) && std::isfinite(sliding_limit)))
        return fail(NSFV1_ERR_ARGUMENT, "sliding factor and limit must be finite");
    int rc = set_boundary_list(h, BND_SLIDING, elem, side, n);
    if (rc) return rc;
    h->slide_k = sliding_factor;
    h->slide_lim = sliding_limit;
    return NSFV1_OK;
}

// Positions of the boundary-face ips of boundary list `kind`, in input
// order: out[b][j][dim] for pair b and the side's j-th corner (j < 4; unused
// slots of triangles and 2D sides are 0).  The FV1 boundary face of corner
// c on a side: 2D [x_c, side midpoint], 3D [x_c, mid(c, next), side centre,
// mid(prev, c)]; its ip is the mean of those points.
int nsfv1_bnd_ip_positions(nsfv1_handle* h, int kind, double* out, int64_t n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (kind < 0 || kind >= BND_KINDS) return fail(NSFV1_ERR_ARGUMENT, "kind: 0 outflow, 1 symmetry, 2 sliding wall");
    const auto& E = h->bnd_elem_in[kind];
    const auto& S = h->bnd_side_in[kind];
    const int dim = h->dim;
    if (!out || n != (int64_t)E.size() * 4 * dim)
        return fail(NSFV1_ERR_ARGUMENT, "expected " + std::to_string(E.size() * 4 * dim) + " values");
    std::fill(out, out + n, 0.0);
    if (E.empty()) return NSFV1_OK;
    std::vector<double> xyz((size_t)h->nv * dim);
    if (hipMemcpy(xyz.data(), h->xyz.p, sizeof(double) * xyz.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "mesh download");
    // the groups' connectivities, downloaded when first needed
    std::vector<std::vector<int32_t>> cn(1 + h->groups.size());
    for (size_t b = 0; b < E.size(); ++b) {
        int64_t e = E[b];
        size_t g = 0;
        int et = h->et;
        int64_t ne = h->ne;
        const DevBuf* cb = &h->conn;
        while (e >= ne) {
            e -= ne;
            const ElemGroup& G = *h->groups[g++];
            et = G.et;
            ne = G.ne;
            cb = &G.conn;
        }
        const int nsh = nsh_count(et);
        if (cn[g].empty()) {
            cn[g].resize((size_t)ne * nsh);
            if (hipMemcpy(cn[g].data(), cb->p, sizeof(int32_t) * cn[g].size(), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(NSFV1_ERR_DEVICE, "connectivity download");
        }
        int k = 0, c[4];
        switch (et) {
        case 0: side_corners<0>(S[b], &k, c); break;
        case 1: side_corners<1>(S[b], &k, c); break;
        case 2: side_corners<2>(S[b], &k, c); break;
        case 3: side_corners<3>(S[b], &k, c); break;
        default: side_corners<4>(S[b], &k, c); break;
        }
        auto X = [&](int corner, int i) { return xyz[(size_t)cn[g][(size_t)e * nsh + corner] * dim + i]; };
        double fc[3] = {0, 0, 0};
        for (int q = 0; q < k; ++q)
            for (int i = 0; i < dim; ++i) fc[i] += X(c[q], i) / k;
        for (int q = 0; q < k; ++q) {
            const int cc = c[q], nx = c[(q + 1) % k], pv = c[(q + k - 1) % k];
            double* o = out + ((int64_t)b * 4 + q) * dim;
            for (int i = 0; i < dim; ++i) {
                double a = 0.0;
                if (dim == 2) {
                    const int other = q == 0 ? c[1] : c[0];
                    a += X(cc, i);
                    a += 0.5 * (X(cc, i) + X(other, i));
                    o[i] = a / 2;
                } else {
                    a += X(cc, i);
                    a += 0.5 * (X(cc, i) + X(nx, i));
                    a += fc[i];
                    a += 0.5 * (X(pv, i) + X(cc, i));
                    o[i] = a / 4;
                }
            }
        }
    }
    return NSFV1_OK;
}

// The viscosity at the boundary-face ips of list `kind` (a function-valued
// viscosity evaluated at nsfv1_bnd_ip_positions): nu[b][j], 4 slots per
// pair; NULL clears (the constant viscosity then applies).
int nsfv1_set_boundary_viscosity(nsfv1_handle* h, int kind, const double* nu, int64_t n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (kind < 0 || kind >= BND_KINDS) return fail(NSFV1_ERR_ARGUMENT, "kind: 0 outflow, 1 symmetry, 2 sliding wall");
    std::vector<BndList*> lists{&h->bnd[kind]};
    for (auto& g : h->groups) lists.push_back(&g->bnd[kind]);
    if (!nu) {
        for (BndList* B : lists) B->nu.release();
        return NSFV1_OK;
    }
    if (n != (int64_t)h->bnd_elem_in[kind].size() * 4)
        return fail(NSFV1_ERR_ARGUMENT, "expected 4 values per boundary side: " +
                                            std::to_string(h->bnd_elem_in[kind].size() * 4));
    for (BndList* B : lists) {
        if (B->nbn == 0) continue;
        std::vector<double> v(B->src.size());
        for (size_t i = 0; i < v.size(); ++i) v[i] = nu[B->src[i]];
        int rc = upload(B->nu, v.data(), (int64_t)v.size());
        if (rc) return rc;
    }
    return NSFV1_OK;
}

int nsfv1_export_size(nsfv1_handle* h, int what, int at, int nip, int64_t* count)
{
    if (!h || !count) return fail(NSFV1_ERR_ARGUMENT, "NULL argument");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (what < NSFV1_EX_VELOCITY || what > NSFV1_EX_PRESSURE_GRAD) return fail(NSFV1_ERR_ARGUMENT, "unknown export");
    if (at < NSFV1_AT_SCVF || at > NSFV1_AT_LOCAL) return fail(NSFV1_ERR_ARGUMENT, "unknown export points");
    if (at == NSFV1_AT_LOCAL && (nip < 1 || nip > EX_MAX_LOCAL))
        return fail(NSFV1_ERR_ARGUMENT, "local points per call must be 1..32");
    if (at == NSFV1_AT_LOCAL && h->mixed())
        return fail(NSFV1_ERR_NOT_IMPLEMENTED, "exports at local points on mixed-element meshes are not provided "
                                               "(local coordinates belong to one reference element)");
    const int dim = h->dim;
    const int ncomp = what == NSFV1_EX_VELOCITY_GRAD ? dim * dim : (what == NSFV1_EX_PRESSURE ? 1 : dim);
    // group after group on a mixed mesh: [ne_g][points_g][ncomp] blocks
    *count = (at == NSFV1_AT_LOCAL ? h->ne * nip : group_points(h, at == NSFV1_AT_SCVF ? 0 : 1)) * ncomp;
    return NSFV1_OK;
}

int nsfv1_export(nsfv1_handle* h, int what, int at, const double* local_ips, int nip, const double* u,
                 const double* u_old, double dt, double* out, void* stream)
{
    int64_t count = 0;
    int rc = nsfv1_export_size(h, what, at, nip, &count);
    if (rc) return rc;

    if (!u || !out) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (at == NSFV1_AT_LOCAL && !local_ips) return fail(NSFV1_ERR_ARGUMENT, "local points are NULL");
    const int dim = (h->et == ET_TRI || h->et == ET_QUAD) ? 2 : 3;
    if (what == NSFV1_EX_VELOCITY_IP && at == NSFV1_AT_SCVF) {
        // the stabilisation's update (ex_velocity_ip :1146-1189)
        if (h->stab_type == STAB_UNSET) return fail(NSFV1_ERR_NO_STABILIZATION, "Stabilization has not been set.");
        if (!h->stokes && h->stab_upwind == UPW_UNSET)
            return fail(NSFV1_ERR_NO_UPWIND, "The stabilization has no upwind object (INavierStokesFV1Stabilization::set_upwind).");
        if (!h->nu_set) return fail(NSFV1_ERR_NO_VISCOSITY, "Kinematic Viscosity has not been set, but is required.");
        if (u_old && !(dt > 0.0)) return fail(NSFV1_ERR_TIME_SERIES, "NavierStokes: time step must be positive for the time-dependent stabilization");
        rc = ensure_geometry(h, stream);
        if (rc) return rc;
        KParams k{};
        k.stab_upwind = h->stab_upwind;
        k.conv_upwind = h->stab_upwind;  // the convective upwind plays no part in the stabilised velocity
        k.stab_type = h->stab_type;
        k.difflen = h->difflen;
        k.stokes = h->stokes;
        k.laplace = h->laplace;
        k.has_source = h->has_source;
        k.has_old = u_old != nullptr;
        k.parts = NSFV1_PART_A;
        k.nu = h->nu;
        k.rho = h->rho;
        k.dt = dt;
        k.inv_dt = u_old ? 1.0 / dt : 0.0;
        k.inv_rho = 1.0 / h->rho;
        for (int i = 0; i < 3; ++i) k.f[i] = h->f[i];
        k.s_a = 1.0;
        k.upw_mode = NSFV1_UPWIND_FRESH;
        KMeshPtrs m{};
        m.xyz = h->xyz.as<double>();
        m.conn = h->conn.as<int32_t>();
        m.l2inv = h->l2inv.as<double>();
        m.vol = h->vol.as<double>();
        m.affine = h->affine.as<uint8_t>();
        m.nu_elem = h->nu_elem.as<double>();
        m.nu_ip = h->nu_ip.as<double>();
        m.f_scv = h->f_scv.as<double>();
        m.f_scvf = h->f_scvf.as<double>();
        m.rho_ip = h->rho_ip.as<double>();
        if (coupled_config(h)) {
            k.conv_upwind = h->conv_upwind;
            k.pac = h->pac;
        }
        auto launch = [&](int et, const KMeshPtrs& mm, int64_t ne, double* o) {
            return coupled_config(h)
                       ? launch_assemble_coupled(et, NSFV1_F64, mm, k, u, u_old, nullptr, nullptr, h->errflag.as<int>(), 0,
                                                 0, ne, nullptr, o, (hipStream_t)stream)
                       : launch_export_stab_vel(et, mm, k, u, u_old, ne, o, h->errflag.as<int>(), (hipStream_t)stream);
        };
        hipError_t e = launch(h->et, m, h->ne, out);
        // further element groups: their elements' blocks follow in order
        int64_t off_ip = h->ne * nip_of(h->et), off_sh = h->ne * nsh_count(h->et);
        for (size_t g = 0; e == hipSuccess && g < h->groups.size(); ++g) {
            const ElemGroup& G = *h->groups[g];
            KMeshPtrs mg = m;
            mg.conn = G.conn.as<int32_t>();
            mg.l2inv = G.l2inv.as<double>();
            mg.vol = G.vol.as<double>();
            mg.affine = G.affine.as<uint8_t>();
            mg.nu_ip = h->nu_ip.p ? h->nu_ip.as<double>() + off_ip : nullptr;
            mg.f_scv = h->f_scv.p ? h->f_scv.as<double>() + off_sh * dim : nullptr;
            mg.f_scvf = h->f_scvf.p ? h->f_scvf.as<double>() + off_ip * dim : nullptr;
            mg.rho_ip = h->rho_ip.p ? h->rho_ip.as<double>() + off_ip : nullptr;
            e = launch(G.et, mg, G.ne, out + off_ip * dim);
            off_ip += G.ne * nip_of(G.et);
            off_sh += G.ne * nsh_count(G.et);
        }
        if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("export launch: ") + hipGetErrorString(e));
        return NSFV1_OK;
    }
    ExportArgs a{};
    a.xyz = h->xyz.as<double>();
    a.what = what;
    a.at = at;
    a.ncomp = what == NSFV1_EX_VELOCITY_GRAD ? dim * dim : (what == NSFV1_EX_PRESSURE ? 1 : dim);
    if (at == NSFV1_AT_LOCAL)
        for (int q = 0; q < nip; ++q)
            for (int i = 0; i < dim; ++i) a.lips[q][i] = local_ips[q * dim + i];
    // group after group (the handle's own elements first)
    auto launch = [&](int et, const DevBuf& conn, int64_t ne, double* o) {
        a.conn = conn.as<int32_t>();
        a.ne = ne;
        a.npt = at == NSFV1_AT_SCVF ? nip_of(et) : (at == NSFV1_AT_SCV ? nsh_count(et) : nip);
        return launch_export_points(et, a, u, o, h->errflag.as<int>(), (hipStream_t)stream);
    };
    hipError_t e = launch(h->et, h->conn, h->ne, out);
    int64_t off = h->ne * (at == NSFV1_AT_SCVF ? nip_of(h->et) : (at == NSFV1_AT_SCV ? nsh_count(h->et) : nip)) * a.ncomp;
    for (size_t g = 0; e == hipSuccess && g < h->groups.size(); ++g) {
        const ElemGroup& G = *h->groups[g];
        e = launch(G.et, G.conn, G.ne, out + off);
        off += G.ne * (at == NSFV1_AT_SCVF ? nip_of(G.et) : nsh_count(G.et)) * a.ncomp;
    }
    if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("export launch: ") + hipGetErrorString(e));
    return NSFV1_OK;
}

int nsfv1_set_assembly_path(nsfv1_handle* h, int path)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (path < 0 || path > 3)
        return fail(NSFV1_ERR_ARGUMENT, "path must be 0 (split where supported), 1 (fused), 2 (split, row lanes) or "
                                        "3 (plane march where supported)");
    if (path != h->path) invalidate_plans(h);  // the row-lane plan packs fewer incidences
    h->path = path;
    return NSFV1_OK;
}

int nsfv1_assembly_path_used(nsfv1_handle* h, int* path)
{
    if (!h || !path) return fail(NSFV1_ERR_ARGUMENT, "NULL argument");
    *path = h->last_path;
    return NSFV1_OK;
}

int nsfv1_set_upwind_decisions(nsfv1_handle* h, int mode)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (mode != NSFV1_UPWIND_FRESH && mode != NSFV1_UPWIND_RECORD && mode != NSFV1_UPWIND_REPLAY)
        return fail(NSFV1_ERR_ARGUMENT, "upwind decision mode must be 0 (fresh), 1 (record) or 2 (replay)");
    h->upw_mode = mode;
    return NSFV1_OK;
}

// Synchronises the stream and reports (then clears) errors raised inside
// the kernels: a skewed/LPS ray that finds no side (upwind.cpp:354) or a
// degenerate element.
int nsfv1_sync_errors(nsfv1_handle* h, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    int flag[KERR_WORDS] = {0, 0, 0, 0};
    if (hipMemcpyAsync(flag, h->errflag.p, sizeof(flag), hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
        hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "error-flag copy");
    if (flag[KERR_CUT_SIDE] || flag[KERR_DEGENERATE] || flag[KERR_SINGULAR_BLOCK]) {
        if (hipMemsetAsync(h->errflag.p, 0, sizeof(flag), (hipStream_t)stream) != hipSuccess ||
            hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
            return fail(NSFV1_ERR_DEVICE, "error-flag reset");
        if (flag[KERR_CUT_SIDE]) return fail(NSFV1_ERR_CUT_SIDE, "GetNodeNextToCut: Cannot find cut side.");
        if (flag[KERR_DEGENERATE]) return fail(NSFV1_ERR_DEGENERATE, "degenerate element (zero Jacobian determinant)");
        return fail(NSFV1_ERR_SINGULAR_BLOCK, "block-Jacobi setup: singular diagonal block in the Jacobian");
    }
    return NSFV1_OK;
}


// Dirichlet DoFs (host array, any order); rows of J become identity rows and
// the defect entries u - g (ugcore DirichletBoundary row treatment).
int nsfv1_set_dirichlet(nsfv1_handle* h, const int64_t* dofs, int64_t n)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    std::vector<int64_t> d(dofs, dofs + (n > 0 ? n : 0));
    for (int64_t x : d)
        if (x < 0 || x >= h->ndof) return fail(NSFV1_ERR_ARGUMENT, "Dirichlet DoF out of range");
    if (h->dir_rows.alloc(sizeof(int64_t) * std::max<int64_t>(n, 1)) != hipSuccess ||
        (n > 0 && hipMemcpy(h->dir_rows.p, d.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice) != hipSuccess))
        return fail(NSFV1_ERR_DEVICE, "Dirichlet upload");
    h->ndir = n;
    return NSFV1_OK;
}

int nsfv1_apply_dirichlet(nsfv1_handle* h, void* jvals, int jprec, double* defect, const double* u,
                          const double* g, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (defect && !u) return fail(NSFV1_ERR_ARGUMENT, "u is NULL");
    hipError_t e = launch_dirichlet(jprec, h->rowptr.as<int64_t>(), h->colidx.as<int32_t>(), h->dir_rows.as<int64_t>(),
                                    h->ndir, jvals, defect, u, g, (hipStream_t)stream);
    if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("dirichlet: ") + hipGetErrorString(e));
    return NSFV1_OK;
}

int nsfv1_spmv(nsfv1_handle* h, const void* jvals, int jprec, const double* x, double* y, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    hipError_t e = launch_spmv(jprec, h->rowptr.as<int64_t>(), h->colidx.as<int32_t>(), h->adj.as<int32_t>(),
                               h->dim + 1, jvals, x, y, h->ndof, h->nnz, (hipStream_t)stream);
    if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("spmv: ") + hipGetErrorString(e));
    return NSFV1_OK;
}


// ---------------------------------------------------------------------
// geometric multigrid preconditioner (V-cycle, damped point-block Jacobi
// smoothing, Galerkin-free: every coarse level re-assembles the same
// discretisation on its own mesh at the injected state)
// ---------------------------------------------------------------------
int nsfv1_set_preconditioner(nsfv1_handle* h, const char* name)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    const std::string n = trim_lower(name);
    if (n == "block-jacobi" || n == "jacobi") h->use_mg = false;
    else if (n == "gmg" || n == "multigrid") h->use_mg = true;
    else return fail(NSFV1_ERR_UNKNOWN_NAME, "preconditioner '" + std::string(name ? name : "") + "' not found. Options are: block-jacobi, gmg");
    return NSFV1_OK;
}

// coarse levels carry the configuration of the fine discretisation
static void mg_sync_settings(nsfv1_handle* h)
{
    for (auto& L : h->mg) {
        nsfv1_handle* c = L->h;
        c->nu = h->nu;
        c->nu_set = h->nu_set;
        c->nu_global_set = h->nu_global_set;
        c->rho = h->rho;
        c->has_source = h->has_source;
        for (int i = 0; i < 3; ++i) c->f[i] = h->f[i];
        c->exact = h->exact;
        c->stokes = h->stokes;
        c->laplace = h->laplace;
        c->peclet = h->peclet;
        c->grad_div = h->grad_div;
        c->stab_type = h->stab_type;
        c->stab_upwind = h->stab_upwind;
        c->difflen = h->difflen;
        c->conv_upwind = h->conv_upwind;
        c->pac = h->pac;
        // imports at the integration points live per level (each level has
        // its own elements); a fine level without them clears the coarse ones
        if (!h->nu_ip.p) c->nu_ip.release();
        if (!h->rho_ip.p) {
            c->rho_ip.release();
            c->rho_scv.release();
        }
        if (!h->f_scv.p) {
            c->f_scv.release();
            c->f_scvf.release();
        }
    }
}

int nsfv1_mg_set_hierarchy(nsfv1_handle* h, int nlev, nsfv1_handle* const* coarse, const int64_t* const* prol_rowptr,
                           const int32_t* const* prol_colidx, const double* const* prol_weights,
                           const int64_t* const* inject, int pre, int post, double omega)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (nlev < 2 || nlev > MG_MAX_LEVELS) return fail(NSFV1_ERR_ARGUMENT, "multigrid needs 2.." + std::to_string(MG_MAX_LEVELS) + " levels");
    if (!coarse || !prol_rowptr || !prol_colidx || !prol_weights || !inject) return fail(NSFV1_ERR_ARGUMENT, "NULL level arrays");
    if (pre < 1 || post < 0 || pre > 16 || post > 16 || !(omega > 0.0 && omega <= 1.5))
        return fail(NSFV1_ERR_ARGUMENT, "smoothing steps must be 1..16 (pre) / 0..16 (post), omega in (0, 1.5]");
    std::vector<std::unique_ptr<MGLevel>> lv;
    int64_t nv_f = h->nv;
    for (int l = 1; l < nlev; ++l) {
        nsfv1_handle* c = coarse[l - 1];
        if (!c || c == h) return fail(NSFV1_ERR_ARGUMENT, "bad coarse handle");
        if (c->et < 0) return fail(NSFV1_ERR_NO_MESH, "coarse level without mesh");
        if (c->dim != h->dim) return fail(NSFV1_ERR_ARGUMENT, "coarse level of another dimension");
        const int64_t nv_c = c->nv;
        const int64_t* rp = prol_rowptr[l - 1];
        const int32_t* ci = prol_colidx[l - 1];
        const double* w = prol_weights[l - 1];
        const int64_t* ij = inject[l - 1];
        if (!rp || !ci || !w || !ij) return fail(NSFV1_ERR_ARGUMENT, "NULL transfer operator");
        if (rp[0] != 0) return fail(NSFV1_ERR_ARGUMENT, "prolongation rowptr must start at 0");
        for (int64_t r = 0; r < nv_f; ++r)
            if (rp[r + 1] < rp[r]) return fail(NSFV1_ERR_ARGUMENT, "prolongation rowptr not monotone");
        const int64_t pnnz = rp[nv_f];
        for (int64_t k = 0; k < pnnz; ++k)
            if (ci[k] < 0 || ci[k] >= nv_c) return fail(NSFV1_ERR_ARGUMENT, "prolongation column out of range");
        for (int64_t k = 0; k < nv_c; ++k)
            if (ij[k] < 0 || ij[k] >= nv_f) return fail(NSFV1_ERR_ARGUMENT, "injection index out of range");
        // restriction = P^T (host transpose, sorted by fine node)
        std::vector<int64_t> rrp(nv_c + 1, 0);
        for (int64_t k = 0; k < pnnz; ++k) ++rrp[ci[k] + 1];
        for (int64_t r = 0; r < nv_c; ++r) rrp[r + 1] += rrp[r];
        std::vector<int32_t> rci(pnnz);
        std::vector<double> rw(pnnz);
        std::vector<int64_t> fill(rrp.begin(), rrp.end() - 1);
        for (int64_t r = 0; r < nv_f; ++r)
            for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
                const int64_t q = fill[ci[k]]++;
                rci[q] = (int32_t)r;
                rw[q] = w[k];
            }
        auto L = std::make_unique<MGLevel>();
        L->h = c;
        L->nv_fine = nv_f;
        auto up = [](DevBuf& b, const void* src, size_t bytes) {
            return b.alloc(bytes) == hipSuccess && (bytes == 0 || hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice) == hipSuccess);
        };
        if (!up(L->P_rp, rp, sizeof(int64_t) * (nv_f + 1)) || !up(L->P_ci, ci, sizeof(int32_t) * pnnz) ||
            !up(L->P_w, w, sizeof(double) * pnnz) || !up(L->R_rp, rrp.data(), sizeof(int64_t) * (nv_c + 1)) ||
            !up(L->R_ci, rci.data(), sizeof(int32_t) * pnnz) || !up(L->R_w, rw.data(), sizeof(double) * pnnz) ||
            !up(L->inj, ij, sizeof(int64_t) * nv_c))
            return fail(NSFV1_ERR_DEVICE, "multigrid transfer upload");
        const int64_t n = c->ndof;
        if (L->Binv.alloc(sizeof(double) * c->nv * (c->dim + 1) * (c->dim + 1)) != hipSuccess ||
            L->u.alloc(sizeof(double) * n) != hipSuccess || L->uold.alloc(sizeof(double) * n) != hipSuccess ||
            L->b.alloc(sizeof(double) * n) != hipSuccess || L->x.alloc(sizeof(double) * n) != hipSuccess ||
            L->r.alloc(sizeof(double) * n) != hipSuccess || L->t.alloc(sizeof(double) * n) != hipSuccess)
            return fail(NSFV1_ERR_DEVICE, "multigrid workspace");
        lv.push_back(std::move(L));
        nv_f = nv_c;
    }
    const int64_t ncoarse = lv.back()->h->ndof;
    if (ncoarse > MG_DENSE_MAX)
        return fail(NSFV1_ERR_ARGUMENT, "coarsest level has " + std::to_string(ncoarse) + " DoFs; at most " +
                                            std::to_string(MG_DENSE_MAX) + " are solved densely");
    if (h->mg_r0.alloc(sizeof(double) * h->ndof) != hipSuccess || h->mg_t0.alloc(sizeof(double) * h->ndof) != hipSuccess ||
        h->mg_dense.alloc(sizeof(double) * ncoarse * ncoarse) != hipSuccess ||
        h->mg_gj.alloc(dense_inverse_work_bytes((int)ncoarse)) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "multigrid workspace");
    h->mg.swap(lv);
    mg_sync_settings(h);
    h->mg_pre = pre;
    h->mg_post = post;
    h->mg_omega = omega;
    h->mg_ncoarse = (int)ncoarse;
    h->mg_ready = false;
    return NSFV1_OK;
}

int nsfv1_mg_setup(nsfv1_handle* h, int jprec, const double* u, const double* u_old, double dt, double s_a,
                   double s_m, int parts, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->mg.empty()) return fail(NSFV1_ERR_ARGUMENT, "no multigrid hierarchy set");
    if (h->nu_elem.p)
        return fail(NSFV1_ERR_NOT_IMPLEMENTED, "multigrid with per-instance viscosity (batched instances) is not provided");
    // the coarse levels take the fine level's current configuration; imports
    // at the integration points must be given on every coarse level too
    // (the caller evaluates its function there), for the same quantities
    mg_sync_settings(h);
    for (auto& Lp : h->mg) {
        const nsfv1_handle* c = Lp->h;
        if ((h->nu_ip.p && !c->nu_ip.p) || (h->f_scv.p && !c->f_scv.p) || (h->rho_ip.p && !c->rho_ip.p))
            return fail(NSFV1_ERR_ARGUMENT, "multigrid with imports at the integration points: every coarse level "
                                            "needs its own imports of the same quantities (nsfv1_set_kinematic_"
                                            "viscosity_ip / nsfv1_set_source_ip / nsfv1_set_density_ip on the "
                                            "coarse handles)");
    }
    if (!u) return fail(NSFV1_ERR_ARGUMENT, "u is NULL");
    if (jprec != NSFV1_F32 && jprec != NSFV1_F64) return fail(NSFV1_ERR_ARGUMENT, "jprec must be 32 or 64");
    hipStream_t s = (hipStream_t)stream;
    const size_t js = jprec == NSFV1_F64 ? 8 : 4;
    const double* uf = u;
    const double* uof = u_old;
    for (auto& Lp : h->mg) {
        MGLevel& L = *Lp;
        nsfv1_handle* c = L.h;
        const int nc = c->dim + 1;
        if (L.jprec != jprec || L.J.p == nullptr) {
            if (L.J.alloc(js * c->nnz) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "multigrid Jacobian");
            L.jprec = jprec;
        }
        hipError_t e = launch_inject(L.inj.as<int64_t>(), uf, L.u.as<double>(), c->nv, nc, s);
        if (e == hipSuccess && uof) e = launch_inject(L.inj.as<int64_t>(), uof, L.uold.as<double>(), c->nv, nc, s);
        if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, "multigrid injection");
        int rc = nsfv1_assemble(c, L.u.as<double>(), uof ? L.uold.as<double>() : nullptr, dt, s_a, s_m, 0.0, parts,
                                L.J.p, jprec, nullptr, stream);
        if (rc) return rc;
        e = launch_dirichlet(jprec, c->rowptr.as<int64_t>(), c->colidx.as<int32_t>(), c->dir_rows.as<int64_t>(), c->ndir,
                             L.J.p, nullptr, nullptr, nullptr, s);
        if (e == hipSuccess)
            e = launch_block_jacobi_setup(jprec, nc, c->rowptr.as<int64_t>(), c->colidx.as<int32_t>(), L.J.p, c->nv,
                                          L.Binv.as<double>(), h->errflag.as<int>(), s);
        if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, "multigrid level setup");
        uf = L.u.as<double>();
        uof = uof ? L.uold.as<double>() : nullptr;
    }
    // coarsest level: dense inverse, Gauss-Jordan on the device
    MGLevel& C = *h->mg.back();
    nsfv1_handle* c = C.h;
    const int n = (int)c->ndof;
    int* gj_err = (int*)((char*)h->mg_gj.p + dense_inverse_work_bytes(n)) - 1;
    if (launch_dense_inverse(jprec, c->rowptr.as<int64_t>(), c->colidx.as<int32_t>(), C.J.p, n, h->mg_gj.p,
                             h->mg_dense.as<double>(), gj_err, s) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "coarse inverse");
    int herr = 0;
    if (hipMemcpyAsync(&herr, gj_err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "coarse inverse");
    if (herr) return fail(NSFV1_ERR_ARGUMENT, "coarsest multigrid level: a column without finite entries (non-finite state or Jacobian)");
    h->mg_jprec = jprec;
    h->mg_ready = true;
    return NSFV1_OK;
}

// x = V-cycle(b) on level l (l = 0: the fine system J0 with block inverse Binv0)
static hipError_t mg_cycle(nsfv1_handle* h, size_t l, const void* J0, const double* Binv0, const double* b, double* x,
                           hipStream_t s)
{
    const bool fine = l == 0;
    nsfv1_handle* c = fine ? h : h->mg[l - 1]->h;
    const int nc = c->dim + 1;
    const int jp = h->mg_jprec;
    if (l == h->mg.size()) return launch_dense_matvec(h->mg_dense.as<double>(), b, x, (int)c->ndof, s);
    const void* J = fine ? J0 : h->mg[l - 1]->J.p;
    const double* Bi = fine ? Binv0 : h->mg[l - 1]->Binv.as<double>();
    double* r = fine ? h->mg_r0.as<double>() : h->mg[l - 1]->r.as<double>();
    double* t = fine ? h->mg_t0.as<double>() : h->mg[l - 1]->t.as<double>();
    const int64_t* rp = c->rowptr.as<int64_t>();
    const int32_t* ci = c->colidx.as<int32_t>();
    const int32_t* ad = c->adj.as<int32_t>();
    const int64_t n = c->ndof, nv = c->nv;
    const double w = h->mg_omega;
    // pre-smoothing: x = w Binv b, then mg_pre - 1 sweeps alternating
    // between x and t; the first iterate goes to t when that count is odd,
    // so that the last one lands in x
    const int pre_sweeps = std::max(h->mg_pre, 1) - 1;
    double* cur = (pre_sweeps & 1) ? t : x;
    double* alt = (pre_sweeps & 1) ? x : t;
    hipError_t e = launch_bj_update(nc, Bi, b, cur, w, 1, nv, s);
    for (int k = 0; k < pre_sweeps && e == hipSuccess; ++k) {
        e = launch_smooth(jp, rp, ci, ad, nc, J, cur, b, Bi, w, r, alt, n, c->nnz, s);
        std::swap(cur, alt);
    }
    if (e != hipSuccess) return e;
    MGLevel& C = *h->mg[l];
    nsfv1_handle* cc = C.h;
    // Dirichlet rows neither send residual to nor take corrections from the
    // coarse level (their equations are decoupled identities)
    e = launch_residual(jp, rp, ci, ad, nc, J, x, b, r, n, c->nnz, s);
    if (e == hipSuccess) e = launch_zero_rows(c->dir_rows.as<int64_t>(), c->ndir, r, s);
    if (e == hipSuccess) e = launch_node_transfer(C.R_rp.as<int64_t>(), C.R_ci.as<int32_t>(), C.R_w.as<double>(), r, C.b.as<double>(), cc->nv, nc, 0, s);
    if (e == hipSuccess) e = launch_zero_rows(cc->dir_rows.as<int64_t>(), cc->ndir, C.b.as<double>(), s);
    if (e == hipSuccess) e = mg_cycle(h, l + 1, J0, Binv0, C.b.as<double>(), C.x.as<double>(), s);
    if (e == hipSuccess) e = launch_node_transfer(C.P_rp.as<int64_t>(), C.P_ci.as<int32_t>(), C.P_w.as<double>(), C.x.as<double>(), r, nv, nc, 0, s);
    if (e == hipSuccess) e = launch_zero_rows(c->dir_rows.as<int64_t>(), c->ndir, r, s);
    if (e == hipSuccess) e = launch_axpby(1.0, r, 1.0, x, n, s);
    cur = x, alt = t;
    for (int k = 0; k < h->mg_post && e == hipSuccess; ++k) {
        e = launch_smooth(jp, rp, ci, ad, nc, J, cur, b, Bi, w, r, alt, n, c->nnz, s);
        std::swap(cur, alt);
    }
    if (e == hipSuccess && cur != x) e = hipMemcpyAsync(x, cur, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    return e;
}


// ---------------------------------------------------------------------
// learned initial guess (MLP) and its training
// ---------------------------------------------------------------------
int nsfv1_mlp_sizes(int64_t n, int64_t* nparam, int64_t* work_floats)
{
    if (n < 0) return fail(NSFV1_ERR_ARGUMENT, "negative node count");
    if (nparam) *nparam = MLP_NPARAM;
    if (work_floats) *work_floats = mlp_work_floats(n);
    return NSFV1_OK;
}

int nsfv1_mlp_pack(const float* params, uint16_t* wbf16, void* stream)
{
    if (!params || !wbf16) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (launch_mlp_pack(params, wbf16, (hipStream_t)stream) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "mlp pack");
    return NSFV1_OK;
}

int nsfv1_mlp_forward(const float* params, const uint16_t* wbf16, const float* x, int64_t n, float* y, float* z1,
                      float* z2, void* stream)
{
    if (!params || !wbf16 || (n > 0 && (!x || !y))) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (((uintptr_t)x | (uintptr_t)y | (uintptr_t)z1 | (uintptr_t)z2) & 15)
        return fail(NSFV1_ERR_ARGUMENT, "MLP buffers must be 16-byte aligned");
    if (launch_mlp_forward(params, wbf16, x, n, y, z1, z2, (hipStream_t)stream) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "mlp forward");
    return NSFV1_OK;
}

int nsfv1_mlp_mse(const float* y, const float* t, int64_t n, float* dy, float* work, double* loss, void* stream)
{
    if (!y || !t || !dy || !work) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (n <= 0) return fail(NSFV1_ERR_ARGUMENT, "empty batch");
    int np = 0;
    hipStream_t s = (hipStream_t)stream;
    if (launch_mlp_mse(y, t, n, dy, work, &np, s) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "mlp mse");
    if (loss) {
        std::vector<float> h(np);
        if (hipMemcpyAsync(h.data(), work, sizeof(float) * np, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(NSFV1_ERR_DEVICE, "mlp loss copy");
        double a = 0;
        for (float v : h) a += v;
        *loss = a / (double)(n * MLP_OUT);
    }
    return NSFV1_OK;
}

int nsfv1_mlp_backward(const uint16_t* wbf16, const float* x, const float* z1, const float* z2, const float* dy,
                       int64_t n, float* grad, float* work, void* stream)
{
    if (!wbf16 || !x || !z1 || !z2 || !dy || !grad || !work) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (launch_mlp_backward(wbf16, x, z1, z2, dy, n, grad, work, (hipStream_t)stream) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "mlp backward");
    return NSFV1_OK;
}

int nsfv1_mlp_adam(float* params, float* m, float* v, const float* grad, int step, double lr, double beta1,
                   double beta2, double eps, uint16_t* wbf16, void* stream)
{
    if (!params || !m || !v || !grad || !wbf16) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (step < 1) return fail(NSFV1_ERR_ARGUMENT, "Adam step counts from 1");
    if (launch_mlp_adam(params, m, v, grad, step, (float)lr, (float)beta1, (float)beta2, (float)eps, wbf16,
                        (hipStream_t)stream) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "mlp adam");
    return NSFV1_OK;
}

static int predict_check(nsfv1_handle* h)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (h->dim != 3) return fail(NSFV1_ERR_ARGUMENT, "the learned predictor takes 3D states (4 components per node)");
    return NSFV1_OK;
}

int nsfv1_predict_features(nsfv1_handle* h, const double* u, const double* u_prev, const double* u_prev2,
                           const double* u_prev3, const float* scale, float* x, double* sigma, void* stream)
{
    int rc = predict_check(h);
    if (rc) return rc;
    if (!u || !u_prev || !u_prev2 || !u_prev3 || !scale || !x || !sigma) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if ((uintptr_t)x & 15) return fail(NSFV1_ERR_ARGUMENT, "feature buffer must be 16-byte aligned");
    if (h->pred_part.n < sizeof(double) * PRED_SCALE_BLOCKS &&
        h->pred_part.alloc(sizeof(double) * PRED_SCALE_BLOCKS) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "predictor scale buffer");
    hipStream_t s = (hipStream_t)stream;
    if (launch_predict_scale(u, u_prev, u_prev2, u_prev3, h->nv, h->pred_part.as<double>(), sigma, s) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "predict scale");
    if (launch_predict_features(h->rowptr.as<int64_t>(), h->colidx.as<int32_t>(), u, u_prev, u_prev2, u_prev3, h->nv,
                                scale, sigma, x, s) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "predict features");
    return NSFV1_OK;
}

int nsfv1_predict_guess(nsfv1_handle* h, const double* u, const double* u_prev, const double* u_prev2, const float* y,
                        const float* out_scale, const double* sigma, double* guess, void* stream)
{
    int rc = predict_check(h);
    if (rc) return rc;
    if (!u || !u_prev || !u_prev2 || !y || !out_scale || !sigma || !guess) return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (launch_predict_guess(u, u_prev, u_prev2, y, out_scale, sigma, h->nv, guess, (hipStream_t)stream) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "predict guess");
    return NSFV1_OK;
}

int nsfv1_predict_target(nsfv1_handle* h, const double* u_next, const double* u, const double* u_prev,
                         const double* u_prev2, const float* out_scale, const double* sigma, float* t, void* stream)
{
    int rc = predict_check(h);
    if (rc) return rc;
    if (!u_next || !u || !u_prev || !u_prev2 || !out_scale || !sigma || !t)
        return fail(NSFV1_ERR_ARGUMENT, "NULL buffer");
    if (launch_predict_target(u_next, u, u_prev, u_prev2, out_scale, sigma, h->nv, t, (hipStream_t)stream) !=
        hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "predict target");
    return NSFV1_OK;
}

// Restarted GMRES(m), right-preconditioned with point-block Jacobi, for
// J x = b with the device Jacobian (jprec) and FP64 vectors.  x is
// overwritten (zero initial guess).  Stops when ||b - J x|| <= rtol ||b||
// or after maxit inner iterations.
int nsfv1_linear_solve(nsfv1_handle* h, const void* jvals, int jprec, const double* b, double* x, double rtol,
                       int maxit, int restart, int* iters, double* relres, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (restart < 1 || restart > GMRES_MAX_RESTART) return fail(NSFV1_ERR_ARGUMENT, "restart must be 1..60");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->ndof, nv = h->nv;
    const int nc = h->dim + 1, m = restart;
    if (h->gm_restart != m) {
        if (h->gm_V.alloc(sizeof(double) * n * (m + 1)) != hipSuccess || h->gm_w.alloc(sizeof(double) * n * 2) != hipSuccess ||
            h->gm_Binv.alloc(sizeof(double) * nv * nc * nc) != hipSuccess ||
            h->gm_part.alloc(sizeof(double) * RED_PARTIALS) != hipSuccess ||
            h->gm_small.alloc(sizeof(double) * 256) != hipSuccess)
            return fail(NSFV1_ERR_DEVICE, "GMRES workspace");
        h->gm_restart = m;
    }
    double* V = h->gm_V.as<double>();
    double* w = h->gm_w.as<double>();
    double* z = w + n;
    double* Binv = h->gm_Binv.as<double>();
    double* part = h->gm_part.as<double>();
    double* dsm = h->gm_small.as<double>();
    const int64_t* rp = h->rowptr.as<int64_t>();
    const int32_t* ci = h->colidx.as<int32_t>();
    const int32_t* ad = h->adj.as<int32_t>();
    auto chk = [&](hipError_t e) { return e == hipSuccess; };
    auto dot = [&](const double* Vb, int k, const double* vec, double* out) -> bool {
        if (!chk(launch_multi_dot(Vb, n, k, vec, n, part, dsm, s))) return false;
        return hipMemcpyAsync(out, dsm, sizeof(double) * k, hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess;
    };
    if (h->use_mg && (!h->mg_ready || h->mg_jprec != jprec))
        return fail(NSFV1_ERR_ARGUMENT, "multigrid preconditioner selected but not set up for this Jacobian precision (call nsfv1_mg_setup)");
    if (!chk(launch_block_jacobi_setup(jprec, nc, rp, ci, jvals, nv, Binv, h->errflag.as<int>(), s)))
        return fail(NSFV1_ERR_DEVICE, "block Jacobi setup");
    auto precond = [&](const double* in, double* out) -> hipError_t {
        if (h->use_mg) return mg_cycle(h, 0, jvals, Binv, in, out, s);
        return launch_block_jacobi_apply(nc, Binv, in, out, nv, s);
    };
    if (hipMemsetAsync(x, 0, sizeof(double) * n, s) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "memset");
    double bn2;
    if (!dot(b, 1, b, &bn2)) return fail(NSFV1_ERR_DEVICE, "dot");
    const double bnorm = std::sqrt(bn2);
    int it = 0;
    double res = bnorm;
    if (bnorm == 0.0) {
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return NSFV1_OK;
    }
    std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    while (it < maxit) {
        // r = b - J x  -> V0
        if (!chk(launch_spmv(jprec, rp, ci, ad, h->dim + 1, jvals, x, w, n, h->nnz, s)) || !chk(launch_axpby(1.0, b, -1.0, w, n, s)))
            return fail(NSFV1_ERR_DEVICE, "residual");
        double rn2;
        if (!dot(w, 1, w, &rn2)) return fail(NSFV1_ERR_DEVICE, "dot");
        const double beta = std::sqrt(rn2);
        res = beta;
        if (beta <= rtol * bnorm) break;
        if (!chk(launch_axpby(1.0 / beta, w, 0.0, V, n, s))) return fail(NSFV1_ERR_DEVICE, "scale");
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && it < maxit; ++j, ++it) {
            if (!chk(precond(V + (int64_t)j * n, z)) || !chk(launch_spmv(jprec, rp, ci, ad, h->dim + 1, jvals, z, w, n, h->nnz, s)))
                return fail(NSFV1_ERR_DEVICE, "arnoldi");
            // classical Gram-Schmidt, twice; the projections stay on the
            // device, and one download brings both and ||w||^2 to the host
            constexpr int HS = GMRES_MAX_RESTART + 1;
            static_assert(128 + 2 * HS + 1 <= 256, "GMRES device scratch holds 256 doubles");
            double* hd = dsm + 128;  // [h1 (HS) | h2 (HS) | ||w||^2]
            for (int pass = 0; pass < 2; ++pass)
                if (!chk(launch_multi_dot(V, n, j + 1, w, n, part, hd + pass * HS, s)) ||
                    !chk(launch_multi_axpy(V, n, j + 1, hd + pass * HS, w, n, s)))
                    return fail(NSFV1_ERR_DEVICE, "orthogonalise");
            double hh[2 * HS + 1];
            if (!chk(launch_multi_dot(w, n, 1, w, n, part, hd + 2 * HS, s)) ||
                hipMemcpyAsync(hh, hd, sizeof(hh), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return fail(NSFV1_ERR_DEVICE, "dot");
            for (int i = 0; i <= j; ++i) H[i * m + j] = hh[i] + hh[HS + i];
            const double hn = std::sqrt(hh[2 * HS]);
            H[(j + 1) * m + j] = hn;
            if (hn > 0 && !chk(launch_axpby(1.0 / hn, w, 0.0, V + (int64_t)(j + 1) * n, n, s)))
                return fail(NSFV1_ERR_DEVICE, "scale");
            // Givens rotations
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
                H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
                H[i * m + j] = t;
            }
            const double a = H[j * m + j], c = H[(j + 1) * m + j];
            const double r = std::hypot(a, c);
            cs[j] = r > 0 ? a / r : 1.0;
            sn[j] = r > 0 ? c / r : 0.0;
            H[j * m + j] = r;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            res = std::fabs(g[j + 1]);
            if (res <= rtol * bnorm || hn == 0.0) { ++j; ++it; break; }
        }
        // x += M^-1 V y
        for (int i = j - 1; i >= 0; --i) {
            double t = g[i];
            for (int k = i + 1; k < j; ++k) t -= H[i * m + k] * y[k];
            y[i] = H[i * m + i] != 0.0 ? t / H[i * m + i] : 0.0;
        }
        if (hipMemcpyAsync(dsm + 128, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemsetAsync(w, 0, sizeof(double) * n, s) != hipSuccess || !chk(launch_multi_add(V, n, j, dsm + 128, w, n, s)) ||
            !chk(precond(w, z)) || !chk(launch_axpby(1.0, z, 1.0, x, n, s)))
            return fail(NSFV1_ERR_DEVICE, "update");
        if (res <= rtol * bnorm) break;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "sync");
    if (iters) *iters = it;
    if (relres) *relres = res / bnorm;
    return NSFV1_OK;
}


// ---------------------------------------------------------------------
// capturable Newton step, its HIP graph, and the launch-shape autotuner
// ---------------------------------------------------------------------
static constexpr int NT_HS = GMRES_MAX_RESTART + 1;
static constexpr size_t NT_DEV_DOUBLES = 2 * NT_HS + 1 + 2 + GMRES_MAX_RESTART + (size_t)NT_HS * GMRES_MAX_RESTART;

static int newton_check(nsfv1_handle* h, double* u, const double* g, void* jvals, int jprec, double* defect, int m)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (!u || !jvals || !defect) return fail(NSFV1_ERR_ARGUMENT, "u, J and d are required");
    if (jprec != NSFV1_F32 && jprec != NSFV1_F64) return fail(NSFV1_ERR_ARGUMENT, "jprec must be 32 or 64");
    if (m < 1 || m > GMRES_MAX_RESTART) return fail(NSFV1_ERR_ARGUMENT, "krylov must be 1..60");
    if (h->ndir && !g) return fail(NSFV1_ERR_ARGUMENT, "Dirichlet values g are required (nsfv1_set_dirichlet rows)");
    if (h->use_mg && (!h->mg_ready || h->mg_jprec != jprec))
        return fail(NSFV1_ERR_ARGUMENT, "multigrid preconditioner selected but not set up for this Jacobian precision (call nsfv1_mg_setup)");
    return NSFV1_OK;
}

// allocations of the step (workspace; plan, records and geometry through one
// assembly, which only writes J and d); never called while a graph captures
static int newton_prepare(nsfv1_handle* h, const double* u, const double* u_old, double dt, double s_m, void* jvals,
                          int jprec, double* defect, int m, hipStream_t s)
{
    const int64_t n = h->ndof, nv = h->nv;
    const int nc = h->dim + 1;
    if (h->nt_m < m) {
        if (h->nt_V.alloc(sizeof(double) * n * (m + 1)) != hipSuccess || h->nt_w.alloc(sizeof(double) * n * 2) != hipSuccess ||
            h->nt_Binv.alloc(sizeof(double) * nv * nc * nc) != hipSuccess ||
            h->nt_part.alloc(sizeof(double) * RED_PARTIALS) != hipSuccess ||
            h->nt_dev.alloc(sizeof(double) * NT_DEV_DOUBLES) != hipSuccess)
            return fail(NSFV1_ERR_DEVICE, "Newton step workspace");
        h->nt_m = m;
    }
    const int parts = NSFV1_PART_A | (s_m != 0.0 ? NSFV1_PART_M : 0);
    return assemble_impl(h, u, u_old, dt, 1.0, s_m, 0.0, parts, jvals, jprec, defect, s, nullptr);
}

// the step: kernels only, on stream s, in a fixed sequence
static int newton_body(nsfv1_handle* h, double* u, const double* u_old, double dt, double s_m, const double* dconst,
                       const double* g, void* jvals, int jprec, double* d, int m, double* res2, hipStream_t s)
{
    const int parts = NSFV1_PART_A | (s_m != 0.0 ? NSFV1_PART_M : 0);
    int rc = assemble_impl(h, u, u_old, dt, 1.0, s_m, 0.0, parts, jvals, jprec, d, s, nullptr);
    if (rc) return rc;
    const int64_t n = h->ndof, nv = h->nv;
    const int nc = h->dim + 1;
    const int64_t* rp = h->rowptr.as<int64_t>();
    const int32_t* ci = h->colidx.as<int32_t>();
    const int32_t* ad = h->adj.as<int32_t>();
    double* V = h->nt_V.as<double>();
    double* w = h->nt_w.as<double>();
    double* z = w + n;
    double* Binv = h->nt_Binv.as<double>();
    double* part = h->nt_part.as<double>();
    double* hd = h->nt_dev.as<double>();  // [h1 (HS) | h2 (HS) | ||w||^2]
    double* n2 = hd + 2 * NT_HS + 1;
    double* beta = n2 + 1;
    double* y = beta + 1;
    double* H = y + GMRES_MAX_RESTART;  // (m+1) x m, row-major
    auto chk = [&](hipError_t e) { return e == hipSuccess; };
    auto precond = [&](const double* in, double* out) -> hipError_t {
        if (h->use_mg) return mg_cycle(h, 0, jvals, Binv, in, out, s);
        return launch_block_jacobi_apply(nc, Binv, in, out, nv, s);
    };
    bool ok = true;
    if (dconst) ok = ok && chk(launch_axpby(1.0, dconst, 1.0, d, n, s));
    if (h->ndir)
        ok = ok && chk(launch_dirichlet(jprec, rp, ci, h->dir_rows.as<int64_t>(), h->ndir, jvals, d, u, g, s));
    ok = ok && chk(launch_block_jacobi_setup(jprec, nc, rp, ci, jvals, nv, Binv, h->errflag.as<int>(), s));
    // GMRES(m) from x = 0: r = d, beta = ||d||, V0 = d / beta
    ok = ok && chk(launch_multi_dot(d, n, 1, d, n, part, n2, s)) && chk(launch_gm_beta(n2, beta, res2, s)) &&
         chk(launch_gm_scale(d, beta, V, n, s));
    for (int j = 0; ok && j < m; ++j) {
        ok = chk(precond(V + (int64_t)j * n, z)) && chk(launch_spmv(jprec, rp, ci, ad, nc, jvals, z, w, n, h->nnz, s));
        for (int pass = 0; ok && pass < 2; ++pass)
            ok = chk(launch_multi_dot(V, n, j + 1, w, n, part, hd + pass * NT_HS, s)) &&
                 chk(launch_multi_axpy(V, n, j + 1, hd + pass * NT_HS, w, n, s));
        ok = ok && chk(launch_multi_dot(w, n, 1, w, n, part, hd + 2 * NT_HS, s)) &&
             chk(launch_gm_hcol(hd, NT_HS, H, m, j, s)) &&
             chk(launch_gm_scale(w, H + (int64_t)(j + 1) * m + j, V + (int64_t)(j + 1) * n, n, s));
    }
    // u -= M^-1 V y
    ok = ok && chk(launch_gm_lsq(H, beta, m, y, s)) && chk(launch_fill(w, 0.0, n, s)) &&
         chk(launch_multi_add(V, n, m, y, w, n, s)) && chk(precond(w, z)) && chk(launch_axpby(-1.0, z, 1.0, u, n, s));
    if (!ok) return fail(NSFV1_ERR_DEVICE, std::string("Newton step launch: ") + hipGetErrorString(hipGetLastError()));
    return NSFV1_OK;
}

int nsfv1_newton_step(nsfv1_handle* h, double* u, const double* u_old, double dt, double s_m, const double* dconst,
                      const double* g, void* jvals, int jprec, double* defect, int krylov, double* res2, void* stream)
{
    int rc = newton_check(h, u, g, jvals, jprec, defect, krylov);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    rc = newton_prepare(h, u, u_old, dt, s_m, jvals, jprec, defect, krylov, s);
    if (rc) return rc;
    return newton_body(h, u, u_old, dt, s_m, dconst, g, jvals, jprec, defect, krylov, res2, s);
}

int nsfv1_newton_graph_capture(nsfv1_handle* h, double* u, const double* u_old, double dt, double s_m,
                               const double* dconst, const double* g, void* jvals, int jprec, double* defect,
                               int krylov, double* res2, void* stream)
{
    int rc = newton_check(h, u, g, jvals, jprec, defect, krylov);
    if (rc) return rc;
    if (!stream) return fail(NSFV1_ERR_ARGUMENT, "graph capture needs a stream other than the null stream");
    hipStream_t s = (hipStream_t)stream;
    rc = newton_prepare(h, u, u_old, dt, s_m, jvals, jprec, defect, krylov, s);
    if (rc) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return fail(NSFV1_ERR_DEVICE, "sync before capture");
    if (h->nt_exec) (void)hipGraphExecDestroy(h->nt_exec);
    if (h->nt_graph) (void)hipGraphDestroy(h->nt_graph);
    h->nt_exec = nullptr;
    h->nt_graph = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "hipStreamBeginCapture");
    rc = newton_body(h, u, u_old, dt, s_m, dconst, g, jvals, jprec, defect, krylov, res2, s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess || !graph) return fail(NSFV1_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(NSFV1_ERR_DEVICE, "hipGraphInstantiate");
    }
    h->nt_graph = graph;
    h->nt_exec = exec;
    return NSFV1_OK;
}

int nsfv1_newton_graph_replay(nsfv1_handle* h, void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (!h->nt_exec) return fail(NSFV1_ERR_ARGUMENT, "no captured Newton step (nsfv1_newton_graph_capture)");
    const hipError_t e = hipGraphLaunch(h->nt_exec, (hipStream_t)stream);
    if (e != hipSuccess) return fail(NSFV1_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return NSFV1_OK;
}

int nsfv1_autotune(nsfv1_handle* h, const double* u, const double* u_old, double dt, double s_m, int parts,
                   void* jvals, int jprec, double* defect, int reps, double* table, int* ntable, int* best,
                   void* stream)
{
    if (!h) return fail(NSFV1_ERR_ARGUMENT, "NULL handle");
    if (h->et < 0) return fail(NSFV1_ERR_NO_MESH, "no mesh set");
    if (reps < 1) return fail(NSFV1_ERR_ARGUMENT, "reps must be positive");
    hipStream_t s = (hipStream_t)stream;
    struct Cand { int path, threads, lds; };
    std::vector<Cand> cand;
    for (int p : {0, 1})
        for (int t : {64, 128, 256})
            for (int l : {16384, 32768, 65536}) cand.push_back({p, t, l});
    cand.push_back({2, 64, 16384});
    cand.push_back({2, 64, 65536});
    if (h->march_ok)
        for (int l : {8192, 16384, 32768}) cand.push_back({3, 256, l});
    const int keep_path = h->path, keep_threads = h->max_threads;
    const size_t keep_lds = h->lds_cap;
    const int keep_zc = h->march_zc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(NSFV1_ERR_DEVICE, "hipEventCreate");
    double best_ms = 1e300;
    Cand bc{-1, 0, 0};
    int nt = 0;
    const int cap = ntable ? *ntable : 0;
    for (const Cand& c : cand) {
        if (nsfv1_set_assembly_path(h, c.path) || nsfv1_set_launch(h, c.threads, c.lds)) continue;
        // a shape the mesh cannot use is refused by the library: skipped
        if (nsfv1_assemble(h, u, u_old, dt, 1.0, s_m, 0.0, parts, jvals, jprec, defect, stream)) continue;
        if (hipEventRecord(e0, s) != hipSuccess) break;
        int rc = 0;
        for (int r = 0; r < reps && !rc; ++r)
            rc = nsfv1_assemble(h, u, u_old, dt, 1.0, s_m, 0.0, parts, jvals, jprec, defect, stream);
        float ms = 0.f;
        if (rc || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            continue;
        const double avg = (double)ms / reps;
        if (table && nt < cap) {
            table[4 * nt] = c.path;
            table[4 * nt + 1] = c.threads;
            table[4 * nt + 2] = c.lds;
            table[4 * nt + 3] = avg;
        }
        ++nt;
        if (avg < best_ms) best_ms = avg, bc = c;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ntable) *ntable = nt;
    if (bc.path < 0) {
        h->path = keep_path;
        h->max_threads = keep_threads;
        h->lds_cap = keep_lds;
        h->march_zc = keep_zc;
        invalidate_plans(h);
        return fail(NSFV1_ERR_ARGUMENT, "no launch shape could assemble this mesh");
    }
    int rc = nsfv1_set_assembly_path(h, bc.path);
    if (!rc) rc = nsfv1_set_launch(h, bc.threads, bc.lds);
    if (rc) return rc;
    if (best) best[0] = bc.path, best[1] = bc.threads, best[2] = bc.lds;
    return NSFV1_OK;
}

}  // extern "C"
