/*
 * nsfv1.h -- C ABI of the MI355X-native FV1 incompressible Navier-Stokes
 * assembly (defect + Jacobian), the drop-in for the element loop of UG4's
 * NavierStokesFV1.
 *
 * The reference boundary is a per-element callback set registered per
 * reference-object id (incompressible/fv1/navier_stokes_fv1.cpp:1553-1579:
 * prep_elem_loop, prep_elem, add_jac_A_elem, add_jac_M_elem, add_def_A_elem,
 * add_def_M_elem, add_rhs_elem) driven by ugcore's element loop.  A
 * per-element callback cannot feed a GPU, so this boundary is the whole-mesh
 * assembly that the ugcore loop performs with those callbacks: one call
 * assembles every element's contributions into a global CSR Jacobian and
 * defect.  The configuration calls mirror the Lua/registry surface
 * (incompressible/fv1/register_fv1.cpp:167-184, register_navier_stokes.cpp:
 * 104-126, incompressible/incompressible_navier_stokes_plugin.cpp:244-266)
 * with the same string keys, defaults and error cases.
 *
 * Conventions
 *   - DoF of (node v, function c) = v*(dim+1) + c; c = 0..dim-1 velocity,
 *     c = dim pressure (the _P_ = dim of navier_stokes_fv1.h:614).
 *   - CSR pattern = node adjacency closure (x) (dim+1)^2, columns sorted.
 *   - Array arguments of set_mesh are HOST pointers (copied); solution,
 *     Jacobian and defect arguments of assemble are DEVICE pointers and the
 *     call is asynchronous on the given hipStream_t (NULL = default stream).
 *   - Every call returns NSFV1_OK (0) or a negative NSFV1_ERR_* code; the
 *     message is available from nsfv1_last_error().
 *   - Handles are independent; one handle must not be used from two host
 *     threads at once.
 */
#ifndef NSFV1_H
#define NSFV1_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nsfv1_handle nsfv1_handle;

/* element types (ugcore ReferenceObjectID subset registered by
   navier_stokes_fv1.cpp:1515-1551 for the 2D and 3D domains).  Prisms
   (UG4 numbering: corners 0-1-2 at the bottom, 3-4-5 above them; edges
   01 12 20 03 14 25 34 45 53; sides 021, 0143, 1254, 2035, 345) assemble on
   the generic path (and the coupled path for the FLOW stabilization and
   the positive upwind), with boundary discs and data exports.  Pyramids are
   refused by nsfv1_set_mesh with NSFV1_ERR_NOT_IMPLEMENTED. */
enum {
    NSFV1_TRIANGLE = 0,
    NSFV1_QUADRILATERAL = 1,
    NSFV1_TETRAHEDRON = 2,
    NSFV1_HEXAHEDRON = 3,
    NSFV1_PRISM = 4,
    NSFV1_PYRAMID = 5
};

/* assembled parts (bit mask) */
enum { NSFV1_PART_A = 1, NSFV1_PART_M = 2, NSFV1_PART_RHS = 4 };

/* Jacobian value precision */
enum { NSFV1_F32 = 32, NSFV1_F64 = 64 };

/* status codes; each mirrors a UG_THROW of the reference path */
#define NSFV1_OK                   0
#define NSFV1_ERR_ARGUMENT        -1  /* invalid argument / NULL handle */
#define NSFV1_ERR_FUNCTION_COUNT  -2  /* navier_stokes_fv1.cpp:66-68 */
#define NSFV1_ERR_TRIAL_SPACE     -3  /* navier_stokes_fv1.cpp:96-103 */
#define NSFV1_ERR_NO_STABILIZATION -4 /* navier_stokes_fv1.cpp:147-148 */
#define NSFV1_ERR_NO_UPWIND       -5  /* navier_stokes_fv1.cpp:156-157 */
#define NSFV1_ERR_NO_VISCOSITY    -6  /* navier_stokes_fv1.cpp:169-171 */
#define NSFV1_ERR_HANGING_NODES   -7  /* navier_stokes_fv1.cpp:1528, :92-93 */
#define NSFV1_ERR_TIME_SERIES     -8  /* navier_stokes_fv1.cpp:272-274, 621-623 */
#define NSFV1_ERR_CUT_SIDE        -9  /* upwind.cpp:354, 422, 550 */
#define NSFV1_ERR_UNKNOWN_NAME   -10  /* upwind_interface.cpp:60-61, stabilization.cpp:55-56, 97-99 */
#define NSFV1_ERR_NOT_IMPLEMENTED -11 /* a reference option this build does not provide */
#define NSFV1_ERR_NO_MESH        -12  /* assemble before set_mesh */
#define NSFV1_ERR_DEVICE         -13  /* HIP runtime error */
#define NSFV1_ERR_DEGENERATE     -14  /* zero Jacobian determinant in an element */
#define NSFV1_ERR_SINGULAR_BLOCK -15  /* singular diagonal block in the block-Jacobi setup */

/* Thread-local message of the last failing call. */
const char* nsfv1_last_error(void);

/* NavierStokesFV1(functions, subsets): replaces the constructors at
   incompressible/fv1/navier_stokes_fv1.cpp:45-59 and init() :62-87.
   `functions` is a comma separated list; it must name exactly dim+1
   functions.  Defaults: density 1.0, exact-Jacobian factor 0, no Stokes,
   Laplace or Peclet blend, no stabilisation, no upwind, no viscosity. */
int nsfv1_create(int dim, const char* functions, const char* subsets, nsfv1_handle** out);
int nsfv1_destroy(nsfv1_handle* h);

/* prepare_setting: navier_stokes_fv1.cpp:89-107.  lfe_type 0 = Lagrange;
   each of the dim+1 spaces must be Lagrange of order 1. */
int nsfv1_prepare_setting(nsfv1_handle* h, const int* lfe_type, const int* order, int n, int non_regular_grid);

/* disc_type(): navier_stokes_fv1.h:241 -- writes "fv1". */
int nsfv1_disc_type(nsfv1_handle* h, char* buf, size_t n);

/* NavierStokesBase setters: navier_stokes_base.h:170-196,
   navier_stokes_base.cpp:72-77, constant user data.  A constant replaces
   an import set below (and the other way round), as a later set_* call
   replaces the user data in the reference. */
int nsfv1_set_kinematic_viscosity(nsfv1_handle* h, double nu);
int nsfv1_set_source(nsfv1_handle* h, const double* f, int n);
int nsfv1_set_exact_jacobian(nsfv1_handle* h, double factor);

/* Function-valued kinematic viscosity and source: the
   set_kinematic_viscosity(CplUserData) / set_source(CplUserData) overloads
   (register_navier_stokes.cpp:109-118, navier_stokes_base.h:170-189).  The
   reference imports them at the integration points
   (navier_stokes_fv1.cpp:188-247: viscosity and source at the SCVF ips,
   source at the SCV ips); here the caller evaluates its function at the
   positions nsfv1_ip_positions returns and hands the values over (host
   arrays, FP64, uploaded once):
     nu    [ne][nip]       at the SCVF ips (ip = the element's edge number);
     f_scv [ne][nsh][dim]  at the SCV ips (the corners): the rhs part;
     f_scvf[ne][nip][dim]  at the SCVF ips: the stabilisation's source.
   NULL clears (the source needs both arrays).  A new mesh clears them.
   These run on the generic path (nsfv1_set_assembly_path 1 behaviour), or
   on the coupled path with the FLOW stabilization and the positive or
   regular upwind; with boundary discs they are refused with
   NSFV1_ERR_NOT_IMPLEMENTED.  With multigrid
   every coarse handle needs its own imports of the same quantities (the
   function evaluated at its ips); nsfv1_mg_setup refuses a hierarchy that
   lacks them with NSFV1_ERR_ARGUMENT and drops coarse imports the fine
   level no longer has. */
int nsfv1_set_kinematic_viscosity_ip(nsfv1_handle* h, const double* nu, int64_t n);
int nsfv1_set_source_ip(nsfv1_handle* h, const double* f_scv, int64_t n_scv, const double* f_scvf, int64_t n_scvf);
/* Function-valued density: replaces set_density(SmartPtr<CplUserData>)
   (navier_stokes_fv1.cpp:118-121), whose two imports the reference
   evaluates at the SCVF ips (m_imDensitySCVF: fluxes, stabilisation,
   navier_stokes_fv1.cpp:336-589, stabilization.cpp:229, 321) and at the SCV
   ips (m_imDensitySCV: mass part and source, navier_stokes_fv1.cpp:805,
   835, 866):
     rho_scvf [ne][nip]  at the SCVF ips (ip = the element's edge number);
     rho_scv  [ne][nsh]  at the SCV ips (the corners).
   Both or neither (NULL, NULL clears; so does nsfv1_set_density and a new
   mesh).  An SCVF value of zero or NaN is refused with NSFV1_ERR_ARGUMENT.
   It runs where the other imports run (generic and coupled paths, mixed
   meshes, multigrid with the coarse levels' own values, the stabilised
   velocity export); boundary discs keep a constant density, so an assembly
   with boundary discs and this import returns NSFV1_ERR_NOT_IMPLEMENTED. */
int nsfv1_set_density_ip(nsfv1_handle* h, const double* rho_scvf, int64_t n_scvf, const double* rho_scv, int64_t n_scv);
/* Global positions of the integration points (host array): kind 0 = SCVF
   ips [ne][nip][dim] (FV1Geometry: the mean of edge midpoint, element
   centre and, in 3D, the centres of the edge's two faces), kind 1 = SCV ips
   [ne][nsh][dim] (the corners). */
int nsfv1_ip_positions(nsfv1_handle* h, int kind, double* out, int64_t n);

/* IncompressibleNavierStokesBase setters:
   incompressible_navier_stokes_base.h:170-227 */
int nsfv1_set_density(nsfv1_handle* h, double rho);
int nsfv1_set_stokes(nsfv1_handle* h, int on);
int nsfv1_set_laplace(nsfv1_handle* h, int on);
int nsfv1_set_peclet_blend(nsfv1_handle* h, int on);
int nsfv1_set_grad_div(nsfv1_handle* h, double factor);

/* NavierStokesFV1::set_upwind(name): navier_stokes_fv1.h:213-215 with the
   factory upwind_interface.cpp:43-62 (keys trimmed, case-insensitive: no,
   full, skewed, linearprofileskewed|lps, positive|pos, regular|reg).  The
   positive and regular upwinds have ip shapes: the assembly then solves the
   stabilisation's numIp x numIp system per element.  The regular upwind is
   2D only (upwind.cpp:806: assembly of a 3D mesh returns
   NSFV1_ERR_NOT_IMPLEMENTED); its ray/SCV intersection, ugcore's
   SCVFofSCVRayIntersection, follows a stated convention: the SCV of corner
   k is [x_k, m_k, c, m_{k-1}], the ray leaves it through the other SCVF,
   side [x_k, m_k] or side [m_{k-1}, x_k] (in that order), with lambda from
   the SCVF's edge midpoint to the centre on an SCVF. */
int nsfv1_set_upwind(nsfv1_handle* h, const char* name);

/* set_stabilization(name[, diffLength]): navier_stokes_fv1.h:189-200 with
   the factories stabilization.cpp:46-57 (fields, flow) and :86-100 (raw,
   fivepoint, cor).  diff_length may be NULL (keeps RAW, stabilization.h:321-325).
   FLOW (stabilization.cpp:436-772) couples the velocity components and uses
   the downwind of its upwind as well. */
int nsfv1_set_stabilization(nsfv1_handle* h, const char* name, const char* diff_length);

/* set_stabilization(NavierStokesFV1WithoutStabilization): the object form
   of register_fv1.cpp:172 with the class of stabilization.h:585-647. */
int nsfv1_set_stabilization_without(nsfv1_handle* h);

/* INavierStokesFV1Stabilization::set_upwind(name) on the current
   stabilisation object: stabilization.h:86-90 (object-sharing semantics of
   navier_stokes_fv1.h:190-215 otherwise apply). */
int nsfv1_set_stabilization_upwind(nsfv1_handle* h, const char* name);

/* set_pac_upwind: navier_stokes_fv1.h:217-226 */
int nsfv1_set_pac_upwind(nsfv1_handle* h, int on);

/* Mesh: node coordinates [nv][dim] and connectivity [ne][corners] with UG4
   corner order; one element type per handle.  Host pointers.  Builds the
   device copies, the CSR pattern and the node->element incidence used by the
   row-owner gather. */
int nsfv1_set_mesh(nsfv1_handle* h, int elem_type, const double* xyz, int64_t nv,
                   const int32_t* conn, int64_t ne);

/* Mixed-element mesh: ngroups (1..8) groups of elements of one type each
   (elem_type[g], conn[g] = [ne[g]][corners], UG4 corner order) over the
   same nodes; elements are numbered group after group.  Replaces ugcore's
   element loop over a subset that holds several element types, which calls
   the disc once per type (prep_elem_loop, navier_stokes_fv1.cpp:136-182,
   for each type registered at :1515-1551).  One CSR pattern covers all
   groups; the assembly runs the generic kernel (the coupled path for the
   FLOW stabilization and the positive/regular upwinds) once per group on
   the stream, the first group writing every row and the others adding
   theirs.  Integration-point imports and nsfv1_ip_positions concatenate
   the groups' blocks in group order ([ne_g][nip_g] ..., each with dim
   where it applies).  Boundary discs take (element, side) pairs with the
   global element numbers.  Data exports at SCVF ips and corners write the
   groups' [ne_g][points_g][components] blocks in order; exports at local
   points and batched instances refuse mixed meshes with
   NSFV1_ERR_NOT_IMPLEMENTED. */
int nsfv1_set_mesh_mixed(nsfv1_handle* h, const double* xyz, int64_t nv, int ngroups, const int* elem_type,
                         const int32_t* const* conn, const int64_t* ne);

int nsfv1_num_dofs(nsfv1_handle* h, int64_t* ndof);
int nsfv1_nnz(nsfv1_handle* h, int64_t* nnz);

/* Copies the CSR pattern to HOST arrays rowptr[ndof+1], colidx[nnz]. */
int nsfv1_get_pattern(nsfv1_handle* h, int64_t* rowptr, int32_t* colidx);

/* Device pointers of the pattern owned by the handle. */
int nsfv1_pattern_device(nsfv1_handle* h, const int64_t** rowptr, const int32_t** colidx);

/* Whole-mesh assembly; replaces ugcore's element loop over
   add_jac_A_elem/add_def_A_elem (navier_stokes_fv1.cpp:250-778),
   add_jac_M_elem/add_def_M_elem (:781-838) and add_rhs_elem (:841-869):
     J = s_a*J_A + s_m*J_M,   d = s_a*d_A + s_m*d_M + s_f*f
   over the parts selected in `parts`.  u, u_old: device [nv][dim+1] FP64;
   u_old != NULL selects the time-dependent stabilisation with time step dt
   (the two-point local time series of navier_stokes_fv1.cpp:268-280).
   jvals: device [nnz] of jprec (NSFV1_F32 or NSFV1_F64), or NULL;
   defect: device [ndof] FP64, or NULL.  Both are overwritten.
   stream: hipStream_t or NULL. */
int nsfv1_assemble(nsfv1_handle* h, const double* u, const double* u_old, double dt,
                   double s_a, double s_m, double s_f, int parts,
                   void* jvals, int jprec, double* defect, void* stream);

/* nsfv1_assemble that allocates no device memory.  No counterpart in the
   reference: it is for callers that cannot call hipMalloc while they
   assemble (code under HIP graph capture, tight serving loops, harnesses
   that forbid allocation).  Same arguments, dispatch and results (bitwise)
   as nsfv1_assemble; the SCVF records that the split and coupled paths
   keep between their two kernels live in a fixed device array of
   NSFV1_NOALLOC_SCRATCH_BYTES (one per device, held until the process
   ends) instead of a buffer the handle allocates.  Returns
   NSFV1_ERR_ARGUMENT, before any launch, when the mesh's records exceed
   that array (split path: 128 bytes per SCVF, so about 87 000 hexahedra,
   174 000 tetrahedra, 262 000 quadrilaterals or 349 000 triangles), or
   when the handle has no launch plan for jprec or, in NSFV1_UPWIND_RECORD
   mode, no decision buffer yet (one nsfv1_assemble call with the same
   settings builds them).  The structured-hexahedra plane march and the
   fused path (FP64 Jacobians, the generic configuration) keep no records
   and have no size limit.
   Nothing in the array has to start at zero.  All calls on one device share
   its array: calls that overlap on one device (other streams, other host
   threads) race. */
#define NSFV1_NOALLOC_SCRATCH_BYTES (128ull << 20)
int nsfv1_assemble_noalloc(nsfv1_handle* h, const double* u, const double* u_old, double dt,
                           double s_a, double s_m, double s_f, int parts,
                           void* jvals, int jprec, double* defect, void* stream);

/* Batched instances: the handle's mesh is the disjoint union of ninst
   independent problems (e.g. cavities of different sizes); instance i owns
   elements [elem_offset[i], elem_offset[i+1]) and has kinematic viscosity
   nu[i], which replaces set_kinematic_viscosity for its elements.  One
   nsfv1_assemble call then assembles every instance in one launch.  Host
   arrays (copied); ninst = 0 returns to the single viscosity.  Per-instance
   Dirichlet values go through nsfv1_set_dirichlet / nsfv1_apply_dirichlet
   on the union.  No counterpart in the plugin (one problem per element
   disc). */
int nsfv1_batch_set_instances(nsfv1_handle* h, int64_t ninst, const int64_t* elem_offset, const double* nu);

/* Skewed-upwind decisions (the corner nearest the ray's cut point,
   upwind.cpp:381-430, or the zero-velocity guard) of later assemblies:
   NSFV1_UPWIND_FRESH (default) decides anew on every assembly;
   NSFV1_UPWIND_RECORD decides anew and stores each SCVF's decision;
   NSFV1_UPWIND_REPLAY reuses the stored decisions (no ray cast).  No
   counterpart in the plugin: the corner choice is a discontinuous function
   of the velocity, so a Newton iteration that lands next to a switching
   point can cycle between two states; replaying the decisions of one
   iterate makes the defect smooth in u for the rest of a solve.  Other
   upwinds are not affected. */
enum { NSFV1_UPWIND_FRESH = 0, NSFV1_UPWIND_RECORD = 1, NSFV1_UPWIND_REPLAY = 2 };
int nsfv1_set_upwind_decisions(nsfv1_handle* h, int mode);

/* NavierStokesNoNormalStressOutflow (incompressible/bnd/no_normal_stress_
   outflow_base.h:70-89 with the FV1 assembly of incompressible/fv1/bnd/
   no_normal_stress_outflow_fv1.cpp:188-426): the element sides (element
   index, UG4 local side index; host arrays) on the outflow boundary.
   nsfv1_assemble then adds, per FV1 boundary face, the tangential viscous
   flux, the outgoing convective flux and the continuity flux (part A, scaled
   by s_a) to the rows of the side's nodes.  n = 0 removes the condition.
   The master's viscosity, density, Laplace and Stokes flags apply (the
   reference copies them from the master disc, :49-56). */
int nsfv1_set_outflow(nsfv1_handle* h, const int32_t* elem, const int32_t* side, int64_t n);

/* NavierStokesSymBCFV1 (incompressible/fv1/bnd/symmetric_boundary_fv1.cpp:
   224-386, Lua name "NavierStokesSymBCFV1", its add() of boundary subsets
   replaced by the side list as in nsfv1_set_outflow): per FV1 boundary face
   the normal viscous stress component-wise times n and the pressure times n
   go into the momentum rows of the face's node (part A, scaled by s_a);
   the continuity rows get nothing.  n = 0 removes the condition.  The
   master's viscosity, density and Laplace flag apply. */
int nsfv1_set_symmetry(nsfv1_handle* h, const int32_t* elem, const int32_t* side, int64_t n);

/* NavierStokesWSBCFV1, the sliding wall (incompressible/fv1/bnd/
   wall_sliding_fv1.cpp:229-425; Lua "NavierStokesWSBCFV1" with add(),
   set_sliding_factor(number), set_sliding_limit(number)): on the listed
   sides, per FV1 boundary face, the tangential convective term, -k times
   the tangential velocity and -limit times its direction go into the
   momentum rows (part A, scaled by s_a), with the reference's approximate
   Jacobian.  The velocity of the k and limit terms is interpolated from the
   element's first dim corners only, as the reference does (:391-397).
   The master's density and Stokes flag apply.  n = 0 removes the condition. */
int nsfv1_set_sliding_wall(nsfv1_handle* h, const int32_t* elem, const int32_t* side, int64_t n,
                           double sliding_factor, double sliding_limit);

/* A function-valued viscosity on the boundary discs: the reference's
   outflow and symmetry discs evaluate the master disc's viscosity at their
   boundary-face ips (no_normal_stress_outflow_fv1.cpp:188-426,
   symmetric_boundary_fv1.cpp:224-386).  nsfv1_bnd_ip_positions returns
   those ips of boundary list `kind` (0 outflow, 1 symmetry, 2 sliding wall)
   in input order, out[b][j][dim] for pair b and the side's j-th corner
   (4 slots per pair, unused ones 0); the caller evaluates its function
   there and hands the values to nsfv1_set_boundary_viscosity (nu[b][j],
   NULL clears).  With nsfv1_set_kinematic_viscosity_ip in force, the
   assembly refuses outflow and symmetry lists without them
   (NSFV1_ERR_ARGUMENT).  A new list clears them. */
int nsfv1_bnd_ip_positions(nsfv1_handle* h, int kind, double* out, int64_t n);
int nsfv1_set_boundary_viscosity(nsfv1_handle* h, int kind, const double* nu, int64_t n);

/* Data exports of the disc (navier_stokes_fv1.cpp:894-1491; the export
   objects of navier_stokes_fv1.cpp:1573-1577): values at points of every
   element, written to the device array out[ne][points][components] (FP64)
   on `stream`.
     what  NSFV1_EX_VELOCITY       ex_nodal_velocity   (dim components)
           NSFV1_EX_VELOCITY_GRAD  ex_velocity_grad    (dim*dim, row d1 = velocity component, column d2)
           NSFV1_EX_VELOCITY_IP    ex_velocity_ip      (dim; at SCVF ips the stabilised velocity, which
                                                        needs the stabilisation as the assembly does and,
                                                        time-dependent, u_old and dt > 0)
           NSFV1_EX_PRESSURE       ex_nodal_pressure   (1)
           NSFV1_EX_PRESSURE_GRAD  ex_pressure_grad    (dim)
     at    NSFV1_AT_SCVF   the element's FV1 SCVF ips (the reference's special case)
           NSFV1_AT_SCV    its SCV ips, i.e. its corners
           NSFV1_AT_LOCAL  the nip (1..32) local points local_ips[nip][dim] (host array), the
                           reference's general case; its velocity gradient there is taken in
                           reference-element coordinates, as ex_velocity_grad :1072-1080 does.
   The derivatives (bDeriv) are the shape values and gradients themselves
   and are not exported.  nsfv1_export_size gives the number of doubles. */
enum { NSFV1_EX_VELOCITY = 0, NSFV1_EX_VELOCITY_GRAD = 1, NSFV1_EX_VELOCITY_IP = 2, NSFV1_EX_PRESSURE = 3,
       NSFV1_EX_PRESSURE_GRAD = 4 };
enum { NSFV1_AT_SCVF = 0, NSFV1_AT_SCV = 1, NSFV1_AT_LOCAL = 2 };
int nsfv1_export_size(nsfv1_handle* h, int what, int at, int nip, int64_t* count);
int nsfv1_export(nsfv1_handle* h, int what, int at, const double* local_ips, int nip, const double* u,
                 const double* u_old, double dt, double* out, void* stream);

/* Synchronises `stream` and reports (then clears) errors raised inside the
   kernels since the last call: NSFV1_ERR_CUT_SIDE when a skewed/LPS upwind
   ray found no side (upwind.cpp:354, 422, 550), NSFV1_ERR_DEGENERATE for an
   element with a zero Jacobian determinant, NSFV1_ERR_SINGULAR_BLOCK for a
   singular diagonal block met by the block-Jacobi setup (linear solver and
   multigrid smoother).  Each kind has its own device flag word, so one kind
   cannot hide another; when several were raised, the first in this order is
   reported. */
int nsfv1_sync_errors(nsfv1_handle* h, void* stream);

/* ---- Newton-loop services (ugcore DirichletBoundary / linear solver, which
   the reference delegates to; not part of the plugin) ---- */

/* Dirichlet DoFs (HOST array of DoF indices).  Replaces the constraint rows
   that ugcore's DirichletBoundary sets for NavierStokesWall /
   NavierStokesInflow (incompressible/bnd/wall_impl.h:57-69,
   incompressible/bnd/inflow_base_impl.h:44-63). */
int nsfv1_set_dirichlet(nsfv1_handle* h, const int64_t* dofs, int64_t n);

/* Rows of the Dirichlet DoFs: J row -> identity, defect -> u - g, where g
   (device, one value per DoF in set_dirichlet order) may be NULL (g = 0).
   jvals or defect may be NULL.  Device pointers, asynchronous. */
int nsfv1_apply_dirichlet(nsfv1_handle* h, void* jvals, int jprec, double* defect, const double* u,
                          const double* g, void* stream);

/* y = J x with the assembled Jacobian (device pointers). */
int nsfv1_spmv(nsfv1_handle* h, const void* jvals, int jprec, const double* x, double* y, void* stream);

/* Restarted GMRES(restart), right-preconditioned with point-block Jacobi
   ((dim+1)x(dim+1) node blocks) or, when selected, one geometric multigrid
   V-cycle, on J x = b: device FP64 b and x (x is overwritten; zero initial
   guess), J in jprec.  Synchronises the stream.  iters/relres (host, may be
   NULL) receive the inner iterations and ||b - J x|| / ||b||.
   Replaces the linear solver ugcore's NewtonSolver calls on the assembled
   operator (not part of the plugin; the plugin's own preconditioner
   incompressible/cr_ilut.h is an ILUT for the Crouzeix-Raviart variant). */
int nsfv1_linear_solve(nsfv1_handle* h, const void* jvals, int jprec, const double* b, double* x, double rtol,
                       int maxit, int restart, int* iters, double* relres, void* stream);

/* One Newton step for d(u) = 0 that runs entirely on the device, with no
   host synchronisation: assemble J (jprec) and d (A part, plus the mass part
   with weight s_m when s_m != 0, time-dependent stabilisation when u_old is
   given), d += dconst (device, or NULL: e.g. the BDF history term), the
   Dirichlet rows of nsfv1_set_dirichlet (values g, device), then `krylov`
   iterations of right-preconditioned GMRES from zero with every reduction
   and the Hessenberg least-squares solve on the device, and u -= J^-1 d so
   obtained.  res2 (device double, or NULL) receives ||d||^2 of this step
   (before the update).  The preconditioner is nsfv1_linear_solve's (block
   Jacobi; the multigrid hierarchy as last set up).  Allocates its
   workspace on the first call.  No counterpart in the plugin: the Newton
   iteration that drives NavierStokesFV1 is ugcore's NewtonSolver, outside
   the reference tree; this is one of its steps with a fixed linear effort,
   for callers that want a step without host round trips or a graph. */
int nsfv1_newton_step(nsfv1_handle* h, double* u, const double* u_old, double dt, double s_m,
                      const double* dconst, const double* g, void* jvals, int jprec, double* defect,
                      int krylov, double* res2, void* stream);

/* Captures nsfv1_newton_step with these arguments into a HIP graph owned
   by the handle (replacing a previous one); stream must not be the null
   stream.  Allocations and the geometry happen eagerly first (one
   assembly, which writes J and d but not u), so the graph holds kernels
   and device-to-device copies only; every pointer must stay valid while
   the graph is replayed.  No counterpart in the plugin. */
int nsfv1_newton_graph_capture(nsfv1_handle* h, double* u, const double* u_old, double dt, double s_m,
                               const double* dconst, const double* g, void* jvals, int jprec, double* defect,
                               int krylov, double* res2, void* stream);

/* Launches the captured Newton step on stream.  No counterpart in the
   plugin. */
int nsfv1_newton_graph_replay(nsfv1_handle* h, void* stream);

/* Launch-shape autotuner: times `reps` assemblies (HIP events on stream)
   of every assembly path and launch shape the handle accepts (paths 0 and
   1 x 64/128/256 threads x 16/32/64 KiB, path 2 x 64 threads, path 3 on
   structured grids), keeps the fastest on the handle and returns it in
   best[3] = {path, threads, lds_bytes}; table (host, 4 doubles per
   candidate: path, threads, lds, ms), with capacity and count in *ntable,
   may be NULL.  Shapes the mesh cannot use are skipped.  No counterpart in
   the plugin (the CPU assembly has no launch shape). */
int nsfv1_autotune(nsfv1_handle* h, const double* u, const double* u_old, double dt, double s_m, int parts,
                   void* jvals, int jprec, double* defect, int reps, double* table, int* ntable, int* best,
                   void* stream);

/* Preconditioner of nsfv1_linear_solve: "block-jacobi" (default) or "gmg"
   (geometric multigrid, needs nsfv1_mg_set_hierarchy and nsfv1_mg_setup).
   No counterpart in the plugin: ugcore's GMG plays this role for it. */
int nsfv1_set_preconditioner(nsfv1_handle* h, const char* name);

/* Multigrid hierarchy: level 0 is h, levels 1..nlev-1 are the handles
   coarse[0..nlev-2] with meshes (and Dirichlet rows) set; they take h's
   configuration.  For level l -> l+1 (array index l): prol_rowptr/colidx/
   weights is a node-level CSR with rows = nodes of level l and columns =
   nodes of level l+1 (applied to every component; restriction is its
   transpose), inject[l][w] = node of level l that coarse node w sits on.
   The coarsest level must have at most 1200 DoFs (dense inverse).  pre/post
   damped (omega) point-block Jacobi sweeps per level.  Coarse handles must
   outlive h's use of them. */
int nsfv1_mg_set_hierarchy(nsfv1_handle* h, int nlev, nsfv1_handle* const* coarse, const int64_t* const* prol_rowptr,
                           const int32_t* const* prol_colidx, const double* const* prol_weights,
                           const int64_t* const* inject, int pre, int post, double omega);

/* Re-assembles every coarse level's Jacobian at the injected state (same
   time-series, scalings and parts as the fine Jacobian), applies the coarse
   Dirichlet rows, factors the point blocks and inverts the coarsest level on
   the device (Gauss-Jordan, four pivots per launch).  Call after each fine
   assembly whose Jacobian the next nsfv1_linear_solve uses.  Synchronises
   the stream once, to report a singular coarsest level. */
int nsfv1_mg_setup(nsfv1_handle* h, int jprec, const double* u, const double* u_old, double dt, double s_a,
                   double s_m, int parts, void* stream);

/* ---- Learned Newton initial guess (no counterpart in the plugin) ----
   A per-node MLP 24 -> 128 -> 128 -> 4 with ReLU, BF16 weights and
   activations, FP32 accumulation (MFMA).  params: FP32 blob of nparam
   (22800) floats, layout W1[128][32] (input columns 24..31 zero), b1[128],
   W2[128][128], b2[128], W3[16][128] (output rows 4..15 zero), b3[16];
   wbf16: the BF16 copy (same offsets) made by nsfv1_mlp_pack or
   nsfv1_mlp_adam.  All buffers are device pointers; x [n][32] and y [n][4]
   FP32, 16-byte aligned. */
int nsfv1_mlp_sizes(int64_t n, int64_t* nparam, int64_t* work_floats);
int nsfv1_mlp_pack(const float* params, uint16_t* wbf16, void* stream);
/* y = MLP(x); z1, z2 (nullable, [n][128]) receive the pre-activations for
   nsfv1_mlp_backward. */
int nsfv1_mlp_forward(const float* params, const uint16_t* wbf16, const float* x, int64_t n, float* y, float* z1,
                      float* z2, void* stream);
/* dy = d(mean squared error)/dy for targets t [n][4]; work holds at least
   ceil(4n/256) floats; loss (host, nullable) receives the MSE (synchronises). */
int nsfv1_mlp_mse(const float* y, const float* t, int64_t n, float* dy, float* work, double* loss, void* stream);
/* grad [nparam] = d loss / d params for the batch (deterministic order);
   work: nsfv1_mlp_sizes(n) floats. */
int nsfv1_mlp_backward(const uint16_t* wbf16, const float* x, const float* z1, const float* z2, const float* dy,
                       int64_t n, float* grad, float* work, void* stream);
/* One Adam step (bias-corrected) on params with moments m, v; refreshes wbf16. */
int nsfv1_mlp_adam(float* params, float* m, float* v, const float* grad, int step, double lr, double beta1,
                   double beta2, double eps, uint16_t* wbf16, void* stream);
/* Learned initial guess, 3D ([nv][4] states, FP64, device pointers).  With
   u1 = u^{n-1}, u2 = u^{n-2}, u3 = u^{n-3}, the backward differences
   d1 = u - u1, d2 = u - 2 u1 + u2, d3 = u - 3 u1 + 3 u2 - u3 (evaluated
   as nested differences of neighbouring states: exactly 0 for equal
   states) and the
   step's scale sigma = max(rms of d3 over all velocity components, 1e-30)
   (a deterministic device reduction, written to sigma[0]):
     x[v] = clamp((u, d1/sigma, d2/sigma, d3/sigma, ring(u) - u,
                   (ring(d1) - d1)/sigma) * scale[24], +-64), padded to 32,
   ring() the mean over the node's 1-ring (its CSR column blocks).  x is
   16-byte aligned, nv * 32 floats. */
int nsfv1_predict_features(nsfv1_handle* h, const double* u, const double* u_prev, const double* u_prev2,
                           const double* u_prev3, const float* scale, float* x, double* sigma, void* stream);
/* guess = 3 u - 3 u1 + u2 + y * out_scale[4] * sigma (quadratic
   extrapolation plus the learned correction). */
int nsfv1_predict_guess(nsfv1_handle* h, const double* u, const double* u_prev, const double* u_prev2, const float* y,
                        const float* out_scale, const double* sigma, double* guess, void* stream);
/* training target t = clamp((u^{n+1} - (3 u - 3 u1 + u2)) / (sigma * out_scale), +-64),
   0 where out_scale is 0. */
int nsfv1_predict_target(nsfv1_handle* h, const double* u_next, const double* u, const double* u_prev,
                         const double* u_prev2, const float* out_scale, const double* sigma, float* t, void* stream);

/* Assembly path: 0 evaluates every SCVF once into a per-SCVF record (FP32
   Jacobian, common configuration: FIELDS, one upwind object, no
   Stokes/PAC/Peclet/exact terms) and gathers rows from the records; 1
   always runs the fused kernel (each SCVF evaluated by both of its rows);
   2 is path 0 with dim+1 gather lanes per (node, element), one per row,
   when the launch has 64 threads (nsfv1_set_launch); 3 (default) marches
   through the element layers of a structured grid of parallelepipeds
   (a lexicographic hexahedral grid, detected by nsfv1_set_mesh) with the
   SCVF records held on chip -- FP32 Jacobian, no/full/skewed upwind, RAW
   diffusion length, one viscosity -- and otherwise behaves as 0.  All are
   deterministic.  The FLOW stabilisation and the positive upwind always
   use the coupled path.  No counterpart in the plugin. */
int nsfv1_set_assembly_path(nsfv1_handle* h, int path);

/* Kernel family that ran the last element assembly of the handle: 3 plane
   march, 0 split (records + row gather), 2 split with row lanes, 1 fused,
   4 coupled (FLOW / positive upwind); -1 before the first assembly.  No
   counterpart in the plugin. */
int nsfv1_assembly_path_used(nsfv1_handle* h, int* path);

/* Launch shape of the assembly: incidences (threads) per workgroup (64..256,
   a multiple of 64) and the LDS budget per workgroup in bytes (4..64 KiB);
   0 keeps the current value.  Set while path 3 is selected, the budget in
   KiB is instead the number of node planes per work unit of the plane
   march.  Used by the autotuner. */
int nsfv1_set_launch(nsfv1_handle* h, int threads, int lds_bytes);

#ifdef __cplusplus
}
#endif
#endif
