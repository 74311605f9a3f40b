/*
 * fv1_oracle.h -- CPU restatement of the UG4 NavierStokes plugin's FV1
 * incompressible element assembly.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline -- never as the thing that is measured or shipped.
 *
 * Arithmetic follows, loop by loop:
 *   incompressible/fv1/navier_stokes_fv1.cpp:250-869   (A, M, rhs parts)
 *   incompressible/fv1/navier_stokes_fv1.cpp:871-892   (peclet_blend)
 *   upwind.cpp:52-80, 133-172, 337-430, 505-575        (no/full/skewed/LPS)
 *   upwind_interface.h:333-358                          (upwind_vel)
 *   incompressible/fv1/stabilization.cpp:102-242        (diff length, FIELDS)
 *   incompressible/fv1/stabilization.cpp:805-850        (no stabilisation)
 *   incompressible/fv1/diffusion_length.h:48-199        (RAW/FIVEPOINT/COR)
 * The FV1 dual-mesh geometry, the side/ray intersection, the DoF numbering
 * and the CSR scatter belong to ugcore, which is not available; they are
 * restated from the published box-method construction (see the comments in
 * fv1_oracle.c) and are the convention the product path shares.
 */
#ifndef FV1_ORACLE_H
#define FV1_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_TRI = 0, OR_QUAD = 1, OR_TET = 2, OR_HEX = 3, OR_PRISM = 4 };
enum { OR_UPW_NONE_SET = -1, OR_UPW_NO = 0, OR_UPW_FULL = 1, OR_UPW_SKEWED = 2, OR_UPW_LPS = 3, OR_UPW_POSITIVE = 4,
       OR_UPW_REGULAR = 5 };
enum { OR_STAB_NONE_SET = -1, OR_STAB_FIELDS = 0, OR_STAB_WITHOUT = 1, OR_STAB_FLOW = 2 };
enum { OR_DL_RAW = 0, OR_DL_FIVEPOINT = 1, OR_DL_COR = 2 };

typedef struct {
    int dim;
    int conv_upwind;   /* upwind object of the convective term (m_spConvUpwind), -1 = unset */
    int stab_type;     /* m_spStab, -1 = unset */
    int stab_upwind;   /* upwind object held by the stabilisation, -1 = unset */
    int pac;           /* 1: the stabilisation is used as upwind (m_spConvStab) */
    int difflen;       /* diffusion-length method of the Schneider-Raw stabilisation */
    int stokes, laplace, peclet;
    double exact;      /* exact-Jacobian factor (m_bFullNewtonFactor) */
    double nu, rho;
    int has_source;
    double f[3];
    int force_matrix;  /* pinning knob: solve the ip system as a matrix even without ip shapes */
    /* imports evaluated at the integration points (NULL: the constants
       above).  Mesh-level calls index them by element: nu_ip [ne][nip] at
       the SCVF ips, f_scv [ne][nsh][dim] at the SCV ips (the corners),
       f_scvf [ne][nip][dim] at the SCVF ips; per-element calls read the
       element's own rows. */
    const double* nu_ip;
    const double* f_scv;
    const double* f_scvf;
    /* viscosity imported at the boundary-face ips (NULL: nu): boundary
       calls on a list of sides read [nb][4] (face j = the side's j-th
       corner), element calls on one side read [4] */
    const double* nu_bf;
    /* density imported at the integration points (NULL: rho), indexed like
       nu_ip: rho_ip [ne][nip] at the SCVF ips, rho_scv [ne][nsh] at the SCV
       ips (the corners) */
    const double* rho_ip;
    const double* rho_scv;
} or_params;

/* FV1 geometry of one element (restated ugcore FV1Geometry). */
typedef struct {
    int et, dim, nsh, nip;
    double co[8][3];
    int from[12], to[12];
    double ipl[12][3], ipg[12][3], n[12][3];
    double phi[12][8], g[12][8][3];
    double c0[12][3], c2[12][3];
    double vol[8];
} or_geo;

int  or_nsh(int et);
int  or_nip(int et);
int  or_dim(int et);

/* geometry of one element; corners row-major [nsh][dim] */
int or_geometry(int et, const double* corners, or_geo* out);

/* Side/ray intersection searching upwind (t <= 0).  Returns 0 on success. */
int or_side_ray(int et, const double* corners, const double* from, const double* dir,
                int* side, double* glob, double* loc);

/* Upwind shapes of one element: corner shapes U[nip][8], ip shapes
   W[nip][12] (NULL: not wanted; nonzero only for the positive upwind),
   convlen[nip]. */
int or_upwind(int type, int et, const double* corners, const double* ipvel /*[nip][dim]*/,
              double* U, double* W, double* convlen);

/* Local A-part defect and Jacobian of one element.
   uloc/uold: [nsh][dim+1] (node-major, DoF interleave); uold may be NULL.
   Jloc: [(nsh*(dim+1))^2] row-major with local index i*(dim+1)+a; dloc: [nsh*(dim+1)].
   Either output may be NULL.  Contributions are ADDED. */
int or_elem_A(const or_params* p, int et, const double* corners, const double* uloc,
              const double* uold, double dt, double* Jloc, double* dloc);
/* Term-magnitude scale of or_elem_A (sum of |terms| per local entry). */
int or_elem_A_abs(const or_params* p, int et, const double* corners, const double* uloc,
                  const double* uold, double dt, double* Jabs, double* dabs);
int or_elem_M(const or_params* p, int et, const double* corners, const double* uloc,
              double* Jloc, double* dloc);
int or_elem_rhs(const or_params* p, int et, const double* corners, double* floc);

/* Global assembly into a CSR pattern with DoF = v*(dim+1)+c.
   parts bit 1: A, bit 2: M, bit 4: rhs (rhs is ADDED with weight s_f).
   J = s_a*J_A + s_m*J_M ; d = s_a*d_A + s_m*d_M + s_f*f.  J, d zeroed first.
   Elements [e0, e1) only (pass 0, ne for all). Returns 0 or a negative error. */
int or_assemble(const or_params* p, int et, const double* xyz, int64_t nv,
                const int32_t* conn, int64_t e0, int64_t e1,
                const double* u, const double* uold, double dt,
                double s_a, double s_m, double s_f, int parts,
                const int64_t* rowptr, const int32_t* colidx,
                double* J, double* d);

/* Same, but returns the term-magnitude scale per entry (see or_elem_A_abs). */
int or_assemble_colored(const or_params* p, int et, const double* xyz, int64_t nv,
                        const int32_t* conn, const int64_t* order, const int64_t* color_ptr, int ncolors,
                        const double* u, const double* uold, double dt,
                        double s_a, double s_m, double s_f, int parts,
                        const int64_t* rowptr, const int32_t* colidx,
                        double* J, double* d, int nthreads);
int or_assemble_abs(const or_params* p, int et, const double* xyz, int64_t nv,
                    const int32_t* conn, int64_t ne,
                    const double* u, const double* uold, double dt,
                    double s_a, double s_m, int parts,
                    const int64_t* rowptr, const int32_t* colidx,
                    double* Jabs, double* dabs);

/* CSR pattern: node adjacency closure (x) (dim+1)^2, sorted columns.
   First call with colidx == NULL returns nnz in *nnz_out (rowptr filled). */
int or_pattern(int et, int64_t nv, const int32_t* conn, int64_t ne,
               int64_t* rowptr, int32_t* colidx, int64_t* nnz_out);

const char* or_last_error(void);

/* FV1 boundary faces of one element side (restated ugcore FV1Geometry BF). */
typedef struct {
    int nbf;
    int node[4];
    double n[4][3], ipl[4][3], ipg[4][3];
    double phi[4][8], g[4][8][3];
} or_bfgeo;
int or_bnd_faces(int et, const double* corners, int side, or_bfgeo* B);

/* No-normal-stress outflow boundary disc
   (incompressible/fv1/bnd/no_normal_stress_outflow_fv1.cpp:188-426):
   element-local contributions of one boundary side, ADDED to Jl/dl. */
int or_outflow_elem(const or_params* p, int et, const double* corners, int side, const double* uloc,
                    double* Jl, double* dl);
/* adds s_a times the contributions of nb boundary sides (element, side) to J/d */
int or_outflow(const or_params* p, int et, const double* xyz, int64_t nv, const int32_t* conn,
               const int32_t* belem, const int32_t* bside, int64_t nb, const double* u, double s_a,
               const int64_t* rowptr, const int32_t* colidx, double* J, double* d);

/* Symmetry boundary disc (incompressible/fv1/bnd/symmetric_boundary_fv1.cpp:
   224-386): element-local contributions of one boundary side, ADDED. */
int or_symmetry_elem(const or_params* p, int et, const double* corners, int side, const double* uloc,
                     double* Jl, double* dl);
int or_symmetry(const or_params* p, int et, const double* xyz, int64_t nv, const int32_t* conn,
                const int32_t* belem, const int32_t* bside, int64_t nb, const double* u, double s_a,
                const int64_t* rowptr, const int32_t* colidx, double* J, double* d);

/* Sliding wall disc (incompressible/fv1/bnd/wall_sliding_fv1.cpp:229-425)
   with constant sliding factor k and sliding limit lim. */
int or_sliding_elem(const or_params* p, int et, const double* corners, int side, const double* uloc, double k,
                    double lim, double* Jl, double* dl);
int or_sliding(const or_params* p, int et, const double* xyz, const int32_t* conn, const int32_t* belem,
               const int32_t* bside, int64_t nb, const double* u, double s_a, double k, double lim,
               const int64_t* rowptr, const int32_t* colidx, double* J, double* d);

/* Data exports (navier_stokes_fv1.cpp:894-1491).  what: 0 velocity,
   1 velocity gradient, 2 stabilised ip velocity, 3 pressure, 4 pressure
   gradient; at: 0 SCVF ips, 1 SCV ips (corners), 2 the local points
   lips[nip][dim].  out: [ne][points][ncomp]. */
int or_export_ncomp(int what, int dim);
int or_export_elem(const or_params* p, int et, const double* corners, const double* uloc, const double* uold,
                   double dt, int what, int at, const double* lips, int nip, double* out);
int or_export(const or_params* p, int et, const double* xyz, const int32_t* conn, int64_t ne, const double* u,
              const double* uold, double dt, int what, int at, const double* lips, int nip, double* out);

#ifdef __cplusplus
}
#endif
#endif
