// nsfv1_internal.hpp -- host-side declarations shared by the C ABI layer and
// the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_device.hpp"

namespace nsfv1 {

// Device arrays of one mesh and one launch plan.
struct KMeshPtrs {
    const double* xyz;        // [nv][dim]
    const int32_t* conn;      // [ne][nsh]
    const int64_t* rowptr;    // [ndof+1]
    const int32_t* colidx;    // [nnz]
    const int32_t* node_inc;  // [nv+1] incidences of each node (node-major)
    const int32_t* inc_elem;  // [ninc] element of the incidence
    const int32_t* inc_node;  // [ninc] node of the incidence
    const int32_t* inc_info;  // [ninc] corner | slot << 8
    const int32_t* wg_node;   // [nwg+1] first node of each workgroup
    const int32_t* wg_rounds; // [nwg] incidence slots of the workgroup
    const double* l2inv;      // [ne][nip] diffusion-length factor per SCVF
    const double* vol;        // [ne][nsh] SCV volumes
    uint8_t* upw_dec;         // [ne][nip] recorded skewed-upwind corner (255: zero velocity), or null
    const uint8_t* affine;    // [ne] 1 = parallelogram / parallelepiped (quadrilaterals, hexahedra)
    const uint64_t* inc_pos;  // [ninc] byte c = column block of corner c in the node's block row
    const int32_t* perm;      // [ninc] lane order of the plan: incidences of a workgroup sorted by slot
    const double* nu_elem;    // [ne] per-element kinematic viscosity (batched instances), or null
    const double* nu_ip;      // [ne][nip] kinematic viscosity at the SCVF ips (function-valued import), or null
    const double* f_scv;      // [ne][nsh][dim] source at the SCV ips (the corners), or null
    const double* f_scvf;     // [ne][nip][dim] source at the SCVF ips (stabilisation), or null
    const double* rho_ip;     // [ne][nip] density at the SCVF ips (function-valued import), or null
    const double* rho_scv;    // [ne][nsh] density at the SCV ips (the corners), or null
    int block;                // threads per workgroup of the launch (64..256)
    const int4* lane_a;       // [ninc] in lane order: {element, node, corner | slot << 8 | row length << 16, LDS row base}
    const uint64_t* lane_pos; // [ninc] in lane order: inc_pos of the lane's incidence
    const int4* wg_a;         // [2 nwg]: {v0, v1, i0, i1}, {jbase low, jbase high, jlen, rounds}
};

hipError_t launch_assemble(int et, int jprec, const KMeshPtrs& m, const KParams& p, const double* u,
                           const double* uold, void* J, double* d, int* err, int nwg, size_t lds_bytes,
                           hipStream_t stream);

// split path (FP32 Jacobian, common configuration): per-SCVF records, then
// the row gather; recs holds split_record_floats(et, ne) floats
bool split_supported(int et, const KParams& p);
int64_t split_record_floats(int et, int64_t ne);
// rowlane: one lane per (node, element, row) in the gather (plan with
// 64 / (dim+1) incidences per 64-lane workgroup)
hipError_t launch_assemble_split(int et, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                                 float* J, double* d, int* err, int nwg, size_t lds, int64_t ne, float* recs,
                                 bool rowlane, hipStream_t stream);

// stabilised SCVF velocities of every element: out[ne][nip][dim]
hipError_t launch_export_stab_vel(int et, const KMeshPtrs& m, const KParams& p, const double* u, const double* uold,
                                  int64_t ne, double* out, int* err, hipStream_t stream);

// coupled path (FLOW stabilisation, upwinds with ip shapes): per-SCVF row
// blocks in `recs` (coupled_record_bytes), then the row gather; with
// vel_out set, only the stabilised SCVF velocities out[ne][nip][dim]
int64_t coupled_record_bytes(int et, int jprec, int64_t ne);
hipError_t launch_assemble_coupled(int et, int jprec, const KMeshPtrs& m, const KParams& p, const double* u,
                                   const double* uold, void* J, double* d, int* err, int nwg, size_t lds, int64_t ne,
                                   void* recs, double* vel_out, hipStream_t stream);

// fixed record scratch of nsfv1_assemble_noalloc (nsfv1_scratch.hip):
// setup looks its device address up once per device (with that device
// current); noalloc_scratch returns it, or null before setup
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
hipError_t noalloc_scratch_setup(int device);
void* noalloc_scratch(int device);

hipError_t launch_geometry_cache(int et, const double* xyz, const int32_t* conn, int64_t ne, int difflen,
                                 double* l2inv, double* vol, uint8_t* affine, int* err, hipStream_t stream);

}  // namespace nsfv1
