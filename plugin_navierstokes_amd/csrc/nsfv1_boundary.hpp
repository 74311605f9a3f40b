// nsfv1_boundary.hpp -- launcher of the FV1 boundary discs (no-normal-stress
// outflow, symmetry, sliding wall) on the rows of the nodes of the listed
// boundary sides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

struct OutflowArgs {
    const double* xyz;
    const int32_t* conn;
    const int64_t* rowptr;
    const int32_t* inc_elem;
    const int32_t* inc_info;
    const uint64_t* inc_pos;
    const int32_t* bn_ptr;     // [nbn+1] boundary faces of each boundary node
    const int64_t* bn_node;    // [nbn] the node
    const int32_t* item_inc;   // incidence (node, element) of the face
    const int32_t* item_side;  // element side of the face
    int64_t nbn;
    double nu, rho, s_a;
    int laplace, stokes;
    int kind;  // BND_OUTFLOW, BND_SYMMETRY or BND_SLIDING
    double slide_k, slide_lim;  // sliding wall: sliding factor and sliding limit
    const double* nu_item;      // [items] viscosity at each face's ip (a function-valued viscosity), or null: nu
};

enum { BND_OUTFLOW = 0, BND_SYMMETRY = 1, BND_SLIDING = 2, BND_KINDS = 3 };

hipError_t launch_outflow(int et, int jprec, const OutflowArgs& a, const double* u, void* J, double* d, int* err,
                          hipStream_t s);

}  // namespace nsfv1
