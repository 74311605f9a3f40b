// nsfv1_internal.hpp -- host-side declarations shared by the C ABI layer and
// the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_device.hpp"

namespace nsfv1 {

struct KMeshPtrs {
    const double* xyz;
    const int32_t* conn;
    const int64_t* rowptr;
    const int32_t* colidx;
    const int32_t* inc_elem;
    const int32_t* inc_node;
    const int32_t* inc_info;
    const int64_t* node_inc;
    const int32_t* wg_node;
    const int32_t* wg_rounds;
};

hipError_t upload_topology(const RefTopo* topo4);

hipError_t launch_assemble(int et, int jprec, const KMeshPtrs& m, const KParams& p,
                           const double* u, const double* uold, void* J, double* d, int* err,
                           int nwg, int threads, size_t lds_bytes, hipStream_t stream);

}  // namespace nsfv1
