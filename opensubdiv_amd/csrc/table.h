// Device stencil table shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/osdhip.h"

// Device copy of a (limit) stencil table: the Far CSR arrays (for the raw
// getters) plus the 64-row slice layout the table kernels stream, built on
// the device from the CSR arrays.
struct osdhip_stencil_table {
    int n = 0, nw = 0, nslices = 0, layout = 0;
    long long nnz = 0, sliceEntries = 0, bytes = 0;
    int *sizes = nullptr, *offsets = nullptr, *indices = nullptr;
    float *w[6] = {};
    long long *sliceOff = nullptr;
    int *sliceWidth = nullptr;
    int *rowSize = nullptr;          // sizes padded to nslices * 64 rows
    unsigned *rec = nullptr;         // layout 0: records of rec_dwords(nw)
    int *sidx = nullptr;             // layout 1: split streams
    float *sw[6] = {};
};

namespace osdhip {

inline int rec_dwords(int nw) { return nw == 1 ? 2 : nw == 3 ? 4 : 8; }

inline int last_error()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? OSDHIP_OK : (int)e;
}

}  // namespace osdhip
