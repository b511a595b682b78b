// Device stencil table shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/osdhip.h"

// Device copy of a (limit) stencil table: the Far CSR arrays (for the raw
// getters) plus the 64-row slice layout the table kernels stream, built on
// the device from the CSR arrays.
struct osdhip_stencil_table {
    int n = 0, nw = 0, nslices = 0, layout = 0;
    int keepCsr = 1;                 // the Far CSR arrays stay on the device
    int group = 8;                   // slice kernels: entries per gather group
    int kernel = 0;                  // table eval: 0 slice layout, 1 CSR
    int order = 0;                   // layout 2 rows: 0 dominant vertex + size windows,
                                     // 1 Far order, 2 dominant vertex only
    long long nnz = 0, sliceEntries = 0, bytes = 0;
    int *sizes = nullptr, *offsets = nullptr, *indices = nullptr;
    float *w[6] = {};
    long long *sliceOff = nullptr;
    int *sliceWidth = nullptr;
    int *rowSize = nullptr;          // sizes padded to nslices * 64 rows
    unsigned *rec = nullptr;         // layouts 0, 2: records of rec_dwords(nw)
    int *sidx = nullptr;             // layout 1: split streams
    float *sw[6] = {};
    // layout 2 (locality-ordered slices): slice row i holds Far row perm[i]
    // (>= n: padding); a slice with nuniq[s] >= 0 lists its distinct control
    // vertices at ctrl[ctrlOff[s] ...] and its records hold positions in
    // that list; nuniq[s] < 0: records hold control indices
    int *perm = nullptr;
    int *nuniq = nullptr;
    long long *ctrlOff = nullptr;
    int *ctrl = nullptr;
    long long ctrlLen = 0;
    int listedSlices = 0;
    int sizeWindow = 0;              // layout 2, order 0: rows sorted by size per window (0: none)
};

namespace osdhip {

inline int rec_dwords(int nw) { return nw == 1 ? 2 : nw == 3 ? 4 : 8; }

inline int last_error()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? OSDHIP_OK : (int)e;
}

}  // namespace osdhip
