// osdhip_eval_patches_sorted_noalloc: patch-ordered EvalPatches for callers
// that cannot allocate device memory (code under graph capture, serving
// loops, harnesses that forbid hipMalloc).  The sort's workspace is one
// file-scope device array, `scratch`, of fixed size; each GPU the library
// runs on holds its own copy until the process ends.  Its address on the
// current device is looked up once per device with hipGetSymbolAddress (the
// array's name in host code is only a host-side handle) and kept; kernels
// get that address.  The call allocates nothing: no hipMalloc/hipFree, no
// allocating library routine.  Every call on a GPU shares that GPU's array
// (the sort clears what it counts in on every call), so calls that overlap
// on one GPU -- on different streams, or from different threads -- race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/osdhip.h"
#include "patch_sort_layout.h"

namespace osdhip {
int eval_patches_in_workspace(const float *src, osdhip_buffer_desc srcDesc, float *const *outs,
                              const osdhip_buffer_desc *ods, int numPatchCoords,
                              const osdhip_patch_coord *coords, const osdhip_patch_array *arrays,
                              const int *patchIndices, const osdhip_patch_param *params,
                              char *ws, void *stream);
}

using namespace osdhip;

namespace {

constexpr size_t kScratchBytes = (size_t)OSDHIP_NOALLOC_SCRATCH_BYTES;
static_assert(PatchSortLayout(OSDHIP_NOALLOC_MAX_PATCH_COORDS).bytes <= kScratchBytes,
              "the sort of the largest accepted call fits the scratch array");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
char *g_scratch[kMaxDevices];        // scratch's address on each device, once looked up

// The current device's address of `scratch`, looked up on the first call on
// that device; NULL (and *rc set) when the lookup fails.
char *device_scratch(int *rc)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *rc = e != hipSuccess ? (int)e : OSDHIP_ERR_UNSUPPORTED;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc = e != hipSuccess ? (int)e : OSDHIP_ERR_UNSUPPORTED;
            return nullptr;
        }
        g_scratch[dev] = (char *)p;
    }
    return g_scratch[dev];
}

}  // namespace

int osdhip_eval_patches_sorted_noalloc(const float *src, osdhip_buffer_desc srcDesc,
                                       float *dst, osdhip_buffer_desc dstDesc,
                                       float *du, osdhip_buffer_desc duDesc,
                                       float *dv, osdhip_buffer_desc dvDesc,
                                       float *duu, osdhip_buffer_desc duuDesc,
                                       float *duv, osdhip_buffer_desc duvDesc,
                                       float *dvv, osdhip_buffer_desc dvvDesc,
                                       int numPatchCoords,
                                       const osdhip_patch_coord *coords,
                                       const osdhip_patch_array *arrays,
                                       const int *patchIndices,
                                       const osdhip_patch_param *params,
                                       void *stream)
{
    if (numPatchCoords > OSDHIP_NOALLOC_MAX_PATCH_COORDS) return OSDHIP_ERR_TOO_LARGE;
    int rc = OSDHIP_OK;
    char *ws = device_scratch(&rc);
    if (!ws) return rc;
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    const osdhip_buffer_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    return eval_patches_in_workspace(src, srcDesc, outs, ods, numPatchCoords, coords, arrays,
                                     patchIndices, params, ws, stream);
}
