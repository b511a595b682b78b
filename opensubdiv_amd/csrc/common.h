// Shared device helpers for the gfx950 (CDNA4, wave64) kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace osdhip {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kNumXcd = 8;

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Index of this wavefront in its block, as a wave-uniform (scalar) value:
// slice indices, widths and loop bounds derived from it then live in SGPRs
// and guard scalar branches, so a wave's record loads issue back to back.
__device__ __forceinline__ int wave_in_block()
{
    return __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
}

// Wave-wide reductions (every lane receives the result).
__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int m = kWave / 2; m; m >>= 1) v = max(v, __shfl_xor(v, m, kWave));
    return v;
}

__device__ __forceinline__ unsigned wave_min_u(unsigned v)
{
#pragma unroll
    for (int m = kWave / 2; m; m >>= 1) v = min(v, (unsigned)__shfl_xor((int)v, m, kWave));
    return v;
}

// Wave-scope LDS hand-off: LDS traffic of one wave is processed in order,
// so only the compiler has to be kept from reordering across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Hardware dispatches work-group b to XCD (b % 8), each XCD with its own L2.
// Remap so XCD x handles the contiguous logical range [x*per, (x+1)*per):
// neighbouring stencils (sharing control points) then share one L2.
// gridDim.x is always a multiple of 8.
__device__ __forceinline__ int xcd_logical_block()
{
    const int per = (int)gridDim.x / kNumXcd;
    return ((int)blockIdx.x % kNumXcd) * per + (int)blockIdx.x / kNumXcd;
}

// Pointers read from device memory (segment tables) are generic to the
// compiler, which then emits flat loads that also count against lgkmcnt and
// force conservative waits; these casts restore global-memory accesses.
#define OSDHIP_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const OSDHIP_GLOBAL T *gptr(const T *p)
{
    return (const OSDHIP_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ OSDHIP_GLOBAL T *gptr_w(T *p)
{
    return (OSDHIP_GLOBAL T *)p;
}

// Read-only launch descriptors (segment tables, item lists): read through
// the constant address space, i.e. with scalar loads into SGPRs, since the
// compiler cannot prove on its own that the kernel's stores leave them
// unchanged.
#define OSDHIP_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const OSDHIP_CONST T *cptr(const T *p)
{
    return (const OSDHIP_CONST T *)p;
}

inline int grid_for_blocks(int nblocks)
{
    return (nblocks + kNumXcd - 1) / kNumXcd * kNumXcd;
}

}  // namespace osdhip
