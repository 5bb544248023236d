// Block-wide staging of operator tables from global memory into LDS.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dg {

// Copies NB bytes (a multiple of 16; both ends 16-byte aligned) as 16-byte vectors with eight loads in flight
// per thread, so a block pays the load latency a few times rather than once per element.  The data are moved
// as bit patterns (nothing is converted).  The caller synchronises the block before reading dst.
template <size_t NB>
__device__ __forceinline__ void stage_to_lds(void* dst, const void* src) {
    static_assert(NB % 16 == 0, "operator tables are whole 16-byte vectors");
    constexpr int U = 8;
    constexpr size_t n16 = NB / 16;
    uint4* d = static_cast<uint4*>(dst);
    const uint4* s = static_cast<const uint4*>(src);
    for (size_t i0 = threadIdx.x; i0 < n16; i0 += size_t(U) * blockDim.x) {
        uint4 v[U];
        // every element read (past the end: the last vector again), so the array lives in registers rather than
        // in private or LDS memory
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + size_t(u) * blockDim.x;
            v[u] = s[i < n16 ? i : n16 - 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = i0 + size_t(u) * blockDim.x;
            if (i < n16) d[i] = v[u];
        }
    }
}

// The same copy by LDS DMA (global_load_lds, 16 bytes per lane): wavefront w of the block moves the 1 KB chunks
// w, w + waves, ... straight into LDS, through no registers, so the copy stays in flight while the block does other
// work.  The caller waits for it (s_waitcnt vmcnt(0), which also waits for every earlier vector load) and then
// synchronises the block before reading dst.
template <size_t NB>
__device__ __forceinline__ void dma_stage_to_lds(void* dst, const void* src) {
    static_assert(NB % 16 == 0, "operator tables are whole 16-byte vectors");
    constexpr int nfull = int(NB / 1024), tail = int(NB % 1024) / 16;
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    const char* s = static_cast<const char*>(src) + lane * 16;
    char* d = static_cast<char*>(dst);
    for (int c = threadIdx.x >> 6; c < nfull; c += waves)
        __builtin_amdgcn_global_load_lds(s + c * 1024, d + c * 1024, 16, 0, 0);
    if (tail > 0 && (threadIdx.x >> 6) == nfull % waves && lane < tail)
        __builtin_amdgcn_global_load_lds(s + nfull * 1024, d + nfull * 1024, 16, 0, 0);
}

}  // namespace dg
