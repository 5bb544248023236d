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

// LDS DMA of NB contiguous bytes (a multiple of 16) from src into one wavefront's LDS buffer dst: each instruction
// moves 64 lanes x 16 B, lane l's bytes landing at offset 16 l.  The wavefront waits for it (s_waitcnt vmcnt(0))
// before reading dst.  A partial last chunk runs on a masked subset of lanes: keep one such copy per code path, since
// the compiler may merge two adjacent masked copies into one instruction whose LDS base (M0, one per wavefront) it
// takes from a per-lane value; dma_to_lds_whole has no masked chunk.
template <size_t NB>
__device__ __forceinline__ void dma_to_lds(const double* src, double* dst, int lane) {
    static_assert(NB % 16 == 0, "whole 16-byte vectors");
    constexpr int nfull = int(NB / 1024), tail = int(NB % 1024) / 16;
    const char* s = reinterpret_cast<const char*>(src) + lane * 16;
#pragma unroll
    for (int c = 0; c < nfull; ++c) __builtin_amdgcn_global_load_lds(s + c * 1024, dst + c * 128, 16, 0, 0);
    if (tail > 0 && lane < tail) __builtin_amdgcn_global_load_lds(s + nfull * 1024, dst + nfull * 128, 16, 0, 0);
}

// Bytes of an LDS buffer that receives NB bytes by the whole-chunk DMA copies below.
constexpr size_t dma_span(size_t nb) { return (nb + 1023) / 1024 * 1024; }

// Whole-chunk LDS DMA (global_load_lds, 16 bytes per lane): every instruction runs with all 64 lanes active and a
// chunk-aligned LDS base; lanes past the NB bytes re-read the last 16 valid bytes and write them into dst's padding
// (dst holds dma_span(NB) bytes).  With no lane masked off, the compiler has no divergent copies it could merge
// into one instruction with a per-lane LDS base (global_load_lds takes its base from M0, one value per wavefront).
// One wavefront's copy; it waits (s_waitcnt vmcnt(0)) before reading dst.
template <size_t NB>
__device__ __forceinline__ void dma_to_lds_whole(const void* src, void* dst, int lane) {
    static_assert(NB % 16 == 0 && NB >= 16, "whole 16-byte vectors");
    constexpr int nc = int(dma_span(NB) / 1024);
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
#pragma unroll
    for (int c = 0; c < nc; ++c) {
        const size_t off = size_t(c) * 1024 + size_t(lane) * 16;
        __builtin_amdgcn_global_load_lds(s + (off < NB ? off : NB - 16), d + c * 1024, 16, 0, 0);
    }
}

// stage_to_lds by whole-chunk LDS DMA: wavefront w of the block moves chunks w, w + waves, ... straight into LDS,
// through no registers, so the copy stays in flight while the block does other work (dst holds dma_span(NB)
// bytes).  The caller waits for it (s_waitcnt vmcnt(0), which also waits for every earlier vector memory access) and
// then synchronises the block before reading dst.
template <size_t NB>
__device__ __forceinline__ void dma_stage_to_lds(void* dst, const void* src) {
    static_assert(NB % 16 == 0 && NB >= 16, "operator tables are whole 16-byte vectors");
    constexpr int nc = int(dma_span(NB) / 1024);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), waves = blockDim.x >> 6;
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    for (int c = wave; c < nc; c += waves) {
        const size_t off = size_t(c) * 1024 + size_t(lane) * 16;
        __builtin_amdgcn_global_load_lds(s + (off < NB ? off : NB - 16), d + c * 1024, 16, 0, 0);
    }
}

}  // namespace dg
