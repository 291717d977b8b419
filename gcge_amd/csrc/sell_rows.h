// Device helper for SELL-64 row kernels.
//
// A SELL-64 slice is 64 consecutive rows stored to the width of its widest
// row, so with one thread per row a slice is exactly one wavefront and the
// slice width is uniform across it.  Looping over a row's entries with a
// per-entry guard (`for j < W: if (j < w) ...`) compiles to one branch per
// entry, each ending in a wait for its own loads, which leaves a single
// entry's gathers in flight per wavefront.  for_width instead turns the
// wave-uniform width into a compile-time count: a scalar branch chain
// selects a straight-line body with exactly w entries, whose gathers are
// all issued before the first wait.  Entries are still summed in order
// 0 .. w-1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace gcge {

// Calls f(std::integral_constant<int, K>{}) with K == w for 0 <= w <= W;
// w must be the same on every active lane (pass it through
// __builtin_amdgcn_readfirstlane).
template <int W, int K = 0, typename F>
__device__ __forceinline__ void for_width(int w, F &&f)
{
    if constexpr (K <= W) {
        if (w == K)
            f(std::integral_constant<int, K>{});
        else
            for_width<W, K + 1>(w, f);
    }
}

// Row block of the calling workgroup.  The hardware places workgroup i on
// XCD i % 8, each XCD with its own L2.  In plain order (XC false) block b
// takes rows b BS ..., so neighbouring row blocks sit on different XCDs
// and a gather of a neighbour row misses in this XCD's L2.  With XC each
// XCD takes a contiguous range of row blocks, which keeps the rows a
// stencil-like matrix gathers (near rows, and rows a grid plane away) in
// the same L2.
template <bool XC>
__device__ __forceinline__ int64_t row_block()
{
    const int64_t bid = blockIdx.x;
    if constexpr (!XC) {
        return bid;
    } else {
        const int64_t nblk = gridDim.x;
        const int64_t q = nblk / 8, r = nblk % 8;
        const int64_t x = bid % 8, k = bid / 8;
        return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
    }
}

// Plane-grouped slices for 256-thread workgroups: with P slices per grid
// plane, workgroup b takes in-plane slice j = b % P of four consecutive
// planes 4 (b / P) .. +3, one per wavefront, so a row's neighbours one
// plane away are mostly rows a sibling wavefront of the same CU gathers at
// the same time.  The grid has ceil(S / 4P) P workgroups for S slices;
// slices past S belong to no row and their wavefronts exit.
// b is the row block (blockIdx.x, or its XCD-contiguous remap, which keeps
// the in-plane neighbour workgroups b +- 1, b +- 2 on one XCD).
__device__ __forceinline__ int64_t plane_slice(int64_t P, int64_t b)
{
    const int64_t w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    return (b / P) * 4 * P + w * P + b % P;
}

// The slice of the calling wavefront for BS-thread workgroups: row block b,
// wavefront v holds slice (BS / 64) b + v.
template <int BS = 256>
__device__ __forceinline__ int64_t wave_slice(int64_t b)
{
    return b * (BS / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Raw buffer loads: the base address and size live in a wave-uniform
// resource (SGPRs) and each lane supplies a 32-bit byte offset, so a gather
// costs one VGPR of address instead of two.  Offsets at or beyond the
// resource's size read as zero.
constexpr int kRsrcFlags = 0x00020000;  // raw 32-bit data, CDNA

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             kRsrcFlags);
}

template <typename V>
__device__ __forceinline__ V buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (sizeof(V) == 8)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    else
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

}  // namespace gcge
