// oibvh_layout.h -- the oibvh implicit-layout index math and the XCD
// workgroup remap, shared by host code and kernels (plain C/C++ when not
// compiled by hipcc).
//
// The oibvh tree of F leaves ("ostensibly-implicit" layout, oibvh.cuh:56-182)
// is a complete binary tree of height L = log2(next_pow2(F)) whose vl =
// next_pow2(F) - F right-most leaves, and every ancestor made only of them,
// are virtual. Real nodes are stored level by level, top-down, left to right
// ("real order"). Level l holds cnt(l) = 2^l - (vl >> (L - l)) real nodes
// starting at real index start(l) = 2^l - 1 - V(l), where V(l) =
// sum_{k<l} (vl >> (L - k)) = 2x - popcount(x), x = vl >> (L - l + 1), is the
// number of virtual nodes above level l. Because the real nodes of a level
// are a left-aligned prefix, cnt(l) = ceil(cnt(l+1) / 2) and node p of level l
// has children 2p and 2p + 1 of level l + 1 (the right one real iff
// 2p + 1 < cnt(l + 1)). An implicit index i = 2^l - 1 + p maps to real index
// i - V(l).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define OI_HD __host__ __device__ __forceinline__
#else
#define OI_HD static inline
#endif

OI_HD uint32_t oi_np2(uint32_t x)
{
    x--;
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    return x + 1u;
}

OI_HD uint32_t oi_ilog2(uint32_t x) { return 31u - (uint32_t)__builtin_clz(x); }
OI_HD uint32_t oi_popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

// number of tree nodes, oibvh_get_size (oibvh.cuh:56-59)
OI_HD uint32_t oi_tree_size(uint32_t F) { return 2u * F + oi_popc(oi_np2(F) - F) - 1u; }

// real nodes at level l (oibvh_level_real_node_count, oibvh.cuh:81-85)
OI_HD uint32_t oi_lvl_cnt(uint32_t l, uint32_t L, uint32_t vl) { return (1u << l) - (vl >> (L - l)); }

// virtual nodes above level l (oibvh_level_all_virtual_node_count of level
// l - 1, oibvh.cuh:94-99)
OI_HD uint32_t oi_virt_above(uint32_t l, uint32_t L, uint32_t vl)
{
    if (l == 0) return 0u;
    uint32_t x = vl >> (L - l + 1u);
    return 2u * x - oi_popc(x);
}

// real index of the first node of level l
OI_HD uint32_t oi_lvl_start(uint32_t l, uint32_t L, uint32_t vl)
{
    return ((1u << l) - 1u) - oi_virt_above(l, L, vl);
}

// right-most real implicit index of level l
// (oibvh_get_most_right_valid_implicitIdx, oibvh.cuh:178-182)
OI_HD uint32_t oi_lvl_last_implicit(uint32_t l, uint32_t L, uint32_t vl)
{
    return ((1u << (l + 1u)) - 2u) - (vl >> (L - l));
}

// Workgroups are dealt to the 8 XCDs of an MI355X round-robin in launch
// order (workgroup b runs on XCD b % 8), and each XCD has its own L2. This maps
// workgroup b to a work index so that each XCD takes one contiguous range of
// work indices: neighbouring Morton ranges, and the vertices they gather,
// stay in one L2. A bijection on [0, n).
OI_HD uint32_t oi_xcd_remap(uint32_t b, uint32_t n)
{
    const uint32_t q = n >> 3, r = n & 7u;
    const uint32_t x = b & 7u, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

