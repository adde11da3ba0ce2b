// Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits.
//
// Per pass: (1) each 2048-key tile counts its digits in LDS; (2) an exclusive
// scan over the digit-major histogram gives every (digit, tile) its output
// base; (3) each tile re-reads its keys in eight 256-key rounds and places
// them stably: inside a wavefront, lanes with the same digit are found with
// eight ballots (one per digit bit) and ranked by popcount of the lower-lane
// peers; waves and rounds are ordered through LDS prefix counts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {
namespace {  // each including translation unit gets its own copy of the kernels

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortWaves = kSortThreads / 64;

__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const uint32_t *__restrict__ keys, int64_t n,
                                                             int shift, int num_tiles,
                                                             uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = base + static_cast<int64_t>(r) * kSortThreads + t;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[static_cast<int64_t>(t) * num_tiles + blockIdx.x] = h[t];
}

// Exclusive scan helpers: per-block scan writing block totals, then a
// single-block scan of the totals, then the add-back.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
    // Wave-level inclusive scan by shuffles, then across waves through LDS.
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < nw; ++w) {
            const uint32_t c = lds[w];
            lds[w] = s;
            s += c;
        }
        lds[nw] = s;
    }
    __syncthreads();
    const uint32_t r = lds[wave] + x - v;
    *total = lds[nw];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(uint32_t *__restrict__ data, int64_t n,
                                                             uint32_t *__restrict__ tile_sums) {
    __shared__ uint32_t lds[kScanThreads / 64 + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? data[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, lds, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) data[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(uint32_t *__restrict__ sums, int n) {
    __shared__ uint32_t lds[kScanThreads / 64 + 1];
    // n <= kScanTile (one block handles all tile sums).
    const int base = threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? sums[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, lds, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) sums[base + k] = run;
        run += v[k];
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t *__restrict__ data, int64_t n,
                                                           const uint32_t *__restrict__ tile_sums) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    const uint32_t add = tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) data[base + k] += add;
}

__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t *__restrict__ keys_in,
                                                                const uint32_t *__restrict__ vals_in, int64_t n,
                                                                int shift, int num_tiles,
                                                                const uint32_t *__restrict__ offsets,
                                                                uint32_t *__restrict__ keys_out,
                                                                uint32_t *__restrict__ vals_out) {
    __shared__ uint32_t digit_base[256];
    __shared__ uint32_t wave_count[kSortWaves][256];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    digit_base[t] = offsets[static_cast<int64_t>(t) * num_tiles + blockIdx.x];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) wave_count[w][t] = 0;
        __syncthreads();
        const int64_t i = base + static_cast<int64_t>(r) * kSortThreads + t;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t digit = (key >> shift) & 0xffu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt_mask));
        if (valid && rank == 0) wave_count[wave][digit] = static_cast<uint32_t>(__popcll(peers));
        __syncthreads();
        // Exclusive prefix over waves for digit t; the round total updates the base.
        {
            uint32_t s = 0;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const uint32_t c = wave_count[w][t];
                wave_count[w][t] = s;
                s += c;
            }
            // Stash this round's total in a register; the base moves after use.
            __syncthreads();
            if (valid) {
                const uint32_t pos = digit_base[digit] + wave_count[wave][digit] + rank;
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
            __syncthreads();
            digit_base[t] += s;
        }
        __syncthreads();
    }
}

// Sorts (keys, vals) in place using the given scratch buffers:
// keys_tmp/vals_tmp (n each), hist (256 * tiles), tile_sums (>= tiles of the
// histogram scan). Passes cover bits [0, key_bits).
inline hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp, int64_t n,
                                   int key_bits, uint32_t *hist, uint32_t *tile_sums, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    const int num_tiles = static_cast<int>((n + kSortTile - 1) / kSortTile);
    const int64_t hn = static_cast<int64_t>(256) * num_tiles;
    const int scan_blocks = static_cast<int>((hn + kScanTile - 1) / kScanTile);
    if (scan_blocks > kScanTile) return hipErrorInvalidValue;
    uint32_t *ki = keys, *vi = vals, *ko = keys_tmp, *vo = vals_tmp;
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(num_tiles), dim3(kSortThreads), 0, s, ki, n, shift, num_tiles, hist);
        hipLaunchKernelGGL(k_scan_tiles, dim3(scan_blocks), dim3(kScanThreads), 0, s, hist, hn, tile_sums);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tile_sums, scan_blocks);
        hipLaunchKernelGGL(k_scan_add, dim3(scan_blocks), dim3(kScanThreads), 0, s, hist, hn, tile_sums);
        hipLaunchKernelGGL(k_radix_scatter, dim3(num_tiles), dim3(kSortThreads), 0, s, ki, vi, n, shift, num_tiles,
                           hist, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
        ++passes;
    }
    if (passes & 1) {
        hipError_t e = hipMemcpyAsync(keys, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(vals, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace
}  // namespace rt
