// Device-wide exclusive prefix sum of uint32 values (two passes, deterministic).
//
// Pass 1 sums each 2048-element block; pass 2 scans each block again, with the
// sum of the block sums before it as its offset, and its last block writes the
// grand total. No atomics, so the result never depends on scheduling. In place
// is allowed (in == out).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"

namespace orr {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kPerBlock = kThreads * kItems;

__device__ std::uint32_t block_exclusive(std::uint32_t v, std::uint32_t* lds, std::uint32_t& total) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const std::uint32_t a = threadIdx.x >= static_cast<unsigned>(off) ? lds[threadIdx.x - off] : 0u;
        __syncthreads();
        lds[threadIdx.x] += a;
        __syncthreads();
    }
    total = lds[kThreads - 1];
    const std::uint32_t inclusive = lds[threadIdx.x];
    __syncthreads();
    return inclusive - v;
}

__global__ __launch_bounds__(kThreads) void scan_block_sums(const std::uint32_t* __restrict__ in,
                                                            std::uint64_t n,
                                                            std::uint32_t* __restrict__ sums) {
    __shared__ std::uint32_t lds[kThreads];
    const std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * kPerBlock + threadIdx.x * kItems;
    std::uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (base + k < n) local += in[base + k];
    }
    std::uint32_t total = 0;
    (void)block_exclusive(local, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Each block first adds up the block sums before it (at most a few thousand
// values, read by the whole block), so no separate pass scans them.
__global__ __launch_bounds__(kThreads) void scan_apply(const std::uint32_t* __restrict__ in,
                                                       std::uint64_t n,
                                                       const std::uint32_t* __restrict__ sums,
                                                       std::uint32_t blocks,
                                                       std::uint32_t* __restrict__ out,
                                                       std::uint32_t* __restrict__ total_out) {
    __shared__ std::uint32_t lds[kThreads];
    std::uint32_t part = 0;
    for (std::uint32_t i = threadIdx.x; i < blockIdx.x; i += kThreads) part += sums[i];
    std::uint32_t offset = 0;
    (void)block_exclusive(part, lds, offset);
    const std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * kPerBlock + threadIdx.x * kItems;
    std::uint32_t v[kItems];
    std::uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        local += v[k];
    }
    std::uint32_t total = 0;
    std::uint32_t run = block_exclusive(local, lds, total) + offset;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (total_out != nullptr && blockIdx.x + 1 == blocks && threadIdx.x == 0) *total_out = offset + total;
}

}  // namespace

std::size_t scan_scratch_words(std::uint64_t n) {
    return static_cast<std::size_t>((n + kPerBlock - 1) / kPerBlock) + 1;
}

void exclusive_scan_u32(orr_ctx* ctx, const std::uint32_t* in, std::uint32_t* out, std::uint64_t n,
                        std::uint32_t* scratch, std::uint32_t* total_out) {
    if (n == 0) return;
    const std::uint32_t blocks = static_cast<std::uint32_t>((n + kPerBlock - 1) / kPerBlock);
    hipLaunchKernelGGL(scan_block_sums, dim3(blocks), dim3(kThreads), 0, ctx->stream, in, n, scratch);
    ORR_LAUNCHED("scan_block_sums");
    hipLaunchKernelGGL(scan_apply, dim3(blocks), dim3(kThreads), 0, ctx->stream, in, n, scratch,
                       blocks, out, total_out);
    ORR_LAUNCHED("scan_apply");
}

}  // namespace orr
