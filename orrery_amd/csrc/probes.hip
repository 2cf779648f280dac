// Machine-ceiling probes (the reference's benchmarks/cuda_probes.cu:15-117 and
// benchmarks/harness/machine_limits, re-done for gfx950).
//
//   fma   independent FP32 FMA chains: the VALU FP32 rate (scalar FP32, the
//         build disables packed FP32)
//   rsq   chains of v_rsq_f32: the transcendental rate the force kernels pay
//         once per interaction
//   mix   one softened-gravity interaction's instruction mix without memory:
//         eleven FP32 FMAs and one add feeding one v_rsq_f32 per unit, the
//         issue-rate ceiling of the direct kernels (gfx950 issues v_rsq_f32
//         and the FMAs from one VALU, so their times add)
//   read  a grid-stride FP32 sum over a buffer: HBM read bandwidth
//   copy  float4 copy of a buffer: HBM read + write, and the known-size kernel
//         against which the FETCH_SIZE / WRITE_SIZE counters are calibrated
//
// Each probe is timed with hipEvents on the context stream and reports its
// algorithmic rate (flop/s, op/s or bytes/s).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"

namespace orr {
namespace {

constexpr int kAccumulators = 8;

__global__ __launch_bounds__(256) void fma_probe(float* __restrict__ out, int iterations, float m) {
    float acc[kAccumulators];
#pragma unroll
    for (int k = 0; k < kAccumulators; ++k) acc[k] = static_cast<float>(threadIdx.x + k);
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int k = 0; k < kAccumulators; ++k) acc[k] = __builtin_fmaf(acc[k], 1.000001f, m);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kAccumulators; ++k) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void rsq_probe(float* __restrict__ out, int iterations) {
    float acc[kAccumulators];
#pragma unroll
    for (int k = 0; k < kAccumulators; ++k) acc[k] = 1.f + static_cast<float>(threadIdx.x + k);
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int k = 0; k < kAccumulators; ++k) acc[k] = __builtin_amdgcn_rsqf(acc[k] + 1.f);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kAccumulators; ++k) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void mix_probe(float* __restrict__ out, int iterations, float m) {
    // Four interactions per trip: 11 FMA chains advance once per interaction and
    // four independent rsq chains once per trip, so no result waits on a rsq.
    constexpr int kChains = 11;
    float acc[kChains];
    float r[4];
#pragma unroll
    for (int k = 0; k < kChains; ++k) acc[k] = static_cast<float>(threadIdx.x + k) * 1e-3f;
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = 1.f + static_cast<float>(threadIdx.x + k);
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) acc[k] = __builtin_fmaf(acc[k], 0.999999f, m);
            r[u] = __builtin_amdgcn_rsqf(r[u] + 1.f);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += acc[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) s += r[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void read_probe(const float4* __restrict__ in, std::uint64_t n4,
                                                  float* __restrict__ out) {
    float s = 0.f;
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n4;
         i += static_cast<std::uint64_t>(gridDim.x) * blockDim.x) {
        const float4 v = in[i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy_probe(const float4* __restrict__ in, std::uint64_t n4,
                                                  float4* __restrict__ out) {
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n4;
         i += static_cast<std::uint64_t>(gridDim.x) * blockDim.x) {
        out[i] = in[i];
    }
}

}  // namespace

// kind: 0 fma, 1 rsq, 2 read, 3 copy, 4 mix. size: iterations (0, 1, 4) or bytes (2, 3).
// Returns the rate in units/s and the kernel time in milliseconds.
void probe(orr_ctx* ctx, int kind, std::uint64_t size, double* rate, double* ms) {
    hipEvent_t start = ctx->take_event(), stop = ctx->take_event();
    Buffer a, b;
    try {
        const std::uint32_t blocks = static_cast<std::uint32_t>(ctx->compute_units) * 32;
        double units = 0;
        if (kind == 0 || kind == 1 || kind == 4) {
            const int iterations = static_cast<int>(size);
            a.reserve(sizeof(float) * blocks * 256);
            ORR_HIP(hipEventRecord(start, ctx->stream));
            if (kind == 0) {
                hipLaunchKernelGGL(fma_probe, dim3(blocks), dim3(256), 0, ctx->stream, a.as<float>(),
                                   iterations, 0.5f);
                units = 2.0 * kAccumulators * static_cast<double>(iterations) * blocks * 256;  // flop
            } else if (kind == 4) {
                hipLaunchKernelGGL(mix_probe, dim3(blocks), dim3(256), 0, ctx->stream, a.as<float>(),
                                   iterations, 0.5f);
                units = 4.0 * static_cast<double>(iterations) * blocks * 256;  // interactions
            } else {
                hipLaunchKernelGGL(rsq_probe, dim3(blocks), dim3(256), 0, ctx->stream, a.as<float>(),
                                   iterations);
                units = 1.0 * kAccumulators * static_cast<double>(iterations) * blocks * 256;  // rsq
            }
            ORR_LAUNCHED("probe");
            ORR_HIP(hipEventRecord(stop, ctx->stream));
        } else {
            const std::uint64_t n4 = size / sizeof(float4);
            if (n4 == 0) fail(ORR_ERR_INVALID_ARGUMENT, "probe: buffer too small");
            a.reserve(n4 * sizeof(float4));
            b.reserve(kind == 3 ? n4 * sizeof(float4) : sizeof(float) * blocks * 256);
            ORR_HIP(hipMemsetAsync(a.ptr, 0, n4 * sizeof(float4), ctx->stream));
            ORR_HIP(hipEventRecord(start, ctx->stream));
            if (kind == 2) {
                hipLaunchKernelGGL(read_probe, dim3(blocks), dim3(256), 0, ctx->stream,
                                   a.as<float4>(), n4, b.as<float>());
                units = static_cast<double>(n4) * sizeof(float4);
            } else {
                hipLaunchKernelGGL(copy_probe, dim3(blocks), dim3(256), 0, ctx->stream,
                                   a.as<float4>(), n4, b.as<float4>());
                units = 2.0 * static_cast<double>(n4) * sizeof(float4);
            }
            ORR_LAUNCHED("probe");
            ORR_HIP(hipEventRecord(stop, ctx->stream));
        }
        ORR_HIP(hipEventSynchronize(stop));
        float t = 0.f;
        ORR_HIP(hipEventElapsedTime(&t, start, stop));
        *ms = t;
        *rate = units / (static_cast<double>(t) * 1e-3);
    } catch (...) {
        a.release();
        b.release();
        ctx->free_events.push_back(start);
        ctx->free_events.push_back(stop);
        throw;
    }
    a.release();
    b.release();
    ctx->free_events.push_back(start);
    ctx->free_events.push_back(stop);
}

}  // namespace orr
