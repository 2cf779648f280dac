// Direct O(N^2) softened gravity on gfx950.
//
// Replaces the reference's CUDA direct kernel (src/solvers/cuda_direct_kernel.cu:87-214),
// which gives one thread one target and stages j-tiles of positions through
// shared memory with two barriers per tile. Here:
//
//   * Sources are packed once per evaluation into 64-particle tiles of float4
//     {dx, dy, dz, m}, the offsets taken in FP64 from the tile's FP64 centre
//     (the midpoint of its bounding box) and only then rounded to FP32. In
//     Morton order a tile is spatially compact, so an offset carries the
//     precision of the tile's size rather than of the system's.
//   * A wave reads a source tile through the scalar unit: the address is
//     wave-uniform, so each {dx, dy, dz, m} arrives in SGPRs and is broadcast
//     to all 64 lanes as an instruction operand. No LDS, no barriers.
//   * Each lane owns R targets. Per target and source tile the FP64
//     centre-to-centre shift is rounded once to FP32, so a pairwise
//     separation is (o_j - (o_i - shift)) with every term small.
//   * The 64 terms of one source tile are summed plainly in FP32 and the tile
//     partial is Kahan-added into the running sum (compensation at tile
//     granularity, three VALU operations per target per tile).
//   * The source range is split over gridDim.y; a fixed-order FP64 reduction
//     kernel folds the splits, multiplies by G and scatters to particle order.
//
// Pairwise arithmetic per interaction: 3 sub, 3 FMA (r^2 + eps^2), 1 rsq,
// 3 mul, 3 FMA: 13 VALU issues for the reference's 20-flop interaction
// (docs/performance.md:103-109 counts 20).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "engine.hpp"

namespace orr {
namespace {

constexpr int kTile = 64;    // sources per tile = one wave of the prepare kernel
constexpr int kBlock = 256;  // four waves

__device__ inline double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// One wave per tile: gather (optionally through `order`), centre, pack.
__global__ __launch_bounds__(64) void direct_prepare(const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ m,
                                                     const std::uint32_t* __restrict__ order,
                                                     std::uint64_t n, float4* __restrict__ packed,
                                                     double4* __restrict__ centres) {
    const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * kTile + threadIdx.x;
    const bool valid = k < n;
    const std::uint64_t src = valid ? (order != nullptr ? order[k] : k) : 0;
    const double px = valid ? x[src] : 0.0;
    const double py = valid ? y[src] : 0.0;
    const double pz = valid ? z[src] : 0.0;
    const double pm = valid ? m[src] : 0.0;
    const double inf = __builtin_huge_val();
    const double cx = 0.5 * (wave_min(valid ? px : inf) + wave_max(valid ? px : -inf));
    const double cy = 0.5 * (wave_min(valid ? py : inf) + wave_max(valid ? py : -inf));
    const double cz = 0.5 * (wave_min(valid ? pz : inf) + wave_max(valid ? pz : -inf));
    packed[k] = valid ? make_float4(static_cast<float>(px - cx), static_cast<float>(py - cy),
                                    static_cast<float>(pz - cz), static_cast<float>(pm))
                      : make_float4(0.f, 0.f, 0.f, 0.f);
    if (threadIdx.x == 0) centres[blockIdx.x] = make_double4(cx, cy, cz, 0.0);
}

// kSafe masks r^2 == 0 (zero softening): the self term and coincident pads
// then contribute m * finite * 0 = 0 instead of NaN, as the reference's CUDA
// kernel masks self and tail terms (cuda_direct_kernel.cu:134-159).
template<int R, bool kSafe>
__global__ __launch_bounds__(kBlock) void direct_tiles(const float4* __restrict__ packed,
                                                       const double4* __restrict__ centres,
                                                       std::uint32_t npad, std::uint32_t tiles,
                                                       std::uint32_t tiles_per_split, float eps2,
                                                       float* __restrict__ partial) {
    const std::uint32_t split = blockIdx.y;
    const std::uint32_t j0 = split * tiles_per_split;
    const std::uint32_t j1 = min(tiles, j0 + tiles_per_split);
    const std::uint32_t base = blockIdx.x * (kBlock * R) + threadIdx.x;

    float ox[R], oy[R], oz[R];
    double cx[R], cy[R], cz[R];
    float sx[R], sy[R], sz[R], kx[R], ky[R], kz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t k = min(base + r * kBlock, npad - 1);
        const float4 o = packed[k];
        const double4 c = centres[k / kTile];
        ox[r] = o.x;
        oy[r] = o.y;
        oz[r] = o.z;
        cx[r] = c.x;
        cy[r] = c.y;
        cz[r] = c.z;
        sx[r] = sy[r] = sz[r] = 0.f;
        kx[r] = ky[r] = kz[r] = 0.f;
    }

    for (std::uint32_t J = j0; J < j1; ++J) {
        const double4 C = centres[J];
        float ux[R], uy[R], uz[R], px[R], py[R], pz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // Target offset re-expressed relative to the source tile's centre.
            ux[r] = ox[r] - static_cast<float>(C.x - cx[r]);
            uy[r] = oy[r] - static_cast<float>(C.y - cy[r]);
            uz[r] = oz[r] - static_cast<float>(C.z - cz[r]);
            px[r] = py[r] = pz[r] = 0.f;
        }
        const float4* __restrict__ src = packed + static_cast<std::size_t>(J) * kTile;
#pragma unroll 16
        for (int j = 0; j < kTile; ++j) {
            const float4 q = src[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = q.x - ux[r];
                const float dy = q.y - uy[r];
                const float dz = q.z - uz[r];
                float r2 = __builtin_fmaf(dx, dx, eps2);
                r2 = __builtin_fmaf(dy, dy, r2);
                r2 = __builtin_fmaf(dz, dz, r2);
                if constexpr (kSafe) r2 = r2 > 0.f ? r2 : 1.f;
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float f = (q.w * inv) * (inv * inv);
                px[r] = __builtin_fmaf(f, dx, px[r]);
                py[r] = __builtin_fmaf(f, dy, py[r]);
                pz[r] = __builtin_fmaf(f, dz, pz[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // Kahan: sum += partial, compensation carries the lost low part.
            float y = px[r] - kx[r];
            float t = sx[r] + y;
            kx[r] = (t - sx[r]) - y;
            sx[r] = t;
            y = py[r] - ky[r];
            t = sy[r] + y;
            ky[r] = (t - sy[r]) - y;
            sy[r] = t;
            y = pz[r] - kz[r];
            t = sz[r] + y;
            kz[r] = (t - sz[r]) - y;
            sz[r] = t;
        }
    }

    float* out = partial + static_cast<std::size_t>(split) * 6 * npad;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t k = base + r * kBlock;
        if (k < npad) {
            out[k] = sx[r];
            out[npad + k] = sy[r];
            out[2 * npad + k] = sz[r];
            out[3 * npad + k] = kx[r];
            out[4 * npad + k] = ky[r];
            out[5 * npad + k] = kz[r];
        }
    }
}

// Fixed-order FP64 fold of the splits; sum - compensation per split.
__global__ __launch_bounds__(256) void direct_reduce(const float* __restrict__ partial,
                                                     std::uint32_t splits, std::uint32_t npad,
                                                     std::uint64_t n,
                                                     const std::uint32_t* __restrict__ order,
                                                     double G, double* __restrict__ ax,
                                                     double* __restrict__ ay,
                                                     double* __restrict__ az) {
    const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (std::uint32_t s = 0; s < splits; ++s) {
        const float* p = partial + static_cast<std::size_t>(s) * 6 * npad;
        a0 += static_cast<double>(p[k]) - static_cast<double>(p[3 * npad + k]);
        a1 += static_cast<double>(p[npad + k]) - static_cast<double>(p[4 * npad + k]);
        a2 += static_cast<double>(p[2 * npad + k]) - static_cast<double>(p[5 * npad + k]);
    }
    const std::uint64_t dst = order != nullptr ? order[k] : k;
    ax[dst] = G * a0;
    ay[dst] = G * a1;
    az[dst] = G * a2;
}

constexpr int kTargetsPerLane = 2;

}  // namespace

void direct_enqueue(orr_ctx* ctx, const orr_direct_args& a, const std::uint32_t* order) {
    const std::uint64_t n = a.count;
    if (n == 0) return;
    if (n > (std::uint64_t{1} << 31)) fail(ORR_ERR_LENGTH, "direct: more than 2^31 particles");
    const std::uint32_t tiles = static_cast<std::uint32_t>((n + kTile - 1) / kTile);
    const std::uint32_t npad = tiles * kTile;
    constexpr int R = kTargetsPerLane;
    const std::uint32_t blocks_i = (npad + kBlock * R - 1) / (kBlock * R);
    // Enough blocks to fill 256 CUs several times over; each split keeps >= 1 tile.
    const std::uint32_t want = 4096;
    std::uint32_t splits = std::max<std::uint32_t>(1, (want + blocks_i - 1) / blocks_i);
    splits = std::min(splits, tiles);
    const std::uint32_t per_split = (tiles + splits - 1) / splits;
    splits = (tiles + per_split - 1) / per_split;

    ctx->direct_packed.reserve(sizeof(float4) * npad);
    ctx->direct_centres.reserve(sizeof(double4) * tiles);
    ctx->direct_partials.reserve(sizeof(float) * 6 * static_cast<std::size_t>(npad) * splits);
    float4* packed = ctx->direct_packed.as<float4>();
    double4* centres = ctx->direct_centres.as<double4>();
    float* partials = ctx->direct_partials.as<float>();

    {
        PhaseTimer t(ctx, kDirectPrepare);
        hipLaunchKernelGGL(direct_prepare, dim3(tiles), dim3(kTile), 0, ctx->stream, a.x, a.y, a.z,
                           a.m, order, n, packed, centres);
        ORR_LAUNCHED("direct_prepare");
    }
    const float eps2 = static_cast<float>(a.softening_squared);
    {
        PhaseTimer t(ctx, kDirectKernel);
        if (a.softening_squared > 0) {
            hipLaunchKernelGGL((direct_tiles<R, false>), dim3(blocks_i, splits), dim3(kBlock), 0,
                               ctx->stream, packed, centres, npad, tiles, per_split, eps2, partials);
        } else {
            hipLaunchKernelGGL((direct_tiles<R, true>), dim3(blocks_i, splits), dim3(kBlock), 0,
                               ctx->stream, packed, centres, npad, tiles, per_split, eps2, partials);
        }
        ORR_LAUNCHED("direct_tiles");
    }
    {
        PhaseTimer t(ctx, kDirectReduce);
        const std::uint32_t blocks = static_cast<std::uint32_t>((n + 255) / 256);
        hipLaunchKernelGGL(direct_reduce, dim3(blocks), dim3(256), 0, ctx->stream, partials, splits,
                           npad, n, order, a.gravitational_constant, a.ax, a.ay, a.az);
        ORR_LAUNCHED("direct_reduce");
    }
}

}  // namespace orr
