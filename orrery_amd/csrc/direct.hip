// Direct O(N^2) softened gravity on gfx950.
//
// Replaces the reference's CUDA direct kernel (src/solvers/cuda_direct_kernel.cu:87-214),
// which gives one thread one target and stages j-tiles of positions through
// shared memory with two barriers per tile. Here:
//
//   * Sources are packed once per evaluation into 64-particle tiles: a float4
//     {ox, oy, oz, m} per particle, the offset o taken in FP64 from the tile's
//     FP64 centre (midpoint of its bounding box) and rounded to FP32, plus the
//     rounding residual o_lo = (x - c) - o as a second float4. In Morton order
//     a tile is spatially compact, so o carries the precision of the tile's
//     size rather than of the system's.
//   * A wave reads a source tile through the scalar unit: the address is
//     wave-uniform, so each {ox, oy, oz, m} arrives in SGPRs and is broadcast
//     to all 64 lanes as an instruction operand. No LDS, no barriers.
//   * Each lane owns R targets. Per target and source tile the target is
//     re-expressed against the source tile's centre as one FP64 difference
//     rounded once; far tiles use the separation o_j - t_hi (one rounding of
//     the separation), 12 VALU operations and one v_rsq_f32 of (r^2+eps^2)^3
//     per interaction.
//   * Near tiles -- boxes whose gap is at most a tenth of their summed
//     half-widths (the mask: one bit per tile pair, written once per
//     evaluation) -- use the double-float separation (o_j - t) + (o_lo_j -
//     t_lo). Pairs among them closer than a quarter of the wider tile's
//     half-width are marked and evaluated afterwards in FP64 (close_pair.hpp):
//     their single term can be many times the net force on either particle.
//   * Far tiles' terms are summed in FP32 in runs of 16, near tiles' in runs
//     of 8, each run Kahan-added into the target's running sum.
//   * The source range is split over gridDim.y; a fixed-order FP64 reduction
//     kernel folds the splits, multiplies by G and scatters to particle order.
//   * By default (direct_pairs, below) each far pair of tiles is evaluated
//     once and applied to both tiles, with the same packing, near test, close
//     pairs and compensation; the one-sided kernel above then covers what the
//     pairs path does not take (configured off, too few tiles, or slots beyond
//     16 GiB).
//
// The reference counts 20 flop per interaction (docs/performance.md:103-109).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "close_pair.hpp"
#include "engine.hpp"

namespace orr {
namespace {

constexpr int kTile = 64;    // sources per tile = one wave of the prepare kernel
constexpr int kBlock = 256;  // four waves
constexpr float kNearGap = 0.1f;

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

// One wave packs one tile: centre its (valid) particles, split each offset
// into an FP32 value and its FP32 residual. centres[tile] = {cx, cy, cz,
// largest half-extent}; extents[tile] = {hx, hy, hz, w}.
__device__ __forceinline__ void pack_tile(double px, double py, double pz, double pm, bool valid,
                                          std::uint64_t k, std::uint32_t tile,
                                          float4* __restrict__ packed,
                                          float4* __restrict__ packed_lo,
                                          double4* __restrict__ centres,
                                          float4* __restrict__ extents,
                                          float2* __restrict__ mass_range = nullptr) {
    const double inf = __builtin_huge_val();
    const double lx = wave_min(valid ? px : inf), hx = wave_max(valid ? px : -inf);
    const double ly = wave_min(valid ? py : inf), hy = wave_max(valid ? py : -inf);
    const double lz = wave_min(valid ? pz : inf), hz = wave_max(valid ? pz : -inf);
    const double cx = 0.5 * (lx + hx), cy = 0.5 * (ly + hy), cz = 0.5 * (lz + hz);
    const double w = 0.5 * fmax(hx - lx, fmax(hy - ly, hz - lz));
    if (valid) {
        const double dx = px - cx, dy = py - cy, dz = pz - cz;
        const float ox = static_cast<float>(dx), oy = static_cast<float>(dy), oz = static_cast<float>(dz);
        packed[k] = make_float4(ox, oy, oz, static_cast<float>(pm));
        packed_lo[k] = make_float4(static_cast<float>(dx - ox), static_cast<float>(dy - oy),
                                   static_cast<float>(dz - oz), 0.f);
    } else {
        packed[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        packed_lo[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (mass_range != nullptr) {
        // The tile's smallest and largest packed mass, its massless pads
        // included: a system with pads never counts as one of equal masses.
        const double fm = valid ? static_cast<double>(static_cast<float>(pm)) : 0.0;
        const double lo = wave_min(fm), hi = wave_max(fm);
        if ((threadIdx.x & (kTile - 1)) == 0) {
            mass_range[tile] = make_float2(static_cast<float>(lo), static_cast<float>(hi));
        }
    }
    if ((threadIdx.x & (kTile - 1)) == 0) {
        centres[tile] = make_double4(cx, cy, cz, w);
        extents[tile] = make_float4(static_cast<float>(0.5 * (hx - lx)), static_cast<float>(0.5 * (hy - ly)),
                                    static_cast<float>(0.5 * (hz - lz)), static_cast<float>(w));
    }
}

// One wave per tile of one system: gather (optionally through `order`) and pack.
__global__ __launch_bounds__(64) void direct_prepare(const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ m,
                                                     const std::uint32_t* __restrict__ order,
                                                     std::uint64_t n, float4* __restrict__ packed,
                                                     float4* __restrict__ packed_lo,
                                                     double4* __restrict__ centres,
                                                     float4* __restrict__ extents,
                                                     float2* __restrict__ mass_range) {
    const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * kTile + threadIdx.x;
    const bool valid = k < n;
    const std::uint64_t src = valid ? (order != nullptr ? order[k] : k) : 0;
    pack_tile(valid ? x[src] : 0.0, valid ? y[src] : 0.0, valid ? z[src] : 0.0,
              valid ? m[src] : 0.0, valid, k, blockIdx.x, packed, packed_lo, centres, extents,
              mass_range);
}

// uniform[0] = the common packed mass when every particle has the same
// positive mass, else 0 (one block over the tiles' ranges).
__global__ __launch_bounds__(256) void mass_uniform(const float2* __restrict__ mass_range,
                                                    std::uint32_t tiles, int enabled,
                                                    float* __restrict__ uniform) {
    __shared__ float lo_s[256], hi_s[256];
    float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
    for (std::uint32_t t = threadIdx.x; t < tiles; t += blockDim.x) {
        const float2 r = mass_range[t];
        lo = fminf(lo, r.x);
        hi = fmaxf(hi, r.y);
    }
    lo_s[threadIdx.x] = lo;
    hi_s[threadIdx.x] = hi;
    __syncthreads();
    for (unsigned half = 128; half > 0; half >>= 1) {
        if (threadIdx.x < half) {
            lo_s[threadIdx.x] = fminf(lo_s[threadIdx.x], lo_s[threadIdx.x + half]);
            hi_s[threadIdx.x] = fmaxf(hi_s[threadIdx.x], hi_s[threadIdx.x + half]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) uniform[0] = (enabled && lo_s[0] == hi_s[0] && lo_s[0] > 0.f) ? lo_s[0] : 0.f;
}

// One pairwise term accumulated into (px, py, pz). kSafe masks r^2 == 0 (zero
// softening): the self term and coincident pads then contribute m * finite * 0
// = 0 instead of NaN, as the reference's CUDA kernel masks self and tail terms
// (cuda_direct_kernel.cu:134-159). kRefine applies one Newton-Raphson step to
// the hardware reciprocal square root y0 before cubing: with e = 1 - x y0^2
// (one rounding, via FMA), y^3 = y0^3 (1 + 3e/2) to second order in e.
template<bool kSafe, bool kRefine>
__device__ __forceinline__ float inv_cubed(float r2) {
    if constexpr (kSafe) r2 = r2 > 0.f ? r2 : 1.f;
    if constexpr (!kSafe && !kRefine) {
        return __builtin_amdgcn_rsqf((r2 * r2) * r2);
    } else {
        const float inv = __builtin_amdgcn_rsqf(r2);
        float inv3 = (inv * inv) * inv;
        if constexpr (kRefine) {
            const float e = __builtin_fmaf(-(r2 * inv), inv, 1.f);
            inv3 = inv3 * __builtin_fmaf(1.5f, e, 1.f);
        }
        return inv3;
    }
}

template<bool kSafe, bool kRefine>
__device__ __forceinline__ void term(float dx, float dy, float dz, float m, float eps2, float& px,
                                     float& py, float& pz) {
    float r2 = __builtin_fmaf(dx, dx, eps2);
    r2 = __builtin_fmaf(dy, dy, r2);
    r2 = __builtin_fmaf(dz, dz, r2);
    const float f = m * inv_cubed<kSafe, kRefine>(r2);
    px = __builtin_fmaf(f, dx, px);
    py = __builtin_fmaf(f, dy, py);
    pz = __builtin_fmaf(f, dz, pz);
}

__device__ __forceinline__ void kahan(float p, float& s, float& c) {
    const float y = p - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Near: the gap between the two tiles' boxes (centre c, half-extents e; the
// source tile's C, E) is at most kNearGap of their summed half-widths. Outside
// it every separation exceeds that gap, so the FP32 offsets round it to
// O(epsilon) relative error. One definition, without contraction, so the mask
// pass and the kernels classify every tile pair alike.
__device__ __forceinline__ bool near_tiles(const double4& C, const float4& E, double cx, double cy,
                                           double cz, const float4& e) {
#pragma clang fp contract(off)
    const float gx = fmaxf(0.f, fabsf(static_cast<float>(C.x - cx)) - e.x - E.x);
    const float gy = fmaxf(0.f, fabsf(static_cast<float>(C.y - cy)) - e.y - E.y);
    const float gz = fmaxf(0.f, fabsf(static_cast<float>(C.z - cz)) - e.z - E.z);
    const float reach = kNearGap * (e.w + E.w);
    return gx * gx + gy * gy + gz * gz <= reach * reach;
}

// Close pairs: closer than kClose times the wider tile's largest half-extent.
// Their terms are evaluated in FP64 (close_pair.hpp); about
// one step in twenty of a near pass has such a pair on some lane.
constexpr float kClose = 0.25f;

// The FP32 r^2 below which a near pair is close (r^2 includes eps^2).
__device__ __forceinline__ float close_limit(double w_target, double w_source, float eps2) {
    const float rc = kClose * static_cast<float>(fmax(w_target, w_source));
    return __builtin_fmaf(rc, rc, eps2);
}

// Adds the double-float value (h, l) to the Kahan pair (s, c) (value s - c).
__device__ __forceinline__ void kahan_df(float h, float l, float& s, float& c) {
    kahan(h, s, c);
    c = c - l;
}

// Targets' sums against one 64-source tile J.
//
//   * Far tiles: each target is re-expressed against the source tile's centre
//     with one rounding (the FP64 difference of its (hi, lo) offset and the
//     centres, rounded once); the terms are summed plainly in FP32 in runs of
//     kRun and each run is Kahan-added per component into (s, k).
//   * Near tiles: double-float separations, runs of kNearRun terms summed in
//     FP32 and Kahan-added into the tile's own pair, close pairs evaluated in
//     FP64 and added to it as (hi, lo); the tile's pair then joins (s, k).
//
// kUnroll sets how many far-path sources are in flight (their SGPR loads
// included).
constexpr int kRun = 16;
// Terms of a near tile summed plainly in FP32 before each Kahan addition.
constexpr int kNearRun = 8;

template<int R, bool kSafe, bool kRefine, bool kGivenNear = false, int kUnroll = 16>
__device__ __forceinline__ void source_tile(
    const float4* __restrict__ packed, const float4* __restrict__ packed_lo,
    const double4* __restrict__ centres, const float4* __restrict__ extents, std::uint32_t J,
    bool given_near, float eps2, const float (&ox)[R], const float (&oy)[R], const float (&oz)[R],
    const float (&lx)[R], const float (&ly)[R], const float (&lz)[R], const double (&cx)[R],
    const double (&cy)[R], const double (&cz)[R], const float4 (&ce)[R], float (&sx)[R],
    float (&sy)[R], float (&sz)[R], float (&kx)[R], float (&ky)[R], float (&kz)[R]) {
    const double4 C = centres[J];
    const float4 E = extents[J];
    const float4* __restrict__ src = packed + static_cast<std::size_t>(J) * kTile;
    // All lanes of a wave share one target tile per r, so the branch is
    // uniform. With kGivenNear the caller read the pair's flag from the mask.
    bool near = false;
    if constexpr (kGivenNear) {
        near = given_near;
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) near = near || near_tiles(C, E, cx[r], cy[r], cz[r], ce[r]);
    }
    // Each target's position relative to the source tile's centre: FP64 from
    // its (hi, lo) offset, split into (hi, lo) floats.
    float th[R][3], tl[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double t[3] = {(static_cast<double>(ox[r]) + lx[r]) - (C.x - cx[r]),
                             (static_cast<double>(oy[r]) + ly[r]) - (C.y - cy[r]),
                             (static_cast<double>(oz[r]) + lz[r]) - (C.z - cz[r])};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            th[r][c] = static_cast<float>(t[c]);
            tl[r][c] = static_cast<float>(t[c] - th[r][c]);
        }
    }
    if (!near) {
#pragma unroll
        for (int run = 0; run < kTile / kRun; ++run) {
            float px[R], py[R], pz[R];
#pragma unroll
            for (int r = 0; r < R; ++r) px[r] = py[r] = pz[r] = 0.f;
#pragma unroll (kUnroll < kRun ? kUnroll : kRun)
            for (int j = run * kRun; j < (run + 1) * kRun; ++j) {
                const float4 q = src[j];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    term<kSafe, kRefine>(q.x - th[r][0], q.y - th[r][1], q.z - th[r][2], q.w, eps2,
                                         px[r], py[r], pz[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                kahan(px[r], sx[r], kx[r]);
                kahan(py[r], sy[r], ky[r]);
                kahan(pz[r], sz[r], kz[r]);
            }
        }
        return;
    }
    const float4* __restrict__ src_lo = packed_lo + static_cast<std::size_t>(J) * kTile;
    float px[R], py[R], pz[R], pkx[R], pky[R], pkz[R], lim[R];
    std::uint64_t close[R];  // per lane: bit j set when source j is a close pair
#pragma unroll
    for (int r = 0; r < R; ++r) {
        px[r] = py[r] = pz[r] = pkx[r] = pky[r] = pkz[r] = 0.f;
        lim[r] = close_limit(ce[r].w, E.w, eps2);
        close[r] = 0;
    }
#pragma unroll 1
    for (int j0 = 0; j0 < kTile; j0 += kNearRun) {
        float qx[R], qy[R], qz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) qx[r] = qy[r] = qz[r] = 0.f;
#pragma unroll
        for (int j = j0; j < j0 + kNearRun; ++j) {
            const float4 q = src[j];
            const float4 ql = src_lo[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = (q.x - th[r][0]) + (ql.x - tl[r][0]);
                const float dy = (q.y - th[r][1]) + (ql.y - tl[r][1]);
                const float dz = (q.z - th[r][2]) + (ql.z - tl[r][2]);
                float r2 = __builtin_fmaf(dx, dx, eps2);
                r2 = __builtin_fmaf(dy, dy, r2);
                r2 = __builtin_fmaf(dz, dz, r2);
                // A close pair's FP32 term is left out here and added in
                // FP64 below.
                const bool near_pair = r2 < lim[r] && r2 > 0.f;
                close[r] |= near_pair ? std::uint64_t{1} << j : std::uint64_t{0};
                const float f = near_pair ? 0.f : q.w * inv_cubed<kSafe, kRefine>(r2);
                qx[r] = __builtin_fmaf(f, dx, qx[r]);
                qy[r] = __builtin_fmaf(f, dy, qy[r]);
                qz[r] = __builtin_fmaf(f, dz, qz[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            kahan(qx[r], px[r], pkx[r]);
            kahan(qy[r], py[r], pky[r]);
            kahan(qz[r], pz[r], pkz[r]);
        }
    }
    // The close pairs, each lane over its own (few): sources read by index.
#pragma unroll
    for (int r = 0; r < R; ++r) {
        std::uint64_t bits = close[r];
        while (bits != 0) {
            const int j = __builtin_ctzll(bits);
            bits &= bits - 1;
            const float4 q = src[j];
            const float4 ql = src_lo[j];
            const float bh[3] = {q.x, q.y, q.z}, bl[3] = {ql.x, ql.y, ql.z};
            double G[3];
            close_pair(th[r], tl[r], bh, bl, eps2, G);
            const double mq = q.w;
            float h, l;
            split(mq * G[0], h, l);
            kahan_df(h, l, px[r], pkx[r]);
            split(mq * G[1], h, l);
            kahan_df(h, l, py[r], pky[r]);
            split(mq * G[2], h, l);
            kahan_df(h, l, pz[r], pkz[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        kahan_df(px[r], -pkx[r], sx[r], kx[r]);
        kahan_df(py[r], -pky[r], sy[r], ky[r]);
        kahan_df(pz[r], -pkz[r], sz[r], kz[r]);
    }
}

// The near flags of every (target tile, source tile) pair, one bit each, rows
// of ceil(tiles / 32) words: computed once per evaluation so that the kernels
// read one word per 32 source tiles through the scalar unit instead of each
// wave repeating the (wave-uniform) box test on its vector lanes.
constexpr int kMaskRows = 8;  // target tiles per block of the mask pass

__global__ __launch_bounds__(256) void near_mask(const double4* __restrict__ centres,
                                                 const float4* __restrict__ extents,
                                                 std::uint32_t tiles, std::uint32_t words,
                                                 std::uint32_t* __restrict__ mask) {
    // kMaskRows target tiles per block (wave-uniform: scalar loads); each
    // thread loads one source tile per chunk and tests it against all of them;
    // a wave's 64 results become two words by ballot.
    const std::uint32_t row0 = blockIdx.x * kMaskRows;
    const int lane = threadIdx.x & 63;
    for (std::uint32_t chunk = 0; chunk < tiles; chunk += blockDim.x) {
        const std::uint32_t J = chunk + threadIdx.x;
        const bool valid = J < tiles;
        const double4 C = valid ? centres[J] : make_double4(0, 0, 0, 0);
        const float4 E = valid ? extents[J] : make_float4(0, 0, 0, 0);
        const std::uint32_t first_word = (chunk + (threadIdx.x & ~63u)) / 32;
#pragma unroll
        for (int r = 0; r < kMaskRows; ++r) {
            const std::uint32_t row = row0 + r;
            if (row >= tiles) break;
            const double4 c = centres[row];
            const float4 e = extents[row];
            const std::uint64_t bits = __ballot(valid && near_tiles(C, E, c.x, c.y, c.z, e));
            if (lane == 0) {
                std::uint32_t* out = mask + static_cast<std::size_t>(row) * words;
                if (first_word < words) out[first_word] = static_cast<std::uint32_t>(bits);
                if (first_word + 1 < words) out[first_word + 1] = static_cast<std::uint32_t>(bits >> 32);
            }
        }
    }
}

template<int R, bool kSafe, bool kRefine, bool kMask>
__global__ __launch_bounds__(kBlock) void direct_tiles(const float4* __restrict__ packed,
                                                       const float4* __restrict__ packed_lo,
                                                       const double4* __restrict__ centres,
                                                       const float4* __restrict__ extents,
                                                       const std::uint32_t* __restrict__ mask,
                                                       std::uint32_t npad, std::uint32_t tiles,
                                                       std::uint32_t tiles_per_split, float eps2,
                                                       float* __restrict__ partial) {
    const std::uint32_t split = blockIdx.y;
    const std::uint32_t j0 = split * tiles_per_split;
    const std::uint32_t j1 = min(tiles, j0 + tiles_per_split);
    const std::uint32_t base = blockIdx.x * (kBlock * R) + threadIdx.x;

    float ox[R], oy[R], oz[R], lx[R], ly[R], lz[R];
    double cx[R], cy[R], cz[R];
    float4 ce[R];
    float sx[R], sy[R], sz[R], kx[R], ky[R], kz[R];
    std::uint32_t row[R];  // first mask word of each target tile's row
    const std::uint32_t words = (tiles + 31) / 32;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t k = min(base + r * kBlock, npad - 1);
        const float4 o = packed[k];
        const float4 l = packed_lo[k];
        // A wave's 64 targets are one whole tile (npad is a multiple of kTile,
        // and a wave past the end clamps every lane to the last tile), so the
        // tile's centre and extents are wave-uniform: scalar loads, SGPRs.
        const std::uint32_t tile = __builtin_amdgcn_readfirstlane(k / kTile);
        const double4 c = centres[tile];
        ox[r] = o.x;
        oy[r] = o.y;
        oz[r] = o.z;
        lx[r] = l.x;
        ly[r] = l.y;
        lz[r] = l.z;
        cx[r] = c.x;
        cy[r] = c.y;
        cz[r] = c.z;
        ce[r] = extents[tile];
        row[r] = tile * words;
        sx[r] = sy[r] = sz[r] = 0.f;
        kx[r] = ky[r] = kz[r] = 0.f;
    }

    std::uint32_t bits = 0;
    for (std::uint32_t J = j0; J < j1; ++J) {
        bool near = false;
        if constexpr (kMask) {
            if (J == j0 || (J & 31u) == 0) {
                bits = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) bits |= mask[row[r] + J / 32];
            }
            near = ((bits >> (J & 31u)) & 1u) != 0;
        }
        source_tile<R, kSafe, kRefine, kMask>(packed, packed_lo, centres, extents, J, near, eps2,
                                              ox, oy, oz, lx, ly, lz, cx, cy, cz, ce, sx, sy, sz, kx,
                                              ky, kz);
    }

    // Component c of target k at partial[(6 split + c) npad + k], in 64-bit
    // offsets (6 npad exceeds 32 bits from npad = 2^30 / 1.5).
    const std::size_t np = npad;
    float* out = partial + static_cast<std::size_t>(split) * 6 * np;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const std::uint32_t k = base + r * kBlock;
        if (k < npad) {
            out[k] = sx[r];
            out[np + k] = sy[r];
            out[2 * np + k] = sz[r];
            out[3 * np + k] = kx[r];
            out[4 * np + k] = ky[r];
            out[5 * np + k] = kz[r];
        }
    }
}

// Fixed-order FP64 fold of the splits; sum - compensation per split.
// The fold's output: a, and the fused kick's v += s a when there is one.
__device__ __forceinline__ void store_field(std::uint64_t dst, double gx, double gy, double gz,
                                            double* __restrict__ ax, double* __restrict__ ay,
                                            double* __restrict__ az, const FusedKick& kick) {
    ax[dst] = gx;
    ay[dst] = gy;
    az[dst] = gz;
    if (kick.vx != nullptr) {
        kick.vx[dst] = kick_update(kick.vx[dst], kick.s, gx);
        kick.vy[dst] = kick_update(kick.vy[dst], kick.s, gy);
        kick.vz[dst] = kick_update(kick.vz[dst], kick.s, gz);
    }
}

__global__ __launch_bounds__(256) void direct_reduce(const float* __restrict__ partial,
                                                     std::uint32_t splits, std::uint32_t npad,
                                                     std::uint64_t n,
                                                     const std::uint32_t* __restrict__ order,
                                                     double G, double* __restrict__ ax,
                                                     double* __restrict__ ay,
                                                     double* __restrict__ az, FusedKick kick) {
    const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (std::uint32_t s = 0; s < splits; ++s) {
        const std::size_t np = npad;
        const float* p = partial + static_cast<std::size_t>(s) * 6 * np;
        a0 += static_cast<double>(p[k]) - static_cast<double>(p[3 * np + k]);
        a1 += static_cast<double>(p[np + k]) - static_cast<double>(p[4 * np + k]);
        a2 += static_cast<double>(p[2 * np + k]) - static_cast<double>(p[5 * np + k]);
    }
    const std::uint64_t dst = order != nullptr ? order[k] : k;
    store_field(dst, G * a0, G * a1, G * a2, ax, ay, az, kick);
}

// ---- batches: many independent systems in one launch ---------------------
//
// Each system of a batch owns whole tiles (its last tile padded with massless
// lanes), so a tile never mixes systems. tile[t] = {first particle, valid
// count, system, first partial slot}. A work item {target tile, first source
// tile, end source tile, partial slot} is one wave's job; the split of a
// system's source range depends on that system's size only, so its result does
// not depend on which other systems share the batch.

// One wave per tile of the batch.
__global__ __launch_bounds__(64) void batch_prepare(const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    const double* __restrict__ m,
                                                    const std::uint32_t* __restrict__ order,
                                                    const uint4* __restrict__ tile,
                                                    float4* __restrict__ packed,
                                                    float4* __restrict__ packed_lo,
                                                    double4* __restrict__ centres,
                                                    float4* __restrict__ extents) {
    const uint4 T = tile[blockIdx.x];
    const bool valid = threadIdx.x < T.y;
    const std::uint64_t k = static_cast<std::uint64_t>(T.x) + threadIdx.x;
    const std::uint64_t src = valid ? (order != nullptr ? order[k] : k) : 0;
    pack_tile(valid ? x[src] : 0.0, valid ? y[src] : 0.0, valid ? z[src] : 0.0,
              valid ? m[src] : 0.0, valid,
              static_cast<std::uint64_t>(blockIdx.x) * kTile + threadIdx.x, blockIdx.x, packed,
              packed_lo, centres, extents);
}

// Persistent: the grid's W waves take the item list (sorted longest first) in
// strides -- wave w runs items w, w + W, w + 2W, ... -- so every wave starts on a
// long item and finishes on short ones, with no shared counter to contend on.
// Each item writes its own partial slot: the result does not depend on which
// wave ran it.
constexpr int kItemUnroll = 8;

// Eight waves per SIMD: without the hint the scheduler spends registers on
// instruction-level parallelism (87 VGPRs, SGPRs spilled to VGPR lanes).
template<bool kSafe, bool kRefine>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void direct_items(const float4* __restrict__ packed,
                                                       const float4* __restrict__ packed_lo,
                                                       const double4* __restrict__ centres,
                                                       const float4* __restrict__ extents,
                                                       const uint4* __restrict__ tile,
                                                       const uint4* __restrict__ item,
                                                       std::uint32_t items,
                                                       const float* __restrict__ eps2_of,
                                                       float* __restrict__ partial) {
    const unsigned lane = threadIdx.x & (kTile - 1);
    const std::uint32_t waves = gridDim.x * (kBlock / kTile);
    const std::uint32_t first = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kTile) + threadIdx.x / kTile);
    for (std::uint32_t next = first; next < items; next += waves) {
        const uint4 it = item[next];
        const float eps2 = eps2_of[tile[it.x].z];
        const std::uint32_t k = it.x * kTile + lane;
        const float4 o = packed[k];
        const float4 l = packed_lo[k];
        const double4 c = centres[it.x];
        const float ox[1] = {o.x}, oy[1] = {o.y}, oz[1] = {o.z};
        const float lx[1] = {l.x}, ly[1] = {l.y}, lz[1] = {l.z};
        const double cx[1] = {c.x}, cy[1] = {c.y}, cz[1] = {c.z};
        const float4 ce[1] = {extents[it.x]};
        float sx[1] = {0.f}, sy[1] = {0.f}, sz[1] = {0.f}, kx[1] = {0.f}, ky[1] = {0.f},
              kz[1] = {0.f};
        for (std::uint32_t J = it.y; J < it.z; ++J)
            source_tile<1, kSafe, kRefine, false, kItemUnroll>(packed, packed_lo, centres, extents, J, false, eps2,
                                                        ox, oy, oz, lx, ly, lz, cx, cy, cz, ce, sx,
                                                        sy, sz, kx, ky, kz);
        float* out = partial + static_cast<std::size_t>(it.w) * 6 * kTile;
        out[lane] = sx[0];
        out[kTile + lane] = sy[0];
        out[2 * kTile + lane] = sz[0];
        out[3 * kTile + lane] = kx[0];
        out[4 * kTile + lane] = ky[0];
        out[5 * kTile + lane] = kz[0];
    }
}

// One thread per particle of the batch (sorted position k): fold its tile's
// slots in split order in FP64, multiply by the system's G, scatter through
// `order`. sys[s] = {first particle, count, first tile, splits}.
__global__ __launch_bounds__(256) void batch_reduce(const float* __restrict__ partial,
                                                    const uint4* __restrict__ sys,
                                                    const uint4* __restrict__ tile,
                                                    const std::uint32_t* __restrict__ seg_of,
                                                    const double* __restrict__ G_of,
                                                    std::uint64_t n,
                                                    const std::uint32_t* __restrict__ order,
                                                    double* __restrict__ ax,
                                                    double* __restrict__ ay,
                                                    double* __restrict__ az) {
    const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const std::uint32_t s = seg_of[k];
    const uint4 S = sys[s];
    const std::uint32_t local = static_cast<std::uint32_t>(k) - S.x;
    const std::uint32_t lane = local & (kTile - 1);
    const std::uint32_t first_slot = tile[S.z + local / kTile].w;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (std::uint32_t i = 0; i < S.w; ++i) {
        const float* p = partial + static_cast<std::size_t>(first_slot + i) * 6 * kTile + lane;
        a0 += static_cast<double>(p[0]) - static_cast<double>(p[3 * kTile]);
        a1 += static_cast<double>(p[kTile]) - static_cast<double>(p[4 * kTile]);
        a2 += static_cast<double>(p[2 * kTile]) - static_cast<double>(p[5 * kTile]);
    }
    const double G = G_of[s];
    const std::uint64_t dst = order != nullptr ? order[k] : k;
    ax[dst] = G * a0;
    ay[dst] = G * a1;
    az[dst] = G * a2;
}

// ---- pairs: each far tile pair once, applied to both tiles ---------------
//
// The one-sided kernel above spends 12 VALU operations and one v_rsq_f32 on
// every interaction. Gravity is symmetric: the term that pulls target i toward
// source j, scaled by m_i / m_j, pulls j toward i. Evaluating each tile pair
// once for both sides costs 16 VALU and one v_rsq_f32 per PAIR, i.e. 8 VALU and
// half a v_rsq_f32 per interaction:
//
//   * Tiles are grouped into blocks of kStay. A wave owns one block: each lane
//     holds one particle of each of its kStay tiles (the stationary side),
//     with their sums and Kahan compensations in registers.
//   * The wave streams the tiles I before its block (a chunk of them per
//     wave). Lane l loads particle l of I; after every step the streamed
//     particles and their force accumulators move one lane around the wave
//     (DPP wave rotate), so in 64 steps every stationary particle meets every
//     streamed one, and the accumulator of each streamed particle comes home
//     holding its sum over the block's 64 kStay particles.
//   * Separations are FP32 offsets from the tile centres, the stationary
//     offset re-expressed against the streamed tile's centre once per tile
//     (as the one-sided kernel re-expresses its target). Near tile pairs
//     (the mask) are skipped in that pass and get a pass of their own with the
//     one-sided kernel's double-float separation. The pairs inside a block are
//     left to the complement kernel, one-sided.
//   * Stationary side: the 64 terms of one streamed tile summed in FP32, then
//     Kahan-added (compensations in LDS), per item written to its own slot.
//     Streamed side: per step the kStay terms summed, then added to the
//     travelling accumulator, which is flushed every kFlush steps into the
//     particle's FP64 sum in LDS; after the block's 64 steps (and the near
//     passes, which flush the same way) the sums are written per (particle,
//     block) to their own FP64 slots. A fixed-order FP64 fold (pairs_reduce)
//     sums a particle's slots.
//   * Close pairs of the near passes are evaluated in FP64 (close_pair.hpp)
//     and applied to both sides.
//
// Items: block Q >= 1 streams tiles [0, Q kStay) in chunks of S tiles, S a
// divisor of kStay (kStay Q / S chunks) or a multiple of it (s = S / kStay:
// ceil(Q / s) chunks); items are numbered block by block, items_before(Q) of
// them before block Q.
constexpr int kStay = 8;
// Steps between flushes of the streamed side's travelling sums (to LDS).
constexpr int kFlush = 8;

__host__ __device__ inline std::uint32_t pair_chunks(std::uint32_t Q, std::uint32_t S) {
    return (kStay * Q + S - 1) / S;
}

__host__ __device__ inline std::uint64_t pair_items_before(std::uint32_t Q, std::uint32_t S) {
    if (Q == 0) return 0;
    const std::uint64_t q = Q;
    if (S < kStay) return (kStay / S) * q * (q - 1) / 2;
    // sum over q' < Q of ceil(q' / s): q' = a s + r
    const std::uint64_t s = S / kStay, m = q - 1, a = m / s, r = m % s;
    return s * a * (a + 1) / 2 + (a + 1) * r;
}

__host__ __device__ inline std::uint64_t pair_slot(std::uint32_t Q, std::uint32_t tile) {
    // Block Q's streamed tiles [0, Q kStay) after those of blocks 1 .. Q - 1.
    return static_cast<std::uint64_t>(kStay) * Q * (Q - 1) / 2 + tile;
}

// Every lane takes the value of its neighbour, around the whole wave (DPP
// wave_ror:1); 64 rotations bring every value home. wave_rotate moves in
// place; rotate_add(a, t) = rotate(a) + t folds the move into the add.
__device__ __forceinline__ float wave_rotate(float v) {
    const int u = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(u, u, 0x13C, 0xf, 0xf, false));
}
__device__ __forceinline__ float rotate_add(float a, float t) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, a), 0x13C,
                                                                 0xf, 0xf, false)) + t;
}

// Four waves per SIMD (128 VGPRs): the stationary side holds 13 values per
// tile per lane (offset re-expressed against the streamed tile, mass, the
// streamed tile's partial sums, Kahan sums and compensations), 104 in all.
//
// kEqual: every particle has one mass (the flag the mass pass leaves). The
// masses then leave the pair term -- both sides add inv^3 d, two multiplies
// fewer per pair and no travelling mass -- and the fold multiplies the pairs'
// slots by the common mass.
template<bool kSafe, bool kRefine, bool kEqual>
__device__ __forceinline__ void pairs_item(const float4* __restrict__ packed,
                                           const float4* __restrict__ packed_lo,
                                           const double4* __restrict__ centres,
                                           const std::uint32_t* __restrict__ mask,
                                           std::uint32_t tiles, std::uint32_t words,
                                           std::uint32_t blocks, std::uint32_t S,
                                           std::uint32_t item, float eps2,
                                           float* __restrict__ jpart,
                                           double* __restrict__ ipart,
                                           double* __restrict__ home,
                                           float* __restrict__ comp) {
    // The item's block: the largest Q with pair_items_before(Q) <= item.
    std::uint32_t lo = 1, hi = blocks - 1;
    while (lo < hi) {
        const std::uint32_t mid = (lo + hi + 1) / 2;
        if (pair_items_before(mid, S) <= item) lo = mid;
        else hi = mid - 1;
    }
    const std::uint32_t Q = lo;
    const std::uint32_t chunk = item - static_cast<std::uint32_t>(pair_items_before(Q, S));
    const std::uint32_t t0 = Q * kStay;
    const std::uint32_t i0 = chunk * S;
    const std::uint32_t i1 = min(i0 + S, t0);
    const unsigned lane = threadIdx.x & (kTile - 1);

    // Stationary tiles past the last (a partial last block) repeat the last
    // tile massless: they pull nothing and their sums are never read.
    // home: the wave's streamed-side sums by particle, in FP64 (LDS); the
    // travelling accumulators are added there every kFlush steps (each lane to
    // the slot of the particle it holds then), so no FP32 sum runs longer than
    // kFlush steps of kStay terms. comp: this lane's stationary-side Kahan
    // compensations (LDS, [(3 r + c) kTile]), touched once per streamed tile,
    // so that the far loop keeps its registers.

    float m[kStay], sx[kStay], sy[kStay], sz[kStay];
#pragma unroll
    for (int r = 0; r < kStay; ++r) {
        if constexpr (kEqual) {
            m[r] = 1.f;
        } else {
            m[r] = t0 + r < tiles ? packed[static_cast<std::size_t>(min(t0 + r, tiles - 1)) * kTile + lane].w : 0.f;
        }
        sx[r] = sy[r] = sz[r] = 0.f;
        comp[(3 * r) * kTile] = comp[(3 * r + 1) * kTile] = comp[(3 * r + 2) * kTile] = 0.f;
    }
    // Adds (h, l) to stationary sum (s, compensation of component c of tile r).
    auto stay_add = [&](int r, int c, float h, float l, float& sum) {
        float k = comp[(3 * r + c) * kTile];
        kahan_df(h, l, sum, k);
        comp[(3 * r + c) * kTile] = k;
    };

    for (std::uint32_t I = i0; I < i1; ++I) {
        const double4 CI = centres[I];
        const float4 qi = packed[static_cast<std::size_t>(I) * kTile + lane];
        float xi = qi.x, yi = qi.y, zi = qi.z, mi = kEqual ? 1.f : qi.w;
        std::uint32_t near = 0;
        float wx[kStay], wy[kStay], wz[kStay];
        // The stationary offsets and centres are re-read for every streamed
        // tile (cache hits) rather than held: the index is hidden from the
        // compiler, which would otherwise keep 9 kStay more VGPRs live.
        std::uint32_t first = t0;
        asm volatile("" : "+s"(first));
#pragma unroll
        for (int r = 0; r < kStay; ++r) {
            const std::uint32_t tile = min(first + r, tiles - 1);
            near |= ((mask[static_cast<std::size_t>(tile) * words + I / 32] >> (I & 31u)) & 1u) << r;
            const double4 C = centres[tile];
            const float4 o = packed[static_cast<std::size_t>(tile) * kTile + lane];
            // The stationary offset re-expressed relative to the streamed
            // tile's centre: the FP64 difference rounded once.
            wx[r] = static_cast<float>(static_cast<double>(o.x) - (CI.x - C.x));
            wy[r] = static_cast<float>(static_cast<double>(o.y) - (CI.y - C.y));
            wz[r] = static_cast<float>(static_cast<double>(o.z) - (CI.z - C.z));
        }
        float px[kStay], py[kStay], pz[kStay];
#pragma unroll
        for (int r = 0; r < kStay; ++r) px[r] = py[r] = pz[r] = 0.f;
        float ax = 0.f, ay = 0.f, az = 0.f;
        // d = x_i - x_j; j gains m_i d / r^3, i gains -m_j d / r^3.
        auto pair = [&](int r, float& tx, float& ty, float& tz) {
            const float dx = xi - wx[r], dy = yi - wy[r], dz = zi - wz[r];
            float r2 = __builtin_fmaf(dx, dx, eps2);
            r2 = __builtin_fmaf(dy, dy, r2);
            r2 = __builtin_fmaf(dz, dz, r2);
            const float inv3 = inv_cubed<kSafe, kRefine>(r2);
            const float fj = kEqual ? inv3 : mi * inv3, fi = kEqual ? inv3 : m[r] * inv3;
            px[r] = __builtin_fmaf(fj, dx, px[r]);
            py[r] = __builtin_fmaf(fj, dy, py[r]);
            pz[r] = __builtin_fmaf(fj, dz, pz[r]);
            tx = __builtin_fmaf(-fi, dx, tx);
            ty = __builtin_fmaf(-fi, dy, ty);
            tz = __builtin_fmaf(-fi, dz, tz);
        };
        auto advance = [&](float tx, float ty, float tz) {
            ax = rotate_add(ax, tx);
            ay = rotate_add(ay, ty);
            az = rotate_add(az, tz);
            xi = wave_rotate(xi);
            yi = wave_rotate(yi);
            zi = wave_rotate(zi);
            if constexpr (!kEqual) mi = wave_rotate(mi);
        };
        // After step s the accumulator on lane l belongs to streamed particle
        // (l - s) mod 64. Flushes after steps kFlush - 1, 2 kFlush - 1, ...
        auto flush_streamed = [&](std::uint32_t last_step) {
            const unsigned slot = (lane - last_step) & (kTile - 1);
            if (last_step == kFlush - 1) {
                home[slot] = ax;
                home[kTile + slot] = ay;
                home[2 * kTile + slot] = az;
            } else {
                atomicAdd(&home[slot], static_cast<double>(ax));
                atomicAdd(&home[kTile + slot], static_cast<double>(ay));
                atomicAdd(&home[2 * kTile + slot], static_cast<double>(az));
            }
            ax = ay = az = 0.f;
        };
        // Runs of kFlush steps, each followed by a flush of the streamed sums.
#pragma unroll 1
        for (std::uint32_t run = 0; run < kTile; run += kFlush) {
            if (near == 0) {
#pragma unroll 4
                for (int u = 0; u < kFlush; ++u) {
                    float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
                    for (int r = 0; r < kStay; ++r) pair(r, tx, ty, tz);
                    advance(tx, ty, tz);
                }
            } else {
#pragma unroll 4
                for (int u = 0; u < kFlush; ++u) {
                    float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll
                    for (int r = 0; r < kStay; ++r) {
                        if (((near >> r) & 1u) == 0) pair(r, tx, ty, tz);
                    }
                    advance(tx, ty, tz);
                }
            }
            flush_streamed(run + kFlush - 1);
        }
#pragma unroll
        for (int r = 0; r < kStay; ++r) {
            stay_add(r, 0, px[r], 0.f, sx[r]);
            stay_add(r, 1, py[r], 0.f, sy[r]);
            stay_add(r, 2, pz[r], 0.f, sz[r]);
        }
        if (near != 0) {
            // Near stationary tiles, one pass of 64 steps each: the one-sided
            // kernel's double-float separation (the streamed offset (hi, lo)
            // against the stationary position taken relative to the streamed
            // tile's centre in FP64 and split into (hi, lo)); both sides sum
            // runs of kFlush terms, the stationary runs Kahan-added, the
            // streamed runs added to the FP64 slots. Close pairs are marked
            // and evaluated afterwards in FP64.
            const float4 il = packed_lo[static_cast<std::size_t>(I) * kTile + lane];
            float xl = il.x, yl = il.y, zl = il.z;
            // The stationary tiles' centres and offsets are read again here
            // (cache hits), from an index hidden afresh: otherwise the
            // compiler keeps the far pass's FP64 centre differences alive
            // through the far loop, and they spill on every streamed tile.
            std::uint32_t again = t0;
            asm volatile("" : "+s"(again));
#pragma unroll
            for (int r = 0; r < kStay; ++r) {
                if (((near >> r) & 1u) == 0) continue;
                const std::uint32_t tile = min(again + r, tiles - 1);
                const double4 C = centres[tile];
                const float4 o = packed[static_cast<std::size_t>(tile) * kTile + lane];
                const float4 l = packed_lo[static_cast<std::size_t>(tile) * kTile + lane];
                const double t[3] = {(static_cast<double>(o.x) + l.x) - (CI.x - C.x),
                                     (static_cast<double>(o.y) + l.y) - (CI.y - C.y),
                                     (static_cast<double>(o.z) + l.z) - (CI.z - C.z)};
                float th[3], tl[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    th[c] = static_cast<float>(t[c]);
                    tl[c] = static_cast<float>(t[c] - th[c]);
                }
                const float lim = close_limit(C.w, CI.w, eps2);
                const float mr = m[r];
                std::uint64_t close = 0;  // bit s: this lane's pair at step s is close
                float qx = 0.f, qy = 0.f, qz = 0.f;  // stationary run
                ax = ay = az = 0.f;                  // streamed run (travelling)
#pragma unroll 1
                for (std::uint32_t run = 0; run < kTile; run += kFlush) {
#pragma unroll 2
                    for (std::uint32_t u = 0; u < kFlush; ++u) {
                        const std::uint32_t step = run + u;
                        const float dx = (xi - th[0]) + (xl - tl[0]);
                        const float dy = (yi - th[1]) + (yl - tl[1]);
                        const float dz = (zi - th[2]) + (zl - tl[2]);
                        float r2 = __builtin_fmaf(dx, dx, eps2);
                        r2 = __builtin_fmaf(dy, dy, r2);
                        r2 = __builtin_fmaf(dz, dz, r2);
                        const bool near_pair = r2 < lim && r2 > 0.f;
                        close |= near_pair ? std::uint64_t{1} << step : std::uint64_t{0};
                        const float inv3 = near_pair ? 0.f : inv_cubed<kSafe, kRefine>(r2);
                        const float fj = kEqual ? inv3 : mi * inv3, fi = kEqual ? inv3 : mr * inv3;
                        qx = __builtin_fmaf(fj, dx, qx);
                        qy = __builtin_fmaf(fj, dy, qy);
                        qz = __builtin_fmaf(fj, dz, qz);
                        ax = rotate_add(ax, -fi * dx);
                        ay = rotate_add(ay, -fi * dy);
                        az = rotate_add(az, -fi * dz);
                        xi = wave_rotate(xi);
                        yi = wave_rotate(yi);
                        zi = wave_rotate(zi);
                        if constexpr (!kEqual) mi = wave_rotate(mi);
                        xl = wave_rotate(xl);
                        yl = wave_rotate(yl);
                        zl = wave_rotate(zl);
                    }
                    const unsigned slot = (lane - (run + kFlush - 1)) & (kTile - 1);
                    atomicAdd(&home[slot], static_cast<double>(ax));
                    atomicAdd(&home[kTile + slot], static_cast<double>(ay));
                    atomicAdd(&home[2 * kTile + slot], static_cast<double>(az));
                    ax = ay = az = 0.f;
                    stay_add(r, 0, qx, 0.f, sx[r]);
                    stay_add(r, 1, qy, 0.f, sy[r]);
                    stay_add(r, 2, qz, 0.f, sz[r]);
                    qx = qy = qz = 0.f;
                }
                // Close pairs: at step s this lane met streamed particle
                // (lane - s) mod 64, read back here by index.
                while (close != 0) {
                    const unsigned step = __builtin_ctzll(close);
                    close &= close - 1;
                    const unsigned p = (lane - step) & (kTile - 1);
                    const std::size_t k = static_cast<std::size_t>(I) * kTile + p;
                    const float4 b = packed[k];
                    const float4 bl4 = packed_lo[k];
                    const float bh[3] = {b.x, b.y, b.z}, bl[3] = {bl4.x, bl4.y, bl4.z};
                    double G[3];
                    close_pair(th, tl, bh, bl, eps2, G);
                    const double mb = kEqual ? 1.0 : static_cast<double>(b.w);
                    const double ms = static_cast<double>(mr);
                    float h, l;
                    split(mb * G[0], h, l);
                    stay_add(r, 0, h, l, sx[r]);
                    split(mb * G[1], h, l);
                    stay_add(r, 1, h, l, sy[r]);
                    split(mb * G[2], h, l);
                    stay_add(r, 2, h, l, sz[r]);
                    atomicAdd(&home[p], -(ms * G[0]));
                    atomicAdd(&home[kTile + p], -(ms * G[1]));
                    atomicAdd(&home[2 * kTile + p], -(ms * G[2]));
                }
            }
        }
        // The streamed particles are home: each lane's sum over the block is in
        // its FP64 slot.
        double* out = ipart + pair_slot(Q, I) * 3 * kTile;
        out[lane] = home[lane];
        out[kTile + lane] = home[kTile + lane];
        out[2 * kTile + lane] = home[2 * kTile + lane];
    }
    float* out = jpart + static_cast<std::size_t>(item) * 6 * kStay * kTile;
    constexpr int kRow = kStay * kTile;
#pragma unroll
    for (int r = 0; r < kStay; ++r) {
        const unsigned k = r * kTile + lane;
        out[k] = sx[r];
        out[kRow + k] = sy[r];
        out[2 * kRow + k] = sz[r];
        out[3 * kRow + k] = comp[(3 * r) * kTile];
        out[4 * kRow + k] = comp[(3 * r + 1) * kTile];
        out[5 * kRow + k] = comp[(3 * r + 2) * kTile];
    }
}

// One launch per mass case, both enqueued; the mass pass's flag (uniform[0] >
// 0: every particle's mass) lets one of them run and the other exit at once,
// so the host never waits for the flag and each case keeps its own registers.
template<bool kSafe, bool kRefine, bool kEqual>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void direct_pairs(const float4* __restrict__ packed,
                                                       const float4* __restrict__ packed_lo,
                                                       const double4* __restrict__ centres,
                                                       const std::uint32_t* __restrict__ mask,
                                                       std::uint32_t tiles, std::uint32_t words,
                                                       std::uint32_t blocks, std::uint32_t S,
                                                       std::uint32_t items, float eps2,
                                                       const float* __restrict__ uniform,
                                                       float* __restrict__ jpart,
                                                       double* __restrict__ ipart) {
    if ((uniform[0] > 0.f) != kEqual) return;
    const std::uint32_t item =
        __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kTile) + threadIdx.x / kTile);
    if (item >= items) return;
    __shared__ double flushed[kBlock / kTile][3][kTile];
    __shared__ float compensation[kBlock / kTile][3 * kStay][kTile];
    const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kTile);
    double* const home = &flushed[wave][0][0];
    float* const comp = &compensation[wave][0][threadIdx.x & (kTile - 1)];
    pairs_item<kSafe, kRefine, kEqual>(packed, packed_lo, centres, mask, tiles, words, blocks, S,
                                       item, eps2, jpart, ipart, home, comp);
}

// What the pairs kernel leaves out: the pairs inside each block, one-sided.
// One wave per target tile over the tiles of its own block (near flags from
// the mask); sums and compensations to cpart[6][npad].
template<bool kSafe, bool kRefine>
__global__ __launch_bounds__(kBlock) void direct_complement(const float4* __restrict__ packed,
                                                            const float4* __restrict__ packed_lo,
                                                            const double4* __restrict__ centres,
                                                            const float4* __restrict__ extents,
                                                            const std::uint32_t* __restrict__ mask,
                                                            std::uint32_t tiles, std::uint32_t words,
                                                            float eps2, float* __restrict__ cpart) {
    const std::uint32_t t =
        __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kTile) + threadIdx.x / kTile);
    if (t >= tiles) return;
    const unsigned lane = threadIdx.x & (kTile - 1);
    const std::size_t k = static_cast<std::size_t>(t) * kTile + lane;
    const float4 o = packed[k];
    const float4 l = packed_lo[k];
    const double4 c = centres[t];
    const float ox[1] = {o.x}, oy[1] = {o.y}, oz[1] = {o.z};
    const float lx[1] = {l.x}, ly[1] = {l.y}, lz[1] = {l.z};
    const double cx[1] = {c.x}, cy[1] = {c.y}, cz[1] = {c.z};
    const float4 ce[1] = {extents[t]};
    float sx[1] = {0.f}, sy[1] = {0.f}, sz[1] = {0.f}, kx[1] = {0.f}, ky[1] = {0.f}, kz[1] = {0.f};
    const std::uint32_t* row = mask + static_cast<std::size_t>(t) * words;
    const std::uint32_t b0 = t / kStay * kStay, b1 = min(b0 + kStay, tiles);
    for (std::uint32_t J = b0; J < b1; ++J) {
        const bool near = ((row[J / 32] >> (J & 31u)) & 1u) != 0;
        source_tile<1, kSafe, kRefine, true>(packed, packed_lo, centres, extents, J, near, eps2, ox,
                                             oy, oz, lx, ly, lz, cx, cy, cz, ce, sx, sy, sz, kx, ky, kz);
    }
    const std::size_t np = static_cast<std::size_t>(tiles) * kTile;
    cpart[k] = sx[0];
    cpart[np + k] = sy[0];
    cpart[2 * np + k] = sz[0];
    cpart[3 * np + k] = kx[0];
    cpart[4 * np + k] = ky[0];
    cpart[5 * np + k] = kz[0];
}

// A particle's slots in a fixed order, in FP64: its block's items (sum minus
// compensation), its tile's streamed sums for every later block, the
// complement; times G, scattered through `order`.
__global__ __launch_bounds__(256) void pairs_reduce(const float* __restrict__ jpart,
                                                    const double* __restrict__ ipart,
                                                    const float* __restrict__ cpart,
                                                    const float* __restrict__ uniform,
                                                    std::uint32_t tiles, std::uint32_t blocks,
                                                    std::uint32_t S, std::uint64_t n,
                                                    const std::uint32_t* __restrict__ order,
                                                    double G, double* __restrict__ ax,
                                                    double* __restrict__ ay,
                                                    double* __restrict__ az, FusedKick kick) {
    const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const std::uint32_t t = static_cast<std::uint32_t>(k / kTile), lane = static_cast<std::uint32_t>(k % kTile);
    const std::uint32_t P = t / kStay, r = t % kStay;
    constexpr std::size_t kRow = kStay * kTile;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const std::uint64_t first = pair_items_before(P, S);
    const std::uint32_t count = pair_chunks(P, S);
#pragma unroll 4
    for (std::uint32_t u = 0; u < count; ++u) {
        const float* p = jpart + (first + u) * 6 * kRow + r * kTile + lane;
        a0 += static_cast<double>(p[0]) - static_cast<double>(p[3 * kRow]);
        a1 += static_cast<double>(p[kRow]) - static_cast<double>(p[4 * kRow]);
        a2 += static_cast<double>(p[2 * kRow]) - static_cast<double>(p[5 * kRow]);
    }
#pragma unroll 4
    for (std::uint32_t Q = P + 1; Q < blocks; ++Q) {
        const double* p = ipart + pair_slot(Q, t) * 3 * kTile + lane;
        a0 += p[0];
        a1 += p[kTile];
        a2 += p[2 * kTile];
    }
    // The pairs kernel's slots leave out a common mass; the complement's not.
    const double scale = uniform[0] > 0.f ? static_cast<double>(uniform[0]) : 1.0;
    a0 *= scale;
    a1 *= scale;
    a2 *= scale;
    const std::size_t np = static_cast<std::size_t>(tiles) * kTile;
    a0 += static_cast<double>(cpart[k]) - static_cast<double>(cpart[3 * np + k]);
    a1 += static_cast<double>(cpart[np + k]) - static_cast<double>(cpart[4 * np + k]);
    a2 += static_cast<double>(cpart[2 * np + k]) - static_cast<double>(cpart[5 * np + k]);
    const std::uint64_t dst = order != nullptr ? order[k] : k;
    store_field(dst, G * a0, G * a1, G * a2, ax, ay, az, kick);
}

template<int R, bool kSafe, bool kRefine, bool kMask>
void launch_tiles(orr_ctx* ctx, std::uint32_t blocks_i, std::uint32_t splits, const float4* packed,
                  const float4* packed_lo, const double4* centres, const float4* extents,
                  const std::uint32_t* mask, std::uint32_t npad, std::uint32_t tiles,
                  std::uint32_t per_split, float eps2, float* partials) {
    hipLaunchKernelGGL((direct_tiles<R, kSafe, kRefine, kMask>), dim3(blocks_i, splits), dim3(kBlock), 0,
                       ctx->stream, packed, packed_lo, centres, extents, mask, npad, tiles, per_split,
                       eps2, partials);
}

using TilesLauncher = void (*)(orr_ctx*, std::uint32_t, std::uint32_t, const float4*,
                               const float4*, const double4*, const float4*, const std::uint32_t*,
                               std::uint32_t, std::uint32_t, std::uint32_t, float, float*);

template<int R, bool kMask> TilesLauncher pick_mask(bool safe, bool refine) {
    if (safe) return refine ? launch_tiles<R, true, true, kMask> : launch_tiles<R, true, false, kMask>;
    return refine ? launch_tiles<R, false, true, kMask> : launch_tiles<R, false, false, kMask>;
}

template<int R> TilesLauncher pick(bool safe, bool refine, bool mask) {
    return mask ? pick_mask<R, true>(safe, refine) : pick_mask<R, false>(safe, refine);
}

// Pairs of tiles flagged by a mask up to this many tiles (a 128 MiB mask);
// beyond it the kernels test each pair themselves.
constexpr std::uint32_t kMaskTiles = 32768;

// The pairs path: blocks of kStay tiles, chunks of S tiles per item, slot
// counts. Taken when the mask exists, there are at least four blocks (pair
// mode 1: at least kPairMinTiles tiles) and the streamed-side slots (3 doubles
// per particle and later block) fit in 16 GiB; otherwise, or when configured
// off, the one-sided kernel runs.
struct PairPlan {
    bool use = false;
    std::uint32_t blocks = 0, S = kStay;  // S: streamed tiles per item
    std::uint64_t items = 0, slots = 0;
    std::size_t jpart_floats = 0, ipart_doubles = 0;
};

// The streamed-side slots grow as N^2 (one per tile and later block), and so
// does the kernel's work: their write and fold cost a fixed few per cent of an
// evaluation at any N. 16 GiB of the 288 GB HBM takes the pairs path to
// 856,064 particles.
constexpr std::size_t kPairSlotBytes = std::size_t{1} << 34;
// Items wanted per evaluation: some sixteen rounds of the 4096 resident waves.
// Short items balance the items that carry near passes (measured at
// N = 262,144: 16384 items 14.8 ms, 32768 14.0, 65536 13.7, 131072 13.5 ms,
// the last with twice the stationary-side slots).
constexpr std::uint64_t kPairItems = 65536;

// Below this many tiles the one-sided kernel is faster (measured per
// evaluation: N = 40,960 0.694 vs 0.702 ms, N = 49,152 0.894 vs 0.881 ms,
// N = 16,384 0.32 vs 0.36 ms): pair mode 1 takes the pairs path from here,
// pair mode 3 wherever it applies.
constexpr std::uint32_t kPairMinTiles = 704;

PairPlan pair_plan(std::uint32_t tiles, bool masked, int mode) {
    PairPlan p;
    if ((mode != 1 && mode != 3) || !masked) return p;
    if (mode == 1 && tiles < kPairMinTiles) return p;
    p.blocks = (tiles + kStay - 1) / kStay;
    if (p.blocks < 4) return p;
    p.slots = pair_slot(p.blocks, 0);
    p.ipart_doubles = static_cast<std::size_t>(p.slots) * 3 * kTile;
    if (p.ipart_doubles * sizeof(double) > kPairSlotBytes) return p;
    // About B^2 kStay / (2 S) items: S = kStay s for large systems, down to
    // kStay / 4 for small ones (more, shorter items; below that the
    // stationary-side slots outgrow the work).
    const std::uint64_t B = p.blocks;
    const std::uint64_t s = B * B / (2 * kPairItems);
    p.S = s >= 1 ? static_cast<std::uint32_t>(s) * kStay
                 : (B * B * 2 >= kPairItems ? kStay / 2 : kStay / 4);
    p.items = pair_items_before(p.blocks, p.S);
    p.jpart_floats = static_cast<std::size_t>(p.items) * 6 * kStay * kTile;
    p.use = true;
    return p;
}

// The complement (one wave per tile: too few waves to fill the device below a
// few hundred thousand particles, and latency-bound there) runs on the
// context's second stream, beside the pairs kernel; the main stream waits for
// it before the fold.
template<bool kSafe, bool kRefine>
void launch_pairs(orr_ctx* ctx, const PairPlan& plan, const float4* packed, const float4* packed_lo,
                  const double4* centres, const float4* extents, const std::uint32_t* mask,
                  std::uint32_t tiles, std::uint32_t words, float eps2, const float* uniform,
                  float* jpart, double* ipart, float* cpart) {
    constexpr std::uint32_t kWaves = kBlock / kTile;
    const std::uint32_t items = static_cast<std::uint32_t>(plan.items);
    ensure_aux(ctx);
    ORR_HIP(hipEventRecord(ctx->aux_fork, ctx->stream));
    ORR_HIP(hipStreamWaitEvent(ctx->aux, ctx->aux_fork, 0));
    hipLaunchKernelGGL((direct_complement<kSafe, kRefine>), dim3((tiles + kWaves - 1) / kWaves),
                       dim3(kBlock), 0, ctx->aux, packed, packed_lo, centres, extents, mask, tiles,
                       words, eps2, cpart);
    ORR_LAUNCHED("direct_complement");
    ORR_HIP(hipEventRecord(ctx->aux_join, ctx->aux));
    hipLaunchKernelGGL((direct_pairs<kSafe, kRefine, true>), dim3((items + kWaves - 1) / kWaves),
                       dim3(kBlock), 0, ctx->stream, packed, packed_lo, centres, mask, tiles, words, plan.blocks,
                       plan.S, items, eps2, uniform, jpart, ipart);
    ORR_LAUNCHED("direct_pairs");
    hipLaunchKernelGGL((direct_pairs<kSafe, kRefine, false>), dim3((items + kWaves - 1) / kWaves),
                       dim3(kBlock), 0, ctx->stream, packed, packed_lo, centres, mask, tiles, words, plan.blocks,
                       plan.S, items, eps2, uniform, jpart, ipart);
    ORR_LAUNCHED("direct_pairs");
    ORR_HIP(hipStreamWaitEvent(ctx->stream, ctx->aux_join, 0));
}

}  // namespace

namespace {

// The shape of one single-system evaluation: tiles, one-sided launch grid,
// near mask and pairs plan, as direct_enqueue chooses them for the context's
// configuration.
struct DirectShape {
    std::uint32_t tiles = 0, npad = 0, blocks_i = 0, splits = 0, per_split = 0, words = 0;
    int R = 1;
    bool masked = false;
    PairPlan plan;
};

DirectShape direct_shape(const orr_ctx* ctx, std::uint64_t n) {
    if (n > (std::uint64_t{1} << 31)) fail(ORR_ERR_LENGTH, "direct: more than 2^31 particles");
    DirectShape d;
    d.tiles = static_cast<std::uint32_t>((n + kTile - 1) / kTile);
    d.npad = d.tiles * kTile;
    d.R = ctx->direct_targets_per_lane;
    d.blocks_i = (d.npad + kBlock * d.R - 1) / (kBlock * d.R);
    // Enough blocks to fill 256 CUs several times over; each split keeps >= 1 tile.
    const std::uint32_t want = ctx->direct_block_target;
    std::uint32_t splits = std::max<std::uint32_t>(1, (want + d.blocks_i - 1) / d.blocks_i);
    splits = std::min(splits, d.tiles);
    d.per_split = (d.tiles + splits - 1) / splits;
    d.splits = (d.tiles + d.per_split - 1) / d.per_split;
    d.masked = d.tiles <= kMaskTiles;
    d.words = (d.tiles + 31) / 32;
    d.plan = pair_plan(d.tiles, d.masked, ctx->direct_pair_mode);
    return d;
}

// Device buffers of one evaluation (DirectShape's sizes).
struct DirectBuffers {
    float4* packed = nullptr;     // [npad] {o, m}, then [npad] residuals
    double4* centres = nullptr;   // [tiles], then extents [tiles], mass ranges [tiles], uniform
    float* partials = nullptr;    // one-sided: [splits][6][npad]; pairs: slots
    std::uint32_t* mask = nullptr;  // [tiles][words] when masked
};

std::size_t packed_buffer_bytes(const DirectShape& d) { return 2 * sizeof(float4) * d.npad; }
std::size_t centre_buffer_bytes(const DirectShape& d) {
    return (sizeof(double4) + sizeof(float4) + sizeof(float2)) * d.tiles + 16;
}
std::size_t partial_buffer_bytes(const DirectShape& d) {
    const std::size_t cpart_floats = 6 * static_cast<std::size_t>(d.npad);
    return d.plan.use ? sizeof(double) * d.plan.ipart_doubles +
                            sizeof(float) * (d.plan.jpart_floats + cpart_floats)
                      : sizeof(float) * 6 * static_cast<std::size_t>(d.npad) * d.splits;
}
std::size_t mask_buffer_bytes(const DirectShape& d) {
    return d.masked ? sizeof(std::uint32_t) * static_cast<std::size_t>(d.words) * d.tiles : 0;
}

// Every launch of one evaluation on given buffers; phases timed when `timed`.
void direct_launch(orr_ctx* ctx, const orr_direct_args& a, const std::uint32_t* order,
                   const DirectShape& d, const DirectBuffers& buf, bool timed,
                   const FusedKick& kick = FusedKick{}) {
    const std::uint64_t n = a.count;
    const std::uint32_t tiles = d.tiles, npad = d.npad, words = d.words;
    const PairPlan& plan = d.plan;
    float4* packed = buf.packed;
    float4* packed_lo = packed + npad;
    double4* centres = buf.centres;
    float4* extents = reinterpret_cast<float4*>(centres + tiles);
    float2* mass_range = reinterpret_cast<float2*>(extents + tiles);
    float* uniform = reinterpret_cast<float*>(mass_range + tiles);
    float* partials = buf.partials;
    std::uint32_t* mask = d.masked ? buf.mask : nullptr;
    {
        PhaseTimer t(ctx, kDirectPrepare, timed);
        hipLaunchKernelGGL(direct_prepare, dim3(tiles), dim3(kTile), 0, ctx->stream, a.x, a.y, a.z,
                           a.m, order, n, packed, packed_lo, centres, extents, mass_range);
        ORR_LAUNCHED("direct_prepare");
        if (plan.use) {
            hipLaunchKernelGGL(mass_uniform, dim3(1), dim3(256), 0, ctx->stream, mass_range, tiles,
                               ctx->direct_equal_masses == 1 ? 1 : 0, uniform);
            ORR_LAUNCHED("mass_uniform");
        }
        if (d.masked) {
            hipLaunchKernelGGL(near_mask, dim3((tiles + kMaskRows - 1) / kMaskRows), dim3(256), 0,
                               ctx->stream, centres, extents, tiles, words, mask);
            ORR_LAUNCHED("near_mask");
        }
    }
    const float eps2 = static_cast<float>(a.softening_squared);
    const bool safe = !(a.softening_squared > 0);
    const bool refine = ctx->direct_refine != 0;
    if (plan.use) {
        // Doubles first: every part stays aligned to its type.
        double* ipart = reinterpret_cast<double*>(partials);
        float* jpart = reinterpret_cast<float*>(ipart + plan.ipart_doubles);
        float* cpart = jpart + plan.jpart_floats;
        {
            PhaseTimer t(ctx, kDirectKernel, timed);
            auto launch = safe ? (refine ? launch_pairs<true, true> : launch_pairs<true, false>)
                               : (refine ? launch_pairs<false, true> : launch_pairs<false, false>);
            launch(ctx, plan, packed, packed_lo, centres, extents, mask, tiles, words, eps2, uniform,
                   jpart, ipart, cpart);
        }
        {
            PhaseTimer t(ctx, kDirectReduce, timed);
            const std::uint32_t blocks = static_cast<std::uint32_t>((n + 255) / 256);
            hipLaunchKernelGGL(pairs_reduce, dim3(blocks), dim3(256), 0, ctx->stream, jpart, ipart,
                               cpart, uniform, tiles, plan.blocks, plan.S, n, order,
                               a.gravitational_constant, a.ax, a.ay, a.az, kick);
            ORR_LAUNCHED("pairs_reduce");
        }
        return;
    }
    {
        PhaseTimer t(ctx, kDirectKernel, timed);
        TilesLauncher launch = d.R == 1 ? pick<1>(safe, refine, d.masked)
                             : d.R == 4 ? pick<4>(safe, refine, d.masked)
                                        : pick<2>(safe, refine, d.masked);
        launch(ctx, d.blocks_i, d.splits, packed, packed_lo, centres, extents, mask, npad, tiles,
               d.per_split, eps2, partials);
        ORR_LAUNCHED("direct_tiles");
    }
    {
        PhaseTimer t(ctx, kDirectReduce, timed);
        const std::uint32_t blocks = static_cast<std::uint32_t>((n + 255) / 256);
        hipLaunchKernelGGL(direct_reduce, dim3(blocks), dim3(256), 0, ctx->stream, partials, d.splits,
                           npad, n, order, a.gravitational_constant, a.ax, a.ay, a.az, kick);
        ORR_LAUNCHED("direct_reduce");
    }
}

}  // namespace

void direct_enqueue(orr_ctx* ctx, const orr_direct_args& a, const std::uint32_t* order,
                    const FusedKick& kick) {
    if (a.count == 0) return;
    const DirectShape d = direct_shape(ctx, a.count);
    ctx->direct_packed.reserve(packed_buffer_bytes(d));
    ctx->direct_centres.reserve(centre_buffer_bytes(d));
    ctx->direct_partials.reserve(partial_buffer_bytes(d));
    if (d.masked) ctx->direct_near.reserve(mask_buffer_bytes(d));
    DirectBuffers buf;
    buf.packed = ctx->direct_packed.as<float4>();
    buf.centres = ctx->direct_centres.as<double4>();
    buf.partials = ctx->direct_partials.as<float>();
    buf.mask = d.masked ? ctx->direct_near.as<std::uint32_t>() : nullptr;
    direct_launch(ctx, a, order, d, buf, true, kick);
}

namespace {
std::size_t packed_bytes(std::uint64_t tiles) { return 2 * sizeof(float4) * kTile * tiles; }
std::size_t centre_bytes(std::uint64_t tiles) { return (sizeof(double4) + sizeof(float4)) * tiles; }
std::size_t slot_bytes(std::uint64_t slots) { return sizeof(float) * 6 * kTile * slots; }
}  // namespace

std::size_t direct_batch_bytes(std::uint64_t tiles, std::uint64_t slots) {
    return packed_bytes(tiles) + centre_bytes(tiles) + slot_bytes(slots);
}

void direct_evaluate(orr_ctx* ctx, const orr_direct_args& a, const FusedKick& kick) {
    if (a.count == 0) return;
    const std::uint32_t* order = nullptr;
    if (a.spatial_order != 0 && a.count > 64) {
        // The Morton order is some nine short launches: replayed as a graph (its
        // phases untimed); the three direct launches stay eager and timed.
        std::vector<unsigned char> key;
        key_add(key, a.x);
        key_add(key, a.y);
        key_add(key, a.z);
        key_add(key, a.count);
        // Compact tiles need only the top 32 bits of the 63-bit codes: the four
        // upper radix passes of eight, or for a small system the same order
        // from one workgroup.
        const bool small = ctx->direct_small_order == 1 && a.count <= kSmallOrder;
        key_add(key, small);
        run_captured(ctx, ctx->direct_graph, std::move(key), [&] {
            if (small) {
                morton_enqueue_small(ctx, a.x, a.y, a.z, a.count, 4);
            } else {
                morton_enqueue(ctx, a.x, a.y, a.z, a.count, 4);
            }
        }, true);
        order = ctx->order_index.as<std::uint32_t>();
    }
    direct_enqueue(ctx, a, order, kick);
}

void direct_batch_enqueue(orr_ctx* ctx, const double* x, const double* y, const double* z,
                          const double* m, const std::uint32_t* order, const DirectBatchView& b,
                          double* ax, double* ay, double* az) {
    if (b.count == 0 || b.tiles == 0) return;
    const std::size_t npad = static_cast<std::size_t>(b.tiles) * kTile;
    ctx->direct_packed.reserve(packed_bytes(b.tiles));
    ctx->direct_centres.reserve(centre_bytes(b.tiles));
    ctx->direct_partials.reserve(slot_bytes(b.slots));
    float4* packed = ctx->direct_packed.as<float4>();
    float4* packed_lo = packed + npad;
    double4* centres = ctx->direct_centres.as<double4>();
    float4* extents = reinterpret_cast<float4*>(centres + b.tiles);
    float* partials = ctx->direct_partials.as<float>();
    {
        PhaseTimer t(ctx, kDirectPrepare);
        hipLaunchKernelGGL(batch_prepare, dim3(b.tiles), dim3(kTile), 0, ctx->stream, x, y, z, m,
                           order, b.tile, packed, packed_lo, centres, extents);
        ORR_LAUNCHED("batch_prepare");
    }
    {
        PhaseTimer t(ctx, kDirectKernel);
        // Eight blocks of four waves per CU; never more waves than items.
        const std::uint32_t cap = static_cast<std::uint32_t>(std::max(1, ctx->compute_units)) * 8;
        const std::uint32_t blocks = std::max<std::uint32_t>(1, std::min(cap, (b.items + 3) / 4));
        const bool refine = ctx->direct_refine != 0;
        auto kernel = b.safe ? (refine ? direct_items<true, true> : direct_items<true, false>)
                             : (refine ? direct_items<false, true> : direct_items<false, false>);
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, ctx->stream, packed, packed_lo,
                           centres, extents, b.tile, b.item, b.items, b.eps2, partials);
        ORR_LAUNCHED("direct_items");
    }
    {
        PhaseTimer t(ctx, kDirectReduce);
        const std::uint32_t blocks = static_cast<std::uint32_t>((b.count + 255) / 256);
        hipLaunchKernelGGL(batch_reduce, dim3(blocks), dim3(256), 0, ctx->stream, partials, b.sys,
                           b.tile, b.seg_of, b.G, b.count, order, ax, ay, az);
        ORR_LAUNCHED("batch_reduce");
    }
}

// ---- the direct sum on the library's static scratch ----------------------
//
// For callers that cannot allocate device memory: every buffer of the
// evaluation -- packed tiles, centres, near mask, the Morton order's codes,
// index and histograms, the one-sided kernel's split partials -- is carved out
// of `scratch` below, a fixed array each device holds from the first call
// until the process ends. Limits: at most kStaticMaxCount particles; the
// one-sided kernel always (the pairs path's slots do not fit); the split count
// of direct_shape, lowered where its partials would not fit (the same bits as
// direct_evaluate with the one-sided kernel whenever it fits, else the same
// terms grouped into fewer Kahan-summed splits). Nothing in the scratch has to
// start at zero: every launch writes what the next reads. Calls on one device
// share the array, so calls that overlap on one device (two contexts, two
// host threads) race.

constexpr std::size_t kStaticScratchBytes = std::size_t{16} << 20;
constexpr std::uint64_t kStaticMaxCount = 32768;

}  // namespace orr

static __device__ __align__(256) unsigned char scratch[orr::kStaticScratchBytes];

namespace {

using orr::kStaticMaxCount;
using orr::kStaticScratchBytes;
constexpr int kMaxDevices = 64;

// The array's device address on each device, looked up once per device with
// that device current (host code must never use the symbol itself as an
// address).
std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) orr::fail(ORR_ERR_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (g_scratch_ptr[device] == nullptr) {
        void* ptr = nullptr;
        ORR_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) orr::fail(ORR_ERR_RUNTIME, "static scratch: no device address");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

// Bump allocation inside the scratch, 256-byte aligned.
struct Carver {
    unsigned char* base;
    std::size_t used = 0;
    template<typename T> T* take(std::size_t bytes) {
        T* p = reinterpret_cast<T*>(base + used);
        used += (bytes + 255) & ~std::size_t{255};
        return p;
    }
};

std::size_t aligned(std::size_t bytes) { return (bytes + 255) & ~std::size_t{255}; }

}  // namespace

namespace orr {

std::size_t direct_static_scratch_bytes() { return kStaticScratchBytes; }
std::uint64_t direct_static_max_count() { return kStaticMaxCount; }

void direct_evaluate_static(orr_ctx* ctx, const orr_direct_args& a) {
    if (a.count == 0) return;
    if (a.count > kStaticMaxCount) {
        fail(ORR_ERR_LENGTH, "direct (static scratch): more than 32768 particles; use orr_direct_eval");
    }
    const std::uint64_t n = a.count;
    DirectShape d = direct_shape(ctx, n);
    d.plan = PairPlan{};  // the one-sided kernel
    const bool ordered = a.spatial_order != 0 && n > 64;
    const OrderSizes oz = order_sizes(n, 1, 0, false);
    const std::size_t fixed = aligned(packed_buffer_bytes(d)) + aligned(centre_buffer_bytes(d)) +
                              aligned(mask_buffer_bytes(d)) +
                              (ordered ? aligned(oz.bbox) + 2 * aligned(oz.codes) +
                                             2 * aligned(oz.index) + aligned(oz.hist)
                                       : 0);
    const std::size_t split_bytes = aligned(sizeof(float) * 6 * static_cast<std::size_t>(d.npad));
    if (fixed + split_bytes > kStaticScratchBytes) {
        fail(ORR_ERR_LENGTH, "direct (static scratch): the evaluation does not fit the scratch");
    }
    const std::uint32_t fit = static_cast<std::uint32_t>((kStaticScratchBytes - fixed) / split_bytes);
    if (d.splits > fit) {
        d.per_split = (d.tiles + fit - 1) / fit;
        d.splits = (d.tiles + d.per_split - 1) / d.per_split;
    }
    Carver c{static_cast<unsigned char*>(scratch_on(ctx->device))};
    DirectBuffers buf;
    buf.packed = c.take<float4>(packed_buffer_bytes(d));
    buf.centres = c.take<double4>(centre_buffer_bytes(d));
    buf.mask = d.masked ? c.take<std::uint32_t>(mask_buffer_bytes(d)) : nullptr;
    const std::uint32_t* order = nullptr;
    if (ordered) {
        OrderScratch ob;
        ob.bbox = c.take<double>(oz.bbox);
        ob.codes[0] = c.take<std::uint64_t>(oz.codes);
        ob.codes[1] = c.take<std::uint64_t>(oz.codes);
        ob.index[0] = c.take<std::uint32_t>(oz.index);
        ob.index[1] = c.take<std::uint32_t>(oz.index);
        ob.hist = c.take<std::uint32_t>(oz.hist);
        // The order direct_evaluate uses: the four upper radix passes.
        const int cur = morton_enqueue_into(ctx, a.x, a.y, a.z, n, SegmentView{}, nullptr, 4, ob, false);
        order = ob.index[cur];
    }
    buf.partials = c.take<float>(split_bytes * d.splits);
    if (c.used > kStaticScratchBytes) fail(ORR_ERR_RUNTIME, "direct (static scratch): layout overflow");
    direct_launch(ctx, a, order, d, buf, false);
}

}  // namespace orr
