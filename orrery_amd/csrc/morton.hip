// Morton ordering on gfx950: bounding cube, 63-bit keys, stable LSD radix sort.
//
// Bit-exact with the reference's MortonOrdering (src/solvers/morton.cpp:60-197):
//   * the cube is the exact min/max box, size = largest extent (1 if not > 0),
//     centred: origin = (lo + hi) / 2 - size / 2 (morton.cpp:60-104);
//   * scale = 2^21 / size, cell = (c - origin) * scale with !(cell > 0) -> 0 and
//     cell >= 2^21 -> 2^21 - 1, then truncation (:128-164) -- a subtract and a
//     multiply, nothing to contract, and the file disables contraction anyway;
//   * the 21-bit dilation masks of :106-122 and x-major interleave of :124-126;
//   * a stable LSD radix sort starting from identity order reproduces the
//     (code, index) lexicographic total order (morton.hpp:155-175).
//
// Sort: 8 passes of 8-bit digits. By default one launch counts the digits of
// every pass and each pass is one scatter that finds the blocks before it by
// decoupled look-back; the multi-kernel form (per pass: histogram -> device-wide
// exclusive scan -> scatter) gives the same permutation and stays selectable.
// Both scatters rank keys stably inside 256-key tiles with wave64
// ballots (8 ballots give each lane the mask of lanes sharing its digit) and an
// LDS table of per-wave digit counts, so no LDS atomics decide an order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <utility>

#include "engine.hpp"

#pragma clang fp contract(off)

namespace orr {
namespace {

constexpr int kBlock = 256;
constexpr int kTilesPerBlock = 8;                   // keys per sort block = 2048
constexpr int kKeysPerBlock = kBlock * kTilesPerBlock;
constexpr unsigned kGrid = 1u << 21;
constexpr int kMaxSortPasses = 10;  // eight code digits and up to two segment digits

// ---- bounding cube -----------------------------------------------------------

__global__ __launch_bounds__(kBlock) void bbox_partial(const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       std::uint64_t n,
                                                       double* __restrict__ partial) {
    __shared__ double s[6][kBlock];
    const double inf = __builtin_huge_val();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<std::uint64_t>(gridDim.x) * kBlock) {
        const double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] = v[c] < lo[c] ? v[c] : lo[c];
            hi[c] = v[c] > hi[c] ? v[c] : hi[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double a = s[c][threadIdx.x + w];
                if (a < s[c][threadIdx.x]) s[c][threadIdx.x] = a;
                const double b = s[3 + c][threadIdx.x + w];
                if (b > s[3 + c][threadIdx.x]) s[3 + c][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// The cube arithmetic of morton.cpp:87-103 from exact extrema: cube = {origin, size}.
__device__ void make_cube(const double lo[3], const double hi[3], double* __restrict__ cube);

// One block folds the block extrema (min and max are exact, so the order does
// not matter), then one thread does the cube arithmetic of morton.cpp:87-103.
// cube = {origin x, y, z, size}.
__global__ __launch_bounds__(kBlock) void bbox_final(const double* __restrict__ partial,
                                                     std::uint32_t blocks,
                                                     double* __restrict__ cube) {
    __shared__ double s[6][kBlock];
    const double inf = __builtin_huge_val();
    double l[3] = {inf, inf, inf}, h[3] = {-inf, -inf, -inf};
    for (std::uint32_t b = threadIdx.x; b < blocks; b += kBlock) {
        for (int c = 0; c < 3; ++c) {
            const double a = partial[b * 6 + c];
            const double e = partial[b * 6 + 3 + c];
            l[c] = a < l[c] ? a : l[c];
            h[c] = e > h[c] ? e : h[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = l[c];
        s[3 + c][threadIdx.x] = h[c];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            for (int c = 0; c < 3; ++c) {
                const double a = s[c][threadIdx.x + w];
                if (a < s[c][threadIdx.x]) s[c][threadIdx.x] = a;
                const double e = s[3 + c][threadIdx.x + w];
                if (e > s[3 + c][threadIdx.x]) s[3 + c][threadIdx.x] = e;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
        lo[c] = s[c][0];
        hi[c] = s[3 + c][0];
    }
    make_cube(lo, hi, cube);
}

// One block per chunk {segment, begin, end} of one segment: its extrema.
__global__ __launch_bounds__(kBlock) void bbox_chunks(const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const std::uint32_t* __restrict__ chunks,
                                                      double* __restrict__ partial) {
    __shared__ double s[6][kBlock];
    const double inf = __builtin_huge_val();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    const std::uint32_t b = chunks[3 * blockIdx.x + 1], e = chunks[3 * blockIdx.x + 2];
    for (std::uint32_t i = b + threadIdx.x; i < e; i += kBlock) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int c = 0; c < 3; ++c) {
            lo[c] = v[c] < lo[c] ? v[c] : lo[c];
            hi[c] = v[c] > hi[c] ? v[c] : hi[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            for (int c = 0; c < 3; ++c) {
                const double a = s[c][threadIdx.x + w];
                if (a < s[c][threadIdx.x]) s[c][threadIdx.x] = a;
                const double h = s[3 + c][threadIdx.x + w];
                if (h > s[3 + c][threadIdx.x]) s[3 + c][threadIdx.x] = h;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// One thread per segment: fold its chunks' extrema, then its cube.
__global__ __launch_bounds__(kBlock) void bbox_segments(const double* __restrict__ partial,
                                                        const std::uint32_t* __restrict__ chunk_first,
                                                        std::uint32_t segments,
                                                        double* __restrict__ cubes) {
    const std::uint32_t seg = blockIdx.x * kBlock + threadIdx.x;
    if (seg >= segments) return;
    const double inf = __builtin_huge_val();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (std::uint32_t c = chunk_first[seg]; c < chunk_first[seg + 1]; ++c) {
        for (int k = 0; k < 3; ++k) {
            const double a = partial[c * 6 + k], h = partial[c * 6 + 3 + k];
            lo[k] = a < lo[k] ? a : lo[k];
            hi[k] = h > hi[k] ? h : hi[k];
        }
    }
    make_cube(lo, hi, cubes + 4 * static_cast<std::size_t>(seg));
}

__device__ void make_cube(const double lo[3], const double hi[3], double* __restrict__ cube) {
    const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    // std::max({ex, ey, ez}) keeps the first of equals; the value is what matters.
    double size = ex;
    if (size < ey) size = ey;
    if (size < ez) size = ez;
    if (!(size > 0)) size = 1;
    for (int c = 0; c < 3; ++c) {
        const double centre = (lo[c] + hi[c]) / 2;
        cube[c] = centre - size / 2;
    }
    cube[3] = size;
}

// ---- encode -----------------------------------------------------------------

__device__ inline std::uint64_t spread_bits(std::uint32_t v) {
    std::uint64_t s = v & 0x1fffffu;
    s = (s | (s << 32)) & 0x1f00000000ffffull;
    s = (s | (s << 16)) & 0x1f0000ff0000ffull;
    s = (s | (s << 8)) & 0x100f00f00f00f00full;
    s = (s | (s << 4)) & 0x10c30c30c30c30c3ull;
    s = (s | (s << 2)) & 0x1249249249249249ull;
    return s;
}

__device__ inline std::uint32_t to_grid(double coordinate, double origin, double scale) {
    const double cell = (coordinate - origin) * scale;
    if (!(cell > 0)) return 0;
    if (cell >= static_cast<double>(kGrid)) return kGrid - 1;
    return static_cast<std::uint32_t>(cell);
}

__global__ __launch_bounds__(kBlock) void encode(const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ m, std::uint64_t n,
                                                 const double* __restrict__ cubes,
                                                 const std::uint32_t* __restrict__ seg_of,
                                                 std::uint64_t* __restrict__ codes,
                                                 std::uint32_t* __restrict__ index,
                                                 double4* __restrict__ aos) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const double* cube = cubes + (seg_of != nullptr ? 4 * static_cast<std::size_t>(seg_of[i]) : 0);
    const double size = cube[3];
    const double px = x[i], py = y[i], pz = z[i];
    std::uint64_t code = 0;
    if (size > 0) {
        const double scale = static_cast<double>(kGrid) / size;
        code = (spread_bits(to_grid(px, cube[0], scale)) << 2) |
               (spread_bits(to_grid(py, cube[1], scale)) << 1) |
               spread_bits(to_grid(pz, cube[2], scale));
    }
    codes[i] = code;
    index[i] = static_cast<std::uint32_t>(i);
    if (aos != nullptr) aos[i] = make_double4(px, py, pz, m[i]);
}

// ---- radix sort -------------------------------------------------------------

// hist is digit-major: hist[d * blocks + b] = count of digit d in block b.
// With seg_of, the digit is taken from the segment of each element's value
// instead of its key: the passes that order by segment after the code passes.
__device__ inline std::uint32_t digit_of(std::uint64_t key, std::uint32_t val, int shift,
                                         const std::uint32_t* __restrict__ seg_of) {
    const std::uint64_t k = seg_of != nullptr ? static_cast<std::uint64_t>(seg_of[val]) : key;
    return static_cast<std::uint32_t>((k >> shift) & 255u);
}

__global__ __launch_bounds__(kBlock) void radix_histogram(const std::uint64_t* __restrict__ keys,
                                                          const std::uint32_t* __restrict__ vals,
                                                          std::uint64_t n, int shift,
                                                          const std::uint32_t* __restrict__ seg_of,
                                                          std::uint32_t* __restrict__ hist) {
    __shared__ std::uint32_t counts[256];
    counts[threadIdx.x] = 0;
    __syncthreads();
    const std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * kKeysPerBlock;
    for (int t = 0; t < kTilesPerBlock; ++t) {
        const std::uint64_t i = base + static_cast<std::uint64_t>(t) * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&counts[digit_of(keys[i], vals[i], shift, seg_of)], 1u);
    }
    __syncthreads();
    hist[static_cast<std::uint64_t>(threadIdx.x) * gridDim.x + blockIdx.x] = counts[threadIdx.x];
}

// Stable scatter of one pass. Each block ranks its 2048 keys tile by tile
// (256 at a time, in order: ballots give each lane the lanes of its wave with
// the same digit, an LDS table the counts of the waves before it) into LDS in
// digit order, then writes them out in that order: the keys of one digit leave
// the block as one contiguous run (16 keys on average) instead of one
// scattered 8-byte store per key. Positions are those of a plain stable
// scatter: offsets (the scanned [digit][block] histogram) + rank within the
// block's digit.
__global__ __launch_bounds__(kBlock) void radix_scatter(const std::uint64_t* __restrict__ keys_in,
                                                        const std::uint32_t* __restrict__ vals_in,
                                                        std::uint64_t n, int shift,
                                                        const std::uint32_t* __restrict__ seg_of,
                                                        const std::uint32_t* __restrict__ offsets,
                                                        std::uint64_t* __restrict__ keys_out,
                                                        std::uint32_t* __restrict__ vals_out) {
    __shared__ std::uint64_t skeys[kKeysPerBlock];
    __shared__ std::uint32_t svals[kKeysPerBlock];
    __shared__ std::uint32_t out_base[256];       // offsets[d][b] - block-local start of d
    __shared__ std::uint32_t running[256];        // next local slot of each digit
    __shared__ std::uint32_t wave_counts[4][256];  // digit counts of each wave in this tile
    __shared__ std::uint32_t wave_sums[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const std::uint32_t grid = gridDim.x;
    const std::uint32_t d0 = threadIdx.x;  // this thread's digit for the per-digit tables
    const std::uint32_t here = d0 * grid + blockIdx.x;
    const std::uint32_t base_d = offsets[here];
    // The block's count of digit d0: the next entry of the digit-major offsets
    // (the one after the last block of digit 255 is n).
    const std::uint64_t next_d = here + 1 < 256u * grid ? offsets[here + 1] : n;
    const std::uint32_t count_d = static_cast<std::uint32_t>(next_d - base_d);
    // Exclusive scan of the 256 counts across the block: within each wave by
    // shuffles, then the four wave sums.
    std::uint32_t incl = count_d;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const std::uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wave_sums[wave] = incl;
    for (int w = 0; w < 4; ++w) wave_counts[w][threadIdx.x] = 0;
    __syncthreads();
    std::uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_sums[w];
    out_base[d0] = base_d - (before + incl - count_d);
    running[d0] = before + incl - count_d;
    __syncthreads();

    const std::uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * kKeysPerBlock;
    const std::uint32_t in_block = static_cast<std::uint32_t>(
        n - base < static_cast<std::uint64_t>(kKeysPerBlock) ? n - base : kKeysPerBlock);
#pragma unroll 1
    for (int t = 0; t < kTilesPerBlock; ++t) {
        const std::uint64_t i = base + static_cast<std::uint64_t>(t) * kBlock + threadIdx.x;
        const bool valid = i < n;
        const std::uint64_t key = valid ? keys_in[i] : 0;
        const std::uint32_t val = valid ? vals_in[i] : 0;
        const std::uint32_t d = valid ? digit_of(key, val, shift, seg_of) : 0u;
        // Lanes of this wave holding a valid key with the same digit.
        std::uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = ((d >> b) & 1u) != 0;
            const std::uint64_t m = __ballot(bit);
            same &= bit ? m : ~m;
        }
        const std::uint32_t rank = static_cast<std::uint32_t>(__popcll(same & lanes_below));
        if (valid && rank == 0) wave_counts[wave][d] = static_cast<std::uint32_t>(__popcll(same));
        __syncthreads();
        if (valid) {
            std::uint32_t slot = running[d] + rank;
            for (int w = 0; w < wave; ++w) slot += wave_counts[w][d];
            skeys[slot] = key;
            svals[slot] = val;
        }
        __syncthreads();
        // Advance each digit's slot past this tile and clear the wave table.
        const std::uint32_t total = wave_counts[0][threadIdx.x] + wave_counts[1][threadIdx.x] +
                                    wave_counts[2][threadIdx.x] + wave_counts[3][threadIdx.x];
        running[threadIdx.x] += total;
        for (int w = 0; w < 4; ++w) wave_counts[w][threadIdx.x] = 0;
        __syncthreads();
    }
    // Out in local (digit) order: slot s of digit d goes to offsets[d][b] + (s - start of d).
#pragma unroll 4
    for (std::uint32_t s = threadIdx.x; s < in_block; s += kBlock) {
        const std::uint64_t key = skeys[s];
        const std::uint32_t val = svals[s];
        const std::uint32_t pos = out_base[digit_of(key, val, shift, seg_of)] + s;
        keys_out[pos] = key;
        vals_out[pos] = val;
    }
}

// ---- one-sweep passes ----------------------------------------------------------
//
// The digit totals of a pass do not depend on the order the keys are in, so one
// launch counts the digits of every pass (per block; a second launch sums them
// and scans them into each digit's first position), and each pass is then one
// scatter launch instead of histogram, scan and scatter. A block counts its own
// keys' digits, publishes the counts, stages its keys in LDS in digit order as
// radix_scatter does, and learns the counts of the blocks before it by
// decoupled look-back over one status word per (block, digit): the count in the
// low 30 bits, above it kAggregate (this block's own count) or kPrefix (the
// count of all blocks up to and including this one). One word carries flag and
// count, so nothing else has to be ordered with it; agent-scope atomic loads
// and stores make it visible across the XCDs' L2 caches. A block waits only on
// blocks with lower indices, and each XCD starts its workgroups in index order,
// so the lowest unfinished block is always running or next to start on its XCD
// (a per-block ticket from one atomic counter, the portable guarantee, cost
// 1024 contended device atomics a pass); every block publishes its counts
// before it waits on anything, and a wait that still outlasts kSpinCap polls
// gives up (a wrong order, never a hang). The positions are those of
// radix_scatter: the same stable permutation.
constexpr std::uint32_t kAggregate = 1u << 30, kPrefix = 2u << 30, kCountMask = kAggregate - 1;
constexpr std::uint32_t kCountBlocks = 256;
constexpr int kCountThreads = 1024;
constexpr int kCountUnroll = 8;  // keys a counting thread loads before it counts them
constexpr std::uint32_t kSpinCap = 1u << 20;
// Blocks whose words one look-back step loads together. At 2,097,152 keys
// (1024 blocks) 32 measured best: 64 took the sort from 0.31 to 0.39 ms, and
// 4096 keys a block (half the blocks) gave no gain at 2M and lost at 262,144.
constexpr int kLookback = 32;

__global__ __launch_bounds__(kCountThreads) void radix_count_all(const std::uint64_t* __restrict__ keys,
                                                                 const std::uint32_t* __restrict__ vals,
                                                                 std::uint64_t n, int first_pass, int passes,
                                                                 const std::uint32_t* __restrict__ seg_of,
                                                                 std::uint32_t* __restrict__ partial,
                                                                 std::uint32_t* __restrict__ clear,
                                                                 std::uint64_t clear_words) {
    __shared__ std::uint32_t counts[kMaxSortPasses * 256];
    for (int w = threadIdx.x; w < kMaxSortPasses * 256; w += kCountThreads) counts[w] = 0;
    __syncthreads();
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kCountThreads;
    for (std::uint64_t i0 = static_cast<std::uint64_t>(blockIdx.x) * kCountThreads + threadIdx.x; i0 < n;
         i0 += stride * kCountUnroll) {
        // Loads first, so that they are in flight together.
        std::uint64_t key[kCountUnroll];
        std::uint32_t val[kCountUnroll];
#pragma unroll
        for (int u = 0; u < kCountUnroll; ++u) {
            const std::uint64_t i = i0 + u * stride;
            key[u] = i < n ? keys[i] : 0;
            val[u] = i < n && seg_of != nullptr ? vals[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kCountUnroll; ++u) {
            if (i0 + u * stride < n) {
                for (int p = first_pass; p < passes; ++p) {
                    const bool by_segment = p >= 8;
                    const std::uint32_t d = digit_of(key[u], val[u], by_segment ? 8 * (p - 8) : 8 * p,
                                                     by_segment ? seg_of : nullptr);
                    atomicAdd(&counts[p * 256 + d], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int w = first_pass * 256 + static_cast<int>(threadIdx.x); w < passes * 256; w += kCountThreads) {
        const int p = w / 256, d = w % 256;
        partial[(static_cast<std::size_t>(p) * gridDim.x + blockIdx.x) * 256 + d] = counts[w];
    }
    // The status words of this sort's passes start at zero.
    for (std::uint64_t w = static_cast<std::uint64_t>(blockIdx.x) * kCountThreads + threadIdx.x; w < clear_words;
         w += stride) {
        clear[w] = 0;
    }
}

// One block per pass: each digit's total over the counting blocks (four
// threads a digit, each over every fourth block), scanned over the digits into
// the position of the digit's first key.
constexpr int kBaseThreads = 1024;
__global__ __launch_bounds__(kBaseThreads) void radix_bases(const std::uint32_t* __restrict__ partial,
                                                            std::uint32_t count_blocks, int first_pass,
                                                            std::uint32_t* __restrict__ bases) {
    __shared__ std::uint32_t quarter[kBaseThreads / 256][256];
    __shared__ std::uint32_t wave_sums[4];
    const int p = first_pass + static_cast<int>(blockIdx.x);
    const std::uint32_t d = threadIdx.x & 255, q = threadIdx.x >> 8;
    const std::uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const std::uint32_t* col = partial + static_cast<std::size_t>(p) * count_blocks * 256 + d;
    std::uint32_t part = 0;
#pragma unroll 16
    for (std::uint32_t b = q; b < count_blocks; b += kBaseThreads / 256) part += col[static_cast<std::size_t>(b) * 256];
    quarter[q][d] = part;
    __syncthreads();
    // The first four waves (one thread a digit) scan the totals.
    std::uint32_t total = 0, incl = 0;
    if (q == 0) {
        for (int k = 0; k < kBaseThreads / 256; ++k) total += quarter[k][d];
        incl = total;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const std::uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= static_cast<std::uint32_t>(off)) incl += v;
        }
        if (lane == 63) wave_sums[wave] = incl;
    }
    __syncthreads();
    if (q == 0) {
        std::uint32_t before = 0;
        for (std::uint32_t w = 0; w < wave; ++w) before += wave_sums[w];
        bases[p * 256 + d] = before + incl - total;
    }
}

__device__ __forceinline__ std::uint32_t status_load(const std::uint32_t* w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(std::uint32_t* w, std::uint32_t v) {
    __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One pass: count, publish, stage in digit order, look back, write out.
__global__ __launch_bounds__(kBlock) void radix_onesweep(const std::uint64_t* __restrict__ keys_in,
                                                         const std::uint32_t* __restrict__ vals_in,
                                                         std::uint64_t n, int shift,
                                                         const std::uint32_t* __restrict__ seg_of,
                                                         const std::uint32_t* __restrict__ bases,
                                                         std::uint32_t* __restrict__ status,
                                                         std::uint64_t* __restrict__ keys_out,
                                                         std::uint32_t* __restrict__ vals_out) {
    __shared__ std::uint64_t skeys[kKeysPerBlock];
    __shared__ std::uint32_t svals[kKeysPerBlock];
    __shared__ std::uint32_t out_base[256];       // global start of d - block-local start of d
    __shared__ std::uint32_t running[256];        // the block's digit counts, then next local slots
    __shared__ std::uint32_t wave_counts[4][256];  // digit counts of each wave in this tile
    __shared__ std::uint32_t wave_sums[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const std::uint32_t d0 = threadIdx.x;
    running[d0] = 0;
    for (int w = 0; w < 4; ++w) wave_counts[w][d0] = 0;
    __syncthreads();
    const std::uint32_t block = blockIdx.x;
    const std::uint64_t base = static_cast<std::uint64_t>(block) * kKeysPerBlock;
    const std::uint32_t in_block = static_cast<std::uint32_t>(
        n - base < static_cast<std::uint64_t>(kKeysPerBlock) ? n - base : kKeysPerBlock);
    std::uint64_t key[kTilesPerBlock];
    std::uint32_t val[kTilesPerBlock], dig[kTilesPerBlock];
#pragma unroll
    for (int t = 0; t < kTilesPerBlock; ++t) {
        const std::uint32_t j = static_cast<std::uint32_t>(t) * kBlock + threadIdx.x;
        const bool valid = j < in_block;
        key[t] = valid ? keys_in[base + j] : 0;
        val[t] = valid ? vals_in[base + j] : 0;
        dig[t] = valid ? digit_of(key[t], val[t], shift, seg_of) : 0u;
        if (valid) atomicAdd(&running[dig[t]], 1u);
    }
    __syncthreads();
    const std::uint32_t count_d = running[d0];
    std::uint32_t* const mine = status + static_cast<std::size_t>(block) * 256 + d0;
    status_store(mine, (block == 0 ? kPrefix : kAggregate) | count_d);
    // Block-local exclusive scan of the counts (as radix_scatter).
    std::uint32_t incl = count_d;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const std::uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    std::uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_sums[w];
    const std::uint32_t local_start = before + incl - count_d;
    running[d0] = local_start;
    __syncthreads();

    const std::uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int t = 0; t < kTilesPerBlock; ++t) {
        const bool valid = static_cast<std::uint32_t>(t) * kBlock + threadIdx.x < in_block;
        const std::uint32_t d = dig[t];
        std::uint64_t same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = ((d >> b) & 1u) != 0;
            const std::uint64_t m = __ballot(bit);
            same &= bit ? m : ~m;
        }
        const std::uint32_t rank = static_cast<std::uint32_t>(__popcll(same & lanes_below));
        if (valid && rank == 0) wave_counts[wave][d] = static_cast<std::uint32_t>(__popcll(same));
        __syncthreads();
        if (valid) {
            std::uint32_t slot = running[d] + rank;
            for (int w = 0; w < wave; ++w) slot += wave_counts[w][d];
            skeys[slot] = key[t];
            svals[slot] = val[t];
        }
        __syncthreads();
        const std::uint32_t total = wave_counts[0][threadIdx.x] + wave_counts[1][threadIdx.x] +
                                    wave_counts[2][threadIdx.x] + wave_counts[3][threadIdx.x];
        running[threadIdx.x] += total;
        for (int w = 0; w < 4; ++w) wave_counts[w][threadIdx.x] = 0;
        __syncthreads();
    }

    // Look back over the blocks before this one, digit d0 per thread: the words
    // of kLookback blocks are loaded together (one memory latency a window, not
    // one a block), then taken nearest first until an inclusive count.
    std::uint32_t prefix = 0;
    if (block > 0) {
        std::int64_t p = static_cast<std::int64_t>(block) - 1;
        bool found = false;
        while (!found) {
            std::uint32_t word[kLookback];
#pragma unroll
            for (int k = 0; k < kLookback; ++k) {
                // Before block 0: an empty inclusive count (block 0 publishes one first anyway).
                word[k] = p - k >= 0 ? status_load(status + static_cast<std::size_t>(p - k) * 256 + d0) : kPrefix;
            }
#pragma unroll
            for (int k = 0; k < kLookback; ++k) {
                if (!found) {
                    std::uint32_t w = word[k];
                    std::uint32_t spins = 0;
                    while ((w & (kAggregate | kPrefix)) == 0 && spins < kSpinCap) {
                        __builtin_amdgcn_s_sleep(1);
                        w = status_load(status + static_cast<std::size_t>(p - k) * 256 + d0);
                        ++spins;
                    }
                    prefix += w & kCountMask;
                    found = (w & kPrefix) != 0 || spins >= kSpinCap;
                }
            }
            p -= kLookback;
        }
        status_store(mine, kPrefix | (prefix + count_d));
    }
    out_base[d0] = bases[d0] + prefix - local_start;
    __syncthreads();
#pragma unroll 4
    for (std::uint32_t s = threadIdx.x; s < in_block; s += kBlock) {
        const std::uint64_t k = skeys[s];
        const std::uint32_t v = svals[s];
        const std::uint32_t pos = out_base[digit_of(k, v, shift, seg_of)] + s;
        keys_out[pos] = k;
        vals_out[pos] = v;
    }
}

// Small systems in one workgroup: extrema, cube, codes and the order by the
// codes' digits from first_digit up, instead of nine short launches.
// The multi-kernel order is a stable sort by those digits, that is a sort by
// (code >> 8 first_digit, index), a key unique per element; a bitonic network
// over the packed keys sorts them to the same permutation. first_digit >= 4
// keeps the digits within 32 bits beside the 32-bit index.
constexpr int kSmallThreads = 1024;
constexpr int kSmallSlots = static_cast<int>(kSmallOrder) / kSmallThreads;  // keys per thread
// Positions: 6 lane bits, 2 slot bits, 4 wave bits.
static_assert(kSmallOrder == 4096 && kSmallThreads == 1024 && kSmallSlots == 4, "small-sort layout");

// Compare-exchange of a key with its partner's: the smaller one where this
// position keeps the minimum, else the larger.
__device__ __forceinline__ unsigned long long keep(unsigned long long mine, unsigned long long other,
                                                   bool keep_min) {
    const bool other_less = other < mine;
    return other_less == keep_min ? other : mine;
}

__global__ __launch_bounds__(kSmallThreads) void morton_small(const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ z,
                                                              std::uint32_t n, int first_digit,
                                                              double* __restrict__ cube_out,
                                                              std::uint32_t* __restrict__ index) {
    __shared__ unsigned long long keys[kSmallOrder];
    __shared__ double wave_ext[kSmallThreads / 64][6];
    __shared__ double cube[4];
    const std::uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const double inf = __builtin_huge_val();
    double e[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (std::uint32_t i = t; i < n; i += kSmallThreads) {
        const double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            e[c] = v[c] < e[c] ? v[c] : e[c];
            e[3 + c] = v[c] > e[3 + c] ? v[c] : e[3 + c];
        }
    }
    // Extrema are exact, so folding them in any order gives the same values.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double a = __shfl_xor(e[c], off, 64);
            const double b = __shfl_xor(e[3 + c], off, 64);
            e[c] = a < e[c] ? a : e[c];
            e[3 + c] = b > e[3 + c] ? b : e[3 + c];
        }
    }
    if (lane == 0) {
        for (int c = 0; c < 6; ++c) wave_ext[wave][c] = e[c];
    }
    __syncthreads();
    if (t == 0) {
        double lo[3], hi[3];
        for (int c = 0; c < 3; ++c) {
            lo[c] = wave_ext[0][c];
            hi[c] = wave_ext[0][3 + c];
            for (std::uint32_t w = 1; w < kSmallThreads / 64; ++w) {
                lo[c] = wave_ext[w][c] < lo[c] ? wave_ext[w][c] : lo[c];
                hi[c] = wave_ext[w][3 + c] > hi[c] ? wave_ext[w][3 + c] : hi[c];
            }
        }
        make_cube(lo, hi, cube);
        for (int c = 0; c < 4; ++c) cube_out[c] = cube[c];
    }
    __syncthreads();
    // Codes as encode computes them, packed (top digits, index); padding last.
    // Thread t holds the positions whose bits 0-5 are its lane, 6-7 the slot
    // r and 8-11 its wave: bit b of a position is exchanged in 12 - b stages,
    // so the low bits, the busiest, stay in the wave or the thread.
    auto pos = [&](int r) { return lane | (static_cast<std::uint32_t>(r) << 6) | (wave << 8); };
    const double size = cube[3];
    const double scale = static_cast<double>(kGrid) / size;
    const int shift = 8 * first_digit;
    unsigned long long v[kSmallSlots];
#pragma unroll
    for (int r = 0; r < kSmallSlots; ++r) {
        const std::uint32_t i = pos(r);
        unsigned long long key = ~0ull;
        if (i < n) {
            std::uint64_t code = 0;
            if (size > 0) {
                code = (spread_bits(to_grid(x[i], cube[0], scale)) << 2) |
                       (spread_bits(to_grid(y[i], cube[1], scale)) << 1) |
                       spread_bits(to_grid(z[i], cube[2], scale));
            }
            key = ((code >> shift) << 32) | i;
        }
        v[r] = key;
    }
    // Bitonic sort, ascending, of all kSmallOrder positions: partners 256 or
    // more apart (another wave) through LDS, 64 or 128 apart in the same
    // thread, under 64 apart by wave shuffles.
    for (std::uint32_t k = 2; k <= kSmallOrder; k <<= 1) {
        std::uint32_t j = k >> 1;
        if (j >= 256) {
            __syncthreads();  // the last round's reads of keys are done
#pragma unroll
            for (int r = 0; r < kSmallSlots; ++r) keys[pos(r)] = v[r];
            __syncthreads();
            for (; j >= 256; j >>= 1) {
#pragma unroll
                for (int r = 0; r < kSmallSlots; ++r) {
                    const std::uint32_t i = pos(r);
                    const std::uint32_t partner = i ^ j;
                    if (partner > i) {
                        const unsigned long long a = keys[i], b = keys[partner];
                        const bool ascending = (i & k) == 0;
                        if ((a > b) == ascending) {
                            keys[i] = b;
                            keys[partner] = a;
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int r = 0; r < kSmallSlots; ++r) v[r] = keys[pos(r)];
        }
        for (; j >= 64; j >>= 1) {
            const int jr = static_cast<int>(j >> 6);
#pragma unroll
            for (int r = 0; r < kSmallSlots; ++r) {
                if ((r & jr) != 0) continue;
                const std::uint32_t i = pos(r);
                const bool ascending = (i & k) == 0;
                const unsigned long long a = v[r], b = v[r | jr];
                if ((a > b) == ascending) {
                    v[r] = b;
                    v[r | jr] = a;
                }
            }
        }
        for (; j > 0; j >>= 1) {
#pragma unroll
            for (int r = 0; r < kSmallSlots; ++r) {
                const std::uint32_t i = pos(r);
                const unsigned long long other = __shfl_xor(v[r], static_cast<int>(j), 64);
                const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
                v[r] = keep(v[r], other, keep_min);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kSmallSlots; ++r) {
        const std::uint32_t i = pos(r);
        if (i < n) index[i] = static_cast<std::uint32_t>(v[r]);
    }
}

}  // namespace

namespace {

// The one-sweep passes' words inside the histogram buffer: one status word
// per (pass, sort block, digit), the counting blocks' digit counts per pass,
// and each pass's digit bases.
struct SweepLayout {
    std::uint32_t count_blocks;
    std::size_t status, partial, bases, words;
};

SweepLayout sweep_layout(std::uint64_t n, std::uint32_t sort_blocks) {
    SweepLayout L{};
    L.count_blocks = static_cast<std::uint32_t>(
        std::min<std::uint64_t>(kCountBlocks, (n + kCountThreads - 1) / kCountThreads));
    L.status = 0;
    L.partial = L.status + static_cast<std::size_t>(kMaxSortPasses) * sort_blocks * 256;
    L.bases = L.partial + static_cast<std::size_t>(kMaxSortPasses) * L.count_blocks * 256;
    L.words = L.bases + static_cast<std::size_t>(kMaxSortPasses) * 256;
    return L;
}

}  // namespace

OrderSizes order_sizes(std::uint64_t n, std::uint32_t segments, std::uint32_t chunks, bool aos) {
    OrderSizes z{};
    z.bbox_blocks = segments > 1
        ? chunks
        : static_cast<std::uint32_t>(std::min<std::uint64_t>(1024, (n + kBlock - 1) / kBlock));
    z.sort_blocks = static_cast<std::uint32_t>((n + kKeysPerBlock - 1) / kKeysPerBlock);
    z.aos = aos ? sizeof(double4) * n : 0;
    z.bbox = sizeof(double) * (6 * static_cast<std::size_t>(z.bbox_blocks) + 4 * static_cast<std::size_t>(segments));
    z.codes = sizeof(std::uint64_t) * n;
    z.index = sizeof(std::uint32_t) * n;
    const std::size_t hist_words = 256 * static_cast<std::size_t>(z.sort_blocks);
    z.hist = sizeof(std::uint32_t) * std::max(hist_words + scan_scratch_words(hist_words),
                                              sweep_layout(n, z.sort_blocks).words);
    z.total = z.aos + z.bbox + 2 * z.codes + 2 * z.index + z.hist;
    return z;
}

void morton_enqueue(orr_ctx* ctx, const double* x, const double* y, const double* z,
                    std::uint64_t n, int first_digit) {
    morton_enqueue_segments(ctx, x, y, z, n, SegmentView{}, nullptr, first_digit);
}

void morton_enqueue_segments(orr_ctx* ctx, const double* x, const double* y, const double* z,
                             std::uint64_t n, const SegmentView& segs, const double* m,
                             int first_digit) {
    if (n == 0) return;
    if (n > (std::uint64_t{1} << 31)) fail(ORR_ERR_LENGTH, "morton: more than 2^31 particles");
    const OrderSizes sz = order_sizes(n, segs.count, segs.chunk_count, m != nullptr);
    ctx->order_bbox.reserve(sz.bbox);
    if (m != nullptr) ctx->order_aos.reserve(sz.aos);
    ctx->order_codes.reserve(sz.codes);
    ctx->order_codes_alt.reserve(sz.codes);
    ctx->order_index.reserve(sz.index);
    ctx->order_index_alt.reserve(sz.index);
    ctx->order_hist.reserve(sz.hist);
    OrderScratch b;
    b.bbox = ctx->order_bbox.as<double>();
    b.aos = m != nullptr ? ctx->order_aos.as<double4>() : nullptr;
    b.codes[0] = ctx->order_codes.as<std::uint64_t>();
    b.codes[1] = ctx->order_codes_alt.as<std::uint64_t>();
    b.index[0] = ctx->order_index.as<std::uint32_t>();
    b.index[1] = ctx->order_index_alt.as<std::uint32_t>();
    b.hist = ctx->order_hist.as<std::uint32_t>();
    if (morton_enqueue_into(ctx, x, y, z, n, segs, m, first_digit, b, true) != 0) {
        // An odd pass count ends in the alternate buffers: swap the roles.
        std::swap(ctx->order_codes, ctx->order_codes_alt);
        std::swap(ctx->order_index, ctx->order_index_alt);
    }
}

int morton_enqueue_into(orr_ctx* ctx, const double* x, const double* y, const double* z,
                        std::uint64_t n, const SegmentView& segs, const double* m,
                        int first_digit, const OrderScratch& b, bool timed) {
    if (n == 0) return 0;
    if (n > (std::uint64_t{1} << 31)) fail(ORR_ERR_LENGTH, "morton: more than 2^31 particles");
    const bool segmented = segs.count > 1;
    const OrderSizes sz = order_sizes(n, segs.count, segs.chunk_count, m != nullptr);
    const std::uint32_t bbox_blocks = sz.bbox_blocks;
    const std::uint32_t sort_blocks = sz.sort_blocks;
    double* cube = b.bbox;  // cubes[4 * segment]
    double* partial = cube + 4 * static_cast<std::size_t>(segs.count);
    {
        PhaseTimer t(ctx, kBbox, timed);
        if (segmented) {
            hipLaunchKernelGGL(bbox_chunks, dim3(segs.chunk_count), dim3(kBlock), 0, ctx->stream, x, y,
                               z, segs.chunks, partial);
            ORR_LAUNCHED("bbox_chunks");
            hipLaunchKernelGGL(bbox_segments, dim3((segs.count + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               ctx->stream, partial, segs.chunk_first, segs.count, cube);
            ORR_LAUNCHED("bbox_segments");
        } else {
            hipLaunchKernelGGL(bbox_partial, dim3(bbox_blocks), dim3(kBlock), 0, ctx->stream, x, y, z,
                               n, partial);
            ORR_LAUNCHED("bbox_partial");
            hipLaunchKernelGGL(bbox_final, dim3(1), dim3(kBlock), 0, ctx->stream, partial,
                               bbox_blocks, cube);
            ORR_LAUNCHED("bbox_final");
        }
    }
    std::uint64_t* const* keys = b.codes;
    std::uint32_t* const* vals = b.index;
    const std::uint32_t* seg_of = segmented ? segs.seg_of : nullptr;
    {
        PhaseTimer t(ctx, kEncode, timed);
        hipLaunchKernelGGL(encode, dim3(static_cast<std::uint32_t>((n + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, ctx->stream, x, y, z, m, n, cube, seg_of, keys[0], vals[0],
                           m != nullptr ? b.aos : nullptr);
        ORR_LAUNCHED("encode");
    }
    int cur = 0;
    {
        PhaseTimer t(ctx, kSort, timed);
        std::uint32_t* hist = b.hist;
        // Eight passes over the 63-bit codes (from first_digit: the order of the
        // top digits only), then (segments > 1) one or two passes over the
        // segment of each element: segment-major, then (code, index).
        int passes = 8;
        if (segmented) passes += segs.count > 256 ? 2 : 1;
        if (ctx->sort_mode == 1 && n <= kCountMask) {
            const SweepLayout L = sweep_layout(n, sort_blocks);
            std::uint32_t* const status = hist + L.status;
            const std::size_t per_pass = static_cast<std::size_t>(sort_blocks) * 256;
            hipLaunchKernelGGL(radix_count_all, dim3(L.count_blocks), dim3(kCountThreads), 0, ctx->stream,
                               keys[0], vals[0], n, first_digit, passes, seg_of, hist + L.partial,
                               status + first_digit * per_pass,
                               static_cast<std::uint64_t>(passes - first_digit) * per_pass);
            ORR_LAUNCHED("radix_count_all");
            hipLaunchKernelGGL(radix_bases, dim3(passes - first_digit), dim3(kBaseThreads), 0, ctx->stream,
                               hist + L.partial, L.count_blocks, first_digit, hist + L.bases);
            ORR_LAUNCHED("radix_bases");
            for (int pass = first_digit; pass < passes; ++pass) {
                const bool by_segment = pass >= 8;
                const int shift = by_segment ? 8 * (pass - 8) : 8 * pass;
                hipLaunchKernelGGL(radix_onesweep, dim3(sort_blocks), dim3(kBlock), 0, ctx->stream,
                                   keys[cur], vals[cur], n, shift, by_segment ? seg_of : nullptr,
                                   hist + L.bases + 256 * pass, status + pass * per_pass, keys[1 - cur],
                                   vals[1 - cur]);
                ORR_LAUNCHED("radix_onesweep");
                cur = 1 - cur;
            }
            return cur;
        }
        for (int pass = first_digit; pass < passes; ++pass) {
            const bool by_segment = pass >= 8;
            const int shift = by_segment ? 8 * (pass - 8) : 8 * pass;
            const std::uint32_t* digits = by_segment ? seg_of : nullptr;
            hipLaunchKernelGGL(radix_histogram, dim3(sort_blocks), dim3(kBlock), 0, ctx->stream,
                               keys[cur], vals[cur], n, shift, digits, hist);
            ORR_LAUNCHED("radix_histogram");
            exclusive_scan_u32(ctx, hist, hist, 256u * sort_blocks, hist + 256u * sort_blocks, nullptr);
            hipLaunchKernelGGL(radix_scatter, dim3(sort_blocks), dim3(kBlock), 0, ctx->stream,
                               keys[cur], vals[cur], n, shift, digits, hist, keys[1 - cur],
                               vals[1 - cur]);
            ORR_LAUNCHED("radix_scatter");
            cur = 1 - cur;
        }
    }
    return cur;
}

void morton_enqueue_small(orr_ctx* ctx, const double* x, const double* y, const double* z,
                          std::uint64_t n, int first_digit) {
    if (n == 0) return;
    if (n > kSmallOrder) fail(ORR_ERR_INVALID_ARGUMENT, "morton_enqueue_small: more than 4096 particles");
    if (first_digit < 4) fail(ORR_ERR_INVALID_ARGUMENT, "morton_enqueue_small: first_digit must be at least 4");
    ctx->order_bbox.reserve(4 * sizeof(double));
    ctx->order_index.reserve(sizeof(std::uint32_t) * n);
    PhaseTimer t(ctx, kSort);
    hipLaunchKernelGGL(morton_small, dim3(1), dim3(kSmallThreads), 0, ctx->stream, x, y, z,
                       static_cast<std::uint32_t>(n), first_digit, ctx->order_bbox.as<double>(),
                       ctx->order_index.as<std::uint32_t>());
    ORR_LAUNCHED("morton_small");
}

void morton_cube(orr_ctx* ctx, double cube[4]) {
    ORR_HIP(hipMemcpyAsync(cube, ctx->order_bbox.ptr, 4 * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    ctx->sync();
}

}  // namespace orr
