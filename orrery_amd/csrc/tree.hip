// Barnes-Hut on gfx950: octree build from sorted Morton keys, FP64 monopole
// moments, and a wave-coherent theta-criterion walk.
//
// The tree is the reference's (src/solvers/octree.cpp:196-468): cells split in
// octant order, empty cells skipped, a leaf when a cell holds <= leaf_capacity
// particles or sits at level 21, nodes in depth-first order with `next` the
// index just past the subtree. The reference builds it by recursion; here every
// node is derived independently from the sorted keys:
//
//   d(k)   the first level at which key k differs from key k-1 (d(0) = 0, 22
//          for a duplicate key): a cell of level l starts at particle k exactly
//          when d(k) <= l.
//   chain  the cells that start at k form a chain of levels d(k), d(k)+1, ...
//          Its first cell is a node iff its parent (level d(k)-1) holds more
//          than leaf_capacity particles (found by scanning +-leaf_capacity keys
//          around k); every deeper cell of the chain is a node while the one
//          above it holds more than leaf_capacity, i.e. down to level
//          min(21, max(d(k), min d over the next leaf_capacity keys)), a leaf.
//   DFS    depth-first order sorts nodes by (first particle, level), so an
//          exclusive scan of chain lengths numbers the nodes, and the node
//          after a subtree is the first node starting at the cell's end:
//          next = first_node[end].
//
// Moments follow octree.cpp:91-185 operation for operation in FP64 (leaf: sums
// over particles in index order; internal: over children in DFS order), with
// contraction disabled, so masses, centres of mass and acceptance radii
// (octree.cpp:374-387) are the reference's bit for bit.
//
// The walk (tree_walk.cpp:20-83 per target) runs one wave over 64 consecutive
// sorted targets. The wave visits node min(next node of each lane) -- each lane
// still follows exactly its own walk, as the reference's coherent CUDA walk
// does (ADR-0029) -- with the node read through the scalar unit. Accepted cells
// use FP32 positions relative to the root centre; leaf particles are stored as
// FP32 offsets from their leaf's geometric centre and a lane interacting with a
// leaf re-expresses its own offset against that centre in FP64, so close pairs
// keep the precision of the leaf size. Per-lane partial sums are Kahan-folded
// after each leaf.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "engine.hpp"

#pragma clang fp contract(off)

namespace orr {
namespace {

constexpr int kBlock = 256;
constexpr unsigned kLevels = 21;
constexpr std::uint8_t kNoBoundary = 22;
constexpr int kMaxChildren = 8;                 // octants
constexpr std::uint32_t kNoChild = 0xffffffffu;
// Spare float4 entries past the last particle of leaf_rel: a leaf's sources are
// read in rounds of four.
constexpr std::size_t kLeafPad = 3;

struct WalkNode {
    float4 a;  // centre of mass relative to the root centre (FP32), mass
    uint4 b;   // acceptance radius^2 (FP32 bits), next, first particle, particle count
};

struct QuadF {
    float4 a;  // xx, xy, xz, yy
    float4 b;  // yz, zz, 0, 0
};

// Device counters, uint32: [0] nodes, [1] leaves, [2] depth.
// Device statistics, uint64: [0] particle-particle, [1] particle-cell, [2] lane visits,
// [3] wave steps.

__device__ inline std::uint32_t grid_stride_start() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline std::uint32_t grid_stride_step() { return gridDim.x * blockDim.x; }

__global__ void clear_counters(std::uint32_t* __restrict__ c32, unsigned long long* __restrict__ c64,
                               std::uint32_t count64) {
    if (threadIdx.x < 4) c32[threadIdx.x] = 0;
    if (c64 != nullptr) {
        for (std::uint32_t k = threadIdx.x; k < count64; k += blockDim.x) c64[k] = 0;
    }
}

// d(k) from adjacent sorted codes; 0 where a system starts (each system of a
// batch is its own tree, and a level-0 boundary ends every window at it).
__global__ __launch_bounds__(kBlock) void tree_levels(const std::uint64_t* __restrict__ codes,
                                                      const std::uint32_t* __restrict__ seg_of,
                                                      std::uint32_t n, std::uint8_t* __restrict__ d) {
    const std::uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    if (k == 0 || (seg_of != nullptr && seg_of[k] != seg_of[k - 1])) {
        d[k] = 0;
        return;
    }
    const std::uint64_t x = codes[k] ^ codes[k - 1];
    if (x == 0) {
        d[k] = kNoBoundary;
        return;
    }
    const unsigned h = 63u - static_cast<unsigned>(__clzll(x));  // highest differing bit
    d[k] = static_cast<std::uint8_t>(kLevels - h / 3u);
}

// Chain length of the nodes that start at particle k; leaf level in `leaf_level`.
__global__ __launch_bounds__(kBlock) void tree_chain(const std::uint8_t* __restrict__ d,
                                                     std::uint32_t n, std::uint32_t cap,
                                                     std::uint32_t* __restrict__ chain,
                                                     std::uint8_t* __restrict__ leaf_level) {
    const std::uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const unsigned dk = d[k];
    std::uint32_t len = 0;
    unsigned lleaf = dk;
    if (dk <= kLevels) {
        bool head = true;
        if (dk > 0) {  // d = 0: the root of its system
            // Parent cell of level dk-1 spans [p, q) around k; it is a leaf (so
            // no node starts at k) iff q - p <= cap.
            const unsigned P = dk - 1;
            const std::uint32_t lo = k + 1 > cap ? k + 1 - cap : 0;
            std::int64_t p = -1;
            for (std::int64_t j = static_cast<std::int64_t>(k) - 1; j >= static_cast<std::int64_t>(lo); --j) {
                if (d[j] <= P) {
                    p = j;
                    break;
                }
            }
            if (p >= 0) {
                const std::uint64_t limit = static_cast<std::uint64_t>(p) + cap;  // q must be <= limit
                std::uint64_t q = n;
                const std::uint64_t stop = limit < n - 1 ? limit : n - 1;
                for (std::uint64_t j = k + 1; j <= stop; ++j) {
                    if (d[j] <= P) {
                        q = j;
                        break;
                    }
                }
                head = !(q <= limit);
            }
        }
        if (head) {
            if (n - k <= cap) {
                lleaf = dk;
            } else {
                unsigned w = kNoBoundary;
                for (std::uint32_t j = k + 1; j <= k + cap; ++j) w = min(w, static_cast<unsigned>(d[j]));
                lleaf = min(kLevels, max(dk, w));
            }
            len = lleaf - dk + 1;
        }
    }
    chain[k] = len;
    leaf_level[k] = static_cast<std::uint8_t>(lleaf);
}

// First index j >= lo whose code prefix of `level` exceeds that of `key`.
__device__ std::uint32_t prefix_end(const std::uint64_t* __restrict__ codes, std::uint32_t n,
                                    std::uint32_t lo, std::uint64_t key, unsigned level) {
    if (level == 0) return n;
    const unsigned s = 3u * (kLevels - level);
    const std::uint64_t want = key >> s;
    if (lo >= n || (codes[lo] >> s) != want) return lo;
    // Gallop, then bisect: codes[lo] shares the prefix.
    std::uint32_t good = lo;
    std::uint32_t step = 1;
    std::uint32_t bad = n;
    while (true) {
        const std::uint64_t probe = static_cast<std::uint64_t>(good) + step;
        if (probe >= n) break;
        if ((codes[probe] >> s) != want) {
            bad = static_cast<std::uint32_t>(probe);
            break;
        }
        good = static_cast<std::uint32_t>(probe);
        step <<= 1;
    }
    while (bad - good > 1) {
        const std::uint32_t mid = good + (bad - good) / 2;
        if ((codes[mid] >> s) != want) bad = mid; else good = mid;
    }
    return bad;
}

// Each chain's nodes: start particle and level (one thread per chain start; no
// searches here, so no thread waits on another's).
__global__ __launch_bounds__(kBlock) void tree_chains(
    const std::uint8_t* __restrict__ d, const std::uint8_t* __restrict__ leaf_level,
    const std::uint32_t* __restrict__ chain, const std::uint32_t* __restrict__ first_node,
    std::uint32_t n, uint2* __restrict__ block_counts, std::uint32_t* __restrict__ node_begin,
    std::uint8_t* __restrict__ node_level) {
    __shared__ unsigned s_leaves[kBlock / 64], s_depth[kBlock / 64];
    const std::uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    const bool owns = k < n && chain[k] != 0;
    unsigned depth = 0;
    if (owns) {
        const unsigned dk = d[k], lleaf = leaf_level[k];
        const std::uint32_t base = first_node[k];
        for (unsigned l = dk; l <= lleaf; ++l) {
            node_begin[base + (l - dk)] = k;
            node_level[base + (l - dk)] = static_cast<std::uint8_t>(l);
        }
        depth = lleaf;
    }
    // One leaf per chain: each block writes its leaf count and deepest level to
    // its own slot (atomics on one address would serialise every wave).
    const unsigned long long leaves = __ballot(owns);
    unsigned deepest = depth;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) deepest = max(deepest, static_cast<unsigned>(__shfl_xor(static_cast<int>(deepest), off, 64)));
    if ((threadIdx.x & 63) == 0) {
        s_leaves[threadIdx.x / 64] = static_cast<unsigned>(__popcll(leaves));
        s_depth[threadIdx.x / 64] = deepest;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned total = 0, deep = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            total += s_leaves[w];
            deep = max(deep, s_depth[w]);
        }
        block_counts[blockIdx.x] = make_uint2(total, deep);
    }
}

// One block: leaf count and depth from the per-block slots.
__global__ __launch_bounds__(kBlock) void tree_counts(const uint2* __restrict__ block_counts,
                                                      std::uint32_t blocks,
                                                      std::uint32_t* __restrict__ counters) {
    __shared__ unsigned s_total[kBlock], s_deep[kBlock];
    unsigned total = 0, deep = 0;
    for (std::uint32_t b = threadIdx.x; b < blocks; b += kBlock) {
        total += block_counts[b].x;
        deep = max(deep, block_counts[b].y);
    }
    s_total[threadIdx.x] = total;
    s_deep[threadIdx.x] = deep;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            s_total[threadIdx.x] += s_total[threadIdx.x + w];
            s_deep[threadIdx.x] = max(s_deep[threadIdx.x], s_deep[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counters[1] = s_total[0];
        counters[2] = s_deep[0];
    }
}

// One thread per node: its cell's end (a prefix search from its start, so a
// node never waits on another), escape index, geometric centre (octree.cpp:41-47:
// one halving per level from the root centre of morton.hpp:93-95) and, for a
// leaf, the leaf id of its particles.
__global__ __launch_bounds__(kBlock) void tree_node_fill(
    const std::uint64_t* __restrict__ codes, const std::uint8_t* __restrict__ leaf_level,
    const std::uint32_t* __restrict__ first_node, std::uint32_t n,
    const std::uint32_t* __restrict__ seg_of, const std::uint32_t* __restrict__ seg_start,
    const double* __restrict__ cubes, const std::uint32_t* __restrict__ counters,
    const std::uint32_t* __restrict__ node_begin, const std::uint8_t* __restrict__ node_level,
    std::uint32_t* __restrict__ node_count, std::uint32_t* __restrict__ node_next,
    double4* __restrict__ node_geo, std::uint32_t* __restrict__ leaf_of) {
    const std::uint32_t M = counters[0];
    for (std::uint32_t node = grid_stride_start(); node < M; node += grid_stride_step()) {
        const std::uint32_t k = node_begin[node];
        const unsigned l = node_level[node];
        const std::uint64_t key = codes[k];
        const std::uint32_t seg = seg_of != nullptr ? seg_of[k] : 0u;
        const std::uint32_t seg_end = seg_start != nullptr ? seg_start[seg + 1] : n;
        const double* cube = cubes + 4 * static_cast<std::size_t>(seg);
        const std::uint32_t end = prefix_end(codes, seg_end, k + 1, key, l);
        node_count[node] = end - k;
        node_next[node] = end >= n ? M : first_node[end];
        double size = cube[3];
        double c[3] = {cube[0] + size / 2, cube[1] + size / 2, cube[2] + size / 2};
        for (unsigned s = 1; s <= l; ++s) {
            const unsigned oct = static_cast<unsigned>((key >> (3u * (kLevels - s))) & 7u);
            const double q = size / 4;
            c[0] = c[0] + ((oct & 4u) != 0 ? q : -q);
            c[1] = c[1] + ((oct & 2u) != 0 ? q : -q);
            c[2] = c[2] + ((oct & 1u) != 0 ? q : -q);
            size = size / 2;
        }
        node_geo[node] = make_double4(c[0], c[1], c[2], size);
        if (l == leaf_level[k]) {
            for (std::uint32_t j = k; j < end; ++j) leaf_of[j] = node;
        }
    }
}

// Sorted FP64 particles for the moments; FP32 walk data: position relative to
// the root centre, and offset from the particle's leaf centre.
__global__ __launch_bounds__(kBlock) void tree_gather(
    const double4* __restrict__ aos, const std::uint32_t* __restrict__ index, std::uint32_t n,
    const double* __restrict__ cubes, const std::uint32_t* __restrict__ seg_of,
    const std::uint32_t* __restrict__ leaf_of,
    const double4* __restrict__ node_geo, double4* __restrict__ sorted,
    float4* __restrict__ root_rel, float4* __restrict__ leaf_rel) {
    const std::uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // One 32-byte record per particle ({x, y, z, m} written in index order by
    // the encode pass) rather than four scattered FP64 loads.
    const double4 p = aos[index[i]];
    const double px = p.x, py = p.y, pz = p.z, pm = p.w;
    sorted[i] = p;
    const double* cube = cubes + (seg_of != nullptr ? 4 * static_cast<std::size_t>(seg_of[i]) : 0);
    const double half = cube[3] / 2;
    const double c0x = cube[0] + half, c0y = cube[1] + half, c0z = cube[2] + half;
    root_rel[i] = make_float4(static_cast<float>(px - c0x), static_cast<float>(py - c0y),
                              static_cast<float>(pz - c0z), static_cast<float>(pm));
    const double4 g = node_geo[leaf_of[i]];
    leaf_rel[i] = make_float4(static_cast<float>(px - g.x), static_cast<float>(py - g.y),
                              static_cast<float>(pz - g.z), static_cast<float>(pm));
}

// The walk's escape index: the DFS successor, or "done" when the node's cell
// ends its system (the successor then belongs to the next system or past M).
__device__ inline std::uint32_t walk_escape(std::uint32_t next, std::uint32_t cell_end,
                                            std::uint32_t system_end) {
    return cell_end >= system_end ? 0xffffffffu : next;
}

__device__ inline void finish_node(std::uint32_t node, double mass, double cx, double cy, double cz,
                                   const double4 geo, double theta, const double* __restrict__ cube,
                                   bool leaf, std::uint32_t begin, std::uint32_t count,
                                   std::uint32_t next, double4* __restrict__ moment,
                                   WalkNode* __restrict__ walk) {
    moment[node] = make_double4(cx, cy, cz, mass);
    // octree.cpp:370-371 and :374-387.
    const double ox = cx - geo.x, oy = cy - geo.y, oz = cz - geo.z;
    const double offset = sqrt((ox * ox) + (oy * oy) + (oz * oz));
    double acc;
    if (!(theta > 0)) {
        acc = __builtin_huge_val();
    } else {
        const double r = (geo.w / theta) + offset;
        acc = r * r;
    }
    const double half = cube[3] / 2;
    WalkNode w;
    w.a = make_float4(static_cast<float>(cx - (cube[0] + half)), static_cast<float>(cy - (cube[1] + half)),
                      static_cast<float>(cz - (cube[2] + half)), static_cast<float>(mass));
    float accf = static_cast<float>(acc);
    std::uint32_t bits;
    std::memcpy(&bits, &accf, sizeof(bits));
    w.b = make_uint4(bits, next, leaf ? begin : 0u, leaf ? count : 0u);
    walk[node] = w;
}

// octree.cpp:75-84: traceless second moment of a point mass displaced by d.
__device__ inline void add_displaced(double q[6], double mass, double dx, double dy, double dz) {
    const double trace = mass * (((dx * dx) + (dy * dy)) + (dz * dz));
    q[0] += (3 * mass * dx * dx) - trace;
    q[1] += 3 * mass * dx * dy;
    q[2] += 3 * mass * dx * dz;
    q[3] += (3 * mass * dy * dy) - trace;
    q[4] += 3 * mass * dy * dz;
    q[5] += (3 * mass * dz * dz) - trace;
}

__device__ inline void store_quad(std::uint32_t node, const double q[6], double* __restrict__ quad64,
                                  QuadF* __restrict__ quadf) {
    for (int k = 0; k < 6; ++k) quad64[6 * static_cast<std::size_t>(node) + k] = q[k];
    QuadF f;
    f.a = make_float4(static_cast<float>(q[0]), static_cast<float>(q[1]), static_cast<float>(q[2]),
                      static_cast<float>(q[3]));
    f.b = make_float4(static_cast<float>(q[4]), static_cast<float>(q[5]), 0.f, 0.f);
    quadf[node] = f;
}

// octree.cpp:91-135 (leaf): index-order sums, fallback to the first particle;
// with quadrupoles, the displaced moments of the particles about the centre of mass.
__global__ __launch_bounds__(kBlock) void leaf_moments(
    const std::uint32_t* __restrict__ counters, const std::uint8_t* __restrict__ node_level,
    const std::uint8_t* __restrict__ leaf_level_of_first, const std::uint32_t* __restrict__ node_begin,
    const std::uint32_t* __restrict__ node_count, const std::uint32_t* __restrict__ node_next,
    const double4* __restrict__ node_geo, const double4* __restrict__ sorted, double theta,
    const double* __restrict__ cubes, const std::uint32_t* __restrict__ seg_of,
    const std::uint32_t* __restrict__ seg_start, std::uint32_t n,
    double4* __restrict__ moment, WalkNode* __restrict__ walk,
    double* __restrict__ quad64, QuadF* __restrict__ quadf) {
    const std::uint32_t M = counters[0];
    for (std::uint32_t node = grid_stride_start(); node < M; node += grid_stride_step()) {
        const std::uint32_t b = node_begin[node];
        if (node_level[node] != leaf_level_of_first[b]) continue;  // internal
        const std::uint32_t seg = seg_of != nullptr ? seg_of[b] : 0u;
        const double* cube = cubes + 4 * static_cast<std::size_t>(seg);
        const std::uint32_t system_end = seg_start != nullptr ? seg_start[seg + 1] : n;
        const std::uint32_t e = b + node_count[node];
        double mass = 0, wx = 0, wy = 0, wz = 0;
        for (std::uint32_t i = b; i < e; ++i) {
            const double4 p = sorted[i];
            mass += p.w;
            wx += p.x * p.w;
            wy += p.y * p.w;
            wz += p.z * p.w;
        }
        double cx, cy, cz;
        if (mass > 0) {
            cx = wx / mass;
            cy = wy / mass;
            cz = wz / mass;
        } else {
            cx = sorted[b].x;
            cy = sorted[b].y;
            cz = sorted[b].z;
        }
        finish_node(node, mass, cx, cy, cz, node_geo[node], theta, cube, true, b, e - b,
                    walk_escape(node_next[node], e, system_end), moment, walk);
        if (quad64 != nullptr) {
            double q[6] = {0, 0, 0, 0, 0, 0};
            for (std::uint32_t i = b; i < e; ++i) {
                const double4 p = sorted[i];
                add_displaced(q, p.w, p.x - cx, p.y - cy, p.z - cz);
            }
            store_quad(node, q, quad64, quadf);
        }
    }
}

// The children of every internal node in DFS order (the next chain from
// node + 1 to the node's own next), written once for all levels in one launch
// so that the per-level moment passes load them side by side instead of
// following the chain one dependent load at a time.
__global__ __launch_bounds__(kBlock) void tree_children(
    const std::uint32_t* __restrict__ counters, const std::uint8_t* __restrict__ node_level,
    const std::uint8_t* __restrict__ leaf_level_of_first, const std::uint32_t* __restrict__ node_begin,
    const std::uint32_t* __restrict__ node_next, std::uint32_t* __restrict__ children) {
    const std::uint32_t M = counters[0];
    for (std::uint32_t node = grid_stride_start(); node < M; node += grid_stride_step()) {
        if (node_level[node] == leaf_level_of_first[node_begin[node]]) continue;  // a leaf
        const std::uint32_t end = node_next[node];
        std::uint32_t* out = children + static_cast<std::size_t>(kMaxChildren) * node;
        int k = 0;
        for (std::uint32_t c = node + 1; c < end && k < kMaxChildren; c = max(node_next[c], c + 1)) out[k++] = c;
        for (; k < kMaxChildren; ++k) out[k] = kNoChild;
    }
}

// octree.cpp:142-185 (internal): children in DFS order via the next chain;
// quadrupoles by the parallel-axis theorem, children's moments first.
__global__ __launch_bounds__(kBlock) void internal_moments(
    unsigned level, const std::uint32_t* __restrict__ counters,
    const std::uint8_t* __restrict__ node_level, const std::uint8_t* __restrict__ leaf_level_of_first,
    const std::uint32_t* __restrict__ node_begin, const std::uint32_t* __restrict__ node_count,
    const std::uint32_t* __restrict__ node_next, const std::uint32_t* __restrict__ children,
    const double4* __restrict__ node_geo, double theta, const double* __restrict__ cubes,
    const std::uint32_t* __restrict__ seg_of, const std::uint32_t* __restrict__ seg_start,
    std::uint32_t n, double4* __restrict__ moment,
    WalkNode* __restrict__ walk, double* __restrict__ quad64, QuadF* __restrict__ quadf) {
    const std::uint32_t M = counters[0];
    for (std::uint32_t node = grid_stride_start(); node < M; node += grid_stride_step()) {
        if (node_level[node] != level) continue;
        const std::uint32_t b = node_begin[node];
        if (leaf_level_of_first[b] == level) continue;  // a leaf
        const std::uint32_t seg = seg_of != nullptr ? seg_of[b] : 0u;
        const double* cube = cubes + 4 * static_cast<std::size_t>(seg);
        const std::uint32_t system_end = seg_start != nullptr ? seg_start[seg + 1] : n;
        // The children's indices, then their moments, loaded side by side; the
        // sums below still run in DFS order.
        const std::uint32_t* list = children + static_cast<std::size_t>(kMaxChildren) * node;
        std::uint32_t child[kMaxChildren];
        double4 cms[kMaxChildren];
#pragma unroll
        for (int k = 0; k < kMaxChildren; ++k) child[k] = list[k];
#pragma unroll
        for (int k = 0; k < kMaxChildren; ++k) {
            if (child[k] != kNoChild) cms[k] = moment[child[k]];
        }
        double mass = 0, wx = 0, wy = 0, wz = 0, mx = 0, my = 0, mz = 0;
        std::uint32_t count = 0;
#pragma unroll
        for (int k = 0; k < kMaxChildren; ++k) {
            if (child[k] == kNoChild) continue;
            const double4 cm = cms[k];
            mass += cm.w;
            wx += cm.x * cm.w;
            wy += cm.y * cm.w;
            wz += cm.z * cm.w;
            mx += cm.x;
            my += cm.y;
            mz += cm.z;
            ++count;
        }
        double cx, cy, cz;
        if (mass > 0) {
            cx = wx / mass;
            cy = wy / mass;
            cz = wz / mass;
        } else {
            const double k = static_cast<double>(count == 0 ? 1 : count);
            cx = mx / k;
            cy = my / k;
            cz = mz / k;
        }
        finish_node(node, mass, cx, cy, cz, node_geo[node], theta, cube, false, b, 0,
                    walk_escape(node_next[node], b + node_count[node], system_end), moment, walk);
        if (quad64 != nullptr) {
            double q[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < kMaxChildren; ++k) {
                if (child[k] == kNoChild) continue;
                const double* cq = quad64 + 6 * static_cast<std::size_t>(child[k]);
                for (int j = 0; j < 6; ++j) q[j] += cq[j];
                const double4 cm = cms[k];
                add_displaced(q, cm.w, cm.x - cx, cm.y - cy, cm.z - cz);
            }
            store_quad(node, q, quad64, quadf);
        }
    }
}

// Minimum over the 64 lanes of a wave (all lanes active). Within each row of
// 16 lanes, four DPP row shifts make lane 15 of the row hold the row minimum
// (shifted-in lanes outside the row read the identity ~0); the four row minima
// are then combined in scalar registers.
template<int kCtrl> __device__ __forceinline__ std::uint32_t dpp_min(std::uint32_t v) {
    const std::uint32_t moved = static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(-1, static_cast<int>(v), kCtrl, 0xf, 0xf, false));
    return min(v, moved);
}

__device__ inline std::uint32_t wave_min_u32(std::uint32_t v) {
    v = dpp_min<0x111>(v);  // row_shr:1
    v = dpp_min<0x112>(v);  // row_shr:2
    v = dpp_min<0x114>(v);  // row_shr:4
    v = dpp_min<0x118>(v);  // row_shr:8
    const std::uint32_t r0 = __builtin_amdgcn_readlane(static_cast<int>(v), 15);
    const std::uint32_t r1 = __builtin_amdgcn_readlane(static_cast<int>(v), 31);
    const std::uint32_t r2 = __builtin_amdgcn_readlane(static_cast<int>(v), 47);
    const std::uint32_t r3 = __builtin_amdgcn_readlane(static_cast<int>(v), 63);
    return min(min(r0, r1), min(r2, r3));
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void kahan(float p, float& s, float& c) {
    const float y = p - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Workgroups are dispatched to the 8 XCDs in turn (block b to XCD b mod 8),
// and each XCD has its own L2. Neighbouring targets in Morton order visit the
// same deep nodes and leaves, so block b walks the logical block that gives
// each XCD runs of kXcdRun consecutive logical blocks, the runs dealt to the
// XCDs in turn: the nodes a wave pulls into an L2 are the ones its
// neighbours on that XCD need, while every XCD still sees all parts of the
// system (one contiguous eighth each would leave the XCD with the densest part
// finishing last). A bijection of [0, blocks): blocks past the last whole
// round of runs keep their own index.
// Up to this many full (64-lane) walk waves -- two per SIMD of the 256 CUs --
// half-lane waves are faster (quadrupole walk, two-disc collision: 1.088 ->
// 0.981 ms at N = 65,536, 1.328 -> 1.265 ms at 131,072); above it they are
// slower (1.813 -> 2.194 ms at 262,144): the walk is bound by issue, not by
// latency, once each SIMD has a few waves.
constexpr std::uint64_t kHalfLaneWaves = 2 * 4 * 256;

constexpr std::uint32_t kXcds = 8;
constexpr std::uint32_t kXcdRun = 16;
__device__ __forceinline__ std::uint32_t xcd_block(std::uint32_t b, std::uint32_t blocks) {
    const std::uint32_t whole = blocks / (kXcds * kXcdRun) * (kXcds * kXcdRun);
    if (b >= whole) return b;
    const std::uint32_t x = b % kXcds, k = b / kXcds;
    return ((k / kXcdRun) * kXcds + x) * kXcdRun + k % kXcdRun;
}

// One target's walk state: FP32 position relative to the root centre and
// offset from its leaf centre, its leaf, the next node it needs, Kahan sums.
struct WalkTarget {
    float4 xi, oi;
    std::uint32_t own, mine;
    float sx, sy, sz, kx, ky, kz, px, py, pz;
    // Per target, pp < N, pc <= M and visits <= M: 32 bits suffice.
    std::uint32_t pp, pc, visits;
};

// The target's step at node `cur` if it is the node it needs (tree_walk.cpp:20-83):
// accept (monopole, optionally quadrupole, then skip the subtree), mark a leaf
// for particle-particle terms (returns true), or open (descend to cur + 1).
template<bool kSafe, bool kQuad>
__device__ __forceinline__ bool visit(WalkTarget& t, std::uint32_t cur, const WalkNode& nd,
                                      const QuadF& Q, std::uint32_t next, float eps2) {
    if (t.mine != cur) return false;
    ++t.visits;
    const float dx = nd.a.x - t.xi.x, dy = nd.a.y - t.xi.y, dz = nd.a.z - t.xi.z;
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    float acc;
    std::memcpy(&acc, &nd.b.x, sizeof(acc));
    if (r2 >= acc) {
        // tree_walk.hpp:129-139 monopole.
        float r2e = r2 + eps2;
        if constexpr (kSafe) r2e = r2e > 0.f ? r2e : 1.f;
        const float inv = __builtin_amdgcn_rsqf(r2e);
        if constexpr (!kQuad) {
            const float f = (nd.a.w * inv) * (inv * inv);
            t.px = __builtin_fmaf(f, dx, t.px);
            t.py = __builtin_fmaf(f, dy, t.py);
            t.pz = __builtin_fmaf(f, dz, t.pz);
        } else {
            // Monopole plus tree_walk.hpp:156-181 in one update:
            // d (m inv^3 + 2.5 (d.Q d) inv^7) - (Q d) inv^5.
            const float i2 = inv * inv;
            const float i3 = i2 * inv;
            const float i5 = i3 * i2;
            const float i7 = i5 * i2;
            const float qx = __builtin_fmaf(Q.a.z, dz, __builtin_fmaf(Q.a.y, dy, Q.a.x * dx));
            const float qy = __builtin_fmaf(Q.b.x, dz, __builtin_fmaf(Q.a.w, dy, Q.a.y * dx));
            const float qz = __builtin_fmaf(Q.b.y, dz, __builtin_fmaf(Q.b.x, dy, Q.a.z * dx));
            const float quadratic = __builtin_fmaf(dz, qz, __builtin_fmaf(dy, qy, dx * qx));
            const float g = __builtin_fmaf(2.5f * quadratic, i7, nd.a.w * i3);
            t.px = __builtin_fmaf(dx, g, __builtin_fmaf(qx, -i5, t.px));
            t.py = __builtin_fmaf(dy, g, __builtin_fmaf(qy, -i5, t.py));
            t.pz = __builtin_fmaf(dz, g, __builtin_fmaf(qz, -i5, t.pz));
        }
        ++t.pc;
        t.mine = next;
        return false;
    }
    if (nd.b.w > 0) {
        t.pp += nd.b.w - (t.own == cur ? 1u : 0u);
        t.mine = next;
        return true;
    }
    t.mine = cur + 1;
    return false;
}

// Particle-particle terms of leaf [b, e) for one target, then the partial is
// Kahan-added. The target is re-expressed against the leaf's centre cb in FP64
// (ca: the centre of the target's own leaf).
template<bool kSafe>
__device__ __forceinline__ void leaf_pairs(WalkTarget& t, const double4& cb, double cax, double cay,
                                           double caz, const float4* __restrict__ leaf_rel,
                                           std::uint32_t b, std::uint32_t e, float eps2) {
    const float tx = t.oi.x - static_cast<float>(cb.x - cax);
    const float ty = t.oi.y - static_cast<float>(cb.y - cay);
    const float tz = t.oi.z - static_cast<float>(cb.z - caz);
    auto pair = [&](const float4 q) {
        const float dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
        float r2 = __builtin_fmaf(dx, dx, eps2);
        r2 = __builtin_fmaf(dy, dy, r2);
        r2 = __builtin_fmaf(dz, dz, r2);
        if constexpr (kSafe) r2 = r2 > 0.f ? r2 : 1.f;
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float f = (q.w * inv) * (inv * inv);
        t.px = __builtin_fmaf(f, dx, t.px);
        t.py = __builtin_fmaf(f, dy, t.py);
        t.pz = __builtin_fmaf(f, dz, t.pz);
    };
    // Four sources per round, addressed from one base so that their scalar
    // loads are issued together and share one wait. A last round of fewer than
    // four still loads four (leaf_rel holds kLeafPad spare entries past the
    // last particle) and uses those in the leaf: most leaves hold under four.
    for (std::uint32_t j = b; j < e; j += 4) {
        const float4* q = leaf_rel + j;
        const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        asm volatile("" ::"s"(q0.x), "s"(q0.y), "s"(q0.z), "s"(q0.w), "s"(q1.x), "s"(q1.y),
                     "s"(q1.z), "s"(q1.w), "s"(q2.x), "s"(q2.y), "s"(q2.z), "s"(q2.w),
                     "s"(q3.x), "s"(q3.y), "s"(q3.z), "s"(q3.w));
        pair(q0);
        if (j + 1 < e) pair(q1);
        if (j + 2 < e) pair(q2);
        if (j + 3 < e) pair(q3);
    }
    kahan(t.px, t.sx, t.kx);
    kahan(t.py, t.sy, t.ky);
    kahan(t.pz, t.sz, t.kz);
    t.px = t.py = t.pz = 0.f;
}

// Wave-coherent walk. Each of the first kLanes lanes owns kR targets (wave w
// owns the kLanes kR sorted particles from kLanes kR w, target r of lane l
// being kLanes r + l; with kLanes = 32 the other half of the wave holds none,
// which doubles the waves a small system gets); the wave visits
// the smallest node any of its targets needs, loads it once, and every target
// that needs it takes its step. A target therefore sees exactly the node
// sequence and arithmetic of a walk of its own, whatever the other targets are.
template<int kR, bool kSafe, bool kQuad, int kLanes>
__global__ __launch_bounds__(kBlock) void tree_walk(
    const WalkNode* __restrict__ nodes, const QuadF* __restrict__ quads,
    const double4* __restrict__ node_geo,
    const float4* __restrict__ root_rel, const float4* __restrict__ leaf_rel,
    const std::uint32_t* __restrict__ leaf_of, const std::uint32_t* __restrict__ index,
    const std::uint32_t* __restrict__ counters, const std::uint32_t* __restrict__ seg_of,
    const std::uint32_t* __restrict__ seg_start, const std::uint32_t* __restrict__ first_node,
    std::uint32_t n, float eps2, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    unsigned long long* __restrict__ stats, unsigned long long* __restrict__ wave_stats) {
    constexpr std::uint32_t kDone = 0xffffffffu;
    const std::uint32_t wave = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + threadIdx.x / 64;
    const std::uint32_t lane = threadIdx.x & 63;
    const std::uint32_t M = counters[0];
    WalkTarget t[kR];
    std::uint32_t seg[kR];
    bool valid[kR];
    // Each lane's targets' own leaf centres, read at every leaf visit: held in
    // LDS rather than registers (12 KB a block at two targets a lane) or
    // re-read from memory at each visit. A lane reads only its own slots.
    __shared__ double s_own[kR][3][kBlock];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        const std::uint32_t i = wave * (kLanes * kR) + kLanes * r + lane;
        valid[r] = lane < kLanes && i < n;
        const std::uint32_t ic = valid[r] ? i : n - 1;
        t[r].xi = root_rel[ic];
        t[r].oi = leaf_rel[ic];
        t[r].own = leaf_of[ic];
        const double4 ca = node_geo[t[r].own];
        s_own[r][0][threadIdx.x] = ca.x;
        s_own[r][1][threadIdx.x] = ca.y;
        s_own[r][2][threadIdx.x] = ca.z;
        // The target's system starts at its root; the walk records' escape
        // indices end the walk at the system's last subtree.
        seg[r] = seg_of != nullptr ? seg_of[ic] : 0u;
        const std::uint32_t root = seg_start != nullptr ? first_node[seg_start[seg[r]]] : 0u;
        t[r].mine = valid[r] ? root : kDone;
        t[r].sx = t[r].sy = t[r].sz = t[r].kx = t[r].ky = t[r].kz = 0.f;
        t[r].px = t[r].py = t[r].pz = 0.f;
        t[r].pp = t[r].pc = t[r].visits = 0;
    }
    std::uint32_t steps = 0;
    // Each iteration advances at least one target, so a walk takes at most
    // 64 kR M iterations; the cap only turns a corrupted tree into a wrong result.
    // 32-bit counts (a 32-bit compare in the loop condition), the cap clamped.
    const std::uint32_t cap = static_cast<std::uint32_t>(min(64ull * kR * M + 64, 0xffffffffull));
    std::uint32_t least = t[0].mine;
#pragma unroll
    for (int r = 1; r < kR; ++r) least = min(least, t[r].mine);
    std::uint32_t cur = __builtin_amdgcn_readfirstlane(wave_min_u32(least));
    while (cur != kDone && steps < cap) {
        ++steps;
        const WalkNode nd = nodes[cur];
        QuadF Q{};
        if constexpr (kQuad) Q = quads[cur];
        // Every scalar load of the iteration is issued here and waited for once;
        // left to itself the compiler sinks them into the branches below, one
        // dependent wait each.
        asm volatile("" ::"s"(nd.a.x), "s"(nd.a.y), "s"(nd.a.z), "s"(nd.a.w), "s"(nd.b.x),
                     "s"(nd.b.y), "s"(nd.b.z), "s"(nd.b.w));
        if constexpr (kQuad) {
            asm volatile("" ::"s"(Q.a.x), "s"(Q.a.y), "s"(Q.a.z), "s"(Q.a.w), "s"(Q.b.x),
                         "s"(Q.b.y));
        }
        // An escape index always lies past its node (or is kDone); the guard only
        // makes a corrupted tree end the walk instead of looping.
        const std::uint32_t next = nd.b.y > cur ? nd.b.y : kDone;
        bool leaf[kR];
        bool any_leaf = false;
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            leaf[r] = visit<kSafe, kQuad>(t[r], cur, nd, Q, next, eps2);
            any_leaf = any_leaf || leaf[r];
        }
        if (__ballot(any_leaf) != 0) {
            const double4 cb = node_geo[cur];
            const std::uint32_t b = nd.b.z, e = nd.b.z + nd.b.w;
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                if (leaf[r]) {
                    leaf_pairs<kSafe>(t[r], cb, s_own[r][0][threadIdx.x], s_own[r][1][threadIdx.x],
                                      s_own[r][2][threadIdx.x], leaf_rel, b, e, eps2);
                }
            }
        }
        // The wave's next node: cur + 1 when any target descends; else, most
        // often, this node's escape `next` (some target is there and none
        // before it), without the wave-wide minimum; else that minimum.
        const std::uint32_t step = cur + 1;
        bool descends = false, at_next = false, before_next = false;
        least = t[0].mine;
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            descends = descends || t[r].mine == step;
            at_next = at_next || t[r].mine == next;
            before_next = before_next || t[r].mine < next;
            least = min(least, t[r].mine);
        }
        if (__ballot(descends) != 0) {
            cur = step;
        } else if (__ballot(before_next) == 0 && __ballot(at_next) != 0) {
            cur = next;
        } else {
            cur = __builtin_amdgcn_readfirstlane(wave_min_u32(least));
        }
    }
    std::uint32_t pp = 0, pc = 0, visits = 0;
    bool mixed = false;
    const std::uint32_t seg0 = __builtin_amdgcn_readfirstlane(seg[0]);
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        kahan(t[r].px, t[r].sx, t[r].kx);
        kahan(t[r].py, t[r].sy, t[r].ky);
        kahan(t[r].pz, t[r].sz, t[r].kz);
        if (valid[r]) {
            const std::uint32_t dst = index[wave * (kLanes * kR) + kLanes * r + lane];
            ax[dst] = G * (static_cast<double>(t[r].sx) - static_cast<double>(t[r].kx));
            ay[dst] = G * (static_cast<double>(t[r].sy) - static_cast<double>(t[r].ky));
            az[dst] = G * (static_cast<double>(t[r].sz) - static_cast<double>(t[r].kz));
            pp += t[r].pp;
            pc += t[r].pc;
            visits += t[r].visits;
        }
        mixed = mixed || (valid[r] && seg[r] != seg0);
    }
    // Counts: a wave inside one system writes its sums to its own slot, summed
    // per system by walk_counts (atomics on one address would serialise every
    // wave); a wave spanning two systems adds per target.
    unsigned long long* mine_slot = wave_stats + 4 * static_cast<std::size_t>(wave);
    if (__ballot(mixed) == 0) {
        const unsigned long long wpp = wave_sum_u64(pp);
        const unsigned long long wpc = wave_sum_u64(pc);
        const unsigned long long wvisits = wave_sum_u64(visits);
        if (lane == 0) {
            mine_slot[0] = wpp;
            mine_slot[1] = wpc;
            mine_slot[2] = wvisits;
            // Lane-target slots, 64 per count.
            mine_slot[3] = static_cast<unsigned long long>(steps) * kR * kLanes / 64;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            if (valid[r]) {
                unsigned long long* st = stats + 4 * static_cast<std::size_t>(seg[r]);
                atomicAdd(&st[0], static_cast<unsigned long long>(t[r].pp));
                atomicAdd(&st[1], static_cast<unsigned long long>(t[r].pc));
                atomicAdd(&st[2], static_cast<unsigned long long>(t[r].visits));
            }
        }
        if (lane == 0) {
            atomicAdd(&stats[4 * static_cast<std::size_t>(seg0) + 3],
                      static_cast<unsigned long long>(steps) * kR * kLanes / 64);
            mine_slot[0] = mine_slot[1] = mine_slot[2] = mine_slot[3] = 0;
        }
    }
}

// One block per system: the wave slots over its particle range, summed (a wave
// spanning two systems left zeros and added its counts directly).
__global__ __launch_bounds__(kBlock) void walk_counts(const unsigned long long* __restrict__ wave_stats,
                                                      std::uint32_t waves, std::uint32_t per_wave,
                                                      const std::uint32_t* __restrict__ seg_start,
                                                      unsigned long long* __restrict__ stats) {
    __shared__ unsigned long long s[4][kBlock];
    const std::uint32_t seg = blockIdx.x;
    std::uint32_t w0 = 0, w1 = waves;
    if (seg_start != nullptr) {
        w0 = seg_start[seg] / per_wave;
        w1 = min(waves, (seg_start[seg + 1] + per_wave - 1) / per_wave);
    }
    unsigned long long acc[4] = {0, 0, 0, 0};
    for (std::uint32_t w = w0 + threadIdx.x; w < w1; w += kBlock) {
        for (int c = 0; c < 4; ++c) acc[c] += wave_stats[4 * static_cast<std::size_t>(w) + c];
    }
    for (int c = 0; c < 4; ++c) s[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) {
            for (int c = 0; c < 4; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) atomicAdd(&stats[4 * static_cast<std::size_t>(seg) + threadIdx.x], s[threadIdx.x][0]);
}

template<typename T> T* carve(char*& p, std::size_t count) {
    T* out = reinterpret_cast<T*>(p);
    p += (sizeof(T) * count + 255) & ~static_cast<std::size_t>(255);
    return out;
}

struct Layout {
    // Per particle.
    std::uint8_t* d;
    std::uint8_t* leaf_level;
    std::uint32_t* chain;
    std::uint32_t* first_node;
    std::uint32_t* leaf_of;
    double4* sorted;
    float4* root_rel;
    float4* leaf_rel;
    std::uint32_t* scan_scratch;
    // Per node.
    std::uint32_t* begin;
    std::uint32_t* count;
    std::uint32_t* next;
    std::uint8_t* level;
    std::uint32_t* children;  // 8 per node: an internal node's children in DFS order, then kNoChild
    double4* geo;
    double4* moment;
    WalkNode* walk;
    double* quad64;  // 6 per node, null without quadrupoles
    QuadF* quadf;
    // Counters.
    std::uint32_t* c32;
    unsigned long long* c64;
    uint2* block_counts;               // per block of tree_chains: {leaves, depth}
    unsigned long long* wave_stats;    // per walk wave: {pp, pc, visits, steps}
};

std::size_t node_bound(std::uint64_t n, std::uint64_t cap) {
    // Leaves partition the particles (<= n); internal nodes of one level are
    // disjoint cells of > cap particles (<= n / (cap + 1) per level, 21 levels).
    const std::uint64_t by_levels = n + 21 * (n / (cap + 1)) + 22;
    const std::uint64_t by_chains = 22 * n;
    return static_cast<std::size_t>(std::min(by_levels, by_chains));
}

// Bytes of one tree arena for n particles, `nodes` node slots and `segments` systems.
std::size_t layout_bytes(std::uint64_t n, std::size_t nodes, bool quadrupole, std::uint32_t segments) {
    // Every carved array is rounded up to 256 bytes: 22 arrays of padding.
    const std::size_t per_particle = 2 * n + 3 * 4 * n + 32 * n + 2 * 16 * n + 16 * kLeafPad +
                                     4 * scan_scratch_words(n);
    const std::size_t per_quad = quadrupole ? 6 * sizeof(double) + sizeof(QuadF) : 0;
    const std::size_t per_node = nodes * (4 * 3 + 1 + 4 * kMaxChildren + 32 + 32 + sizeof(WalkNode) + per_quad);
    const std::size_t counters = 16 + 8 * 4 * static_cast<std::size_t>(segments) +
                                 8 * ((n + kBlock - 1) / kBlock) +
                                 32 * (kBlock / 64) * ((2 * n + kBlock - 1) / kBlock) + 512;
    return per_particle + per_node + counters + 22 * 256 + 64;
}

Layout layout(orr_ctx* ctx, std::uint64_t n, std::size_t nodes, bool quadrupole, std::uint32_t segments,
              bool allocate) {
    const std::size_t bytes = layout_bytes(n, nodes, quadrupole, segments);
    if (allocate) ctx->tree_work.reserve(bytes);
    char* p = static_cast<char*>(ctx->tree_work.ptr);
    Layout L{};
    L.d = carve<std::uint8_t>(p, n);
    L.leaf_level = carve<std::uint8_t>(p, n);
    L.chain = carve<std::uint32_t>(p, n);
    L.first_node = carve<std::uint32_t>(p, n);
    L.leaf_of = carve<std::uint32_t>(p, n);
    L.sorted = carve<double4>(p, n);
    L.root_rel = carve<float4>(p, n);
    L.leaf_rel = carve<float4>(p, n + kLeafPad);
    L.scan_scratch = carve<std::uint32_t>(p, scan_scratch_words(n));
    L.begin = carve<std::uint32_t>(p, nodes);
    L.count = carve<std::uint32_t>(p, nodes);
    L.next = carve<std::uint32_t>(p, nodes);
    L.level = carve<std::uint8_t>(p, nodes);
    L.children = carve<std::uint32_t>(p, kMaxChildren * nodes);
    L.geo = carve<double4>(p, nodes);
    L.moment = carve<double4>(p, nodes);
    L.walk = carve<WalkNode>(p, nodes);
    if (quadrupole) {
        L.quad64 = carve<double>(p, 6 * nodes);
        L.quadf = carve<QuadF>(p, nodes);
    }
    L.c32 = carve<std::uint32_t>(p, 4);
    L.c64 = carve<unsigned long long>(p, 4 * static_cast<std::size_t>(segments));
    L.block_counts = carve<uint2>(p, (n + kBlock - 1) / kBlock);
    // Half-lane walks (kLanes = 32) launch twice the waves of full ones.
    L.wave_stats = carve<unsigned long long>(p, 4 * (kBlock / 64) * ((2 * n + kBlock - 1) / kBlock));
    return L;
}

}  // namespace

std::size_t tree_bytes(std::uint64_t n, std::uint64_t leaf_capacity, bool quadrupole,
                       std::uint32_t segments) {
    if (n == 0) return 0;
    // As tree_evaluate_segments sizes its arena.
    const std::uint64_t cap = std::min<std::uint64_t>(std::max<std::uint64_t>(leaf_capacity, 1), n);
    const std::size_t bound = node_bound(n, cap) + 22 * static_cast<std::size_t>(segments);
    return layout_bytes(n, bound, quadrupole, segments);
}

namespace {

// Every launch of one evaluation on ctx->stream, no host synchronisation: the
// part a graph captures.
void tree_enqueue(orr_ctx* ctx, const orr_tree_args& a, const SegmentView& segs,
                  unsigned long long* seg_stats) {
    const std::uint64_t n = a.count;
    ctx->tree_last_n = n;
    ctx->tree_last_segments = segs.count;
    if (n == 0) return;
    if (n >= (std::uint64_t{1} << 31)) fail(ORR_ERR_LENGTH, "tree: 2^31 or more particles");
    // octree.cpp:189-194 parameter correction.
    const double theta = std::clamp(a.opening_angle, 0.0, 1.0);
    const std::uint64_t cap64 = std::max<std::uint64_t>(a.leaf_capacity, 1);
    const std::uint32_t cap = static_cast<std::uint32_t>(std::min<std::uint64_t>(cap64, n));
    const std::size_t bound = node_bound(n, cap) + 22 * static_cast<std::size_t>(segs.count);
    ctx->tree_last_bound = bound;
    ctx->tree_last_theta = theta;
    ctx->tree_last_quadrupole = a.quadrupole != 0;
    const Layout L = layout(ctx, n, bound, a.quadrupole != 0, segs.count, true);
    const std::uint32_t n32 = static_cast<std::uint32_t>(n);
    const std::uint32_t blocks = (n32 + kBlock - 1) / kBlock;
    const std::uint32_t node_blocks = static_cast<std::uint32_t>(std::min<std::size_t>(4096, (bound + kBlock - 1) / kBlock));
    const std::uint32_t* seg_of = segs.count > 1 ? segs.seg_of : nullptr;
    const std::uint32_t* seg_start = segs.count > 1 ? segs.start : nullptr;
    unsigned long long* counts64 = seg_stats != nullptr ? seg_stats : L.c64;

    morton_enqueue_segments(ctx, a.x, a.y, a.z, n, segs, a.m);
    const std::uint64_t* codes = ctx->order_codes.as<std::uint64_t>();
    const std::uint32_t* index = ctx->order_index.as<std::uint32_t>();
    const double* cube = ctx->order_bbox.as<double>();
    {
        PhaseTimer t(ctx, kBuild);
        hipLaunchKernelGGL(clear_counters, dim3(1), dim3(kBlock), 0, ctx->stream, L.c32,
                           seg_stats != nullptr ? nullptr : L.c64, 4 * segs.count);
        ORR_LAUNCHED("clear_counters");
        hipLaunchKernelGGL(tree_levels, dim3(blocks), dim3(kBlock), 0, ctx->stream, codes, seg_of, n32,
                           L.d);
        ORR_LAUNCHED("tree_levels");
        hipLaunchKernelGGL(tree_chain, dim3(blocks), dim3(kBlock), 0, ctx->stream, L.d, n32, cap,
                           L.chain, L.leaf_level);
        ORR_LAUNCHED("tree_chain");
        exclusive_scan_u32(ctx, L.chain, L.first_node, n, L.scan_scratch, &L.c32[0]);
        hipLaunchKernelGGL(tree_chains, dim3(blocks), dim3(kBlock), 0, ctx->stream, L.d,
                           L.leaf_level, L.chain, L.first_node, n32, L.block_counts, L.begin,
                           L.level);
        ORR_LAUNCHED("tree_chains");
        hipLaunchKernelGGL(tree_counts, dim3(1), dim3(kBlock), 0, ctx->stream, L.block_counts,
                           blocks, L.c32);
        ORR_LAUNCHED("tree_counts");
        hipLaunchKernelGGL(tree_node_fill, dim3(node_blocks), dim3(kBlock), 0, ctx->stream, codes,
                           L.leaf_level, L.first_node, n32, seg_of, seg_start, cube, L.c32,
                           L.begin, L.level, L.count, L.next, L.geo, L.leaf_of);
        ORR_LAUNCHED("tree_node_fill");
    }
    {
        PhaseTimer t(ctx, kGather);
        hipLaunchKernelGGL(tree_gather, dim3(blocks), dim3(kBlock), 0, ctx->stream,
                           ctx->order_aos.as<double4>(), index, n32, cube, seg_of, L.leaf_of, L.geo,
                           L.sorted, L.root_rel, L.leaf_rel);
        ORR_LAUNCHED("tree_gather");
    }
    {
        PhaseTimer t(ctx, kMoments);
        hipLaunchKernelGGL(leaf_moments, dim3(node_blocks), dim3(kBlock), 0, ctx->stream, L.c32,
                           L.level, L.leaf_level, L.begin, L.count, L.next, L.geo, L.sorted, theta,
                           cube, seg_of, seg_start, n32, L.moment, L.walk, L.quad64, L.quadf);
        ORR_LAUNCHED("leaf_moments");
        hipLaunchKernelGGL(tree_children, dim3(node_blocks), dim3(kBlock), 0, ctx->stream, L.c32,
                           L.level, L.leaf_level, L.begin, L.next, L.children);
        ORR_LAUNCHED("tree_children");
        for (int level = static_cast<int>(kLevels) - 1; level >= 0; --level) {
            hipLaunchKernelGGL(internal_moments, dim3(node_blocks), dim3(kBlock), 0, ctx->stream,
                               static_cast<unsigned>(level), L.c32, L.level, L.leaf_level, L.begin,
                               L.count, L.next, L.children, L.geo, theta, cube, seg_of, seg_start,
                               n32, L.moment, L.walk, L.quad64, L.quadf);
            ORR_LAUNCHED("internal_moments");
        }
    }
    {
        PhaseTimer t(ctx, kWalk);
        const float eps2 = static_cast<float>(a.softening_squared);
        const bool safe = !(a.softening_squared > 0);
        const bool quad = a.quadrupole != 0;
        // Two targets per lane pay off for one large system (measured: 4.5 % at
        // N = 2^21); zero in the context means choose by that rule.
        const int R = ctx->walk_targets_per_lane != 0 ? ctx->walk_targets_per_lane
                      : (segs.count == 1 && n >= (std::uint64_t{1} << 20) ? 2 : 1);
        // Half the lanes of each wave for a launch of at most kHalfLaneWaves
        // full waves (twice the waves, each over a tighter group of targets);
        // zero in the context means choose by that rule. A target's walk and
        // sums do not depend on which targets share its wave, so the results
        // are the same bits either way.
        const std::uint64_t full_waves = (n + 64u * R - 1) / (64u * R);
        const int lanes = ctx->walk_lanes != 0 ? ctx->walk_lanes
                          : (full_waves <= kHalfLaneWaves ? 32 : 64);
        using Walk = decltype(&tree_walk<1, true, true, 64>);
        auto pick = [&](auto r, auto l) -> Walk {
            constexpr int kRr = decltype(r)::value, kLl = decltype(l)::value;
            return safe ? (quad ? tree_walk<kRr, true, true, kLl> : tree_walk<kRr, true, false, kLl>)
                        : (quad ? tree_walk<kRr, false, true, kLl> : tree_walk<kRr, false, false, kLl>);
        };
        using One = std::integral_constant<int, 1>;
        using Two = std::integral_constant<int, 2>;
        using Half = std::integral_constant<int, 32>;
        using Full = std::integral_constant<int, 64>;
        const Walk walk = R == 2 ? (lanes == 32 ? pick(Two{}, Half{}) : pick(Two{}, Full{}))
                                 : (lanes == 32 ? pick(One{}, Half{}) : pick(One{}, Full{}));
        const std::uint32_t per_wave = static_cast<std::uint32_t>(lanes * R);
        const std::uint32_t per_block = (kBlock / 64) * per_wave;
        const std::uint32_t walk_blocks = (n32 + per_block - 1) / per_block;
        hipLaunchKernelGGL(walk, dim3(walk_blocks), dim3(kBlock), 0, ctx->stream, L.walk, L.quadf,
                           L.geo, L.root_rel, L.leaf_rel, L.leaf_of, index, L.c32, seg_of,
                           seg_start, L.first_node, n32, eps2, a.gravitational_constant, a.ax,
                           a.ay, a.az, counts64, L.wave_stats);
        ORR_LAUNCHED("tree_walk");
        hipLaunchKernelGGL(walk_counts, dim3(segs.count), dim3(kBlock), 0, ctx->stream,
                           L.wave_stats, walk_blocks * (kBlock / 64), per_wave, seg_start, counts64);
        ORR_LAUNCHED("walk_counts");
    }
}

// The counts and tree summary of the evaluation last enqueued on ctx (syncs).
void tree_read_stats(orr_ctx* ctx, unsigned long long* seg_stats, orr_tree_stats* stats) {
    const std::uint64_t n = ctx->tree_last_n;
    const std::uint32_t segments = ctx->tree_last_segments;
    if (n == 0) return;
    const Layout L = layout(ctx, n, ctx->tree_last_bound, ctx->tree_last_quadrupole, segments, false);
    const unsigned long long* counts64 = seg_stats != nullptr ? seg_stats : L.c64;
    std::uint32_t c32[4];
    std::vector<unsigned long long> c64(4 * static_cast<std::size_t>(segments));
    ORR_HIP(hipMemcpyAsync(c32, L.c32, sizeof(c32), hipMemcpyDeviceToHost, ctx->stream));
    ORR_HIP(hipMemcpyAsync(c64.data(), counts64, c64.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    stats->node_count = c32[0];
    stats->leaf_count = c32[1];
    stats->depth = c32[2];
    for (std::uint32_t sgm = 0; sgm < segments; ++sgm) {
        stats->particle_particle += c64[4 * sgm];
        stats->particle_cell += c64[4 * sgm + 1];
        stats->node_visits += c64[4 * sgm + 2];
        stats->wave_steps += c64[4 * sgm + 3];
    }
    ctx->last_tree = *stats;
}

}  // namespace

void tree_evaluate_segments(orr_ctx* ctx, const orr_tree_args& a, const SegmentView& segs,
                            unsigned long long* seg_stats, orr_tree_stats* stats) {
    if (stats != nullptr) std::memset(stats, 0, sizeof(*stats));
    ctx->last_tree = orr_tree_stats{};
    tree_enqueue(ctx, a, segs, seg_stats);
    if (stats != nullptr) tree_read_stats(ctx, seg_stats, stats);
}

// One system: replayed as a captured graph where run_captured allows (the
// key: every argument and the walk setting).
void tree_evaluate(orr_ctx* ctx, const orr_tree_args& a, orr_tree_stats* stats) {
    if (stats != nullptr) std::memset(stats, 0, sizeof(*stats));
    ctx->last_tree = orr_tree_stats{};
    if (a.count == 0) {
        tree_enqueue(ctx, a, SegmentView{}, nullptr);
        return;
    }
    std::vector<unsigned char> key;
    for (const void* p : {static_cast<const void*>(a.x), static_cast<const void*>(a.y),
                          static_cast<const void*>(a.z), static_cast<const void*>(a.m),
                          static_cast<const void*>(a.ax), static_cast<const void*>(a.ay),
                          static_cast<const void*>(a.az)}) {
        key_add(key, p);
    }
    key_add(key, a.count);
    key_add(key, a.softening_squared);
    key_add(key, a.gravitational_constant);
    key_add(key, a.opening_angle);
    key_add(key, a.leaf_capacity);
    key_add(key, a.quadrupole);
    key_add(key, ctx->walk_targets_per_lane);
    key_add(key, ctx->walk_lanes);
    run_captured(ctx, ctx->tree_graph, std::move(key),
                 [&] { tree_enqueue(ctx, a, SegmentView{}, nullptr); });
    if (stats != nullptr) tree_read_stats(ctx, nullptr, stats);
}

void tree_copy_segment_counts(orr_ctx* ctx, std::uint64_t* out) {
    const std::uint64_t n = ctx->tree_last_n;
    const std::size_t words = 4 * static_cast<std::size_t>(ctx->tree_last_segments);
    if (n == 0 || ctx->tree_work.ptr == nullptr) {
        for (std::size_t k = 0; k < words; ++k) out[k] = 0;
        return;
    }
    const Layout L = layout(ctx, n, ctx->tree_last_bound, ctx->tree_last_quadrupole,
                            ctx->tree_last_segments, false);
    static_assert(sizeof(unsigned long long) == sizeof(std::uint64_t));
    ORR_HIP(hipMemcpyAsync(out, L.c64, words * sizeof(std::uint64_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    ctx->sync();
}

void tree_copy_quadrupoles(orr_ctx* ctx, double* out) {
    const std::uint64_t n = ctx->tree_last_n;
    if (n == 0 || ctx->tree_work.ptr == nullptr) return;
    if (!ctx->tree_last_quadrupole) fail(ORR_ERR_INVALID_ARGUMENT, "the last tree had no quadrupoles");
    const Layout L = layout(ctx, n, ctx->tree_last_bound, true, ctx->tree_last_segments, false);
    std::uint32_t M = 0;
    ORR_HIP(hipMemcpyAsync(&M, L.c32, sizeof(M), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (M == 0) return;
    ORR_HIP(hipMemcpyAsync(out, L.quad64, 6 * sizeof(double) * M, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
}

void tree_copy_nodes(orr_ctx* ctx, double* mass, double* cx, double* cy, double* cz, double* acc,
                     std::uint64_t* next, std::uint64_t* first, std::uint64_t* count) {
    const std::uint64_t n = ctx->tree_last_n;
    if (n == 0 || ctx->tree_work.ptr == nullptr) return;
    const Layout L = layout(ctx, n, ctx->tree_last_bound, ctx->tree_last_quadrupole, ctx->tree_last_segments, false);
    std::uint32_t M = 0;
    ORR_HIP(hipMemcpyAsync(&M, L.c32, sizeof(M), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    std::vector<double4> mom(M), geo(M);
    std::vector<WalkNode> walk(M);
    std::vector<std::uint32_t> nx(M), b(M), c(M);
    if (M > 0) {
        ORR_HIP(hipMemcpyAsync(mom.data(), L.moment, M * sizeof(double4), hipMemcpyDeviceToHost, ctx->stream));
        ORR_HIP(hipMemcpyAsync(geo.data(), L.geo, M * sizeof(double4), hipMemcpyDeviceToHost, ctx->stream));
        ORR_HIP(hipMemcpyAsync(walk.data(), L.walk, M * sizeof(WalkNode), hipMemcpyDeviceToHost, ctx->stream));
        ORR_HIP(hipMemcpyAsync(nx.data(), L.next, M * sizeof(std::uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    }
    const double theta = ctx->tree_last_theta;
    for (std::uint32_t k = 0; k < M; ++k) {
        if (mass) mass[k] = mom[k].w;
        if (cx) cx[k] = mom[k].x;
        if (cy) cy[k] = mom[k].y;
        if (cz) cz[k] = mom[k].z;
        if (acc) {
            // The FP64 radius of octree.cpp:374-387 (the walk holds it rounded to FP32).
            const double ox = mom[k].x - geo[k].x, oy = mom[k].y - geo[k].y, oz = mom[k].z - geo[k].z;
            const double offset = std::sqrt((ox * ox) + (oy * oy) + (oz * oz));
            if (!(theta > 0)) {
                acc[k] = std::numeric_limits<double>::infinity();
            } else {
                const double r = (geo[k].w / theta) + offset;
                acc[k] = r * r;
            }
        }
        if (next) next[k] = nx[k];
        if (first) first[k] = walk[k].b.z;
        if (count) count[k] = walk[k].b.w;
    }
}

}  // namespace orr
