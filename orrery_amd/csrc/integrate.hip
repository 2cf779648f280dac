// FP64 integrator update and diagnostics on device-resident state.
//
// kick / drift follow src/integrators/velocity_verlet.cpp:20-55 (v += s a,
// x += s v, one multiply and one add each, never fused: the file disables
// contraction so the update rounds as the reference's portable build does).
// RK4 follows src/integrators/runge_kutta4.cpp:19-139 with the same unfused
// multiply-then-add updates.
// The diagnostics follow src/core/diagnostics.cpp:85-200: kinetic energy,
// momentum and angular momentum as Neumaier-compensated sums, potential
// energy as rows i of plain sums over j > i (in column chunks folded in order),
// the rows then compensated.
// Device reductions are deterministic: fixed chunks, each summed in index
// order by one lane, then the chunk results summed in order by one lane.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "engine.hpp"

#pragma clang fp contract(off)

namespace orr {
namespace {

__global__ __launch_bounds__(256) void kick_kernel(std::uint64_t n, double* __restrict__ vx,
                                                   double* __restrict__ vy,
                                                   double* __restrict__ vz,
                                                   const double* __restrict__ ax,
                                                   const double* __restrict__ ay,
                                                   const double* __restrict__ az, double s) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] = vx[i] + s * ax[i];
    vy[i] = vy[i] + s * ay[i];
    vz[i] = vz[i] + s * az[i];
}

// A kick, then a drift with the kicked velocity: each element's arithmetic is
// kick_kernel's twice over, so the result is that of the two launches.
__global__ __launch_bounds__(256) void kick_drift_kernel(std::uint64_t n, double* __restrict__ x,
                                                         double* __restrict__ y, double* __restrict__ z,
                                                         double* __restrict__ vx, double* __restrict__ vy,
                                                         double* __restrict__ vz,
                                                         const double* __restrict__ ax,
                                                         const double* __restrict__ ay,
                                                         const double* __restrict__ az, double s,
                                                         double t) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ux = vx[i] + s * ax[i];
    const double uy = vy[i] + s * ay[i];
    const double uz = vz[i] + s * az[i];
    vx[i] = ux;
    vy[i] = uy;
    vz[i] = uz;
    x[i] = x[i] + t * ux;
    y[i] = y[i] + t * uy;
    z[i] = z[i] + t * uz;
}

// runge_kutta4.cpp:19-56, one kernel for its four element-wise loops:
// op 0 begins a step (x0 = x, v0 = v, Sx = v, Sv = a); op 1 moves to a stage
// (x = x0 + s v, v = v0 + s a, reading v first); op 2 accumulates a stage's
// derivatives (Sx += s v, Sv += s a); op 3 combines (x = x0 + s Sx, v = v0 + s Sv).
__global__ __launch_bounds__(256) void rk4_kernel(std::uint64_t n, Rk4Arrays r, int op, double s) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (op == 0) {
            r.x0[c][i] = r.x[c][i];
            r.v0[c][i] = r.v[c][i];
            r.sx[c][i] = r.v[c][i];
            r.sv[c][i] = r.a[c][i];
        } else if (op == 1) {
            const double stage_velocity = r.v[c][i];
            r.x[c][i] = r.x0[c][i] + (s * stage_velocity);
            r.v[c][i] = r.v0[c][i] + (s * r.a[c][i]);
        } else if (op == 2) {
            r.sx[c][i] = r.sx[c][i] + (s * r.v[c][i]);
            r.sv[c][i] = r.sv[c][i] + (s * r.a[c][i]);
        } else {
            r.x[c][i] = r.x0[c][i] + (s * r.sx[c][i]);
            r.v[c][i] = r.v0[c][i] + (s * r.sv[c][i]);
        }
    }
}

// Neumaier step (core/diagnostics.cpp:36-59).
struct Compensated {
    double total = 0.0, correction = 0.0;
    __device__ void add(double v) {
        const double u = total + v;
        if (fabs(total) >= fabs(v)) {
            correction += (total - u) + v;
        } else {
            correction += (v - u) + total;
        }
        total = u;
    }
    __device__ double value() const { return total + correction; }
};

constexpr int kTerms = 8;  // ke, px, py, pz, lx, ly, lz, pe

// Columns of the potential's rows are summed in at most kColumnChunks chunks
// of at least kColumnMin columns (whole tiles of kPotTile), so a system of a
// few thousand particles still spreads over many waves.
constexpr int kPotTile = 256;  // columns staged in LDS per step = threads per block
constexpr std::uint64_t kColumnMin = 2048;
constexpr std::uint64_t kColumnChunks = 16;

__host__ __device__ inline std::uint64_t column_width(std::uint64_t n) {
    const std::uint64_t even = (n + kColumnChunks - 1) / kColumnChunks;
    const std::uint64_t w = even > kColumnMin ? even : kColumnMin;
    return (w + kPotTile - 1) / kPotTile * kPotTile;
}

// 1 / sqrt(v) in FP64: the hardware reciprocal square root refined by two
// Newton steps, y <- y (3/2 - (v/2) y^2) with the residual by FMA (each step
// squares the relative error, so the result is within an ulp or two).
// Branch-free: v = 0 gives +inf and NaN stays NaN, as 1 / sqrt(v) would.
__device__ inline double inverse_sqrt(double v) {
    const double h = 0.5 * v;
    double y = __builtin_amdgcn_rsq(v);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-h, y * y, 0.5);
        y = fma(y, e, y);
    }
    return v > 0.0 ? y : (v == 0.0 ? __builtin_huge_val() : __builtin_nan(""));
}

// m / sqrt(|d|^2 + e2) for the separation d = (dx, dy, dz).
__device__ inline double pair_potential(double m, double dx, double dy, double dz, double e2) {
    return m * inverse_sqrt(fma(dz, dz, fma(dy, dy, fma(dx, dx, e2))));
}

// Row i of the potential over column chunk blockIdx.y: a plain FP64 sum over
// j > i of m_j / sqrt(r^2 + e^2) in index order (r^2 + e^2 by FMA) (diagnostics.cpp:129-158 sums
// the whole row plainly; here the chunks' sums are folded in chunk order).
// A block holds kPotTile rows; the chunk's columns pass through LDS a tile at
// a time and every lane reads each staged column (a broadcast). Blocks wholly
// below the diagonal write zeros and leave.
__global__ __launch_bounds__(kPotTile) void potential_rows(std::uint64_t n, std::uint64_t width,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ z,
                                                           const double* __restrict__ m, double e2,
                                                           double* __restrict__ rows) {
    __shared__ double sx[kPotTile], sy[kPotTile], sz[kPotTile], sm[kPotTile];
    const std::uint64_t row0 = static_cast<std::uint64_t>(blockIdx.x) * kPotTile;
    const std::uint64_t i = row0 + threadIdx.x;
    const std::uint64_t c0 = static_cast<std::uint64_t>(blockIdx.y) * width;
    const std::uint64_t c1 = c0 + width < n ? c0 + width : n;
    // First column any row of the block needs: row0 + 1.
    const std::uint64_t j_begin = c0 > row0 + 1 ? c0 : row0 + 1;
    double inter = 0.0;
    if (j_begin < c1) {
        const bool valid_row = i < n;
        const double xi = valid_row ? x[i] : 0.0, yi = valid_row ? y[i] : 0.0;
        const double zi = valid_row ? z[i] : 0.0;
        // Tiles aligned to kPotTile from the first one holding j_begin.
        for (std::uint64_t t0 = j_begin / kPotTile * kPotTile; t0 < c1; t0 += kPotTile) {
            __syncthreads();
            const std::uint64_t j = t0 + threadIdx.x;
            const bool in = j < c1;
            sx[threadIdx.x] = in ? x[j] : 0.0;
            sy[threadIdx.x] = in ? y[j] : 0.0;
            sz[threadIdx.x] = in ? z[j] : 0.0;
            sm[threadIdx.x] = in ? m[j] : 0.0;
            __syncthreads();
            const std::uint64_t t1 = t0 + kPotTile < c1 ? t0 + kPotTile : c1;
            const int count = static_cast<int>(t1 - t0);
            if (t0 > row0 + kPotTile - 1 && count == kPotTile) {
                // Every column of the tile lies above every row of the block.
#pragma unroll 4
                for (int k = 0; k < kPotTile; ++k) {
                    inter += pair_potential(sm[k], sx[k] - xi, sy[k] - yi, sz[k] - zi, e2);
                }
            } else {
                for (int k = 0; k < count; ++k) {
                    if (t0 + k <= i) continue;
                    inter += pair_potential(sm[k], sx[k] - xi, sy[k] - yi, sz[k] - zi, e2);
                }
            }
        }
    }
    if (i < n) rows[static_cast<std::uint64_t>(blockIdx.y) * n + i] = inter;
}

// Per-particle terms, term-major [kTerms][n].
__global__ __launch_bounds__(256) void diagnostic_terms(std::uint64_t n,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const double* __restrict__ vx,
                                                        const double* __restrict__ vy,
                                                        const double* __restrict__ vz,
                                                        const double* __restrict__ m,
                                                        const double* __restrict__ rows,
                                                        std::uint32_t chunks, double G,
                                                        double* __restrict__ terms) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mi = m[i];
    const double ux = vx[i], uy = vy[i], uz = vz[i];
    const double xi = x[i], yi = y[i], zi = z[i];
    terms[i] = mi * ((ux * ux) + (uy * uy) + (uz * uz));
    terms[n + i] = ux * mi;
    terms[2 * n + i] = uy * mi;
    terms[3 * n + i] = uz * mi;
    terms[4 * n + i] = ((yi * uz) - (zi * uy)) * mi;
    terms[5 * n + i] = ((zi * ux) - (xi * uz)) * mi;
    terms[6 * n + i] = ((xi * uy) - (yi * ux)) * mi;
    double inter = rows[i];
    for (std::uint32_t c = 1; c < chunks; ++c) inter += rows[static_cast<std::uint64_t>(c) * n + i];
    terms[7 * n + i] = -G * mi * inter;
}

// One lane per (term, chunk): compensated sum of the chunk in index order.
__global__ __launch_bounds__(64) void chunk_sums(std::uint64_t n, const double* __restrict__ terms,
                                                 std::uint32_t chunks,
                                                 double* __restrict__ out /* [kTerms][chunks][2] */) {
    const std::uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= kTerms * chunks) return;
    const std::uint32_t term = id / chunks, chunk = id % chunks;
    const std::uint64_t per = (n + chunks - 1) / chunks;
    const std::uint64_t b = chunk * per;
    const std::uint64_t e = b + per < n ? b + per : n;
    Compensated s;
    for (std::uint64_t i = b; i < e; ++i) s.add(terms[term * n + i]);
    out[2 * id] = s.total;
    out[2 * id + 1] = s.correction;
}

// One block, 32 lanes per term: each lane folds a strided subset of the chunk
// results (compensated), then the lanes' (total, correction) pairs are folded
// in a fixed tree order; written in the output order {kinetic, potential, p,
// L}, kinetic halved (diagnostics.cpp:197).
__global__ __launch_bounds__(256) void final_sums(std::uint32_t chunks,
                                                  const double* __restrict__ parts,
                                                  double* __restrict__ result) {
    __shared__ double s_total[256], s_corr[256];
    const int term = threadIdx.x / 32, lane = threadIdx.x % 32;
    Compensated acc;
    for (std::uint32_t c = lane; c < chunks; c += 32) {
        acc.add(parts[2 * (term * chunks + c)]);
        acc.add(parts[2 * (term * chunks + c) + 1]);
    }
    s_total[threadIdx.x] = acc.total;
    s_corr[threadIdx.x] = acc.correction;
    __syncthreads();
    for (int h = 16; h > 0; h >>= 1) {
        if (lane < h) {
            Compensated pair;
            pair.add(s_total[threadIdx.x]);
            pair.add(s_corr[threadIdx.x]);
            pair.add(s_total[threadIdx.x + h]);
            pair.add(s_corr[threadIdx.x + h]);
            s_total[threadIdx.x] = pair.total;
            s_corr[threadIdx.x] = pair.correction;
        }
        __syncthreads();
    }
    if (lane == 0) {
        const double v = s_total[threadIdx.x] + s_corr[threadIdx.x];
        if (term == 0) {
            result[0] = v / 2;
        } else if (term == 7) {
            result[1] = v;
        } else {
            result[1 + term] = v;
        }
    }
}

}  // namespace

void kick(orr_ctx* ctx, std::uint64_t n, double* vx, double* vy, double* vz, const double* ax,
          const double* ay, const double* az, double s) {
    if (n == 0) return;
    PhaseTimer t(ctx, kIntegrate);
    hipLaunchKernelGGL(kick_kernel, dim3(static_cast<std::uint32_t>((n + 255) / 256)), dim3(256),
                       0, ctx->stream, n, vx, vy, vz, ax, ay, az, s);
    ORR_LAUNCHED("kick");
}

void kick_drift(orr_ctx* ctx, std::uint64_t n, double* x, double* y, double* z, double* vx, double* vy,
                double* vz, const double* ax, const double* ay, const double* az, double kick_scale,
                double drift_scale) {
    if (n == 0) return;
    PhaseTimer t(ctx, kIntegrate);
    hipLaunchKernelGGL(kick_drift_kernel, dim3(static_cast<std::uint32_t>((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, n, x, y, z, vx, vy, vz, ax, ay, az, kick_scale, drift_scale);
    ORR_LAUNCHED("kick_drift");
}

void rk4(orr_ctx* ctx, std::uint64_t n, const Rk4Arrays& arrays, int op, double scale) {
    if (n == 0) return;
    PhaseTimer t(ctx, kIntegrate);
    hipLaunchKernelGGL(rk4_kernel, dim3(static_cast<std::uint32_t>((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, n, arrays, op, scale);
    ORR_LAUNCHED("rk4");
}

void drift(orr_ctx* ctx, std::uint64_t n, double* x, double* y, double* z, const double* vx,
           const double* vy, const double* vz, double s) {
    // Same arithmetic as a kick with the roles of the arrays exchanged.
    kick(ctx, n, x, y, z, vx, vy, vz, s);
}

std::size_t diagnostics_scratch_bytes(std::uint64_t n) {
    const std::uint64_t chunks = n < 1024 ? n : 1024;
    const std::uint64_t columns = (n + column_width(n) - 1) / column_width(n);
    return sizeof(double) * (kTerms * n + columns * n + 2 * kTerms * chunks + kTerms);
}

void diagnostics_enqueue(orr_ctx* ctx, std::uint64_t n, const double* x, const double* y,
                         const double* z, const double* vx, const double* vy, const double* vz,
                         const double* m, double e2, double G, double* out) {
    if (n == 0) {
        for (int k = 0; k < 8; ++k) out[k] = 0.0;
        return;
    }
    const std::uint32_t chunks = static_cast<std::uint32_t>(n < 1024 ? n : 1024);
    ctx->diag_scratch.reserve(diagnostics_scratch_bytes(n));
    const std::uint64_t width = column_width(n);
    const std::uint32_t columns = static_cast<std::uint32_t>((n + width - 1) / width);
    double* terms = ctx->diag_scratch.as<double>();
    double* rows = terms + kTerms * n;
    double* parts = rows + static_cast<std::size_t>(columns) * n;
    double* result = parts + 2 * kTerms * chunks;
    PhaseTimer t(ctx, kEnergy);
    const std::uint32_t row_blocks = static_cast<std::uint32_t>((n + 255) / 256);
    hipLaunchKernelGGL(potential_rows, dim3(static_cast<std::uint32_t>((n + kPotTile - 1) / kPotTile), columns),
                       dim3(kPotTile), 0, ctx->stream, n, width, x, y, z, m, e2, rows);
    ORR_LAUNCHED("potential_rows");
    hipLaunchKernelGGL(diagnostic_terms, dim3(row_blocks), dim3(256), 0, ctx->stream, n, x, y, z,
                       vx, vy, vz, m, rows, columns, G, terms);
    ORR_LAUNCHED("diagnostic_terms");
    hipLaunchKernelGGL(chunk_sums, dim3((kTerms * chunks + 63) / 64), dim3(64), 0, ctx->stream, n,
                       terms, chunks, parts);
    ORR_LAUNCHED("chunk_sums");
    hipLaunchKernelGGL(final_sums, dim3(1), dim3(256), 0, ctx->stream, chunks, parts, result);
    ORR_LAUNCHED("final_sums");
    ORR_HIP(hipMemcpyAsync(out, result, sizeof(double) * kTerms, hipMemcpyDeviceToHost, ctx->stream));
}

void diagnostics(orr_ctx* ctx, std::uint64_t n, const double* x, const double* y,
                 const double* z, const double* vx, const double* vy, const double* vz,
                 const double* m, double e2, double G, double out[8]) {
    diagnostics_enqueue(ctx, n, x, y, z, vx, vy, vz, m, e2, G, out);
    ctx->sync();
}

}  // namespace orr
