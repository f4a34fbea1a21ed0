// Learned vertex-block preconditioner: graph-network training and inference
// kernels and the learned polynomial apply.  No reference counterpart (the
// reference defers learned preconditioners, REVIEW_AND_IMPROVEMENT_PLAN.md:
// 377-379); it plugs into the same preconditioner interface as Jacobi, SSOR
// and ILU (krylov_solvers.hpp:52-60).
//
// Preconditioner.  For a component-diagonal split matrix A = D + A_off,
//     u_0 = D^-1 r,   u_(k+1) = -D^-1 A_off u_k,   z = sum_k gamma_k * u_k,
// a truncated Neumann series of the Jacobi splitting whose terms carry
// per-vertex, per-component weights gamma_k.  gamma = 1 is the plain series.
// An apply is K off-diagonal SpMVs with a fused epilogue: HBM-bound, no
// triangular dependency chains.
//
// Network (weights BF16, arithmetic FP32), one thread per vertex:
//   features  x_i = [rho_ic, log(max(1 - rho_ic, 1e-6)) / 8,
//                    log(deg_i max_e |a_ec| / sum_e |a_ec|) / 2, deg_i / 8],
//             rho_ic = sum_e |a_ec| / d_ic (row dominance); the third group is
//             the row's edge anisotropy (0 when all its edges are equal)
//   weights   w_e = (1/b) sum_c |a_ec| / d_ic
//   encoder   h0 = tanh(W_in x + b_in)
//   3 layers  agg_i = sum_e w_e h_col(e),  h = tanh(W_s h + W_n agg + b)
//   decoder   o = W_out h3 + b_out,  gamma_kic = g_kc + o_(kb+c),i
//             (additive and unbounded; a zero decoder leaves the global
//             polynomial g)
//
// Training loss over (residual, error) pairs that the solver produced:
//   L = (1/P) sum_p |D^-1/2 A (z_p - e_p)|^2 / |D^-1/2 A e_p|^2,  z_p = M r_p.
// The basis vectors u_k(r_p) do not depend on the network, so a training
// problem keeps them resident in HBM and a step costs one SpMV and one
// transposed SpMV per pair plus the network's forward and backward passes.
// Weight gradients are per-block partial outer products summed in a fixed
// order by a finisher kernel: deterministic, no atomics.
#include "learned.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

using namespace owtk;

namespace owtk {
namespace {

using dev::kBlock;
constexpr int W = kGnnWidth;
constexpr int kTile = 64;            // vertices per LDS tile in the gradient reductions
constexpr int kReduceGrid = 240;     // blocks of the gradient reductions
constexpr int kMaxOuterOutputs = kGnnMaxOutputs * W + kGnnMaxOutputs;
constexpr int kOutPerThread = (kMaxOuterOutputs + kBlock - 1) / kBlock;

unsigned int rows_grid(std::int64_t n)
{
    const std::int64_t g = (n + kBlock - 1) / kBlock;
    if (g > 0x7fffffffLL) throw std::invalid_argument("too many vertices");
    return static_cast<unsigned int>(std::max<std::int64_t>(g, 1));
}

// ------------------------------------------------------------------ kernels --

// Matrix values -> graph arrays (FP32).  slot: device slot of each CSR entry.
__global__ __launch_bounds__(256) void k_gnn_gather(std::int64_t n, int b, const std::int32_t* __restrict__ offsets,
                                                    const std::int64_t* __restrict__ slot, int sell,
                                                    const double* __restrict__ v64, const double* __restrict__ d64,
                                                    float* __restrict__ vals, float* __restrict__ diag)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < b; ++c) {
        const std::int64_t di = sell ? (i & ~std::int64_t(63)) * b + 64 * c + (i & 63) : i * b + c;
        diag[i * b + c] = static_cast<float>(d64[di]);
    }
    for (std::int32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
        const std::int64_t s = slot[e];
        for (int c = 0; c < b; ++c) {
            const std::int64_t vi = sell ? (s & ~std::int64_t(63)) * b + 64 * c + (s & 63) : s * b + c;
            vals[static_cast<std::int64_t>(e) * b + c] = static_cast<float>(v64[vi]);
        }
    }
}

__global__ __launch_bounds__(256) void k_gnn_features(std::int64_t n, int b, int f0,
                                                      const std::int32_t* __restrict__ offsets,
                                                      const float* __restrict__ vals, const float* __restrict__ diag,
                                                      float* __restrict__ x, float* __restrict__ w)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    float mx[4] = {0.f, 0.f, 0.f, 0.f};
    float inv[4];
    for (int c = 0; c < b; ++c) inv[c] = 1.0f / diag[i * b + c];
    const std::int32_t e0 = offsets[i], e1 = offsets[i + 1];
    for (std::int32_t e = e0; e < e1; ++e) {
        float we = 0.f;
        for (int c = 0; c < b; ++c) {
            const float a = __builtin_fabsf(vals[static_cast<std::int64_t>(e) * b + c]);
            s[c] += a;
            mx[c] = fmaxf(mx[c], a);
            we += a * inv[c];
        }
        w[e] = we / static_cast<float>(b);
    }
    const float deg = static_cast<float>(e1 - e0);
    for (int c = 0; c < b; ++c) {
        const float rho = s[c] * inv[c];
        x[i * f0 + c] = rho;
        x[i * f0 + b + c] = logf(fmaxf(1.0f - rho, 1e-6f)) * 0.125f;
        x[i * f0 + 2 * b + c] = s[c] > 0.f ? logf(fmaxf(fmaxf(deg, 1.f) * mx[c] / s[c], 1.0f)) * 0.5f : 0.f;
    }
    x[i * f0 + 3 * b] = deg * 0.125f;
}

// k_gnn_gather + k_gnn_features for a SELL-layout matrix, read in its own
// slot order: row i's k-th off-diagonal entry sits in slot
// slice_ptr[i / 64] + 64 k + i % 64, so each k is one coalesced run, and the
// entries come in the CSR order of the row.  Same arithmetic, same values;
// the edge weights are stored per slot and the features feature-major,
// x[f n + i], so every store is coalesced.  Four slots per pass, all loads
// issued first (a slot past the slice's width is clamped to its last run,
// in bounds; one past the row's length is dropped).
__global__ __launch_bounds__(256) void k_gnn_features_sell(std::int64_t n, int b, int f0,
                                                           const std::int32_t* __restrict__ slice_ptr,
                                                           const std::int32_t* __restrict__ row_len,
                                                           const double* __restrict__ v64,
                                                           const double* __restrict__ d64, float* __restrict__ x,
                                                           float* __restrict__ w)
{
    constexpr int U = 4;
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const std::int64_t lane = i & 63;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    float mx[4] = {0.f, 0.f, 0.f, 0.f};
    float inv[4];
    for (int c = 0; c < b; ++c) inv[c] = 1.0f / static_cast<float>(d64[(i - lane) * b + 64 * c + lane]);
    const std::int64_t base = slice_ptr[i >> 6];
    const int width = static_cast<int>((slice_ptr[(i >> 6) + 1] - base) >> 6);
    const int len = row_len[i];
    for (int k0 = 0; k0 < len; k0 += U) {
        double va[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = k0 + u < width ? k0 + u : width - 1;
            const std::int64_t run = base + 64 * kk;
#pragma unroll
            for (int c = 0; c < 4; ++c) va[u][c] = c < b ? v64[run * b + 64 * c + lane] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + u >= len) break;
            float we = 0.f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (c >= b) break;
                const float a = __builtin_fabsf(static_cast<float>(va[u][c]));
                s[c] += a;
                mx[c] = fmaxf(mx[c], a);
                we += a * inv[c];
            }
            w[base + 64 * (k0 + u) + lane] = we / static_cast<float>(b);
        }
    }
    const float deg = static_cast<float>(len);
    for (int c = 0; c < b; ++c) {
        const float rho = s[c] * inv[c];
        x[c * n + i] = rho;
        x[(b + c) * n + i] = logf(fmaxf(1.0f - rho, 1e-6f)) * 0.125f;
        x[(2 * b + c) * n + i] = s[c] > 0.f ? logf(fmaxf(fmaxf(deg, 1.f) * mx[c] / s[c], 1.0f)) * 0.5f : 0.f;
    }
    x[3 * b * n + i] = deg * 0.125f;
}

// BF16 parameters -> LDS floats.
__device__ inline void load_weights(float* dst, const std::uint16_t* __restrict__ p16, std::int64_t off, int count)
{
    for (int t = threadIdx.x; t < count; t += blockDim.x) dst[t] = dev::bf16_to_float(p16[off + t]);
}

__global__ __launch_bounds__(256) void k_gnn_encode(std::int64_t n, int f0, const float* __restrict__ x,
                                                    const std::uint16_t* __restrict__ p16, std::int64_t off_w,
                                                    std::int64_t off_b, float* __restrict__ h)
{
    __shared__ float sw[W * kGnnMaxFeatures];
    __shared__ float sb[W];
    load_weights(sw, p16, off_w, W * f0);
    load_weights(sb, p16, off_b, W);
    __syncthreads();
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float xi[kGnnMaxFeatures];
#pragma unroll
    for (int f = 0; f < kGnnMaxFeatures; ++f) xi[f] = f < f0 ? x[i * f0 + f] : 0.f;
    float4* out = reinterpret_cast<float4*>(h + i * W);
    for (int q = 0; q < W / 4; ++q) {
        float r[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 4 * q + j;
            float acc = sb[o];
#pragma unroll
            for (int f = 0; f < kGnnMaxFeatures; ++f)
                if (f < f0) acc = __builtin_fmaf(sw[o * f0 + f], xi[f], acc);
            r[j] = tanhf(acc);
        }
        out[q] = make_float4(r[0], r[1], r[2], r[3]);
    }
}

__device__ inline void load_row(float (&v)[W], const float* __restrict__ p)
{
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int j = 0; j < W / 4; ++j) {
        const float4 t = q[j];
        v[4 * j] = t.x;
        v[4 * j + 1] = t.y;
        v[4 * j + 2] = t.z;
        v[4 * j + 3] = t.w;
    }
}

__device__ inline void store_row(float* __restrict__ p, const float (&v)[W])
{
    float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
    for (int j = 0; j < W / 4; ++j) q[j] = make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
}

// One message-passing layer.  agg_out (training) keeps the aggregate.
__global__ __launch_bounds__(256) void k_gnn_layer(std::int64_t n, const std::int32_t* __restrict__ offsets,
                                                   const std::int32_t* __restrict__ cols, const float* __restrict__ w,
                                                   const float* __restrict__ hin,
                                                   const std::uint16_t* __restrict__ p16, std::int64_t off_ws,
                                                   std::int64_t off_wn, std::int64_t off_b,
                                                   float* __restrict__ hout, float* __restrict__ agg_out)
{
    __shared__ float sws[W * W];
    __shared__ float swn[W * W];
    __shared__ float sb[W];
    load_weights(sws, p16, off_ws, W * W);
    load_weights(swn, p16, off_wn, W * W);
    load_weights(sb, p16, off_b, W);
    __syncthreads();
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) a[k] = 0.f;
    for (std::int32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
        const float we = w[e];
        const float4* hj = reinterpret_cast<const float4*>(hin + static_cast<std::int64_t>(cols[e]) * W);
#pragma unroll
        for (int q = 0; q < W / 4; ++q) {
            const float4 t = hj[q];
            a[4 * q] = __builtin_fmaf(we, t.x, a[4 * q]);
            a[4 * q + 1] = __builtin_fmaf(we, t.y, a[4 * q + 1]);
            a[4 * q + 2] = __builtin_fmaf(we, t.z, a[4 * q + 2]);
            a[4 * q + 3] = __builtin_fmaf(we, t.w, a[4 * q + 3]);
        }
    }
    if (agg_out) store_row(agg_out + i * W, a);
    float h[W];
    load_row(h, hin + i * W);
    float4* out = reinterpret_cast<float4*>(hout + i * W);
    for (int q = 0; q < W / 4; ++q) {
        float r[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 4 * q + j;
            float acc = sb[o];
#pragma unroll
            for (int k = 0; k < W; ++k) acc = __builtin_fmaf(sws[o * W + k], h[k], acc);
#pragma unroll
            for (int k = 0; k < W; ++k) acc = __builtin_fmaf(swn[o * W + k], a[k], acc);
            r[j] = tanhf(acc);
        }
        out[q] = make_float4(r[0], r[1], r[2], r[3]);
    }
}

using f32x16 = float __attribute__((ext_vector_type(16)));
constexpr int kMfmaEdges = 2;  // edges gathered per pass of the aggregation

// One message-passing layer over a SELL-layout graph on the matrix cores
// (inference).  A wave computes a tile of 32 vertices:
//     out = tanh([h_v | a_v] Wcat^T + b),  Wcat = [Ws | Wn] (32 x 64),
// as 32 steps of v_mfma_f32_32x32x2_f32.  Lane l supplies row v = l % 32 of
// the 32 x 2 A slice and column o = l % 32 of the 2 x 32 B slice; the two
// half-waves (h = l / 32) split the 64 inputs, step t taking
//     t < 16: self feature 16 h + t,    t >= 16: aggregate 16 h + t - 16,
// so each lane gathers and sums only its half of its vertex's features, and
// its 32 weights stay in registers for the whole launch.  The aggregate is
// summed in edge order exactly as k_gnn_layer; the products are accumulated
// in the matrix cores' order.  D: lane l holds column o, rows
// 8 (r / 4) + 4 h + r % 4 of register r.
// The input encoding on the matrix cores (inference): a wave computes a
// tile of 32 vertices, h = tanh(x W_in^T + b_in) with x feature-major
// (x[f n + v], as k_gnn_features_sell writes it), the f0 <= 16 features
// padded with zeros to 16 and split over the half-waves (step t takes
// feature t + 8 h), 8 steps of v_mfma_f32_32x32x2_f32; output layout as in
// k_gnn_layer_mfma, so each store instruction writes two whole rows of h.
static_assert(kGnnMaxFeatures <= 16, "the matrix-core encoding takes at most 16 features");
__global__ __launch_bounds__(256) void k_gnn_encode_mfma(std::int64_t n, int f0, const float* __restrict__ x,
                                                         const std::uint16_t* __restrict__ p16, std::int64_t off_w,
                                                         std::int64_t off_b, float* __restrict__ h)
{
    const int l = static_cast<int>(threadIdx.x & 63);
    const int o = l & 31, half = l >> 5;
    float bw[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int f = 8 * half + t;
        bw[t] = f < f0 ? dev::bf16_to_float(p16[off_w + o * f0 + f]) : 0.f;
    }
    const float bias = dev::bf16_to_float(p16[off_b + o]);
    const std::int64_t tiles = (n + 31) >> 5;
    const std::int64_t waves = static_cast<std::int64_t>(gridDim.x) * (blockDim.x >> 6);
    for (std::int64_t tile = static_cast<std::int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
         tile < tiles; tile += waves) {
        const std::int64_t v = tile * 32 + o;
        const bool valid = v < n;
        float av[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int f = 8 * half + t;
            av[t] = valid && f < f0 ? x[f * n + v] : 0.f;  // feature-major
        }
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = bias;
#pragma unroll
        for (int t = 0; t < 8; ++t) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], bw[t], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const std::int64_t vv = tile * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            if (vv < n) h[vv * W + o] = tanhf(acc[r]);
        }
    }
}

// Three waves per SIMD (at most 168 registers): a tile waits on its gathers,
// and 2 edges per pass with 3 waves measured faster than 4 edges with 2.
__global__ __launch_bounds__(256, 3) void k_gnn_layer_mfma(std::int64_t n, const std::int32_t* __restrict__ slice_ptr,
                                                        const std::int32_t* __restrict__ row_len,
                                                        const std::int32_t* __restrict__ cols,
                                                        const float* __restrict__ w, const float* __restrict__ hin,
                                                        const std::uint16_t* __restrict__ p16, std::int64_t off_ws,
                                                        std::int64_t off_wn, std::int64_t off_b,
                                                        float* __restrict__ hout)
{
    const int l = static_cast<int>(threadIdx.x & 63);
    const int o = l & 31, half = l >> 5;
    float bw[32];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        bw[t] = dev::bf16_to_float(p16[off_ws + o * W + 16 * half + t]);
        bw[16 + t] = dev::bf16_to_float(p16[off_wn + o * W + 16 * half + t]);
    }
    const float bias = dev::bf16_to_float(p16[off_b + o]);
    const std::int64_t tiles = (n + 31) >> 5;
    const std::int64_t waves = static_cast<std::int64_t>(gridDim.x) * (blockDim.x >> 6);
    for (std::int64_t tile = static_cast<std::int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
         tile < tiles; tile += waves) {
        const std::int64_t v = tile * 32 + o;
        const bool valid = v < n;
        const std::int64_t vc = valid ? v : n - 1;  // in-bounds stand-in for a tail lane
        float av[32];
        {
            const float4* hs = reinterpret_cast<const float4*>(hin + vc * W + 16 * half);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 t = hs[q];
                av[4 * q] = t.x;
                av[4 * q + 1] = t.y;
                av[4 * q + 2] = t.z;
                av[4 * q + 3] = t.w;
            }
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) av[16 + j] = 0.f;
        const std::int64_t base = slice_ptr[vc >> 6];
        const int width = static_cast<int>((slice_ptr[(vc >> 6) + 1] - base) >> 6);
        const int len = valid ? row_len[vc] : 0;
        // kMfmaEdges edges per pass, every load issued before the sums: a
        // slot past the slice's width is clamped to its last run (in
        // bounds) and an edge past the row's length is dropped by a select
        for (int k0 = 0; k0 < len; k0 += kMfmaEdges) {
            int ck[kMfmaEdges];
            float wk[kMfmaEdges];
#pragma unroll
            for (int u = 0; u < kMfmaEdges; ++u) {
                const int kk = k0 + u < width ? k0 + u : width - 1;
                const std::int64_t slot = base + 64 * kk + (vc & 63);
                ck[u] = cols[slot];
                wk[u] = w[slot];
            }
            float4 hk[kMfmaEdges][4];
#pragma unroll
            for (int u = 0; u < kMfmaEdges; ++u) {
                const float4* hj = reinterpret_cast<const float4*>(hin + static_cast<std::int64_t>(ck[u]) * W + 16 * half);
#pragma unroll
                for (int q = 0; q < 4; ++q) hk[u][q] = hj[q];
            }
#pragma unroll
            for (int u = 0; u < kMfmaEdges; ++u) {
                const bool on = k0 + u < len;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 t = hk[u][q];
                    float* a = av + 16 + 4 * q;
                    const float f0 = __builtin_fmaf(wk[u], t.x, a[0]);
                    const float f1 = __builtin_fmaf(wk[u], t.y, a[1]);
                    const float f2 = __builtin_fmaf(wk[u], t.z, a[2]);
                    const float f3 = __builtin_fmaf(wk[u], t.w, a[3]);
                    a[0] = on ? f0 : a[0];
                    a[1] = on ? f1 : a[1];
                    a[2] = on ? f2 : a[2];
                    a[3] = on ? f3 : a[3];
                }
            }
        }
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = bias;
#pragma unroll
        for (int t = 0; t < 32; ++t) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], bw[t], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const std::int64_t vv = tile * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            if (vv < n) hout[vv * W + o] = tanhf(acc[r]);
        }
    }
}

__global__ __launch_bounds__(256) void k_gnn_decode(std::int64_t n, int b, int outs, const float* __restrict__ h3,
                                                    const std::uint16_t* __restrict__ p16, std::int64_t off_w,
                                                    std::int64_t off_b, std::int64_t off_g,
                                                    float* __restrict__ gamma)
{
    __shared__ float sw[kGnnMaxOutputs * W];
    __shared__ float sb[kGnnMaxOutputs];
    __shared__ float sg[kGnnMaxOutputs];
    load_weights(sw, p16, off_w, outs * W);
    load_weights(sb, p16, off_b, outs);
    load_weights(sg, p16, off_g, outs);
    __syncthreads();
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float h[W];
    load_row(h, h3 + i * W);
    const std::int64_t nb = n * b;
    for (int o = 0; o < outs; ++o) {
        float acc = sb[o];
#pragma unroll
        for (int k = 0; k < W; ++k) acc = __builtin_fmaf(sw[o * W + k], h[k], acc);
        const int kk = o / b, c = o - kk * b;
        gamma[kk * nb + i * b + c] = sg[o] + acc;
    }
}

// ------------------------------------------------------- backward kernels --

// d gamma -> d o (do_out, which is also each vertex's contribution to d g)
// and d h3 = W_out^T d o.
__global__ __launch_bounds__(256) void k_gnn_decode_bwd(std::int64_t n, int b, int outs,
                                                        const float* __restrict__ dgamma,
                                                        const std::uint16_t* __restrict__ p16, std::int64_t off_w,
                                                        float* __restrict__ do_out, float* __restrict__ dh)
{
    __shared__ float sw[kGnnMaxOutputs * W];
    load_weights(sw, p16, off_w, outs * W);
    __syncthreads();
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const std::int64_t nb = n * b;
    float acc[W];
#pragma unroll
    for (int k = 0; k < W; ++k) acc[k] = 0.f;
    for (int o = 0; o < outs; ++o) {
        const int kk = o / b, c = o - kk * b;
        const float d = dgamma[kk * nb + i * b + c];
        do_out[i * outs + o] = d;
#pragma unroll
        for (int k = 0; k < W; ++k) acc[k] = __builtin_fmaf(sw[o * W + k], d, acc[k]);
    }
    store_row(dh + i * W, acc);
}

// du = dh (1 - h^2); s = W_s^T du (written to dprev); v = W_n^T du.
__global__ __launch_bounds__(256) void k_gnn_layer_bwd(std::int64_t n, const float* __restrict__ h,
                                                       const float* __restrict__ dh,
                                                       const std::uint16_t* __restrict__ p16, std::int64_t off_ws,
                                                       std::int64_t off_wn, float* __restrict__ du_out,
                                                       float* __restrict__ dprev, float* __restrict__ v_out)
{
    __shared__ float sws[W * W];
    __shared__ float swn[W * W];
    load_weights(sws, p16, off_ws, W * W);
    load_weights(swn, p16, off_wn, W * W);
    __syncthreads();
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float hv[W], du[W];
    load_row(hv, h + i * W);
    load_row(du, dh + i * W);
#pragma unroll
    for (int k = 0; k < W; ++k) du[k] *= 1.0f - hv[k] * hv[k];
    store_row(du_out + i * W, du);
    float s[W], v[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        s[k] = 0.f;
        v[k] = 0.f;
    }
#pragma unroll
    for (int o = 0; o < W; ++o) {
        const float d = du[o];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            s[k] = __builtin_fmaf(sws[o * W + k], d, s[k]);
            v[k] = __builtin_fmaf(swn[o * W + k], d, v[k]);
        }
    }
    store_row(dprev + i * W, s);
    store_row(v_out + i * W, v);
}

// dprev_j += sum_(e in row j) w_mirror(e) v_col(e): the transposed aggregation
// (pattern symmetric, weight of entry (col, j) found through the mirror).
__global__ __launch_bounds__(256) void k_gnn_agg_t(std::int64_t n, const std::int32_t* __restrict__ offsets,
                                                   const std::int32_t* __restrict__ cols,
                                                   const std::int32_t* __restrict__ mirror,
                                                   const float* __restrict__ w, const float* __restrict__ v,
                                                   float* __restrict__ dprev)
{
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float acc[W];
    load_row(acc, dprev + j * W);
    for (std::int32_t e = offsets[j]; e < offsets[j + 1]; ++e) {
        const float we = w[mirror[e]];
        const float4* vi = reinterpret_cast<const float4*>(v + static_cast<std::int64_t>(cols[e]) * W);
#pragma unroll
        for (int q = 0; q < W / 4; ++q) {
            const float4 t = vi[q];
            acc[4 * q] = __builtin_fmaf(we, t.x, acc[4 * q]);
            acc[4 * q + 1] = __builtin_fmaf(we, t.y, acc[4 * q + 1]);
            acc[4 * q + 2] = __builtin_fmaf(we, t.z, acc[4 * q + 2]);
            acc[4 * q + 3] = __builtin_fmaf(we, t.w, acc[4 * q + 3]);
        }
    }
    store_row(dprev + j * W, acc);
}

__global__ __launch_bounds__(256) void k_tanh_bwd(std::int64_t count, const float* __restrict__ h,
                                                  float* __restrict__ d)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < count; t += step)
        d[t] *= 1.0f - h[t] * h[t];
}

// Partial sums over vertex tiles of A_i B_i^T (R x C) and A_i (R):
// partials[block][R*C + R].  B may be null (C = 0).
__global__ __launch_bounds__(256) void k_outer_partials(std::int64_t n, const float* __restrict__ A, int R,
                                                        const float* __restrict__ B, int C,
                                                        float* __restrict__ partials)
{
    __shared__ float sa[kTile * kGnnMaxOutputs];
    __shared__ float sb[kTile * W];
    const int outs = R * C + R;
    float acc[kOutPerThread];
#pragma unroll
    for (int j = 0; j < kOutPerThread; ++j) acc[j] = 0.f;
    for (std::int64_t tile = static_cast<std::int64_t>(blockIdx.x) * kTile; tile < n;
         tile += static_cast<std::int64_t>(gridDim.x) * kTile) {
        const int cnt = static_cast<int>(std::min<std::int64_t>(kTile, n - tile));
        for (int q = threadIdx.x; q < cnt * R; q += blockDim.x) sa[q] = A[tile * R + q];
        for (int q = threadIdx.x; q < cnt * C; q += blockDim.x) sb[q] = B[tile * C + q];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kOutPerThread; ++j) {
            const int o = threadIdx.x + kBlock * j;
            if (o < R * C) {
                const int r = o / C, c = o - (o / C) * C;
                float s = 0.f;
                for (int t = 0; t < cnt; ++t) s = __builtin_fmaf(sa[t * R + r], sb[t * C + c], s);
                acc[j] += s;
            } else if (o < outs) {
                const int r = o - R * C;
                float s = 0.f;
                for (int t = 0; t < cnt; ++t) s += sa[t * R + r];
                acc[j] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < kOutPerThread; ++j) {
        const int o = threadIdx.x + kBlock * j;
        if (o < outs) partials[static_cast<std::int64_t>(blockIdx.x) * outs + o] = acc[j];
    }
}

// grad[mat + r*C + c] += sum of partials (and grad[vec + r] for the sums).
__global__ __launch_bounds__(256) void k_outer_finish(const float* __restrict__ partials, int blocks, int R, int C,
                                                      float* __restrict__ mat, float* __restrict__ vec)
{
    const int outs = R * C + R;
    const int o = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= outs) return;
    double s = 0.0;
    for (int k = 0; k < blocks; ++k) s += static_cast<double>(partials[static_cast<std::int64_t>(k) * outs + o]);
    if (o < R * C) {
        if (mat) mat[o] += static_cast<float>(s);
    } else if (vec) {
        vec[o - R * C] += static_cast<float>(s);
    }
}

// ------------------------------------------------------------ loss kernels --

// z[p] = sum_k gamma_k * u[p][k]
__global__ __launch_bounds__(256) void k_loss_z(std::int64_t nb, int pairs, int k1, const float* __restrict__ gamma,
                                                const float* __restrict__ u, float* __restrict__ z)
{
    const std::int64_t total = nb * pairs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t p = t / nb, q = t - p * nb;
        float s = 0.f;
        for (int k = 0; k < k1; ++k) s = __builtin_fmaf(gamma[k * nb + q], u[(p * k1 + k) * nb + q], s);
        z[t] = s;
    }
}

// y = D^-1/2 A z_p - target_p; ys = D^-1/2 y; partial sums of coef_p y^2.
// Grid: (row blocks, pairs).
__global__ __launch_bounds__(256) void k_loss_y(std::int64_t n, int b, const std::int32_t* __restrict__ offsets,
                                                const std::int32_t* __restrict__ cols,
                                                const float* __restrict__ vals, const float* __restrict__ diag,
                                                const float* __restrict__ dis, const float* __restrict__ z,
                                                const float* __restrict__ target, const float* __restrict__ coef,
                                                float* __restrict__ ys, double* __restrict__ partials)
{
    const std::int64_t nb = n * b;
    const int p = blockIdx.y;
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double local = 0.0;
    if (i < n) {
        const float* zp = z + p * nb;
        for (int c = 0; c < b; ++c) {
            float acc = diag[i * b + c] * zp[i * b + c];
            for (std::int32_t e = offsets[i]; e < offsets[i + 1]; ++e)
                acc = __builtin_fmaf(vals[static_cast<std::int64_t>(e) * b + c], zp[static_cast<std::int64_t>(cols[e]) * b + c], acc);
            const float d = dis[i * b + c];
            const float y = d * acc - target[p * nb + i * b + c];
            ys[p * nb + i * b + c] = d * y;
            local += static_cast<double>(y) * static_cast<double>(y);
        }
        local *= static_cast<double>(coef[p]);
    }
    const double s = dev::block_sum_all(local);
    if (threadIdx.x == 0) partials[static_cast<std::int64_t>(p) * gridDim.x + blockIdx.x] = s;
}

__global__ void k_loss_finish(const double* __restrict__ partials, int count, double scale, double* __restrict__ loss)
{
    double s = 0.0;
    for (int t = threadIdx.x; t < count; t += blockDim.x) s += partials[t];
    s = dev::block_sum_all(s);
    if (threadIdx.x == 0) loss[0] += scale * s;
}

// g_p = gcoef_p A^T ys_p   (transposed SpMV through the mirror index)
__global__ __launch_bounds__(256) void k_loss_g(std::int64_t n, int b, const std::int32_t* __restrict__ offsets,
                                                const std::int32_t* __restrict__ cols,
                                                const std::int32_t* __restrict__ mirror,
                                                const float* __restrict__ vals, const float* __restrict__ diag,
                                                const float* __restrict__ ys, const float* __restrict__ gcoef,
                                                float* __restrict__ gout)
{
    const std::int64_t nb = n * b;
    const int p = blockIdx.y;
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float* yp = ys + p * nb;
    for (int c = 0; c < b; ++c) {
        float g = diag[j * b + c] * yp[j * b + c];
        for (std::int32_t e = offsets[j]; e < offsets[j + 1]; ++e)
            g = __builtin_fmaf(vals[static_cast<std::int64_t>(mirror[e]) * b + c],
                               yp[static_cast<std::int64_t>(cols[e]) * b + c], g);
        gout[p * nb + j * b + c] = g * gcoef[p];
    }
}

// dgamma_k = sum_p g_p * u[p][k]
__global__ __launch_bounds__(256) void k_loss_dgamma(std::int64_t nb, int pairs, int k1, const float* __restrict__ g,
                                                     const float* __restrict__ u, float* __restrict__ dgamma)
{
    const std::int64_t total = nb * k1;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t k = t / nb, q = t - k * nb;
        float s = 0.f;
        for (int p = 0; p < pairs; ++p) s = __builtin_fmaf(g[p * nb + q], u[(p * k1 + k) * nb + q], s);
        dgamma[t] = s;
    }
}

// ------------------------------------------------- training data kernels --

// u0 = r / d, target base (FP32 copies of one pair's FP64 vectors).
__global__ __launch_bounds__(256) void k_basis_init(std::int64_t nb, const double* __restrict__ r,
                                                    const float* __restrict__ diag, float* __restrict__ u0)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < nb; t += step)
        u0[t] = static_cast<float>(r[t]) / diag[t];
}

// u_next = -D^-1 A_off u   (graph arrays)
__global__ __launch_bounds__(256) void k_basis_step(std::int64_t n, int b, const std::int32_t* __restrict__ offsets,
                                                    const std::int32_t* __restrict__ cols,
                                                    const float* __restrict__ vals, const float* __restrict__ diag,
                                                    const float* __restrict__ u, float* __restrict__ un)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < b; ++c) {
        float acc = 0.f;
        for (std::int32_t e = offsets[i]; e < offsets[i + 1]; ++e)
            acc = __builtin_fmaf(vals[static_cast<std::int64_t>(e) * b + c], u[static_cast<std::int64_t>(cols[e]) * b + c], acc);
        un[i * b + c] = -acc / diag[i * b + c];
    }
}

// target = D^-1/2 A e  (FP64 arithmetic from the FP32 graph values)
__global__ __launch_bounds__(256) void k_target(std::int64_t n, int b, const std::int32_t* __restrict__ offsets,
                                                const std::int32_t* __restrict__ cols,
                                                const float* __restrict__ vals, const float* __restrict__ diag,
                                                const double* __restrict__ e, float* __restrict__ target,
                                                float* __restrict__ dis)
{
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < b; ++c) {
        double acc = static_cast<double>(diag[i * b + c]) * e[i * b + c];
        for (std::int32_t en = offsets[i]; en < offsets[i + 1]; ++en)
            acc = __builtin_fma(static_cast<double>(vals[static_cast<std::int64_t>(en) * b + c]),
                                e[static_cast<std::int64_t>(cols[en]) * b + c], acc);
        const double s = 1.0 / __builtin_sqrt(static_cast<double>(diag[i * b + c]));
        target[i * b + c] = static_cast<float>(s * acc);
        dis[i * b + c] = static_cast<float>(s);
    }
}

// ------------------------------------------------------------------- Adam --
__global__ __launch_bounds__(256) void k_adam(std::int64_t count, float* __restrict__ params, float* __restrict__ grad,
                                              float* __restrict__ m, float* __restrict__ v,
                                              std::uint16_t* __restrict__ p16, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2)
{
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float g = grad[t];
    const float mt = b1 * m[t] + (1.0f - b1) * g;
    const float vt = b2 * v[t] + (1.0f - b2) * g * g;
    m[t] = mt;
    v[t] = vt;
    const float p = params[t] - lr * (mt / bc1) / (__builtin_sqrtf(vt / bc2) + eps);
    params[t] = p;
    p16[t] = dev::float_to_bf16(p);
    grad[t] = 0.0f;
}

__global__ __launch_bounds__(256) void k_round_bf16(std::int64_t count, const float* __restrict__ params,
                                                    std::uint16_t* __restrict__ p16)
{
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < count) p16[t] = dev::float_to_bf16(params[t]);
}

// ------------------------------------------------------------ apply kernels --

// w_0 = r and z = gamma'_0 r (the apply's first step run on its own: degree 0,
// or an in-place apply, whose first step may not gather the input it overwrites).
template<class X>
__global__ __launch_bounds__(256) void k_series_init(std::int64_t nb, const X* __restrict__ in,
                                                     const X* __restrict__ gamma0, X* __restrict__ w0,
                                                     X* __restrict__ z, const int* stop)
{
    if (stop != nullptr && *stop != 0) return;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < nb; t += step) {
        const X r = in[t];
        if (w0) w0[t] = r;
        z[t] = gamma0[t] * r;
    }
}

// 1 / d for every diagonal entry, in the diagonal's own layout: the IEEE
// quotient the column and coefficient scalings below multiply by, formed once
// per entry instead of once per off-diagonal value.
__global__ __launch_bounds__(256) void k_inv_diag(std::int64_t count, const double* __restrict__ d64,
                                                  double* __restrict__ dinv)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < count; t += step)
        dinv[t] = 1.0 / d64[t];
}

// Column-scaled off-diagonal values a_ec * (1 / d_(col e, c)) in the matrix's
// value layout (SELL-64 or CSR), rounded once from the FP64 product.  A
// thread takes one slot (its column loaded once) and all its components.
template<class V>
__global__ __launch_bounds__(256) void k_scale_columns(std::int64_t slots, int b, int sell,
                                                       const std::int32_t* __restrict__ cols,
                                                       const double* __restrict__ v64,
                                                       const double* __restrict__ dinv, V* __restrict__ out)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t sl = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; sl < slots; sl += step) {
        const std::int64_t j = cols[sl];
        for (int c = 0; c < b; ++c) {
            // SELL: value (slot, c) at (slot & ~63) b + 64 c + slot % 64, diagonal (j, c) likewise
            const std::int64_t t = sell ? (sl & ~std::int64_t(63)) * b + 64 * c + (sl & 63) : sl * b + c;
            const std::int64_t dj = sell ? (j & ~std::int64_t(63)) * b + 64 * c + (j & 63) : j * b + c;
            const double a = v64[t] * dinv[dj];
            if constexpr (sizeof(V) == 2) out[t] = dev::float_to_bf16(static_cast<float>(a));
            else out[t] = static_cast<V>(a);
        }
    }
}

// gamma'_k = gamma_k * (1 / d), rounded once from the FP64 product, k = 0..K.
// One thread per (vertex, component), consecutive threads on consecutive
// entries; the vertex index by a 32-bit division when n b fits.
template<class X>
__global__ __launch_bounds__(256) void k_scale_gamma(std::int64_t nb, int b, int sell, int k1,
                                                     const float* __restrict__ gamma, const double* __restrict__ dinv,
                                                     X* __restrict__ out)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    const bool narrow = nb < (std::int64_t(1) << 31);
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < nb; t += step) {
        const std::int64_t i = narrow ? static_cast<std::int64_t>(static_cast<std::uint32_t>(t) / static_cast<std::uint32_t>(b))
                                      : t / b;
        const int c = static_cast<int>(t - i * b);
        const std::int64_t di = sell ? (i & ~std::int64_t(63)) * b + 64 * c + (i & 63) : t;
        const double inv = dinv[di];
        for (int k = 0; k < k1; ++k)
            out[k * nb + t] = static_cast<X>(static_cast<double>(gamma[k * nb + t]) * inv);
    }
}

} // namespace
} // namespace owtk

namespace owtk::dev {

// w_next = -(A_off D^-1) w, z += gamma'_k w_next.  The last step of an apply
// does not store w_next (nothing reads it).
template<class X>
struct EpiSeries : EpiNone {
    static constexpr bool kKeepMatrix = true;  // the apply's next step rereads it
    X* w_next;
    const X* gamma;
    X* z;
    bool store_w = true;
    template<class A>
    __device__ void elem(std::int64_t i, A acc, double*) const
    {
        const X w = -static_cast<X>(acc);
        if (store_w) w_next[i] = w;
        z[i] = dfma(gamma[i], w, z[i]);
    }
    template<int BS, class A>
    __device__ void elem_row(std::int64_t i0, const A (&acc)[BS], double*) const
    {
        X g[BS], zz[BS], w[BS];
        load_row<BS>(gamma + i0, g);
        load_row<BS>(z + i0, zz);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            w[c] = -static_cast<X>(acc[c]);
            zz[c] = dfma(g[c], w[c], zz[c]);
        }
        if (store_w) store_row<BS>(w_next + i0, w);
        store_row<BS>(z + i0, zz);
    }
};

// First step with the initialisation folded in: w_0 = r is gathered as it is,
// and z = gamma'_0 r + gamma'_1 w_1 is written without a prior z, with the
// roundings of z = gamma'_0 * r, z = fma(gamma'_1, w_1, z) (k_series_init
// followed by an EpiSeries step).
template<class X>
struct EpiSeriesFirst : EpiNone {
    static constexpr bool kKeepMatrix = true;
    X* w_next;
    const X* r;
    const X* gamma0;
    const X* gamma1;
    X* z;
    bool store_w = true;
    template<class A>
    __device__ void elem(std::int64_t i, A acc, double*) const
    {
        const X w = -static_cast<X>(acc);
        if (store_w) w_next[i] = w;
        z[i] = dfma(gamma1[i], w, gamma0[i] * r[i]);
    }
    template<int BS, class A>
    __device__ void elem_row(std::int64_t i0, const A (&acc)[BS], double*) const
    {
        X rr[BS], g0[BS], g1[BS], w[BS], zz[BS];
        load_row<BS>(r + i0, rr);
        load_row<BS>(gamma0 + i0, g0);
        load_row<BS>(gamma1 + i0, g1);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            w[c] = -static_cast<X>(acc[c]);
            zz[c] = dfma(g1[c], w[c], g0[c] * rr[c]);
        }
        if (store_w) store_row<BS>(w_next + i0, w);
        store_row<BS>(z + i0, zz);
    }
};

} // namespace owtk::dev

namespace owtk {
namespace {

// Off-diagonal SpMV of a split matrix with epilogue e.
template<class V, class X, class E>
void offdiag_spmv(owtk_matrix m, const V* vals, const X* x, E e)
{
    using namespace owtk::kern;
    const hipStream_t st = m->ctx->stream;
    const auto scratch = m->ctx->scratch();
    const auto& s = m->s;
    const std::int64_t rows = m->owned;
    if (s.layout == Layout::sell) {
        const unsigned int grid = rows_grid(rows);
        launch_sell_cd_bs<V, X, false, E>(m->bs, s.max_width, grid, st, s.slice_ptr.get(), s.row_len.get(),
                                          s.columns.get(), vals, static_cast<const V*>(nullptr), x, rows, e, scratch);
    } else {
        std::int64_t blocks = (rows + dev::kWavesPerBlock - 1) / dev::kWavesPerBlock;
        blocks = std::max<std::int64_t>(1, std::min<std::int64_t>(blocks, dev::kMaxReduceBlocks));
        spmv_csr_cd<V, X, false, E><<<static_cast<unsigned int>(blocks), kBlock, 0, st>>>(
            s.offsets.get(), s.columns.get(), vals, nullptr, x, rows, m->bs, e, scratch);
    }
    OWTK_LAUNCH_CHECK();
}

// Host-side mirror index of a structurally symmetric pattern.
std::vector<std::int32_t> mirror_index(std::int64_t n, const std::vector<std::int32_t>& off,
                                       const std::vector<std::int32_t>& col)
{
    const std::int64_t nnz = off[n];
    // per row: entries sorted by column
    std::vector<std::int32_t> order(nnz);
    std::iota(order.begin(), order.end(), 0);
    for (std::int64_t i = 0; i < n; ++i)
        std::sort(order.begin() + off[i], order.begin() + off[i + 1],
                  [&](std::int32_t a, std::int32_t b) { return col[a] < col[b]; });
    std::vector<std::int32_t> mir(nnz);
    for (std::int64_t i = 0; i < n; ++i) {
        for (std::int32_t e = off[i]; e < off[i + 1]; ++e) {
            const std::int32_t j = col[e];
            auto first = order.begin() + off[j], last = order.begin() + off[j + 1];
            auto it = std::lower_bound(first, last, static_cast<std::int32_t>(i),
                                       [&](std::int32_t a, std::int32_t v) { return col[a] < v; });
            if (it == last || col[*it] != i)
                throw std::invalid_argument("learned preconditioner training needs a structurally symmetric pattern");
            mir[e] = *it;
        }
    }
    return mir;
}

template<class T>
void upload_vec(DeviceArray<T>& d, const std::vector<T>& h)
{
    d.allocate(std::max<std::size_t>(h.size(), 1));
    if (!h.empty()) OWTK_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

void check_matrix(owtk_gnn g, owtk_matrix m)
{
    if (!m) throw std::invalid_argument("null matrix");
    if (m->kind != OWTK_MATRIX_SPLIT) throw unsupported_error("learned preconditioner needs a split matrix");
    if (m->ghost != 0) throw unsupported_error("learned preconditioner needs a matrix without ghost nodes");
    if (m->bs != g->b) throw std::invalid_argument("matrix block size differs from the model's");
    if (m->ctx != g->ctx) throw std::invalid_argument("matrix and model belong to different contexts");
}

// The graph's structure (CSR, mirror, the matrix slot of each entry) and
// its arrays, built once per matrix pattern.
void build_graph_structure(owtk_gnn model, owtk_matrix m, GnnGraph& g, bool with_mirror)
{
    check_matrix(model, m);
    g.n = m->owned;
    g.b = m->bs;
    g.nnz = m->h_offsets[g.n];
    if (!with_mirror && m->s.layout == Layout::sell) {
        g.sell = true;
        g.slice_ptr = m->s.slice_ptr.get();
        g.row_len = m->s.row_len.get();
        g.slot_cols = m->s.columns.get();
        g.slot_w.allocate(std::max<std::int64_t>(m->s.slots, 1));
        g.x.allocate(g.n * model->f0);
        return;
    }
    upload_vec(g.offsets, m->h_offsets);
    upload_vec(g.cols, m->h_columns);
    if (with_mirror) upload_vec(g.mirror, mirror_index(g.n, m->h_offsets, m->h_columns));
    std::vector<std::int64_t> slot(m->s.slot_of_entry.begin(), m->s.slot_of_entry.end());
    if (slot.empty()) {
        slot.resize(g.nnz);
        std::iota(slot.begin(), slot.end(), 0);
    }
    upload_vec(g.slot, slot);
    g.vals.allocate(std::max<std::int64_t>(g.nnz * g.b, 1));
    g.diag.allocate(g.n * g.b);
    g.x.allocate(g.n * model->f0);
    g.w.allocate(std::max<std::int64_t>(g.nnz, 1));
}

// The graph's values and features from the matrix's current FP64 values.
void graph_values(owtk_gnn model, owtk_matrix m, GnnGraph& g)
{
    const hipStream_t st = m->ctx->stream;
    const unsigned int grid = rows_grid(g.n);
    if (g.sell) {
        k_gnn_features_sell<<<grid, kBlock, 0, st>>>(g.n, g.b, model->f0, g.slice_ptr, g.row_len, m->v64.get(),
                                                     m->d64.get(), g.x.get(), g.slot_w.get());
        OWTK_LAUNCH_CHECK();
        return;
    }
    k_gnn_gather<<<grid, kBlock, 0, st>>>(g.n, g.b, g.offsets.get(), g.slot.get(), m->s.layout == Layout::sell ? 1 : 0,
                                          m->v64.get(), m->d64.get(), g.vals.get(), g.diag.get());
    OWTK_LAUNCH_CHECK();
    k_gnn_features<<<grid, kBlock, 0, st>>>(g.n, g.b, model->f0, g.offsets.get(), g.vals.get(), g.diag.get(),
                                            g.x.get(), g.w.get());
    OWTK_LAUNCH_CHECK();
}

void build_graph(owtk_gnn model, owtk_matrix m, GnnGraph& g, bool with_mirror)
{
    build_graph_structure(model, m, g, with_mirror);
    graph_values(model, m, g);
}

// Forward pass.  h: (layers+1) * n * W floats when `train`, else 2 * n * W
// (ping-pong); agg (train only): layers * n * W.
void forward(owtk_gnn model, const GnnGraph& g, float* h, float* agg, float* gamma, bool train)
{
    const hipStream_t st = model->ctx->stream;
    const std::int64_t n = g.n;
    const unsigned int grid = rows_grid(n);
    const std::uint16_t* p16 = model->params16.get();
    if (g.sell && train) throw std::logic_error("training runs over the CSR graph");
    // the matrix-core kernels keep each wave's weights in registers across a
    // grid-stride loop over the 32-vertex tiles
    const std::int64_t tiles = (n + 31) / 32;
    const unsigned int mgrid = static_cast<unsigned int>(std::max<std::int64_t>(1, std::min<std::int64_t>((tiles + 3) / 4, 4096)));
    if (g.sell)
        k_gnn_encode_mfma<<<mgrid, kBlock, 0, st>>>(n, model->f0, g.x.get(), p16, model->off_w_in, model->off_b_in, h);
    else
        k_gnn_encode<<<grid, kBlock, 0, st>>>(n, model->f0, g.x.get(), p16, model->off_w_in, model->off_b_in, h);
    OWTK_LAUNCH_CHECK();
    float* cur = h;
    for (int l = 0; l < kGnnLayers; ++l) {
        float* next = train ? h + static_cast<std::int64_t>(l + 1) * n * W : h + ((l + 1) & 1) * n * W;
        if (g.sell) {
            k_gnn_layer_mfma<<<mgrid, kBlock, 0, st>>>(n, g.slice_ptr, g.row_len, g.slot_cols, g.slot_w.get(), cur, p16,
                                                       model->off_ws[l], model->off_wn[l], model->off_b[l], next);
            OWTK_LAUNCH_CHECK();
            cur = next;
            continue;
        }
        k_gnn_layer<<<grid, kBlock, 0, st>>>(n, g.offsets.get(), g.cols.get(), g.w.get(), cur, p16, model->off_ws[l],
                                             model->off_wn[l], model->off_b[l], next,
                                             train ? agg + static_cast<std::int64_t>(l) * n * W : nullptr);
        OWTK_LAUNCH_CHECK();
        cur = next;
    }
    k_gnn_decode<<<grid, kBlock, 0, st>>>(n, g.b, model->outs, cur, p16, model->off_w_out, model->off_b_out,
                                          model->off_g, gamma);
    OWTK_LAUNCH_CHECK();
}

void outer_reduce(owtk_gnn model, std::int64_t n, const float* A, int R, const float* B, int C, float* mat, float* vec)
{
    const hipStream_t st = model->ctx->stream;
    const int blocks = static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(kReduceGrid, (n + kTile - 1) / kTile)));
    k_outer_partials<<<blocks, kBlock, 0, st>>>(n, A, R, B, C, model->partials.get());
    OWTK_LAUNCH_CHECK();
    const int outs = R * C + R;
    k_outer_finish<<<(outs + kBlock - 1) / kBlock, kBlock, 0, st>>>(model->partials.get(), blocks, R, C, mat, vec);
    OWTK_LAUNCH_CHECK();
}

// Workspace layout (floats per vertex): h (layers+1)*W | agg layers*W |
// do outs | dh W | dprev W | du W | v W
struct TrainViews {
    float *h, *agg, *dout, *dh, *dprev, *du, *v;
};

TrainViews ensure_workspace(owtk_gnn model, std::int64_t n, int pairs, int k1, int b)
{
    const std::int64_t per = (kGnnLayers + 1) * W + kGnnLayers * W + model->outs + 4 * W;
    if (n > model->ws_nodes || pairs > model->ws_pairs) {
        const std::int64_t nn = std::max(n, model->ws_nodes);
        const std::int64_t pp = std::max<std::int64_t>(pairs, model->ws_pairs);
        model->act.allocate(nn * per);
        model->gamma.allocate(nn * b * k1);
        model->dgamma.allocate(nn * b * k1);
        model->z.allocate(nn * b * pp);
        model->ys.allocate(nn * b * pp);
        model->loss_partials.allocate(rows_grid(nn) * pp);
        model->coef.allocate(2 * pp);
        model->ws_nodes = nn;
        model->ws_pairs = pp;
    }
    TrainViews v{};
    float* p = model->act.get();
    v.h = p; p += (kGnnLayers + 1) * W * n;
    v.agg = p; p += kGnnLayers * W * n;
    v.dout = p; p += model->outs * n;
    v.dh = p; p += W * n;
    v.dprev = p; p += W * n;
    v.du = p; p += W * n;
    v.v = p;
    return v;
}

void accumulate(owtk_gnn model, owtk_gnn_problem prob, double weight)
{
    const GnnGraph& g = prob->g;
    const hipStream_t st = model->ctx->stream;
    const std::int64_t n = g.n, nb = g.n * g.b;
    const int k1 = model->degree + 1, P = prob->pairs;
    TrainViews v = ensure_workspace(model, n, P, k1, g.b);
    // per-pair coefficients: loss weight, gradient weight
    std::vector<float> coef(2 * P);
    for (int p = 0; p < P; ++p) {
        coef[p] = static_cast<float>(1.0 / prob->norms[p]);
        coef[P + p] = static_cast<float>(2.0 * weight / (P * prob->norms[p]));
    }
    OWTK_HIP(hipMemcpyAsync(model->coef.get(), coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice, st));
    OWTK_HIP(hipStreamSynchronize(st));  // coef is a host temporary

    forward(model, g, v.h, v.agg, model->gamma.get(), true);
    const unsigned int grid = rows_grid(n);
    k_loss_z<<<dev::stream_grid(nb * P, 4), kBlock, 0, st>>>(nb, P, k1, model->gamma.get(), prob->u.get(), model->z.get());
    OWTK_LAUNCH_CHECK();
    k_loss_y<<<dim3(grid, P), kBlock, 0, st>>>(n, g.b, g.offsets.get(), g.cols.get(), g.vals.get(), g.diag.get(),
                                               prob->dis.get(), model->z.get(), prob->target.get(), model->coef.get(),
                                               model->ys.get(), model->loss_partials.get());
    OWTK_LAUNCH_CHECK();
    k_loss_finish<<<1, kBlock, 0, st>>>(model->loss_partials.get(), static_cast<int>(grid) * P, weight / P,
                                        model->loss.get());
    OWTK_LAUNCH_CHECK();
    // z is free again: it holds g_p = gcoef_p A^T ys_p
    k_loss_g<<<dim3(grid, P), kBlock, 0, st>>>(n, g.b, g.offsets.get(), g.cols.get(), g.mirror.get(), g.vals.get(),
                                               g.diag.get(), model->ys.get(), model->coef.get() + P, model->z.get());
    OWTK_LAUNCH_CHECK();
    k_loss_dgamma<<<dev::stream_grid(nb * k1, 4), kBlock, 0, st>>>(nb, P, k1, model->z.get(), prob->u.get(),
                                                                   model->dgamma.get());
    OWTK_LAUNCH_CHECK();

    float* grad = model->grad.get();
    const std::uint16_t* p16 = model->params16.get();
    const float* h3 = v.h + static_cast<std::int64_t>(kGnnLayers) * n * W;
    k_gnn_decode_bwd<<<grid, kBlock, 0, st>>>(n, g.b, model->outs, model->dgamma.get(), p16, model->off_w_out,
                                              v.dout, v.dh);
    OWTK_LAUNCH_CHECK();
    outer_reduce(model, n, v.dout, model->outs, h3, W, grad + model->off_w_out, grad + model->off_b_out);
    outer_reduce(model, n, v.dout, model->outs, nullptr, 0, nullptr, grad + model->off_g);
    float* dh = v.dh;
    float* dprev = v.dprev;
    for (int l = kGnnLayers - 1; l >= 0; --l) {
        const float* hout = v.h + static_cast<std::int64_t>(l + 1) * n * W;
        const float* hin = v.h + static_cast<std::int64_t>(l) * n * W;
        const float* agg = v.agg + static_cast<std::int64_t>(l) * n * W;
        k_gnn_layer_bwd<<<grid, kBlock, 0, st>>>(n, hout, dh, p16, model->off_ws[l], model->off_wn[l], v.du, dprev, v.v);
        OWTK_LAUNCH_CHECK();
        k_gnn_agg_t<<<grid, kBlock, 0, st>>>(n, g.offsets.get(), g.cols.get(), g.mirror.get(), g.w.get(), v.v, dprev);
        OWTK_LAUNCH_CHECK();
        outer_reduce(model, n, v.du, W, hin, W, grad + model->off_ws[l], grad + model->off_b[l]);
        outer_reduce(model, n, v.du, W, agg, W, grad + model->off_wn[l], nullptr);
        std::swap(dh, dprev);
    }
    k_tanh_bwd<<<dev::stream_grid(n * W, 4), kBlock, 0, st>>>(n * W, v.h, dh);
    OWTK_LAUNCH_CHECK();
    outer_reduce(model, n, dh, W, g.x.get(), model->f0, grad + model->off_w_in, grad + model->off_b_in);
}

void refresh_bf16(owtk_gnn g)
{
    k_round_bf16<<<rows_grid(g->count), kBlock, 0, g->ctx->stream>>>(g->count, g->params.get(), g->params16.get());
    OWTK_LAUNCH_CHECK();
}

} // namespace

// ------------------------------------------------ preconditioner hooks --

void learned_setup(owtk_precond pc)
{
    if (!pc->learned) throw std::invalid_argument("learned preconditioner needs a model");
    auto* model = static_cast<owtk_gnn>(pc->learned);
    check_matrix(model, pc->m);
    auto st = std::make_unique<LearnedPc>();
    st->model = model;
    const std::int64_t nb = pc->n();
    const std::size_t k1 = static_cast<std::size_t>(model->degree + 1);
    const std::size_t sv = pc->storage == OWTK_F64 ? 8 : (pc->storage == OWTK_F32 ? 4 : 2);
    st->scaled.allocate(static_cast<std::size_t>(std::max<std::int64_t>(pc->m->value_count, 1)) * sv);
    if (pc->storage == OWTK_F64) {
        st->u64a.allocate(nb);
        st->u64b.allocate(nb);
        st->sgamma64.allocate(nb * k1);
    } else {
        st->u32a.allocate(nb);
        st->u32b.allocate(nb);
        st->sgamma32.allocate(nb * k1);
    }
    st->gamma.allocate(nb * k1);
    build_graph_structure(model, pc->m, st->graph, false);
    st->h.allocate(static_cast<std::size_t>(2 * st->graph.n * W));
    pc->learned_state = st.release();
}

void learned_numeric(owtk_precond pc)
{
    auto* st = static_cast<LearnedPc*>(pc->learned_state);
    owtk_gnn model = st->model;
    owtk_matrix m = pc->m;
    const hipStream_t s = m->ctx->stream;
    GnnGraph& g = st->graph;
    graph_values(model, m, g);
    forward(model, g, st->h.get(), nullptr, st->gamma.get(), false);
    const int sell = m->s.layout == Layout::sell ? 1 : 0;
    const int k1 = model->degree + 1;
    const std::int64_t slots = m->s.slots;
    const std::int64_t n = m->owned;
    const std::int64_t dcount = m->diag_count;
    if (st->dinv.count != static_cast<std::size_t>(std::max<std::int64_t>(dcount, 1)))
        st->dinv.allocate(std::max<std::int64_t>(dcount, 1));
    k_inv_diag<<<dev::stream_grid(dcount, 4), kBlock, 0, s>>>(dcount, m->d64.get(), st->dinv.get());
    OWTK_LAUNCH_CHECK();
    const std::int64_t nb = n * m->bs;
    const unsigned int gv = dev::stream_grid(slots, 4);
    const unsigned int gg = dev::stream_grid(nb, 4);
    const double* dinv = st->dinv.get();
    switch (pc->storage) {
    case OWTK_F64:
        k_scale_columns<double><<<gv, kBlock, 0, s>>>(slots, m->bs, sell, m->s.columns.get(), m->v64.get(), dinv,
                                                      reinterpret_cast<double*>(st->scaled.get()));
        k_scale_gamma<double><<<gg, kBlock, 0, s>>>(nb, m->bs, sell, k1, st->gamma.get(), dinv, st->sgamma64.get());
        break;
    case OWTK_F32:
        k_scale_columns<float><<<gv, kBlock, 0, s>>>(slots, m->bs, sell, m->s.columns.get(), m->v64.get(), dinv,
                                                     reinterpret_cast<float*>(st->scaled.get()));
        k_scale_gamma<float><<<gg, kBlock, 0, s>>>(nb, m->bs, sell, k1, st->gamma.get(), dinv, st->sgamma32.get());
        break;
    default:
        k_scale_columns<std::uint16_t><<<gv, kBlock, 0, s>>>(slots, m->bs, sell, m->s.columns.get(), m->v64.get(),
                                                             dinv, reinterpret_cast<std::uint16_t*>(st->scaled.get()));
        k_scale_gamma<float><<<gg, kBlock, 0, s>>>(nb, m->bs, sell, k1, st->gamma.get(), dinv, st->sgamma32.get());
        break;
    }
    OWTK_LAUNCH_CHECK();
}

// Algorithmic bytes of one series step: off-diagonal values + columns + row
// lengths, w gathered (once), gamma'_k, z read and written, w_next written.
std::int64_t learned_step_bytes(owtk_precond pc)
{
    const std::int64_t nb = pc->n();
    const int sv = pc->storage == OWTK_F64 ? 8 : (pc->storage == OWTK_F32 ? 4 : 2);
    const int sx = pc->storage == OWTK_F64 ? 8 : 4;
    return pc->m->entries * (pc->m->bs * sv + 4) + pc->m->owned * 4 + nb * (5 * sx);
}

// First step (init folded in): matrix; r, gamma'_0, gamma'_1 read; w_1
// (unless it is also the last step) and z written.  Last step: a middle step
// without the w_next write.
std::int64_t learned_end_step_bytes(owtk_precond pc, bool first, bool last)
{
    const std::int64_t nb = pc->n();
    const int sx = pc->storage == OWTK_F64 ? 8 : 4;
    const std::int64_t matrix = learned_step_bytes(pc) - nb * (5 * sx);
    if (first) return matrix + nb * ((last ? 4 : 5) * sx);
    return matrix + nb * (4 * sx);
}

// Apply: degree 0 is the init kernel alone (r, gamma'_0 read, z written);
// otherwise first, middle and last steps.
std::int64_t learned_apply_bytes(owtk_precond pc)
{
    auto* st = static_cast<LearnedPc*>(pc->learned_state);
    const std::int64_t nb = pc->n();
    const int sx = pc->storage == OWTK_F64 ? 8 : 4;
    const std::int64_t K = st ? st->model->degree : 0;
    if (K == 0) return nb * (3 * sx);
    if (K == 1) return learned_end_step_bytes(pc, true, true);
    return learned_end_step_bytes(pc, true, false) + (K - 2) * learned_step_bytes(pc)
         + learned_end_step_bytes(pc, false, true);
}

template<class V, class X>
void learned_apply_t(owtk_precond pc, const X* in, X* out, const int* stop, X* wa, X* wb, const X* sgamma)
{
    auto* st = static_cast<LearnedPc*>(pc->learned_state);
    owtk_matrix m = pc->m;
    const V* vals = reinterpret_cast<const V*>(st->scaled.get());
    const std::int64_t nb = pc->n();
    const std::int64_t step_bytes = learned_step_bytes(pc);
    const int degree = st->model->degree;
    // The fused first step gathers `in` while it writes `out`, so an in-place
    // apply takes the separate initialisation kernel instead.
    const bool fuse = degree > 0 && in != out;
    if (!fuse) {
        k_series_init<X><<<dev::stream_grid(nb, 4), kBlock, 0, m->ctx->stream>>>(nb, in, sgamma, degree > 0 ? wa : nullptr,
                                                                                 out, stop);
        OWTK_LAUNCH_CHECK();
    }
    if (fuse) {
        dev::EpiSeriesFirst<X> e;
        e.stop = stop;
        e.w_next = wb;
        e.r = in;
        e.gamma0 = sgamma;
        e.gamma1 = sgamma + nb;
        e.z = out;
        e.store_w = degree > 1;
        ProfileScope prof(m->ctx, prof_learned_ends, learned_end_step_bytes(pc, true, degree == 1));
        offdiag_spmv<V, X, dev::EpiSeriesFirst<X>>(m, vals, in, e);
        std::swap(wa, wb);
    }
    for (int k = fuse ? 2 : 1; k <= degree; ++k) {
        dev::EpiSeries<X> e;
        e.stop = stop;
        e.w_next = wb;
        e.gamma = sgamma + static_cast<std::int64_t>(k) * nb;
        e.z = out;
        e.store_w = k < degree;
        // middle steps are the series-step kernel proper; the last step is an end step
        ProfileScope prof(m->ctx, k < degree ? prof_learned : prof_learned_ends,
                          k < degree ? step_bytes : learned_end_step_bytes(pc, false, true));
        offdiag_spmv<V, X, dev::EpiSeries<X>>(m, vals, wa, e);
        std::swap(wa, wb);
    }
}

void learned_apply_f64(owtk_precond pc, const double* in, double* out, const int* stop)
{
    auto* st = static_cast<LearnedPc*>(pc->learned_state);
    if (pc->storage != OWTK_F64) throw std::invalid_argument("FP64 apply needs F64 preconditioner storage");
    learned_apply_t<double, double>(pc, in, out, stop, st->u64a.get(), st->u64b.get(), st->sgamma64.get());
}

void learned_apply_f32(owtk_precond pc, const float* in, float* out, const int* stop)
{
    auto* st = static_cast<LearnedPc*>(pc->learned_state);
    if (pc->storage == OWTK_F32)
        learned_apply_t<float, float>(pc, in, out, stop, st->u32a.get(), st->u32b.get(), st->sgamma32.get());
    else if (pc->storage == OWTK_BF16)
        learned_apply_t<std::uint16_t, float>(pc, in, out, stop, st->u32a.get(), st->u32b.get(), st->sgamma32.get());
    else
        throw std::invalid_argument("FP32 apply needs F32 or BF16 preconditioner storage");
}

void learned_release(owtk_precond pc)
{
    delete static_cast<LearnedPc*>(pc->learned_state);
    pc->learned_state = nullptr;
}

} // namespace owtk

// ------------------------------------------------------------------- C ABI --
extern "C" {

int owtk_gnn_create(owtk_context ctx, int32_t block_size, int32_t degree, owtk_gnn* out)
{
    return guarded([&] {
        if (!ctx || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        if (block_size < 1 || block_size > 4) throw std::invalid_argument("block size must be 1..4");
        if (degree < 0 || (degree + 1) * block_size > kGnnMaxOutputs)
            throw std::invalid_argument("(degree + 1) * block size must not exceed 32");
        ctx->activate();
        auto g = std::make_unique<owtk_gnn_s>();
        g->ctx = ctx;
        g->b = block_size;
        g->degree = degree;
        g->f0 = 3 * block_size + 1;
        g->outs = (degree + 1) * block_size;
        std::int64_t pos = 0;
        g->off_w_in = pos; pos += W * g->f0;
        g->off_b_in = pos; pos += W;
        for (int l = 0; l < kGnnLayers; ++l) {
            g->off_ws[l] = pos; pos += W * W;
            g->off_wn[l] = pos; pos += W * W;
            g->off_b[l] = pos; pos += W;
        }
        g->off_w_out = pos; pos += g->outs * W;
        g->off_b_out = pos; pos += g->outs;
        g->off_g = pos; pos += g->outs;
        g->count = pos;
        g->params.allocate(pos);
        g->grad.allocate(pos);
        g->adam_m.allocate(pos);
        g->adam_v.allocate(pos);
        g->params16.allocate(pos);
        g->loss.allocate(1);
        g->partials.allocate(static_cast<std::size_t>(kReduceGrid) * kMaxOuterOutputs);
        // default parameters: Glorot-uniform encoder and layer weights (a
        // fixed SplitMix64 stream), zero biases, zero decoder and g = 1, so a
        // fresh model is the plain Neumann series and every weight receives a
        // gradient from the first training step on
        std::vector<float> init(pos, 0.0f);
        std::uint64_t state = 0x6a09e667f3bcc909ULL;
        auto uniform = [&state](float lim) {
            std::uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            z ^= z >> 31;
            const double u = static_cast<double>(z >> 11) * 0x1.0p-53;  // [0, 1)
            return static_cast<float>((2.0 * u - 1.0) * lim);
        };
        auto glorot = [&](std::int64_t off, int rows, int cols) {
            const float lim = std::sqrt(6.0f / static_cast<float>(rows + cols));
            for (std::int64_t t = 0; t < static_cast<std::int64_t>(rows) * cols; ++t) init[off + t] = uniform(lim);
        };
        glorot(g->off_w_in, W, g->f0);
        for (int l = 0; l < kGnnLayers; ++l) {
            glorot(g->off_ws[l], W, W);
            glorot(g->off_wn[l], W, W);
        }
        for (int o = 0; o < g->outs; ++o) init[g->off_g + o] = 1.0f;
        OWTK_HIP(hipMemcpy(g->params.get(), init.data(), pos * sizeof(float), hipMemcpyHostToDevice));
        OWTK_HIP(hipMemset(g->grad.get(), 0, pos * sizeof(float)));
        OWTK_HIP(hipMemset(g->adam_m.get(), 0, pos * sizeof(float)));
        OWTK_HIP(hipMemset(g->adam_v.get(), 0, pos * sizeof(float)));
        OWTK_HIP(hipMemset(g->loss.get(), 0, sizeof(double)));
        refresh_bf16(g.get());
        ctx->sync();
        *out = g.release();
    });
}

int owtk_gnn_destroy(owtk_gnn g)
{
    return guarded([&] {
        if (!g) return;
        g->ctx->activate();
        g->ctx->sync();
        delete g;
    });
}

int owtk_gnn_info(owtk_gnn g, int32_t* block_size, int32_t* degree, int64_t* parameter_count)
{
    return guarded([&] {
        if (!g) throw std::invalid_argument("null model");
        if (block_size) *block_size = g->b;
        if (degree) *degree = g->degree;
        if (parameter_count) *parameter_count = g->count;
    });
}

int owtk_gnn_set_parameters(owtk_gnn g, const float* params)
{
    return guarded([&] {
        if (!g || !params) throw std::invalid_argument("null argument");
        g->ctx->activate();
        const std::size_t bytes = g->count * sizeof(float);
        OWTK_HIP(hipMemcpy(g->params.get(), params, bytes, hipMemcpyHostToDevice));
        OWTK_HIP(hipMemset(g->adam_m.get(), 0, bytes));
        OWTK_HIP(hipMemset(g->adam_v.get(), 0, bytes));
        OWTK_HIP(hipMemset(g->grad.get(), 0, bytes));
        g->adam_steps = 0;
        refresh_bf16(g);
        g->ctx->sync();
    });
}

int owtk_gnn_get_parameters(owtk_gnn g, float* params)
{
    return guarded([&] {
        if (!g || !params) throw std::invalid_argument("null argument");
        g->ctx->activate();
        g->ctx->sync();
        OWTK_HIP(hipMemcpy(params, g->params.get(), g->count * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int owtk_gnn_coefficients(owtk_gnn model, owtk_matrix m, float* gamma)
{
    return guarded([&] {
        if (!model || !gamma) throw std::invalid_argument("null argument");
        model->ctx->activate();
        GnnGraph g;
        build_graph(model, m, g, false);
        const std::int64_t nb = g.n * g.b;
        DeviceArray<float> h(2 * g.n * W);
        DeviceArray<float> gam(nb * (model->degree + 1));
        forward(model, g, h.get(), nullptr, gam.get(), false);
        model->ctx->sync();
        OWTK_HIP(hipMemcpy(gamma, gam.get(), gam.bytes(), hipMemcpyDeviceToHost));
    });
}

int owtk_gnn_problem_create(owtk_gnn model, owtk_matrix m, int32_t pairs, const double* residuals,
                            const double* errors, owtk_gnn_problem* out)
{
    return guarded([&] {
        if (!model || !out || !residuals || !errors) throw std::invalid_argument("null argument");
        if (pairs < 1) throw std::invalid_argument("a training problem needs at least one pair");
        *out = nullptr;
        model->ctx->activate();
        auto p = std::make_unique<owtk_gnn_problem_s>();
        p->ctx = model->ctx;
        p->b = model->b;
        p->degree = model->degree;
        p->pairs = pairs;
        build_graph(model, m, p->g, true);
        const GnnGraph& g = p->g;
        const hipStream_t st = model->ctx->stream;
        const std::int64_t n = g.n, nb = g.n * g.b;
        const int k1 = model->degree + 1;
        p->u.allocate(static_cast<std::int64_t>(pairs) * k1 * nb);
        p->target.allocate(static_cast<std::int64_t>(pairs) * nb);
        p->dis.allocate(nb);
        const unsigned int grid = rows_grid(n);
        for (int q = 0; q < pairs; ++q) {
            float* u = p->u.get() + static_cast<std::int64_t>(q) * k1 * nb;
            k_basis_init<<<dev::stream_grid(nb, 4), kBlock, 0, st>>>(nb, residuals + q * nb, g.diag.get(), u);
            OWTK_LAUNCH_CHECK();
            for (int k = 1; k < k1; ++k) {
                k_basis_step<<<grid, kBlock, 0, st>>>(n, g.b, g.offsets.get(), g.cols.get(), g.vals.get(), g.diag.get(),
                                                      u + (k - 1) * nb, u + k * nb);
                OWTK_LAUNCH_CHECK();
            }
            k_target<<<grid, kBlock, 0, st>>>(n, g.b, g.offsets.get(), g.cols.get(), g.vals.get(), g.diag.get(),
                                              errors + q * nb, p->target.get() + q * nb, p->dis.get());
            OWTK_LAUNCH_CHECK();
        }
        std::vector<float> tgt(static_cast<std::size_t>(pairs) * nb);
        OWTK_HIP(hipStreamSynchronize(st));
        OWTK_HIP(hipMemcpy(tgt.data(), p->target.get(), tgt.size() * sizeof(float), hipMemcpyDeviceToHost));
        p->norms.resize(pairs);
        for (int q = 0; q < pairs; ++q) {
            double s = 0.0;
            for (std::int64_t i = 0; i < nb; ++i) s += static_cast<double>(tgt[q * nb + i]) * tgt[q * nb + i];
            if (!(s > 0.0) || !std::isfinite(s)) throw std::invalid_argument("training pair with zero or non-finite error");
            p->norms[q] = s;
        }
        *out = p.release();
    });
}

int owtk_gnn_problem_destroy(owtk_gnn_problem p)
{
    return guarded([&] {
        if (!p) return;
        p->ctx->activate();
        p->ctx->sync();
        delete p;
    });
}

int owtk_gnn_accumulate(owtk_gnn g, owtk_gnn_problem p, double weight)
{
    return guarded([&] {
        if (!g || !p) throw std::invalid_argument("null argument");
        if (p->ctx != g->ctx || p->b != g->b || p->degree != g->degree)
            throw std::invalid_argument("problem was built for another context, block size or degree");
        g->ctx->activate();
        accumulate(g, p, weight);
    });
}

int owtk_gnn_read(owtk_gnn g, double* loss, float* grad)
{
    return guarded([&] {
        if (!g) throw std::invalid_argument("null model");
        g->ctx->activate();
        g->ctx->sync();
        if (loss) OWTK_HIP(hipMemcpy(loss, g->loss.get(), sizeof(double), hipMemcpyDeviceToHost));
        if (grad) OWTK_HIP(hipMemcpy(grad, g->grad.get(), g->count * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int owtk_gnn_zero_grad(owtk_gnn g)
{
    return guarded([&] {
        if (!g) throw std::invalid_argument("null model");
        g->ctx->activate();
        OWTK_HIP(hipMemsetAsync(g->grad.get(), 0, g->count * sizeof(float), g->ctx->stream));
        OWTK_HIP(hipMemsetAsync(g->loss.get(), 0, sizeof(double), g->ctx->stream));
    });
}

int owtk_gnn_adam_step(owtk_gnn g, double lr, double beta1, double beta2, double eps, double* loss)
{
    return guarded([&] {
        if (!g) throw std::invalid_argument("null model");
        if (!(lr > 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps > 0.0))
            throw std::invalid_argument("invalid Adam hyper-parameters");
        g->ctx->activate();
        ++g->adam_steps;
        const double bc1 = 1.0 - std::pow(beta1, static_cast<double>(g->adam_steps));
        const double bc2 = 1.0 - std::pow(beta2, static_cast<double>(g->adam_steps));
        k_adam<<<rows_grid(g->count), kBlock, 0, g->ctx->stream>>>(
            g->count, g->params.get(), g->grad.get(), g->adam_m.get(), g->adam_v.get(), g->params16.get(),
            static_cast<float>(lr), static_cast<float>(beta1), static_cast<float>(beta2), static_cast<float>(eps),
            static_cast<float>(bc1), static_cast<float>(bc2));
        OWTK_LAUNCH_CHECK();
        if (loss) {
            g->ctx->sync();
            OWTK_HIP(hipMemcpy(loss, g->loss.get(), sizeof(double), hipMemcpyDeviceToHost));
        }
        OWTK_HIP(hipMemsetAsync(g->loss.get(), 0, sizeof(double), g->ctx->stream));
    });
}

} // extern "C"
