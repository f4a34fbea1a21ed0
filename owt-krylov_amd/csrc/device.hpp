// Device utilities for gfx950 (wave64): block reductions with a deterministic
// "last block finishes" step, number formats, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace owtk::dev {

constexpr int kWave = 64;
constexpr int kBlock = 256;          // threads per block for streaming kernels
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxReduceValues = 64; // values one reduction kernel may carry
constexpr int kMaxReduceBlocks = 16384;

// ------------------------------------------------------------ number types --
struct bf16 {
    std::uint16_t bits;
};

__host__ __device__ inline float bf16_to_float(std::uint16_t b)
{
    union {
        std::uint32_t u;
        float f;
    } v;
    v.u = static_cast<std::uint32_t>(b) << 16;
    return v.f;
}

// Round to nearest even; NaN stays NaN.
__host__ __device__ inline std::uint16_t float_to_bf16(float f)
{
    union {
        std::uint32_t u;
        float f;
    } v;
    v.f = f;
    if ((v.u & 0x7fffffffU) > 0x7f800000U) return static_cast<std::uint16_t>((v.u >> 16) | 0x40U);
    const std::uint32_t rounding = 0x7fffU + ((v.u >> 16) & 1U);
    return static_cast<std::uint16_t>((v.u + rounding) >> 16);
}

// Dense B x B blocks in the SELL-64 layout, elements in lane-interleaved
// pairs: elements 2g and 2g+1 (j = r*B + c) of the 64 slots of a wave sit in
// one contiguous run of 128 values, each lane's pair adjacent, so one load
// of a wave reads both elements of a pair (16 bytes a lane in FP64).  With
// B*B odd the last element is a run of 64 single values.  Element j of slot
// s is at (s & ~63) * B*B + dense_offset(s & 63, j, B*B).
__host__ __device__ constexpr std::int64_t dense_offset(std::int64_t lane, int j, int bb)
{
    return j < (bb & ~1) ? 128 * (j >> 1) + 2 * lane + (j & 1) : 64 * j + lane;
}
__host__ __device__ constexpr std::int64_t dense_index(std::int64_t slot, int j, int bb)
{
    return (slot & ~std::int64_t(63)) * bb + dense_offset(slot & 63, j, bb);
}
// The inverse within one wave's run of 64 * bb values: offset o holds
// element j of lane `lane`.
__host__ __device__ constexpr void dense_element(std::int64_t o, int bb, int& lane, int& j)
{
    if (o < 64 * (bb & ~1)) {
        lane = static_cast<int>((o & 127) >> 1);
        j = static_cast<int>(2 * (o >> 7) + (o & 1));
    } else {
        lane = static_cast<int>(o & 63);
        j = bb - 1;
    }
}

template<class V> __device__ inline float load_as_float(const V* p, std::int64_t i);
template<> __device__ inline float load_as_float<float>(const float* p, std::int64_t i) { return p[i]; }
template<> __device__ inline float load_as_float<std::uint16_t>(const std::uint16_t* p, std::int64_t i)
{
    return bf16_to_float(p[i]);
}

__device__ inline double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ------------------------------------------------------------- reductions --
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, kWave);
    return v;
}

__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v = fmax(v, __shfl_xor(v, offset, kWave));
    return v;
}

// How a reduced value is combined: sum, or max of magnitudes (NaN-aware: a NaN
// anywhere yields NaN).
enum class Combine { sum, max };

template<Combine C>
__device__ inline double combine(double a, double b)
{
    if constexpr (C == Combine::sum) {
        return a + b;
    } else {
        if (a != a || b != b) return __builtin_nan("");
        return a > b ? a : b;
    }
}

template<Combine C>
__device__ inline double wave_combine(double v)
{
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v = combine<C>(v, __shfl_xor(v, offset, kWave));
    return v;
}

// Reduce v[0..K) over the block; thread 0 gets the block totals in v.
// Bit k of max_mask set: value k is a max of magnitudes, else a sum.
template<int K>
__device__ inline void block_reduce(double (&v)[K], std::uint64_t max_mask)
{
    __shared__ double lds[kWavesPerBlock][K > 0 ? K : 1];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        v[k] = ((max_mask >> k) & 1) ? wave_combine<Combine::max>(v[k])
                                     : wave_combine<Combine::sum>(v[k]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double acc = lds[0][k];
            for (int w = 1; w < kWavesPerBlock; ++w)
                acc = ((max_mask >> k) & 1) ? combine<Combine::max>(acc, lds[w][k])
                                            : combine<Combine::sum>(acc, lds[w][k]);
            v[k] = acc;
        }
    }
    __syncthreads();
}

// Partials of cross-block reductions: value k of block b lives at
// partials[k * stride + b], so a reduction over G blocks needs stride >= G.
// A context's buffer holds kMaxReduceValues rows of `stride` blocks and grows
// (at matrix creation) for kernels that launch more blocks than that.
struct ReduceScratch {
    double* partials;
    std::int64_t stride;
};

// Cross-block reductions run in two launches.  The streaming kernel calls
// store_partials: each block reduces its values and writes one partial per
// value.  A one-block finisher (k_finish, launched right after on the same
// stream) then sums the partials in a fixed order, so results are
// deterministic for a fixed grid, and runs the caller's finalizer with all of
// its threads (totals in LDS, identical on every thread); single-thread state
// updates inside a finalizer are guarded by threadIdx.x == 0.  The kernel
// boundary makes the partials visible across XCDs without any fence or atomic
// inside the streaming kernel.
template<int K>
__device__ inline void store_partials(double (&v)[K], std::uint64_t max_mask, ReduceScratch scratch)
{
    block_reduce<K>(v, max_mask);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) scratch.partials[k * scratch.stride + blockIdx.x] = v[k];
    }
}

template<int K, class Fin>
__global__ __launch_bounds__(256) void k_finish(Fin fin, unsigned int nblocks, std::uint64_t max_mask,
                                                ReduceScratch scratch, const int* stop, const int* need)
{
    if (stop != nullptr && *stop != 0) return;
    if (need != nullptr && *need == 0) return;
    __shared__ double totals[K > 0 ? K : 1];
    // The partials were written by other XCDs, so each load waits a full
    // memory latency.  A thread issues all loads of a chunk (U blocks x K
    // values) before its first combine; the combine order per thread stays
    // b = t, t + 256, ..., so totals do not depend on U.
    constexpr int U = K >= 32 ? 1 : (K >= 16 ? 2 : 4);
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (unsigned int base = threadIdx.x; base < nblocks; base += U * blockDim.x) {
        double p[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned int b = base + u * blockDim.x;
#pragma unroll
            for (int k = 0; k < K; ++k) p[u][k] = b < nblocks ? scratch.partials[k * scratch.stride + b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + u * blockDim.x >= nblocks) break;
#pragma unroll
            for (int k = 0; k < K; ++k)
                acc[k] = ((max_mask >> k) & 1) ? combine<Combine::max>(acc[k], p[u][k])
                                               : combine<Combine::sum>(acc[k], p[u][k]);
        }
    }
    block_reduce<K>(acc, max_mask);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) totals[k] = acc[k];
    }
    __syncthreads();
    fin(static_cast<const double*>(totals));
}

// stop: skip when *stop != 0; need (optional): skip when *need == 0.
template<int K, class Fin>
inline void launch_finish(hipStream_t st, const Fin& fin, unsigned int nblocks, std::uint64_t max_mask,
                          ReduceScratch scratch, const int* stop, const int* need = nullptr)
{
    k_finish<K, Fin><<<1, kBlock, 0, st>>>(fin, nblocks, max_mask, scratch, stop, need);
}

// ------------------------------------------- double-double (compensated) --
// Compensated reductions carry each sum as an unevaluated pair hi + lo.
// TwoSum (Knuth) gives s + e == a + b exactly; TwoProd by FMA gives the
// product's rounding error.  The kernels are built with -ffp-contract=off,
// so none of these operations is fused or reassociated.
__device__ inline void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// (hi, lo) += a * b
__device__ inline void dd_fma(double& hi, double& lo, double a, double b)
{
    const double p = a * b;
    const double pe = __builtin_fma(a, b, -p);
    double s, e;
    two_sum(hi, p, s, e);
    hi = s;
    lo += e + pe;
}

// (h1, l1) += (h2, l2)
__device__ inline void dd_combine(double& h1, double& l1, double h2, double l2)
{
    double s, e;
    two_sum(h1, h2, s, e);
    h1 = s;
    l1 = (l1 + l2) + e;
}

// Block reduction of K pairs (hi[k], lo[k]); bit k of max_mask: value k is a
// max of magnitudes (carried in hi, lo stays 0).  Thread 0 gets the totals.
template<int K>
__device__ inline void block_reduce_dd(double (&hi)[K], double (&lo)[K], std::uint64_t max_mask)
{
    __shared__ double lds[kWavesPerBlock][2 * (K > 0 ? K : 1)];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int offset = kWave / 2; offset > 0; offset >>= 1) {
            const double h2 = __shfl_xor(hi[k], offset, kWave);
            const double l2 = __shfl_xor(lo[k], offset, kWave);
            if ((max_mask >> k) & 1) hi[k] = combine<Combine::max>(hi[k], h2);
            else dd_combine(hi[k], lo[k], h2, l2);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            lds[wave][k] = hi[k];
            lds[wave][K + k] = lo[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double h = lds[0][k], l = lds[0][K + k];
            for (int w = 1; w < kWavesPerBlock; ++w) {
                if ((max_mask >> k) & 1) h = combine<Combine::max>(h, lds[w][k]);
                else dd_combine(h, l, lds[w][k], lds[w][K + k]);
            }
            hi[k] = h;
            lo[k] = l;
        }
    }
    __syncthreads();
}

// Compensated partials: value k of block b as hi at row k, lo at row K + k
// (2K <= kMaxReduceValues rows).
template<int K>
__device__ inline void store_partials_dd(double (&hi)[K], double (&lo)[K], std::uint64_t max_mask,
                                         ReduceScratch scratch)
{
    static_assert(2 * K <= kMaxReduceValues, "too many compensated values");
    block_reduce_dd<K>(hi, lo, max_mask);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            scratch.partials[k * scratch.stride + blockIdx.x] = hi[k];
            scratch.partials[(K + k) * scratch.stride + blockIdx.x] = lo[k];
        }
    }
}

// Finisher of compensated partials: pairs combined in a fixed order (block
// b = t, t + 256, ... per thread, then the block tree); the finalizer gets
// hi + lo rounded once.
template<int K, class Fin>
__global__ __launch_bounds__(256) void k_finish_dd(Fin fin, unsigned int nblocks, std::uint64_t max_mask,
                                                   ReduceScratch scratch, const int* stop)
{
    if (stop != nullptr && *stop != 0) return;
    __shared__ double totals[K > 0 ? K : 1];
    double hi[K], lo[K];
#pragma unroll
    for (int k = 0; k < K; ++k) hi[k] = lo[k] = 0.0;
    for (unsigned int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        double ph[K], pl[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ph[k] = scratch.partials[k * scratch.stride + b];
            pl[k] = scratch.partials[(K + k) * scratch.stride + b];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if ((max_mask >> k) & 1) hi[k] = combine<Combine::max>(hi[k], ph[k]);
            else dd_combine(hi[k], lo[k], ph[k], pl[k]);
        }
    }
    block_reduce_dd<K>(hi, lo, max_mask);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) totals[k] = ((max_mask >> k) & 1) ? hi[k] : hi[k] + lo[k];
    }
    __syncthreads();
    fin(static_cast<const double*>(totals));
}

template<int K, class Fin>
inline void launch_finish_dd(hipStream_t st, const Fin& fin, unsigned int nblocks, std::uint64_t max_mask,
                             ReduceScratch scratch, const int* stop)
{
    k_finish_dd<K, Fin><<<1, kBlock, 0, st>>>(fin, nblocks, max_mask, scratch, stop);
}

// Sum over the block, result on every thread.
__device__ inline double block_sum_all(double v)
{
    __shared__ double lds[kWavesPerBlock];
    __shared__ double total;
    v = wave_combine<Combine::sum>(v);
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = lds[0];
        for (int w = 1; w < kWavesPerBlock; ++w) acc += lds[w];
        total = acc;
    }
    __syncthreads();
    return total;
}

// 2-norm from a fused (sum of squares, max magnitude) pair, overflow and
// underflow safe with the semantics of detail::scaled_norm
// (reduction.hpp:64-89): NaN anywhere -> NaN, otherwise an infinite entry ->
// inf, all zero -> 0.  When the plain sum of squares over- or underflowed,
// the calling block re-reads v once and sums (v_i / max)^2.  Call with all
// threads of one block; every thread gets the result.
constexpr double kSafeSumOfSquares = 0x1.0p-960;

template<class X>
__device__ inline double block_norm(double sum_squares, double max_abs, const X* v, std::int64_t n)
{
    if (max_abs != max_abs) return __builtin_nan("");
    if (max_abs == 0.0) return 0.0;
    if (__builtin_isinf(max_abs)) return max_abs;
    if (__builtin_isfinite(sum_squares) && sum_squares >= kSafeSumOfSquares) return __builtin_sqrt(sum_squares);
    double acc = 0.0;
    for (std::int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double r = static_cast<double>(v[i]) / max_abs;
        acc = __builtin_fma(r, r, acc);
    }
    return max_abs * __builtin_sqrt(block_sum_all(acc));
}

// Grid for streaming kernels over n items with `per_thread` items per thread.
inline unsigned int stream_grid(std::int64_t n, int per_thread = 1)
{
    std::int64_t blocks = (n + static_cast<std::int64_t>(kBlock) * per_thread - 1)
        / (static_cast<std::int64_t>(kBlock) * per_thread);
    if (blocks < 1) blocks = 1;
    if (blocks > kMaxReduceBlocks) blocks = kMaxReduceBlocks;
    return static_cast<unsigned int>(blocks);
}

} // namespace owtk::dev
