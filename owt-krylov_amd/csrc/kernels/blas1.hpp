// Streaming vector kernels: fills, conversions, dot/norm with FP64
// accumulation, axpy, axpy+dot and the batched multi-dot of classical
// Gram-Schmidt.
#pragma once

#include "../device.hpp"
#include "../internal.hpp"

namespace owtk::kern {
inline namespace {  // kernels get internal linkage in every including TU

using namespace owtk::dev;

constexpr int kMaxMultidot = 62;
constexpr std::int64_t kMultidotMaxBlocks = 1024;  // grid cap of the multi-dot

__global__ __launch_bounds__(256) void k_fill_zero16(uint4* p, std::int64_t n)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = make_uint4(0, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_fill_zero1(unsigned char* p, std::int64_t n)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 0;
}

inline void fill_zero_bytes(hipStream_t st, void* p, std::size_t bytes)
{
    if (bytes == 0) return;
    const auto addr = reinterpret_cast<std::uintptr_t>(p);
    if (addr % 16 == 0 && bytes % 16 == 0) {
        const std::int64_t n = static_cast<std::int64_t>(bytes / 16);
        k_fill_zero16<<<stream_grid(n, 4), kBlock, 0, st>>>(static_cast<uint4*>(p), n);
    } else {
        const std::int64_t n = static_cast<std::int64_t>(bytes);
        k_fill_zero1<<<stream_grid(n, 4), kBlock, 0, st>>>(static_cast<unsigned char*>(p), n);
    }
    OWTK_LAUNCH_CHECK();
}

template<class D, class S>
__global__ __launch_bounds__(256) void k_convert(D* dst, const S* src, std::int64_t n)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = static_cast<D>(src[i]);
}

template<class D, class S>
inline void convert(hipStream_t st, D* dst, const S* src, std::size_t n)
{
    k_convert<D, S><<<stream_grid(static_cast<std::int64_t>(n), 4), kBlock, 0, st>>>(dst, src, static_cast<std::int64_t>(n));
    OWTK_LAUNCH_CHECK();
}

// FP64 -> BF16 (round to nearest even) and FP32 copies of a value array.
__global__ __launch_bounds__(256) void k_to_bf16(std::uint16_t* dst, const double* src, std::int64_t n)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = float_to_bf16(static_cast<float>(src[i]));
}

// Dot product with FP64 accumulation; CMP: compensated (double-double).
template<class X, bool CMP = false>
__global__ __launch_bounds__(256) void k_dot(const X* __restrict__ a, const X* __restrict__ b, std::int64_t n,
                                             ReduceScratch scratch, double* out)
{
    double red[1] = {0.0}, lo[1] = {0.0};
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        if constexpr (CMP) dd_fma(red[0], lo[0], static_cast<double>(a[i]), static_cast<double>(b[i]));
        else red[0] = __builtin_fma(static_cast<double>(a[i]), static_cast<double>(b[i]), red[0]);
    }
    if constexpr (CMP) store_partials_dd<1>(red, lo, 0, scratch);
    else store_partials<1>(red, 0, scratch);
}

struct FinStore {
    double* out;
    int count;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x < count) out[threadIdx.x] = t[threadIdx.x];
    }
};

template<class X>
inline void dot_to(hipStream_t st, ReduceScratch scratch, const X* a, const X* b, std::int64_t n, double* out,
                   bool compensated = false)
{
    const unsigned int g = stream_grid(n, 4);
    if (compensated) {
        k_dot<X, true><<<g, kBlock, 0, st>>>(a, b, n, scratch, out);
        launch_finish_dd<1>(st, FinStore{out, 1}, g, 0, scratch, nullptr);
    } else {
        k_dot<X, false><<<g, kBlock, 0, st>>>(a, b, n, scratch, out);
        launch_finish<1>(st, FinStore{out, 1}, g, 0, scratch, nullptr);
    }
    OWTK_LAUNCH_CHECK();
}

template<class X, bool CMP = false>
__global__ __launch_bounds__(256) void k_norm(const X* __restrict__ a, std::int64_t n, ReduceScratch scratch,
                                              double* out)
{
    double red[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = static_cast<double>(a[i]);
        if constexpr (CMP) dd_fma(red[0], lo[0], v, v);
        else red[0] = __builtin_fma(v, v, red[0]);
        red[1] = combine<Combine::max>(red[1], __builtin_fabs(v));
    }
    if constexpr (CMP) store_partials_dd<2>(red, lo, 2, scratch);
    else store_partials<2>(red, 2, scratch);
}

template<class X>
struct FinNorm {
    const X* a;
    std::int64_t n;
    double* out;
    __device__ void operator()(const double* t) const
    {
        const double nrm = block_norm(t[0], t[1], a, n);
        if (threadIdx.x == 0) *out = nrm;
    }
};

template<class X>
inline void norm_to(hipStream_t st, ReduceScratch scratch, const X* a, std::int64_t n, double* out,
                    bool compensated = false)
{
    const unsigned int g = stream_grid(n, 4);
    if (compensated) {
        k_norm<X, true><<<g, kBlock, 0, st>>>(a, n, scratch, out);
        launch_finish_dd<2>(st, FinNorm<X>{a, n, out}, g, 2, scratch, nullptr);
    } else {
        k_norm<X, false><<<g, kBlock, 0, st>>>(a, n, scratch, out);
        launch_finish<2>(st, FinNorm<X>{a, n, out}, g, 2, scratch, nullptr);
    }
    OWTK_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_axpy(std::int64_t n, double alpha, const double* __restrict__ x,
                                              double* __restrict__ y)
{
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        y[i] = __builtin_fma(alpha, x[i], y[i]);
}

// 16-byte form: each thread moves 4 pairs per trip with every load issued
// before the first store (full trips run without bounds tests; the tail
// takes one pair per thread).  The per-element arithmetic is k_axpy's.
typedef double double2_v __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_axpy2(std::int64_t n2, double alpha, const double2_v* __restrict__ x,
                                               double2_v* __restrict__ y)
{
    constexpr int U = 4;
    const std::int64_t T = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * T < n2; i += U * T) {
        double2_v xv[U], yv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xv[u] = __builtin_nontemporal_load(x + i + u * T);
            yv[u] = y[i + u * T];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double2_v r;
            r.x = __builtin_fma(alpha, xv[u].x, yv[u].x);
            r.y = __builtin_fma(alpha, xv[u].y, yv[u].y);
            y[i + u * T] = r;
        }
    }
    for (; i < n2; i += T) {
        const double2_v xv = x[i];
        double2_v r = y[i];
        r.x = __builtin_fma(alpha, xv.x, r.x);
        r.y = __builtin_fma(alpha, xv.y, r.y);
        y[i] = r;
    }
}

inline void axpy(hipStream_t st, std::int64_t n, double alpha, const double* x, double* y)
{
    const bool vec = (reinterpret_cast<std::uintptr_t>(x) | reinterpret_cast<std::uintptr_t>(y)) % 16 == 0;
    if (vec && n >= 2) {
        const std::int64_t n2 = n / 2;
        k_axpy2<<<stream_grid(n2, 8), kBlock, 0, st>>>(n2, alpha, reinterpret_cast<const double2_v*>(x),
                                                       reinterpret_cast<double2_v*>(y));
        OWTK_LAUNCH_CHECK();
        if (n % 2) {
            k_axpy<<<1, kBlock, 0, st>>>(1, alpha, x + n - 1, y + n - 1);
            OWTK_LAUNCH_CHECK();
        }
        return;
    }
    k_axpy<<<stream_grid(n, 4), kBlock, 0, st>>>(n, alpha, x, y);
    OWTK_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_axpy_dot(std::int64_t n, double alpha, const double* __restrict__ x,
                                                  double* __restrict__ y, const double* __restrict__ w,
                                                  ReduceScratch scratch, double* out)
{
    double red[1] = {0.0};
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = __builtin_fma(alpha, x[i], y[i]);
        y[i] = v;
        red[0] = __builtin_fma(w[i], v, red[0]);
    }
    store_partials<1>(red, 0, scratch);
}

inline void axpy_dot_to(hipStream_t st, ReduceScratch scratch, std::int64_t n, double alpha, const double* x,
                        double* y, const double* w, double* out)
{
    const unsigned int g = stream_grid(n, 4);
    k_axpy_dot<<<g, kBlock, 0, st>>>(n, alpha, x, y, w, scratch, out);
    launch_finish<1>(st, FinStore{out, 1}, g, 0, scratch, nullptr);
    OWTK_LAUNCH_CHECK();
}

struct BasisPointers {
    const double* v[kMaxMultidot];
};

// out[j] = <w, v_j> (j < k), out[k] = <w, w>; K = reduced-value bucket >= k+1.
template<int K>
__global__ __launch_bounds__(256) void k_multidot(std::int64_t n, int k, BasisPointers basis,
                                                  const double* __restrict__ w, ReduceScratch scratch,
                                                  double* out)
{
    double red[K];
#pragma unroll
    for (int j = 0; j < K; ++j) red[j] = 0.0;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j < K - 1; ++j)
            if (j < k) red[j] = __builtin_fma(wi, basis.v[j][i], red[j]);
        red[K - 1] = __builtin_fma(wi, wi, red[K - 1]);
    }
    store_partials<K>(red, 0, scratch);
}

template<int K>
struct FinMulti {
    double* out;
    int k;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x == 0) {
            for (int j = 0; j < k; ++j) out[j] = t[j];
            out[k] = t[K - 1];
        }
    }
};

inline void multidot_to(hipStream_t st, ReduceScratch scratch, std::int64_t n, int k, const double* const* vs,
                        const double* w, double* out)
{
    BasisPointers bp{};
    for (int j = 0; j < k; ++j) bp.v[j] = vs[j];
    std::int64_t blocks = (n + 4 * kBlock - 1) / (4 * kBlock);
    if (blocks > kMultidotMaxBlocks) blocks = kMultidotMaxBlocks;
    if (blocks < 1) blocks = 1;
    const unsigned int g = static_cast<unsigned int>(blocks);
    if (k + 1 <= 4) {
        k_multidot<4><<<g, kBlock, 0, st>>>(n, k, bp, w, scratch, out);
        launch_finish<4>(st, FinMulti<4>{out, k}, g, 0, scratch, nullptr);
    } else if (k + 1 <= 8) {
        k_multidot<8><<<g, kBlock, 0, st>>>(n, k, bp, w, scratch, out);
        launch_finish<8>(st, FinMulti<8>{out, k}, g, 0, scratch, nullptr);
    } else if (k + 1 <= 16) {
        k_multidot<16><<<g, kBlock, 0, st>>>(n, k, bp, w, scratch, out);
        launch_finish<16>(st, FinMulti<16>{out, k}, g, 0, scratch, nullptr);
    } else if (k + 1 <= 32) {
        k_multidot<32><<<g, kBlock, 0, st>>>(n, k, bp, w, scratch, out);
        launch_finish<32>(st, FinMulti<32>{out, k}, g, 0, scratch, nullptr);
    } else {
        k_multidot<63><<<g, kBlock, 0, st>>>(n, k, bp, w, scratch, out);
        launch_finish<63>(st, FinMulti<63>{out, k}, g, 0, scratch, nullptr);
    }
    OWTK_LAUNCH_CHECK();
}

} // namespace
} // namespace owtk::kern
