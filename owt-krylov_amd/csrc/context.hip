// Context, device memory helpers and the standalone vector kernels of the C ABI
// (dot, norm, axpy, axpy+dot, batched multi-dot, conversions).
#include "internal.hpp"
#include "kernels/blas1.hpp"

#include <algorithm>
#include <string>

namespace owtk {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const char* message) { g_last_error = message ? message : ""; }

void throw_hip(hipError_t err, const char* what)
{
    std::string msg = std::string(what) + ": " + hipGetErrorString(err);
    if (err == hipErrorOutOfMemory) throw std::bad_alloc();
    if (err == hipErrorNoDevice || err == hipErrorInvalidDevice) throw no_device_error(msg);
    throw hip_error(msg);
}

void PinnedBuffer::ensure(std::size_t n)
{
    if (n <= bytes) return;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    OWTK_HIP(hipHostMalloc(&ptr, n, hipHostMallocDefault));
    bytes = n;
}

PinnedBuffer::~PinnedBuffer()
{
    if (ptr) (void)hipHostFree(ptr);
}

namespace {

void check_ctx(owtk_context ctx)
{
    if (!ctx) throw std::invalid_argument("null context");
}

template<class T>
const T* as(const void* p)
{
    return static_cast<const T*>(p);
}

} // namespace
} // namespace owtk

void owtk_context_s::activate() const { OWTK_HIP(hipSetDevice(device)); }

void owtk_context_s::sync() const { OWTK_HIP(hipStreamSynchronize(stream)); }

void owtk_context_s::reserve_partials(std::int64_t blocks)
{
    if (blocks <= partial_stride) return;
    if (partials.get() != nullptr) sync();
    partials.allocate(static_cast<std::size_t>(owtk::dev::kMaxReduceValues) * static_cast<std::size_t>(blocks));
    partial_stride = blocks;
}

hipEvent_t owtk_context_s::take_event()
{
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    OWTK_HIP(hipEventCreate(&e));
    return e;
}

namespace owtk {

ProfileScope::ProfileScope(owtk_context ctx, int category, std::int64_t bytes) : ctx_(ctx)
{
    if (!ctx_ || !ctx_->profiling) return;
    ProfileRecord r{category, ctx_->take_event(), ctx_->take_event(), bytes};
    OWTK_HIP(hipEventRecord(r.start, ctx_->stream));
    ctx_->records.push_back(r);
    index_ = ctx_->records.size() - 1;
}

ProfileScope::~ProfileScope()
{
    if (index_ != static_cast<std::size_t>(-1)) (void)hipEventRecord(ctx_->records[index_].stop, ctx_->stream);
}

} // namespace owtk

using namespace owtk;

extern "C" {

const char* owtk_last_error(void) { return g_last_error.c_str(); }

const char* owtk_version(void) { return "owtk-hip 0.1 (gfx950)"; }

const char* owtk_policy_name(void) { return "hip-gfx950"; }

int owtk_device_count(int* count)
{
    return guarded([&] {
        if (!count) throw std::invalid_argument("null count");
        int n = 0;
        const hipError_t err = hipGetDeviceCount(&n);
        *count = (err == hipSuccess) ? n : 0;
    });
}

int owtk_context_create(int device, owtk_context* out)
{
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw no_device_error("no HIP device");
        if (device < 0 || device >= n) throw std::invalid_argument("device index out of range");
        auto ctx = std::make_unique<owtk_context_s>();
        ctx->device = device;
        ctx->activate();
        hipDeviceProp_t prop;
        OWTK_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        OWTK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->reserve_partials(dev::kMaxReduceBlocks);
        ctx->scalars.allocate(1024);
        kern::fill_zero_bytes(ctx->stream, ctx->scalars.get(), ctx->scalars.bytes());
        ctx->pinned.ensure(1 << 16);
        ctx->sync();
        *out = ctx.release();
    });
}

int owtk_profile_enable(owtk_context ctx, int enable)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->profiling = enable != 0;
    });
}

int owtk_profile_reset(owtk_context ctx)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        ctx->sync();
        for (auto& r : ctx->records) {
            ctx->event_pool.push_back(r.start);
            ctx->event_pool.push_back(r.stop);
        }
        ctx->records.clear();
    });
}

int owtk_profile_read(owtk_context ctx, int category, uint64_t* launches, double* total_ms, int64_t* bytes)
{
    return guarded([&] {
        check_ctx(ctx);
        if (!launches || !total_ms || !bytes) throw std::invalid_argument("null output");
        ctx->activate();
        ctx->sync();
        *launches = 0;
        *total_ms = 0.0;
        *bytes = 0;
        for (const auto& r : ctx->records) {
            if (r.category != category) continue;
            float ms = 0.0f;
            OWTK_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
            *launches += 1;
            *total_ms += ms;
            *bytes += r.bytes;
        }
    });
}

int owtk_profile_region(owtk_context ctx, int begin, double* ms)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (!ctx->region_start) {
            OWTK_HIP(hipEventCreate(&ctx->region_start));
            OWTK_HIP(hipEventCreate(&ctx->region_stop));
        }
        if (begin) {
            OWTK_HIP(hipEventRecord(ctx->region_start, ctx->stream));
            return;
        }
        if (!ms) throw std::invalid_argument("null output");
        OWTK_HIP(hipEventRecord(ctx->region_stop, ctx->stream));
        OWTK_HIP(hipEventSynchronize(ctx->region_stop));
        float v = 0.0f;
        OWTK_HIP(hipEventElapsedTime(&v, ctx->region_start, ctx->region_stop));
        *ms = v;
    });
}

int owtk_context_destroy(owtk_context ctx)
{
    return guarded([&] {
        if (!ctx) return;
        ctx->activate();
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& r : ctx->records) {
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.stop);
        }
        for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
        if (ctx->region_start) (void)hipEventDestroy(ctx->region_start);
        if (ctx->region_stop) (void)hipEventDestroy(ctx->region_stop);
        hipStream_t st = ctx->stream;
        delete ctx;
        (void)hipStreamDestroy(st);
    });
}

int owtk_context_synchronize(owtk_context ctx)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->sync();
    });
}

int owtk_malloc(owtk_context ctx, size_t bytes, void** device_pointer)
{
    return guarded([&] {
        check_ctx(ctx);
        if (!device_pointer) throw std::invalid_argument("null output");
        ctx->activate();
        *device_pointer = nullptr;
        if (bytes == 0) return;
        const hipError_t err = hipMalloc(device_pointer, bytes);
        if (err == hipErrorOutOfMemory) throw std::bad_alloc();
        if (err != hipSuccess) throw_hip(err, "hipMalloc");
    });
}

int owtk_free(owtk_context ctx, void* device_pointer)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (device_pointer) {
            ctx->sync();
            OWTK_HIP(hipFree(device_pointer));
        }
    });
}

int owtk_memcpy_htod(owtk_context ctx, void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (bytes == 0) return;
        if (!dst || !src) throw std::invalid_argument("null pointer");
        OWTK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        ctx->sync();
    });
}

int owtk_memcpy_dtoh(owtk_context ctx, void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (bytes == 0) return;
        if (!dst || !src) throw std::invalid_argument("null pointer");
        OWTK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

int owtk_memcpy_dtod(owtk_context ctx, void* dst, const void* src, size_t bytes)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (bytes == 0) return;
        if (!dst || !src) throw std::invalid_argument("null pointer");
        OWTK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}

int owtk_fill_zero(owtk_context ctx, void* device_pointer, size_t bytes)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (bytes == 0) return;
        if (!device_pointer) throw std::invalid_argument("null pointer");
        kern::fill_zero_bytes(ctx->stream, device_pointer, bytes);
    });
}

int owtk_convert(owtk_context ctx, size_t count, owtk_dtype dst_type, void* dst, owtk_dtype src_type,
                 const void* src)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (count == 0) return;
        if (!dst || !src) throw std::invalid_argument("null pointer");
        if (dst_type == OWTK_F32 && src_type == OWTK_F64)
            kern::convert(ctx->stream, static_cast<float*>(dst), as<double>(src), count);
        else if (dst_type == OWTK_F64 && src_type == OWTK_F32)
            kern::convert(ctx->stream, static_cast<double*>(dst), as<float>(src), count);
        else if (dst_type == src_type)
            OWTK_HIP(hipMemcpyAsync(dst, src, count * (dst_type == OWTK_F64 ? 8 : 4), hipMemcpyDeviceToDevice,
                                    ctx->stream));
        else
            throw std::invalid_argument("unsupported vector conversion");
    });
}

int owtk_dot(owtk_context ctx, owtk_dtype dtype, size_t n, const void* a, const void* b, double* out)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (!out) throw std::invalid_argument("null output");
        if (n == 0) {
            *out = 0.0;
            return;
        }
        if (!a || !b) throw std::invalid_argument("null vector");
        double* d_out = ctx->scalars.get();
        if (dtype == OWTK_F64)
            kern::dot_to(ctx->stream, ctx->scratch(), as<double>(a), as<double>(b), static_cast<std::int64_t>(n), d_out);
        else
            kern::dot_to(ctx->stream, ctx->scratch(), as<float>(a), as<float>(b), static_cast<std::int64_t>(n), d_out);
        OWTK_HIP(hipMemcpyAsync(ctx->pinned.ptr, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        *out = *static_cast<double*>(ctx->pinned.ptr);
    });
}

int owtk_norm(owtk_context ctx, owtk_dtype dtype, size_t n, const void* a, double* out)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (!out) throw std::invalid_argument("null output");
        if (n == 0) {
            *out = 0.0;
            return;
        }
        if (!a) throw std::invalid_argument("null vector");
        double* d_out = ctx->scalars.get();
        if (dtype == OWTK_F64)
            kern::norm_to(ctx->stream, ctx->scratch(), as<double>(a), static_cast<std::int64_t>(n), d_out);
        else
            kern::norm_to(ctx->stream, ctx->scratch(), as<float>(a), static_cast<std::int64_t>(n), d_out);
        OWTK_HIP(hipMemcpyAsync(ctx->pinned.ptr, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        *out = *static_cast<double*>(ctx->pinned.ptr);
    });
}

// Reduction::dot and Reduction::norm (reduction.hpp:114-168, 176-299) with
// the reduction selected: serial, compensated or mixed precision (FP64
// vectors: double-double sums; FP32 vectors: FP64 sums, as serial).
int owtk_reduction_dot(owtk_context ctx, int32_t reduction, owtk_dtype dtype, size_t n, const void* a,
                       const void* b, double* out)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (reduction < OWTK_REDUCTION_SERIAL || reduction > OWTK_REDUCTION_MIXED_PRECISION)
            throw std::invalid_argument("unknown reduction");
        if (!out) throw std::invalid_argument("null output");
        if (n == 0) {
            *out = 0.0;
            return;
        }
        if (!a || !b) throw std::invalid_argument("null vector");
        double* d_out = ctx->scalars.get();
        if (dtype == OWTK_F64)
            kern::dot_to(ctx->stream, ctx->scratch(), as<double>(a), as<double>(b), static_cast<std::int64_t>(n), d_out,
                         reduction != OWTK_REDUCTION_SERIAL);
        else
            kern::dot_to(ctx->stream, ctx->scratch(), as<float>(a), as<float>(b), static_cast<std::int64_t>(n), d_out,
                         reduction == OWTK_REDUCTION_COMPENSATED);
        OWTK_HIP(hipMemcpyAsync(ctx->pinned.ptr, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        *out = *static_cast<double*>(ctx->pinned.ptr);
    });
}

int owtk_reduction_norm(owtk_context ctx, int32_t reduction, owtk_dtype dtype, size_t n, const void* a, double* out)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (reduction < OWTK_REDUCTION_SERIAL || reduction > OWTK_REDUCTION_MIXED_PRECISION)
            throw std::invalid_argument("unknown reduction");
        if (!out) throw std::invalid_argument("null output");
        if (n == 0) {
            *out = 0.0;
            return;
        }
        if (!a) throw std::invalid_argument("null vector");
        double* d_out = ctx->scalars.get();
        if (dtype == OWTK_F64)
            kern::norm_to(ctx->stream, ctx->scratch(), as<double>(a), static_cast<std::int64_t>(n), d_out,
                          reduction != OWTK_REDUCTION_SERIAL);
        else
            kern::norm_to(ctx->stream, ctx->scratch(), as<float>(a), static_cast<std::int64_t>(n), d_out,
                          reduction == OWTK_REDUCTION_COMPENSATED);
        OWTK_HIP(hipMemcpyAsync(ctx->pinned.ptr, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        *out = *static_cast<double*>(ctx->pinned.ptr);
    });
}

int owtk_axpy(owtk_context ctx, size_t n, double alpha, const double* x, double* y)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (n == 0) return;
        if (!x || !y) throw std::invalid_argument("null vector");
        kern::axpy(ctx->stream, static_cast<std::int64_t>(n), alpha, x, y);
    });
}

int owtk_axpy_dot(owtk_context ctx, size_t n, double alpha, const double* x, double* y, const double* w,
                  double* out)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (!out) throw std::invalid_argument("null output");
        if (n == 0) {
            *out = 0.0;
            return;
        }
        if (!x || !y || !w) throw std::invalid_argument("null vector");
        double* d_out = ctx->scalars.get();
        kern::axpy_dot_to(ctx->stream, ctx->scratch(), static_cast<std::int64_t>(n), alpha, x, y, w, d_out);
        OWTK_HIP(hipMemcpyAsync(ctx->pinned.ptr, d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        *out = *static_cast<double*>(ctx->pinned.ptr);
    });
}

int owtk_multidot(owtk_context ctx, size_t n, size_t k, const double* const* vs, const double* w, double* out)
{
    return guarded([&] {
        check_ctx(ctx);
        ctx->activate();
        if (!out || (k > 0 && !vs)) throw std::invalid_argument("null argument");
        if (k + 1 > static_cast<size_t>(kern::kMaxMultidot) + 1)
            throw std::invalid_argument("multidot supports at most 62 basis vectors");
        if (n == 0) {
            std::fill(out, out + k + 1, 0.0);
            return;
        }
        if (!w) throw std::invalid_argument("null vector");
        for (size_t j = 0; j < k; ++j)
            if (!vs[j]) throw std::invalid_argument("null basis vector");
        double* d_out = ctx->scalars.get();
        kern::multidot_to(ctx->stream, ctx->scratch(), static_cast<std::int64_t>(n), static_cast<int>(k), vs, w, d_out);
        OWTK_HIP(hipMemcpyAsync(ctx->pinned.ptr, d_out, (k + 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        std::copy_n(static_cast<double*>(ctx->pinned.ptr), k + 1, out);
    });
}

} // extern "C"
