// Objects behind the opaque C handles and the host-side launch interfaces
// shared between the translation units of libowtk_hip.
#pragma once

#include "device.hpp"
#include "owtk_hip_api.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

namespace owtk {

[[noreturn]] void throw_hip(hipError_t err, const char* what);

#define OWTK_HIP(call)                                                  \
    do {                                                                \
        hipError_t owtk_hip_status_ = (call);                           \
        if (owtk_hip_status_ != hipSuccess) ::owtk::throw_hip(owtk_hip_status_, #call); \
    } while (0)

#define OWTK_LAUNCH_CHECK() OWTK_HIP(hipGetLastError())

// Owning device allocation.
template<class T>
struct DeviceArray {
    T* ptr = nullptr;
    std::size_t count = 0;

    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) { allocate(n); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), count(o.count)
    {
        o.ptr = nullptr;
        o.count = 0;
    }
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DeviceArray() { release(); }

    void allocate(std::size_t n)
    {
        release();
        if (n == 0) return;
        void* p = nullptr;
        const hipError_t err = hipMalloc(&p, n * sizeof(T));
        if (err == hipErrorOutOfMemory) throw std::bad_alloc();
        if (err != hipSuccess) throw_hip(err, "hipMalloc");
        ptr = static_cast<T*>(p);
        count = n;
    }
    void release() noexcept
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    std::size_t bytes() const { return count * sizeof(T); }
    T* get() const { return ptr; }
};

// Pinned host staging for scalar read-back.
struct PinnedBuffer {
    void* ptr = nullptr;
    std::size_t bytes = 0;
    void ensure(std::size_t n);
    ~PinnedBuffer();
};

} // namespace owtk

namespace owtk {

// Live kernel timing: HIP events around selected launches on the context
// stream, read back after the timed region (owtk_profile_*).
struct ProfileRecord {
    int category;
    hipEvent_t start, stop;
    std::int64_t bytes;
};

} // namespace owtk

struct owtk_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    owtk::DeviceArray<double> partials;      // [kMaxReduceValues][partial_stride]
    std::int64_t partial_stride = 0;         // blocks one reduction may span
    owtk::DeviceArray<double> scalars;       // small device scalar results
    owtk::PinnedBuffer pinned;
    bool profiling = false;
    std::vector<owtk::ProfileRecord> records;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t region_start = nullptr, region_stop = nullptr;  // owtk_profile_region
    // IDR(s) shadow draws by (owned entries, s, restart epoch): a fixed
    // function of those three (seeded std::mt19937_64 + normal_distribution),
    // kept on the device for later solves of the same size
    std::map<std::tuple<std::int64_t, int, std::uint64_t>, owtk::DeviceArray<double>> idr_draws;
    std::size_t idr_draw_bytes = 0;
    // FGMRES workspaces (basis, state, flags) kept between solves, one per
    // element size ([0] FP32, [1] FP64): the restart-30 basis at 2M vertices
    // is 1.6 GB, whose allocation and zero fill would otherwise recur on
    // every call.  Calls on a context run one at a time and FGMRES engines
    // never nest, so one workspace per element size serves every solve.
    // Type-erased because the workspace type is private to the FGMRES code.
    std::shared_ptr<void> fg_work[2];
    // FP64 iterative-refinement vectors, kept between solves for the same
    // reason; refine_layout and refine_owned are the local and owned lengths
    // their zeroed ghosts fit.
    owtk::DeviceArray<unsigned char> refine_ws;
    std::int64_t refine_layout = -1, refine_owned = -1;

    owtk::dev::ReduceScratch scratch() const { return {partials.get(), partial_stride}; }
    // Grow the reduction partials to hold `blocks` blocks per value (waits for
    // the stream first: queued kernels may still use the old buffer).
    void reserve_partials(std::int64_t blocks);
    void activate() const;  // hipSetDevice(device)
    void sync() const;
    hipEvent_t take_event();
};

namespace owtk {

// Records start/stop events around one launch when profiling is enabled.
class ProfileScope {
public:
    ProfileScope(owtk_context ctx, int category, std::int64_t bytes);
    ~ProfileScope();
    ProfileScope(const ProfileScope&) = delete;
    ProfileScope& operator=(const ProfileScope&) = delete;

private:
    owtk_context ctx_;
    std::size_t index_ = static_cast<std::size_t>(-1);
};

} // namespace owtk

namespace owtk {

// Device layout of one value storage of a matrix.
//
// Small blocks (block size <= 4): SELL-64 ("sliced ELLPACK", slice height = one
// wave).  Slice s holds rows 64s..64s+63; its width w_s is the longest row of
// the slice; slot (s, k, lane) lives at slice_ptr[s] + 64k + lane, so the k-th
// entry of 64 consecutive rows is one contiguous 64-wide load.  Component-
// diagonal values sit at [slot & ~63] * B + 64c + (slot & 63) (per entry k and
// component c a contiguous 64-lane run); dense blocks likewise, with elements
// 2g and 2g+1 of a slot adjacent in a 128-value run per pair (dense_index).
// Row entries keep the CSR order, so the
// per-row accumulation order is the reference's.
//
// Larger component-diagonal blocks keep the reference's CSR [entry][comp]
// layout (lanes run over components, which is already contiguous).
enum class Layout { sell, csr };

struct MatrixStructure {
    Layout layout = Layout::sell;
    std::int64_t rows = 0;
    std::int64_t slices = 0;
    std::int64_t slots = 0;            // SELL slots incl. padding (or CSR entries)
    int max_width = 0;
    DeviceArray<std::int32_t> slice_ptr;
    DeviceArray<std::int32_t> row_len;   // SELL: entries per row
    DeviceArray<std::int32_t> columns;   // per slot (or per CSR entry)
    DeviceArray<std::int32_t> offsets;   // CSR layout only
    std::vector<std::int64_t> slot_of_entry; // host: CSR entry -> slot
};

} // namespace owtk

struct owtk_matrix_s {
    owtk_context ctx = nullptr;
    int kind = OWTK_MATRIX_SPLIT;
    int bs = 1;
    std::int64_t owned = 0;
    std::int64_t ghost = 0;
    std::int64_t entries = 0;   // stored block entries (split: off-diagonal)
    // host copy of the structure (symbolic analyses of preconditioners)
    std::vector<std::int32_t> h_offsets;
    std::vector<std::int32_t> h_columns;
    owtk::MatrixStructure s;
    // values in slot layout (value_count per storage)
    std::int64_t value_count = 0;
    owtk::DeviceArray<double> v64;
    owtk::DeviceArray<float> v32;
    owtk::DeviceArray<std::uint16_t> v16;
    // split diagonal, [slice][comp][lane] (SELL) or [row][comp] (CSR layout)
    std::int64_t diag_count = 0;
    owtk::DeviceArray<double> d64;
    owtk::DeviceArray<float> d32;
    owtk::DeviceArray<std::uint16_t> d16;
    std::uint64_t value_epoch = 0; // bumped by update_values

    std::int64_t local() const { return owned + ghost; }
    bool dense() const { return kind == OWTK_MATRIX_BSR; }
    bool has_diag_array() const { return kind == OWTK_MATRIX_SPLIT; }
    // index of (slot, component c) in the value array (component-diagonal)
    std::int64_t value_index(std::int64_t slot, int c) const
    {
        if (s.layout == owtk::Layout::csr) return slot * bs + c;
        return (slot & ~std::int64_t(63)) * bs + 64 * c + (slot & 63);
    }
    std::int64_t diag_index(std::int64_t row, int c) const
    {
        if (s.layout == owtk::Layout::csr) return row * bs + c;
        return (row & ~std::int64_t(63)) * bs + 64 * c + (row & 63);
    }
};


namespace owtk {

enum ProfileCategory : int {
    prof_spmv_f64 = 0,
    prof_spmv_f32 = 1,
    prof_spmv_bf16 = 2,
    prof_precond_f64 = 3,
    prof_precond_low = 4,
    prof_learned = 5,        // learned preconditioner: middle series steps
    prof_learned_ends = 6,   // learned preconditioner: first (init folded in) and last series steps
    prof_categories = 8,
};

// Algorithmic HBM bytes of one SpMV with value size sv and vector size sx.
std::int64_t spmv_bytes(owtk_matrix m, int sv, int sx);

} // namespace owtk
