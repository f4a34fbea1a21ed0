// The batched Gram-Schmidt multi-dot for callers that cannot allocate
// (code under HIP graph capture, tight serving loops): the same kernels and
// grid as owtk_multidot, with the cross-block partials in a file-scope device
// array instead of a context's buffer, and the results left on the device.
//
// Each device holds its own copy of `scratch` (516,096 bytes) until the
// process ends.  Host code never passes the array's name to a kernel: the
// per-device address comes from hipGetSymbolAddress on the entry point's
// first call on that device, and every later call uses that address.  Every
// partial a finisher reads is written by the streaming kernel of the same
// call, so nothing in `scratch` has to start at zero; calls that overlap on
// one device share the array and race.
#include "internal.hpp"
#include "kernels/blas1.hpp"
#include "owtk_hip_api.hpp"

#include <mutex>

namespace {

using namespace owtk;

// The multi-dot launches at most 1024 blocks and reduces at most 63 values.
constexpr std::int64_t kNoAllocBlocks = kern::kMultidotMaxBlocks;
constexpr std::int64_t kNoAllocValues = kern::kMaxMultidot + 1;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocValues * kNoAllocBlocks];

std::mutex scratch_mutex;
double* scratch_address[kMaxDevices] = {};

// Address of this device's `scratch` (looked up once per device).
double* device_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_address[device] == nullptr) {
        void* p = nullptr;
        OWTK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) throw hip_error("hipGetSymbolAddress returned a null address");
        scratch_address[device] = static_cast<double*>(p);
    }
    return scratch_address[device];
}

__global__ __launch_bounds__(64) void k_zero_out(double* out, int count)
{
    const int i = static_cast<int>(threadIdx.x);
    if (i < count) out[i] = 0.0;
}

} // namespace

extern "C" {

int owtk_multidot_noalloc(void* stream, size_t n, size_t k, const double* const* vs, const double* w, double* out)
{
    return guarded([&] {
        if (!out || (k > 0 && !vs)) throw std::invalid_argument("null argument");
        if (k > static_cast<size_t>(kern::kMaxMultidot))
            throw std::invalid_argument("multidot supports at most 62 basis vectors");
        if (n > 0) {
            if (!w) throw std::invalid_argument("null vector");
            for (size_t j = 0; j < k; ++j)
                if (!vs[j]) throw std::invalid_argument("null basis vector");
        }
        int device = -1;
        OWTK_HIP(hipGetDevice(&device));
        double* partials = device_scratch(device);
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (n == 0) {
            k_zero_out<<<1, 64, 0, st>>>(out, static_cast<int>(k + 1));
            OWTK_LAUNCH_CHECK();
            return;
        }
        kern::multidot_to(st, dev::ReduceScratch{partials, kNoAllocBlocks}, static_cast<std::int64_t>(n),
                          static_cast<int>(k), vs, w, out);
    });
}

} // extern "C"
