// Internal helpers shared by every C ABI entry point: exception-to-status
// translation and the thread-local last-error message.
#pragma once

#include "../../include/owtk_hip.h"

#include <new>
#include <stdexcept>
#include <string>

namespace owtk {

// Thrown for a numerically singular ILU pivot; maps to OWTK_ERROR_ZERO_PIVOT
// (the reference throws std::runtime_error("zero pivot in ILU(0)"),
// preconditioner.hpp:451, 480).
struct zero_pivot_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct no_device_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct unsupported_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const char* message);

template<class F>
int guarded(F&& body) noexcept
{
    try {
        body();
        set_last_error("");
        return OWTK_OK;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return OWTK_ERROR_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        set_last_error(e.what());
        return OWTK_ERROR_INVALID_ARGUMENT;
    } catch (const zero_pivot_error& e) {
        set_last_error(e.what());
        return OWTK_ERROR_ZERO_PIVOT;
    } catch (const hip_error& e) {
        set_last_error(e.what());
        return OWTK_ERROR_HIP;
    } catch (const no_device_error& e) {
        set_last_error(e.what());
        return OWTK_ERROR_NO_DEVICE;
    } catch (const unsupported_error& e) {
        set_last_error(e.what());
        return OWTK_ERROR_UNSUPPORTED;
    } catch (const std::bad_alloc&) {
        set_last_error("out of memory");
        return OWTK_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return OWTK_ERROR_RUNTIME;
    } catch (...) {
        set_last_error("unknown error");
        return OWTK_ERROR_RUNTIME;
    }
}

} // namespace owtk
