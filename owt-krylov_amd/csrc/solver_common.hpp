// Host-side pieces shared by the device solvers: the reference's option
// validation, stopping threshold, result bookkeeping and enum values.
#pragma once

#include "precond.hpp"

#include <cmath>
#include <cstring>
#include <limits>
#include <string>

namespace owtk {

// SolverStatus (core.hpp:24-30)
enum : int {
    status_converged = 0,
    status_maximum_iterations = 1,
    status_breakdown = 2,
    status_diverged = 3,
    status_invalid_input = 4,
};

// BreakdownReason (core.hpp:41-52)
enum : int {
    reason_none = 0,
    reason_arnoldi_invariant_subspace = 1,
    reason_singular_projected_system = 2,
    reason_biorthogonality_loss = 3,
    reason_alpha_denominator = 4,
    reason_omega_denominator = 5,
    reason_omega_zero = 6,
    reason_idr_shadow_space = 7,
    reason_idr_small_system = 8,
    reason_non_finite_scalar = 9,
};

inline void init_result(owtk_solver_result* r)
{
    std::memset(r, 0, sizeof(*r));
    r->status = status_invalid_input;
    r->breakdown_reason = reason_none;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    r->initial_residual_norm = r->recursive_residual_norm = r->true_residual_norm = r->relative_residual_norm = nan;
}

// valid_options (core.hpp:536-547)
inline bool valid_options(const owtk_solver_options* o)
{
    return std::isfinite(o->relative_tolerance) && std::isfinite(o->absolute_tolerance)
        && std::isfinite(o->breakdown_tolerance) && o->relative_tolerance >= 0 && o->absolute_tolerance >= 0
        && o->breakdown_tolerance > 0 && o->maximum_iterations > 0 && o->convergence_check_interval > 0;
}

// Options only some solvers implement on the device: a non-serial reduction
// (owtk_bicgstab) and the application convergence test (owtk_bicgstab,
// owtk_fgmres).  Other solvers reject them instead of silently ignoring them.
inline void require_serial_reduction(const owtk_solver_options* o, const char* solver)
{
    if (o->reduction != OWTK_REDUCTION_SERIAL)
        throw unsupported_error(std::string(solver)
                                + ": compensated and mixed-precision reductions are implemented by owtk_bicgstab only");
}

inline void require_plain_options(const owtk_solver_options* o, const char* solver)
{
    require_serial_reduction(o, solver);
    if (o->convergence_test != nullptr)
        throw unsupported_error(std::string(solver) + ": the application convergence test is implemented by "
                                "owtk_bicgstab and owtk_fgmres only");
}

// convergence_threshold (core.hpp:549-560), without the application callback
inline double threshold_of(double rhs_norm, const owtk_solver_options* o)
{
    const double rel = o->relative_tolerance * (rhs_norm > 0 ? rhs_norm : 1.0);
    if (!std::isfinite(rhs_norm) || !std::isfinite(rel)) return std::numeric_limits<double>::quiet_NaN();
    return std::max(o->absolute_tolerance, rel);
}

// detail::set_residual_result (krylov_solvers.hpp:110-120)
inline void set_residual(owtk_solver_result* r, double recursive, double true_norm, double rhs_norm)
{
    r->recursive_residual_norm = recursive;
    r->true_residual_norm = true_norm;
    r->relative_residual_norm = true_norm / (rhs_norm > 0 ? rhs_norm : 1.0);
}

// Copy that honours a solver's stop flag (identity preconditioner).
void copy_stop_f64(hipStream_t st, const double* in, double* out, std::int64_t n, const int* stop);
void copy_stop_f32(hipStream_t st, const float* in, float* out, std::int64_t n, const int* stop);

inline void apply_pc_f64(owtk_precond pc, const double* in, double* out, std::int64_t n, hipStream_t st,
                         const int* stop)
{
    if (pc == nullptr || pc->kind == OWTK_PC_IDENTITY) {
        copy_stop_f64(st, in, out, n, stop);
        return;
    }
    precond_apply_f64(pc, in, out, stop);
}

inline void apply_pc_f32(owtk_precond pc, const float* in, float* out, std::int64_t n, hipStream_t st,
                         const int* stop)
{
    if (pc == nullptr || pc->kind == OWTK_PC_IDENTITY) {
        copy_stop_f32(st, in, out, n, stop);
        return;
    }
    precond_apply_f32(pc, in, out, stop);
}

int bicgstab_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, const owtk_solver_options* o,
                 owtk_solver_result* res);

} // namespace owtk
