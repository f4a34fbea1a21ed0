// Header-only C++ host layer over the C ABI of libowtk_hip.so: a HIP execution
// policy with the interface of the reference's policies (execution.hpp:17-46:
// apply(matrix, input, output), name(), accelerator_available()), device
// mirrors of the three block-CSR forms, preconditioners and device-resident
// solvers.  Reference types are used duck-typed (through the accessors the
// reference exposes), so this header does not include the reference.
//
// Error behaviour follows the reference: std::invalid_argument for layout or
// argument errors, std::runtime_error("zero pivot ...") for ILU pivots
// (preconditioner.hpp:451, 480); numeric breakdown comes back in the result.
#pragma once

#include "../owtk_hip.h"

#include <concepts>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace owtk::hip {

inline void check(int code)
{
    if (code == OWTK_OK) return;
    const std::string msg = owtk_last_error();
    if (code == OWTK_ERROR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);  // zero pivot, HIP error, no device, unsupported
}

// One GPU and one HIP stream.
class Context {
public:
    explicit Context(int device = 0) { check(owtk_context_create(device, &ctx_)); }
    ~Context()
    {
        if (ctx_) (void)owtk_context_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    owtk_context get() const noexcept { return ctx_; }
    void synchronize() const { check(owtk_context_synchronize(ctx_)); }

private:
    owtk_context ctx_ = nullptr;
};

// Device vector of T (double or float), [node][comp] like BlockVector.
template<class T>
class DeviceVector {
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>);

public:
    DeviceVector(Context& ctx, std::size_t count) : ctx_(&ctx), count_(count)
    {
        check(owtk_malloc(ctx.get(), std::max<std::size_t>(count, 1) * sizeof(T), &ptr_));
        check(owtk_fill_zero(ctx.get(), ptr_, count * sizeof(T)));
    }
    ~DeviceVector()
    {
        if (ptr_) (void)owtk_free(ctx_->get(), ptr_);
    }
    DeviceVector(const DeviceVector&) = delete;
    DeviceVector& operator=(const DeviceVector&) = delete;
    DeviceVector(DeviceVector&& o) noexcept : ctx_(o.ctx_), count_(o.count_), ptr_(std::exchange(o.ptr_, nullptr)) {}

    std::size_t size() const noexcept { return count_; }
    T* data() noexcept { return static_cast<T*>(ptr_); }
    const T* data() const noexcept { return static_cast<const T*>(ptr_); }
    void upload(const T* host, std::size_t count)
    {
        if (count > count_) throw std::invalid_argument("upload larger than the device vector");
        check(owtk_memcpy_htod(ctx_->get(), ptr_, host, count * sizeof(T)));
    }
    void download(T* host, std::size_t count) const
    {
        if (count > count_) throw std::invalid_argument("download larger than the device vector");
        check(owtk_memcpy_dtoh(ctx_->get(), host, ptr_, count * sizeof(T)));
    }

private:
    Context* ctx_;
    std::size_t count_;
    void* ptr_ = nullptr;
};

namespace detail {

template<class M>
concept SplitView = requires(const M& m) {
    m.off_diagonal_row_offsets();
    m.off_diagonal_columns();
    m.diagonal_values();
    m.off_diagonal_values();
};

template<class M>
concept DenseBlocks = requires(const M& m) { m.block_values(); };

template<class M>
concept ComponentCsr = requires(const M& m) {
    m.row_offsets();
    m.column_indices();
    m.values();
} && !DenseBlocks<M>;

template<class Range>
std::vector<std::int32_t> to_int32(const Range& r)
{
    std::vector<std::int32_t> out(r.size());
    for (std::size_t i = 0; i < r.size(); ++i) {
        const auto v = r[i];
        if (v < 0 || static_cast<std::uint64_t>(v) > 0x7fffffffULL)
            throw std::invalid_argument("index does not fit the device's int32 indices");
        out[i] = static_cast<std::int32_t>(v);
    }
    return out;
}

template<class Range>
std::vector<double> to_double(const Range& r)
{
    return std::vector<double>(r.begin(), r.end());
}

} // namespace detail

// Host structure of a reference matrix (BlockCsrMatrix,
// SplitBlockCsrMatrixView or DenseBlockCsrMatrix, read through their
// accessors) as the device's int32 indices, and its values in the reference
// layout.
struct HostStructure {
    std::int64_t owned = 0, ghost = 0, block = 0;
    std::vector<std::int32_t> offsets, columns;
    bool operator==(const HostStructure&) const = default;
};

template<class Matrix>
HostStructure host_structure(const Matrix& m)
{
    HostStructure h;
    h.owned = static_cast<std::int64_t>(m.owned_nodes());
    h.ghost = static_cast<std::int64_t>(m.ghost_nodes());
    h.block = static_cast<std::int64_t>(m.block_size());
    if constexpr (detail::SplitView<Matrix>) {
        h.offsets = detail::to_int32(m.off_diagonal_row_offsets());
        h.columns = detail::to_int32(m.off_diagonal_columns());
    } else {
        h.offsets = detail::to_int32(m.row_offsets());
        h.columns = detail::to_int32(m.column_indices());
    }
    return h;
}

// Device copy of a reference matrix.
class DeviceMatrix {
public:
    template<class Matrix>
    DeviceMatrix(Context& ctx, const Matrix& m) : DeviceMatrix(ctx, m, host_structure(m))
    {
    }
    template<class Matrix>
    DeviceMatrix(Context& ctx, const Matrix& m, const HostStructure& h)
        : ctx_(&ctx), owned_(h.owned), ghost_(h.ghost), block_(h.block)
    {
        if constexpr (detail::SplitView<Matrix>) {
            const auto diag = detail::to_double(m.diagonal_values());
            const auto val = detail::to_double(m.off_diagonal_values());
            check(owtk_matrix_create_split(ctx.get(), owned_, ghost_, block_, h.offsets.data(), h.columns.data(),
                                           diag.data(), val.data(), &m_));
        } else if constexpr (detail::DenseBlocks<Matrix>) {
            const auto val = detail::to_double(m.block_values());
            check(owtk_matrix_create_bsr(ctx.get(), owned_, ghost_, block_, h.offsets.data(), h.columns.data(),
                                         val.data(), &m_));
        } else {
            static_assert(detail::ComponentCsr<Matrix>, "unsupported matrix type");
            const auto val = detail::to_double(m.values());
            check(owtk_matrix_create_csr(ctx.get(), owned_, ghost_, block_, h.offsets.data(), h.columns.data(),
                                         val.data(), &m_));
        }
    }
    ~DeviceMatrix()
    {
        if (m_) (void)owtk_matrix_destroy(m_);
    }
    DeviceMatrix(const DeviceMatrix&) = delete;
    DeviceMatrix& operator=(const DeviceMatrix&) = delete;

    owtk_matrix get() const noexcept { return m_; }
    Context& context() const noexcept { return *ctx_; }
    std::int64_t owned_nodes() const noexcept { return owned_; }
    std::int64_t ghost_nodes() const noexcept { return ghost_; }
    std::int64_t block_size() const noexcept { return block_; }
    std::size_t owned_size() const noexcept { return static_cast<std::size_t>(owned_ * block_); }
    std::size_t local_size() const noexcept { return static_cast<std::size_t>((owned_ + ghost_) * block_); }

    // Numeric refresh (same pattern): values / off_diagonal in the reference's layout.
    void update_values(const double* values, const double* off_diagonal = nullptr)
    {
        check(owtk_matrix_update_values(m_, values, off_diagonal));
    }
    // Numeric refresh from a reference matrix of the same structure.
    template<class Matrix>
    void update_values_from(const Matrix& m)
    {
        if constexpr (detail::SplitView<Matrix>) {
            const auto diag = detail::to_double(m.diagonal_values());
            const auto val = detail::to_double(m.off_diagonal_values());
            update_values(diag.data(), val.data());
        } else if constexpr (detail::DenseBlocks<Matrix>) {
            const auto val = detail::to_double(m.block_values());
            update_values(val.data());
        } else {
            const auto val = detail::to_double(m.values());
            update_values(val.data());
        }
    }
    void add_storage(owtk_dtype dtype) { check(owtk_matrix_add_storage(m_, dtype)); }

    // y = A x on device vectors (the operator interface, krylov_solvers.hpp:42-50).
    void apply(const DeviceVector<double>& x, DeviceVector<double>& y) const
    {
        check(owtk_spmv(m_, OWTK_F64, x.data(), y.data()));
    }

private:
    Context* ctx_;
    owtk_matrix m_ = nullptr;
    std::int64_t owned_, ghost_, block_;
};

// The HIP execution policy: apply(matrix, input, output) for host vectors
// with the reference's BlockVector accessors (data(), local_size(),
// owned_size()).  Like the reference's OpenMP Target policy, which maps the
// matrix on every apply (execution.hpp:66-69, 115-119), every apply sees the
// matrix's current values: a device copy is kept per matrix address, checked
// against the matrix's sizes and index arrays on every apply (rebuilt when
// they differ, e.g. a different matrix at a reused address) and its values
// uploaded again.  Transfers are sized from the checked matrix.
class HipExecutionPolicy {
public:
    explicit HipExecutionPolicy(int device = 0) : state_(std::make_shared<State>(device)) {}

    [[nodiscard]] static constexpr const char* name() noexcept { return "hip-gfx950"; }
    [[nodiscard]] static bool accelerator_available() noexcept
    {
        int n = 0;
        return owtk_device_count(&n) == OWTK_OK && n > 0;
    }

    template<class Matrix, class Vector>
    void apply(const Matrix& matrix, const Vector& input, Vector& output) const
    {
        const std::size_t local = static_cast<std::size_t>((matrix.owned_nodes() + matrix.ghost_nodes())
                                                           * matrix.block_size());
        const std::size_t owned = static_cast<std::size_t>(matrix.owned_nodes() * matrix.block_size());
        if (input.local_size() != local || output.owned_size() != owned)
            throw std::invalid_argument("execution-policy vector layout mismatch");
        using T = std::remove_cvref_t<decltype(*input.data())>;
        static_assert(std::is_same_v<T, double>, "the HIP policy's apply is FP64");
        std::lock_guard<std::mutex> lock(state_->mutex);
        DeviceMatrix& dm = state_->mirror(matrix);
        if (dm.local_size() != local || dm.owned_size() != owned)
            throw std::logic_error("device copy does not match the matrix");
        auto& x = state_->buffer(0, local);
        auto& y = state_->buffer(1, owned);
        x.upload(input.data(), local);
        dm.apply(x, y);
        y.download(output.data(), owned);
    }

    // Drops the device copy of `matrix` (the next apply rebuilds it).
    template<class Matrix>
    void refresh(const Matrix& matrix) const
    {
        std::lock_guard<std::mutex> lock(state_->mutex);
        state_->mirrors.erase(static_cast<const void*>(&matrix));
    }

    Context& context() const noexcept { return state_->ctx; }

private:
    struct Mirror {
        HostStructure structure;
        std::unique_ptr<DeviceMatrix> matrix;
    };
    struct State {
        explicit State(int device) : ctx(device) {}
        Context ctx;
        std::mutex mutex;
        std::map<const void*, Mirror> mirrors;
        std::unique_ptr<DeviceVector<double>> buffers[2];

        template<class Matrix>
        DeviceMatrix& mirror(const Matrix& m)
        {
            Mirror& slot = mirrors[static_cast<const void*>(&m)];
            HostStructure h = host_structure(m);
            if (slot.matrix && slot.structure == h) {
                slot.matrix->update_values_from(m);
            } else {
                slot.matrix.reset();
                slot.matrix = std::make_unique<DeviceMatrix>(ctx, m, h);
                slot.structure = std::move(h);
            }
            return *slot.matrix;
        }
        DeviceVector<double>& buffer(int k, std::size_t n)
        {
            if (!buffers[k] || buffers[k]->size() < n) buffers[k] = std::make_unique<DeviceVector<double>>(ctx, n);
            return *buffers[k];
        }
    };
    std::shared_ptr<State> state_;
};

// Device preconditioner (Jacobi, SSOR, ILU(k), learned).
// The learned preconditioner's graph network (owtk_gnn, no reference
// counterpart): block size and series degree fixed at creation, parameters
// one flat FP32 vector of parameter_count() values in the order
// owtk_gnn_set_parameters documents.  A fresh network gives gamma = 1, the
// plain truncated Neumann series.  Preconditioner::params(OWTK_PC_LEARNED)
// with learned_model = model.get() builds the preconditioner from it.
class GnnModel {
public:
    GnnModel(Context& ctx, int block_size, int degree) { check(owtk_gnn_create(ctx.get(), block_size, degree, &g_)); }
    ~GnnModel()
    {
        if (g_) (void)owtk_gnn_destroy(g_);
    }
    GnnModel(const GnnModel&) = delete;
    GnnModel& operator=(const GnnModel&) = delete;
    owtk_gnn get() const noexcept { return g_; }
    std::size_t parameter_count() const
    {
        std::int32_t b = 0, d = 0;
        std::int64_t n = 0;
        check(owtk_gnn_info(g_, &b, &d, &n));
        return static_cast<std::size_t>(n);
    }
    std::vector<float> parameters() const
    {
        std::vector<float> p(parameter_count());
        check(owtk_gnn_get_parameters(g_, p.data()));
        return p;
    }
    void set_parameters(const std::vector<float>& p)
    {
        if (p.size() != parameter_count()) throw std::invalid_argument("network parameter count mismatch");
        check(owtk_gnn_set_parameters(g_, p.data()));
    }

private:
    owtk_gnn g_ = nullptr;
};

class Preconditioner {
public:
    Preconditioner(DeviceMatrix& m, const owtk_precond_params& params) { check(owtk_precond_create(m.get(), &params, &pc_)); }
    static owtk_precond_params params(owtk_precond_kind kind)
    {
        owtk_precond_params p;
        owtk_precond_params_default(&p, kind);
        return p;
    }
    ~Preconditioner()
    {
        if (pc_) (void)owtk_precond_destroy(pc_);
    }
    Preconditioner(const Preconditioner&) = delete;
    Preconditioner& operator=(const Preconditioner&) = delete;
    owtk_precond get() const noexcept { return pc_; }
    void update_values() { check(owtk_precond_update_values(pc_)); }
    std::uint64_t numeric_updates() const
    {
        std::uint64_t n = 0;
        check(owtk_precond_numeric_updates(pc_, &n));
        return n;
    }
    void apply(const DeviceVector<double>& in, DeviceVector<double>& out) const
    {
        check(owtk_precond_apply(pc_, in.data(), out.data()));
    }

private:
    owtk_precond pc_ = nullptr;
};

// ---------------------------------------------------- reference-shaped solvers
// SolverStatus, GmresOrthogonalization and BreakdownReason (core.hpp:24-52)
// with the reference's values: the C ABI carries them as integers.
enum class SolverStatus : std::int32_t { converged, maximum_iterations, breakdown, diverged, invalid_input };
enum class GmresOrthogonalization : std::int32_t {
    iterated_classical_gram_schmidt,
    adaptive_classical_gram_schmidt,
    one_synchronization_classical_gram_schmidt,
};
enum class BreakdownReason : std::int32_t {
    none,
    arnoldi_invariant_subspace,
    singular_projected_system,
    biorthogonality_loss,
    alpha_denominator,
    omega_denominator,
    omega_zero,
    idr_shadow_space,
    idr_small_system,
    non_finite_scalar,
};

// SolverOptions<T> (core.hpp:88-111): the same fields and defaults.  The
// application test receives the device solution (count owned FP64 values).
template<class T = double>
struct SolverOptions {
    T relative_tolerance = T(1e-8);
    T absolute_tolerance = T(0);
    T breakdown_tolerance = T(64) * std::numeric_limits<T>::epsilon();
    std::size_t maximum_iterations = 1000;
    std::size_t convergence_check_interval = 10;
    std::size_t restart = 30;
    std::size_t residual_replacement_interval = 50;
    std::size_t idr_shadow_space = 4;
    T relaxation = T(1);
    bool verify_true_residual = true;
    bool collect_timings = false;
    std::function<bool(std::size_t, const T*, std::size_t)> convergence_test;
    GmresOrthogonalization gmres_orthogonalization = GmresOrthogonalization::iterated_classical_gram_schmidt;
};

// SolverResult<T> (core.hpp:118-146): the same fields.  It converts to the
// reference's SolverResult<T> (or any type with those fields), so
// `owt::krylov::SolverResult<double> r = bicgstab(...)` compiles unchanged.
template<class T = double>
struct SolverResult {
    SolverStatus status = SolverStatus::invalid_input;
    BreakdownReason breakdown_reason = BreakdownReason::none;
    std::size_t iterations = 0;
    std::size_t operator_applications = 0;
    std::size_t preconditioner_applications = 0;
    std::size_t global_reductions = 0;
    std::size_t reorthogonalizations = 0;
    std::size_t arnoldi_norm_verifications = 0;
    T initial_residual_norm = std::numeric_limits<T>::quiet_NaN();
    T recursive_residual_norm = std::numeric_limits<T>::quiet_NaN();
    T true_residual_norm = std::numeric_limits<T>::quiet_NaN();
    T relative_residual_norm = std::numeric_limits<T>::quiet_NaN();
    bool converged_by_application = false;
    double seconds = 0;
    // iterative refinement (fgmres_refined) only
    std::size_t refinement_steps = 0, inner_iterations = 0, fallback_refinements = 0;

    [[nodiscard]] constexpr bool converged() const noexcept { return status == SolverStatus::converged; }
    [[nodiscard]] double solve_seconds() const noexcept { return seconds; }

    template<class R>
        requires(!std::is_same_v<R, SolverResult> && requires(R r) {
            r.status;
            r.breakdown_reason;
            r.iterations;
            r.relative_residual_norm;
        })
    operator R() const
    {
        R r{};
        r.status = static_cast<decltype(r.status)>(status);
        r.breakdown_reason = static_cast<decltype(r.breakdown_reason)>(breakdown_reason);
        r.iterations = iterations;
        r.operator_applications = operator_applications;
        r.preconditioner_applications = preconditioner_applications;
        r.global_reductions = global_reductions;
        r.reorthogonalizations = reorthogonalizations;
        r.arnoldi_norm_verifications = arnoldi_norm_verifications;
        r.initial_residual_norm = initial_residual_norm;
        r.recursive_residual_norm = recursive_residual_norm;
        r.true_residual_norm = true_residual_norm;
        r.relative_residual_norm = relative_residual_norm;
        if constexpr (requires { r.converged_by_application; }) r.converged_by_application = converged_by_application;
        if constexpr (requires { r.timings->solve_seconds; }) {
            if (seconds > 0) {
                using Timings = typename std::remove_cvref_t<decltype(r.timings)>::element_type;
                r.timings = std::make_shared<Timings>();
                r.timings->solve_seconds = seconds;
            }
        }
        return r;
    }
};

inline SolverResult<double> from_c(const owtk_solver_result& c)
{
    SolverResult<double> r;
    r.status = static_cast<SolverStatus>(c.status);
    r.breakdown_reason = static_cast<BreakdownReason>(c.breakdown_reason);
    r.iterations = c.iterations;
    r.operator_applications = c.operator_applications;
    r.preconditioner_applications = c.preconditioner_applications;
    r.global_reductions = c.global_reductions;
    r.reorthogonalizations = c.reorthogonalizations;
    r.arnoldi_norm_verifications = c.arnoldi_norm_verifications;
    r.initial_residual_norm = c.initial_residual_norm;
    r.recursive_residual_norm = c.recursive_residual_norm;
    r.true_residual_norm = c.true_residual_norm;
    r.relative_residual_norm = c.relative_residual_norm;
    r.converged_by_application = c.converged_by_application != 0;
    r.seconds = c.solve_seconds;
    r.refinement_steps = c.refinement_steps;
    r.inner_iterations = c.inner_iterations;
    r.fallback_refinements = c.fallback_refinements;
    return r;
}

// The reductions (reduction.hpp:103-299): how the device reduces a solve's
// inner products and norms (owtk_reduction).
struct SerialReduction {
    static constexpr owtk_reduction kind = OWTK_REDUCTION_SERIAL;
};
struct CompensatedSerialReduction {
    static constexpr owtk_reduction kind = OWTK_REDUCTION_COMPENSATED;
};
struct MixedPrecisionSerialReduction {
    static constexpr owtk_reduction kind = OWTK_REDUCTION_MIXED_PRECISION;
};

// IdentityPreconditioner (preconditioner.hpp:17-23)
struct IdentityPreconditioner {};

// ArnoldiSnapshot<double> (core.hpp:462-534) kept on the device: the last
// FGMRES cycle's unrotated Hessenberg columns and preconditioned basis
// vectors.  preconditioned_basis(column) is the device address of the
// vector (owned + ghost values, ghosts zero).
class ArnoldiSnapshot {
public:
    explicit ArnoldiSnapshot(Context& ctx) { check(owtk_arnoldi_snapshot_create(ctx.get(), &snap_)); }
    ~ArnoldiSnapshot()
    {
        if (snap_) (void)owtk_arnoldi_snapshot_destroy(snap_);
    }
    ArnoldiSnapshot(const ArnoldiSnapshot&) = delete;
    ArnoldiSnapshot& operator=(const ArnoldiSnapshot&) = delete;
    owtk_arnoldi_snapshot get() const noexcept { return snap_; }
    void clear() { check(owtk_arnoldi_snapshot_clear(snap_)); }
    std::size_t columns() const
    {
        std::uint64_t c = 0;
        check(owtk_arnoldi_snapshot_columns(snap_, &c));
        return static_cast<std::size_t>(c);
    }
    double h(std::size_t row, std::size_t column) const
    {
        const std::size_t c = columns();
        if (column >= c || row > c) throw std::out_of_range("Arnoldi snapshot entry");
        double v = 0;
        check(owtk_arnoldi_snapshot_h(snap_, row, column, &v));
        return v;
    }
    const double* preconditioned_basis(std::size_t column) const
    {
        if (column >= columns()) throw std::out_of_range("Arnoldi snapshot basis");
        const double* p = nullptr;
        check(owtk_arnoldi_snapshot_basis(snap_, column, &p));
        return p;
    }
    // begin_cycle / record_column / finish_cycle (core.hpp:465-502) for a
    // snapshot an application assembles itself; `layout` fixes the vector
    // layout and `basis` holds device vectors.
    void begin_cycle(const DeviceMatrix& layout, std::size_t maximum_columns)
    {
        check(owtk_arnoldi_snapshot_begin_cycle(snap_, layout.get(), maximum_columns));
    }
    void record_column(std::size_t column, std::span<const double> values)
    {
        check(owtk_arnoldi_snapshot_record_column(snap_, column, values.data(), values.size()));
    }
    void finish_cycle(std::span<const DeviceVector<double>> basis, std::size_t columns)
    {
        if (basis.size() < columns) throw std::invalid_argument("invalid Arnoldi snapshot basis");
        std::vector<const double*> p(columns);
        for (std::size_t c = 0; c < columns; ++c) p[c] = basis[c].data();
        check(owtk_arnoldi_snapshot_finish_cycle(snap_, p.data(), columns));
    }

private:
    owtk_arnoldi_snapshot snap_ = nullptr;
};

// RecycleSpace<double> (recycling.hpp:24-187) kept on the device: up to
// `capacity` pairs (U_i, C_i = A U_i), C orthonormal, in the vector layout of
// `layout`.  The device reduces with the serial reduction.
class RecycleSpace {
public:
    RecycleSpace(const DeviceMatrix& layout, std::size_t capacity) : capacity_(capacity)
    {
        check(owtk_recycle_space_create(layout.get(), capacity, &space_));
    }
    ~RecycleSpace()
    {
        if (space_) (void)owtk_recycle_space_destroy(space_);
    }
    RecycleSpace(const RecycleSpace&) = delete;
    RecycleSpace& operator=(const RecycleSpace&) = delete;
    owtk_recycle_space get() const noexcept { return space_; }
    std::size_t capacity() const noexcept { return capacity_; }
    std::size_t size() const
    {
        std::int64_t n = 0;
        check(owtk_recycle_space_size(space_, &n));
        return static_cast<std::size_t>(n);
    }
    bool empty() const { return size() == 0; }
    void clear() { check(owtk_recycle_space_clear(space_)); }
    // add_candidate (recycling.hpp:76-135): false when the candidate is
    // numerically contained in the space.
    template<class Telemetry = std::nullptr_t>
    bool add_candidate(DeviceMatrix& op, const DeviceVector<double>& candidate, Telemetry* telemetry = nullptr,
                       double dependence_tolerance = 128 * std::numeric_limits<double>::epsilon())
    {
        owtk_solver_result t{};
        std::int32_t added = 0;
        check(owtk_recycle_space_add_candidate(space_, op.get(), candidate.data(), dependence_tolerance, &added, &t));
        if (telemetry) {
            telemetry->operator_applications += t.operator_applications;
            telemetry->global_reductions += t.global_reductions;
        }
        return added != 0;
    }

private:
    owtk_recycle_space space_ = nullptr;
    std::size_t capacity_;
};

namespace detail {

inline owtk_precond pc_handle(const Preconditioner& p) noexcept { return p.get(); }
inline owtk_precond pc_handle(const Preconditioner* p) noexcept { return p ? p->get() : nullptr; }
inline owtk_precond pc_handle(IdentityPreconditioner) noexcept { return nullptr; }
inline owtk_precond pc_handle(std::nullptr_t) noexcept { return nullptr; }

template<class Snapshot>
owtk_arnoldi_snapshot snapshot_handle(Snapshot s)
{
    static_assert(std::is_same_v<Snapshot, std::nullptr_t> || std::is_same_v<Snapshot, ArnoldiSnapshot*>,
                  "the device solvers record into an owtk::hip::ArnoldiSnapshot");
    if constexpr (std::is_same_v<Snapshot, ArnoldiSnapshot*>) return s ? s->get() : nullptr;
    else return nullptr;
}

template<class R>
constexpr owtk_reduction reduction_of()
{
    static_assert(requires { R::kind; },
                  "reduction: owtk::hip::SerialReduction, CompensatedSerialReduction or MixedPrecisionSerialReduction");
    return R::kind;
}

template<class F>
struct function_args;
template<class R, class A1, class A2>
struct function_args<std::function<R(A1, A2)>> {
    using second = std::remove_cvref_t<A2>;
};

// The application test behind the C callback.  A test over a device
// pointer (owtk::hip::SolverOptions) gets the device solution; a test over
// a host vector type (the reference's BlockVector<T>, constructed as
// (owned, ghost, block)) gets a host copy of the owned values.
template<class Options>
struct TestBridge {
    const Options* options = nullptr;
    Context* ctx = nullptr;
    std::size_t owned_nodes = 0, ghost_nodes = 0, block = 0;
    std::exception_ptr error;

    static int call(void* self_, std::uint64_t iteration, const double* device_solution, std::size_t count)
    {
        auto* self = static_cast<TestBridge*>(self_);
        try {
            const auto& test = self->options->convergence_test;
            using F = std::remove_cvref_t<decltype(test)>;
            if constexpr (std::is_invocable_r_v<bool, F, std::size_t, const double*, std::size_t>) {
                return test(static_cast<std::size_t>(iteration), device_solution, count) ? 1 : 0;
            } else {
                using HostVector = typename function_args<F>::second;
                HostVector host(self->owned_nodes, self->ghost_nodes, self->block);
                check(owtk_memcpy_dtoh(self->ctx->get(), host.data(), device_solution, count * sizeof(double)));
                return test(static_cast<std::size_t>(iteration), host) ? 1 : 0;
            }
        } catch (...) {
            self->error = std::current_exception();
            return 1;  // stop the solve; the exception is rethrown after it
        }
    }
};

// SolverOptions (ours or the reference's) to the C struct.
template<class Options>
owtk_solver_options to_c(const Options& o, owtk_reduction reduction = OWTK_REDUCTION_SERIAL)
{
    owtk_solver_options c;
    owtk_solver_options_default(&c);
    c.relative_tolerance = static_cast<double>(o.relative_tolerance);
    c.absolute_tolerance = static_cast<double>(o.absolute_tolerance);
    c.breakdown_tolerance = static_cast<double>(o.breakdown_tolerance);
    c.maximum_iterations = o.maximum_iterations;
    c.convergence_check_interval = o.convergence_check_interval;
    c.restart = o.restart;
    c.residual_replacement_interval = o.residual_replacement_interval;
    c.gmres_orthogonalization = static_cast<std::int32_t>(o.gmres_orthogonalization);
    c.collect_timings = o.collect_timings ? 1 : 0;
    if constexpr (requires { o.idr_shadow_space; }) c.idr_shadow_space = o.idr_shadow_space;
    if constexpr (requires { o.relaxation; }) c.relaxation = static_cast<double>(o.relaxation);
    if constexpr (requires { o.verify_true_residual; }) c.verify_true_residual = o.verify_true_residual ? 1 : 0;
    c.reduction = reduction;
    return c;
}

// One solve through a C entry point with the options' test bridged.
template<class Options, class Run>
SolverResult<double> run_solver(DeviceMatrix& op, const DeviceVector<double>& rhs, DeviceVector<double>& solution,
                                const Options& options, owtk_reduction reduction, Run&& run)
{
    if (rhs.size() < op.owned_size() || solution.size() < op.local_size())
        throw std::invalid_argument("solver vector layout mismatch");
    owtk_solver_options c = to_c(options, reduction);
    TestBridge<Options> bridge;
    if constexpr (requires { static_cast<bool>(options.convergence_test); }) {
        if (static_cast<bool>(options.convergence_test)) {
            bridge.options = &options;
            bridge.ctx = &op.context();
            bridge.owned_nodes = static_cast<std::size_t>(op.owned_nodes());
            bridge.ghost_nodes = static_cast<std::size_t>(op.ghost_nodes());
            bridge.block = static_cast<std::size_t>(op.block_size());
            c.convergence_test = &TestBridge<Options>::call;
            c.convergence_test_data = &bridge;
        }
    }
    owtk_solver_result r;
    check(run(c, r));
    if (bridge.error) std::rethrow_exception(bridge.error);
    return from_c(r);
}

} // namespace detail

// bicgstab(op, b, x, options, pc, reduction, workspace)
// (krylov_solvers.hpp:542-553) on device vectors, found by argument-dependent
// lookup on the owtk::hip types: a caller's unqualified
// `bicgstab(op, b, x, options, pc)` keeps its argument order.  The device
// keeps its own workspace, so `workspace` is accepted and not used.
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t>
SolverResult<double> bicgstab(DeviceMatrix& op, const DeviceVector<double>& rhs, DeviceVector<double>& solution,
                              const Options& options = Options{}, Pc&& preconditioner = Pc{},
                              Reduction = Reduction{}, Workspace = nullptr)
{
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_bicgstab(op.get(), pc, rhs.data(), solution.data(), &c, &r);
                              });
}

// bicgstab_stable / bicgstab_mixed_precision (krylov_solvers.hpp:766-809):
// bicgstab with MixedPrecisionSerialReduction (true residual verified).
template<class Options, class Pc>
SolverResult<double> bicgstab_stable(DeviceMatrix& op, const DeviceVector<double>& rhs,
                                     DeviceVector<double>& solution, const Options& options, Pc&& preconditioner)
{
    return bicgstab(op, rhs, solution, options, std::forward<Pc>(preconditioner), MixedPrecisionSerialReduction{});
}

template<class Options, class Pc>
SolverResult<double> bicgstab_mixed_precision(DeviceMatrix& op, const DeviceVector<double>& rhs,
                                              DeviceVector<double>& solution, const Options& options,
                                              Pc&& preconditioner)
{
    return bicgstab(op, rhs, solution, options, std::forward<Pc>(preconditioner), MixedPrecisionSerialReduction{});
}

// fgmres(op, b, x, options, pc, reduction, workspace, arnoldi_snapshot)
// (krylov_solvers.hpp:159-171).  The device FGMRES reduces with the serial
// reduction (another reduction is rejected); `arnoldi_snapshot` is null or an
// owtk::hip::ArnoldiSnapshot*.
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t, class Snapshot = std::nullptr_t>
SolverResult<double> fgmres(DeviceMatrix& op, const DeviceVector<double>& rhs, DeviceVector<double>& solution,
                            const Options& options = Options{}, Pc&& preconditioner = Pc{},
                            Reduction = Reduction{}, Workspace = nullptr, Snapshot arnoldi_snapshot = nullptr)
{
    const owtk_arnoldi_snapshot snap = detail::snapshot_handle(arnoldi_snapshot);
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_fgmres_with_snapshot(op.get(), pc, rhs.data(), solution.data(), &c, snap,
                                                                   &r);
                              });
}

// recycled_fgmres(op, b, x, space, options, pc, reduction, update,
// workspace, arnoldi_snapshot) (recycling.hpp:291-310): FGMRES augmented
// with the recycle space; with update_recycle_space the converged correction
// becomes the next candidate.  Serial reduction only.
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t, class Snapshot = std::nullptr_t>
SolverResult<double> recycled_fgmres(DeviceMatrix& op, const DeviceVector<double>& rhs,
                                     DeviceVector<double>& solution, RecycleSpace& recycle_space,
                                     const Options& options = Options{}, Pc&& preconditioner = Pc{},
                                     Reduction = Reduction{}, bool update_recycle_space = true,
                                     Workspace = nullptr, Snapshot arnoldi_snapshot = nullptr)
{
    const owtk_arnoldi_snapshot snap = detail::snapshot_handle(arnoldi_snapshot);
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_recycled_fgmres_with_snapshot(op.get(), pc, rhs.data(),
                                                                            solution.data(), recycle_space.get(), &c,
                                                                            update_recycle_space ? 1 : 0, snap, &r);
                              });
}

// extract_harmonic_ritz(op, snapshot, space, reduction, telemetry)
// (gcrodr.hpp:86-181): replaces the space with the harmonic Ritz vectors of
// the snapshot's basis, smallest |theta| first; returns the new size.
template<class Reduction = SerialReduction, class Telemetry = std::nullptr_t>
std::size_t extract_harmonic_ritz(DeviceMatrix& op, const ArnoldiSnapshot& snapshot, RecycleSpace& recycle_space,
                                  Reduction = Reduction{}, Telemetry* telemetry = nullptr)
{
    static_assert(detail::reduction_of<Reduction>() == OWTK_REDUCTION_SERIAL,
                  "the device harmonic-Ritz extraction reduces with owtk::hip::SerialReduction");
    owtk_solver_result t{};
    std::uint64_t size = 0;
    check(owtk_extract_harmonic_ritz(op.get(), snapshot.get(), recycle_space.get(), &t, &size));
    if (telemetry) {
        telemetry->operator_applications += t.operator_applications;
        telemetry->global_reductions += t.global_reductions;
    }
    return static_cast<std::size_t>(size);
}

// gcrodr(op, b, x, space, options, pc, reduction, workspace, snapshot)
// (gcrodr.hpp:185-212): recycled_fgmres without the correction update,
// recording its last cycle, then extract_harmonic_ritz into the space.
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t, class Snapshot = std::nullptr_t>
SolverResult<double> gcrodr(DeviceMatrix& op, const DeviceVector<double>& rhs, DeviceVector<double>& solution,
                            RecycleSpace& recycle_space, const Options& options = Options{},
                            Pc&& preconditioner = Pc{}, Reduction = Reduction{}, Workspace = nullptr,
                            Snapshot supplied_snapshot = nullptr)
{
    const owtk_arnoldi_snapshot snap = detail::snapshot_handle(supplied_snapshot);
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_gcrodr(op.get(), pc, rhs.data(), solution.data(), recycle_space.get(),
                                                     &c, snap, &r);
                              });
}

// pipelined_bicgstab(op, b, x, options, pc, reduction, workspace)
// (pipelined_bicgstab.hpp:20-35)
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t>
SolverResult<double> pipelined_bicgstab(DeviceMatrix& op, const DeviceVector<double>& rhs,
                                        DeviceVector<double>& solution, const Options& options = Options{},
                                        Pc&& preconditioner = Pc{}, Reduction = Reduction{}, Workspace = nullptr)
{
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_pipelined_bicgstab(op.get(), pc, rhs.data(), solution.data(), &c, &r);
                              });
}

// communication_hiding_bicgstab (pipelined_bicgstab.hpp:317-597)
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t>
SolverResult<double> communication_hiding_bicgstab(DeviceMatrix& op, const DeviceVector<double>& rhs,
                                                   DeviceVector<double>& solution,
                                                   const Options& options = Options{}, Pc&& preconditioner = Pc{},
                                                   Reduction = Reduction{}, Workspace = nullptr)
{
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_communication_hiding_bicgstab(op.get(), pc, rhs.data(),
                                                                            solution.data(), &c, &r);
                              });
}

// idrs(op, b, x, options, pc, reduction, workspace) (idrs.hpp:128-140),
// s = options.idr_shadow_space
template<class Options = SolverOptions<double>, class Pc = IdentityPreconditioner, class Reduction = SerialReduction,
         class Workspace = std::nullptr_t>
SolverResult<double> idrs(DeviceMatrix& op, const DeviceVector<double>& rhs, DeviceVector<double>& solution,
                          const Options& options = Options{}, Pc&& preconditioner = Pc{},
                          Reduction = Reduction{}, Workspace = nullptr)
{
    const owtk_precond pc = detail::pc_handle(preconditioner);
    return detail::run_solver(op, rhs, solution, options, detail::reduction_of<Reduction>(),
                              [&](const owtk_solver_options& c, owtk_solver_result& r) {
                                  return owtk_idrs(op.get(), pc, rhs.data(), solution.data(), &c, &r);
                              });
}

// FP64 iterative refinement around an FP32/BF16 inner FGMRES (no reference
// counterpart): the outer tolerances and limits come from `options` (a
// SolverOptions), the inner solve from `inner` (owtk_refinement_options
// without its outer part); `fallback` (may be null) takes over when a
// refinement step with `preconditioner` fails.
template<class Options, class Pc>
SolverResult<double> fgmres_refined(DeviceMatrix& op, const DeviceVector<double>& rhs,
                                    DeviceVector<double>& solution, const Options& options, Pc&& preconditioner,
                                    owtk_refinement_options inner, const Preconditioner* fallback = nullptr)
{
    if (rhs.size() < op.owned_size() || solution.size() < op.local_size())
        throw std::invalid_argument("solver vector layout mismatch");
    inner.outer = detail::to_c(options);
    const owtk_precond pc = detail::pc_handle(preconditioner);
    owtk_solver_result r;
    if (fallback)
        check(owtk_fgmres_refined_fallback(op.get(), pc, fallback->get(), rhs.data(), solution.data(), &inner, &r));
    else
        check(owtk_fgmres_refined(op.get(), pc, rhs.data(), solution.data(), &inner, &r));
    return from_c(r);
}

inline owtk_refinement_options refinement_options()
{
    owtk_refinement_options o;
    owtk_refinement_options_default(&o);
    return o;
}

// Host-vector routes: the reference's bicgstab / fgmres on a matrix and
// host vectors (data(), owned_size(), local_size()), with the whole solve on
// the device.  The preconditioner is described by owtk_precond_params (built
// on the device copy of the matrix); x holds the initial guess and receives
// the solution.
template<class Matrix, class Vector, class Options>
SolverResult<double> solve_bicgstab(HipExecutionPolicy& policy, const Matrix& matrix, const Vector& b, Vector& x,
                                    const Options& options, const owtk_precond_params& pc_params)
{
    Context& ctx = policy.context();
    DeviceMatrix a(ctx, matrix);
    Preconditioner pc(a, pc_params);
    DeviceVector<double> db(ctx, a.local_size()), dx(ctx, a.local_size());
    db.upload(b.data(), a.owned_size());
    dx.upload(x.data(), a.local_size());
    SolverResult<double> r = bicgstab(a, db, dx, options, pc);
    dx.download(x.data(), a.owned_size());
    return r;
}

template<class Matrix, class Vector, class Options>
SolverResult<double> solve_fgmres(HipExecutionPolicy& policy, const Matrix& matrix, const Vector& b, Vector& x,
                                  const Options& options, const owtk_precond_params& pc_params)
{
    Context& ctx = policy.context();
    DeviceMatrix a(ctx, matrix);
    Preconditioner pc(a, pc_params);
    DeviceVector<double> db(ctx, a.local_size()), dx(ctx, a.local_size());
    db.upload(b.data(), a.owned_size());
    dx.upload(x.data(), a.local_size());
    SolverResult<double> r = fgmres(a, db, dx, options, pc);
    dx.download(x.data(), a.owned_size());
    return r;
}

} // namespace owtk::hip
