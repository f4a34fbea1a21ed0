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
#include <map>
#include <memory>
#include <mutex>
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

// Options/result conversion from and to the reference's SolverOptions /
// SolverResult (core.hpp:24-146): same field names, same enum values.
template<class Options>
owtk_solver_options to_c(const Options& o)
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
    return c;
}

template<class Result>
void from_c(const owtk_solver_result& c, Result& r)
{
    using Status = decltype(r.status);
    using Reason = decltype(r.breakdown_reason);
    r.status = static_cast<Status>(c.status);
    r.breakdown_reason = static_cast<Reason>(c.breakdown_reason);
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
}

// Device-resident solves on device vectors.
inline owtk_solver_result bicgstab(DeviceMatrix& a, const Preconditioner* pc, const DeviceVector<double>& b,
                                   DeviceVector<double>& x, const owtk_solver_options& o)
{
    owtk_solver_result r;
    check(owtk_bicgstab(a.get(), pc ? pc->get() : nullptr, b.data(), x.data(), &o, &r));
    return r;
}

// pipelined_bicgstab (pipelined_bicgstab.hpp:20-244)
inline owtk_solver_result pipelined_bicgstab(DeviceMatrix& a, const Preconditioner* pc, const DeviceVector<double>& b,
                                             DeviceVector<double>& x, const owtk_solver_options& o)
{
    owtk_solver_result r;
    check(owtk_pipelined_bicgstab(a.get(), pc ? pc->get() : nullptr, b.data(), x.data(), &o, &r));
    return r;
}

// communication_hiding_bicgstab (pipelined_bicgstab.hpp:317-597)
inline owtk_solver_result communication_hiding_bicgstab(DeviceMatrix& a, const Preconditioner* pc,
                                                        const DeviceVector<double>& b, DeviceVector<double>& x,
                                                        const owtk_solver_options& o)
{
    owtk_solver_result r;
    check(owtk_communication_hiding_bicgstab(a.get(), pc ? pc->get() : nullptr, b.data(), x.data(), &o, &r));
    return r;
}

// idrs (idrs.hpp:128-505), s = o.idr_shadow_space
inline owtk_solver_result idrs(DeviceMatrix& a, const Preconditioner* pc, const DeviceVector<double>& b,
                               DeviceVector<double>& x, const owtk_solver_options& o)
{
    owtk_solver_result r;
    check(owtk_idrs(a.get(), pc ? pc->get() : nullptr, b.data(), x.data(), &o, &r));
    return r;
}

inline owtk_solver_result fgmres(DeviceMatrix& a, const Preconditioner* pc, const DeviceVector<double>& b,
                                 DeviceVector<double>& x, const owtk_solver_options& o)
{
    owtk_solver_result r;
    check(owtk_fgmres(a.get(), pc ? pc->get() : nullptr, b.data(), x.data(), &o, &r));
    return r;
}

inline owtk_solver_result fgmres_refined(DeviceMatrix& a, const Preconditioner* pc, const DeviceVector<double>& b,
                                         DeviceVector<double>& x, const owtk_refinement_options& o,
                                         const Preconditioner* fallback = nullptr)
{
    owtk_solver_result r;
    if (fallback)
        check(owtk_fgmres_refined_fallback(a.get(), pc ? pc->get() : nullptr, fallback->get(), b.data(), x.data(), &o, &r));
    else
        check(owtk_fgmres_refined(a.get(), pc ? pc->get() : nullptr, b.data(), x.data(), &o, &r));
    return r;
}

// Host-vector convenience: the reference's bicgstab(op, b, x, options, pc)
// shape for a matrix solved on the device.  b and x are host vectors with
// data()/owned_size()/local_size(); x holds the initial guess and receives
// the solution; Options/Result are the reference's SolverOptions /
// SolverResult (or any type with the same fields).
template<class Result, class Matrix, class Vector, class Options>
Result solve_bicgstab(HipExecutionPolicy& policy, const Matrix& matrix, const Vector& b, Vector& x,
                      const Options& options, owtk_precond_kind kind = OWTK_PC_JACOBI, std::int64_t block_nodes = 0)
{
    Context& ctx = policy.context();
    DeviceMatrix a(ctx, matrix);
    auto p = Preconditioner::params(kind);
    p.block_nodes = block_nodes;
    Preconditioner pc(a, p);
    DeviceVector<double> db(ctx, a.local_size()), dx(ctx, a.local_size());
    db.upload(b.data(), a.owned_size());
    dx.upload(x.data(), a.local_size());
    const owtk_solver_result c = bicgstab(a, &pc, db, dx, to_c(options));
    dx.download(x.data(), a.owned_size());
    Result r{};
    from_c(c, r);
    return r;
}

} // namespace owtk::hip
