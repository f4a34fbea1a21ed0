// Drives include/owtk/hip_policy.hpp with stand-in types that carry the
// reference's names and accessors (core.hpp:24-146, 197-395;
// split_block_csr.hpp:38-124), in namespace owt::krylov as a caller's code
// would see them:
//   - HipExecutionPolicy::apply bitwise against a host loop, a different
//     matrix at a reused address and an in-place value change;
//   - device solves called unqualified in the reference's argument order
//     (krylov_solvers.hpp:159-171, 542-553; pipelined_bicgstab.hpp:20-35):
//     FGMRES + block-Jacobi ILU(0), FP64-refined FGMRES with FP32 storage,
//     BiCGSTAB with the compensated reduction, an application convergence
//     test over a host BlockVector, pipelined BiCGSTAB + SSOR, and the
//     host-vector route, GCRO-DR with its recycle space and Arnoldi
//     snapshot (gcrodr.hpp:86-212; recycling.hpp:291-310), and the learned
//     preconditioner (no reference counterpart) in refined FGMRES;
//   each checked by its true residual recomputed on the host.
#include <owtk/hip_policy.hpp>

#include <cmath>
#include <cstdio>
#include <functional>
#include <memory>
#include <optional>
#include <vector>

namespace owt::krylov {

enum class SolverStatus { converged, maximum_iterations, breakdown, diverged, invalid_input };
enum class GmresOrthogonalization {
    iterated_classical_gram_schmidt,
    adaptive_classical_gram_schmidt,
    one_synchronization_classical_gram_schmidt,
};
enum class BreakdownReason { none, arnoldi_invariant_subspace, singular_projected_system, biorthogonality_loss };

template<class T>
class BlockVector {
public:
    BlockVector(std::size_t owned, std::size_t ghost, std::size_t block)
        : owned_(owned), local_(owned + ghost), block_(block), v_((owned + ghost) * block, T(0))
    {
    }
    T* data() { return v_.data(); }
    const T* data() const { return v_.data(); }
    std::size_t owned_size() const { return owned_ * block_; }
    std::size_t local_size() const { return local_ * block_; }
    std::vector<T>& values() { return v_; }
    const std::vector<T>& values() const { return v_; }

private:
    std::size_t owned_, local_, block_;
    std::vector<T> v_;
};

template<class T>
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
    std::function<bool(std::size_t, const BlockVector<T>&)> convergence_test;
    GmresOrthogonalization gmres_orthogonalization = GmresOrthogonalization::iterated_classical_gram_schmidt;
};

template<class T>
struct SolverTimings {
    double solve_seconds = 0;
};

template<class T>
struct SolverResult {
    SolverStatus status = SolverStatus::invalid_input;
    BreakdownReason breakdown_reason = BreakdownReason::none;
    std::size_t iterations = 0, operator_applications = 0, preconditioner_applications = 0, global_reductions = 0,
                reorthogonalizations = 0, arnoldi_norm_verifications = 0;
    T initial_residual_norm = 0, recursive_residual_norm = 0, true_residual_norm = 0, relative_residual_norm = 0;
    bool converged_by_application = false;
    std::shared_ptr<SolverTimings<T>> timings;
    bool converged() const { return status == SolverStatus::converged; }
};

} // namespace owt::krylov

namespace {

using Vec = owt::krylov::BlockVector<double>;

struct Split {
    std::size_t n, b;
    std::vector<std::int64_t> off;
    std::vector<std::int64_t> col;
    std::vector<double> diag, vals;
    std::size_t owned_nodes() const { return n; }
    std::size_t ghost_nodes() const { return 0; }
    std::size_t block_size() const { return b; }
    const std::vector<std::int64_t>& off_diagonal_row_offsets() const { return off; }
    const std::vector<std::int64_t>& off_diagonal_columns() const { return col; }
    const std::vector<double>& diagonal_values() const { return diag; }
    const std::vector<double>& off_diagonal_values() const { return vals; }
};

// 1-D chain of n vertices, b components: d = 4 + c, neighbours -1 - 0.1c
Split chain(std::size_t n, std::size_t b)
{
    Split s{n, b, {0}, {}, {}, {}};
    for (std::size_t i = 0; i < n; ++i) {
        for (std::size_t c = 0; c < b; ++c) s.diag.push_back(4.0 + c);
        for (long j : {long(i) - 1, long(i) + 1}) {
            if (j < 0 || j >= long(n)) continue;
            s.col.push_back(j);
            for (std::size_t c = 0; c < b; ++c) s.vals.push_back(-1.0 - 0.1 * c - 0.05 * static_cast<double>(i % 7));
        }
        s.off.push_back(static_cast<std::int64_t>(s.col.size()));
    }
    return s;
}

void host_apply(const Split& s, const double* x, double* y)
{
    for (std::size_t i = 0; i < s.n; ++i)
        for (std::size_t c = 0; c < s.b; ++c) {
            double acc = s.diag[i * s.b + c] * x[i * s.b + c];
            for (auto e = s.off[i]; e < s.off[i + 1]; ++e)
                acc = std::fma(s.vals[e * s.b + c], x[s.col[e] * s.b + c], acc);
            y[i * s.b + c] = acc;
        }
}

double true_rel(const Split& s, const std::vector<double>& b, const std::vector<double>& x)
{
    std::vector<double> ax(b.size());
    host_apply(s, x.data(), ax.data());
    double rn = 0, bn = 0;
    for (std::size_t i = 0; i < b.size(); ++i) {
        rn += (b[i] - ax[i]) * (b[i] - ax[i]);
        bn += b[i] * b[i];
    }
    return std::sqrt(rn / bn);
}

int failures = 0;
void expect(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        ++failures;
    }
}

void check_hip_copy(owtk::hip::Context& ctx, double* host, const double* device, std::size_t count)
{
    owtk::hip::check(owtk_memcpy_dtoh(ctx.get(), host, device, count * sizeof(double)));
}

} // namespace

int main()
{
    using owtk::hip::HipExecutionPolicy;
    if (!HipExecutionPolicy::accelerator_available()) {
        std::printf("no device\n");
        return 2;
    }
    const Split s = chain(1000, 3);
    const std::size_t n = 3000;
    HipExecutionPolicy policy;
    Vec x(1000, 0, 3), y(1000, 0, 3), yref(1000, 0, 3);
    for (std::size_t i = 0; i < n; ++i) x.data()[i] = std::sin(0.1 * i);
    policy.apply(s, x, y);
    host_apply(s, x.data(), yref.data());
    for (std::size_t i = 0; i < n; ++i)
        if (y.data()[i] != yref.data()[i]) {
            std::printf("apply mismatch at %zu: %.17g vs %.17g\n", i, y.data()[i], yref.data()[i]);
            return 1;
        }
    // A different matrix at a reused address (smaller), then a value change
    // without refresh(): both must be seen by the next apply.
    {
        std::optional<Split> slot;
        slot.emplace(chain(700, 3));
        Vec x1(700, 0, 3), y1(700, 0, 3);
        for (std::size_t i = 0; i < 2100; ++i) x1.data()[i] = std::cos(0.05 * i);
        policy.apply(*slot, x1, y1);
        const void* first = &*slot;
        slot.reset();
        slot.emplace(chain(400, 3));
        if (static_cast<const void*>(&*slot) != first) {
            std::printf("address not reused\n");
            return 1;
        }
        Vec x2(400, 0, 3), y2(400, 0, 3), r2(400, 0, 3);
        for (std::size_t i = 0; i < 1200; ++i) x2.data()[i] = std::sin(0.3 * i);
        policy.apply(*slot, x2, y2);
        host_apply(*slot, x2.data(), r2.data());
        for (std::size_t i = 0; i < 1200; ++i)
            if (y2.data()[i] != r2.data()[i]) {
                std::printf("reused-address apply mismatch at %zu\n", i);
                return 1;
            }
        for (auto& v : slot->vals) v *= 1.5;
        slot->diag[7] = 11.0;
        policy.apply(*slot, x2, y2);
        host_apply(*slot, x2.data(), r2.data());
        for (std::size_t i = 0; i < 1200; ++i)
            if (y2.data()[i] != r2.data()[i]) {
                std::printf("value change not seen at %zu\n", i);
                return 1;
            }
        std::printf("reused address and in-place value change ok\n");
    }
    bool bad_layout_caught = false;
    try {
        Vec shortv(10, 0, 1);
        policy.apply(s, shortv, y);
    } catch (const std::invalid_argument&) {
        bad_layout_caught = true;
    }
    expect(bad_layout_caught, "layout mismatch rejected");
    std::printf("policy %s: apply bitwise ok\n", HipExecutionPolicy::name());

    std::vector<double> b(n);
    for (std::size_t i = 0; i < n; ++i) b[i] = std::cos(0.37 * i);

    // device solves in the reference's argument order, unqualified
    owtk::hip::Context& ctx = policy.context();
    owtk::hip::DeviceMatrix dm(ctx, s);
    auto ilu_params = owtk::hip::Preconditioner::params(OWTK_PC_ILU);
    ilu_params.block_nodes = 256;
    owtk::hip::Preconditioner ilu(dm, ilu_params);
    owtk::hip::DeviceVector<double> db(ctx, n), dx(ctx, n);
    db.upload(b.data(), n);
    std::vector<double> hx(n);
    owt::krylov::SolverOptions<double> opts;
    opts.relative_tolerance = 1e-10;
    opts.collect_timings = true;

    owt::krylov::SolverResult<double> r1 = fgmres(dm, db, dx, opts, ilu);
    dx.download(hx.data(), n);
    const double rel1 = true_rel(s, b, hx);
    std::printf("fgmres + block-Jacobi ILU(0): status %d iterations %zu true rel %.3e timed %d\n",
                static_cast<int>(r1.status), r1.iterations, rel1, r1.timings != nullptr);
    expect(r1.converged() && rel1 <= 1e-10 && r1.timings != nullptr, "fgmres + ILU(0)");

    dm.add_storage(OWTK_F32);
    auto ilu32_params = ilu_params;
    ilu32_params.storage = OWTK_F32;
    owtk::hip::Preconditioner ilu32(dm, ilu32_params);
    dx.upload(std::vector<double>(n, 0.0).data(), n);
    owt::krylov::SolverResult<double> r2 =
        fgmres_refined(dm, db, dx, opts, ilu32, owtk::hip::refinement_options());
    dx.download(hx.data(), n);
    const double rel2 = true_rel(s, b, hx);
    std::printf("fgmres_refined (FP32 storage) + ILU(0): status %d iterations %zu true rel %.3e\n",
                static_cast<int>(r2.status), r2.iterations, rel2);
    expect(r2.converged() && rel2 <= 1e-10, "refined fgmres");

    dx.upload(std::vector<double>(n, 0.0).data(), n);
    owt::krylov::SolverResult<double> r3s = bicgstab(dm, db, dx, opts, ilu);
    dx.upload(std::vector<double>(n, 0.0).data(), n);
    owt::krylov::SolverResult<double> r3 = bicgstab(dm, db, dx, opts, ilu, owtk::hip::CompensatedSerialReduction{});
    dx.download(hx.data(), n);
    const double rel3 = true_rel(s, b, hx);
    std::printf("bicgstab + ILU(0), compensated reduction: status %d iterations %zu (serial %zu) reductions %zu "
                "(serial %zu) true rel %.3e\n",
                static_cast<int>(r3.status), r3.iterations, r3s.iterations, r3.global_reductions,
                r3s.global_reductions, rel3);
    expect(r3.converged() && rel3 <= 1e-10, "compensated bicgstab");
    expect(r3.global_reductions == r3s.global_reductions || r3.iterations != r3s.iterations,
           "compensated bicgstab counts the same reductions");

    // application convergence test over a host BlockVector: stop after the
    // third iterate; the test sees iterations 0, 1, 2, 3
    std::vector<std::size_t> seen;
    double seen_norm = 0;
    owt::krylov::SolverOptions<double> app = opts;
    app.convergence_test = [&](std::size_t it, const Vec& sol) {
        seen.push_back(it);
        double a = 0;
        for (std::size_t i = 0; i < sol.owned_size(); ++i) a += sol.data()[i] * sol.data()[i];
        seen_norm = std::sqrt(a);
        return it >= 3;
    };
    dx.upload(std::vector<double>(n, 0.0).data(), n);
    owt::krylov::SolverResult<double> r4 = bicgstab(dm, db, dx, app, ilu);
    dx.download(hx.data(), n);
    double xn = 0;
    for (double v : hx) xn += v * v;
    std::printf("bicgstab + application test: status %d iterations %zu by application %d, test saw %zu calls, "
                "last |x| %.6e (solution %.6e)\n",
                static_cast<int>(r4.status), r4.iterations, r4.converged_by_application, seen.size(), seen_norm,
                std::sqrt(xn));
    expect(r4.converged() && r4.converged_by_application && r4.iterations == 3, "application test stops at 3");
    expect(seen == std::vector<std::size_t>({0, 1, 2, 3}), "application test sees iterations 0..3");
    expect(std::abs(seen_norm - std::sqrt(xn)) <= 1e-12 * std::sqrt(xn), "application test sees the solution");
    expect(std::abs(r4.relative_residual_norm - true_rel(s, b, hx)) <= 1e-9 * true_rel(s, b, hx),
           "application stop reports the true residual");

    // fgmres with an application test and an Arnoldi snapshot, unqualified
    // with the reference's trailing arguments: a stop at iteration 1 keeps a
    // one-column snapshot, a stop at 0 clears it
    {
        owtk::hip::ArnoldiSnapshot snap(ctx);
        owt::krylov::SolverOptions<double> at1 = opts;
        at1.convergence_test = [](std::size_t it, const Vec&) { return it == 1; };
        dx.upload(std::vector<double>(n, 0.0).data(), n);
        owt::krylov::SolverResult<double> r7 =
            fgmres(dm, db, dx, at1, owtk::hip::IdentityPreconditioner{}, owtk::hip::SerialReduction{}, nullptr, &snap);
        std::printf("fgmres + application test + snapshot: status %d iterations %zu by application %d, "
                    "snapshot columns %zu, h(0,0) %.6e h(1,0) %.6e\n",
                    static_cast<int>(r7.status), r7.iterations, r7.converged_by_application, snap.columns(),
                    snap.h(0, 0), snap.h(1, 0));
        expect(r7.converged_by_application && r7.iterations == 1 && snap.columns() == 1,
               "application stop keeps its Arnoldi snapshot");
        // Z_0 = b / |b| for the identity preconditioner and x0 = 0
        std::vector<double> z0(n);
        check_hip_copy(ctx, z0.data(), snap.preconditioned_basis(0), n);
        double bn = 0;
        for (double v : b) bn += v * v;
        bn = std::sqrt(bn);
        double dz = 0;
        for (std::size_t i = 0; i < n; ++i) dz = std::max(dz, std::abs(z0[i] - b[i] / bn));
        expect(dz <= 1e-15, "snapshot basis column 0 is b / |b|");
        bool out_of_range = false;
        try {
            (void)snap.h(0, 1);
        } catch (const std::out_of_range&) {
            out_of_range = true;
        }
        expect(out_of_range, "snapshot entry beyond its columns is out of range");
        owt::krylov::SolverOptions<double> at0 = opts;
        at0.convergence_test = [](std::size_t, const Vec&) { return true; };
        owt::krylov::SolverResult<double> r8 =
            fgmres(dm, db, dx, at0, owtk::hip::IdentityPreconditioner{}, owtk::hip::SerialReduction{}, nullptr, &snap);
        expect(r8.converged_by_application && r8.iterations == 0 && snap.columns() == 0,
               "initial application stop clears a stale snapshot");
    }

    // GCRO-DR (gcrodr.hpp:185-212) called unqualified with the reference's
    // trailing arguments, as test_core.cpp:611-634: two harmonic Ritz
    // vectors kept with a recorded snapshot, then a related system converges
    // keeping two
    {
        owtk::hip::Preconditioner jacobi(dm, owtk::hip::Preconditioner::params(OWTK_PC_JACOBI));
        owt::krylov::SolverOptions<double> gopts;
        gopts.relative_tolerance = 1e-10;
        gopts.restart = 20;
        owtk::hip::RecycleSpace space(dm, 2);
        owtk::hip::ArnoldiSnapshot snap(ctx);
        dx.upload(std::vector<double>(n, 0.0).data(), n);
        owt::krylov::SolverResult<double> g1 =
            gcrodr(dm, db, dx, space, gopts, jacobi, owtk::hip::SerialReduction{}, nullptr, &snap);
        dx.download(hx.data(), n);
        const double relg1 = true_rel(s, b, hx);
        std::vector<double> b2(b);
        for (std::size_t i = 0; i < n; i += 97) b2[i] += 0.25;
        owtk::hip::DeviceVector<double> db2(ctx, n);
        db2.upload(b2.data(), n);
        dx.upload(std::vector<double>(n, 0.0).data(), n);
        owt::krylov::SolverResult<double> g2 = gcrodr(dm, db2, dx, space, gopts, jacobi);
        dx.download(hx.data(), n);
        const double relg2 = true_rel(s, b2, hx);
        std::printf("gcrodr + Jacobi: first %zu iterations (true rel %.3e, snapshot columns %zu), related %zu "
                    "iterations (true rel %.3e), space %zu\n",
                    g1.iterations, relg1, snap.columns(), g2.iterations, relg2, space.size());
        expect(g1.converged() && relg1 <= 1e-10 && space.size() == 2 && snap.columns() > 0,
               "GCRO-DR keeps two harmonic Ritz vectors");
        expect(g2.converged() && relg2 <= 1e-10 && space.size() == 2, "GCRO-DR on a related system");

        // recycled_fgmres recording its last cycle, then the extraction by
        // hand: the same space as gcrodr's replacement
        owtk::hip::RecycleSpace manual(dm, 2);
        owtk::hip::ArnoldiSnapshot rec(ctx);
        dx.upload(std::vector<double>(n, 0.0).data(), n);
        owt::krylov::SolverResult<double> rr = recycled_fgmres(dm, db, dx, manual, gopts, jacobi,
                                                               owtk::hip::SerialReduction{}, false, nullptr, &rec);
        const std::size_t kept = extract_harmonic_ritz(dm, rec, manual, owtk::hip::SerialReduction{}, &rr);
        expect(rr.converged() && rec.columns() == snap.columns() && kept == 2 && rr.iterations == g1.iterations
                   && rr.operator_applications == g1.operator_applications
                   && rr.global_reductions == g1.global_reductions,
               "recycled_fgmres + extract_harmonic_ritz is gcrodr");

        // an application-assembled snapshot (core.hpp:465-502): Z = e_0, the
        // diagonal-dominant chain gives one real harmonic value
        owtk::hip::ArnoldiSnapshot one(ctx);
        one.begin_cycle(dm, 1);
        const double col0[2] = {4.0, 1.0};
        one.record_column(0, col0);
        std::vector<owtk::hip::DeviceVector<double>> basis;
        basis.emplace_back(ctx, n);
        std::vector<double> e0(n, 0.0);
        e0[0] = 1.0;
        basis[0].upload(e0.data(), n);
        one.finish_cycle(basis, 1);
        owtk::hip::RecycleSpace single(dm, 1);
        owt::krylov::SolverResult<double> tel;
        const std::size_t got = extract_harmonic_ritz(dm, one, single, owtk::hip::SerialReduction{}, &tel);
        expect(got == 1 && one.columns() == 1 && one.h(1, 0) == 1.0 && tel.operator_applications == 2
                   && tel.global_reductions == 3,
               "extraction from an assembled snapshot");
        bool refused_record = false;
        try {
            one.record_column(0, std::span<const double>(col0, 1));
        } catch (const std::invalid_argument&) {
            refused_record = true;
        }
        expect(refused_record, "record_column takes column + 2 values");
    }

    // the learned preconditioner through C++: a fresh network (gamma = 1,
    // the truncated Neumann series of degree 3) in BF16 storage inside the
    // FP64-refined FP32 FGMRES, then the same network's parameters round trip
    {
        owtk::hip::GnnModel model(ctx, 3, 3);
        auto lp = owtk::hip::Preconditioner::params(OWTK_PC_LEARNED);
        lp.learned_model = model.get();
        lp.storage = OWTK_BF16;
        dm.add_storage(OWTK_BF16);
        owtk::hip::Preconditioner learned(dm, lp);
        dx.upload(std::vector<double>(n, 0.0).data(), n);
        owt::krylov::SolverResult<double> rl =
            fgmres_refined(dm, db, dx, opts, learned, owtk::hip::refinement_options());
        dx.download(hx.data(), n);
        const double rell = true_rel(s, b, hx);
        std::printf("fgmres_refined + learned (fresh network, BF16): status %d iterations %zu true rel %.3e\n",
                    static_cast<int>(rl.status), rl.iterations, rell);
        expect(rl.converged() && rell <= 1e-10, "refined fgmres + learned preconditioner");
        std::vector<float> params = model.parameters();
        params.back() = 0.5f;  // one series coefficient g: a changed model
        model.set_parameters(params);
        expect(model.parameters() == params, "network parameters round trip");
        bool wrong_size = false;
        try {
            model.set_parameters(std::vector<float>(3));
        } catch (const std::invalid_argument&) {
            wrong_size = true;
        }
        expect(wrong_size, "network parameter count is checked");
    }

    // the production selection: split SSOR + pipelined BiCGSTAB
    owtk::hip::Preconditioner ssor(dm, owtk::hip::Preconditioner::params(OWTK_PC_SSOR));
    dx.upload(std::vector<double>(n, 0.0).data(), n);
    owt::krylov::SolverResult<double> r5 = pipelined_bicgstab(dm, db, dx, opts, ssor);
    dx.download(hx.data(), n);
    const double rel5 = true_rel(s, b, hx);
    std::printf("pipelined bicgstab + ssor: status %d iterations %zu true rel %.3e reductions %zu\n",
                static_cast<int>(r5.status), r5.iterations, rel5, r5.global_reductions);
    expect(r5.converged() && rel5 <= 1e-10, "pipelined bicgstab + ssor");

    // a reduction or test the device solver does not implement is refused
    bool refused = false;
    try {
        (void)fgmres(dm, db, dx, opts, ilu, owtk::hip::CompensatedSerialReduction{});
    } catch (const std::runtime_error&) {
        refused = true;
    }
    expect(refused, "fgmres refuses a compensated reduction");

    // host vectors, whole solve on the device
    Vec hb(1000, 0, 3), hsol(1000, 0, 3);
    std::copy(b.begin(), b.end(), hb.data());
    owt::krylov::SolverResult<double> r6 = owtk::hip::solve_bicgstab(
        policy, s, hb, hsol, opts, owtk::hip::Preconditioner::params(OWTK_PC_JACOBI));
    const double rel6 = true_rel(s, b, hsol.values());
    std::printf("host-vector bicgstab + Jacobi: status %d iterations %zu true rel %.3e\n",
                static_cast<int>(r6.status), r6.iterations, rel6);
    expect(r6.converged() && rel6 <= 1e-10, "host-vector bicgstab");
    std::printf("%s\n", failures == 0 ? "all C++ checks passed" : "C++ checks FAILED");
    return failures == 0 ? 0 : 1;
}
