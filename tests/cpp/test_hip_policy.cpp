// Drives include/owtk/hip_policy.hpp with stand-in types shaped like the
// reference's BlockVector and SplitBlockCsrMatrixView accessors
// (core.hpp:197-395, split_block_csr.hpp:38-124): a split SpMV through
// HipExecutionPolicy::apply checked against a host loop, and a BiCGSTAB +
// Jacobi solve through solve_bicgstab checked by its true residual.
#include <owtk/hip_policy.hpp>

#include <cmath>
#include <cstdio>
#include <optional>
#include <vector>

namespace {

struct Vec {
    std::vector<double> v;
    std::size_t owned, local;
    Vec(std::size_t o, std::size_t l) : v(l, 0.0), owned(o), local(l) {}
    double* data() { return v.data(); }
    const double* data() const { return v.data(); }
    std::size_t owned_size() const { return owned; }
    std::size_t local_size() const { return local; }
};

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

struct Options {
    double relative_tolerance = 1e-10, absolute_tolerance = 0, breakdown_tolerance = 1e-14;
    std::size_t maximum_iterations = 1000, convergence_check_interval = 10, restart = 30,
                residual_replacement_interval = 50;
    int gmres_orthogonalization = 0;
    bool collect_timings = false;
};

enum class Status { converged, maximum_iterations, breakdown, diverged, invalid_input };
enum class Reason { none };
struct Result {
    Status status;
    Reason breakdown_reason;
    std::size_t iterations, operator_applications, preconditioner_applications, global_reductions,
        reorthogonalizations, arnoldi_norm_verifications;
    double initial_residual_norm, recursive_residual_norm, true_residual_norm, relative_residual_norm;
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

void host_apply(const Split& s, const Vec& x, Vec& y)
{
    for (std::size_t i = 0; i < s.n; ++i)
        for (std::size_t c = 0; c < s.b; ++c) {
            double acc = s.diag[i * s.b + c] * x.v[i * s.b + c];
            for (auto e = s.off[i]; e < s.off[i + 1]; ++e)
                acc = std::fma(s.vals[e * s.b + c], x.v[s.col[e] * s.b + c], acc);
            y.v[i * s.b + c] = acc;
        }
}

} // namespace

int main()
{
    using namespace owtk::hip;
    if (!HipExecutionPolicy::accelerator_available()) {
        std::printf("no device\n");
        return 2;
    }
    const Split s = chain(1000, 3);
    HipExecutionPolicy policy;
    Vec x(3000, 3000), y(3000, 3000), yref(3000, 3000);
    for (std::size_t i = 0; i < 3000; ++i) x.v[i] = std::sin(0.1 * i);
    policy.apply(s, x, y);
    host_apply(s, x, yref);
    for (std::size_t i = 0; i < 3000; ++i)
        if (y.v[i] != yref.v[i]) {
            std::printf("apply mismatch at %zu: %.17g vs %.17g\n", i, y.v[i], yref.v[i]);
            return 1;
        }
    // A different matrix at a reused address (smaller), then a value change
    // without refresh(): both must be seen by the next apply.
    {
        std::optional<Split> slot;
        slot.emplace(chain(700, 3));
        Vec x1(2100, 2100), y1(2100, 2100), r1(2100, 2100);
        for (std::size_t i = 0; i < 2100; ++i) x1.v[i] = std::cos(0.05 * i);
        policy.apply(*slot, x1, y1);
        const void* first = &*slot;
        slot.reset();
        slot.emplace(chain(400, 3));
        if (static_cast<const void*>(&*slot) != first) {
            std::printf("address not reused\n");
            return 1;
        }
        Vec x2(1200, 1200), y2(1200, 1200), r2(1200, 1200);
        for (std::size_t i = 0; i < 1200; ++i) x2.v[i] = std::sin(0.3 * i);
        policy.apply(*slot, x2, y2);
        host_apply(*slot, x2, r2);
        for (std::size_t i = 0; i < 1200; ++i)
            if (y2.v[i] != r2.v[i]) {
                std::printf("reused-address apply mismatch at %zu\n", i);
                return 1;
            }
        for (auto& v : slot->vals) v *= 1.5;
        slot->diag[7] = 11.0;
        policy.apply(*slot, x2, y2);
        host_apply(*slot, x2, r2);
        for (std::size_t i = 0; i < 1200; ++i)
            if (y2.v[i] != r2.v[i]) {
                std::printf("value change not seen at %zu\n", i);
                return 1;
            }
        std::printf("reused address and in-place value change ok\n");
    }
    Vec b(3000, 3000), sol(3000, 3000), r(3000, 3000);
    for (std::size_t i = 0; i < 3000; ++i) b.v[i] = std::cos(0.37 * i);
    const Result res = solve_bicgstab<Result>(policy, s, b, sol, Options{});
    host_apply(s, sol, r);
    double rn = 0, bn = 0;
    for (std::size_t i = 0; i < 3000; ++i) {
        rn += (b.v[i] - r.v[i]) * (b.v[i] - r.v[i]);
        bn += b.v[i] * b.v[i];
    }
    const double rel = std::sqrt(rn / bn);
    std::printf("policy %s: apply bitwise ok; bicgstab status %d iterations %zu true rel %.3e\n",
                HipExecutionPolicy::name(), static_cast<int>(res.status), res.iterations, rel);
    bool bad_layout_caught = false;
    try {
        Vec shortv(10, 10);
        policy.apply(s, shortv, y);
    } catch (const std::invalid_argument&) {
        bad_layout_caught = true;
    }
    if (!bad_layout_caught) {
        std::printf("layout mismatch not rejected\n");
        return 1;
    }
    // the production selection on device vectors: split SSOR + pipelined BiCGSTAB
    Context ctx(0);
    DeviceMatrix dm(ctx, s);
    auto pp = Preconditioner::params(OWTK_PC_SSOR);
    Preconditioner ssor(dm, pp);
    DeviceVector<double> db(ctx, 3000), dx(ctx, 3000);
    db.upload(b.v.data(), 3000);
    owtk_solver_options o;
    owtk_solver_options_default(&o);
    o.relative_tolerance = 1e-10;
    const owtk_solver_result pr = pipelined_bicgstab(dm, &ssor, db, dx, o);
    Vec px(3000, 3000), pr_ax(3000, 3000);
    dx.download(px.v.data(), 3000);
    host_apply(s, px, pr_ax);
    double prn = 0;
    for (std::size_t i = 0; i < 3000; ++i) prn += (b.v[i] - pr_ax.v[i]) * (b.v[i] - pr_ax.v[i]);
    const double prel = std::sqrt(prn / bn);
    std::printf("pipelined bicgstab + ssor: status %d iterations %llu true rel %.3e reductions %llu\n", pr.status,
                static_cast<unsigned long long>(pr.iterations), prel,
                static_cast<unsigned long long>(pr.global_reductions));
    if (pr.status != 0 || prel > 1e-10) return 1;
    return (res.status == Status::converged && rel <= 1e-10) ? 0 : 1;
}
