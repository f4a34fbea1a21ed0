// Device IDR(s) (idrs, idrs.hpp:128-505): right-preconditioned IDR(s) with
// the Sonneveld-space recurrence.  The n-vectors (residual, the s residual
// and solution directions G and Z, the s shadow vectors P) live on the GPU;
// the s x s moment matrix, the small dense solves, the breakdown and
// restart logic run on the host, which reads each reduction as the
// reference's serial reduction would return it.  Vector updates that the
// reference performs as a sequence of axpys over one output run as one pass
// with the same per-element operation order.  The shadow vectors are the
// reference's draws (std::mt19937_64 + std::normal_distribution<double>,
// seed 42 + restart * 0x9e3779b97f4a7c15), orthonormalised on the device;
// the draws of a size and restart are generated once per context.  IDR(1)
// is the BiCGSTAB recurrence, as in the reference.
#include "precond.hpp"
#include "solver_common.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <random>
#include <vector>

using namespace owtk;

namespace owtk {
namespace {

using dev::kBlock;
constexpr int kMaxShadow = 16;

struct Combo {
    const double* v[kMaxShadow];
    double c[kMaxShadow];
    int m;
};

// out = in (times `scale` when `scaled`), then out = fma(c_j, v_j, out) for
// j = 0..m-1 in order: copy/scale followed by m axpys, element by element.
// in may alias out.
__global__ __launch_bounds__(256) void k_idr_combine(double* out, const double* in, double scale, int scaled, Combo cb,
                                                     std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        double v = in[i];
        if (scaled) v = v * scale;
        for (int j = 0; j < cb.m; ++j) v = __builtin_fma(cb.c[j], cb.v[j][i], v);
        out[i] = v;
    }
}

// y1 = fma(a, x1, y1); y2 = fma(a, x2, y2)
__global__ __launch_bounds__(256) void k_idr_axpy2(double a, const double* __restrict__ x1, double* __restrict__ y1,
                                                   const double* __restrict__ x2, double* __restrict__ y2,
                                                   std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        y1[i] = __builtin_fma(a, x1[i], y1[i]);
        y2[i] = __builtin_fma(a, x2[i], y2[i]);
    }
}

// solve_dense_system (idrs.hpp:34-77): partial pivoting, false on a pivot
// that is non-finite or at most tol
bool solve_dense(std::vector<double> a, std::vector<double> rhs, std::vector<double>& out, double tol)
{
    const std::size_t n = rhs.size();
    out.assign(n, 0.0);
    for (std::size_t col = 0; col < n; ++col) {
        std::size_t piv = col;
        for (std::size_t row = col + 1; row < n; ++row)
            if (std::abs(a[row * n + col]) > std::abs(a[piv * n + col])) piv = row;
        if (!std::isfinite(a[piv * n + col]) || std::abs(a[piv * n + col]) <= tol) return false;
        if (piv != col) {
            for (std::size_t e = col; e < n; ++e) std::swap(a[col * n + e], a[piv * n + e]);
            std::swap(rhs[col], rhs[piv]);
        }
        for (std::size_t row = col + 1; row < n; ++row) {
            const double f = a[row * n + col] / a[col * n + col];
            for (std::size_t e = col; e < n; ++e) a[row * n + e] -= f * a[col * n + e];
            rhs[row] -= f * rhs[col];
        }
    }
    for (std::size_t k = n; k-- > 0;) {
        double v = rhs[k];
        for (std::size_t col = k + 1; col < n; ++col) v -= a[k * n + col] * out[col];
        out[k] = v / a[k * n + k];
    }
    return true;
}

} // namespace

int idrs_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, const owtk_solver_options* o,
             owtk_solver_result* res)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    const std::int64_t n = m->owned * m->bs;
    const std::uint64_t s64 = o->idr_shadow_space;
    if (!valid_options(o) || s64 == 0 || s64 > static_cast<std::uint64_t>(n)) return OWTK_OK;
    if (s64 == 1) return bicgstab_run(m, pc, b, x, o, res);  // IDR(1) is BiCGSTAB (idrs.hpp:165-172)
    if (s64 > static_cast<std::uint64_t>(kMaxShadow)) throw unsupported_error("IDR(s) on the device supports s <= 16");
    const int s = static_cast<int>(s64);
    if (pc && pc->m->owned * pc->m->bs != n) throw std::invalid_argument("preconditioner layout mismatch");
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage != OWTK_F64)
        throw std::invalid_argument("FP64 solve needs an FP64 preconditioner");
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const std::int64_t L = m->local() * m->bs;
    DeviceArray<double> vecs;
    vecs.allocate(static_cast<std::size_t>((4 + 3 * s) * L));
    kern::fill_zero_bytes(st, vecs.get(), vecs.bytes());
    auto vec = [&](int k) { return vecs.get() + static_cast<std::int64_t>(k) * L; };
    double* r = vec(0);
    double* v = vec(1);   // projected residual
    double* z = vec(2);   // preconditioned / solution direction
    double* w = vec(3);   // operator output
    auto G = [&](int k) { return vec(4 + k); };
    auto Z = [&](int k) { return vec(4 + s + k); };
    auto P = [&](int k) { return vec(4 + 2 * s + k); };
    DeviceArray<double> dsc;
    dsc.allocate(kMaxShadow + 2);
    ctx->pinned.ensure((kMaxShadow + 2) * sizeof(double));
    auto* hsc = static_cast<double*>(ctx->pinned.ptr);
    std::uint64_t ops = 0, pcs = 0, reds = 0;
    const unsigned int g = dev::stream_grid(n, 4);

    auto read = [&](int count) {
        OWTK_HIP(hipMemcpyAsync(hsc, dsc.get(), count * sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
    };
    auto dotv = [&](const double* a, const double* bb) {
        kern::dot_to(st, scratch, a, bb, n, dsc.get());
        read(1);
        ++reds;
        return hsc[0];
    };
    auto nrmv = [&](const double* a) {
        kern::norm_to(st, scratch, a, n, dsc.get());
        read(1);
        ++reds;
        return hsc[0];
    };
    // out[j] = <P_{first+j}, a>, j < count, one pass; count reductions as the reference's dots
    auto pdots = [&](int first, int count, const double* a, double* out) {
        const double* vs[kMaxShadow];
        for (int j = 0; j < count; ++j) vs[j] = P(first + j);
        kern::multidot_to(st, scratch, n, count, vs, a, dsc.get());
        read(count);
        for (int j = 0; j < count; ++j) out[j] = hsc[j];
        reds += static_cast<std::uint64_t>(count);
    };
    // out = in (scaled) + sum_j coef_j D_{first+j}, D = Z (use_z) or G
    auto combine = [&](double* out, const double* in, double scale, bool scaled, const double* coef, int first,
                       int count, bool use_z) {
        Combo cb{};
        cb.m = count;
        for (int j = 0; j < count; ++j) {
            cb.v[j] = use_z ? Z(first + j) : G(first + j);
            cb.c[j] = coef[j];
        }
        k_idr_combine<<<g, kBlock, 0, st>>>(out, in, scale, scaled ? 1 : 0, cb, n);
        OWTK_LAUNCH_CHECK();
    };
    auto apply_op = [&](const double* in, double* out) {
        dev::EpiStore<double> e;
        e.y = out;
        spmv_f64(m, in, e);
        ++ops;
    };
    auto apply_pc = [&](const double* in, double* out) {
        apply_pc_f64(pc, in, out, n, st, nullptr);
        ++pcs;
    };
    // r = b - A x and its norm: true_residual followed by norm (one op, one reduction)
    auto true_residual_norm = [&] {
        dev::EpiResidual e;
        e.y = r;
        e.b = b;
        e.n = n;
        e.norm_out = dsc.get();
        spmv_f64(m, x, e);
        read(1);
        ++ops;
        ++reds;
        return hsc[0];
    };
    auto finish = [&](int status) {
        res->status = status;
        res->operator_applications = ops;
        res->preconditioner_applications = pcs;
        res->global_reductions = reds;
        if (o->collect_timings)
            res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return OWTK_OK;
    };
    auto breakdown = [&](int reason) {
        res->breakdown_reason = reason;
        return finish(status_breakdown);
    };

    // true_residual, norm(rhs), norm(residual): one operator application, two reductions
    kern::norm_to(st, scratch, b, n, dsc.get() + 1);
    {
        dev::EpiResidual e;
        e.y = r;
        e.b = b;
        e.n = n;
        e.norm_out = dsc.get();
        spmv_f64(m, x, e);
    }
    read(2);
    ops += 1;
    reds += 2;
    double rnorm = hsc[0];
    const double bnorm = hsc[1];
    const double thr = threshold_of(bnorm, o);
    res->initial_residual_norm = rnorm;
    set_residual(res, rnorm, rnorm, bnorm);
    if (!std::isfinite(rnorm) || !std::isfinite(thr)) return breakdown(reason_non_finite_scalar);
    if (rnorm <= thr) return finish(status_converged);

    std::vector<double> M(static_cast<std::size_t>(s) * s, 0.0), f(s, 0.0), am, arhs, c;
    std::uint64_t epoch = 0;
    int restarts = 0;
    double omega = 1.0;
    // The raw draws of one restart epoch: s x n normals from the seeded
    // generator, on the host (the sequence is serial), then kept in the
    // context for later solves of the same size (up to kDrawCacheBytes).
    constexpr std::size_t kDrawCacheBytes = std::size_t(4) << 30;
    auto draws_of = [&](std::uint64_t e) -> const double* {
        const auto key = std::make_tuple(n, s, e);
        auto it = ctx->idr_draws.find(key);
        if (it != ctx->idr_draws.end()) return it->second.get();
        std::vector<double> h(static_cast<std::size_t>(s) * n);
        std::mt19937_64 gen(42ULL + e * 0x9e3779b97f4a7c15ULL);
        std::normal_distribution<double> dist(0.0, 1.0);
        for (auto& d : h) d = dist(gen);
        const std::size_t bytes = h.size() * sizeof(double);
        if (ctx->idr_draw_bytes + bytes > kDrawCacheBytes) {
            ctx->sync();  // earlier copies out of the cached draws have completed
            ctx->idr_draws.clear();
            ctx->idr_draw_bytes = 0;
        }
        DeviceArray<double> d(h.size());
        OWTK_HIP(hipMemcpy(d.get(), h.data(), bytes, hipMemcpyHostToDevice));
        ctx->idr_draw_bytes += bytes;
        return ctx->idr_draws.emplace(key, std::move(d)).first->second.get();
    };
    // initialize_idr_shadow (idrs.hpp:79-118), rank 0 of a serial reduction
    auto shadow = [&]() -> bool {
        const double* drawn = draws_of(epoch++);
        for (int k = 0; k < s; ++k) {
            OWTK_HIP(hipMemcpyAsync(P(k), drawn + static_cast<std::int64_t>(k) * n, n * sizeof(double),
                                    hipMemcpyDeviceToDevice, st));
            for (int p = 0; p < k; ++p) {
                const double proj = dotv(P(p), P(k));
                kern::axpy(st, n, -proj, P(p), P(k));
            }
            const double nv = nrmv(P(k));
            if (nv <= o->breakdown_tolerance) return false;
            Combo none{};
            k_idr_combine<<<g, kBlock, 0, st>>>(P(k), P(k), 1.0 / nv, 1, none, n);
            OWTK_LAUNCH_CHECK();
        }
        return true;
    };
    auto restart_basis = [&]() -> bool {
        if (!shadow()) return false;
        std::fill(M.begin(), M.end(), 0.0);
        for (int k = 0; k < s; ++k) M[k * s + k] = 1.0;
        kern::fill_zero_bytes(st, G(0), static_cast<std::size_t>(2 * s * L) * sizeof(double));
        pdots(0, s, r, f.data());
        omega = 1.0;
        return true;
    };
    auto restart_or_fail = [&]() -> bool {
        ++restarts;
        return restarts <= 4 && restart_basis();
    };
    auto pivot_small = [&](int col) -> bool {
        double scale = 0.0;
        for (int row = col; row < s; ++row) scale = std::max(scale, std::abs(M[row * s + col]));
        const double piv = M[col * s + col];
        const double tol = std::max(std::numeric_limits<double>::min(), o->breakdown_tolerance * scale);
        return !std::isfinite(piv) || std::abs(piv) <= tol;
    };
    auto iterations = [&] { return res->iterations; };
    if (!restart_basis()) return breakdown(reason_idr_shadow_space);

    const std::uint64_t maxit = o->maximum_iterations;
    const std::uint64_t replace = o->residual_replacement_interval;
    while (iterations() < maxit) {
        bool restarted = false;
        for (int k = 0; k < s && iterations() < maxit; ++k) {
            const int act = s - k;
            am.assign(static_cast<std::size_t>(act) * act, 0.0);
            arhs.assign(act, 0.0);
            double mscale = 0.0;
            for (int row = 0; row < act; ++row) {
                arhs[row] = f[k + row];
                for (int col = 0; col < act; ++col) {
                    const double val = M[(k + row) * s + (k + col)];
                    am[row * act + col] = val;
                    mscale = std::max(mscale, std::abs(val));
                }
            }
            const double dtol =
                std::max(std::numeric_limits<double>::min(), 64 * std::numeric_limits<double>::epsilon() * mscale);
            if (!solve_dense(am, arhs, c, dtol)) {
                if (!restart_or_fail()) return breakdown(reason_idr_small_system);
                restarted = true;
                break;
            }
            // v = r - G_k:s c; z = omega M^-1 v + Z_k:s c
            std::vector<double> negc(act);
            for (int j = 0; j < act; ++j) negc[j] = -c[j];
            combine(v, r, 1.0, false, negc.data(), k, act, false);
            apply_pc(v, z);
            combine(z, z, omega, true, c.data(), k, act, true);
            apply_op(z, w);
            for (int prev = 0; prev < k; ++prev) {
                const double piv = M[prev * s + prev];
                if (pivot_small(prev)) {
                    if (!restart_or_fail()) return breakdown(reason_idr_small_system);
                    restarted = true;
                    break;
                }
                const double alpha = dotv(P(prev), w) / piv;
                k_idr_axpy2<<<g, kBlock, 0, st>>>(-alpha, G(prev), w, Z(prev), z, n);
                OWTK_LAUNCH_CHECK();
            }
            if (restarted) break;
            for (int row = 0; row < k; ++row) M[row * s + k] = 0.0;
            {
                double col[kMaxShadow];
                pdots(k, s - k, w, col);
                for (int row = k; row < s; ++row) M[row * s + k] = col[row - k];
            }
            const double piv = M[k * s + k];
            if (pivot_small(k)) {
                if (!restart_or_fail()) return breakdown(reason_idr_small_system);
                restarted = true;
                break;
            }
            OWTK_HIP(hipMemcpyAsync(G(k), w, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            OWTK_HIP(hipMemcpyAsync(Z(k), z, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            const double beta = f[k] / piv;
            kern::axpy(st, n, beta, Z(k), x);
            kern::axpy(st, n, -beta, G(k), r);
            for (int row = k + 1; row < s; ++row) f[row] -= beta * M[row * s + k];
            f[k] = 0.0;
            ++res->iterations;
            restarts = 0;
            rnorm = nrmv(r);
            res->recursive_residual_norm = rnorm;
            if (!std::isfinite(rnorm)) return breakdown(reason_non_finite_scalar);
            if (rnorm <= thr) {
                const double tn = true_residual_norm();
                set_residual(res, rnorm, tn, bnorm);
                if (tn <= thr) return finish(status_converged);
                rnorm = tn;
                if (!restart_or_fail()) return finish(status_diverged);
                restarted = true;
                break;
            }
            if (replace > 0 && res->iterations % replace == 0) {
                rnorm = true_residual_norm();
                set_residual(res, rnorm, rnorm, bnorm);
                if (rnorm <= thr) return finish(status_converged);
                if (!restart_basis()) return breakdown(reason_idr_shadow_space);
                restarted = true;
                break;
            }
        }
        if (restarted) continue;
        if (iterations() >= maxit) break;
        // stabilising minimum-residual step with the angle guard (idrs.hpp:428-446)
        apply_pc(r, z);
        apply_op(z, w);
        {
            const double* vs[1] = {r};
            kern::multidot_to(st, scratch, n, 1, vs, w, dsc.get());  // <w, r>, <w, w>
            read(2);
            reds += 2;
        }
        const double num = hsc[0], den = hsc[1];
        if (!std::isfinite(den) || den <= std::numeric_limits<double>::min())
            return breakdown(reason_omega_denominator);
        omega = num / den;
        const double adn = std::sqrt(std::max(0.0, den)) * std::max(rnorm, std::numeric_limits<double>::min());
        const double angle = adn > 0 ? std::abs(num) / adn : 1.0;
        constexpr double min_angle = 0.7;
        if (angle > 0 && angle < min_angle) omega *= min_angle / angle;
        if (!std::isfinite(omega) || std::abs(omega) <= std::numeric_limits<double>::min())
            return breakdown(reason_omega_denominator);
        kern::axpy(st, n, omega, z, x);
        kern::axpy(st, n, -omega, w, r);
        ++res->iterations;
        restarts = 0;
        rnorm = nrmv(r);
        res->recursive_residual_norm = rnorm;
        if (!std::isfinite(rnorm)) return breakdown(reason_non_finite_scalar);
        if (rnorm <= thr) {
            const double tn = true_residual_norm();
            set_residual(res, rnorm, tn, bnorm);
            if (tn <= thr) return finish(status_converged);
            rnorm = tn;
            if (!restart_or_fail()) return finish(status_diverged);
            continue;
        }
        if (replace > 0 && res->iterations % replace == 0) {
            rnorm = true_residual_norm();
            set_residual(res, rnorm, rnorm, bnorm);
            if (rnorm <= thr) return finish(status_converged);
            if (!restart_basis()) return breakdown(reason_idr_shadow_space);
            continue;
        }
        pdots(0, s, r, f.data());
    }
    const double tn = true_residual_norm();
    set_residual(res, rnorm, tn, bnorm);
    return finish(status_maximum_iterations);
}

} // namespace owtk

extern "C" int owtk_idrs(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                         const owtk_solver_options* options, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        if (options->idr_shadow_space != 1) require_plain_options(options, "idrs");  // IDR(1) runs bicgstab
        m->ctx->activate();
        idrs_run(m, pc, b, x, options, result);
    });
}
