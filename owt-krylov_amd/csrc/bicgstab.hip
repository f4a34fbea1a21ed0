// Device-resident BiCGSTAB (bicgstab, krylov_solvers.hpp:542-745).
//
// Every vector and scalar of the recurrence stays on the GPU.  The seven
// launches of an iteration carry their reductions in the same pass:
//   p   = r + beta (p - omega v)                      (beta from the state)
//   z   = M^-1 p
//   v   = A z            + <shadow, v>   -> alpha
//   s   = r - alpha v    + ||s||         -> early-exit test
//   y   = M^-1 s
//   t   = A y            + <t, s>, <t,t> -> omega
//   x  += alpha z; x += omega y; r = s - omega t + <shadow, r>, ||r||
//                                        -> next rho, cadence test
// Scalar logic (breakdown guards, cadence, counters) runs in the one-block
// finisher launched after each reducing kernel.  Once a stop condition is
// recorded every later
// launch returns immediately, so the device state equals the reference's at
// the point it would have returned.  The host synchronises only at the
// reference's convergence-check/replacement cadence (every
// convergence_check_interval iterations by default) and on stops.
#include "precond.hpp"
#include "solver_common.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"

#include <chrono>
#include <cmath>
#include <limits>

using namespace owtk;

namespace owtk {
namespace {

using dev::kBlock;
using dev::ReduceScratch;

enum Stop : int { stop_none = 0, stop_breakdown = 1, stop_early = 2, stop_cadence = 3, stop_max = 4 };

struct BicgState {
    double rho, rho_prev, alpha, omega, beta;
    double threshold, tiny, breakdown_tolerance;
    double snorm, rnorm, recursive;
    double norm_out;
    long long iteration, max_iterations, check_interval, replace_interval;
    int first;
    int stop;
    int stage;
    int reason;
    long long stop_iteration;
    long long extra_reductions;  // pipelined: fused squared norms that fell back to a scaled norm
    int app_test;                // an application convergence test replaces the cadence test
};

// norm_from_squared (krylov_solvers.hpp:95-108) charges one more reduction
// when the fused squared norm is zero, underflows or is not finite
__device__ __host__ inline bool norm_fallback(double squared)
{
    return !(__builtin_isfinite(squared) && squared >= 2.2250738585072014e-308);
}

__device__ void fin_rho(BicgState* st, double rho, long long it)
{
    if (!__builtin_isfinite(rho) || __builtin_fabs(rho) <= st->tiny) {
        st->stop = stop_breakdown;
        st->reason = __builtin_isfinite(rho) ? reason_biorthogonality_loss : reason_non_finite_scalar;
        st->stage = 0;
        st->stop_iteration = it;
        return;
    }
    if (!st->first) {
        if (__builtin_fabs(st->omega) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = reason_omega_zero;
            st->stage = 0;
            st->stop_iteration = it;
            return;
        }
        st->beta = (rho / st->rho_prev) * (st->alpha / st->omega);
    }
    st->rho = rho;
}

// <shadow, r> at (re)start.  CMP: compensated (double-double) accumulation.
template<bool CMP>
__global__ __launch_bounds__(256) void k_rho(const double* __restrict__ shadow, const double* __restrict__ r,
                                             std::int64_t n, BicgState* st, ReduceScratch scratch)
{
    double red[1] = {0.0}, lo[1] = {0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        if constexpr (CMP) dev::dd_fma(red[0], lo[0], shadow[i], r[i]);
        else red[0] = __builtin_fma(shadow[i], r[i], red[0]);
    }
    if constexpr (CMP) dev::store_partials_dd<1>(red, lo, 0, scratch);
    else dev::store_partials<1>(red, 0, scratch);
}

struct FinRho {
    BicgState* st;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x == 0) fin_rho(st, t[0], st->iteration);
    }
};

// p = r (first) or p = r + beta (p - omega v)
__global__ __launch_bounds__(256) void k_p_update(double* __restrict__ p, const double* __restrict__ r,
                                                  const double* __restrict__ v, std::int64_t n, const BicgState* st)
{
    if (st->stop) return;
    const bool first = st->first != 0;
    const double beta = st->beta, omega = st->omega;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        p[i] = first ? r[i] : __builtin_fma(beta, __builtin_fma(-omega, v[i], p[i]), r[i]);
}

// s = r - alpha v, ||s||
template<bool CMP>
__global__ __launch_bounds__(256) void k_s_update(double* __restrict__ s, const double* __restrict__ r,
                                                  const double* __restrict__ v, std::int64_t n, BicgState* st,
                                                  ReduceScratch scratch)
{
    if (st->stop) return;
    const double alpha = st->alpha;
    double red[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        const double si = __builtin_fma(-alpha, v[i], r[i]);
        s[i] = si;
        if constexpr (CMP) dev::dd_fma(red[0], lo[0], si, si);
        else red[0] = __builtin_fma(si, si, red[0]);
        red[1] = dev::combine<dev::Combine::max>(red[1], __builtin_fabs(si));
    }
    if constexpr (CMP) dev::store_partials_dd<2>(red, lo, 2, scratch);
    else dev::store_partials<2>(red, 2, scratch);
}

struct FinSnorm {
    BicgState* st;
    const double* s;
    std::int64_t n;
    __device__ void operator()(const double* t) const
    {
        const double nrm = dev::block_norm(t[0], t[1], s, n);
        if (threadIdx.x == 0) {
            st->snorm = nrm;
            if (nrm <= st->threshold) {
                st->stop = stop_early;
                st->stop_iteration = st->iteration;
            }
        }
    }
};

// x += alpha z; x += omega y; r = s - omega t; <shadow, r>, ||r||
template<bool CMP>
__global__ __launch_bounds__(256) void k_xr_update(double* __restrict__ x, double* __restrict__ r,
                                                   const double* __restrict__ z, const double* __restrict__ y,
                                                   const double* __restrict__ s, const double* __restrict__ t,
                                                   const double* __restrict__ shadow, std::int64_t n, BicgState* st,
                                                   ReduceScratch scratch)
{
    if (st->stop) return;
    const double alpha = st->alpha, omega = st->omega;
    double red[3] = {0.0, 0.0, 0.0}, lo[3] = {0.0, 0.0, 0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        double xi = __builtin_fma(alpha, z[i], x[i]);
        xi = __builtin_fma(omega, y[i], xi);
        x[i] = xi;
        const double ri = __builtin_fma(-omega, t[i], s[i]);
        r[i] = ri;
        if constexpr (CMP) {
            dev::dd_fma(red[0], lo[0], shadow[i], ri);
            dev::dd_fma(red[1], lo[1], ri, ri);
        } else {
            red[0] = __builtin_fma(shadow[i], ri, red[0]);
            red[1] = __builtin_fma(ri, ri, red[1]);
        }
        red[2] = dev::combine<dev::Combine::max>(red[2], __builtin_fabs(ri));
    }
    if constexpr (CMP) dev::store_partials_dd<3>(red, lo, 4, scratch);
    else dev::store_partials<3>(red, 4, scratch);
}

// Compensated-mode reductions that the FP64 path fuses into SpMV epilogues:
// the SpMV stores its output, then one of these passes reduces it.
//   k_dot_dd<1>: <a, w>;  k_dot_dd<2>: <a, w>, <a, a>
template<int K>
__global__ __launch_bounds__(256) void k_dot_dd(const double* __restrict__ a, const double* __restrict__ w,
                                                std::int64_t n, const int* stop, ReduceScratch scratch)
{
    if (stop != nullptr && *stop != 0) return;
    double hi[K], lo[K];
#pragma unroll
    for (int k = 0; k < K; ++k) hi[k] = lo[k] = 0.0;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        const double ai = a[i];
        dev::dd_fma(hi[0], lo[0], ai, w[i]);
        if constexpr (K > 1) dev::dd_fma(hi[1], lo[1], ai, ai);
    }
    dev::store_partials_dd<K>(hi, lo, 0, scratch);
}

// r = b - r (r holds A x), compensated sum of squares and max |r|; with b
// null: sum of squares and max of r itself (a norm).
__global__ __launch_bounds__(256) void k_residual_dd(double* __restrict__ r, const double* __restrict__ b,
                                                     std::int64_t n, ReduceScratch scratch)
{
    double hi[2] = {0.0, 0.0}, lo[2] = {0.0, 0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        double ri = r[i];
        if (b != nullptr) {
            ri = b[i] - ri;
            r[i] = ri;
        }
        dev::dd_fma(hi[0], lo[0], ri, ri);
        hi[1] = dev::combine<dev::Combine::max>(hi[1], __builtin_fabs(ri));
    }
    dev::store_partials_dd<2>(hi, lo, 2, scratch);
}

struct FinNormTo {
    const double* v;
    std::int64_t n;
    double* out;
    __device__ void operator()(const double* t) const
    {
        const double nrm = dev::block_norm(t[0], t[1], v, n);
        if (threadIdx.x == 0) *out = nrm;
    }
};

template<class Fin>
struct FinDot1 {
    Fin fin;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x == 0) fin(t[0]);
    }
};

template<class Fin>
struct FinDot2 {
    Fin fin;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x == 0) fin(t[0], t[1]);
    }
};

struct FinXr {
    BicgState* st;
    const double* r;
    std::int64_t n;
    __device__ void operator()(const double* tot) const
    {
        const long long it = st->iteration;
        const bool replace = st->replace_interval > 0 && it % st->replace_interval == 0;
        const bool check = !st->app_test && it % st->check_interval == 0;
        double rn = 0.0;
        if (replace || check) rn = dev::block_norm(tot[1], tot[2], r, n);
        __syncthreads();  // every thread has read the state before thread 0 updates it
        if (threadIdx.x != 0) return;
        if (replace || check) {
            st->rnorm = rn;
            if (replace || rn <= st->threshold) {
                st->stop = stop_cadence;
                st->stop_iteration = it;
                return;
            }
            st->recursive = rn;
        }
        st->rho_prev = st->rho;
        if (it >= st->max_iterations) {
            st->stop = stop_max;
            st->stop_iteration = it;
            return;
        }
        st->iteration = it + 1;
        fin_rho(st, tot[0], it + 1);
    }
};

// v = A z epilogue: <shadow, v> -> alpha
struct FinAlpha {
    BicgState* st;
    __device__ void operator()(double sv) const
    {
        if (!__builtin_isfinite(sv) || __builtin_fabs(sv) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = __builtin_isfinite(sv) ? reason_alpha_denominator : reason_non_finite_scalar;
            st->stage = 1;
            st->stop_iteration = st->iteration;
            return;
        }
        st->alpha = st->rho / sv;
        st->first = 0;
    }
};

// t = A y epilogue: <t, s>, <t, t> -> omega
struct FinOmega {
    BicgState* st;
    __device__ void operator()(double num, double den) const
    {
        if (!__builtin_isfinite(den) || __builtin_fabs(den) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = __builtin_isfinite(den) ? reason_omega_denominator : reason_non_finite_scalar;
            st->stage = 3;
            st->stop_iteration = st->iteration;
            return;
        }
        const double omega = num / den;
        st->omega = omega;
        const double angle = __builtin_fabs(num) / __builtin_sqrt(den) / st->snorm;
        if (!__builtin_isfinite(omega) || omega == 0.0 || angle <= st->breakdown_tolerance) {
            st->stop = stop_breakdown;
            st->reason = __builtin_isfinite(omega) ? reason_omega_zero : reason_non_finite_scalar;
            st->stage = 3;
            st->stop_iteration = st->iteration;
        }
    }
};

// ---------------------------------------------------------------------------
// Pipelined BiCGSTAB (pipelined_bicgstab.hpp:20-244).  Same vectors; the
// scalar products are fused as in the reference: <shadow, v>;
// <t, s>, <t, t>, <s, s> in the second SpMV's epilogue; <shadow, r>, <r, r>
// with the x/r update.  The recursive residual is tested every iteration.
enum PStop : int { pstop_replace = 5 };

// s = r - alpha v (no reduction: |s| comes with the omega products)
__global__ __launch_bounds__(256) void k_s_plain(double* __restrict__ s, const double* __restrict__ r,
                                                 const double* __restrict__ v, std::int64_t n, const BicgState* st)
{
    if (st->stop) return;
    const double alpha = st->alpha;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        s[i] = __builtin_fma(-alpha, v[i], r[i]);
}

// t = A y epilogue: <t, s>, <t, t>, <s, s>, max |s| -> |s|, early exit, omega
struct EpiOmegaP : dev::EpiNone {
    static constexpr int K = 4;
    static constexpr std::uint64_t kMaxMask = 8;
    double* y;
    const double* s;
    std::int64_t n;
    BicgState* st;
    __device__ void elem(std::int64_t i, double tv, double* red) const
    {
        y[i] = tv;
        const double si = s[i];
        red[0] = __builtin_fma(tv, si, red[0]);
        red[1] = __builtin_fma(tv, tv, red[1]);
        red[2] = __builtin_fma(si, si, red[2]);
        red[3] = dev::combine<dev::Combine::max>(red[3], __builtin_fabs(si));
    }
    template<int BS>
    __device__ void elem_row(std::int64_t i0, const double (&tv)[BS], double* red) const
    {
        double ss[BS];
        dev::load_row<BS>(s + i0, ss);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            red[0] = __builtin_fma(tv[c], ss[c], red[0]);
            red[1] = __builtin_fma(tv[c], tv[c], red[1]);
            red[2] = __builtin_fma(ss[c], ss[c], red[2]);
            red[3] = dev::combine<dev::Combine::max>(red[3], __builtin_fabs(ss[c]));
        }
        dev::store_row<BS>(y + i0, tv);
    }
    __device__ void finalize(const double* t) const
    {
        // norm_from_squared: the fused square unless it over/underflowed
        const double snorm = dev::block_norm(t[2], t[3], s, n);
        __syncthreads();
        if (threadIdx.x != 0) return;
        st->snorm = snorm;
        if (norm_fallback(t[2])) st->extra_reductions += 1;
        const double num = t[0], den = t[1];
        if (!__builtin_isfinite(snorm)) {
            st->stop = stop_breakdown;
            st->reason = reason_non_finite_scalar;
            st->stage = 3;
            st->stop_iteration = st->iteration;
            return;
        }
        if (snorm <= st->threshold) {
            st->stop = stop_early;
            st->stop_iteration = st->iteration;
            return;
        }
        if (!__builtin_isfinite(den) || __builtin_fabs(den) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = reason_omega_denominator;
            st->stage = 3;
            st->stop_iteration = st->iteration;
            return;
        }
        const double omega = num / den;
        st->omega = omega;
        const double angle = __builtin_fabs(num) / __builtin_sqrt(den) / snorm;
        if (!__builtin_isfinite(omega) || omega == 0.0 || angle <= st->breakdown_tolerance) {
            st->stop = stop_breakdown;
            st->reason = reason_omega_zero;
            st->stage = 3;
            st->stop_iteration = st->iteration;
        }
    }
};

// after x/r: next rho, recursive |r| (tested every iteration), beta, cadence
struct FinNextP {
    BicgState* st;
    const double* r;
    std::int64_t n;
    __device__ void operator()(const double* tot) const
    {
        const double rn = dev::block_norm(tot[1], tot[2], r, n);
        __syncthreads();  // every thread has read the state before thread 0 updates it
        if (threadIdx.x != 0) return;
        const long long it = st->iteration;
        const double prev = st->rho;
        st->rho = tot[0];
        if (norm_fallback(tot[1])) st->extra_reductions += 1;
        st->recursive = rn;
        st->rnorm = rn;
        if (rn <= st->threshold) {
            st->stop = stop_cadence;
            st->stop_iteration = it;
            return;
        }
        if (!__builtin_isfinite(prev) || __builtin_fabs(prev) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = reason_biorthogonality_loss;
            st->stage = 4;
            st->stop_iteration = it;
            return;
        }
        st->beta = (tot[0] / prev) * (st->alpha / st->omega);
        st->first = 0;
        if (st->replace_interval > 0 && it % st->replace_interval == 0) {
            st->stop = pstop_replace;
            st->stop_iteration = it;
            return;
        }
        if (it >= st->max_iterations) {
            st->stop = stop_max;
            st->stop_iteration = it;
            return;
        }
        st->iteration = it + 1;
    }
};

// rho = <shadow, r> at (re)start, without the BiCGSTAB start-of-iteration guard
struct FinRhoP {
    BicgState* st;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x == 0) st->rho = t[0];
    }
};

// ---------------------------------------------------------------------------
// Communication-hiding BiCGSTAB (pipelined_bicgstab.hpp:317-597).  Fifteen
// vectors; per iteration one streaming pass updates the four auxiliary
// directions and forms q, q2, y with their three products, and one pass
// updates x, r, r2, w with five products.  The preconditioner + operator
// pair that the reference overlaps with each reduction runs right after its
// finisher, on the same stream.
enum ChStop : int { chstop_init = 6 };

struct ChState {
    double rho, alpha, beta, omega;
    double threshold, tiny, breakdown_tolerance;
    double qnorm, rnorm, recursive;
    long long iteration, max_iterations, replace_interval;
    int first;
    int stop;
    int stage;  // 1: first reduction, 2: second reduction, 10/11: restart rho / <w, rp>
    int reason;
    long long stop_iteration;
    long long extra_reductions;
};

struct ChVecs {
    double *r, *rp, *s, *s2, *p2, *r2, *w, *w2, *z, *z2, *t, *v, *q, *q2, *y;
};

// restart: rp = r and <r, rp>
__global__ __launch_bounds__(256) void k_ch_rho(const double* __restrict__ r, double* __restrict__ rp, std::int64_t n,
                                                ReduceScratch scratch)
{
    double red[1] = {0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        const double ri = r[i];
        rp[i] = ri;
        red[0] = __builtin_fma(ri, ri, red[0]);
    }
    dev::store_partials<1>(red, 0, scratch);
}

struct FinChRho {
    ChState* st;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x != 0) return;
        const double rho = t[0];
        st->rho = rho;
        if (!__builtin_isfinite(rho) || __builtin_fabs(rho) <= st->tiny) {
            st->stop = chstop_init;
            st->stage = 10;
            st->stop_iteration = st->iteration;
        }
    }
};

// restart: <A M r, rp> -> alpha
struct FinChD2 {
    ChState* st;
    __device__ void operator()(double d2) const
    {
        if (!__builtin_isfinite(d2) || __builtin_fabs(d2) <= st->tiny) {
            st->stop = chstop_init;
            st->stage = 11;
            st->stop_iteration = st->iteration;
            return;
        }
        st->alpha = st->rho / d2;
        st->beta = 0.0;
        st->omega = 0.0;
        st->first = 1;
    }
};

// direction updates (or the restart copies), q, q2, y; <q,y>, <y,y>, <q,q>, max|q|
__global__ __launch_bounds__(256) void k_ch_first(ChVecs V, std::int64_t n, const ChState* st, ReduceScratch scratch)
{
    if (st->stop) return;
    const bool first = st->first != 0;
    const double alpha = st->alpha, beta = st->beta, omega = st->omega;
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        const double r2 = V.r2[i], w = V.w[i], w2 = V.w2[i], t = V.t[i];
        double p2, s, s2, z;
        if (first) {
            p2 = r2;
            s = w;
            s2 = w2;
            z = t;
        } else {
            p2 = __builtin_fma(beta, __builtin_fma(-omega, V.s2[i], V.p2[i]), r2);
            s = __builtin_fma(beta, __builtin_fma(-omega, V.z[i], V.s[i]), w);
            s2 = __builtin_fma(beta, __builtin_fma(-omega, V.z2[i], V.s2[i]), w2);
            z = __builtin_fma(beta, __builtin_fma(-omega, V.v[i], V.z[i]), t);
        }
        V.p2[i] = p2;
        V.s[i] = s;
        V.s2[i] = s2;
        V.z[i] = z;
        const double q = __builtin_fma(-alpha, s, V.r[i]);
        const double y = __builtin_fma(-alpha, z, w);
        V.q[i] = q;
        V.q2[i] = __builtin_fma(-alpha, s2, r2);
        V.y[i] = y;
        red[0] = __builtin_fma(q, y, red[0]);
        red[1] = __builtin_fma(y, y, red[1]);
        red[2] = __builtin_fma(q, q, red[2]);
        red[3] = dev::combine<dev::Combine::max>(red[3], __builtin_fabs(q));
    }
    dev::store_partials<4>(red, 8, scratch);
}

struct FinChFirst {
    ChState* st;
    const double* q;
    std::int64_t n;
    __device__ void operator()(const double* t) const
    {
        const double qn = dev::block_norm(t[2], t[3], q, n);
        __syncthreads();
        if (threadIdx.x != 0) return;
        st->qnorm = qn;
        if (norm_fallback(t[2])) st->extra_reductions += 1;
        const long long it = st->iteration;
        const double qy = t[0], yy = t[1];
        if (!__builtin_isfinite(qn)) {
            st->stop = stop_breakdown;
            st->reason = reason_non_finite_scalar;
            st->stage = 1;
            st->stop_iteration = it;
            return;
        }
        if (qn <= st->threshold) {
            st->stop = stop_early;
            st->stop_iteration = it;
            return;
        }
        if (!__builtin_isfinite(yy) || __builtin_fabs(yy) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = __builtin_isfinite(yy) ? reason_omega_denominator : reason_non_finite_scalar;
            st->stage = 1;
            st->stop_iteration = it;
            return;
        }
        const double omega = qy / yy;
        st->omega = omega;
        const double angle = __builtin_fabs(qy) / __builtin_sqrt(yy) / qn;
        if (!__builtin_isfinite(omega) || omega == 0.0 || angle <= st->breakdown_tolerance) {
            st->stop = stop_breakdown;
            st->reason = reason_omega_zero;
            st->stage = 1;
            st->stop_iteration = it;
            return;
        }
        st->first = 0;
    }
};

// x += alpha p2; x += omega q2; r, r2, w; <r,r>, max|r|, <r,rp>, <s,rp>, <w,rp>, <z,rp>
__global__ __launch_bounds__(256) void k_ch_second(ChVecs V, double* __restrict__ x, std::int64_t n,
                                                   const ChState* st, ReduceScratch scratch)
{
    if (st->stop) return;
    const double alpha = st->alpha, omega = st->omega;
    double red[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        double xi = __builtin_fma(alpha, V.p2[i], x[i]);
        x[i] = __builtin_fma(omega, V.q2[i], xi);
        const double y = V.y[i];
        const double r = __builtin_fma(-omega, y, V.q[i]);
        V.r[i] = r;
        V.r2[i] = __builtin_fma(-omega, __builtin_fma(-alpha, V.z2[i], V.w2[i]), V.q2[i]);
        const double w = __builtin_fma(-omega, __builtin_fma(-alpha, V.v[i], V.t[i]), y);
        V.w[i] = w;
        const double rp = V.rp[i];
        red[0] = __builtin_fma(r, r, red[0]);
        red[1] = dev::combine<dev::Combine::max>(red[1], __builtin_fabs(r));
        red[2] = __builtin_fma(r, rp, red[2]);
        red[3] = __builtin_fma(V.s[i], rp, red[3]);
        red[4] = __builtin_fma(w, rp, red[4]);
        red[5] = __builtin_fma(V.z[i], rp, red[5]);
    }
    dev::store_partials<6>(red, 2, scratch);
}

struct FinChSecond {
    ChState* st;
    const double* r;
    std::int64_t n;
    __device__ void operator()(const double* t) const
    {
        const double rn = dev::block_norm(t[0], t[1], r, n);
        __syncthreads();
        if (threadIdx.x != 0) return;
        if (norm_fallback(t[0])) st->extra_reductions += 1;
        const long long it = st->iteration;
        const double prev = st->rho;
        const double rho = t[2];
        st->rho = rho;
        st->rnorm = rn;
        st->recursive = rn;
        if (rn <= st->threshold) {
            st->stop = stop_cadence;
            st->stop_iteration = it;
            return;
        }
        if (st->replace_interval > 0 && it % st->replace_interval == 0) {
            st->stop = pstop_replace;
            st->stop_iteration = it;
            return;
        }
        if (!__builtin_isfinite(rho) || !__builtin_isfinite(prev) || __builtin_fabs(rho) <= st->tiny
            || __builtin_fabs(prev) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = __builtin_isfinite(rho) && __builtin_isfinite(prev) ? reason_biorthogonality_loss
                                                                             : reason_non_finite_scalar;
            st->stage = 2;
            st->stop_iteration = it;
            return;
        }
        const double beta = (rho / prev) * (st->alpha / st->omega);
        st->beta = beta;
        const double den = t[4] + beta * t[3] - beta * st->omega * t[5];
        if (!__builtin_isfinite(den) || __builtin_fabs(den) <= st->tiny) {
            st->stop = stop_breakdown;
            st->reason = __builtin_isfinite(den) ? reason_alpha_denominator : reason_non_finite_scalar;
            st->stage = 2;
            st->stop_iteration = it;
            return;
        }
        st->alpha = rho / den;
        if (it >= st->max_iterations) {
            st->stop = stop_max;
            st->stop_iteration = it;
            return;
        }
        st->iteration = it + 1;
    }
};

struct Workspace {
    DeviceArray<double> vecs;  // 8 owned-size vectors (+ ghosts for SpMV inputs)
    DeviceArray<BicgState> state;
    std::int64_t local = 0;
    double* v(int k) const { return vecs.get() + static_cast<std::int64_t>(k) * local; }
};

} // namespace

int bicgstab_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, const owtk_solver_options* o,
                 owtk_solver_result* res)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    if (!valid_options(o)) return OWTK_OK;
    if (o->reduction < OWTK_REDUCTION_SERIAL || o->reduction > OWTK_REDUCTION_MIXED_PRECISION)
        throw std::invalid_argument("unknown reduction");
    // compensated and mixed-precision reductions: double-double sums (the
    // vectors are FP64, so both widen the accumulation beyond FP64)
    const bool cmp = o->reduction != OWTK_REDUCTION_SERIAL;
    const bool app = o->convergence_test != nullptr;
    if (pc && pc->m->owned * pc->m->bs != m->owned * m->bs)
        throw std::invalid_argument("preconditioner layout mismatch");
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage != OWTK_F64)
        throw std::invalid_argument("FP64 solve needs an FP64 preconditioner");
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const std::int64_t n = m->owned * m->bs;
    const std::int64_t L = m->local() * m->bs;
    Workspace w;
    w.local = L;
    w.vecs.allocate(static_cast<std::size_t>(9 * L));
    kern::fill_zero_bytes(st, w.vecs.get(), w.vecs.bytes());
    w.state.allocate(1);
    double* r = w.v(0);
    double* shadow = w.v(1);
    double* p = w.v(2);
    double* z = w.v(3);
    double* v = w.v(4);
    double* s = w.v(5);
    double* yv = w.v(6);
    double* t = w.v(7);
    BicgState* dst = w.state.get();
    auto* hst = static_cast<BicgState*>(ctx->pinned.ptr);
    ctx->pinned.ensure(sizeof(BicgState) + 64);
    hst = static_cast<BicgState*>(ctx->pinned.ptr);
    std::uint64_t ops = 0, pcs = 0, reds = 0;

    const unsigned int g = dev::stream_grid(n, 4);
    // ||v|| into *out (compensated mode; the FP64 path fuses its norms)
    auto norm_dd = [&](double* v_in, const double* b_in, double* out) {
        k_residual_dd<<<g, kBlock, 0, st>>>(v_in, b_in, n, scratch);
        dev::launch_finish_dd<2>(st, FinNormTo{v_in, n, out}, g, 2, scratch, nullptr);
        OWTK_LAUNCH_CHECK();
    };
    auto true_residual = [&](double* norm_slot) {
        if (cmp) {
            dev::EpiStore<double> e;
            e.y = r;
            spmv_f64(m, x, e);
            norm_dd(r, b, norm_slot);
        } else {
            dev::EpiResidual e;
            e.y = r;
            e.b = b;
            e.n = n;
            e.norm_out = norm_slot;
            spmv_f64(m, x, e);
        }
        ++ops;
        ++reds;
    };
    auto read_state = [&] {
        OWTK_HIP(hipMemcpyAsync(hst, dst, sizeof(BicgState), hipMemcpyDeviceToHost, st));
        ctx->sync();
    };
    auto write_state = [&] { OWTK_HIP(hipMemcpyAsync(dst, hst, sizeof(BicgState), hipMemcpyHostToDevice, st)); };

    double* d_norms = ctx->scalars.get();
    double norms[2];
    auto rhs_norm_to = [&](double* slot) {
        if (cmp) {
            norm_dd(const_cast<double*>(b), nullptr, slot);  // read only: no b_in
        } else {
            kern::norm_to(st, scratch, b, n, slot);
        }
        ++reds;
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
    // application_convergence (krylov_solvers.hpp:122-143): the test sees the
    // current solution; on success the true residual is measured
    auto app_converged = [&](long long iteration) {
        ctx->sync();
        if (!o->convergence_test(o->convergence_test_data, static_cast<std::uint64_t>(iteration), x,
                                 static_cast<std::size_t>(n)))
            return false;
        true_residual(&d_norms[2]);
        rhs_norm_to(&d_norms[3]);
        OWTK_HIP(hipMemcpyAsync(norms, &d_norms[2], 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
        set_residual(res, res->recursive_residual_norm, norms[0], norms[1]);
        if (!std::isfinite(norms[0]) || !std::isfinite(norms[1])) {
            res->status = status_breakdown;
            res->breakdown_reason = reason_non_finite_scalar;
        } else {
            res->status = status_converged;
            res->converged_by_application = 1;
        }
        return true;
    };
    if (app && app_converged(0)) {
        finish(res->status);
        return OWTK_OK;
    }

    // r = b - A x; shadow = r; ||b||, ||r||
    std::memset(hst, 0, sizeof(BicgState));
    write_state();
    true_residual(&d_norms[1]);
    OWTK_HIP(hipMemcpyAsync(shadow, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    rhs_norm_to(&d_norms[0]);
    OWTK_HIP(hipMemcpyAsync(norms, d_norms, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    const double rhs_norm = norms[0], initial = norms[1];
    // with an application test only an exactly solved system stops on the
    // residual (convergence_threshold, core.hpp:549-560)
    double thr = threshold_of(rhs_norm, o);
    if (app && std::isfinite(thr)) thr = 0.0;
    res->initial_residual_norm = initial;
    set_residual(res, initial, initial, rhs_norm);
    if (!std::isfinite(initial) || !std::isfinite(thr)) {
        res->breakdown_reason = reason_non_finite_scalar;
        return finish(status_breakdown);
    }
    if (initial <= thr) return finish(status_converged);

    auto restart_state = [&](long long it) {
        hst->rho_prev = 1;
        hst->alpha = 1;
        hst->omega = 1;
        hst->first = 1;
        hst->stop = stop_none;
        hst->iteration = it;
        write_state();
        OWTK_HIP(hipMemcpyAsync(shadow, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        kern::fill_zero_bytes(st, p, n * sizeof(double));
        kern::fill_zero_bytes(st, v, n * sizeof(double));
        if (cmp) {
            k_rho<true><<<g, kBlock, 0, st>>>(shadow, r, n, dst, scratch);
            dev::launch_finish_dd<1>(st, FinRho{dst}, g, 0, scratch, nullptr);
        } else {
            k_rho<false><<<g, kBlock, 0, st>>>(shadow, r, n, dst, scratch);
            dev::launch_finish<1>(st, FinRho{dst}, g, 0, scratch, nullptr);
        }
        OWTK_LAUNCH_CHECK();
    };
    std::memset(hst, 0, sizeof(BicgState));
    hst->threshold = thr;
    hst->tiny = 1024 * std::numeric_limits<double>::min();
    hst->breakdown_tolerance = o->breakdown_tolerance;
    hst->recursive = initial;
    hst->max_iterations = static_cast<long long>(o->maximum_iterations);
    hst->check_interval = static_cast<long long>(o->convergence_check_interval);
    hst->replace_interval = static_cast<long long>(o->residual_replacement_interval);
    hst->app_test = app ? 1 : 0;
    restart_state(1);

    const int* stop = &dst->stop;
    auto enqueue_iteration = [&] {
        k_p_update<<<g, kBlock, 0, st>>>(p, r, v, n, dst);
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, p, z, n, st, stop);
        if (cmp) {
            dev::EpiStore<double> e;
            e.y = v;
            e.stop = stop;
            spmv_f64(m, z, e);
            k_dot_dd<1><<<g, kBlock, 0, st>>>(v, shadow, n, stop, scratch);
            dev::launch_finish_dd<1>(st, FinDot1<FinAlpha>{FinAlpha{dst}}, g, 0, scratch, stop);
            k_s_update<true><<<g, kBlock, 0, st>>>(s, r, v, n, dst, scratch);
            dev::launch_finish_dd<2>(st, FinSnorm{dst, s, n}, g, 2, scratch, stop);
        } else {
            dev::EpiDot1<FinAlpha> e;
            e.y = v;
            e.w = shadow;
            e.stop = stop;
            e.fin = FinAlpha{dst};
            spmv_f64(m, z, e);
            k_s_update<false><<<g, kBlock, 0, st>>>(s, r, v, n, dst, scratch);
            dev::launch_finish<2>(st, FinSnorm{dst, s, n}, g, 2, scratch, stop);
        }
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, s, yv, n, st, stop);
        if (cmp) {
            dev::EpiStore<double> e;
            e.y = t;
            e.stop = stop;
            spmv_f64(m, yv, e);
            k_dot_dd<2><<<g, kBlock, 0, st>>>(t, s, n, stop, scratch);
            dev::launch_finish_dd<2>(st, FinDot2<FinOmega>{FinOmega{dst}}, g, 0, scratch, stop);
            k_xr_update<true><<<g, kBlock, 0, st>>>(x, r, z, yv, s, t, shadow, n, dst, scratch);
            dev::launch_finish_dd<3>(st, FinXr{dst, r, n}, g, 4, scratch, stop);
        } else {
            dev::EpiDot2<FinOmega> e;
            e.y = t;
            e.w = s;
            e.stop = stop;
            e.fin = FinOmega{dst};
            spmv_f64(m, yv, e);
            k_xr_update<false><<<g, kBlock, 0, st>>>(x, r, z, yv, s, t, shadow, n, dst, scratch);
            dev::launch_finish<3>(st, FinXr{dst, r, n}, g, 4, scratch, stop);
        }
        OWTK_LAUNCH_CHECK();
    };
    const long long maxit = static_cast<long long>(o->maximum_iterations);
    const long long check = static_cast<long long>(o->convergence_check_interval);
    const long long replace = static_cast<long long>(o->residual_replacement_interval);
    auto is_cadence = [&](long long it) { return (replace > 0 && it % replace == 0) || it % check == 0; };
    auto count_full = [&](long long a, long long b_incl) {
        for (long long k = a; k <= b_incl; ++k) {
            pcs += 2;
            ops += 2;
            reds += 5 + (is_cadence(k) ? 1 : 0);
        }
    };

    // with an application test: one iteration per host round trip, the test
    // after every complete iterate (krylov_solvers.hpp:658, 710)
    auto app_stop = [&]() { return finish(res->status); };
    long long it = 1;
    while (true) {
        long long end = it;
        if (!app)
            while (end < maxit && !is_cadence(end)) ++end;
        for (long long k = it; k <= end; ++k) enqueue_iteration();
        read_state();
        const int code = hst->stop;
        const long long j = hst->stop_iteration;
        if (code == stop_none) {
            // cadence iteration passed without a stop
            count_full(it, end);
            res->iterations = static_cast<std::uint64_t>(end);
            res->recursive_residual_norm = hst->recursive;
            if (app && app_converged(end)) return app_stop();
            it = end + 1;
            continue;
        }
        if (code == stop_breakdown) {
            // a rho breakdown (stage 0) follows the complete iterate j - 1
            if (app && hst->stage == 0 && j == end + 1) {
                count_full(it, j - 1);
                res->iterations = static_cast<std::uint64_t>(j - 1);
                res->recursive_residual_norm = hst->recursive;
                if (app_converged(j - 1)) return app_stop();
                it = j;  // counted already
            }
            count_full(it, j - 1);
            if (hst->stage == 0) reds += 1;
            else if (hst->stage == 1) { pcs += 1; ops += 1; reds += 2; }
            else { pcs += 2; ops += 2; reds += 5; }
            res->iterations = static_cast<std::uint64_t>(j);
            res->breakdown_reason = hst->reason;
            res->recursive_residual_norm = hst->recursive;
            return finish(status_breakdown);
        }
        if (code == stop_max) {
            count_full(it, j);
            res->iterations = static_cast<std::uint64_t>(j);
            res->recursive_residual_norm = hst->recursive;
            if (app && app_converged(j)) return app_stop();
            break;
        }
        // early (s small) or cadence: x/true residual then converge or restart
        double rec;
        if (code == stop_early) {
            count_full(it, j - 1);
            pcs += 1;
            ops += 1;
            reds += 3;
            kern::axpy(st, n, hst->alpha, z, x);
            rec = hst->snorm;
            res->iterations = static_cast<std::uint64_t>(j);
            if (app && app_converged(j)) return app_stop();
        } else {
            count_full(it, j);
            rec = hst->rnorm;
            res->iterations = static_cast<std::uint64_t>(j);
            res->recursive_residual_norm = hst->recursive;
            if (app && app_converged(j)) return app_stop();
        }
        res->iterations = static_cast<std::uint64_t>(j);
        true_residual(&d_norms[2]);
        double tn;
        OWTK_HIP(hipMemcpyAsync(&norms[0], &d_norms[2], sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
        tn = norms[0];
        set_residual(res, rec, tn, rhs_norm);
        hst->recursive = rec;
        if (tn <= thr) return finish(status_converged);
        it = j + 1;
        if (it > maxit) break;
        restart_state(it);
    }
    // final true residual (krylov_solvers.hpp:735-744)
    true_residual(&d_norms[3]);
    OWTK_HIP(hipMemcpyAsync(&norms[0], &d_norms[3], sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    const double tn = norms[0];
    set_residual(res, res->recursive_residual_norm, tn, rhs_norm);
    int status = tn <= thr ? status_converged : status_maximum_iterations;
    if (!std::isfinite(tn)) {
        status = status_breakdown;
        res->breakdown_reason = reason_non_finite_scalar;
    }
    return finish(status);
}


// pipelined_bicgstab (pipelined_bicgstab.hpp:20-244) with the reference's
// counters: per full iteration 2 preconditioner applies, 2 operator applies
// and 3 fused reductions.
int pipelined_bicgstab_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, const owtk_solver_options* o,
                           owtk_solver_result* res)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    if (!valid_options(o)) return OWTK_OK;
    if (pc && pc->m->owned * pc->m->bs != m->owned * m->bs)
        throw std::invalid_argument("preconditioner layout mismatch");
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage != OWTK_F64)
        throw std::invalid_argument("FP64 solve needs an FP64 preconditioner");
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const std::int64_t n = m->owned * m->bs;
    const std::int64_t L = m->local() * m->bs;
    Workspace w;
    w.local = L;
    w.vecs.allocate(static_cast<std::size_t>(9 * L));
    kern::fill_zero_bytes(st, w.vecs.get(), w.vecs.bytes());
    w.state.allocate(1);
    double* r = w.v(0);
    double* shadow = w.v(1);
    double* p = w.v(2);
    double* z = w.v(3);
    double* v = w.v(4);
    double* s = w.v(5);
    double* yv = w.v(6);
    double* t = w.v(7);
    BicgState* dst = w.state.get();
    ctx->pinned.ensure(sizeof(BicgState) + 64);
    auto* hst = static_cast<BicgState*>(ctx->pinned.ptr);
    std::uint64_t ops = 0, pcs = 0, reds = 0;
    auto read_state = [&] {
        OWTK_HIP(hipMemcpyAsync(hst, dst, sizeof(BicgState), hipMemcpyDeviceToHost, st));
        ctx->sync();
    };
    auto write_state = [&] { OWTK_HIP(hipMemcpyAsync(dst, hst, sizeof(BicgState), hipMemcpyHostToDevice, st)); };
    double* d_norms = ctx->scalars.get();
    double norms[2];
    auto true_residual = [&](double* norm_slot) {
        dev::EpiResidual e;
        e.y = r;
        e.b = b;
        e.n = n;
        e.norm_out = norm_slot;
        spmv_f64(m, x, e);
        ++ops;
    };
    auto read_norm = [&](double* slot) {
        OWTK_HIP(hipMemcpyAsync(&norms[0], slot, sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
        return norms[0];
    };
    std::uint64_t initial_extra = 0;
    auto finish = [&](int status) {
        res->status = status;
        res->operator_applications = ops;
        res->preconditioner_applications = pcs;
        res->global_reductions = reds + initial_extra + static_cast<std::uint64_t>(hst->extra_reductions);
        if (o->collect_timings)
            res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return OWTK_OK;
    };

    std::memset(hst, 0, sizeof(BicgState));
    write_state();
    true_residual(&d_norms[1]);
    kern::norm_to(st, scratch, b, n, &d_norms[0]);
    ++reds;  // |b|^2 and |r|^2 are one fused reduction in the reference
    OWTK_HIP(hipMemcpyAsync(norms, d_norms, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    const double rhs_norm = norms[0], initial = norms[1];
    // the squares the reference reduces, recovered from the norms
    initial_extra = static_cast<std::uint64_t>(norm_fallback(rhs_norm * rhs_norm))
                  + static_cast<std::uint64_t>(norm_fallback(initial * initial));
    const double thr = threshold_of(rhs_norm, o);
    res->initial_residual_norm = initial;
    set_residual(res, initial, initial, rhs_norm);
    if (!std::isfinite(initial) || !std::isfinite(thr)) {
        res->breakdown_reason = reason_non_finite_scalar;
        return finish(status_breakdown);
    }
    if (initial <= thr) return finish(status_converged);

    const unsigned int g = dev::stream_grid(n, 4);
    // shadow = r, p = r (first), rho = <shadow, r>
    auto restart = [&](long long it) {
        hst->alpha = 1;
        hst->omega = 1;
        hst->first = 1;
        hst->stop = stop_none;
        hst->iteration = it;
        write_state();
        OWTK_HIP(hipMemcpyAsync(shadow, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        k_rho<false><<<g, kBlock, 0, st>>>(shadow, r, n, dst, scratch);
        dev::launch_finish<1>(st, FinRhoP{dst}, g, 0, scratch, nullptr);
        OWTK_LAUNCH_CHECK();
        ++reds;
    };
    std::memset(hst, 0, sizeof(BicgState));
    hst->threshold = thr;
    hst->tiny = 1024 * std::numeric_limits<double>::min();
    hst->breakdown_tolerance = o->breakdown_tolerance;
    hst->recursive = initial;
    hst->max_iterations = static_cast<long long>(o->maximum_iterations);
    hst->check_interval = 1;
    hst->replace_interval = static_cast<long long>(o->residual_replacement_interval);
    restart(1);

    const int* stop = &dst->stop;
    auto enqueue_iteration = [&] {
        k_p_update<<<g, kBlock, 0, st>>>(p, r, v, n, dst);  // p = r (first) or r + beta (p - omega v)
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, p, z, n, st, stop);
        {
            dev::EpiDot1<FinAlpha> e;
            e.y = v;
            e.w = shadow;
            e.stop = stop;
            e.fin = FinAlpha{dst};
            spmv_f64(m, z, e);
        }
        k_s_plain<<<g, kBlock, 0, st>>>(s, r, v, n, dst);
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, s, yv, n, st, stop);
        {
            EpiOmegaP e;
            e.y = t;
            e.s = s;
            e.n = n;
            e.st = dst;
            e.stop = stop;
            spmv_f64(m, yv, e);
        }
        k_xr_update<false><<<g, kBlock, 0, st>>>(x, r, z, yv, s, t, shadow, n, dst, scratch);
        dev::launch_finish<3>(st, FinNextP{dst, r, n}, g, 4, scratch, stop);
        OWTK_LAUNCH_CHECK();
    };
    const long long maxit = static_cast<long long>(o->maximum_iterations);
    constexpr long long kBatch = 8;  // iterations between looks at the device state
    auto count_full = [&](long long a, long long b_incl) {
        if (b_incl < a) return;
        const auto k = static_cast<std::uint64_t>(b_incl - a + 1);
        pcs += 2 * k;
        ops += 2 * k;
        reds += 3 * k;
    };
    long long it = 1;
    while (true) {
        const long long end = std::min(maxit, it + kBatch - 1);
        for (long long k = it; k <= end; ++k) enqueue_iteration();
        read_state();
        const int code = hst->stop;
        const long long j = hst->stop_iteration;
        if (code == stop_none) {
            count_full(it, end);
            res->iterations = static_cast<std::uint64_t>(end);
            res->recursive_residual_norm = hst->recursive;
            it = end + 1;
            continue;
        }
        res->iterations = static_cast<std::uint64_t>(j);
        if (code == stop_breakdown) {
            count_full(it, j - 1);
            if (hst->stage == 1) { pcs += 1; ops += 1; reds += 1; }
            else if (hst->stage == 3) { pcs += 2; ops += 2; reds += 2; }
            else count_full(j, j);
            res->breakdown_reason = hst->reason;
            res->recursive_residual_norm = hst->recursive;
            return finish(status_breakdown);
        }
        if (code == stop_max) {
            count_full(it, j);
            res->recursive_residual_norm = hst->recursive;
            break;
        }
        if (code == pstop_replace) {
            count_full(it, j);
            res->recursive_residual_norm = hst->recursive;
            true_residual(&d_norms[2]);  // replacement: no convergence test, restart
            it = j + 1;
            restart(it);  // the reference restarts (one reduction) even after the last iteration
            if (it > maxit) break;
            continue;
        }
        double rec;
        if (code == stop_early) {
            count_full(it, j - 1);
            pcs += 2;
            ops += 2;
            reds += 2;
            kern::axpy(st, n, hst->alpha, z, x);
            rec = hst->snorm;
        } else {
            count_full(it, j);
            rec = hst->rnorm;
        }
        true_residual(&d_norms[2]);
        ++reds;
        const double tn = read_norm(&d_norms[2]);
        set_residual(res, rec, tn, rhs_norm);
        if (!std::isfinite(tn)) {
            res->breakdown_reason = reason_non_finite_scalar;
            return finish(status_breakdown);
        }
        if (tn <= thr) return finish(status_converged);
        it = j + 1;
        restart(it);
        if (it > maxit) break;
    }
    true_residual(&d_norms[3]);
    ++reds;
    const double tn = read_norm(&d_norms[3]);
    set_residual(res, hst->recursive, tn, rhs_norm);  // the last recursive |r| (not an early-exit |s|)
    int status = tn <= thr ? status_converged : status_maximum_iterations;
    if (!std::isfinite(tn)) {
        status = status_breakdown;
        res->breakdown_reason = reason_non_finite_scalar;
    }
    return finish(status);
}

// communication_hiding_bicgstab (pipelined_bicgstab.hpp:317-597) with the
// reference's counters: per full iteration 2 preconditioner applies, 2
// operator applies and 2 fused reductions; a restart of the recurrence 2 of
// each.
int communication_hiding_bicgstab_run(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                                      const owtk_solver_options* o, owtk_solver_result* res)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    if (!valid_options(o)) return OWTK_OK;
    if (pc && pc->m->owned * pc->m->bs != m->owned * m->bs)
        throw std::invalid_argument("preconditioner layout mismatch");
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage != OWTK_F64)
        throw std::invalid_argument("FP64 solve needs an FP64 preconditioner");
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const std::int64_t n = m->owned * m->bs;
    const std::int64_t L = m->local() * m->bs;
    DeviceArray<double> vecs;
    vecs.allocate(static_cast<std::size_t>(15 * L));
    kern::fill_zero_bytes(st, vecs.get(), vecs.bytes());
    auto vec = [&](int k) { return vecs.get() + static_cast<std::int64_t>(k) * L; };
    ChVecs V{vec(0), vec(1), vec(2), vec(3), vec(4), vec(5), vec(6), vec(7),
             vec(8), vec(9), vec(10), vec(11), vec(12), vec(13), vec(14)};
    DeviceArray<ChState> state;
    state.allocate(1);
    ChState* dst = state.get();
    ctx->pinned.ensure(sizeof(ChState) + 64);
    auto* hst = static_cast<ChState*>(ctx->pinned.ptr);
    std::uint64_t ops = 0, pcs = 0, reds = 0, initial_extra = 0;
    auto read_state = [&] {
        OWTK_HIP(hipMemcpyAsync(hst, dst, sizeof(ChState), hipMemcpyDeviceToHost, st));
        ctx->sync();
    };
    auto write_state = [&] { OWTK_HIP(hipMemcpyAsync(dst, hst, sizeof(ChState), hipMemcpyHostToDevice, st)); };
    double* d_norms = ctx->scalars.get();
    double norms[2];
    auto true_residual = [&](double* norm_slot) {
        dev::EpiResidual e;
        e.y = V.r;
        e.b = b;
        e.n = n;
        e.norm_out = norm_slot;
        spmv_f64(m, x, e);
        ++ops;
    };
    auto read_norm = [&](double* slot) {
        OWTK_HIP(hipMemcpyAsync(&norms[0], slot, sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
        return norms[0];
    };
    auto finish = [&](int status) {
        res->status = status;
        res->operator_applications = ops;
        res->preconditioner_applications = pcs;
        res->global_reductions = reds + initial_extra + static_cast<std::uint64_t>(hst->extra_reductions);
        if (o->collect_timings)
            res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return OWTK_OK;
    };

    std::memset(hst, 0, sizeof(ChState));
    write_state();
    true_residual(&d_norms[1]);
    kern::norm_to(st, scratch, b, n, &d_norms[0]);
    ++reds;  // |b|^2 and |r|^2 are one fused reduction in the reference
    OWTK_HIP(hipMemcpyAsync(norms, d_norms, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    const double rhs_norm = norms[0], initial = norms[1];
    initial_extra = static_cast<std::uint64_t>(norm_fallback(rhs_norm * rhs_norm))
                  + static_cast<std::uint64_t>(norm_fallback(initial * initial));
    const double thr = threshold_of(rhs_norm, o);
    res->initial_residual_norm = initial;
    set_residual(res, initial, initial, rhs_norm);
    if (!std::isfinite(initial) || !std::isfinite(thr)) {
        res->breakdown_reason = reason_non_finite_scalar;
        return finish(status_breakdown);
    }
    if (initial <= thr) return finish(status_converged);

    std::memset(hst, 0, sizeof(ChState));
    hst->threshold = thr;
    hst->tiny = 1024 * std::numeric_limits<double>::min();
    hst->breakdown_tolerance = o->breakdown_tolerance;
    hst->recursive = initial;
    hst->max_iterations = static_cast<long long>(o->maximum_iterations);
    hst->replace_interval = static_cast<long long>(o->residual_replacement_interval);
    const unsigned int g = dev::stream_grid(n, 4);
    const int* stop = &dst->stop;
    auto plain_op = [&](const double* in, double* out) {
        dev::EpiStore<double> e;
        e.y = out;
        e.stop = stop;
        spmv_f64(m, in, e);
    };
    // initialize_recurrence: rp = r, rho = <r, rp>, r2 = M r, w = A r2,
    // alpha = rho / <w, rp>, w2 = M w, t = A w2.  Returns false on a
    // breakdown, with the reference's counts of what ran.
    auto initialize = [&](long long it) {
        hst->stop = stop_none;
        hst->iteration = it;
        write_state();
        k_ch_rho<<<g, kBlock, 0, st>>>(V.r, V.rp, n, scratch);
        dev::launch_finish<1>(st, FinChRho{dst}, g, 0, scratch, nullptr);
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, V.r, V.r2, n, st, stop);
        {
            dev::EpiDot1<FinChD2> e;
            e.y = V.w;
            e.w = V.rp;
            e.stop = stop;
            e.fin = FinChD2{dst};
            spmv_f64(m, V.r2, e);
        }
        apply_pc_f64(pc, V.w, V.w2, n, st, stop);
        plain_op(V.w2, V.t);
        read_state();
        if (hst->stop == chstop_init) {
            reds += hst->stage == 10 ? 1 : 2;
            if (hst->stage == 11) {
                pcs += 1;
                ops += 1;
            }
            return false;
        }
        reds += 2;
        pcs += 2;
        ops += 2;
        return true;
    };
    if (!initialize(1)) {
        res->breakdown_reason = reason_biorthogonality_loss;
        return finish(status_breakdown);
    }

    auto enqueue_iteration = [&] {
        k_ch_first<<<g, kBlock, 0, st>>>(V, n, dst, scratch);
        dev::launch_finish<4>(st, FinChFirst{dst, V.q, n}, g, 8, scratch, stop);
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, V.z, V.z2, n, st, stop);
        plain_op(V.z2, V.v);
        k_ch_second<<<g, kBlock, 0, st>>>(V, x, n, dst, scratch);
        dev::launch_finish<6>(st, FinChSecond{dst, V.r, n}, g, 2, scratch, stop);
        OWTK_LAUNCH_CHECK();
        apply_pc_f64(pc, V.w, V.w2, n, st, stop);
        plain_op(V.w2, V.t);
    };
    auto count_full = [&](long long a, long long b_incl) {
        if (b_incl < a) return;
        const auto k = static_cast<std::uint64_t>(b_incl - a + 1);
        pcs += 2 * k;
        ops += 2 * k;
        reds += 2 * k;
    };
    // restart after a false convergence or a replacement (the reference's
    // `continue`): a failed restart is a biorthogonality breakdown
    auto restart_or_fail = [&](long long next) { return initialize(next); };
    const long long maxit = static_cast<long long>(o->maximum_iterations);
    constexpr long long kBatch = 8;  // iterations between looks at the device state
    long long it = 1;
    while (true) {
        const long long end = std::min(maxit, it + kBatch - 1);
        for (long long k = it; k <= end; ++k) enqueue_iteration();
        read_state();
        const int code = hst->stop;
        const long long j = hst->stop_iteration;
        if (code == stop_none) {
            count_full(it, end);
            res->iterations = static_cast<std::uint64_t>(end);
            res->recursive_residual_norm = hst->recursive;
            it = end + 1;
            continue;
        }
        res->iterations = static_cast<std::uint64_t>(j);
        res->recursive_residual_norm = hst->recursive;
        if (code == stop_breakdown) {
            count_full(it, j - 1);
            if (hst->stage == 1) {
                pcs += 1;
                ops += 1;
                reds += 1;
            } else {
                count_full(j, j);
            }
            res->breakdown_reason = hst->reason;
            return finish(status_breakdown);
        }
        if (code == stop_max) {
            count_full(it, j);
            break;
        }
        if (code == pstop_replace) {
            count_full(it, j);
            true_residual(&d_norms[2]);  // replacement: no convergence test
            if (!restart_or_fail(j + 1)) {
                res->breakdown_reason = reason_biorthogonality_loss;
                return finish(status_breakdown);
            }
            it = j + 1;
            if (it > maxit) break;
            continue;
        }
        double rec;
        if (code == stop_early) {
            count_full(it, j - 1);
            pcs += 1;
            ops += 1;
            reds += 1;
            kern::axpy(st, n, hst->alpha, V.p2, x);
            rec = hst->qnorm;
        } else {
            count_full(it, j);
            rec = hst->rnorm;
        }
        true_residual(&d_norms[2]);
        ++reds;
        const double tn = read_norm(&d_norms[2]);
        set_residual(res, rec, tn, rhs_norm);
        if (!std::isfinite(tn)) {
            res->breakdown_reason = reason_non_finite_scalar;
            return finish(status_breakdown);
        }
        if (tn <= thr) return finish(status_converged);
        if (!restart_or_fail(j + 1)) {
            res->breakdown_reason = reason_biorthogonality_loss;
            return finish(status_breakdown);
        }
        it = j + 1;
        if (it > maxit) break;
    }
    true_residual(&d_norms[3]);
    ++reds;
    const double tn = read_norm(&d_norms[3]);
    set_residual(res, hst->recursive, tn, rhs_norm);  // the last recursive |r| (not an early-exit |s|)
    int status = tn <= thr ? status_converged : status_maximum_iterations;
    if (!std::isfinite(tn)) {
        status = status_breakdown;
        res->breakdown_reason = reason_non_finite_scalar;
    }
    return finish(status);
}
} // namespace owtk

extern "C" int owtk_bicgstab(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                             const owtk_solver_options* options, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        m->ctx->activate();
        bicgstab_run(m, pc, b, x, options, result);
    });
}

extern "C" int owtk_pipelined_bicgstab(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                                       const owtk_solver_options* options, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        require_plain_options(options, "pipelined_bicgstab");
        m->ctx->activate();
        pipelined_bicgstab_run(m, pc, b, x, options, result);
    });
}

extern "C" int owtk_communication_hiding_bicgstab(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                                                  const owtk_solver_options* options, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        require_plain_options(options, "communication_hiding_bicgstab");
        m->ctx->activate();
        communication_hiding_bicgstab_run(m, pc, b, x, options, result);
    });
}
