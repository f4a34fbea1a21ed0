// Device-resident flexible restarted GMRES (fgmres, krylov_solvers.hpp:159-486)
// and the mixed-precision iterative-refinement driver built on it.
//
// One Arnoldi column is five launches:
//   z_k = M^-1 v_k ; w = A z_k
//   A: <w, v_0..v_k>, <w, w>                  -> h(:,k), projected norm
//   B: w -= sum_j h_j v_j  (+ second CGS pass dots when re-orthogonalising)
//   C: w -= sum_j c_j v_j, then v_{k+1} = w / h_{k+1,k} (or the verified-norm
//      path of the reference), Givens rotation, residual estimate
//   D: scale after a verified norm (no-op otherwise)
// Hessenberg matrix, rotations and flags live in a device state; the host
// reads it once per restart cycle for the back-substitution.  The engine is a
// template on the vector type: double for the FP64 solver, float for the inner
// solve of the refinement driver (inner products and the Hessenberg stay FP64).
#include "precond.hpp"
#include "solver_common.hpp"
#include "host/host.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"

#include <chrono>
#include <deque>
#include <cmath>
#include <functional>
#include <limits>
#include <memory>
#include <vector>

using namespace owtk;

// ArnoldiSnapshot (core.hpp:463-534) on the device: the last cycle's
// unrotated Hessenberg columns and its preconditioned basis vectors.  The
// columns are written by the Givens finisher of each Arnoldi step (FP64,
// column stride kStride) and read back once when the cycle ends; the basis
// vectors keep the solver's owned + ghost layout with zero ghosts, so each
// one is a valid SpMV input.
struct owtk_arnoldi_snapshot_s {
    static constexpr int kStride = 61;   // restart limit + 1
    owtk_context ctx = nullptr;
    std::int64_t maximum_columns = 0, columns = 0;
    std::int64_t ld = -1, owned = -1;    // layout of the basis vectors
    owtk::DeviceArray<double> h;         // kStride * (kStride - 1)
    owtk::DeviceArray<double> basis;     // maximum_columns * ld
    std::vector<double> host_h;          // h as of the last finished cycle
};

namespace owtk {

namespace {

using dev::kBlock;
using dev::ReduceScratch;

constexpr int kMaxRestart = 60;
constexpr int kLd = kMaxRestart + 1;

struct FgState {
    double H[kLd * kMaxRestart];
    double cs[kMaxRestart], sn[kMaxRestart], g[kLd];
    double corr[kLd];
    double original, projected, nsq, inv_h, threshold, breakdown_tolerance, recursive;
    int mode;
    int halt;
    int stop;
    int reason;
    int cycle_done;
    int abreak;
    int columns;
    int reorth_count, verif_count;
    int do_reorth, verify, direct, need_scale;
};

__device__ inline double& Hd(FgState* s, int row, int col) { return s->H[col * kLd + row]; }

static_assert(owtk_arnoldi_snapshot_s::kStride == kLd, "snapshot stride follows the Hessenberg stride");

// ArnoldiSnapshot::clear / begin_cycle / finish_cycle (core.hpp:465-508).
void snapshot_clear(owtk_arnoldi_snapshot sn)
{
    if (!sn) return;
    sn->columns = 0;
    sn->maximum_columns = 0;
}

void snapshot_begin(owtk_arnoldi_snapshot sn, int m, std::int64_t ld, std::int64_t owned, hipStream_t st)
{
    sn->maximum_columns = m;
    sn->columns = 0;
    if (!sn->h.get()) sn->h.allocate(static_cast<std::size_t>(kLd) * kMaxRestart);
    const std::size_t need = static_cast<std::size_t>(m) * static_cast<std::size_t>(ld);
    if (sn->basis.count < need || sn->ld != ld || sn->owned != owned) {
        if (sn->basis.count < need) sn->basis.allocate(need);
        // ghost entries are never written afterwards
        kern::fill_zero_bytes(st, sn->basis.get(), sn->basis.count * sizeof(double));
        sn->ld = ld;
        sn->owned = owned;
    }
}

void snapshot_finish(owtk_arnoldi_snapshot sn, const double* Z, int columns, owtk_context ctx)
{
    const hipStream_t st = ctx->stream;
    if (columns > 0)
        OWTK_HIP(hipMemcpyAsync(sn->basis.get(), Z, static_cast<std::size_t>(columns) * sn->ld * sizeof(double),
                                hipMemcpyDeviceToDevice, st));
    sn->host_h.resize(sn->h.count);
    OWTK_HIP(hipMemcpyAsync(sn->host_h.data(), sn->h.get(), sn->h.count * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    sn->columns = columns;
}

// Post-process h(k+1,k): finiteness, breakdown test, reciprocal
// (krylov_solvers.hpp:346-359).
__device__ void settle_h(FgState* s, int k, double h, bool direct)
{
    Hd(s, k + 1, k) = h;
    if (!__builtin_isfinite(h)) {
        s->stop = 1;
        s->reason = reason_non_finite_scalar;
        s->halt = 1;
        return;
    }
    double scale = 0.0;
    for (int row = 0; row <= k + 1; ++row) scale = fmax(scale, __builtin_fabs(Hd(s, row, k)));
    if (h > s->breakdown_tolerance * scale) {
        s->inv_h = 1.0 / h;
        s->direct = direct ? 1 : 0;
        s->need_scale = direct ? 0 : 1;
    } else {
        s->abreak = 1;
        s->direct = 0;
        s->need_scale = 0;
    }
}

// Norm decision after the projections (krylov_solvers.hpp:332-345).
__device__ void decide(FgState* s, int k)
{
    const double dmin = 2.2250738585072014e-308;
    const double eps = 2.220446049250313e-16;
    const double nsq = s->nsq;
    if (s->original < dmin || !__builtin_isfinite(nsq) || nsq <= eps * s->original || (nsq > 0 && nsq < dmin)) {
        s->verify = 1;
        s->direct = 0;
        return;
    }
    s->verify = 0;
    settle_h(s, k, __builtin_sqrt(fmax(nsq, 0.0)), true);
}

template<class T>
__global__ __launch_bounds__(256) void k_fg_v0(T* __restrict__ v0, const T* __restrict__ r, std::int64_t n, double beta)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        v0[i] = static_cast<T>(static_cast<double>(r[i]) / beta);
}

// L consecutive elements moved as one 16-byte access (L = 1: scalar).  The
// passes walk packets; the host picks L > 1 only when n and the basis stride
// are multiples of L (so every packet is 16-byte aligned).
template<class T, int L>
struct alignas(sizeof(T) * L) Packet {
    T v[L];
};

template<class T, int L>
__device__ inline Packet<T, L> load_packet(const T* p)
{
    return *reinterpret_cast<const Packet<T, L>*>(p);
}

// Basis vectors are streamed once per pass: FP32 bases are read with the
// non-temporal hint so they do not displace what the preconditioner and the
// operator reread every iteration (the learned preconditioner's matrix stays
// in the MALL).  FP64 bases are read normally: measured, the hint made the
// FP64 FGMRES + ILU(0) solve 3% slower.
template<class T, int L>
__device__ inline Packet<T, L> load_packet_nt(const T* p)
{
    Packet<T, L> out;
    if constexpr (sizeof(T) == 8) {
        out = load_packet<T, L>(p);
    } else if constexpr (L == 1) {
        out.v[0] = __builtin_nontemporal_load(p);
    } else {
        typedef T vec __attribute__((ext_vector_type(L)));
        const vec v = __builtin_nontemporal_load(reinterpret_cast<const vec*>(p));
#pragma unroll
        for (int l = 0; l < L; ++l) out.v[l] = v[l];
    }
    return out;
}

template<class T, int L>
__device__ inline void store_packet(T* p, const Packet<T, L>& x)
{
    *reinterpret_cast<Packet<T, L>*>(p) = x;
}

// Pass A: h_j = <w, v_j>, j <= k, and <w, w>.
template<int K, class T, int L>
__global__ __launch_bounds__(256) void k_fg_pass_a(const T* __restrict__ V, std::int64_t ld, int count,
                                                   const T* __restrict__ w, std::int64_t n, FgState* s, int k,
                                                   ReduceScratch scratch)
{
    if (s->halt) return;
    double red[K];
#pragma unroll
    for (int j = 0; j < K; ++j) red[j] = 0.0;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    const std::int64_t packets = n / L;
    for (std::int64_t ip = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; ip < packets; ip += step) {
        const std::int64_t i = ip * L;
        const Packet<T, L> wp = load_packet<T, L>(w + i);
#pragma unroll
        for (int j = 0; j < K - 1; ++j) {
            if (j < count) {
                const Packet<T, L> vp = load_packet_nt<T, L>(V + j * ld + i);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    red[j] = __builtin_fma(static_cast<double>(wp.v[l]), static_cast<double>(vp.v[l]), red[j]);
            }
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const double wi = static_cast<double>(wp.v[l]);
            red[K - 1] = __builtin_fma(wi, wi, red[K - 1]);
        }
    }
    dev::store_partials<K>(red, 0, scratch);
}

template<int K>
struct FinA {
    FgState* s;
    int k;
    int count;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x != 0) return;
        s->original = t[K - 1];
        double projected = s->original;
        for (int j = 0; j < count; ++j) {
            Hd(s, j, k) = t[j];
            projected -= t[j] * t[j];
        }
        s->projected = projected;
        const double nonneg = fmax(projected, 0.0);
        const bool forced = s->mode == 0;
        const bool adaptive = s->mode == 1 && (projected <= 0.0 || nonneg < 0.5 * s->original);
        s->do_reorth = (forced || adaptive) ? 1 : 0;
        s->nsq = nonneg;
        s->verify = 0;
        s->direct = 0;
        s->need_scale = 0;
        if (!s->do_reorth) decide(s, k);
    }
};

// Pass B: w -= sum h_j v_j; with re-orthogonalisation also c_j = <w, v_j>, <w, w>.
// Without re-orthogonalisation and with h(k+1,k) already settled from the
// projected norm, the new basis vector is final here: it is written straight
// to v_{k+1} = w / h (rounded through T exactly as pass C would) and pass C
// skips its vector work.
template<int K, class T, int L>
__global__ __launch_bounds__(256) void k_fg_pass_b(const T* __restrict__ V, std::int64_t ld, int count,
                                                   T* __restrict__ w, T* __restrict__ vnext, std::int64_t n,
                                                   FgState* s, int k, ReduceScratch scratch)
{
    if (s->halt) return;
    __shared__ double h[K];
    if (threadIdx.x < count) h[threadIdx.x] = Hd(s, threadIdx.x, k);
    __syncthreads();
    const bool reorth = s->do_reorth != 0;
    const bool final_here = !reorth && s->direct != 0;
    const double inv = s->inv_h;
    double red[K];
#pragma unroll
    for (int j = 0; j < K; ++j) red[j] = 0.0;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    const std::int64_t packets = n / L;
    for (std::int64_t ip = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; ip < packets; ip += step) {
        const std::int64_t i = ip * L;
        const Packet<T, L> wp = load_packet<T, L>(w + i);
        double wi[L];
#pragma unroll
        for (int l = 0; l < L; ++l) wi[l] = static_cast<double>(wp.v[l]);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (j < count) {
                const Packet<T, L> vp = load_packet_nt<T, L>(V + j * ld + i);
#pragma unroll
                for (int l = 0; l < L; ++l) wi[l] = __builtin_fma(-h[j], static_cast<double>(vp.v[l]), wi[l]);
            }
        }
        Packet<T, L> wt;
#pragma unroll
        for (int l = 0; l < L; ++l) wt.v[l] = static_cast<T>(wi[l]);
        if (final_here) {
            Packet<T, L> vn;
#pragma unroll
            for (int l = 0; l < L; ++l) vn.v[l] = static_cast<T>(static_cast<double>(wt.v[l]) * inv);
            store_packet<T, L>(vnext + i, vn);
            continue;
        }
        store_packet<T, L>(w + i, wt);
        if (reorth) {
#pragma unroll
            for (int j = 0; j < K - 1; ++j) {
                if (j < count) {
                    const Packet<T, L> vp = load_packet_nt<T, L>(V + j * ld + i);
#pragma unroll
                    for (int l = 0; l < L; ++l)
                        red[j] = __builtin_fma(static_cast<double>(wt.v[l]), static_cast<double>(vp.v[l]), red[j]);
                }
            }
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const double wr = static_cast<double>(wt.v[l]);
                red[K - 1] = __builtin_fma(wr, wr, red[K - 1]);
            }
        }
    }
    dev::store_partials<K>(red, 0, scratch);
}

template<int K>
struct FinB {
    FgState* s;
    int k;
    int count;
    __device__ void operator()(const double* t) const
    {
        if (threadIdx.x != 0 || !s->do_reorth) return;
        double nsq = t[K - 1];
        for (int j = 0; j < count; ++j) {
            s->corr[j] = t[j];
            Hd(s, j, k) += t[j];
            nsq -= t[j] * t[j];
        }
        s->nsq = nsq;
        s->reorth_count += 1;
        decide(s, k);
    }
};

// Pass C: second projection, then v_{k+1} (direct) or the verified norm; then
// the Givens rotation and residual estimate (krylov_solvers.hpp:369-433).
template<int K, class T, int L>
__global__ __launch_bounds__(256) void k_fg_pass_c(const T* __restrict__ V, std::int64_t ld, int count,
                                                   T* __restrict__ w, T* __restrict__ vnext, std::int64_t n,
                                                   FgState* s, int k, ReduceScratch scratch)
{
    if (s->halt) return;
    __shared__ double c[K];
    const bool reorth = s->do_reorth != 0;
    if (threadIdx.x < count) c[threadIdx.x] = reorth ? s->corr[threadIdx.x] : 0.0;
    __syncthreads();
    const bool direct = s->direct != 0;
    const bool verify = s->verify != 0;
    const double inv = s->inv_h;
    double red[2] = {0.0, 0.0};
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    // pass B already wrote v_{k+1} when there was nothing left to project
    const std::int64_t packets = (!reorth && direct) ? 0 : n / L;
    for (std::int64_t ip = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; ip < packets; ip += step) {
        const std::int64_t i = ip * L;
        const Packet<T, L> wp = load_packet<T, L>(w + i);
        double wi[L];
#pragma unroll
        for (int l = 0; l < L; ++l) wi[l] = static_cast<double>(wp.v[l]);
        if (reorth) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j < count) {
                    const Packet<T, L> vp = load_packet_nt<T, L>(V + j * ld + i);
#pragma unroll
                    for (int l = 0; l < L; ++l) wi[l] = __builtin_fma(-c[j], static_cast<double>(vp.v[l]), wi[l]);
                }
            }
        }
        Packet<T, L> wt;
#pragma unroll
        for (int l = 0; l < L; ++l) wt.v[l] = static_cast<T>(wi[l]);
        if (direct) {
            Packet<T, L> vn;
#pragma unroll
            for (int l = 0; l < L; ++l) vn.v[l] = static_cast<T>(static_cast<double>(wt.v[l]) * inv);
            store_packet<T, L>(vnext + i, vn);
        } else {
            store_packet<T, L>(w + i, wt);
            if (verify) {
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    const double wr = static_cast<double>(wt.v[l]);
                    red[0] = __builtin_fma(wr, wr, red[0]);
                    red[1] = dev::combine<dev::Combine::max>(red[1], __builtin_fabs(wr));
                }
            }
        }
    }
    dev::store_partials<2>(red, 2, scratch);
}

template<class T>
struct FinC {
    FgState* s;
    int k;
    const T* w;
    std::int64_t n;
    double* snap_h;  // Arnoldi snapshot columns (stride kLd), or null
    __device__ void operator()(const double* t) const
    {
        double verified = 0.0;
        if (s->verify) verified = dev::block_norm(t[0], t[1], w, n);
        __syncthreads();  // every thread has read the state before thread 0 updates it
        if (threadIdx.x != 0) return;
        if (s->verify) {
            s->verif_count += 1;
            settle_h(s, k, verified, false);
            if (s->stop) return;
        }
        // the unrotated column, as ArnoldiSnapshot::record_column stores it
        // before the Givens rotations (krylov_solvers.hpp:361-366)
        if (snap_h)
            for (int row = 0; row <= k + 1; ++row) snap_h[k * kLd + row] = Hd(s, row, k);
        for (int row = 0; row < k; ++row) {
            const double first = s->cs[row] * Hd(s, row, k) + s->sn[row] * Hd(s, row + 1, k);
            Hd(s, row + 1, k) = -s->sn[row] * Hd(s, row, k) + s->cs[row] * Hd(s, row + 1, k);
            Hd(s, row, k) = first;
        }
        const double den = hypot(Hd(s, k, k), Hd(s, k + 1, k));
        if (!__builtin_isfinite(den)) {
            s->stop = 1;
            s->reason = reason_non_finite_scalar;
            s->halt = 1;
            return;
        }
        if (den == 0.0) {
            s->abreak = 1;
            s->cs[k] = 1.0;
            s->sn[k] = 0.0;
        } else {
            s->cs[k] = Hd(s, k, k) / den;
            s->sn[k] = Hd(s, k + 1, k) / den;
        }
        Hd(s, k, k) = s->cs[k] * Hd(s, k, k) + s->sn[k] * Hd(s, k + 1, k);
        Hd(s, k + 1, k) = 0.0;
        s->g[k + 1] = -s->sn[k] * s->g[k];
        s->g[k] *= s->cs[k];
        const double est = __builtin_fabs(s->g[k + 1]);
        s->recursive = est;
        s->columns = k + 1;
        if (est <= s->threshold || s->abreak) {
            s->cycle_done = 1;
            s->halt = 1;
        }
    }
};

// Pass D: v_{k+1} = w / h after a verified norm.
template<class T>
__global__ __launch_bounds__(256) void k_fg_pass_d(const T* __restrict__ w, T* __restrict__ vnext, std::int64_t n,
                                                   const FgState* s)
{
    // after the cycle has stopped the next basis vector is not needed
    if (s->halt || s->stop || !s->need_scale) return;
    const double inv = s->inv_h;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        vnext[i] = static_cast<T>(static_cast<double>(w[i]) * inv);
}

struct Coefficients {
    double y[kMaxRestart];
};

// x = base + (sum_j y_j Z_j with Neumaier compensation) (detail::gmres_update,
// krylov_solvers.hpp:23-40).
template<class T>
__global__ __launch_bounds__(256) void k_fg_update(T* __restrict__ x, const T* __restrict__ Z, std::int64_t ld,
                                                   int count, Coefficients coef, std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        double sum = 0.0, comp = 0.0;
        // eight basis values loaded before the sequential compensated sums
        // (an index past count re-reads the last column and is not summed)
        constexpr int C = 8;
        for (int j0 = 0; j0 < count; j0 += C) {
            T zv[C];
#pragma unroll
            for (int u = 0; u < C; ++u) {
                const std::int64_t jj = j0 + u < count ? j0 + u : count - 1;
                if constexpr (sizeof(T) == 8) zv[u] = Z[jj * ld + i];
                else zv[u] = __builtin_nontemporal_load(&Z[jj * ld + i]);
            }
#pragma unroll
            for (int u = 0; u < C; ++u) {
                if (j0 + u >= count) break;
                const double term = coef.y[j0 + u] * static_cast<double>(zv[u]);
                const double updated = sum + term;
                comp += __builtin_fabs(sum) >= __builtin_fabs(term) ? (sum - updated) + term : (term - updated) + sum;
                sum = updated;
            }
        }
        x[i] = static_cast<T>(static_cast<double>(x[i]) + (sum + comp));
    }
}

// r = b - A x with its norm, for either vector type.
template<class X>
struct EpiResidualT : dev::EpiNone {
    static constexpr int K = 2;
    static constexpr std::uint64_t kMaxMask = 2;
    X* y;
    const X* b;
    std::int64_t n;
    double* norm_out;
    template<class A>
    __device__ void elem(std::int64_t i, A ax, double* red) const
    {
        const X r = static_cast<X>(b[i] - static_cast<X>(ax));
        y[i] = r;
        const double rd = static_cast<double>(r);
        red[0] = __builtin_fma(rd, rd, red[0]);
        red[1] = dev::combine<dev::Combine::max>(red[1], __builtin_fabs(rd));
    }
    template<int BS, class A>
    __device__ void elem_row(std::int64_t i0, const A (&ax)[BS], double* red) const
    {
        X r[BS];
        dev::load_row<BS>(b + i0, r);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            r[c] = static_cast<X>(r[c] - static_cast<X>(ax[c]));
            const double rd = static_cast<double>(r[c]);
            red[0] = __builtin_fma(rd, rd, red[0]);
            red[1] = dev::combine<dev::Combine::max>(red[1], __builtin_fabs(rd));
        }
        dev::store_row<BS>(y + i0, r);
    }
    __device__ void finalize(const double* t) const
    {
        const double nrm = dev::block_norm(t[0], t[1], y, n);
        if (threadIdx.x == 0 && norm_out) *norm_out = nrm;
    }
};

__global__ void k_stop_copy64(const double* __restrict__ in, double* __restrict__ out, std::int64_t n, const int* stop)
{
    if (stop && *stop) return;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) out[i] = in[i];
}

__global__ void k_stop_copy32(const float* __restrict__ in, float* __restrict__ out, std::int64_t n, const int* stop)
{
    if (stop && *stop) return;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) out[i] = in[i];
}

__global__ void k_add_f32_to_f64(double* __restrict__ x, const float* __restrict__ d, std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        x[i] = x[i] + static_cast<double>(d[i]);
}

struct Counters {
    std::uint64_t ops = 0, pcs = 0, reds = 0, reorth = 0, verif = 0;
};

// The FGMRES algorithm over vectors of type T.  op(x, y, stop) computes y = A x;
// residual(b, x, r, norm_slot) computes r = b - A x and its norm; pc(in, out,
// stop) applies the preconditioner.
// Per-engine device workspace (Krylov basis, state) kept across runs, plus
// the pinned flags and events of the throttled launch loop.
struct EngineWork {
    DeviceArray<unsigned char> vecs;
    std::int64_t ld = -1, owned = -1;  // vector stride and owned length the zeroed ghosts fit
    DeviceArray<FgState> dstate;
    PinnedBuffer flags;
    hipEvent_t ev[2] = {nullptr, nullptr};
    EngineWork() = default;
    EngineWork(const EngineWork&) = delete;
    EngineWork& operator=(const EngineWork&) = delete;
    ~EngineWork()
    {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// The context's FGMRES workspace for elements of `bytes` bytes.
std::shared_ptr<EngineWork> engine_work(owtk_context ctx, std::size_t bytes)
{
    std::shared_ptr<void>& slot = ctx->fg_work[bytes == 8 ? 1 : 0];
    if (!slot) slot = std::make_shared<EngineWork>();
    return std::static_pointer_cast<EngineWork>(slot);
}

// Columns launched between two looks at the device's stop flag.  The host
// waits for the check two blocks back, so the GPU always has a block queued
// while at most two blocks of no-op launches follow a stop.
constexpr int kCheckEvery = 2;

template<class T>
struct FgEngine {
    owtk_context ctx;
    std::shared_ptr<EngineWork> work;  // the context's workspace for T
    bool may_reorth = true;             // Gram-Schmidt modes 0 and 1 can re-orthogonalise
    std::int64_t n = 0;       // owned size
    std::int64_t local = 0;   // local size (SpMV inputs)
    std::function<void(const T*, T*, const int*)> op;
    std::function<void(const T*, const T*, T*, double*)> residual;
    std::function<void(const T*, T*, const int*)> pc;
    double* snap_h = nullptr;           // snapshot Hessenberg columns, when one is recorded

    // `snap` (optional) receives the last cycle's Arnoldi relation; an
    // application convergence test in `o` sees every iterate (one host round
    // trip per column).  Both need FP64 vectors.
    int run(const T* b, T* x, const owtk_solver_options* o, owtk_solver_result* res, Counters& cnt,
            owtk_arnoldi_snapshot snap = nullptr)
    {
        init_result(res);
        snap_h = nullptr;
        if (!valid_options(o) || o->restart == 0) {
            snapshot_clear(snap);
            return OWTK_OK;
        }
        if (o->restart > static_cast<std::uint64_t>(kMaxRestart))
            throw unsupported_error("restart above 60 is not supported on the device path");
        const bool app = o->convergence_test != nullptr;
        if ((app || snap) && sizeof(T) != 8)
            throw unsupported_error("Arnoldi snapshots and application tests need FP64 vectors");
        const hipStream_t st = ctx->stream;
        const auto scratch = ctx->scratch();
        const int m = static_cast<int>(o->restart);
        const std::int64_t ld = local;
        may_reorth = o->gmres_orthogonalization == 0 || o->gmres_orthogonalization == 1;
        if (!work) work = engine_work(ctx, sizeof(T));
        const std::size_t need = static_cast<std::size_t>((2 * m + 3) * ld) * sizeof(T);
        if (work->vecs.bytes() < need || work->ld != ld || work->owned != n) {
            // ghost entries of the basis are never written, so after one zero
            // fill they stay zero for every later run with the same layout
            if (work->vecs.bytes() < need) work->vecs.allocate(need);
            kern::fill_zero_bytes(st, work->vecs.get(), work->vecs.bytes());
            work->ld = ld;
            work->owned = n;
        }
        if (!work->dstate.get()) work->dstate.allocate(1);
        if (!work->ev[0]) {
            for (auto& e : work->ev) OWTK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            work->flags.ensure(2 * sizeof(int));
        }
        T* r = reinterpret_cast<T*>(work->vecs.get());
        T* w = r + ld;
        T* V = w + ld;                  // m+1 vectors
        T* Z = V + static_cast<std::int64_t>(m + 1) * ld;  // m vectors
        DeviceArray<FgState>& dstate = work->dstate;
        int* flag = static_cast<int*>(work->flags.ptr);
        ctx->pinned.ensure(sizeof(FgState) + 256);
        auto* hs = static_cast<FgState*>(ctx->pinned.ptr);
        double* d_norm = ctx->scalars.get() + 16;
        const int* halt = &dstate.get()->halt;
        auto read_norm = [&]() {
            double v;
            OWTK_HIP(hipMemcpyAsync(&v, d_norm, sizeof(double), hipMemcpyDeviceToHost, st));
            ctx->sync();
            return v;
        };
        auto mark = [&](int reason) {
            res->status = status_breakdown;
            res->breakdown_reason = reason;
        };
        auto finish = [&](int status) {
            if (status >= 0) res->status = status;
            res->operator_applications = cnt.ops;
            res->preconditioner_applications = cnt.pcs;
            res->global_reductions = cnt.reds;
            res->reorthogonalizations = cnt.reorth;
            res->arnoldi_norm_verifications = cnt.verif;
            return OWTK_OK;
        };
        // application_convergence (krylov_solvers.hpp:122-143): the test sees
        // `cand`; when it accepts, the true residual and ||b|| are measured
        auto app_check = [&](const T* cand, std::uint64_t iteration) {
            ctx->sync();
            if (!o->convergence_test(o->convergence_test_data, iteration, reinterpret_cast<const double*>(cand),
                                     static_cast<std::size_t>(n)))
                return false;
            residual(b, cand, r, d_norm);
            ++cnt.ops;
            if constexpr (sizeof(T) == 8) kern::norm_to(st, scratch, reinterpret_cast<const double*>(b), n, d_norm + 1);
            double two[2];
            OWTK_HIP(hipMemcpyAsync(two, d_norm, sizeof(two), hipMemcpyDeviceToHost, st));
            ctx->sync();
            cnt.reds += 2;
            set_residual(res, res->recursive_residual_norm, two[0], two[1]);
            if (!std::isfinite(two[0]) || !std::isfinite(two[1])) {
                mark(reason_non_finite_scalar);
            } else {
                res->status = status_converged;
                res->converged_by_application = 1;
            }
            return true;
        };
        if (app && app_check(x, 0)) {
            snapshot_clear(snap);
            return finish(-1);
        }
        // ||b - A x|| and ||b|| in two slots, read back together
        residual(b, x, r, d_norm);
        ++cnt.ops;
        if constexpr (sizeof(T) == 8) kern::norm_to(st, scratch, reinterpret_cast<const double*>(b), n, d_norm + 1);
        else kern::norm_to(st, scratch, reinterpret_cast<const float*>(b), n, d_norm + 1);
        double pair[2];
        OWTK_HIP(hipMemcpyAsync(pair, d_norm, sizeof(pair), hipMemcpyDeviceToHost, st));
        ctx->sync();
        const double initial = pair[0], rhs_norm = pair[1];
        cnt.reds += 2;
        // with an application test only an exactly solved system stops on the
        // residual (convergence_threshold, core.hpp:549-560)
        double thr = threshold_of(rhs_norm, o);
        if (app && std::isfinite(thr)) thr = 0.0;
        res->initial_residual_norm = initial;
        set_residual(res, initial, initial, rhs_norm);
        if (!std::isfinite(initial) || !std::isfinite(thr)) {
            mark(reason_non_finite_scalar);
            return finish(-1);
        }
        if (initial <= thr) {
            snapshot_clear(snap);
            return finish(status_converged);
        }
        // the candidate iterate x + Z y formed for the application test
        DeviceArray<T> cand;
        if (app) cand.allocate(static_cast<std::size_t>(ld));
        auto account = [&](int ran) {
            cnt.pcs += ran;
            cnt.ops += ran;
            cnt.reds += ran + hs->reorth_count + hs->verif_count;
            cnt.reorth += hs->reorth_count;
            cnt.verif += hs->verif_count;
        };
        double beta = initial;
        const unsigned int g = dev::stream_grid(n, 4);
        // 1024 blocks: more would multiply the per-block K-value reductions and
        // the finisher's partial sums faster than they add bandwidth (the
        // learned+mixed solve on the 2M-vertex grid: 512 / 768 / 1024 / 2048 /
        // 4096 blocks took 9.43 / 8.93 / 8.72 / 8.95 / 9.36 ms)
        std::int64_t red_blocks = (n + 4 * kBlock - 1) / (4 * kBlock);
        if (red_blocks > 1024) red_blocks = 1024;
        if (red_blocks < 1) red_blocks = 1;
        const unsigned int gr = static_cast<unsigned int>(red_blocks);
        while (res->iterations < o->maximum_iterations) {
            if (snap) {
                snapshot_begin(snap, m, ld, n, st);
                snap_h = snap->h.get();
            }
            // beta = ||r|| (the reference re-reduces it; same vector, same value)
            ++cnt.reds;
            k_fg_v0<T><<<g, kBlock, 0, st>>>(V, r, n, beta);
            OWTK_LAUNCH_CHECK();
            std::memset(hs, 0, sizeof(FgState));
            hs->g[0] = beta;
            hs->threshold = thr;
            hs->breakdown_tolerance = o->breakdown_tolerance;
            hs->mode = o->gmres_orthogonalization;
            hs->recursive = res->recursive_residual_norm;
            OWTK_HIP(hipMemcpyAsync(dstate.get(), hs, sizeof(FgState), hipMemcpyHostToDevice, st));
            const std::uint64_t left = o->maximum_iterations - res->iterations;
            const int cols = static_cast<int>(std::min<std::uint64_t>(static_cast<std::uint64_t>(m), left));
            bool have_prev = false;
            for (int k = 0; k < cols; ++k) {
                T* vk = V + static_cast<std::int64_t>(k) * ld;
                T* zk = Z + static_cast<std::int64_t>(k) * ld;
                T* vnext = V + static_cast<std::int64_t>(k + 1) * ld;
                pc(vk, zk, halt);
                op(zk, w, halt);
                launch_passes(k, V, ld, w, vnext, dstate.get(), gr, st, scratch);
                if (app) {
                    // the iterate after column k, formed without touching x
                    // (krylov_solvers.hpp:398-427)
                    OWTK_HIP(hipMemcpyAsync(hs, dstate.get(), sizeof(FgState), hipMemcpyDeviceToHost, st));
                    ctx->sync();
                    if (hs->stop) break;
                    res->recursive_residual_norm = hs->recursive;
                    Coefficients cc{};
                    bool nonsingular = true;
                    for (int row = k + 1; row-- > 0;) {
                        double value = hs->g[row];
                        for (int j = row + 1; j <= k; ++j) value -= hs->H[j * kLd + row] * cc.y[j];
                        if (hs->H[row * kLd + row] == 0.0) {
                            nonsingular = false;
                            break;
                        }
                        cc.y[row] = value / hs->H[row * kLd + row];
                    }
                    if (nonsingular) {
                        OWTK_HIP(hipMemcpyAsync(cand.get(), x, static_cast<std::size_t>(ld) * sizeof(T),
                                                hipMemcpyDeviceToDevice, st));
                        k_fg_update<T><<<g, kBlock, 0, st>>>(cand.get(), Z, ld, k + 1, cc, n);
                        OWTK_LAUNCH_CHECK();
                        const std::uint64_t iteration = res->iterations + static_cast<std::uint64_t>(k) + 1;
                        if (app_check(cand.get(), iteration)) {
                            account(k + 1);
                            res->iterations = iteration;
                            if constexpr (sizeof(T) == 8)
                                if (snap) snapshot_finish(snap, Z, k + 1, ctx);
                            OWTK_HIP(hipMemcpyAsync(x, cand.get(), static_cast<std::size_t>(n) * sizeof(T),
                                                    hipMemcpyDeviceToDevice, st));
                            return finish(-1);
                        }
                    }
                    if (hs->cycle_done) break;
                    continue;
                }
                if ((k + 1) % kCheckEvery == 0 && k + 1 < cols) {
                    // the launches after a stop are no-ops: stop issuing them
                    const int slot = ((k + 1) / kCheckEvery) & 1;
                    OWTK_HIP(hipMemcpyAsync(&flag[slot], halt, sizeof(int), hipMemcpyDeviceToHost, st));
                    OWTK_HIP(hipEventRecord(work->ev[slot], st));
                    if (have_prev) {
                        OWTK_HIP(hipEventSynchronize(work->ev[slot ^ 1]));
                        if (flag[slot ^ 1] != 0) break;
                    }
                    have_prev = true;
                }
            }
            OWTK_HIP(hipMemcpyAsync(hs, dstate.get(), sizeof(FgState), hipMemcpyDeviceToHost, st));
            ctx->sync();
            const int done = hs->columns;
            // columns whose launches ran: all up to `done`, and on a hard stop
            // the failing column too
            const int ran = hs->stop ? std::max(done + 1, 1) : done;
            account(ran);
            res->iterations += static_cast<std::uint64_t>(hs->stop ? done : done);
            res->recursive_residual_norm = hs->recursive;
            if (hs->stop) {
                mark(hs->reason);
                return finish(-1);
            }
            const bool abreak = hs->abreak != 0;
            if constexpr (sizeof(T) == 8)
                if (snap) snapshot_finish(snap, Z, done, ctx);
            // back-substitution (krylov_solvers.hpp:439-457)
            Coefficients coef{};
            auto h = [&](int row, int col) { return hs->H[col * kLd + row]; };
            for (int rev = done; rev-- > 0;) {
                double value = hs->g[rev];
                for (int c = rev + 1; c < done; ++c) value -= h(rev, c) * coef.y[c];
                double scale = 0;
                for (int row = 0; row <= rev; ++row) scale = std::max(scale, std::abs(h(row, rev)));
                if (std::abs(h(rev, rev)) <= o->breakdown_tolerance * scale) {
                    mark(reason_singular_projected_system);
                    return finish(-1);
                }
                coef.y[rev] = value / h(rev, rev);
            }
            k_fg_update<T><<<g, kBlock, 0, st>>>(x, Z, ld, done, coef, n);
            OWTK_LAUNCH_CHECK();
            residual(b, x, r, d_norm);
            ++cnt.ops;
            ++cnt.reds;
            const double tn = read_norm();
            set_residual(res, res->recursive_residual_norm, tn, rhs_norm);
            if (!std::isfinite(tn)) {
                mark(reason_non_finite_scalar);
                return finish(-1);
            }
            if (tn <= thr) return finish(status_converged);
            if (abreak && !(tn < beta)) {
                mark(reason_arnoldi_invariant_subspace);
                return finish(-1);
            }
            beta = tn;
        }
        return finish(status_maximum_iterations);
    }

    template<int K>
    void passes(int k, T* V, std::int64_t ld, T* w, T* vnext, FgState* s, unsigned int gr, hipStream_t st,
                ReduceScratch scratch)
    {
        constexpr int L = 16 / static_cast<int>(sizeof(T));
        if (n % L == 0 && ld % L == 0) passes_l<K, L>(k, V, ld, w, vnext, s, gr, st, scratch);
        else passes_l<K, 1>(k, V, ld, w, vnext, s, gr, st, scratch);
    }

    template<int K, int L>
    void passes_l(int k, T* V, std::int64_t ld, T* w, T* vnext, FgState* s, unsigned int gr, hipStream_t st,
                  ReduceScratch scratch)
    {
        const int count = k + 1;
        const int* halt = &s->halt;
        k_fg_pass_a<K, T, L><<<gr, kBlock, 0, st>>>(V, ld, count, w, n, s, k, scratch);
        dev::launch_finish<K>(st, FinA<K>{s, k, count}, gr, 0, scratch, halt);
        k_fg_pass_b<K, T, L><<<gr, kBlock, 0, st>>>(V, ld, count, w, vnext, n, s, k, scratch);
        // FinB only acts after a re-orthogonalisation, which the single-pass
        // mode never does
        if (may_reorth) dev::launch_finish<K>(st, FinB<K>{s, k, count}, gr, 0, scratch, halt, &s->do_reorth);
        k_fg_pass_c<K, T, L><<<gr, kBlock, 0, st>>>(V, ld, count, w, vnext, n, s, k, scratch);
        dev::launch_finish<2>(st, FinC<T>{s, k, w, n, snap_h}, gr, 2, scratch, halt);
        k_fg_pass_d<T><<<dev::stream_grid(n, 4), kBlock, 0, st>>>(w, vnext, n, s);
        OWTK_LAUNCH_CHECK();
    }

    void launch_passes(int k, T* V, std::int64_t ld, T* w, T* vnext, FgState* s, unsigned int gr, hipStream_t st,
                       ReduceScratch scratch)
    {
        const int need = k + 2;
        if (need <= 4) passes<4>(k, V, ld, w, vnext, s, gr, st, scratch);
        else if (need <= 8) passes<8>(k, V, ld, w, vnext, s, gr, st, scratch);
        else if (need <= 16) passes<16>(k, V, ld, w, vnext, s, gr, st, scratch);
        else if (need <= 32) passes<32>(k, V, ld, w, vnext, s, gr, st, scratch);
        else passes<62>(k, V, ld, w, vnext, s, gr, st, scratch);
    }
};

FgEngine<double> fp64_engine(owtk_matrix m, owtk_precond pc)
{
    FgEngine<double> e;
    e.ctx = m->ctx;
    e.n = m->owned * m->bs;
    e.local = m->local() * m->bs;
    const std::int64_t n = e.n;
    const hipStream_t st = m->ctx->stream;
    e.op = [m](const double* x, double* y, const int* stop) {
        dev::EpiStore<double> ep;
        ep.y = y;
        ep.stop = stop;
        spmv_f64(m, x, ep);
    };
    e.residual = [m, n](const double* b, const double* x, double* r, double* norm_slot) {
        EpiResidualT<double> ep;
        ep.y = r;
        ep.b = b;
        ep.n = n;
        ep.norm_out = norm_slot;
        spmv_f64(m, x, ep);
    };
    e.pc = [pc, n, st](const double* in, double* out, const int* stop) { apply_pc_f64(pc, in, out, n, st, stop); };
    return e;
}

FgEngine<float> low_engine(owtk_matrix m, owtk_precond pc, owtk_dtype dtype)
{
    FgEngine<float> e;
    e.ctx = m->ctx;
    e.n = m->owned * m->bs;
    e.local = m->local() * m->bs;
    const std::int64_t n = e.n;
    const hipStream_t st = m->ctx->stream;
    e.op = [m, dtype](const float* x, float* y, const int* stop) {
        dev::EpiStore<float> ep;
        ep.y = y;
        ep.stop = stop;
        spmv_low(m, dtype, x, ep);
    };
    e.residual = [m, n, dtype](const float* b, const float* x, float* r, double* norm_slot) {
        EpiResidualT<float> ep;
        ep.y = r;
        ep.b = b;
        ep.n = n;
        ep.norm_out = norm_slot;
        spmv_low(m, dtype, x, ep);
    };
    e.pc = [pc, n, st](const float* in, float* out, const int* stop) { apply_pc_f32(pc, in, out, n, st, stop); };
    return e;
}

} // namespace

void copy_stop_f64(hipStream_t st, const double* in, double* out, std::int64_t n, const int* stop)
{
    if (in == out) return;
    k_stop_copy64<<<dev::stream_grid(n, 4), kBlock, 0, st>>>(in, out, n, stop);
    OWTK_LAUNCH_CHECK();
}

void copy_stop_f32(hipStream_t st, const float* in, float* out, std::int64_t n, const int* stop)
{
    if (in == out) return;
    k_stop_copy32<<<dev::stream_grid(n, 4), kBlock, 0, st>>>(in, out, n, stop);
    OWTK_LAUNCH_CHECK();
}

int fgmres_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, const owtk_solver_options* o,
               owtk_solver_result* res, owtk_arnoldi_snapshot snap = nullptr)
{
    const auto t0 = std::chrono::steady_clock::now();
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage != OWTK_F64)
        throw std::invalid_argument("FP64 solve needs an FP64 preconditioner");
    auto e = fp64_engine(m, pc);
    Counters cnt;
    e.run(b, x, o, res, cnt, snap);
    if (o->collect_timings)
        res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return OWTK_OK;
}

// FP64 iterative refinement around an inner low-precision FGMRES.
int fgmres_refined_run(owtk_matrix m, owtk_precond pc, owtk_precond fallback, const double* b, double* x,
                       const owtk_refinement_options* ro, owtk_solver_result* res)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    const owtk_solver_options* o = &ro->outer;
    if (!valid_options(o)) return OWTK_OK;
    const owtk_dtype dtype = static_cast<owtk_dtype>(ro->inner_storage);
    if (dtype != OWTK_F32 && dtype != OWTK_BF16) throw std::invalid_argument("inner storage must be F32 or BF16");
    if ((dtype == OWTK_F32 && !m->v32.get()) || (dtype == OWTK_BF16 && !m->v16.get())) {
        const int rc = owtk_matrix_add_storage(m, dtype);
        if (rc != OWTK_OK) throw std::runtime_error("could not add low-precision matrix storage");
    }
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage == OWTK_F64)
        throw std::invalid_argument("inner preconditioner needs F32 or BF16 storage");
    if (fallback && fallback->kind != OWTK_PC_IDENTITY && fallback->storage == OWTK_F64)
        throw std::invalid_argument("fallback preconditioner needs F32 or BF16 storage");
    if (fallback && !(ro->fallback_ratio > 0.0 && ro->fallback_ratio <= 1.0))
        throw std::invalid_argument("fallback_ratio must be in (0, 1]");
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const std::int64_t n = m->owned * m->bs;
    const std::int64_t L = m->local() * m->bs;
    // r (FP64), rf (r in FP32) and the correction d (FP32), each local-length
    // with zero ghosts, plus the fallback's saved x: the context keeps them
    // between solves and zero-fills them when the layout changes
    const auto round256 = [](std::size_t b) { return (b + 255) & ~static_cast<std::size_t>(255); };
    const std::size_t r_bytes = round256(static_cast<std::size_t>(L) * 8);
    const std::size_t f_bytes = round256(static_cast<std::size_t>(L) * 4);
    const std::size_t ws_bytes = r_bytes + 2 * f_bytes + (fallback ? round256(static_cast<std::size_t>(n) * 8) : 0);
    if (ctx->refine_ws.bytes() < ws_bytes || ctx->refine_layout != L || ctx->refine_owned != n) {
        if (ctx->refine_ws.bytes() < ws_bytes) ctx->refine_ws.allocate(ws_bytes);
        kern::fill_zero_bytes(st, ctx->refine_ws.get(), ctx->refine_ws.bytes());
        ctx->refine_layout = L;
        ctx->refine_owned = n;
    }
    unsigned char* ws = ctx->refine_ws.get();
    double* r = reinterpret_cast<double*>(ws);
    float* rf = reinterpret_cast<float*>(ws + r_bytes);
    float* d = reinterpret_cast<float*>(ws + r_bytes + f_bytes);
    double* x_prev = fallback ? reinterpret_cast<double*>(ws + r_bytes + 2 * f_bytes) : nullptr;
    double* d_norm = ctx->scalars.get() + 24;
    auto read = [&]() {
        double v;
        OWTK_HIP(hipMemcpyAsync(&v, d_norm, sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
        return v;
    };
    Counters cnt;
    auto launch_outer_residual = [&](double* slot) {
        dev::EpiResidual e;
        e.y = r;
        e.b = b;
        e.n = n;
        e.norm_out = slot;
        spmv_f64(m, x, e);
        ++cnt.ops;
        ++cnt.reds;
    };
    auto outer_residual = [&]() {
        launch_outer_residual(d_norm);
        return read();
    };
    // ||b|| and the initial ||b - A x||, read back together
    kern::norm_to(st, scratch, b, n, d_norm);
    ++cnt.reds;
    launch_outer_residual(d_norm + 1);
    double pair[2];
    OWTK_HIP(hipMemcpyAsync(pair, d_norm, sizeof(pair), hipMemcpyDeviceToHost, st));
    ctx->sync();
    const double rhs_norm = pair[0];
    const double thr = threshold_of(rhs_norm, o);
    double rn = pair[1];
    res->initial_residual_norm = rn;
    set_residual(res, rn, rn, rhs_norm);
    auto finish = [&](int status) {
        res->status = status;
        res->operator_applications = cnt.ops;
        res->preconditioner_applications = cnt.pcs;
        res->global_reductions = cnt.reds;
        res->reorthogonalizations = cnt.reorth;
        res->arnoldi_norm_verifications = cnt.verif;
        if (o->collect_timings)
            res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return OWTK_OK;
    };
    if (!std::isfinite(rn) || !std::isfinite(thr)) {
        res->breakdown_reason = reason_non_finite_scalar;
        return finish(status_breakdown);
    }
    auto inner = low_engine(m, pc, dtype);
    std::unique_ptr<FgEngine<float>> backup;  // built on the first switch
    bool use_fallback = false;
    owtk_solver_options io = *o;
    io.relative_tolerance = ro->inner_relative_tolerance;
    io.absolute_tolerance = 0.0;
    io.restart = static_cast<std::uint64_t>(ro->inner_restart);
    io.maximum_iterations = ro->inner_maximum_iterations;
    std::uint64_t steps = 0;
    while (!(rn <= thr)) {
        if (steps >= ro->maximum_refinements || res->inner_iterations >= o->maximum_iterations)
            return finish(status_maximum_iterations);
        kern::convert(st, rf, r, static_cast<std::size_t>(n));
        kern::fill_zero_bytes(st, d, f_bytes);
        owtk_solver_result ir;
        const bool guarded_step = fallback && !use_fallback;
        if (guarded_step)
            OWTK_HIP(hipMemcpyAsync(x_prev, x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (use_fallback) backup->run(rf, d, &io, &ir, cnt);
        else inner.run(rf, d, &io, &ir, cnt);
        res->inner_iterations += ir.iterations;
        res->iterations = res->inner_iterations;
        ++steps;
        res->refinement_steps = steps;
        if (use_fallback) ++res->fallback_refinements;
        const bool inner_failed = ir.status == status_breakdown && ir.iterations == 0;
        if (inner_failed && !guarded_step) {
            res->breakdown_reason = ir.breakdown_reason;
            return finish(status_breakdown);
        }
        if (!inner_failed) {
            k_add_f32_to_f64<<<dev::stream_grid(n, 4), kBlock, 0, st>>>(x, d, n);
            OWTK_LAUNCH_CHECK();
        }
        const double previous = rn;
        rn = outer_residual();
        res->recursive_residual_norm = ir.true_residual_norm;
        set_residual(res, ir.true_residual_norm, rn, rhs_norm);
        if (guarded_step && (inner_failed || !std::isfinite(rn) || !(rn <= ro->fallback_ratio * previous))) {
            // the true residual judges the step: undo it unless it helped,
            // then continue with the fallback preconditioner
            if (!(rn < previous)) {
                OWTK_HIP(hipMemcpyAsync(x, x_prev, n * sizeof(double), hipMemcpyDeviceToDevice, st));
                rn = outer_residual();
                set_residual(res, rn, rn, rhs_norm);
            }
            use_fallback = true;
            backup = std::make_unique<FgEngine<float>>(low_engine(m, fallback, dtype));
            continue;
        }
        if (!std::isfinite(rn)) {
            res->breakdown_reason = reason_non_finite_scalar;
            return finish(status_breakdown);
        }
        if (!(rn < previous)) {
            // no progress: the low-precision solve cannot reduce this residual further
            res->breakdown_reason = reason_none;
            return finish(status_diverged);
        }
    }
    return finish(status_converged);
}

// ------------------------------------------------------------ recycling --
// RecycleSpace and recycled_fgmres (recycling.hpp:24-360): U and C = A U
// (C orthonormal) on the device.  Inside the solve the projected
// preconditioner forms z = M v and A z, takes the coefficients C^T A z with
// one multi-dot pass and projects both, all on the device (no host read);
// the operator then returns the projected image.
struct RecycleVectors {
    std::deque<DeviceArray<double>> U, C;  // owned + ghost length, ghosts zero
};

namespace {

// out = fma(-c_i, U_i, out), img = fma(-c_i, C_i, img) for i in order,
// c read from device memory; sign = +1 forms x = fma(c_i, U_i, x) only
__global__ __launch_bounds__(256) void k_rc_project(double* __restrict__ out, double* __restrict__ img,
                                                    kern::BasisPointers u, kern::BasisPointers c,
                                                    const double* __restrict__ coef, int count, std::int64_t n,
                                                    double sign, const int* stop)
{
    if (stop && *stop) return;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += step) {
        double o = out[t];
        double g = img ? img[t] : 0.0;
        for (int i = 0; i < count; ++i) {
            const double a = sign * coef[i];
            o = __builtin_fma(a, u.v[i][t], o);
            if (img) g = __builtin_fma(a, c.v[i][t], g);
        }
        out[t] = o;
        if (img) img[t] = g;
    }
}

__global__ __launch_bounds__(256) void k_rc_scale2(double* __restrict__ a, double* __restrict__ b, double s,
                                                   std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += step) {
        a[t] = a[t] * s;
        b[t] = b[t] * s;
    }
}

void count(owtk_solver_result* tel, std::uint64_t ops, std::uint64_t reds)
{
    if (!tel) return;
    tel->operator_applications += ops;
    tel->global_reductions += reds;
}

} // namespace

} // namespace owtk

struct owtk_recycle_space_s {
    owtk_context ctx = nullptr;
    std::int64_t n = 0, local = 0;
    std::size_t capacity = 0;
    owtk::RecycleVectors v;
    owtk::DeviceArray<double> coef;  // capacity + 1 coefficients
    // vectors of `local` doubles kept for reuse: add_candidate takes its two
    // work vectors from here and rejected, dropped or replaced pairs return
    // here, so a steady sequence of solves allocates (and frees, which
    // synchronises the device) nothing
    std::vector<owtk::DeviceArray<double>> spare;
};

namespace owtk {
namespace {

kern::BasisPointers pointers(const std::deque<DeviceArray<double>>& list)
{
    kern::BasisPointers p{};
    for (std::size_t i = 0; i < list.size(); ++i) p.v[i] = list[i].get();
    return p;
}

double read_scalar(owtk_context ctx, const double* d)
{
    double v;
    OWTK_HIP(hipMemcpyAsync(&v, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    return v;
}

// A vector of the space's layout from its spares, or a new one.
DeviceArray<double> rc_take(owtk_recycle_space sp)
{
    if (sp->spare.empty()) return DeviceArray<double>(static_cast<std::size_t>(sp->local));
    DeviceArray<double> v = std::move(sp->spare.back());
    sp->spare.pop_back();
    return v;
}

// Back to the spares (at most two per recycled pair plus two).  Every use is
// on the context's one stream, so work already launched on a vector stays
// ordered before its next user's.
void rc_give(owtk_recycle_space sp, DeviceArray<double>&& v)
{
    if (v.get() && sp->spare.size() < 2 * sp->capacity + 2) sp->spare.push_back(std::move(v));
}

// add_candidate (recycling.hpp:79-135)
bool rc_add(owtk_recycle_space sp, owtk_matrix m, const double* cand, double tol, owtk_solver_result* tel)
{
    owtk_context ctx = sp->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const std::int64_t n = sp->n;
    DeviceArray<double> vec = rc_take(sp), img = rc_take(sp);
    auto reject = [&] {
        rc_give(sp, std::move(vec));
        rc_give(sp, std::move(img));
        return false;
    };
    kern::fill_zero_bytes(st, vec.get(), vec.bytes());
    kern::fill_zero_bytes(st, img.get(), img.bytes());
    OWTK_HIP(hipMemcpyAsync(vec.get(), cand, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    {
        dev::EpiStore<double> e;
        e.y = img.get();
        spmv_f64(m, vec.get(), e);
    }
    double* d_norm = ctx->scalars.get() + 24;
    kern::norm_to(st, scratch, img.get(), n, d_norm);
    const double orig = read_scalar(ctx, d_norm);
    count(tel, 1, 1);
    if (!std::isfinite(orig) || orig == 0.0) return reject();
    const int k = static_cast<int>(sp->v.C.size());
    const unsigned int g = dev::stream_grid(n, 4);
    for (int pass = 0; pass < 2 && k > 0; ++pass) {
        const auto cp = pointers(sp->v.C);
        kern::multidot_to(st, scratch, n, k, cp.v, img.get(), sp->coef.get());
        count(tel, 0, 1);
        k_rc_project<<<g, kBlock, 0, st>>>(vec.get(), img.get(), pointers(sp->v.U), cp, sp->coef.get(), k, n, -1.0,
                                           nullptr);
        OWTK_LAUNCH_CHECK();
    }
    kern::norm_to(st, scratch, img.get(), n, d_norm);
    const double nv = read_scalar(ctx, d_norm);
    count(tel, 0, 1);
    if (!std::isfinite(nv) || nv == 0.0 || nv / orig <= tol) return reject();
    k_rc_scale2<<<g, kBlock, 0, st>>>(img.get(), vec.get(), 1.0 / nv, n);
    OWTK_LAUNCH_CHECK();
    if (sp->v.C.size() == sp->capacity) {
        rc_give(sp, std::move(sp->v.C.front()));
        rc_give(sp, std::move(sp->v.U.front()));
        sp->v.C.pop_front();
        sp->v.U.pop_front();
    }
    sp->v.C.push_back(std::move(img));
    sp->v.U.push_back(std::move(vec));
    return true;
}

// refresh (recycling.hpp:65-73): C = A U rebuilt from the stored U
void rc_refresh(owtk_recycle_space sp, owtk_matrix m, owtk_solver_result* tel)
{
    RecycleVectors old = std::move(sp->v);
    sp->v = RecycleVectors{};
    for (const auto& u : old.U) rc_add(sp, m, u.get(), 128 * std::numeric_limits<double>::epsilon(), tel);
    for (auto& u : old.U) rc_give(sp, std::move(u));
    for (auto& c : old.C) rc_give(sp, std::move(c));
}

// project_initial_guess (recycling.hpp:138-163): x += U C^T r
void rc_project(owtk_recycle_space sp, const double* r, double* x, owtk_solver_result* tel)
{
    const int k = static_cast<int>(sp->v.C.size());
    if (k == 0) return;
    const hipStream_t st = sp->ctx->stream;
    const auto cp = pointers(sp->v.C);
    kern::multidot_to(st, sp->ctx->scratch(), sp->n, k, cp.v, r, sp->coef.get());
    count(tel, 0, 1);
    k_rc_project<<<dev::stream_grid(sp->n, 4), kBlock, 0, st>>>(x, nullptr, pointers(sp->v.U), cp, sp->coef.get(), k,
                                                                 sp->n, 1.0, nullptr);
    OWTK_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_rc_residual(const double* __restrict__ b, const double* __restrict__ ax,
                                                     double* __restrict__ r, std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += step)
        r[t] = b[t] - ax[t];
}

__global__ __launch_bounds__(256) void k_rc_diff(const double* __restrict__ x, const double* __restrict__ x0,
                                                 double* __restrict__ out, std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += step)
        out[t] = x[t] - x0[t];
}

// The projected matrices of the harmonic-Ritz extraction in one pass:
// P[i][j] = <A_i, B_j> for i < na, j < nb, with A_i = a + i lda and
// B_j = j < nb1 ? b1 + j ldb1 : b2 + (j - nb1) ldb2 (here A = AZ, B = [AZ | Z],
// so P = [(AZ)^T AZ | (AZ)^T Z]).  A multi-dot per column would read every
// vector once per column; here each is read once.  A block stages `rows`
// rows of all na + nb vectors in LDS ([row][vector]), then accumulates 4 x 4
// tiles of P over them: with fewer tiles than threads, the threads of a tile
// split the staged rows and are summed in thread order at the end; with
// more, each thread takes several tiles.  Per-block partials are summed in
// block order by k_gram_finish, so the result does not depend on scheduling.
constexpr int kGramBlocks = 2048;
constexpr int kGramLds = 48 * 1024;                         // staging bytes per block
constexpr int kGramMaxTilesPerThread = 2;                   // (60 / 4) x (120 / 4) = 450 tiles, 256 threads

int gram_rows(int vecs) { return std::max(1, std::min(128, kGramLds / (8 * vecs))); }

__global__ __launch_bounds__(256) void k_gram(const double* __restrict__ a, std::int64_t lda, int na,
                                              const double* __restrict__ b1, std::int64_t ldb1,
                                              const double* __restrict__ b2, std::int64_t ldb2, int nb1, int nb,
                                              std::int64_t n, int rows, double* __restrict__ partials)
{
    extern __shared__ double stage[];  // [rows][na + nb], later the per-thread tiles
    const int vecs = na + nb;
    const int ta = (na + 3) / 4, tb = (nb + 3) / 4, tiles = ta * tb;
    const int nt = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
    const int per = tiles >= nt ? 1 : nt / tiles;    // threads sharing a tile
    const int sub = tiles >= nt ? 0 : tid % per;     // this thread's share of the staged rows
    const int first = tiles >= nt ? tid : tid / per; // this thread's first tile
    const bool active = tiles >= nt || tid < tiles * per;
    double acc[kGramMaxTilesPerThread][16];
#pragma unroll
    for (int u = 0; u < kGramMaxTilesPerThread; ++u)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[u][e] = 0.0;
    for (std::int64_t r0 = static_cast<std::int64_t>(blockIdx.x) * rows; r0 < n;
         r0 += static_cast<std::int64_t>(gridDim.x) * rows) {
        // rows r0 .. r0 + rows - 1 of every vector (coalesced along the rows), zero past n
        for (int e = tid; e < rows * vecs; e += nt) {
            const int v = e / rows, r = e - v * rows;
            const std::int64_t row = r0 + r;
            const double* src = v < na ? a + v * lda
                                       : (v - na < nb1 ? b1 + (v - na) * ldb1 : b2 + (v - na - nb1) * ldb2);
            stage[r * vecs + v] = row < n ? src[row] : 0.0;
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int u = 0; u < kGramMaxTilesPerThread; ++u) {
                const int tile = first + u * nt;
                if (tile >= tiles || (u > 0 && tiles < nt)) break;
                const int i0 = 4 * (tile / tb), j0 = 4 * (tile % tb);
                for (int r = sub; r < rows; r += per) {
                    const double* row = stage + r * vecs;
                    double av[4], bv[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        av[q] = i0 + q < na ? row[i0 + q] : 0.0;
                        bv[q] = j0 + q < nb ? row[na + j0 + q] : 0.0;
                    }
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[u][4 * p + q] = __builtin_fma(av[p], bv[q], acc[u][4 * p + q]);
                }
            }
        }
        __syncthreads();
    }
    double* out = partials + static_cast<std::int64_t>(blockIdx.x) * na * nb;
    if (tiles >= nt) {
#pragma unroll
        for (int u = 0; u < kGramMaxTilesPerThread; ++u) {
            const int tile = first + u * nt;
            if (tile >= tiles) break;
            const int i0 = 4 * (tile / tb), j0 = 4 * (tile % tb);
            for (int p = 0; p < 4; ++p)
                for (int q = 0; q < 4; ++q)
                    if (i0 + p < na && j0 + q < nb) out[(i0 + p) * nb + j0 + q] = acc[u][4 * p + q];
        }
        return;
    }
    // the threads of a tile, summed in thread order (256 x 16 doubles fit the staging area)
    for (int e = 0; e < 16; ++e) stage[tid * 16 + e] = acc[0][e];
    __syncthreads();
    for (int t = tid; t < tiles * 16; t += nt) {
        const int tile = t >> 4, e = t & 15;
        double s = 0.0;
        for (int k = 0; k < per; ++k) s += stage[(tile * per + k) * 16 + e];
        const int i = 4 * (tile / tb) + (e >> 2), j = 4 * (tile % tb) + (e & 3);
        if (i < na && j < nb) out[i * nb + j] = s;
    }
}

__global__ __launch_bounds__(256) void k_gram_finish(const double* __restrict__ partials, int blocks, int count,
                                                     double* __restrict__ out)
{
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= count) return;
    double s = 0.0;
    for (int blk = 0; blk < blocks; ++blk) s += partials[static_cast<std::int64_t>(blk) * count + t];
    out[t] = s;
}

} // namespace

// recycled_fgmres (recycling.hpp:291-360)
int recycled_fgmres_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, owtk_recycle_space sp,
                        const owtk_solver_options* o, bool update, owtk_solver_result* res,
                        owtk_arnoldi_snapshot snap = nullptr)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    if (!valid_options(o) || o->restart == 0) return OWTK_OK;
    if (pc && pc->kind != OWTK_PC_IDENTITY && pc->storage != OWTK_F64)
        throw std::invalid_argument("FP64 solve needs an FP64 preconditioner");
    if (sp->n != m->owned * m->bs || sp->local != m->local() * m->bs)
        throw std::invalid_argument("recycle space vector layout mismatch");
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const std::int64_t n = sp->n;
    const unsigned int g = dev::stream_grid(n, 4);
    owtk_solver_result tel{};
    DeviceArray<double> x0(static_cast<std::size_t>(n));
    OWTK_HIP(hipMemcpyAsync(x0.get(), x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (!sp->v.C.empty()) {
        rc_refresh(sp, m, &tel);
        DeviceArray<double> work(static_cast<std::size_t>(sp->local)), r(static_cast<std::size_t>(n));
        dev::EpiStore<double> e;
        e.y = work.get();
        spmv_f64(m, x, e);
        ++tel.operator_applications;
        k_rc_residual<<<g, kBlock, 0, st>>>(b, work.get(), r.get(), n);
        OWTK_LAUNCH_CHECK();
        rc_project(sp, r.get(), x, &tel);
        ctx->sync();
    }
    auto eng = fp64_engine(m, pc);
    const int k = static_cast<int>(sp->v.C.size());
    DeviceArray<double> cached(static_cast<std::size_t>(sp->local));
    kern::fill_zero_bytes(st, cached.get(), cached.bytes());
    const auto up = pointers(sp->v.U), cp = pointers(sp->v.C);
    double* dcoef = sp->coef.get();
    const auto scratch = ctx->scratch();
    auto base_pc = eng.pc;
    auto base_op = eng.op;
    // RecycledPreconditioner::apply (recycling.hpp:236-262)
    eng.pc = [&, base_pc, base_op](const double* in, double* out, const int* stop) {
        base_pc(in, out, stop);
        base_op(out, cached.get(), stop);
        if (k > 0) {
            kern::multidot_to(st, scratch, n, k, cp.v, cached.get(), dcoef);
            k_rc_project<<<g, kBlock, 0, st>>>(out, cached.get(), up, cp, dcoef, k, n, -1.0, stop);
            OWTK_LAUNCH_CHECK();
        }
    };
    // RecycledOperator::apply (recycling.hpp:206-214): the cached projected image
    eng.op = [&](const double*, double* out, const int* stop) { copy_stop_f64(st, cached.get(), out, n, stop); };
    Counters cnt;
    eng.run(b, x, o, res, cnt, snap);
    res->operator_applications += tel.operator_applications;
    res->global_reductions += tel.global_reductions + (k > 0 ? res->preconditioner_applications : 0);
    if (update && res->status == status_converged) {
        DeviceArray<double> corr(static_cast<std::size_t>(n));
        k_rc_diff<<<g, kBlock, 0, st>>>(x, x0.get(), corr.get(), n);
        OWTK_LAUNCH_CHECK();
        rc_add(sp, m, corr.get(), o->breakdown_tolerance, res);
    }
    ctx->sync();
    if (o->collect_timings)
        res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return OWTK_OK;
}


// extract_harmonic_ritz (gcrodr.hpp:86-181): images A Z, the projected
// matrices (AZ)^T AZ and (AZ)^T Z in one pass over the vectors (k_gram), the
// eigenproblem on the host, then the candidates Z y in ascending |theta|,
// each group committed only when all its directions pass add_candidate.
std::uint64_t extract_harmonic_ritz(owtk_matrix m, owtk_arnoldi_snapshot sn, owtk_recycle_space sp,
                                    owtk_solver_result* tel)
{
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const int n = static_cast<int>(sn->columns);
    if (sp->n != m->owned * m->bs || sp->local != m->local() * m->bs)
        throw std::invalid_argument("recycle space vector layout mismatch");
    if (n == 0) {
        ctx->sync();
        sp->v = RecycleVectors{};
        return 0;
    }
    if (sn->ld != sp->local || sn->owned != sp->n) throw std::invalid_argument("snapshot vector layout mismatch");
    if (n > kMaxRestart) throw std::invalid_argument("Arnoldi snapshot holds too many columns");
    const std::int64_t local = sp->local, owned = sp->n;
    std::vector<const double*> Z(n), AZ(n);
    for (int c = 0; c < n; ++c) Z[c] = sn->basis.get() + static_cast<std::int64_t>(c) * local;
    DeviceArray<double> images(static_cast<std::size_t>(n) * local);
    kern::fill_zero_bytes(st, images.get(), images.bytes());
    for (int c = 0; c < n; ++c) {
        AZ[c] = images.get() + static_cast<std::int64_t>(c) * local;
        dev::EpiStore<double> e;
        e.y = const_cast<double*>(AZ[c]);
        spmv_f64(m, Z[c], e);
    }
    count(tel, static_cast<std::uint64_t>(n), 1);
    // P = [(AZ)^T AZ | (AZ)^T Z] (n x 2n, row-major) in one pass over the vectors
    const int grows = gram_rows(3 * n);
    const int blocks = static_cast<int>(std::max<std::int64_t>(1, std::min<std::int64_t>(
        (owned + grows - 1) / grows, kGramBlocks)));
    const int pairs = 2 * n * n;
    DeviceArray<double> partials(static_cast<std::size_t>(blocks) * pairs), dots(static_cast<std::size_t>(pairs));
    k_gram<<<blocks, kBlock, kGramLds, st>>>(images.get(), local, n, images.get(), local, sn->basis.get(), sn->ld, n,
                                             2 * n, owned, grows, partials.get());
    OWTK_LAUNCH_CHECK();
    k_gram_finish<<<(pairs + kBlock - 1) / kBlock, kBlock, 0, st>>>(partials.get(), blocks, pairs, dots.get());
    OWTK_LAUNCH_CHECK();
    std::vector<double> h_dots(dots.count);
    OWTK_HIP(hipMemcpyAsync(h_dots.data(), dots.get(), dots.bytes(), hipMemcpyDeviceToHost, st));
    ctx->sync();
    std::vector<double> harmonic(static_cast<std::size_t>(n) * n), metric(harmonic.size());
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) {
            harmonic[static_cast<std::size_t>(c) * n + r] = h_dots[static_cast<std::size_t>(r) * 2 * n + c];
            metric[static_cast<std::size_t>(c) * n + r] = h_dots[static_cast<std::size_t>(r) * 2 * n + n + c];
        }
    for (double v : harmonic)
        if (!std::isfinite(v)) throw std::runtime_error("nonfinite harmonic-Ritz projected matrix");
    for (double v : metric)
        if (!std::isfinite(v)) throw std::runtime_error("nonfinite harmonic-Ritz projected matrix");
    std::vector<double> re(n), im(n), be(n), right(static_cast<std::size_t>(n) * n);
    host::harmonic_ggev(n, harmonic.data(), metric.data(), re.data(), im.data(), be.data(), right.data());
    struct Group {
        double magnitude;
        int first, width;
    };
    std::vector<Group> groups;
    for (int e = 0; e < n;) {
        const int width = im[e] > 0.0 ? 2 : 1;
        const double magnitude = std::hypot(re[e], im[e]) / std::abs(be[e]);
        if (im[e] >= 0.0 && e + width <= n && be[e] != 0.0 && std::isfinite(magnitude))
            groups.push_back({magnitude, e, width});
        e += width;
    }
    std::stable_sort(groups.begin(), groups.end(),
                     [](const Group& a, const Group& b) { return a.magnitude < b.magnitude; });
    auto replacement = std::make_unique<owtk_recycle_space_s>();
    replacement->ctx = ctx;
    replacement->n = owned;
    replacement->local = local;
    replacement->capacity = sp->capacity;
    replacement->coef.allocate(sp->capacity + 1);
    replacement->spare = std::move(sp->spare);  // the replacement draws on the space's spares
    kern::BasisPointers zp{};
    for (int c = 0; c < n; ++c) zp.v[c] = Z[c];
    DeviceArray<double> coef(static_cast<std::size_t>(n)), cand(static_cast<std::size_t>(owned));
    const unsigned int g = dev::stream_grid(owned, 4);
    for (const Group& grp : groups) {
        if (replacement->v.C.size() + static_cast<std::size_t>(grp.width) > replacement->capacity) continue;
        // The group is tried in place: within the capacity add_candidate only
        // appends (a rejected candidate leaves the space as it was), so
        // taking back what a rejected group appended restores the space, as
        // discarding the reference's trial copy does.
        const std::size_t before = replacement->v.C.size();
        bool accepted = true;
        for (int part = 0; part < grp.width; ++part) {
            // candidate = sum_b y_b Z_b, added in basis order from zero
            OWTK_HIP(hipMemcpyAsync(coef.get(), right.data() + static_cast<std::size_t>(grp.first + part) * n,
                                    static_cast<std::size_t>(n) * sizeof(double), hipMemcpyHostToDevice, st));
            kern::fill_zero_bytes(st, cand.get(), cand.bytes());
            k_rc_project<<<g, kBlock, 0, st>>>(cand.get(), nullptr, zp, kern::BasisPointers{}, coef.get(), n, owned,
                                               1.0, nullptr);
            OWTK_LAUNCH_CHECK();
            accepted = rc_add(replacement.get(), m, cand.get(), 128 * std::numeric_limits<double>::epsilon(), tel)
                && accepted;
        }
        ctx->sync();  // the candidate and coefficients are reused by the next group
        if (!accepted) {
            while (replacement->v.C.size() > before) {
                rc_give(replacement.get(), std::move(replacement->v.C.back()));
                rc_give(replacement.get(), std::move(replacement->v.U.back()));
                replacement->v.C.pop_back();
                replacement->v.U.pop_back();
            }
        }
    }
    ctx->sync();
    for (auto& u : sp->v.U) rc_give(replacement.get(), std::move(u));
    for (auto& c : sp->v.C) rc_give(replacement.get(), std::move(c));
    sp->v = std::move(replacement->v);
    sp->spare = std::move(replacement->spare);
    return static_cast<std::uint64_t>(sp->v.C.size());
}

// gcrodr (gcrodr.hpp:185-212)
int gcrodr_run(owtk_matrix m, owtk_precond pc, const double* b, double* x, owtk_recycle_space sp,
               const owtk_solver_options* o, owtk_arnoldi_snapshot snap, owtk_solver_result* res)
{
    owtk_arnoldi_snapshot_s local_snap;
    local_snap.ctx = m->ctx;
    owtk_arnoldi_snapshot s = snap ? snap : &local_snap;
    recycled_fgmres_run(m, pc, b, x, sp, o, false, res, s);
    if (s->columns != 0) extract_harmonic_ritz(m, s, sp, res);
    return OWTK_OK;
}

} // namespace owtk

extern "C" {

void owtk_solver_options_default(owtk_solver_options* o)
{
    if (!o) return;
    o->relative_tolerance = 1e-8;
    o->absolute_tolerance = 0.0;
    o->breakdown_tolerance = 64.0 * std::numeric_limits<double>::epsilon();
    o->maximum_iterations = 1000;
    o->convergence_check_interval = 10;
    o->restart = 30;
    o->residual_replacement_interval = 50;
    o->gmres_orthogonalization = 0;
    o->collect_timings = 0;
    o->idr_shadow_space = 4;
    o->relaxation = 1.0;
    o->verify_true_residual = 1;
    o->reduction = OWTK_REDUCTION_SERIAL;
    o->convergence_test = nullptr;
    o->convergence_test_data = nullptr;
}

void owtk_refinement_options_default(owtk_refinement_options* o)
{
    if (!o) return;
    owtk_solver_options_default(&o->outer);
    o->outer.maximum_iterations = 100000;
    o->inner_storage = OWTK_F32;
    o->inner_restart = 30;
    o->inner_relative_tolerance = 1e-4;
    o->inner_maximum_iterations = 300;
    o->maximum_refinements = 50;
    o->fallback_ratio = 0.5;
}

int owtk_fgmres(owtk_matrix m, owtk_precond pc, const double* b, double* x, const owtk_solver_options* options,
                owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        require_serial_reduction(options, "fgmres");
        m->ctx->activate();
        fgmres_run(m, pc, b, x, options, result);
    });
}

int owtk_fgmres_with_snapshot(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                              const owtk_solver_options* options, owtk_arnoldi_snapshot snapshot,
                              owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        if (snapshot && snapshot->ctx != m->ctx) throw std::invalid_argument("snapshot belongs to another context");
        require_serial_reduction(options, "fgmres");
        m->ctx->activate();
        fgmres_run(m, pc, b, x, options, result, snapshot);
    });
}

int owtk_arnoldi_snapshot_create(owtk_context ctx, owtk_arnoldi_snapshot* out)
{
    return guarded([&] {
        if (!ctx || !out) throw std::invalid_argument("null argument");
        auto sn = std::make_unique<owtk_arnoldi_snapshot_s>();
        sn->ctx = ctx;
        *out = sn.release();
    });
}

int owtk_arnoldi_snapshot_destroy(owtk_arnoldi_snapshot snapshot)
{
    return guarded([&] {
        if (!snapshot) return;
        snapshot->ctx->activate();
        snapshot->ctx->sync();
        delete snapshot;
    });
}

int owtk_arnoldi_snapshot_clear(owtk_arnoldi_snapshot snapshot)
{
    return guarded([&] {
        if (!snapshot) throw std::invalid_argument("null argument");
        snapshot_clear(snapshot);
    });
}

int owtk_arnoldi_snapshot_columns(owtk_arnoldi_snapshot snapshot, uint64_t* columns)
{
    return guarded([&] {
        if (!snapshot || !columns) throw std::invalid_argument("null argument");
        *columns = static_cast<uint64_t>(snapshot->columns);
    });
}

int owtk_arnoldi_snapshot_h(owtk_arnoldi_snapshot snapshot, uint64_t row, uint64_t column, double* value)
{
    return guarded([&] {
        if (!snapshot || !value) throw std::invalid_argument("null argument");
        const auto cols = static_cast<uint64_t>(snapshot->columns);
        if (column >= cols || row > cols) throw std::out_of_range("Arnoldi snapshot entry");
        // rows below the recorded subdiagonal are the zeros begin_cycle wrote
        *value = row > column + 1 ? 0.0 : snapshot->host_h[column * owtk_arnoldi_snapshot_s::kStride + row];
    });
}

int owtk_arnoldi_snapshot_begin_cycle(owtk_arnoldi_snapshot snapshot, owtk_matrix layout, uint64_t maximum_columns)
{
    return guarded([&] {
        if (!snapshot || !layout) throw std::invalid_argument("null argument");
        if (snapshot->ctx != layout->ctx) throw std::invalid_argument("snapshot belongs to another context");
        if (maximum_columns > static_cast<uint64_t>(owtk_arnoldi_snapshot_s::kStride - 1))
            throw unsupported_error("Arnoldi snapshots hold at most 60 columns on the device path");
        layout->ctx->activate();
        owtk::snapshot_begin(snapshot, static_cast<int>(maximum_columns), layout->local() * layout->bs,
                             layout->owned * layout->bs, layout->ctx->stream);
        snapshot->host_h.assign(static_cast<std::size_t>(owtk_arnoldi_snapshot_s::kStride)
                                    * (owtk_arnoldi_snapshot_s::kStride - 1),
                                0.0);
        layout->ctx->sync();
    });
}

int owtk_arnoldi_snapshot_record_column(owtk_arnoldi_snapshot snapshot, uint64_t column, const double* values,
                                        size_t count)
{
    return guarded([&] {
        if (!snapshot || (count && !values)) throw std::invalid_argument("null argument");
        if (column >= static_cast<uint64_t>(snapshot->maximum_columns) || count != column + 2)
            throw std::invalid_argument("invalid Arnoldi snapshot column");
        for (std::size_t row = 0; row < count; ++row)
            snapshot->host_h[column * owtk_arnoldi_snapshot_s::kStride + row] = values[row];
    });
}

int owtk_arnoldi_snapshot_finish_cycle(owtk_arnoldi_snapshot snapshot, const double* const* basis, uint64_t columns)
{
    return guarded([&] {
        if (!snapshot || (columns && !basis)) throw std::invalid_argument("null argument");
        if (columns > static_cast<uint64_t>(snapshot->maximum_columns))
            throw std::invalid_argument("invalid Arnoldi snapshot basis");
        snapshot->ctx->activate();
        const hipStream_t st = snapshot->ctx->stream;
        for (uint64_t c = 0; c < columns; ++c) {
            if (!basis[c]) throw std::invalid_argument("invalid Arnoldi snapshot basis");
            OWTK_HIP(hipMemcpyAsync(snapshot->basis.get() + static_cast<std::int64_t>(c) * snapshot->ld, basis[c],
                                    static_cast<std::size_t>(snapshot->owned) * sizeof(double),
                                    hipMemcpyDeviceToDevice, st));
        }
        snapshot->ctx->sync();
        snapshot->columns = static_cast<std::int64_t>(columns);
    });
}

int owtk_extract_harmonic_ritz(owtk_matrix m, owtk_arnoldi_snapshot snapshot, owtk_recycle_space space,
                               owtk_solver_result* telemetry, uint64_t* size)
{
    return guarded([&] {
        if (!m || !snapshot || !space) throw std::invalid_argument("null argument");
        if (snapshot->ctx != m->ctx || space->ctx != m->ctx) throw std::invalid_argument("objects of another context");
        m->ctx->activate();
        const uint64_t k = owtk::extract_harmonic_ritz(m, snapshot, space, telemetry);
        if (size) *size = k;
    });
}

int owtk_gcrodr(owtk_matrix m, owtk_precond pc, const double* b, double* x, owtk_recycle_space space,
                const owtk_solver_options* options, owtk_arnoldi_snapshot snapshot, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !space || !options || !result) throw std::invalid_argument("null argument");
        if (space->ctx != m->ctx || (snapshot && snapshot->ctx != m->ctx))
            throw std::invalid_argument("objects of another context");
        require_plain_options(options, "gcrodr");
        m->ctx->activate();
        owtk::gcrodr_run(m, pc, b, x, space, options, snapshot, result);
    });
}

int owtk_arnoldi_snapshot_basis(owtk_arnoldi_snapshot snapshot, uint64_t column, const double** device_vector)
{
    return guarded([&] {
        if (!snapshot || !device_vector) throw std::invalid_argument("null argument");
        if (column >= static_cast<uint64_t>(snapshot->columns)) throw std::out_of_range("Arnoldi snapshot basis");
        *device_vector = snapshot->basis.get() + static_cast<std::int64_t>(column) * snapshot->ld;
    });
}

int owtk_fgmres_refined(owtk_matrix m, owtk_precond inner_pc, const double* b, double* x,
                        const owtk_refinement_options* options, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        require_plain_options(&options->outer, "fgmres_refined");
        m->ctx->activate();
        fgmres_refined_run(m, inner_pc, nullptr, b, x, options, result);
    });
}

int owtk_fgmres_refined_fallback(owtk_matrix m, owtk_precond inner_pc, owtk_precond fallback_pc, const double* b,
                                 double* x, const owtk_refinement_options* options, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !options || !result || !fallback_pc) throw std::invalid_argument("null argument");
        require_plain_options(&options->outer, "fgmres_refined_fallback");
        m->ctx->activate();
        fgmres_refined_run(m, inner_pc, fallback_pc, b, x, options, result);
    });
}

} // extern "C"

extern "C" {

int owtk_recycle_space_create(owtk_matrix layout, uint64_t capacity, owtk_recycle_space* out)
{
    return guarded([&] {
        if (!layout || !out) throw std::invalid_argument("null argument");
        if (capacity == 0) throw std::invalid_argument("recycle space capacity must be positive");
        if (capacity > static_cast<std::uint64_t>(kern::kMaxMultidot))
            throw unsupported_error("recycle space capacity above 62 is not supported on the device path");
        layout->ctx->activate();
        auto sp = std::make_unique<owtk_recycle_space_s>();
        sp->ctx = layout->ctx;
        sp->n = layout->owned * layout->bs;
        sp->local = layout->local() * layout->bs;
        sp->capacity = capacity;
        sp->coef.allocate(capacity + 2);
        *out = sp.release();
    });
}

int owtk_recycle_space_destroy(owtk_recycle_space sp)
{
    return guarded([&] {
        if (!sp) return;
        sp->ctx->activate();
        sp->ctx->sync();
        delete sp;
    });
}

int owtk_recycle_space_size(owtk_recycle_space sp, int64_t* size)
{
    return guarded([&] {
        if (!sp || !size) throw std::invalid_argument("null argument");
        *size = static_cast<int64_t>(sp->v.C.size());
    });
}

int owtk_recycle_space_clear(owtk_recycle_space sp)
{
    return guarded([&] {
        if (!sp) throw std::invalid_argument("null argument");
        sp->ctx->activate();
        sp->ctx->sync();
        sp->v = owtk::RecycleVectors{};
        sp->spare.clear();
    });
}

int owtk_recycle_space_get(owtk_recycle_space sp, int64_t index, double* u, double* c)
{
    return guarded([&] {
        if (!sp || !u || !c) throw std::invalid_argument("null argument");
        if (index < 0 || index >= static_cast<int64_t>(sp->v.C.size())) throw std::invalid_argument("index out of range");
        sp->ctx->activate();
        const hipStream_t st = sp->ctx->stream;
        OWTK_HIP(hipMemcpyAsync(u, sp->v.U[index].get(), sp->n * sizeof(double), hipMemcpyDeviceToDevice, st));
        OWTK_HIP(hipMemcpyAsync(c, sp->v.C[index].get(), sp->n * sizeof(double), hipMemcpyDeviceToDevice, st));
        sp->ctx->sync();
    });
}

int owtk_recycle_space_add_candidate(owtk_recycle_space sp, owtk_matrix op, const double* candidate,
                                     double dependence_tolerance, int32_t* added, owtk_solver_result* telemetry)
{
    return guarded([&] {
        if (!sp || !op || !candidate || !added) throw std::invalid_argument("null argument");
        if (sp->n != op->owned * op->bs || sp->local != op->local() * op->bs)
            throw std::invalid_argument("recycle space vector layout mismatch");
        op->ctx->activate();
        *added = owtk::rc_add(sp, op, candidate, dependence_tolerance, telemetry) ? 1 : 0;
        sp->ctx->sync();
    });
}

int owtk_recycle_space_refresh(owtk_recycle_space sp, owtk_matrix op, owtk_solver_result* telemetry)
{
    return guarded([&] {
        if (!sp || !op) throw std::invalid_argument("null argument");
        if (sp->n != op->owned * op->bs || sp->local != op->local() * op->bs)
            throw std::invalid_argument("recycle space vector layout mismatch");
        op->ctx->activate();
        owtk::rc_refresh(sp, op, telemetry);
    });
}

int owtk_recycle_space_project(owtk_recycle_space sp, const double* residual, double* x, owtk_solver_result* telemetry)
{
    return guarded([&] {
        if (!sp || !residual || !x) throw std::invalid_argument("null argument");
        sp->ctx->activate();
        owtk::rc_project(sp, residual, x, telemetry);
        sp->ctx->sync();
    });
}

int owtk_recycled_fgmres(owtk_matrix m, owtk_precond pc, const double* b, double* x, owtk_recycle_space space,
                         const owtk_solver_options* options, int32_t update_recycle_space, owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !space || !options || !result) throw std::invalid_argument("null argument");
        require_plain_options(options, "recycled_fgmres");
        m->ctx->activate();
        owtk::recycled_fgmres_run(m, pc, b, x, space, options, update_recycle_space != 0, result);
    });
}

int owtk_recycled_fgmres_with_snapshot(owtk_matrix m, owtk_precond pc, const double* b, double* x,
                                       owtk_recycle_space space, const owtk_solver_options* options,
                                       int32_t update_recycle_space, owtk_arnoldi_snapshot snapshot,
                                       owtk_solver_result* result)
{
    return guarded([&] {
        if (!m || !b || !x || !space || !options || !result) throw std::invalid_argument("null argument");
        if (space->ctx != m->ctx || (snapshot && snapshot->ctx != m->ctx))
            throw std::invalid_argument("objects of another context");
        require_plain_options(options, "recycled_fgmres");
        m->ctx->activate();
        owtk::recycled_fgmres_run(m, pc, b, x, space, options, update_recycle_space != 0, result, snapshot);
    });
}

} // extern "C"
