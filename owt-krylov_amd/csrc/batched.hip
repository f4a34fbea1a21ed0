// Batched BiCGSTAB + Jacobi over many independent component-diagonal systems
// of uneven size.  No reference counterpart (the reference solves one system
// per call); each system follows the recurrence, breakdown guards, check
// cadence and residual replacement of bicgstab (krylov_solvers.hpp:542-745)
// exactly, with right Jacobi preconditioning (preconditioner.hpp:25-117).
//
// Layout.  The systems' rows are concatenated; each system's rows are padded
// to a multiple of 64 in a SELL-64 matrix (slot (s, k, lane) at
// slice_ptr[s] + 64k + lane, values [slot & ~63][c][lane]) so that a wave's
// k-th entries are one contiguous load and no slice straddles two systems.
// Vectors stay unpadded.  A tile of T rows (T = 64..512, one row per thread)
// never straddles two systems either: tile -> system is a table, so a block
// reads its system's mask first and returns at once once that system stopped.
//
// Iteration.  Twelve launches: 5 streaming kernels (update p, v = A M p with
// <r^,v>, s = r - alpha v with |s|^2, t = A M s with <t,s> and <t,t>, update
// x and r with <r^,r> and |r|^2) each followed by a per-system finisher (one
// wave per system summing that system's tile partials in a fixed order:
// deterministic), plus the masked true-residual kernel of the convergence and
// replacement checks, and kb_live, which lists the tiles of the systems still
// running for the next iteration's kernels (their blocks past the list return
// at once).  G iterations are captured once into a hipGraph and the graph is
// replayed until every system has stopped; the host reads one counter per
// replay.  The tile size is chosen by owtk_batch_autotune.
//
// Small systems.  Systems of at most a set number of vertices can instead be
// solved whole by one workgroup each (kb_solve_group): one launch on a second
// stream, beside the graph replays, loops over the iterations with the same
// passes and scalar stages, its reductions summed inside the workgroup.
// owtk_batch_autotune_groups picks the limit and the workgroup size by timing
// whole solves.
#include "internal.hpp"
#include "solver_common.hpp"
#include "kernels/blas1.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <vector>

using namespace owtk;

namespace {

constexpr int kMaxTile = 512;
constexpr int kPartials = 3;

// per-system state
struct SysState {
    double rho, rho_prev, alpha, omega, beta;
    double thr, bnorm, snorm, tiny, breakdown_tolerance;
    double rr, true_norm;
    int phase;        // 1 running, 0 stopped
    int first;        // next iteration starts the recurrence (p = r)
    int restart;      // next iteration also resets r^ = r
    int early;        // |s| <= threshold this iteration
    int need_true;    // true residual this iteration
    int final;        // iteration limit reached
    int iters, status, reason;
    int check_interval, replace_interval, max_iterations;
    int owner;        // 0: graph path; 1: one workgroup (kb_solve_group)
};

// A tile of the graph path: its system, first padded row, the offset from
// padded row to vertex (vertex = prow + vshift) and its number of real rows.
// One 16-byte load.
struct alignas(16) TileRef {
    std::int32_t sys, prow0, vshift, valid;
};

struct Tiling {
    int tile = 0;
    std::int64_t blocks = 0;
    DeviceArray<TileRef> refs;             // tile -> system, rows
    DeviceArray<std::int32_t> sys_blk0;    // system -> first tile (size systems + 1)
    DeviceArray<double> partials;          // [kPartials][blocks]
    DeviceArray<TileRef> live;             // [blocks] tiles of the running systems (kb_live)
    DeviceArray<std::int32_t> live_count;
    DeviceArray<std::int32_t> live_first;  // [systems] first entry of each running system in `live`
    int ranks = 0;                         // graph-path systems
    DeviceArray<std::int32_t> blk_rank;    // tile -> rank of its system among the graph-path systems
    DeviceArray<std::int32_t> rank_sys;    // rank -> system
    DeviceArray<std::int32_t> rank_blk0;   // rank -> first tile (size ranks + 1)
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_iterations = 0;
    const double* graph_b = nullptr;  // the graph's kernels bind b and x
    double* graph_x = nullptr;
    ~Tiling()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

} // namespace

struct owtk_batch_s {
    owtk_context ctx = nullptr;
    int systems = 0;
    int bs = 1;
    std::int64_t vertices = 0;
    std::vector<std::int64_t> h_vprefix;   // unpadded vertex offset per system
    std::vector<std::int64_t> h_pprefix;   // padded row offset per system
    std::int64_t padded_rows = 0;
    DeviceArray<std::int64_t> vprefix, pprefix;
    DeviceArray<std::int32_t> slice_ptr, row_len, cols;
    DeviceArray<double> vals, dinv;
    std::int64_t entries = 0;
    // work vectors (unpadded)
    DeviceArray<double> r, rhat, p, v, s, t, phat, shat;
    DeviceArray<SysState> state;
    DeviceArray<int> active_count;
    std::unique_ptr<Tiling> tiling;
    int tile = 256;
    // Systems of at most group_limit vertices are solved by one workgroup of
    // group_threads threads each (kb_solve_group) on `side`, beside the graph
    // replays of the others.
    std::int64_t group_limit = 0;
    int group_threads = 512;
    std::vector<int> h_owner;              // per system: 0 graph, 1 workgroup
    int graph_systems = 0, group_systems = 0;
    DeviceArray<std::int32_t> group_order; // workgroup systems, largest first
    DeviceArray<unsigned long long> clocks;  // [systems] stop times, [systems] start
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    ~owtk_batch_s()
    {
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace {

struct Ptrs {
    const std::int64_t* vprefix;
    const std::int64_t* pprefix;
    const std::int32_t* slice_ptr;
    const std::int32_t* row_len;
    const std::int32_t* cols;
    const double* vals;
    const double* dinv;
    SysState* st;
    double* partials;
    const TileRef* refs;             // every graph-path tile, in tile order
    const TileRef* live;             // the running systems' tiles, in tile order
    const std::int32_t* live_count;
    const std::int32_t* live_first;  // system -> its first entry in `live`
    std::int64_t blocks;
    int bs;
};

// Row of this thread: system, unpadded vertex (or -1 past the system's end)
// and padded row.
struct RowRef {
    int sys;
    std::int64_t vertex;  // -1: padding
    std::int64_t prow;
};

__device__ inline RowRef row_of(const TileRef& T)
{
    RowRef rr;
    rr.sys = T.sys;
    rr.prow = static_cast<std::int64_t>(T.prow0) + threadIdx.x;
    rr.vertex = static_cast<int>(threadIdx.x) < T.valid ? rr.prow + T.vshift : -1;
    return rr;
}

// y_c = sum over the row's entries (diagonal included) of a_c x_col,c
template<int BS, int U = 8>
__device__ inline void row_spmv(const Ptrs& P, std::int64_t prow, const double* __restrict__ x, double (&y)[BS])
{
    const std::int64_t slice = prow >> 6;
    const int lane = static_cast<int>(prow & 63);
    const std::int64_t base = P.slice_ptr[slice];
    const int width = static_cast<int>((P.slice_ptr[slice + 1] - base) >> 6);
    const int len = P.row_len[prow];
#pragma unroll
    for (int c = 0; c < BS; ++c) y[c] = 0.0;
    // The first U slots load with no branch (index clamped into the slice,
    // product kept for k < len), so all of them are in flight at once.
    if (width > 0) {
        std::int32_t colk[U];
#pragma unroll
        for (int k = 0; k < U; ++k) colk[k] = P.cols[base + 64 * (k < width ? k : width - 1) + lane];
        double v[U][BS], xv[U][BS];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const double* vk = P.vals + (base + 64 * (k < width ? k : width - 1)) * BS + lane;
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                v[k][c] = vk[64 * c];
                xv[k][c] = x[static_cast<std::int64_t>(colk[k]) * BS + c];
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const double f = __builtin_fma(v[k][c], xv[k][c], y[c]);
                y[c] = k < len ? f : y[c];
            }
        }
    }
    for (int k = U; k < len; ++k) {
        const std::int64_t run = base + 64 * k;
        const std::int64_t col = P.cols[run + lane];
        const double* vk = P.vals + run * BS + lane;
#pragma unroll
        for (int c = 0; c < BS; ++c) y[c] = __builtin_fma(vk[64 * c], x[col * BS + c], y[c]);
    }
}

// Block sum of K values; thread 0 writes partials[k][slot].  An iteration
// kernel's slot is its block's place in the running systems' tile list, the
// initial kernel's its tile.
template<int K>
__device__ inline void store_tile_partials(double (&v)[K], const Ptrs& P, int slot)
{
    __shared__ double lds[kMaxTile / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = dev::wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int waves = static_cast<int>(blockDim.x >> 6);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double acc = lds[0][k];
            for (int w = 1; w < waves; ++w) acc += lds[w][k];
            P.partials[k * P.blocks + slot] = acc;
        }
    }
}

// ------------------------------------------------------ streaming kernels --
// The iteration kernels' block q takes entry q of the running systems' tile
// list (blocks past its end return at once).  Each loads its system's state
// beside its first data loads, not before them, and tests the stop flag once
// the data are in registers: the empty asm statements keep the compiler from
// sinking the loads below that test, which would put a second round trip on
// every row.

template<typename T>
__device__ inline void keep(const T& v)
{
    asm volatile("" ::"v"(v));
}

// Initial residual: r = b - A x, r^ = r, p = v = 0; partials |r|^2, |b|^2.
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_init(Ptrs P, const double* __restrict__ b, const double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ rhat,
                                                    double* __restrict__ p, double* __restrict__ v)
{
    const int tile = static_cast<int>(blockIdx.x);
    const RowRef rr = row_of(P.refs[tile]);
    double red[2] = {0.0, 0.0};
    if (rr.vertex >= 0) {
        double ax[BS];
        row_spmv<BS>(P, rr.prow, x, ax);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const std::int64_t i = rr.vertex * BS + c;
            const double ri = b[i] - ax[c];
            r[i] = ri;
            rhat[i] = ri;
            p[i] = 0.0;
            v[i] = 0.0;
            red[0] = __builtin_fma(ri, ri, red[0]);
            red[1] = __builtin_fma(b[i], b[i], red[1]);
        }
    }
    store_tile_partials<2>(red, P, tile);
}

// p = r (first) or r + beta (p - omega v); r^ = r on restart; phat = M p.
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_update_p(Ptrs P, const double* __restrict__ r, double* __restrict__ rhat,
                                                        double* __restrict__ p, const double* __restrict__ v,
                                                        double* __restrict__ phat)
{
    const int q = static_cast<int>(blockIdx.x);
    if (q >= *P.live_count) return;
    const RowRef rr = row_of(P.live[q]);
    if (rr.vertex < 0) return;
    const SysState& st = P.st[rr.sys];
    const int phase = st.phase;
    const bool first = st.first != 0;
    const bool restart = st.restart != 0;
    const double beta = st.beta, omega = st.omega;
    // On a first iteration (after the initial kernel or a restart) the p and
    // v values loaded are discarded by the select.
    const std::int64_t i0 = rr.vertex * BS;
    double rv[BS], pv[BS], vv[BS], dv[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        rv[c] = r[i0 + c];
        pv[c] = p[i0 + c];
        vv[c] = v[i0 + c];
        dv[c] = P.dinv[i0 + c];
    }
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        keep(rv[c]);
        keep(pv[c]);
        keep(vv[c]);
        keep(dv[c]);
    }
    if (phase != 1) return;
    double pn[BS], ph[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        const double upd = rv[c] + beta * (pv[c] - omega * vv[c]);
        pn[c] = first ? rv[c] : upd;
        ph[c] = dv[c] * pn[c];
    }
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        p[i0 + c] = pn[c];
        phat[i0 + c] = ph[c];
    }
    if (restart) {
#pragma unroll
        for (int c = 0; c < BS; ++c) rhat[i0 + c] = rv[c];
    }
}

// v = A phat; partial <r^, v>
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_spmv_v(Ptrs P, const double* __restrict__ phat,
                                                      const double* __restrict__ rhat, double* __restrict__ v)
{
    const int q = static_cast<int>(blockIdx.x);
    if (q >= *P.live_count) return;
    const RowRef rr = row_of(P.live[q]);
    const int phase = P.st[rr.sys].phase;
    double y[BS], rh[BS];
    if (rr.vertex >= 0) {
        row_spmv<BS>(P, rr.prow, phat, y);
#pragma unroll
        for (int c = 0; c < BS; ++c) rh[c] = rhat[rr.vertex * BS + c];
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            keep(y[c]);
            keep(rh[c]);
        }
    }
    if (phase != 1) return;
    double red[1] = {0.0};
    if (rr.vertex >= 0) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            v[rr.vertex * BS + c] = y[c];
            red[0] = __builtin_fma(rh[c], y[c], red[0]);
        }
    }
    store_tile_partials<1>(red, P, q);
}

// s = r - alpha v; shat = M s; partial |s|^2
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_update_s(Ptrs P, const double* __restrict__ r, const double* __restrict__ v,
                                                        double* __restrict__ s, double* __restrict__ shat)
{
    const int q = static_cast<int>(blockIdx.x);
    if (q >= *P.live_count) return;
    const RowRef rr = row_of(P.live[q]);
    const SysState& st = P.st[rr.sys];
    const int phase = st.phase;
    const double alpha = st.alpha;
    double rv[BS], vv[BS], dv[BS];
    if (rr.vertex >= 0) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const std::int64_t i = rr.vertex * BS + c;
            rv[c] = r[i];
            vv[c] = v[i];
            dv[c] = P.dinv[i];
        }
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            keep(rv[c]);
            keep(vv[c]);
            keep(dv[c]);
        }
    }
    if (phase != 1) return;
    double red[1] = {0.0};
    if (rr.vertex >= 0) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const std::int64_t i = rr.vertex * BS + c;
            const double si = rv[c] - alpha * vv[c];
            s[i] = si;
            shat[i] = dv[c] * si;
            red[0] = __builtin_fma(si, si, red[0]);
        }
    }
    store_tile_partials<1>(red, P, q);
}

// t = A shat; partials <t, s>, <t, t>.  Early exit: x += alpha phat only.
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_spmv_t(Ptrs P, const double* __restrict__ shat,
                                                      const double* __restrict__ s, double* __restrict__ t,
                                                      const double* __restrict__ phat, double* __restrict__ x)
{
    const int q = static_cast<int>(blockIdx.x);
    if (q >= *P.live_count) return;
    const RowRef rr = row_of(P.live[q]);
    const SysState& st = P.st[rr.sys];
    const int phase = st.phase, early = st.early;
    const double alpha = st.alpha;
    // The product is formed before the flags are known: an early exit (at
    // most once per system) discards it.
    double y[BS], sv[BS];
    if (rr.vertex >= 0) {
        row_spmv<BS>(P, rr.prow, shat, y);
#pragma unroll
        for (int c = 0; c < BS; ++c) sv[c] = s[rr.vertex * BS + c];
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            keep(y[c]);
            keep(sv[c]);
        }
    }
    if (phase != 1) return;
    if (early) {
        if (rr.vertex >= 0) {
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const std::int64_t i = rr.vertex * BS + c;
                x[i] = __builtin_fma(alpha, phat[i], x[i]);
            }
        }
        return;
    }
    double red[2] = {0.0, 0.0};
    if (rr.vertex >= 0) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            t[rr.vertex * BS + c] = y[c];
            red[0] = __builtin_fma(y[c], sv[c], red[0]);
            red[1] = __builtin_fma(y[c], y[c], red[1]);
        }
    }
    store_tile_partials<2>(red, P, q);
}

// x += alpha phat + omega shat; r = s - omega t; partials <r^, r>, |r|^2
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_update_x(Ptrs P, const double* __restrict__ phat,
                                                        const double* __restrict__ shat, const double* __restrict__ s,
                                                        const double* __restrict__ t, const double* __restrict__ rhat,
                                                        double* __restrict__ x, double* __restrict__ r)
{
    const int q = static_cast<int>(blockIdx.x);
    if (q >= *P.live_count) return;
    const RowRef rr = row_of(P.live[q]);
    const SysState& st = P.st[rr.sys];
    const int phase = st.phase, early = st.early;
    const double alpha = st.alpha, omega = st.omega;
    double ph[BS], sh[BS], sv[BS], tv[BS], rh[BS], xv[BS];
    if (rr.vertex >= 0) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const std::int64_t i = rr.vertex * BS + c;
            ph[c] = phat[i];
            sh[c] = shat[i];
            sv[c] = s[i];
            tv[c] = t[i];
            rh[c] = rhat[i];
            xv[c] = x[i];
        }
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            keep(ph[c]);
            keep(sh[c]);
            keep(sv[c]);
            keep(tv[c]);
            keep(rh[c]);
            keep(xv[c]);
        }
    }
    if (phase != 1 || early) return;
    double red[2] = {0.0, 0.0};
    if (rr.vertex >= 0) {
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const std::int64_t i = rr.vertex * BS + c;
            x[i] = __builtin_fma(omega, sh[c], __builtin_fma(alpha, ph[c], xv[c]));
            const double ri = sv[c] - omega * tv[c];
            r[i] = ri;
            red[0] = __builtin_fma(rh[c], ri, red[0]);
            red[1] = __builtin_fma(ri, ri, red[1]);
        }
    }
    store_tile_partials<2>(red, P, q);
}

// True residual of the systems that asked for it: r = b - A x; partial |r|^2.
// Most iterations ask for none, so the flags are read first here.
template<int BS>
__global__ __launch_bounds__(kMaxTile) void kb_true_residual(Ptrs P, const double* __restrict__ b,
                                                             const double* __restrict__ x, double* __restrict__ r)
{
    const int q = static_cast<int>(blockIdx.x);
    if (q >= *P.live_count) return;
    const RowRef rr = row_of(P.live[q]);
    const SysState& st = P.st[rr.sys];
    if (st.phase != 1 || !st.need_true) return;
    double red[1] = {0.0};
    if (rr.vertex >= 0) {
        double ax[BS];
        row_spmv<BS>(P, rr.prow, x, ax);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const std::int64_t i = rr.vertex * BS + c;
            const double ri = b[i] - ax[c];
            r[i] = ri;
            red[0] = __builtin_fma(ri, ri, red[0]);
        }
    }
    store_tile_partials<1>(red, P, q);
}

// ------------------------------------------------------------- finishers --
// One wave per system: sum the system's tile partials in a fixed order.

struct Sums {
    double v[kPartials];
};

// The initial kernel's partials sit at the system's tiles, the iteration
// kernels' at the system's entries of the running tiles' list: the same
// tiles in the same order.
template<int K, bool INIT>
__device__ inline Sums system_sums(const Ptrs& P, const std::int32_t* sys_blk0, int sys)
{
    Sums out;
    const int lane = threadIdx.x;
    const std::int32_t tiles = sys_blk0[sys + 1] - sys_blk0[sys];
    const std::int32_t b0 = INIT ? sys_blk0[sys] : P.live_first[sys], b1 = b0 + tiles;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (std::int32_t b = b0 + lane; b < b1; b += 64) acc += P.partials[k * P.blocks + b];
        out.v[k] = dev::wave_sum(acc);
    }
    return out;
}

enum Stage : int { fin_init = 0, fin_alpha = 1, fin_s = 2, fin_omega = 3, fin_rho = 4, fin_true = 5 };

__device__ inline void stop(SysState& st, int status, int reason)
{
    st.phase = 0;
    st.status = status;
    st.reason = reason;
}

// Breakdown check of rho at the start of iteration `it` (krylov_solvers.hpp:611-620).
__device__ inline bool rho_ok(SysState& st, double rho, int it)
{
    if (!__builtin_isfinite(rho) || __builtin_fabs(rho) <= st.tiny) {
        st.iters = it;
        stop(st, status_breakdown, __builtin_isfinite(rho) ? reason_biorthogonality_loss : reason_non_finite_scalar);
        return false;
    }
    return true;
}

__device__ inline void restart_recurrence(SysState& st)
{
    st.first = 1;
    st.restart = 1;
    st.rho_prev = 1.0;
    st.alpha = 1.0;
    st.omega = 1.0;
}

// Whether stage STAGE needs the system's sums; when it does not, stage_skip
// makes the one state change a skipped stage makes.
template<int STAGE>
__device__ inline bool stage_needs_sums(const SysState& st)
{
    if (STAGE != fin_init && st.phase != 1) return false;
    if (STAGE == fin_true && !st.need_true) return false;
    if ((STAGE == fin_omega || STAGE == fin_rho) && st.early) return false;
    return true;
}

template<int STAGE>
__device__ inline void stage_skip(SysState& st)
{
    if (STAGE == fin_rho && st.phase == 1 && st.early) st.need_true = 1;
}

template<int STAGE>
constexpr int stage_sums = (STAGE == fin_init || STAGE == fin_omega || STAGE == fin_rho) ? 2 : 1;

// The scalar step of stage STAGE from the system's sums.
template<int STAGE>
__device__ inline void finish_stage(SysState& st, const Sums& sm)
{
    if constexpr (STAGE == fin_init) {
        const double rn = __builtin_sqrt(sm.v[0]);
        st.bnorm = __builtin_sqrt(sm.v[1]);
        st.thr = fmax(st.thr, st.rr * st.bnorm);  // thr holds atol, rr holds rtol on entry
        st.rr = rn;
        st.true_norm = rn;
        st.iters = 0;
        st.early = st.need_true = st.final = st.restart = 0;
        st.first = 1;
        st.rho_prev = st.alpha = st.omega = 1.0;
        st.beta = 0.0;
        if (!__builtin_isfinite(rn) || !__builtin_isfinite(st.thr)) {
            stop(st, status_breakdown, reason_non_finite_scalar);
        } else if (rn <= st.thr) {
            stop(st, status_converged, reason_none);
        } else {
            st.phase = 1;
            st.status = status_maximum_iterations;
            st.reason = reason_none;
            st.rho = sm.v[0];  // <r^, r> with r^ = r
            rho_ok(st, st.rho, 1);
        }
    } else if constexpr (STAGE == fin_alpha) {
        st.iters += 1;
        st.first = 0;
        st.restart = 0;
        const double sv = sm.v[0];
        if (!__builtin_isfinite(sv) || __builtin_fabs(sv) <= st.tiny) {
            stop(st, status_breakdown, __builtin_isfinite(sv) ? reason_alpha_denominator : reason_non_finite_scalar);
            return;
        }
        st.alpha = st.rho / sv;
    } else if constexpr (STAGE == fin_s) {
        st.snorm = __builtin_sqrt(sm.v[0]);
        st.early = st.snorm <= st.thr ? 1 : 0;
    } else if constexpr (STAGE == fin_omega) {
        const double num = sm.v[0], den = sm.v[1];
        if (!__builtin_isfinite(den) || __builtin_fabs(den) <= st.tiny) {
            stop(st, status_breakdown, __builtin_isfinite(den) ? reason_omega_denominator : reason_non_finite_scalar);
            return;
        }
        const double omega = num / den;
        const double angle = __builtin_fabs(num) / __builtin_sqrt(den) / st.snorm;
        if (!__builtin_isfinite(omega) || omega == 0.0 || angle <= st.breakdown_tolerance) {
            stop(st, status_breakdown, __builtin_isfinite(omega) ? reason_omega_zero : reason_non_finite_scalar);
            return;
        }
        st.omega = omega;
    } else if constexpr (STAGE == fin_rho) {
        const int it = st.iters;
        const double rho_new = sm.v[0];
        const double rec = __builtin_sqrt(sm.v[1]);
        st.rr = rec;
        const bool replace = st.replace_interval > 0 && it % st.replace_interval == 0;
        const bool check = st.check_interval > 0 && it % st.check_interval == 0;
        if (replace || (check && rec <= st.thr)) {
            st.need_true = 1;
            return;
        }
        if (it >= st.max_iterations) {
            st.need_true = 1;
            st.final = 1;
            return;
        }
        st.rho_prev = st.rho;
        if (!rho_ok(st, rho_new, it + 1)) return;
        if (__builtin_fabs(st.omega) <= st.tiny) {
            st.iters = it + 1;
            stop(st, status_breakdown, reason_omega_zero);
            return;
        }
        st.beta = (rho_new / st.rho_prev) * (st.alpha / st.omega);
        st.rho = rho_new;
    } else {  // fin_true
        const double tn = __builtin_sqrt(sm.v[0]);
        st.true_norm = tn;
        st.need_true = 0;
        st.early = 0;
        if (!__builtin_isfinite(tn)) {
            stop(st, status_breakdown, reason_non_finite_scalar);
            return;
        }
        if (tn <= st.thr) {
            stop(st, status_converged, reason_none);
            return;
        }
        if (st.final || st.iters >= st.max_iterations) {
            stop(st, status_maximum_iterations, reason_none);
            return;
        }
        restart_recurrence(st);
        st.rho = sm.v[0];  // <r^, r> with r^ = r
        if (!rho_ok(st, st.rho, st.iters + 1)) return;
    }
}

template<int STAGE>
__global__ __launch_bounds__(64) void kb_finish(Ptrs P, const std::int32_t* __restrict__ sys_blk0)
{
    const int sys = blockIdx.x;
    SysState& st = P.st[sys];
    if (st.owner != 0) return;
    if (!stage_needs_sums<STAGE>(st)) {
        if (threadIdx.x == 0) stage_skip<STAGE>(st);
        return;
    }
    const Sums sm = system_sums<stage_sums<STAGE>, STAGE == fin_init>(P, sys_blk0, sys);
    if (threadIdx.x != 0) return;
    finish_stage<STAGE>(st, sm);
}

// Exclusive prefix sum of one int per thread over a 256-thread block; `total`
// receives the block's sum.  wave_tot: 4 ints of LDS.
__device__ inline int block_exclusive_scan256(int v, int* wave_tot, int& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(inc, d, 64);
        if (lane >= d) inc += u;
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int t = wave_tot[w];
        before += w < wave ? t : 0;
        total += t;
    }
    __syncthreads();  // wave_tot is rewritten by the next call
    return before + inc - v;
}

__device__ inline bool graph_running(const SysState& st)
{
    return st.owner == 0 && st.phase == 1;
}

// The tiles of the running graph-path systems, in tile order, and their
// number (and each running system's first entry): the iteration kernels'
// blocks take live[blockIdx.x], and the blocks
// past the count return at once, so a system that stopped costs its tiles
// nothing in later iterations.  The graph-path systems are ranked 0..G-1 in
// system order (every rank has at least one tile).  Block b writes the
// entries of tiles [256 b, 256 b + 256), which belong to at most 256
// consecutive ranks: a tile's place is the running tiles of the ranks before
// its own (each block scans the ranks' counts) plus its index in its system.
// Launched with 256 threads and ceil(tiles / 256) blocks.
__global__ __launch_bounds__(256) void kb_live(Ptrs P, const std::int32_t* __restrict__ blk_rank,
                                               const std::int32_t* __restrict__ rank_sys,
                                               const std::int32_t* __restrict__ rank_blk0, int ranks,
                                               TileRef* __restrict__ live, std::int32_t* __restrict__ live_count,
                                               std::int32_t* __restrict__ live_first)
{
    __shared__ int wave_tot[4];
    __shared__ int first_place[256];  // running tiles before each rank of this block
    const std::int64_t t0 = static_cast<std::int64_t>(blockIdx.x) * 256;
    const std::int64_t t = t0 + threadIdx.x;
    const std::int64_t t_end = t0 + 256 < P.blocks ? t0 + 256 : P.blocks;
    const int r_lo = blk_rank[t0], r_hi = blk_rank[t_end - 1];  // r_hi - r_lo < 256
    int carry = 0;
    for (int c = 0; c < ranks; c += 256) {
        const int r = c + static_cast<int>(threadIdx.x);
        int cnt = 0;
        if (r < ranks && graph_running(P.st[rank_sys[r]])) cnt = rank_blk0[r + 1] - rank_blk0[r];
        int total = 0;
        const int before = block_exclusive_scan256(cnt, wave_tot, total);
        if (r >= r_lo && r <= r_hi) first_place[r - r_lo] = carry + before;
        carry += total;
    }
    __syncthreads();
    if (t < P.blocks) {
        const int r = blk_rank[t];
        const int sy = rank_sys[r];
        if (graph_running(P.st[sy])) {
            const int place = first_place[r - r_lo];
            live[place + static_cast<int>(t - rank_blk0[r])] = P.refs[t];
            if (t == rank_blk0[r]) live_first[sy] = place;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *live_count = carry;
}

__global__ void kb_count_active(const SysState* __restrict__ st, int systems, int* __restrict__ count)
{
    int c = 0;
    for (int s = threadIdx.x; s < systems; s += blockDim.x) c += (st[s].owner == 0 && st[s].phase == 1) ? 1 : 0;
    const double total = dev::block_sum_all(static_cast<double>(c));
    if (threadIdx.x == 0) *count = static_cast<int>(total);
}

// ------------------------------------------------- one-workgroup solves --
// Systems at or below the group limit are solved whole by one workgroup each,
// in one launch that loops over the iterations: the same streaming passes and
// the same scalar stages as the graph path, with the five reductions summed
// inside the workgroup (per wave, then the waves in order: deterministic) and
// workgroup barriers where the graph path has kernel boundaries.  No
// workgroup waits on another, and the stages stop every system by its
// iteration limit, so every workgroup reaches the end.

__global__ void kb_stamp(unsigned long long* __restrict__ out)
{
    *out = wall_clock64();
}

// Sum of K values over the workgroup, valid in thread 0.  Ends with the
// barrier that makes the wave sums visible; the caller's next barrier
// separates this use of `lds` from the next.
template<int K, int NT>
__device__ inline Sums group_sums(double (&v)[K], double (*lds)[kPartials])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = dev::wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave][k] = v[k];
    }
    __syncthreads();
    Sums out{};
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double acc = lds[0][k];
            for (int w = 1; w < NT / 64; ++w) acc += lds[w][k];
            out.v[k] = acc;
        }
    }
    return out;
}

struct GroupVectors {
    const double* b;
    double *x, *r, *rhat, *p, *v, *s, *t, *phat, *shat;
};

template<int BS, int NT>
__global__ __launch_bounds__(NT) void kb_solve_group(Ptrs P, const std::int32_t* __restrict__ order, GroupVectors V,
                                                     unsigned long long* __restrict__ stop_clock)
{
    // 1024 threads leave 128 registers a lane: four slots in flight then
    constexpr int kU = NT >= 1024 ? 4 : 8;
    __shared__ SysState sst;
    __shared__ double lds[NT / 64][kPartials];
    const int sys = order[blockIdx.x];
    const int tid = threadIdx.x;
    if (tid == 0) sst = P.st[sys];
    const std::int64_t v0 = P.vprefix[sys];
    const std::int64_t n = P.vprefix[sys + 1] - v0;
    const std::int64_t p0 = P.pprefix[sys];
    __syncthreads();

    // r = b - A x, r^ = r, p = v = 0; |r|^2, |b|^2
    {
        double red[2] = {0.0, 0.0};
        for (std::int64_t l = tid; l < n; l += NT) {
            double ax[BS];
            row_spmv<BS, kU>(P, p0 + l, V.x, ax);
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const std::int64_t i = (v0 + l) * BS + c;
                const double ri = V.b[i] - ax[c];
                V.r[i] = ri;
                V.rhat[i] = ri;
                V.p[i] = 0.0;
                V.v[i] = 0.0;
                red[0] = __builtin_fma(ri, ri, red[0]);
                red[1] = __builtin_fma(V.b[i], V.b[i], red[1]);
            }
        }
        const Sums sm = group_sums<2, NT>(red, lds);
        if (tid == 0) finish_stage<fin_init>(sst, sm);
        __syncthreads();
    }

    // Every stage stops the system by max_iterations + 1 iterations; the
    // bound below only guards the loop itself.
    const int cap = sst.max_iterations + 2;
    for (int guard = 0; guard < cap && sst.phase == 1; ++guard) {
        // p = r or r + beta (p - omega v); r^ = r on restart; phat = M p
        {
            const bool first = sst.first != 0, restart = sst.restart != 0;
            const double beta = sst.beta, omega = sst.omega;
            for (std::int64_t l = tid; l < n; l += NT) {
                const std::int64_t i0 = (v0 + l) * BS;
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const double rv = V.r[i0 + c];
                    const double pn = first ? rv : rv + beta * (V.p[i0 + c] - omega * V.v[i0 + c]);
                    V.p[i0 + c] = pn;
                    V.phat[i0 + c] = P.dinv[i0 + c] * pn;
                    if (restart) V.rhat[i0 + c] = rv;
                }
            }
            __syncthreads();
        }
        // v = A phat; <r^, v> -> alpha
        {
            double red[1] = {0.0};
            for (std::int64_t l = tid; l < n; l += NT) {
                double y[BS];
                row_spmv<BS, kU>(P, p0 + l, V.phat, y);
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const std::int64_t i = (v0 + l) * BS + c;
                    V.v[i] = y[c];
                    red[0] = __builtin_fma(V.rhat[i], y[c], red[0]);
                }
            }
            const Sums sm = group_sums<1, NT>(red, lds);
            if (tid == 0) finish_stage<fin_alpha>(sst, sm);
            __syncthreads();
        }
        if (sst.phase != 1) break;
        // s = r - alpha v; shat = M s; |s|^2 -> early exit test
        {
            const double alpha = sst.alpha;
            double red[1] = {0.0};
            for (std::int64_t l = tid; l < n; l += NT) {
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const std::int64_t i = (v0 + l) * BS + c;
                    const double si = V.r[i] - alpha * V.v[i];
                    V.s[i] = si;
                    V.shat[i] = P.dinv[i] * si;
                    red[0] = __builtin_fma(si, si, red[0]);
                }
            }
            const Sums sm = group_sums<1, NT>(red, lds);
            if (tid == 0) finish_stage<fin_s>(sst, sm);
            __syncthreads();
        }
        // early exit: x += alpha phat; else t = A shat; <t, s>, <t, t> -> omega
        if (sst.early) {
            const double alpha = sst.alpha;
            for (std::int64_t l = tid; l < n; l += NT) {
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const std::int64_t i = (v0 + l) * BS + c;
                    V.x[i] = __builtin_fma(alpha, V.phat[i], V.x[i]);
                }
            }
            if (tid == 0) stage_skip<fin_rho>(sst);  // fin_omega skips with no change
            __syncthreads();
        } else {
            double red[2] = {0.0, 0.0};
            for (std::int64_t l = tid; l < n; l += NT) {
                double y[BS];
                row_spmv<BS, kU>(P, p0 + l, V.shat, y);
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const std::int64_t i = (v0 + l) * BS + c;
                    V.t[i] = y[c];
                    red[0] = __builtin_fma(y[c], V.s[i], red[0]);
                    red[1] = __builtin_fma(y[c], y[c], red[1]);
                }
            }
            const Sums sm = group_sums<2, NT>(red, lds);
            if (tid == 0) finish_stage<fin_omega>(sst, sm);
            __syncthreads();
            if (sst.phase != 1) break;
            // x += alpha phat + omega shat; r = s - omega t; <r^, r>, |r|^2
            const double alpha = sst.alpha, omega = sst.omega;
            double red2[2] = {0.0, 0.0};
            for (std::int64_t l = tid; l < n; l += NT) {
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const std::int64_t i = (v0 + l) * BS + c;
                    V.x[i] = __builtin_fma(omega, V.shat[i], __builtin_fma(alpha, V.phat[i], V.x[i]));
                    const double ri = V.s[i] - omega * V.t[i];
                    V.r[i] = ri;
                    red2[0] = __builtin_fma(V.rhat[i], ri, red2[0]);
                    red2[1] = __builtin_fma(ri, ri, red2[1]);
                }
            }
            const Sums sm2 = group_sums<2, NT>(red2, lds);
            if (tid == 0) finish_stage<fin_rho>(sst, sm2);
            __syncthreads();
        }
        // true residual of the check, replacement or final iteration
        if (sst.phase == 1 && sst.need_true) {
            double red[1] = {0.0};
            for (std::int64_t l = tid; l < n; l += NT) {
                double ax[BS];
                row_spmv<BS, kU>(P, p0 + l, V.x, ax);
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const std::int64_t i = (v0 + l) * BS + c;
                    const double ri = V.b[i] - ax[c];
                    V.r[i] = ri;
                    red[0] = __builtin_fma(ri, ri, red[0]);
                }
            }
            const Sums sm = group_sums<1, NT>(red, lds);
            if (tid == 0) finish_stage<fin_true>(sst, sm);
            __syncthreads();
        }
    }
    if (tid == 0) {
        P.st[sys] = sst;
        stop_clock[sys] = wall_clock64();
    }
}

// ------------------------------------------------------------------ host --

Ptrs make_ptrs(owtk_batch bt, const Tiling& tl)
{
    Ptrs P{};
    P.refs = tl.refs.get();
    P.vprefix = bt->vprefix.get();
    P.pprefix = bt->pprefix.get();
    P.slice_ptr = bt->slice_ptr.get();
    P.row_len = bt->row_len.get();
    P.cols = bt->cols.get();
    P.vals = bt->vals.get();
    P.dinv = bt->dinv.get();
    P.st = bt->state.get();
    P.partials = tl.partials.get();
    P.live = tl.live.get();
    P.live_count = tl.live_count.get();
    P.live_first = tl.live_first.get();
    P.blocks = tl.blocks;
    P.bs = bt->bs;
    return P;
}

std::unique_ptr<Tiling> make_tiling(owtk_batch bt, int tile)
{
    auto tl = std::make_unique<Tiling>();
    tl->tile = tile;
    std::vector<std::int32_t> brank, rsys, rb0, sb0(bt->systems + 1, 0);
    std::vector<TileRef> bref;
    for (int s = 0; s < bt->systems; ++s) {
        sb0[s] = static_cast<std::int32_t>(bref.size());
        if (bt->h_owner[s] != 0) continue;  // solved by one workgroup
        const int rank = static_cast<int>(rsys.size());
        rsys.push_back(s);
        rb0.push_back(static_cast<std::int32_t>(bref.size()));
        const std::int64_t n = bt->h_vprefix[s + 1] - bt->h_vprefix[s];
        for (std::int64_t r0 = 0; r0 < n; r0 += tile) {
            TileRef t{};
            t.sys = s;
            t.prow0 = static_cast<std::int32_t>(bt->h_pprefix[s] + r0);
            t.vshift = static_cast<std::int32_t>(bt->h_vprefix[s] - bt->h_pprefix[s]);
            t.valid = static_cast<std::int32_t>(std::min<std::int64_t>(tile, n - r0));
            bref.push_back(t);
            brank.push_back(rank);
        }
    }
    sb0[bt->systems] = static_cast<std::int32_t>(bref.size());
    rb0.push_back(static_cast<std::int32_t>(bref.size()));
    tl->blocks = static_cast<std::int64_t>(bref.size());
    tl->sys_blk0.allocate(sb0.size());
    OWTK_HIP(hipMemcpy(tl->sys_blk0.get(), sb0.data(), sb0.size() * 4, hipMemcpyHostToDevice));
    if (bref.empty()) return tl;  // every system is solved by a workgroup
    tl->refs.allocate(bref.size());
    tl->partials.allocate(kPartials * bref.size());
    tl->live.allocate(bref.size());
    tl->live_count.allocate(1);
    tl->live_first.allocate(bt->systems);
    tl->ranks = static_cast<int>(rsys.size());
    tl->blk_rank.allocate(brank.size());
    tl->rank_sys.allocate(rsys.size());
    tl->rank_blk0.allocate(rb0.size());
    OWTK_HIP(hipMemcpy(tl->blk_rank.get(), brank.data(), brank.size() * 4, hipMemcpyHostToDevice));
    OWTK_HIP(hipMemcpy(tl->rank_sys.get(), rsys.data(), rsys.size() * 4, hipMemcpyHostToDevice));
    OWTK_HIP(hipMemcpy(tl->rank_blk0.get(), rb0.data(), rb0.size() * 4, hipMemcpyHostToDevice));
    OWTK_HIP(hipMemcpy(tl->refs.get(), bref.data(), bref.size() * sizeof(TileRef), hipMemcpyHostToDevice));
    return tl;
}

// Rebuild the running systems' tile list from the current states.
void launch_live(owtk_batch bt, Tiling& tl)
{
    const unsigned int grid = static_cast<unsigned int>((tl.blocks + 255) / 256);
    kb_live<<<grid, 256, 0, bt->ctx->stream>>>(make_ptrs(bt, tl), tl.blk_rank.get(), tl.rank_sys.get(),
                                               tl.rank_blk0.get(), tl.ranks, tl.live.get(), tl.live_count.get(),
                                               tl.live_first.get());
}

template<int BS>
void launch_iteration(owtk_batch bt, Tiling& tl, const double* b, double* x)
{
    const hipStream_t st = bt->ctx->stream;
    const Ptrs P = make_ptrs(bt, tl);
    const unsigned int grid = static_cast<unsigned int>(tl.blocks);
    const unsigned int T = static_cast<unsigned int>(tl.tile);
    const unsigned int S = static_cast<unsigned int>(bt->systems);
    const std::int32_t* sb0 = tl.sys_blk0.get();
    kb_update_p<BS><<<grid, T, 0, st>>>(P, bt->r.get(), bt->rhat.get(), bt->p.get(), bt->v.get(), bt->phat.get());
    kb_spmv_v<BS><<<grid, T, 0, st>>>(P, bt->phat.get(), bt->rhat.get(), bt->v.get());
    kb_finish<fin_alpha><<<S, 64, 0, st>>>(P, sb0);
    kb_update_s<BS><<<grid, T, 0, st>>>(P, bt->r.get(), bt->v.get(), bt->s.get(), bt->shat.get());
    kb_finish<fin_s><<<S, 64, 0, st>>>(P, sb0);
    kb_spmv_t<BS><<<grid, T, 0, st>>>(P, bt->shat.get(), bt->s.get(), bt->t.get(), bt->phat.get(), x);
    kb_finish<fin_omega><<<S, 64, 0, st>>>(P, sb0);
    kb_update_x<BS><<<grid, T, 0, st>>>(P, bt->phat.get(), bt->shat.get(), bt->s.get(), bt->t.get(), bt->rhat.get(), x,
                                        bt->r.get());
    kb_finish<fin_rho><<<S, 64, 0, st>>>(P, sb0);
    kb_true_residual<BS><<<grid, T, 0, st>>>(P, b, x, bt->r.get());
    kb_finish<fin_true><<<S, 64, 0, st>>>(P, sb0);
    launch_live(bt, tl);
    OWTK_LAUNCH_CHECK();
}

void iteration(owtk_batch bt, Tiling& tl, const double* b, double* x)
{
    switch (bt->bs) {
    case 1: launch_iteration<1>(bt, tl, b, x); break;
    case 2: launch_iteration<2>(bt, tl, b, x); break;
    case 3: launch_iteration<3>(bt, tl, b, x); break;
    default: launch_iteration<4>(bt, tl, b, x); break;
    }
}

template<int BS>
void launch_init(owtk_batch bt, Tiling& tl, const double* b, const double* x)
{
    const hipStream_t st = bt->ctx->stream;
    const Ptrs P = make_ptrs(bt, tl);
    kb_init<BS><<<static_cast<unsigned int>(tl.blocks), static_cast<unsigned int>(tl.tile), 0, st>>>(
        P, b, x, bt->r.get(), bt->rhat.get(), bt->p.get(), bt->v.get());
    kb_finish<fin_init><<<static_cast<unsigned int>(bt->systems), 64, 0, st>>>(P, tl.sys_blk0.get());
    launch_live(bt, tl);
    OWTK_LAUNCH_CHECK();
}

void init(owtk_batch bt, Tiling& tl, const double* b, const double* x)
{
    switch (bt->bs) {
    case 1: launch_init<1>(bt, tl, b, x); break;
    case 2: launch_init<2>(bt, tl, b, x); break;
    case 3: launch_init<3>(bt, tl, b, x); break;
    default: launch_init<4>(bt, tl, b, x); break;
    }
}

// Graph of G iterations on (b, x); rebuilt when b, x or G change.
void ensure_graph(owtk_batch bt, Tiling& tl, const double* b, double* x, int g)
{
    if (tl.exec && tl.graph_b == b && tl.graph_x == x && tl.graph_iterations == g) return;
    if (tl.exec) {
        OWTK_HIP(hipGraphExecDestroy(tl.exec));
        tl.exec = nullptr;
    }
    if (tl.graph) {
        OWTK_HIP(hipGraphDestroy(tl.graph));
        tl.graph = nullptr;
    }
    const hipStream_t st = bt->ctx->stream;
    OWTK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        for (int k = 0; k < g; ++k) iteration(bt, tl, b, x);
    } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(st, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
    }
    OWTK_HIP(hipStreamEndCapture(st, &tl.graph));
    OWTK_HIP(hipGraphInstantiate(&tl.exec, tl.graph, nullptr, nullptr, 0));
    tl.graph_iterations = g;
    tl.graph_b = b;
    tl.graph_x = x;
}

void set_options(owtk_batch bt, const owtk_solver_options* o)
{
    std::vector<SysState> h(bt->systems);
    for (auto& s : h) {
        std::memset(&s, 0, sizeof(s));
        s.thr = o->absolute_tolerance;   // combined with rtol * |b| by the init finisher
        s.rr = o->relative_tolerance;
        s.tiny = 1024.0 * std::numeric_limits<double>::min();
        s.breakdown_tolerance = o->breakdown_tolerance;
        s.check_interval = static_cast<int>(std::min<std::uint64_t>(o->convergence_check_interval, 1u << 30));
        s.replace_interval = static_cast<int>(std::min<std::uint64_t>(o->residual_replacement_interval, 1u << 30));
        s.max_iterations = static_cast<int>(std::min<std::uint64_t>(o->maximum_iterations, 1u << 30));
        s.status = status_invalid_input;
    }
    for (int k = 0; k < bt->systems; ++k) h[k].owner = bt->h_owner[k];
    OWTK_HIP(hipMemcpyAsync(bt->state.get(), h.data(), h.size() * sizeof(SysState), hipMemcpyHostToDevice,
                            bt->ctx->stream));
    OWTK_HIP(hipStreamSynchronize(bt->ctx->stream));  // h is a host temporary
}

int read_active(owtk_batch bt)
{
    const hipStream_t st = bt->ctx->stream;
    kb_count_active<<<1, 256, 0, st>>>(bt->state.get(), bt->systems, bt->active_count.get());
    OWTK_LAUNCH_CHECK();
    int n = 0;
    OWTK_HIP(hipMemcpyAsync(&n, bt->active_count.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    OWTK_HIP(hipStreamSynchronize(st));
    return n;
}

template<int BS, int NT>
void launch_group_bs(owtk_batch bt, const Ptrs& P, const GroupVectors& V)
{
    kb_solve_group<BS, NT><<<static_cast<unsigned int>(bt->group_systems), NT, 0, bt->side>>>(
        P, bt->group_order.get(), V, bt->clocks.get());
}

template<int NT>
void launch_group_nt(owtk_batch bt, const Ptrs& P, const GroupVectors& V)
{
    switch (bt->bs) {
    case 1: launch_group_bs<1, NT>(bt, P, V); break;
    case 2: launch_group_bs<2, NT>(bt, P, V); break;
    case 3: launch_group_bs<3, NT>(bt, P, V); break;
    default: launch_group_bs<4, NT>(bt, P, V); break;
    }
}

// The workgroup solves, on the side stream after everything queued so far
// on the context's stream; stamps the start time first.
void launch_groups(owtk_batch bt, const double* b, double* x)
{
    const hipStream_t st = bt->ctx->stream;
    kb_stamp<<<1, 1, 0, st>>>(bt->clocks.get() + bt->systems);
    OWTK_HIP(hipEventRecord(bt->fork, st));
    OWTK_HIP(hipStreamWaitEvent(bt->side, bt->fork, 0));
    Ptrs P = make_ptrs(bt, *bt->tiling);
    const GroupVectors V{b, x, bt->r.get(), bt->rhat.get(), bt->p.get(), bt->v.get(), bt->s.get(), bt->t.get(),
                         bt->phat.get(), bt->shat.get()};
    switch (bt->group_threads) {
    case 256: launch_group_nt<256>(bt, P, V); break;
    case 512: launch_group_nt<512>(bt, P, V); break;
    default: launch_group_nt<1024>(bt, P, V); break;
    }
    OWTK_LAUNCH_CHECK();
    OWTK_HIP(hipEventRecord(bt->join, bt->side));
}

// One batched solve; stop_ms[s] (may be null): ms from the start to the end
// of the graph replay in which system s stopped, or, for a system solved by a
// workgroup, to the device clock's reading when it stopped.
void solve(owtk_batch bt, const double* b, double* x, const owtk_solver_options* o, int g,
           std::vector<double>* stop_ms, std::vector<int>* stop_iters)
{
    Tiling& tl = *bt->tiling;
    set_options(bt, o);
    const bool graph = bt->graph_systems > 0, groups = bt->group_systems > 0;
    if (graph) ensure_graph(bt, tl, b, x, g);
    const auto t0 = std::chrono::steady_clock::now();
    if (groups) launch_groups(bt, b, x);
    std::vector<SysState> h(bt->systems);
    if (stop_ms) stop_ms->assign(bt->systems, 0.0);
    std::vector<char> done(bt->systems, 0);
    // Reads every state; counts the running graph-path systems (the
    // workgroup systems' entries may be mid-update and are read again after
    // the join).
    auto snapshot = [&]() {
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        OWTK_HIP(hipMemcpyAsync(h.data(), bt->state.get(), h.size() * sizeof(SysState), hipMemcpyDeviceToHost,
                                bt->ctx->stream));
        OWTK_HIP(hipStreamSynchronize(bt->ctx->stream));
        int active = 0;
        for (int s = 0; s < bt->systems; ++s) {
            if (bt->h_owner[s] != 0) continue;
            if (h[s].phase == 1) {
                ++active;
            } else if (!done[s]) {
                done[s] = 1;
                if (stop_ms) (*stop_ms)[s] = ms;
            }
        }
        return active;
    };
    int active = 0;
    if (graph) {
        init(bt, tl, b, x);
        active = stop_ms ? snapshot() : read_active(bt);
    }
    const std::uint64_t cap = o->maximum_iterations + 2;
    std::uint64_t replays = 0;
    while (active > 0 && replays * static_cast<std::uint64_t>(g) <= cap) {
        OWTK_HIP(hipGraphLaunch(tl.exec, bt->ctx->stream));
        ++replays;
        active = stop_ms ? snapshot() : read_active(bt);
    }
    if (groups) {
        OWTK_HIP(hipStreamWaitEvent(bt->ctx->stream, bt->join, 0));
        OWTK_HIP(hipStreamSynchronize(bt->ctx->stream));
    }
    if (active > 0) throw std::runtime_error("batched BiCGSTAB did not stop within its iteration limit");
    if (groups || stop_iters) {
        OWTK_HIP(hipMemcpy(h.data(), bt->state.get(), h.size() * sizeof(SysState), hipMemcpyDeviceToHost));
        for (int s = 0; s < bt->systems; ++s)
            if (bt->h_owner[s] != 0 && h[s].phase == 1)
                throw std::runtime_error("batched BiCGSTAB did not stop within its iteration limit");
    }
    if (groups && stop_ms) {
        std::vector<unsigned long long> clk(bt->systems + 1);
        OWTK_HIP(hipMemcpy(clk.data(), bt->clocks.get(), clk.size() * sizeof(clk[0]), hipMemcpyDeviceToHost));
        int khz = 0;
        OWTK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, bt->ctx->device));
        if (khz <= 0) throw std::runtime_error("device wall clock rate unavailable");
        for (int s = 0; s < bt->systems; ++s)
            if (bt->h_owner[s] != 0)
                (*stop_ms)[s] = static_cast<double>(clk[s] - clk[bt->systems]) / khz;
    }
    if (stop_iters) {
        stop_iters->resize(bt->systems);
        for (int s = 0; s < bt->systems; ++s) (*stop_iters)[s] = h[s].iters;
    }
}

// Partition the systems at `limit` vertices (0: all on the graph path) and
// rebuild the tiling of the graph systems.
void set_groups(owtk_batch bt, std::int64_t limit, int threads)
{
    // the same split: the tiling (and its captured graph) stays
    if (bt->tiling && limit == bt->group_limit && (limit == 0 || threads == bt->group_threads)) return;
    bt->ctx->sync();
    std::vector<int> owner(bt->systems, 0);
    std::vector<std::int32_t> order;
    for (int s = 0; s < bt->systems; ++s) {
        if (bt->h_vprefix[s + 1] - bt->h_vprefix[s] <= limit) {
            owner[s] = 1;
            order.push_back(s);
        }
    }
    std::stable_sort(order.begin(), order.end(), [&](std::int32_t a, std::int32_t c) {
        return bt->h_vprefix[a + 1] - bt->h_vprefix[a] > bt->h_vprefix[c + 1] - bt->h_vprefix[c];
    });
    bt->h_owner = owner;
    bt->group_systems = static_cast<int>(order.size());
    bt->graph_systems = bt->systems - bt->group_systems;
    bt->group_limit = limit;
    bt->group_threads = threads;
    if (!order.empty()) {
        bt->group_order.allocate(order.size());
        OWTK_HIP(hipMemcpy(bt->group_order.get(), order.data(), order.size() * 4, hipMemcpyHostToDevice));
        if (!bt->side) {
            OWTK_HIP(hipStreamCreateWithFlags(&bt->side, hipStreamNonBlocking));
            OWTK_HIP(hipEventCreateWithFlags(&bt->fork, hipEventDisableTiming));
            OWTK_HIP(hipEventCreateWithFlags(&bt->join, hipEventDisableTiming));
        }
    }
    bt->tiling = make_tiling(bt, bt->tile);
}

} // namespace

extern "C" {

int owtk_batch_create(owtk_context ctx, int32_t systems, int32_t block_size, const int64_t* vertex_offsets,
                      const int64_t* row_offsets, const int32_t* columns, const double* values, owtk_batch* out)
{
    return guarded([&] {
        if (!ctx || !vertex_offsets || !row_offsets || !columns || !values || !out)
            throw std::invalid_argument("null argument");
        if (systems < 1) throw std::invalid_argument("a batch needs at least one system");
        if (block_size < 1 || block_size > 4) throw std::invalid_argument("batched solver supports block size 1..4");
        *out = nullptr;
        ctx->activate();
        auto bt = std::make_unique<owtk_batch_s>();
        bt->ctx = ctx;
        bt->systems = systems;
        bt->bs = block_size;
        bt->h_vprefix.assign(vertex_offsets, vertex_offsets + systems + 1);
        if (bt->h_vprefix[0] != 0) throw std::invalid_argument("vertex_offsets must start at 0");
        for (int s = 0; s < systems; ++s)
            if (bt->h_vprefix[s + 1] <= bt->h_vprefix[s]) throw std::invalid_argument("every system needs a vertex");
        const std::int64_t N = bt->h_vprefix[systems];
        bt->vertices = N;
        const int B = block_size;
        // padded row space and SELL slices
        bt->h_pprefix.assign(systems + 1, 0);
        for (int s = 0; s < systems; ++s)
            bt->h_pprefix[s + 1] = bt->h_pprefix[s] + ((bt->h_vprefix[s + 1] - bt->h_vprefix[s] + 63) / 64) * 64;
        const std::int64_t PR = bt->h_pprefix[systems];
        bt->padded_rows = PR;
        const std::int64_t slices = PR / 64;
        std::vector<std::int32_t> slice_ptr(slices + 1, 0), row_len(PR, 0);
        std::vector<std::int64_t> prow_vertex(PR, -1);
        for (int s = 0; s < systems; ++s)
            for (std::int64_t v = bt->h_vprefix[s]; v < bt->h_vprefix[s + 1]; ++v)
                prow_vertex[bt->h_pprefix[s] + (v - bt->h_vprefix[s])] = v;
        std::int64_t slots = 0;
        for (std::int64_t sl = 0; sl < slices; ++sl) {
            int width = 0;
            for (int lane = 0; lane < 64; ++lane) {
                const std::int64_t v = prow_vertex[sl * 64 + lane];
                if (v < 0) continue;
                const std::int64_t len = row_offsets[v + 1] - row_offsets[v];
                if (len < 0) throw std::invalid_argument("row offsets must be non-decreasing");
                row_len[sl * 64 + lane] = static_cast<std::int32_t>(len);
                width = std::max<int>(width, static_cast<int>(len));
            }
            slots += static_cast<std::int64_t>(width) * 64;
            if (slots >= (std::int64_t(1) << 31)) throw std::invalid_argument("batch too large for int32 slots");
            slice_ptr[sl + 1] = static_cast<std::int32_t>(slots);
        }
        std::vector<std::int32_t> cols(std::max<std::int64_t>(slots, 1), 0);
        std::vector<double> vals(std::max<std::int64_t>(slots, 1) * B, 0.0);
        std::vector<double> dinv(N * B, 0.0);
        int sys = 0;
        for (std::int64_t p = 0; p < PR; ++p) {
            const std::int64_t v = prow_vertex[p];
            if (v < 0) continue;
            while (v >= bt->h_vprefix[sys + 1]) ++sys;
            const std::int64_t sl = p / 64;
            const int lane = static_cast<int>(p % 64);
            bool diag = false;
            for (std::int64_t e = row_offsets[v]; e < row_offsets[v + 1]; ++e) {
                const std::int64_t k = e - row_offsets[v];
                const std::int64_t slot = slice_ptr[sl] + 64 * k + lane;
                const std::int32_t c = columns[e];
                if (c < bt->h_vprefix[sys] || c >= bt->h_vprefix[sys + 1])
                    throw std::invalid_argument("column outside its system");
                cols[slot] = c;
                for (int q = 0; q < B; ++q) vals[(slot & ~std::int64_t(63)) * B + 64 * q + (slot & 63)] = values[e * B + q];
                if (c == v) {
                    diag = true;
                    for (int q = 0; q < B; ++q) {
                        const double d = values[e * B + q];
                        // JacobiPreconditioner rejects zero and non-finite diagonals (preconditioner.hpp:62-68)
                        if (d == 0.0 || !std::isfinite(d)) throw std::invalid_argument("singular diagonal");
                        dinv[v * B + q] = 1.0 / d;
                    }
                }
            }
            if (!diag) throw std::invalid_argument("row without diagonal entry");
        }
        bt->entries = row_offsets[N];
        auto up = [](auto& d, const auto& h) {
            d.allocate(h.size());
            OWTK_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
        };
        up(bt->slice_ptr, slice_ptr);
        up(bt->row_len, row_len);
        up(bt->cols, cols);
        up(bt->vals, vals);
        up(bt->dinv, dinv);
        up(bt->vprefix, bt->h_vprefix);
        up(bt->pprefix, bt->h_pprefix);
        const std::size_t n = static_cast<std::size_t>(N * B);
        for (auto* a : {&bt->r, &bt->rhat, &bt->p, &bt->v, &bt->s, &bt->t, &bt->phat, &bt->shat}) {
            a->allocate(n);
            kern::fill_zero_bytes(ctx->stream, a->get(), a->bytes());
        }
        bt->state.allocate(systems);
        bt->active_count.allocate(1);
        bt->clocks.allocate(static_cast<std::size_t>(systems) + 1);
        bt->h_owner.assign(systems, 0);
        bt->graph_systems = systems;
        bt->tiling = make_tiling(bt.get(), bt->tile);
        ctx->sync();
        *out = bt.release();
    });
}

int owtk_batch_destroy(owtk_batch bt)
{
    return guarded([&] {
        if (!bt) return;
        bt->ctx->activate();
        bt->ctx->sync();
        delete bt;
    });
}

int owtk_batch_set_tile(owtk_batch bt, int32_t tile_rows)
{
    return guarded([&] {
        if (!bt) throw std::invalid_argument("null batch");
        if (tile_rows != 64 && tile_rows != 128 && tile_rows != 256 && tile_rows != 512)
            throw std::invalid_argument("tile rows must be 64, 128, 256 or 512");
        bt->ctx->activate();
        if (tile_rows == bt->tile) return;
        bt->ctx->sync();
        bt->tiling = make_tiling(bt, tile_rows);
        bt->tile = tile_rows;
    });
}

int owtk_batch_autotune(owtk_batch bt, const double* b, int32_t iterations, int32_t* tile_rows, double* ms_per_iteration)
{
    return guarded([&] {
        if (!bt || !b) throw std::invalid_argument("null argument");
        if (iterations < 1) throw std::invalid_argument("autotune needs at least one iteration");
        bt->ctx->activate();
        DeviceArray<double> x(static_cast<std::size_t>(bt->vertices * bt->bs));
        owtk_solver_options o;
        owtk_solver_options_default(&o);
        o.relative_tolerance = 0.0;   // nothing converges: every system stays active
        o.absolute_tolerance = 0.0;
        o.convergence_check_interval = 1u << 30;
        o.residual_replacement_interval = 0;
        o.maximum_iterations = 1u << 30;
        int best = bt->tile;
        double best_ms = std::numeric_limits<double>::infinity();
        if (bt->graph_systems == 0) {  // no system on the graph path: nothing to tune
            if (tile_rows) *tile_rows = best;
            if (ms_per_iteration) *ms_per_iteration = 0.0;
            return;
        }
        for (int tile : {64, 128, 256, 512}) {
            owtk_batch_set_tile(bt, tile);
            Tiling& tl = *bt->tiling;
            double ms_best_trial = std::numeric_limits<double>::infinity();
            for (int trial = 0; trial < 2; ++trial) {
                kern::fill_zero_bytes(bt->ctx->stream, x.get(), x.bytes());
                set_options(bt, &o);
                init(bt, tl, b, x.get());
                ensure_graph(bt, tl, b, x.get(), iterations);
                OWTK_HIP(hipStreamSynchronize(bt->ctx->stream));
                const auto t0 = std::chrono::steady_clock::now();
                OWTK_HIP(hipGraphLaunch(tl.exec, bt->ctx->stream));
                OWTK_HIP(hipStreamSynchronize(bt->ctx->stream));
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                ms_best_trial = std::min(ms_best_trial, ms / iterations);
            }
            if (ms_best_trial < best_ms) {
                best_ms = ms_best_trial;
                best = tile;
            }
        }
        owtk_batch_set_tile(bt, best);
        bt->tiling->graph_b = nullptr;  // the trial graphs bound the scratch x
        if (tile_rows) *tile_rows = best;
        if (ms_per_iteration) *ms_per_iteration = best_ms;
    });
}

int owtk_batch_set_group_limit(owtk_batch bt, int64_t max_vertices, int32_t threads)
{
    return guarded([&] {
        if (!bt) throw std::invalid_argument("null batch");
        if (max_vertices < 0) throw std::invalid_argument("group limit must be >= 0");
        if (threads != 256 && threads != 512 && threads != 1024)
            throw std::invalid_argument("group threads must be 256, 512 or 1024");
        bt->ctx->activate();
        set_groups(bt, max_vertices, threads);
    });
}

int owtk_batch_autotune_groups(owtk_batch bt, const double* b, const owtk_solver_options* options,
                               int32_t graph_iterations, int64_t* max_vertices, int32_t* threads, double* ms)
{
    return guarded([&] {
        if (!bt || !b || !options) throw std::invalid_argument("null argument");
        if (graph_iterations < 1 || graph_iterations > 1024) throw std::invalid_argument("graph_iterations must be 1..1024");
        if (!valid_options(options)) throw std::invalid_argument("invalid solver options");
        require_plain_options(options, "batched bicgstab");
        bt->ctx->activate();
        DeviceArray<double> x(static_cast<std::size_t>(bt->vertices * bt->bs));
        auto trial = [&](std::int64_t limit, int nt) {
            set_groups(bt, limit, nt);
            double best = std::numeric_limits<double>::infinity();
            for (int k = 0; k < 2; ++k) {
                kern::fill_zero_bytes(bt->ctx->stream, x.get(), x.bytes());
                OWTK_HIP(hipStreamSynchronize(bt->ctx->stream));
                const auto t0 = std::chrono::steady_clock::now();
                solve(bt, b, x.get(), options, graph_iterations, nullptr, nullptr);
                best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
            }
            return best;
        };
        std::int64_t largest = 0;
        for (int s = 0; s < bt->systems; ++s) largest = std::max(largest, bt->h_vprefix[s + 1] - bt->h_vprefix[s]);
        std::int64_t best_limit = 0;
        int best_threads = 1024;
        double best_ms = trial(0, 1024);
        for (std::int64_t limit = 1024; limit <= std::min<std::int64_t>(largest, 65536); limit *= 2) {
            const double t = trial(limit, 1024);
            if (t < best_ms) {
                best_ms = t;
                best_limit = limit;
            }
        }
        if (best_limit > 0) {
            for (int nt : {256, 512}) {
                const double t = trial(best_limit, nt);
                if (t < best_ms) {
                    best_ms = t;
                    best_threads = nt;
                }
            }
        }
        set_groups(bt, best_limit, best_threads);
        bt->tiling->graph_b = nullptr;  // the trial graphs bound the scratch x
        if (max_vertices) *max_vertices = best_limit;
        if (threads) *threads = best_threads;
        if (ms) *ms = best_ms;
    });
}

int owtk_batch_bicgstab(owtk_batch bt, const double* b, double* x, const owtk_solver_options* options,
                        int32_t graph_iterations, owtk_batch_result* results)
{
    return guarded([&] {
        if (!bt || !b || !x || !options) throw std::invalid_argument("null argument");
        if (graph_iterations < 1 || graph_iterations > 1024) throw std::invalid_argument("graph_iterations must be 1..1024");
        if (!valid_options(options)) throw std::invalid_argument("invalid solver options");
        require_plain_options(options, "batched bicgstab");
        bt->ctx->activate();
        std::vector<double> ms;
        std::vector<int> its;
        solve(bt, b, x, options, graph_iterations, results ? &ms : nullptr, results ? &its : nullptr);
        if (results) {
            std::vector<SysState> h(bt->systems);
            OWTK_HIP(hipMemcpy(h.data(), bt->state.get(), h.size() * sizeof(SysState), hipMemcpyDeviceToHost));
            for (int s = 0; s < bt->systems; ++s) {
                results[s].status = h[s].status;
                results[s].breakdown_reason = h[s].reason;
                results[s].iterations = static_cast<std::uint64_t>(h[s].iters);
                results[s].true_residual_norm = h[s].true_norm;
                results[s].rhs_norm = h[s].bnorm;
                results[s].relative_residual_norm = h[s].bnorm > 0 ? h[s].true_norm / h[s].bnorm : h[s].true_norm;
                results[s].stop_ms = ms[s];
            }
        }
    });
}

} // extern "C"
