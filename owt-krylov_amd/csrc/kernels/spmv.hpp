// SpMV kernels for the three block-CSR forms, with fused epilogues.
//
// Accumulation order per output element is the reference's:
//   component-diagonal CSR: out = 0, then out = fma(a_e, x_col(e), out) over the
//     row's entries in CSR order (block_csr.hpp:198-215);
//   split: out = d * x_row, then the same fma chain over the edges
//     (split_block_csr.hpp:95-119);
//   dense BSR: per block row t = fma chain over the block's columns starting
//     at 0, then out += t (dense_block_csr.hpp:179-205).
// The build compiles with -ffp-contract=off and every fused multiply-add is an
// explicit fma, so these kernels and the CPU oracle round identically.
//
// An epilogue E receives every finished output element and may fold it into
// up to E::K reduced values; the last block then runs E::finalize.
#pragma once

#include "../device.hpp"

namespace owtk::dev {

template<class X> struct acc_type { using type = double; };
template<> struct acc_type<float> { using type = float; };

template<class V> __device__ inline auto load_value(const V* p, std::int64_t i)
{
    if constexpr (sizeof(V) == 2) return bf16_to_float(reinterpret_cast<const std::uint16_t*>(p)[i]);
    else return p[i];
}

// Matrix values are read once per SpMV: stream them past the caches so the
// gathered vector stays resident.
template<class V> __device__ inline auto load_value_nt(const V* p, std::int64_t i)
{
    if constexpr (sizeof(V) == 2)
        return bf16_to_float(__builtin_nontemporal_load(reinterpret_cast<const std::uint16_t*>(p) + i));
    else return __builtin_nontemporal_load(p + i);
}

// ----------------------------------------------------------------- epilogues
// A row's BS contiguous components as one load / store (the compiler emits
// one wide access per lane: 64 lanes x BS elements, fully coalesced).
template<int BS, class X>
__device__ inline void load_row(const X* __restrict__ p, X (&v)[BS])
{
#pragma unroll
    for (int c = 0; c < BS; ++c) v[c] = p[c];
}
template<int BS, class X>
__device__ inline void store_row(X* __restrict__ p, const X (&v)[BS])
{
#pragma unroll
    for (int c = 0; c < BS; ++c) p[c] = v[c];
}

struct EpiNone {
    static constexpr int K = 0;
    static constexpr std::uint64_t kMaxMask = 0;
    // The matrix is streamed with the non-temporal hint unless an epilogue
    // belongs to a kernel that rereads the same matrix right away.
    static constexpr bool kKeepMatrix = false;
    const int* stop = nullptr;
    __device__ bool stopped() const { return stop != nullptr && *stop != 0; }
    __device__ void finalize(const double*) const {}
    // the operand value the kernel multiplies at vector index i (an epilogue
    // may form it on the fly, e.g. a diagonally scaled input)
    template<class X>
    __device__ X xval(const X* x, std::int64_t i) const { return x[i]; }
};

template<class X>
struct EpiStore : EpiNone {
    X* y;
    template<class A>
    __device__ void elem(std::int64_t i, A value, double*) const { y[i] = static_cast<X>(value); }
    // all components of a row in one contiguous store
    template<int BS, class A>
    __device__ void elem_row(std::int64_t i0, const A (&acc)[BS], double*) const
    {
        X v[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) v[c] = static_cast<X>(acc[c]);
        store_row<BS>(y + i0, v);
    }
};

// Epilogue of one row's BS components: the epilogue's elem_row when it has
// one (its loads and stores then cover the row's components contiguously),
// else elem per component.
template<int BS, class E, class A>
__device__ inline void epi_row(const E& e, std::int64_t i0, const A (&acc)[BS], double* red)
{
    if constexpr (requires { e.template elem_row<BS>(i0, acc, red); }) {
        e.template elem_row<BS>(i0, acc, red);
    } else {
#pragma unroll
        for (int c = 0; c < BS; ++c) e.elem(i0 + c, acc[c], red);
    }
}

// y = b - A x, reduced: sum y^2, max |y|; norm written by finalize.
struct EpiResidual : EpiNone {
    static constexpr int K = 2;
    static constexpr std::uint64_t kMaxMask = 2;
    double* y;
    const double* b;
    std::int64_t n;       // owned size
    double* norm_out;     // may be null
    __device__ void elem(std::int64_t i, double ax, double* red) const
    {
        const double r = b[i] - ax;
        y[i] = r;
        red[0] = __builtin_fma(r, r, red[0]);
        red[1] = combine<Combine::max>(red[1], __builtin_fabs(r));
    }
    template<int BS>
    __device__ void elem_row(std::int64_t i0, const double (&ax)[BS], double* red) const
    {
        double r[BS];
        load_row<BS>(b + i0, r);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            r[c] = r[c] - ax[c];
            red[0] = __builtin_fma(r[c], r[c], red[0]);
            red[1] = combine<Combine::max>(red[1], __builtin_fabs(r[c]));
        }
        store_row<BS>(y + i0, r);
    }
    __device__ void finalize(const double* t) const
    {
        const double nrm = block_norm(t[0], t[1], y, n);
        if (threadIdx.x == 0 && norm_out) *norm_out = nrm;
    }
};

// y = A x, reduced <w, y>; fin(total) on thread 0 of the last block.
template<class Fin>
struct EpiDot1 : EpiNone {
    static constexpr int K = 1;
    static constexpr std::uint64_t kMaxMask = 0;
    double* y;
    const double* w;
    Fin fin;
    __device__ void elem(std::int64_t i, double v, double* red) const
    {
        y[i] = v;
        red[0] = __builtin_fma(w[i], v, red[0]);
    }
    template<int BS>
    __device__ void elem_row(std::int64_t i0, const double (&v)[BS], double* red) const
    {
        double ww[BS];
        load_row<BS>(w + i0, ww);
#pragma unroll
        for (int c = 0; c < BS; ++c) red[0] = __builtin_fma(ww[c], v[c], red[0]);
        store_row<BS>(y + i0, v);
    }
    __device__ void finalize(const double* t) const
    {
        if (threadIdx.x == 0) fin(t[0]);
    }
};

// y = A x, reduced <y, w>, <y, y>; fin(t0, t1) on thread 0 of the last block.
template<class Fin>
struct EpiDot2 : EpiNone {
    static constexpr int K = 2;
    static constexpr std::uint64_t kMaxMask = 0;
    double* y;
    const double* w;
    Fin fin;
    __device__ void elem(std::int64_t i, double v, double* red) const
    {
        y[i] = v;
        red[0] = __builtin_fma(v, w[i], red[0]);
        red[1] = __builtin_fma(v, v, red[1]);
    }
    template<int BS>
    __device__ void elem_row(std::int64_t i0, const double (&v)[BS], double* red) const
    {
        double ww[BS];
        load_row<BS>(w + i0, ww);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            red[0] = __builtin_fma(v[c], ww[c], red[0]);
            red[1] = __builtin_fma(v[c], v[c], red[1]);
        }
        store_row<BS>(y + i0, v);
    }
    __device__ void finalize(const double* t) const
    {
        if (threadIdx.x == 0) fin(t[0], t[1]);
    }
};

// Finisher adaptor: runs an epilogue's finalize in the one-block finisher.
template<class E>
struct EpiFinish {
    E e;
    __device__ void operator()(const double* t) const { e.finalize(t); }
};

} // namespace owtk::dev

namespace owtk::kern {

using namespace owtk::dev;

// Physical block -> logical row block.  Blocks are dealt round-robin over the
// 8 XCDs; giving each XCD a contiguous eighth of the rows keeps the x-gather
// working set of an XCD inside its own L2.
__device__ inline unsigned int xcd_block(unsigned int b, unsigned int n)
{
    const unsigned int per = n / 8;
    if (per == 0 || n % 8 != 0) return b;
    return (b % 8) * per + b / 8;
}

// Slots a SELL row kernel loads without a branch: the smallest of 4, 6 and 8
// covering the widest slice (rows longer than 8 finish in a loop).
inline int sell_unroll(int max_width) { return max_width <= 4 ? 4 : (max_width <= 6 ? 6 : 8); }

// Component-diagonal (and split) SpMV, SELL-64 layout, one lane per row.
//
// The first U slots of a row are loaded unconditionally: a slot index past
// the slice's width is clamped to its last run (same cache lines, always in
// bounds), and each product is kept only for k < len by a select.  A load
// under a divergent branch makes the compiler wait for it at the branch's
// end, which serialised the row's memory round trips; with no branch every
// column, value and gather load of the row is in flight before the first FMA.
// The FMAs still run in CSR order over k < len, so results are unchanged.
template<int BS, class V, class X, bool DIAG, class E, int U>
__global__ __launch_bounds__(256) void spmv_sell_cd(const std::int32_t* __restrict__ slice_ptr,
                                                    const std::int32_t* __restrict__ row_len,
                                                    const std::int32_t* __restrict__ cols,
                                                    const V* __restrict__ vals,
                                                    const V* __restrict__ diag,
                                                    const X* __restrict__ x, std::int64_t rows,
                                                    E e, ReduceScratch scratch)
{
    using A = typename acc_type<X>::type;
    if (e.stopped()) return;
    double red[E::K > 0 ? E::K : 1];
#pragma unroll
    for (int k = 0; k < (E::K > 0 ? E::K : 1); ++k) red[k] = 0.0;
    const std::int64_t row = static_cast<std::int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
    if (row < rows) {
        const int lane = static_cast<int>(row & 63);
        const std::int64_t slice = row >> 6;
        A acc[BS];
        if constexpr (DIAG) {
#pragma unroll
            for (int c = 0; c < BS; ++c)
                acc[c] = static_cast<A>(load_value(diag, slice * 64 * BS + c * 64 + lane))
                    * static_cast<A>(x[row * BS + c]);
        } else {
#pragma unroll
            for (int c = 0; c < BS; ++c) acc[c] = A(0);
        }
        const std::int64_t base = slice_ptr[slice];
        const int width = static_cast<int>((slice_ptr[slice + 1] - base) >> 6);
        const int len = row_len[row];
        if (width > 0) {
            std::int32_t colk[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int kk = k < width ? k : width - 1;
                if constexpr (E::kKeepMatrix) colk[k] = cols[base + 64 * kk + lane];
                else colk[k] = __builtin_nontemporal_load(&cols[base + 64 * kk + lane]);
            }
            A v[U][BS], xv[U][BS];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int kk = k < width ? k : width - 1;
                const V* vk = vals + (base + 64 * kk) * BS + lane;
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    if constexpr (E::kKeepMatrix) v[k][c] = static_cast<A>(load_value(vk, 64 * c));
                    else v[k][c] = static_cast<A>(load_value_nt(vk, 64 * c));
                    xv[k][c] = static_cast<A>(e.xval(x, static_cast<std::int64_t>(colk[k]) * BS + c));
                }
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
#pragma unroll
                for (int c = 0; c < BS; ++c) {
                    const A f = dfma(v[k][c], xv[k][c], acc[c]);
                    acc[c] = k < len ? f : acc[c];
                }
            }
        }
        for (int k = U; k < len; ++k) {
            const std::int64_t run = base + 64 * k;
            const std::int64_t col = cols[run + lane];
            const V* vk = vals + run * BS + lane;
#pragma unroll
            for (int c = 0; c < BS; ++c)
                acc[c] = dfma(static_cast<A>(load_value(vk, 64 * c)), static_cast<A>(e.xval(x, col * BS + c)),
                              acc[c]);
        }
        epi_row<BS>(e, row * BS, acc, red);
    }
    if constexpr (E::K > 0) store_partials<E::K>(red, E::kMaxMask, scratch);
}

// Launch spmv_sell_cd with the unroll that covers the matrix's slices.
template<int BS, class V, class X, bool DIAG, class E>
void launch_sell_cd(int max_width, unsigned int grid, hipStream_t st, const std::int32_t* slice_ptr,
                    const std::int32_t* row_len, const std::int32_t* cols, const V* vals, const V* diag,
                    const X* x, std::int64_t rows, E e, ReduceScratch scratch)
{
    switch (sell_unroll(max_width)) {
    case 4: spmv_sell_cd<BS, V, X, DIAG, E, 4><<<grid, kBlock, 0, st>>>(slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    case 6: spmv_sell_cd<BS, V, X, DIAG, E, 6><<<grid, kBlock, 0, st>>>(slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    default: spmv_sell_cd<BS, V, X, DIAG, E, 8><<<grid, kBlock, 0, st>>>(slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    }
}

template<class V, class X, bool DIAG, class E>
void launch_sell_cd_bs(int bs, int max_width, unsigned int grid, hipStream_t st, const std::int32_t* slice_ptr,
                       const std::int32_t* row_len, const std::int32_t* cols, const V* vals, const V* diag,
                       const X* x, std::int64_t rows, E e, ReduceScratch scratch)
{
    switch (bs) {
    case 1: launch_sell_cd<1, V, X, DIAG, E>(max_width, grid, st, slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    case 2: launch_sell_cd<2, V, X, DIAG, E>(max_width, grid, st, slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    case 3: launch_sell_cd<3, V, X, DIAG, E>(max_width, grid, st, slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    default: launch_sell_cd<4, V, X, DIAG, E>(max_width, grid, st, slice_ptr, row_len, cols, vals, diag, x, rows, e, scratch); break;
    }
}

// Component-diagonal (and split) SpMV for any block size, reference CSR
// layout: one wave per row, lanes over components.
template<class V, class X, bool DIAG, class E>
__global__ __launch_bounds__(256) void spmv_csr_cd(const std::int32_t* __restrict__ offsets,
                                                   const std::int32_t* __restrict__ cols,
                                                   const V* __restrict__ vals,
                                                   const V* __restrict__ diag,
                                                   const X* __restrict__ x, std::int64_t rows,
                                                   int bs, E e, ReduceScratch scratch)
{
    using A = typename acc_type<X>::type;
    if (e.stopped()) return;
    double red[E::K > 0 ? E::K : 1];
#pragma unroll
    for (int k = 0; k < (E::K > 0 ? E::K : 1); ++k) red[k] = 0.0;
    const int lane = threadIdx.x & 63;
    const std::int64_t waves = static_cast<std::int64_t>(gridDim.x) * kWavesPerBlock;
    for (std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / 64;
         row < rows; row += waves) {
        const std::int32_t begin = offsets[row];
        const std::int32_t end = offsets[row + 1];
        for (int c = lane; c < bs; c += 64) {
            A acc = A(0);
            if constexpr (DIAG)
                acc = static_cast<A>(load_value(diag, row * bs + c)) * static_cast<A>(x[row * bs + c]);
            for (std::int32_t en = begin; en < end; ++en) {
                const std::int64_t col = cols[en];
                acc = dfma(static_cast<A>(load_value(vals, static_cast<std::int64_t>(en) * bs + c)),
                           static_cast<A>(e.xval(x, col * bs + c)), acc);
            }
            e.elem(row * bs + c, acc, red);
        }
    }
    if constexpr (E::K > 0) store_partials<E::K>(red, E::kMaxMask, scratch);
}

// Dense BSR, SELL-64 layout, one lane per row, BS x BS blocks.
template<int BS, class V, class X, class E, int U>
__global__ __launch_bounds__(256) void spmv_sell_bsr(const std::int32_t* __restrict__ slice_ptr,
                                                     const std::int32_t* __restrict__ row_len,
                                                     const std::int32_t* __restrict__ cols,
                                                     const V* __restrict__ vals,
                                                     const X* __restrict__ x, std::int64_t rows,
                                                     E e, ReduceScratch scratch)
{
    using A = typename acc_type<X>::type;
    if (e.stopped()) return;
    double red[E::K > 0 ? E::K : 1];
#pragma unroll
    for (int k = 0; k < (E::K > 0 ? E::K : 1); ++k) red[k] = 0.0;
    const std::int64_t row = static_cast<std::int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kBlock + threadIdx.x;
    if (row < rows) {
        const int lane = static_cast<int>(row & 63);
        const std::int64_t slice = row >> 6;
        A acc[BS];
#pragma unroll
        for (int r = 0; r < BS; ++r) acc[r] = A(0);
        const std::int64_t base = slice_ptr[slice];
        const int width = static_cast<int>((slice_ptr[slice + 1] - base) >> 6);
        const int len = row_len[row];
        // U slots per pass, all loads issued before the arithmetic; a slot
        // past the slice's width is clamped to its last run (in bounds) and
        // a slot past the row's length is dropped by a select, so each
        // block row still adds its blocks in slot order.  The next pass's
        // columns are loaded one pass ahead, so a pass's x gathers issue with
        // its value loads (one memory round trip per pass, not two).
        std::int32_t coln[U];
#pragma unroll
        for (int u = 0; u < U; ++u) coln[u] = cols[base + 64 * (u < width ? u : width - 1) + lane];
        for (int k0 = 0; k0 < len; k0 += U) {
            std::int32_t colk[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                colk[u] = coln[u];
                const int kn = k0 + U + u < width ? k0 + U + u : width - 1;
                coln[u] = cols[base + 64 * kn + lane];
            }
            // keeps the prefetch in this pass: left to itself the compiler
            // sinks it to its use in the next pass, after this pass's wait
            __asm__ volatile("" ::: "memory");
            A v[U][BS * BS], xc[U][BS];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k0 + u < width ? k0 + u : width - 1;
                const V* blk = vals + (base + 64 * kk) * (BS * BS);  // element j at blk[dense_offset(lane, j)]
                // FP64 values stream past the caches (4x4 blocks of a 1M-vertex
                // grid are 1 GB), so the gathered x stays resident: 191 -> 174
                // us there.  BF16 values measured the same either way.
#pragma unroll
                for (int j = 0; j < BS * BS; ++j) {
                    if constexpr (sizeof(V) == 8)
                        v[u][j] = static_cast<A>(load_value_nt(blk, dense_offset(lane, j, BS * BS)));
                    else
                        v[u][j] = static_cast<A>(load_value(blk, dense_offset(lane, j, BS * BS)));
                }
#pragma unroll
                for (int c = 0; c < BS; ++c) xc[u][c] = static_cast<A>(x[static_cast<std::int64_t>(colk[u]) * BS + c]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int r = 0; r < BS; ++r) {
                    A t = A(0);
#pragma unroll
                    for (int c = 0; c < BS; ++c) t = dfma(v[u][r * BS + c], xc[u][c], t);
                    acc[r] = k0 + u < len ? acc[r] + t : acc[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < BS; ++r) e.elem(row * BS + r, acc[r], red);
    }
    if constexpr (E::K > 0) store_partials<E::K>(red, E::kMaxMask, scratch);
}

// Dense BSR for any block size, reference CSR layout: one thread per
// (row, block row).
template<class V, class X, class E>
__global__ __launch_bounds__(256) void spmv_csr_bsr(const std::int32_t* __restrict__ offsets,
                                                    const std::int32_t* __restrict__ cols,
                                                    const V* __restrict__ vals,
                                                    const X* __restrict__ x, std::int64_t rows,
                                                    int bs, E e, ReduceScratch scratch)
{
    using A = typename acc_type<X>::type;
    if (e.stopped()) return;
    double red[E::K > 0 ? E::K : 1];
#pragma unroll
    for (int k = 0; k < (E::K > 0 ? E::K : 1); ++k) red[k] = 0.0;
    const std::int64_t total = rows * bs;
    const std::int64_t stride = static_cast<std::int64_t>(gridDim.x) * kBlock;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < total; i += stride) {
        const std::int64_t row = i / bs;
        const int r = static_cast<int>(i % bs);
        A acc = A(0);
        for (std::int32_t en = offsets[row]; en < offsets[row + 1]; ++en) {
            const std::int64_t col = cols[en];
            const std::int64_t blk = static_cast<std::int64_t>(en) * bs * bs + static_cast<std::int64_t>(r) * bs;
            A t = A(0);
            for (int c = 0; c < bs; ++c)
                t = dfma(static_cast<A>(load_value(vals, blk + c)), static_cast<A>(x[col * bs + c]), t);
            acc = acc + t;
        }
        e.elem(i, acc, red);
    }
    if constexpr (E::K > 0) store_partials<E::K>(red, E::kMaxMask, scratch);
}

} // namespace owtk::kern

namespace owtk {

// Launch the SpMV of m with value storage `dtype` and epilogue e.  V/X are
// deduced from dtype by the caller through the typed entry below.
template<class V, class X, class E>
void spmv_dispatch(owtk_matrix m, const V* vals, const V* diag, const X* x, E e)
{
    using namespace owtk::kern;
    const hipStream_t st = m->ctx->stream;
    const auto scratch = m->ctx->scratch();
    const auto& s = m->s;
    const std::int64_t rows = m->owned;
    constexpr int sv = static_cast<int>(sizeof(V));
    constexpr int category = sv == 8 ? prof_spmv_f64 : (sv == 4 ? prof_spmv_f32 : prof_spmv_bf16);
    ProfileScope prof(m->ctx, category, spmv_bytes(m, sv, static_cast<int>(sizeof(X))));
    unsigned int launched = 0;
    if (s.layout == Layout::sell) {
        const std::int64_t blocks = (rows + kBlock - 1) / kBlock;
        if (E::K > 0 && blocks > scratch.stride)
            throw std::logic_error("reduction partials smaller than the SpMV grid");
        if (blocks > 0x7fffffff) throw std::invalid_argument("too many rows");
        const unsigned int grid = static_cast<unsigned int>(blocks);
        launched = grid;
        if (m->dense()) {
            // slots per pass: 2 for FP32/BF16 values (4x4 BF16 blocks of a
            // 1M-vertex grid: 54.6 -> 50.7 us), 1 for FP64, whose 16 values
            // per block already keep enough loads in flight (2 and 4 measured
            // slower)
            constexpr int kBsrU = sizeof(V) == 8 ? 1 : 2;
            switch (m->bs) {
            case 2: spmv_sell_bsr<2, V, X, E, kBsrU><<<grid, kBlock, 0, st>>>(s.slice_ptr.get(), s.row_len.get(), s.columns.get(), vals, x, rows, e, scratch); break;
            case 3: spmv_sell_bsr<3, V, X, E, kBsrU><<<grid, kBlock, 0, st>>>(s.slice_ptr.get(), s.row_len.get(), s.columns.get(), vals, x, rows, e, scratch); break;
            case 4: spmv_sell_bsr<4, V, X, E, kBsrU><<<grid, kBlock, 0, st>>>(s.slice_ptr.get(), s.row_len.get(), s.columns.get(), vals, x, rows, e, scratch); break;
            default: spmv_sell_bsr<1, V, X, E, kBsrU><<<grid, kBlock, 0, st>>>(s.slice_ptr.get(), s.row_len.get(), s.columns.get(), vals, x, rows, e, scratch); break;
            }
        } else if (m->has_diag_array()) {
            launch_sell_cd_bs<V, X, true, E>(m->bs, s.max_width, grid, st, s.slice_ptr.get(), s.row_len.get(),
                                             s.columns.get(), vals, diag, x, rows, e, scratch);
        } else {
            launch_sell_cd_bs<V, X, false, E>(m->bs, s.max_width, grid, st, s.slice_ptr.get(), s.row_len.get(),
                                              s.columns.get(), vals, diag, x, rows, e, scratch);
        }
    } else {
        if (m->dense()) {
            const unsigned int grid = stream_grid(rows * m->bs);
            launched = grid;
            spmv_csr_bsr<V, X, E><<<grid, kBlock, 0, st>>>(s.offsets.get(), s.columns.get(), vals, x, rows, m->bs, e, scratch);
        } else {
            std::int64_t blocks = (rows + kWavesPerBlock - 1) / kWavesPerBlock;
            if (blocks > kMaxReduceBlocks) blocks = kMaxReduceBlocks;
            launched = static_cast<unsigned int>(blocks);
            if (m->has_diag_array())
                spmv_csr_cd<V, X, true, E><<<static_cast<unsigned int>(blocks), kBlock, 0, st>>>(s.offsets.get(), s.columns.get(), vals, diag, x, rows, m->bs, e, scratch);
            else
                spmv_csr_cd<V, X, false, E><<<static_cast<unsigned int>(blocks), kBlock, 0, st>>>(s.offsets.get(), s.columns.get(), vals, diag, x, rows, m->bs, e, scratch);
        }
    }
    OWTK_LAUNCH_CHECK();
    if constexpr (E::K > 0) {
        launch_finish<E::K>(st, EpiFinish<E>{e}, launched, E::kMaxMask, scratch, e.stop);
        OWTK_LAUNCH_CHECK();
    }
}

// FP64 SpMV with an epilogue (the solver path).
template<class E>
void spmv_f64(owtk_matrix m, const double* x, E e)
{
    spmv_dispatch<double, double, E>(m, m->v64.get(), m->d64.get(), x, e);
}

// Low-precision SpMV (FP32 or BF16 values, FP32 vectors).
template<class E>
void spmv_low(owtk_matrix m, owtk_dtype dtype, const float* x, E e)
{
    if (dtype == OWTK_F32) {
        if (!m->v32.get()) throw std::invalid_argument("matrix has no FP32 storage");
        spmv_dispatch<float, float, E>(m, m->v32.get(), m->d32.get(), x, e);
    } else {
        if (!m->v16.get()) throw std::invalid_argument("matrix has no BF16 storage");
        spmv_dispatch<std::uint16_t, float, E>(m, m->v16.get(), m->d16.get(), x, e);
    }
}

} // namespace owtk
