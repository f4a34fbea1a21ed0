// Jacobi, SSOR and ILU(k) preconditioners on the device.
//
// Symbolic work (block partition, ILU fill pattern, update lists, level
// schedules) runs once on the host at creation; numeric updates (value
// gather, ILU factorization, inverse diagonals) and every apply run on the
// GPU.  This is the reference's "symbolic once, numeric many" contract
// (IluLevelPreconditioner::update_values, preconditioner.hpp:269-311).
#include "precond.hpp"
#include "learned.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"
#include "host/host.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <numeric>

using namespace owtk;

namespace owtk {
namespace {

using dev::kBlock;

constexpr double kTol = 64.0 * std::numeric_limits<double>::epsilon();
// LDS per chain workgroup above which the block's solution stays in global memory
constexpr std::size_t kChainLdsBudget = 24 * 1024;

enum Flag : int { flag_ok = 0, flag_zero_pivot = 1, flag_singular_diag = 2, flag_no_diag = 3 };

// ------------------------------------------------------------------ kernels --

// Gather matrix values through encoded sources: src >= 0 -> values[src + c*stride],
// src <= -2 -> diag[-2-src + c*stride], src == -1 -> 0.  out is [c][count].
__global__ __launch_bounds__(256) void k_gather(double* __restrict__ out, const std::int64_t* __restrict__ src,
                                                std::int64_t count, int bs, std::int64_t stride,
                                                const double* __restrict__ values, const double* __restrict__ diag,
                                                double scale)
{
    const std::int64_t total = count * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t c = t / count;
        const std::int64_t e = t - c * count;
        const std::int64_t s = src[e];
        double v = 0.0;
        if (s >= 0) v = values[s + c * stride];
        else if (s <= -2) v = diag[(-2 - s) + c * stride];
        out[t] = scale == 1.0 ? v : scale * v;
    }
}

// Gather from a [c][nsrc] array by index (index < 0 -> 0).
__global__ __launch_bounds__(256) void k_gather_cm(double* __restrict__ out, const std::int64_t* __restrict__ src,
                                                   std::int64_t count, int bs, const double* __restrict__ in,
                                                   std::int64_t nsrc)
{
    const std::int64_t total = count * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t c = t / count;
        const std::int64_t e = t - c * count;
        const std::int64_t s = src[e];
        out[t] = s >= 0 ? in[c * nsrc + s] : 0.0;
    }
}

// SSOR diagonal check (LocalSsorPreconditioner::update_values,
// preconditioner.hpp:142-174): finite and |d| > 64 eps.
__global__ __launch_bounds__(256) void k_check_diag(const double* __restrict__ d, std::int64_t n, double tol, int* flag)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        const double v = d[i];
        if (!(__builtin_isfinite(v) && __builtin_fabs(v) > tol)) atomicExch(flag, static_cast<int>(flag_singular_diag));
    }
}

// Jacobi inverse (JacobiPreconditioner::update_values, preconditioner.hpp:55-71).
__global__ __launch_bounds__(256) void k_jacobi_inverse(double* __restrict__ inv, const std::int64_t* __restrict__ src,
                                                        std::int64_t rows, int bs, std::int64_t stride,
                                                        const double* __restrict__ values,
                                                        const double* __restrict__ diag, double minimum, int* flag)
{
    const std::int64_t total = rows * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t row = t / bs;
        const std::int64_t c = t - row * bs;
        const std::int64_t s = src[row];
        const double v = s >= 0 ? values[s + c * stride] : diag[(-2 - s) + c * stride];
        if (!(__builtin_isfinite(v) && __builtin_fabs(v) > minimum)) atomicExch(flag, static_cast<int>(flag_singular_diag));
        inv[t] = 1.0 / v;
    }
}

template<class X, class V>
__global__ __launch_bounds__(256) void k_jacobi_apply(const V* __restrict__ inv, const X* __restrict__ in,
                                                      X* __restrict__ out, std::int64_t n, const int* stop)
{
    if (stop && *stop) return;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        out[i] = static_cast<X>(inv[i]) * in[i];
}

// Dense block-Jacobi: invert each bs x bs diagonal block by Gauss-Jordan with
// partial pivoting (thread per row).  No reference counterpart: the dense
// BSR form has no diagonal() (dense_block_csr.hpp).
template<int BS>
__global__ __launch_bounds__(256) void k_block_inverse(double* __restrict__ inv, const std::int64_t* __restrict__ slot,
                                                       std::int64_t rows, const double* __restrict__ values,
                                                       int* flag)
{
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    double a[BS][BS], b[BS][BS];
    const std::int64_t ds = slot[row];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            a[r][c] = values[dev::dense_index(ds, r * BS + c, BS * BS)];
            b[r][c] = r == c ? 1.0 : 0.0;
        }
    for (int k = 0; k < BS; ++k) {
        int p = k;
        for (int r = k + 1; r < BS; ++r)
            if (__builtin_fabs(a[r][k]) > __builtin_fabs(a[p][k])) p = r;
        if (!(__builtin_fabs(a[p][k]) > 0.0) || !__builtin_isfinite(a[p][k])) {
            atomicExch(flag, static_cast<int>(flag_singular_diag));
            return;
        }
        if (p != k)
            for (int c = 0; c < BS; ++c) {
                double t = a[k][c]; a[k][c] = a[p][c]; a[p][c] = t;
                t = b[k][c]; b[k][c] = b[p][c]; b[p][c] = t;
            }
        const double inv_p = 1.0 / a[k][k];
        for (int c = 0; c < BS; ++c) {
            a[k][c] *= inv_p;
            b[k][c] *= inv_p;
        }
        for (int r = 0; r < BS; ++r) {
            if (r == k) continue;
            const double f = a[r][k];
            for (int c = 0; c < BS; ++c) {
                a[r][c] = __builtin_fma(-f, a[k][c], a[r][c]);
                b[r][c] = __builtin_fma(-f, b[k][c], b[r][c]);
            }
        }
    }
    double* out = inv + row * (BS * BS);
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) out[r * BS + c] = b[r][c];
}

template<int BS, class X, class V>
__global__ __launch_bounds__(256) void k_block_jacobi_apply(const V* __restrict__ inv, const X* __restrict__ in,
                                                            X* __restrict__ out, std::int64_t rows, const int* stop)
{
    if (stop && *stop) return;
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    X r[BS];
#pragma unroll
    for (int c = 0; c < BS; ++c) r[c] = in[row * BS + c];
    const V* blk = inv + row * (BS * BS);
#pragma unroll
    for (int i = 0; i < BS; ++i) {
        X t = X(0);
#pragma unroll
        for (int c = 0; c < BS; ++c) t = dev::dfma(static_cast<X>(blk[i * BS + c]), r[c], t);
        out[row * BS + i] = t;
    }
}

// Dense-block SSOR in multicolour order (no reference counterpart: the
// reference's DenseBlockCsrMatrix has no SSOR).  Rows of one colour are
// independent, so each colour is one parallel launch:
//   forward  y_i = D_i^-1 (r_i - omega sum_(colour j < colour i) A_ij y_j)
//   backward z_i = D_i^-1 (s D_i y_i - omega sum_(colour j > colour i) A_ij z_j),
// s = omega (2 - omega), i.e. z = s (D + omega U)^-1 D (D + omega L)^-1 r in
// the colour-permuted order.  Ghost columns are dropped (rank-local, like the
// component-diagonal SSOR).  Entries are visited in the row's stored order.
template<int BS, class V, class X, class I, bool FWD>
__global__ __launch_bounds__(256) void k_mc_ssor(const std::int32_t* __restrict__ prow, std::int64_t seg0,
                                                 std::int64_t count, const std::int32_t* __restrict__ slice_ptr,
                                                 const std::int32_t* __restrict__ row_len,
                                                 const std::int32_t* __restrict__ cols, const V* __restrict__ vals,
                                                 const V* __restrict__ dvals, double omega, double scale,
                                                 const X* __restrict__ in, X* __restrict__ y, X* __restrict__ out,
                                                 const I* __restrict__ inv, const int* stop)
{
    using A = typename dev::acc_type<X>::type;
    if (stop && *stop) return;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const std::int64_t q = seg0 + t;  // copy row: consecutive threads read consecutive slots
    const std::int64_t i = prow[q];
    const A w = static_cast<A>(omega);
    A acc[BS];
    if constexpr (FWD) {
#pragma unroll
        for (int r = 0; r < BS; ++r) acc[r] = static_cast<A>(in[i * BS + r]);
    } else {
        A yi[BS];
#pragma unroll
        for (int r = 0; r < BS; ++r) yi[r] = static_cast<A>(y[i * BS + r]);
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            A d = A(0);
#pragma unroll
            for (int c = 0; c < BS; ++c)
                d = dev::dfma(static_cast<A>(dev::load_value(dvals, dev::dense_index(q, r * BS + c, BS * BS))),
                              yi[c], d);
            acc[r] = d * static_cast<A>(scale);
        }
    }
    const std::int64_t base = slice_ptr[q >> 6];
    const int lane = static_cast<int>(q & 63);
    const int len = row_len[q];
    const X* src = FWD ? y : out;
    for (int k = 0; k < len; ++k) {
        const std::int64_t slot = base + 64 * k + lane;
        const std::int64_t j = cols[slot];
        A xj[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) xj[c] = static_cast<A>(src[j * BS + c]);
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < BS; ++c)
                acc[r] = dev::dfma(-(w * static_cast<A>(dev::load_value(vals, dev::dense_index(slot, r * BS + c,
                                                                                              BS * BS)))),
                                   xj[c], acc[r]);
    }
    X* dst = FWD ? y : out;
    const I* blk = inv + i * (BS * BS);
#pragma unroll
    for (int r = 0; r < BS; ++r) {
        A sum = A(0);
#pragma unroll
        for (int c = 0; c < BS; ++c) sum = dev::dfma(static_cast<A>(blk[r * BS + c]), acc[c], sum);
        dst[i * BS + r] = static_cast<X>(sum);
    }
}

// Copy slots' block values from the matrix: dst[dense_index(s, j, bb)] =
// src[dense_index(map[s], j, bb)] (0 for padding).  Thread t writes dst[t].
template<class V>
__global__ __launch_bounds__(256) void k_mc_gather(V* __restrict__ dst, const V* __restrict__ src,
                                                   const std::int64_t* __restrict__ map, std::int64_t slots, int bb)
{
    const std::int64_t total = slots * bb;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t slice = t / (64 * bb);
        int lane = 0, j = 0;
        dev::dense_element(t - slice * 64 * bb, bb, lane, j);
        const std::int64_t sl = slice * 64 + lane;
        const std::int64_t from = map[sl];
        dst[t] = from >= 0 ? src[dev::dense_index(from, j, bb)] : V(0);
    }
}

struct TriArgs {
    const std::int32_t* row0;
    const std::int32_t* nrows;
    const std::int32_t* flev;
    const std::int32_t* blev;
    const std::int32_t* flev_pos;
    const std::int32_t* blev_pos;
    const std::int32_t* frow;
    const std::int32_t* brow;
    const std::int64_t* fptr;
    const std::int64_t* bptr;
    const std::int32_t* fcol;
    const std::int32_t* bcol;
    const std::int32_t* fnout;
    const std::int32_t* bnout;
    std::int64_t fnnz, bnnz, rows;
    int bs;
};

// ILU(k) numeric factorization (IluLevelPreconditioner::factorize,
// preconditioner.hpp:439-485), one workgroup per (block, component), rows in
// forward level order.  F is the FP64 factor, [c][entry].
__global__ __launch_bounds__(256) void k_ilu_factor(TriArgs a, const std::int32_t* __restrict__ lvl_rows,
                                                    const std::int64_t* __restrict__ offsets,
                                                    const std::int64_t* __restrict__ diag_pos,
                                                    const std::int64_t* __restrict__ upd_ptr,
                                                    const std::int64_t* __restrict__ upd_target,
                                                    const std::int64_t* __restrict__ upd_upper,
                                                    const std::int64_t* __restrict__ columns_global,
                                                    double* __restrict__ F, std::int64_t nnz,
                                                    double* __restrict__ inv, int* flag)
{
    const int c = blockIdx.x % a.bs;
    const int b = blockIdx.x / a.bs;
    const std::int64_t r0 = a.row0[b];
    double* Fc = F + static_cast<std::int64_t>(c) * nnz;
    const int l_begin = a.flev[b], l_end = a.flev[b + 1] - 1;
    for (int l = l_begin; l < l_end; ++l) {
        const int p0 = a.flev_pos[l], p1 = a.flev_pos[l + 1];
        for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
            const std::int64_t row = r0 + lvl_rows[r0 + p];
            const std::int64_t dpos = diag_pos[row];
            for (std::int64_t e = offsets[row]; e < dpos; ++e) {
                const std::int64_t k = columns_global[e];
                const double pivot = Fc[diag_pos[k]];
                if (__builtin_fabs(pivot) <= kTol) atomicExch(flag, static_cast<int>(flag_zero_pivot));
                const double lv = Fc[e] / pivot;
                Fc[e] = lv;
                for (std::int64_t u = upd_ptr[e]; u < upd_ptr[e + 1]; ++u) {
                    const std::int64_t t = upd_target[u];
                    Fc[t] = __builtin_fma(-lv, Fc[upd_upper[u]], Fc[t]);
                }
            }
        }
        __syncthreads();
    }
    const std::int64_t n = a.nrows[b];
    for (std::int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const std::int64_t row = r0 + i;
        const double d = Fc[diag_pos[row]];
        if (!(__builtin_isfinite(d) && __builtin_fabs(d) > kTol)) atomicExch(flag, static_cast<int>(flag_zero_pivot));
        inv[static_cast<std::int64_t>(c) * a.rows + row] = 1.0 / d;
    }
}

// ILU(k) numeric factorization in 64-row chunks (the IKJ order of
// IluLevelPreconditioner::factorize, preconditioner.hpp:439-485, per row),
// one wave per (block, component), rows in natural order, one row per lane:
//   1. a lane loads its row's entries into LDS (rv[slot][lane]), and the
//      chunk's static program (IluChunkProgram) is staged in LDS;
//   2. it gathers the final values of earlier chunks that its lower entries
//      need (pivots u_kk and sources u_kj) in one batch (ga[t][lane]) and
//      applies those lower entries in column order;
//   3. the chunk's own lower entries run in 64 steps: at step s row s is
//      final, and every lane whose row has an entry in column s divides by
//      row s's pivot and updates its targets from row s's entries (read from
//      LDS), which keeps each row's column order.  With a bound NUB on the
//      updates per op, a step's reads go out together and its updates take
//      no branch (an ILU(0) factor of a triangulated grid has NUB = 3);
//   4. the chunk's rows are stored, fenced for the next chunks' gathers.
// Each row sees the same operations in the same order as the reference's
// loop, so the factors are bitwise those of the level-scheduled kernel.
struct IluChunkArgs {
    const std::int32_t* chunk_first;
    const std::int64_t* hdr;
    const std::int32_t* src_a;
    const std::uint32_t* op_a;
    const std::uint32_t* op_b;
    const std::uint16_t* upd_a;
    const std::uint16_t* upd_b;
    int kmax, max_src, max_a, max_nua, max_b, max_nub;
};

inline std::size_t ilu_chunk_lds(const IluChunkProgram& P)
{
    return static_cast<std::size_t>(64) * (8 * (P.kmax + P.max_src) + 4 * (P.max_a + P.max_b)
                                           + 2 * (P.max_a * P.max_nua + P.max_b * P.max_nub));
}

// count values load(j) -> store(j, v), eight loads in flight before their
// stores; load() gets only indices below count
template<class T, class Load, class Store>
__device__ inline void copy_batched(int count, Load load, Store store)
{
    for (int j0 = 0; j0 < count; j0 += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = load(j0 + u < count ? j0 + u : count - 1);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < count) store(j0 + u, v[u]);
    }
}

// A chunk's header, row extents, program words and row values held in
// registers (the prefetching kernel loads chunk ch + 1's during chunk ch).
// Its caps bound the chunk's row length and program sizes.
constexpr int kPfRows = 16, kPfOps = 16, kPfUpdates = 48;
struct ChunkPrefetch {
    std::int64_t row_first, src_off, base;
    int cnt, nsrc, na, nua, nb, nub, len;
    double vd[kPfRows];
    std::uint32_t v32[kPfOps];      // op words: earlier-chunk ops, then this chunk's
    std::uint16_t v16[kPfUpdates];  // update words, likewise
};

// Header, row extents and program words of chunk ch (no row values).
__device__ inline void prefetch_program(const IluChunkArgs& P, const std::int64_t* __restrict__ offsets, int ch,
                                        int lane, ChunkPrefetch& n)
{
    const std::int64_t* h = P.hdr + static_cast<std::int64_t>(ch) * 12;
    n.row_first = h[0];
    n.cnt = static_cast<int>(h[1]);
    n.src_off = h[2];
    n.nsrc = static_cast<int>(h[3]);
    const std::int64_t opa_off = h[4], upa_off = h[6], opb_off = h[8], upb_off = h[10];
    n.na = static_cast<int>(h[5]);
    n.nua = static_cast<int>(h[7]);
    n.nb = static_cast<int>(h[9]);
    n.nub = static_cast<int>(h[11]);
    n.base = 0;
    n.len = 0;
    if (lane < n.cnt) {
        n.base = offsets[n.row_first + lane];
        n.len = static_cast<int>(offsets[n.row_first + lane + 1] - n.base);
    }
    const int n32 = n.na + n.nb, na16 = n.na * n.nua, n16 = na16 + n.nb * n.nub;
#pragma unroll
    for (int j = 0; j < kPfOps; ++j)
        if (j < n32) n.v32[j] = j < n.na ? P.op_a[opa_off + j * 64 + lane] : P.op_b[opb_off + (j - n.na) * 64 + lane];
#pragma unroll
    for (int j = 0; j < kPfUpdates; ++j)
        if (j < n16) n.v16[j] = j < na16 ? P.upd_a[upa_off + j * 64 + lane] : P.upd_b[upb_off + (j - na16) * 64 + lane];
}

// Row values of a prefetched chunk (after its row extents have arrived).
__device__ inline void prefetch_rows(const double* __restrict__ Fc, int kmax, ChunkPrefetch& n)
{
    const int lastq = n.len > 0 ? n.len - 1 : 0;
#pragma unroll
    for (int q = 0; q < kPfRows; ++q)
        if (q < kmax) n.vd[q] = Fc[n.base + (q < lastq ? q : lastq)];
}

inline bool ilu_chunks_prefetchable(const IluChunkProgram& P)
{
    return P.kmax <= kPfRows && P.max_a + P.max_b <= kPfOps
           && P.max_a * P.max_nua + P.max_b * P.max_nub <= kPfUpdates;
}

template<int NUB, bool PF>
__global__ __launch_bounds__(64) void k_ilu_factor_chunks(IluChunkArgs P, const std::int64_t* __restrict__ offsets,
                                                          const std::int64_t* __restrict__ diag_pos,
                                                          const std::int32_t* __restrict__ row0,
                                                          const std::int32_t* __restrict__ nrows, int bs,
                                                          double* __restrict__ F, std::int64_t nnz,
                                                          std::int64_t rows, double* __restrict__ inv, int* flag)
{
    extern __shared__ double lds_d[];
    const int lane = static_cast<int>(threadIdx.x);
    const int c = static_cast<int>(blockIdx.x) % bs;
    const int b = static_cast<int>(blockIdx.x) / bs;
    double* Fc = F + static_cast<std::int64_t>(c) * nnz;
    double* rv = lds_d;                                   // [kmax][64]
    double* ga = rv + P.kmax * 64;                        // [max_src][64]
    auto* opa = reinterpret_cast<std::uint32_t*>(ga + P.max_src * 64);  // [max_a][64]
    std::uint32_t* opb = opa + P.max_a * 64;              // [max_b][64]
    auto* upa = reinterpret_cast<std::uint16_t*>(opb + P.max_b * 64);  // [max_a*max_nua][64]
    std::uint16_t* upb = upa + P.max_a * P.max_nua * 64;  // [max_b*max_nub][64]
    const int ch_end = P.chunk_first[b + 1];
    ChunkPrefetch cur, nxt;
    if constexpr (PF) {
        if (P.chunk_first[b] < ch_end) {
            prefetch_program(P, offsets, P.chunk_first[b], lane, cur);
            prefetch_rows(Fc, P.kmax, cur);
        }
    }
    for (int ch = P.chunk_first[b]; ch < ch_end; ++ch) {
        const std::int64_t* h = P.hdr + static_cast<std::int64_t>(ch) * 12;
        const int cnt = PF ? cur.cnt : static_cast<int>(h[1]);
        const std::int64_t src_off = PF ? cur.src_off : h[2];
        const int nsrc = PF ? cur.nsrc : static_cast<int>(h[3]);
        const int na = PF ? cur.na : static_cast<int>(h[5]), nua = PF ? cur.nua : static_cast<int>(h[7]);
        const int nb = PF ? cur.nb : static_cast<int>(h[9]), nub = PF ? cur.nub : static_cast<int>(h[11]);
        std::int64_t base = 0;
        int len = 0;
        if constexpr (PF) {
            // 1. the prefetched row values and program into LDS; then chunk
            //    ch + 1's header, row extents and program start loading
            base = cur.base;
            len = cur.len;
#pragma unroll
            for (int q = 0; q < kPfRows; ++q)
                if (q < P.kmax) rv[q * 64 + lane] = q < len ? cur.vd[q] : 0.0;
            const int na16 = na * nua;
#pragma unroll
            for (int j = 0; j < kPfOps; ++j) {
                if (j < na) opa[j * 64 + lane] = cur.v32[j];
                else if (j < na + nb) opb[(j - na) * 64 + lane] = cur.v32[j];
            }
#pragma unroll
            for (int j = 0; j < kPfUpdates; ++j) {
                if (j < na16) upa[j * 64 + lane] = cur.v16[j];
                else if (j < na16 + nb * nub) upb[(j - na16) * 64 + lane] = cur.v16[j];
            }
            if (ch + 1 < ch_end) prefetch_program(P, offsets, ch + 1, lane, nxt);
        } else {
        const std::int64_t row_first = h[0];
        const std::int64_t opa_off = h[4], upa_off = h[6], opb_off = h[8], upb_off = h[10];
        if (lane < cnt) {
            base = offsets[row_first + lane];
            len = static_cast<int>(offsets[row_first + lane + 1] - base);
        }
        // 1. row values and the chunk's program, all loads independent and
        //    issued in batches (a plain loop would wait for each load before
        //    its LDS store)
        const int lastq = len > 0 ? len - 1 : 0;
        copy_batched<double>(P.kmax, [&](int q) { return Fc[base + (q < lastq ? q : lastq)]; },
                             [&](int q, double v) { rv[q * 64 + lane] = q < len ? v : 0.0; });
        copy_batched<std::uint32_t>(na, [&](int j) { return P.op_a[opa_off + j * 64 + lane]; },
                                    [&](int j, std::uint32_t v) { opa[j * 64 + lane] = v; });
        copy_batched<std::uint16_t>(na * nua, [&](int j) { return P.upd_a[upa_off + j * 64 + lane]; },
                                    [&](int j, std::uint16_t v) { upa[j * 64 + lane] = v; });
        copy_batched<std::uint32_t>(nb, [&](int j) { return P.op_b[opb_off + j * 64 + lane]; },
                                    [&](int j, std::uint32_t v) { opb[j * 64 + lane] = v; });
        copy_batched<std::uint16_t>(nb * nub, [&](int j) { return P.upd_b[upb_off + j * 64 + lane]; },
                                    [&](int j, std::uint16_t v) { upb[j * 64 + lane] = v; });
        }
        // 2. earlier chunks: gather (source indices, then values, each batch
        //    in flight together), then the lower entries in column order
        for (int t0 = 0; t0 < nsrc; t0 += 8) {
            std::int32_t idx[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = t0 + u < nsrc ? t0 + u : nsrc - 1;
                idx[u] = P.src_a[src_off + t * 64 + lane];
            }
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = Fc[idx[u] >= 0 ? idx[u] : 0];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (t0 + u < nsrc) ga[(t0 + u) * 64 + lane] = idx[u] >= 0 ? v[u] : 0.0;
        }
        // chunk ch + 1's row values (their extents have arrived by now) load
        // during this chunk's LDS work; no row of chunk ch + 1 is written
        // before its own store
        if constexpr (PF) {
            if (ch + 1 < ch_end) prefetch_rows(Fc, P.kmax, nxt);
        }
        for (int j = 0; j < na; ++j) {
            const std::uint32_t d = opa[j * 64 + lane];
            if (d == 0xffffffffu) continue;
            const int es = static_cast<int>(d & 255u), pt = static_cast<int>((d >> 8) & 255u);
            const int nu = static_cast<int>((d >> 16) & 255u);
            const double pivot = ga[pt * 64 + lane];
            if (__builtin_fabs(pivot) <= kTol) atomicExch(flag, static_cast<int>(flag_zero_pivot));
            const double lv = rv[es * 64 + lane] / pivot;
            rv[es * 64 + lane] = lv;
            for (int u = 0; u < nu; ++u) {
                const std::uint16_t w = upa[(j * nua + u) * 64 + lane];
                const int ts = w & 255, st = w >> 8;
                rv[ts * 64 + lane] = __builtin_fma(-lv, ga[st * 64 + lane], rv[ts * 64 + lane]);
            }
        }
        // 3. the chunk's own lower entries, one pivot row per step
        if constexpr (NUB > 0) {
            // At most NUB updates per op: an op's update words load with the
            // op, and a step issues all its reads (pivot, the lower entry,
            // row s's sources, its own targets) before its division, then
            // its stores, with no branch per update (an update past the op's
            // count reads a valid slot and writes the lane's trash slot).
            // The targets are distinct slots of the lane's own row, neither
            // the lower entry nor row s, so the values are those of the
            // one-at-a-time order.
            __shared__ double trash[64];
            int jb = 0;
            std::uint32_t d = nb > 0 ? opb[lane] : 0xffffffffu;
            std::uint16_t w[NUB];
#pragma unroll
            for (int u = 0; u < NUB; ++u) w[u] = 0;
            if (nb > 0 && nub > 0) {
#pragma unroll
                for (int u = 0; u < NUB; ++u) w[u] = upb[(u < nub ? u : nub - 1) * 64 + lane];
            }
            for (int s = 0; s < cnt; ++s) {
                __builtin_amdgcn_wave_barrier();
                if (d != 0xffffffffu && static_cast<int>(d & 255u) == s) {
                    const int es = static_cast<int>((d >> 8) & 255u), ps = static_cast<int>((d >> 16) & 255u);
                    const int nu = static_cast<int>(d >> 24);
                    const double pivot = rv[ps * 64 + s];
                    const double own = rv[es * 64 + lane];
                    double src[NUB], tgt[NUB];
                    double* dst[NUB];
#pragma unroll
                    for (int u = 0; u < NUB; ++u) {
                        const bool on = u < nu;
                        dst[u] = on ? &rv[(w[u] & 255) * 64 + lane] : &trash[lane];
                        src[u] = rv[(on ? (w[u] >> 8) : 0) * 64 + s];
                        tgt[u] = *dst[u];
                    }
                    if (__builtin_fabs(pivot) <= kTol) atomicExch(flag, static_cast<int>(flag_zero_pivot));
                    const double lv = own / pivot;
                    rv[es * 64 + lane] = lv;
#pragma unroll
                    for (int u = 0; u < NUB; ++u) *dst[u] = __builtin_fma(-lv, src[u], tgt[u]);
                    ++jb;
                    const int jn = jb < nb ? jb : nb - 1;
                    d = jb < nb ? opb[jn * 64 + lane] : 0xffffffffu;
                    if (nub > 0) {
#pragma unroll
                        for (int u = 0; u < NUB; ++u) w[u] = upb[(jn * nub + (u < nub ? u : nub - 1)) * 64 + lane];
                    }
                }
            }
        } else {
            int jb = 0;
            std::uint32_t d = nb > 0 ? opb[lane] : 0xffffffffu;
            for (int s = 0; s < cnt; ++s) {
                __builtin_amdgcn_wave_barrier();
                if (d != 0xffffffffu && static_cast<int>(d & 255u) == s) {
                    const int es = static_cast<int>((d >> 8) & 255u), ps = static_cast<int>((d >> 16) & 255u);
                    const int nu = static_cast<int>(d >> 24);
                    const double pivot = rv[ps * 64 + s];
                    if (__builtin_fabs(pivot) <= kTol) atomicExch(flag, static_cast<int>(flag_zero_pivot));
                    const double lv = rv[es * 64 + lane] / pivot;
                    rv[es * 64 + lane] = lv;
                    for (int u = 0; u < nu; ++u) {
                        const std::uint16_t w = upb[(jb * nub + u) * 64 + lane];
                        const int ts = w & 255, qs = w >> 8;
                        rv[ts * 64 + lane] = __builtin_fma(-lv, rv[qs * 64 + s], rv[ts * 64 + lane]);
                    }
                    ++jb;
                    d = jb < nb ? opb[jb * 64 + lane] : 0xffffffffu;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // 4. store; later chunks of this wave gather these values
        for (int q = 0; q < len; ++q) Fc[base + q] = rv[q * 64 + lane];
        __threadfence_block();
        __builtin_amdgcn_wave_barrier();
        if constexpr (PF) cur = nxt;
    }
    const std::int64_t r0 = row0[b], n = nrows[b];
    for (std::int64_t i = lane; i < n; i += 64) {
        const std::int64_t row = r0 + i;
        const double dg = Fc[diag_pos[row]];
        if (!(__builtin_isfinite(dg) && __builtin_fabs(dg) > kTol)) atomicExch(flag, static_cast<int>(flag_zero_pivot));
        inv[static_cast<std::int64_t>(c) * rows + row] = 1.0 / dg;
    }
}

enum class TriMode { ilu, ssor };

template<class V> __device__ inline auto vload(const V* p, std::int64_t i)
{
    if constexpr (sizeof(V) == 2) return dev::bf16_to_float(p[i]);
    else return p[i];
}

// Level-scheduled triangular apply, one workgroup per (block, component).
//  ILU : forward y_i = r_i - sum_j<i l_ij y_j; backward x_i = (y_i - sum_j>i u_ij x_j) * inv_i
//        (IluLevelPreconditioner::apply, preconditioner.hpp:320-363)
//  SSOR: forward y_i = (r_i - sum_j<i (w a_ij) y_j) / d_i; backward
//        z_i = (d_i y_i - sum_j>i (w a_ij) z_j) / d_i, then * w(2-w) if w != 1
//        (preconditioner.hpp:181-239, split_block_csr.hpp:243-309)
// With LDS the block's component vector lives in LDS; otherwise it is kept in
// `out` (strided by bs).
template<TriMode MODE, class X, class V, bool LDS>
__global__ __launch_bounds__(256) void k_tri_apply(TriArgs a, const V* __restrict__ fval, const V* __restrict__ bval,
                                                   const X* __restrict__ dval, const X* in, X* out, X scale,
                                                   const int* stop)
{
    extern __shared__ unsigned char smem_raw[];
    if (stop && *stop) return;
    const int bs = a.bs;
    const int c = blockIdx.x % bs;
    const int b = blockIdx.x / bs;
    const std::int64_t r0 = a.row0[b];
    const std::int64_t nb = a.nrows[b];
    X* lds = reinterpret_cast<X*>(smem_raw);
    auto Y = [&](std::int64_t i) -> X& {
        if constexpr (LDS) return lds[i];
        else return out[(r0 + i) * bs + c];
    };
    if constexpr (LDS) {
        for (std::int64_t i = threadIdx.x; i < nb; i += blockDim.x) lds[i] = in[(r0 + i) * bs + c];
    } else {
        if (in != out)
            for (std::int64_t i = threadIdx.x; i < nb; i += blockDim.x) out[(r0 + i) * bs + c] = in[(r0 + i) * bs + c];
    }
    __syncthreads();
    const V* fv = fval + static_cast<std::int64_t>(c) * a.fnnz;
    const V* bv = bval + static_cast<std::int64_t>(c) * a.bnnz;
    const X* dv = dval + static_cast<std::int64_t>(c) * a.rows + r0;
    for (int l = a.flev[b]; l < a.flev[b + 1] - 1; ++l) {
        const int p0 = a.flev_pos[l], p1 = a.flev_pos[l + 1];
        for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
            const std::int64_t gp = r0 + p;
            const std::int32_t i = a.frow[gp];
            X acc = Y(i);
            for (std::int64_t e = a.fptr[gp]; e < a.fptr[gp + 1]; ++e)
                acc = dev::dfma(-static_cast<X>(vload(fv, e)), Y(a.fcol[e]), acc);
            if constexpr (MODE == TriMode::ssor) acc = acc / dv[i];
            Y(i) = acc;
        }
        __syncthreads();
    }
    for (int l = a.blev[b]; l < a.blev[b + 1] - 1; ++l) {
        const int p0 = a.blev_pos[l], p1 = a.blev_pos[l + 1];
        for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
            const std::int64_t gp = r0 + p;
            const std::int32_t i = a.brow[gp];
            X acc;
            if constexpr (MODE == TriMode::ssor) acc = dv[i] * Y(i);
            else acc = Y(i);
            for (std::int64_t e = a.bptr[gp]; e < a.bptr[gp + 1]; ++e)
                acc = dev::dfma(-static_cast<X>(vload(bv, e)), Y(a.bcol[e]), acc);
            if constexpr (MODE == TriMode::ssor) acc = acc / dv[i];
            else acc = acc * dv[i];
            Y(i) = acc;
        }
        __syncthreads();
    }
    if constexpr (LDS) {
        for (std::int64_t i = threadIdx.x; i < nb; i += blockDim.x) {
            X v = lds[i];
            if constexpr (MODE == TriMode::ssor)
                if (scale != X(1)) v = v * scale;
            out[(r0 + i) * bs + c] = v;
        }
    } else if constexpr (MODE == TriMode::ssor) {
        if (scale != X(1))
            for (std::int64_t i = threadIdx.x; i < nb; i += blockDim.x) out[(r0 + i) * bs + c] *= scale;
    }
}

__device__ inline double lane_value(double v, int lane)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

__device__ inline float lane_value(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// One sweep of a chain-scheduled triangular solve by one wave.  Rows are
// taken in chunks of 64 in sweep order (ascending rows forward, descending
// backward), lane l owning the chunk's l-th row.  Each row's entries are
// stored as [entries whose row lies before the chunk] then [entries inside
// the chunk, by the step at which that row is finished].  The first group is
// final already and read from LDS.  The second group is scattered into an
// LDS tile, tile[step][lane] = coefficient (zero elsewhere); step s then
// broadcasts lane s's finished value with v_readlane and every lane folds in
// its tile entry for step s.  A lane's entries all precede its own step, so
// its value is final from its step on and is stored once at the chunk end.
// Forward sweeps of sorted rows accumulate in the reference's order; backward
// sweeps take their in-chunk entries in the order they become available.
template<TriMode MODE, bool FWD, class X, class V, int EI>
__device__ inline void chain_sweep(const std::int64_t* __restrict__ ptr, const std::int32_t* __restrict__ nout,
                                   const std::int32_t* __restrict__ cols, const V* __restrict__ vals,
                                   const X* __restrict__ dv, X* y, int ys, X* tile, int nb, int lane)
{
    for (int cs = 0; cs < nb; cs += 64) {
        const int p = cs + lane;
        const bool valid = p < nb;
        const int i = FWD ? p : nb - 1 - p;
        X acc = X(0);
        X dmine = X(1);
        int st[EI];
#pragma unroll
        for (int k = 0; k < EI; ++k) st[k] = -1;
        if (valid) {
            const std::int64_t e0 = ptr[p];
            const std::int64_t e1 = ptr[p + 1];
            const std::int64_t em = e0 + nout[p];
            if constexpr (MODE == TriMode::ssor || !FWD) dmine = dv[i];
            if constexpr (MODE == TriMode::ssor && !FWD) acc = dmine * y[i * ys];
            else acc = y[i * ys];
            for (std::int64_t e = e0; e < em; ++e)
                acc = dev::dfma(-static_cast<X>(vload(vals, e)), y[static_cast<std::int64_t>(cols[e]) * ys], acc);
#pragma unroll
            for (int k = 0; k < EI; ++k) {
                if (em + k < e1) {
                    const int c = cols[em + k];
                    st[k] = FWD ? c - cs : nb - 1 - cs - c;
                    tile[st[k] * 64 + lane] = static_cast<X>(vload(vals, em + k));
                }
            }
        }
        __syncthreads();
        const int steps = nb - cs < 64 ? nb - cs : 64;
#pragma unroll
        for (int s = 0; s < 64; ++s) {
            if (s < steps) {
                const X coef = tile[s * 64 + lane];
                X fin;
                if constexpr (MODE == TriMode::ssor) fin = acc / dmine;
                else if constexpr (!FWD) fin = acc * dmine;
                else fin = acc;
                const X v = lane_value(fin, s);
                acc = coef != X(0) ? dev::dfma(-coef, v, acc) : acc;
            }
        }
        X mine;
        if constexpr (MODE == TriMode::ssor) mine = acc / dmine;
        else if constexpr (!FWD) mine = acc * dmine;
        else mine = acc;
        if (valid) y[i * ys] = mine;
#pragma unroll
        for (int k = 0; k < EI; ++k)
            if (st[k] >= 0) tile[st[k] * 64 + lane] = X(0);
        __syncthreads();
    }
}

// GLOBAL_Y: the block's solution lives in `out` itself (L2-resident, strided by
// the block size) instead of LDS, so a workgroup needs only the 64 x 64 tile
// and more of them fit on a CU.  The block's rows belong to this one wave, so
// no other wave reads or writes them.
template<TriMode MODE, class X, class V, int EI, bool GLOBAL_Y>
__global__ __launch_bounds__(64) void k_tri_chain(TriArgs a, const V* __restrict__ fval, const V* __restrict__ bval,
                                                  const X* __restrict__ dval, const X* in, X* out, X scale,
                                                  const int* stop)
{
    extern __shared__ unsigned char smem_raw[];
    if (stop && *stop) return;
    const int bs = a.bs;
    const int c = blockIdx.x % bs;
    const int b = blockIdx.x / bs;
    const std::int64_t r0 = a.row0[b];
    const int nb = a.nrows[b];
    const int lane = threadIdx.x;
    X* tile = reinterpret_cast<X*>(smem_raw);  // [64 steps][64 lanes]
    for (int s = 0; s < 64; ++s) tile[s * 64 + lane] = X(0);
    X* y;
    int ys;
    if constexpr (GLOBAL_Y) {
        y = out + r0 * bs + c;
        ys = bs;
        if (in != out)
            for (int i = lane; i < nb; i += 64) y[i * ys] = in[(r0 + i) * bs + c];
    } else {
        y = tile + 64 * 64;  // [nb]
        ys = 1;
        for (int i = lane; i < nb; i += 64) y[i] = in[(r0 + i) * bs + c];
    }
    __syncthreads();
    const X* dv = dval + static_cast<std::int64_t>(c) * a.rows + r0;
    chain_sweep<MODE, true, X, V, EI>(a.fptr + r0, a.fnout + r0, a.fcol, fval + static_cast<std::int64_t>(c) * a.fnnz,
                                      dv, y, ys, tile, nb, lane);
    chain_sweep<MODE, false, X, V, EI>(a.bptr + r0, a.bnout + r0, a.bcol,
                                       bval + static_cast<std::int64_t>(c) * a.bnnz, dv, y, ys, tile, nb, lane);
    const bool scaled = MODE == TriMode::ssor && scale != X(1);
    if (GLOBAL_Y && !scaled) return;
    for (int i = lane; i < nb; i += 64) {
        X v = y[i * ys];
        if (scaled) v = v * scale;
        out[(r0 + i) * bs + c] = v;
    }
}

// Chain sweeps over an ELL copy of the factors.  A sweep's rows are cut into
// the same 64-row chunks; for chunk g and slot k < K the 64 rows' k-th entry
// (chain order) is contiguous: idx[(g*K + k)*64 + lane] and
// val[c][(g*K + k)*64 + lane].  idx >= 0: block-local column of an entry on a
// row of an earlier chunk; idx <= -2: in-chunk entry finished at step
// -idx - 2; -1: padding.  A chunk's description is then K coalesced 16-bit
// loads with no pointer chase: the next chunk's indices are fetched at the
// start of a chunk and its values (only the slots that hold an entry) at its
// end, so both are in flight during the chunk's 64 steps.  In-chunk
// coefficients go to a packed strictly-lower-triangular LDS tile (position p
// only depends on steps s < p), half the dense tile, so more one-wave
// workgroups fit per CU.
__host__ __device__ constexpr int tri_base(int s) { return (s * (127 - s)) / 2; }  // first packed entry of step s

// The sweep's workgroup is one wave.  LDS instructions of a wave complete in
// issue order, and so do its vector memory instructions to one address, so
// lanes see each other's earlier tile and solution writes without a
// workgroup barrier; only the compiler must keep the order.  (__syncthreads
// would also wait for every outstanding load, including the next chunk's
// prefetch.)
__device__ inline void wave_order() { __asm__ volatile("" ::: "memory"); }

template<class X>
__device__ inline X chain_finish(X acc, X, X, std::integral_constant<int, 0>)  // ILU forward: unit diagonal
{
    return acc;
}
template<class X>
__device__ inline X chain_finish(X acc, X d, X, std::integral_constant<int, 1>)  // ILU backward: inverse diagonal
{
    return acc * d;
}
// SSOR: acc / d from the reciprocal and one fma-based correction (Markstein:
// with rinv = RN(1/d) and q = RN(acc rinv), the residual acc - q d is exact
// and RN(q + residual rinv) is the correctly rounded quotient), which keeps
// the division's long latency off the step-to-step dependency chain.
template<class X>
__device__ inline X chain_finish(X acc, X d, X rinv, std::integral_constant<int, 2>)
{
    const X q = acc * rinv;
    const X r = dev::dfma(-q, d, acc);
    return dev::dfma(r, rinv, q);
}

template<TriMode MODE, bool FWD>
using FinishKind = std::integral_constant<int, MODE == TriMode::ssor ? 2 : (FWD ? 0 : 1)>;

// A 64-bit value from another lane chosen per lane (ds_bpermute); every lane
// of the wave must execute it.
template<class X>
__device__ inline X lane_pick(X v, int src)
{
    if constexpr (sizeof(X) == 8) {
        const long long bits = __double_as_longlong(v);
        const int lo = __shfl(static_cast<int>(bits), src, 64);
        const int hi = __shfl(static_cast<int>(bits >> 32), src, 64);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    } else {
        return __shfl(v, src, 64);
    }
}

// One sweep over the ELL copy.  The host classifies every slot (see
// build_ell), so the chunk set-up is branch-free:
//   idx >= 0              entry on a chunk before q-1: block-local row of y,
//                         loaded one chunk ahead (it is final by then);
//   -2018 <= idx <= -3    in-chunk entry: its value goes to packed tile
//                         entry -idx - 2;
//   idx <= -2019          entry on chunk q-1: lane -idx - 2019 of that
//                         chunk, taken from the lanes' final values by a
//                         cross-lane shuffle;
//   idx == -1             padding (value 0, contributes nothing).
// Every slot's index and value is loaded (padding values are zeros), tile
// writes of other kinds go to a spare LDS entry, and out-of-chunk
// contributions are selected, not branched on.  Indices are loaded two
// chunks ahead, values, own row value, diagonal and older y one chunk ahead.
// Each row folds its out-of-chunk entries in slot order, then the in-chunk
// steps.  The steps read their tile coefficients at per-lane addresses
// computed once per kernel (taddr), so a step is one LDS read, the
// broadcast and one FMA.
template<TriMode MODE, bool FWD, class X, class V, int E>
__device__ __attribute__((always_inline)) inline void ell_sweep(const std::int16_t* __restrict__ idx, const V* __restrict__ val, std::int64_t g0,
                                 const X* __restrict__ dv, X* y, int ys, X* tri, const int (&taddr)[64], int nb,
                                 int lane)
{
    constexpr int kTrash = 1 + tri_base(64);  // spare tile entry for writes that are not in-chunk
    const int chunks = (nb + 63) >> 6;
    constexpr std::int64_t stride = static_cast<std::int64_t>(E) * 64;  // one chunk's slots
    auto row_of = [&](int pos) { return FWD ? pos : nb - 1 - pos; };
    auto load_idx = [&](int q, int (&out)[E]) {
        const std::int64_t base = (g0 + (q < chunks ? q : chunks - 1)) * stride + lane;
#pragma unroll
        for (int k = 0; k < E; ++k) out[k] = static_cast<int>(__builtin_nontemporal_load(&idx[base + k * 64]));
    };
    auto load_val = [&](int q, X (&out)[E]) {
        const std::int64_t base = (g0 + (q < chunks ? q : chunks - 1)) * stride + lane;
#pragma unroll
        for (int k = 0; k < E; ++k) out[k] = static_cast<X>(vload(val, base + k * 64));
    };
    // older entries' y, for the slots of the chunk whose indices are `ix`
    auto load_old = [&](const int (&ix)[E], X (&out)[E]) {
#pragma unroll
        for (int k = 0; k < E; ++k) out[k] = y[static_cast<std::int64_t>(ix[k] > 0 ? ix[k] : 0) * ys];
    };
    int ci[E], ni[E];
    X cv[E], pyo[E];
    load_idx(0, ci);
    load_idx(1, ni);
    load_val(0, cv);
#pragma unroll
    for (int k = 0; k < E; ++k) pyo[k] = X(0);  // chunk 0 has no older chunks
    const int p0 = lane < nb ? lane : nb - 1;
    X cd = X(1);
    if constexpr (MODE == TriMode::ssor || !FWD) cd = dv[row_of(p0)];
    X cacc = lane < nb ? y[row_of(p0) * ys] : X(0);
    X yprev = X(0);  // chunk q-1's final value of this lane's row
    for (int q = 0; q < chunks; ++q) {
        const int p = q * 64 + lane;
        const bool valid = p < nb;
        // chunk q+1 (indices of q+2), in flight during this chunk's steps
        int nni[E];
        load_idx(q + 2, nni);
        X ncv[E], npyo[E];
        load_val(q + 1, ncv);
        const int pn = p + 64 < nb ? p + 64 : nb - 1;
        X nacc = y[row_of(pn) * ys], nd = X(1);
        if constexpr (MODE == TriMode::ssor || !FWD) nd = dv[row_of(pn)];
        // in-chunk coefficients to the tile
#pragma unroll
        for (int k = 0; k < E; ++k) tri[ci[k] <= -3 && ci[k] >= -2018 ? -ci[k] - 2 : kTrash] = cv[k];
        X acc = cacc;
        if constexpr (MODE == TriMode::ssor && !FWD) acc = cd * acc;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const X sh = lane_pick(yprev, (-ci[k] - 2019) & 63);
            const X yo = ci[k] >= 0 ? pyo[k] : sh;
            const X upd = dev::dfma(-cv[k], yo, acc);
            acc = ci[k] >= 0 || ci[k] <= -2019 ? upd : acc;
        }
        // Chunk q-1's results are stored only now, after the wait for the
        // loads above: stores count with the loads in vmcnt, so a store
        // issued before them would be waited for there.  The loads of chunk
        // q+1's entries on chunks before q (q-1 among them) follow the store
        // (one wave's accesses to one address stay in order); the next wait
        // for them is a whole chunk of steps later.
        if (q > 0) y[row_of(p - 64) * ys] = yprev;
        load_old(ni, npyo);
        X rinv = X(1);
        if constexpr (MODE == TriMode::ssor) rinv = X(1) / cd;
        wave_order();
        // step st: lane st's row is final; every later lane folds it in with
        // its tile coefficient (lanes at or before st read the zero guard, so
        // a finished row keeps its value)
        const int steps = nb - q * 64 < 64 ? nb - q * 64 : 64;
        if (steps == 64) {
#pragma unroll
            for (int st = 0; st < 64; ++st) {
                const X coef = tri[taddr[st]];
                const X v = lane_value(chain_finish(acc, cd, rinv, FinishKind<MODE, FWD>{}), st);
                acc = dev::dfma(-coef, v, acc);
            }
        } else {
            for (int st = 0; st < steps; ++st) {
                const X coef = tri[lane > st ? tri_base(st) + lane - st : 0];
                const X v = lane_value(chain_finish(acc, cd, rinv, FinishKind<MODE, FWD>{}), st);
                acc = dev::dfma(-coef, v, acc);
            }
        }
        const X fin = chain_finish(acc, cd, rinv, FinishKind<MODE, FWD>{});
        yprev = valid ? fin : X(0);
        if (q + 1 == chunks && valid) y[row_of(p) * ys] = fin;  // the last chunk's results
#pragma unroll
        for (int k = 0; k < E; ++k) tri[ci[k] <= -3 && ci[k] >= -2018 ? -ci[k] - 2 : kTrash] = X(0);
#pragma unroll
        for (int k = 0; k < E; ++k) {
            ci[k] = ni[k];
            ni[k] = nni[k];
            cv[k] = ncv[k];
            pyo[k] = npyo[k];
        }
        cacc = nacc;
        cd = nd;
        wave_order();
    }
}

struct EllArgs {
    const std::int32_t* chunk0;  // [blocks] first global chunk of each block
    const std::int16_t* fidx;
    const std::int16_t* bidx;
    std::int64_t fslots, bslots;  // values per component: chunks * K * 64
    int fk, bk;                   // slots per row of each sweep (both the kernel's E)
};

// Two waves per SIMD at least (at most 256 VGPRs), so all 1,536 one-wave
// workgroups of a 2M-vertex, 3-component apply in 4096-vertex blocks are
// resident at once.
template<TriMode MODE, class X, class V, int E, bool GLOBAL_Y>
__global__ __launch_bounds__(64, 2) void k_tri_ell(TriArgs a, EllArgs l, const V* __restrict__ fval,
                                                const V* __restrict__ bval, const X* __restrict__ dval, const X* in,
                                                X* out, X scale, const int* stop)
{
    extern __shared__ unsigned char smem_raw[];
    if (stop && *stop) return;
    const int bs = a.bs;
    // the bs component waves of one block on one XCD: they read the same indices
    const unsigned int wg = kern::xcd_block(blockIdx.x, gridDim.x);
    const int c = static_cast<int>(wg % bs);
    const int b = static_cast<int>(wg / bs);
    const std::int64_t r0 = a.row0[b];
    const int nb = a.nrows[b];
    const int lane = threadIdx.x;
    constexpr int kTri = 2 + tri_base(64);  // guard entry, packed triangle, spare entry
    X* tri = reinterpret_cast<X*>(smem_raw);
    for (int s = lane; s < kTri; s += 64) tri[s] = X(0);
    X* y;
    int ys;
    if constexpr (GLOBAL_Y) {
        y = out + r0 * bs + c;
        ys = bs;
    } else {
        y = tri + kTri;  // [nb]
        ys = 1;
    }
    // the right-hand side into y, 16 loads in flight per lane (in and out
    // may be one vector, so a plain loop would wait for each load in turn)
    if (!GLOBAL_Y || in != out) {
        for (int i0 = 0; i0 < nb; i0 += 16 * 64) {
            const bool full = i0 + 16 * 64 <= nb;  // whole batches store without per-lane branches
            X v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = i0 + u * 64 + lane;
                v[u] = in[(r0 + (i < nb ? i : nb - 1)) * bs + c];
            }
            if (full) {
#pragma unroll
                for (int u = 0; u < 16; ++u) y[(i0 + u * 64 + lane) * ys] = v[u];
            } else {
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int i = i0 + u * 64 + lane;
                    if (i < nb) y[i * ys] = v[u];
                }
            }
        }
    }
    __syncthreads();
    const X* dv = dval + static_cast<std::int64_t>(c) * a.rows + r0;
    const std::int64_t g0 = l.chunk0[b];
    // tile entry of step st for this lane (the zero guard at or before st);
    // kept in registers for both sweeps
    int taddr[64];
#pragma unroll
    for (int st = 0; st < 64; ++st) {
        taddr[st] = lane > st ? tri_base(st) + lane - st : 0;
        __asm__ volatile("" : "+v"(taddr[st]));
    }
    ell_sweep<MODE, true, X, V, E>(l.fidx, fval + static_cast<std::int64_t>(c) * l.fslots, g0, dv, y, ys, tri, taddr,
                                   nb, lane);
    __syncthreads();
    ell_sweep<MODE, false, X, V, E>(l.bidx, bval + static_cast<std::int64_t>(c) * l.bslots, g0, dv, y, ys, tri,
                                    taddr, nb, lane);
    const bool scaled = MODE == TriMode::ssor && scale != X(1);
    if (GLOBAL_Y && !scaled) return;
    __syncthreads();
    for (int i0 = 0; i0 < nb; i0 += 16 * 64) {
        const bool full = i0 + 16 * 64 <= nb;
        X v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int i = i0 + u * 64 + lane;
            v[u] = y[(i < nb ? i : nb - 1) * ys];
            if (scaled) v[u] = v[u] * scale;
        }
        if (full) {
#pragma unroll
            for (int u = 0; u < 16; ++u) out[(r0 + i0 + u * 64 + lane) * bs + c] = v[u];
        } else {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = i0 + u * 64 + lane;
                if (i < nb) out[(r0 + i) * bs + c] = v[u];
            }
        }
    }
}

// ELL values from the chain-order values: dst[c][slot] = src[c][map[slot]] (0 for padding)
template<class V>
__global__ __launch_bounds__(256) void k_ell_gather(V* __restrict__ dst, const V* __restrict__ src,
                                                    const std::int64_t* __restrict__ map, std::int64_t slots,
                                                    std::int64_t nnz, int bs)
{
    const std::int64_t total = slots * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t c = t / slots;
        const std::int64_t sl = t - c * slots;
        const std::int64_t e = map[sl];
        dst[t] = e >= 0 ? src[c * nnz + e] : V(0);
    }
}

// ---------------------------------------------------------------- symbolic --

// Encoded gather source of matrix value (row, CSR entry e): >= 0 index of
// component 0 in the value array; <= -2: diagonal array.
std::int64_t entry_source(owtk_matrix m, std::int64_t e) { return m->value_index(m->s.slot_of_entry[e], 0); }
std::int64_t diag_source(owtk_matrix m, std::int64_t row) { return -2 - m->diag_index(row, 0); }
std::int64_t comp_stride(owtk_matrix m) { return m->s.layout == Layout::sell ? 64 : 1; }

// Block-restricted local rows in the original CSR order (ghost columns and
// cross-block columns dropped), with each row's diagonal source.
struct LocalRows {
    std::vector<std::int64_t> off;
    std::vector<std::int64_t> col;
    std::vector<std::int64_t> src;
    std::vector<std::int64_t> diag_src; // -1 when the row has no diagonal
};

LocalRows local_rows(owtk_matrix m, std::int64_t block_nodes)
{
    LocalRows L;
    const std::int64_t n = m->owned;
    L.off.assign(n + 1, 0);
    L.diag_src.assign(n, -1);
    L.col.reserve(m->h_columns.size() + (m->has_diag_array() ? n : 0));
    L.src.reserve(L.col.capacity());
    for (std::int64_t row = 0; row < n; ++row) {
        const std::int64_t blk = block_nodes > 0 ? row / block_nodes : 0;
        if (m->has_diag_array()) L.diag_src[row] = diag_source(m, row);
        for (std::int32_t e = m->h_offsets[row]; e < m->h_offsets[row + 1]; ++e) {
            const std::int64_t c = m->h_columns[e];
            if (c >= n) continue;
            if (block_nodes > 0 && c / block_nodes != blk) continue;
            if (!m->has_diag_array() && c == row && L.diag_src[row] == -1) L.diag_src[row] = entry_source(m, e);
            L.col.push_back(c);
            L.src.push_back(entry_source(m, e));
        }
        if (m->has_diag_array()) {
            // the split diagonal is a separate array; keep it out of the edge list
        }
        L.off[row + 1] = static_cast<std::int64_t>(L.col.size());
    }
    return L;
}

template<class T>
void upload(DeviceArray<T>& d, const std::vector<T>& h)
{
    d.allocate(std::max<std::size_t>(h.size(), 1));
    if (!h.empty()) OWTK_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// Build the level schedule of one triangular pair.  lower(i)/upper(i) list
// the block-local entries of row i (global index space) in apply order.
struct TriHost {
    // per row: forward entries (col, src) and backward entries
    std::vector<std::int64_t> f_off, f_col, f_src, b_off, b_col, b_src;
};

std::size_t lds_limit(owtk_precond pc);

// ELL copy of a chain layout (see k_tri_ell): slot k of position p holds the
// p-th row's k-th chain-order entry, encoded by where its value comes from:
// its block-local column (entry on a chunk before the previous one),
// -(t + 2) (in-chunk entry finished at step s, t = tri_base(s) + lane - s its
// packed tile entry, 1..2016), -(l + 2019) (entry on the previous chunk,
// finished by lane l) or -1 (padding).  Needs block-local columns below 2^15
// and at most 16 entries per row.
void build_ell(TriSchedule& tri, const std::vector<std::int64_t>& fptr, const std::vector<std::int32_t>& fnout,
               const std::vector<std::int32_t>& fcol, const std::vector<std::int64_t>& bptr,
               const std::vector<std::int32_t>& bnout, const std::vector<std::int32_t>& bcol)
{
    tri.ell = false;
    if (!tri.chain || tri.max_block_rows > 32767) return;
    std::vector<std::int32_t> chunk0(tri.blocks);
    std::int64_t chunks = 0;
    for (std::int64_t b = 0; b < tri.blocks; ++b) {
        chunk0[b] = static_cast<std::int32_t>(chunks);
        chunks += (tri.h_nrows[b] + 63) / 64;
    }
    const std::int64_t n = tri.rows;
    auto kmax = [&](const std::vector<std::int64_t>& ptr) {
        std::int64_t k = 1;
        for (std::int64_t i = 0; i < n; ++i) k = std::max(k, ptr[i + 1] - ptr[i]);
        return k;
    };
    const std::int64_t kneed = std::max(kmax(fptr), kmax(bptr));
    if (kneed > 16 || chunks > 0x7fffffff) return;
    // both sweeps padded to the kernel's slot count (4, 6, 8 or 16)
    const std::int64_t fk = kneed <= 4 ? 4 : kneed <= 6 ? 6 : kneed <= 8 ? 8 : 16, bk = fk;
    for (int dir = 0; dir < 2; ++dir) {
        const auto& ptr = dir == 0 ? fptr : bptr;
        const auto& nout = dir == 0 ? fnout : bnout;
        const auto& col = dir == 0 ? fcol : bcol;
        const std::int64_t K = dir == 0 ? fk : bk;
        std::vector<std::int16_t> idx(static_cast<std::size_t>(chunks * K * 64), -1);
        std::vector<std::int64_t> map(idx.size(), -1);
        for (std::int64_t b = 0; b < tri.blocks; ++b) {
            const std::int64_t r0 = tri.h_row0[b];
            const std::int64_t nb = tri.h_nrows[b];
            for (std::int64_t p = 0; p < nb; ++p) {
                const std::int64_t q = p / 64, lane = p % 64, cs = q * 64;
                const std::int64_t g = chunk0[b] + q;
                const std::int64_t e0 = ptr[r0 + p], e1 = ptr[r0 + p + 1], no = nout[r0 + p];
                for (std::int64_t e = e0; e < e1; ++e) {
                    const std::int64_t c = col[e];
                    const std::int64_t pc = dir == 0 ? c : nb - 1 - c;  // position of the column in the sweep
                    std::int64_t code = c;
                    if (e - e0 >= no) {
                        const std::int64_t step = pc - cs;
                        if (step < 0 || step >= lane) throw std::logic_error("chain layout: in-chunk entry order");
                        code = -(tri_base(static_cast<int>(step)) + lane - step + 2);
                    } else if (pc >= cs) {
                        throw std::logic_error("chain layout: out-of-chunk entry inside the chunk");
                    } else if (pc >= cs - 64) {
                        code = -(pc - (cs - 64) + 2019);
                    }
                    const std::size_t slot = static_cast<std::size_t>((g * K + (e - e0)) * 64 + lane);
                    idx[slot] = static_cast<std::int16_t>(code);
                    map[slot] = e;
                }
            }
        }
        upload(dir == 0 ? tri.fidx_ell : tri.bidx_ell, idx);
        upload(dir == 0 ? tri.fmap_ell : tri.bmap_ell, map);
        (dir == 0 ? tri.fslots : tri.bslots) = static_cast<std::int64_t>(idx.size());
    }
    tri.fk = static_cast<int>(fk);
    tri.bk = static_cast<int>(bk);
    upload(tri.chunk0, chunk0);
    tri.ell = true;
}

void build_schedule(owtk_precond pc, const TriHost& t)
{
    auto& tri = pc->tri;
    const std::int64_t n = pc->m->owned;
    const std::int64_t bn = pc->block_nodes > 0 ? pc->block_nodes : n;
    tri.rows = n;
    tri.bs = pc->m->bs;
    tri.blocks = (n + bn - 1) / bn;
    tri.h_row0.assign(tri.blocks, 0);
    tri.h_nrows.assign(tri.blocks, 0);
    tri.h_flev.assign(tri.blocks + 1, 0);
    tri.h_blev.assign(tri.blocks + 1, 0);
    tri.h_flev_pos.clear();
    tri.h_blev_pos.clear();
    tri.max_block_rows = tri.max_flevels = tri.max_blevels = tri.max_level_width = 0;
    // pass 1: levels (forward from the top, backward from the bottom), depths, widths
    std::vector<std::int32_t> flevel(n, 0), blevel(n, 0);
    std::vector<int> fdepth(tri.blocks, 0), bdepth(tri.blocks, 0);
    int emax = 0;
    for (std::int64_t i = 0; i < n; ++i)
        emax = std::max<int>(emax, static_cast<int>(std::max(t.f_off[i + 1] - t.f_off[i], t.b_off[i + 1] - t.b_off[i])));
    std::int64_t depth_sum = 0;
    for (std::int64_t b = 0; b < tri.blocks; ++b) {
        const std::int64_t r0 = b * bn;
        const std::int64_t nb = std::min(bn, n - r0);
        tri.h_row0[b] = static_cast<std::int32_t>(r0);
        tri.h_nrows[b] = static_cast<std::int32_t>(nb);
        tri.max_block_rows = std::max(tri.max_block_rows, nb);
        for (std::int64_t i = r0; i < r0 + nb; ++i) {
            int lv = 0;
            for (std::int64_t e = t.f_off[i]; e < t.f_off[i + 1]; ++e) lv = std::max(lv, flevel[t.f_col[e]] + 1);
            flevel[i] = lv;
            fdepth[b] = std::max(fdepth[b], lv + 1);
        }
        for (std::int64_t i = r0 + nb; i-- > r0;) {
            int lv = 0;
            for (std::int64_t e = t.b_off[i]; e < t.b_off[i + 1]; ++e) lv = std::max(lv, blevel[t.b_col[e]] + 1);
            blevel[i] = lv;
            bdepth[b] = std::max(bdepth[b], lv + 1);
        }
        tri.max_flevels = std::max<std::int64_t>(tri.max_flevels, fdepth[b]);
        tri.max_blevels = std::max<std::int64_t>(tri.max_blevels, bdepth[b]);
        depth_sum += std::max(fdepth[b], bdepth[b]);
    }
    // Chain sweeps win when levels are narrow (average width below 16 rows):
    // a level then costs a barrier for a handful of rows.
    tri.emax = emax;
    const bool chain_ok = emax <= 16
        && (static_cast<std::size_t>(tri.max_block_rows) + 64 * 64) * sizeof(double) <= lds_limit(pc);
    if (pc->schedule == 2 && !chain_ok)
        throw unsupported_error("chain sweeps need at most 16 entries per row and a block that fits in LDS");
    tri.chain = pc->schedule == 2 || (pc->schedule == 0 && chain_ok && depth_sum * 16 > n);
    // pass 2: level lists and apply arrays
    std::vector<std::int32_t> frow(n), brow(n), frow_lvl(n), fnout(n, 0), bnout(n, 0);
    tri.ei_max = 0;
    std::vector<std::int64_t> fptr(n + 1, 0), bptr(n + 1, 0);
    std::vector<std::int32_t> fcol, bcol;
    std::vector<std::int64_t> fsrc, bsrc;
    fcol.reserve(t.f_col.size());
    bcol.reserve(t.b_col.size());
    for (std::int64_t b = 0; b < tri.blocks; ++b) {
        const std::int64_t r0 = tri.h_row0[b];
        const std::int64_t nb = tri.h_nrows[b];
        for (int dir = 0; dir < 2; ++dir) {
            const auto& off = dir == 0 ? t.f_off : t.b_off;
            const auto& col = dir == 0 ? t.f_col : t.b_col;
            const auto& src = dir == 0 ? t.f_src : t.b_src;
            const auto& level = dir == 0 ? flevel : blevel;
            const int depth = dir == 0 ? fdepth[b] : bdepth[b];
            // counting sort by level, stable in row order (forward) / reverse row order (backward)
            std::vector<std::int64_t> count(depth + 1, 0);
            for (std::int64_t i = r0; i < r0 + nb; ++i) ++count[level[i] + 1];
            for (int l = 0; l < depth; ++l) count[l + 1] += count[l];
            auto& lev_pos = dir == 0 ? tri.h_flev_pos : tri.h_blev_pos;
            auto& lev = dir == 0 ? tri.h_flev : tri.h_blev;
            lev[b] = static_cast<std::int32_t>(lev_pos.size());
            for (int l = 0; l <= depth; ++l) lev_pos.push_back(static_cast<std::int32_t>(count[l]));
            for (int l = 0; l < depth; ++l) tri.max_level_width = std::max(tri.max_level_width, count[l + 1] - count[l]);
            lev[b + 1] = static_cast<std::int32_t>(lev_pos.size());
            std::vector<std::int64_t> next(count.begin(), count.end() - 1);
            std::vector<std::int32_t> lvl_map(nb);
            auto place = [&](std::int64_t i) { lvl_map[next[level[i]]++] = static_cast<std::int32_t>(i - r0); };
            if (dir == 0)
                for (std::int64_t i = r0; i < r0 + nb; ++i) place(i);
            else
                for (std::int64_t i = r0 + nb; i-- > r0;) place(i);
            if (dir == 0) std::copy(lvl_map.begin(), lvl_map.end(), frow_lvl.begin() + r0);
            auto& rowmap = dir == 0 ? frow : brow;
            for (std::int64_t p = 0; p < nb; ++p) {
                if (tri.chain) rowmap[r0 + p] = static_cast<std::int32_t>(dir == 0 ? p : nb - 1 - p);
                else rowmap[r0 + p] = lvl_map[p];
            }
            auto& ptr = dir == 0 ? fptr : bptr;
            auto& ocol = dir == 0 ? fcol : bcol;
            auto& osrc = dir == 0 ? fsrc : bsrc;
            auto& onout = dir == 0 ? fnout : bnout;
            std::vector<std::pair<int, std::int64_t>> inside;
            for (std::int64_t p = 0; p < nb; ++p) {
                const std::int64_t i = r0 + rowmap[r0 + p];
                if (!tri.chain) {
                    for (std::int64_t e = off[i]; e < off[i + 1]; ++e) {
                        ocol.push_back(static_cast<std::int32_t>(col[e] - r0));
                        osrc.push_back(src[e]);
                    }
                } else {
                    // chain layout: entries from rows before this chunk, then the
                    // in-chunk entries by the step that finishes their row
                    const std::int64_t cs = (p / 64) * 64;
                    inside.clear();
                    std::int32_t outside = 0;
                    for (std::int64_t e = off[i]; e < off[i + 1]; ++e) {
                        const std::int64_t c = col[e] - r0;
                        const bool in_chunk = dir == 0 ? c >= cs : c <= nb - 1 - cs;
                        if (in_chunk) {
                            inside.push_back({static_cast<int>(dir == 0 ? c - cs : nb - 1 - cs - c), e});
                        } else {
                            ocol.push_back(static_cast<std::int32_t>(c));
                            osrc.push_back(src[e]);
                            ++outside;
                        }
                    }
                    std::sort(inside.begin(), inside.end());
                    for (const auto& [step, e] : inside) {
                        ocol.push_back(static_cast<std::int32_t>(col[e] - r0));
                        osrc.push_back(src[e]);
                    }
                    onout[r0 + p] = outside;
                    tri.ei_max = std::max<int>(tri.ei_max, static_cast<int>(inside.size()));
                }
                ptr[r0 + p + 1] = static_cast<std::int64_t>(ocol.size());
            }
        }
    }
    tri.fnnz = static_cast<std::int64_t>(fcol.size());
    tri.bnnz = static_cast<std::int64_t>(bcol.size());
    build_ell(tri, fptr, fnout, fcol, bptr, bnout, bcol);
    upload(tri.row0, tri.h_row0);
    upload(tri.nrows, tri.h_nrows);
    upload(tri.flev, tri.h_flev);
    upload(tri.blev, tri.h_blev);
    upload(tri.flev_pos, tri.h_flev_pos);
    upload(tri.blev_pos, tri.h_blev_pos);
    upload(tri.frow, frow);
    upload(tri.brow, brow);
    upload(tri.frow_lvl, frow_lvl);
    upload(tri.fnout, fnout);
    upload(tri.bnout, bnout);
    upload(tri.fptr, fptr);
    upload(tri.bptr, bptr);
    upload(tri.fcol, fcol);
    upload(tri.bcol, bcol);
    upload(tri.fsrc, fsrc);
    upload(tri.bsrc, bsrc);
}

TriArgs tri_args(const TriSchedule& t)
{
    return TriArgs{t.row0.get(), t.nrows.get(), t.flev.get(), t.blev.get(), t.flev_pos.get(), t.blev_pos.get(),
                   t.frow.get(), t.brow.get(), t.fptr.get(), t.bptr.get(), t.fcol.get(), t.bcol.get(),
                   t.fnout.get(), t.bnout.get(), t.fnnz, t.bnnz, t.rows, t.bs};
}

std::size_t lds_limit(owtk_precond pc)
{
    // device LDS limit per workgroup (cached per device)
    static thread_local int cached_device = -1;
    static thread_local std::size_t cached = 0;
    if (cached_device != pc->m->ctx->device) {
        hipDeviceProp_t prop;
        OWTK_HIP(hipGetDeviceProperties(&prop, pc->m->ctx->device));
        cached = prop.sharedMemPerBlock;
        cached_device = pc->m->ctx->device;
    }
    return cached;
}

void read_flag(owtk_precond pc, int* value)
{
    OWTK_HIP(hipMemcpyAsync(value, pc->flag.get(), sizeof(int), hipMemcpyDeviceToHost, pc->m->ctx->stream));
    pc->m->ctx->sync();
}

void clear_flag(owtk_precond pc) { kern::fill_zero_bytes(pc->m->ctx->stream, pc->flag.get(), sizeof(int) * 4); }

unsigned int grid_for(std::int64_t n) { return dev::stream_grid(n, 4); }

// Value arrays in the preconditioner's storage precision.
void refresh_storage(owtk_precond pc)
{
    auto& t = pc->tri;
    const hipStream_t st = pc->m->ctx->stream;
    const int bs = pc->m->bs;
    if (t.ell) {
        // ELL values, FP64 from the chain-order factor, then each storage's copy
        if (!t.fell64.get()) t.fell64.allocate(std::max<std::int64_t>(t.fslots * bs, 1));
        if (!t.bell64.get()) t.bell64.allocate(std::max<std::int64_t>(t.bslots * bs, 1));
        k_ell_gather<double><<<grid_for(t.fslots * bs), kBlock, 0, st>>>(t.fell64.get(), t.fval64.get(),
                                                                         t.fmap_ell.get(), t.fslots, t.fnnz, bs);
        k_ell_gather<double><<<grid_for(t.bslots * bs), kBlock, 0, st>>>(t.bell64.get(), t.bval64.get(),
                                                                         t.bmap_ell.get(), t.bslots, t.bnnz, bs);
        OWTK_LAUNCH_CHECK();
        if (pc->storage == OWTK_F32) {
            if (!t.fell32.get()) t.fell32.allocate(std::max<std::int64_t>(t.fslots * bs, 1));
            if (!t.bell32.get()) t.bell32.allocate(std::max<std::int64_t>(t.bslots * bs, 1));
            kern::convert(st, t.fell32.get(), t.fell64.get(), static_cast<std::size_t>(t.fslots * bs));
            kern::convert(st, t.bell32.get(), t.bell64.get(), static_cast<std::size_t>(t.bslots * bs));
        } else if (pc->storage == OWTK_BF16) {
            if (!t.fell16.get()) t.fell16.allocate(std::max<std::int64_t>(t.fslots * bs, 1));
            if (!t.bell16.get()) t.bell16.allocate(std::max<std::int64_t>(t.bslots * bs, 1));
            kern::k_to_bf16<<<grid_for(t.fslots * bs), kBlock, 0, st>>>(t.fell16.get(), t.fell64.get(), t.fslots * bs);
            kern::k_to_bf16<<<grid_for(t.bslots * bs), kBlock, 0, st>>>(t.bell16.get(), t.bell64.get(), t.bslots * bs);
            OWTK_LAUNCH_CHECK();
        }
    }
    if (pc->storage == OWTK_F32 || pc->storage == OWTK_BF16) {
        const std::int64_t nd = t.rows * bs;
        if (!t.dval32.get()) t.dval32.allocate(std::max<std::int64_t>(nd, 1));
        kern::convert(st, t.dval32.get(), t.dval64.get(), static_cast<std::size_t>(nd));
    }
    if (pc->storage == OWTK_F32) {
        if (!t.fval32.get()) t.fval32.allocate(std::max<std::int64_t>(t.fnnz * bs, 1));
        if (!t.bval32.get()) t.bval32.allocate(std::max<std::int64_t>(t.bnnz * bs, 1));
        if (t.fnnz) kern::convert(st, t.fval32.get(), t.fval64.get(), static_cast<std::size_t>(t.fnnz * bs));
        if (t.bnnz) kern::convert(st, t.bval32.get(), t.bval64.get(), static_cast<std::size_t>(t.bnnz * bs));
    } else if (pc->storage == OWTK_BF16) {
        if (!t.fval16.get()) t.fval16.allocate(std::max<std::int64_t>(t.fnnz * bs, 1));
        if (!t.bval16.get()) t.bval16.allocate(std::max<std::int64_t>(t.bnnz * bs, 1));
        if (t.fnnz)
            kern::k_to_bf16<<<grid_for(t.fnnz * bs), kBlock, 0, st>>>(t.fval16.get(), t.fval64.get(), t.fnnz * bs);
        if (t.bnnz)
            kern::k_to_bf16<<<grid_for(t.bnnz * bs), kBlock, 0, st>>>(t.bval16.get(), t.bval64.get(), t.bnnz * bs);
        OWTK_LAUNCH_CHECK();
    }
}

// ------------------------------------------------------------------- Jacobi --
void jacobi_setup(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    if (m->dense()) {
        if (m->bs > 8 || m->s.layout != Layout::sell)
            throw unsupported_error("dense block-Jacobi supports block sizes up to 4");
        std::vector<std::int64_t> slot(m->owned);
        for (std::int64_t row = 0; row < m->owned; ++row) {
            std::int64_t s = -1;
            for (std::int32_t e = m->h_offsets[row]; e < m->h_offsets[row + 1]; ++e)
                if (m->h_columns[e] == row) {
                    s = m->s.slot_of_entry[e];
                    break;
                }
            if (s < 0) throw std::invalid_argument("dense-block CSR row has no diagonal block");
            slot[row] = s;
        }
        upload(pc->tri.dsrc, slot);
        pc->inv64.allocate(static_cast<std::size_t>(m->owned) * m->bs * m->bs);
        return;
    }
    std::vector<std::int64_t> src(m->owned);
    for (std::int64_t row = 0; row < m->owned; ++row) {
        if (m->has_diag_array()) {
            src[row] = diag_source(m, row);
            continue;
        }
        std::int64_t s = std::numeric_limits<std::int64_t>::min();
        for (std::int32_t e = m->h_offsets[row]; e < m->h_offsets[row + 1]; ++e)
            if (m->h_columns[e] == row) {
                s = entry_source(m, e);
                break;
            }
        if (s == std::numeric_limits<std::int64_t>::min())
            throw std::invalid_argument("BlockCsrMatrix row has no diagonal");
        src[row] = s;
    }
    upload(pc->tri.dsrc, src);
    pc->inv64.allocate(static_cast<std::size_t>(m->owned) * m->bs);
}

void jacobi_numeric(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    const hipStream_t st = m->ctx->stream;
    clear_flag(pc);
    if (m->dense()) {
        const unsigned int g = static_cast<unsigned int>((m->owned + 255) / 256);
        switch (m->bs) {
        case 1: k_block_inverse<1><<<g, kBlock, 0, st>>>(pc->inv64.get(), pc->tri.dsrc.get(), m->owned, m->v64.get(), pc->flag.get()); break;
        case 2: k_block_inverse<2><<<g, kBlock, 0, st>>>(pc->inv64.get(), pc->tri.dsrc.get(), m->owned, m->v64.get(), pc->flag.get()); break;
        case 3: k_block_inverse<3><<<g, kBlock, 0, st>>>(pc->inv64.get(), pc->tri.dsrc.get(), m->owned, m->v64.get(), pc->flag.get()); break;
        default: k_block_inverse<4><<<g, kBlock, 0, st>>>(pc->inv64.get(), pc->tri.dsrc.get(), m->owned, m->v64.get(), pc->flag.get()); break;
        }
        OWTK_LAUNCH_CHECK();
    } else {
        const double minimum = pc->minimum_diagonal > 0 ? pc->minimum_diagonal : kTol;
        k_jacobi_inverse<<<grid_for(m->owned * m->bs), kBlock, 0, st>>>(pc->inv64.get(), pc->tri.dsrc.get(),
                                                                       m->owned, m->bs, comp_stride(m),
                                                                       m->v64.get(), m->d64.get(), minimum,
                                                                       pc->flag.get());
        OWTK_LAUNCH_CHECK();
    }
    int flag = 0;
    read_flag(pc, &flag);
    if (flag != flag_ok) {
        if (m->dense()) throw std::invalid_argument("block-Jacobi preconditioner has a singular diagonal block");
        throw std::invalid_argument("Jacobi preconditioner has a singular diagonal");
    }
    if (pc->storage != OWTK_F64) {
        if (!pc->inv32.get()) pc->inv32.allocate(pc->inv64.count);
        kern::convert(st, pc->inv32.get(), pc->inv64.get(), pc->inv64.count);
    }
}

// --------------------------------------------------------------------- SSOR --
// Greedy colouring of the symmetrised owned-row graph in row order: each row
// takes the smallest colour none of its neighbours holds.
std::vector<std::int32_t> greedy_colours(owtk_matrix m)
{
    const std::int64_t n = m->owned;
    std::vector<std::vector<std::int32_t>> adj(n);
    for (std::int64_t i = 0; i < n; ++i)
        for (std::int32_t e = m->h_offsets[i]; e < m->h_offsets[i + 1]; ++e) {
            const std::int64_t j = m->h_columns[e];
            if (j == i || j >= n) continue;
            adj[i].push_back(static_cast<std::int32_t>(j));
            adj[j].push_back(static_cast<std::int32_t>(i));
        }
    std::vector<std::int32_t> colour(n, -1);
    std::vector<std::int64_t> seen;
    for (std::int64_t i = 0; i < n; ++i) {
        seen.assign(adj[i].size() + 2, -1);
        for (std::int32_t j : adj[i])
            if (colour[j] >= 0 && colour[j] < static_cast<std::int32_t>(seen.size())) seen[colour[j]] = i;
        std::int32_t c = 0;
        while (seen[c] == i) ++c;
        colour[i] = c;
    }
    return colour;
}

void dense_ssor_setup(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    if (m->bs > 4 || m->s.layout != Layout::sell) throw unsupported_error("dense-block SSOR supports block sizes up to 4");
    jacobi_setup(pc);  // diagonal block slots + inverse storage
    const std::vector<std::int32_t> colour = greedy_colours(m);
    const std::int32_t colours = colour.empty() ? 0 : *std::max_element(colour.begin(), colour.end()) + 1;
    pc->h_mc_ptr.assign(colours + 1, 0);
    for (std::int32_t c : colour) ++pc->h_mc_ptr[c + 1];
    for (std::int32_t c = 0; c < colours; ++c) pc->h_mc_ptr[c + 1] += pc->h_mc_ptr[c];
    std::vector<std::int32_t> rows(colour.size());
    std::vector<std::int64_t> fill(pc->h_mc_ptr.begin(), pc->h_mc_ptr.end() - 1);
    for (std::int64_t i = 0; i < static_cast<std::int64_t>(colour.size()); ++i) rows[fill[colour[i]]++] = static_cast<std::int32_t>(i);
    // Colour-ordered copies: colour c's rows (ascending) fill copy rows
    // h_mc_seg[c] .. h_mc_seg[c] + count - 1, padded to whole slices.  The
    // forward copy keeps each row's blocks on earlier colours, the backward
    // copy those on later colours, both in the row's slot order (the FMA
    // order of the sweeps); the diagonal copy has one block per copy row.
    pc->h_mc_seg.assign(colours + 1, 0);
    for (std::int32_t c = 0; c < colours; ++c) {
        const std::int64_t cnt = pc->h_mc_ptr[c + 1] - pc->h_mc_ptr[c];
        pc->h_mc_seg[c + 1] = pc->h_mc_seg[c] + (cnt + 63) / 64 * 64;
    }
    const std::int64_t padded = pc->h_mc_seg[colours];
    std::vector<std::int32_t> prow(static_cast<std::size_t>(std::max<std::int64_t>(padded, 1)), -1);
    for (std::int32_t c = 0; c < colours; ++c)
        for (std::int64_t t = 0; t < pc->h_mc_ptr[c + 1] - pc->h_mc_ptr[c]; ++t)
            prow[pc->h_mc_seg[c] + t] = rows[pc->h_mc_ptr[c] + t];
    upload(pc->mc_prow, prow);
    std::vector<std::int64_t> dslot(m->owned);
    OWTK_HIP(hipMemcpy(dslot.data(), pc->tri.dsrc.get(), dslot.size() * sizeof(std::int64_t), hipMemcpyDeviceToHost));
    const int bb = m->bs * m->bs;
    const std::size_t sv = pc->storage == OWTK_F64 ? 8 : (pc->storage == OWTK_F32 ? 4 : 2);
    auto build = [&](McSell& out, int dir) {  // dir -1 forward, +1 backward, 0 diagonal
        std::vector<std::int32_t> len(static_cast<std::size_t>(std::max<std::int64_t>(padded, 1)), 0);
        std::vector<std::vector<std::pair<std::int32_t, std::int64_t>>> ent(static_cast<std::size_t>(padded));
        for (std::int64_t q = 0; q < padded; ++q) {
            const std::int32_t i = prow[q];
            if (i < 0) continue;
            if (dir == 0) {
                ent[q].push_back({i, dslot[i]});
            } else {
                for (std::int32_t e = m->h_offsets[i]; e < m->h_offsets[i + 1]; ++e) {
                    const std::int32_t j = m->h_columns[e];
                    if (j >= m->owned || j == i) continue;
                    if ((dir < 0) == (colour[j] < colour[i])) ent[q].push_back({j, m->s.slot_of_entry[e]});
                }
            }
            len[q] = static_cast<std::int32_t>(ent[q].size());
        }
        const std::int64_t slices = padded / 64;
        std::vector<std::int32_t> sp(static_cast<std::size_t>(slices + 1), 0);
        for (std::int64_t sl = 0; sl < slices; ++sl) {
            int wd = 0;
            for (int l = 0; l < 64; ++l) wd = std::max(wd, len[sl * 64 + l]);
            sp[sl + 1] = sp[sl] + 64 * wd;
        }
        const std::int64_t slots = sp[slices];
        std::vector<std::int32_t> cols(static_cast<std::size_t>(std::max<std::int64_t>(slots, 1)), 0);
        std::vector<std::int64_t> map(static_cast<std::size_t>(std::max<std::int64_t>(slots, 1)), -1);
        for (std::int64_t q = 0; q < padded; ++q)
            for (std::size_t k = 0; k < ent[q].size(); ++k) {
                const std::int64_t slot = sp[q / 64] + 64 * static_cast<std::int64_t>(k) + (q % 64);
                cols[slot] = ent[q][k].first;
                map[slot] = ent[q][k].second;
            }
        if (dir == 0) {  // one block per copy row: slot q
            map.assign(static_cast<std::size_t>(std::max<std::int64_t>(padded, 1)), -1);
            for (std::int64_t q = 0; q < padded; ++q)
                if (prow[q] >= 0) map[q] = dslot[prow[q]];
            out.slots = padded;
        } else {
            upload(out.slice_ptr, sp);
            upload(out.row_len, len);
            upload(out.cols, cols);
            out.slots = slots;
        }
        upload(out.map, map);
        out.vals.allocate(static_cast<std::size_t>(std::max<std::int64_t>(out.slots, 1)) * bb * sv);
    };
    build(pc->mc_fwd, -1);
    build(pc->mc_bwd, 1);
    build(pc->mc_diag, 0);
    const std::size_t nb = static_cast<std::size_t>(m->owned * m->bs);
    if (pc->storage == OWTK_F64) pc->work64.allocate(std::max<std::size_t>(nb, 1));
    else pc->work32.allocate(std::max<std::size_t>(nb, 1));
}

// The copies' values from the matrix, in the preconditioner's storage type.
void dense_ssor_gather(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    const hipStream_t st = m->ctx->stream;
    const int bb = m->bs * m->bs;
    for (McSell* c : {&pc->mc_fwd, &pc->mc_bwd, &pc->mc_diag}) {
        if (c->slots == 0) continue;
        const unsigned int g = grid_for(c->slots * bb);
        switch (pc->storage) {
        case OWTK_F64:
            k_mc_gather<double><<<g, kBlock, 0, st>>>(reinterpret_cast<double*>(c->vals.get()), m->v64.get(),
                                                      c->map.get(), c->slots, bb);
            break;
        case OWTK_F32:
            k_mc_gather<float><<<g, kBlock, 0, st>>>(reinterpret_cast<float*>(c->vals.get()), m->v32.get(),
                                                     c->map.get(), c->slots, bb);
            break;
        default:
            k_mc_gather<std::uint16_t><<<g, kBlock, 0, st>>>(reinterpret_cast<std::uint16_t*>(c->vals.get()),
                                                             m->v16.get(), c->map.get(), c->slots, bb);
            break;
        }
        OWTK_LAUNCH_CHECK();
    }
}

template<class V, class X, class I>
void dense_ssor_apply(owtk_precond pc, const X* in, X* out, const I* inv, X* y, const int* stop)
{
    owtk_matrix m = pc->m;
    const hipStream_t st = m->ctx->stream;
    const double scale = pc->omega * (2.0 - pc->omega);
    const int colours = static_cast<int>(pc->h_mc_ptr.size()) - 1;
    const V* dvals = reinterpret_cast<const V*>(pc->mc_diag.vals.get());
    auto launch = [&](int c, auto fwd) {
        constexpr bool FWD = decltype(fwd)::value;
        const std::int64_t count = pc->h_mc_ptr[c + 1] - pc->h_mc_ptr[c];
        if (count == 0) return;
        const unsigned int g = static_cast<unsigned int>((count + kBlock - 1) / kBlock);
        const McSell& cp = FWD ? pc->mc_fwd : pc->mc_bwd;
        const V* vals = reinterpret_cast<const V*>(cp.vals.get());
        const std::int64_t seg0 = pc->h_mc_seg[c];
        const std::int32_t* prow = pc->mc_prow.get();
        switch (m->bs) {
        case 1: k_mc_ssor<1, V, X, I, FWD><<<g, kBlock, 0, st>>>(prow, seg0, count, cp.slice_ptr.get(), cp.row_len.get(), cp.cols.get(), vals, dvals, pc->omega, scale, in, y, out, inv, stop); break;
        case 2: k_mc_ssor<2, V, X, I, FWD><<<g, kBlock, 0, st>>>(prow, seg0, count, cp.slice_ptr.get(), cp.row_len.get(), cp.cols.get(), vals, dvals, pc->omega, scale, in, y, out, inv, stop); break;
        case 3: k_mc_ssor<3, V, X, I, FWD><<<g, kBlock, 0, st>>>(prow, seg0, count, cp.slice_ptr.get(), cp.row_len.get(), cp.cols.get(), vals, dvals, pc->omega, scale, in, y, out, inv, stop); break;
        default: k_mc_ssor<4, V, X, I, FWD><<<g, kBlock, 0, st>>>(prow, seg0, count, cp.slice_ptr.get(), cp.row_len.get(), cp.cols.get(), vals, dvals, pc->omega, scale, in, y, out, inv, stop); break;
        }
        OWTK_LAUNCH_CHECK();
    };
    const int category = sizeof(X) == 8 ? prof_precond_f64 : prof_precond_low;
    ProfileScope prof(m->ctx, category, precond_apply_bytes(pc));
    for (int c = 0; c < colours; ++c) launch(c, std::true_type{});
    for (int c = colours - 1; c >= 0; --c) launch(c, std::false_type{});
}

void ssor_setup(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    if (!(pc->omega > 0.0 && pc->omega < 2.0)) throw std::invalid_argument("SSOR omega must be in (0,2)");
    if (m->dense()) {
        dense_ssor_setup(pc);
        return;
    }
    const LocalRows L = local_rows(m, pc->block_nodes);
    TriHost t;
    const std::int64_t n = m->owned;
    t.f_off.assign(n + 1, 0);
    t.b_off.assign(n + 1, 0);
    std::vector<std::int64_t> dsrc(n);
    for (std::int64_t row = 0; row < n; ++row) {
        if (L.diag_src[row] == -1) throw std::invalid_argument("SSOR preconditioner row has no diagonal");
        dsrc[row] = L.diag_src[row];
        for (std::int64_t e = L.off[row]; e < L.off[row + 1]; ++e) {
            if (L.col[e] < row) {
                t.f_col.push_back(L.col[e]);
                t.f_src.push_back(L.src[e]);
            } else if (L.col[e] > row) {
                t.b_col.push_back(L.col[e]);
                t.b_src.push_back(L.src[e]);
            }
        }
        t.f_off[row + 1] = static_cast<std::int64_t>(t.f_col.size());
        t.b_off[row + 1] = static_cast<std::int64_t>(t.b_col.size());
    }
    build_schedule(pc, t);
    upload(pc->tri.dsrc, dsrc);
    const int bs = m->bs;
    pc->tri.fval64.allocate(std::max<std::int64_t>(pc->tri.fnnz * bs, 1));
    pc->tri.bval64.allocate(std::max<std::int64_t>(pc->tri.bnnz * bs, 1));
    pc->tri.dval64.allocate(std::max<std::int64_t>(n * bs, 1));
}

void ssor_numeric(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    if (m->dense()) {
        try {
            jacobi_numeric(pc);  // diagonal-block inverses (FP64, and FP32 copy for low storage)
        } catch (const std::invalid_argument&) {
            throw std::invalid_argument("SSOR preconditioner has a singular diagonal block");
        }
        dense_ssor_gather(pc);
        return;
    }
    auto& t = pc->tri;
    const hipStream_t st = m->ctx->stream;
    const int bs = m->bs;
    const std::int64_t stride = comp_stride(m);
    clear_flag(pc);
    if (t.fnnz)
        k_gather<<<grid_for(t.fnnz * bs), kBlock, 0, st>>>(t.fval64.get(), t.fsrc.get(), t.fnnz, bs, stride,
                                                           m->v64.get(), m->d64.get(), pc->omega);
    if (t.bnnz)
        k_gather<<<grid_for(t.bnnz * bs), kBlock, 0, st>>>(t.bval64.get(), t.bsrc.get(), t.bnnz, bs, stride,
                                                           m->v64.get(), m->d64.get(), pc->omega);
    k_gather<<<grid_for(t.rows * bs), kBlock, 0, st>>>(t.dval64.get(), t.dsrc.get(), t.rows, bs, stride,
                                                       m->v64.get(), m->d64.get(), 1.0);
    k_check_diag<<<grid_for(t.rows * bs), kBlock, 0, st>>>(t.dval64.get(), t.rows * bs, kTol, pc->flag.get());
    OWTK_LAUNCH_CHECK();
    int flag = 0;
    read_flag(pc, &flag);
    if (flag != flag_ok) throw std::invalid_argument("SSOR preconditioner has a singular diagonal");
    refresh_storage(pc);
}

// ---------------------------------------------------------------------- ILU --
// Program of k_ilu_factor_chunks (IluChunkProgram) from the factor pattern
// and the update lists.  Left disabled (the level-scheduled k_ilu_factor
// runs) when a row has more than 64 entries or a count does not fit the
// program's byte fields.
void build_ilu_chunks(owtk_precond pc, const std::vector<std::int64_t>& upd_ptr,
                      const std::vector<std::int64_t>& upd_target, const std::vector<std::int64_t>& upd_upper)
{
    auto& f = pc->ilu;
    auto& P = f.chunks;
    const auto& tri = pc->tri;
    P.enabled = false;
    if (f.nnz > 0x7fffffffLL) return;
    int kmax = 1;
    for (std::int64_t row = 0; row < tri.rows; ++row)
        kmax = std::max<int>(kmax, static_cast<int>(f.h_offsets[row + 1] - f.h_offsets[row]));
    if (kmax > 64) return;
    std::vector<std::int32_t> chunk_first(tri.blocks + 1, 0);
    std::vector<std::int64_t> hdr;
    std::vector<std::int32_t> src_a;
    std::vector<std::uint32_t> op_a, op_b;
    std::vector<std::uint16_t> upd_a, upd_b;
    int max_src = 0, max_a = 0, max_nua = 0, max_b = 0, max_nub = 0;
    struct OpA { int es, pt; std::vector<std::pair<int, int>> upd; };
    struct OpB { int step, es, ps; std::vector<std::pair<int, int>> upd; };
    std::int64_t chunks = 0;
    for (std::int64_t b = 0; b < tri.blocks; ++b) {
        chunk_first[b] = static_cast<std::int32_t>(chunks);
        const std::int64_t r0 = tri.h_row0[b], nb = tri.h_nrows[b];
        for (std::int64_t c0 = 0; c0 < nb; c0 += 64, ++chunks) {
            const int cnt = static_cast<int>(std::min<std::int64_t>(64, nb - c0));
            const std::int64_t first = r0 + c0;
            std::vector<std::vector<std::int32_t>> srcs(64);
            std::vector<std::vector<OpA>> as(64);
            std::vector<std::vector<OpB>> bs_(64);
            for (int lane = 0; lane < cnt; ++lane) {
                const std::int64_t row = first + lane;
                const std::int64_t rb = f.h_offsets[row];
                for (std::int64_t e = rb; e < f.h_diag[row]; ++e) {
                    const std::int64_t k = f.h_columns[e];
                    if (k < first) {
                        OpA op{static_cast<int>(e - rb), static_cast<int>(srcs[lane].size()), {}};
                        srcs[lane].push_back(static_cast<std::int32_t>(f.h_diag[k]));
                        for (std::int64_t u = upd_ptr[e]; u < upd_ptr[e + 1]; ++u) {
                            op.upd.push_back({static_cast<int>(upd_target[u] - rb), static_cast<int>(srcs[lane].size())});
                            srcs[lane].push_back(static_cast<std::int32_t>(upd_upper[u]));
                        }
                        as[lane].push_back(std::move(op));
                    } else {
                        const std::int64_t kb = f.h_offsets[k];
                        OpB op{static_cast<int>(k - first), static_cast<int>(e - rb), static_cast<int>(f.h_diag[k] - kb), {}};
                        for (std::int64_t u = upd_ptr[e]; u < upd_ptr[e + 1]; ++u)
                            op.upd.push_back({static_cast<int>(upd_target[u] - rb), static_cast<int>(upd_upper[u] - kb)});
                        bs_[lane].push_back(std::move(op));
                    }
                }
            }
            int nsrc = 0, na = 0, nua = 0, nbo = 0, nub = 0;
            for (int lane = 0; lane < cnt; ++lane) {
                nsrc = std::max<int>(nsrc, static_cast<int>(srcs[lane].size()));
                na = std::max<int>(na, static_cast<int>(as[lane].size()));
                nbo = std::max<int>(nbo, static_cast<int>(bs_[lane].size()));
                for (const auto& o : as[lane]) nua = std::max<int>(nua, static_cast<int>(o.upd.size()));
                for (const auto& o : bs_[lane]) nub = std::max<int>(nub, static_cast<int>(o.upd.size()));
            }
            if (nsrc > 255 || nua > 255 || nub > 255) return;
            hdr.insert(hdr.end(), {first, cnt, static_cast<std::int64_t>(src_a.size()), nsrc,
                                   static_cast<std::int64_t>(op_a.size()), na, static_cast<std::int64_t>(upd_a.size()),
                                   nua, static_cast<std::int64_t>(op_b.size()), nbo,
                                   static_cast<std::int64_t>(upd_b.size()), nub});
            const std::size_t s0 = src_a.size(), a0 = op_a.size(), ua0 = upd_a.size(), b0 = op_b.size(),
                              ub0 = upd_b.size();
            src_a.resize(s0 + static_cast<std::size_t>(nsrc) * 64, -1);
            op_a.resize(a0 + static_cast<std::size_t>(na) * 64, 0xffffffffu);
            upd_a.resize(ua0 + static_cast<std::size_t>(na) * nua * 64, 0);
            op_b.resize(b0 + static_cast<std::size_t>(nbo) * 64, 0xffffffffu);
            upd_b.resize(ub0 + static_cast<std::size_t>(nbo) * nub * 64, 0);
            for (int lane = 0; lane < cnt; ++lane) {
                for (std::size_t t2 = 0; t2 < srcs[lane].size(); ++t2) src_a[s0 + t2 * 64 + lane] = srcs[lane][t2];
                for (std::size_t j = 0; j < as[lane].size(); ++j) {
                    const auto& o = as[lane][j];
                    op_a[a0 + j * 64 + lane] = static_cast<std::uint32_t>(o.es) | (static_cast<std::uint32_t>(o.pt) << 8)
                        | (static_cast<std::uint32_t>(o.upd.size()) << 16);
                    for (std::size_t u = 0; u < o.upd.size(); ++u)
                        upd_a[ua0 + (j * nua + u) * 64 + lane] =
                            static_cast<std::uint16_t>(o.upd[u].first | (o.upd[u].second << 8));
                }
                for (std::size_t j = 0; j < bs_[lane].size(); ++j) {
                    const auto& o = bs_[lane][j];
                    op_b[b0 + j * 64 + lane] = static_cast<std::uint32_t>(o.step) | (static_cast<std::uint32_t>(o.es) << 8)
                        | (static_cast<std::uint32_t>(o.ps) << 16) | (static_cast<std::uint32_t>(o.upd.size()) << 24);
                    for (std::size_t u = 0; u < o.upd.size(); ++u)
                        upd_b[ub0 + (j * nub + u) * 64 + lane] =
                            static_cast<std::uint16_t>(o.upd[u].first | (o.upd[u].second << 8));
                }
            }
            max_src = std::max(max_src, nsrc);
            max_a = std::max(max_a, na);
            max_nua = std::max(max_nua, nua);
            max_b = std::max(max_b, nbo);
            max_nub = std::max(max_nub, nub);
        }
    }
    chunk_first[tri.blocks] = static_cast<std::int32_t>(chunks);
    if (std::max({src_a.size(), op_a.size(), upd_a.size(), op_b.size(), upd_b.size()}) > 0x7fffffffULL) return;
    P.kmax = kmax;
    P.max_src = max_src;
    P.max_a = max_a;
    P.max_nua = max_nua;
    P.max_b = max_b;
    P.max_nub = max_nub;
    const std::size_t lds = ilu_chunk_lds(P);
    if (lds > 64 * 1024) return;
    upload(P.chunk_first, chunk_first);
    upload(P.hdr, hdr);
    upload(P.src_a, src_a);
    upload(P.op_a, op_a);
    upload(P.op_b, op_b);
    upload(P.upd_a, upd_a);
    upload(P.upd_b, upd_b);
    P.enabled = true;
}

void ilu_setup(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    if (m->dense()) throw unsupported_error("ILU is defined for component-diagonal matrices");
    const LocalRows L = local_rows(m, pc->block_nodes);
    const std::int64_t n = m->owned;
    const std::size_t fill = static_cast<std::size_t>(std::max(0, pc->fill_level));
    auto& f = pc->ilu;
    f.h_offsets.assign(1, 0);
    f.h_columns.clear();
    f.h_diag.assign(n, 0);
    std::vector<std::int64_t> src;
    std::vector<std::size_t> levels;
    constexpr std::int64_t no_source = -1;
    // build_local_pattern (preconditioner.hpp:373-427), block-restricted
    for (std::int64_t row = 0; row < n; ++row) {
        std::map<std::int64_t, std::pair<std::size_t, std::int64_t>> entries;
        if (m->has_diag_array()) entries.emplace(row, std::pair{std::size_t(0), L.diag_src[row]});
        for (std::int64_t e = L.off[row]; e < L.off[row + 1]; ++e) {
            if (!entries.emplace(L.col[e], std::pair{std::size_t(0), L.src[e]}).second)
                throw std::invalid_argument("ILU requires unique matrix columns");
        }
        for (auto it = entries.begin(); it != entries.end() && it->first < row; ++it) {
            const std::int64_t pivot = it->first;
            const std::size_t lower_level = it->second.first;
            if (lower_level >= fill) continue;
            for (std::int64_t upper = f.h_diag[pivot] + 1; upper < f.h_offsets[pivot + 1]; ++upper) {
                if (levels[upper] > fill - lower_level - 1) continue;
                const std::size_t level = lower_level + levels[upper] + 1;
                auto [target, inserted] = entries.emplace(f.h_columns[upper], std::pair{level, no_source});
                if (!inserted) target->second.first = std::min(target->second.first, level);
            }
        }
        bool diagonal_found = false;
        for (const auto& [column, entry] : entries) {
            const std::int64_t destination = static_cast<std::int64_t>(f.h_columns.size());
            f.h_columns.push_back(column);
            levels.push_back(entry.first);
            src.push_back(entry.second);
            if (column == row) {
                f.h_diag[row] = destination;
                diagonal_found = true;
            }
        }
        if (!diagonal_found) throw std::invalid_argument("ILU(0) row has no diagonal");
        f.h_offsets.push_back(static_cast<std::int64_t>(f.h_columns.size()));
    }
    f.nnz = static_cast<std::int64_t>(f.h_columns.size());
    // update lists: for each lower entry (i,k), targets (i,j) of the pivot row's upper entries (k,j)
    std::vector<std::int64_t> upd_ptr(f.nnz + 1, 0), upd_target, upd_upper;
    for (std::int64_t row = 0; row < n; ++row) {
        for (std::int64_t e = f.h_offsets[row]; e < f.h_offsets[row + 1]; ++e) {
            if (e < f.h_diag[row]) {
                const std::int64_t k = f.h_columns[e];
                for (std::int64_t u = f.h_diag[k] + 1; u < f.h_offsets[k + 1]; ++u) {
                    const std::int64_t j = f.h_columns[u];
                    // find (row, j)
                    const auto begin = f.h_columns.begin() + f.h_offsets[row];
                    const auto end = f.h_columns.begin() + f.h_offsets[row + 1];
                    const auto it = std::lower_bound(begin, end, j);
                    if (it != end && *it == j) {
                        upd_target.push_back(static_cast<std::int64_t>(it - f.h_columns.begin()));
                        upd_upper.push_back(u);
                    }
                }
            }
            upd_ptr[e + 1] = static_cast<std::int64_t>(upd_target.size());
        }
    }
    upload(f.offsets, f.h_offsets);
    upload(f.diag_pos, f.h_diag);
    upload(f.src, src);
    upload(f.upd_ptr, upd_ptr);
    upload(f.upd_target, upd_target);
    upload(f.upd_upper, upd_upper);
    upload(f.columns, f.h_columns);
    f.values.allocate(std::max<std::int64_t>(f.nnz * m->bs, 1));
    // apply schedule: lower entries before the diagonal, upper after, factor order
    TriHost t;
    t.f_off.assign(n + 1, 0);
    t.b_off.assign(n + 1, 0);
    for (std::int64_t row = 0; row < n; ++row) {
        for (std::int64_t e = f.h_offsets[row]; e < f.h_diag[row]; ++e) {
            t.f_col.push_back(f.h_columns[e]);
            t.f_src.push_back(e);
        }
        for (std::int64_t e = f.h_diag[row] + 1; e < f.h_offsets[row + 1]; ++e) {
            t.b_col.push_back(f.h_columns[e]);
            t.b_src.push_back(e);
        }
        t.f_off[row + 1] = static_cast<std::int64_t>(t.f_col.size());
        t.b_off[row + 1] = static_cast<std::int64_t>(t.b_col.size());
    }
    build_schedule(pc, t);
    build_ilu_chunks(pc, upd_ptr, upd_target, upd_upper);
    const int bs = m->bs;
    pc->tri.fval64.allocate(std::max<std::int64_t>(pc->tri.fnnz * bs, 1));
    pc->tri.bval64.allocate(std::max<std::int64_t>(pc->tri.bnnz * bs, 1));
    pc->tri.dval64.allocate(std::max<std::int64_t>(n * bs, 1));
}

void ilu_numeric(owtk_precond pc)
{
    owtk_matrix m = pc->m;
    auto& f = pc->ilu;
    auto& t = pc->tri;
    const hipStream_t st = m->ctx->stream;
    const int bs = m->bs;
    clear_flag(pc);
    k_gather<<<grid_for(f.nnz * bs), kBlock, 0, st>>>(f.values.get(), f.src.get(), f.nnz, bs, comp_stride(m),
                                                      m->v64.get(), m->d64.get(), 1.0);
    OWTK_LAUNCH_CHECK();
    const unsigned int groups = static_cast<unsigned int>(t.blocks * bs);
    if (f.chunks.enabled) {
        const auto& P = f.chunks;
        const IluChunkArgs args{P.chunk_first.get(), P.hdr.get(), P.src_a.get(), P.op_a.get(), P.op_b.get(),
                                P.upd_a.get(), P.upd_b.get(), P.kmax, P.max_src, P.max_a, P.max_nua, P.max_b,
                                P.max_nub};
        // ILU(0) on a triangulated grid has at most 3 updates per op
        auto launch = [&](auto kernel) {
            kernel<<<groups, 64, ilu_chunk_lds(P), st>>>(args, f.offsets.get(), f.diag_pos.get(), t.row0.get(),
                                                         t.nrows.get(), bs, f.values.get(), f.nnz, t.rows,
                                                         t.dval64.get(), pc->flag.get());
        };
        const bool pf = ilu_chunks_prefetchable(P);
        if (P.max_nub <= 3) pf ? launch(k_ilu_factor_chunks<3, true>) : launch(k_ilu_factor_chunks<3, false>);
        else if (P.max_nub <= 6) pf ? launch(k_ilu_factor_chunks<6, true>) : launch(k_ilu_factor_chunks<6, false>);
        else pf ? launch(k_ilu_factor_chunks<0, true>) : launch(k_ilu_factor_chunks<0, false>);
    } else {
        k_ilu_factor<<<groups, 64, 0, st>>>(tri_args(t), t.frow_lvl.get(), f.offsets.get(), f.diag_pos.get(),
                                            f.upd_ptr.get(), f.upd_target.get(), f.upd_upper.get(), f.columns.get(),
                                            f.values.get(), f.nnz, t.dval64.get(), pc->flag.get());
    }
    OWTK_LAUNCH_CHECK();
    int flag = 0;
    read_flag(pc, &flag);
    if (flag != flag_ok) throw zero_pivot_error("zero pivot in ILU(0)");
    if (t.fnnz)
        k_gather_cm<<<grid_for(t.fnnz * bs), kBlock, 0, st>>>(t.fval64.get(), t.fsrc.get(), t.fnnz, bs,
                                                              f.values.get(), f.nnz);
    if (t.bnnz)
        k_gather_cm<<<grid_for(t.bnnz * bs), kBlock, 0, st>>>(t.bval64.get(), t.bsrc.get(), t.bnnz, bs,
                                                              f.values.get(), f.nnz);
    OWTK_LAUNCH_CHECK();
    refresh_storage(pc);
}

// -------------------------------------------------------------------- apply --
template<TriMode MODE, class X, class V>
void tri_launch(owtk_precond pc, const V* fv, const V* bv, const X* dv, const X* in, X* out, const int* stop)
{
    const auto& t = pc->tri;
    const unsigned int groups = static_cast<unsigned int>(t.blocks * t.bs);
    const std::size_t need = static_cast<std::size_t>(t.max_block_rows) * sizeof(X);
    const int threads = t.max_level_width > 128 ? 256 : (t.max_level_width > 64 ? 128 : 64);
    const X scale = static_cast<X>(pc->omega * (2.0 - pc->omega));
    const hipStream_t st = pc->m->ctx->stream;
    ProfileScope prof(pc->m->ctx, sizeof(X) == 8 ? prof_precond_f64 : prof_precond_low, precond_apply_bytes(pc));
    if (t.chain && t.ell) {
        const TriArgs args = tri_args(t);
        const EllArgs ea{t.chunk0.get(), t.fidx_ell.get(), t.bidx_ell.get(), t.fslots, t.bslots, t.fk, t.bk};
        const V* efv;
        const V* ebv;
        if constexpr (std::is_same_v<V, double>) {
            efv = t.fell64.get();
            ebv = t.bell64.get();
        } else if constexpr (std::is_same_v<V, float>) {
            efv = t.fell32.get();
            ebv = t.bell32.get();
        } else {
            efv = t.fell16.get();
            ebv = t.bell16.get();
        }
        const std::size_t tri_lds = static_cast<std::size_t>(2 + 64 * 63 / 2) * sizeof(X);
        const bool global_y = need + tri_lds > kChainLdsBudget;
        const std::size_t lds = global_y ? tri_lds : need + tri_lds;
        const int k = std::max(t.fk, t.bk);
        auto launch = [&](auto e, auto gy) {
            k_tri_ell<MODE, X, V, decltype(e)::value, decltype(gy)::value>
                <<<groups, 64, lds, st>>>(args, ea, efv, ebv, dv, in, out, scale, stop);
        };
        auto pick = [&](auto e) {
            if (global_y) launch(e, std::true_type{});
            else launch(e, std::false_type{});
        };
        if (k == 4) pick(std::integral_constant<int, 4>{});
        else if (k == 6) pick(std::integral_constant<int, 6>{});
        else if (k == 8) pick(std::integral_constant<int, 8>{});
        else pick(std::integral_constant<int, 16>{});
    } else if (t.chain) {
        const TriArgs args = tri_args(t);
        // Large blocks keep their solution in global memory (L2): the LDS
        // then holds only the tile, and more one-wave workgroups fit per CU.
        const std::size_t tile_lds = 64 * 64 * sizeof(X);
        const bool global_y = need + tile_lds > kChainLdsBudget;
        const std::size_t chain_lds = global_y ? tile_lds : need + tile_lds;
        auto launch = [&](auto ei, auto gy) {
            k_tri_chain<MODE, X, V, decltype(ei)::value, decltype(gy)::value>
                <<<groups, 64, chain_lds, st>>>(args, fv, bv, dv, in, out, scale, stop);
        };
        auto pick = [&](auto ei) {
            if (global_y) launch(ei, std::true_type{});
            else launch(ei, std::false_type{});
        };
        if (t.ei_max <= 2) pick(std::integral_constant<int, 2>{});
        else if (t.ei_max <= 4) pick(std::integral_constant<int, 4>{});
        else if (t.ei_max <= 8) pick(std::integral_constant<int, 8>{});
        else pick(std::integral_constant<int, 16>{});
    } else if (need <= lds_limit(pc)) {
        k_tri_apply<MODE, X, V, true><<<groups, threads, need, st>>>(tri_args(t), fv, bv, dv, in, out, scale, stop);
    } else {
        k_tri_apply<MODE, X, V, false><<<groups, threads, 0, st>>>(tri_args(t), fv, bv, dv, in, out, scale, stop);
    }
    OWTK_LAUNCH_CHECK();
}

} // namespace

void precond_apply_f64(owtk_precond pc, const double* in, double* out, const int* stop)
{
    owtk_matrix m = pc->m;
    const hipStream_t st = m->ctx->stream;
    const std::int64_t n = pc->n();
    switch (pc->kind) {
    case OWTK_PC_IDENTITY:
        if (in != out) OWTK_HIP(hipMemcpyAsync(out, in, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        return;
    case OWTK_PC_JACOBI:
        if (m->dense()) {
            const unsigned int g = static_cast<unsigned int>((m->owned + 255) / 256);
            switch (m->bs) {
            case 1: k_block_jacobi_apply<1, double, double><<<g, kBlock, 0, st>>>(pc->inv64.get(), in, out, m->owned, stop); break;
            case 2: k_block_jacobi_apply<2, double, double><<<g, kBlock, 0, st>>>(pc->inv64.get(), in, out, m->owned, stop); break;
            case 3: k_block_jacobi_apply<3, double, double><<<g, kBlock, 0, st>>>(pc->inv64.get(), in, out, m->owned, stop); break;
            default: k_block_jacobi_apply<4, double, double><<<g, kBlock, 0, st>>>(pc->inv64.get(), in, out, m->owned, stop); break;
            }
        } else {
            k_jacobi_apply<double, double><<<grid_for(n), kBlock, 0, st>>>(pc->inv64.get(), in, out, n, stop);
        }
        OWTK_LAUNCH_CHECK();
        return;
    case OWTK_PC_SSOR:
        if (m->dense()) {
            dense_ssor_apply<double, double, double>(pc, in, out, pc->inv64.get(), pc->work64.get(), stop);
            return;
        }
        tri_launch<TriMode::ssor, double, double>(pc, pc->tri.fval64.get(), pc->tri.bval64.get(),
                                                  pc->tri.dval64.get(), in, out, stop);
        return;
    case OWTK_PC_ILU:
        tri_launch<TriMode::ilu, double, double>(pc, pc->tri.fval64.get(), pc->tri.bval64.get(),
                                                 pc->tri.dval64.get(), in, out, stop);
        return;
    case OWTK_PC_LEARNED:
        learned_apply_f64(pc, in, out, stop);
        return;
    default:
        throw unsupported_error("preconditioner kind not available on the FP64 path");
    }
}

void precond_apply_f32(owtk_precond pc, const float* in, float* out, const int* stop)
{
    owtk_matrix m = pc->m;
    const hipStream_t st = m->ctx->stream;
    const std::int64_t n = pc->n();
    if (pc->kind != OWTK_PC_IDENTITY && pc->storage == OWTK_F64)
        throw std::invalid_argument("FP32 apply needs F32 or BF16 preconditioner storage");
    switch (pc->kind) {
    case OWTK_PC_IDENTITY:
        if (in != out) OWTK_HIP(hipMemcpyAsync(out, in, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        return;
    case OWTK_PC_JACOBI:
        if (m->dense()) {
            const unsigned int g = static_cast<unsigned int>((m->owned + 255) / 256);
            switch (m->bs) {
            case 1: k_block_jacobi_apply<1, float, float><<<g, kBlock, 0, st>>>(pc->inv32.get(), in, out, m->owned, stop); break;
            case 2: k_block_jacobi_apply<2, float, float><<<g, kBlock, 0, st>>>(pc->inv32.get(), in, out, m->owned, stop); break;
            case 3: k_block_jacobi_apply<3, float, float><<<g, kBlock, 0, st>>>(pc->inv32.get(), in, out, m->owned, stop); break;
            default: k_block_jacobi_apply<4, float, float><<<g, kBlock, 0, st>>>(pc->inv32.get(), in, out, m->owned, stop); break;
            }
        } else {
            k_jacobi_apply<float, float><<<grid_for(n), kBlock, 0, st>>>(pc->inv32.get(), in, out, n, stop);
        }
        OWTK_LAUNCH_CHECK();
        return;
    case OWTK_PC_SSOR:
    case OWTK_PC_ILU: {
        if (m->dense()) {  // dense-block SSOR (ILU is rejected at creation)
            if (pc->storage == OWTK_F32) {
                if (!m->v32.get()) throw std::invalid_argument("matrix has no FP32 storage");
                dense_ssor_apply<float, float, float>(pc, in, out, pc->inv32.get(), pc->work32.get(), stop);
            } else {
                if (!m->v16.get()) throw std::invalid_argument("matrix has no BF16 storage");
                dense_ssor_apply<std::uint16_t, float, float>(pc, in, out, pc->inv32.get(), pc->work32.get(), stop);
            }
            return;
        }
        const auto& t = pc->tri;
        const bool ilu = pc->kind == OWTK_PC_ILU;
        if (pc->storage == OWTK_F32) {
            if (ilu) tri_launch<TriMode::ilu, float, float>(pc, t.fval32.get(), t.bval32.get(), t.dval32.get(), in, out, stop);
            else tri_launch<TriMode::ssor, float, float>(pc, t.fval32.get(), t.bval32.get(), t.dval32.get(), in, out, stop);
        } else {
            if (ilu) tri_launch<TriMode::ilu, float, std::uint16_t>(pc, t.fval16.get(), t.bval16.get(), t.dval32.get(), in, out, stop);
            else tri_launch<TriMode::ssor, float, std::uint16_t>(pc, t.fval16.get(), t.bval16.get(), t.dval32.get(), in, out, stop);
        }
        return;
    }
    case OWTK_PC_LEARNED:
        learned_apply_f32(pc, in, out, stop);
        return;
    default:
        throw unsupported_error("preconditioner kind not available on the FP32 path");
    }
}

std::int64_t precond_apply_bytes(owtk_precond pc)
{
    const std::int64_t n = pc->n();
    const int sv = pc->storage == OWTK_F64 ? 8 : (pc->storage == OWTK_F32 ? 4 : 2);
    const int sx = pc->storage == OWTK_F64 ? 8 : 4;
    switch (pc->kind) {
    case OWTK_PC_IDENTITY: return 2 * n * sx;
    case OWTK_PC_JACOBI:
        return pc->m->dense() ? n * pc->m->bs * sv + 2 * n * sx : n * sv + 2 * n * sx;
    case OWTK_PC_LEARNED: return learned_apply_bytes(pc);
    case OWTK_PC_SSOR:
        if (pc->m->dense()) {
            // both sweeps read every block once (lower or upper), the diagonal
            // block and inverse once each; in, y, out
            const std::int64_t b2 = std::int64_t(pc->m->bs) * pc->m->bs;
            return pc->m->entries * (b2 * sv + 4) + pc->m->owned * (b2 * sv + b2 * sv + 8) + 4 * n * sx;
        }
        [[fallthrough]];
    default: {
        const auto& t = pc->tri;
        const int bs = pc->m->bs;
        // factor values + block-local columns + row positions/pointers + diagonal + in/out
        return (t.fnnz + t.bnnz) * (bs * sv + 4) + 2 * t.rows * (4 + 8) + n * sx + 2 * n * sx;
    }
    }
}

} // namespace owtk

extern "C" {

void owtk_precond_params_default(owtk_precond_params* p, int32_t kind)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->storage = OWTK_F64;
    p->omega = 1.0;
    p->minimum_diagonal = 0.0;
    p->block_nodes = 0;
    p->fill_level = 0;
    p->learned_model = nullptr;
}

int owtk_precond_create(owtk_matrix m, const owtk_precond_params* params, owtk_precond* out)
{
    return guarded([&] {
        if (!m || !params || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        m->ctx->activate();
        auto pc = std::make_unique<owtk_precond_s>();
        pc->m = m;
        pc->ctx = m->ctx;
        pc->kind = params->kind;
        pc->storage = params->storage;
        pc->omega = params->omega;
        pc->minimum_diagonal = params->minimum_diagonal;
        pc->block_nodes = params->block_nodes;
        pc->fill_level = params->fill_level;
        pc->schedule = params->schedule;
        pc->learned = params->learned_model;
        if (pc->schedule < 0 || pc->schedule > 2) throw std::invalid_argument("unknown triangular schedule");
        if (pc->storage < OWTK_F64 || pc->storage > OWTK_BF16) throw std::invalid_argument("unknown storage type");
        if (pc->block_nodes < 0) throw std::invalid_argument("negative block size");
        pc->flag.allocate(4);
        switch (pc->kind) {
        case OWTK_PC_IDENTITY: break;
        case OWTK_PC_JACOBI:
            if (pc->storage == OWTK_BF16) pc->storage = OWTK_F32;
            jacobi_setup(pc.get());
            jacobi_numeric(pc.get());
            break;
        case OWTK_PC_SSOR:
            ssor_setup(pc.get());
            ssor_numeric(pc.get());
            break;
        case OWTK_PC_ILU:
            ilu_setup(pc.get());
            ilu_numeric(pc.get());
            break;
        case OWTK_PC_LEARNED:
            learned_setup(pc.get());
            try {
                learned_numeric(pc.get());
            } catch (...) {
                learned_release(pc.get());
                throw;
            }
            break;
        default: throw unsupported_error("preconditioner kind not available");
        }
        pc->updates = 1;
        m->ctx->sync();
        *out = pc.release();
    });
}

int owtk_precond_update_values(owtk_precond pc)
{
    return guarded([&] {
        if (!pc) throw std::invalid_argument("null preconditioner");
        pc->m->ctx->activate();
        switch (pc->kind) {
        case OWTK_PC_JACOBI: jacobi_numeric(pc); break;
        case OWTK_PC_SSOR: ssor_numeric(pc); break;
        case OWTK_PC_ILU: ilu_numeric(pc); break;
        case OWTK_PC_LEARNED: learned_numeric(pc); break;
        default: break;
        }
        ++pc->updates;
        pc->m->ctx->sync();
    });
}

int owtk_precond_apply(owtk_precond pc, const void* in, void* out)
{
    return guarded([&] {
        if (!pc) throw std::invalid_argument("null preconditioner");
        if (!in || !out) throw std::invalid_argument("null vector");
        pc->m->ctx->activate();
        if (pc->storage == OWTK_F64) precond_apply_f64(pc, static_cast<const double*>(in), static_cast<double*>(out), nullptr);
        else precond_apply_f32(pc, static_cast<const float*>(in), static_cast<float*>(out), nullptr);
    });
}

int owtk_precond_numeric_updates(owtk_precond pc, uint64_t* count)
{
    return guarded([&] {
        if (!pc || !count) throw std::invalid_argument("null argument");
        *count = pc->updates;
    });
}

int owtk_precond_info_get(owtk_precond pc, owtk_precond_info* info)
{
    return guarded([&] {
        if (!pc || !info) throw std::invalid_argument("null argument");
        info->kind = pc->kind;
        info->storage = pc->storage;
        info->blocks = pc->tri.blocks;
        info->forward_levels = pc->tri.max_flevels;
        info->backward_levels = pc->tri.max_blevels;
        info->factor_entries = pc->ilu.nnz;
        info->apply_bytes = precond_apply_bytes(pc);
        info->rows = pc->m->owned;
        info->block_size = pc->m->bs;
    });
}

int owtk_precond_ilu_factors(owtk_precond pc, int64_t* entries, int64_t* row_offsets, int64_t* columns,
                             double* values, double* inverse_diagonal)
{
    return guarded([&] {
        if (!pc || !entries) throw std::invalid_argument("null argument");
        if (pc->kind != OWTK_PC_ILU) throw std::invalid_argument("not an ILU preconditioner");
        pc->m->ctx->activate();
        const auto& f = pc->ilu;
        *entries = f.nnz;
        if (row_offsets) std::copy(f.h_offsets.begin(), f.h_offsets.end(), row_offsets);
        if (columns) std::copy(f.h_columns.begin(), f.h_columns.end(), columns);
        const int bs = pc->m->bs;
        if (values) {
            std::vector<double> cm(static_cast<std::size_t>(f.nnz * bs));
            OWTK_HIP(hipMemcpy(cm.data(), f.values.get(), cm.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (std::int64_t e = 0; e < f.nnz; ++e)
                for (int c = 0; c < bs; ++c) values[e * bs + c] = cm[c * f.nnz + e];
        }
        if (inverse_diagonal) {
            const std::int64_t n = pc->m->owned;
            std::vector<double> cm(static_cast<std::size_t>(n * bs));
            OWTK_HIP(hipMemcpy(cm.data(), pc->tri.dval64.get(), cm.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (std::int64_t r = 0; r < n; ++r)
                for (int c = 0; c < bs; ++c) inverse_diagonal[r * bs + c] = cm[c * n + r];
        }
    });
}

int owtk_precond_destroy(owtk_precond pc)
{
    return guarded([&] {
        if (!pc) return;
        pc->ctx->activate();
        pc->ctx->sync();
        if (pc->kind == OWTK_PC_LEARNED) learned_release(pc);
        delete pc;
    });
}

} // extern "C"
