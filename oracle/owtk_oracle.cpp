// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the OWT-Krylov block-CSR solve path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// It is never linked into, loaded by or called from the product library.
//
// Each function cites the reference file:line it restates.  Operation order
// per output element follows the reference; wherever the reference writes a
// multiply-add statement (out += a*b, y -= a*b, ...) this file uses an
// explicit std::fma, which is what GCC emits for those statements when FMA is
// available (-O3 -march=native, default -ffp-contract=fast).  Built with
// -ffp-contract=off so that no other contraction happens.
//
// Pinning (see tests/test_oracle.py): SpMV by the reference's closed-form
// known answers (test_core.cpp:31-58, 176-195, 233-279) and a long-double
// dense matvec; Jacobi/SSOR by closed forms (test_core.cpp:434-437,
// review_cases.cpp:95-113, review_contracts.cpp:29-96); ILU(k) by a dense
// dot-product Doolittle on the retained pattern and LU basis columns
// (test_bicgstab_ilu_audit.cpp:14-91); BiCGSTAB by a long-double dense
// recurrence (ilu_audit_reference.hpp:119-156) and the exact-preconditioner
// one-iteration case; FGMRES by orthonormality/Arnoldi invariants and a
// direct dense solve; norms by the extreme-magnitude cases
// (test_foundation.cpp:71-107).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <omp.h>

namespace {

thread_local std::string g_error;

// Threads of the all-cores CPU baseline (oracle_set_threads).  With the
// default of 1 every loop below runs serially in the reference's order, which
// is what the tests check; with more threads the row, block and vector loops
// run under OpenMP (results per row unchanged) and the dot products and
// norms sum per-thread partials (a different summation order).
int g_threads = 1;
#define OMP_ROWS _Pragma("omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)")

// ------------------------------------------------------------------ matrix --
enum Kind { kCsr = 0, kSplit = 1, kBsr = 2 };

struct Matrix {
    int kind;
    std::int64_t owned, ghost;
    int bs;
    const std::int32_t* off;
    const std::int32_t* col;
    const double* val;   // csr: entries*bs; split: off-diagonal; bsr: entries*bs*bs
    const double* diag;  // split only
    std::int64_t local() const { return owned + ghost; }
};

// BlockCsrMatrix::apply_rows (block_csr.hpp:198-215)
void spmv_csr(const Matrix& a, const double* x, double* y)
{
    const int bs = a.bs;
    OMP_ROWS
    for (std::int64_t row = 0; row < a.owned; ++row) {
        double* out = y + row * bs;
        std::fill(out, out + bs, 0.0);
        for (std::int32_t e = a.off[row]; e < a.off[row + 1]; ++e) {
            const double* in = x + static_cast<std::int64_t>(a.col[e]) * bs;
            const double* coef = a.val + static_cast<std::int64_t>(e) * bs;
            for (int c = 0; c < bs; ++c) out[c] = std::fma(coef[c], in[c], out[c]);
        }
    }
}

// SplitBlockCsrMatrixView::apply_rows (split_block_csr.hpp:95-119)
void spmv_split(const Matrix& a, const double* x, double* y)
{
    const int bs = a.bs;
    OMP_ROWS
    for (std::int64_t row = 0; row < a.owned; ++row) {
        double* out = y + row * bs;
        for (int c = 0; c < bs; ++c) out[c] = a.diag[row * bs + c] * x[row * bs + c];
        for (std::int32_t e = a.off[row]; e < a.off[row + 1]; ++e) {
            const double* in = x + static_cast<std::int64_t>(a.col[e]) * bs;
            const double* coef = a.val + static_cast<std::int64_t>(e) * bs;
            for (int c = 0; c < bs; ++c) out[c] = std::fma(coef[c], in[c], out[c]);
        }
    }
}

// DenseBlockCsrMatrix::apply_rows (dense_block_csr.hpp:179-205)
void spmv_bsr(const Matrix& a, const double* x, double* y)
{
    const int bs = a.bs;
    OMP_ROWS
    for (std::int64_t row = 0; row < a.owned; ++row) {
        double* dst = y + row * bs;
        std::fill(dst, dst + bs, 0.0);
        for (std::int32_t e = a.off[row]; e < a.off[row + 1]; ++e) {
            const double* src = x + static_cast<std::int64_t>(a.col[e]) * bs;
            const double* blk = a.val + static_cast<std::int64_t>(e) * bs * bs;
            for (int r = 0; r < bs; ++r) {
                double t = 0.0;
                for (int c = 0; c < bs; ++c) t = std::fma(blk[r * bs + c], src[c], t);
                dst[r] += t;
            }
        }
    }
}

void apply(const Matrix& a, const double* x, double* y)
{
    if (a.kind == kCsr) spmv_csr(a, x, y);
    else if (a.kind == kSplit) spmv_split(a, x, y);
    else spmv_bsr(a, x, y);
}

// Merged component-diagonal CSR (diagonal in sorted position) of a split or
// csr matrix, optionally restricted to the block-diagonal of `block_nodes`
// consecutive node ranges (the rank-local restriction the reference applies to
// non-owned columns, preconditioner.hpp:389-391).  Ghost columns dropped.
struct Csr {
    std::int64_t n = 0;
    int bs = 1;
    std::vector<std::int32_t> off, col;
    std::vector<double> val;
};

Csr local_csr(const Matrix& a, std::int64_t block_nodes, bool sorted = true)
{
    Csr out;
    out.n = a.owned;
    out.bs = a.bs;
    out.off.push_back(0);
    const int bs = a.bs;
    auto same_block = [&](std::int64_t r, std::int64_t c) {
        if (c >= a.owned) return false;
        if (block_nodes <= 0) return true;
        return r / block_nodes == c / block_nodes;
    };
    for (std::int64_t row = 0; row < a.owned; ++row) {
        std::vector<std::pair<std::int32_t, std::vector<double>>> entries;
        if (a.kind == kSplit) {
            entries.push_back({static_cast<std::int32_t>(row),
                               std::vector<double>(a.diag + row * bs, a.diag + row * bs + bs)});
        }
        for (std::int32_t e = a.off[row]; e < a.off[row + 1]; ++e) {
            if (!same_block(row, a.col[e])) continue;
            const double* v = a.val + static_cast<std::int64_t>(e) * bs;
            entries.push_back({a.col[e], std::vector<double>(v, v + bs)});
        }
        if (sorted)
            std::stable_sort(entries.begin(), entries.end(),
                             [](const auto& l, const auto& r) { return l.first < r.first; });
        for (auto& [c, v] : entries) {
            out.col.push_back(c);
            out.val.insert(out.val.end(), v.begin(), v.end());
        }
        out.off.push_back(static_cast<std::int32_t>(out.col.size()));
    }
    return out;
}

// -------------------------------------------------------------- reductions --
// SerialReduction::dot (reduction.hpp:114-145): eight lane accumulators.
double dot_serial(std::int64_t n, const double* a, const double* b)
{
    double partial[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::int64_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int lane = 0; lane < 8; ++lane) partial[lane] = std::fma(a[i + lane], b[i + lane], partial[lane]);
    double sum = (partial[0] + partial[1]) + (partial[2] + partial[3]) + (partial[4] + partial[5])
        + (partial[6] + partial[7]);
    for (; i < n; ++i) sum = std::fma(a[i], b[i], sum);
    return sum;
}

// All-cores baseline: contiguous per-thread chunks, each summed as above,
// the chunk sums added in thread order.
double dot(std::int64_t n, const double* a, const double* b)
{
    if (g_threads <= 1 || n < 65536) return dot_serial(n, a, b);
    std::vector<double> part(static_cast<std::size_t>(g_threads), 0.0);
#pragma omp parallel num_threads(g_threads)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        const std::int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        part[t] = dot_serial(hi - lo, a + lo, b + lo);
    }
    double sum = 0;
    for (double v : part) sum += v;
    return sum;
}

double norm_serial(std::int64_t n, const double* v);

// detail::scaled_norm (reduction.hpp:64-89); all-cores baseline: the maximum
// and the scaled compensated sums per thread chunk, then combined.
double norm(std::int64_t n, const double* v)
{
    if (g_threads <= 1 || n < 65536) return norm_serial(n, v);
    const int nt = g_threads;
    std::vector<double> scale(nt, 0.0), sum(nt, 0.0), corr(nt, 0.0);
    std::vector<int> nan(nt, 0);
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num(), nth = omp_get_num_threads();
        const std::int64_t lo = n * t / nth, hi = n * (t + 1) / nth;
        for (std::int64_t i = lo; i < hi; ++i) {
            const double m = std::abs(v[i]);
            if (std::isnan(m)) nan[t] = 1;
            scale[t] = std::max(scale[t], m);
        }
    }
    double s = 0;
    for (int t = 0; t < nt; ++t) {
        if (nan[t]) return std::numeric_limits<double>::quiet_NaN();
        s = std::max(s, scale[t]);
    }
    if (s == 0.0 || std::isinf(s)) return s;
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num(), nth = omp_get_num_threads();
        const std::int64_t lo = n * t / nth, hi = n * (t + 1) / nth;
        double acc = 0, c = 0;
        for (std::int64_t i = lo; i < hi; ++i) {
            const double ratio = v[i] / s;
            const double term = ratio * ratio;
            const double updated = acc + term;
            c += acc >= term ? (acc - updated) + term : (term - updated) + acc;
            acc = updated;
        }
        sum[t] = acc;
        corr[t] = c;
    }
    double total = 0, c = 0;
    for (int t = 0; t < nt; ++t) {
        total += sum[t];
        c += corr[t];
    }
    return s * std::sqrt(total + c);
}

double norm_serial(std::int64_t n, const double* v)
{
    double scale = 0;
    for (std::int64_t i = 0; i < n; ++i) {
        const double m = std::abs(v[i]);
        if (std::isnan(m)) return std::numeric_limits<double>::quiet_NaN();
        scale = std::max(scale, m);
    }
    if (scale == 0.0 || std::isinf(scale)) return scale;
    double sum = 0, correction = 0;
    for (std::int64_t i = 0; i < n; ++i) {
        const double ratio = v[i] / scale;
        const double term = ratio * ratio;
        const double updated = sum + term;
        correction += sum >= term ? (sum - updated) + term : (term - updated) + sum;
        sum = updated;
    }
    return scale * std::sqrt(sum + correction);
}

// --------------------------------------------------------- preconditioners --
struct Pc {
    virtual ~Pc() = default;
    virtual void apply(const double* in, double* out) = 0;
    virtual void update() {}
    std::uint64_t updates = 0;
};

struct Identity : Pc {
    std::int64_t n;
    explicit Identity(std::int64_t size) : n(size) {}
    void apply(const double* in, double* out) override { std::copy(in, in + n, out); }  // preconditioner.hpp:17-23
};

// JacobiPreconditioner (preconditioner.hpp:25-117)
struct Jacobi : Pc {
    const Matrix* a;
    double minimum;
    std::vector<double> inv;
    Jacobi(const Matrix* m, double min_diag) : a(m), minimum(min_diag) { update(); }
    void update() override
    {
        const int bs = a->bs;
        inv.assign(a->owned * bs, 0.0);
        std::vector<double> d(a->owned * bs, 0.0);
        if (a->kind == kSplit) {
            std::copy(a->diag, a->diag + a->owned * bs, d.begin());
        } else if (a->kind == kCsr) {
            for (std::int64_t row = 0; row < a->owned; ++row) {
                bool found = false;
                for (std::int32_t e = a->off[row]; e < a->off[row + 1]; ++e) {
                    if (a->col[e] == row) {
                        std::copy_n(a->val + static_cast<std::int64_t>(e) * bs, bs, d.begin() + row * bs);
                        found = true;
                        break;
                    }
                }
                if (!found) throw std::invalid_argument("BlockCsrMatrix row has no diagonal");
            }
        } else {
            throw std::invalid_argument("dense-block CSR has no component diagonal");
        }
        for (std::size_t i = 0; i < d.size(); ++i) {
            if (!std::isfinite(d[i]) || std::abs(d[i]) <= minimum)
                throw std::invalid_argument("Jacobi preconditioner has a singular diagonal");
            inv[i] = 1.0 / d[i];
        }
        ++updates;
    }
    void apply(const double* in, double* out) override
    {
        for (std::size_t i = 0; i < inv.size(); ++i) out[i] = inv[i] * in[i];
    }
};

// LocalSsorPreconditioner (preconditioner.hpp:124-247) on the merged local CSR
// (optionally block-restricted), and SplitLocalSsorPreconditioner
// (split_block_csr.hpp:204-332) which skips the output clear and the omega=1
// scaling; both give the same values.
struct Ssor : Pc {
    const Matrix* a;
    double omega;
    std::int64_t block_nodes;
    Csr c;
    std::vector<double> diag, work;
    Ssor(const Matrix* m, double w, std::int64_t blocks) : a(m), omega(w), block_nodes(blocks)
    {
        if (!(w > 0 && w < 2)) throw std::invalid_argument("SSOR omega must be in (0,2)");
        update();
    }
    void update() override
    {
        // the sweeps visit each row's entries in the matrix's own CSR order
        c = local_csr(*a, block_nodes, false);
        const int bs = a->bs;
        diag.assign(a->owned * bs, 0.0);
        work.assign(a->owned * bs, 0.0);
        const double tol = 64 * std::numeric_limits<double>::epsilon();
        for (std::int64_t row = 0; row < c.n; ++row) {
            bool found = false;
            for (std::int32_t e = c.off[row]; e < c.off[row + 1]; ++e) {
                if (c.col[e] != row) continue;
                for (int k = 0; k < bs; ++k) {
                    const double v = c.val[static_cast<std::int64_t>(e) * bs + k];
                    if (!std::isfinite(v) || std::abs(v) <= tol)
                        throw std::invalid_argument("SSOR preconditioner has a singular diagonal");
                    diag[row * bs + k] = v;
                }
                found = true;
                break;
            }
            if (!found) throw std::invalid_argument("SSOR preconditioner row has no diagonal");
        }
        ++updates;
    }
    void apply(const double* in, double* out) override
    {
        const int bs = a->bs;
        const std::int64_t n = c.n;
        for (std::int64_t row = 0; row < n; ++row) {
            double* y = work.data() + row * bs;
            std::copy_n(in + row * bs, bs, y);
            for (std::int32_t e = c.off[row]; e < c.off[row + 1]; ++e) {
                const std::int64_t column = c.col[e];
                if (column >= row) continue;
                const double* coef = c.val.data() + static_cast<std::int64_t>(e) * bs;
                const double* prev = work.data() + column * bs;
                for (int k = 0; k < bs; ++k) y[k] = std::fma(-(omega * coef[k]), prev[k], y[k]);
            }
            for (int k = 0; k < bs; ++k) y[k] /= diag[row * bs + k];
        }
        std::vector<double> z(n * bs);
        for (std::int64_t row = n; row-- > 0;) {
            double* zr = z.data() + row * bs;
            const double* y = work.data() + row * bs;
            for (int k = 0; k < bs; ++k) zr[k] = diag[row * bs + k] * y[k];
            for (std::int32_t e = c.off[row]; e < c.off[row + 1]; ++e) {
                const std::int64_t column = c.col[e];
                if (column <= row) continue;
                const double* coef = c.val.data() + static_cast<std::int64_t>(e) * bs;
                const double* prev = z.data() + column * bs;
                for (int k = 0; k < bs; ++k) zr[k] = std::fma(-(omega * coef[k]), prev[k], zr[k]);
            }
            for (int k = 0; k < bs; ++k) zr[k] /= diag[row * bs + k];
        }
        const double s = omega * (2.0 - omega);
        if (omega != 1.0)
            for (std::int64_t i = 0; i < n * bs; ++i) z[i] *= s;
        std::copy(z.begin(), z.end(), out);
    }
};

// Dense-block SSOR in multicolour order.  No reference counterpart: the
// reference's DenseBlockCsrMatrix has no diagonal() and no SSOR
// (dense_block_csr.hpp; preconditioner.hpp:28-34).  This restates the
// project's definition: greedy colouring of the symmetrised owned graph in
// row order; forward y_i = D_i^-1 (r_i - omega sum_(colour j < colour i)
// A_ij y_j), backward z_i = D_i^-1 (s D_i y_i - omega sum_(colour j >
// colour i) A_ij z_j), s = omega (2 - omega); D_i^-1 by Gauss-Jordan with
// partial pivoting.  Pinned in tests by the textbook identity: the result
// equals s (D + omega U)^-1 D (D + omega L)^-1 r of the colour-permuted
// matrix, solved densely.
struct DenseSsor : Pc {
    const Matrix* a;
    double omega;
    int bs;
    std::vector<std::int32_t> colour, order;
    std::vector<double> inv, dblk, work;
    DenseSsor(const Matrix* m, double w) : a(m), omega(w), bs(m->bs)
    {
        if (!(w > 0 && w < 2)) throw std::invalid_argument("SSOR omega must be in (0,2)");
        const std::int64_t n = a->owned;
        std::vector<std::vector<std::int32_t>> adj(n);
        for (std::int64_t i = 0; i < n; ++i)
            for (std::int32_t e = a->off[i]; e < a->off[i + 1]; ++e) {
                const std::int64_t j = a->col[e];
                if (j == i || j >= n) continue;
                adj[i].push_back(static_cast<std::int32_t>(j));
                adj[j].push_back(static_cast<std::int32_t>(i));
            }
        colour.assign(n, -1);
        for (std::int64_t i = 0; i < n; ++i) {
            std::vector<char> used(adj[i].size() + 2, 0);
            for (std::int32_t j : adj[i])
                if (colour[j] >= 0 && colour[j] < static_cast<std::int32_t>(used.size())) used[colour[j]] = 1;
            std::int32_t c = 0;
            while (used[c]) ++c;
            colour[i] = c;
        }
        order.resize(n);
        for (std::int64_t i = 0; i < n; ++i) order[i] = static_cast<std::int32_t>(i);
        std::stable_sort(order.begin(), order.end(), [&](std::int32_t x, std::int32_t y) { return colour[x] < colour[y]; });
        update();
    }
    void update() override
    {
        const std::int64_t n = a->owned;
        const int b2 = bs * bs;
        inv.assign(n * b2, 0.0);
        dblk.assign(n * b2, 0.0);
        work.assign(n * bs, 0.0);
        for (std::int64_t i = 0; i < n; ++i) {
            const double* blk = nullptr;
            for (std::int32_t e = a->off[i]; e < a->off[i + 1]; ++e)
                if (a->col[e] == i) blk = a->val + static_cast<std::int64_t>(e) * b2;
            if (!blk) throw std::invalid_argument("dense-block CSR row has no diagonal block");
            std::copy(blk, blk + b2, dblk.begin() + i * b2);
            std::vector<double> m(blk, blk + b2), r(b2, 0.0);
            for (int k = 0; k < bs; ++k) r[k * bs + k] = 1.0;
            for (int k = 0; k < bs; ++k) {
                int p = k;
                for (int q = k + 1; q < bs; ++q)
                    if (std::abs(m[q * bs + k]) > std::abs(m[p * bs + k])) p = q;
                if (!(std::abs(m[p * bs + k]) > 0.0) || !std::isfinite(m[p * bs + k]))
                    throw std::invalid_argument("SSOR preconditioner has a singular diagonal block");
                if (p != k)
                    for (int c = 0; c < bs; ++c) {
                        std::swap(m[k * bs + c], m[p * bs + c]);
                        std::swap(r[k * bs + c], r[p * bs + c]);
                    }
                const double ip = 1.0 / m[k * bs + k];
                for (int c = 0; c < bs; ++c) {
                    m[k * bs + c] *= ip;
                    r[k * bs + c] *= ip;
                }
                for (int q = 0; q < bs; ++q) {
                    if (q == k) continue;
                    const double f = m[q * bs + k];
                    for (int c = 0; c < bs; ++c) {
                        m[q * bs + c] = std::fma(-f, m[k * bs + c], m[q * bs + c]);
                        r[q * bs + c] = std::fma(-f, r[k * bs + c], r[q * bs + c]);
                    }
                }
            }
            std::copy(r.begin(), r.end(), inv.begin() + i * b2);
        }
        ++updates;
    }
    void sweep(std::int64_t i, bool forward, const double* in, double* out)
    {
        const int b2 = bs * bs;
        std::vector<double> acc(bs), xj(bs);
        if (forward) {
            for (int r = 0; r < bs; ++r) acc[r] = in[i * bs + r];
        } else {
            const double s = omega * (2.0 - omega);
            for (int r = 0; r < bs; ++r) {
                double d = 0.0;
                for (int q = 0; q < bs; ++q) d = std::fma(dblk[i * b2 + r * bs + q], work[i * bs + q], d);
                acc[r] = d * s;
            }
        }
        const double* src = forward ? work.data() : out;
        for (std::int32_t e = a->off[i]; e < a->off[i + 1]; ++e) {
            const std::int64_t j = a->col[e];
            if (j >= a->owned) continue;
            if (forward ? !(colour[j] < colour[i]) : !(colour[j] > colour[i])) continue;
            const double* blk = a->val + static_cast<std::int64_t>(e) * b2;
            for (int q = 0; q < bs; ++q) xj[q] = src[j * bs + q];
            for (int r = 0; r < bs; ++r)
                for (int q = 0; q < bs; ++q) acc[r] = std::fma(-(omega * blk[r * bs + q]), xj[q], acc[r]);
        }
        double* dst = forward ? work.data() : out;
        for (int r = 0; r < bs; ++r) {
            double t = 0.0;
            for (int q = 0; q < bs; ++q) t = std::fma(inv[i * b2 + r * bs + q], acc[q], t);
            dst[i * bs + r] = t;
        }
    }
    void apply(const double* in, double* out) override
    {
        const std::int64_t n = a->owned;
        for (std::int64_t k = 0; k < n; ++k) sweep(order[k], true, in, out);
        for (std::int64_t k = n; k-- > 0;) sweep(order[k], false, in, out);
    }
};

// IluLevelPreconditioner (preconditioner.hpp:253-500) on the merged local CSR,
// optionally block-restricted (block-Jacobi ILU).
struct Ilu : Pc {
    const Matrix* a;
    std::int64_t block_nodes;
    std::size_t fill;
    std::int64_t n = 0;
    int bs = 1;
    static constexpr std::size_t no_source = std::numeric_limits<std::size_t>::max();
    std::vector<std::size_t> row_offsets, columns, diagonal_positions, source_positions;
    std::vector<double> values, inverse_diagonal;

    Ilu(const Matrix* m, std::int64_t blocks, int fill_level) : a(m), block_nodes(blocks), fill(fill_level)
    {
        const Csr c = local_csr(*a, block_nodes);
        n = c.n;
        bs = c.bs;
        build_pattern(c);
        factorize();
        updates = 1;
    }
    // build_local_pattern (preconditioner.hpp:373-427)
    void build_pattern(const Csr& c)
    {
        row_offsets.push_back(0);
        diagonal_positions.resize(n);
        std::vector<std::size_t> levels;
        for (std::int64_t row = 0; row < n; ++row) {
            bool diagonal_found = false;
            std::map<std::size_t, std::pair<std::size_t, std::size_t>> entries;
            for (std::int32_t s = c.off[row]; s < c.off[row + 1]; ++s) {
                const std::size_t column = c.col[s];
                if (!entries.emplace(column, std::pair{std::size_t(0), static_cast<std::size_t>(s)}).second)
                    throw std::invalid_argument("ILU requires unique matrix columns");
            }
            for (auto it = entries.begin(); it != entries.end() && it->first < static_cast<std::size_t>(row); ++it) {
                const auto pivot = it->first, lower_level = it->second.first;
                if (lower_level >= fill) continue;
                for (auto upper = diagonal_positions[pivot] + 1; upper < row_offsets[pivot + 1]; ++upper) {
                    if (levels[upper] > fill - lower_level - 1) continue;
                    const auto level = lower_level + levels[upper] + 1;
                    auto [target, inserted] = entries.emplace(columns[upper], std::pair{level, no_source});
                    if (!inserted) target->second.first = std::min(target->second.first, level);
                }
            }
            for (const auto& [column, entry] : entries) {
                const auto [level, source] = entry;
                const auto destination = columns.size();
                columns.push_back(column);
                levels.push_back(level);
                source_positions.push_back(source);
                if (source == no_source) values.insert(values.end(), bs, 0.0);
                else values.insert(values.end(), c.val.begin() + source * bs, c.val.begin() + source * bs + bs);
                if (column == static_cast<std::size_t>(row)) {
                    diagonal_positions[row] = destination;
                    diagonal_found = true;
                }
            }
            if (!diagonal_found) throw std::invalid_argument("ILU(0) row has no diagonal");
            row_offsets.push_back(columns.size());
        }
    }
    std::size_t find_entry(std::size_t row, std::size_t column) const
    {
        for (std::size_t e = row_offsets[row]; e < row_offsets[row + 1]; ++e)
            if (columns[e] == column) return e;
        return values.size();
    }
    // factorize (preconditioner.hpp:439-485)
    void factorize()
    {
        const double tol = 64 * std::numeric_limits<double>::epsilon();
        for (std::int64_t row = 0; row < n; ++row) {
            for (std::size_t lower = row_offsets[row]; lower < diagonal_positions[row]; ++lower) {
                const std::size_t pivot_row = columns[lower];
                double* lv = values.data() + lower * bs;
                const double* pivot = values.data() + diagonal_positions[pivot_row] * bs;
                for (int k = 0; k < bs; ++k) {
                    if (std::abs(pivot[k]) <= tol) throw std::runtime_error("zero pivot in ILU(0)");
                    lv[k] /= pivot[k];
                }
                for (std::size_t upper = diagonal_positions[pivot_row] + 1; upper < row_offsets[pivot_row + 1];
                     ++upper) {
                    const std::size_t target = find_entry(row, columns[upper]);
                    if (target == values.size()) continue;
                    double* tv = values.data() + target * bs;
                    const double* uv = values.data() + upper * bs;
                    for (int k = 0; k < bs; ++k) tv[k] = std::fma(-lv[k], uv[k], tv[k]);
                }
            }
        }
        inverse_diagonal.assign(n * bs, 0.0);
        for (std::int64_t row = 0; row < n; ++row) {
            const double* d = values.data() + diagonal_positions[row] * bs;
            for (int k = 0; k < bs; ++k) {
                if (!std::isfinite(d[k]) || std::abs(d[k]) <= tol) throw std::runtime_error("zero pivot in ILU(0)");
                inverse_diagonal[row * bs + k] = 1.0 / d[k];
            }
        }
    }
    // update_values (preconditioner.hpp:269-311): re-gather through the
    // recorded source positions of the (block-)restricted matrix.
    void update() override
    {
        const Csr c = local_csr(*a, block_nodes);
        for (std::size_t d = 0; d < columns.size(); ++d) {
            const std::size_t s = source_positions[d];
            if (s == no_source) std::fill_n(values.begin() + d * bs, bs, 0.0);
            else std::copy_n(c.val.begin() + s * bs, bs, values.begin() + d * bs);
        }
        factorize();
        ++updates;
    }
    // apply (preconditioner.hpp:320-363); in == out allowed.  All-cores
    // baseline: the independent block-Jacobi blocks in parallel, each swept
    // forward then backward in the same row order.
    void apply(const double* in, double* out) override
    {
        if (g_threads > 1 && block_nodes > 0 && n > block_nodes) {
            const std::int64_t blocks = (n + block_nodes - 1) / block_nodes;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
            for (std::int64_t k = 0; k < blocks; ++k)
                sweep(in, out, k * block_nodes, std::min(n, (k + 1) * block_nodes));
            return;
        }
        sweep(in, out, 0, n);
    }
    void sweep(const double* in, double* out, std::int64_t first, std::int64_t last) const
    {
        for (std::int64_t row = first; row < last; ++row) {
            double* y = out + row * bs;
            if (in != out) std::copy_n(in + row * bs, bs, y);
            for (std::size_t e = row_offsets[row]; e < diagonal_positions[row]; ++e) {
                const double* f = values.data() + e * bs;
                const double* prev = out + columns[e] * bs;
                for (int k = 0; k < bs; ++k) y[k] = std::fma(-f[k], prev[k], y[k]);
            }
        }
        for (std::int64_t row = last; row-- > first;) {
            double* x = out + row * bs;
            for (std::size_t e = diagonal_positions[row] + 1; e < row_offsets[row + 1]; ++e) {
                const double* f = values.data() + e * bs;
                const double* prev = out + columns[e] * bs;
                for (int k = 0; k < bs; ++k) x[k] = std::fma(-f[k], prev[k], x[k]);
            }
            const double* inv = inverse_diagonal.data() + row * bs;
            for (int k = 0; k < bs; ++k) x[k] *= inv[k];
        }
    }
};

// ----------------------------------------------------------------- solvers --
enum Status { converged = 0, maximum_iterations = 1, breakdown = 2, diverged = 3, invalid_input = 4 };
enum Reason {
    none = 0,
    arnoldi_invariant_subspace = 1,
    singular_projected_system = 2,
    biorthogonality_loss = 3,
    alpha_denominator = 4,
    omega_denominator = 5,
    omega_zero = 6,
    idr_shadow_space = 7,
    idr_small_system = 8,
    non_finite_scalar = 9,
};

} // namespace

extern "C" {

typedef struct oracle_options {
    double relative_tolerance, absolute_tolerance, breakdown_tolerance;
    std::uint64_t maximum_iterations, convergence_check_interval, restart, residual_replacement_interval;
    std::int32_t gmres_orthogonalization, collect_timings;
    std::uint64_t idr_shadow_space;
} oracle_options;

typedef struct oracle_result {
    std::int32_t status, breakdown_reason;
    std::uint64_t iterations, operator_applications, preconditioner_applications, global_reductions,
        reorthogonalizations, arnoldi_norm_verifications;
    double initial_residual_norm, recursive_residual_norm, true_residual_norm, relative_residual_norm;
    double solve_seconds;
    std::uint64_t refinement_steps, inner_iterations;
} oracle_result;

} // extern "C"

namespace {

struct Solve {
    const Matrix& a;
    Pc& pc;
    oracle_result& r;
    std::int64_t n;  // owned size
    std::int64_t local;
    Solve(const Matrix& m, Pc& p, oracle_result& res) : a(m), pc(p), r(res), n(m.owned * m.bs), local(m.local() * m.bs) {}
    // recycled_fgmres's RecycledOperator (recycling.hpp:195-220): right after a
    // projected preconditioner apply, the operator returns its cached image
    double* cached = nullptr;
    bool* cache_ready = nullptr;
    void op(const double* x, double* y)
    {
        ++r.operator_applications;
        if (cache_ready && *cache_ready) {
            std::copy_n(cached, n, y);
            *cache_ready = false;
            return;
        }
        apply(a, x, y);
    }
    void prec(const double* in, double* out)
    {
        pc.apply(in, out);
        ++r.preconditioner_applications;
    }
    double d(const double* x, const double* y)
    {
        ++r.global_reductions;
        return dot(n, x, y);
    }
    double nrm(const double* x)
    {
        ++r.global_reductions;
        return norm(n, x);
    }
    // detail::norm_from_squared (krylov_solvers.hpp:95-108): the fused
    // squared norm unless it over/underflowed, then an extra scaled norm
    double nfs(const double* x, double squared)
    {
        if (std::isfinite(squared) && squared >= std::numeric_limits<double>::min()) return std::sqrt(squared);
        return nrm(x);
    }
    // detail::true_residual (krylov_solvers.hpp:62-74)
    void true_residual(const double* b, const double* x, double* res, double* work)
    {
        op(x, work);
        OMP_ROWS
        for (std::int64_t i = 0; i < n; ++i) res[i] = b[i] - work[i];
    }
    void set_residual(double recursive, double truen, double rhs_norm)
    {
        r.recursive_residual_norm = recursive;
        r.true_residual_norm = truen;
        r.relative_residual_norm = truen / (rhs_norm > 0 ? rhs_norm : 1.0);
    }
};

bool valid_options(const oracle_options& o)
{
    return std::isfinite(o.relative_tolerance) && std::isfinite(o.absolute_tolerance)
        && std::isfinite(o.breakdown_tolerance) && o.relative_tolerance >= 0 && o.absolute_tolerance >= 0
        && o.breakdown_tolerance > 0 && o.maximum_iterations > 0 && o.convergence_check_interval > 0;
}

// convergence_threshold (core.hpp:549-560)
double threshold_of(double rhs_norm, const oracle_options& o)
{
    const double rel = o.relative_tolerance * (rhs_norm > 0 ? rhs_norm : 1.0);
    if (!std::isfinite(rhs_norm) || !std::isfinite(rel)) return std::numeric_limits<double>::quiet_NaN();
    return std::max(o.absolute_tolerance, rel);
}

void init_result(oracle_result& r)
{
    std::memset(&r, 0, sizeof(r));
    r.status = invalid_input;
    r.initial_residual_norm = r.recursive_residual_norm = r.true_residual_norm = r.relative_residual_norm =
        std::numeric_limits<double>::quiet_NaN();
}

// pipelined_bicgstab (pipelined_bicgstab.hpp:20-244): fused scalar products
// (<shadow, v>; <t, s>, <t, t>, <s, s>; <shadow, r>, <r, r>), a convergence
// test on the recursive residual every iteration, and residual replacement
// every residual_replacement_interval iterations (which only restarts).
void pipelined_bicgstab(const Matrix& a, Pc& pc, const double* b, double* x, const oracle_options& o,
                        oracle_result& r)
{
    init_result(r);
    if (!valid_options(o)) return;
    Solve s(a, pc, r);
    const std::int64_t L = s.local;
    const std::int64_t n = s.n;
    std::vector<double> residual(L), shadow(L), search(L), osearch(L), inter(L), ointer(L), psearch(L), pinter(L),
        work(L);
    auto mark = [&](Reason reason) {
        r.status = breakdown;
        r.breakdown_reason = reason;
    };
    s.true_residual(b, x, residual.data(), work.data());
    std::copy_n(residual.begin(), n, shadow.begin());
    std::copy_n(residual.begin(), n, search.begin());
    // one fused reduction: |b|^2, |r|^2
    ++r.global_reductions;
    const double rhs_norm = s.nfs(b, dot(n, b, b));
    double rnorm = s.nfs(residual.data(), dot(n, residual.data(), residual.data()));
    const double thr = threshold_of(rhs_norm, o);
    const double tiny = 1024 * std::numeric_limits<double>::min();
    r.initial_residual_norm = rnorm;
    s.set_residual(rnorm, rnorm, rhs_norm);
    if (!std::isfinite(rnorm) || !std::isfinite(thr)) {
        mark(non_finite_scalar);
        return;
    }
    if (rnorm <= thr) {
        r.status = converged;
        return;
    }
    double rho = s.d(shadow.data(), residual.data());
    double alpha = 1, omega = 1;
    auto restart = [&] {
        std::copy_n(residual.begin(), n, shadow.begin());
        std::copy_n(residual.begin(), n, search.begin());
        rho = s.d(shadow.data(), residual.data());
        alpha = omega = 1;
    };
    for (std::uint64_t it = 1; it <= o.maximum_iterations; ++it) {
        r.iterations = it;
        s.prec(search.data(), psearch.data());
        s.op(psearch.data(), osearch.data());
        const double sv = s.d(shadow.data(), osearch.data());
        if (!std::isfinite(sv) || std::abs(sv) <= tiny) {
            mark(alpha_denominator);
            return;
        }
        alpha = rho / sv;
        for (std::int64_t i = 0; i < n; ++i) inter[i] = residual[i] - alpha * osearch[i];
        s.prec(inter.data(), pinter.data());
        s.op(pinter.data(), ointer.data());
        ++r.global_reductions;  // one fused reduction of three products
        const double num = dot(n, ointer.data(), inter.data());
        const double den = dot(n, ointer.data(), ointer.data());
        const double inorm = s.nfs(inter.data(), dot(n, inter.data(), inter.data()));
        if (!std::isfinite(inorm)) {
            mark(non_finite_scalar);
            return;
        }
        if (inorm <= thr || !std::isfinite(den) || std::abs(den) <= tiny) {
            if (inorm <= thr) {
                for (std::int64_t i = 0; i < n; ++i) x[i] = std::fma(alpha, psearch[i], x[i]);
                s.true_residual(b, x, residual.data(), work.data());
                const double tn = s.nrm(residual.data());
                s.set_residual(inorm, tn, rhs_norm);
                if (tn <= thr) {
                    r.status = converged;
                    return;
                }
                restart();
                continue;
            }
            mark(omega_denominator);
            return;
        }
        omega = num / den;
        const double angle = std::abs(num) / std::sqrt(den) / inorm;
        if (!std::isfinite(omega) || omega == 0.0 || angle <= o.breakdown_tolerance) {
            mark(omega_zero);
            return;
        }
        for (std::int64_t i = 0; i < n; ++i) {
            x[i] = std::fma(alpha, psearch[i], x[i]);
            x[i] = std::fma(omega, pinter[i], x[i]);
            residual[i] = inter[i] - omega * ointer[i];
        }
        ++r.global_reductions;  // <shadow, r>, <r, r> fused
        const double prev = rho;
        rho = dot(n, shadow.data(), residual.data());
        rnorm = s.nfs(residual.data(), dot(n, residual.data(), residual.data()));
        r.recursive_residual_norm = rnorm;
        if (rnorm <= thr) {
            s.true_residual(b, x, residual.data(), work.data());
            const double tn = s.nrm(residual.data());
            s.set_residual(rnorm, tn, rhs_norm);
            if (tn <= thr) {
                r.status = converged;
                return;
            }
            restart();
            continue;
        }
        if (!std::isfinite(prev) || std::abs(prev) <= tiny) {
            mark(biorthogonality_loss);
            return;
        }
        const double beta = (rho / prev) * (alpha / omega);
        for (std::int64_t i = 0; i < n; ++i) search[i] = residual[i] + beta * (search[i] - omega * osearch[i]);
        if (o.residual_replacement_interval > 0 && it % o.residual_replacement_interval == 0) {
            s.true_residual(b, x, residual.data(), work.data());
            restart();
        }
    }
    s.true_residual(b, x, residual.data(), work.data());
    const double tn = s.nrm(residual.data());
    s.set_residual(rnorm, tn, rhs_norm);
    r.status = tn <= thr ? converged : maximum_iterations;
    if (!std::isfinite(tn)) mark(non_finite_scalar);
}

// communication_hiding_bicgstab (pipelined_bicgstab.hpp:317-597): the
// 15-vector Cools/PETSc recurrence with two fused reductions per iteration
// (<q,y>, <y,y>, <q,q>; <r,r>, <r,rp>, <s,rp>, <w,rp>, <z,rp>), each followed
// by one preconditioner + operator application, and a restart of the
// recurrence (two reductions, two applications) after a false convergence or
// a residual replacement.
void communication_hiding_bicgstab(const Matrix& a, Pc& pc, const double* b, double* x, const oracle_options& o,
                                   oracle_result& res)
{
    init_result(res);
    if (!valid_options(o)) return;
    Solve sv(a, pc, res);
    const std::int64_t L = sv.local;
    const std::int64_t n = sv.n;
    std::vector<double> r(L), rp(L), s(L), s2(L), p2(L), r2(L), w(L), w2(L), z(L), z2(L), t(L), v(L), q(L), q2(L),
        y(L), work(L);
    auto mark = [&](Reason reason) {
        res.status = breakdown;
        res.breakdown_reason = reason;
    };
    sv.true_residual(b, x, r.data(), work.data());
    std::copy_n(r.begin(), n, rp.begin());
    ++res.global_reductions;  // |b|^2, |r|^2 fused
    const double rhs_norm = sv.nfs(b, dot(n, b, b));
    double rnorm = sv.nfs(r.data(), dot(n, r.data(), r.data()));
    const double thr = threshold_of(rhs_norm, o);
    const double tiny = 1024 * std::numeric_limits<double>::min();
    res.initial_residual_norm = rnorm;
    sv.set_residual(rnorm, rnorm, rhs_norm);
    if (!std::isfinite(rnorm) || !std::isfinite(thr)) {
        mark(non_finite_scalar);
        return;
    }
    if (rnorm <= thr) {
        res.status = converged;
        return;
    }
    double rho = 0, alpha = 0, beta = 0, omega = 0;
    auto initialize = [&]() -> bool {
        std::copy_n(r.begin(), n, rp.begin());
        rho = sv.d(r.data(), rp.data());
        if (!std::isfinite(rho) || std::abs(rho) <= tiny) return false;
        sv.prec(r.data(), r2.data());
        sv.op(r2.data(), w.data());
        const double d2 = sv.d(w.data(), rp.data());
        if (!std::isfinite(d2) || std::abs(d2) <= tiny) return false;
        sv.prec(w.data(), w2.data());
        sv.op(w2.data(), t.data());
        alpha = rho / d2;
        beta = 0;
        omega = 0;
        return true;
    };
    if (!initialize()) {
        mark(biorthogonality_loss);
        return;
    }
    bool first = true;
    for (std::uint64_t it = 1; it <= o.maximum_iterations; ++it) {
        res.iterations = it;
        if (first) {
            std::copy_n(r2.begin(), n, p2.begin());
            std::copy_n(w.begin(), n, s.begin());
            std::copy_n(w2.begin(), n, s2.begin());
            std::copy_n(t.begin(), n, z.begin());
            first = false;
        } else {
            for (std::int64_t i = 0; i < n; ++i) {
                p2[i] = r2[i] + beta * (p2[i] - omega * s2[i]);
                s[i] = w[i] + beta * (s[i] - omega * z[i]);
                s2[i] = w2[i] + beta * (s2[i] - omega * z2[i]);
                z[i] = t[i] + beta * (z[i] - omega * v[i]);
            }
        }
        for (std::int64_t i = 0; i < n; ++i) {
            q[i] = r[i] - alpha * s[i];
            q2[i] = r2[i] - alpha * s2[i];
            y[i] = w[i] - alpha * z[i];
        }
        ++res.global_reductions;  // <q,y>, <y,y>, <q,q> fused
        const double qy = dot(n, q.data(), y.data());
        const double yy = dot(n, y.data(), y.data());
        const double qq = dot(n, q.data(), q.data());
        sv.prec(z.data(), z2.data());
        sv.op(z2.data(), v.data());
        const double inorm = sv.nfs(q.data(), qq);
        if (!std::isfinite(inorm)) {
            mark(non_finite_scalar);
            return;
        }
        if (inorm <= thr || !std::isfinite(yy) || std::abs(yy) <= tiny) {
            if (inorm <= thr) {
                for (std::int64_t i = 0; i < n; ++i) x[i] = std::fma(alpha, p2[i], x[i]);
                sv.true_residual(b, x, r.data(), work.data());
                const double tn = sv.nrm(r.data());
                sv.set_residual(inorm, tn, rhs_norm);
                if (tn <= thr) {
                    res.status = converged;
                    return;
                }
                if (!initialize()) {
                    mark(biorthogonality_loss);
                    return;
                }
                first = true;
                continue;
            }
            mark(std::isfinite(yy) ? omega_denominator : non_finite_scalar);
            return;
        }
        omega = qy / yy;
        const double angle = std::abs(qy) / std::sqrt(yy) / inorm;
        if (!std::isfinite(omega) || omega == 0.0 || angle <= o.breakdown_tolerance) {
            mark(omega_zero);
            return;
        }
        for (std::int64_t i = 0; i < n; ++i) x[i] = std::fma(alpha, p2[i], x[i]);
        for (std::int64_t i = 0; i < n; ++i) x[i] = std::fma(omega, q2[i], x[i]);
        for (std::int64_t i = 0; i < n; ++i) {
            r[i] = q[i] - omega * y[i];
            r2[i] = q2[i] - omega * (w2[i] - alpha * z2[i]);
            w[i] = y[i] - omega * (t[i] - alpha * v[i]);
        }
        const double prev = rho;
        ++res.global_reductions;  // five products fused
        const double rr = dot(n, r.data(), r.data());
        rho = dot(n, r.data(), rp.data());
        const double srp = dot(n, s.data(), rp.data());
        const double wrp = dot(n, w.data(), rp.data());
        const double zrp = dot(n, z.data(), rp.data());
        sv.prec(w.data(), w2.data());
        sv.op(w2.data(), t.data());
        rnorm = sv.nfs(r.data(), rr);
        res.recursive_residual_norm = rnorm;
        if (rnorm <= thr) {
            sv.true_residual(b, x, r.data(), work.data());
            const double tn = sv.nrm(r.data());
            sv.set_residual(rnorm, tn, rhs_norm);
            if (tn <= thr) {
                res.status = converged;
                return;
            }
            if (!initialize()) {
                mark(biorthogonality_loss);
                return;
            }
            first = true;
            continue;
        }
        if (o.residual_replacement_interval > 0 && it % o.residual_replacement_interval == 0) {
            sv.true_residual(b, x, r.data(), work.data());
            if (!initialize()) {
                mark(biorthogonality_loss);
                return;
            }
            first = true;
            continue;
        }
        if (!std::isfinite(rho) || !std::isfinite(prev) || std::abs(rho) <= tiny || std::abs(prev) <= tiny) {
            mark(std::isfinite(rho) && std::isfinite(prev) ? biorthogonality_loss : non_finite_scalar);
            return;
        }
        beta = (rho / prev) * (alpha / omega);
        const double den = wrp + beta * srp - beta * omega * zrp;
        if (!std::isfinite(den) || std::abs(den) <= tiny) {
            mark(std::isfinite(den) ? alpha_denominator : non_finite_scalar);
            return;
        }
        alpha = rho / den;
    }
    sv.true_residual(b, x, r.data(), work.data());
    const double tn = sv.nrm(r.data());
    sv.set_residual(rnorm, tn, rhs_norm);
    res.status = tn <= thr ? converged : maximum_iterations;
    if (!std::isfinite(tn)) mark(non_finite_scalar);
}

// bicgstab (krylov_solvers.hpp:542-745)
void bicgstab(const Matrix& a, Pc& pc, const double* b, double* x, const oracle_options& o, oracle_result& r)
{
    init_result(r);
    if (!valid_options(o)) return;
    Solve s(a, pc, r);
    const std::int64_t L = s.local;
    std::vector<double> residual(L), shadow(L), search(L), psearch(L), osearch(L), inter(L), pinter(L), ointer(L),
        work(L);
    auto mark = [&](Reason reason) {
        r.status = breakdown;
        r.breakdown_reason = reason;
    };
    s.true_residual(b, x, residual.data(), work.data());
    std::copy_n(residual.begin(), s.n, shadow.begin());
    const double rhs_norm = s.nrm(b);
    const double initial = s.nrm(residual.data());
    const double thr = threshold_of(rhs_norm, o);
    const double tiny = 1024 * std::numeric_limits<double>::min();
    r.initial_residual_norm = initial;
    s.set_residual(initial, initial, rhs_norm);
    if (!std::isfinite(initial) || !std::isfinite(thr)) {
        mark(non_finite_scalar);
        return;
    }
    if (initial <= thr) {
        r.status = converged;
        return;
    }
    double rho_prev = 1, alpha = 1, omega = 1;
    bool first = true;
    auto restart = [&] {
        std::copy_n(residual.begin(), s.n, shadow.begin());
        std::fill_n(search.begin(), s.n, 0.0);
        std::fill_n(osearch.begin(), s.n, 0.0);
        rho_prev = alpha = omega = 1;
        first = true;
    };
    for (std::uint64_t it = 1; it <= o.maximum_iterations; ++it) {
        r.iterations = it;
        const double rho = s.d(shadow.data(), residual.data());
        if (!std::isfinite(rho) || std::abs(rho) <= tiny) {
            mark(std::isfinite(rho) ? biorthogonality_loss : non_finite_scalar);
            return;
        }
        if (first) {
            std::copy_n(residual.begin(), s.n, search.begin());
            first = false;
        } else {
            if (std::abs(omega) <= tiny) {
                mark(omega_zero);
                return;
            }
            const double beta = (rho / rho_prev) * (alpha / omega);
            OMP_ROWS
            for (std::int64_t i = 0; i < s.n; ++i)
                search[i] = std::fma(beta, std::fma(-omega, osearch[i], search[i]), residual[i]);
        }
        s.prec(search.data(), psearch.data());
        s.op(psearch.data(), osearch.data());
        const double sv = s.d(shadow.data(), osearch.data());
        if (!std::isfinite(sv) || std::abs(sv) <= tiny) {
            mark(std::isfinite(sv) ? alpha_denominator : non_finite_scalar);
            return;
        }
        alpha = rho / sv;
        OMP_ROWS
        for (std::int64_t i = 0; i < s.n; ++i) inter[i] = std::fma(-alpha, osearch[i], residual[i]);
        const double inorm = s.nrm(inter.data());
        if (inorm <= thr) {
            OMP_ROWS
            for (std::int64_t i = 0; i < s.n; ++i) x[i] = std::fma(alpha, psearch[i], x[i]);
            s.true_residual(b, x, residual.data(), work.data());
            const double tn = s.nrm(residual.data());
            s.set_residual(inorm, tn, rhs_norm);
            if (tn <= thr) {
                r.status = converged;
                return;
            }
            restart();
            continue;
        }
        s.prec(inter.data(), pinter.data());
        s.op(pinter.data(), ointer.data());
        const double num = s.d(ointer.data(), inter.data());
        const double den = s.d(ointer.data(), ointer.data());
        if (!std::isfinite(den) || std::abs(den) <= tiny) {
            mark(std::isfinite(den) ? omega_denominator : non_finite_scalar);
            return;
        }
        omega = num / den;
        const double angle = std::abs(num) / std::sqrt(den) / inorm;
        if (!std::isfinite(omega) || omega == 0.0 || angle <= o.breakdown_tolerance) {
            mark(std::isfinite(omega) ? omega_zero : non_finite_scalar);
            return;
        }
        OMP_ROWS
        for (std::int64_t i = 0; i < s.n; ++i) x[i] = std::fma(alpha, psearch[i], x[i]);
        OMP_ROWS
        for (std::int64_t i = 0; i < s.n; ++i) x[i] = std::fma(omega, pinter[i], x[i]);
        OMP_ROWS
        for (std::int64_t i = 0; i < s.n; ++i) residual[i] = std::fma(-omega, ointer[i], inter[i]);
        const bool replace = o.residual_replacement_interval > 0 && it % o.residual_replacement_interval == 0;
        const bool check = it % o.convergence_check_interval == 0;
        if (replace || check) {
            const double rn = s.nrm(residual.data());
            if (replace || rn <= thr) {
                s.true_residual(b, x, residual.data(), work.data());
                const double tn = s.nrm(residual.data());
                s.set_residual(rn, tn, rhs_norm);
                if (tn <= thr) {
                    r.status = converged;
                    return;
                }
                restart();
                continue;
            } else {
                r.recursive_residual_norm = rn;
            }
        }
        rho_prev = rho;
    }
    s.true_residual(b, x, residual.data(), work.data());
    const double tn = s.nrm(residual.data());
    s.set_residual(r.recursive_residual_norm, tn, rhs_norm);
    r.status = tn <= thr ? converged : maximum_iterations;
    if (!std::isfinite(tn)) mark(non_finite_scalar);
}

// gmres_update (krylov_solvers.hpp:23-40)
void gmres_update(std::int64_t n, const double* base, const std::vector<std::vector<double>>& dirs,
                  const std::vector<double>& coef, std::size_t count, double* out)
{
    for (std::int64_t i = 0; i < n; ++i) {
        double sum = 0, comp = 0;
        for (std::size_t j = 0; j < count; ++j) {
            const double term = coef[j] * dirs[j][i];
            const double updated = sum + term;
            comp += std::abs(sum) >= std::abs(term) ? (sum - updated) + term : (term - updated) + sum;
            sum = updated;
        }
        out[i] = base[i] + (sum + comp);
    }
}

// solve_dense_system (idrs.hpp:34-77): Gaussian elimination with partial
// pivoting, false when a pivot is non-finite or at most `tol`.
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

// initialize_idr_shadow (idrs.hpp:79-118), serial reduction (rank 0): the
// draws of std::mt19937_64 seeded 42 + epoch * 0x9e3779b97f4a7c15 through
// std::normal_distribution<double>, Gram-Schmidt orthonormalised; with s = 1
// the residual itself, normalised.
bool idr_shadow(std::vector<std::vector<double>>& shadow, const double* residual, std::int64_t n, Solve& sv,
                double tol, std::uint64_t epoch)
{
    std::mt19937_64 gen(42ULL + epoch * 0x9e3779b97f4a7c15ULL);
    std::normal_distribution<double> dist(0.0, 1.0);
    const std::size_t first = shadow.size() == 1 ? 1 : 0;
    if (shadow.size() == 1) {
        std::copy_n(residual, n, shadow[0].begin());
        const double nv = sv.nrm(shadow[0].data());
        if (nv <= tol) return false;
        for (std::int64_t i = 0; i < n; ++i) shadow[0][i] *= 1.0 / nv;
    }
    for (std::size_t k = first; k < shadow.size(); ++k) {
        for (std::int64_t i = 0; i < n; ++i) shadow[k][i] = dist(gen);
        for (std::size_t p = 0; p < k; ++p) {
            const double proj = sv.d(shadow[p].data(), shadow[k].data());
            for (std::int64_t i = 0; i < n; ++i) shadow[k][i] = std::fma(-proj, shadow[p][i], shadow[k][i]);
        }
        const double nv = sv.nrm(shadow[k].data());
        if (nv <= tol) return false;
        const double inv = 1.0 / nv;
        for (std::int64_t i = 0; i < n; ++i) shadow[k][i] *= inv;
    }
    return true;
}

void bicgstab(const Matrix& a, Pc& pc, const double* b, double* x, const oracle_options& o, oracle_result& r);

// idrs (idrs.hpp:128-505): right-preconditioned IDR(s), serial reduction.
void idrs(const Matrix& a, Pc& pc, const double* b, double* x, const oracle_options& o, oracle_result& r)
{
    init_result(r);
    const std::size_t s = static_cast<std::size_t>(o.idr_shadow_space);
    const std::int64_t n = a.owned * a.bs;
    if (!valid_options(o) || s == 0 || s > static_cast<std::size_t>(n)) return;
    if (s == 1) {  // IDR(1) is BiCGSTAB's recurrence
        bicgstab(a, pc, b, x, o, r);
        return;
    }
    Solve sv(a, pc, r);
    const std::int64_t L = sv.local;
    std::vector<double> residual(L), projected(L), pre(L), work(L), tmp(L);
    std::vector<std::vector<double>> G(s, std::vector<double>(L)), Z(s, std::vector<double>(L)),
        P(s, std::vector<double>(L));
    auto mark = [&](Reason reason) {
        r.status = breakdown;
        r.breakdown_reason = reason;
    };
    auto axpy = [&](double alpha, const std::vector<double>& xv, std::vector<double>& yv) {
        for (std::int64_t i = 0; i < n; ++i) yv[i] = std::fma(alpha, xv[i], yv[i]);
    };
    auto axpy_x = [&](double alpha, const std::vector<double>& xv) {
        for (std::int64_t i = 0; i < n; ++i) x[i] = std::fma(alpha, xv[i], x[i]);
    };
    sv.true_residual(b, x, residual.data(), tmp.data());
    const double rhs_norm = sv.nrm(b);
    double rnorm = sv.nrm(residual.data());
    const double thr = threshold_of(rhs_norm, o);
    r.initial_residual_norm = rnorm;
    sv.set_residual(rnorm, rnorm, rhs_norm);
    if (!std::isfinite(rnorm) || !std::isfinite(thr)) {
        mark(non_finite_scalar);
        return;
    }
    if (rnorm <= thr) {
        r.status = converged;
        return;
    }
    std::vector<double> M(s * s, 0.0), f(s, 0.0), am, arhs, c;
    std::size_t epoch = 0, restarts = 0;
    double omega = 1.0;
    auto restart_basis = [&]() -> bool {
        if (!idr_shadow(P, residual.data(), n, sv, o.breakdown_tolerance, epoch++)) return false;
        std::fill(M.begin(), M.end(), 0.0);
        for (std::size_t k = 0; k < s; ++k) {
            M[k * s + k] = 1.0;
            std::fill_n(G[k].begin(), n, 0.0);
            std::fill_n(Z[k].begin(), n, 0.0);
            f[k] = sv.d(P[k].data(), residual.data());
        }
        omega = 1.0;
        return true;
    };
    auto restart_or_fail = [&]() -> bool {
        ++restarts;
        return restarts <= 4 && restart_basis();
    };
    auto pivot_small = [&](std::size_t col) -> bool {
        double scale = 0.0;
        for (std::size_t row = col; row < s; ++row) scale = std::max(scale, std::abs(M[row * s + col]));
        const double piv = M[col * s + col];
        const double tol = std::max(std::numeric_limits<double>::min(), o.breakdown_tolerance * scale);
        return !std::isfinite(piv) || std::abs(piv) <= tol;
    };
    if (!restart_basis()) {
        mark(idr_shadow_space);
        return;
    }
    while (r.iterations < o.maximum_iterations) {
        bool restarted = false;
        for (std::size_t k = 0; k < s && r.iterations < o.maximum_iterations; ++k) {
            const std::size_t act = s - k;
            am.assign(act * act, 0.0);
            arhs.assign(act, 0.0);
            double mscale = 0.0;
            for (std::size_t row = 0; row < act; ++row) {
                arhs[row] = f[k + row];
                for (std::size_t col = 0; col < act; ++col) {
                    const double v = M[(k + row) * s + (k + col)];
                    am[row * act + col] = v;
                    mscale = std::max(mscale, std::abs(v));
                }
            }
            const double dtol = std::max(std::numeric_limits<double>::min(),
                                         64 * std::numeric_limits<double>::epsilon() * mscale);
            if (!solve_dense(am, arhs, c, dtol)) {
                if (!restart_or_fail()) {
                    mark(idr_small_system);
                    return;
                }
                restarted = true;
                break;
            }
            // v = r - G_k:s c; z = omega M^-1 v + Z_k:s c
            std::copy_n(residual.begin(), n, projected.begin());
            for (std::size_t col = k; col < s; ++col) axpy(-c[col - k], G[col], projected);
            sv.prec(projected.data(), pre.data());
            for (std::int64_t i = 0; i < n; ++i) pre[i] *= omega;
            for (std::size_t col = k; col < s; ++col) axpy(c[col - k], Z[col], pre);
            sv.op(pre.data(), work.data());
            for (std::size_t prev = 0; prev < k; ++prev) {
                const double piv = M[prev * s + prev];
                if (pivot_small(prev)) {
                    if (!restart_or_fail()) {
                        mark(idr_small_system);
                        return;
                    }
                    restarted = true;
                    break;
                }
                const double alpha = sv.d(P[prev].data(), work.data()) / piv;
                axpy(-alpha, G[prev], work);
                axpy(-alpha, Z[prev], pre);
            }
            if (restarted) break;
            for (std::size_t row = 0; row < k; ++row) M[row * s + k] = 0.0;
            for (std::size_t row = k; row < s; ++row) M[row * s + k] = sv.d(P[row].data(), work.data());
            const double piv = M[k * s + k];
            if (pivot_small(k)) {
                if (!restart_or_fail()) {
                    mark(idr_small_system);
                    return;
                }
                restarted = true;
                break;
            }
            std::copy_n(work.begin(), n, G[k].begin());
            std::copy_n(pre.begin(), n, Z[k].begin());
            const double beta = f[k] / piv;
            axpy_x(beta, Z[k]);
            axpy(-beta, G[k], residual);
            for (std::size_t row = k + 1; row < s; ++row) f[row] -= beta * M[row * s + k];
            f[k] = 0.0;
            ++r.iterations;
            restarts = 0;
            rnorm = sv.nrm(residual.data());
            r.recursive_residual_norm = rnorm;
            if (!std::isfinite(rnorm)) {
                mark(non_finite_scalar);
                return;
            }
            if (rnorm <= thr) {
                sv.true_residual(b, x, residual.data(), tmp.data());
                const double tn = sv.nrm(residual.data());
                sv.set_residual(rnorm, tn, rhs_norm);
                if (tn <= thr) {
                    r.status = converged;
                    return;
                }
                rnorm = tn;
                if (!restart_or_fail()) {
                    r.status = diverged;
                    return;
                }
                restarted = true;
                break;
            }
            if (o.residual_replacement_interval > 0 && r.iterations % o.residual_replacement_interval == 0) {
                sv.true_residual(b, x, residual.data(), tmp.data());
                rnorm = sv.nrm(residual.data());
                sv.set_residual(rnorm, rnorm, rhs_norm);
                if (rnorm <= thr) {
                    r.status = converged;
                    return;
                }
                if (!restart_basis()) {
                    mark(idr_shadow_space);
                    return;
                }
                restarted = true;
                break;
            }
        }
        if (restarted) continue;
        if (r.iterations >= o.maximum_iterations) break;
        // stabilising minimum-residual step, angle-guarded (omega >= 0.7 of the angle)
        sv.prec(residual.data(), pre.data());
        sv.op(pre.data(), work.data());
        const double num = sv.d(work.data(), residual.data());
        const double den = sv.d(work.data(), work.data());
        if (!std::isfinite(den) || den <= std::numeric_limits<double>::min()) {
            mark(omega_denominator);
            return;
        }
        omega = num / den;
        const double adn = std::sqrt(std::max(0.0, den)) * std::max(rnorm, std::numeric_limits<double>::min());
        const double angle = adn > 0 ? std::abs(num) / adn : 1.0;
        constexpr double min_angle = 0.7;
        if (angle > 0 && angle < min_angle) omega *= min_angle / angle;
        if (!std::isfinite(omega) || std::abs(omega) <= std::numeric_limits<double>::min()) {
            mark(omega_denominator);
            return;
        }
        axpy_x(omega, pre);
        axpy(-omega, work, residual);
        ++r.iterations;
        restarts = 0;
        rnorm = sv.nrm(residual.data());
        r.recursive_residual_norm = rnorm;
        if (!std::isfinite(rnorm)) {
            mark(non_finite_scalar);
            return;
        }
        if (rnorm <= thr) {
            sv.true_residual(b, x, residual.data(), tmp.data());
            const double tn = sv.nrm(residual.data());
            sv.set_residual(rnorm, tn, rhs_norm);
            if (tn <= thr) {
                r.status = converged;
                return;
            }
            rnorm = tn;
            if (!restart_or_fail()) {
                r.status = diverged;
                return;
            }
            continue;
        }
        if (o.residual_replacement_interval > 0 && r.iterations % o.residual_replacement_interval == 0) {
            sv.true_residual(b, x, residual.data(), tmp.data());
            rnorm = sv.nrm(residual.data());
            sv.set_residual(rnorm, rnorm, rhs_norm);
            if (rnorm <= thr) {
                r.status = converged;
                return;
            }
            if (!restart_basis()) {
                mark(idr_shadow_space);
                return;
            }
            continue;
        }
        for (std::size_t row = 0; row < s; ++row) f[row] = sv.d(P[row].data(), residual.data());
    }
    sv.true_residual(b, x, residual.data(), tmp.data());
    const double tn = sv.nrm(residual.data());
    sv.set_residual(rnorm, tn, rhs_norm);
    r.status = maximum_iterations;
}

// AggregationMultigrid (multigrid.hpp:41-421): greedy aggregates, the
// row-averaged coarse matrix, Jacobi / red-black Gauss-Seidel / Chebyshev
// smoothing, V, W and full cycles.  The fine matrix is a component-diagonal
// BlockCsrMatrix (kind CSR).
struct MgOptions {
    std::uint64_t pre, post, coarse_steps, aggregate;
    double jacobi_weight;
    std::int32_t smoother;  // 0 jacobi, 1 red-black Gauss-Seidel, 2 Chebyshev
    std::uint64_t chebyshev_levels;
};

struct Multigrid {
    Matrix fine;
    MgOptions opt;
    int bs;
    std::int64_t n;
    std::vector<std::int64_t> f2c, sizes;
    std::vector<std::int32_t> coff, ccol;  // coarse matrix, CSR kind
    std::vector<double> cval;
    Matrix coarse;
    std::vector<double> fine_inv, coarse_inv;
    std::vector<std::int64_t> fine_red, fine_black, coarse_red, coarse_black;

    Multigrid(const Matrix& a, const MgOptions& o) : fine(a), opt(o), bs(a.bs), n(a.owned)
    {
        if (a.kind != 0) throw std::invalid_argument("multigrid needs a component-diagonal BlockCsrMatrix");
        if (o.aggregate == 0) throw std::invalid_argument("multigrid aggregate size must be positive");
        // build_aggregates (multigrid.hpp:237-262)
        constexpr std::int64_t none = -1;
        f2c.assign(n, none);
        for (std::int64_t seed = 0; seed < n; ++seed) {
            if (f2c[seed] != none) continue;
            const std::int64_t agg = static_cast<std::int64_t>(sizes.size());
            std::int64_t size = 1;
            f2c[seed] = agg;
            for (std::int32_t e = a.off[seed]; e < a.off[seed + 1] && static_cast<std::uint64_t>(size) < o.aggregate;
                 ++e) {
                const std::int64_t nb = a.col[e];
                if (nb < n && f2c[nb] == none) {
                    f2c[nb] = agg;
                    ++size;
                }
            }
            sizes.push_back(size);
        }
        // build_coarse_matrix (multigrid.hpp:264-305)
        const std::int64_t nc = static_cast<std::int64_t>(sizes.size());
        std::vector<std::map<std::int64_t, std::vector<double>>> rows(nc);
        for (std::int64_t r = 0; r < n; ++r) {
            const std::int64_t cr = f2c[r];
            for (std::int32_t e = a.off[r]; e < a.off[r + 1]; ++e) {
                const std::int64_t fc = a.col[e];
                if (fc >= n) continue;
                auto [pos, inserted] = rows[cr].try_emplace(f2c[fc], bs, 0.0);
                const double* coef = a.val + static_cast<std::int64_t>(e) * bs;
                for (int c = 0; c < bs; ++c) pos->second[c] += coef[c] / static_cast<double>(sizes[cr]);
            }
        }
        coff.assign(nc + 1, 0);
        for (std::int64_t r = 0; r < nc; ++r) {
            for (const auto& [col, block] : rows[r]) {
                ccol.push_back(static_cast<std::int32_t>(col));
                cval.insert(cval.end(), block.begin(), block.end());
            }
            coff[r + 1] = static_cast<std::int32_t>(ccol.size());
        }
        coarse = Matrix{0, nc, 0, bs, coff.data(), ccol.data(), cval.data(), nullptr};
        fine_inv = inverse_diagonal(fine);
        coarse_inv = inverse_diagonal(coarse);
        coloring(fine, fine_red, fine_black);
        coloring(coarse, coarse_red, coarse_black);
    }
    // BlockCsrMatrix::diagonal, inverted (multigrid.hpp:223-235)
    static std::vector<double> inverse_diagonal(const Matrix& m)
    {
        std::vector<double> d(static_cast<std::size_t>(m.owned) * m.bs, 0.0);
        for (std::int64_t r = 0; r < m.owned; ++r)
            for (std::int32_t e = m.off[r]; e < m.off[r + 1]; ++e)
                if (m.col[e] == r)
                    for (int c = 0; c < m.bs; ++c) d[r * m.bs + c] = m.val[static_cast<std::int64_t>(e) * m.bs + c];
        for (double& v : d) {
            if (v == 0.0) throw std::runtime_error("zero diagonal in multigrid smoother");
            v = 1.0 / v;
        }
        return d;
    }
    // build_coloring (multigrid.hpp:345-363)
    static void coloring(const Matrix& m, std::vector<std::int64_t>& red, std::vector<std::int64_t>& black)
    {
        std::vector<int> colors(m.owned, -1);
        for (std::int64_t r = 0; r < m.owned; ++r) {
            bool nred = false;
            for (std::int32_t e = m.off[r]; e < m.off[r + 1]; ++e) {
                const std::int64_t nb = m.col[e];
                nred = nred || (nb < m.owned && colors[nb] == 0);
            }
            colors[r] = nred ? 1 : 0;
            (colors[r] == 0 ? red : black).push_back(r);
        }
    }
    double weight(std::uint64_t step) const
    {
        if (opt.smoother != 2) return opt.jacobi_weight;
        const std::uint64_t levels = std::max<std::uint64_t>(1, opt.chebyshev_levels);
        const double pi = std::acos(-1.0);
        const double rho = 0.99;
        const double angle = pi * static_cast<double>(2 * (step % levels) + 1) / static_cast<double>(2 * levels);
        const double cs = std::cos(angle);
        return 2.0 / (1.0 + std::sqrt(1.0 - rho * rho * cs * cs));
    }
    // gauss_seidel_rows (multigrid.hpp:365-395)
    static void gs_rows(const Matrix& m, const double* rhs, double* x, const std::vector<std::int64_t>& rows)
    {
        std::vector<double> rr(m.bs);
        for (const std::int64_t r : rows) {
            std::copy_n(rhs + r * m.bs, m.bs, rr.data());
            const double* diag = nullptr;
            for (std::int32_t e = m.off[r]; e < m.off[r + 1]; ++e) {
                const std::int64_t c = m.col[e];
                const double* coef = m.val + static_cast<std::int64_t>(e) * m.bs;
                if (c == r) {
                    diag = coef;
                    continue;
                }
                for (int k = 0; k < m.bs; ++k) rr[k] -= coef[k] * x[c * m.bs + k];
            }
            if (!diag) throw std::runtime_error("multigrid smoother row has no diagonal");
            for (int k = 0; k < m.bs; ++k) x[r * m.bs + k] = rr[k] / diag[k];
        }
    }
    // smooth (multigrid.hpp:307-331), the coarse level
    void smooth_coarse(const double* rhs, double* x, std::uint64_t steps) const
    {
        const std::int64_t nn = coarse.owned * bs;
        std::vector<double> applied(nn);
        for (std::uint64_t st = 0; st < steps; ++st) {
            if (opt.smoother == 1) {
                gs_rows(coarse, rhs, x, coarse_red);
                gs_rows(coarse, rhs, x, coarse_black);
            } else {
                spmv_csr(coarse, x, applied.data());
                const double w = opt.smoother == 2 ? weight(st) : opt.jacobi_weight;
                for (std::int64_t i = 0; i < nn; ++i) x[i] += w * coarse_inv[i] * (rhs[i] - applied[i]);
            }
        }
    }
    // restrict_average / prolong_add (multigrid.hpp:397-418)
    void restrict_avg(const double* f, double* c) const
    {
        std::fill_n(c, coarse.owned * bs, 0.0);
        for (std::int64_t node = 0; node < n; ++node) {
            const std::int64_t agg = f2c[node];
            for (int k = 0; k < bs; ++k) c[agg * bs + k] += f[node * bs + k] / static_cast<double>(sizes[agg]);
        }
    }
    void prolong_add(const double* c, double* f) const
    {
        for (std::int64_t node = 0; node < n; ++node)
            for (int k = 0; k < bs; ++k) f[node * bs + k] += c[f2c[node] * bs + k];
    }
    // solve (multigrid.hpp:74-221), fine operator = the fine matrix
    void solve(const double* b, double* x, const oracle_options& o, int cycle, oracle_result& r) const
    {
        init_result(r);
        if (!valid_options(o)) return;
        Identity id(n * bs);
        Solve sv(fine, id, r);
        const std::int64_t nn = n * bs;
        std::vector<double> residual(fine.local() * bs), work(fine.local() * bs);
        const double rhs_norm = sv.nrm(b);
        if (!std::isfinite(rhs_norm)) {
            r.status = breakdown;
            r.breakdown_reason = non_finite_scalar;
            return;
        }
        const double thr = threshold_of(rhs_norm, o);
        sv.true_residual(b, x, residual.data(), work.data());
        double nv = sv.nrm(residual.data());
        r.initial_residual_norm = nv;
        sv.set_residual(nv, nv, rhs_norm);
        if (!std::isfinite(nv)) {
            r.status = breakdown;
            r.breakdown_reason = non_finite_scalar;
            return;
        }
        if (nv <= thr) {
            r.status = converged;
            return;
        }
        auto smooth_fine = [&](std::uint64_t steps) {
            for (std::uint64_t st = 0; st < steps; ++st) {
                if (opt.smoother == 1) {
                    for (const auto* rows : {&fine_red, &fine_black}) {
                        sv.op(x, work.data());
                        for (const std::int64_t row : *rows)
                            for (int k = 0; k < bs; ++k) {
                                const std::int64_t i = row * bs + k;
                                x[i] += fine_inv[i] * (b[i] - work[i]);
                            }
                    }
                } else {
                    sv.op(x, work.data());
                    const double w = weight(st);
                    for (std::int64_t i = 0; i < nn; ++i) x[i] += w * fine_inv[i] * (b[i] - work[i]);
                }
            }
        };
        const std::int64_t ncn = coarse.owned * bs;
        std::vector<double> crhs(ncn), csol(ncn);
        if (cycle == 2) {  // full
            restrict_avg(b, crhs.data());
            std::fill(csol.begin(), csol.end(), 0.0);
            smooth_coarse(crhs.data(), csol.data(), opt.coarse_steps);
            std::fill_n(x, nn, 0.0);
            prolong_add(csol.data(), x);
        }
        for (std::uint64_t it = 1; it <= o.maximum_iterations; ++it) {
            r.iterations = it;
            smooth_fine(opt.pre);
            const int corrections = cycle == 1 ? 2 : 1;
            for (int k = 0; k < corrections; ++k) {
                sv.true_residual(b, x, residual.data(), work.data());
                restrict_avg(residual.data(), crhs.data());
                std::fill(csol.begin(), csol.end(), 0.0);
                smooth_coarse(crhs.data(), csol.data(), opt.coarse_steps);
                prolong_add(csol.data(), x);
            }
            smooth_fine(opt.post);
            sv.true_residual(b, x, residual.data(), work.data());
            nv = sv.nrm(residual.data());
            sv.set_residual(nv, nv, rhs_norm);
            if (!std::isfinite(nv)) {
                r.status = breakdown;
                r.breakdown_reason = non_finite_scalar;
                return;
            }
            if (nv <= thr) {
                r.status = converged;
                return;
            }
        }
        r.status = maximum_iterations;
    }
};

// ArnoldiSnapshot<double> (core.hpp:462-534): unrotated Hessenberg columns
// and the preconditioned basis of the last cycle.
struct Snapshot {
    std::size_t maximum_columns = 0, columns = 0;
    std::vector<double> H;                 // column stride maximum_columns + 1
    std::vector<std::vector<double>> Z;    // owned values, ghosts zero
    void clear()
    {
        columns = maximum_columns = 0;
        H.clear();
        Z.clear();
    }
};

// fgmres (krylov_solvers.hpp:159-486); `snap` as the arnoldi_snapshot argument
void fgmres(const Matrix& a, Pc& pc, const double* b, double* x, const oracle_options& o, oracle_result& r,
            double* cached = nullptr, bool* cache_ready = nullptr, Snapshot* snap = nullptr)
{
    init_result(r);
    if (!valid_options(o) || o.restart == 0) {
        if (snap) snap->clear();
        return;
    }
    Solve s(a, pc, r);
    s.cached = cached;
    s.cache_ready = cache_ready;
    const std::size_t m = o.restart;
    const std::int64_t L = s.local;
    std::vector<double> residual(L), work(L);
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(L)), Z(m, std::vector<double>(L));
    std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), glob(m + 2), coef(m + 1);
    auto h = [&](std::size_t row, std::size_t col) -> double& { return H[col * (m + 1) + row]; };
    auto mark = [&](Reason reason) {
        r.status = breakdown;
        r.breakdown_reason = reason;
    };
    s.true_residual(b, x, residual.data(), work.data());
    const double rhs_norm = s.nrm(b);
    const double initial = s.nrm(residual.data());
    const double thr = threshold_of(rhs_norm, o);
    r.initial_residual_norm = initial;
    s.set_residual(initial, initial, rhs_norm);
    if (!std::isfinite(initial) || !std::isfinite(thr)) {
        mark(non_finite_scalar);
        return;
    }
    if (initial <= thr) {
        if (snap) snap->clear();
        r.status = converged;
        return;
    }
    const double eps = std::numeric_limits<double>::epsilon();
    const double dmin = std::numeric_limits<double>::min();
    while (r.iterations < o.maximum_iterations) {
        if (snap) {  // begin_cycle (core.hpp:465-480)
            snap->maximum_columns = m;
            snap->columns = 0;
            snap->H.assign((m + 1) * m, 0.0);
            snap->Z.assign(m, std::vector<double>(L, 0.0));
        }
        const double beta = s.nrm(residual.data());
        for (std::int64_t i = 0; i < s.n; ++i) V[0][i] = residual[i] / beta;
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        std::size_t cols = 0;
        bool abreak = false;
        for (; cols < m && r.iterations < o.maximum_iterations; ++cols, ++r.iterations) {
            s.prec(V[cols].data(), Z[cols].data());
            s.op(Z[cols].data(), work.data());
            const std::size_t count = cols + 1;
            for (std::size_t row = 0; row < count; ++row) glob[row] = dot(s.n, work.data(), V[row].data());
            glob[count] = dot(s.n, work.data(), work.data());
            ++r.global_reductions;
            const double original = glob[count];
            double projected = original;
            for (std::size_t row = 0; row < count; ++row) {
                h(row, cols) = glob[row];
                for (std::int64_t i = 0; i < s.n; ++i) work[i] = std::fma(-h(row, cols), V[row][i], work[i]);
                projected -= h(row, cols) * h(row, cols);
            }
            const double nonneg = std::max(projected, 0.0);
            const bool forced = o.gmres_orthogonalization == 0;
            const bool adaptive = o.gmres_orthogonalization == 1 && (projected <= 0 || nonneg < 0.5 * original);
            double nsq = nonneg;
            if (forced || adaptive) {
                for (std::size_t row = 0; row < count; ++row) glob[row] = dot(s.n, work.data(), V[row].data());
                glob[count] = dot(s.n, work.data(), work.data());
                ++r.global_reductions;
                ++r.reorthogonalizations;
                nsq = glob[count];
                for (std::size_t row = 0; row < count; ++row) {
                    const double corr = glob[row];
                    h(row, cols) += corr;
                    for (std::int64_t i = 0; i < s.n; ++i) work[i] = std::fma(-corr, V[row][i], work[i]);
                    nsq -= corr * corr;
                }
            }
            if (original < dmin || !std::isfinite(nsq) || nsq <= eps * original || (nsq > 0 && nsq < dmin)) {
                h(cols + 1, cols) = s.nrm(work.data());
                ++r.arnoldi_norm_verifications;
            } else {
                h(cols + 1, cols) = std::sqrt(std::max(nsq, 0.0));
            }
            if (!std::isfinite(h(cols + 1, cols))) {
                mark(non_finite_scalar);
                return;
            }
            double cscale = 0;
            for (std::size_t row = 0; row <= cols + 1; ++row) cscale = std::max(cscale, std::abs(h(row, cols)));
            if (h(cols + 1, cols) > o.breakdown_tolerance * cscale) {
                const double inv = 1.0 / h(cols + 1, cols);
                for (std::int64_t i = 0; i < s.n; ++i) V[cols + 1][i] = work[i] * inv;
            } else {
                abreak = true;
            }
            if (snap)  // record_column (krylov_solvers.hpp:361-366)
                for (std::size_t row = 0; row <= cols + 1; ++row) snap->H[cols * (m + 1) + row] = h(row, cols);
            for (std::size_t row = 0; row < cols; ++row) {
                const double first = cs[row] * h(row, cols) + sn[row] * h(row + 1, cols);
                h(row + 1, cols) = -sn[row] * h(row, cols) + cs[row] * h(row + 1, cols);
                h(row, cols) = first;
            }
            const double den = std::hypot(h(cols, cols), h(cols + 1, cols));
            if (!std::isfinite(den)) {
                mark(non_finite_scalar);
                return;
            }
            if (den == 0) {
                abreak = true;
                cs[cols] = 1;
                sn[cols] = 0;
            } else {
                cs[cols] = h(cols, cols) / den;
                sn[cols] = h(cols + 1, cols) / den;
            }
            h(cols, cols) = cs[cols] * h(cols, cols) + sn[cols] * h(cols + 1, cols);
            h(cols + 1, cols) = 0;
            g[cols + 1] = -sn[cols] * g[cols];
            g[cols] *= cs[cols];
            const double est = std::abs(g[cols + 1]);
            r.recursive_residual_norm = est;
            if (est <= thr || abreak) {
                ++cols;
                ++r.iterations;
                break;
            }
        }
        if (snap) {  // finish_cycle (krylov_solvers.hpp:435-437): owned values
            for (std::size_t c = 0; c < cols; ++c) std::copy(Z[c].begin(), Z[c].begin() + s.n, snap->Z[c].begin());
            snap->columns = cols;
        }
        std::fill(coef.begin(), coef.end(), 0.0);
        for (std::size_t rev = cols; rev-- > 0;) {
            double value = g[rev];
            for (std::size_t c = rev + 1; c < cols; ++c) value -= h(rev, c) * coef[c];
            double cscale = 0;
            for (std::size_t row = 0; row <= rev; ++row) cscale = std::max(cscale, std::abs(h(row, rev)));
            if (std::abs(h(rev, rev)) <= o.breakdown_tolerance * cscale) {
                mark(singular_projected_system);
                return;
            }
            coef[rev] = value / h(rev, rev);
        }
        gmres_update(s.n, x, Z, coef, cols, x);
        s.true_residual(b, x, residual.data(), work.data());
        const double tn = s.nrm(residual.data());
        s.set_residual(r.recursive_residual_norm, tn, rhs_norm);
        if (!std::isfinite(tn)) {
            mark(non_finite_scalar);
            return;
        }
        if (tn <= thr) {
            r.status = converged;
            return;
        }
        if (abreak && !(tn < beta)) {
            mark(arnoldi_invariant_subspace);
            return;
        }
    }
    r.status = maximum_iterations;
}

// RecycleSpace (recycling.hpp:24-187): U and C = A U with C orthonormal,
// at most `capacity` pairs (the oldest dropped first).
struct RecycleSpace {
    std::size_t capacity;
    std::vector<std::vector<double>> U, C;
    explicit RecycleSpace(std::uint64_t cap) : capacity(cap)
    {
        if (cap == 0) throw std::invalid_argument("recycle space capacity must be positive");
    }
    // add_candidate (recycling.hpp:79-135): A-image CGS2 against C
    bool add_candidate(const Matrix& a, const double* cand, oracle_result* tel, double tol)
    {
        const std::int64_t n = a.owned * a.bs;
        std::vector<double> v(cand, cand + n), img(a.local() * a.bs, 0.0), vin(a.local() * a.bs, 0.0);
        std::copy(cand, cand + n, vin.begin());
        apply(a, vin.data(), img.data());
        if (tel) ++tel->operator_applications;
        const double orig = norm(n, img.data());
        if (tel) ++tel->global_reductions;
        if (!std::isfinite(orig) || orig == 0.0) return false;
        for (int pass = 0; pass < 2 && !C.empty(); ++pass) {
            std::vector<double> coef(C.size());
            for (std::size_t i = 0; i < C.size(); ++i) coef[i] = dot(n, C[i].data(), img.data());
            if (tel) ++tel->global_reductions;
            for (std::size_t i = 0; i < C.size(); ++i) {
                for (std::int64_t k = 0; k < n; ++k) img[k] = std::fma(-coef[i], C[i][k], img[k]);
                for (std::int64_t k = 0; k < n; ++k) v[k] = std::fma(-coef[i], U[i][k], v[k]);
            }
        }
        const double nv = norm(n, img.data());
        if (tel) ++tel->global_reductions;
        if (!std::isfinite(nv) || nv == 0.0 || nv / orig <= tol) return false;
        const double inv = 1.0 / nv;
        img.resize(n);
        for (auto& e : img) e *= inv;
        for (auto& e : v) e *= inv;
        if (C.size() == capacity) {
            C.erase(C.begin());
            U.erase(U.begin());
        }
        C.push_back(std::move(img));
        U.push_back(std::move(v));
        return true;
    }
    // refresh (recycling.hpp:65-73): rebuild C = A U from the stored U
    void refresh(const Matrix& a, oracle_result* tel)
    {
        RecycleSpace fresh(capacity);
        for (const auto& u : U) fresh.add_candidate(a, u.data(), tel, 128 * std::numeric_limits<double>::epsilon());
        *this = std::move(fresh);
    }
    // project_initial_guess (recycling.hpp:138-163): x += U C^T r
    void project(const double* res, double* x, std::int64_t n, oracle_result* tel)
    {
        if (C.empty()) return;
        std::vector<double> coef(C.size());
        for (std::size_t i = 0; i < C.size(); ++i) coef[i] = dot(n, C[i].data(), res);
        if (tel) ++tel->global_reductions;
        for (std::size_t i = 0; i < C.size(); ++i)
            for (std::int64_t k = 0; k < n; ++k) x[k] = std::fma(coef[i], U[i][k], x[k]);
    }
};

// RecycledPreconditioner (recycling.hpp:222-289): z = M v, then the operator
// image A z is projected off C (and z off U) and cached for the next operator
// application.
struct RecycledPc : Pc {
    const Matrix* a;
    Pc* inner;
    const RecycleSpace* space;
    std::vector<double> cached;
    bool ready = false;
    std::uint64_t extra_reductions = 0;
    void apply(const double* in, double* out) override
    {
        const std::int64_t n = a->owned * a->bs;
        inner->apply(in, out);
        std::vector<double> xin(a->local() * a->bs, 0.0);
        std::copy(out, out + n, xin.begin());
        ::apply(*a, xin.data(), cached.data());
        if (!space->C.empty()) {
            std::vector<double> coef(space->C.size());
            for (std::size_t i = 0; i < space->C.size(); ++i) coef[i] = dot(n, space->C[i].data(), cached.data());
            ++extra_reductions;
            for (std::size_t i = 0; i < space->C.size(); ++i) {
                for (std::int64_t k = 0; k < n; ++k) out[k] = std::fma(-coef[i], space->U[i][k], out[k]);
                for (std::int64_t k = 0; k < n; ++k) cached[k] = std::fma(-coef[i], space->C[i][k], cached[k]);
            }
        }
        ready = true;
    }
};

// recycled_fgmres (recycling.hpp:291-360)
void recycled_fgmres(const Matrix& a, Pc& pc, const double* b, double* x, RecycleSpace& space,
                     const oracle_options& o, bool update, oracle_result& r, Snapshot* snap = nullptr)
{
    oracle_result tel;
    std::memset(&tel, 0, sizeof(tel));
    init_result(r);
    if (!valid_options(o) || o.restart == 0) return;
    const std::int64_t n = a.owned * a.bs;
    std::vector<double> x0(x, x + n);
    if (!space.C.empty()) {
        space.refresh(a, &tel);
        std::vector<double> work(a.local() * a.bs), res(n);
        apply(a, x, work.data());
        ++tel.operator_applications;
        for (std::int64_t i = 0; i < n; ++i) res[i] = b[i] - work[i];
        space.project(res.data(), x, n, &tel);
    }
    RecycledPc rpc;
    rpc.a = &a;
    rpc.inner = &pc;
    rpc.space = &space;
    rpc.cached.assign(a.local() * a.bs, 0.0);
    fgmres(a, rpc, b, x, o, r, rpc.cached.data(), &rpc.ready, snap);
    r.operator_applications += tel.operator_applications;
    r.global_reductions += tel.global_reductions + rpc.extra_reductions;
    if (update && r.status == converged) {
        std::vector<double> corr(n);
        for (std::int64_t i = 0; i < n; ++i) corr[i] = x[i] - x0[i];
        space.add_candidate(a, corr.data(), &r, o.breakdown_tolerance);
    }
}

template<class F>
int guard(F&& f)
{
    try {
        f();
        g_error.clear();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return -1;
    } catch (const std::runtime_error& e) {
        g_error = e.what();
        return std::string(e.what()).find("zero pivot") != std::string::npos ? -2 : -7;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -7;
    }
}

} // namespace

// ------------------------------------------------------------------- C ABI --
extern "C" {

typedef struct oracle_matrix {
    std::int32_t kind;
    std::int32_t block_size;
    std::int64_t owned_nodes, ghost_nodes;
    const std::int32_t* offsets;
    const std::int32_t* columns;
    const double* values;
    const double* diagonal;
} oracle_matrix;

static Matrix to_matrix(const oracle_matrix* m)
{
    if (!m || m->owned_nodes <= 0 || m->block_size <= 0) throw std::invalid_argument("invalid matrix");
    return Matrix{m->kind, m->owned_nodes, m->ghost_nodes, m->block_size, m->offsets, m->columns, m->values, m->diagonal};
}

const char* oracle_last_error(void) { return g_error.c_str(); }

// Threads of the all-cores CPU baseline (1: the serial restatement).
void oracle_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }
int oracle_get_threads(void) { return g_threads; }

int oracle_spmv(const oracle_matrix* m, const double* x, double* y)
{
    return guard([&] { apply(to_matrix(m), x, y); });
}

// Low-precision SpMV: values (and split diagonal) already rounded to the
// storage type and given as float, vectors float, float fma accumulation.
int oracle_spmv_f32(const oracle_matrix* m, const float* values, const float* diagonal, const float* x, float* y)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        const int bs = a.bs;
        for (std::int64_t row = 0; row < a.owned; ++row) {
            if (a.kind == kBsr) {
                float acc[64] = {0};
                for (std::int32_t e = a.off[row]; e < a.off[row + 1]; ++e)
                    for (int r = 0; r < bs; ++r) {
                        float t = 0.0f;
                        for (int c = 0; c < bs; ++c)
                            t = std::fma(values[(static_cast<std::int64_t>(e) * bs + r) * bs + c],
                                         x[static_cast<std::int64_t>(a.col[e]) * bs + c], t);
                        acc[r] += t;
                    }
                for (int r = 0; r < bs; ++r) y[row * bs + r] = acc[r];
                continue;
            }
            for (int c = 0; c < bs; ++c) {
                float acc = a.kind == kSplit ? diagonal[row * bs + c] * x[row * bs + c] : 0.0f;
                for (std::int32_t e = a.off[row]; e < a.off[row + 1]; ++e)
                    acc = std::fma(values[static_cast<std::int64_t>(e) * bs + c],
                                   x[static_cast<std::int64_t>(a.col[e]) * bs + c], acc);
                y[row * bs + c] = acc;
            }
        }
    });
}

double oracle_dot(std::int64_t n, const double* a, const double* b) { return dot(n, a, b); }
double oracle_norm(std::int64_t n, const double* a) { return norm(n, a); }

typedef struct oracle_pc_params {
    std::int32_t kind;     // 0 identity, 1 jacobi, 2 ssor, 3 ilu
    std::int32_t fill_level;
    double omega;
    double minimum_diagonal;
    std::int64_t block_nodes;
} oracle_pc_params;

struct OraclePcHandle {
    Matrix matrix;
    std::unique_ptr<Pc> pc;
};

int oracle_pc_create(const oracle_matrix* m, const oracle_pc_params* p, void** out)
{
    return guard([&] {
        if (!p || !out) throw std::invalid_argument("null argument");
        auto h = std::make_unique<OraclePcHandle>();
        h->matrix = to_matrix(m);
        const double mind = p->minimum_diagonal > 0 ? p->minimum_diagonal : 64 * std::numeric_limits<double>::epsilon();
        switch (p->kind) {
        case 0: h->pc = std::make_unique<Identity>(h->matrix.owned * h->matrix.bs); break;
        case 1: h->pc = std::make_unique<Jacobi>(&h->matrix, mind); break;
        case 2:
            if (h->matrix.kind == kBsr) h->pc = std::make_unique<DenseSsor>(&h->matrix, p->omega);
            else h->pc = std::make_unique<Ssor>(&h->matrix, p->omega, p->block_nodes);
            break;
        case 3: h->pc = std::make_unique<Ilu>(&h->matrix, p->block_nodes, p->fill_level); break;
        default: throw std::invalid_argument("unknown preconditioner");
        }
        *out = h.release();
    });
}

int oracle_pc_update(void* handle)
{
    return guard([&] { static_cast<OraclePcHandle*>(handle)->pc->update(); });
}

int oracle_pc_apply(void* handle, const double* in, double* out)
{
    return guard([&] { static_cast<OraclePcHandle*>(handle)->pc->apply(in, out); });
}

void oracle_pc_destroy(void* handle) { delete static_cast<OraclePcHandle*>(handle); }

// ILU factors in the reference's accessor layout (preconditioner.hpp:365-370).
int oracle_ilu_factors(void* handle, std::int64_t* entries, std::int64_t* offsets, std::int64_t* columns,
                       double* values, double* inverse_diagonal)
{
    return guard([&] {
        auto* ilu = dynamic_cast<Ilu*>(static_cast<OraclePcHandle*>(handle)->pc.get());
        if (!ilu) throw std::invalid_argument("not an ILU preconditioner");
        *entries = static_cast<std::int64_t>(ilu->columns.size());
        if (offsets) std::copy(ilu->row_offsets.begin(), ilu->row_offsets.end(), offsets);
        if (columns) std::copy(ilu->columns.begin(), ilu->columns.end(), columns);
        if (values) std::copy(ilu->values.begin(), ilu->values.end(), values);
        if (inverse_diagonal) std::copy(ilu->inverse_diagonal.begin(), ilu->inverse_diagonal.end(), inverse_diagonal);
    });
}

int oracle_bicgstab(const oracle_matrix* m, void* pc_handle, const double* b, double* x, const oracle_options* o,
                    oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        bicgstab(a, pc, b, x, *o, *r);
    });
}

int oracle_pipelined_bicgstab(const oracle_matrix* m, void* pc_handle, const double* b, double* x,
                              const oracle_options* o, oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        pipelined_bicgstab(a, pc, b, x, *o, *r);
    });
}

int oracle_communication_hiding_bicgstab(const oracle_matrix* m, void* pc_handle, const double* b, double* x,
                                         const oracle_options* o, oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        communication_hiding_bicgstab(a, pc, b, x, *o, *r);
    });
}

int oracle_idrs(const oracle_matrix* m, void* pc_handle, const double* b, double* x, const oracle_options* o,
                oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        idrs(a, pc, b, x, *o, *r);
    });
}

int oracle_recycle_create(std::uint64_t capacity, void** out)
{
    return guard([&] { *out = new RecycleSpace(capacity); });
}

void oracle_recycle_destroy(void* h) { delete static_cast<RecycleSpace*>(h); }

std::int64_t oracle_recycle_size(void* h) { return static_cast<std::int64_t>(static_cast<RecycleSpace*>(h)->C.size()); }

// U_i and C_i (owned entries)
int oracle_recycle_get(void* h, std::int64_t i, double* u, double* c)
{
    return guard([&] {
        auto* sp = static_cast<RecycleSpace*>(h);
        std::copy(sp->U.at(i).begin(), sp->U.at(i).end(), u);
        std::copy(sp->C.at(i).begin(), sp->C.at(i).end(), c);
    });
}

int oracle_recycle_add(void* h, const oracle_matrix* m, const double* candidate, double tol, int* added,
                       oracle_result* telemetry)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        *added = static_cast<RecycleSpace*>(h)->add_candidate(a, candidate, telemetry, tol) ? 1 : 0;
    });
}

// recycled_fgmres with an ArnoldiSnapshot (recycling.hpp:300-342), outputs as
// oracle_fgmres_snapshot's
int oracle_recycled_fgmres_snapshot(const oracle_matrix* m, void* pc_handle, const double* b, double* x, void* space,
                                    const oracle_options* o, int update, std::uint64_t* columns, double* h, double* z,
                                    oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        Snapshot snap;
        recycled_fgmres(a, pc, b, x, *static_cast<RecycleSpace*>(space), *o, update != 0, *r, &snap);
        *columns = snap.columns;
        const std::size_t mm = o->restart;
        const std::size_t L = static_cast<std::size_t>((a.owned + a.ghost) * a.bs);
        std::fill(h, h + (mm + 1) * mm, 0.0);
        std::fill(z, z + mm * L, 0.0);
        if (snap.columns == 0) return;
        std::copy(snap.H.begin(), snap.H.end(), h);
        for (std::size_t c = 0; c < snap.columns; ++c) std::copy(snap.Z[c].begin(), snap.Z[c].end(), z + c * L);
    });
}

int oracle_recycled_fgmres(const oracle_matrix* m, void* pc_handle, const double* b, double* x, void* space,
                           const oracle_options* o, int update, oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        recycled_fgmres(a, pc, b, x, *static_cast<RecycleSpace*>(space), *o, update != 0, *r);
    });
}

int oracle_mg_create(const oracle_matrix* m, const MgOptions* o, void** out)
{
    return guard([&] { *out = new Multigrid(to_matrix(m), *o); });
}

int oracle_mg_solve(void* h, const double* b, double* x, const oracle_options* o, int cycle, oracle_result* r)
{
    return guard([&] { static_cast<Multigrid*>(h)->solve(b, x, *o, cycle, *r); });
}

// coarse matrix and aggregate map (NULL arrays: sizes only)
int oracle_mg_coarse(void* h, std::int64_t* nodes, std::int64_t* entries, std::int32_t* off, std::int32_t* col,
                     double* val, std::int64_t* f2c)
{
    return guard([&] {
        const auto* mg = static_cast<Multigrid*>(h);
        *nodes = mg->coarse.owned;
        *entries = static_cast<std::int64_t>(mg->ccol.size());
        if (off) std::copy(mg->coff.begin(), mg->coff.end(), off);
        if (col) std::copy(mg->ccol.begin(), mg->ccol.end(), col);
        if (val) std::copy(mg->cval.begin(), mg->cval.end(), val);
        if (f2c) std::copy(mg->f2c.begin(), mg->f2c.end(), f2c);
    });
}

void oracle_mg_destroy(void* h) { delete static_cast<Multigrid*>(h); }

int oracle_fgmres(const oracle_matrix* m, void* pc_handle, const double* b, double* x, const oracle_options* o,
                  oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        fgmres(a, pc, b, x, *o, *r);
    });
}

// fgmres with an ArnoldiSnapshot: *columns receives its column count, h the
// (restart + 1) x restart unrotated Hessenberg (column-major, zeros beyond),
// z the restart preconditioned basis vectors (local length each, ghosts 0).
int oracle_fgmres_snapshot(const oracle_matrix* m, void* pc_handle, const double* b, double* x,
                           const oracle_options* o, std::uint64_t* columns, double* h, double* z, oracle_result* r)
{
    return guard([&] {
        const Matrix a = to_matrix(m);
        Identity id(a.owned * a.bs);
        Pc& pc = pc_handle ? *static_cast<OraclePcHandle*>(pc_handle)->pc : static_cast<Pc&>(id);
        Snapshot snap;
        fgmres(a, pc, b, x, *o, *r, nullptr, nullptr, &snap);
        *columns = snap.columns;
        const std::size_t mm = o->restart;
        const std::size_t L = static_cast<std::size_t>((a.owned + a.ghost) * a.bs);
        std::fill(h, h + (mm + 1) * mm, 0.0);
        std::fill(z, z + mm * L, 0.0);
        if (snap.columns == 0) return;
        std::copy(snap.H.begin(), snap.H.end(), h);
        for (std::size_t c = 0; c < snap.columns; ++c) std::copy(snap.Z[c].begin(), snap.Z[c].end(), z + c * L);
    });
}

} // extern "C"
