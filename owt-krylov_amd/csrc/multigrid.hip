// Device two-level aggregation multigrid (AggregationMultigrid,
// multigrid.hpp:41-421).  Set-up follows the reference on the host: greedy
// aggregates seeded in row order, the row-averaged coarse matrix (sums in
// fine-row, entry order), inverse diagonals and the greedy red/black
// colourings.  Every operation of a cycle runs on the device:
//   Jacobi / Chebyshev smoothing  SpMV + one update pass
//   red-black Gauss-Seidel (fine) full residual per colour + colour update
//   red-black Gauss-Seidel (coarse, true Gauss-Seidel rows, gauss_seidel_rows
//     multigrid.hpp:365-395) level schedule of each colour's rows: rows of
//     a level touch no earlier row of the same colour, so each level is one
//     parallel pass, one workgroup walks the levels, and the result is the
//     sequential sweep's
//   restriction                   per aggregate, its members in ascending
//                                 order (the reference's summation order)
//   prolongation                  elementwise add
// with the reference's arithmetic in each update.
#include "precond.hpp"
#include "solver_common.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <vector>

using namespace owtk;

struct owtk_multigrid_s {
    owtk_matrix fine = nullptr;
    owtk_context ctx = nullptr;  // destroy may run after the fine matrix is gone
    owtk_matrix coarse = nullptr;  // owned
    owtk_multigrid_options opt{};
    int bs = 1;
    std::int64_t n = 0, nc = 0;
    std::vector<std::int64_t> h_f2c;
    std::vector<std::int32_t> h_coff, h_ccol;
    std::vector<double> h_cval;
    owtk::DeviceArray<double> fine_inv, coarse_inv, agg_size;
    owtk::DeviceArray<std::int32_t> f2c, agg_ptr, agg_members, fine_red, fine_black;
    std::int64_t n_red = 0, n_black = 0;
    // coarse Gauss-Seidel: CSR on the device and the level schedule of each colour
    owtk::DeviceArray<std::int32_t> coff, ccol, gs_rows;
    owtk::DeviceArray<std::int64_t> gs_levels;  // device copy of gs_level_ptr
    owtk::DeviceArray<double> cval;
    std::vector<std::int64_t> gs_level_ptr;  // red levels, then black levels
    std::int64_t red_levels = 0;
    // the same rows in sweep (level) order as ELL: position p's k-th
    // off-diagonal entry (CSR order) at k * positions + p, its values at
    // (k * positions + p) * bs + c; padding slots hold the row itself and 0
    owtk::DeviceArray<std::int32_t> gs_len, gs_col;
    owtk::DeviceArray<double> gs_val, gs_diag;
    std::int64_t gs_positions = 0;
    int gs_k = 0;
    ~owtk_multigrid_s()
    {
        if (coarse) (void)owtk_matrix_destroy(coarse);
    }
};

namespace owtk {
namespace {

using dev::kBlock;

// x += (w * inv) * (b - work)
__global__ __launch_bounds__(256) void k_mg_jacobi(double* __restrict__ x, const double* __restrict__ b,
                                                   const double* __restrict__ work, const double* __restrict__ inv,
                                                   double w, std::int64_t n)
{
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
        x[i] = x[i] + (w * inv[i]) * (b[i] - work[i]);
}

// rows of one colour: x += inv * (b - work)
__global__ __launch_bounds__(256) void k_mg_colour(const std::int32_t* __restrict__ rows, std::int64_t count, int bs,
                                                   double* __restrict__ x, const double* __restrict__ b,
                                                   const double* __restrict__ work, const double* __restrict__ inv)
{
    const std::int64_t total = count * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t i = static_cast<std::int64_t>(rows[t / bs]) * bs + t % bs;
        x[i] = x[i] + inv[i] * (b[i] - work[i]);
    }
}

// coarse[a] = sum over members m (ascending) of fine[m] / size[a], from zero
__global__ __launch_bounds__(256) void k_mg_restrict(const double* __restrict__ fine, double* __restrict__ coarse,
                                                     const std::int32_t* __restrict__ ptr,
                                                     const std::int32_t* __restrict__ members,
                                                     const double* __restrict__ size, std::int64_t nc, int bs)
{
    const std::int64_t total = nc * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step) {
        const std::int64_t a = t / bs;
        const int k = static_cast<int>(t % bs);
        const double sz = size[a];
        double acc = 0.0;
        for (std::int32_t e = ptr[a]; e < ptr[a + 1]; ++e)
            acc = acc + fine[static_cast<std::int64_t>(members[e]) * bs + k] / sz;
        coarse[t] = acc;
    }
}

// fine[node] += coarse[f2c[node]]
__global__ __launch_bounds__(256) void k_mg_prolong(const double* __restrict__ coarse, double* __restrict__ fine,
                                                    const std::int32_t* __restrict__ f2c, std::int64_t n, int bs)
{
    const std::int64_t total = n * bs;
    const std::int64_t step = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += step)
        fine[t] = fine[t] + coarse[static_cast<std::int64_t>(f2c[t / bs]) * bs + t % bs];
}

// One colour's Gauss-Seidel sweep, all levels in one workgroup: the rows of
// a level in parallel, a barrier between levels (a level's reads of earlier
// levels' writes stay inside the workgroup).  The rows come from a
// level-ordered ELL copy: a work item's row, row length, diagonal and
// right-hand side, and then eight of its slots at a time, load with no
// dependence on one another (padding slots read the row's
// own column, and their products are dropped by a select), so a level costs
// one round trip for its structure and one for the gathered solution values
// instead of a pointer chase through rows, offsets and columns.  The
// arithmetic and its order are the CSR sweep's.
struct GsArgs {
    const std::int32_t* rows;
    const std::int32_t* len;
    const std::int32_t* col;
    const double* val;
    const double* diag;
    const double* rhs;
    std::int64_t positions;
    int bs, K;
};

// One colour's Gauss-Seidel sweep, all levels in one workgroup: the rows of
// a level in parallel, a barrier between levels (a level's reads of earlier
// levels' writes stay inside the workgroup).  The rows come from a
// level-ordered ELL copy: a work item's row, row length, diagonal and
// right-hand side, and then eight of its slots at a time, load with no
// dependence on one another (padding slots name the row itself, and their
// products are dropped by a select), so a level costs one round trip for its
// structure and one for the gathered solution values instead of a pointer
// chase through rows, offsets and columns.  The arithmetic and its order are
// the CSR sweep's.  (Loading the next level's structure during the current
// one did not help: vector loads retire in issue order, so the gathers then
// waited for the prefetch as well.)
__global__ __launch_bounds__(1024) void k_mg_gs_ell(const std::int64_t* __restrict__ level_ptr, std::int64_t first,
                                                    std::int64_t last, GsArgs g, double* x)
{
    constexpr int C = 8;
    for (std::int64_t l = first; l < last; ++l) {
        const std::int64_t a = level_ptr[l];
        const std::int64_t total = (level_ptr[l + 1] - a) * g.bs;
        for (std::int64_t t = threadIdx.x; t < total; t += blockDim.x) {
            const std::int64_t p = a + t / g.bs;
            const int k = static_cast<int>(t % g.bs);
            const std::int64_t r = g.rows[p];
            const int n = g.len[p];
            const double d = g.diag[p * g.bs + k];
            double rr = g.rhs[r * g.bs + k];
            for (int k0 = 0; k0 < g.K; k0 += C) {
                std::int64_t cj[C];
                double v[C], xv[C];
#pragma unroll
                for (int u = 0; u < C; ++u) {
                    const std::int64_t slot = static_cast<std::int64_t>(k0 + u < g.K ? k0 + u : g.K - 1) * g.positions + p;
                    cj[u] = g.col[slot];
                    v[u] = g.val[slot * g.bs + k];
                }
#pragma unroll
                for (int u = 0; u < C; ++u) xv[u] = x[cj[u] * g.bs + k];
#pragma unroll
                for (int u = 0; u < C; ++u) {
                    const double f = rr - v[u] * xv[u];
                    rr = k0 + u < n ? f : rr;
                }
            }
            x[r * g.bs + k] = rr / d;
        }
        __syncthreads();
    }
}

template<class T>
void upload(DeviceArray<T>& d, const std::vector<T>& h)
{
    d.allocate(std::max<std::size_t>(h.size(), 1));
    if (!h.empty()) OWTK_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

// inverse of a CSR-kind matrix's diagonal entries (BlockCsrMatrix::diagonal)
std::vector<double> inverse_diagonal(std::int64_t n, int bs, const std::vector<std::int32_t>& off,
                                     const std::vector<std::int32_t>& col, const std::vector<double>& val)
{
    std::vector<double> d(static_cast<std::size_t>(n) * bs, 0.0);
    for (std::int64_t r = 0; r < n; ++r)
        for (std::int32_t e = off[r]; e < off[r + 1]; ++e)
            if (col[e] == r)
                for (int c = 0; c < bs; ++c) d[r * bs + c] = val[static_cast<std::size_t>(e) * bs + c];
    for (double& v : d) {
        if (v == 0.0) throw std::runtime_error("zero diagonal in multigrid smoother");
        v = 1.0 / v;
    }
    return d;
}

// build_coloring (multigrid.hpp:345-363)
void colouring(std::int64_t n, const std::vector<std::int32_t>& off, const std::vector<std::int32_t>& col,
               std::vector<std::int32_t>& red, std::vector<std::int32_t>& black)
{
    std::vector<int> colours(n, -1);
    for (std::int64_t r = 0; r < n; ++r) {
        bool nred = false;
        for (std::int32_t e = off[r]; e < off[r + 1]; ++e) {
            const std::int64_t c = col[e];
            nred = nred || (c < n && colours[c] == 0);
        }
        colours[r] = nred ? 1 : 0;
        (colours[r] == 0 ? red : black).push_back(static_cast<std::int32_t>(r));
    }
}

// Levels of one colour's sequential sweep: a row waits for every earlier row
// of the list it reads or that reads it (in-place Gauss-Seidel order).
void gs_levels(std::int64_t n, const std::vector<std::int32_t>& off, const std::vector<std::int32_t>& col,
               const std::vector<std::int32_t>& tptr, const std::vector<std::int32_t>& tcol,
               const std::vector<std::int32_t>& list, std::vector<std::int32_t>& rows_out,
               std::vector<std::int64_t>& level_ptr)
{
    std::vector<std::int64_t> pos(n, -1);
    for (std::size_t p = 0; p < list.size(); ++p) pos[list[p]] = static_cast<std::int64_t>(p);
    std::vector<int> level(n, 0);
    int depth = 0;
    for (std::size_t p = 0; p < list.size(); ++p) {
        const std::int32_t r = list[p];
        int lv = 0;
        auto visit = [&](std::int64_t q) {
            if (q != r && q < n && pos[q] >= 0 && pos[q] < static_cast<std::int64_t>(p)) lv = std::max(lv, level[q] + 1);
        };
        for (std::int32_t e = off[r]; e < off[r + 1]; ++e) visit(col[e]);
        for (std::int32_t e = tptr[r]; e < tptr[r + 1]; ++e) visit(tcol[e]);
        level[r] = lv;
        depth = std::max(depth, lv + 1);
    }
    std::vector<std::vector<std::int32_t>> by(depth);
    for (const std::int32_t r : list) by[level[r]].push_back(r);
    for (const auto& l : by) {
        rows_out.insert(rows_out.end(), l.begin(), l.end());
        level_ptr.push_back(static_cast<std::int64_t>(rows_out.size()));
    }
}

void mg_setup(owtk_multigrid_s* mg)
{
    owtk_matrix m = mg->fine;
    const auto& o = mg->opt;
    const int bs = m->bs;
    const std::int64_t n = m->owned;
    mg->bs = bs;
    mg->n = n;
    // fine values on the host, [entry][comp] (the reference's BlockCsrMatrix layout)
    std::vector<double> slots(static_cast<std::size_t>(m->value_count));
    OWTK_HIP(hipMemcpy(slots.data(), m->v64.get(), slots.size() * sizeof(double), hipMemcpyDeviceToHost));
    const auto& off = m->h_offsets;
    const auto& col = m->h_columns;
    std::vector<double> val(static_cast<std::size_t>(off[n]) * bs);
    for (std::int64_t e = 0; e < off[n]; ++e)
        for (int c = 0; c < bs; ++c) val[e * bs + c] = slots[m->value_index(m->s.slot_of_entry[e], c)];
    // build_aggregates (multigrid.hpp:237-262)
    constexpr std::int64_t none = -1;
    std::vector<std::int64_t> f2c(n, none), sizes;
    for (std::int64_t seed = 0; seed < n; ++seed) {
        if (f2c[seed] != none) continue;
        const std::int64_t agg = static_cast<std::int64_t>(sizes.size());
        std::int64_t size = 1;
        f2c[seed] = agg;
        for (std::int32_t e = off[seed]; e < off[seed + 1] && static_cast<std::uint64_t>(size) < o.target_aggregate_size;
             ++e) {
            const std::int64_t nb = col[e];
            if (nb < n && f2c[nb] == none) {
                f2c[nb] = agg;
                ++size;
            }
        }
        sizes.push_back(size);
    }
    const std::int64_t nc = static_cast<std::int64_t>(sizes.size());
    mg->nc = nc;
    // build_coarse_matrix (multigrid.hpp:264-305)
    std::vector<std::map<std::int64_t, std::vector<double>>> rows(nc);
    for (std::int64_t r = 0; r < n; ++r) {
        const std::int64_t cr = f2c[r];
        for (std::int32_t e = off[r]; e < off[r + 1]; ++e) {
            const std::int64_t fc = col[e];
            if (fc >= n) continue;
            auto [pos, inserted] = rows[cr].try_emplace(f2c[fc], bs, 0.0);
            for (int c = 0; c < bs; ++c) pos->second[c] += val[static_cast<std::size_t>(e) * bs + c] / static_cast<double>(sizes[cr]);
        }
    }
    mg->h_coff.assign(nc + 1, 0);
    mg->h_ccol.clear();
    mg->h_cval.clear();
    for (std::int64_t r = 0; r < nc; ++r) {
        for (const auto& [c, block] : rows[r]) {
            mg->h_ccol.push_back(static_cast<std::int32_t>(c));
            mg->h_cval.insert(mg->h_cval.end(), block.begin(), block.end());
        }
        mg->h_coff[r + 1] = static_cast<std::int32_t>(mg->h_ccol.size());
    }
    mg->h_f2c = f2c;
    const int rc = owtk_matrix_create_csr(m->ctx, nc, 0, bs, mg->h_coff.data(), mg->h_ccol.data(), mg->h_cval.data(),
                                          &mg->coarse);
    if (rc != OWTK_OK) throw std::runtime_error(std::string("coarse matrix: ") + owtk_last_error());
    upload(mg->fine_inv, inverse_diagonal(n, bs, off, col, val));
    upload(mg->coarse_inv, inverse_diagonal(nc, bs, mg->h_coff, mg->h_ccol, mg->h_cval));
    // aggregate members in ascending order, sizes, fine -> coarse
    std::vector<std::int32_t> aptr(nc + 1, 0), members(n), f2c32(n);
    for (std::int64_t r = 0; r < n; ++r) {
        ++aptr[f2c[r] + 1];
        f2c32[r] = static_cast<std::int32_t>(f2c[r]);
    }
    for (std::int64_t a = 0; a < nc; ++a) aptr[a + 1] += aptr[a];
    {
        std::vector<std::int32_t> next(aptr.begin(), aptr.end() - 1);
        for (std::int64_t r = 0; r < n; ++r) members[next[f2c[r]]++] = static_cast<std::int32_t>(r);
    }
    std::vector<double> sz(sizes.begin(), sizes.end());
    upload(mg->agg_ptr, aptr);
    upload(mg->agg_members, members);
    upload(mg->agg_size, sz);
    upload(mg->f2c, f2c32);
    std::vector<std::int32_t> red, black;
    colouring(n, off, col, red, black);
    mg->n_red = static_cast<std::int64_t>(red.size());
    mg->n_black = static_cast<std::int64_t>(black.size());
    upload(mg->fine_red, red);
    upload(mg->fine_black, black);
    // coarse Gauss-Seidel: device CSR and level schedules
    std::vector<std::int32_t> cred, cblack;
    colouring(nc, mg->h_coff, mg->h_ccol, cred, cblack);
    std::vector<std::int32_t> tptr(nc + 1, 0), tcol(mg->h_ccol.size());
    for (const std::int32_t c : mg->h_ccol) ++tptr[c + 1];
    for (std::int64_t a = 0; a < nc; ++a) tptr[a + 1] += tptr[a];
    {
        std::vector<std::int32_t> next(tptr.begin(), tptr.end() - 1);
        for (std::int64_t r = 0; r < nc; ++r)
            for (std::int32_t e = mg->h_coff[r]; e < mg->h_coff[r + 1]; ++e)
                tcol[next[mg->h_ccol[e]]++] = static_cast<std::int32_t>(r);
    }
    std::vector<std::int32_t> gs_rows;
    mg->gs_level_ptr.assign(1, 0);
    gs_levels(nc, mg->h_coff, mg->h_ccol, tptr, tcol, cred, gs_rows, mg->gs_level_ptr);
    mg->red_levels = static_cast<std::int64_t>(mg->gs_level_ptr.size()) - 1;
    gs_levels(nc, mg->h_coff, mg->h_ccol, tptr, tcol, cblack, gs_rows, mg->gs_level_ptr);
    upload(mg->gs_rows, gs_rows);
    upload(mg->gs_levels, mg->gs_level_ptr);
    upload(mg->coff, mg->h_coff);
    upload(mg->ccol, mg->h_ccol);
    upload(mg->cval, mg->h_cval);
    // level-ordered ELL copy of the coarse rows for the sweep kernel
    const std::int64_t P = static_cast<std::int64_t>(gs_rows.size());
    int K = 0;
    for (const std::int32_t r : gs_rows) {
        int off = 0;
        for (std::int32_t e = mg->h_coff[r]; e < mg->h_coff[r + 1]; ++e) off += mg->h_ccol[e] != r;
        K = std::max(K, off);
    }
    std::vector<std::int32_t> glen(static_cast<std::size_t>(std::max<std::int64_t>(P, 1)), 0);
    std::vector<std::int32_t> gcol(static_cast<std::size_t>(std::max<std::int64_t>(P * K, 1)), 0);
    std::vector<double> gval(static_cast<std::size_t>(std::max<std::int64_t>(P * K * bs, 1)), 0.0);
    std::vector<double> gdiag(static_cast<std::size_t>(std::max<std::int64_t>(P * bs, 1)), 0.0);
    for (std::int64_t p = 0; p < P; ++p) {
        const std::int32_t r = gs_rows[p];
        int k = 0;
        for (int q = 0; q < K; ++q) gcol[static_cast<std::size_t>(q * P + p)] = r;
        for (std::int32_t e = mg->h_coff[r]; e < mg->h_coff[r + 1]; ++e) {
            const std::int32_t c = mg->h_ccol[e];
            if (c == r) {
                for (int j = 0; j < bs; ++j) gdiag[static_cast<std::size_t>(p * bs + j)] = mg->h_cval[static_cast<std::size_t>(e) * bs + j];
                continue;
            }
            gcol[static_cast<std::size_t>(k * P + p)] = c;
            for (int j = 0; j < bs; ++j)
                gval[static_cast<std::size_t>((k * P + p) * bs + j)] = mg->h_cval[static_cast<std::size_t>(e) * bs + j];
            ++k;
        }
        glen[static_cast<std::size_t>(p)] = k;
    }
    mg->gs_positions = P;
    mg->gs_k = K;
    upload(mg->gs_len, glen);
    upload(mg->gs_col, gcol);
    upload(mg->gs_val, gval);
    upload(mg->gs_diag, gdiag);
}

double smoother_weight(const owtk_multigrid_options& o, std::uint64_t step)
{
    if (o.smoother != OWTK_MG_CHEBYSHEV) return o.jacobi_weight;
    const std::uint64_t levels = std::max<std::uint64_t>(1, o.chebyshev_levels);
    const double pi = std::acos(-1.0);
    const double rho = 0.99;
    const double angle = pi * static_cast<double>(2 * (step % levels) + 1) / static_cast<double>(2 * levels);
    const double cs = std::cos(angle);
    return 2.0 / (1.0 + std::sqrt(1.0 - rho * rho * cs * cs));
}

int mg_solve(owtk_multigrid_s* mg, const double* b, double* x, const owtk_solver_options* o, int cycle,
             owtk_solver_result* res)
{
    const auto t0 = std::chrono::steady_clock::now();
    init_result(res);
    if (!valid_options(o)) return OWTK_OK;
    if (cycle < OWTK_MG_V || cycle > OWTK_MG_FULL) throw std::invalid_argument("unknown multigrid cycle");
    owtk_matrix m = mg->fine;
    owtk_context ctx = m->ctx;
    const hipStream_t st = ctx->stream;
    const auto scratch = ctx->scratch();
    const int bs = mg->bs;
    const std::int64_t n = mg->n * bs, nc = mg->nc * bs;
    const std::int64_t L = m->local() * bs;
    DeviceArray<double> vecs;
    vecs.allocate(static_cast<std::size_t>(2 * L + 3 * nc));
    kern::fill_zero_bytes(st, vecs.get(), vecs.bytes());
    double* residual = vecs.get();
    double* work = residual + L;
    double* crhs = work + L;
    double* csol = crhs + nc;
    double* capplied = csol + nc;
    double* d_norm = ctx->scalars.get();
    std::uint64_t ops = 0, reds = 0;
    const auto& opt = mg->opt;
    auto grid = [](std::int64_t k) { return dev::stream_grid(std::max<std::int64_t>(k, 1), 4); };
    auto read_norm = [&] {
        double v;
        OWTK_HIP(hipMemcpyAsync(&v, d_norm, sizeof(double), hipMemcpyDeviceToHost, st));
        ctx->sync();
        return v;
    };
    auto finish = [&](int status) {
        res->status = status;
        res->operator_applications = ops;
        res->global_reductions = reds;
        if (o->collect_timings)
            res->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return OWTK_OK;
    };
    auto fine_op = [&](const double* in, double* out) {
        dev::EpiStore<double> e;
        e.y = out;
        spmv_f64(m, in, e);
        ++ops;
    };
    // true_residual (one operator application); with_norm adds detail::norm
    auto true_residual = [&](bool with_norm) {
        dev::EpiResidual e;
        e.y = residual;
        e.b = b;
        e.n = n;
        e.norm_out = with_norm ? d_norm : nullptr;
        spmv_f64(m, x, e);
        ++ops;
        if (!with_norm) return 0.0;
        ++reds;
        return read_norm();
    };
    auto smooth_fine = [&](std::uint64_t steps) {
        for (std::uint64_t s = 0; s < steps; ++s) {
            if (opt.smoother == OWTK_MG_RED_BLACK) {
                fine_op(x, work);
                k_mg_colour<<<grid(mg->n_red * bs), kBlock, 0, st>>>(mg->fine_red.get(), mg->n_red, bs, x, b, work,
                                                                      mg->fine_inv.get());
                fine_op(x, work);
                k_mg_colour<<<grid(mg->n_black * bs), kBlock, 0, st>>>(mg->fine_black.get(), mg->n_black, bs, x, b,
                                                                        work, mg->fine_inv.get());
            } else {
                fine_op(x, work);
                k_mg_jacobi<<<grid(n), kBlock, 0, st>>>(x, b, work, mg->fine_inv.get(), smoother_weight(opt, s), n);
            }
            OWTK_LAUNCH_CHECK();
        }
    };
    // smooth (multigrid.hpp:307-331) on the coarse level; its matrix applies are not counted
    auto smooth_coarse = [&](std::uint64_t steps) {
        for (std::uint64_t s = 0; s < steps; ++s) {
            if (opt.smoother == OWTK_MG_RED_BLACK) {
                // red levels, then black levels (red rows read the previous black values)
                const std::int64_t levels = static_cast<std::int64_t>(mg->gs_level_ptr.size()) - 1;
                const GsArgs ga{mg->gs_rows.get(), mg->gs_len.get(), mg->gs_col.get(), mg->gs_val.get(),
                                mg->gs_diag.get(), crhs, mg->gs_positions, bs, mg->gs_k};
                k_mg_gs_ell<<<1, 1024, 0, st>>>(mg->gs_levels.get(), 0, mg->red_levels, ga, csol);
                k_mg_gs_ell<<<1, 1024, 0, st>>>(mg->gs_levels.get(), mg->red_levels, levels, ga, csol);
            } else {
                dev::EpiStore<double> e;
                e.y = capplied;
                spmv_f64(mg->coarse, csol, e);
                const double w = opt.smoother == OWTK_MG_CHEBYSHEV ? smoother_weight(opt, s) : opt.jacobi_weight;
                k_mg_jacobi<<<grid(nc), kBlock, 0, st>>>(csol, crhs, capplied, mg->coarse_inv.get(), w, nc);
            }
            OWTK_LAUNCH_CHECK();
        }
    };
    auto coarse_correction = [&](const double* fine_rhs, bool zero_x) {
        k_mg_restrict<<<grid(nc), kBlock, 0, st>>>(fine_rhs, crhs, mg->agg_ptr.get(), mg->agg_members.get(),
                                                   mg->agg_size.get(), mg->nc, bs);
        kern::fill_zero_bytes(st, csol, nc * sizeof(double));
        OWTK_LAUNCH_CHECK();
        smooth_coarse(opt.coarse_smoothing_steps);
        if (zero_x) kern::fill_zero_bytes(st, x, n * sizeof(double));
        k_mg_prolong<<<grid(n), kBlock, 0, st>>>(csol, x, mg->f2c.get(), mg->n, bs);
        OWTK_LAUNCH_CHECK();
    };

    kern::norm_to(st, scratch, b, n, d_norm);
    ++reds;
    const double rhs_norm = read_norm();
    if (!std::isfinite(rhs_norm)) {
        res->breakdown_reason = reason_non_finite_scalar;
        return finish(status_breakdown);
    }
    const double thr = threshold_of(rhs_norm, o);
    double nv = true_residual(true);
    res->initial_residual_norm = nv;
    set_residual(res, nv, nv, rhs_norm);
    if (!std::isfinite(nv)) {
        res->breakdown_reason = reason_non_finite_scalar;
        return finish(status_breakdown);
    }
    if (nv <= thr) return finish(status_converged);
    if (cycle == OWTK_MG_FULL) coarse_correction(b, true);
    for (std::uint64_t it = 1; it <= o->maximum_iterations; ++it) {
        res->iterations = it;
        smooth_fine(opt.pre_smoothing_steps);
        const int corrections = cycle == OWTK_MG_W ? 2 : 1;
        for (int k = 0; k < corrections; ++k) {
            true_residual(false);
            coarse_correction(residual, false);
        }
        smooth_fine(opt.post_smoothing_steps);
        nv = true_residual(true);
        set_residual(res, nv, nv, rhs_norm);
        if (!std::isfinite(nv)) {
            res->breakdown_reason = reason_non_finite_scalar;
            return finish(status_breakdown);
        }
        if (nv <= thr) return finish(status_converged);
    }
    return finish(status_maximum_iterations);
}

} // namespace
} // namespace owtk

extern "C" {

void owtk_multigrid_options_default(owtk_multigrid_options* o)
{
    if (!o) return;
    o->pre_smoothing_steps = 2;
    o->post_smoothing_steps = 2;
    o->coarse_smoothing_steps = 24;
    o->target_aggregate_size = 4;
    o->jacobi_weight = 0.8;
    o->smoother = OWTK_MG_JACOBI;
    o->chebyshev_levels = 8;
}

int owtk_multigrid_create(owtk_matrix fine, const owtk_multigrid_options* options, owtk_multigrid* out)
{
    return guarded([&] {
        if (!fine || !options || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        if (fine->kind != OWTK_MATRIX_CSR)
            throw std::invalid_argument("multigrid needs a component-diagonal BlockCsrMatrix (CSR kind)");
        if (options->target_aggregate_size == 0) throw std::invalid_argument("multigrid aggregate size must be positive");
        if (options->smoother < OWTK_MG_JACOBI || options->smoother > OWTK_MG_CHEBYSHEV)
            throw std::invalid_argument("unknown multigrid smoother");
        fine->ctx->activate();
        auto mg = std::make_unique<owtk_multigrid_s>();
        mg->fine = fine;
        mg->ctx = fine->ctx;
        mg->opt = *options;
        mg_setup(mg.get());
        *out = mg.release();
    });
}

int owtk_multigrid_destroy(owtk_multigrid mg)
{
    return guarded([&] {
        if (!mg) return;
        mg->ctx->activate();
        mg->ctx->sync();
        delete mg;
    });
}

int owtk_multigrid_coarse(owtk_multigrid mg, int64_t* nodes, int64_t* entries, int32_t* row_offsets,
                          int32_t* columns, double* values, int64_t* fine_to_coarse)
{
    return guarded([&] {
        if (!mg || !nodes || !entries) throw std::invalid_argument("null argument");
        *nodes = mg->nc;
        *entries = static_cast<int64_t>(mg->h_ccol.size());
        if (row_offsets) std::copy(mg->h_coff.begin(), mg->h_coff.end(), row_offsets);
        if (columns) std::copy(mg->h_ccol.begin(), mg->h_ccol.end(), columns);
        if (values) std::copy(mg->h_cval.begin(), mg->h_cval.end(), values);
        if (fine_to_coarse) std::copy(mg->h_f2c.begin(), mg->h_f2c.end(), fine_to_coarse);
    });
}

int owtk_multigrid_solve(owtk_multigrid mg, const double* b, double* x, const owtk_solver_options* options,
                         int32_t cycle, owtk_solver_result* result)
{
    return guarded([&] {
        if (!mg || !b || !x || !options || !result) throw std::invalid_argument("null argument");
        require_plain_options(options, "multigrid");
        mg->fine->ctx->activate();
        mg_solve(mg, b, x, options, cycle, result);
    });
}

} // extern "C"
