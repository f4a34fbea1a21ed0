// Matrix upload (reference CSR forms -> device SELL-64 / CSR layouts),
// numeric updates, low-precision shadows and the SpMV entry points.
#include "internal.hpp"
#include "kernels/blas1.hpp"
#include "kernels/spmv.hpp"
#include "host/host.hpp"

#include <algorithm>
#include <cmath>

using namespace owtk;

namespace owtk {
namespace {

void validate_structure(std::int64_t owned, std::int64_t ghost, std::int64_t bs, const std::int32_t* offsets,
                        const std::int32_t* columns, const char* what)
{
    // Mirrors the checks of BlockCsrMatrix::validate (block_csr.hpp:281-300).
    if (owned <= 0 || bs <= 0 || ghost < 0 || !offsets)
        throw std::invalid_argument(std::string("invalid ") + what + " storage");
    if (owned + ghost >= (std::int64_t(1) << 31))
        throw std::invalid_argument(std::string(what) + ": node count exceeds int32 indexing");
    if (offsets[0] != 0) throw std::invalid_argument(std::string("invalid ") + what + " storage");
    for (std::int64_t row = 0; row < owned; ++row) {
        if (offsets[row] > offsets[row + 1])
            throw std::invalid_argument(std::string("nonmonotone ") + what + " offsets");
    }
    const std::int64_t entries = offsets[owned];
    if (entries > 0 && !columns) throw std::invalid_argument(std::string("invalid ") + what + " storage");
    for (std::int64_t e = 0; e < entries; ++e) {
        if (columns[e] < 0) throw std::invalid_argument(std::string("negative ") + what + " index");
        if (columns[e] >= owned + ghost)
            throw std::invalid_argument(std::string(what) + " column outside local layout");
    }
}

// Host construction of the SELL-64 structure.
void build_sell(owtk_matrix m)
{
    auto& s = m->s;
    const std::int64_t rows = m->owned;
    s.layout = Layout::sell;
    s.rows = rows;
    s.slices = (rows + 63) / 64;
    std::vector<std::int32_t> slice_ptr(static_cast<std::size_t>(s.slices + 1), 0);
    std::vector<std::int32_t> row_len(static_cast<std::size_t>(s.slices * 64), 0);
    int max_width = 0;
    std::int64_t slots = 0;
    for (std::int64_t sl = 0; sl < s.slices; ++sl) {
        int width = 0;
        for (int lane = 0; lane < 64; ++lane) {
            const std::int64_t r = sl * 64 + lane;
            if (r >= rows) break;
            const int len = m->h_offsets[r + 1] - m->h_offsets[r];
            row_len[r] = len;
            width = std::max(width, len);
        }
        max_width = std::max(max_width, width);
        slots += static_cast<std::int64_t>(width) * 64;
        if (slots >= (std::int64_t(1) << 31)) throw std::invalid_argument("matrix too large for int32 slots");
        slice_ptr[sl + 1] = static_cast<std::int32_t>(slots);
    }
    s.slots = slots;
    s.max_width = max_width;
    std::vector<std::int32_t> cols(static_cast<std::size_t>(std::max<std::int64_t>(slots, 1)), 0);
    s.slot_of_entry.assign(static_cast<std::size_t>(m->entries), 0);
    for (std::int64_t r = 0; r < rows; ++r) {
        const std::int64_t sl = r / 64;
        const int lane = static_cast<int>(r % 64);
        for (std::int32_t e = m->h_offsets[r]; e < m->h_offsets[r + 1]; ++e) {
            const std::int64_t k = e - m->h_offsets[r];
            const std::int64_t slot = slice_ptr[sl] + 64 * k + lane;
            cols[slot] = m->h_columns[e];
            s.slot_of_entry[e] = slot;
        }
    }
    s.slice_ptr.allocate(slice_ptr.size());
    s.row_len.allocate(row_len.size());
    s.columns.allocate(cols.size());
    OWTK_HIP(hipMemcpy(s.slice_ptr.get(), slice_ptr.data(), slice_ptr.size() * 4, hipMemcpyHostToDevice));
    OWTK_HIP(hipMemcpy(s.row_len.get(), row_len.data(), row_len.size() * 4, hipMemcpyHostToDevice));
    OWTK_HIP(hipMemcpy(s.columns.get(), cols.data(), cols.size() * 4, hipMemcpyHostToDevice));
}

void build_csr(owtk_matrix m)
{
    auto& s = m->s;
    s.layout = Layout::csr;
    s.rows = m->owned;
    s.slots = m->entries;
    s.offsets.allocate(m->h_offsets.size());
    OWTK_HIP(hipMemcpy(s.offsets.get(), m->h_offsets.data(), m->h_offsets.size() * 4, hipMemcpyHostToDevice));
    s.columns.allocate(std::max<std::size_t>(m->h_columns.size(), 1));
    if (!m->h_columns.empty())
        OWTK_HIP(hipMemcpy(s.columns.get(), m->h_columns.data(), m->h_columns.size() * 4, hipMemcpyHostToDevice));
    s.slot_of_entry.resize(static_cast<std::size_t>(m->entries));
    for (std::int64_t e = 0; e < m->entries; ++e) s.slot_of_entry[e] = e;
    int max_width = 0;
    for (std::int64_t r = 0; r < m->owned; ++r)
        max_width = std::max(max_width, m->h_offsets[r + 1] - m->h_offsets[r]);
    s.max_width = max_width;
}

std::int64_t block_values(owtk_matrix m) { return m->dense() ? std::int64_t(m->bs) * m->bs : m->bs; }

// Scatter reference-layout values into the device layout and upload them.
void upload_values(owtk_matrix m, const double* values, const double* diagonal)
{
    const std::int64_t bv = block_values(m);
    const std::int64_t count = m->s.layout == Layout::sell ? std::max<std::int64_t>(m->s.slots, 1) * bv
                                                           : std::max<std::int64_t>(m->entries, 1) * bv;
    std::vector<double> staged(static_cast<std::size_t>(count), 0.0);
    if (m->entries > 0 && !values) throw std::invalid_argument("null matrix values");
    host::parallel_for(static_cast<std::size_t>(m->entries), [&](std::size_t b, std::size_t e) {
        for (std::size_t en = b; en < e; ++en) {
            const std::int64_t slot = m->s.slot_of_entry[en];
            if (m->dense()) {
                if (m->s.layout == Layout::sell)
                    for (std::int64_t j = 0; j < bv; ++j)
                        staged[dev::dense_index(slot, static_cast<int>(j), static_cast<int>(bv))] = values[en * bv + j];
                else
                    std::copy_n(values + en * bv, bv, staged.data() + slot * bv);
            } else {
                for (int c = 0; c < m->bs; ++c) staged[m->value_index(slot, c)] = values[en * bv + c];
            }
        }
    });
    m->value_count = count;
    if (m->v64.count != static_cast<std::size_t>(count)) m->v64.allocate(static_cast<std::size_t>(count));
    OWTK_HIP(hipMemcpy(m->v64.get(), staged.data(), staged.size() * sizeof(double), hipMemcpyHostToDevice));
    if (m->has_diag_array()) {
        if (!diagonal) throw std::invalid_argument("null split diagonal");
        const std::int64_t dcount = m->s.layout == Layout::sell ? m->s.slices * 64 * m->bs : m->owned * m->bs;
        std::vector<double> dstaged(static_cast<std::size_t>(dcount), 0.0);
        for (std::int64_t r = 0; r < m->owned; ++r)
            for (int c = 0; c < m->bs; ++c) dstaged[m->diag_index(r, c)] = diagonal[r * m->bs + c];
        m->diag_count = dcount;
        if (m->d64.count != static_cast<std::size_t>(dcount)) m->d64.allocate(static_cast<std::size_t>(dcount));
        OWTK_HIP(hipMemcpy(m->d64.get(), dstaged.data(), dstaged.size() * sizeof(double), hipMemcpyHostToDevice));
    }
}

void refresh_shadows(owtk_matrix m)
{
    const hipStream_t st = m->ctx->stream;
    if (m->v32.get()) {
        kern::convert(st, m->v32.get(), m->v64.get(), static_cast<std::size_t>(m->value_count));
        if (m->d64.get()) kern::convert(st, m->d32.get(), m->d64.get(), static_cast<std::size_t>(m->diag_count));
    }
    if (m->v16.get()) {
        const std::int64_t n = m->value_count;
        kern::k_to_bf16<<<dev::stream_grid(n, 4), dev::kBlock, 0, st>>>(m->v16.get(), m->v64.get(), n);
        OWTK_LAUNCH_CHECK();
        if (m->d64.get()) {
            kern::k_to_bf16<<<dev::stream_grid(m->diag_count, 4), dev::kBlock, 0, st>>>(m->d16.get(), m->d64.get(),
                                                                                       m->diag_count);
            OWTK_LAUNCH_CHECK();
        }
    }
    m->ctx->sync();
}

owtk_matrix create_matrix(owtk_context ctx, int kind, std::int64_t owned, std::int64_t ghost, std::int64_t bs,
                          const std::int32_t* offsets, const std::int32_t* columns, const double* values,
                          const double* diagonal)
{
    if (!ctx) throw std::invalid_argument("null context");
    const char* what = kind == OWTK_MATRIX_SPLIT ? "split CSR" : (kind == OWTK_MATRIX_BSR ? "dense-block CSR" : "BlockCsrMatrix");
    validate_structure(owned, ghost, bs, offsets, columns, what);
    if (bs > 4096) throw std::invalid_argument("block size too large");
    ctx->activate();
    auto m = std::make_unique<owtk_matrix_s>();
    m->ctx = ctx;
    m->kind = kind;
    m->bs = static_cast<int>(bs);
    m->owned = owned;
    m->ghost = ghost;
    m->entries = offsets[owned];
    m->h_offsets.assign(offsets, offsets + owned + 1);
    m->h_columns.assign(columns, columns + m->entries);
    if (kind == OWTK_MATRIX_BSR && bs > 16) throw unsupported_error("dense block size above 16");
    const bool sell = (kind == OWTK_MATRIX_BSR) ? (bs <= 4) : (bs <= 4);
    if (sell) build_sell(m.get());
    else build_csr(m.get());
    // The SELL SpMVs launch one block per 256 rows and fuse the solvers'
    // reductions into the same pass: the context's partials must span them.
    if (sell) ctx->reserve_partials((owned + dev::kBlock - 1) / dev::kBlock);
    upload_values(m.get(), values, diagonal);
    return m.release();
}

} // namespace

std::int64_t spmv_bytes(owtk_matrix m, int sv, int sx)
{
    const std::int64_t n = m->owned, b = m->bs;
    if (m->dense()) return m->entries * (b * b * sv + 4) + (n + 1) * 4 + 2 * n * b * sx;
    const std::int64_t diag = m->has_diag_array() ? n * b * sv : 0;
    return m->entries * (b * sv + 4) + diag + n * 4 + 2 * n * b * sx;
}

} // namespace owtk

extern "C" {

int owtk_matrix_create_csr(owtk_context ctx, int64_t owned_nodes, int64_t ghost_nodes, int64_t block_size,
                           const int32_t* row_offsets, const int32_t* columns, const double* values, owtk_matrix* out)
{
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output");
        *out = create_matrix(ctx, OWTK_MATRIX_CSR, owned_nodes, ghost_nodes, block_size, row_offsets, columns,
                             values, nullptr);
    });
}

int owtk_matrix_create_split(owtk_context ctx, int64_t owned_nodes, int64_t ghost_nodes, int64_t block_size,
                             const int32_t* off_diagonal_row_offsets, const int32_t* off_diagonal_columns,
                             const double* diagonal, const double* off_diagonal, owtk_matrix* out)
{
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output");
        *out = create_matrix(ctx, OWTK_MATRIX_SPLIT, owned_nodes, ghost_nodes, block_size, off_diagonal_row_offsets,
                             off_diagonal_columns, off_diagonal, diagonal);
    });
}

int owtk_matrix_create_bsr(owtk_context ctx, int64_t owned_nodes, int64_t ghost_nodes, int64_t block_size,
                           const int32_t* row_offsets, const int32_t* columns, const double* values, owtk_matrix* out)
{
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output");
        *out = create_matrix(ctx, OWTK_MATRIX_BSR, owned_nodes, ghost_nodes, block_size, row_offsets, columns, values,
                             nullptr);
    });
}

int owtk_matrix_update_values(owtk_matrix m, const double* values, const double* off_diagonal)
{
    return guarded([&] {
        if (!m) throw std::invalid_argument("null matrix");
        m->ctx->activate();
        m->ctx->sync();
        if (m->has_diag_array()) upload_values(m, off_diagonal, values);
        else upload_values(m, values, nullptr);
        refresh_shadows(m);
        ++m->value_epoch;
    });
}

int owtk_matrix_add_storage(owtk_matrix m, owtk_dtype dtype)
{
    return guarded([&] {
        if (!m) throw std::invalid_argument("null matrix");
        m->ctx->activate();
        if (dtype == OWTK_F32) {
            if (!m->v32.get()) {
                m->v32.allocate(static_cast<std::size_t>(m->value_count));
                if (m->d64.get()) m->d32.allocate(static_cast<std::size_t>(m->diag_count));
            }
        } else if (dtype == OWTK_BF16) {
            if (!m->v16.get()) {
                m->v16.allocate(static_cast<std::size_t>(m->value_count));
                if (m->d64.get()) m->d16.allocate(static_cast<std::size_t>(m->diag_count));
            }
        } else if (dtype != OWTK_F64) {
            throw std::invalid_argument("unknown storage type");
        }
        refresh_shadows(m);
    });
}

int owtk_matrix_info_get(owtk_matrix m, owtk_matrix_info* info)
{
    return guarded([&] {
        if (!m || !info) throw std::invalid_argument("null argument");
        info->kind = m->kind;
        info->block_size = m->bs;
        info->owned_nodes = m->owned;
        info->ghost_nodes = m->ghost;
        info->entries = m->entries;
        info->device_bytes = static_cast<int64_t>(m->v64.bytes() + m->v32.bytes() + m->v16.bytes() + m->d64.bytes()
                                                  + m->d32.bytes() + m->d16.bytes() + m->s.columns.bytes()
                                                  + m->s.row_len.bytes() + m->s.slice_ptr.bytes()
                                                  + m->s.offsets.bytes());
        info->spmv_bytes_f64 = spmv_bytes(m, 8, 8);
        info->spmv_bytes_f32 = spmv_bytes(m, 4, 4);
        info->spmv_bytes_bf16 = spmv_bytes(m, 2, 4);
        info->has_f32 = m->v32.get() != nullptr;
        info->has_bf16 = m->v16.get() != nullptr;
    });
}

int owtk_matrix_destroy(owtk_matrix m)
{
    return guarded([&] {
        if (!m) return;
        m->ctx->activate();
        m->ctx->sync();
        delete m;
    });
}

int owtk_spmv(owtk_matrix m, owtk_dtype dtype, const void* x, void* y)
{
    return guarded([&] {
        if (!m) throw std::invalid_argument("null matrix");
        if (!x || !y) throw std::invalid_argument("null vector");
        m->ctx->activate();
        if (dtype == OWTK_F64) {
            dev::EpiStore<double> e;
            e.y = static_cast<double*>(y);
            spmv_f64(m, static_cast<const double*>(x), e);
        } else if (dtype == OWTK_F32 || dtype == OWTK_BF16) {
            dev::EpiStore<float> e;
            e.y = static_cast<float*>(y);
            spmv_low(m, dtype, static_cast<const float*>(x), e);
        } else {
            throw std::invalid_argument("unknown dtype");
        }
    });
}

int owtk_residual(owtk_matrix m, const double* b, const double* x, double* r)
{
    return guarded([&] {
        if (!m) throw std::invalid_argument("null matrix");
        if (!b || !x || !r) throw std::invalid_argument("null vector");
        m->ctx->activate();
        dev::EpiResidual e;
        e.y = r;
        e.b = b;
        e.n = m->owned * m->bs;
        e.norm_out = nullptr;
        spmv_f64(m, x, e);
    });
}

} // extern "C"
