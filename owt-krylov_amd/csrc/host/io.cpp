// Binary matrix and factor files in the reference's fixture format
// (benchmarks/matrix_analysis/ilu_fill.cpp:20-44): little-endian
//   system:  u64 {n, bins, nnz}, u64 offsets[n+1], u64 columns[nnz],
//            f64 values[nnz * bins]            (BlockCsrMatrix, [entry][comp])
//   factors: u64 {n, bins, factor_nnz}, u64 factor_row_offsets[n+1],
//            u64 factor_columns[factor_nnz], f64 factor_values[factor_nnz * bins],
//            f64 inverse_diagonal[n * bins]
// so application-exported systems run on the device path and the device ILU
// factors can be compared with the reference's own dumps; and the
// application's per-rank final state files (SWSTATE1).
#include "../owtk_hip_api.hpp"
#include "host.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct File {
    std::FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode))
    {
        if (!f) throw std::invalid_argument(std::string("cannot open ") + path);
    }
    ~File()
    {
        if (f) std::fclose(f);
    }
    template<class T>
    void read(T* out, std::size_t count)
    {
        if (count && std::fread(out, sizeof(T), count, f) != count) throw std::invalid_argument("truncated matrix file");
    }
    template<class T>
    void write(const T* in, std::size_t count)
    {
        if (count && std::fwrite(in, sizeof(T), count, f) != count) throw std::runtime_error("short write");
    }
    long size()
    {
        const long here = std::ftell(f);
        std::fseek(f, 0, SEEK_END);
        const long end = std::ftell(f);
        std::fseek(f, here, SEEK_SET);
        return end;
    }
};

} // namespace

extern "C" {

int owtk_read_block_csr(const char* path, int64_t* nodes, int64_t* block_size, int64_t* entries,
                        int32_t* row_offsets, int32_t* columns, double* values)
{
    return owtk::guarded([&] {
        if (!path || !nodes || !block_size || !entries) throw std::invalid_argument("null argument");
        File file(path, "rb");
        std::uint64_t header[3];
        file.read(header, 3);
        const std::uint64_t n = header[0], bins = header[1], nnz = header[2];
        if (!n || !bins || nnz < n) throw std::invalid_argument("invalid matrix file dimensions");
        if (n > 0x7fffffffULL || nnz > 0x7fffffffULL) throw std::invalid_argument("matrix exceeds int32 indices");
        const unsigned long long expect = 24ULL + 8ULL * (n + 1) + 8ULL * nnz + 8ULL * nnz * bins;
        if (static_cast<unsigned long long>(file.size()) != expect)
            throw std::invalid_argument("matrix file size does not match its header");
        *nodes = static_cast<int64_t>(n);
        *block_size = static_cast<int64_t>(bins);
        *entries = static_cast<int64_t>(nnz);
        if (!row_offsets && !columns && !values) return;  // size query
        if (!row_offsets || !columns || !values) throw std::invalid_argument("null output array");
        std::vector<std::uint64_t> buf(n + 1);
        file.read(buf.data(), n + 1);
        if (buf[0] != 0 || buf[n] != nnz) throw std::invalid_argument("row offsets do not span the entries");
        for (std::uint64_t i = 0; i <= n; ++i) {
            if (i && buf[i] < buf[i - 1]) throw std::invalid_argument("row offsets are not monotone");
            row_offsets[i] = static_cast<int32_t>(buf[i]);
        }
        buf.resize(nnz);
        file.read(buf.data(), nnz);
        for (std::uint64_t e = 0; e < nnz; ++e) {
            if (buf[e] >= n) throw std::invalid_argument("column index out of range");
            columns[e] = static_cast<int32_t>(buf[e]);
        }
        file.read(values, nnz * bins);
    });
}

int owtk_write_block_csr(const char* path, int64_t nodes, int64_t block_size, const int32_t* row_offsets,
                         const int32_t* columns, const double* values)
{
    return owtk::guarded([&] {
        if (!path || !row_offsets || !columns || !values) throw std::invalid_argument("null argument");
        if (nodes <= 0 || block_size <= 0) throw std::invalid_argument("invalid dimensions");
        const std::uint64_t n = static_cast<std::uint64_t>(nodes), bins = static_cast<std::uint64_t>(block_size);
        const std::uint64_t nnz = static_cast<std::uint64_t>(row_offsets[nodes]);
        File file(path, "wb");
        const std::uint64_t header[3] = {n, bins, nnz};
        file.write(header, 3);
        std::vector<std::uint64_t> buf(n + 1);
        for (std::uint64_t i = 0; i <= n; ++i) buf[i] = static_cast<std::uint64_t>(row_offsets[i]);
        file.write(buf.data(), n + 1);
        buf.resize(nnz);
        for (std::uint64_t e = 0; e < nnz; ++e) buf[e] = static_cast<std::uint64_t>(columns[e]);
        file.write(buf.data(), nnz);
        file.write(values, nnz * bins);
    });
}

int owtk_write_ilu_factors(owtk_precond pc, const char* path)
{
    return owtk::guarded([&] {
        if (!pc || !path) throw std::invalid_argument("null argument");
        int64_t entries = 0;
        int rc = owtk_precond_ilu_factors(pc, &entries, nullptr, nullptr, nullptr, nullptr);
        if (rc != OWTK_OK) throw std::invalid_argument("not an ILU preconditioner");
        owtk_precond_info info;
        rc = owtk_precond_info_get(pc, &info);
        if (rc != OWTK_OK) throw std::runtime_error("preconditioner info");
        const int64_t n = info.rows;
        const int64_t bins = info.block_size;
        std::vector<int64_t> off(n + 1), col(entries);
        std::vector<double> val(entries * bins), inv(n * bins);
        rc = owtk_precond_ilu_factors(pc, &entries, off.data(), col.data(), val.data(), inv.data());
        if (rc != OWTK_OK) throw std::runtime_error("reading the ILU factors failed");
        File file(path, "wb");
        const std::uint64_t header[3] = {static_cast<std::uint64_t>(n), static_cast<std::uint64_t>(bins),
                                         static_cast<std::uint64_t>(entries)};
        file.write(header, 3);
        std::vector<std::uint64_t> buf(off.begin(), off.end());
        file.write(buf.data(), buf.size());
        buf.assign(col.begin(), col.end());
        file.write(buf.data(), buf.size());
        file.write(val.data(), val.size());
        file.write(inv.data(), inv.size());
    });
}

// Per-rank final state files of the application (specwave_limon.py:66-100):
//   "SWSTATE1", u64 {width, owned, ns, nd, count}, u64 ids[owned],
//   f{width} values[owned][ns*nd]
// The files of all ranks together must cover global vertices 0..count-1
// exactly once; values must be finite and non-negative (a spectrum).  The
// output is ordered by global id, FP32 files widened to FP64.
int owtk_read_state(const char* const* paths, int32_t files, int64_t ns, int64_t nd, int64_t* vertices,
                    double* values)
{
    return owtk::guarded([&] {
        if (!paths || !vertices || files < 0) throw std::invalid_argument("null argument");
        if (files == 0) throw std::invalid_argument("no final state files");
        if (ns <= 0 || nd <= 0) throw std::invalid_argument("invalid state dimensions");
        const std::uint64_t comps = static_cast<std::uint64_t>(ns) * static_cast<std::uint64_t>(nd);
        std::uint64_t global = 0, width = 0;
        std::vector<std::uint64_t> ids;
        std::vector<double> vals;
        for (int32_t f = 0; f < files; ++f) {
            if (!paths[f]) throw std::invalid_argument("null path");
            File file(paths[f], "rb");
            char magic[8];
            if (std::fread(magic, 1, 8, file.f) != 8 || std::memcmp(magic, "SWSTATE1", 8) != 0)
                throw std::invalid_argument(std::string("invalid state magic: ") + paths[f]);
            std::uint64_t h[5];
            if (std::fread(h, sizeof(std::uint64_t), 5, file.f) != 5)
                throw std::invalid_argument(std::string("truncated state header: ") + paths[f]);
            const std::uint64_t w = h[0], owned = h[1], count = h[4];
            if ((w != 4 && w != 8) || h[2] != static_cast<std::uint64_t>(ns) || h[3] != static_cast<std::uint64_t>(nd)
                || count == 0 || owned > count || (global && count != global) || (width && w != width))
                throw std::invalid_argument(std::string("inconsistent state dimensions: ") + paths[f]);
            const unsigned long long expect = 48ULL + owned * (8ULL + comps * w);
            if (static_cast<unsigned long long>(file.size()) != expect)
                throw std::invalid_argument(std::string("invalid state file size: ") + paths[f]);
            global = count;
            width = w;
            const std::size_t first = ids.size();
            ids.resize(first + owned);
            file.read(ids.data() + first, owned);
            const std::size_t vfirst = vals.size();
            vals.resize(vfirst + owned * comps);
            if (w == 8) {
                file.read(vals.data() + vfirst, owned * comps);
            } else {
                std::vector<float> tmp(owned * comps);
                file.read(tmp.data(), tmp.size());
                for (std::size_t i = 0; i < tmp.size(); ++i) vals[vfirst + i] = tmp[i];
            }
        }
        if (ids.size() != global) throw std::invalid_argument("final state does not cover every global vertex exactly once");
        std::vector<std::int64_t> where(global, -1);
        for (std::size_t i = 0; i < ids.size(); ++i) {
            if (ids[i] >= global || where[ids[i]] >= 0)
                throw std::invalid_argument("final state does not cover every global vertex exactly once");
            where[ids[i]] = static_cast<std::int64_t>(i);
        }
        for (const double v : vals)
            if (!std::isfinite(v) || v < 0) throw std::invalid_argument("nonfinite or negative final spectrum");
        *vertices = static_cast<int64_t>(global);
        if (!values) return;  // size query
        for (std::uint64_t g = 0; g < global; ++g)
            std::memcpy(values + g * comps, vals.data() + static_cast<std::uint64_t>(where[g]) * comps,
                        comps * sizeof(double));
    });
}

} // extern "C"
