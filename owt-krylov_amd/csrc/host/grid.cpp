// Synthetic triangulated grids and node orderings (host only).
//
// Grid: nx x ny nodes, natural id u = j*nx + i, spacing h = 1.  Every quad is
// split along its (i,j)-(i+1,j+1) diagonal, so node u couples to the six
// neighbours E, W, N, S, NE, SW (a P1 7-point vertex stencil).  Nodes are
// jittered by U(-jitter, jitter) h, optionally permuted by a seeded random
// permutation, then renumbered by RCM or Hilbert order.
//
// Every random value is a counter-based draw keyed by the natural ids of the
// nodes it belongs to, so the generated operator is the same physical
// operator under any numbering: A_final = P A_natural P^T.
#include "host.hpp"
#include "../owtk_hip_api.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <utility>

namespace owtk::host {

double draw_normal(std::uint64_t seed, std::uint64_t stream, std::uint64_t counter)
{
    const double u1 = (static_cast<double>(draw_bits(seed, stream, 2 * counter) >> 11) + 0.5)
        * 0x1.0p-53;
    const double u2 = static_cast<double>(draw_bits(seed, stream, 2 * counter + 1) >> 11)
        * 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
}

void parallel_for(std::size_t n, const std::function<void(std::size_t, std::size_t)>& body)
{
    const std::size_t hw = std::max<std::size_t>(1, std::thread::hardware_concurrency());
    const std::size_t workers = std::min<std::size_t>(std::min<std::size_t>(hw, 16),
                                                      std::max<std::size_t>(1, n / 4096));
    if (workers <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> threads;
    const std::size_t chunk = (n + workers - 1) / workers;
    for (std::size_t w = 0; w < workers; ++w) {
        const std::size_t begin = w * chunk;
        const std::size_t end = std::min(n, begin + chunk);
        if (begin >= end) break;
        threads.emplace_back([&body, begin, end] { body(begin, end); });
    }
    for (auto& t : threads) t.join();
}

// Classic xy -> d Hilbert index on an extent x extent grid (extent a power of
// two), the same curve as detail::hilbert_index (ordering.hpp:64-77).
std::uint64_t hilbert_index(std::uint32_t x, std::uint32_t y, std::uint32_t extent)
{
    std::uint64_t d = 0;
    for (std::uint32_t s = extent / 2; s > 0; s /= 2) {
        const std::uint32_t rx = (x & s) ? 1U : 0U;
        const std::uint32_t ry = (y & s) ? 1U : 0U;
        d += static_cast<std::uint64_t>(s) * s * ((3U * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

// hilbert_order (ordering.hpp:231-261): normalise to the bounding box, quantise
// to a 2^16 grid, sort by (curve index, node).
std::vector<std::int64_t> hilbert_order(std::int64_t n, const double* xy)
{
    std::vector<std::int64_t> order;
    if (n <= 0) return order;
    double x_min = xy[0], x_max = xy[0], y_min = xy[1], y_max = xy[1];
    for (std::int64_t i = 0; i < n; ++i) {
        x_min = std::min(x_min, xy[2 * i]);
        x_max = std::max(x_max, xy[2 * i]);
        y_min = std::min(y_min, xy[2 * i + 1]);
        y_max = std::max(y_max, xy[2 * i + 1]);
    }
    constexpr std::uint32_t extent = 1U << 16U;
    const double eps = 2.220446049250313080847e-16;
    const double wx = std::max(x_max - x_min, eps);
    const double wy = std::max(y_max - y_min, eps);
    std::vector<std::pair<std::uint64_t, std::int64_t>> keyed(static_cast<std::size_t>(n));
    parallel_for(static_cast<std::size_t>(n), [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) {
            const double nx = std::clamp((xy[2 * i] - x_min) / wx, 0.0, 1.0);
            const double ny = std::clamp((xy[2 * i + 1] - y_min) / wy, 0.0, 1.0);
            const auto qx = static_cast<std::uint32_t>(nx * (extent - 1));
            const auto qy = static_cast<std::uint32_t>(ny * (extent - 1));
            keyed[i] = {hilbert_index(qx, qy, extent), static_cast<std::int64_t>(i)};
        }
    });
    std::sort(keyed.begin(), keyed.end());
    order.resize(static_cast<std::size_t>(n));
    for (std::size_t i = 0; i < keyed.size(); ++i) order[i] = keyed[i].second;
    return order;
}

// reverse_cuthill_mckee_order (ordering.hpp:138-180).  Graph: distinct owned
// off-diagonal neighbours.  Seed: lowest-numbered unvisited node of minimum
// degree.  Unvisited neighbours are enqueued by (degree, index): the
// reference sorts an index-ascending list by degree with std::sort, which for
// the short lists of a mesh stencil (< 16 entries, insertion sort in
// libstdc++) keeps equal degrees in index order; this definition fixes that
// tie-break for any list length.
std::vector<std::int64_t> rcm_order(std::int64_t n, const std::int32_t* offsets,
                                    const std::int32_t* columns)
{
    std::vector<std::vector<std::int64_t>> graph(static_cast<std::size_t>(n));
    for (std::int64_t row = 0; row < n; ++row) {
        auto& g = graph[static_cast<std::size_t>(row)];
        for (std::int32_t e = offsets[row]; e < offsets[row + 1]; ++e) {
            const std::int64_t c = columns[e];
            if (c < n && c != row) g.push_back(c);
        }
        std::sort(g.begin(), g.end());
        g.erase(std::unique(g.begin(), g.end()), g.end());
    }
    std::vector<char> visited(static_cast<std::size_t>(n), 0);
    std::vector<std::int64_t> order;
    order.reserve(static_cast<std::size_t>(n));
    std::vector<std::int64_t> neighbors;
    while (static_cast<std::int64_t>(order.size()) < n) {
        std::int64_t seed = n;
        for (std::int64_t node = 0; node < n; ++node) {
            if (!visited[node]
                && (seed == n || graph[node].size() < graph[seed].size())) {
                seed = node;
            }
        }
        std::queue<std::int64_t> queue;
        queue.push(seed);
        visited[seed] = 1;
        while (!queue.empty()) {
            const std::int64_t node = queue.front();
            queue.pop();
            order.push_back(node);
            neighbors.clear();
            for (const std::int64_t nb : graph[node]) {
                if (!visited[nb]) neighbors.push_back(nb);
            }
            std::stable_sort(neighbors.begin(), neighbors.end(),
                             [&](std::int64_t a, std::int64_t b) {
                                 return graph[a].size() < graph[b].size();
                             });
            for (const std::int64_t nb : neighbors) {
                if (!visited[nb]) {
                    visited[nb] = 1;
                    queue.push(nb);
                }
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

namespace {

constexpr int kSlots = 6;
constexpr int kDi[kSlots] = {1, -1, 0, 0, 1, -1};
constexpr int kDj[kSlots] = {0, 0, 1, -1, 1, -1};

struct Grid {
    owtk_grid_spec spec;
    std::int64_t n = 0;
    std::vector<std::int64_t> final_to_natural;
    std::vector<std::int64_t> natural_to_final;
    std::vector<double> coords; // natural order
    std::vector<double> field;  // natural order, lognormal-law Gaussian field
};

bool neighbor(const owtk_grid_spec& s, std::int64_t u, int slot, std::int64_t* v)
{
    const std::int64_t i = u % s.nx + kDi[slot];
    const std::int64_t j = u / s.nx + kDj[slot];
    if (i < 0 || j < 0 || i >= s.nx || j >= s.ny) return false;
    *v = j * s.nx + i;
    return true;
}

// Undirected edge key: positive-direction slots (0, 2, 4) own the edge.
std::uint64_t undirected_key(std::int64_t u, int slot, std::int64_t v)
{
    return (slot % 2 == 0) ? static_cast<std::uint64_t>(u) * 8 + slot
                           : static_cast<std::uint64_t>(v) * 8 + (slot - 1);
}

// Unit-variance Gaussian field with Gaussian correlation of length `corr`
// (standard deviation of the correlation function, in units of h): white
// noise filtered by a separable Gaussian of width corr/sqrt(2), truncated at
// three widths, normalised node by node to unit variance.
std::vector<double> gaussian_field(const owtk_grid_spec& s)
{
    const std::int64_t nx = s.nx, ny = s.ny;
    const double width = std::max(1e-3, s.field_correlation / std::sqrt(2.0));
    const int radius = std::max(1, static_cast<int>(std::ceil(3.0 * width)));
    std::vector<double> kernel(2 * radius + 1);
    for (int k = -radius; k <= radius; ++k)
        kernel[k + radius] = std::exp(-0.5 * (k / width) * (k / width));
    std::vector<double> noise(static_cast<std::size_t>(nx * ny));
    parallel_for(noise.size(), [&](std::size_t b, std::size_t e) {
        for (std::size_t u = b; u < e; ++u) noise[u] = draw_normal(s.seed, stream_field, u);
    });
    std::vector<double> tmp(noise.size()), out(noise.size());
    std::vector<double> varx(static_cast<std::size_t>(nx)), vary(static_cast<std::size_t>(ny));
    for (std::int64_t i = 0; i < nx; ++i) {
        double v = 0;
        for (int k = -radius; k <= radius; ++k)
            if (i + k >= 0 && i + k < nx) v += kernel[k + radius] * kernel[k + radius];
        varx[i] = v;
    }
    for (std::int64_t j = 0; j < ny; ++j) {
        double v = 0;
        for (int k = -radius; k <= radius; ++k)
            if (j + k >= 0 && j + k < ny) v += kernel[k + radius] * kernel[k + radius];
        vary[j] = v;
    }
    parallel_for(static_cast<std::size_t>(ny), [&](std::size_t b, std::size_t e) {
        for (std::size_t j = b; j < e; ++j)
            for (std::int64_t i = 0; i < nx; ++i) {
                double acc = 0;
                for (int k = -radius; k <= radius; ++k)
                    if (i + k >= 0 && i + k < nx)
                        acc += kernel[k + radius] * noise[j * nx + i + k];
                tmp[j * nx + i] = acc;
            }
    });
    parallel_for(static_cast<std::size_t>(ny), [&](std::size_t b, std::size_t e) {
        for (std::size_t j = b; j < e; ++j)
            for (std::int64_t i = 0; i < nx; ++i) {
                double acc = 0;
                for (int k = -radius; k <= radius; ++k) {
                    const std::int64_t jj = static_cast<std::int64_t>(j) + k;
                    if (jj >= 0 && jj < ny) acc += kernel[k + radius] * tmp[jj * nx + i];
                }
                out[j * nx + i] = acc / std::sqrt(varx[i] * vary[j]);
            }
    });
    return out;
}

void build_numbering(Grid& g)
{
    const owtk_grid_spec& s = g.spec;
    g.n = static_cast<std::int64_t>(s.nx) * s.ny;
    const std::size_t n = static_cast<std::size_t>(g.n);
    g.coords.resize(2 * n);
    parallel_for(n, [&](std::size_t b, std::size_t e) {
        for (std::size_t u = b; u < e; ++u) {
            const double i = static_cast<double>(u % s.nx);
            const double j = static_cast<double>(u / s.nx);
            g.coords[2 * u] = i + s.jitter * (2.0 * draw_uniform(s.seed, stream_jitter_x, u) - 1.0);
            g.coords[2 * u + 1] = j + s.jitter * (2.0 * draw_uniform(s.seed, stream_jitter_y, u) - 1.0);
        }
    });
    // permuted index p -> natural id
    std::vector<std::int64_t> perm_to_nat(n);
    std::iota(perm_to_nat.begin(), perm_to_nat.end(), 0);
    if (s.permute) {
        std::vector<std::pair<std::uint64_t, std::int64_t>> keys(n);
        for (std::size_t u = 0; u < n; ++u)
            keys[u] = {draw_bits(s.seed, stream_permutation, u), static_cast<std::int64_t>(u)};
        std::sort(keys.begin(), keys.end());
        for (std::size_t p = 0; p < n; ++p) perm_to_nat[p] = keys[p].second;
    }
    std::vector<std::int64_t> nat_to_perm(n);
    for (std::size_t p = 0; p < n; ++p) nat_to_perm[perm_to_nat[p]] = static_cast<std::int64_t>(p);

    std::vector<std::int64_t> new_to_old; // over permuted indices
    if (s.ordering == OWTK_ORDER_RCM) {
        std::vector<std::int32_t> offsets(n + 1, 0), columns;
        columns.reserve(n * kSlots);
        for (std::size_t p = 0; p < n; ++p) {
            const std::int64_t u = perm_to_nat[p];
            for (int slot = 0; slot < kSlots; ++slot) {
                std::int64_t v;
                if (neighbor(s, u, slot, &v)) columns.push_back(static_cast<std::int32_t>(nat_to_perm[v]));
            }
            std::sort(columns.begin() + offsets[p], columns.end());
            offsets[p + 1] = static_cast<std::int32_t>(columns.size());
        }
        new_to_old = rcm_order(g.n, offsets.data(), columns.data());
    } else if (s.ordering == OWTK_ORDER_HILBERT) {
        std::vector<double> xy(2 * n);
        for (std::size_t p = 0; p < n; ++p) {
            xy[2 * p] = g.coords[2 * perm_to_nat[p]];
            xy[2 * p + 1] = g.coords[2 * perm_to_nat[p] + 1];
        }
        new_to_old = hilbert_order(g.n, xy.data());
    } else {
        new_to_old.resize(n);
        std::iota(new_to_old.begin(), new_to_old.end(), 0);
    }
    g.final_to_natural.resize(n);
    g.natural_to_final.resize(n);
    for (std::size_t f = 0; f < n; ++f) {
        g.final_to_natural[f] = perm_to_nat[new_to_old[f]];
        g.natural_to_final[g.final_to_natural[f]] = static_cast<std::int64_t>(f);
    }
    if (s.coefficient_law == OWTK_LAW_LOGNORMAL_FIELD) g.field = gaussian_field(s);
}

// Off-diagonal coefficient of directed edge u -> v (slot) and component c.
double off_diagonal_value(const Grid& g, std::int64_t u, int slot, std::int64_t v, int c)
{
    const owtk_grid_spec& s = g.spec;
    const std::uint64_t key = (static_cast<std::uint64_t>(u) * 8 + slot) * 64 + c;
    const double magnitude = 0.1 + 0.9 * draw_uniform(s.seed, stream_magnitude, key);
    const double perturb = 1.0 + 0.3 * (2.0 * draw_uniform(s.seed, stream_perturbation, key) - 1.0);
    double value = -magnitude * perturb;
    if (s.coefficient_law == OWTK_LAW_LOGNORMAL_FIELD) {
        value *= std::exp(s.field_sigma * 0.5 * (g.field[u] + g.field[v]));
    } else if (s.coefficient_law == OWTK_LAW_EDGE_CONTRAST) {
        value *= std::exp(s.field_sigma
                          * draw_normal(s.seed, stream_edge_contrast, undirected_key(u, slot, v)));
    }
    return value;
}

struct RowEntry {
    std::int64_t column;
    std::int64_t natural;
    int slot;
};

void sorted_row(const Grid& g, std::int64_t f, RowEntry* out, int* count)
{
    const std::int64_t u = g.final_to_natural[f];
    int k = 0;
    for (int slot = 0; slot < kSlots; ++slot) {
        std::int64_t v;
        if (neighbor(g.spec, u, slot, &v)) out[k++] = {g.natural_to_final[v], v, slot};
    }
    std::sort(out, out + k, [](const RowEntry& a, const RowEntry& b) { return a.column < b.column; });
    *count = k;
}

std::int64_t edge_count(const owtk_grid_spec& s)
{
    const std::int64_t nx = s.nx, ny = s.ny;
    return (nx - 1) * ny + nx * (ny - 1) + (nx - 1) * (ny - 1);
}

bool valid_spec(const owtk_grid_spec* s)
{
    return s != nullptr && s->nx >= 2 && s->ny >= 2 && s->block_size >= 1
        && s->block_size <= 64 && static_cast<std::int64_t>(s->nx) * s->ny < (1LL << 30)
        && 2 * edge_count(*s) < (1LL << 31)
        && s->ordering >= 0 && s->ordering <= 2 && s->coefficient_law >= 0
        && s->coefficient_law <= 2;
}

} // namespace
} // namespace owtk::host

using namespace owtk::host;

extern "C" {

void owtk_grid_spec_default(owtk_grid_spec* spec, int32_t nx, int32_t ny, int32_t block_size,
                            uint64_t seed)
{
    if (!spec) return;
    spec->nx = nx;
    spec->ny = ny;
    spec->block_size = block_size;
    spec->ordering = OWTK_ORDER_RCM;
    spec->seed = seed;
    spec->permute = 1;
    spec->coefficient_law = OWTK_LAW_UNIFORM;
    spec->jitter = 0.2;
    spec->diagonal_shift = 0.05;
    spec->field_sigma = 0.0;
    spec->field_correlation = 16.0;
}

int owtk_grid_sizes(const owtk_grid_spec* spec, int64_t* nodes, int64_t* off_diagonal_entries)
{
    return owtk::guarded([&] {
        if (!valid_spec(spec)) throw std::invalid_argument("invalid grid specification");
        if (nodes) *nodes = static_cast<int64_t>(spec->nx) * spec->ny;
        if (off_diagonal_entries) *off_diagonal_entries = 2 * edge_count(*spec);
    });
}

int owtk_grid_generate_split(const owtk_grid_spec* spec, int32_t* row_offsets,
                             int32_t* columns, double* diagonal, double* off_diagonal,
                             double* rhs, double* coordinates, int64_t* final_to_natural)
{
    return owtk::guarded([&] {
        if (!valid_spec(spec)) throw std::invalid_argument("invalid grid specification");
        if (!row_offsets || !columns || !diagonal || !off_diagonal)
            throw std::invalid_argument("grid output arrays must not be null");
        Grid g;
        g.spec = *spec;
        build_numbering(g);
        const std::size_t n = static_cast<std::size_t>(g.n);
        const int bs = spec->block_size;
        // Row offsets from row lengths (degree does not depend on numbering).
        row_offsets[0] = 0;
        for (std::size_t f = 0; f < n; ++f) {
            RowEntry row[kSlots];
            int count = 0;
            const std::int64_t u = g.final_to_natural[f];
            for (int slot = 0; slot < kSlots; ++slot) {
                std::int64_t v;
                count += neighbor(*spec, u, slot, &v) ? 1 : 0;
            }
            (void)row;
            row_offsets[f + 1] = row_offsets[f] + count;
        }
        parallel_for(n, [&](std::size_t b, std::size_t e) {
            for (std::size_t f = b; f < e; ++f) {
                RowEntry row[kSlots];
                int count = 0;
                sorted_row(g, static_cast<std::int64_t>(f), row, &count);
                const std::int64_t u = g.final_to_natural[f];
                double sums[64] = {0};
                for (int k = 0; k < count; ++k) {
                    const std::int64_t entry = row_offsets[f] + k;
                    columns[entry] = static_cast<int32_t>(row[k].column);
                    for (int c = 0; c < bs; ++c)
                        off_diagonal[entry * bs + c] = off_diagonal_value(g, u, row[k].slot, row[k].natural, c);
                }
                // diagonal sums in neighbour-slot order: independent of the numbering
                for (int slot = 0; slot < kSlots; ++slot) {
                    std::int64_t v;
                    if (!neighbor(*spec, u, slot, &v)) continue;
                    for (int c = 0; c < bs; ++c) sums[c] += std::fabs(off_diagonal_value(g, u, slot, v, c));
                }
                for (int c = 0; c < bs; ++c) {
                    diagonal[f * bs + c] = sums[c] + spec->diagonal_shift;
                    if (rhs)
                        rhs[f * bs + c] = 2.0 * draw_uniform(spec->seed, stream_rhs,
                                                             static_cast<std::uint64_t>(u) * 64 + c) - 1.0;
                }
                if (coordinates) {
                    coordinates[2 * f] = g.coords[2 * u];
                    coordinates[2 * f + 1] = g.coords[2 * u + 1];
                }
                if (final_to_natural) final_to_natural[f] = u;
            }
        });
    });
}

int owtk_grid_generate_bsr(const owtk_grid_spec* spec, int32_t* row_offsets, int32_t* columns,
                           double* values, double* rhs)
{
    return owtk::guarded([&] {
        if (!valid_spec(spec)) throw std::invalid_argument("invalid grid specification");
        if (!row_offsets || !columns || !values)
            throw std::invalid_argument("grid output arrays must not be null");
        Grid g;
        g.spec = *spec;
        build_numbering(g);
        const std::size_t n = static_cast<std::size_t>(g.n);
        const int bs = spec->block_size;
        const int bb = bs * bs;
        row_offsets[0] = 0;
        for (std::size_t f = 0; f < n; ++f) {
            int count = 1;
            const std::int64_t u = g.final_to_natural[f];
            for (int slot = 0; slot < kSlots; ++slot) {
                std::int64_t v;
                count += neighbor(*spec, u, slot, &v) ? 1 : 0;
            }
            row_offsets[f + 1] = row_offsets[f] + count;
        }
        parallel_for(n, [&](std::size_t b, std::size_t e) {
            for (std::size_t f = b; f < e; ++f) {
                RowEntry row[kSlots + 1];
                int count = 0;
                sorted_row(g, static_cast<std::int64_t>(f), row, &count);
                const std::int64_t u = g.final_to_natural[f];
                row[count++] = {static_cast<std::int64_t>(f), u, -1};
                std::sort(row, row + count,
                          [](const RowEntry& a, const RowEntry& b) { return a.column < b.column; });
                for (int k = 0; k < count; ++k) {
                    const std::int64_t entry = row_offsets[f] + k;
                    columns[entry] = static_cast<int32_t>(row[k].column);
                    double* block = values + entry * bb;
                    if (row[k].slot < 0) {
                        for (int q = 0; q < bb; ++q) {
                            block[q] = draw_normal(spec->seed, stream_bsr_diagonal,
                                                   static_cast<std::uint64_t>(u) * 256 + q);
                        }
                        for (int d = 0; d < bs; ++d) block[d * bs + d] += 6.0;
                    } else {
                        const std::uint64_t key = static_cast<std::uint64_t>(u) * 8 + row[k].slot;
                        const double scale = 0.1 + 0.9 * draw_uniform(spec->seed, stream_bsr_block_scale, key);
                        for (int q = 0; q < bb; ++q) {
                            block[q] = 0.15 * scale
                                * draw_normal(spec->seed, stream_bsr_off_diagonal, key * 256 + q);
                        }
                    }
                }
                if (rhs) {
                    for (int c = 0; c < bs; ++c)
                        rhs[f * bs + c] = 2.0 * draw_uniform(spec->seed, stream_rhs,
                                                             static_cast<std::uint64_t>(u) * 64 + c) - 1.0;
                }
            }
        });
    });
}

int owtk_split_to_csr(int64_t owned_nodes, int64_t block_size, const int32_t* offsets,
                      const int32_t* columns, const double* diagonal, const double* off_diagonal,
                      int32_t* csr_offsets, int32_t* csr_columns, double* csr_values)
{
    return owtk::guarded([&] {
        if (owned_nodes <= 0 || block_size <= 0 || !offsets || !columns || !diagonal
            || !off_diagonal || !csr_offsets || !csr_columns || !csr_values)
            throw std::invalid_argument("invalid split layout");
        const std::int64_t bs = block_size;
        std::int64_t out = 0;
        csr_offsets[0] = 0;
        for (std::int64_t row = 0; row < owned_nodes; ++row) {
            bool placed = false;
            for (std::int32_t e = offsets[row]; e <= offsets[row + 1]; ++e) {
                const bool at_end = e == offsets[row + 1];
                if (!placed && (at_end || columns[e] > row)) {
                    csr_columns[out] = static_cast<int32_t>(row);
                    for (std::int64_t c = 0; c < bs; ++c)
                        csr_values[out * bs + c] = diagonal[row * bs + c];
                    ++out;
                    placed = true;
                }
                if (at_end) break;
                csr_columns[out] = columns[e];
                for (std::int64_t c = 0; c < bs; ++c)
                    csr_values[out * bs + c] = off_diagonal[static_cast<std::int64_t>(e) * bs + c];
                ++out;
            }
            csr_offsets[row + 1] = static_cast<int32_t>(out);
        }
    });
}

int owtk_order_rcm(int64_t n, const int32_t* row_offsets, const int32_t* columns,
                   int64_t* new_to_old)
{
    return owtk::guarded([&] {
        if (n <= 0 || !row_offsets || !columns || !new_to_old)
            throw std::invalid_argument("invalid graph");
        const auto order = rcm_order(n, row_offsets, columns);
        std::copy(order.begin(), order.end(), new_to_old);
    });
}

int owtk_order_hilbert(int64_t n, const double* coordinates, int64_t* new_to_old)
{
    return owtk::guarded([&] {
        if (n <= 0 || !coordinates || !new_to_old) throw std::invalid_argument("invalid coordinates");
        const auto order = hilbert_order(n, coordinates);
        std::copy(order.begin(), order.end(), new_to_old);
    });
}

} // extern "C"
