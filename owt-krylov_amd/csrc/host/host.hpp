// Host-side set-up code of libowtk_hip: synthetic grids, node orderings and the
// symbolic analysis (triangular level schedules, ILU patterns) that the device
// kernels consume.  Nothing here touches the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace owtk::host {

// ---------------------------------------------------------------------------
// Counter-based random numbers.  Every draw is a pure function of
// (seed, stream, counter), so a value never depends on generation order and
// CPU/GPU or serial/threaded generators agree bit for bit.
// ---------------------------------------------------------------------------
inline std::uint64_t splitmix64(std::uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

inline std::uint64_t draw_bits(std::uint64_t seed, std::uint64_t stream,
                               std::uint64_t counter)
{
    return splitmix64(splitmix64(seed ^ (stream << 56))
                      ^ splitmix64(counter * 0xd1b54a32d192ed03ULL + stream));
}

// Uniform in [0, 1) with 53 random bits.
inline double draw_uniform(std::uint64_t seed, std::uint64_t stream,
                           std::uint64_t counter)
{
    return static_cast<double>(draw_bits(seed, stream, counter) >> 11)
        * 0x1.0p-53;
}

// Standard normal by Box-Muller from two independent counters.
double draw_normal(std::uint64_t seed, std::uint64_t stream,
                   std::uint64_t counter);

enum Stream : std::uint64_t {
    stream_jitter_x = 1,
    stream_jitter_y = 2,
    stream_permutation = 3,
    stream_magnitude = 4,
    stream_perturbation = 5,
    stream_rhs = 6,
    stream_field = 7,
    stream_edge_contrast = 8,
    stream_bsr_diagonal = 9,
    stream_bsr_off_diagonal = 10,
    stream_bsr_block_scale = 11,
};

// Run body(begin, end) over [0, n) on up to hardware_concurrency threads.
void parallel_for(std::size_t n, const std::function<void(std::size_t, std::size_t)>& body);

// ---------------------------------------------------------------------------
// Orderings (new_to_old permutations).
// ---------------------------------------------------------------------------
std::vector<std::int64_t> rcm_order(std::int64_t n, const std::int32_t* offsets,
                                    const std::int32_t* columns);
std::vector<std::int64_t> hilbert_order(std::int64_t n, const double* xy);
std::uint64_t hilbert_index(std::uint32_t x, std::uint32_t y, std::uint32_t extent);

// ---------------------------------------------------------------------------
// Triangular schedules.
// ---------------------------------------------------------------------------
struct LevelSchedule {
    // rows grouped by level; level l holds rows[level_ptr[l] .. level_ptr[l+1])
    std::vector<std::int32_t> level_ptr;
    std::vector<std::int32_t> rows;
};

// GCRO-DR's harmonic-Ritz eigenproblem harmonic y = theta metric y
// (harmonic SPD, n x n column-major) in LAPACK dggev's output convention
// (eig.cpp).
void harmonic_ggev(int n, const double* harmonic, const double* metric, double* real, double* imag, double* beta,
                   double* right);

} // namespace owtk::host
