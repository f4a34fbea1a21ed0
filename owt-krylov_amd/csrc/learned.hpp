// Learned vertex-block preconditioner: model, per-matrix graph data and the
// preconditioner hooks used by precond.hip.
#pragma once

#include "precond.hpp"

#include <cstdint>
#include <vector>

namespace owtk {

constexpr int kGnnWidth = 32;        // hidden width
constexpr int kGnnLayers = 3;        // message-passing layers
constexpr int kGnnMaxOutputs = 32;   // (degree + 1) * block size
constexpr int kGnnMaxFeatures = 13;  // 3 * block size + 1

// Graph view of a component-diagonal split matrix for the network: CSR over
// vertices (the matrix's own entry order), FP32 copies of the values, the
// per-vertex features and per-entry aggregation weights.
struct GnnGraph {
    std::int64_t n = 0, nnz = 0;
    int b = 1;
    DeviceArray<std::int32_t> offsets, cols, mirror;  // mirror: entry of (col, row); training only
    DeviceArray<float> vals, diag;                     // [e][c], [i][c]
    DeviceArray<float> x, w;                           // features [i][f0] ([f][i] when sell), weights [e]
    DeviceArray<std::int64_t> slot;                    // matrix value slot of each CSR entry
    // Inference over a SELL-layout matrix reads the matrix's own structure
    // (edges in slot order, the CSR order of each row) and keeps the edge
    // weights per slot; the CSR arrays above stay empty.
    bool sell = false;
    const std::int32_t* slice_ptr = nullptr;
    const std::int32_t* row_len = nullptr;
    const std::int32_t* slot_cols = nullptr;
    DeviceArray<float> slot_w;                         // weight of the edge in each slot
};

} // namespace owtk

struct owtk_gnn_s {
    owtk_context ctx = nullptr;
    int b = 1, degree = 0, f0 = 0, outs = 0;
    std::int64_t count = 0;
    // offsets of the parameter blocks in the flat vector
    std::int64_t off_w_in = 0, off_b_in = 0, off_w_out = 0, off_b_out = 0, off_g = 0;
    std::int64_t off_ws[owtk::kGnnLayers] = {}, off_wn[owtk::kGnnLayers] = {}, off_b[owtk::kGnnLayers] = {};
    owtk::DeviceArray<float> params, grad, adam_m, adam_v;
    owtk::DeviceArray<std::uint16_t> params16;     // BF16 weights used by every kernel
    owtk::DeviceArray<double> loss;                // accumulated weighted loss
    std::uint64_t adam_steps = 0;
    // training workspace, sized to the largest problem seen
    std::int64_t ws_nodes = 0, ws_pairs = 0;
    owtk::DeviceArray<float> act, gamma, dgamma, z, ys, partials;
    owtk::DeviceArray<double> loss_partials;
    owtk::DeviceArray<float> coef;                  // per pair loss / gradient scale
};

// A training problem depends on the block size and the degree only (its
// basis vectors u_k and targets), so every model of its context with that
// block size and degree can train on it.
struct owtk_gnn_problem_s {
    owtk_context ctx = nullptr;
    int b = 1, degree = 0;
    int pairs = 0;
    owtk::GnnGraph g;
    owtk::DeviceArray<float> u;       // [p][k][n*b] basis vectors
    owtk::DeviceArray<float> target;  // [p][n*b]  D^-1/2 A e_p
    owtk::DeviceArray<float> dis;     // [n*b]     D^-1/2
    std::vector<double> norms;        // |target_p|^2
};

namespace owtk {

// Preconditioner hooks (OWTK_PC_LEARNED).
// The apply evaluates z = sum_k gamma_k u_k, u_0 = D^-1 r,
// u_(k+1) = -D^-1 A_off u_k, through w_k = D u_k:
//   w_0 = r,  w_(k+1) = -(A_off D^-1) w_k,  z = sum_k (gamma_k D^-1) w_k,
// so no step reads the inverse diagonal.  The column-scaled off-diagonal
// values (matrix slot layout, storage type) and the scaled coefficients are
// formed once per numeric set-up.
struct LearnedPc {
    owtk_gnn model = nullptr;
    GnnGraph graph;                       // structure built once; values refreshed per update
    DeviceArray<float> h;                 // network activations, 2 * n * width

    DeviceArray<float> gamma;             // [k][n*b], the network's coefficients
    DeviceArray<unsigned char> scaled;    // A_off D^-1 values, value_count entries
    DeviceArray<double> sgamma64;         // [k][n*b] gamma_k D^-1 (FP64 storage)
    DeviceArray<float> sgamma32;          // [k][n*b] gamma_k D^-1 (FP32 and BF16 storage)
    DeviceArray<double> dinv;             // 1 / d in the diagonal's layout
    DeviceArray<double> u64a, u64b;       // w_k work vectors
    DeviceArray<float> u32a, u32b;
};

void learned_setup(owtk_precond pc);
void learned_numeric(owtk_precond pc);
void learned_apply_f64(owtk_precond pc, const double* in, double* out, const int* stop);
void learned_apply_f32(owtk_precond pc, const float* in, float* out, const int* stop);
std::int64_t learned_apply_bytes(owtk_precond pc);
void learned_release(owtk_precond pc);

} // namespace owtk
