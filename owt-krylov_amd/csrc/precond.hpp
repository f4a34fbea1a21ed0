// Preconditioner objects shared by the C ABI and the device solvers.
#pragma once

#include "internal.hpp"

#include <cstdint>
#include <vector>

namespace owtk {

// Level-scheduled block triangular system (one per preconditioner).
//
// The owned rows are cut into blocks of consecutive rows (block-Jacobi
// partition; one block when block_nodes == 0).  Inside a block, the forward
// sweep visits rows level by level, where a row's level is one more than the
// deepest lower-neighbour level; the backward sweep does the same from the
// bottom.  Rows within one level are independent, so one workgroup per
// (block, component) runs a level in parallel and separates levels with a
// barrier.  The per-row accumulation keeps the reference's entry order.
struct TriSchedule {
    std::int64_t blocks = 0;
    std::int64_t rows = 0;
    int bs = 1;
    // per block
    std::vector<std::int32_t> h_row0, h_nrows, h_flev, h_blev; // h_flev/h_blev: first level index of block
    DeviceArray<std::int32_t> row0, nrows, flev, blev;         // flev/blev size blocks+1
    // level -> first position (size levels+1), positions are block-local
    DeviceArray<std::int32_t> flev_pos, blev_pos;
    std::vector<std::int32_t> h_flev_pos, h_blev_pos;
    // apply position (global) -> block-local row: level order, or plain row
    // order (forward) / reverse row order (backward) for chain sweeps
    DeviceArray<std::int32_t> frow, brow;
    DeviceArray<std::int32_t> frow_lvl;  // forward level order (numeric factorization)
    bool chain = false;  // one-wave chain sweeps instead of level barriers
    int emax = 0;        // most entries of one row in one direction
    int ei_max = 0;      // chain layout: most in-chunk entries of one row
    DeviceArray<std::int32_t> fnout, bnout;  // chain layout: entries before the chunk
    // entry pointers per position (global), block-local columns
    DeviceArray<std::int64_t> fptr, bptr;
    DeviceArray<std::int32_t> fcol, bcol;
    std::int64_t fnnz = 0, bnnz = 0;
    std::int64_t max_block_rows = 0;
    std::int64_t max_flevels = 0, max_blevels = 0;
    std::int64_t max_level_width = 0;
    // values, component-major: [c][entry]
    DeviceArray<double> fval64, bval64, dval64; // dval: ILU inverse diagonal / SSOR diagonal, [c][row]
    DeviceArray<float> fval32, bval32, dval32;
    DeviceArray<std::uint16_t> fval16, bval16;
    // gather sources of the apply values (numeric update)
    DeviceArray<std::int64_t> fsrc, bsrc, dsrc;
    // ELL copy of the chain layout: per 64-row chunk and slot, the rows'
    // entries contiguous (k_tri_ell); values [c][slot] in each storage
    bool ell = false;
    int fk = 0, bk = 0;                       // slots per row, forward / backward sweep
    std::int64_t fslots = 0, bslots = 0;      // chunks * slots * 64
    DeviceArray<std::int32_t> chunk0;         // first chunk of each block
    DeviceArray<std::int16_t> fidx_ell, bidx_ell;
    DeviceArray<std::int64_t> fmap_ell, bmap_ell;  // chain-order entry of each slot, -1 padding
    DeviceArray<double> fell64, bell64;
    DeviceArray<float> fell32, bell32;
    DeviceArray<std::uint16_t> fell16, bell16;
};

// Program of the chunked numeric factorization (k_ilu_factor_chunks): rows
// in 64-row chunks, one per lane.  For each chunk:
//   srcA  [nsrc][64]  factor entries on earlier chunks a lane reads (-1 pad)
//   opA   [nA][64]    lower entries whose pivot row lies on an earlier chunk,
//                     in column order: eslot | pivot source << 8 | updates << 16
//   updA  [nA*nuA][64] updates of those: target slot | source << 8
//   opB   [nB][64]    lower entries whose pivot row lies in the chunk, in column
//                     order: step | eslot << 8 | pivot slot << 16 | updates << 24
//   updB  [nB*nuB][64] target slot | source slot (in the pivot row) << 8
// (slots are positions in a row's factor entries; 0xffffffff pads opA/opB)
// hdr per chunk: row of lane 0, rows, and offset and count of each array.
struct IluChunkProgram {
    bool enabled = false;
    int kmax = 0, max_src = 0, max_a = 0, max_nua = 0, max_b = 0, max_nub = 0;
    DeviceArray<std::int32_t> chunk_first;  // per block (blocks + 1)
    DeviceArray<std::int64_t> hdr;          // per chunk: 12 values (see k_ilu_factor_chunks)
    DeviceArray<std::int32_t> src_a;
    DeviceArray<std::uint32_t> op_a, op_b;
    DeviceArray<std::uint16_t> upd_a, upd_b;
};

struct IluFactor {
    // factor pattern, reference accessor layout (global rows, sorted columns)
    std::vector<std::int64_t> h_offsets, h_columns, h_diag;
    std::int64_t nnz = 0;
    IluChunkProgram chunks;
    DeviceArray<std::int64_t> diag_pos;      // factor entry of each row's diagonal
    DeviceArray<std::int64_t> src;           // matrix gather source per factor entry (-1: fill)
    DeviceArray<std::int64_t> columns;       // global column per factor entry
    DeviceArray<std::int64_t> upd_ptr;       // per factor entry (lower entries): update list range
    DeviceArray<std::int64_t> upd_target, upd_upper;
    DeviceArray<std::int64_t> offsets;       // factor row offsets (global)
    DeviceArray<double> values;              // [c][entry] FP64 factor
};

} // namespace owtk

namespace owtk {
// A SELL-64 copy of some of a dense-block matrix's blocks, rows in a
// different order: slot (s, k, lane) at slice_ptr[s] + 64k + lane, block
// values at dense_index(slot, j, bb) in the preconditioner's storage type
// (vals), map = matrix slot of each copy slot (-1 for padding).
struct McSell {
    DeviceArray<std::int32_t> slice_ptr, row_len, cols;
    DeviceArray<std::int64_t> map;
    DeviceArray<unsigned char> vals;
    std::int64_t slots = 0;
};
} // namespace owtk

struct owtk_precond_s {
    owtk_matrix m = nullptr;
    // the context, kept apart from the matrix: destroy may run after the
    // matrix is gone (a garbage collector finalizes in no fixed order)
    owtk_context ctx = nullptr;
    int kind = OWTK_PC_IDENTITY;
    int storage = OWTK_F64;
    double omega = 1.0;
    double minimum_diagonal = 0.0;
    std::int64_t block_nodes = 0;
    int fill_level = 0;
    int schedule = 0;  // 0 auto, 1 level sets, 2 chains
    std::uint64_t updates = 0;
    // Jacobi: inverse diagonal (component-diagonal) or inverse blocks (dense)
    owtk::DeviceArray<double> inv64;
    owtk::DeviceArray<float> inv32;
    owtk::TriSchedule tri;
    owtk::IluFactor ilu;
    owtk::DeviceArray<int> flag;   // numeric failure flag (zero pivot / singular)
    // dense-block SSOR: rows of colour c are h_mc_ptr[c] .. h_mc_ptr[c+1]-1
    // in colour order; colour-ordered SELL-64 copies of the blocks each
    // sweep reads: colour c's rows occupy copy rows h_mc_seg[c].. (padded to
    // whole slices), mc_prow maps a copy row to its matrix row (-1 padding)
    std::vector<std::int64_t> h_mc_ptr;
    std::vector<std::int64_t> h_mc_seg;
    owtk::DeviceArray<std::int32_t> mc_prow;
    owtk::McSell mc_fwd, mc_bwd, mc_diag;
    owtk::DeviceArray<double> work64;
    owtk::DeviceArray<float> work32;
    void* learned = nullptr;       // owtk_gnn model for OWTK_PC_LEARNED
    void* learned_state = nullptr; // owtk::LearnedPc (coefficients, inverse diagonal, work vectors)
    std::int64_t n() const { return m->owned * m->bs; }
};

namespace owtk {

// Apply z = M^{-1} r for the solver path.  `stop` (device, may be null) turns
// the launches into no-ops once a solver has stopped.
void precond_apply_f64(owtk_precond pc, const double* in, double* out, const int* stop);
void precond_apply_f32(owtk_precond pc, const float* in, float* out, const int* stop);
// Bytes one apply moves (algorithmic).
std::int64_t precond_apply_bytes(owtk_precond pc);

} // namespace owtk
