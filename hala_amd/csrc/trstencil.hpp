// trstencil.hpp -- the ILU(0) sweeps of a stencil matrix: pencil schedule.
//
// A matrix whose rows are the points (i, j, k) of an nx x ny x nz grid in
// natural order (row = i + nx j + nx ny k) and whose entries couple each point
// only to points of its 3 x 3 x 3 neighbourhood (5- and 9-point stencils in 2-D,
// 7- and 27-point in 3-D, Dirichlet-trimmed or with missing entries) is swept
// line by line: one lane owns the line (j, k) and runs along i, one row per
// step, the row (i, j, k) at step i + a j + b k, where the weights make every
// lower dependency an earlier step.  The i-1 term is in the lane's registers,
// terms of other lines of the lane's 8 x 8 tile (64 x 1 in 2-D) arrive by
// cross-lane moves from the previous steps' values, and terms of other tiles
// arrive through global memory, polled by a loader wave ahead of the compute
// wave.  No barrier is needed, input and output stay in natural order, and each
// row subtracts its terms in CSR order, as the reference's apply_ilu_array.
#pragma once
#include "common.hpp"

struct hala_ilu_s;

namespace hala_amd {

enum stencil_kind : int { SK_NONE = 0, SK_5 = 1, SK_9 = 2, SK_7 = 3, SK_27 = 4 };

// One triangle's schedule and packed operands.
struct stencil_tri {
    int ntiles = 0;
    int64_t nprow = 0;   // packed rows: (tile, step, lane) slots
    dbuf tile_of_ticket;  // tile per ticket, in dependency order
    dbuf meta;            // int4 per tile: (j0, k0, first global step, number of steps)
    dbuf base;            // int64 per tile: first packed row
    dbuf pos;             // int per packed coefficient: CSR position, -1 absent
    dbuf dpos;            // U: int per packed row: CSR position of the diagonal, -1 idle
    dbuf mask;            // uint16 per packed row: bit q set when slot q's entry is present
    dbuf rec64, rec32;    // step records (coefficients, U's diagonal, masks) in FP64 / FP32
    dbuf ticket;          // one counter (reset by the other triangle's sweep)
};

struct stencil_plan {
    int kind = SK_NONE;
    int nx = 0, ny = 0, nz = 0, P = 0;
    int ntj = 0, ntk = 0;
    stencil_tri L, U;
    dbuf prow;  // natural row of each packed row (L order), -1 for padding
    // packed vectors (FP64 / FP32): the input, L's output (U's input) and U's
    // output; the outputs hold sentinels between applies
    dbuf xp64, y64, u64, xp32, y32, u32;
    bool loaded64 = false, loaded32 = false;
};

// Detects a supported stencil pattern and builds the plan (false: not one).
bool stencil_build(hala_ilu_s *M);
// Packs the factor's values (FP64, and FP32 when f32) into the plan's streams.
void stencil_load(hala_ilu_s *M, bool f32);
// r = (LU)^-1 x by the two pencil sweeps in precision V (x: TIN, r: TOUT).
template <class V, class TIN, class TOUT>
void stencil_apply(hala_ilu_s *M, const TIN *x, TOUT *r, const int *gate);

}  // namespace hala_amd
