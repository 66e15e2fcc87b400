// trsolve.hpp -- level-scheduled sparse triangular solves on gfx950.
//
// A plan describes one triangular solve by its dependency lists: row r needs
// x[dep] for every dep listed in dptr[r]..dptr[r+1] (all deps below r for a
// forward plan, all above r for a backward plan), multiplied by the factor
// value at position src.  Building a plan on the device computes
//   lev[r]      = 0 without deps, else 1 + max lev[dep]   (sync-free kernel),
//   level_ptr   = rows per level (exclusive scan),
//   perm        = rows sorted by (level, row),
//   chunks      = each level cut into 64-row chunks (one wavefront each),
//   entries     = per chunk, k-th dep of every lane stored contiguously
//                 (slot-major "sliced" layout: loads are coalesced).
// Executing a plan launches one wavefront per chunk in level order (tickets
// taken with an atomic counter, so a wave only waits on waves that started
// earlier).  A lane waits for each dependency by polling its value with
// agent-scope loads until it differs from a signalling-NaN sentinel: the
// value is its own ready flag, no fences are needed.  Every wait is bounded.
#pragma once
#include <vector>

#include "util.hpp"

namespace hala_amd {

enum tri_mode : int {
    TRI_SUB = 0,  // acc = in; acc -= v * x_dep (term by term); [acc /= diag]
    TRI_SUM = 1   // s = 0; s += v * x_dep; acc = alpha*in - s; [acc /= diag]
};

// Column-blocked layout of a plan (see trcol.hip): rows cut into ncol
// blocks (contiguous, or pieces of slabs of planes), each swept level by
// level by one workgroup.
struct tri_colplan {
    bool ok = false;
    int ncol = 0;     // columns (one workgroup each)
    int split = 1;    // pieces per slab of planes (1: contiguous row blocks)
    int plane = 0;    // rows per plane of the split (0 when not split)
    int K = 0;        // levels kept in the LDS window
    int CP = 0;       // rows of the largest cell, rounded up to 64 (= compute threads per workgroup)
    int L = 0;        // dependency record width (longest row)
    int nlevcol = 0;  // most levels spanned by one column
    int64_t nrec = 0;
    int next = 0;     // external-list entries
    int maxext = 0;   // most distinct external values one cell stages
    int EL = 0;       // external values per lane of the loader wave (power of two, >= maxext / 64)
    int why = 0;      // when not built: 1 small, 2 long rows, 3 cells too wide, 4 LDS, 5/6 externals
    dbuf ci, colrange;           // cell records (column-major, 5 ints each), level range per column
    dbuf code, src, ext_slot;    // dependency codes and factor positions (k-major per level); external slots
    dbuf spar;                   // per slot: parity of its level within its column (which external buffer it reads)
    dbuf val64, val32;           // packed row records (see col_pack)
    dbuf xs64, xs32;             // input, gathered into slot order per sweep
    bool hand64 = false;         // val64 carries hand-off slots (col_set_handoff)
};

struct tri_plan {
    int n = 0;
    bool backward = false;
    bool unit = true;
    bool perm_sorted = true;  // perm in (level, row) order (else rows of a level in any order)
    bool laid_out = false;    // slot layout built
    int nlev = 0;
    int nchunk = 0;
    int64_t nent = 0;  // padded entries (64 per slab)
    // dependency lists
    dbuf dptr, ddep, dsrc, diag_src;
    // schedule
    dbuf lev, level_ptr, perm;
    std::vector<int> h_level_ptr;
    dbuf chunk_row, chunk_len, chunk_off, slot_of_row;
    dbuf lvl_chunk;  // first chunk of every level (nlev + 1)

    dbuf ent_dep, ent_src, dg_src;
    // values
    dbuf ent_val64, ent_val32, dg_val64, dg_val32;
    bool has64 = false, has32 = false;
    // polled outputs in slot order (re-armed to sentinels at the start of every sweep)
    dbuf out64, out32;
    tri_colplan col;
};

// dependency selection for build_deps()
enum dep_sel : int {
    DEP_LOWER = 0,    // entries pntr[i] .. diag[i]-1           (ascending)
    DEP_UPPER = 1,    // entries diag[i]+1 .. pntr[i+1]-1       (ascending)
    DEP_T_LOWER = 2,  // transposed row k: source rows i < k    (ascending)
    DEP_T_UPPER = 3,  // transposed row k: source rows i > k    (descending)
    DEP_ROW_BUT_LAST = 4,   // entries pntr[i] .. pntr[i+1]-2   (one-triangle L, diag last)
    DEP_ROW_BUT_FIRST = 5,  // entries pntr[i]+1 .. pntr[i+1]-1 (one-triangle U, diag first)
    DEP_T_UPPER_ASC = 6     // transposed row k: source rows i > k (ascending)
};

// Fills plan.dptr/ddep/dsrc from a pattern.  diagpos[i] (nullable for the
// one-triangle selections) gives the diagonal position of row i; tp is the
// transposed pattern for the DEP_T_* selections.  diag_src of each row is
// taken from diagpos (or -1 when unit).
void build_deps(hala_ctx c, tri_plan &p, int n, dep_sel sel, const int *pntr, const int *indx,
                const int *diagpos, const csr_transpose *tp, bool unit, bool backward);

// Computes levels, permutation, chunks and entry layout; fills out64 with sentinels.
void build_schedule(hala_ctx c, tri_plan &p, bool want32, bool have_levels = false);
// The level of every row (p.lev, p.nlev) by the sync-free level-set kernel;
// build_schedule calls it unless told the levels are there.
void compute_levels(hala_ctx c, tri_plan &p);
// The levels of p as nlev_other - 1 - lev_other (another triangle's levels
// mirrored), if they satisfy p's own level recurrence, which makes them equal
// to compute_levels' result; false (p.lev unspecified) otherwise.  For an ILU
// factor's U, with L's levels, this holds on stencil patterns whose rows all
// lie on a longest dependency chain.
bool mirror_levels(hala_ctx c, tri_plan &p, const int *lev_other, int nlev_other);

// perm in (level, row) order (sorts it if the analysis left a level's rows unsorted).
void sort_perm(hala_ctx c, tri_plan &p);

// Sorts the rows by (level, key[r], r), key in [0, nkey): order receives the
// rows, bptr[l * nkey + k] the start of bucket (l, k) in it (nlev * nkey + 1
// entries).  The plan is not changed.
void sort_level_keys(hala_ctx c, const tri_plan &p, const int *key, int nkey, int *bptr, int *order);
// Makes a row order from sort_level_keys the plan's level order and rebuilds
// the slot layout; perm takes the new order (sort_perm restores (level, row)),
// levels and level_ptr are unchanged.  Called before any values are loaded.
void apply_level_order(hala_ctx c, tri_plan &p, const int *order);

// Gathers plan values from a factor value array (positions ent_src / diag_src).
void load_values64(hala_ctx c, tri_plan &p, const double *factor);
void load_values32(hala_ctx c, tri_plan &p, const double *factor);

// Solves with plan p.  in: natural-order input (TIN), out_nat: natural-order
// output (TOUT).  The plan's polled output of this precision is re-armed first,
// so sweeps of any kind may follow each other on the same plans.
// gate (nullable, device): when *gate != 0 at launch the solve does nothing.
template <class T, class TIN, class TOUT>
void tri_exec(hala_ctx c, tri_plan &p, int mode, T alpha, const TIN *in, TOUT *out_nat, const int *gate = nullptr);

// Column-blocked layout (trcol.hip): built after build_schedule when the
// pattern allows it (returns false and leaves p.col.ok unset otherwise);
// ncol_req = 0 picks the column count.  tri_exec uses it when present and
// enabled (option tri_col).
bool build_col_plan(hala_ctx c, tri_plan &p, int ncol_req);
void col_load_values(hala_ctx c, tri_plan &p, const double *factor, bool f32);
bool col_usable(hala_ctx c, const tri_plan &p, bool is64);
// next (nullable; FP64, TRI_SUB, unit p with hand-off slots, see
// col_set_handoff): the sweep also writes next's input in next's slot order,
// and its helper workgroups re-arm next's polled output.
template <class T, class TOUT>
void tri_col_exec(hala_ctx c, tri_plan &p, int mode, T alpha, const TOUT *in, TOUT *out_nat, const int *gate,
                  tri_plan *next = nullptr);

// Pieces of the FP64 column sweep for callers that fuse its ends: the column
// kernel alone (input in slot order in col.xs64; the result in slot order in
// out64 and, with next, in next's input as above; arm (nullable): a plan whose
// polled output the sweep's helper workgroups re-arm meanwhile); the
// slot-order output scattered to natural order.
// ax (nullable): x[i] = x[i] + (*alpha) * p[i] over n rows, also done by the
// helper workgroups (a solver's vector update riding on idle CUs; skipped,
// like the sweep, when the gate is set).
struct col_axpy {
    double *x;
    const double *p;
    const double *alpha;
    int n;
};
void col_sweep64(hala_ctx c, tri_plan &p, int mode, double alpha, const int *gate, tri_plan *next = nullptr,
                 tri_plan *arm = nullptr, const col_axpy *ax = nullptr);
void col_scatter64(hala_ctx c, tri_plan &p, double *out_nat, const int *gate);
// Writes into the FP64 records of the unit triangle src the slot of each of
// its rows in dst's input (sets src.col.hand64; left unset when either plan
// has no column layout).  Needed after every col_load_values of src.
void col_set_handoff(hala_ctx c, tri_plan &src, tri_plan &dst);

// Refills the polled output buffer(s) with sentinels (both precisions, or one).
void tri_rearm(hala_ctx c, tri_plan &p);
void tri_rearm_prec(hala_ctx c, tri_plan &p, bool is64);

constexpr int WAVE = 64;
constexpr int SLOT_PAD = 1024;             // slot-indexed arrays run this far past their end (idle-lane reads)
constexpr int MAXD = 16;                   // dependencies held in registers per row (longer rows loop)
constexpr unsigned SPIN_LIMIT = 1u << 22;  // polls before a wait gives up (a few seconds)

}  // namespace hala_amd
