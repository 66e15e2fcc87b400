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

struct tri_plan {
    int n = 0;
    bool backward = false;
    bool unit = true;
    int nlev = 0;
    int nchunk = 0;
    int64_t nent = 0;  // padded entries (64 per slab)
    // dependency lists
    dbuf dptr, ddep, dsrc, diag_src;
    // schedule
    dbuf lev, level_ptr, perm;
    std::vector<int> h_level_ptr;
    dbuf chunk_row, chunk_len, chunk_off, slot_of_row;

    dbuf ent_dep, ent_src, dg_src;
    // values
    dbuf ent_val64, ent_val32, dg_val64, dg_val32;
    bool has64 = false, has32 = false;
    // polled outputs in slot order (all-sentinel between solves)
    dbuf out64, out32;
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
void build_schedule(hala_ctx c, tri_plan &p, bool want32);

// Gathers plan values from a factor value array (positions ent_src / diag_src).
void load_values64(hala_ctx c, tri_plan &p, const double *factor);
void load_values32(hala_ctx c, tri_plan &p, const double *factor);

// Solves with plan p.  in: natural-order input (TIN), out_nat: natural-order
// output (TOUT, nullable).  reset/reset_slot: the partner plan's polled buffer
// and slot map, set back to sentinels for its next solve (nullable).
// gate (nullable, device): when *gate != 0 at launch the solve does nothing.
template <class T, class TIN, class TOUT>
void tri_exec(hala_ctx c, tri_plan &p, int mode, T alpha, const TIN *in, TOUT *out_nat, T *reset,
              const int *reset_slot, const int *gate = nullptr);

// Refills the polled output buffer(s) with sentinels.
void tri_rearm(hala_ctx c, tri_plan &p);

// Device error flag raised by a dependency wait that gave up; read and cleared by check_wait_error.
int *wait_error_flag(hala_ctx c);

constexpr int WAVE = 64;
constexpr int MAXD = 16;                   // dependencies held in registers per row (longer rows loop)
constexpr unsigned SPIN_LIMIT = 1u << 22;  // polls before a wait gives up (a few seconds)
void check_wait_error(hala_ctx c);

}  // namespace hala_amd
