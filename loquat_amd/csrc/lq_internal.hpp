// lq_internal.hpp -- host-side launchers shared between the module ABI and
// the engine.  Every launcher is asynchronous on the given stream, allocates
// nothing and records no memset (so it can be captured into a hipGraph).
#pragma once

#include <algorithm>
#include <string>

#include "lq_common.hpp"

namespace lq {

// Thread-local error text for lq_last_error().
void set_error(const std::string& msg);
#define LQ_HIP_TRY(expr)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::lq::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));          \
            return LQ_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)
#define LQ_TRY(expr)                      \
    do {                                  \
        lq_status s_ = (expr);            \
        if (s_ != LQ_OK) return s_;       \
    } while (0)

// Host-side derivations.
float neighbour_r2_cut(float i_h);
void fill_desc_static(RunDesc& d, const lq_params& par, const lq_vw_params* vw);

// ---------------------------------------------------------------- grid
// Radix-sort geometry.
constexpr int RS_THREADS = 256;
constexpr int RS_IPT = 8;
constexpr int RS_TILE = RS_THREADS * RS_IPT;  // 2048 keys per tile
constexpr int RS_BINS = 256;

inline int rs_tiles(int n) { return (n + RS_TILE - 1) / RS_TILE; }
inline int rs_bits(int64_t ncell) {
    int b = 1;
    while ((int64_t(1) << b) < ncell) ++b;
    return b;
}

// Scratch owned by a grid: keys/values ping-pong, histograms, cell table.
struct GridBuffers {
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;    // RS_BINS * tiles
    uint32_t* totals;  // RS_BINS
    int32_t* cell_beg;
    int32_t* cell_end;
    int64_t np_cap;
    int64_t ncell_cap;
    int sorted_slot;  // which keys/vals slot holds the sorted result
};

lq_status alloc_grid(GridBuffers& g, int64_t np_cap, int64_t ncell_cap);
void free_grid(GridBuffers& g);

// Neighbour grid of N particles (one run by value, or a table of runs with
// zbits z-bin bits in the key): cell keys, stable LSD radix sort of
// (key, index) over ceil(bits/8) 8-bit passes (result in slot
// g.sorted_slot), and cell_beg / cell_end (-1 for empty cells).
void launch_grid_build_one(const RunDesc& d, const float4* pospres, GridBuffers& g, hipStream_t s);
// which (GridRekey): the runs whose particles get fresh keys; the others keep
// the keys of the last build (g.keys[g.sorted_slot], in the current order).
void launch_grid_build_table(const RunDesc* d, const int32_t* run_of, const float4* pospres, int n,
                             int64_t ncell_total, int zbits, GridBuffers& g, hipStream_t s,
                             int which = REKEY_ALL);
// dst[i] = src[i] + shift, i < n: kept keys moved with their run's segment.
void launch_shift_keys(const uint32_t* src, uint32_t* dst, int n, uint32_t shift, hipStream_t s);
// The grid build of one run into caller-owned arrays (part_cel, part_idx:
// np entries; cell_beg, cell_end: ncell), the sort's scratch in a
// file-scope device array of lq_grid.hip (np <= NOALLOC_MAX_PARTICLES).  No
// allocation; calls that overlap on one GPU share that scratch and race.
constexpr int NOALLOC_MAX_PARTICLES = LQ_NOALLOC_MAX_PARTICLES;
lq_status launch_grid_build_noalloc(const RunDesc& d, const float4* pospres, int32_t* part_cel, int32_t* part_idx,
                                    int32_t* cell_beg, int32_t* cell_end, hipStream_t s);
// dst[k] = src[perm[k]] for all state arrays (and ini when both set).  When
// str_dst is non-null the permuted stress goes there instead of dst.str.
// Engine states also carry table positions: those of the runs `which`
// re-keyed (desc / run_of) become the positions just sorted.
// half (engine, with desc): also writes the walls' positions, FP32 mirror,
// velocity and stress into that half-step state.
void launch_permute(const uint32_t* perm, int n, const DevState& src, const DevState& dst, const DevIni* ini_src,
                    const DevIni* ini_dst, tensor2d* str_dst, hipStream_t s, const RunDesc* desc = nullptr,
                    const int32_t* run_of = nullptr, int which = REKEY_ALL, const DevState* half = nullptr);

// ---------------------------------------------------------------- physics
enum Stage { ST_ADAMI = 0, ST_INTERACT = 1, ST_PREDICT = 2, ST_CORRECT = 3, ST_SHEPARD = 4 };

// _one: a single run by value (walls [0, npb), material [npb, np)).
// _table: all resident runs; walls / mats map a compact thread index to the
// particle index of the t-th wall / material particle.
// variant >= 3 selects the list-based neighbour scan for Adami and Shepard
// (same sums in the same order as the direct scan).
// stale (nullable): the cell table was built before the particles last moved
// (positions then and a displacement bound); the scans widen their reach by
// the bound and test the current cells, so that the neighbour sets are those
// of a fresh table (list-based kernels only; nbr: per-particle counts).
void launch_adami_one(const RunDesc& d, const DevState& st, const int32_t* beg, const int32_t* end, int variant,
                      const StaleIn* stale, int32_t* nbr, hipStream_t s);
// str2 (variant 19 only): also writes the walls' stress there.
void launch_adami_table(const RunDesc* d, const int32_t* run_of, const DevState& st, const int32_t* beg,
                        const int32_t* end, IndexMap walls, int nwalls, int variant, const StaleIn* stale,
                        hipStream_t s, tensor2d* str2 = nullptr);
void launch_interaction_one(const RunDesc& d, const DevState& st, const DevRates& rt, const int32_t* beg,
                            const int32_t* end, int32_t* nbr, int variant, const StaleIn* stale, hipStream_t s);
// The last nart of the nmat material indices belong to runs with
// artificial-pressure = 1; they get a kernel of their own so that the other
// runs keep the exact arithmetic of the plain kernel.
// fuse (variant 19 in split mode, nart = 0): the predictor's (corr: the
// corrector's) half step is done by the interaction kernel itself, as
// launch_integrate_table would do it, and rt is left untouched.
struct FuseIn {
    // the scans read st: the step's start in the predictor half, the half
    // step in the corrector half; other is the other one
    DevState other;
    int32_t* fail;              // corrector: per-run failure flags
    const uint32_t* dmax_pred;  // corrector: the predictor's folded bound
    uint32_t* dmax;             // DMAX_SLOTS displacement slots per run
    uint32_t* nout;             // per run: a particle ended outside the grid
};
void launch_interaction_table(const RunDesc* d, const int32_t* run_of, const DevState& st, const DevRates& rt,
                              const int32_t* beg, const int32_t* end, IndexMap mats, int nmat, int nart,
                              int32_t* nbr, int variant, const StaleIn* stale, hipStream_t s,
                              const FuseIn* fuse = nullptr, bool corr = false);
void launch_integrate_one(bool corrector, const RunDesc& d, const DevState& st, const DevIni& ini,
                          const DevRates& rt, hipStream_t s);
// fail: per-run flags raised by the corrector on a non-finite state
// (nullable).  dmax (nullable, per
// run, zeroed by launch_step_begin): the predictor records each run's largest
// displacement (and nout, per run, whether a particle ended outside the
// grid); the corrector raises fail when the displacement exceeded what a
// stale table tolerates.
// dmaxb / noutb (nullable, corrector): the same for where the step leaves the
// particles relative to their run's cell table (engine states).
// half (non-null: split mode): the predictor leaves st as the step's start
// and writes the half step's positions, velocity and stress to *half; the
// corrector reads its ini set from st and the half step's pressure from
// *half (same arithmetic, no ini copy).
void launch_integrate_table(bool corrector, const RunDesc* d, const int32_t* run_of, const DevState& st,
                            const DevIni& ini, const DevRates& rt, IndexMap mats, int nmat, int32_t* fail,
                            uint32_t* dmax, uint32_t* nout, hipStream_t s, uint32_t* dmaxb = nullptr,
                            uint32_t* noutb = nullptr, const DevState* half = nullptr);

// Shepard filter: reads stress from str_src, writes str_dst (walls and runs
// not due are copied through).
void launch_shepard_one(const RunDesc& d, const DevState& st, const tensor2d* str_src, tensor2d* str_dst,
                        const int32_t* beg, const int32_t* end, int n, int variant, hipStream_t s);
void launch_shepard_table(const RunDesc* d, const int32_t* run_of, const DevState& st, const tensor2d* str_src,
                          tensor2d* str_dst, const int32_t* beg, const int32_t* end, int n, int variant,
                          hipStream_t s);
// Per-run step bookkeeping on the device.
// sort_every > 0: sets each run's sort_due for this step, and stops (fail)
// a run whose kept cell table is further behind than a stale scan covers.
void launch_step_begin(RunDesc* d, int nruns, uint32_t* dmax, uint32_t* nout, hipStream_t s, int sort_every = 0,
                       int32_t* fail = nullptr);
// Folds the predictor's per-run displacement slots (DMAX_SLOTS per run).
void launch_dmax_fold(uint32_t* dmax, int nruns, hipStream_t s);
// dmaxb / noutb (nullable): fold the corrector's bounds into each run's dtab
// / out_tab and advance since_sort (grid_ran, shep_ran: which builds ran).
void launch_step_end(RunDesc* d, int nruns, const int32_t* fail, hipStream_t s, uint32_t* dmaxb = nullptr,
                     uint32_t* noutb = nullptr, bool grid_ran = true, bool shep_ran = false);
// Dynamic time step of the runs with dynamic-time-step = 1: per-run maxima of
// |v| (state) and |a| (rates of the previous corrector) into red (2 x
// DMAX_SLOTS per run, zero on entry, cleared again on exit), then each such
// run's par.dt.
void launch_dynamic_dt(RunDesc* d, const int32_t* run_of, const DevState& st, const DevRates& rt, IndexMap mats,
                       int nmat, int nruns, uint32_t* red, hipStream_t s);
// Copies material stress back (reference's D2D copy, src/field_smoothing.cu:144).
void launch_copy_tensor(const tensor2d* src, tensor2d* dst, int n, hipStream_t s);


}  // namespace lq
