// lq_engine.hip -- the C ABI: module entry points (one run, caller-owned
// arrays) and the engine that runs the reference's time loop for one or many
// runs packed in one segmented layout.
//
// Engine step (src/loquat.cu:322-409), every stage one launch over all runs:
//   step_begin                       per-run Shepard flag (src/loquat.cu:400),
//                                    dynamic time step when a run asks for it
//   [predict] grid build (hash, sort, cell table) and permute Primary
//             (src/loquat.cu:345-347), Adami, interaction, predictor (+ the
//             fused *_ini copy and each run's largest displacement)
//   [correct] no second neighbour search: Adami and the interaction scan the
//             predictor half's cell table with their reach widened by the
//             run's largest displacement (the neighbour sets of the reference's
//             second search, src/loquat.cu:362-375), then the corrector
//   [shepard] if any run is due: grid build, permute Primary with the stress
//             routed to scratch, Shepard filter back into the state
//   step_end                         per-run time += dt, step++, finish flag
// The state ping-pongs between two buffers; each (buffer, shepard) pair has
// its own captured hipGraph, rebuilt when runs join or leave.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <limits>
#include <vector>

#include "lq_internal.hpp"

namespace lq {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// Interaction variant of module-level calls made by this thread, and the
// initial variant of engines this thread creates (each engine keeps its own).
static thread_local int g_variant = LQ_DEFAULT_INTERACTION_VARIANT;

// Smallest FP32 value T with sqrtf(T) * i_h >= 2 (both correctly rounded):
// r2 < T is then exactly the predicate q = sqrtf(r2) * i_h < 2 for every
// FP32 r2, since both operations are monotone.
float neighbour_r2_cut(float i_h) {
    uint32_t lo = 0, hi = 0x7f800000u;  // [0, +inf]
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float v;
        std::memcpy(&v, &mid, 4);
        volatile float r = sqrtf(v);
        volatile float q = r * i_h;
        if (q >= 2.0f)
            hi = mid;
        else
            lo = mid + 1;
    }
    float t;
    std::memcpy(&t, &lo, 4);
    return t;
}

static double deg2rad(double x) { return x * 3.14159265358979323846 / 180.0; }

void fill_desc_static(RunDesc& d, const lq_params& par, const lq_vw_params* vw) {
    std::memset(&d, 0, sizeof(d));
    d.par = par;
    d.np = par.np;
    d.npb = par.npb;
    d.npm = par.npm;
    d.r2_cut = neighbour_r2_cut(par.i_h);
    // until a step has seen where the particles are, assume some may lie
    // outside the grid (the kept-table scans then test every edge hit)
    d.out_tab = 1;
    d.active = 1;
    d.shepard_due = 1;
    d.nsteps = 0x7fffffff;
    d.time = (double)par.time;
    if (vw) {
        d.vw = *vw;
        if (par.constitutive_model == 3 && vw->phi_c > 0.0f) {
            const double sp = sin(deg2rad(vw->phi_c));
            const double a = sqrt(3.0) * (3.0 - sp) / (2.0 * sqrt(2.0) * sp);
            const double ratio = (double)(vw->ei0 - vw->ed0) / (double)(vw->ec0 - vw->ed0);
            const double den = 3.0 + a * a - a * sqrt(3.0) * pow(ratio, (double)vw->alpha);
            d.vw_a = (float)a;
            d.vw_fb0 = (float)((double)vw->hs / (double)vw->n * pow((double)vw->ei0 / (double)vw->ec0, (double)vw->beta) / den);
        }
    }
}

}  // namespace lq

using namespace lq;

// ------------------------------------------------------------------ helpers
static lq_status check_params(const lq_params* par) {
    if (!par) {
        set_error("null params");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->np < 0 || par->npb < 0 || par->npm < 0 || par->np != par->npb + par->npm) {
        set_error("invalid particle counts: np must equal npb + npm");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->np > LQ_MAX_PARTICLES) {
        set_error("more than LQ_MAX_PARTICLES particles (32-bit gather offsets)");
        return LQ_ERR_CAPACITY;
    }
    if (par->ncell <= 0 || par->ncell != 2 * par->ncell_one_layer + 1 ||
        (int64_t)par->cell_num_x * par->cell_num_y * par->cell_num_z != par->ncell_one_layer) {
        set_error("cell counts are not derived (call lq_params_derive)");
        return LQ_ERR_INVALID_ARG;
    }
    if (!(par->i_h > 0.0f) || !(par->m > 0.0f)) {
        set_error("h / m not derived (call lq_params_derive)");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->artificial_pressure == 1 && !(par->kernel_dr > 0.0f)) {
        set_error("kernel-dr not derived for artificial-pressure = 1 (call lq_params_derive)");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->constitutive_model < 1 || par->constitutive_model > 3) {
        set_error("constitutive-model must be 1 (Drucker-Prager), 2 (Wu hypoplastic) or 3 (von Wolffersdorff)");
        return LQ_ERR_UNSUPPORTED;
    }
    return LQ_OK;
}

static bool state_ok(const lq_state* s) {
    return s && s->xd && s->yd && s->zd && s->pospres && s->velrhop && s->idep && s->str && s->stn && s->orig;
}
static bool ini_ok(const lq_state_ini* s) {
    return s && s->xd && s->yd && s->zd && s->velrhop && s->epvr && s->str && s->stn;
}
static bool rates_ok(const lq_rates* r) { return r && r->accad && r->spin && r->acestn; }

static RunDesc one_desc(const lq_params* par, const lq_vw_params* vw) {
    RunDesc d;
    fill_desc_static(d, *par, vw);
    return d;
}

static lq_status last_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return LQ_ERR_HIP;
    }
    return LQ_OK;
}

extern "C" const char* lq_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* lq_version(void) { return "loquat_amd 1.0 (gfx950)"; }

extern "C" lq_status lq_neighbour_r2_cut(const lq_params* par, float* r2_cut) {
    if (!par || !r2_cut || !(par->i_h > 0.0f)) {
        set_error("lq_neighbour_r2_cut: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    *r2_cut = neighbour_r2_cut(par->i_h);
    return LQ_OK;
}

extern "C" lq_status lq_set_interaction_variant(int32_t v) {
    if (v < 0 || v >= LQ_INTERACTION_VARIANTS) {
        set_error("interaction variant must be 0 .. LQ_INTERACTION_VARIANTS - 1");
        return LQ_ERR_INVALID_ARG;
    }
    g_variant = v;
    return LQ_OK;
}
extern "C" int32_t lq_get_interaction_variant(void) { return g_variant; }

// ------------------------------------------------------------- grid handle
struct lq_grid {
    GridBuffers g;
    int n = -1;
};

extern "C" lq_status lq_grid_create(int64_t np_cap, int64_t ncell_cap, lq_grid** out) {
    if (!out || np_cap < 0 || ncell_cap < 0 || ncell_cap > 0x7fffffff) {
        set_error("lq_grid_create: invalid capacity");
        return LQ_ERR_INVALID_ARG;
    }
    if (np_cap > LQ_MAX_PARTICLES) {
        set_error("lq_grid_create: np_cap exceeds LQ_MAX_PARTICLES (32-bit gather offsets)");
        return LQ_ERR_CAPACITY;
    }
    lq_grid* g = new lq_grid();
    lq_status st = alloc_grid(g->g, np_cap, ncell_cap);
    if (st != LQ_OK) {
        delete g;
        return st;
    }
    *out = g;
    return LQ_OK;
}

extern "C" lq_status lq_grid_destroy(lq_grid* grid) {
    if (!grid) return LQ_OK;
    free_grid(grid->g);
    delete grid;
    return LQ_OK;
}

extern "C" lq_status lq_grid_build(lq_grid* grid, const lq_params* par, const float* pospres, lq_stream stream) {
    if (!grid || !pospres) {
        set_error("lq_grid_build: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(check_params(par));
    if (par->np > grid->g.np_cap || par->ncell > grid->g.ncell_cap) {
        set_error("lq_grid_build: np or ncell exceeds the grid capacity");
        return LQ_ERR_CAPACITY;
    }
    hipStream_t s = (hipStream_t)stream;
    const RunDesc d = one_desc(par, nullptr);
    launch_grid_build_one(d, (const float4*)pospres, grid->g, s);
    grid->n = par->np;
    return last_launch("lq_grid_build");
}

extern "C" lq_status lq_grid_build_noalloc(const lq_params* par, const float* pospres, int32_t* part_cel,
                                           int32_t* part_idx, int32_t* cell_beg, int32_t* cell_end,
                                           lq_stream stream) {
    if (!pospres || !part_cel || !part_idx || !cell_beg || !cell_end) {
        set_error("lq_grid_build_noalloc: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(check_params(par));
    if (par->np > LQ_NOALLOC_MAX_PARTICLES) {
        set_error("lq_grid_build_noalloc: np exceeds LQ_NOALLOC_MAX_PARTICLES");
        return LQ_ERR_CAPACITY;
    }
    LQ_TRY(launch_grid_build_noalloc(one_desc(par, nullptr), (const float4*)pospres, part_cel, part_idx, cell_beg,
                                     cell_end, (hipStream_t)stream));
    return last_launch("lq_grid_build_noalloc");
}

extern "C" lq_status lq_grid_permute(lq_grid* grid, int32_t np, const lq_state* src, const lq_state* dst,
                                     const lq_state_ini* ini_src, const lq_state_ini* ini_dst, lq_stream stream) {
    if (!grid || !state_ok(src) || !state_ok(dst) || (ini_src && !ini_ok(ini_src)) ||
        (ini_dst && !ini_ok(ini_dst)) || (!ini_src) != (!ini_dst)) {
        set_error("lq_grid_permute: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (grid->n < 0 || np != grid->n) {
        set_error("lq_grid_permute: np differs from the last lq_grid_build");
        return LQ_ERR_STATE;
    }
    DevIni a, b;
    if (ini_src) {
        a = to_dev(*ini_src);
        b = to_dev(*ini_dst);
    }
    launch_permute(grid->g.vals[grid->g.sorted_slot], np, to_dev(*src), to_dev(*dst), ini_src ? &a : nullptr,
                   ini_dst ? &b : nullptr, nullptr, (hipStream_t)stream);
    return last_launch("lq_grid_permute");
}

extern "C" lq_status lq_grid_arrays(const lq_grid* grid, const int32_t** part_cel, const int32_t** part_idx,
                                    const int32_t** cell_beg, const int32_t** cell_end) {
    if (!grid) {
        set_error("lq_grid_arrays: null grid");
        return LQ_ERR_INVALID_ARG;
    }
    if (grid->n < 0) {
        set_error("lq_grid_arrays: no build yet");
        return LQ_ERR_STATE;
    }
    if (part_cel) *part_cel = (const int32_t*)grid->g.keys[grid->g.sorted_slot];
    if (part_idx) *part_idx = (const int32_t*)grid->g.vals[grid->g.sorted_slot];
    if (cell_beg) *cell_beg = grid->g.cell_beg;
    if (cell_end) *cell_end = grid->g.cell_end;
    return LQ_OK;
}

// ------------------------------------------------------- module kernels
extern "C" lq_status lq_adami_boundary(const lq_params* par, const lq_state* st, const int32_t* cell_beg,
                                       const int32_t* cell_end, lq_stream stream) {
    LQ_TRY(check_params(par));
    if (!state_ok(st) || !cell_beg || !cell_end) {
        set_error("lq_adami_boundary: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->npb == 0) return LQ_OK;  // src/adami_boundary.cu:135
    launch_adami_one(one_desc(par, nullptr), to_dev(*st), cell_beg, cell_end, g_variant, nullptr, nullptr,
                     (hipStream_t)stream);
    return last_launch("lq_adami_boundary");
}

extern "C" lq_status lq_particle_interaction(const lq_params* par, const lq_state* st, const lq_rates* rates,
                                             const int32_t* cell_beg, const int32_t* cell_end, int32_t* nbr_count,
                                             lq_stream stream) {
    LQ_TRY(check_params(par));
    if (!state_ok(st) || !rates_ok(rates) || !cell_beg || !cell_end) {
        set_error("lq_particle_interaction: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->npm == 0) return LQ_OK;
    launch_interaction_one(one_desc(par, nullptr), to_dev(*st), to_dev(*rates), cell_beg, cell_end, nbr_count,
                           g_variant, nullptr, (hipStream_t)stream);
    return last_launch("lq_particle_interaction");
}

static lq_status check_disp(const lq_params* par, float disp, const char* who) {
    if (!(disp >= 0.0f) || !(2.0f * disp <= STALE_LIMIT * (2.0f / par->i_h))) {
        set_error(std::string(who) + ": displacement must lie in [0, 0.45 * 2h]");
        return LQ_ERR_INVALID_ARG;
    }
    return LQ_OK;
}

extern "C" lq_status lq_particle_interaction_stale(const lq_params* par, const lq_state* st, const lq_rates* rates,
                                                   const double* x_then, const double* y_then, const double* z_then,
                                                   float disp, const int32_t* cell_beg, const int32_t* cell_end,
                                                   int32_t* nbr_count, lq_stream stream) {
    LQ_TRY(check_params(par));
    if (!state_ok(st) || !rates_ok(rates) || !x_then || !y_then || !z_then || !cell_beg || !cell_end) {
        set_error("lq_particle_interaction_stale: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(check_disp(par, disp, "lq_particle_interaction_stale"));
    if (par->npm == 0) return LQ_OK;
    const StaleIn stale{x_then, y_then, z_then, nullptr, disp, nullptr};
    launch_interaction_one(one_desc(par, nullptr), to_dev(*st), to_dev(*rates), cell_beg, cell_end, nbr_count,
                           g_variant, &stale, (hipStream_t)stream);
    return last_launch("lq_particle_interaction_stale");
}

extern "C" lq_status lq_adami_boundary_stale(const lq_params* par, const lq_state* st, float disp,
                                             const int32_t* cell_beg, const int32_t* cell_end, int32_t* nbr_count,
                                             lq_stream stream) {
    LQ_TRY(check_params(par));
    if (!state_ok(st) || !cell_beg || !cell_end) {
        set_error("lq_adami_boundary_stale: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(check_disp(par, disp, "lq_adami_boundary_stale"));
    if (par->npb == 0) return LQ_OK;
    const StaleIn stale{nullptr, nullptr, nullptr, nullptr, disp, nullptr};
    launch_adami_one(one_desc(par, nullptr), to_dev(*st), cell_beg, cell_end, g_variant, &stale, nbr_count,
                     (hipStream_t)stream);
    return last_launch("lq_adami_boundary_stale");
}

static lq_status integrate(bool corr, const lq_params* par, const lq_vw_params* vw, const lq_state* st,
                           const lq_state_ini* ini, const lq_rates* rates, lq_stream stream) {
    LQ_TRY(check_params(par));
    if (!state_ok(st) || !ini_ok(ini) || !rates_ok(rates)) {
        set_error("lq_predictor/lq_corrector: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->constitutive_model == 3 && !vw) {
        set_error("constitutive model 3 needs lq_vw_params");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->npm == 0) return LQ_OK;  // src/predictor_corrector.cu:238,327
    launch_integrate_one(corr, one_desc(par, vw), to_dev(*st), to_dev(*ini), to_dev(*rates), (hipStream_t)stream);
    return last_launch(corr ? "lq_corrector" : "lq_predictor");
}

extern "C" lq_status lq_predictor(const lq_params* par, const lq_vw_params* vw, const lq_state* st,
                                  const lq_state_ini* ini, const lq_rates* rates, lq_stream stream) {
    return integrate(false, par, vw, st, ini, rates, stream);
}
extern "C" lq_status lq_corrector(const lq_params* par, const lq_vw_params* vw, const lq_state* st,
                                  const lq_state_ini* ini, const lq_rates* rates, lq_stream stream) {
    return integrate(true, par, vw, st, ini, rates, stream);
}

extern "C" lq_status lq_field_smoothing(const lq_params* par, const lq_state* st, float* str_cache,
                                        const int32_t* cell_beg, const int32_t* cell_end, lq_stream stream) {
    LQ_TRY(check_params(par));
    if (!state_ok(st) || !str_cache || !cell_beg || !cell_end) {
        set_error("lq_field_smoothing: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (par->npm == 0) return LQ_OK;  // src/field_smoothing.cu:134
    hipStream_t s = (hipStream_t)stream;
    DevState ds = to_dev(*st);
    launch_shepard_one(one_desc(par, nullptr), ds, ds.str, (tensor2d*)str_cache, cell_beg, cell_end, par->np,
                       g_variant, s);
    launch_copy_tensor((const tensor2d*)str_cache, ds.str, par->np, s);
    return last_launch("lq_field_smoothing");
}

// ------------------------------------------------------------------ engine
namespace {

struct RunSlot {
    int32_t id;
    RunDesc desc;  // host mirror; dynamic fields advanced in lockstep with k_step_end
    int64_t nsteps;
    int32_t failed;    // non-finite state seen by the corrector (lq_engine_poll)
    int32_t np_input;  // particles the run was added with (deleted ones included)
};

struct Buffers {
    DevState S[2];
    DevIni I;  // written by the predictor, read by the corrector
    // Variant 19 (split integration): the half step's state is I's positions,
    // velocity and stress with hpos as its FP32 mirror; the step's start stays
    // in the state and is the corrector's ini set.  Its wall entries are
    // written by each build's permute.
    float4* hpos;
    DevRates R;
    tensor2d* T;
    int32_t* run_of;
    RunDesc* desc;
    int32_t* fail;  // per-run flag, raised on the device by the corrector
    uint32_t* dtred;  // per-run |v|max, |a|max of the dynamic time step (2 x DMAX_SLOTS per run, zero between steps)
    uint32_t* dmax;   // per-run largest displacement of the predictor half (float bits, DMAX_SLOTS per run)
    uint32_t* nout;   // per run: some particle ended the predictor half outside the grid
    uint32_t* dmaxb;  // per run: the corrector's largest distance from the run's cell table (DMAX_SLOTS per run)
    uint32_t* noutb;  // per run: some particle ended the corrector outside the grid
    int32_t* mat_idx;   // particle index of the t-th material particle (all runs)
    int32_t* wall_idx;  // particle index of the t-th wall particle (all runs)
    GridBuffers grid;
};

template <class T>
bool dmalloc(T*& p, int64_t n, std::vector<void*>& owned) {
    void* q = nullptr;
    if (hipMalloc(&q, (size_t)(n > 0 ? n : 1) * sizeof(T)) != hipSuccess) return false;
    owned.push_back(q);
    p = (T*)q;
    return true;
}

}  // namespace

struct lq_engine {
    hipStream_t stream = nullptr;
    int64_t np_cap = 0, ncell_cap = 0;
    int32_t max_runs = 0;
    Buffers b{};
    std::vector<void*> owned;
    std::vector<RunSlot> runs;  // layout order
    int cur = 0;
    int64_t np_total = 0, ncell_total = 0;
    int64_t nmat_total = 0, nwall_total = 0;
    int32_t next_id = 0;
    bool use_graphs = true;
    int variant = LQ_DEFAULT_INTERACTION_VARIANT;  // interaction variant (set at creation, by the autotuner or
                                                   // lq_engine_set_interaction_variant)
    int zbits = 4;     // z-bins per cell (log2) in the engine's sort key
    int xtile = LQ_DEFAULT_COLUMN_STRIP;  // x-rows per strip of the (x, y) column order in the sort key
    int sort_every = LQ_DEFAULT_SORT_INTERVAL;  // a run's cell table is rebuilt every sort_every of its steps
    bool any_dynamic = false;  // some resident run has dynamic-time-step = 1
    int64_t nmat_art = 0;      // material particles of runs with artificial-pressure = 1 (last in mat_idx)
    // per (buffer, Shepard step, grid build: none / some runs / all runs)
    hipGraphExec_t graph[2][2][3] = {};
};

namespace {

// z-bin bits that fit beside the cell id in a 32-bit sort key.
int zbits_eff(const lq_engine* e) { return std::max(0, std::min(e->zbits, 32 - rs_bits(e->ncell_total))); }

void drop_graphs(lq_engine* e) {
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 3; ++k)
                if (e->graph[i][j][k]) {
                    (void)hipGraphExecDestroy(e->graph[i][j][k]);
                    e->graph[i][j][k] = nullptr;
                }
}

struct Prof {
    hipEvent_t ev[2];
    double ms[5];
};

struct StageTimer {
    Prof* p;
    hipStream_t s;
    int stage;
    StageTimer(Prof* p_, hipStream_t s_, int st) : p(p_), s(s_), stage(st) {
        if (p) (void)hipEventRecord(p->ev[0], s);
    }
    ~StageTimer() {
        if (!p) return;
        (void)hipEventRecord(p->ev[1], s);
        (void)hipEventSynchronize(p->ev[1]);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p->ev[0], p->ev[1]);
        p->ms[stage] += ms;
    }
};

enum { PG_GRID = 0, PG_ADAMI = 1, PG_INTER = 2, PG_INTEG = 3, PG_SHEP = 4 };

// Grid build of the resident state (which: the runs re-keyed, GridRekey) and
// the permute into the other state buffer; returns that buffer.
int enqueue_sort(lq_engine* e, int cur, int which, tensor2d* str_dst, hipStream_t s) {
    Buffers& b = e->b;
    const int n = (int)e->np_total;
    launch_grid_build_table(b.desc, b.run_of, b.S[cur].pospres, n, e->ncell_total, zbits_eff(e), b.grid, s, which);
    // the walls also go to the half-step state of split integration
    DevState half{};
    half.xd = b.I.xd;
    half.yd = b.I.yd;
    half.zd = b.I.zd;
    half.pospres = b.hpos;
    half.velrhop = b.I.velrhop;
    half.str = b.I.str;
    launch_permute(b.grid.vals[b.grid.sorted_slot], n, b.S[cur], b.S[1 - cur], nullptr, nullptr, str_dst, s, b.desc,
                   b.run_of, which, &half);
    return 1 - cur;
}

// Sorts every resident run afresh, so that the neighbour table is that of the
// current positions (the runs' recorded bounds dtab stay: conservative).
void sort_resident(lq_engine* e) {
    e->cur = enqueue_sort(e, e->cur, REKEY_ALL, nullptr, e->stream);
}

// Host mirror of k_step_begin's sort_due.
bool sort_due(const lq_engine* e, const RunDesc& d) { return d.active && d.since_sort >= e->sort_every; }

// Which grid build a step needs: 0 none (every active run keeps its table),
// 1 the runs due, 2 all runs.
int sort_state(const lq_engine* e) {
    int due = 0, active = 0;
    for (const RunSlot& r : e->runs) {
        if (!r.desc.active) continue;
        ++active;
        due += sort_due(e, r.desc) ? 1 : 0;
    }
    return due == 0 ? 0 : (due == active ? 2 : 1);
}

// One full step for every resident run, starting from buffer `cur`; returns
// the buffer that holds the state afterwards (step_out).
//
// A run's cell table is rebuilt at the start of every sort_every-th step of
// its own (sort_state: none, some or all of the resident runs), and by a
// Shepard filter; the runs not due keep their sort keys, so their particles
// stay in place and their table is the one of the last build.  Between
// builds every scan reads the kept table: the particles have moved by at
// most the run's recorded bound since (dtab at step start, the predictor's
// dmax in the corrector half), and the stale scans widen their reach by it
// and test each hit against its current cell, so every neighbour set is the
// one a fresh sort gives (the reference sorts in both half steps,
// src/loquat.cu:345-366).  The ini set is in the order of the state, so the
// corrector reads it directly.
int enqueue_step(lq_engine* e, int cur, bool shepard, int sorts, Prof* prof) {
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    const int nr = (int)e->runs.size();
    const int var = e->variant;
    launch_step_begin(b.desc, nr, b.dmax, b.nout, s, e->sort_every, b.fail);
    if (e->any_dynamic)
        launch_dynamic_dt(b.desc, b.run_of, b.S[cur], b.R, IndexMap{b.mat_idx, 0}, (int)e->nmat_total, nr, b.dtred, s);
    int c = cur;
    const IndexMap walls{b.wall_idx, 0}, mats{b.mat_idx, 0};
    const int nw = (int)e->nwall_total, nm = (int)e->nmat_total;
    // variant 19: the predictor writes the half step beside the step's start
    // (no ini copy); the corrector half's scans read it.  Without artificial
    // pressure or a dynamic time step (which reads the rates), the
    // interaction kernel integrates its particles itself.
    const bool split = var == 19;
    const bool fuse = split && e->nmat_art == 0 && !e->any_dynamic;
    for (int half = 0; half < 2; ++half) {
        const bool corr = half == 1;
        if (!corr && sorts > 0) {
            StageTimer t(prof, s, PG_GRID);
            c = enqueue_sort(e, c, sorts == 2 ? REKEY_ALL : REKEY_SORT_DUE, nullptr, s);
        }
        DevState hs = b.S[c];
        hs.xd = b.I.xd;
        hs.yd = b.I.yd;
        hs.zd = b.I.zd;
        hs.pospres = b.hpos;
        hs.velrhop = b.I.velrhop;
        hs.str = b.I.str;
        const DevState& sc = (split && corr) ? hs : b.S[c];  // what this half's scans read
        // predictor half: the tables kept from earlier steps (from_desc 1:
        // the runs re-keyed just now have bound 0; 2: none was); corrector
        // half: the table of this step's start and the predictor's bound
        StaleIn stale{};
        stale.po = b.S[c].xo;
        if (corr) {
            stale.dmax = b.dmax;
            stale.nout = b.nout;
        } else {
            stale.from_desc = sorts == 1 ? 1 : 2;
        }
        const StaleIn* sp = (corr || sorts != 2) ? &stale : nullptr;
        {
            StageTimer t(prof, s, PG_ADAMI);
            launch_adami_table(b.desc, b.run_of, sc, b.grid.cell_beg, b.grid.cell_end, walls, nw, var, sp, s,
                               (split && corr) ? b.S[c].str : nullptr);
        }
        FuseIn fz{};
        if (corr) {
            fz.other = b.S[c];
            fz.fail = b.fail;
            fz.dmax_pred = b.dmax;
            fz.dmax = b.dmaxb;
            fz.nout = b.noutb;
        } else {
            fz.other = hs;
            fz.dmax = b.dmax;
            fz.nout = b.nout;
        }
        {
            StageTimer t(prof, s, PG_INTER);
            launch_interaction_table(b.desc, b.run_of, sc, b.R, b.grid.cell_beg, b.grid.cell_end, mats, nm,
                                     (int)e->nmat_art, nullptr, var, sp, s, fuse ? &fz : nullptr, corr);
        }
        {
            StageTimer t(prof, s, PG_INTEG);
            if (!fuse)
                launch_integrate_table(corr, b.desc, b.run_of, b.S[c], b.I, b.R, mats, nm, b.fail, b.dmax, b.nout,
                                       s, b.dmaxb, b.noutb, split ? &hs : nullptr);
            if (!corr) launch_dmax_fold(b.dmax, nr, s);
        }
    }
    if (shepard) {
        {
            StageTimer t(prof, s, PG_GRID);
            c = enqueue_sort(e, c, REKEY_SHEPARD_DUE, b.T, s);
        }
        StageTimer t(prof, s, PG_SHEP);
        launch_shepard_table(b.desc, b.run_of, b.S[c], b.T, b.S[c].str, b.grid.cell_beg, b.grid.cell_end,
                             (int)e->np_total, var, s);
    }
    launch_step_end(b.desc, nr, b.fail, s, b.dmaxb, b.noutb, sorts > 0, shepard);
    return c;
}

// Buffer holding the state after a step from `cur` (one permute per grid
// build, and one for a Shepard filter).
inline int step_out(int cur, bool shepard, int sorts) { return (shepard != (sorts > 0)) ? 1 - cur : cur; }

// Host mirror of k_step_begin's Shepard test over all active runs.
bool any_shepard_due(const lq_engine* e) {
    for (const RunSlot& r : e->runs) {
        const RunDesc& d = r.desc;
        const int f = d.par.stress_shepard_filter;
        if (d.active && f != 0 && d.step % f == 0 && d.step != 0) return true;
    }
    return false;
}

// Host mirror of k_step_end (sorts: the step's grid build, shep: whether it
// ran a Shepard filter).
void host_step_end(lq_engine* e, int sorts, bool shep) {
    for (RunSlot& r : e->runs) {
        RunDesc& d = r.desc;
        if (!d.active) continue;
        const int f = d.par.stress_shepard_filter;
        const bool shep_sorted = shep && f != 0 && d.step % f == 0 && d.step != 0;
        d.since_sort = shep_sorted ? 0 : ((sorts > 0 && sort_due(e, d)) ? 1 : d.since_sort + 1);
        d.time += (double)d.par.dt;
        d.step += 1;
        if (d.step >= d.nsteps) d.active = 0;
    }
}

// Uploads the run table, the failure flags and the wall / material index maps
// of the current layout.
// Refreshes the host mirror of what the device advances by itself (step
// counter, time, dt, active flag): required before the mirror is uploaded
// again, and for reporting runs whose dt the device chooses.
lq_status sync_descs(lq_engine* e) {
    const size_t nr = e->runs.size();
    if (nr == 0) return LQ_OK;
    std::vector<RunDesc> h(nr);
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    LQ_HIP_TRY(hipMemcpy(h.data(), e->b.desc, nr * sizeof(RunDesc), hipMemcpyDeviceToHost));
    std::vector<int32_t> f(nr, 0);
    LQ_HIP_TRY(hipMemcpy(f.data(), e->b.fail, nr * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nr; ++i) {
        RunDesc& d = e->runs[i].desc;
        d.step = h[i].step;
        d.time = h[i].time;
        d.par.dt = h[i].par.dt;
        d.active = h[i].active;
        d.shepard_due = h[i].shepard_due;
        d.since_sort = h[i].since_sort;
        d.sort_due = h[i].sort_due;
        d.dtab = h[i].dtab;
        d.out_tab = h[i].out_tab;
        // the corrector's failure flag, before any upload writes the host
        // copy back (lq_engine_poll reads it the same way)
        if (f[i] && !e->runs[i].failed) {
            e->runs[i].failed = 1;
            d.active = 0;
        }
    }
    return LQ_OK;
}

lq_status upload_descs(lq_engine* e) {
    std::vector<RunDesc> h(e->runs.size());
    std::vector<int32_t> f(e->runs.size());
    std::vector<int32_t> walls, mats, mats_art;
    const int zb = zbits_eff(e);
    e->any_dynamic = false;
    for (size_t i = 0; i < e->runs.size(); ++i) {
        RunDesc& d = e->runs[i].desc;
        d.zbits = zb;
        d.xtile = e->xtile;
        e->any_dynamic = e->any_dynamic || d.par.dynamic_time_step == 1;
        h[i] = d;
        f[i] = e->runs[i].failed;
        for (int k = 0; k < d.npb; ++k) walls.push_back(d.p_off + k);
        std::vector<int32_t>& m = d.par.artificial_pressure == 1 ? mats_art : mats;
        for (int k = d.npb; k < d.np; ++k) m.push_back(d.p_off + k);
    }
    e->nmat_art = (int64_t)mats_art.size();
    mats.insert(mats.end(), mats_art.begin(), mats_art.end());
    e->nwall_total = (int64_t)walls.size();
    e->nmat_total = (int64_t)mats.size();
    if (!h.empty()) {
        LQ_HIP_TRY(hipMemcpyAsync(e->b.desc, h.data(), h.size() * sizeof(RunDesc), hipMemcpyHostToDevice, e->stream));
        LQ_HIP_TRY(hipMemcpyAsync(e->b.fail, f.data(), f.size() * 4, hipMemcpyHostToDevice, e->stream));
    }
    if (!walls.empty())
        LQ_HIP_TRY(hipMemcpyAsync(e->b.wall_idx, walls.data(), walls.size() * 4, hipMemcpyHostToDevice, e->stream));
    if (!mats.empty())
        LQ_HIP_TRY(hipMemcpyAsync(e->b.mat_idx, mats.data(), mats.size() * 4, hipMemcpyHostToDevice, e->stream));
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    return LQ_OK;
}

// Recomputes offsets and cell bases of all runs, moving every segment of
// buffer cur to its new place in the other buffer.
lq_status relayout(lq_engine* e) {
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    LQ_HIP_TRY(hipStreamSynchronize(s));
    const DevState& src = b.S[e->cur];
    const DevState& dst = b.S[1 - e->cur];
    // the sort keys each run keeps move with its segment and cell range
    const int ks = b.grid.sorted_slot;
    const int zb = zbits_eff(e);
    int64_t off = 0, cbase = 0;
    std::vector<int32_t> run_of;
    for (size_t i = 0; i < e->runs.size(); ++i) {
        RunDesc& d = e->runs[i].desc;
        const int64_t o = d.p_off, np = d.np;
        auto cp = [&](void* dp, const void* sp, size_t elem) {
            return hipMemcpyAsync((char*)dp + off * elem, (const char*)sp + o * elem, np * elem,
                                  hipMemcpyDeviceToDevice, s);
        };
        if (np > 0) {
            LQ_HIP_TRY(cp(dst.xd, src.xd, 8));
            LQ_HIP_TRY(cp(dst.yd, src.yd, 8));
            LQ_HIP_TRY(cp(dst.zd, src.zd, 8));
            LQ_HIP_TRY(cp(dst.pospres, src.pospres, 16));
            LQ_HIP_TRY(cp(dst.velrhop, src.velrhop, 16));
            LQ_HIP_TRY(cp(dst.idep, src.idep, 16));
            LQ_HIP_TRY(cp(dst.str, src.str, 24));
            LQ_HIP_TRY(cp(dst.stn, src.stn, 24));
            LQ_HIP_TRY(cp(dst.orig, src.orig, 4));
            LQ_HIP_TRY(cp(dst.xo, src.xo, 16));
            if (ks >= 0)
                launch_shift_keys(b.grid.keys[ks] + o, b.grid.keys[1 - ks] + off, (int)np,
                                  (uint32_t)((cbase - (int64_t)d.cell_base) * ((int64_t)1 << zb)), s);
        }
        d.p_off = (int32_t)off;
        d.cell_base = (int32_t)cbase;
        run_of.insert(run_of.end(), (size_t)np, (int32_t)i);
        off += np;
        cbase += d.par.ncell;
    }
    e->np_total = off;
    e->ncell_total = cbase;
    e->cur = 1 - e->cur;
    if (!run_of.empty())
        LQ_HIP_TRY(hipMemcpyAsync(b.run_of, run_of.data(), run_of.size() * 4, hipMemcpyHostToDevice, s));
    LQ_TRY(upload_descs(e));
    drop_graphs(e);
    // The cell table of the new layout: every run keeps its keys (moved
    // above), unless none were kept or their format changed (the z-bin bits
    // depend on the total cell count): then every run is sorted afresh.
    if (e->np_total > 0) {
        const bool keep = ks >= 0 && zbits_eff(e) == zb;
        if (keep) b.grid.sorted_slot = 1 - ks;
        e->cur = enqueue_sort(e, e->cur, keep ? REKEY_NONE : REKEY_ALL, nullptr, s);
        LQ_TRY(last_launch("relayout"));
    }
    return LQ_OK;
}

}  // namespace

extern "C" int64_t lq_engine_bytes_per_run(int64_t np, int64_t ncell) {
    // 2 states (124 B) + ini (96 B) + rates (56 B) + Shepard scratch (24 B)
    // + run id (4 B) + wall/material index (4 B) + sort keys/values (16 B)
    // + histograms, + table 8 B/cell.
    const int64_t per_particle = 2 * 124 + 96 + 56 + 24 + 4 + 4 + 16 + 1;
    return np * per_particle + ncell * 8 + (int64_t)sizeof(RunDesc);
}

extern "C" lq_status lq_engine_create(int64_t np_cap, int64_t ncell_cap, int32_t max_runs, lq_engine** out) {
    if (!out || np_cap <= 0 || ncell_cap <= 0 || max_runs <= 0 || ncell_cap > 0x7fffffff) {
        set_error("lq_engine_create: invalid capacity");
        return LQ_ERR_INVALID_ARG;
    }
    if (np_cap > LQ_MAX_PARTICLES) {
        set_error("lq_engine_create: np_cap exceeds LQ_MAX_PARTICLES (32-bit gather offsets)");
        return LQ_ERR_CAPACITY;
    }
    lq_engine* e = new lq_engine();
    e->variant = g_variant;
    e->np_cap = np_cap;
    e->ncell_cap = ncell_cap;
    e->max_runs = max_runs;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        set_error("hipStreamCreate failed");
        return LQ_ERR_HIP;
    }
    Buffers& b = e->b;
    bool ok = true;
    for (int i = 0; i < 2 && ok; ++i) {
        DevState& S = b.S[i];
        ok = ok && dmalloc(S.xd, np_cap, e->owned) && dmalloc(S.yd, np_cap, e->owned) &&
             dmalloc(S.zd, np_cap, e->owned) && dmalloc(S.pospres, np_cap + SCAN_PAD, e->owned) &&
             dmalloc(S.velrhop, np_cap, e->owned) && dmalloc(S.idep, np_cap, e->owned) &&
             dmalloc(S.str, np_cap, e->owned) && dmalloc(S.stn, np_cap, e->owned) &&
             dmalloc(S.orig, np_cap, e->owned) && dmalloc(S.xo, np_cap, e->owned);
    }
    {
        DevIni& I = b.I;
        ok = ok && dmalloc(I.xd, np_cap, e->owned) && dmalloc(I.yd, np_cap, e->owned) &&
             dmalloc(I.zd, np_cap, e->owned) && dmalloc(I.velrhop, np_cap, e->owned) &&
             dmalloc(I.epvr, np_cap, e->owned) && dmalloc(I.str, np_cap, e->owned) && dmalloc(I.stn, np_cap, e->owned) &&
             dmalloc(b.hpos, np_cap + SCAN_PAD, e->owned);
    }
    ok = ok && dmalloc(b.R.accad, np_cap, e->owned) && dmalloc(b.R.spin, np_cap, e->owned) &&
         dmalloc(b.R.acestn, np_cap, e->owned) && dmalloc(b.T, np_cap, e->owned) &&
         dmalloc(b.run_of, np_cap, e->owned) && dmalloc(b.desc, max_runs, e->owned) &&
         dmalloc(b.fail, max_runs, e->owned) && dmalloc(b.mat_idx, np_cap, e->owned) &&
         dmalloc(b.wall_idx, np_cap, e->owned) && dmalloc(b.dtred, 2 * (int64_t)max_runs * DMAX_SLOTS, e->owned) &&
         dmalloc(b.dmax, (int64_t)max_runs * DMAX_SLOTS, e->owned) && dmalloc(b.nout, max_runs, e->owned) &&
         dmalloc(b.dmaxb, (int64_t)max_runs * DMAX_SLOTS, e->owned) && dmalloc(b.noutb, max_runs, e->owned);
    ok = ok && hipMemset(b.dtred, 0, 2 * (size_t)max_runs * DMAX_SLOTS * sizeof(uint32_t)) == hipSuccess &&
         hipMemset(b.dmaxb, 0, (size_t)max_runs * DMAX_SLOTS * sizeof(uint32_t)) == hipSuccess &&
         hipMemset(b.noutb, 0, (size_t)max_runs * sizeof(uint32_t)) == hipSuccess;
    if (!ok || alloc_grid(b.grid, np_cap, ncell_cap) != LQ_OK) {
        for (void* p : e->owned) (void)hipFree(p);
        (void)hipStreamDestroy(e->stream);
        delete e;
        set_error("lq_engine_create: device allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    *out = e;
    return LQ_OK;
}

extern "C" lq_status lq_engine_destroy(lq_engine* e) {
    if (!e) return LQ_OK;
    (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    free_grid(e->b.grid);
    for (void* p : e->owned) (void)hipFree(p);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return LQ_OK;
}

extern "C" lq_stream lq_engine_stream(const lq_engine* e) { return e ? (lq_stream)e->stream : nullptr; }

extern "C" lq_status lq_engine_set_graphs(lq_engine* e, int32_t use_graphs) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    e->use_graphs = use_graphs != 0;
    return LQ_OK;
}

extern "C" lq_status lq_engine_set_interaction_variant(lq_engine* e, int32_t variant) {
    if (!e || variant < 0 || variant >= LQ_INTERACTION_VARIANTS) {
        set_error("lq_engine_set_interaction_variant: variant must be 0 .. LQ_INTERACTION_VARIANTS - 1");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    e->variant = variant;
    drop_graphs(e);
    return LQ_OK;
}

extern "C" int32_t lq_engine_get_interaction_variant(const lq_engine* e) { return e ? e->variant : -1; }

extern "C" lq_status lq_engine_set_sort_interval(lq_engine* e, int32_t steps) {
    if (!e || steps < 1 || steps > LQ_MAX_SORT_INTERVAL) {
        set_error("lq_engine_set_sort_interval: steps must be 1 .. LQ_MAX_SORT_INTERVAL");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    e->sort_every = steps;
    drop_graphs(e);
    return LQ_OK;
}

extern "C" int32_t lq_engine_get_sort_interval(const lq_engine* e) { return e ? e->sort_every : -1; }

extern "C" lq_status lq_engine_set_cell_order(lq_engine* e, int32_t zbin_bits) {
    if (!e || zbin_bits < 0 || zbin_bits > 8) {
        set_error("lq_engine_set_cell_order: zbin_bits must be 0..8");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(sync_descs(e));
    e->zbits = zbin_bits;
    LQ_TRY(upload_descs(e));
    drop_graphs(e);
    // the kept sort keys carry the z-bins: every run is sorted afresh
    sort_resident(e);
    return last_launch("lq_engine_set_cell_order");
}

extern "C" lq_status lq_engine_set_column_order(lq_engine* e, int32_t strip_rows) {
    if (!e || strip_rows < 0 || strip_rows > LQ_MAX_COLUMN_STRIP) {
        set_error("lq_engine_set_column_order: strip_rows must be 0 .. LQ_MAX_COLUMN_STRIP");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(sync_descs(e));
    e->xtile = strip_rows;
    LQ_TRY(upload_descs(e));
    drop_graphs(e);
    // the kept sort keys carry the column ranks: every run is sorted afresh
    sort_resident(e);
    return last_launch("lq_engine_set_column_order");
}
extern "C" int32_t lq_engine_get_column_order(const lq_engine* e) { return e ? e->xtile : -1; }

static int64_t steps_until(double t, double tmax, double dt) {
    int64_t n = 0;
    while (t < tmax) {
        t += dt;
        ++n;
    }
    return n;
}

namespace lq {
// Everything lq_engine_add_run checks about a run before it touches the
// device: arguments, derived parameters, model-3 constants, and that every
// wall particle lies inside the cell grid (the wall/material split is by
// index after each sort; a wall sent to ncell_delete would land past the
// material block).  par_out receives the derived parameters.  Host only, so
// the server validates requests at submission with it.
lq_status check_run_input(const lq_run_input* in, lq_params* par_out) {
    if (!in || !in->host.pospres || !in->host.velrhop || !in->host.idep || !in->host.str || !in->host.stn) {
        set_error("run input: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    lq_params par = in->par;
    LQ_TRY(lq_params_derive(&par));
    LQ_TRY(check_params(&par));
    if (par.constitutive_model == 3 && !(in->vw.phi_c > 0.0f && in->vw.hs > 0.0f && in->vw.n > 0.0f)) {
        set_error("constitutive model 3 needs phi_c, hs, n > 0");
        return LQ_ERR_INVALID_ARG;
    }
    for (int i = 0; i < par.npb; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float v = in->posd ? (float)in->posd[3 * (size_t)i + a] : in->host.pospres[4 * (size_t)i + a];
            const int n = a == 0 ? par.cell_num_x : (a == 1 ? par.cell_num_y : par.cell_num_z);
            const float f = floorf((0.5f * v) * par.i_h);
            if (!(f >= 0.0f) || !(f < (float)n)) {
                set_error("wall particle " + std::to_string(i) + " lies outside the cell grid");
                return LQ_ERR_DOMAIN;
            }
        }
    }
    *par_out = par;
    return LQ_OK;
}
}  // namespace lq

extern "C" lq_status lq_engine_add_run(lq_engine* e, const lq_run_input* in, int32_t* run_id) {
    if (!e || !run_id) {
        set_error("lq_engine_add_run: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    lq_params par;
    LQ_TRY(check_run_input(in, &par));
    LQ_TRY(sync_descs(e));
    if ((int64_t)e->runs.size() >= e->max_runs) {
        set_error("engine run table is full");
        return LQ_ERR_CAPACITY;
    }
    if (e->np_total + par.np > e->np_cap || e->ncell_total + par.ncell > e->ncell_cap) {
        set_error("engine particle or cell capacity exceeded");
        return LQ_ERR_CAPACITY;
    }
    const int np = par.np;
    const float* pp = in->host.pospres;
    std::vector<double> xd(np), yd(np), zd(np);
    std::vector<float> pospres(pp, pp + 4 * (size_t)np);
    for (int i = 0; i < np; ++i) {
        if (in->posd) {
            xd[i] = in->posd[3 * (size_t)i];
            yd[i] = in->posd[3 * (size_t)i + 1];
            zd[i] = in->posd[3 * (size_t)i + 2];
            pospres[4 * (size_t)i] = (float)xd[i];
            pospres[4 * (size_t)i + 1] = (float)yd[i];
            pospres[4 * (size_t)i + 2] = (float)zd[i];
        } else {
            xd[i] = pp[4 * (size_t)i];
            yd[i] = pp[4 * (size_t)i + 1];
            zd[i] = pp[4 * (size_t)i + 2];
        }
    }
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    RunSlot slot;
    slot.id = e->next_id++;
    fill_desc_static(slot.desc, par, &in->vw);
    slot.desc.p_off = (int32_t)e->np_total;
    slot.desc.cell_base = (int32_t)e->ncell_total;
    if (in->nsteps > 0) {
        slot.nsteps = in->nsteps;
    } else if (par.dynamic_time_step == 1) {
        // dt changes every step: the device ends the run at time_max
        slot.nsteps = 0x7fffffff;
        slot.desc.until_time = 1;
    } else {
        slot.nsteps = steps_until((double)par.time, (double)par.time_max, (double)par.dt);
    }
    slot.desc.nsteps = (int32_t)std::min<int64_t>(slot.nsteps, 0x7fffffff);
    slot.desc.active = slot.desc.nsteps > 0 && (!slot.desc.until_time || par.time < par.time_max) ? 1 : 0;
    slot.failed = 0;
    slot.np_input = par.np;
    const DevState& S = e->b.S[e->cur];
    const int64_t o = e->np_total;
    std::vector<int32_t> orig(np), run_of(np, (int32_t)e->runs.size());
    for (int i = 0; i < np; ++i) orig[i] = i;
    if (np > 0) {
        LQ_HIP_TRY(hipMemcpy(S.xd + o, xd.data(), np * 8, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.yd + o, yd.data(), np * 8, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.zd + o, zd.data(), np * 8, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.pospres + o, pospres.data(), np * 16, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.velrhop + o, in->host.velrhop, np * 16, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.idep + o, in->host.idep, np * 16, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.str + o, in->host.str, np * 24, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.stn + o, in->host.stn, np * 24, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(S.orig + o, orig.data(), np * 4, hipMemcpyHostToDevice));
        LQ_HIP_TRY(hipMemcpy(e->b.run_of + o, run_of.data(), np * 4, hipMemcpyHostToDevice));
    }
    // Initial neighbour search + Primary sort (src/loquat.cu:297-298) of the
    // new run only: the resident runs keep their keys and tables (their
    // segments do not move), unless the key format changes with the larger
    // cell count.
    const int zb_before = zbits_eff(e);
    const bool had_keys = !e->runs.empty() && e->b.grid.sorted_slot >= 0;
    for (RunSlot& r : e->runs) r.desc.sort_due = 0;
    slot.desc.sort_due = 1;
    slot.desc.since_sort = 0;
    e->runs.push_back(slot);
    e->np_total += np;
    e->ncell_total += par.ncell;
    LQ_TRY(upload_descs(e));
    drop_graphs(e);
    const bool keep = had_keys && zbits_eff(e) == zb_before;
    e->cur = enqueue_sort(e, e->cur, keep ? REKEY_SORT_DUE : REKEY_ALL, nullptr, e->stream);
    e->runs.back().desc.sort_due = 0;
    LQ_TRY(last_launch("lq_engine_add_run"));
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    *run_id = slot.id;
    return LQ_OK;
}

static int find_run(const lq_engine* e, int32_t id) {
    for (size_t i = 0; i < e->runs.size(); ++i)
        if (e->runs[i].id == id) return (int)i;
    return -1;
}

extern "C" lq_status lq_engine_remove_run(lq_engine* e, int32_t run_id) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    const int i = find_run(e, run_id);
    if (i < 0) {
        set_error("unknown run id");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_TRY(sync_descs(e));
    e->runs.erase(e->runs.begin() + i);
    return relayout(e);
}

extern "C" lq_status lq_engine_step(lq_engine* e, int32_t nsteps) {
    if (!e || nsteps < 0) {
        set_error("lq_engine_step: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    for (int it = 0; it < nsteps; ++it) {
        bool any_active = false;
        for (const RunSlot& r : e->runs) any_active = any_active || r.desc.active;
        if (!any_active) break;
        const bool shep = any_shepard_due(e);
        const int sorts = sort_state(e);
        if (e->use_graphs) {
            hipGraphExec_t& g = e->graph[e->cur][shep ? 1 : 0][sorts];
            if (!g) {
                hipGraph_t graph = nullptr;
                LQ_HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                (void)enqueue_step(e, e->cur, shep, sorts, nullptr);
                LQ_HIP_TRY(hipStreamEndCapture(e->stream, &graph));
                hipError_t err = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (err != hipSuccess) {
                    g = nullptr;
                    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
                    return LQ_ERR_HIP;
                }
            }
            LQ_HIP_TRY(hipGraphLaunch(g, e->stream));
        } else {
            (void)enqueue_step(e, e->cur, shep, sorts, nullptr);
            LQ_TRY(last_launch("lq_engine_step"));
        }
        e->cur = step_out(e->cur, shep, sorts);
        host_step_end(e, sorts, shep);
    }
    return LQ_OK;
}


extern "C" lq_status lq_engine_neighbour_counts(lq_engine* e, int32_t run_id, int32_t* counts) {
    if (!e || !counts) {
        set_error("lq_engine_neighbour_counts: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    const int i = find_run(e, run_id);
    if (i < 0) {
        set_error("unknown run id");
        return LQ_ERR_INVALID_ARG;
    }
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    LQ_HIP_TRY(hipStreamSynchronize(s));
    sort_resident(e);
    int32_t* nbr = nullptr;
    if (hipMalloc(&nbr, (size_t)std::max<int64_t>(e->np_total, 1) * 4) != hipSuccess) {
        (void)hipGetLastError();
        set_error("lq_engine_neighbour_counts: device allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    const IndexMap mats{b.mat_idx, 0};
    const int var = e->variant;
    launch_interaction_table(b.desc, b.run_of, b.S[e->cur], b.R, b.grid.cell_beg, b.grid.cell_end, mats,
                             (int)e->nmat_total, (int)e->nmat_art, nbr, var, nullptr, s);
    const RunDesc& d = e->runs[i].desc;
    std::vector<int32_t> c(d.np), orig(d.np);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess && d.np > 0) err = hipMemcpyAsync(c.data(), nbr + d.p_off, d.np * 4, hipMemcpyDeviceToHost, s);
    if (err == hipSuccess && d.np > 0)
        err = hipMemcpyAsync(orig.data(), b.S[e->cur].orig + d.p_off, d.np * 4, hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    (void)hipFree(nbr);
    if (err != hipSuccess) {
        set_error(std::string("lq_engine_neighbour_counts: ") + hipGetErrorString(err));
        return LQ_ERR_HIP;
    }
    for (int k = 0; k < d.np; ++k) {
        const int t = orig[k];
        if (t < 0 || t >= d.np) {
            set_error("corrupt original-index array");
            return LQ_ERR_STATE;
        }
        counts[t] = k < d.npb ? 0 : c[k];
    }
    return LQ_OK;
}

extern "C" lq_status lq_engine_neighbour_counts_stale(lq_engine* e, int32_t run_id, int32_t* counts) {
    if (!e || !counts) {
        set_error("lq_engine_neighbour_counts_stale: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    const int i = find_run(e, run_id);
    if (i < 0) {
        set_error("unknown run id");
        return LQ_ERR_INVALID_ARG;
    }
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    LQ_HIP_TRY(hipStreamSynchronize(s));
    int32_t* nbr = nullptr;
    if (hipMalloc(&nbr, (size_t)std::max<int64_t>(e->np_total, 1) * 4) != hipSuccess) {
        (void)hipGetLastError();
        set_error("lq_engine_neighbour_counts_stale: device allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    const IndexMap mats{b.mat_idx, 0};
    // the scan of a step's predictor half on the tables the runs keep
    StaleIn stale{};
    stale.po = b.S[e->cur].xo;
    stale.from_desc = 2;
    const int var = e->variant;
    launch_interaction_table(b.desc, b.run_of, b.S[e->cur], b.R, b.grid.cell_beg, b.grid.cell_end, mats,
                             (int)e->nmat_total, (int)e->nmat_art, nbr, var, &stale, s);
    const RunDesc& d = e->runs[i].desc;
    std::vector<int32_t> c(d.np), orig(d.np);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess && d.np > 0) err = hipMemcpyAsync(c.data(), nbr + d.p_off, d.np * 4, hipMemcpyDeviceToHost, s);
    if (err == hipSuccess && d.np > 0)
        err = hipMemcpyAsync(orig.data(), b.S[e->cur].orig + d.p_off, d.np * 4, hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    (void)hipFree(nbr);
    if (err != hipSuccess) {
        set_error(std::string("lq_engine_neighbour_counts_stale: ") + hipGetErrorString(err));
        return LQ_ERR_HIP;
    }
    for (int k = 0; k < d.np; ++k) {
        const int t = orig[k];
        if (t < 0 || t >= d.np) {
            set_error("corrupt original-index array");
            return LQ_ERR_STATE;
        }
        counts[t] = k < d.npb ? 0 : c[k];
    }
    return LQ_OK;
}

extern "C" lq_status lq_engine_autotune(lq_engine* e, int32_t nrep, int32_t* best, double* ms_per_variant) {
    if (!e || nrep <= 0) {
        set_error("lq_engine_autotune: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (e->runs.empty() || e->nmat_total == 0) {
        set_error("lq_engine_autotune: no material particles resident");
        return LQ_ERR_STATE;
    }
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    sort_resident(e);
    drop_graphs(e);
    hipEvent_t ev[2];
    LQ_HIP_TRY(hipEventCreate(&ev[0]));
    LQ_HIP_TRY(hipEventCreate(&ev[1]));
    const IndexMap mats{b.mat_idx, 0};
    // A run's steps launch the interaction 2K times per sort interval of K
    // steps: once on the fresh table of its sort, 2K - 1 times on the kept
    // table.  Both are timed (the kept-table scans with each run's recorded
    // displacement bound) and weighted so.
    StaleIn kept{};
    kept.po = b.S[e->cur].xo;
    kept.from_desc = 2;
    const double w_fresh = 1.0 / (2.0 * e->sort_every), w_kept = 1.0 - w_fresh;
    auto time_launches = [&](int v, const StaleIn* st, double& ms_out) -> lq_status {
        launch_interaction_table(b.desc, b.run_of, b.S[e->cur], b.R, b.grid.cell_beg, b.grid.cell_end, mats,
                                 (int)e->nmat_total, (int)e->nmat_art, nullptr, v, st, s);  // warm-up
        LQ_HIP_TRY(hipEventRecord(ev[0], s));
        for (int r = 0; r < nrep; ++r)
            launch_interaction_table(b.desc, b.run_of, b.S[e->cur], b.R, b.grid.cell_beg, b.grid.cell_end, mats,
                                     (int)e->nmat_total, (int)e->nmat_art, nullptr, v, st, s);
        LQ_HIP_TRY(hipEventRecord(ev[1], s));
        LQ_HIP_TRY(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        LQ_HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        ms_out = ms / nrep;
        return LQ_OK;
    };
    double best_ms = 1e30, start_ms = -1.0;
    int best_v = 3;
    const int start = e->variant;
    for (int v = 0; v < LQ_INTERACTION_VARIANTS; ++v) {
        double t_fresh = 0.0, t_kept = 0.0;
        LQ_TRY(time_launches(v, nullptr, t_fresh));
        if (e->sort_every > 1) LQ_TRY(time_launches(v, &kept, t_kept));
        const double per = e->sort_every > 1 ? w_fresh * t_fresh + w_kept * t_kept : t_fresh;
        if (ms_per_variant) ms_per_variant[v] = per;
        if (v == start) start_ms = per;
        if (per < best_ms) {
            best_ms = per;
            best_v = v;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    LQ_TRY(last_launch("lq_engine_autotune"));
    // the engine's variant stays unless another is clearly faster (1 %,
    // above the timing noise), so that calls do not flip between near-equal
    // variants
    if (start_ms > 0.0 && start_ms <= 1.01 * best_ms) best_v = start;
    e->variant = best_v;
    drop_graphs(e);
    if (best) *best = best_v;
    return LQ_OK;
}

extern "C" lq_status lq_engine_autotune_cell_order(lq_engine* e, int32_t nrep, int32_t* best_bits,
                                                   double* ms_per_order) {
    if (!e || nrep <= 0) {
        set_error("lq_engine_autotune_cell_order: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (e->runs.empty() || e->nmat_total == 0) {
        set_error("lq_engine_autotune_cell_order: no material particles resident");
        return LQ_ERR_STATE;
    }
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    LQ_HIP_TRY(hipStreamSynchronize(s));
    const IndexMap walls{b.wall_idx, 0}, mats{b.mat_idx, 0};
    const int var = e->variant;
    // what the in-cell order changes: the sort (key width), Adami and the
    // interaction (z-window pruning); both scans on a fresh table
    auto once = [&]() {
        sort_resident(e);
        launch_adami_table(b.desc, b.run_of, b.S[e->cur], b.grid.cell_beg, b.grid.cell_end, walls,
                           (int)e->nwall_total, var, nullptr, s);
        launch_interaction_table(b.desc, b.run_of, b.S[e->cur], b.R, b.grid.cell_beg, b.grid.cell_end, mats,
                                 (int)e->nmat_total, (int)e->nmat_art, nullptr, var, nullptr, s);
    };
    hipEvent_t ev[2];
    LQ_HIP_TRY(hipEventCreate(&ev[0]));
    LQ_HIP_TRY(hipEventCreate(&ev[1]));
    const int start = e->zbits;
    double ms_all[LQ_CELL_ORDER_CANDIDATES];
    double best_ms = 1e30;
    int best_b = start;
    lq_status st = LQ_OK;
    for (int bits = 0; bits < LQ_CELL_ORDER_CANDIDATES && st == LQ_OK; ++bits) {
        st = lq_engine_set_cell_order(e, bits);
        if (st != LQ_OK) break;
        once();  // warm-up, and the resident state in this order
        hipError_t err = hipEventRecord(ev[0], s);
        for (int r = 0; r < nrep && err == hipSuccess; ++r) once();
        if (err == hipSuccess) err = hipEventRecord(ev[1], s);
        if (err == hipSuccess) err = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (err != hipSuccess) {
            set_error(std::string("lq_engine_autotune_cell_order: ") + hipGetErrorString(err));
            st = LQ_ERR_HIP;
            break;
        }
        const double per = ms / nrep;
        ms_all[bits] = per;
        if (ms_per_order) ms_per_order[bits] = per;
        if (per < best_ms) {
            best_ms = per;
            best_b = bits;
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (st == LQ_OK) st = last_launch("lq_engine_autotune_cell_order");
    if (st != LQ_OK) {
        (void)lq_engine_set_cell_order(e, start);  // leave the engine as it was
        return st;
    }
    // the order the engine had stays unless another is clearly faster
    // (1 %, above the timing noise), so that calls do not flip between
    // near-equal orders
    if (start >= 0 && start < LQ_CELL_ORDER_CANDIDATES && ms_all[start] <= 1.01 * best_ms) best_b = start;
    LQ_TRY(lq_engine_set_cell_order(e, best_b));
    sort_resident(e);
    LQ_HIP_TRY(hipStreamSynchronize(s));
    if (best_bits) *best_bits = best_b;
    return LQ_OK;
}

extern "C" lq_status lq_engine_autotune_column_order(lq_engine* e, int32_t nrep, int32_t* best_rows,
                                                     double* ms_per_candidate) {
    static constexpr int kRows[LQ_COLUMN_ORDER_CANDIDATES] = {0, 2, 4, 8, 16};
    if (!e || nrep <= 0) {
        set_error("lq_engine_autotune_column_order: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    if (e->runs.empty() || e->nmat_total == 0) {
        set_error("lq_engine_autotune_column_order: no material particles resident");
        return LQ_ERR_STATE;
    }
    Buffers& b = e->b;
    hipStream_t s = e->stream;
    LQ_HIP_TRY(hipStreamSynchronize(s));
    const IndexMap walls{b.wall_idx, 0}, mats{b.mat_idx, 0};
    const int var = e->variant;
    // what the column order changes: where the sorted particles lie, hence
    // the L2 reuse of Adami's and the interaction's gathers (both scans on a
    // fresh table), and the sort itself
    auto once = [&]() {
        sort_resident(e);
        launch_adami_table(b.desc, b.run_of, b.S[e->cur], b.grid.cell_beg, b.grid.cell_end, walls,
                           (int)e->nwall_total, var, nullptr, s);
        launch_interaction_table(b.desc, b.run_of, b.S[e->cur], b.R, b.grid.cell_beg, b.grid.cell_end, mats,
                                 (int)e->nmat_total, (int)e->nmat_art, nullptr, var, nullptr, s);
    };
    hipEvent_t ev[2];
    LQ_HIP_TRY(hipEventCreate(&ev[0]));
    LQ_HIP_TRY(hipEventCreate(&ev[1]));
    const int start = e->xtile;
    double best_ms = 1e30, start_ms = -1.0;
    int best_r = start;
    lq_status st = LQ_OK;
    for (int c = 0; c < LQ_COLUMN_ORDER_CANDIDATES && st == LQ_OK; ++c) {
        st = lq_engine_set_column_order(e, kRows[c]);
        if (st != LQ_OK) break;
        once();  // warm-up, and the resident state in this order
        hipError_t err = hipEventRecord(ev[0], s);
        for (int r = 0; r < nrep && err == hipSuccess; ++r) once();
        if (err == hipSuccess) err = hipEventRecord(ev[1], s);
        if (err == hipSuccess) err = hipEventSynchronize(ev[1]);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (err != hipSuccess) {
            set_error(std::string("lq_engine_autotune_column_order: ") + hipGetErrorString(err));
            st = LQ_ERR_HIP;
            break;
        }
        const double per = ms / nrep;
        if (ms_per_candidate) ms_per_candidate[c] = per;
        if (kRows[c] == start) start_ms = per;
        if (per < best_ms) {
            best_ms = per;
            best_r = kRows[c];
        }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (st == LQ_OK) st = last_launch("lq_engine_autotune_column_order");
    if (st != LQ_OK) {
        (void)lq_engine_set_column_order(e, start);  // leave the engine as it was
        return st;
    }
    // the order the engine had stays unless another is clearly faster (1 %)
    if (start_ms > 0.0 && start_ms <= 1.01 * best_ms) best_r = start;
    LQ_TRY(lq_engine_set_column_order(e, best_r));
    LQ_HIP_TRY(hipStreamSynchronize(s));
    if (best_rows) *best_rows = best_r;
    return LQ_OK;
}

extern "C" lq_status lq_engine_profile(lq_engine* e, int32_t nsteps, double* ms_stage5) {
    if (!e || nsteps <= 0 || !ms_stage5) {
        set_error("lq_engine_profile: invalid argument");
        return LQ_ERR_INVALID_ARG;
    }
    Prof p;
    std::memset(&p, 0, sizeof(p));
    LQ_HIP_TRY(hipEventCreate(&p.ev[0]));
    LQ_HIP_TRY(hipEventCreate(&p.ev[1]));
    int done = 0;
    for (int it = 0; it < nsteps; ++it) {
        bool any_active = false;
        for (const RunSlot& r : e->runs) any_active = any_active || r.desc.active;
        if (!any_active) break;
        const bool shep = any_shepard_due(e);
        const int sorts = sort_state(e);
        e->cur = enqueue_step(e, e->cur, shep, sorts, &p);
        host_step_end(e, sorts, shep);
        ++done;
    }
    (void)hipEventDestroy(p.ev[0]);
    (void)hipEventDestroy(p.ev[1]);
    LQ_TRY(last_launch("lq_engine_profile"));
    for (int i = 0; i < 5; ++i) ms_stage5[i] = done ? p.ms[i] / done : 0.0;
    return LQ_OK;
}

extern "C" lq_status lq_engine_synchronize(lq_engine* e) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    return LQ_OK;
}

extern "C" lq_status lq_engine_run_info(const lq_engine* e, int32_t run_id, int64_t* steps_done, double* time,
                                        int32_t* finished, int32_t* np) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    const int i = find_run(e, run_id);
    if (i < 0) {
        set_error("unknown run id");
        return LQ_ERR_INVALID_ARG;
    }
    // the device advances step, time and dt on its own: report its values
    LQ_TRY(sync_descs(const_cast<lq_engine*>(e)));
    const RunDesc& d = e->runs[i].desc;
    if (steps_done) *steps_done = d.step;
    if (time) *time = d.time;
    if (finished) *finished = d.active ? 0 : 1;
    if (np) *np = e->runs[i].np_input;
    return LQ_OK;
}

extern "C" lq_status lq_engine_poll(lq_engine* e, int32_t* nfailed) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t nr = e->runs.size();
    std::vector<int32_t> f(nr, 0);
    if (nr) LQ_HIP_TRY(hipMemcpy(f.data(), e->b.fail, nr * 4, hipMemcpyDeviceToHost));
    int32_t total = 0;
    for (size_t i = 0; i < nr; ++i) {
        if (f[i] && !e->runs[i].failed) {
            e->runs[i].failed = 1;
            e->runs[i].desc.active = 0;
        }
        total += e->runs[i].failed;
    }
    if (nfailed) *nfailed = total;
    return LQ_OK;
}

extern "C" lq_status lq_engine_run_failed(const lq_engine* e, int32_t run_id, int32_t* failed) {
    if (!e || !failed) {
        set_error("lq_engine_run_failed: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    const int i = find_run(e, run_id);
    if (i < 0) {
        set_error("unknown run id");
        return LQ_ERR_INVALID_ARG;
    }
    *failed = e->runs[i].failed;
    return LQ_OK;
}

extern "C" lq_status lq_engine_totals(const lq_engine* e, int64_t* npm, int64_t* np, int64_t* ncell, int32_t* nruns) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    int64_t m = 0;
    for (const RunSlot& r : e->runs) m += r.desc.npm;
    if (npm) *npm = m;
    if (np) *np = e->np_total;
    if (ncell) *ncell = e->ncell_total;
    if (nruns) *nruns = (int32_t)e->runs.size();
    return LQ_OK;
}

extern "C" lq_status lq_engine_download(lq_engine* e, int32_t run_id, lq_host_particles* out, double* posd) {
    if (!e || !out) {
        set_error("lq_engine_download: null argument");
        return LQ_ERR_INVALID_ARG;
    }
    const int i = find_run(e, run_id);
    if (i < 0) {
        set_error("unknown run id");
        return LQ_ERR_INVALID_ARG;
    }
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    const RunDesc& d = e->runs[i].desc;
    const int np = d.np;
    const int64_t o = d.p_off;
    const DevState& S = e->b.S[e->cur];
    std::vector<double> xd(np), yd(np), zd(np);
    std::vector<float4> pp(np), vr(np), id(np);
    std::vector<tensor2d> str(np), stn(np);
    std::vector<int32_t> orig(np);
    if (np > 0) {
        LQ_HIP_TRY(hipMemcpy(xd.data(), S.xd + o, np * 8, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(yd.data(), S.yd + o, np * 8, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(zd.data(), S.zd + o, np * 8, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(pp.data(), S.pospres + o, np * 16, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(vr.data(), S.velrhop + o, np * 16, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(id.data(), S.idep + o, np * 16, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(str.data(), S.str + o, np * 24, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(stn.data(), S.stn + o, np * 24, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(orig.data(), S.orig + o, np * 4, hipMemcpyDeviceToHost));
    }
    const int n_in = e->runs[i].np_input;
    std::vector<char> seen((size_t)n_in, 0);
    for (int k = 0; k < np; ++k) {
        const int t = orig[k];
        if (t < 0 || t >= n_in) {
            set_error("corrupt original-index array");
            return LQ_ERR_STATE;
        }
        seen[t] = 1;
        if (out->pospres) std::memcpy(out->pospres + 4 * (size_t)t, &pp[k], 16);
        if (out->velrhop) std::memcpy(out->velrhop + 4 * (size_t)t, &vr[k], 16);
        if (out->idep) std::memcpy(out->idep + 4 * (size_t)t, &id[k], 16);
        if (out->str) std::memcpy(out->str + 6 * (size_t)t, &str[k], 24);
        if (out->stn) std::memcpy(out->stn + 6 * (size_t)t, &stn[k], 24);
        if (posd) {
            posd[3 * (size_t)t] = xd[k];
            posd[3 * (size_t)t + 1] = yd[k];
            posd[3 * (size_t)t + 2] = zd[k];
        }
    }
    // particles removed by lq_engine_compact: NaN positions, other fields untouched
    const float fnan = std::numeric_limits<float>::quiet_NaN();
    for (int t = 0; t < n_in; ++t) {
        if (seen[t]) continue;
        if (out->pospres)
            for (int a = 0; a < 3; ++a) out->pospres[4 * (size_t)t + a] = fnan;
        if (posd)
            for (int a = 0; a < 3; ++a) posd[3 * (size_t)t + a] = std::numeric_limits<double>::quiet_NaN();
    }
    return LQ_OK;
}

extern "C" lq_status lq_engine_compact(lq_engine* e, int64_t* removed) {
    if (!e) {
        set_error("null engine");
        return LQ_ERR_INVALID_ARG;
    }
    if (removed) *removed = 0;
    if (e->runs.empty()) return LQ_OK;
    LQ_TRY(sync_descs(e));
    // the table of each run's last sort: its particles outside the grid then
    // (cell ncell - 1) are the tail of its segment, and the others keep their
    // keys, order and schedule
    LQ_HIP_TRY(hipStreamSynchronize(e->stream));
    int64_t total = 0;
    for (RunSlot& r : e->runs) {
        RunDesc& d = r.desc;
        const int64_t cell = (int64_t)d.cell_base + d.par.ncell - 1;
        int32_t be[2] = {-1, -1};
        LQ_HIP_TRY(hipMemcpy(&be[0], e->b.grid.cell_beg + cell, 4, hipMemcpyDeviceToHost));
        LQ_HIP_TRY(hipMemcpy(&be[1], e->b.grid.cell_end + cell, 4, hipMemcpyDeviceToHost));
        if (be[0] < 0) continue;
        const int32_t ndel = be[1] - be[0];
        if (be[1] != d.p_off + d.np || be[0] < d.p_off + d.npb) {
            set_error("lq_engine_compact: particles outside the grid are not the material tail of their run");
            return LQ_ERR_STATE;
        }
        d.np -= ndel;
        d.npm -= ndel;
        d.par.np -= ndel;
        d.par.npm -= ndel;
        d.par.np_delete += ndel;
        total += ndel;
    }
    if (removed) *removed = total;
    if (total > 0) LQ_TRY(relayout(e));
    return LQ_OK;
}
