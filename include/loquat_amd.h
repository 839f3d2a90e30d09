/*
 * loquat_amd.h -- C ABI of the MI355X-native LOQUAT SPH time step.
 *
 * Plain C: pointers, sizes and status codes; no HIP or torch types.  Every
 * entry point returns an lq_status and never exits the process (the reference
 * exits through checkCudaErrors, include/helper_cuda.h:985-1016).
 *
 * Module boundaries follow the reference:
 *   loquat_grid          -> lq_grid_*                (src/loquat_grid.h:28-66)
 *   particle_interaction -> lq_particle_interaction  (src/particle_interaction.h:29-41)
 *   adami_boundary       -> lq_adami_boundary        (src/adami_boundary.h:29-36)
 *   predictor_corrector  -> lq_predictor/lq_corrector(src/predictor_corrector.h:34-67)
 *   field_smoothing      -> lq_field_smoothing       (src/field_smoothing.h:29-37)
 *   LoquatSph time loop  -> lq_engine_*              (src/loquat.cu:216-475)
 *   LoquatIo input side  -> lq_load_project          (src/loquat_io.cu:69-237)
 *
 * Device state is structure-of-arrays, sorted by cell, with FP64 positions
 * beside the reference's FP32 float4/tensor2d arrays.  Module functions
 * operate on caller-owned device arrays; all launches go to the given stream
 * and are asynchronous, as in the reference (the caller synchronises).
 */
#ifndef LOQUAT_AMD_H
#define LOQUAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lq_status {
    LQ_OK = 0,
    LQ_ERR_INVALID_ARG = 1,   /* null pointer, negative size, np != npb + npm, ... */
    LQ_ERR_HIP = 2,           /* a HIP runtime call failed */
    LQ_ERR_OUT_OF_MEMORY = 3, /* device allocation failed */
    LQ_ERR_IO = 4,            /* project file missing or unreadable */
    LQ_ERR_PARSE = 5,         /* malformed XML / .dat, wrong value type */
    LQ_ERR_MISSING_KEY = 6,   /* mandatory XML key absent (src/loquat_io.cu:56-66) */
    LQ_ERR_UNSUPPORTED = 7,   /* option the path does not implement */
    LQ_ERR_CAPACITY = 8,      /* more particles / cells / runs than the handle holds */
    LQ_ERR_DOMAIN = 9,        /* a wall particle lies outside the cell grid */
    LQ_ERR_STATE = 10         /* call made in the wrong order (e.g. permute before build) */
} lq_status;

/* Largest particle count of any call, grid or engine.  Neighbour gathers use
 * 32-bit byte offsets from each array's base (one shift per load instead of
 * 64-bit address arithmetic); the widest gathered element is a 24-byte
 * tensor2d, so indices must stay below 2^32 / 24.  Larger capacities return
 * LQ_ERR_CAPACITY before anything is allocated (about 97 GB of engine state;
 * a larger ensemble is split across engines, which share nothing). */
#define LQ_MAX_PARTICLES 178956970

/* hipStream_t passed opaquely; NULL is the null stream. */
typedef void* lq_stream;

/* Run-time constants.  Field for field, and byte for byte (208 B), the
 * reference's Parameters (include/types.h:89-155), so the same XML fills it
 * and a reference Parameters* can be passed in unchanged. */
typedef struct lq_params {
    int32_t np, npb, npm, np_delete;
    float domain_size_x, domain_size_y, domain_size_z;
    float domain_min_x, domain_min_y, domain_min_z;
    int32_t cell_num_x, cell_num_y, cell_num_z;
    int32_t ncell, ncell_one_layer, ncell_delete;
    float h, i_h, eta, dr, hdr;
    float m, cs, cfl, rho0;
    int32_t constitutive_model;   /* 1 Drucker-Prager, 2 Wu hypoplastic, 3 von Wolffersdorff */
    float model_par_1, model_par_2, model_par_3, model_par_4, model_par_5;
    int32_t density_shepard_filter, stress_shepard_filter;
    int32_t dynamic_time_step, integration_method, viscosity_type;
    float artificial_viscosity_alpha;
    int32_t artificial_pressure;
    float artificial_pressure_coefficient;
    float acc_x, acc_y, acc_z;
    float time_max, time, dt, output_frequency;
    int32_t output_number;
    float kernel_normalization_par, kernel_gradient_normalization_par;
    float kernel_zero, kernel_dr;
    int32_t output_format;
} lq_params;

/* Constants of constitutive model 3 (von Wolffersdorff hypoplasticity).  The
 * reference has no such model; its five model_par slots cannot hold eight
 * constants.  Angles in degrees, hs in Pa.  The update evaluates the rate
 * with p floored at 1 Pa and tan(psi) capped just below sqrt(2) (the bound
 * of stresses without a tensile principal value), returns a tensile mean
 * stress to zero and holds the void ratio (idep.z) in [e_d, e_i] at the new
 * pressure. */
typedef struct lq_vw_params {
    float phi_c, hs, n, ed0, ec0, ei0, alpha, beta;
} lq_vw_params;

/* Particle state on the device, sorted by cell.  np entries per array.
 *   xd, yd, zd : FP64 positions (the precise copy, integrated in FP64)
 *   pospres    : float4 (float)x, (float)y, (float)z, pressure -- always the
 *                FP32 rounding of xd/yd/zd; the grid and neighbour tests read it
 *   velrhop    : float4 vx, vy, vz, density
 *   idep       : float4 object id, equivalent plastic strain, void ratio, unused
 *   str, stn   : tensor2d (6 floats: xx yy zz xy xz yz) stress, total strain
 *   orig       : int32 index of the particle in the input file */
typedef struct lq_state {
    double* xd;
    double* yd;
    double* zd;
    float* pospres;
    float* velrhop;
    float* idep;
    float* str;
    float* stn;
    int32_t* orig;
} lq_state;

/* Start-of-step copy read by the corrector (the reference's *_ini arrays,
 * src/loquat.h:66-71, reduced to the fields Corrector_cuk reads,
 * src/predictor_corrector.cu:269-290).  epvr is float2 (ep, void ratio). */
typedef struct lq_state_ini {
    double* xd;
    double* yd;
    double* zd;
    float* velrhop;
    float* epvr;
    float* str;
    float* stn;
} lq_state_ini;

/* Rates produced by the interaction (src/loquat.h:73-76):
 *   accad  float4 ax, ay, az, drho/dt;  spin float4 (W_yx, W_zx, W_zy, 0);
 *   acestn tensor2d strain rate. */
typedef struct lq_rates {
    float* accad;
    float* spin;
    float* acestn;
} lq_rates;

/* ------------------------------------------------------------------ params */

/* InitializeParameters (src/loquat.cu:111-139): h, i_h, eta, m, cell counts,
 * ncell, Wendland constants, and dt = cfl*h/cs (src/loquat.cu:331-333). */
lq_status lq_params_derive(lq_params* par);

/* FP32 squared-distance threshold T: r2 < T  <=>  sqrtf(r2) * i_h < 2
 * (neighbour predicate q < 2 of src/adami_boundary.cu:41-43). */
lq_status lq_neighbour_r2_cut(const lq_params* par, float* r2_cut);

/* ---------------------------------------------------------------- project */

/* Host particle arrays in the reference's AoS layout (src/loquat_io.cu:207-213):
 * pospres/velrhop/idep float4[np], str/stn tensor2d[np]. */
typedef struct lq_host_particles {
    float* pospres;
    float* velrhop;
    float* idep;
    float* str;
    float* stn;
} lq_host_particles;

/* LoquatIo::LoadParameters (src/loquat_io.cu:69-201): reads <prefix>.xml.
 * Missing keys, wrong types and np != npb + npm return errors. */
lq_status lq_load_parameters(const char* prefix, lq_params* par, lq_vw_params* vw);

/* LoquatIo::LoadParticles (src/loquat_io.cu:204-237): reads par->np rows of 22
 * columns from <prefix>.dat into caller-allocated host arrays. */
lq_status lq_load_particles(const char* prefix, const lq_params* par, lq_host_particles* out);

/* LoquatIo::SaveXml + SaveParticles (src/loquat_io.cu:240-549): writes
 * <dir>/output_<N>.xml and <dir>/<N>.dat (N = par->output_number; the
 * reference's 22-column format), a binary <dir>/<N>.rst with what the .dat
 * cannot hold (FP64 positions -- posd, or the FP32 ones when posd is NULL --
 * and the model-3 void ratio idep.z), and with output_format 2 one
 * <dir>/<id>_<N>.vtu per object id.  h holds the particles in input order.
 * Every output is a restart (lq_load_parameters("<dir>/output_<N>"),
 * lq_load_particles("<dir>/<N>"), lq_load_restart("<dir>/<N>")), which the
 * reference's is not; the differences:
 *   - the DP shear modulus is written as elastic-G, the key the reader
 *     expects (the reference writes poisson-G, src/loquat_io.cu:272 vs :153);
 *   - the .dat name has no trailing newline (src/loquat_io.cu:354);
 *   - the .dat is written for output_format 2 as well;
 *   - numbers carry 9 significant digits, so they read back as the same
 *     floats (the reference prints 6);
 *   - model 3 writes its vw-* constants (vw may be NULL otherwise). */
lq_status lq_save_output(const char* dir, const lq_params* par, const lq_vw_params* vw, const lq_host_particles* h,
                         const double* posd);

/* Reads <prefix>.rst written by lq_save_output: posd (3*np doubles) and, when
 * idep is not NULL, idep[4*i+2] (void ratio) of each particle. */
lq_status lq_load_restart(const char* prefix, const lq_params* par, double* posd, float* idep);

/* Last error message of the calling thread (empty if none). */
const char* lq_last_error(void);

/* ------------------------------------------------------------ loquat_grid */

typedef struct lq_grid lq_grid;

/* LoquatGrid() + GridConfig + ResizeArrayLength (src/loquat_grid.h:47-58). */
lq_status lq_grid_create(int64_t np_cap, int64_t ncell_cap, lq_grid** out);
lq_status lq_grid_destroy(lq_grid* grid);

/* NeighborSearch (src/loquat_grid.h:60): cell hash of every particle from
 * pospres, stable radix sort of (cell, index), cell_beg/cell_end table
 * (cell_beg = -1 and cell_end = -1 for empty cells). */
lq_status lq_grid_build(lq_grid* grid, const lq_params* par, const float* pospres, lq_stream stream);

/* No counterpart in the reference.  lq_grid_build for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same cell hash, stable radix sort and
 * cell table, with results in caller-owned device arrays -- part_cel and
 * part_idx (int32[np], as lq_grid_arrays gives them) and cell_beg, cell_end
 * (int32[ncell]) -- and the sort's scratch in a fixed device array inside the
 * library (about 17 MB per GPU, held until the process ends).  It allocates
 * nothing.  np above LQ_NOALLOC_MAX_PARTICLES returns LQ_ERR_CAPACITY.  All
 * calls on one GPU share that scratch: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
#define LQ_NOALLOC_MAX_PARTICLES 2097152
lq_status lq_grid_build_noalloc(const lq_params* par, const float* pospres, int32_t* part_cel, int32_t* part_idx,
                                int32_t* cell_beg, int32_t* cell_end, lq_stream stream);

/* SortArraysPrimary / SortArraysAll (src/loquat.cu:190-213) through
 * SortArrayFloat4 / SortArrayTensor2d (src/loquat_grid.h:62-64):
 * dst[k] = src[part_idx[k]] for every state array; ini_src/ini_dst may be
 * NULL (Primary) or both set (All).  src and dst must not alias. */
lq_status lq_grid_permute(lq_grid* grid, int32_t np,
                          const lq_state* src, const lq_state* dst,
                          const lq_state_ini* ini_src, const lq_state_ini* ini_dst,
                          lq_stream stream);

/* GetPartCel / GetPartIdx / GetCellBeg / GetCellEnd (src/loquat_grid.h:51-54):
 * device pointers valid until the next build or destroy. */
lq_status lq_grid_arrays(const lq_grid* grid, const int32_t** part_cel, const int32_t** part_idx,
                         const int32_t** cell_beg, const int32_t** cell_end);

/* ------------------------------------------------------- module kernels */

/* AdamiBoundary (src/adami_boundary.cu:126-143): stress of wall particles
 * [0, npb) extrapolated from material neighbours. */
lq_status lq_adami_boundary(const lq_params* par, const lq_state* st,
                            const int32_t* cell_beg, const int32_t* cell_end, lq_stream stream);

/* ParticleInteraction (src/particle_interaction.h:29-41): accad, spin and
 * acestn of material particles [npb, np).  nbr_count (nullable, int32[np])
 * receives each material particle's neighbour count (j != i, q < 2).
 * artificial-pressure = 1 adds Monaghan's tensile-instability term (J. Comput.
 * Phys. 159, 2000, exponent 4) to the pair's viscous pressure:
 * (R_i + R_j) (W(r_ij) / W(dr))^4 with R = coefficient * max(-p, 0) / rho^2,
 * p = pospres.w, W(dr) = par->kernel_dr (src/loquat.cu:133-138). */
lq_status lq_particle_interaction(const lq_params* par, const lq_state* st, const lq_rates* rates,
                                  const int32_t* cell_beg, const int32_t* cell_end,
                                  int32_t* nbr_count, lq_stream stream);

/* Stale-table variants (no reference counterpart; the engine's corrector
 * half uses them instead of a second neighbour search).  cell_beg / cell_end
 * come from a grid built when the particles were elsewhere: the state st is
 * in that build's sorted order, (x_then, y_then, z_then) are the material
 * particles' FP64 positions at build time (walls do not move), and no
 * particle has moved more than disp since, 0 <= 2 disp <= 0.9 * 2h
 * (LQ_ERR_INVALID_ARG otherwise).  Results use the neighbour sets a fresh
 * grid of the current positions gives (the 27-cell stencil of each
 * particle's current cell, q < 2): the scans reach disp further and test
 * every hit that could differ against its current cell.  Only the order of
 * the FP32 sums differs from the fresh-grid call.
 * lq_adami_boundary_stale: nbr_count (nullable, int32[np]) receives each
 * wall's material neighbour count. */
lq_status lq_particle_interaction_stale(const lq_params* par, const lq_state* st, const lq_rates* rates,
                                        const double* x_then, const double* y_then, const double* z_then,
                                        float disp, const int32_t* cell_beg, const int32_t* cell_end,
                                        int32_t* nbr_count, lq_stream stream);
lq_status lq_adami_boundary_stale(const lq_params* par, const lq_state* st, float disp,
                                  const int32_t* cell_beg, const int32_t* cell_end, int32_t* nbr_count,
                                  lq_stream stream);

/* Predictor (src/predictor_corrector.cu:221-253): copies the material state
 * into ini, then advances it half a step with the constitutive update.
 * vw is read only when constitutive_model == 3 (may be NULL otherwise). */
lq_status lq_predictor(const lq_params* par, const lq_vw_params* vw, const lq_state* st,
                       const lq_state_ini* ini, const lq_rates* rates, lq_stream stream);

/* Corrector (src/predictor_corrector.cu:310-336): midpoint update from ini. */
lq_status lq_corrector(const lq_params* par, const lq_vw_params* vw, const lq_state* st,
                       const lq_state_ini* ini, const lq_rates* rates, lq_stream stream);

/* StressRegularization (src/field_smoothing.cu:124-145): Shepard filter of
 * material stress; str_cache is tensor2d[np] scratch; result lands in st->str. */
lq_status lq_field_smoothing(const lq_params* par, const lq_state* st, float* str_cache,
                             const int32_t* cell_beg, const int32_t* cell_end, lq_stream stream);

/* ----------------------------------------------------------------- engine
 * The time loop of LoquatSph::RunSimulation (src/loquat.cu:322-409) for one
 * or many independent runs packed in one segmented layout: every stage is one
 * launch over all resident runs, each step is replayed from a hipGraph, and
 * runs join and leave between steps (continuous batching).  A single run is a
 * batch of one. */

typedef struct lq_engine lq_engine;

typedef struct lq_run_input {
    lq_params par;               /* as loaded; lq_params_derive is applied by add_run */
    lq_vw_params vw;             /* model 3 only */
    lq_host_particles host;      /* reference-layout host arrays, np rows */
    const double* posd;          /* optional FP64 positions, double[3*np] xyz; NULL = widen pospres */
    int64_t nsteps;              /* steps to run; <= 0: until time >= time_max (src/loquat.cu:322) */
} lq_run_input;

/* Time step.  par.dynamic_time_step = 0 (the reference's only working mode,
 * src/loquat.cu:331-333): dt = cfl*h/cs every step.  = 1: each step's dt is
 * chosen on the device before the step as
 *     dt = cfl * min(h / (cs + max|v|), sqrt(h / max|a|))
 * over the run's material particles (a = the previous step's corrector
 * rates; no |a| term on the run's first step).  The reference parses the flag
 * but its call is commented out (src/loquat.cu:325-327) and its body is not
 * shipped, so this rule is a specification decision.  A dynamic run added
 * with nsteps <= 0 ends on the device once time >= time_max;
 * lq_engine_run_info reports the device's step count and time. */

/* Capacities bound the packed layout: particles (at most LQ_MAX_PARTICLES),
 * cells (sum over runs), runs. */
lq_status lq_engine_create(int64_t np_cap, int64_t ncell_cap, int32_t max_runs, lq_engine** out);
lq_status lq_engine_destroy(lq_engine* eng);

/* Uploads a run, derives its parameters, performs the initial neighbour
 * search and Primary sort (src/loquat.cu:297-298).  run_id is stable for the
 * run's lifetime. */
lq_status lq_engine_add_run(lq_engine* eng, const lq_run_input* in, int32_t* run_id);

/* Removes a run and compacts the layout (the other runs' results are not
 * affected: each run's arithmetic depends only on its own particles). */
lq_status lq_engine_remove_run(lq_engine* eng, int32_t run_id);

/* Advances every unfinished run by up to nsteps steps (runs stop at their own
 * nsteps).  Asynchronous on the engine stream.  A run's neighbours are
 * searched afresh (sort and cell table) every sort-interval steps of its own
 * (lq_engine_set_sort_interval) and by each Shepard filter; every other scan
 * reuses the run's last table with the reach widened by how far its
 * particles have moved since, which gives the neighbour sets of a fresh
 * search (the reference searches in both half steps, src/loquat.cu:345-366). */
lq_status lq_engine_step(lq_engine* eng, int32_t nsteps);
lq_status lq_engine_synchronize(lq_engine* eng);

/* Progress of one run as the device has it: steps done, FP64 time, finished
 * flag, and np = the number of particles the run was added with (the row
 * count lq_engine_download writes).  Synchronizes the engine stream. */
lq_status lq_engine_run_info(const lq_engine* eng, int32_t run_id, int64_t* steps_done,
                             double* time, int32_t* finished, int32_t* np);

/* Copies one run's state back in input-file order (reference layout host
 * arrays, np rows as run_info reports; posd double[3*np] optional).  Rows of
 * particles removed by lq_engine_compact get NaN positions (pospres xyz and
 * posd) and are otherwise left as they were.  Synchronous. */
lq_status lq_engine_download(lq_engine* eng, int32_t run_id, lq_host_particles* out, double* posd);

/* Particle deletion (Parameters::np_delete / ncell_delete, include/types.h:93,108,
 * and the LoquatGrid::ResizeArrayLength hook, src/loquat_grid.h:58, which the
 * reference never calls): removes from every run the material particles that
 * lay outside its cell grid when the run was last sorted (cell ncell - 1,
 * where the reference's index rule sends them and where no neighbour search
 * finds them; they are the tail of the run's segment), shrinks np and npm,
 * adds them to par.np_delete, and packs the runs again.  Particles that left
 * the grid since are removed by a call after the run's next sort.  The
 * surviving particles' states, order and sort schedule are unchanged, so
 * their results are the same as without compaction.  removed (nullable):
 * particles removed in total. */
lq_status lq_engine_compact(lq_engine* eng, int64_t* removed);

/* Collects the per-run failure flags the corrector raises on the device when
 * a particle's new state is not finite (NaN/Inf), or when the run's
 * particles moved more than 0.45 cell (0.9 h) in half a step, beyond what
 * the corrector half's reuse of the cell table covers (no run within a CFL
 * bound comes near it).  A failed run stops stepping at once; the other runs
 * are unaffected.  Synchronous; nfailed
 * (nullable) receives the number of failed resident runs. */
lq_status lq_engine_poll(lq_engine* eng, int32_t* nfailed);
lq_status lq_engine_run_failed(const lq_engine* eng, int32_t run_id, int32_t* failed);

/* Sum over resident runs: material particles, all particles, cells. */
lq_status lq_engine_totals(const lq_engine* eng, int64_t* npm, int64_t* np, int64_t* ncell, int32_t* nruns);

/* Device bytes one run of np particles and ncell cells occupies in an engine
 * (used for admission by HBM footprint). */
int64_t lq_engine_bytes_per_run(int64_t np, int64_t ncell);

/* Engine stream (hipStream_t) and options. */
lq_stream lq_engine_stream(const lq_engine* eng);
/* use_graphs: 1 = replay each step from a captured hipGraph (default), 0 = direct launches */
lq_status lq_engine_set_graphs(lq_engine* eng, int32_t use_graphs);
/* Order of particles inside a cell after each step's sort: zbin_bits = 0
 * keeps the reference's stable order (by cell only, SortParticleByCellIndex,
 * src/loquat_grid.h:43);
 * b > 0 (default 4) sorts each cell by 2^b height bins, so that every stencil
 * column is ordered by z and the neighbour scans read only the z-window of
 * the kernel support.  Cell tables and neighbour sets are the same either way;
 * only the order of the FP32 pair sums changes.  No reference counterpart. */
lq_status lq_engine_set_cell_order(lq_engine* eng, int32_t zbin_bits);

/* Order of the (x, y) cell columns in the engine's sort (no reference
 * counterpart; module-level grids keep the reference's cell id order).
 * strip_rows = 0 or 1: cell id order, x-major; T > 1: strips of T x-rows,
 * each strip's columns y-major -- (x, y) ranks as (x / T, y, x % T) -- so
 * that consecutive particles, and the workgroups an XCD runs together,
 * cover a patch T rows wide instead of a single row, and the neighbour
 * gathers of a patch share more of the XCD's L2.  The z cells of a column
 * stay consecutive, cell ids, cell tables and neighbour sets are unchanged,
 * and each particle's pairs are summed in the same order, so a run gives the
 * same bits under every column order.  Re-sorts the resident runs (mid-run
 * that re-sort, like any extra one, moves particles within their cells to
 * the z-bin order of their current positions: the FP32 sums change order).
 * 0 .. LQ_MAX_COLUMN_STRIP; default LQ_DEFAULT_COLUMN_STRIP. */
#define LQ_DEFAULT_COLUMN_STRIP 8
#define LQ_MAX_COLUMN_STRIP 64
lq_status lq_engine_set_column_order(lq_engine* eng, int32_t strip_rows);
int32_t lq_engine_get_column_order(const lq_engine* eng);
/* Picks the column order by timing, on the resident state, the sort, Adami
 * and the interaction (fresh table) for strips of 0, 2, 4, 8 and 16 rows
 * (ms_per_candidate, nullable, LQ_COLUMN_ORDER_CANDIDATES entries, in that
 * order); the order the engine had stays unless another is more than 1 %
 * faster.  Results do not depend on it.  No reference counterpart. */
#define LQ_COLUMN_ORDER_CANDIDATES 5
lq_status lq_engine_autotune_column_order(lq_engine* eng, int32_t nrep, int32_t* best_rows, double* ms_per_candidate);

/* Neighbour count of every particle of a run, in the run's input order:
 * material particles get the number of particles (walls and material, not
 * itself) the interaction pairs them with on the resident state, walls 0.
 * Sorts the resident state first (every run afresh).  counts: host, np
 * ints.  Diagnostic; the count semantics are lq_particle_interaction's. */
lq_status lq_engine_neighbour_counts(lq_engine* eng, int32_t run_id, int32_t* counts);
/* The same counts from the cell tables the runs keep between builds, as the
 * next step's predictor half scans them (no sort; reach widened by each
 * run's recorded displacement since its table was built): they equal the
 * fresh table's at any point of the sort interval.  Overwrites the engine's
 * last interaction rates (the next step recomputes them).  Diagnostic. */
lq_status lq_engine_neighbour_counts_stale(lq_engine* eng, int32_t run_id, int32_t* counts);

/* Sort interval (engine setting; no counterpart in the reference, which
 * sorts in both half steps of every step, src/loquat.cu:345-366).  A run's
 * particles are sorted and its cell table rebuilt at the start of every
 * `steps`-th step of its own and by each Shepard filter; in between they keep
 * their order and table, and every scan widens its reach by how far the
 * run's particles have moved since the table was built (tracked exactly from
 * each particle's position at the build), so that every neighbour set is
 * that of a fresh table.  The schedule counts the run's own steps, so a run
 * gives the same bits alone and in any batch.  A run whose particles move
 * more than 0.45 cell between builds is stopped as failed.  Only the order of
 * the FP32 pair sums differs between intervals.  1 .. LQ_MAX_SORT_INTERVAL;
 * default LQ_DEFAULT_SORT_INTERVAL. */
#define LQ_DEFAULT_SORT_INTERVAL 12
#define LQ_MAX_SORT_INTERVAL 16
lq_status lq_engine_set_sort_interval(lq_engine* eng, int32_t steps);
int32_t lq_engine_get_sort_interval(const lq_engine* eng);

/* Mean device time per step of each stage over the last lq_engine_profile
 * call: grid (hash+sort+table+permute), adami, interaction, integrate,
 * shepard, in milliseconds.  Runs nsteps steps with events between stages
 * (no graphs). */
lq_status lq_engine_profile(lq_engine* eng, int32_t nsteps, double* ms_stage5);

/* Autotuner: times nrep launches of each interaction variant on the engine's
 * resident state, on the fresh table of a sort and on the kept table (with
 * each run's recorded displacement bound), weights them as a step does (one
 * fresh launch in 2K for sort interval K), and pins the fastest for this
 * engine; the engine's variant stays unless another is more than 1 % faster.
 * All variants produce the same neighbour sets; their sum orders differ only
 * as stated below.  ms_per_variant (nullable, LQ_INTERACTION_VARIANTS
 * entries) receives each variant's weighted mean launch time. */
#define LQ_INTERACTION_VARIANTS 20
#define LQ_DEFAULT_INTERACTION_VARIANT 19
lq_status lq_engine_autotune(lq_engine* eng, int32_t nrep, int32_t* best, double* ms_per_variant);

/* Picks the in-cell order (lq_engine_set_cell_order: 2^b height bins per
 * cell, b = 0 .. LQ_CELL_ORDER_CANDIDATES - 1; the 2h cell itself is fixed
 * by the reference's cell hash) by timing, on the resident state, what the
 * order changes: the sort and permute, Adami and the interaction with the
 * engine's variant.  ms_per_order (nullable, LQ_CELL_ORDER_CANDIDATES
 * entries) receives the mean time of those four per candidate.  Leaves the
 * engine in the fastest order (the order it had, unless another is more than
 * 1 % faster) with the resident runs sorted by it (results change only in
 * the order of FP32 sums).  No reference counterpart. */
#define LQ_CELL_ORDER_CANDIDATES 6
lq_status lq_engine_autotune_cell_order(lq_engine* eng, int32_t nrep, int32_t* best_bits, double* ms_per_order);

/* Interaction kernel variant.  lq_set_interaction_variant sets it for the
 * module-level calls (lq_particle_interaction, lq_adami_boundary, ...) made by
 * the calling thread, and as the initial variant of engines that thread
 * creates afterwards; each engine then keeps its own
 * (lq_engine_set_interaction_variant, lq_engine_autotune), so handles on
 * different threads or streams do not share it.
 * 0 = per-thread direct loop, 1 = LDS-compacted neighbour lists,
 * 2 = LDS lists with a 4-wide unrolled candidate scan,
 * 3 = variant 2 with offset loads, pruned column scans, XCD-contiguous
 *     blocks,
 * 4 = variant 3 evaluating the lists two pairs at a time,
 * 5, 6 = variants 3, 4 where a lane past the end of its list loads nothing
 *        (exec-masked gathers) instead of re-reading its own particle,
 * 7 - 11 = variant 6's code in other shapes, for the autotuner: 7 with
 *        32-entry neighbour lists, 8 with 48-entry lists (three workgroups
 *        per CU instead of four), 9 testing eight candidates per load round,
 *        10 in 128-thread workgroups, 11 in 64-thread workgroups (one wave
 *        per workgroup, so a CU holds as many as its registers allow and a
 *        finished wave frees its slot at once),
 * 12 - 15 = variant 11 with partial flushes (a full list evaluates only its
 *        first PF entries and keeps the rest) and the stencil columns in an
 *        order that keeps the lanes' list lengths alike (opposite columns in
 *        pairs): 12 PF = 12 with opposite columns pairwise, own column last;
 *        13 PF = 16 with the edge columns first, then the corners, own
 *        column last; 14 as 13 with PF = 8; 15 as 13 with full flushes,
 * 16 = 13 with two lanes per particle (32 particles per wave, the lane pair
 *        taking a column's candidates alternately, the two partial sums
 *        added at the end),
 * 17 = 13 with 32-entry lists and the candidates of each stencil column
 *        staged in LDS by the wave (chunks of 128 consecutive particles from
 *        its lanes' lowest pending candidate, loaded coalesced; chunk bounds
 *        by wavefront-shuffle minimum / maximum) and tested from there,
 * 18, 19 = 13 and 17 with unmasked gathers (a lane past the end of its list
 *        re-reads its own particle, whose pair terms are exactly zero,
 *        instead of substituting it from registers); 19 stages only
 *        launches of at least 524,288 material particles and runs as 18
 *        below (the same bits either way: small launches are latency-bound,
 *        where staging costs); 19 is the default.
 *        Variants 3 - 11 sum each particle's pairs in one order, 13 - 15 and
 *        17 - 19 in another, 12 and 16 in their own (all within FP32
 *        rounding).
 * Variants 3 and up also select the list-based Adami and Shepard scans; the
 * scans of a kept (stale) cell table and the artificial-pressure kernel use
 * variant 6's shape (in the workgroup size of variants 10 to 19, with their
 * flushes, column orders, lanes per particle, staging and gathers). */
lq_status lq_set_interaction_variant(int32_t variant);
int32_t lq_get_interaction_variant(void);
lq_status lq_engine_set_interaction_variant(lq_engine* eng, int32_t variant);
int32_t lq_engine_get_interaction_variant(const lq_engine* eng);

/* ----------------------------------------------------------------- server
 * Continuous batching of independent runs (no counterpart in the reference,
 * which runs one simulation per process, src/loquat.cu:216-475).  Requests
 * queue in arrival order; each tick admits the queue head while its device
 * footprint fits the HBM budget and the capacities (strict FIFO), advances all
 * resident runs steps_per_tick steps in shared launches, then evicts finished
 * and failed runs.  Each run's result is bit-identical to running it alone. */

typedef struct lq_server lq_server;

typedef struct lq_server_config {
    int64_t np_cap;           /* engine particle capacity (sum over resident runs) */
    int64_t ncell_cap;        /* engine cell capacity */
    int32_t max_runs;         /* resident runs at most */
    int64_t hbm_budget;       /* bytes resident runs may occupy (lq_engine_bytes_per_run) */
    int32_t steps_per_tick;   /* steps between admission/eviction points */
    int32_t fake_backend;     /* 1: no GPU, runs advance on the host (scheduler tests) */
    int32_t keep_results;     /* 1: download each finished run's state */
} lq_server_config;

typedef enum lq_request_state {
    LQ_REQ_QUEUED = 0,
    LQ_REQ_RUNNING = 1,
    LQ_REQ_DONE = 2,
    LQ_REQ_FAILED = 3,        /* non-finite state: evicted, others unaffected */
    LQ_REQ_REJECTED = 4       /* larger than the budget or the capacities, or failing the
                                 engine's checks (lq_engine_add_run's errors) */
} lq_request_state;

lq_status lq_server_create(const lq_server_config* cfg, lq_server** out);
lq_status lq_server_destroy(lq_server* srv);
/* Queues a run (host arrays are copied).  The checks lq_engine_add_run makes
 * (derived parameters, constitutive model and its constants, walls inside the
 * grid) are made here, so an invalid request is rejected at once instead of
 * blocking the queue.  request_id is set even when the request is rejected
 * (state LQ_REQ_REJECTED). */
lq_status lq_server_submit(lq_server* srv, const lq_run_input* in, int32_t* request_id);
/* One admission / step / eviction round.  finished_ids receives up to cap ids
 * of requests that completed, failed or were rejected at admission in this
 * tick; nfinished the count.  A request the engine refuses at admission is
 * rejected and dropped from the queue (the requests behind it go on); a
 * device error there is returned after the request is dropped. */
lq_status lq_server_tick(lq_server* srv, int32_t* finished_ids, int32_t cap, int32_t* nfinished);
lq_status lq_server_request_info(const lq_server* srv, int32_t request_id, int32_t* state, int64_t* steps_done,
                                 int64_t* nsteps, int64_t* bytes);
/* Final state of a finished request in input order (keep_results = 1). */
lq_status lq_server_result(lq_server* srv, int32_t request_id, lq_host_particles* out, double* posd);
lq_status lq_server_stats(const lq_server* srv, int32_t* queued, int32_t* running, int64_t* bytes_in_use,
                          int64_t* npm_running);

/* Build/version string. */
const char* lq_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LOQUAT_AMD_H */
