// lq_common.hpp -- device-side types and helpers shared by the HIP kernels.
//
// Layout: every resident run owns one contiguous particle segment
// [p_off, p_off + np) of the packed SoA arrays -- walls first, then material
// particles, as the reference's index convention requires (walls are [0, npb),
// src/adami_boundary.cu:73, src/predictor_corrector.cu:182-184) -- and one
// contiguous range of cell ids [cell_base, cell_base + ncell).  A single run
// is a batch of one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/loquat_amd.h"

namespace lq {

// Symmetric second-order tensor, stored xx yy zz xy xz yz (include/types.h:36-38).
struct alignas(8) tensor2d {
    float xx, yy, zz, xy, xz, yz;
};
static_assert(sizeof(tensor2d) == 24, "tensor2d must stay 24 bytes");

// Per-run constants and counters, device resident.  par is the reference's
// Parameters (already derived); the rest describes the run's segment.
struct RunDesc {
    lq_params par;        // 208 B
    int32_t p_off;        // first particle of the run in the packed arrays
    int32_t np, npb, npm; // particles, walls, material
    int32_t cell_base;    // first global cell id of the run
    float r2_cut;         // neighbour threshold on FP32 squared distance
    int32_t active;       // 1 while the run still steps
    int32_t shepard_due;  // this step applies the stress Shepard filter
    int32_t step;         // steps completed
    int32_t nsteps;       // steps requested
    int32_t zbits;        // sort key = cell << zbits | z-bin in cell (0: cell only)
    // Engine sort order of the (x, y) cell columns: 0 = cell id order;
    // T > 1 = strips of T x-rows, each strip's columns y-major (x, y) ->
    // (x / T, y, x % T), so that consecutive particles cover a T-row patch
    // of the grid (cell_rank / cell_of_rank).  Cell ids, the cell table and
    // every neighbour set are unchanged.
    int32_t xtile;
    int32_t until_time;   // 1: the run ends when time >= time_max (dynamic dt), else after nsteps
    // The run's cell table is rebuilt every sort-interval steps of its own
    // (engine setting) and after a Shepard filter; in between its particles
    // keep the sort keys of the last build and its scans read that table.
    int32_t since_sort;   // steps completed since the run's table was built
    int32_t sort_due;     // this step's grid build gives the run fresh keys (k_step_begin)
    float dtab;           // bound on how far a particle of the run has moved since the table was built (last step end)
    int32_t out_tab;      // some particle of the run was outside the grid at the last step end
    double time;          // FP64 simulated time (src/loquat.cu:305,455)
    lq_vw_params vw;      // constitutive model 3 constants
    float vw_a, vw_fb0;   // derived model-3 constants (a, and the f_b prefactor)
};

// Descriptor access for one run passed by value (module-level calls).
// `padded`: the particle arrays carry SCAN_PAD spare entries past the last
// particle, so a neighbour scan reading blocks of up to SCAN_PAD candidates
// may read past the end of a range without clamping (engine buffers; module
// calls get caller arrays).
constexpr int SCAN_PAD = 8;

struct DescOne {
    static constexpr bool padded = false;
    RunDesc d;
    __device__ __forceinline__ int run(int) const { return 0; }
    __device__ __forceinline__ const RunDesc& get(int) const { return d; }
};

// Descriptor access through a device table and a per-particle run id.
struct DescTable {
    static constexpr bool padded = true;
    const RunDesc* d;
    const int32_t* run_of;
    __device__ __forceinline__ int run(int k) const { return run_of[k]; }
    __device__ __forceinline__ const RunDesc& get(int r) const { return d[r]; }
};

// Maps a compact thread index t (over the material particles, or over the
// walls, of all resident runs) to a particle index: idx[t] when a table is
// given, else base + t (one run).  Launching over exactly the particles a
// stage works on keeps every XCD's slice of blocks equally loaded.
struct IndexMap {
    const int32_t* idx;
    int32_t base;
    __device__ __forceinline__ int operator()(int t) const { return idx ? idx[t] : base + t; }
};

// A cell table built before the particles last moved ("stale"): the
// positions of the particles being updated when it was built -- FP32 (po) or
// FP64 (xo, yo, zo) -- and a bound on how far any particle of a run has
// moved since: per run as float bits (dmax[run * DMAX_SLOTS]), or one value
// (disp) when dmax is null.  nout (nullable, per run): nonzero when some
// particle of the run lies outside the grid now (null: assume so).
// from_desc: the bound and the outside flag are the run descriptor's dtab
// and out_tab (a table from an earlier step); 1: a run whose table was
// rebuilt this step (sort_due) has bound 0.
struct StaleIn {
    const double* xo;
    const double* yo;
    const double* zo;
    const uint32_t* dmax;
    float disp;
    const uint32_t* nout;
    const float4* po;
    int from_desc;
};

// Which runs a grid build gives fresh keys: all, those with sort_due or
// shepard_due set, or none (every particle keeps the key of the last build).
enum GridRekey { REKEY_ALL = 0, REKEY_SORT_DUE = 1, REKEY_SHEPARD_DUE = 2, REKEY_NONE = 3 };

// Per-run displacement maxima are accumulated in DMAX_SLOTS slots per run
// (workgroup b uses slot b % DMAX_SLOTS, so that atomics do not pile up on
// one address) and folded into slot 0 before they are read.
constexpr int DMAX_SLOTS = 64;

// Largest displacement a stale table tolerates, as a fraction of the cell
// width: 2 disp <= STALE_LIMIT * (2 h).  A run that moves further in half a
// step has violated any CFL bound and is stopped as diverging.
constexpr float STALE_LIMIT = 0.9f;

struct DevState {
    double* xd;
    double* yd;
    double* zd;
    float4* pospres;
    float4* velrhop;
    float4* idep;
    tensor2d* str;
    tensor2d* stn;
    int32_t* orig;
    float4* xo;  // engine: each particle's FP32 position when its run's table was built (null in module calls)
};

struct DevIni {
    double* xd;
    double* yd;
    double* zd;
    float4* velrhop;
    float2* epvr;
    tensor2d* str;
    tensor2d* stn;
};

struct DevRates {
    float4* accad;
    float4* spin;
    tensor2d* acestn;
};

inline DevState to_dev(const lq_state& s) {
    return DevState{s.xd, s.yd, s.zd, (float4*)s.pospres, (float4*)s.velrhop, (float4*)s.idep,
                    (tensor2d*)s.str, (tensor2d*)s.stn, s.orig};
}
inline DevIni to_dev(const lq_state_ini& s) {
    return DevIni{s.xd, s.yd, s.zd, (float4*)s.velrhop, (float2*)s.epvr, (tensor2d*)s.str, (tensor2d*)s.stn};
}
inline DevRates to_dev(const lq_rates& r) {
    return DevRates{(float4*)r.accad, (float4*)r.spin, (tensor2d*)r.acestn};
}

// Cell coordinate of one axis: floorf(0.5f * x * i_h) evaluated left to right
// in FP32 exactly as the lookups of src/adami_boundary.cu:80-82 do.  The
// result is clamped to [-2, 2^30] (NaN -> -2) so that out-of-domain particles
// give a well-defined, empty or partial stencil instead of an undefined
// float->int conversion.
__device__ __forceinline__ int cell_coord(float x, float i_h) {
#pragma clang fp contract(off)
    float f = floorf((0.5f * x) * i_h);
    if (!(f >= -2.0f)) return -2;
    if (f > 1073741824.0f) return 1073741824;
    return (int)f;
}

// z-bin of a particle inside its cell: floor(s * 2^zbits) - cz * 2^zbits with
// s = (0.5 z) / h the cell-unit coordinate and cz = cell_coord(z).  Scaling by
// a power of two is exact, so the global bin (cz << zbits) + bin is
// non-decreasing in z: sorting a cell by it sorts every column of cells by z
// to within one bin (cell height / 2^zbits).
__device__ __forceinline__ uint32_t zbin_in_cell(float z, float i_h, int zbits, int cz) {
#pragma clang fp contract(off)
    const float nb = (float)(1 << zbits);
    const float f = floorf(((0.5f * z) * i_h) * nb) - (float)cz * nb;
    return (uint32_t)fminf(fmaxf(f, 0.0f), nb - 1.0f);
}

// FP32 squared distance with a fixed operation order: dx * dx rounded, then
// two fused multiply-adds (the same rounding on every implementation); the
// neighbour predicate is r2 < RunDesc::r2_cut.
__device__ __forceinline__ float r2_fp32(float dx, float dy, float dz) {
#pragma clang fp contract(off)
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Calls f(jb, je) for the contiguous particle range of each (x, y) column of
// the 3x3x3 stencil around (cix, ciy, ciz), for cell layers [layer_lo,
// layer_hi].  Cells of one column are consecutive ids (cell = (cx*ny + cy)*nz
// + cz, src/adami_boundary.cu:98), so the up-to-three z cells of a column map
// to one range of the sorted particles.  Cells outside the grid are skipped
// as in src/adami_boundary.cu:90-96.
template <class F>
__device__ __forceinline__ void for_each_column(const RunDesc& d, int cix, int ciy, int ciz,
                                                int layer_lo, int layer_hi, const int32_t* __restrict__ beg,
                                                const int32_t* __restrict__ end, F&& f) {
    const int nx = d.par.cell_num_x, ny = d.par.cell_num_y, nz = d.par.cell_num_z;
    const int czlo = max(ciz - 1, 0);
    const int czhi = min(ciz + 1, nz - 1);
    if (czlo > czhi) return;
    for (int layer = layer_lo; layer <= layer_hi; ++layer) {
        const int lbase = d.cell_base + layer * d.par.ncell_one_layer;
        for (int ox = -1; ox <= 1; ++ox) {
            const int cx = cix + ox;
            if (cx < 0 || cx >= nx) continue;
            for (int oy = -1; oy <= 1; ++oy) {
                const int cy = ciy + oy;
                if (cy < 0 || cy >= ny) continue;
                const int c0 = lbase + (cx * ny + cy) * nz;
                int jb = -1, je = -1;
                for (int cz = czlo; cz <= czhi; ++cz) {
                    const int b = beg[c0 + cz];
                    if (b != -1) {
                        if (jb < 0) jb = b;
                        je = end[c0 + cz];
                    }
                }
                if (jb >= 0) f(jb, je);
            }
        }
    }
}

}  // namespace lq
