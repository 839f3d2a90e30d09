// lq_grid.hip -- neighbour grid for gfx950: cell hash, stable LSD radix sort,
// cell begin/end table and the state permute.
//
// Replaces LoquatGrid (src/loquat_grid.h:28-66), whose body ships only as a
// prebuilt library, and the cub::DeviceRadixSort::SortPairs call it is built
// around (stable ascending sort of (cell, index) pairs,
// include/cub/device/device_radix_sort.cuh:147-175).  The result is the unique
// stable order, so it matches any stable sort of the same keys bit for bit.
//
// Radix sort, per 8-bit digit pass (ceil(bits/8) passes, bits = ceil(log2
// ncell) + the z-bin bits of the engine's key):
//   hash    : the first pass's count is fused with the key computation and
//             with clearing the cell table (k_hash_count).
//   count   : one workgroup per 2048-key tile; each wave ranks 64 keys per
//             round with 8 ballots (wave-wide match of the digit), the lowest
//             lane of each peer group bumps the wave's LDS counter -- no LDS
//             atomics, so runs of equal digits (sorted particles share cells)
//             cost nothing extra.
//   scan    : one workgroup per digit scans that digit's tile counts.
//   scatter : re-ranks the tile the same way, stages keys/values in LDS in
//             digit order and writes each digit's run contiguously.
#include <mutex>

#include "lq_internal.hpp"

namespace lq {

namespace {

__device__ __forceinline__ unsigned long long match8(uint32_t d, bool valid) {
    const unsigned long long vm = __ballot(valid);
    unsigned long long peers = vm;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers & vm;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan over the 256 threads of a workgroup; total of all values out.
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t* sm4, uint32_t& total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) sm4[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t x = sm4[i];
        if (i < w) pre += x;
        tot += x;
    }
    total = tot;
    __syncthreads();
    return pre + inc - v;
}

constexpr int WAVE_CHUNK = RS_TILE / (RS_THREADS / 64);  // keys per wave per tile

__global__ __launch_bounds__(RS_THREADS) void k_radix_count(const uint32_t* __restrict__ keys, int n, int shift,
                                                             uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t cnt[4][RS_BINS];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < 4 * RS_BINS; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE + w * WAVE_CHUNK;
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll 4
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & 255u : 0u;
        const unsigned long long peers = match8(d, valid);
        if (valid && (peers & lt) == 0ull) cnt[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    hist[(size_t)t * ntiles + blockIdx.x] = cnt[0][t] + cnt[1][t] + cnt[2][t] + cnt[3][t];
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scan(uint32_t* __restrict__ hist, int ntiles,
                                                            uint32_t* __restrict__ totals) {
    __shared__ uint32_t sm[4];
    uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t running = 0;
    for (int c = 0; c < ntiles; c += RS_THREADS) {
        const int i = c + threadIdx.x;
        const uint32_t v = i < ntiles ? row[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan_256(v, sm, tot);
        if (i < ntiles) row[i] = ex + running;
        running += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = running;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin,
                                                               uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout, int n, int shift,
                                                               const uint32_t* __restrict__ hist,
                                                               const uint32_t* __restrict__ totals, int ntiles) {
    __shared__ uint32_t cnt[4][RS_BINS];
    __shared__ uint32_t lbase[RS_BINS];
    __shared__ uint32_t goff[RS_BINS];
    __shared__ uint32_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t sm[4];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < 4 * RS_BINS; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE + w * WAVE_CHUNK;
    const unsigned long long lt = (1ull << lane) - 1ull;
    uint32_t key[RS_IPT], val[RS_IPT], rank[RS_IPT];
#pragma unroll
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * 64 + lane;
        const bool valid = idx < n;
        key[r] = valid ? kin[idx] : 0u;
        val[r] = valid ? (vin ? vin[idx] : (uint32_t)idx) : 0u;
        const uint32_t d = (key[r] >> shift) & 255u;
        const unsigned long long peers = match8(d, valid);
        rank[r] = 0;
        if (valid) {
            const uint32_t b = cnt[w][d];
            const unsigned long long below = peers & lt;
            rank[r] = b + (uint32_t)__popcll(below);
            if (below == 0ull) cnt[w][d] = b + (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t c = cnt[i][t];
        cnt[i][t] = run;
        run += c;
    }
    uint32_t tot_local, tot_global;
    const uint32_t ex = block_excl_scan_256(run, sm, tot_local);
    lbase[t] = ex;
    const uint32_t gex = block_excl_scan_256(totals[t], sm, tot_global);
    goff[t] = gex + hist[(size_t)t * ntiles + blockIdx.x] - ex;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[r] >> shift) & 255u;
            const uint32_t pos = lbase[d] + cnt[w][d] + rank[r];
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
    }
    __syncthreads();
    const int nvalid = min(RS_TILE, n - (int)blockIdx.x * RS_TILE);
    for (int p = t; p < nvalid; p += RS_THREADS) {
        const uint32_t k = sk[p];
        const uint32_t g = goff[(k >> shift) & 255u] + (uint32_t)p;
        kout[g] = k;
        vout[g] = sv[p];
    }
}

// Position of (x, y) column (cx, cy) in the run's sort order (RunDesc::xtile):
// strips of T x-rows, each strip's columns y-major.  A bijection of
// [0, nx ny); the last strip holds the nx mod T remaining rows.
__device__ __forceinline__ int column_rank(const RunDesc& d, int cx, int cy) {
    const int ny = d.par.cell_num_y, T = d.xtile;
    if (T <= 1) return cx * ny + cy;
    const int tx = cx / T, te = min(T, d.par.cell_num_x - tx * T);
    return tx * T * ny + cy * te + (cx - tx * T);
}

// Sort key of particle k: its cell's rank (layer 0 for walls, 1 for
// material, index rule of src/adami_boundary.cu:73 and src/loquat.cu:119-124;
// any axis outside the grid sends the particle to ncell_delete = ncell - 1),
// shifted left by zbits with the particle's z-bin inside the cell below it.
// The rank is the cell id when xtile <= 1 (the reference's order); the z
// cells of a column stay consecutive either way.
template <class DP>
__device__ __forceinline__ uint32_t cell_key(const DP& dp, int k, float4 p) {
    const RunDesc& d = dp.get(dp.run(k));
    const int layer = (k - d.p_off) < d.npb ? 0 : 1;
    const float i_h = d.par.i_h;
    const int cx = cell_coord(p.x, i_h), cy = cell_coord(p.y, i_h), cz = cell_coord(p.z, i_h);
    const int nx = d.par.cell_num_x, ny = d.par.cell_num_y, nz = d.par.cell_num_z;
    if (cx < 0 || cx >= nx || cy < 0 || cy >= ny || cz < 0 || cz >= nz)
        return (uint32_t)(d.cell_base + d.par.ncell - 1) << d.zbits;
    return ((uint32_t)(d.cell_base + layer * d.par.ncell_one_layer + column_rank(d, cx, cy) * nz + cz) << d.zbits) |
           (d.zbits ? zbin_in_cell(p.z, i_h, d.zbits, cz) : 0u);
}

// Cell id of a global cell rank (inverse of cell_key's rank).
__device__ __forceinline__ uint32_t cell_of_rank(const RunDesc& d, uint32_t rank) {
    const int T = d.xtile;
    const int local = (int)rank - d.cell_base;
    const int nol = d.par.ncell_one_layer, nz = d.par.cell_num_z, ny = d.par.cell_num_y;
    if (T <= 1 || local >= 2 * nol) return rank;  // cell id order, or the outside cell
    const int layer = local / nol, r = local - layer * nol;
    const int col = r / nz, cz = r - col * nz;
    const int tx = col / (T * ny), rem = col - tx * T * ny;
    const int te = min(T, d.par.cell_num_x - tx * T);
    const int cy = rem / te, cx = tx * T + (rem - cy * te);
    return (uint32_t)(d.cell_base + layer * nol + (cx * ny + cy) * nz + cz);
}

// Whether a build re-keys particle k's run (`which`, see GridRekey); the
// others keep the key of the last build, kept[k].
template <class DP>
__device__ __forceinline__ bool rekeyed(const DP& dp, int k, int which) {
    if (which == REKEY_ALL) return true;
    if (which == REKEY_NONE) return false;
    const RunDesc& d = dp.get(dp.run(k));
    return which == REKEY_SORT_DUE ? d.sort_due != 0 : d.shepard_due != 0;
}

// Hash fused with the first radix pass's count: one workgroup per tile
// computes the keys of its tile (stored for the scatter), counts their low
// digit as k_radix_count does, and clears its share of the cell table
// (cell_beg = cell_end = -1; the occupied cells are written after the sort).
// Particles of runs that `which` does not re-key take their key from kept
// (the sorted keys of the last build, in the current particle order): their
// stable order and their cells are then those of the last build.
template <class DP>
__global__ __launch_bounds__(RS_THREADS) void k_hash_count(DP dp, const float4* __restrict__ pospres, int n,
                                                            uint32_t* __restrict__ keys, uint32_t* __restrict__ hist,
                                                            int ntiles, int32_t* __restrict__ cbeg,
                                                            int32_t* __restrict__ cend, int64_t ncell,
                                                            const uint32_t* kept, int which) {
    __shared__ uint32_t cnt[4][RS_BINS];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < 4 * RS_BINS; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    {
        const int64_t per = (ncell + ntiles - 1) / ntiles;
        const int64_t c0 = (int64_t)blockIdx.x * per, c1 = min(c0 + per, ncell);
        for (int64_t c = c0 + t; c < c1; c += RS_THREADS) {
            cbeg[c] = -1;
            cend[c] = -1;
        }
    }
    __syncthreads();
    const int base = blockIdx.x * RS_TILE + w * WAVE_CHUNK;
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll 4
    for (int r = 0; r < RS_IPT; ++r) {
        const int idx = base + r * 64 + lane;
        const bool valid = idx < n;
        uint32_t key = 0u;
        if (valid) {
            // kept may alias keys (the same slot): each thread reads its own
            // entry before writing it
            key = (kept && !rekeyed(dp, idx, which)) ? kept[idx] : cell_key(dp, idx, pospres[idx]);
            keys[idx] = key;
        }
        const uint32_t d = key & 255u;
        const unsigned long long peers = match8(d, valid);
        if (valid && (peers & lt) == 0ull) cnt[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    hist[(size_t)t * ntiles + blockIdx.x] = cnt[0][t] + cnt[1][t] + cnt[2][t] + cnt[3][t];
}

__global__ void k_fill2(int32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = v;
        b[i] = v;
    }
}

// Begin/end of every occupied cell; the cell's rank is the key without its
// low zbits (the z-bin inside the cell), its id that of the rank in the
// run's order.
template <class DP>
__global__ void k_cell_table(DP dp, const uint32_t* __restrict__ keys, int n, int zbits, int32_t* __restrict__ beg,
                             int32_t* __restrict__ end) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t rank = keys[k] >> zbits;
    const bool first = k == 0 || (keys[k - 1] >> zbits) != rank;
    const bool last = k == n - 1 || (keys[k + 1] >> zbits) != rank;
    if (!first && !last) return;
    const uint32_t cell = cell_of_rank(dp.get(dp.run(k)), rank);
    if (first) beg[cell] = k;
    if (last) end[cell] = k + 1;
}

// With table positions (engine states): particles of re-keyed runs record
// the position their keys came from; the others keep theirs.
// half (engine, nullable arrays): the walls' positions, FP32 mirror,
// velocity and stress are also written there, in the new order (walls do not
// move between builds, so the half-step state of split integration keeps
// them from one build to the next).
__global__ __launch_bounds__(256) void k_permute(const uint32_t* __restrict__ perm, int n, DevState src, DevState dst,
                                                 int with_ini, DevIni isrc, DevIni idst, tensor2d* str_dst,
                                                 DescTable dp, int which, DevState half) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = (int)perm[k];
    if (dst.xo) dst.xo[k] = rekeyed(dp, s, which) ? src.pospres[s] : src.xo[s];
    dst.xd[k] = src.xd[s];
    dst.yd[k] = src.yd[s];
    dst.zd[k] = src.zd[s];
    dst.pospres[k] = src.pospres[s];
    dst.velrhop[k] = src.velrhop[s];
    dst.idep[k] = src.idep[s];
    (str_dst ? str_dst : dst.str)[k] = src.str[s];
    dst.stn[k] = src.stn[s];
    dst.orig[k] = src.orig[s];
    if (half.xd) {
        const RunDesc& d = dp.get(dp.run(k));
        if (k - d.p_off < d.npb) {
            half.xd[k] = src.xd[s];
            half.yd[k] = src.yd[s];
            half.zd[k] = src.zd[s];
            half.pospres[k] = src.pospres[s];
            half.velrhop[k] = src.velrhop[s];
            half.str[k] = src.str[s];
        }
    }
    if (with_ini) {
        idst.xd[k] = isrc.xd[s];
        idst.yd[k] = isrc.yd[s];
        idst.zd[k] = isrc.zd[s];
        idst.velrhop[k] = isrc.velrhop[s];
        idst.epvr[k] = isrc.epvr[s];
        idst.str[k] = isrc.str[s];
        idst.stn[k] = isrc.stn[s];
    }
}

inline int blocks_for(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

lq_status alloc_grid(GridBuffers& g, int64_t np_cap, int64_t ncell_cap) {
    g = GridBuffers{};
    g.np_cap = np_cap;
    g.ncell_cap = ncell_cap;
    const int64_t np_alloc = np_cap > 0 ? np_cap : 1;
    const int64_t nc_alloc = ncell_cap > 0 ? ncell_cap : 1;
    const int64_t tiles = rs_tiles((int)np_alloc);
    void* p = nullptr;
    for (int i = 0; i < 2; ++i) {
        if (hipMalloc(&p, np_alloc * 4) != hipSuccess) goto oom;
        g.keys[i] = (uint32_t*)p;
        if (hipMalloc(&p, np_alloc * 4) != hipSuccess) goto oom;
        g.vals[i] = (uint32_t*)p;
    }
    if (hipMalloc(&p, RS_BINS * tiles * 4) != hipSuccess) goto oom;
    g.hist = (uint32_t*)p;
    if (hipMalloc(&p, RS_BINS * 4) != hipSuccess) goto oom;
    g.totals = (uint32_t*)p;
    if (hipMalloc(&p, nc_alloc * 4) != hipSuccess) goto oom;
    g.cell_beg = (int32_t*)p;
    if (hipMalloc(&p, nc_alloc * 4) != hipSuccess) goto oom;
    g.cell_end = (int32_t*)p;
    g.sorted_slot = -1;
    return LQ_OK;
oom:
    free_grid(g);
    set_error("hipMalloc failed for the neighbour grid");
    return LQ_ERR_OUT_OF_MEMORY;
}

void free_grid(GridBuffers& g) {
    for (int i = 0; i < 2; ++i) {
        if (g.keys[i]) (void)hipFree(g.keys[i]);
        if (g.vals[i]) (void)hipFree(g.vals[i]);
        g.keys[i] = g.vals[i] = nullptr;
    }
    if (g.hist) (void)hipFree(g.hist);
    if (g.totals) (void)hipFree(g.totals);
    if (g.cell_beg) (void)hipFree(g.cell_beg);
    if (g.cell_end) (void)hipFree(g.cell_end);
    g.hist = g.totals = nullptr;
    g.cell_beg = g.cell_end = nullptr;
}

// Radix passes after the fused first count; pass 0 reads hist from
// k_hash_count.
static void radix_passes(GridBuffers& g, int n, int bits, hipStream_t s) {
    const int ntiles = rs_tiles(n);
    const int passes = (bits + 7) / 8;
    int src = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (p > 0) k_radix_count<<<ntiles, RS_THREADS, 0, s>>>(g.keys[src], n, shift, g.hist, ntiles);
        k_radix_scan<<<RS_BINS, RS_THREADS, 0, s>>>(g.hist, ntiles, g.totals);
        k_radix_scatter<<<ntiles, RS_THREADS, 0, s>>>(g.keys[src], p == 0 ? nullptr : g.vals[src],
                                                       g.keys[1 - src], g.vals[1 - src], n, shift, g.hist,
                                                       g.totals, ntiles);
        src = 1 - src;
    }
    g.sorted_slot = src;
}

template <class DP>
static void grid_build(const DP& dp, const float4* pospres, int n, int bits, int64_t ncell_total, int zbits,
                       GridBuffers& g, hipStream_t s, int which = REKEY_ALL) {
    if (n <= 0) {
        g.sorted_slot = 0;
        const int fill_blocks = (int)std::min<int64_t>(blocks_for(ncell_total, 256), 4096);
        if (ncell_total > 0) k_fill2<<<fill_blocks, 256, 0, s>>>(g.cell_beg, g.cell_end, ncell_total, -1);
        return;
    }
    const int ntiles = rs_tiles(n);
    const uint32_t* kept = which != REKEY_ALL && g.sorted_slot >= 0 ? g.keys[g.sorted_slot] : nullptr;
    k_hash_count<DP><<<ntiles, RS_THREADS, 0, s>>>(dp, pospres, n, g.keys[0], g.hist, ntiles, g.cell_beg,
                                                   g.cell_end, ncell_total, kept, which);
    radix_passes(g, n, bits, s);
    k_cell_table<DP><<<blocks_for(n, 256), 256, 0, s>>>(dp, g.keys[g.sorted_slot], n, zbits, g.cell_beg, g.cell_end);
}

void launch_grid_build_one(const RunDesc& d, const float4* pospres, GridBuffers& g, hipStream_t s) {
    grid_build(DescOne{d}, pospres, d.np, rs_bits(d.par.ncell) + d.zbits, d.par.ncell, d.zbits, g, s);
}

void launch_grid_build_table(const RunDesc* d, const int32_t* run_of, const float4* pospres, int n,
                             int64_t ncell_total, int zbits, GridBuffers& g, hipStream_t s, int which) {
    grid_build(DescTable{d, run_of}, pospres, n, rs_bits(ncell_total) + zbits, ncell_total, zbits, g, s, which);
}

// dst[i] = src[i] + shift for i < n (a run's kept keys moved to its new
// segment and cell range).
__global__ void k_shift_keys(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int n, uint32_t shift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + shift;
}

void launch_shift_keys(const uint32_t* src, uint32_t* dst, int n, uint32_t shift, hipStream_t s) {
    if (n > 0) k_shift_keys<<<blocks_for(n, 256), 256, 0, s>>>(src, dst, n, shift);
}

// ------------------------------------------------- grid build without allocation
// Scratch of lq_grid_build_noalloc, a file-scope device array: the radix
// sort's second key/value buffer (the caller's part_cel / part_idx are the
// first) and its tile histograms and digit totals, for up to
// NOALLOC_MAX_PARTICLES particles.  Every GPU the library runs on holds its
// own copy until the process ends.  Nothing in it needs clearing: each pass
// writes every histogram entry and digit total before reading them.
constexpr int NOALLOC_TILES = (NOALLOC_MAX_PARTICLES + RS_TILE - 1) / RS_TILE;
constexpr int64_t NOALLOC_WORDS = 2 * (int64_t)NOALLOC_MAX_PARTICLES + (int64_t)RS_BINS * NOALLOC_TILES + RS_BINS;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[64] = {};  // per device: the device address of scratch
}  // namespace

// Device address of this GPU's scratch, looked up once per device with that
// device current.  The host-side name of the array is not a device address.
static lq_status noalloc_scratch(uint32_t** out) {
    int dev = 0;
    LQ_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("lq_grid_build_noalloc: device index beyond 63");
        return LQ_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        LQ_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (uint32_t*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return LQ_OK;
}

lq_status launch_grid_build_noalloc(const RunDesc& d, const float4* pospres, int32_t* part_cel, int32_t* part_idx,
                                    int32_t* cell_beg, int32_t* cell_end, hipStream_t s) {
    if (d.np > NOALLOC_MAX_PARTICLES) {
        set_error("lq_grid_build_noalloc: np exceeds LQ_NOALLOC_MAX_PARTICLES");
        return LQ_ERR_CAPACITY;
    }
    uint32_t* ptr = nullptr;
    LQ_TRY(noalloc_scratch(&ptr));
    // The sorted pairs end in slot passes % 2: make that slot the caller's
    // arrays and the other one the scratch.
    const int bits = rs_bits(d.par.ncell) + d.zbits;
    const int out = ((bits + 7) / 8) % 2;
    GridBuffers g{};
    g.keys[out] = (uint32_t*)part_cel;
    g.vals[out] = (uint32_t*)part_idx;
    g.keys[1 - out] = ptr;
    g.vals[1 - out] = ptr + NOALLOC_MAX_PARTICLES;
    g.hist = ptr + 2 * (int64_t)NOALLOC_MAX_PARTICLES;
    g.totals = g.hist + (int64_t)RS_BINS * NOALLOC_TILES;
    g.cell_beg = cell_beg;
    g.cell_end = cell_end;
    g.np_cap = NOALLOC_MAX_PARTICLES;
    g.ncell_cap = d.par.ncell;
    launch_grid_build_one(d, pospres, g, s);
    if (d.np > 0 && g.sorted_slot != out) {
        set_error("lq_grid_build_noalloc: sorted result not in the caller's arrays");
        return LQ_ERR_STATE;
    }
    return LQ_OK;
}

void launch_permute(const uint32_t* perm, int n, const DevState& src, const DevState& dst, const DevIni* ini_src,
                    const DevIni* ini_dst, tensor2d* str_dst, hipStream_t s, const RunDesc* desc,
                    const int32_t* run_of, int which, const DevState* half) {
    if (n <= 0) return;
    const int with_ini = (ini_src && ini_dst) ? 1 : 0;
    DevIni a = with_ini ? *ini_src : DevIni{};
    DevIni b = with_ini ? *ini_dst : DevIni{};
    k_permute<<<blocks_for(n, 256), 256, 0, s>>>(perm, n, src, dst, with_ini, a, b, str_dst,
                                                 DescTable{desc, run_of}, which,
                                                 (half && desc) ? *half : DevState{});
}

}  // namespace lq
