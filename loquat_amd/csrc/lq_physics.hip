// lq_physics.hip -- per-particle SPH stages for gfx950.
//
//   Adami wall stress         src/adami_boundary.cu:27-123
//   particle interaction      src/particle_interaction.h:29-41 (body not shipped;
//                             formulation fixed by its consumers,
//                             src/predictor_corrector.cu:186-211,276-300)
//   predictor / corrector     src/predictor_corrector.cu:169-218, 256-307
//   Drucker-Prager            src/predictor_corrector.cu:30-105
//   Wu hypoplastic            src/predictor_corrector.cu:111-165
//   von Wolffersdorff         constitutive model 3, no reference counterpart
//   Shepard stress filter     src/field_smoothing.cu:27-121
//
// One thread per particle of the packed multi-run layout.  Neighbour loops
// walk the nine (x, y) columns of the 3x3x3 cell stencil per layer; the
// neighbour test uses the FP32 mirror pospres (bit-identical predicate on
// every implementation), pair terms use the FP64 position difference rounded
// to FP32.  All sums and stress rates are FP32; positions and time are FP64;
// the Drucker-Prager return mapping is FP64.
#include "lq_internal.hpp"

namespace lq {

namespace {

inline int blocks_for(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Block index remapped so that each XCD works on one contiguous slice of the
// launch's index space.  Workgroups are dealt to the 8 XCDs round-robin;
// without the remap every XCD touches the whole domain and spatial
// neighbours are fetched into all eight L2s.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int xcd = b & 7, idx = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Compact thread index of this thread in an XCD-contiguous launch.
__device__ __forceinline__ int xcd_thread() { return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; }

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// Minimum / maximum of v over the 64 lanes (every lane must call).
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
// Slots per wave of the staged scan (scan_compact, STG): a chunk stages
// STG_SLOTS consecutive candidates.  128 x 16 B with 32-entry lists keeps a
// one-wave workgroup at 10 KB of LDS: sixteen per CU, as many as registers
// allow (with 4 more slots, or 40-entry lists, fifteen or thirteen fit, and
// the launch slows by 9 % and 15 %).
constexpr int STG_SLOTS = 128;
// Variant 19 stages the scan from this many material particles per launch
// (single columns, ms per step, variant 18 unstaged / 19 staged: 0.288 /
// 0.339 at 65k, 0.403 / 0.423 at 221k, 0.726 / 0.721 at 524k, 1.264 / 1.211
// at 1.02 M).
constexpr int STG_MIN_LAUNCH = 1 << 19;

// The value of lane ^ 1 (DPP quad permutation 1, 0, 3, 2).
__device__ __forceinline__ float swap_pair(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ tensor2d tadd(const tensor2d& a, const tensor2d& b) {
    return tensor2d{a.xx + b.xx, a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}
__device__ __forceinline__ tensor2d tscale(const tensor2d& a, float s) {
    return tensor2d{a.xx * s, a.yy * s, a.zz * s, a.xy * s, a.xz * s, a.yz * s};
}
// Double contraction with off-diagonals weighted twice (J2 of include/types.h:77-85).
__device__ __forceinline__ float tdot(const tensor2d& a, const tensor2d& b) {
    return a.xx * b.xx + a.yy * b.yy + a.zz * b.zz + 2.0f * (a.xy * b.xy + a.xz * b.xz + a.yz * b.yz);
}

// Jaumann co-rotational term W.sigma - sigma.W with spin (W_yx, W_zx, W_zy)
// (src/predictor_corrector.cu:39-45 and :118-124).
__device__ __forceinline__ tensor2d jaumann(const float4 w, const tensor2d& s) {
    tensor2d r;
    r.xx = -2.0f * (w.x * s.xy + w.y * s.xz);
    r.yy = 2.0f * (w.x * s.xy - w.z * s.yz);
    r.zz = 2.0f * (w.y * s.xz + w.z * s.yz);
    r.xy = w.x * (s.xx - s.yy) - w.z * s.xz - w.y * s.yz;
    r.xz = w.y * (s.xx - s.zz) + w.z * s.xy - w.x * s.yz;
    r.yz = w.z * (s.yy - s.zz) + w.x * s.xz + w.y * s.xy;
    return r;
}

// ------------------------------------------------------------ Adami walls
// n threads over the walls of all runs, walls(t) = particle index.  Blocks
// keep the hardware's round-robin XCD order: only walls next to material have
// work, and contiguous XCD slices would leave most XCDs idle.
template <class DP>
__global__ __launch_bounds__(256) void k_adami(DP dp, DevState st, const int32_t* __restrict__ beg,
                                               const int32_t* __restrict__ end, IndexMap walls, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = walls(t);
    const RunDesc& d = dp.get(dp.run(k));
    if ((k - d.p_off) >= d.npb || !d.active) return;
    const float4 pi = st.pospres[k];
    const double xi = st.xd[k], yi = st.yd[k], zi = st.zd[k];
    const float i_h = d.par.i_h, cw = d.par.kernel_normalization_par, m = d.par.m, r2cut = d.r2_cut;
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
    float px = 0.f, py = 0.f, pz = 0.f, w = 0.f;
    const int cix = cell_coord(pi.x, i_h), ciy = cell_coord(pi.y, i_h), ciz = cell_coord(pi.z, i_h);
    for_each_column(d, cix, ciy, ciz, 1, 1, beg, end, [&](int jb, int je) {
        for (int j = jb; j < je; ++j) {
            const float4 pj = st.pospres[j];
            if (r2_fp32(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < r2cut) {
                const float dx = (float)(xi - st.xd[j]), dy = (float)(yi - st.yd[j]), dz = (float)(zi - st.zd[j]);
                const float q = fsqrt(dx * dx + dy * dy + dz * dz) * i_h;
                const float t = fmaxf(2.0f - q, 0.0f);
                const float t2 = t * t;
                const float kv = cw * t2 * t2 * (q + 0.5f) * m * frcp(st.velrhop[j].w);
                const tensor2d s = st.str[j];
                sxx += s.xx * kv;
                syy += s.yy * kv;
                szz += s.zz * kv;
                sxy += s.xy * kv;
                sxz += s.xz * kv;
                syz += s.yz * kv;
                px += dx * kv;
                py += dy * kv;
                pz += dz * kv;
                w += kv;
            }
        }
    });
    if (w > 0.0f) {
        const float iw = 1.0f / w;
        const float rho0 = d.par.rho0;
        tensor2d o;
        o.xx = (sxx - rho0 * d.par.acc_x * px) * iw;
        o.yy = (syy - rho0 * d.par.acc_y * py) * iw;
        o.zz = (szz - rho0 * d.par.acc_z * pz) * iw;
        o.xy = sxy * iw;
        o.xz = sxz * iw;
        o.yz = syz * iw;
        st.str[k] = o;
    }
}

// ------------------------------------------------------------ interaction
struct IPart {
    double x, y, z;
    float vx, vy, vz, rho;
    tensor2d T;  // sigma_i / rho_i^2
    float R;     // artificial pressure eps * max(-p_i, 0) / rho_i^2 (0 when off)
};
struct IAcc {
    float ax, ay, az, dvg;
    float L[9];  // sum_j (v_j - v_i)_a / rho_j * dW_b  (times m at the end)
};
struct IConst {
    float i_h, cg, h, eta, alpha_cs;
    float eps, cw, inv_wdr;  // artificial pressure: coefficient (0 when off), C_w, 1 / W(dr)
};

// One pair i-j: Wendland C2 gradient dW = C_g (2 - q)^3 (x_i - x_j), SPH
// velocity gradient L_ab += (v_j - v_i)_a / rho_j dW_b, continuity
// drho += (v_i - v_j).dW, momentum a += (sigma_i/rho_i^2 + sigma_j/rho_j^2
// - Pi_ij I).dW with Monaghan's viscosity Pi_ij for approaching pairs (the
// mass m multiplies the sums when they are stored).
__device__ __forceinline__ void pair_eval(const IPart& pi, int j, const DevState& st, const IConst& c, IAcc& a) {
    const float dx = (float)(pi.x - st.xd[j]), dy = (float)(pi.y - st.yd[j]), dz = (float)(pi.z - st.zd[j]);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float q = fsqrt(r2) * c.i_h;
    const float t = fmaxf(2.0f - q, 0.0f);
    const float fac = c.cg * t * t * t;
    const float gx = fac * dx, gy = fac * dy, gz = fac * dz;
    const float4 vr = st.velrhop[j];
    const tensor2d sj = st.str[j];
    const float ir = frcp(vr.w);
    const float dvx = vr.x - pi.vx, dvy = vr.y - pi.vy, dvz = vr.z - pi.vz;
    const float ux = dvx * ir, uy = dvy * ir, uz = dvz * ir;
    a.L[0] += ux * gx;
    a.L[1] += ux * gy;
    a.L[2] += ux * gz;
    a.L[3] += uy * gx;
    a.L[4] += uy * gy;
    a.L[5] += uy * gz;
    a.L[6] += uz * gx;
    a.L[7] += uz * gy;
    a.L[8] += uz * gz;
    a.dvg += dvx * gx + dvy * gy + dvz * gz;
    const float vdx = -(dvx * dx + dvy * dy + dvz * dz);  // (v_i - v_j).(x_i - x_j)
    float pij = 0.0f;
    if (vdx < 0.0f) {
        const float mu = c.h * vdx * frcp(r2 + c.eta);
        pij = -2.0f * c.alpha_cs * mu * frcp(pi.rho + vr.w);
    }
    const float ir2 = ir * ir;
    const float sxx = pi.T.xx + sj.xx * ir2 - pij;
    const float syy = pi.T.yy + sj.yy * ir2 - pij;
    const float szz = pi.T.zz + sj.zz * ir2 - pij;
    const float sxy = pi.T.xy + sj.xy * ir2;
    const float sxz = pi.T.xz + sj.xz * ir2;
    const float syz = pi.T.yz + sj.yz * ir2;
    a.ax += sxx * gx + sxy * gy + sxz * gz;
    a.ay += sxy * gx + syy * gy + syz * gz;
    a.az += sxz * gx + syz * gy + szz * gz;
}

__device__ __forceinline__ void load_part(const DevState& st, int k, IPart& p) {
    p.x = st.xd[k];
    p.y = st.yd[k];
    p.z = st.zd[k];
    const float4 vr = st.velrhop[k];
    p.vx = vr.x;
    p.vy = vr.y;
    p.vz = vr.z;
    p.rho = vr.w;
    const float ir = 1.0f / vr.w;
    p.T = tscale(st.str[k], ir * ir);
}

__device__ __forceinline__ IConst make_iconst(const RunDesc& d) {
    IConst c;
    c.i_h = d.par.i_h;
    c.cg = d.par.kernel_gradient_normalization_par;
    c.h = d.par.h;
    c.eta = d.par.eta;
    c.alpha_cs = d.par.viscosity_type == 1 ? d.par.artificial_viscosity_alpha * d.par.cs : 0.0f;
    const bool art = d.par.artificial_pressure == 1 && d.par.kernel_dr > 0.0f;
    c.eps = art ? d.par.artificial_pressure_coefficient : 0.0f;
    c.cw = d.par.kernel_normalization_par;
    c.inv_wdr = art ? 1.0f / d.par.kernel_dr : 0.0f;
    return c;
}

// One particle's rates: acceleration and density rate, spin, strain rate.
struct RateVals {
    float4 accad;
    float4 spin;
    tensor2d acestn;
};

__device__ __forceinline__ RateVals rates_of(const RunDesc& d, const IAcc& a) {
    const float m = d.par.m;
    RateVals r;
    r.accad = make_float4(m * a.ax + d.par.acc_x, m * a.ay + d.par.acc_y, m * a.az + d.par.acc_z, -m * a.dvg);
    const float hm = 0.5f * m;
    // spin = (W_yx, W_zx, W_zy): sign fixed by the Jaumann term of
    // src/predictor_corrector.cu:40-45 (rigid rotation about z gives spin.x = Omega).
    r.spin = make_float4(hm * (a.L[3] - a.L[1]), hm * (a.L[6] - a.L[2]), hm * (a.L[7] - a.L[5]), 0.0f);
    r.acestn.xx = m * a.L[0];
    r.acestn.yy = m * a.L[4];
    r.acestn.zz = m * a.L[8];
    r.acestn.xy = hm * (a.L[1] + a.L[3]);
    r.acestn.xz = hm * (a.L[2] + a.L[6]);
    r.acestn.yz = hm * (a.L[5] + a.L[7]);
    return r;
}

__device__ __forceinline__ void store_rates(const RunDesc& d, const DevRates& rt, int k, const IAcc& a) {
    const RateVals r = rates_of(d, a);
    rt.accad[k] = r.accad;
    rt.spin[k] = r.spin;
    rt.acestn[k] = r.acestn;
}

__device__ __forceinline__ RateVals load_rates(const DevRates& rt, int k) {
    return RateVals{rt.accad[k], rt.spin[k], rt.acestn[k]};
}

// All interaction variants run n threads over the material particles of all
// runs, mats(t) = particle index.
// Variant 0: each thread walks its own candidates and evaluates hits inline.
template <class DP>
__global__ __launch_bounds__(256) void k_interaction_direct(DP dp, DevState st, DevRates rt,
                                                            const int32_t* __restrict__ beg,
                                                            const int32_t* __restrict__ end, IndexMap mats,
                                                            int n, int32_t* __restrict__ nbr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = mats(t);
    const RunDesc& d = dp.get(dp.run(k));
    if ((k - d.p_off) < d.npb || !d.active) return;
    const IConst c = make_iconst(d);
    const float r2cut = d.r2_cut;
    const float4 ps = st.pospres[k];
    IPart pi;
    load_part(st, k, pi);
    IAcc a = {};
    int count = 0;
    const int cix = cell_coord(ps.x, c.i_h), ciy = cell_coord(ps.y, c.i_h), ciz = cell_coord(ps.z, c.i_h);
    for_each_column(d, cix, ciy, ciz, 0, 1, beg, end, [&](int jb, int je) {
        for (int j = jb; j < je; ++j) {
            if (j == k) continue;
            const float4 pj = st.pospres[j];
            if (r2_fp32(ps.x - pj.x, ps.y - pj.y, ps.z - pj.z) < r2cut) {
                ++count;
                pair_eval(pi, j, st, c, a);
            }
        }
    });
    store_rates(d, rt, k, a);
    if (nbr) nbr[k] = count;
}

// Variant 1: the candidate scan appends hits to a per-lane list in LDS; the
// wave evaluates the lists when any lane's list is full and at the end, so
// the ~60-flop pair body runs with most lanes busy instead of once per
// candidate with ~15% of lanes busy.  Summation order per particle is the
// candidate order, identical to variant 0.
constexpr int IC_CAP = 32;

template <class DP>
__global__ __launch_bounds__(256) void k_interaction_lds(DP dp, DevState st, DevRates rt,
                                                         const int32_t* __restrict__ beg,
                                                         const int32_t* __restrict__ end, IndexMap mats,
                                                         int n, int32_t* __restrict__ nbr) {
    __shared__ int32_t lists[4][IC_CAP][64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = t < n ? mats(t) : -1;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int32_t(*my)[64] = lists[w];
    bool mat = false;
    const RunDesc* dptr = nullptr;
    if (t < n) {
        const RunDesc& d = dp.get(dp.run(k));
        dptr = &d;
        mat = (k - d.p_off) >= d.npb && d.active;
    }
    IConst c = {};
    float r2cut = 0.f;
    float4 ps = make_float4(0.f, 0.f, 0.f, 0.f);
    IPart pi = {};
    int cix = 0, ciy = 0, ciz = 0, nx = 0, ny = 0, nz = 0, nol = 0, cbase = 0;
    if (mat) {
        const RunDesc& d = *dptr;
        c = make_iconst(d);
        r2cut = d.r2_cut;
        ps = st.pospres[k];
        load_part(st, k, pi);
        cix = cell_coord(ps.x, c.i_h);
        ciy = cell_coord(ps.y, c.i_h);
        ciz = cell_coord(ps.z, c.i_h);
        nx = d.par.cell_num_x;
        ny = d.par.cell_num_y;
        nz = d.par.cell_num_z;
        nol = d.par.ncell_one_layer;
        cbase = d.cell_base;
    }
    IAcc a = {};
    int count = 0, cnt = 0;
    const int czlo = max(ciz - 1, 0), czhi = min(ciz + 1, nz - 1);
    for (int col = 0; col < 18; ++col) {
        const int layer = col / 9, ox = (col % 9) / 3 - 1, oy = col % 3 - 1;
        int jb = 0, je = 0;
        if (mat && czlo <= czhi) {
            const int cx = cix + ox, cy = ciy + oy;
            if (cx >= 0 && cx < nx && cy >= 0 && cy < ny) {
                const int c0 = cbase + layer * nol + (cx * ny + cy) * nz;
                int b0 = -1, e0 = -1;
                for (int cz = czlo; cz <= czhi; ++cz) {
                    const int b = beg[c0 + cz];
                    if (b != -1) {
                        if (b0 < 0) b0 = b;
                        e0 = end[c0 + cz];
                    }
                }
                if (b0 >= 0) {
                    jb = b0;
                    je = e0;
                }
            }
        }
        int j = jb;
        while (__ballot(j < je)) {
            if (j < je) {
                if (j != k) {
                    const float4 pj = st.pospres[j];
                    if (r2_fp32(ps.x - pj.x, ps.y - pj.y, ps.z - pj.z) < r2cut) {
                        my[cnt][lane] = j;
                        ++cnt;
                    }
                }
                ++j;
            }
            if (__ballot(cnt == IC_CAP)) {
                for (int e = 0; __ballot(e < cnt); ++e) {
                    if (e < cnt) pair_eval(pi, my[e][lane], st, c, a);
                }
                count += cnt;
                cnt = 0;
            }
        }
    }
    for (int e = 0; __ballot(e < cnt); ++e) {
        if (e < cnt) pair_eval(pi, my[e][lane], st, c, a);
    }
    count += cnt;
    if (mat) {
        store_rates(*dptr, rt, k, a);
        if (nbr) nbr[k] = count;
    }
}

// Variant 2: as variant 1, restructured for latency and issue rate.  The
// candidate loop tests U = 4 candidates per iteration (four loads in flight
// per lane), appends branch-free (the slot is always written, the count only
// advances on a hit), and never touches the accumulators: lists are evaluated
// between column segments, so the loop carries no accumulator registers.
constexpr int IC2_CAP = 32;
constexpr int IC2_U = 4;
// List length of variants 3 to 6: 4 x 40 x 64 x 4 B = 40 KB per 256-thread
// workgroup, so four workgroups (the register limit) fill the CU's 160 KB.
constexpr int IC3_CAP = 40;

template <class DP>
__global__ __launch_bounds__(256) void k_interaction_unrolled(DP dp, DevState st, DevRates rt,
                                                              const int32_t* __restrict__ beg,
                                                              const int32_t* __restrict__ end, IndexMap mats,
                                                              int n, int32_t* __restrict__ nbr) {
    __shared__ int32_t lists[4][IC2_CAP][64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = t < n ? mats(t) : -1;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int32_t(*my)[64] = lists[w];
    bool mat = false;
    const RunDesc* dptr = nullptr;
    if (t < n) {
        const RunDesc& d = dp.get(dp.run(k));
        dptr = &d;
        mat = (k - d.p_off) >= d.npb && d.active;
    }
    IConst c = {};
    float r2cut = 0.f;
    float4 ps = make_float4(0.f, 0.f, 0.f, 0.f);
    IPart pi = {};
    int cix = 0, ciy = 0, ciz = 0, nx = 0, ny = 0, nz = 0, nol = 0, cbase = 0;
    if (mat) {
        const RunDesc& d = *dptr;
        c = make_iconst(d);
        r2cut = d.r2_cut;
        ps = st.pospres[k];
        load_part(st, k, pi);
        cix = cell_coord(ps.x, c.i_h);
        ciy = cell_coord(ps.y, c.i_h);
        ciz = cell_coord(ps.z, c.i_h);
        nx = d.par.cell_num_x;
        ny = d.par.cell_num_y;
        nz = d.par.cell_num_z;
        nol = d.par.ncell_one_layer;
        cbase = d.cell_base;
    }
    IAcc a = {};
    int count = 0, cnt = 0;
    const int czlo = max(ciz - 1, 0), czhi = min(ciz + 1, nz - 1);
    const float4* __restrict__ pp = st.pospres;
    for (int col = 0; col < 18; ++col) {
        const int layer = col / 9, ox = (col % 9) / 3 - 1, oy = col % 3 - 1;
        int j = 0, je = 0;
        if (mat && czlo <= czhi) {
            const int cx = cix + ox, cy = ciy + oy;
            if (cx >= 0 && cx < nx && cy >= 0 && cy < ny) {
                const int c0 = cbase + layer * nol + (cx * ny + cy) * nz;
                int b0 = -1, e0 = -1;
                for (int cz = czlo; cz <= czhi; ++cz) {
                    const int b = beg[c0 + cz];
                    if (b != -1) {
                        if (b0 < 0) b0 = b;
                        e0 = end[c0 + cz];
                    }
                }
                if (b0 >= 0) {
                    j = b0;
                    je = e0;
                }
            }
        }
        while (true) {
            while (__ballot(j < je) && !__ballot(cnt > IC2_CAP - IC2_U)) {
                float4 q[IC2_U];
#pragma unroll
                for (int u = 0; u < IC2_U; ++u) {
                    const int jj = min(j + u, je - 1);
                    q[u] = pp[max(jj, 0)];
                }
#pragma unroll
                for (int u = 0; u < IC2_U; ++u) {
                    const int jj = j + u;
                    const bool hit = jj < je && jj != k &&
                                     r2_fp32(ps.x - q[u].x, ps.y - q[u].y, ps.z - q[u].z) < r2cut;
                    my[cnt][lane] = jj;
                    cnt += hit ? 1 : 0;
                }
                j += IC2_U;
            }
            if (!__ballot(j < je)) break;
            for (int e = 0; __ballot(e < cnt); ++e) {
                if (e < cnt) pair_eval(pi, my[e][lane], st, c, a);
            }
            count += cnt;
            cnt = 0;
        }
        if (__ballot(cnt > IC2_CAP - 3 * IC2_U)) {
            for (int e = 0; __ballot(e < cnt); ++e) {
                if (e < cnt) pair_eval(pi, my[e][lane], st, c, a);
            }
            count += cnt;
            cnt = 0;
        }
    }
    for (int e = 0; __ballot(e < cnt); ++e) {
        if (e < cnt) pair_eval(pi, my[e][lane], st, c, a);
    }
    count += cnt;
    if (mat) {
        store_rates(*dptr, rt, k, a);
        if (nbr) nbr[k] = count;
    }
}

// 32-bit byte offset loads from a uniform base (saddr addressing: one shift
// per load instead of 64-bit address arithmetic).
template <class T>
__device__ __forceinline__ T ldo(const T* __restrict__ base, uint32_t idx) {
    return *(const T*)((const char*)base + (idx * (uint32_t)sizeof(T)));
}

// What a pair reads of particle j.
struct PairIn {
    double x, y, z;
    float4 vr;
    tensor2d s;
    float p;  // pressure of j, read only with artificial pressure
};

// The value at byte offset off from base.
template <class T>
__device__ __forceinline__ T ldb(const T* __restrict__ base, uint32_t off) {
    return *(const T*)((const char*)base + off);
}

// Byte offsets of particle j in the 8-, 16- and 24-byte arrays from one
// shift: 24 j = 8 j + 16 j (the empty asm keeps the compiler from folding
// it back into a 32-bit multiply, a quarter-rate instruction).
struct Offs {
    uint32_t o8, o16, o24;
};
__device__ __forceinline__ Offs offs_of(uint32_t j) {
    uint32_t o8 = j << 3;
    asm("" : "+v"(o8));
    const uint32_t o16 = o8 << 1;
    return Offs{o8, o16, o8 + o16};
}

template <bool ART = false>
__device__ __forceinline__ PairIn pair_load(const DevState& st, uint32_t j) {
    const float pj = ART ? ldo(reinterpret_cast<const float*>(st.pospres) + 3, 4u * j) : 0.0f;
    const Offs o = offs_of(j);
    return PairIn{ldb(st.xd, o.o8), ldb(st.yd, o.o8), ldb(st.zd, o.o8), ldb(st.velrhop, o.o16), ldb(st.str, o.o24), pj};
}

// Pair body of variants 3 to 6: as pair_eval, with 32-bit offset loads and
// the viscosity term evaluated branch-free.  ART adds Monaghan's artificial
// pressure (R_i + R_j) f^4, f = W(r) / W(dr), R = eps max(-p, 0) / rho^2, to
// the viscous pressure (enabled by artificial-pressure = 1).
template <bool ART = false>
__device__ __forceinline__ void pair_compute(const IPart& pi, const PairIn& pj, const IConst& c, IAcc& a) {
    const float dx = (float)(pi.x - pj.x), dy = (float)(pi.y - pj.y), dz = (float)(pi.z - pj.z);
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float q = fsqrt(r2) * c.i_h;
    const float t = fmaxf(2.0f - q, 0.0f);
    const float fac = c.cg * t * t * t;
    const float gx = fac * dx, gy = fac * dy, gz = fac * dz;
    const float4 vr = pj.vr;
    const tensor2d sj = pj.s;
    const float ir = frcp(vr.w);
    const float dvx = vr.x - pi.vx, dvy = vr.y - pi.vy, dvz = vr.z - pi.vz;
    const float ux = dvx * ir, uy = dvy * ir, uz = dvz * ir;
    a.L[0] += ux * gx;
    a.L[1] += ux * gy;
    a.L[2] += ux * gz;
    a.L[3] += uy * gx;
    a.L[4] += uy * gy;
    a.L[5] += uy * gz;
    a.L[6] += uz * gx;
    a.L[7] += uz * gy;
    a.L[8] += uz * gz;
    const float vdx = -(dvx * dx + dvy * dy + dvz * dz);
    // (v_j - v_i).grad W = -fac (v_i - v_j).(x_i - x_j)
    a.dvg -= fac * vdx;
    const float mu = c.h * fminf(vdx, 0.0f) * frcp(r2 + c.eta);
    float pij = -2.0f * c.alpha_cs * mu * frcp(pi.rho + vr.w);
    const float ir2 = ir * ir;
    if constexpr (ART) {
        const float t2 = t * t;
        const float f = c.cw * t2 * t2 * (q + 0.5f) * c.inv_wdr;
        const float f2 = f * f;
        pij += (pi.R + c.eps * fmaxf(-pj.p, 0.0f) * ir2) * (f2 * f2);
    }
    const float sxx = pi.T.xx + sj.xx * ir2 - pij;
    const float syy = pi.T.yy + sj.yy * ir2 - pij;
    const float szz = pi.T.zz + sj.zz * ir2 - pij;
    const float sxy = pi.T.xy + sj.xy * ir2;
    const float sxz = pi.T.xz + sj.xz * ir2;
    const float syz = pi.T.yz + sj.yz * ir2;
    a.ax += sxx * gx + sxy * gy + sxz * gz;
    a.ay += sxy * gx + syy * gy + syz * gz;
    a.az += sxz * gx + syz * gy + szz * gz;
}

__device__ __forceinline__ void pair_eval3(const IPart& pi, uint32_t j, const DevState& st, const IConst& c,
                                           IAcc& a) {
    pair_compute(pi, pair_load(st, j), c, a);
}

__device__ __forceinline__ float stale_disp(const StaleIn& sn, const RunDesc& d, int run) {
    if (sn.from_desc) return (sn.from_desc == 1 && d.sort_due) ? 0.0f : d.dtab;
    return sn.dmax ? __uint_as_float(sn.dmax[run * DMAX_SLOTS]) : sn.disp;
}
__device__ __forceinline__ bool stale_out_now(const StaleIn& sn, const RunDesc& d, int run) {
    if (sn.from_desc) return d.out_tab != 0;
    return !sn.nout || sn.nout[run] != 0u;
}
// Where the table put particle k.
__device__ __forceinline__ float4 stale_pos(const StaleIn& sn, int k) {
    return sn.po ? sn.po[k] : make_float4((float)sn.xo[k], (float)sn.yo[k], (float)sn.zo[k], 0.0f);
}

// Wave-cooperative neighbour scan shared by the list-based kernels.  For a
// lane with `active` set it walks the z-merged (x, y) columns of cell layers
// [layer_lo, layer_hi] around cell (cix, ciy, ciz) -- (x, y) major, layers
// inside -- tests candidates four at a
// time with the FP32 predicate (branch-free, four loads in flight), appends
// hits to the lane's LDS list `my`, and calls eval(j, true) for every listed
// j, in candidate order, whenever some lane's list is nearly full and at the
// end.  Lanes with a shorter list meanwhile call eval(dummy, false), so that
// the accumulators are updated unconditionally (no copies around a masked
// update); eval must make that call add exactly nothing.  Every lane of the
// wave must call it (it votes); returns the lane's hits.
//
// Pruning (never drops a pair the predicate accepts): a column whose nearest
// (x, y) distance already exceeds the support is skipped, and when cells are
// sorted by z-bin (zbits > 0) each column is read only over the z-window
// |dz| <= sqrt(R^2 - dxy_min^2) widened by one bin: blocks of four lying wholly
// below it are stepped over (four blocks per load round), and the scan stops
// at the first block whose last particle lies above it.  The kernel is bound
// by the latency of dependent load rounds, so each column's set-up issues its
// loads together: both layers' cell-table entries in one round, the z-window
// start in (usually) one more.  R is the support radius inflated by eps and
// dxy_min is shrunk by eps of a cell (eps = 1e-3, more on grids over ~2000
// cells per axis), far beyond the FP32 rounding of the coordinates, so every
// skipped candidate has r2 > r2_cut.
//
// STALE: the table was built before the particles moved (make_geom_stale);
// reach and z-window grow by the run's slack, a ring of columns two cells
// out and the run's outside cell are read when some lane needs them, and
// hits that could lie outside the current cell's stencil are tested against
// it, so the neighbour sets are those of a freshly built table.
//
// B = 2 evaluates the list two entries at a time, eval(j0, act0, j1, act1),
// so that both gathers are in flight before either pair is computed.
//
// PF > 0: when some lane's list is full, only the first PF entries of every
// list are evaluated and the rest moved to the front (each lane still sums
// its pairs in candidate order); without it the whole lists are evaluated,
// every lane padding to the longest.  ORD: the order of the nine (x, y)
// columns (below); with opposite columns in pairs the lanes' list lengths
// stay alike, so partial flushes pad little and the lanes' gathers stay in
// step.  On the 1M-particle column (variant 11's shape): 0.562 ms per
// launch; PF = 8 alone 0.583 (the lanes' gathers drift apart: more
// vector-memory traffic); ORD = 1 alone 0.547; ORD = 2 with PF = 8 / 12 / 16
// / 20: 0.516 / 0.510 / 0.509 / 0.508.
struct ScanGeom {
    int cix, ciy, ciz, nx, ny, nz, nol, cbase;  // stencil cell and grid
    int czlo, czhi;  // z cells read in every column
    int zbits;
    float i_h;
    float u, v;     // position inside the stencil cell along x and y, in cell units
    float cw;       // cell width (2 h)
    float rw2;      // ((1 + eps) R)^2, R^2 = r2_cut
    float eps;      // rounding allowance, in cells (and relative to R)
    float z, binw;  // particle z and the z-bin width widened by 1 %
    // Stale tables: 0, or twice the largest displacement of any particle of
    // the run since the table was built; whether the particle's current cell
    // lies on (or beyond) the grid's edge, and whether some particle of the
    // run lies outside the grid now.
    float slack;
    float ext2;  // ((1 - eps) cell)^2
    bool edge;
    bool out_now;
};

__device__ __forceinline__ ScanGeom make_geom(const RunDesc& d, float4 ps) {
    const float i_h = d.par.i_h;
    ScanGeom g;
    g.cix = cell_coord(ps.x, i_h);
    g.ciy = cell_coord(ps.y, i_h);
    g.ciz = cell_coord(ps.z, i_h);
    g.nx = d.par.cell_num_x;
    g.ny = d.par.cell_num_y;
    g.nz = d.par.cell_num_z;
    g.nol = d.par.ncell_one_layer;
    g.cbase = d.cell_base;
    g.czlo = max(g.ciz - 1, 0);
    g.czhi = min(g.ciz + 1, g.nz - 1);
    g.zbits = d.zbits;
    g.i_h = i_h;
    const float sx = (0.5f * ps.x) * i_h, sy = (0.5f * ps.y) * i_h;
    g.u = sx - floorf(sx);
    g.v = sy - floorf(sy);
    g.cw = 2.0f / i_h;
    // FP32 rounding of a cell coordinate (at most a few ulps of the largest
    // one) and of the squared distance stay far below eps = 1e-3 for grids of
    // up to ~2000 cells per axis; larger grids get a proportionally larger eps
    const int nmax = max(max(g.nx, g.ny), g.nz);
    g.eps = fmaxf(1e-3f, 4.8e-7f * (float)nmax);
    const float rw = (1.0f + g.eps) * sqrtf(d.r2_cut);
    g.rw2 = rw * rw;
    g.z = ps.z;
    g.binw = 1.01f * g.cw / (float)(1 << d.zbits);
    g.slack = 0.0f;
    g.ext2 = ((1.0f - g.eps) * g.cw) * ((1.0f - g.eps) * g.cw);
    g.edge = false;
    g.out_now = false;
    return g;
}

// Geometry for a table built when the particle was at `po` (ps: now), every
// particle of the run having moved at most `disp` since.  The (x, y) stencil
// is taken around the old cell; the z cells are those a neighbour's old z can
// lie in, |z_old - z| <= (1 + eps) R + disp (plus rounding).
__device__ __forceinline__ ScanGeom make_geom_stale(const RunDesc& d, float4 ps, float4 po, float disp, bool out_now) {
    ScanGeom g = make_geom(d, po);
    const float i_h = g.i_h;
    const int ccx = cell_coord(ps.x, i_h), ccy = cell_coord(ps.y, i_h), ccz = cell_coord(ps.z, i_h);
    g.z = ps.z;
    g.slack = 2.0f * disp;
    const float reach = sqrtf(g.rw2) + disp + g.eps * g.cw;
    g.czlo = max(cell_coord(ps.z - reach, i_h), 0);
    g.czhi = min(cell_coord(ps.z + reach, i_h), g.nz - 1);
    g.edge = ccx <= 0 || ccx >= g.nx - 1 || ccy <= 0 || ccy >= g.ny - 1 || ccz <= 0 || ccz >= g.nz - 1;
    g.out_now = out_now;
    return g;
}

// A candidate q is in the stencil of the current cell of the particle at
// ps: q's current cell lies inside the grid and within one cell of it along
// every axis -- the neighbour set of a freshly sorted table.
__device__ __forceinline__ bool in_current_stencil(const ScanGeom& g, float4 ps, float4 q) {
    const int x = cell_coord(q.x, g.i_h), y = cell_coord(q.y, g.i_h), z = cell_coord(q.z, g.i_h);
    return x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz &&
           abs(x - cell_coord(ps.x, g.i_h)) <= 1 && abs(y - cell_coord(ps.y, g.i_h)) <= 1 &&
           abs(z - cell_coord(ps.z, g.i_h)) <= 1;
}

template <int B, bool PADDED, int CAP = IC2_CAP, bool STALE = false, int UU = IC2_U, int PF = 0, int ORD = 0,
          int LPP = 1, bool STG = false, class Eval>
__device__ __forceinline__ int scan_compact(bool active, int layer_lo, int layer_hi, const ScanGeom& g, float4 ps,
                                            float r2cut, const int32_t* __restrict__ beg,
                                            const int32_t* __restrict__ end, const float4* __restrict__ pp,
                                            int32_t (*my)[64], int lane, uint32_t dummy, Eval&& eval,
                                            float4* stage = nullptr) {
    static_assert(B == 1 || B == 2, "batch of 1 or 2");
    static_assert(LPP == 1 || LPP == 2, "one or two lanes per particle");
    static_assert(UU * LPP <= SCAN_PAD, "a block of candidates may run SCAN_PAD entries past a range");
    constexpr int U = UU;  // candidates per load round (per lane)
    // LPP = 2: the two lanes of a particle (lane ^ 1) take a column's
    // candidates alternately (sub = 0: even offsets from the block start,
    // 1: odd), so a wave holds 32 particles
    constexpr int S = U * LPP;  // candidates per round of a particle's lanes
    const int sub = LPP == 2 ? (lane & 1) : 0;
    static_assert(PF == 0 || (PF % B == 0 && PF >= UU && PF <= CAP - UU), "partial flush size");
    int count = 0, cnt = 0;
    // Evaluates the first min(cnt, m) entries of every lane's list, in list
    // order, and moves the rest to the front (m = CAP: the whole list).
    auto flush_m = [&](int m) {
        const int lim = min(cnt, m);
        // bottom-tested loop: the accumulators flow straight from one
        // iteration to the next (a top-tested loop makes the compiler copy
        // them every iteration)
        if (__ballot(lim > 0)) {
            int e = 0;
            do {
                if constexpr (B == 1) {
                    const bool act = e < lim;
                    const uint32_t jl = (uint32_t)my[e][lane];
                    eval(act ? jl : dummy, act);
                } else {
                    const bool a0 = e < lim, a1 = e + 1 < lim;
                    const uint32_t j0 = (uint32_t)my[e][lane];
                    const uint32_t j1 = (uint32_t)my[min(e + 1, CAP - 1)][lane];
                    eval(a0 ? j0 : dummy, a0, a1 ? j1 : dummy, a1);
                }
                e += B;
            } while (__ballot(e < lim));
        }
        count += lim;
        cnt -= lim;
        if (m < CAP) {
            // entries [m, m + cnt) -> [0, cnt); m is the same in every lane
            for (int e = 0; __ballot(e < cnt); ++e) my[e][lane] = my[min(e + m, CAP - 1)][lane];
        }
    };
    auto flush = [&]() { flush_m(CAP); };
    // a full list: with PF > 0 only its first PF entries are evaluated (lanes
    // whose lists are shorter pad less than they would to the longest list)
    auto flush_full = [&]() {
        if constexpr (PF > 0)
            flush_m(PF);
        else
            flush();
    };
    const float* __restrict__ pz = reinterpret_cast<const float*>(pp) + 2;  // pz[4 j] = z of particle j
    const int czlo = g.czlo, czhi = g.czhi;
    const int nl = layer_hi - layer_lo + 1;
    const float margin = g.eps * g.cw;
    // Stale tables: a neighbour's old position may lie one ring of columns
    // further out (read only when some lane's particle is that close to its
    // old cell's side), and one that was outside the grid when the table was
    // built sits in the run's outside cell (read by lanes on the grid's edge).
    int nxy = 9;
    bool scan_out = false;
    // a hit that lies outside the grid now is possible only for a lane on
    // the grid's edge of a run with particles outside
    const unsigned long long edge_lanes = STALE ? __ballot(active && g.edge && g.out_now) : 0ull;
    int out_b = -1, out_e = -1;
    if constexpr (STALE) {
        const float fm = fminf(fminf(g.u, 1.0f - g.u), fminf(g.v, 1.0f - g.v));
        if (__ballot(active && (fm + 1.0f) * g.cw < sqrtf(g.rw2) + margin + g.slack)) nxy = 25;
        if (active && g.edge) {
            const int co = g.cbase + 2 * g.nol;
            out_b = beg[co];
            out_e = end[co];
        }
        scan_out = __ballot(out_b != -1) != 0ull;
    }
    const int ncol = nxy * nl + (scan_out ? 1 : 0);
    // [b0, e0) of the column's cells czlo..czhi starting at cell id c0 (b0 =
    // -1 when all are empty).  The (up to) three cells' six table entries are
    // loaded together; particles of consecutive cells are consecutive, so the
    // range runs from the first non-empty cell's begin to the last one's end.
    auto col_range = [&](int c0, int& b0, int& e0) {
        if constexpr (STALE) {
            // up to four cells (the old z of a neighbour spans a little more
            // than three cells)
            const int ca = c0 + czlo, cb = c0 + min(czlo + 1, czhi), cc = c0 + min(czlo + 2, czhi), cd = c0 + czhi;
            const int ba = beg[ca], bb = beg[cb], bc = beg[cc];
            const int ea = end[ca], eb = end[cb], ec = end[cc];
            int bd = -1, ed = -1;
            if (czhi - czlo == 3) {  // rare: the fourth cell
                bd = beg[cd];
                ed = end[cd];
            }
            b0 = ba != -1 ? ba : (bb != -1 ? bb : (bc != -1 ? bc : bd));
            e0 = bd != -1 ? ed : (bc != -1 ? ec : (bb != -1 ? eb : ea));
        } else {
            const int ca = c0 + czlo, cb = c0 + min(czlo + 1, czhi), cc = c0 + czhi;
            const int ba = beg[ca], bb = beg[cb], bc = beg[cc];
            const int ea = end[ca], eb = end[cb], ec = end[cc];
            b0 = ba != -1 ? ba : (bb != -1 ? bb : bc);
            e0 = bc != -1 ? ec : (bb != -1 ? eb : ea);
        }
    };
    // First candidate of the non-empty range [b0, e0) that may lie in the
    // z-window.  A block of U whose last particle is below zlo lies wholly
    // below the window, and so does everything before it (z-bins ascend
    // along the column); four blocks are tested per round, their last z
    // loaded together.
    auto skip_below = [&](int b0, int e0, float zlo) {
        int j = b0;
        while (true) {
            int adv = 0;
#pragma unroll
            for (int m = 1; m <= 4; ++m) {
                const int last = j + m * S - 1;
                const float zm = ldo(pz, 4u * (uint32_t)min(last, e0 - 1));
                if (last < e0 && zm < zlo) adv = m * S;
            }
            j += adv;
            if (adv < 4 * S) break;
        }
        return j;
    };
    // Columns are visited (x, y) major with the layers inside, so that one
    // set-up (geometry, table loads, z-window) serves both layers: the
    // second layer's range is prepared with the first and kept meanwhile.
    int jn = 0, jen = 0;
    float zhin = INFINITY;
    for (int col = 0; col < ncol; ++col) {
        const int xy = col / nl, layer = layer_lo + col % nl;
        int ox, oy;
        if (xy < 9) {
            // Column order (xy = 3 (ox + 1) + oy + 1, one hex digit per
            // step).  ORD = 0: x major.  ORD = 1: opposite columns one after
            // the other, the particle's own column last -- (-1,-1) (1,1)
            // (-1,0) (1,0) (-1,1) (1,-1) (0,-1) (0,1) (0,0).  ORD = 2: the
            // four edge columns in opposite pairs, the four corners in
            // opposite pairs, the own column -- (-1,0) (1,0) (0,-1) (0,1)
            // (-1,-1) (1,1) (-1,1) (1,-1) (0,0).  A lane near one side of
            // its cell gains in one column what it loses in the opposite
            // one, so at any point of the scan the lanes of a wave have
            // listed similar numbers of hits and flushes pad less.
            constexpr unsigned long long kOrd = ORD == 1 ? 0x453627180ull
                                                : ORD == 2 ? 0x462805371ull : 0x876543210ull;
            const int o = (int)((kOrd >> (4 * xy)) & 0xF);
            ox = o / 3 - 1;
            oy = o % 3 - 1;
        } else {
            // the ring of 16 columns two cells out
            const int r = xy - 9;
            ox = r < 5 ? -2 : (r < 10 ? 2 : (r < 13 ? r - 11 : r - 14));
            oy = r < 5 ? r - 2 : (r < 10 ? r - 7 : (r < 13 ? -2 : 2));
        }
        int j = 0, je = 0;
        float zhi = INFINITY;
        if (STALE && col == nxy * nl) {
            // the run's outside cell, no z order
            if (out_b != -1) {
                j = out_b;
                je = out_e;
            }
        } else if (layer != layer_lo) {
            j = jn;
            je = jen;
            zhi = zhin;
        } else {
            jn = jen = 0;
            zhin = INFINITY;
            if (active && czlo <= czhi) {
                const int cx = g.cix + ox, cy = g.ciy + oy;
                // distance from the particle's (old) position to the column,
                // less the rounding margin and the slack
                const float sh = margin + g.slack;
                const float ex =
                    ox == 0 ? 0.0f
                            : fmaxf(((ox < 0 ? g.u : 1.0f - g.u) + (float)(abs(ox) - 1)) * g.cw - sh, 0.0f);
                const float ey =
                    oy == 0 ? 0.0f
                            : fmaxf(((oy < 0 ? g.v : 1.0f - g.v) + (float)(abs(oy) - 1)) * g.cw - sh, 0.0f);
                const float hz2 = g.rw2 - (ex * ex + ey * ey);
                if (hz2 > 0.0f && cx >= 0 && cx < g.nx && cy >= 0 && cy < g.ny) {
                    const int c0 = g.cbase + layer * g.nol + (cx * g.ny + cy) * g.nz;
                    int b0, e0, b1 = -1, e1 = -1;
                    col_range(c0, b0, e0);
                    if (nl == 2) col_range(c0 + g.nol, b1, e1);
                    float zlo = -INFINITY;
                    if (g.zbits) {
                        // the slack also covers the order of the z-bins,
                        // which is that of the old positions
                        const float hz = sqrtf(hz2);
                        zlo = g.z - hz - g.binw - g.slack;
                        zhi = g.z + hz + g.binw + g.slack;
                    }
                    if (b1 >= 0) {
                        jn = g.zbits ? skip_below(b1, e1, zlo) : b1;
                        jen = e1;
                        zhin = zhi;
                    }
                    if (b0 >= 0) {
                        j = g.zbits ? skip_below(b0, e0, zlo) : b0;
                        je = e0;
                    } else {
                        zhi = INFINITY;
                    }
                }
            }
        }
        const uint32_t jlast = (uint32_t)max(je - 1, 0);
        if constexpr (STG && PADDED && LPP == 1) {
            // Staged scan: the wave copies the candidates of its lanes'
            // windows in chunks of STG_SLOTS consecutive particles (the
            // chunk starts at the lowest pending candidate of any lane) into
            // LDS with coalesced loads, and each lane tests its own blocks of
            // the chunk from there; a block whose candidates reach past the
            // chunk waits for the next one.  Same candidates, same order, same
            // lists.
            while (true) {
                if (__ballot(cnt > CAP - U)) {
                    flush_full();
                    continue;
                }
                const bool pend = j < je;
                if (!__ballot(pend)) break;
                const int base = wave_min_i32(pend ? j : INT_MAX);
                const int top = wave_max_i32(pend ? je : 0);  // every loaded index stays below it
#pragma unroll
                for (int h = 0; h < STG_SLOTS / 64; ++h)
                    stage[h * 64 + lane] = pp[min(base + h * 64 + lane, top - 1)];
                // the wave's own LDS writes precede its reads (one wave per slot)
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // a block is tested from this chunk when it starts within the
                // first STG_SLOTS - U + 1 slots: its four reads
                // (j - base + 3 < STG_SLOTS) then stay in the chunk; jlim =
                // -1 marks a column that ended in this chunk
                int jlim = min(je, base + (STG_SLOTS - U) + 1);
                // the chunk slot of j and the list entry my[cnt][lane] as LDS
                // addresses, advanced in place
                const float4* sq = stage + (j - base);
                int32_t* wp = &my[cnt][lane];
                int32_t* const wfull = &my[CAP - U][lane];
                while (true) {
                    if (__ballot(wp > wfull)) {
                        cnt = (int)((wp - &my[0][lane]) >> 6);
                        flush_full();
                        wp = &my[cnt][lane];
                    }
                    const bool can = j < jlim;
                    if (!__ballot(can)) break;
                    if (can) {
                        float4 q[U];
#pragma unroll
                        for (int u = 0; u < U; ++u) q[u] = sq[u];
                        int hit[U];  // 256 (the byte stride of list entries) or 0
                        bool edge = false;  // a hit past ext2: the exact cell test decides
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const float r2 = r2_fp32(ps.x - q[u].x, ps.y - q[u].y, ps.z - q[u].z);
                            const bool in = r2 < r2cut;
                            hit[u] = (in && (u == 0 || j + u < je)) ? 256 : 0;  // j < je holds
                            if constexpr (STALE) edge = edge | (in & (r2 >= g.ext2));
                        }
                        if constexpr (STALE) {
                            if (__ballot(edge) | edge_lanes) {
#pragma unroll
                                for (int u = 0; u < U; ++u)
                                    if (hit[u] && !in_current_stencil(g, ps, q[u])) hit[u] = 0;
                            }
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            *wp = j + u;
                            wp = (int32_t*)((char*)wp + hit[u]);
                        }
                        j += U;
                        sq += U;
                        if (q[U - 1].z > zhi) jlim = -1;
                    }
                }
                if (jlim < 0) je = min(je, j);
                cnt = (int)((wp - &my[0][lane]) >> 6);
                // the chunk's slots are rewritten only after every lane's reads
                __builtin_amdgcn_wave_barrier();
            }
        } else
        while (true) {
            while (__ballot(j < je) && !__ballot(cnt > CAP - U)) {
                // a lane whose column is done issues no loads (module calls
                // keep every lane in step: their loads are clamped instead)
                if (PADDED && j >= je) continue;
                float4 q[U];
                if (PADDED) {
                    // one offset for the four loads; a finished lane re-reads
                    // [je, je + S), inside the padded arrays
                    const uint32_t jj = (uint32_t)(min(j, je) + sub);
#pragma unroll
                    for (int u = 0; u < U; ++u) q[u] = ldo(pp + LPP * u, jj);
                } else {
#pragma unroll
                    for (int u = 0; u < U; ++u) q[u] = ldo(pp, min((uint32_t)(j + sub + LPP * u), jlast));
                }
                int hit[U];
                bool edge = false;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    // both tests evaluated, combined without a branch
                    const float r2 = r2_fp32(ps.x - q[u].x, ps.y - q[u].y, ps.z - q[u].z);
                    const bool in = r2 < r2cut;
                    hit[u] = (in & (j + sub + LPP * u < je)) ? 1 : 0;
                    // Closer than (1 - eps) cell (so along every axis), two
                    // particles' cells differ by at most one (the FP32
                    // rounding of their cell coordinates is well below eps),
                    // so only hits at that distance or of a lane on the
                    // grid's edge need the exact cell test.  (Bitwise, not
                    // short-circuit: the lane masks combine in scalar
                    // registers.)
                    if constexpr (STALE) edge = edge | (in & (r2 >= g.ext2));
                }
                if constexpr (STALE) {
                    if (__ballot(edge) | edge_lanes) {
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (hit[u] && !in_current_stencil(g, ps, q[u])) hit[u] = 0;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    my[cnt][lane] = j + sub + LPP * u;
                    cnt += hit[u];
                }
                j += S;
                // every later particle of the column lies above the window
                if (q[U - 1].z > zhi) je = min(je, j);
            }
            if (!__ballot(j < je)) break;
            flush_full();
        }
        // flush between columns when another column might not fit (with
        // partial flushes a full list is flushed where it fills instead)
        if (PF == 0 && __ballot(cnt > CAP - 3 * U)) flush();
    }
    flush();
    return count;
}

// Adami wall stress with the list-based scan (material layer only; a wall is
// never in its own candidate stream).  Same sums as k_adami, same order.
// WPB waves per workgroup; PF and ORD as in scan_compact (partial flushes,
// column order).  B = 2 gathers two listed neighbours' data before summing
// either (the same sums in the same order): a wave of walls next to the
// material is bound by its chain of dependent gathers, one round per flush
// iteration.
// str2 (nullable): a second copy of the walls' stress (split integration:
// the step's state beside the half-step state the scans read).
template <class DP, bool STALE = false, int WPB = 4, int PF = 0, int ORD = 0, int B = 1, int U = IC2_U>
__global__ __launch_bounds__(256) void k_adami_lds(DP dp, DevState st, const int32_t* __restrict__ beg,
                                                   const int32_t* __restrict__ end, IndexMap walls, int n,
                                                   StaleIn sn = {}, int32_t* __restrict__ nbr = nullptr,
                                                   tensor2d* __restrict__ str2 = nullptr) {
    __shared__ int32_t lists[WPB][IC2_CAP][64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // round-robin XCDs, as k_adami
    const int k = t < n ? walls(t) : -1;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    bool on = false;
    const RunDesc* dptr = nullptr;
    if (t < n) {
        const RunDesc& d = dp.get(dp.run(k));
        dptr = &d;
        on = (k - d.p_off) < d.npb && d.active;
    }
    float i_h = 0.f, cw = 0.f, m = 0.f, r2cut = 0.f;
    float4 ps = make_float4(0.f, 0.f, 0.f, 0.f);
    double xi = 0.0, yi = 0.0, zi = 0.0;
    ScanGeom g = {};
    if (on) {
        const RunDesc& d = *dptr;
        i_h = d.par.i_h;
        cw = d.par.kernel_normalization_par;
        m = d.par.m;
        r2cut = d.r2_cut;
        ps = st.pospres[k];
        xi = st.xd[k];
        yi = st.yd[k];
        zi = st.zd[k];
        // walls do not move: the table's position is the current one
        g = STALE ? make_geom_stale(d, ps, ps, stale_disp(sn, d, dp.run(k)), stale_out_now(sn, d, dp.run(k)))
                  : make_geom(d, ps);
    }
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
    float px = 0.f, py = 0.f, pz = 0.f, wsum = 0.f;
    // one pair's sums from its gathered data
    auto add = [&](double xj, double yj, double zj, float rhoj, const tensor2d& s, bool act) {
        const float dx = (float)(xi - xj), dy = (float)(yi - yj), dz = (float)(zi - zj);
        const float q = fsqrt(dx * dx + dy * dy + dz * dz) * i_h;
        const float tt = fmaxf(2.0f - q, 0.0f);
        const float t2 = tt * tt;
        const float kv0 = cw * t2 * t2 * (q + 0.5f) * m * frcp(rhoj);
        const float kv = act ? kv0 : 0.0f;
        sxx += s.xx * kv;
        syy += s.yy * kv;
        szz += s.zz * kv;
        sxy += s.xy * kv;
        sxz += s.xz * kv;
        syz += s.yz * kv;
        px += dx * kv;
        py += dy * kv;
        pz += dz * kv;
        wsum += kv;
    };
    int cnt;
    if constexpr (B == 1) {
        cnt = scan_compact<1, DP::padded, IC2_CAP, STALE, U, PF, ORD>(
            on, 1, 1, g, ps, r2cut, beg, end, st.pospres, lists[w], lane, on ? (uint32_t)k : 0u,
            [&](uint32_t j, bool act) {
                const Offs o = offs_of(j);
                add(ldb(st.xd, o.o8), ldb(st.yd, o.o8), ldb(st.zd, o.o8), ldb(st.velrhop, o.o16).w, ldb(st.str, o.o24),
                    act);
            });
    } else {
        cnt = scan_compact<2, DP::padded, IC2_CAP, STALE, U, PF, ORD>(
            on, 1, 1, g, ps, r2cut, beg, end, st.pospres, lists[w], lane, on ? (uint32_t)k : 0u,
            [&](uint32_t j0, bool a0, uint32_t j1, bool a1) {
                const Offs o0 = offs_of(j0), o1 = offs_of(j1);
                const double x0 = ldb(st.xd, o0.o8), y0 = ldb(st.yd, o0.o8), z0 = ldb(st.zd, o0.o8);
                const double x1 = ldb(st.xd, o1.o8), y1 = ldb(st.yd, o1.o8), z1 = ldb(st.zd, o1.o8);
                const float r0 = ldb(st.velrhop, o0.o16).w, r1 = ldb(st.velrhop, o1.o16).w;
                const tensor2d s0 = ldb(st.str, o0.o24), s1 = ldb(st.str, o1.o24);
                add(x0, y0, z0, r0, s0, a0);
                add(x1, y1, z1, r1, s1, a1);
            });
    }
    if (on && nbr) nbr[k] = cnt;
    if (on && wsum > 0.0f) {
        const RunDesc& d = *dptr;
        const float iw = 1.0f / wsum;
        const float rho0 = d.par.rho0;
        tensor2d o;
        o.xx = (sxx - rho0 * d.par.acc_x * px) * iw;
        o.yy = (syy - rho0 * d.par.acc_y * py) * iw;
        o.zz = (szz - rho0 * d.par.acc_z * pz) * iw;
        o.xy = sxy * iw;
        o.xz = sxz * iw;
        o.yz = syz * iw;
        st.str[k] = o;
        if (str2) str2[k] = o;
    }
}

// ------------------------------------------------------ constitutive models
// Drucker-Prager (src/predictor_corrector.cu:30-105): FP32 Jaumann + elastic
// trial, FP64 return mapping (tensile cut-off, then shear return with
// non-associated dilatancy).
__device__ __forceinline__ void mat_dp(const RunDesc& d, float dt, const float4 w, const tensor2d& D, float& p,
                                       float& ep, tensor2d& s) {
    const float K = d.par.model_par_1, G = d.par.model_par_2;
    tensor2d r = jaumann(w, s);
    r = tadd(r, tscale(D, 2.0f * G));
    const float ace = (K - 2.0f / 3.0f * G) * (D.xx + D.yy + D.zz);
    r.xx += ace;
    r.yy += ace;
    r.zz += ace;
    s = tadd(s, tscale(r, dt));

    const double Kd = K, Gd = G, kphi = d.par.model_par_3, kpsi = d.par.model_par_4, kc = d.par.model_par_5;
    double sm = ((double)s.xx + (double)s.yy + (double)s.zz) / 3.0;
    double xx = (double)s.xx - sm, yy = (double)s.yy - sm, zz = (double)s.zz - sm;
    double xy = s.xy, xz = s.xz, yz = s.yz;
    double epd = ep;
    const double tenf = (d.par.model_par_3 == 0.0f) ? 0.0 : kc / kphi;
    const double dpti = sm - tenf;
    if (dpti >= 0.0) {
        epd += dpti / Kd * (1.0 / 3.0) * 1.4142135623730951;
        sm = tenf;
    }
    const double tau = sqrt(0.5 * (xx * xx + yy * yy + zz * zz + 2.0 * (xy * xy + xz * xz + yz * yz)));
    const double dpfi = tau + kphi * sm - kc;
    if (dpfi > 0.0) {
        const double dl = dpfi / (Gd + Kd * kphi * kpsi);
        sm -= Kd * kpsi * dl;
        const double ratio = tau < 1e-10 ? 1.0 : (kc - kphi * sm) / tau;
        xx *= ratio;
        yy *= ratio;
        zz *= ratio;
        xy *= ratio;
        xz *= ratio;
        yz *= ratio;
        epd += dl * sqrt(1.0 / 3.0 + 2.0 / 9.0 * kpsi * kpsi);
    }
    s = tensor2d{(float)(xx + sm), (float)(yy + sm), (float)(zz + sm), (float)xy, (float)xz, (float)yz};
    p = (float)(-sm);
    ep = (float)epd;
}

// Wu-type 4-constant hypoplastic rate (src/predictor_corrector.cu:111-165),
// FP32, including the reference's small-trace guard exactly as written there
// (it resets xy twice and leaves xz, :140-141).
__device__ __forceinline__ void mat_wu(const RunDesc& d, float dt, const tensor2d& D, const float4 w, tensor2d& s) {
    const tensor2d jau = jaumann(w, s);
    const float c1 = d.par.model_par_1, c2 = d.par.model_par_2, c3 = d.par.model_par_3, c4 = d.par.model_par_4,
                coh = d.par.model_par_5;
    tensor2d t = s;
    if (coh > 0.0f) {
        t.xx -= coh;
        t.yy -= coh;
        t.zz -= coh;
    }
    float t1 = t.xx + t.yy + t.zz;
    if (fabsf(t1) < 4.0f) {
        t1 = -4.0f;
        t.xx = -1.0f;
        t.yy = -1.0f;
        t.zz = -2.0f;
        t.xy = 0.0f;
        t.yz = 0.0f;
    }
    tensor2d ts = t;
    ts.xx -= t1 / 3.0f;
    ts.yy -= t1 / 3.0f;
    ts.zz -= t1 / 3.0f;
    const float d1 = D.xx + D.yy + D.zz;
    const float t1d1 = tdot(t, D);
    const float nd = sqrtf(tdot(D, D));
    tensor2d rate = tadd(jau, tscale(D, c1 * t1));
    rate = tadd(rate, tscale(t, c2 * d1 + c3 * t1d1 / t1));
    rate = tadd(rate, tscale(tadd(t, ts), c4 * nd));
    s = tadd(s, tscale(rate, dt));
    const float pres = (s.xx + s.yy + s.zz) / 3.0f;
    if (pres > coh) {
        s.xx -= (pres - coh);
        s.yy -= (pres - coh);
        s.zz -= (pres - coh);
    }
}

// von Wolffersdorff hypoplasticity (constitutive model 3; von Wolffersdorff
// 1996, Mech. Cohes.-Frict. Mater. 1:251-271, with the Bauer compression law
// of Herle & Gudehus 1999).  Tension-positive stress; p = -tr(T)/3 is floored
// at VW_PMIN for the rate evaluation.  Updates stress and void ratio.
constexpr float VW_PMIN = 1.0f;
constexpr float VW_TPSI_CAP = 1.4142135f * (1.0f - 1e-3f);

__device__ __forceinline__ void mat_vw(const RunDesc& d, float dt, const tensor2d& D, const float4 w, tensor2d& s,
                                       float& e) {
    const tensor2d jau = jaumann(w, s);
    const float trT = s.xx + s.yy + s.zz;
    float p;
    tensor2d th;  // T / tr T
    if (trT > -3.0f * VW_PMIN) {
        p = VW_PMIN;
        th = tensor2d{1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 0.f, 0.f, 0.f};
    } else {
        p = -trT / 3.0f;
        th = tscale(s, 1.0f / trT);
    }
    tensor2d ts = th;  // deviator of th
    ts.xx -= 1.0f / 3.0f;
    ts.yy -= 1.0f / 3.0f;
    ts.zz -= 1.0f / 3.0f;
    const float tt = tdot(th, th);
    const float ss = tdot(ts, ts);
    const float tpsi = 1.7320508f * sqrtf(ss);
    float c3t = 0.0f;
    if (ss > 1e-12f) {
        const float det = ts.xx * (ts.yy * ts.zz - ts.yz * ts.yz) - ts.xy * (ts.xy * ts.zz - ts.yz * ts.xz) +
                          ts.xz * (ts.xy * ts.yz - ts.yy * ts.xz);
        c3t = -2.4494897f * 3.0f * det / (ss * sqrtf(ss));
        c3t = fminf(fmaxf(c3t, -1.0f), 1.0f);
    }
    // F with tan(psi) capped just below sqrt(2): a stress without a tensile
    // principal value has tan(psi) <= sqrt(2) (T / tr T is then positive
    // semidefinite with unit trace), and beyond it the denominator
    // 2 + sqrt(2) tan(psi) cos(3 theta) can vanish, which makes F unbounded.
    const float tpf = fminf(tpsi, VW_TPSI_CAP);
    const float farg = tpf * tpf / 8.0f + (2.0f - tpf * tpf) / (2.0f + 1.4142135f * tpf * c3t);
    const float F = sqrtf(fmaxf(farg, 0.0f)) - tpf / (2.0f * 1.4142135f);
    const float a = d.vw_a;
    const float hs = d.vw.hs, nn = d.vw.n;
    const float x = 3.0f * p / hs;
    const float ex = expf(-powf(x, nn));
    const float ei = d.vw.ei0 * ex, ec = d.vw.ec0 * ex, ed = d.vw.ed0 * ex;
    const float fe = powf(ec / e, d.vw.beta);
    const float fd = powf(fmaxf((e - ed) / (ec - ed), 0.0f), d.vw.alpha);
    const float fb = d.vw_fb0 * (1.0f + ei) / ei * powf(x, 1.0f - nn);
    const float fs = fb * fe;
    const float nd = sqrtf(tdot(D, D));
    const float thd = tdot(th, D);
    tensor2d rate = tscale(D, F * F);
    rate = tadd(rate, tscale(th, a * a * thd));
    rate = tadd(rate, tscale(tadd(th, ts), fd * a * F * nd));
    rate = tscale(rate, fs / tt);
    s = tadd(s, tscale(tadd(jau, rate), dt));
    e = e + dt * (1.0f + e) * (D.xx + D.yy + D.zz);
    // A cohesionless material carries no tensile mean stress: such a state
    // returns to zero stress (what the Drucker-Prager tension cut and apex
    // return give for c = 0).  The void ratio stays between the densest and
    // the loosest state at the new pressure, e_d <= e <= e_i.  Both tests
    // let a NaN through, for the corrector's failure check.
    const float trn = s.xx + s.yy + s.zz;
    if (trn >= 0.0f) s = tensor2d{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float exn = expf(-powf(fmaxf(-trn, 0.0f) / hs, nn));
    const float elo = d.vw.ed0 * exn, ehi = d.vw.ei0 * exn;
    e = e < elo ? elo : (e > ehi ? ehi : e);
}

__device__ __forceinline__ void constitutive(const RunDesc& d, float dt, const float4 w, const tensor2d& D,
                                             const tensor2d& stn_new, float& p, float& ep, float& voidr,
                                             tensor2d& s) {
    const int model = d.par.constitutive_model;
    if (model == 1) {
        mat_dp(d, dt, w, D, p, ep, s);
    } else if (model == 2) {
        mat_wu(d, dt, D, w, s);
        ep = sqrtf(2.0f / 3.0f * tdot(stn_new, stn_new));
    } else {
        mat_vw(d, dt, D, w, s, voidr);
        ep = sqrtf(2.0f / 3.0f * tdot(stn_new, stn_new));
        p = -(s.xx + s.yy + s.zz) / 3.0f;
    }
}

// Maximum of v over the workgroup's threads with run >= 0, into the run's
// slot blockIdx % DMAX_SLOTS of field f: red[(run * DMAX_SLOTS + slot) * nf +
// f] (float bits, v >= 0).  One atomic per workgroup when all its
// contributing threads share a run, so that a large run does not serialise
// one atomic per wave on one address.  Every thread of the workgroup (at most
// 256) must call it.
__device__ __forceinline__ void block_max_by_run(int run, float v, uint32_t* red, int nf, int f) {
    __shared__ float wmax[4];
    __shared__ int wrun[4];
    const int slot = (int)(blockIdx.x % DMAX_SLOTS);
    auto at = [&](int r) { return &red[((size_t)r * DMAX_SLOTS + slot) * nf + f]; };
    // wave maximum when the wave's contributing lanes share a run (-1: none;
    // -2: mixed, the lanes then add their own)
    const unsigned long long has = __ballot(run >= 0);
    int wr = -1;
    if (has != 0ull) {
        const int r0 = __shfl(run, __ffsll((long long)has) - 1, 64);
        if (__ballot(run >= 0 && run != r0) == 0ull) {
            for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
            wr = r0;
        } else {
            wr = -2;
            if (run >= 0) atomicMax(at(run), __float_as_uint(v));
        }
    }
    __syncthreads();  // the arrays may hold a previous call's values
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wmax[w] = v;
        wrun[w] = wr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x + 63) / 64;
        int r0 = -1;
        bool same = true;
        for (int i = 0; i < nw; ++i) {
            if (wrun[i] < 0) continue;
            if (r0 < 0) r0 = wrun[i];
            if (wrun[i] != r0) same = false;
        }
        if (same && r0 >= 0) {
            float m = 0.0f;
            for (int i = 0; i < nw; ++i)
                if (wrun[i] == r0) m = fmaxf(m, wmax[i]);
            atomicMax(at(r0), __float_as_uint(m));
        } else {
            for (int i = 0; i < nw; ++i)
                if (wrun[i] >= 0) atomicMax(at(wrun[i]), __float_as_uint(wmax[i]));
        }
    }
}

// Distance of the FP32 position of (x, y, z) from where the cell table put
// the particle (po), inflated far beyond the rounding of the FP32 arithmetic:
// the bound the stale scans widen their reach by.
__device__ __forceinline__ float moved_since_table(float4 po, double x, double y, double z) {
    const float dx = (float)x - po.x, dy = (float)y - po.y, dz = (float)z - po.z;
    return sqrtf(dx * dx + dy * dy + dz * dz) * 1.0001f + 1e-12f;
}

// Predictor (src/predictor_corrector.cu:169-218) fused with the *_ini copy
// that the reference does as five device-to-device copies (:240-246), for
// one material particle; returns how far it moved (rounded up).
// SPLIT (engine): the step's start stays in st, which the corrector then
// reads as its ini set, and the half step's positions, velocity and stress
// go to h (nothing is copied to ini, idep and stn are left as they were).
template <bool SPLIT = false>
__device__ __forceinline__ float predict_one(const RunDesc& d, int k, const DevState& st, const DevIni& ini,
                                             const RateVals& rv, bool& outside, const DevState& h = DevState{}) {
    const float dt = d.par.dt;
    const float hdt = 0.5f * dt;
    const double hdtd = (double)hdt;
    const float4 a = rv.accad;
    float4 vr = st.velrhop[k];
    float4 idep = st.idep[k];
    const tensor2d sig0 = st.str[k];
    const tensor2d stn0 = st.stn[k];
    double x = st.xd[k], y = st.yd[k], z = st.zd[k];
    if constexpr (!SPLIT) {
        ini.xd[k] = x;
        ini.yd[k] = y;
        ini.zd[k] = z;
        ini.velrhop[k] = vr;
        ini.epvr[k] = make_float2(idep.y, idep.z);
        ini.str[k] = sig0;
        ini.stn[k] = stn0;
    }

    vr.x += hdt * a.x;
    vr.y += hdt * a.y;
    vr.z += hdt * a.z;
    vr.w += hdt * a.w;
    x += hdtd * (double)vr.x;
    y += hdtd * (double)vr.y;
    z += hdtd * (double)vr.z;
    const tensor2d D = rv.acestn;
    const tensor2d stn = tadd(stn0, tscale(D, hdt));
    float p = st.pospres[k].w;
    float ep = idep.y, voidr = idep.z;
    tensor2d s = sig0;
    constitutive(d, hdt, rv.spin, D, stn, p, ep, voidr, s);

    const DevState& o = SPLIT ? h : st;
    o.xd[k] = x;
    o.yd[k] = y;
    o.zd[k] = z;
    o.pospres[k] = make_float4((float)x, (float)y, (float)z, p);
    o.velrhop[k] = vr;
    if constexpr (!SPLIT) {
        idep.y = ep;
        idep.z = voidr;
        st.idep[k] = idep;
        st.stn[k] = stn;
    }
    o.str[k] = s;
    const float i_h = d.par.i_h;
    const int cx = cell_coord((float)x, i_h), cy = cell_coord((float)y, i_h), cz = cell_coord((float)z, i_h);
    outside = cx < 0 || cx >= d.par.cell_num_x || cy < 0 || cy >= d.par.cell_num_y || cz < 0 ||
              cz >= d.par.cell_num_z;
    if (st.xo) return moved_since_table(st.xo[k], x, y, z);
    // |x_new - x_old| = hdt |v|, inflated far beyond the rounding of either
    const float vn = sqrtf(vr.x * vr.x + vr.y * vr.y + vr.z * vr.z);
    return hdt * vn * 1.0001f + 1e-12f;
}

// dmax (nullable, DMAX_SLOTS per run): the largest displacement of the run's
// particles, for the corrector half's scans over this half step's (now
// stale) table, folded into slot 0 by k_dmax_fold.  nout (with dmax, per
// run): set when a particle of the run ends outside the grid.
template <class DP, bool SPLIT>
__global__ __launch_bounds__(256) void k_predictor(DP dp, DevState st, DevIni ini, DevRates rt, IndexMap mats,
                                                   int n, uint32_t* __restrict__ dmax, uint32_t* __restrict__ nout,
                                                   DevState h) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int run = -1;
    float disp = 0.0f;
    if (t < n) {
        const int k = mats(t);
        const int r = dp.run(k);
        const RunDesc& d = dp.get(r);
        if ((k - d.p_off) >= d.npb && d.active) {
            run = r;
            bool outside = false;
            disp = predict_one<SPLIT>(d, k, st, ini, load_rates(rt, k), outside, h);
            if (outside && nout) atomicOr(&nout[r], 1u);
        }
    }
    if (dmax) block_max_by_run(run, disp, dmax, 1, 0);
}

// Folds each run's displacement slots into slot 0.
__global__ void k_dmax_fold(uint32_t* __restrict__ dmax, int nruns) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    uint32_t m = 0u;
    for (int i = 0; i < DMAX_SLOTS; ++i) m = max(m, dmax[r * DMAX_SLOTS + i]);
    dmax[r * DMAX_SLOTS] = m;
}

// Corrector (src/predictor_corrector.cu:256-307): midpoint update from ini,
// which is in the order of the state (no neighbour search separates the
// predictor from the corrector, so nothing needs the reference's
// SortArraysAll of the *_ini arrays, src/loquat.cu:200-213).
// fail (nullable, one flag per run) is raised when a particle's new state is
// not finite, so a diverging run can be stopped without touching the others.
// Returns how far the particle is from where its cell table put it (engine
// states; 0 otherwise) and whether it ended outside the grid.
// SPLIT (engine): the predictor left the step's start in st and wrote the
// half step to h, so the ini values are read from st and the half step's
// pressure from h.
template <bool SPLIT = false>
__device__ __forceinline__ float correct_one(const RunDesc& d, int run, int k, const DevState& st, const DevIni& ini,
                                             const RateVals& rv, int32_t* __restrict__ fail,
                                             const uint32_t* __restrict__ dmax, bool& outside,
                                             const DevState& h = DevState{}) {
    const float dt = d.par.dt;
    const double dtd = (double)dt;
    const float4 a = rv.accad;
    float4 idep = st.idep[k];
    const float4 v0 = SPLIT ? st.velrhop[k] : ini.velrhop[k];
    float4 vr;
    vr.x = v0.x + dt * a.x;
    vr.y = v0.y + dt * a.y;
    vr.z = v0.z + dt * a.z;
    vr.w = v0.w + dt * a.w;
    const double x = 0.5 * ((double)vr.x + (double)v0.x) * dtd + (SPLIT ? st.xd[k] : ini.xd[k]);
    const double y = 0.5 * ((double)vr.y + (double)v0.y) * dtd + (SPLIT ? st.yd[k] : ini.yd[k]);
    const double z = 0.5 * ((double)vr.z + (double)v0.z) * dtd + (SPLIT ? st.zd[k] : ini.zd[k]);
    const tensor2d D = rv.acestn;
    const tensor2d stn = tadd(SPLIT ? st.stn[k] : ini.stn[k], tscale(D, dt));
    float p = SPLIT ? h.pospres[k].w : st.pospres[k].w;
    const float2 epvr = SPLIT ? make_float2(idep.y, idep.z) : ini.epvr[k];
    float ep = epvr.x, voidr = epvr.y;
    tensor2d s = SPLIT ? st.str[k] : ini.str[k];
    constitutive(d, dt, rv.spin, D, stn, p, ep, voidr, s);
    if (fail) {
        const float chk = (vr.x + vr.y + vr.z + vr.w) + (s.xx + s.yy + s.zz) + (s.xy + s.xz + s.yz);
        if (!isfinite(chk) || !isfinite(x + y + z)) fail[run] = 1;
        // the corrector half scanned a stale table: its neighbour sets are
        // exact only within the displacement bound
        if (dmax && !(2.0f * __uint_as_float(dmax[run * DMAX_SLOTS]) <= STALE_LIMIT * (2.0f / d.par.i_h)))
            fail[run] = 1;
    }

    st.xd[k] = x;
    st.yd[k] = y;
    st.zd[k] = z;
    st.pospres[k] = make_float4((float)x, (float)y, (float)z, p);
    st.velrhop[k] = vr;
    idep.y = ep;
    idep.z = voidr;
    st.idep[k] = idep;
    st.str[k] = s;
    st.stn[k] = stn;
    const float i_h = d.par.i_h;
    const int cx = cell_coord((float)x, i_h), cy = cell_coord((float)y, i_h), cz = cell_coord((float)z, i_h);
    outside = cx < 0 || cx >= d.par.cell_num_x || cy < 0 || cy >= d.par.cell_num_y || cz < 0 ||
              cz >= d.par.cell_num_z;
    return st.xo ? moved_since_table(st.xo[k], x, y, z) : 0.0f;
}

// dmaxb / noutb (nullable, as dmax / nout of k_predictor): where the step
// leaves each run's particles relative to its cell table, for the next
// step's scans when the table is kept.
template <class DP, bool SPLIT>
__global__ __launch_bounds__(256) void k_corrector(DP dp, DevState st, DevIni ini, DevRates rt, IndexMap mats,
                                                   int n, int32_t* __restrict__ fail,
                                                   const uint32_t* __restrict__ dmax, uint32_t* __restrict__ dmaxb,
                                                   uint32_t* __restrict__ noutb, DevState h) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int run = -1;
    float disp = 0.0f;
    if (t < n) {
        const int k = mats(t);
        const int r = dp.run(k);
        const RunDesc& d = dp.get(r);
        if ((k - d.p_off) >= d.npb && d.active) {
            run = r;
            bool outside = false;
            disp = correct_one<SPLIT>(d, r, k, st, ini, load_rates(rt, k), fail, dmax, outside, h);
            if (outside && noutb) atomicOr(&noutb[r], 1u);
        }
    }
    if (dmaxb) block_max_by_run(run, disp, dmaxb, 1, 0);
}

// Largest value per run over one wave (workgroups of one wave, no shared
// memory), into slot blockIdx % DMAX_SLOTS of red: when the contributing
// lanes share a run the wave's maximum is added once, otherwise each lane
// adds its own.  v >= 0; lanes with run < 0 contribute nothing.  Every lane
// of the wave must call it.
__device__ __forceinline__ void wave_max_by_run(int run, float v, uint32_t* red) {
    const int slot = (int)(blockIdx.x % DMAX_SLOTS);
    const unsigned long long has = __ballot(run >= 0);
    if (has == 0ull) return;
    const int r0 = __shfl(run, __ffsll((long long)has) - 1, 64);
    if (__ballot(run >= 0 && run != r0) == 0ull) {
        if (run < 0) v = 0.0f;
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) atomicMax(&red[(size_t)r0 * DMAX_SLOTS + slot], __float_as_uint(v));
    } else if (run >= 0) {
        atomicMax(&red[(size_t)run * DMAX_SLOTS + slot], __float_as_uint(v));
    }
}

// The integration half step fused into the interaction's epilogue (FUSE 1:
// predictor, 2: corrector; FuseIn): the particle's rates go from registers
// into predict_one / correct_one instead of through the rates arrays.
// Neither integrator writes anything the scan reads (the predictor writes
// only the half step, the corrector only the step's start, each at its own
// particle), so the fused kernel computes what the two kernels compute.
template <int FUSE>
struct FuseArg : FuseIn {};
template <>
struct FuseArg<0> {};

// Variant 3 (B = 1): variant 2's structure with fewer instructions per
// candidate and per pair, pruned scans, and XCD-contiguous blocks.  The
// particle itself stays in its candidate stream (it is always a hit when it
// lies inside the grid): all its pair terms are exactly zero (x_i - x_i = 0
// zeroes the kernel gradient and v_i - v_i the velocity terms), so it changes
// no sum, and one is taken off the neighbour count.  Variant 4 (B = 2)
// evaluates the lists two pairs at a time; same pairs, same order.  ART adds
// the artificial pressure; STALE scans a table built before the particles
// last moved (sn).  Kernels without artificial pressure are held to four
// waves per SIMD (at most 128 VGPRs).
// Shape parameters searched by the autotuner: CAP list entries per lane,
// UU candidates per load round, WPB waves per workgroup (lists are per wave,
// so a workgroup holds WPB x CAP x 64 entries of LDS).  FUSE (one-wave
// workgroups, one lane per particle): the integration fused in (FuseIn),
// nothing written to rt.
template <class DP, int B, bool MASKED = false, bool ART = false, bool STALE = false, int CAP = IC3_CAP,
          int UU = IC2_U, int WPB = 4, int PF = 0, int ORD = 0, int LPP = 1, bool STG = false, int FUSE = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 256), amdgpu_waves_per_eu((STALE || !ART) ? 4 : 1))) void
k_interaction_v3(DP dp, DevState st, DevRates rt, const int32_t* __restrict__ beg, const int32_t* __restrict__ end,
                 IndexMap mats, int n, int32_t* __restrict__ nbr, StaleIn sn = {}, FuseArg<FUSE> fz = {}) {
    static_assert(FUSE == 0 || (WPB == 1 && LPP == 1), "fused integration needs one-wave workgroups");
    __shared__ int32_t lists[WPB][CAP][64];
    // STG: each wave's chunk of staged candidates
    __shared__ float4 staged[STG ? WPB : 1][STG ? STG_SLOTS : 1];
    // LPP lanes per particle: adjacent lanes of one wave
    const int t = xcd_thread() / LPP;
    const int k = t < n ? mats(t) : -1;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    bool mat = false;
    const RunDesc* dptr = nullptr;
    if (t < n) {
        const RunDesc& d = dp.get(dp.run(k));
        dptr = &d;
        mat = (k - d.p_off) >= d.npb && d.active;
    }
    IConst c = {};
    float r2cut = 0.f;
    float4 ps = make_float4(0.f, 0.f, 0.f, 0.f);
    IPart pi = {};
    ScanGeom g = {};
    if (mat) {
        const RunDesc& d = *dptr;
        c = make_iconst(d);
        r2cut = d.r2_cut;
        ps = st.pospres[k];
        load_part(st, k, pi);
        if constexpr (STALE) {
            g = make_geom_stale(d, ps, stale_pos(sn, k), stale_disp(sn, d, dp.run(k)), stale_out_now(sn, d, dp.run(k)));
        } else {
            g = make_geom(d, ps);
        }
    }
    // the particle is its own candidate exactly when its current cell lies
    // inside the grid
    const int ccx = cell_coord(ps.x, c.i_h), ccy = cell_coord(ps.y, c.i_h), ccz = cell_coord(ps.z, c.i_h);
    const bool self_in = mat && ccx >= 0 && ccx < g.nx && ccy >= 0 && ccy < g.ny && ccz >= 0 && ccz < g.nz;
    IAcc a = {};
    // dummy = the particle itself: its pair terms are exactly zero
    int count;
    // MASKED: a lane past the end of its list loads nothing and computes the
    // self pair from registers (positions and velocity of i, zero stress),
    // instead of gathering particle i's data again
    const PairIn self = PairIn{pi.x, pi.y, pi.z, make_float4(pi.vx, pi.vy, pi.vz, pi.rho), tensor2d{}, ps.w};
    if constexpr (ART) pi.R = c.eps * fmaxf(-ps.w, 0.0f) / (pi.rho * pi.rho);
    auto load = [&](uint32_t j, bool act) {
        if constexpr (MASKED) {
            PairIn q = self;
            if (act) q = pair_load<ART>(st, j);
            return q;
        } else {
            return pair_load<ART>(st, j);
        }
    };
    if constexpr (B == 1) {
        count = scan_compact<1, DP::padded, CAP, STALE, UU, PF, ORD, LPP, STG>(mat, 0, 1, g, ps, r2cut, beg, end, st.pospres, lists[w], lane,
                                            mat ? (uint32_t)k : 0u,
                                            [&](uint32_t j, bool act) { pair_compute<ART>(pi, load(j, act), c, a); },
                                            staged[STG ? w : 0]);
    } else {
        count = scan_compact<2, DP::padded, CAP, STALE, UU, PF, ORD, LPP, STG>(mat, 0, 1, g, ps, r2cut, beg, end, st.pospres, lists[w], lane,
                                            mat ? (uint32_t)k : 0u,
                                            [&](uint32_t j0, bool a0, uint32_t j1, bool a1) {
                                                const PairIn q0 = load(j0, a0);
                                                const PairIn q1 = load(j1, a1);
                                                pair_compute<ART>(pi, q0, c, a);
                                                pair_compute<ART>(pi, q1, c, a);
                                            }, staged[STG ? w : 0]);
    }
    if constexpr (LPP == 2) {
        // the particle's two partial sums: each lane adds its neighbour's
        // (lane ^ 1), so both hold the same total
        a.ax += swap_pair(a.ax);
        a.ay += swap_pair(a.ay);
        a.az += swap_pair(a.az);
        a.dvg += swap_pair(a.dvg);
#pragma unroll
        for (int q = 0; q < 9; ++q) a.L[q] += swap_pair(a.L[q]);
        count += __float_as_int(swap_pair(__int_as_float(count)));
    }
    if constexpr (FUSE == 0) {
        if (mat && (LPP == 1 || (lane & 1) == 0)) {
            store_rates(*dptr, rt, k, a);
            if (nbr) nbr[k] = count - (self_in ? 1 : 0);
        }
    } else {
        int run = -1;
        float disp = 0.0f;
        if (mat) {
            const RunDesc& d = *dptr;
            run = dp.run(k);
            bool outside = false;
            const RateVals rv = rates_of(d, a);
            if constexpr (FUSE == 1)
                disp = predict_one<true>(d, k, st, DevIni{}, rv, outside, fz.other);
            else
                disp = correct_one<true>(d, run, k, fz.other, DevIni{}, rv, fz.fail, fz.dmax_pred, outside, st);
            if (outside && fz.nout) atomicOr(&fz.nout[run], 1u);
            if (nbr) nbr[k] = count - (self_in ? 1 : 0);
        }
        if (fz.dmax) wave_max_by_run(run, disp, fz.dmax);
    }
}

// ----------------------------------------------------------------- Shepard
// src/field_smoothing.cu:65-121: material neighbours only, self included.
// A particle with no neighbour at all (only possible outside the grid) keeps
// its stress instead of dividing by zero.
template <class DP>
__global__ __launch_bounds__(256) void k_shepard(DP dp, DevState st, const tensor2d* __restrict__ src,
                                                 tensor2d* __restrict__ dst, const int32_t* __restrict__ beg,
                                                 const int32_t* __restrict__ end, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const RunDesc& d = dp.get(dp.run(k));
    if ((k - d.p_off) < d.npb || !d.active || !d.shepard_due) {
        dst[k] = src[k];
        return;
    }
    const float4 pi = st.pospres[k];
    const double xi = st.xd[k], yi = st.yd[k], zi = st.zd[k];
    const float i_h = d.par.i_h, cw = d.par.kernel_normalization_par, m = d.par.m, r2cut = d.r2_cut;
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f, w = 0.f;
    const int cix = cell_coord(pi.x, i_h), ciy = cell_coord(pi.y, i_h), ciz = cell_coord(pi.z, i_h);
    for_each_column(d, cix, ciy, ciz, 1, 1, beg, end, [&](int jb, int je) {
        for (int j = jb; j < je; ++j) {
            const float4 pj = st.pospres[j];
            if (r2_fp32(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < r2cut) {
                const float dx = (float)(xi - st.xd[j]), dy = (float)(yi - st.yd[j]), dz = (float)(zi - st.zd[j]);
                const float q = fsqrt(dx * dx + dy * dy + dz * dz) * i_h;
                const float t = fmaxf(2.0f - q, 0.0f);
                const float t2 = t * t;
                const float kv = cw * t2 * t2 * (q + 0.5f) * m * frcp(st.velrhop[j].w);
                const tensor2d s = src[j];
                sxx += s.xx * kv;
                syy += s.yy * kv;
                szz += s.zz * kv;
                sxy += s.xy * kv;
                sxz += s.xz * kv;
                syz += s.yz * kv;
                w += kv;
            }
        }
    });
    if (w > 0.0f) {
        const float iw = 1.0f / w;
        dst[k] = tensor2d{sxx * iw, syy * iw, szz * iw, sxy * iw, sxz * iw, syz * iw};
    } else {
        dst[k] = src[k];
    }
}

// Shepard filter with the list-based scan; same sums as k_shepard, same order.
// WPB waves per workgroup; PF (partial flushes), B (gathers per flush round)
// and U (candidates per round) as in scan_compact, none of which changes a
// sum; XCD: XCD-contiguous blocks (otherwise round-robin, which balances the
// walls, who only copy).
template <class DP, int WPB = 4, int PF = 0, int B = 1, int U = IC2_U, bool XCD = false>
__global__ __launch_bounds__(256) void k_shepard_lds(DP dp, DevState st, const tensor2d* __restrict__ src,
                                                     tensor2d* __restrict__ dst, const int32_t* __restrict__ beg,
                                                     const int32_t* __restrict__ end, int n) {
    __shared__ int32_t lists[WPB][IC2_CAP][64];
    const int k = XCD ? xcd_thread() : (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    bool on = false;
    float i_h = 0.f, cw = 0.f, m = 0.f, r2cut = 0.f;
    float4 ps = make_float4(0.f, 0.f, 0.f, 0.f);
    double xi = 0.0, yi = 0.0, zi = 0.0;
    ScanGeom g = {};
    if (k < n) {
        const RunDesc& d = dp.get(dp.run(k));
        on = (k - d.p_off) >= d.npb && d.active && d.shepard_due;
        if (on) {
            i_h = d.par.i_h;
            cw = d.par.kernel_normalization_par;
            m = d.par.m;
            r2cut = d.r2_cut;
            ps = st.pospres[k];
            xi = st.xd[k];
            yi = st.yd[k];
            zi = st.zd[k];
            g = make_geom(d, ps);
        }
    }
    float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f, wsum = 0.f;
    // one pair's sums from its gathered data
    auto add = [&](double xj, double yj, double zj, float rhoj, const tensor2d& s, bool act) {
        const float dx = (float)(xi - xj), dy = (float)(yi - yj), dz = (float)(zi - zj);
        const float q = fsqrt(dx * dx + dy * dy + dz * dz) * i_h;
        const float tt = fmaxf(2.0f - q, 0.0f);
        const float t2 = tt * tt;
        const float kv0 = cw * t2 * t2 * (q + 0.5f) * m * frcp(rhoj);
        const float kv = act ? kv0 : 0.0f;
        sxx += s.xx * kv;
        syy += s.yy * kv;
        szz += s.zz * kv;
        sxy += s.xy * kv;
        sxz += s.xz * kv;
        syz += s.yz * kv;
        wsum += kv;
    };
    if constexpr (B == 1) {
        scan_compact<1, DP::padded, IC2_CAP, false, U, PF>(
            on, 1, 1, g, ps, r2cut, beg, end, st.pospres, lists[w], lane, on ? (uint32_t)k : 0u,
            [&](uint32_t j, bool act) {
                const Offs o = offs_of(j);
                add(ldb(st.xd, o.o8), ldb(st.yd, o.o8), ldb(st.zd, o.o8), ldb(st.velrhop, o.o16).w, ldb(src, o.o24),
                    act);
            });
    } else {
        scan_compact<2, DP::padded, IC2_CAP, false, U, PF>(
            on, 1, 1, g, ps, r2cut, beg, end, st.pospres, lists[w], lane, on ? (uint32_t)k : 0u,
            [&](uint32_t j0, bool a0, uint32_t j1, bool a1) {
                const Offs o0 = offs_of(j0), o1 = offs_of(j1);
                const double x0 = ldb(st.xd, o0.o8), y0 = ldb(st.yd, o0.o8), z0 = ldb(st.zd, o0.o8);
                const double x1 = ldb(st.xd, o1.o8), y1 = ldb(st.yd, o1.o8), z1 = ldb(st.zd, o1.o8);
                const float r0 = ldb(st.velrhop, o0.o16).w, r1 = ldb(st.velrhop, o1.o16).w;
                const tensor2d s0 = ldb(src, o0.o24), s1 = ldb(src, o1.o24);
                add(x0, y0, z0, r0, s0, a0);
                add(x1, y1, z1, r1, s1, a1);
            });
    }
    if (k >= n) return;
    if (on && wsum > 0.0f) {
        const float iw = 1.0f / wsum;
        dst[k] = tensor2d{sxx * iw, syy * iw, szz * iw, sxy * iw, sxz * iw, syz * iw};
    } else {
        dst[k] = src[k];
    }
}

// sort_every > 0: marks the runs whose cell table is rebuilt this step
// (sort_due: sort_every steps since the last build), and stops a run whose
// kept table is further behind its particles than a stale scan covers.
__global__ void k_step_begin(RunDesc* d, int nruns, uint32_t* __restrict__ dmax, uint32_t* __restrict__ nout,
                             int sort_every, int32_t* __restrict__ fail) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    if (nout) nout[r] = 0u;
    if (dmax)
        for (int i = 0; i < DMAX_SLOTS; ++i) dmax[r * DMAX_SLOTS + i] = 0u;
    if (sort_every > 0) {
        d[r].sort_due = (d[r].active && d[r].since_sort >= sort_every) ? 1 : 0;
        if (d[r].active && !d[r].sort_due && !(2.0f * d[r].dtab <= STALE_LIMIT * (2.0f / d[r].par.i_h))) {
            if (fail) fail[r] = 1;
            d[r].active = 0;
        }
    }
    const int f = d[r].par.stress_shepard_filter;
    const int step = d[r].step;
    d[r].shepard_due = (d[r].active && f != 0 && step % f == 0 && step != 0) ? 1 : 0;
}

// dmaxb / noutb (nullable): the corrector's bound on how far each run's
// particles are from its cell table, folded into the run's dtab / out_tab
// for the next step (and the slots cleared); grid_ran / shep_ran: this step
// rebuilt the table of the runs with sort_due / shepard_due.
__global__ void k_step_end(RunDesc* d, int nruns, const int32_t* __restrict__ fail, uint32_t* __restrict__ dmaxb,
                           uint32_t* __restrict__ noutb, int grid_ran, int shep_ran) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    uint32_t moved = 0u, out = 0u;
    if (dmaxb) {
        for (int i = 0; i < DMAX_SLOTS; ++i) {
            moved = max(moved, dmaxb[r * DMAX_SLOTS + i]);
            dmaxb[r * DMAX_SLOTS + i] = 0u;
        }
        out = noutb[r];
        noutb[r] = 0u;
    }
    if (!d[r].active) return;
    if (fail && fail[r]) {
        d[r].active = 0;
        return;
    }
    if (dmaxb) {
        // a Shepard filter re-sorted the run after the corrector: its table
        // is that of the current positions
        const bool shep_sorted = shep_ran && d[r].shepard_due;
        d[r].dtab = shep_sorted ? 0.0f : __uint_as_float(moved);
        d[r].out_tab = out != 0u ? 1 : 0;
        d[r].since_sort = shep_sorted ? 0 : ((grid_ran && d[r].sort_due) ? 1 : d[r].since_sort + 1);
    }
    d[r].time += (double)d[r].par.dt;
    d[r].step += 1;
    if (d[r].step >= d[r].nsteps) d[r].active = 0;
    // the reference's loop condition, src/loquat.cu:322
    if (d[r].until_time && d[r].time >= (double)d[r].par.time_max) d[r].active = 0;
}

// Dynamic time step, part 1: per-run maxima of |v| over the material
// particles of the current state and of |a| over the rates of the previous
// step's corrector (ignored at a run's first step, when there are none).
// Non-negative floats order as their bit patterns, so the maxima are
// unsigned atomics: one per wave when the wave lies in one run.
__global__ __launch_bounds__(256) void k_dt_reduce(const RunDesc* __restrict__ d, const int32_t* __restrict__ run_of,
                                                   const float4* __restrict__ velrhop,
                                                   const float4* __restrict__ accad, IndexMap mats, int n,
                                                   uint32_t* __restrict__ red) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int r = -1;
    float vm = 0.0f, am = 0.0f;
    if (t < n) {
        const int k = mats(t);
        r = run_of[k];
        const RunDesc& dr = d[r];
        if (dr.active && dr.par.dynamic_time_step == 1) {
            const float4 v = velrhop[k];
            vm = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
            if (dr.step > 0) {
                const float4 acc = accad[k];
                am = sqrtf(acc.x * acc.x + acc.y * acc.y + acc.z * acc.z);
            }
        } else {
            r = -1;
        }
    }
    block_max_by_run(r, vm, red, 2, 0);
    block_max_by_run(r, am, red, 2, 1);
}

// Dynamic time step, part 2: dt = cfl * min(h / (cs + |v|max), sqrt(h / |a|max))
// for each run with dynamic-time-step = 1 (no term for |a|max = 0), then
// clears the maxima for the next step.
__global__ void k_dt_update(RunDesc* d, int nruns, uint32_t* __restrict__ red) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    RunDesc& dr = d[r];
    uint32_t vb = 0u, ab = 0u;
    for (int i = 0; i < DMAX_SLOTS; ++i) {
        uint32_t* sl = &red[((size_t)r * DMAX_SLOTS + i) * 2];
        vb = max(vb, sl[0]);
        ab = max(ab, sl[1]);
        sl[0] = 0u;
        sl[1] = 0u;
    }
    if (dr.active && dr.par.dynamic_time_step == 1) {
        const float vmax = __uint_as_float(vb), amax = __uint_as_float(ab);
        const float h = dr.par.h;
        float dt = h / (dr.par.cs + vmax);
        if (amax > 0.0f) dt = fminf(dt, sqrtf(h / amax));
        dr.par.dt = dr.par.cfl * dt;
    }
}

__global__ void k_copy_tensor(const tensor2d* __restrict__ src, tensor2d* __restrict__ dst, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

}  // namespace

// One run: walls are [0, npb), material particles [npb, np).
// Variants 3 and up use the list-based neighbour scan, the others the direct one.
template <class DP>
static void launch_adami(const DP& dp, const DevState& st, const int32_t* beg, const int32_t* end, IndexMap walls,
                         int n, int variant, const StaleIn* stale, int32_t* nbr, hipStream_t s,
                         tensor2d* str2 = nullptr) {
    if (n <= 0) return;
    if (variant == 19) {
        // variant 18's shape gathering two pairs at a time (same sums)
        if (stale)
            k_adami_lds<DP, true, 1, 16, 2, 2, 8>
                <<<blocks_for(n, 64), 64, 0, s>>>(dp, st, beg, end, walls, n, *stale, nbr, str2);
        else
            k_adami_lds<DP, false, 1, 16, 2, 2, 8>
                <<<blocks_for(n, 64), 64, 0, s>>>(dp, st, beg, end, walls, n, StaleIn{}, nbr, str2);
        return;
    }
    if (variant >= 12) {
        // variants 12 to 18: partial flushes, the edge-first column order
        // and one wave per workgroup
        if (stale)
            k_adami_lds<DP, true, 1, 16, 2><<<blocks_for(n, 64), 64, 0, s>>>(dp, st, beg, end, walls, n, *stale, nbr);
        else
            k_adami_lds<DP, false, 1, 16, 2><<<blocks_for(n, 64), 64, 0, s>>>(dp, st, beg, end, walls, n, StaleIn{}, nbr);
        return;
    }
    if (stale)
        k_adami_lds<DP, true><<<blocks_for(n, 256), 256, 0, s>>>(dp, st, beg, end, walls, n, *stale, nbr);
    else if (variant >= 3)
        k_adami_lds<DP><<<blocks_for(n, 256), 256, 0, s>>>(dp, st, beg, end, walls, n, StaleIn{}, nbr);
    else
        k_adami<DP><<<blocks_for(n, 256), 256, 0, s>>>(dp, st, beg, end, walls, n);
}
void launch_adami_one(const RunDesc& d, const DevState& st, const int32_t* beg, const int32_t* end, int variant,
                      const StaleIn* stale, int32_t* nbr, hipStream_t s) {
    launch_adami(DescOne{d}, st, beg, end, IndexMap{nullptr, 0}, d.npb, variant, stale, nbr, s);
}
void launch_adami_table(const RunDesc* d, const int32_t* run_of, const DevState& st, const int32_t* beg,
                        const int32_t* end, IndexMap walls, int nwalls, int variant, const StaleIn* stale,
                        hipStream_t s, tensor2d* str2) {
    launch_adami(DescTable{d, run_of}, st, beg, end, walls, nwalls, variant, stale, nullptr, s, str2);
}

// Variant 11's shape (64-thread workgroups, 40-entry lists, four candidates
// per round) with partial flushes of PF entries and column order ORD.
// F: the fused-integration kernels are built (fz non-null selects them, the
// predictor's or, with corr, the corrector's; runs without artificial
// pressure only).
template <int PF, int ORD, int LPP = 1, int P = IC3_CAP, bool STG = false, bool M = true, bool F = false, class DP>
static void launch_v11_shape(const DP& dp, const DevState& st, const DevRates& rt, const int32_t* beg,
                             const int32_t* end, IndexMap mats, int n, int32_t* nbr, bool art, const StaleIn* stale,
                             hipStream_t s, const FuseIn* fz = nullptr, bool corr = false) {
    constexpr int U = IC2_U;
    const int nb = blocks_for((int64_t)LPP * n, 64);
    if constexpr (F) {
        if (fz) {
            const StaleIn sn = stale ? *stale : StaleIn{};
            if (stale && corr)
                k_interaction_v3<DP, 2, M, false, true, P, U, 1, PF, ORD, LPP, STG, 2>
                    <<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, sn, FuseArg<2>{*fz});
            else if (corr)
                k_interaction_v3<DP, 2, M, false, false, P, U, 1, PF, ORD, LPP, STG, 2>
                    <<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, sn, FuseArg<2>{*fz});
            else if (stale)
                k_interaction_v3<DP, 2, M, false, true, P, U, 1, PF, ORD, LPP, STG, 1>
                    <<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, sn, FuseArg<1>{*fz});
            else
                k_interaction_v3<DP, 2, M, false, false, P, U, 1, PF, ORD, LPP, STG, 1>
                    <<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, sn, FuseArg<1>{*fz});
            return;
        }
    }
    if (stale && art)
        k_interaction_v3<DP, 2, M, true, true, P, U, 1, PF, ORD, LPP, STG><<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (stale)
        k_interaction_v3<DP, 2, M, false, true, P, U, 1, PF, ORD, LPP, STG><<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (art)
        k_interaction_v3<DP, 2, M, true, false, P, U, 1, PF, ORD, LPP, STG><<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else
        k_interaction_v3<DP, 2, M, false, false, P, U, 1, PF, ORD, LPP, STG><<<nb, 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
}

template <class DP>
static void launch_interaction(const DP& dp, const DevState& st, const DevRates& rt, const int32_t* beg,
                               const int32_t* end, IndexMap mats, int n, int32_t* nbr, int variant, bool art,
                               const StaleIn* stale, hipStream_t s, const FuseIn* fz = nullptr, bool corr = false) {
    if (n <= 0) return;
    const int nb = blocks_for(n, 256);
    // variants 12 to 19: variant 11 with partial flushes of PF entries and
    // the stencil columns in order ORD (see scan_compact); 16 as 13 with two
    // lanes per particle (32 particles per wave); 17 as 13 with the
    // candidates staged in LDS (32-entry lists, to keep sixteen workgroups
    // per CU); 18 and 19 as 13 and 17 with unmasked gathers, 19 staged only
    // on launches of at least STG_MIN_LAUNCH particles
    switch (variant) {
        case 12: return launch_v11_shape<12, 1>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 13: return launch_v11_shape<16, 2>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 14: return launch_v11_shape<8, 2>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 15: return launch_v11_shape<0, 2>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 16: return launch_v11_shape<16, 2, 2>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 17: return launch_v11_shape<16, 2, 1, 32, true>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 18: return launch_v11_shape<16, 2, 1, IC3_CAP, false, false>(dp, st, rt, beg, end, mats, n, nbr, art, stale, s);
        case 19:
            // 18 and 19 give the same bits: staged only where it pays, on
            // launches large enough to fill the GPU many times over (a small
            // launch's time is one wave's latency, which staging lengthens)
            if (n >= STG_MIN_LAUNCH)
                return launch_v11_shape<16, 2, 1, 32, true, false, true>(dp, st, rt, beg, end, mats, n, nbr, art,
                                                                         stale, s, fz, corr);
            return launch_v11_shape<16, 2, 1, IC3_CAP, false, false, true>(dp, st, rt, beg, end, mats, n, nbr, art,
                                                                           stale, s, fz, corr);
        default: break;
    }
    // stale tables: variant 6's shape, in the workgroup size of variants 10
    // and 11 (12 to 15 above)
    if (stale && art && variant == 11)
        k_interaction_v3<DP, 2, true, true, true, IC3_CAP, IC2_U, 1>
            <<<blocks_for(n, 64), 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (stale && variant == 11)
        k_interaction_v3<DP, 2, true, false, true, IC3_CAP, IC2_U, 1>
            <<<blocks_for(n, 64), 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (stale && art && variant == 10)
        k_interaction_v3<DP, 2, true, true, true, IC3_CAP, IC2_U, 2>
            <<<blocks_for(n, 128), 128, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (stale && variant == 10)
        k_interaction_v3<DP, 2, true, false, true, IC3_CAP, IC2_U, 2>
            <<<blocks_for(n, 128), 128, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (stale && art)
        k_interaction_v3<DP, 2, true, true, true><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (stale)
        k_interaction_v3<DP, 2, true, false, true><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr, *stale);
    else if (art && variant == 11)  // artificial pressure: variant 6's shape carries the extra term
        k_interaction_v3<DP, 2, true, true, false, IC3_CAP, IC2_U, 1>
            <<<blocks_for(n, 64), 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (art)
        k_interaction_v3<DP, 2, true, true><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 0)
        k_interaction_direct<DP><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 1)
        k_interaction_lds<DP><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 2)
        k_interaction_unrolled<DP><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 3)
        k_interaction_v3<DP, 1><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 4)
        k_interaction_v3<DP, 2><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 5)
        k_interaction_v3<DP, 1, true><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 7)  // variant 6 with 32-entry lists
        k_interaction_v3<DP, 2, true, false, false, 32><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 8)  // 48-entry lists (three workgroups per CU)
        k_interaction_v3<DP, 2, true, false, false, 48><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 9)  // eight candidates per load round
        k_interaction_v3<DP, 2, true, false, false, IC3_CAP, 8><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 10)  // 128-thread workgroups
        k_interaction_v3<DP, 2, true, false, false, IC3_CAP, IC2_U, 2>
            <<<blocks_for(n, 128), 128, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else if (variant == 11)  // 64-thread workgroups
        k_interaction_v3<DP, 2, true, false, false, IC3_CAP, IC2_U, 1>
            <<<blocks_for(n, 64), 64, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
    else
        k_interaction_v3<DP, 2, true><<<nb, 256, 0, s>>>(dp, st, rt, beg, end, mats, n, nbr);
}

void launch_interaction_one(const RunDesc& d, const DevState& st, const DevRates& rt, const int32_t* beg,
                            const int32_t* end, int32_t* nbr, int variant, const StaleIn* stale, hipStream_t s) {
    launch_interaction(DescOne{d}, st, rt, beg, end, IndexMap{nullptr, d.npb}, d.np - d.npb, nbr, variant,
                       d.par.artificial_pressure == 1, stale, s);
}
void launch_interaction_table(const RunDesc* d, const int32_t* run_of, const DevState& st, const DevRates& rt,
                              const int32_t* beg, const int32_t* end, IndexMap mats, int nmat, int nart,
                              int32_t* nbr, int variant, const StaleIn* stale, hipStream_t s, const FuseIn* fuse,
                              bool corr) {
    const DescTable dt{d, run_of};
    launch_interaction(dt, st, rt, beg, end, mats, nmat - nart, nbr, variant, false, stale, s, fuse, corr);
    if (nart > 0)
        launch_interaction(dt, st, rt, beg, end, IndexMap{mats.idx + (nmat - nart), 0}, nart, nbr, variant, true,
                           stale, s);
}

void launch_integrate_one(bool corrector, const RunDesc& d, const DevState& st, const DevIni& ini,
                          const DevRates& rt, hipStream_t s) {
    const int n = d.np - d.npb;
    if (n <= 0) return;
    const IndexMap mats{nullptr, d.npb};
    if (corrector)
        k_corrector<DescOne, false><<<blocks_for(n, 256), 256, 0, s>>>(DescOne{d}, st, ini, rt, mats, n, nullptr,
                                                                       nullptr, nullptr, nullptr, DevState{});
    else
        k_predictor<DescOne, false><<<blocks_for(n, 256), 256, 0, s>>>(DescOne{d}, st, ini, rt, mats, n, nullptr,
                                                                       nullptr, DevState{});
}
void launch_integrate_table(bool corrector, const RunDesc* d, const int32_t* run_of, const DevState& st,
                            const DevIni& ini, const DevRates& rt, IndexMap mats, int nmat, int32_t* fail,
                            uint32_t* dmax, uint32_t* nout, hipStream_t s, uint32_t* dmaxb, uint32_t* noutb,
                            const DevState* half) {
    if (nmat <= 0) return;
    const int nb = blocks_for(nmat, 256);
    const DescTable dp{d, run_of};
    if (corrector) {
        if (half)
            k_corrector<DescTable, true><<<nb, 256, 0, s>>>(dp, st, ini, rt, mats, nmat, fail, dmax, dmaxb, noutb,
                                                            *half);
        else
            k_corrector<DescTable, false><<<nb, 256, 0, s>>>(dp, st, ini, rt, mats, nmat, fail, dmax, dmaxb, noutb,
                                                             DevState{});
    } else {
        if (half)
            k_predictor<DescTable, true><<<nb, 256, 0, s>>>(dp, st, ini, rt, mats, nmat, dmax, nout, *half);
        else
            k_predictor<DescTable, false><<<nb, 256, 0, s>>>(dp, st, ini, rt, mats, nmat, dmax, nout, DevState{});
    }
}

void launch_dmax_fold(uint32_t* dmax, int nruns, hipStream_t s) {
    if (nruns > 0) k_dmax_fold<<<blocks_for(nruns, 64), 64, 0, s>>>(dmax, nruns);
}

template <class DP>
static void launch_shepard(const DP& dp, const DevState& st, const tensor2d* str_src, tensor2d* str_dst,
                           const int32_t* beg, const int32_t* end, int n, int variant, hipStream_t s) {
    if (n <= 0) return;
    if (variant == 19)
        // one-wave workgroups, partial flushes, two gathers per round and
        // eight candidates per round (the sums of variant 18's shape); blocks
        // round-robin over the XCDs (XCD-contiguous blocks left the XCDs
        // holding walls idle: 0.041 against 0.029 ms per C2 step)
        k_shepard_lds<DP, 1, 16, 2, 8, false><<<blocks_for(n, 64), 64, 0, s>>>(dp, st, str_src, str_dst, beg, end, n);
    else if (variant >= 3)
        k_shepard_lds<DP><<<blocks_for(n, 256), 256, 0, s>>>(dp, st, str_src, str_dst, beg, end, n);
    else
        k_shepard<DP><<<blocks_for(n, 256), 256, 0, s>>>(dp, st, str_src, str_dst, beg, end, n);
}
void launch_shepard_one(const RunDesc& d, const DevState& st, const tensor2d* str_src, tensor2d* str_dst,
                        const int32_t* beg, const int32_t* end, int n, int variant, hipStream_t s) {
    launch_shepard(DescOne{d}, st, str_src, str_dst, beg, end, n, variant, s);
}
void launch_shepard_table(const RunDesc* d, const int32_t* run_of, const DevState& st, const tensor2d* str_src,
                          tensor2d* str_dst, const int32_t* beg, const int32_t* end, int n, int variant,
                          hipStream_t s) {
    launch_shepard(DescTable{d, run_of}, st, str_src, str_dst, beg, end, n, variant, s);
}

void launch_step_begin(RunDesc* d, int nruns, uint32_t* dmax, uint32_t* nout, hipStream_t s, int sort_every,
                       int32_t* fail) {
    if (nruns > 0) k_step_begin<<<blocks_for(nruns, 64), 64, 0, s>>>(d, nruns, dmax, nout, sort_every, fail);
}
void launch_dynamic_dt(RunDesc* d, const int32_t* run_of, const DevState& st, const DevRates& rt, IndexMap mats,
                       int nmat, int nruns, uint32_t* red, hipStream_t s) {
    if (nmat > 0) k_dt_reduce<<<blocks_for(nmat, 256), 256, 0, s>>>(d, run_of, st.velrhop, rt.accad, mats, nmat, red);
    if (nruns > 0) k_dt_update<<<blocks_for(nruns, 64), 64, 0, s>>>(d, nruns, red);
}
void launch_step_end(RunDesc* d, int nruns, const int32_t* fail, hipStream_t s, uint32_t* dmaxb, uint32_t* noutb,
                     bool grid_ran, bool shep_ran) {
    if (nruns > 0)
        k_step_end<<<blocks_for(nruns, 64), 64, 0, s>>>(d, nruns, fail, dmaxb, noutb, grid_ran ? 1 : 0,
                                                        shep_ran ? 1 : 0);
}
void launch_copy_tensor(const tensor2d* src, tensor2d* dst, int n, hipStream_t s) {
    if (n > 0) k_copy_tensor<<<blocks_for(n, 256), 256, 0, s>>>(src, dst, n);
}

}  // namespace lq
