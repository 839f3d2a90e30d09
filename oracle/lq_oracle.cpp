// lq_oracle.cpp -- FP64 CPU restatement of LOQUAT's SPH time step.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it.
// The product library never links or calls it.
//
// Every stage follows the reference's per-particle loops, in FP64 arithmetic
// on FP64 state:
//   cell hash / sort / table   src/loquat_grid.h:28-66 (body not shipped) as
//                              pinned by src/adami_boundary.cu:80-106 and
//                              src/loquat.cu:119-124; stable ascending sort
//                              (cub SortPairs contract,
//                              include/cub/device/device_radix_sort.cuh:125-175)
//   permute                    src/loquat.cu:190-213
//   Adami                      src/adami_boundary.cu:27-123
//   interaction                src/particle_interaction.h:29-41 (body not
//                              shipped; formulation fixed by its consumers,
//                              src/predictor_corrector.cu:40-53,186-211)
//   predictor / corrector      src/predictor_corrector.cu:169-218, 256-307
//   Drucker-Prager             src/predictor_corrector.cu:30-105
//   Wu hypoplastic             src/predictor_corrector.cu:111-165
//   Shepard                    src/field_smoothing.cu:27-121, 144
//   time loop                  src/loquat.cu:322-409
// Two choices are shared with the HIP path because they define identity, not
// accuracy: the neighbour test is the FP32 predicate sqrtf(r2) * i_h < 2 on the
// FP32 roundings of the positions, and the cell hash is computed in FP32.
// Pair terms use the FP64 position differences.
//
// Build: make -C oracle (g++ -O2 -fopenmp -ffp-contract=off -mfma).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/loquat_amd.h"

extern "C" {

typedef struct lqo_state {
    double *x, *y, *z, *p;
    double *vx, *vy, *vz, *rho;
    double *id, *ep, *voidr;
    double* s[6];  // stress xx yy zz xy xz yz
    double* e[6];  // total strain
    int32_t* orig;
} lqo_state;

typedef struct lqo_ini {
    double *x, *y, *z;
    double *vx, *vy, *vz, *rho;
    double *ep, *voidr;
    double* s[6];
    double* e[6];
} lqo_ini;

typedef struct lqo_rates {
    double *ax, *ay, *az, *drho;
    double *wx, *wy, *wz;  // spin (W_yx, W_zx, W_zy)
    double* d[6];          // strain rate
} lqo_rates;

}  // extern "C"

namespace {

struct T6 {
    double xx, yy, zz, xy, xz, yz;
};

inline T6 load6(double* const a[6], int i) { return T6{a[0][i], a[1][i], a[2][i], a[3][i], a[4][i], a[5][i]}; }
inline void store6(double* const a[6], int i, const T6& t) {
    a[0][i] = t.xx;
    a[1][i] = t.yy;
    a[2][i] = t.zz;
    a[3][i] = t.xy;
    a[4][i] = t.xz;
    a[5][i] = t.yz;
}
inline T6 add(const T6& a, const T6& b) {
    return T6{a.xx + b.xx, a.yy + b.yy, a.zz + b.zz, a.xy + b.xy, a.xz + b.xz, a.yz + b.yz};
}
inline T6 mul(const T6& a, double s) { return T6{a.xx * s, a.yy * s, a.zz * s, a.xy * s, a.xz * s, a.yz * s}; }
// J2 of include/types.h:77-85 (off-diagonals counted twice).
inline double j2(const T6& a, const T6& b) {
    return a.xx * b.xx + a.yy * b.yy + a.zz * b.zz + 2.0 * a.xy * b.xy + 2.0 * a.xz * b.xz + 2.0 * a.yz * b.yz;
}

// Cell coordinate, FP32: floorf(0.5f * x * i_h) (src/adami_boundary.cu:80-82),
// clamped to [-2, 2^30] (NaN -> -2).
inline int cell_coord(float x, float i_h) {
    float f = floorf((0.5f * x) * i_h);
    if (!(f >= -2.0f)) return -2;
    if (f > 1073741824.0f) return 1073741824;
    return (int)f;
}

// Neighbour predicate q = r / h < 2 (src/adami_boundary.cu:41-43), evaluated
// in FP32 on the FP32 position roundings: r2 = fma(dz, dz, fma(dy, dy, dx dx))
// with dx dx rounded, sqrt correctly rounded.
inline bool is_nbr(float xi, float yi, float zi, float xj, float yj, float zj, float i_h) {
    const float dx = xi - xj, dy = yi - yj, dz = zi - zj;
    const float r2 = std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
    return sqrtf(r2) * i_h < 2.0f;
}

struct Grid {
    int nx, ny, nz, nol;
};

inline Grid grid_of(const lq_params* p) {
    return Grid{p->cell_num_x, p->cell_num_y, p->cell_num_z, p->ncell_one_layer};
}

// Visits the cells of the 3x3x3 stencil of layer `layer`, reference loop
// order cz, cy, cx (src/adami_boundary.cu:84-110).
template <class F>
inline void stencil(const Grid& g, int cix, int ciy, int ciz, int layer, const int32_t* beg, const int32_t* end,
                    F&& f) {
    for (int cz = -1; cz <= 1; ++cz)
        for (int cy = -1; cy <= 1; ++cy)
            for (int cx = -1; cx <= 1; ++cx) {
                const int x = cix + cx, y = ciy + cy, z = ciz + cz;
                if (x < 0 || x > g.nx - 1 || y < 0 || y > g.ny - 1 || z < 0 || z > g.nz - 1) continue;
                const int c = x * g.ny * g.nz + y * g.nz + z + layer * g.nol;
                const int b = beg[c];
                if (b != -1) f(b, end[c]);
            }
}

inline T6 jaumann(double wx, double wy, double wz, const T6& s) {
    T6 r;
    r.xx = -2.0 * (wx * s.xy + wy * s.xz);
    r.yy = 2.0 * (wx * s.xy - wz * s.yz);
    r.zz = 2.0 * (wy * s.xz + wz * s.yz);
    r.xy = wx * (s.xx - s.yy) - wz * s.xz - wy * s.yz;
    r.xz = wy * (s.xx - s.zz) + wz * s.xy - wx * s.yz;
    r.yz = wz * (s.yy - s.zz) + wx * s.xz + wy * s.xy;
    return r;
}

// src/predictor_corrector.cu:30-105 in FP64.
void mat_dp(const lq_params* par, double dt, double wx, double wy, double wz, const T6& D, double& p, double& ep,
            T6& s) {
    const double K = par->model_par_1, G = par->model_par_2, kphi = par->model_par_3, kpsi = par->model_par_4,
                 kc = par->model_par_5;
    T6 r = jaumann(wx, wy, wz, s);
    r = add(r, mul(D, 2.0 * G));
    const double ace = (K - 2.0 / 3.0 * G) * (D.xx + D.yy + D.zz);
    r.xx += ace;
    r.yy += ace;
    r.zz += ace;
    s = add(s, mul(r, dt));
    double sm = (s.xx + s.yy + s.zz) / 3.0;
    s.xx -= sm;
    s.yy -= sm;
    s.zz -= sm;
    const double tenf = (par->model_par_3 == 0.0f) ? 0.0 : kc / kphi;
    const double dpti = sm - tenf;
    if (dpti >= 0.0) {
        ep += dpti / K * (1.0 / 3.0) * std::sqrt(2.0);
        sm = tenf;
    }
    const double tau = std::sqrt(0.5 * j2(s, s));
    const double dpfi = tau + kphi * sm - kc;
    if (dpfi > 0.0) {
        const double dl = dpfi / (G + K * kphi * kpsi);
        sm -= K * kpsi * dl;
        const double ratio = tau < 1e-10 ? 1.0 : (kc - kphi * sm) / tau;
        s = mul(s, ratio);
        ep += dl * std::sqrt(1.0 / 3.0 + 2.0 / 9.0 * kpsi * kpsi);
    }
    s.xx += sm;
    s.yy += sm;
    s.zz += sm;
    p = -sm;
}

// src/predictor_corrector.cu:111-165 in FP64, guard kept as written there
// (xy reset twice, xz kept, :140-141).
void mat_wu(const lq_params* par, double dt, const T6& D, double wx, double wy, double wz, T6& s) {
    const T6 jau = jaumann(wx, wy, wz, s);
    const double c1 = par->model_par_1, c2 = par->model_par_2, c3 = par->model_par_3, c4 = par->model_par_4,
                 coh = par->model_par_5;
    T6 t = s;
    if (par->model_par_5 > 0.0f) {
        t.xx -= coh;
        t.yy -= coh;
        t.zz -= coh;
    }
    double t1 = t.xx + t.yy + t.zz;
    if (std::fabs(t1) < 4.0) {
        t1 = -4.0;
        t.xx = -1.0;
        t.yy = -1.0;
        t.zz = -2.0;
        t.xy = 0.0;
        t.yz = 0.0;
    }
    T6 ts = t;
    ts.xx -= t1 / 3.0;
    ts.yy -= t1 / 3.0;
    ts.zz -= t1 / 3.0;
    const double d1 = D.xx + D.yy + D.zz;
    const double t1d1 = j2(t, D);
    const double nd = std::sqrt(j2(D, D));
    T6 rate = add(jau, mul(D, c1 * t1));
    rate = add(rate, mul(t, c2 * d1 + c3 * t1d1 / t1));
    rate = add(rate, mul(add(t, ts), c4 * nd));
    s = add(s, mul(rate, dt));
    const double pres = (s.xx + s.yy + s.zz) / 3.0;
    if (pres > coh) {
        s.xx -= (pres - coh);
        s.yy -= (pres - coh);
        s.zz -= (pres - coh);
    }
}

// von Wolffersdorff (1996) hypoplasticity with the Bauer compression law,
// FP64 (constitutive model 3; no reference counterpart).  p floored at 1 Pa.
void mat_vw(const lq_vw_params* vw, double dt, const T6& D, double wx, double wy, double wz, T6& s, double& e) {
    const double pi = 3.14159265358979323846;
    const T6 jau = jaumann(wx, wy, wz, s);
    const double trT = s.xx + s.yy + s.zz;
    double p;
    T6 th;
    if (trT > -3.0) {
        p = 1.0;
        th = T6{1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0, 0, 0};
    } else {
        p = -trT / 3.0;
        th = mul(s, 1.0 / trT);
    }
    T6 ts = th;
    ts.xx -= 1.0 / 3.0;
    ts.yy -= 1.0 / 3.0;
    ts.zz -= 1.0 / 3.0;
    const double tt = j2(th, th), ss = j2(ts, ts);
    const double tpsi = std::sqrt(3.0) * std::sqrt(ss);
    double c3t = 0.0;
    if (ss > 1e-12) {
        const double det = ts.xx * (ts.yy * ts.zz - ts.yz * ts.yz) - ts.xy * (ts.xy * ts.zz - ts.yz * ts.xz) +
                           ts.xz * (ts.xy * ts.yz - ts.yy * ts.xz);
        c3t = -std::sqrt(6.0) * 3.0 * det / (ss * std::sqrt(ss));
        c3t = std::min(1.0, std::max(-1.0, c3t));
    }
    // tan(psi) capped just below sqrt(2), the bound of stresses without a
    // tensile principal value, where the denominator can vanish
    const double tpf = std::min(tpsi, (double)(1.4142135f * (1.0f - 1e-3f)));
    const double farg = tpf * tpf / 8.0 + (2.0 - tpf * tpf) / (2.0 + std::sqrt(2.0) * tpf * c3t);
    const double F = std::sqrt(std::max(farg, 0.0)) - tpf / (2.0 * std::sqrt(2.0));
    const double sp = std::sin((double)vw->phi_c * pi / 180.0);
    const double a = std::sqrt(3.0) * (3.0 - sp) / (2.0 * std::sqrt(2.0) * sp);
    const double hs = vw->hs, n = vw->n;
    const double x = 3.0 * p / hs;
    const double ex = std::exp(-std::pow(x, n));
    const double ei = vw->ei0 * ex, ec = vw->ec0 * ex, ed = vw->ed0 * ex;
    const double fe = std::pow(ec / e, (double)vw->beta);
    const double fd = std::pow(std::max((e - ed) / (ec - ed), 0.0), (double)vw->alpha);
    const double ratio = ((double)vw->ei0 - vw->ed0) / ((double)vw->ec0 - vw->ed0);
    const double den = 3.0 + a * a - a * std::sqrt(3.0) * std::pow(ratio, (double)vw->alpha);
    const double fb = hs / n * std::pow((double)vw->ei0 / vw->ec0, (double)vw->beta) / den * (1.0 + ei) / ei *
                      std::pow(x, 1.0 - n);
    const double fs = fb * fe;
    const double nd = std::sqrt(j2(D, D));
    const double thd = j2(th, D);
    T6 rate = mul(D, F * F);
    rate = add(rate, mul(th, a * a * thd));
    rate = add(rate, mul(add(th, ts), fd * a * F * nd));
    rate = mul(rate, fs / tt);
    s = add(s, mul(add(jau, rate), dt));
    e = e + dt * (1.0 + e) * (D.xx + D.yy + D.zz);
    // Tension cut (zero stress for a tensile mean stress, as the DP tension
    // cut plus apex return give for c = 0) and e_d <= e <= e_i at the new
    // pressure; NaN passes both.
    const double trn = s.xx + s.yy + s.zz;
    if (trn >= 0.0) s = T6{0, 0, 0, 0, 0, 0};
    const double exn = std::exp(-std::pow(std::max(-trn, 0.0) / hs, n));
    const double elo = vw->ed0 * exn, ehi = vw->ei0 * exn;
    e = e < elo ? elo : (e > ehi ? ehi : e);
}

void constitutive(const lq_params* par, const lq_vw_params* vw, double dt, double wx, double wy, double wz,
                  const T6& D, const T6& stn_new, double& p, double& ep, double& voidr, T6& s) {
    if (par->constitutive_model == 1) {
        mat_dp(par, dt, wx, wy, wz, D, p, ep, s);
    } else if (par->constitutive_model == 2) {
        mat_wu(par, dt, D, wx, wy, wz, s);
        ep = std::sqrt(2.0 / 3.0 * j2(stn_new, stn_new));
    } else {
        mat_vw(vw, dt, D, wx, wy, wz, s, voidr);
        ep = std::sqrt(2.0 / 3.0 * j2(stn_new, stn_new));
        p = -(s.xx + s.yy + s.zz) / 3.0;
    }
}

}  // namespace

extern "C" {

// Cell id of each particle (layer 0 for i < npb, 1 otherwise;
// ncell_delete = ncell - 1 outside the grid).
void lqo_cell_keys(const lq_params* par, int n, const double* x, const double* y, const double* z, int32_t* keys) {
    const Grid g = grid_of(par);
    const float i_h = par->i_h;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int cx = cell_coord((float)x[i], i_h), cy = cell_coord((float)y[i], i_h), cz = cell_coord((float)z[i], i_h);
        if (cx < 0 || cx > g.nx - 1 || cy < 0 || cy > g.ny - 1 || cz < 0 || cz > g.nz - 1)
            keys[i] = par->ncell - 1;
        else
            keys[i] = cx * g.ny * g.nz + cy * g.nz + cz + (i < par->npb ? 0 : 1) * g.nol;
    }
}

// Stable ascending counting sort: perm[k] = index of the k-th smallest key
// (ties in input order), sorted[k] = its key.
void lqo_stable_sort(int n, const int32_t* keys, int ncell, int32_t* perm, int32_t* sorted) {
    std::vector<int64_t> start((size_t)ncell + 1, 0);
    for (int i = 0; i < n; ++i) start[(size_t)keys[i] + 1]++;
    for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
    for (int i = 0; i < n; ++i) {
        const int64_t k = start[keys[i]]++;
        perm[k] = i;
        sorted[k] = keys[i];
    }
}

// cell_beg[c] = first sorted index with key c or -1; cell_end[c] = last + 1 or -1.
void lqo_cell_table(int n, const int32_t* sorted, int ncell, int32_t* beg, int32_t* end) {
    for (int c = 0; c < ncell; ++c) beg[c] = end[c] = -1;
    for (int k = 0; k < n; ++k) {
        const int c = sorted[k];
        if (k == 0 || sorted[k - 1] != c) beg[c] = k;
        if (k == n - 1 || sorted[k + 1] != c) end[c] = k + 1;
    }
}

void lqo_grid_build(const lq_params* par, int n, const double* x, const double* y, const double* z, int32_t* perm,
                    int32_t* sorted, int32_t* beg, int32_t* end) {
    std::vector<int32_t> keys(n);
    lqo_cell_keys(par, n, x, y, z, keys.data());
    lqo_stable_sort(n, keys.data(), par->ncell, perm, sorted);
    lqo_cell_table(n, sorted, par->ncell, beg, end);
}

static void permute_arr(int n, const int32_t* perm, double* a, std::vector<double>& tmp) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) tmp[k] = a[k];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) a[k] = tmp[perm[k]];
}

// SortArraysPrimary / SortArraysAll (src/loquat.cu:190-213), in place.
void lqo_permute(int n, const int32_t* perm, lqo_state* st, lqo_ini* ini) {
    std::vector<double> tmp(n);
    double* arrs[] = {st->x, st->y, st->z, st->p, st->vx, st->vy, st->vz, st->rho, st->id, st->ep, st->voidr};
    for (double* a : arrs) permute_arr(n, perm, a, tmp);
    for (int c = 0; c < 6; ++c) {
        permute_arr(n, perm, st->s[c], tmp);
        permute_arr(n, perm, st->e[c], tmp);
    }
    std::vector<int32_t> ti(st->orig, st->orig + n);
    for (int k = 0; k < n; ++k) st->orig[k] = ti[perm[k]];
    if (ini) {
        double* ia[] = {ini->x, ini->y, ini->z, ini->vx, ini->vy, ini->vz, ini->rho, ini->ep, ini->voidr};
        for (double* a : ia) permute_arr(n, perm, a, tmp);
        for (int c = 0; c < 6; ++c) {
            permute_arr(n, perm, ini->s[c], tmp);
            permute_arr(n, perm, ini->e[c], tmp);
        }
    }
}

// AdamiBoundary_cuk (src/adami_boundary.cu:65-123), wall particles [0, npb).
void lqo_adami(const lq_params* par, lqo_state* st, const int32_t* beg, const int32_t* end) {
    const Grid g = grid_of(par);
    const float i_hf = par->i_h;
    const double i_h = par->i_h, cw = par->kernel_normalization_par, m = par->m;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < par->npb; ++i) {
        const float xf = (float)st->x[i], yf = (float)st->y[i], zf = (float)st->z[i];
        T6 acc{0, 0, 0, 0, 0, 0};
        double px = 0, py = 0, pz = 0, w = 0;
        stencil(g, cell_coord(xf, i_hf), cell_coord(yf, i_hf), cell_coord(zf, i_hf), 1, beg, end, [&](int b, int e) {
            for (int j = b; j < e; ++j) {
                if (!is_nbr(xf, yf, zf, (float)st->x[j], (float)st->y[j], (float)st->z[j], i_hf)) continue;
                const double dx = st->x[i] - st->x[j], dy = st->y[i] - st->y[j], dz = st->z[i] - st->z[j];
                const double q = std::sqrt(dx * dx + dy * dy + dz * dz) * i_h;
                const double t = std::max(2.0 - q, 0.0);
                const double kv = cw * t * t * t * t * (q + 0.5) * (m / st->rho[j]);
                acc = add(acc, mul(load6(st->s, j), kv));
                px += dx * kv;
                py += dy * kv;
                pz += dz * kv;
                w += kv;
            }
        });
        if (w > 0.0) {
            const double rho0 = par->rho0;
            T6 o;
            o.xx = (acc.xx - rho0 * par->acc_x * px) / w;
            o.yy = (acc.yy - rho0 * par->acc_y * py) / w;
            o.zz = (acc.zz - rho0 * par->acc_z * pz) / w;
            o.xy = acc.xy / w;
            o.xz = acc.xz / w;
            o.yz = acc.yz / w;
            store6(st->s, i, o);
        }
    }
}

// ParticleInteraction for material particles [npb, n): neighbours j != i in
// both layers of the 27-cell stencil.
//   grad_i W_ij = C_g (2 - q)^3 (x_i - x_j)      (C_g, src/loquat.cu:131)
//   L_ab   = sum_j m/rho_j (v_j - v_i)_a d_b W
//   acestn = sym(L),  spin = (W_yx, W_zx, W_zy) with W = skew(L)
//   drho   = sum_j m (v_i - v_j) . grad W
//   a      = sum_j m (s_i/rho_i^2 + s_j/rho_j^2 - Pi_ij I) . grad W + g
//   Pi_ij  = -alpha cs mu_ij / rho_bar,  mu_ij = h v_ij.x_ij / (r^2 + eta)
//            when v_ij.x_ij < 0, else 0.
}  // extern "C"

namespace {

// The interaction loop.  ABS = false: the rates.  ABS = true: for ax, ay,
// az and drho the sum over the particle's pairs of each pair term's absolute
// value, plus |g| for the accelerations -- the scale against which the
// rounding of a sum of cancelling terms is measured (other outputs zero).
template <bool ABS>
void interaction_impl(const lq_params* par, const lqo_state* st, lqo_rates* rt, const int32_t* beg,
                      const int32_t* end, int32_t* nbr) {
    auto acc = [](double& sum, double term) { sum += ABS ? std::fabs(term) : term; };
    const Grid g = grid_of(par);
    const float i_hf = par->i_h;
    const double i_h = par->i_h, cg = par->kernel_gradient_normalization_par, m = par->m, h = par->h, eta = par->eta;
    const double alpha_cs =
        par->viscosity_type == 1 ? (double)par->artificial_viscosity_alpha * (double)par->cs : 0.0;
    // Artificial pressure (artificial-pressure = 1).  Specification decision:
    // the reference derives W(dr) for it (src/loquat.cu:133-138) but the term
    // lives in the unshipped interaction body.  Monaghan's form (J. Comput.
    // Phys. 159, 2000) with n = 4: R_i = eps * max(-p_i, 0) / rho_i^2 (p =
    // pospres.w, tension negative), f_ij = W(r_ij) / W(dr), and
    // (R_i + R_j) f_ij^4 is subtracted from the diagonal of the momentum
    // sum like a pressure.
    const bool art = par->artificial_pressure == 1;
    const double eps = art ? (double)par->artificial_pressure_coefficient : 0.0;
    const double cw = par->kernel_normalization_par, inv_wdr = art ? 1.0 / (double)par->kernel_dr : 0.0;
    const int n = par->np;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = par->npb; i < n; ++i) {
        const float xf = (float)st->x[i], yf = (float)st->y[i], zf = (float)st->z[i];
        const double ri = st->rho[i];
        const T6 Ti = mul(load6(st->s, i), 1.0 / (ri * ri));
        const double Ri = eps * std::max(-st->p[i], 0.0) / (ri * ri);
        double ax = 0, ay = 0, az = 0, drho = 0, L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int count = 0;
        const int cix = cell_coord(xf, i_hf), ciy = cell_coord(yf, i_hf), ciz = cell_coord(zf, i_hf);
        for (int layer = 0; layer <= 1; ++layer) {
            stencil(g, cix, ciy, ciz, layer, beg, end, [&](int b, int e) {
                for (int j = b; j < e; ++j) {
                    if (j == i) continue;
                    if (!is_nbr(xf, yf, zf, (float)st->x[j], (float)st->y[j], (float)st->z[j], i_hf)) continue;
                    ++count;
                    const double dx = st->x[i] - st->x[j], dy = st->y[i] - st->y[j], dz = st->z[i] - st->z[j];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double q = std::sqrt(r2) * i_h;
                    const double t = std::max(2.0 - q, 0.0);
                    const double fac = cg * t * t * t;
                    const double gx = fac * dx, gy = fac * dy, gz = fac * dz;
                    const double rj = st->rho[j];
                    const double vj = m / rj;
                    const double dv[3] = {st->vx[j] - st->vx[i], st->vy[j] - st->vy[i], st->vz[j] - st->vz[i]};
                    const double gr[3] = {gx, gy, gz};
                    for (int a = 0; a < 3; ++a)
                        for (int bb = 0; bb < 3; ++bb) L[3 * a + bb] += vj * dv[a] * gr[bb];
                    acc(drho, m * (-(dv[0] * gx + dv[1] * gy + dv[2] * gz)));
                    const double vdx = -(dv[0] * dx + dv[1] * dy + dv[2] * dz);
                    double pij = 0.0;
                    if (vdx < 0.0) {
                        const double mu = h * vdx / (r2 + eta);
                        pij = -alpha_cs * mu / (0.5 * (ri + rj));
                    }
                    if (art) {
                        const double Rj = eps * std::max(-st->p[j], 0.0) / (rj * rj);
                        const double f = cw * t * t * t * t * (q + 0.5) * inv_wdr;
                        pij += (Ri + Rj) * (f * f) * (f * f);
                    }
                    const T6 S = add(Ti, mul(load6(st->s, j), 1.0 / (rj * rj)));
                    const double sxx = S.xx - pij, syy = S.yy - pij, szz = S.zz - pij;
                    acc(ax, m * (sxx * gx + S.xy * gy + S.xz * gz));
                    acc(ay, m * (S.xy * gx + syy * gy + S.yz * gz));
                    acc(az, m * (S.xz * gx + S.yz * gy + szz * gz));
                }
            });
        }
        if (ABS) {
            rt->ax[i] = ax + std::fabs((double)par->acc_x);
            rt->ay[i] = ay + std::fabs((double)par->acc_y);
            rt->az[i] = az + std::fabs((double)par->acc_z);
            rt->drho[i] = drho;
            rt->wx[i] = rt->wy[i] = rt->wz[i] = 0.0;
            for (int c = 0; c < 6; ++c) rt->d[c][i] = 0.0;
            if (nbr) nbr[i] = count;
            continue;
        }
        rt->ax[i] = ax + par->acc_x;
        rt->ay[i] = ay + par->acc_y;
        rt->az[i] = az + par->acc_z;
        rt->drho[i] = drho;
        rt->wx[i] = 0.5 * (L[3] - L[1]);
        rt->wy[i] = 0.5 * (L[6] - L[2]);
        rt->wz[i] = 0.5 * (L[7] - L[5]);
        rt->d[0][i] = L[0];
        rt->d[1][i] = L[4];
        rt->d[2][i] = L[8];
        rt->d[3][i] = 0.5 * (L[1] + L[3]);
        rt->d[4][i] = 0.5 * (L[2] + L[6]);
        rt->d[5][i] = 0.5 * (L[5] + L[7]);
        if (nbr) nbr[i] = count;
    }
}

}  // namespace

extern "C" {

void lqo_interaction(const lq_params* par, const lqo_state* st, lqo_rates* rt, const int32_t* beg, const int32_t* end,
                     int32_t* nbr) {
    interaction_impl<false>(par, st, rt, beg, end, nbr);
}

// Per material particle, sum_j |term_ij| of ax, ay, az (plus |g|) and drho:
// the scale of the per-particle error of a sum whose terms cancel (an
// oracle diagnostic).
void lqo_interaction_abs(const lq_params* par, const lqo_state* st, lqo_rates* rt, const int32_t* beg,
                         const int32_t* end) {
    interaction_impl<true>(par, st, rt, beg, end, nullptr);
}

// Predictor (src/predictor_corrector.cu:169-253): *_ini copy of the material
// state, then the half step with the constitutive update.
void lqo_predictor(const lq_params* par, const lq_vw_params* vw, lqo_state* st, lqo_ini* ini, const lqo_rates* rt) {
    const double dt = par->dt, hdt = 0.5 * dt;
    const int n = par->np;
#pragma omp parallel for schedule(static)
    for (int i = par->npb; i < n; ++i) {
        ini->x[i] = st->x[i];
        ini->y[i] = st->y[i];
        ini->z[i] = st->z[i];
        ini->vx[i] = st->vx[i];
        ini->vy[i] = st->vy[i];
        ini->vz[i] = st->vz[i];
        ini->rho[i] = st->rho[i];
        ini->ep[i] = st->ep[i];
        ini->voidr[i] = st->voidr[i];
        for (int c = 0; c < 6; ++c) {
            ini->s[c][i] = st->s[c][i];
            ini->e[c][i] = st->e[c][i];
        }
        st->vx[i] += hdt * rt->ax[i];
        st->vy[i] += hdt * rt->ay[i];
        st->vz[i] += hdt * rt->az[i];
        st->rho[i] += hdt * rt->drho[i];
        st->x[i] += hdt * st->vx[i];
        st->y[i] += hdt * st->vy[i];
        st->z[i] += hdt * st->vz[i];
        const T6 D = load6(rt->d, i);
        const T6 stn = add(load6(st->e, i), mul(D, hdt));
        store6(st->e, i, stn);
        double p = st->p[i], ep = st->ep[i], voidr = st->voidr[i];
        T6 s = load6(st->s, i);
        constitutive(par, vw, hdt, rt->wx[i], rt->wy[i], rt->wz[i], D, stn, p, ep, voidr, s);
        st->p[i] = p;
        st->ep[i] = ep;
        st->voidr[i] = voidr;
        store6(st->s, i, s);
    }
}

// Corrector (src/predictor_corrector.cu:256-336).
void lqo_corrector(const lq_params* par, const lq_vw_params* vw, lqo_state* st, const lqo_ini* ini,
                   const lqo_rates* rt) {
    const double dt = par->dt;
    const int n = par->np;
#pragma omp parallel for schedule(static)
    for (int i = par->npb; i < n; ++i) {
        st->vx[i] = ini->vx[i] + dt * rt->ax[i];
        st->vy[i] = ini->vy[i] + dt * rt->ay[i];
        st->vz[i] = ini->vz[i] + dt * rt->az[i];
        st->rho[i] = ini->rho[i] + dt * rt->drho[i];
        st->x[i] = 0.5 * (st->vx[i] + ini->vx[i]) * dt + ini->x[i];
        st->y[i] = 0.5 * (st->vy[i] + ini->vy[i]) * dt + ini->y[i];
        st->z[i] = 0.5 * (st->vz[i] + ini->vz[i]) * dt + ini->z[i];
        const T6 D = load6(rt->d, i);
        const T6 stn = add(load6(ini->e, i), mul(D, dt));
        store6(st->e, i, stn);
        double p = st->p[i], ep = ini->ep[i], voidr = ini->voidr[i];
        T6 s = load6(ini->s, i);
        constitutive(par, vw, dt, rt->wx[i], rt->wy[i], rt->wz[i], D, stn, p, ep, voidr, s);
        st->p[i] = p;
        st->ep[i] = ep;
        st->voidr[i] = voidr;
        store6(st->s, i, s);
    }
}

// StressRegularization (src/field_smoothing.cu:65-145): material stress
// replaced by its Shepard average over material neighbours, self included.
// A particle with no neighbour keeps its stress.
void lqo_shepard(const lq_params* par, lqo_state* st, const int32_t* beg, const int32_t* end) {
    const Grid g = grid_of(par);
    const float i_hf = par->i_h;
    const double i_h = par->i_h, cw = par->kernel_normalization_par, m = par->m;
    const int n = par->np;
    std::vector<T6> out((size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = par->npb; i < n; ++i) {
        const float xf = (float)st->x[i], yf = (float)st->y[i], zf = (float)st->z[i];
        T6 acc{0, 0, 0, 0, 0, 0};
        double w = 0;
        stencil(g, cell_coord(xf, i_hf), cell_coord(yf, i_hf), cell_coord(zf, i_hf), 1, beg, end, [&](int b, int e) {
            for (int j = b; j < e; ++j) {
                if (!is_nbr(xf, yf, zf, (float)st->x[j], (float)st->y[j], (float)st->z[j], i_hf)) continue;
                const double dx = st->x[i] - st->x[j], dy = st->y[i] - st->y[j], dz = st->z[i] - st->z[j];
                const double q = std::sqrt(dx * dx + dy * dy + dz * dz) * i_h;
                const double t = std::max(2.0 - q, 0.0);
                const double kv = cw * t * t * t * t * (q + 0.5) * (m / st->rho[j]);
                acc = add(acc, mul(load6(st->s, j), kv));
                w += kv;
            }
        });
        out[i] = w > 0.0 ? mul(acc, 1.0 / w) : load6(st->s, i);
    }
    for (int i = par->npb; i < n; ++i) store6(st->s, i, out[i]);
}

// One neighbour search + Primary sort (src/loquat.cu:297-298), grid arrays
// kept by the caller.
static void build_and_sort(const lq_params* par, lqo_state* st, lqo_ini* ini, std::vector<int32_t>& perm,
                           std::vector<int32_t>& sorted, std::vector<int32_t>& beg, std::vector<int32_t>& end) {
    const int n = par->np;
    lqo_grid_build(par, n, st->x, st->y, st->z, perm.data(), sorted.data(), beg.data(), end.data());
    lqo_permute(n, perm.data(), st, ini);
}

void lqo_initial_sort(const lq_params* par, lqo_state* st) {
    const int n = par->np;
    std::vector<int32_t> perm(n), sorted(n), beg(par->ncell), end(par->ncell);
    build_and_sort(par, st, nullptr, perm, sorted, beg, end);
}

// nsteps steps of the time loop (src/loquat.cu:322-409) starting at step
// index step0 (the Shepard cadence counts from there).
// Dynamic time step (dynamic-time-step = 1).  Specification decision: the
// reference's call is commented out (src/loquat.cu:325-327) and its body is
// not shipped.  dt = cfl * min(h / (cs + max|v|), sqrt(h / max|a|)) over the
// material particles, with a = the previous corrector's rates; the |a| term
// is absent when use_rates is 0 (a run's first step) or max|a| = 0.  The
// result is rounded to FP32, the type of Parameters::dt.
double lqo_dynamic_dt(const lq_params* par, const lqo_state* st, const lqo_rates* rt, int use_rates) {
    double vmax = 0.0, amax = 0.0;
    for (int i = par->npb; i < par->np; ++i) {
        vmax = std::max(vmax, std::sqrt(st->vx[i] * st->vx[i] + st->vy[i] * st->vy[i] + st->vz[i] * st->vz[i]));
        if (use_rates)
            amax = std::max(amax, std::sqrt(rt->ax[i] * rt->ax[i] + rt->ay[i] * rt->ay[i] + rt->az[i] * rt->az[i]));
    }
    const double h = par->h;
    double dt = h / ((double)par->cs + vmax);
    if (amax > 0.0) dt = std::min(dt, std::sqrt(h / amax));
    return (double)(float)((double)par->cfl * dt);
}

// lqo_run, also writing each step's dt to dts (nullable).
void lqo_run_dt(const lq_params* par_in, const lq_vw_params* vw, lqo_state* st, int nsteps, int step0,
                double* dts) {
    lq_params pl = *par_in;  // dt changes per step with dynamic-time-step = 1
    const lq_params* par = &pl;
    const int n = par->np;
    std::vector<int32_t> perm(n), sorted(n), beg(par->ncell), end(par->ncell);
    std::vector<double> ib((size_t)21 * n), rb((size_t)13 * n);
    lqo_ini ini;
    double** ip[] = {&ini.x, &ini.y, &ini.z, &ini.vx, &ini.vy, &ini.vz, &ini.rho, &ini.ep, &ini.voidr};
    int k = 0;
    for (double** p : ip) *p = ib.data() + (size_t)(k++) * n;
    for (int c = 0; c < 6; ++c) {
        ini.s[c] = ib.data() + (size_t)(k++) * n;
        ini.e[c] = ib.data() + (size_t)(k++) * n;
    }
    lqo_rates rt;
    double** rp[] = {&rt.ax, &rt.ay, &rt.az, &rt.drho, &rt.wx, &rt.wy, &rt.wz};
    k = 0;
    for (double** p : rp) *p = rb.data() + (size_t)(k++) * n;
    for (int c = 0; c < 6; ++c) rt.d[c] = rb.data() + (size_t)(k++) * n;
    for (int it = 0; it < nsteps; ++it) {
        const int step = step0 + it;
        if (pl.dynamic_time_step == 1) pl.dt = (float)lqo_dynamic_dt(&pl, st, &rt, it > 0 ? 1 : 0);
        if (dts) dts[it] = (double)pl.dt;
        build_and_sort(par, st, nullptr, perm, sorted, beg, end);
        lqo_adami(par, st, beg.data(), end.data());
        lqo_interaction(par, st, &rt, beg.data(), end.data(), nullptr);
        lqo_predictor(par, vw, st, &ini, &rt);
        build_and_sort(par, st, &ini, perm, sorted, beg, end);
        lqo_adami(par, st, beg.data(), end.data());
        lqo_interaction(par, st, &rt, beg.data(), end.data(), nullptr);
        lqo_corrector(par, vw, st, &ini, &rt);
        const int f = par->stress_shepard_filter;
        if (f != 0 && step % f == 0 && step != 0) {
            build_and_sort(par, st, nullptr, perm, sorted, beg, end);
            lqo_shepard(par, st, beg.data(), end.data());
        }
    }
}

void lqo_run(const lq_params* par, const lq_vw_params* vw, lqo_state* st, int nsteps, int step0) {
    lqo_run_dt(par, vw, st, nsteps, step0, nullptr);
}

// One constitutive update at a single material point (the constitutive
// branch of Predictor_cuk / Corrector_cuk): stress s[6], ep, voidr, p in/out.
void lqo_constitutive(const lq_params* par, const lq_vw_params* vw, double dt, const double* spin3,
                      const double* d6, const double* stn6, double* p, double* ep, double* voidr, double* s6) {
    T6 D{d6[0], d6[1], d6[2], d6[3], d6[4], d6[5]};
    T6 stn{stn6[0], stn6[1], stn6[2], stn6[3], stn6[4], stn6[5]};
    T6 s{s6[0], s6[1], s6[2], s6[3], s6[4], s6[5]};
    constitutive(par, vw, dt, spin3[0], spin3[1], spin3[2], D, stn, *p, *ep, *voidr, s);
    s6[0] = s.xx;
    s6[1] = s.yy;
    s6[2] = s.zz;
    s6[3] = s.xy;
    s6[4] = s.xz;
    s6[5] = s.yz;
}

// Pairs whose neighbour decision depends on the precision of the predicate
// (an oracle diagnostic, no reference counterpart): over the material
// particles' 27-cell stencils (both layers, j != i), the reference's FP32
// test is_nbr against q = r / h < 2 evaluated in FP64 on the FP64 positions.
// out[0]: FP32 accepts and FP64 rejects; out[1]: FP64 accepts and FP32
// rejects.  Counts are per ordered pair (a material-material pair counts
// from both sides).
void lqo_borderline_pairs(const lq_params* par, const lqo_state* st, const int32_t* beg, const int32_t* end,
                          int64_t* out) {
    const Grid g = grid_of(par);
    const float i_hf = par->i_h;
    const double i_h = par->i_h;
    const int n = par->np;
    int64_t only32 = 0, only64 = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : only32, only64)
    for (int i = par->npb; i < n; ++i) {
        const float xf = (float)st->x[i], yf = (float)st->y[i], zf = (float)st->z[i];
        const int cix = cell_coord(xf, i_hf), ciy = cell_coord(yf, i_hf), ciz = cell_coord(zf, i_hf);
        for (int layer = 0; layer <= 1; ++layer) {
            stencil(g, cix, ciy, ciz, layer, beg, end, [&](int b, int e) {
                for (int j = b; j < e; ++j) {
                    if (j == i) continue;
                    const bool f32 = is_nbr(xf, yf, zf, (float)st->x[j], (float)st->y[j], (float)st->z[j], i_hf);
                    const double dx = st->x[i] - st->x[j], dy = st->y[i] - st->y[j], dz = st->z[i] - st->z[j];
                    const bool f64 = std::sqrt(dx * dx + dy * dy + dz * dz) * i_h < 2.0;
                    only32 += (f32 && !f64) ? 1 : 0;
                    only64 += (f64 && !f32) ? 1 : 0;
                }
            });
        }
    }
    out[0] = only32;
    out[1] = only64;
}

void lqo_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int lqo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
