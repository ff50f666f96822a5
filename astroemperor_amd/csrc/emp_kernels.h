// Likelihood building blocks shared by the batched-lnL, model-eval and PT
// kernels.
//
// Launch geometry: one workgroup = 8 wave64s.  The workgroup owns SLOTS =
// 8 / WPT parameter vectors; each vector is processed by WPT waves whose
// 64*WPT lanes stride over the epochs.  Epoch data (t, y, yerr^2, MA time
// step, instrument tag) is staged into LDS in tiles of TILE epochs and shared
// by all slots of the workgroup.
#pragma once
#include "emp_internal.h"

namespace emp {

constexpr double LN2 = 0.69314718055994530942;

// Prior of one full parameter vector, sequentially and in the reference's
// order so that the sum is bit-identical to the generated my_prior
// (emp/emp.py:182-254): per block, add each free parameter's prior, return
// -inf as soon as a block leaves the sum at -inf, then add the block's
// derived eccentricity prior.
__device__ inline double prior_term(const EmpPriorTerm& p, double x) {
#pragma clang fp contract(off)
    if (!(x >= p.lo && x <= p.hi)) return -__builtin_inf();
    switch (p.kind) {
        case EMP_PRIOR_UNIFORM:
        case EMP_PRIOR_JEFFREYS:
            return p.a0;
        case EMP_PRIOR_NORMAL: {
            double z = (x - p.a0) / p.a1;
            return -0.5 * (z * z) - p.a3 - p.a2;
        }
        case EMP_PRIOR_BETA: {
            double v = -p.a2;
            if (p.a0 != 1.0) v += (p.a0 - 1.0) * log(x);
            if (p.a1 != 1.0) v += (p.a1 - 1.0) * log1p(-x);
            return v;
        }
        default:
            return 0.0;
    }
}

__device__ inline double prior_seq(const ModelDev& md, const double* th) {
#pragma clang fp contract(off)
    double lp = 0.0;
    for (int b = 0; b < md.nblocks; ++b) {
        for (int j = md.block_start[b]; j < md.block_start[b + 1]; ++j) {
            if (md.inv_free[j] < 0) continue;
            lp += prior_term(md.prior[j], th[j]);
        }
        if (lp == -__builtin_inf()) return lp;
        if (b < md.K && md.ecc_prior[b].kind != EMP_PRIOR_NONE) {
            const double* s = th + md.off_kep + 5 * b;
            double x = s[3] * s[3] + s[4] * s[4];
            lp += prior_term(md.ecc_prior[b], x);
        }
    }
    return lp;
}

// Per-slot constants (planets, instruments, acceleration) into LDS.
// Called by the slot's first wave; the caller synchronises afterwards.
template <class SL>
__device__ inline void slot_setup(const ModelDev& md, SL& sl, int lane) {
    int kf = 0;
    if (lane < md.K) {
        PlanetC pc = planet_setup(sl.theta + md.off_kep + 5 * lane, md.param, md.use_c);
        pc.ktab = md.N >= KT_MIN_EPOCHS && pc.e <= KT_EMAX;
        kf = kt_flags(pc.ktab != 0, pc.e);
        sl.planet[lane] = pc;
    }
    // the whole wave calls these: flag bit b of planet p to mask bit 4 p + b
    unsigned m = 0;
    for (int b = 0; b < 3; ++b) {
        const unsigned long long bits = __ballot((kf >> b) & 1);
        for (int p = 0; p < KMAX; ++p) m |= (unsigned)((bits >> p) & 1ull) << (4 * p + b);
    }
    if (lane == 0) sl.ktab_mask = (int)m;
    if (lane < md.I) {
        sl.gam[lane] = sl.theta[md.off_off + lane];
        double j = md.has_jit ? sl.theta[md.off_jit + lane] : 0.0;
        sl.jit[lane] = j;
        sl.jit2[lane] = j * j;
        if (md.ma_active) {
            int set = md.ma_global ? 0 : lane;
            sl.ma_c[lane] = sl.theta[md.off_ma + 2 * md.ma_q * set];
            sl.ma_tau[lane] = sl.theta[md.off_ma + 2 * md.ma_q * set + 1];
        }
    }
    if (lane < md.acc) sl.acc[lane] = sl.theta[md.off_acc + lane];
    if (lane < md.n_cos) {
        // Term 0 is the sinusoid when there is one (sinusoid00.model:2-5:
        // freq = 2 pi / P); the magnetic cycle's terms are (A1, pi / P, phase1)
        // and (A2, 2 pi / P, phase2) (magneticcycle00.model:2-8).
        const int c = md.off_sin >= 0 ? lane - 1 : lane;
        double A, f, ph;
        if (c < 0) {
            const double* th = sl.theta + md.off_sin;
            A = th[1]; f = (2.0 * PI_D) / th[0]; ph = th[2];
        } else {
            const double* th = sl.theta + md.off_mag;
            A = th[1 + c]; f = (c == 0 ? PI_D : 2.0 * PI_D) / th[0]; ph = th[3 + c];
        }
        sl.cosw[lane][0] = A;
        sl.cosw[lane][1] = f;
        sl.cosw[lane][2] = ph;
    }
}

struct TileLDS {
    double t[TILE], y[TILE], s2[TILE], dt[TILE];
    int ins[TILE];
    SinCosTab tab;
};

// The sin/cos table is read by every planet-epoch; load it once per workgroup.
__device__ __forceinline__ void tab_load(const ModelDev& md, TileLDS& tl) {
    for (int j = threadIdx.x; j <= SC_N; j += blockDim.x) tl.tab.v[j] = md.sctab[j];
}

__device__ __forceinline__ void tile_load(const ModelDev& md, TileLDS& tl, int tb, int tn) {
    for (int j = threadIdx.x; j < tn; j += blockDim.x) {
        tl.t[j] = md.t[tb + j];
        tl.y[j] = md.y[tb + j];
        tl.s2[j] = md.s2[tb + j];
        tl.ins[j] = md.ins[tb + j];
        if (md.ma_active) tl.dt[j] = md.dt[tb + j];
    }
}

// One epoch of a tile in registers.  Every launch that stages tiles has at
// least TILE threads, so thread j holds epoch tb + j of the tile at tb: its
// global loads are issued a tile ahead and stored to LDS when the tile's
// turn comes.
struct EpochRegs {
    double t, y, s2, dt;
    int ins;
};

__device__ __forceinline__ void epoch_prefetch(const ModelDev& md, int tb, EpochRegs& r) {
    const int j = threadIdx.x, n = tb + j;
    if (j < TILE && n < md.N) {
        r.t = md.t[n];
        r.y = md.y[n];
        r.s2 = md.s2[n];
        r.ins = md.ins[n];
        r.dt = md.ma_active ? md.dt[n] : 0.0;
    }
}

__device__ __forceinline__ void epoch_store(const ModelDev& md, TileLDS& tl, int tb, const EpochRegs& r) {
    const int j = threadIdx.x;
    if (j < TILE && tb + j < md.N) {
        tl.t[j] = r.t;
        tl.y[j] = r.y;
        tl.s2[j] = r.s2;
        tl.ins[j] = r.ins;
        if (md.ma_active) tl.dt[j] = r.dt;
    }
}

// Per-slot constants, then (mixed path) the per-vector Kepler tables, one
// planet per wave of the slot.  The sin/cos table must already be in LDS.
// All threads of the workgroup call this (it holds two barriers).
template <int PREC, int WPT, class SL>
__device__ __forceinline__ void slot_prepare(const ModelDev& md, SL& sl, const TileLDS& tl, bool live,
                                             int wsub, int lane) {
    if (live && wsub == 0) slot_setup(md, sl, lane);
    __syncthreads();
    if (PREC == 1 && live) {
        for (int p = wsub; p < md.K; p += WPT)
            if (sl.planet[p].ktab) ktab_fill(sl.planet[p], lane, sl.ktab[p][lane]);
    }
    __syncthreads();
}

// A value that is the same in every lane of the wave (every kernel keeps
// one parameter vector per wave), read into a scalar register so that a
// branch on it is a scalar branch.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The model's planets are parameterised by the time of periastron
// (block_repo.py:23-31, parameterisations 3 and 4).
__host__ __device__ __forceinline__ bool model_tp(const ModelDev& md) { return md.param == 3 || md.param == 4; }

// Model value (planets + offset + acceleration) of one epoch.  kmask: the
// slot's SlotLDS::ktab_mask as a wave-uniform scalar.
template <int PREC, class SL>
__device__ __forceinline__ double epoch_mu(const ModelDev& md, const SL& sl, const SinCosTab& tb, double t, int k,
                                           int kmask) {
    double mu = sl.gam[k];
    if (md.acc) {
        const double x = t - md.t0;
        double v = 0.0;
        for (int a = 0; a < md.acc; ++a) v = __builtin_fma(v, x, sl.acc[a]);
        mu += v * x;
    }
    // The planet loop twice, one copy per parameterisation (a model-wide
    // choice), so that no planet-epoch branches on it.
    if (model_tp(md)) {
        for (int p = 0; p < md.K; ++p) mu += planet_mu<PREC>(sl.planet[p], tb, sl.ktab[p], t, true, kt_mode(kmask, p));
    } else {
        for (int p = 0; p < md.K; ++p) mu += planet_mu<PREC>(sl.planet[p], tb, sl.ktab[p], t, false, kt_mode(kmask, p));
    }
    return mu;
}

// Phase freq * t + phase of cosine term c, rounded as numpy rounds
// `freq * X_ + phase` (no contraction).
template <class SL>
__device__ __forceinline__ double cos_phase(const SL& sl, int c, double t) {
#pragma clang fp contract(off)
    return sl.cosw[c][1] * t + sl.cosw[c][2];
}

// Terms added after the moving average: in the reference's block order
// StellarActivity, Sinusoid and MagneticCycle follow MOAV (emp/emp.py:
// 2628-2650), so the MA residuals do not contain them.
__host__ __device__ inline bool has_post(const ModelDev& md) { return md.n_cos > 0 || md.sai_w > 0; }

// Model features the common models do not use: the terms added after the MA
// and MA orders above 1.  Kernels come in two instances, with and without
// them, so that the common instance keeps its register budget.
__host__ __device__ inline bool needs_ext(const ModelDev& md) {
    return has_post(md) || (md.ma_active && md.ma_q > 1);
}

// Their sum at epoch n (device order; time t, instrument k): the activity
// columns of the epoch's instrument times their coefficients (sai00.model),
// then the cosine terms, whose angle is reduced in FP64 and whose cosine
// comes from the FP64 table on both precision paths.
template <class SL>
__device__ __forceinline__ double epoch_post(const ModelDev& md, const SL& sl, const SinCosTab& tb, double t,
                                             int n, int k) {
    double v = 0.0;
    if (md.sai_w) {
        const int b0 = md.sai_base[k], nc = md.sai_base[k + 1] - b0;
        const double* zeta = sl.theta + md.off_sai + b0;
#pragma unroll 1
        for (int c = 0; c < nc; ++c) v = __builtin_fma(zeta[c], md.sai[(size_t)c * md.N + n], v);
    }
#pragma unroll 1
    for (int c = 0; c < md.n_cos; ++c) {
        double sn, cs;
        sincos_tab(tb, __builtin_fabs(reduce_2pi(cos_phase(sl, c, t))), sn, cs);
        v = __builtin_fma(sl.cosw[c][0], cs, v);
    }
    return v;
}

// Affine map composition for the MA(1) recurrence eps_i = a_i eps_{i-1} + b_i:
// applying (a1, b1) then (a2, b2) gives (a2 a1, a2 b1 + b2).  Inclusive scan
// over the wave by DPP: within rows of 16 lanes, then across rows; lanes
// without a source compose with the identity (1, 0).  Whole wave active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void affine_step(double& a, double& b) {
    double pa, pb;
    if constexpr (ROW_MASK == 0xf && CTRL > 0x110 && CTRL < 0x120) {
        pa = dpp_shr_one_f64<CTRL>(a);
        pb = dpp_shr_zero_f64<CTRL>(b);
    } else {
        pa = dpp_f64<CTRL, ROW_MASK>(1.0, a);
        pb = dpp_f64<CTRL, ROW_MASK>(0.0, b);
    }
    b = __builtin_fma(a, pb, b);
    a = a * pa;
}

__device__ __forceinline__ void wave_affine_scan(double& a, double& b, int lane) {
    (void)lane;
    affine_step<0x111, 0xf>(a, b);
    affine_step<0x112, 0xf>(a, b);
    affine_step<0x114, 0xf>(a, b);
    affine_step<0x118, 0xf>(a, b);
    affine_step<0x142, 0xa>(a, b);
    affine_step<0x143, 0xc>(a, b);
}

// MA(q), q >= 2: the order-q recurrence (moav01.model:7-15)
//   eps_i = r_i + sum_{c<q} a_{i,c} eps_{i-1-c},
//   a_{i,c} = -phi_c exp(-|t_i - t_{i-1-c}| / tau_c)  (0 when epoch i-1-c is
//   not in the segment),
// as an affine map on the state (eps_i, ..., eps_{i-Q+1}): a companion matrix
// whose first row is (a_{i,0}, ..., a_{i,Q-1}) plus (r_i, 0, ...).  Maps are
// composed over the wave by DPP exactly as the order-1 scan composes its
// (a, b) pairs; lanes without a source compose with the identity.
template <int Q, int CTRL, int ROW_MASK>
__device__ __forceinline__ void affq_step(double (&m)[Q][Q], double (&v)[Q]) {
    double pm[Q][Q], pv[Q];
#pragma unroll
    for (int a = 0; a < Q; ++a) {
#pragma unroll
        for (int b = 0; b < Q; ++b) pm[a][b] = dpp_f64<CTRL, ROW_MASK>(a == b ? 1.0 : 0.0, m[a][b]);
        pv[a] = dpp_f64<CTRL, ROW_MASK>(0.0, v[a]);
    }
    double nm[Q][Q], nv[Q];
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        double acc = v[a];
#pragma unroll
        for (int b = 0; b < Q; ++b) acc = __builtin_fma(m[a][b], pv[b], acc);
        nv[a] = acc;
#pragma unroll
        for (int c = 0; c < Q; ++c) {
            double e = 0.0;
#pragma unroll
            for (int b = 0; b < Q; ++b) e = __builtin_fma(m[a][b], pm[b][c], e);
            nm[a][c] = e;
        }
    }
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        v[a] = nv[a];
#pragma unroll
        for (int c = 0; c < Q; ++c) m[a][c] = nm[a][c];
    }
}

// One chunk of the order-q MA: the lane's epoch (valid: j of the tile at tb,
// time tj, instrument k, residual r) and the state carried in from the
// previous chunk (`carry`, updated).  Reads the tile only before the first
// barrier.  Returns eps of the lane's epoch.  Whole wave active; with WPT > 1
// all waves of the workgroup call this together (one barrier).
template <int Q, int WPT, class SL>
__device__ __forceinline__ double ma_chunk_q(const ModelDev& md, const SL& sl, const TileLDS& tl, bool valid,
                                             int j, int tb, double tj, int k, double r, int lane, int wave,
                                             int wsub, int& parity, double (&carry)[4]) {
    __shared__ double qtot[2][MAX_WAVES][Q * Q + Q];
    double m[Q][Q], v[Q];
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        v[a] = 0.0;
#pragma unroll
        for (int b = 0; b < Q; ++b) m[a][b] = a == b ? 1.0 : 0.0;
    }
    if (valid) {
        const int q = md.ma_q;
        const double* ph = sl.theta + md.off_ma + 2 * q * (md.ma_global ? 0 : k);
        const int n = tb + j;
#pragma unroll
        for (int c = 0; c < Q; ++c) {
            double a = 0.0;
            const int n1 = n - 1 - c, j1 = j - 1 - c;
            if (c < q && n1 >= 0) {
                const int k1 = j1 >= 0 ? tl.ins[j1] : md.ins[n1];
                if (md.ma_global || k1 == k) {
                    const double t1 = j1 >= 0 ? tl.t[j1] : md.t[n1];
                    a = -(ph[2 * c] * exp(-__builtin_fabs(tj - t1) / ph[2 * c + 1]));
                }
            }
            m[0][c] = a;
        }
#pragma unroll
        for (int a = 1; a < Q; ++a)
#pragma unroll
            for (int b = 0; b < Q; ++b) m[a][b] = b == a - 1 ? 1.0 : 0.0;
        v[0] = r;
    }
    affq_step<Q, 0x111, 0xf>(m, v);
    affq_step<Q, 0x112, 0xf>(m, v);
    affq_step<Q, 0x114, 0xf>(m, v);
    affq_step<Q, 0x118, 0xf>(m, v);
    affq_step<Q, 0x142, 0xa>(m, v);
    affq_step<Q, 0x143, 0xc>(m, v);
    double sin_[Q];
    if (WPT == 1) {
#pragma unroll
        for (int a = 0; a < Q; ++a) sin_[a] = carry[a];
    } else {
        if (lane == 63) {
#pragma unroll
            for (int a = 0; a < Q; ++a) {
#pragma unroll
                for (int b = 0; b < Q; ++b) qtot[parity][wave][a * Q + b] = m[a][b];
                qtot[parity][wave][Q * Q + a] = v[a];
            }
        }
        __syncthreads();
        const int w0 = wave - wsub;
        double s_all[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a) s_all[a] = carry[a];
        for (int w = 0; w < WPT; ++w) {
            if (w == wsub) {
#pragma unroll
                for (int a = 0; a < Q; ++a) sin_[a] = s_all[a];
            }
            const double* T = qtot[parity][w0 + w];
            double ns[Q];
#pragma unroll
            for (int a = 0; a < Q; ++a) {
                double acc = T[Q * Q + a];
#pragma unroll
                for (int b = 0; b < Q; ++b) acc = __builtin_fma(T[a * Q + b], s_all[b], acc);
                ns[a] = acc;
            }
#pragma unroll
            for (int a = 0; a < Q; ++a) s_all[a] = ns[a];
        }
#pragma unroll
        for (int a = 0; a < Q; ++a) carry[a] = s_all[a];
        parity ^= 1;
    }
    double st[Q];
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        double acc = v[a];
#pragma unroll
        for (int b = 0; b < Q; ++b) acc = __builtin_fma(m[a][b], sin_[b], acc);
        st[a] = acc;
    }
    if (WPT == 1) {
#pragma unroll
        for (int a = 0; a < Q; ++a) carry[a] = readlane_f64(st[a], 63);
    }
    return st[0];
}

// One suffix step within rows of 16 lanes: lane i composes with lane i + n
// of its row (row_shl:n), the last n lanes of a row with the identity.
template <int CTRL>
__device__ __forceinline__ void affine_suffix_step(double& a, double& b) {
    const double pa = dpp_shr_one_f64<CTRL>(a);
    const double pb = dpp_shr_zero_f64<CTRL>(b);
    b = __builtin_fma(a, pb, b);
    a = a * pa;
}

// Suffix composition for the adjoint of that recurrence, lambda_i = b_i +
// a_i lambda_{i+1}: afterwards lane i holds the map of lanes i..63, so that
// lambda_i = b + a lambda_in with lambda_in the value after lane 63.  Within
// rows by DPP, after which lane i holds the map of lanes i to the end of its
// row and the first lane of a row that of the whole row; then across rows as
// the prefix scan does it mirrored: rows 0 and 2 compose with the first lane
// of the next row, rows 0 and 1 with lane 32 (by then rows 2 and 3).  Whole
// wave active.
__device__ __forceinline__ void wave_affine_suffix_scan(double& a, double& b, int lane) {
    affine_suffix_step<0x101>(a, b);
    affine_suffix_step<0x102>(a, b);
    affine_suffix_step<0x104>(a, b);
    affine_suffix_step<0x108>(a, b);
    {
        const int src = (lane & 48) + 16;
        const double pa = __shfl(a, src, 64), pb = __shfl(b, src, 64);
        if (!(lane & 16)) {
            b = __builtin_fma(a, pb, b);
            a = a * pa;
        }
    }
    {
        const double pa = readlane_f64(a, 32), pb = readlane_f64(b, 32);
        if (lane < 32) {
            b = __builtin_fma(a, pb, b);
            a = a * pa;
        }
    }
}

struct ScanLDS {
    double tot[2][MAX_WAVES][2];
    double red[MAX_WAVES][2];
};

// lnL of the slot's parameter vector.  All threads of the workgroup must
// call this together (it contains workgroup barriers); `live` == false makes
// a slot skip the arithmetic while still taking part in the barriers.  The
// slot must be prepared (slot_prepare) and the sin/cos table loaded.
// With WRITE the per-epoch model and err2 are stored (device order); with
// WEPS the per-epoch MA-filtered residual eps (= r without MA; before the
// terms added after the MA are taken off).
// `resident`: the caller has already staged all epochs (md.N <= TILE) in tl.
// `first`: the first tile's epochs, already loaded by the caller (else loaded
// here).  With PF the next tile's loads are issued a tile ahead (launches
// have at least TILE threads); without, each tile is loaded when its turn
// comes (fewer registers).  POST = false compiles out the features of
// needs_ext: the terms added after the moving average (activity and cosine
// terms) and MA orders above 1.
template <int PREC, int WPT, bool WRITE, bool WEPS = false, bool PF = true, bool POST = true, class SL>
__device__ double block_lnl(const ModelDev& md, const SL& sl, TileLDS& tl, ScanLDS& sc,
                            bool live, int wsub, double* out_model, double* out_err2, bool resident = false,
                            double* out_eps = nullptr, const EpochRegs* first = nullptr) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int G = 64 * WPT;
    const int g = wsub * 64 + lane;
    const int kmask = wave_uniform(sl.ktab_mask);
    double chi = 0.0, prod = 1.0;
    int pexp = 0;
    double carry = 0.0;
    double qcarry[4] = {0.0, 0.0, 0.0, 0.0};   // MA(q >= 2) state carried between chunks
    int parity = 0;
    const bool staged = resident && md.N <= TILE;
    // With the moving average over several waves, every chunk ends with the
    // carry's barrier, after the chunk's last read of the tile.
    const bool ma_sync = md.ma_active && WPT > 1;
    EpochRegs nx{};
    if (PF && !staged) {
        if (first) nx = *first;
        else epoch_prefetch(md, 0, nx);
    }
    for (int tb = 0; tb < md.N; tb += TILE) {
        const int tn = min(TILE, md.N - tb);
        if (!staged) {
            if (tb == 0 || !ma_sync) __syncthreads();     // the tile's previous readers are done
            if (PF) {
                epoch_store(md, tl, tb, nx);
                __syncthreads();
                if (tb + TILE < md.N) epoch_prefetch(md, tb + TILE, nx);
            } else {
                tile_load(md, tl, tb, tn);
                __syncthreads();
            }
        }
        for (int cb = 0; cb < tn; cb += G) {
            const int j = cb + g;
            const bool valid = live && (j < tn);
            // Everything this chunk needs from the tile is read before the MA
            // scan: with the MA over several waves the scan's barrier is the
            // last one before the next tile overwrites this one.
            double r = 0.0, err2 = 1.0, mu = 0.0, tj = 0.0;
            int k = 0;
            if (valid) {
                k = tl.ins[j];
                tj = tl.t[j];
                mu = epoch_mu<PREC>(md, sl, tl.tab, tj, k, kmask);
                err2 = tl.s2[j] + sl.jit2[k];
                r = tl.y[j] - mu;
            }
            double eps = r;
            if (POST && md.ma_active && md.ma_q > 1) {
                eps = md.ma_q == 2
                          ? ma_chunk_q<2, WPT>(md, sl, tl, valid, j, tb, tj, k, r, lane, wave, wsub, parity, qcarry)
                          : ma_chunk_q<4, WPT>(md, sl, tl, valid, j, tb, tj, k, r, lane, wave, wsub, parity, qcarry);
            } else if (md.ma_active) {
                double a = 1.0, b = 0.0;
                if (valid) {
                    const double dtj = tl.dt[j];
                    if (md.ma_wfixed) a = -(sl.ma_c[k] * dtj);
                    else a = dtj < 0.0 ? 0.0 : -(sl.ma_c[k] * exp(-dtj / sl.ma_tau[k]));
                    b = r;
                }
                wave_affine_scan(a, b, lane);
                if (WPT == 1) {
                    eps = __builtin_fma(a, carry, b);
                    carry = readlane_f64(eps, 63);
                } else {
                    if (lane == 63) { sc.tot[parity][wave][0] = a; sc.tot[parity][wave][1] = b; }
                    __syncthreads();
                    const int w0 = wave - wsub;
                    // (pa, pb): the maps of waves 0 .. w of the vector composed
                    if (wsub == 0) eps = __builtin_fma(a, carry, b);
                    double pa = sc.tot[parity][w0][0], pb = sc.tot[parity][w0][1];
                    for (int w = 1; w < WPT; ++w) {
                        if (w == wsub) {
                            eps = __builtin_fma(a, __builtin_fma(pa, carry, pb), b);
                        }
                        const double ta = sc.tot[parity][w0 + w][0], tbv = sc.tot[parity][w0 + w][1];
                        pb = __builtin_fma(ta, pb, tbv);
                        pa = ta * pa;
                    }
                    carry = __builtin_fma(pa, carry, pb);
                    parity ^= 1;
                }
            }
            if (valid) {
                const double cs = (POST && has_post(md)) ? epoch_post(md, sl, tl.tab, tj, tb + j, k) : 0.0;
                const double res = eps - cs;
                chi = __builtin_fma(res * res, PREC == 1 ? rcp64_1(err2) : rcp64(err2), chi);
                prod *= err2;
                pexp += __builtin_amdgcn_frexp_exp(prod);
                prod = __builtin_amdgcn_frexp_mant(prod);
                if (WRITE) {
                    out_model[tb + j] = mu + (r - eps) + cs;
                    out_err2[tb + j] = err2;
                }
                if (WEPS) out_eps[tb + j] = eps;
            }
        }
    }
    double ld = live ? log(prod) + (double)pexp * LN2 : 0.0;
    chi = wave_sum(chi);
    ld = wave_sum(ld);
    if (WPT > 1) {
        __syncthreads();
        if (lane == 0) { sc.red[wave][0] = chi; sc.red[wave][1] = ld; }
        __syncthreads();
        chi = 0.0; ld = 0.0;
        const int w0 = wave - wsub;
        for (int w = 0; w < WPT; ++w) { chi += sc.red[w0 + w][0]; ld += sc.red[w0 + w][1]; }
    }
    return -0.5 * (chi + ld) + md.lnl_const;
}

// Expand the free vector (global memory, row `row`) into the slot's full
// vector in LDS, re-inserting fixed values (emp/emp_model.py:709-711).
template <class SL>
__device__ __forceinline__ void expand_theta(const ModelDev& md, SL& sl, const double* row, int g, int G) {
    for (int j = g; j < md.Dfull; j += G) {
        const int f = md.inv_free[j];
        sl.theta[j] = f >= 0 ? row[f] : md.fixed_val[j];
    }
}

}  // namespace emp
