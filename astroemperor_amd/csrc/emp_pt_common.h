// Device pieces of one parallel-tempering ensemble, shared by the
// single-system engine (emp_pt.hip) and the batched-systems engine
// (emp_systems.hip).
//
// Random numbers: Philox-4x32-10 keyed by (key0, key1) with counters
//   stretch:  (t W + k, sweep_lo, sweep_hi ^ sysmix, 2 step + half [| 2^31])
//   swap:     (i, sweep_lo, sweep_hi ^ sysmix, 2^30 | h)
// and the swap pairing a keyed Feistel permutation of the walkers.  sysmix is
// 0 for a single system, so one system of a batch draws exactly what the
// single-system engine draws.
#pragma once
#include "emp_kernels.h"

namespace emp {

struct PTView {
    int T, W, W2, D;
    double a;
    uint32_t key0, key1, sysmix;
    int adapt_mode;
    double adapt_tau, adapt_nu;
    double* X;                  // [T*W slots][D]
    double* lnl;                // [slots]
    double* lnp;                // [slots]
    int* map;                   // [T][W] -> slot
    double* betas;              // [T]
    long long* accepted;        // [T][W]
};

__host__ __device__ inline PTView pt_view(int T, int W, int D, double a, uint32_t key0, uint32_t key1,
                                          uint32_t sysmix, int adapt_mode, double adapt_tau, double adapt_nu,
                                          double* X, double* lnl, double* lnp, int* map, double* betas,
                                          long long* accepted) {
    PTView v;
    v.T = T; v.W = W; v.W2 = W / 2; v.D = D; v.a = a;
    v.key0 = key0; v.key1 = key1; v.sysmix = sysmix;
    v.adapt_mode = adapt_mode; v.adapt_tau = adapt_tau; v.adapt_nu = adapt_nu;
    v.X = X; v.lnl = lnl; v.lnp = lnp; v.map = map; v.betas = betas; v.accepted = accepted;
    return v;
}

// One slot's task of a fused stretch half-step: walker k = half W/2 + (b mod
// W/2) of temperature t = b / (W/2).  Proposal (Goodman & Weare 2010 eq. 9,
// partner from the complementary half), prior, likelihood when the prior is
// finite, Metropolis acceptance of
//   (D - 1) ln z + beta_t (lnL' - lnL) + ln pi' - ln pi.
// Every evaluated likelihood adds one to *n_lnl (global or LDS).  The
// sin/cos table must be in `tile`; with `resident` all epochs are too.
// All threads of the workgroup call this together (it contains barriers).
template <int PREC, int WPT>
__device__ void pt_half_task(const ModelDev& md, const PTView& pv, long long sweep, int half, int step,
                             int64_t b, SlotLDS& sl, TileLDS& tile, ScanLDS& scan, double* s_lp, int* s_acc,
                             int slot, int wsub, bool resident, unsigned long long* n_lnl) {
    const int lane = threadIdx.x & 63;
    const int g = wsub * 64 + lane, G = 64 * WPT;
    const bool task = b < (int64_t)pv.T * pv.W2;
    int t = 0, k = 0, sk = 0;
    double z = 1.0, uacc = 0.5;
    const uint32_t shi = (uint32_t)((unsigned long long)sweep >> 32) ^ pv.sysmix;
    if (task) {
        t = (int)(b / pv.W2);
        k = half * pv.W2 + (int)(b % pv.W2);
        const uint32_t wid = (uint32_t)(t * pv.W + k);
        U4 c1{wid, (uint32_t)sweep, shi, (uint32_t)(step * 2 + half)};
        U4 r1 = philox4x32_10(c1, pv.key0, pv.key1);
        U4 c2 = c1;
        c2.w |= 0x80000000u;
        U4 r2 = philox4x32_10(c2, pv.key0, pv.key1);
        const double uz = u01(r1.x, r1.y), uj = u01(r1.z, r1.w);
        uacc = u01(r2.x, r2.y);
        const double zz = (pv.a - 1.0) * uz + 1.0;
        z = zz * zz / pv.a;
        const int j = (1 - half) * pv.W2 + (int)(uj * (double)pv.W2);
        sk = pv.map[t * pv.W + k];
        const int sj = pv.map[t * pv.W + j];
        const double* Xk = pv.X + (size_t)sk * pv.D;
        const double* Xj = pv.X + (size_t)sj * pv.D;
        for (int d = g; d < md.D; d += G) sl.theta[md.free_idx[d]] = Xj[d] + z * (Xk[d] - Xj[d]);
        for (int q = g; q < md.Dfull; q += G)
            if (md.inv_free[q] < 0) sl.theta[q] = md.fixed_val[q];
    }
    __syncthreads();
    if (task && wsub == 0) {
        if (lane == 0) s_lp[slot] = prior_seq(md, sl.theta);
        slot_setup(md, sl, lane);
    }
    __syncthreads();
    const double lp = task ? s_lp[slot] : -__builtin_inf();
    const bool live = task && lp == lp && lp != -__builtin_inf();
    if (PREC == 1 && live) {       // per-vector Kepler tables, one planet per wave
        for (int p = wsub; p < md.K; p += WPT)
            if (sl.planet[p].e <= KT_EMAX) sl.ktab[p][lane] = ktab_node(sl.planet[p], tile.tab, lane);
    }
    __syncthreads();
    double ll = block_lnl<PREC, WPT, false>(md, sl, tile, scan, live, wsub, nullptr, nullptr, resident);
    if (task && wsub == 0 && lane == 0) {
        int acc = 0;
        if (live) {
            atomicAdd(n_lnl, 1ull);
            if (!(ll == ll) || ll == __builtin_inf()) ll = -__builtin_inf();
            const double beta = pv.betas[t];
            const double tn = beta == 0.0 ? 0.0 : beta * ll;
            const double to = beta == 0.0 ? 0.0 : beta * pv.lnl[sk];
            const double lnalpha = (double)(md.D - 1) * log(z) + (tn - to) + (lp - pv.lnp[sk]);
            if (log(uacc) < lnalpha) {
                pv.lnl[sk] = ll;
                pv.lnp[sk] = lp;
                pv.accepted[t * pv.W + k] += 1;
                acc = 1;
            }
        }
        s_acc[slot] = acc;
    }
    __syncthreads();
    if (task && s_acc[slot]) {
        double* Xk = pv.X + (size_t)sk * pv.D;
        for (int d = g; d < md.D; d += G) Xk[d] = sl.theta[md.free_idx[d]];
    }
}

__device__ __forceinline__ uint32_t swap_partner(const PTView& pv, long long sweep, int h, int i) {
    return feistel_perm((uint32_t)i, (uint32_t)pv.W, pv.key0 ^ (uint32_t)sweep ^ 0x5bd1e995u,
                        pv.key1 ^ ((uint32_t)h * 0x27d4eb2fu) ^ (uint32_t)((unsigned long long)sweep >> 32) ^
                            pv.sysmix);
}

__device__ __forceinline__ double swap_logu(const PTView& pv, long long sweep, int h, int i) {
    U4 cu{(uint32_t)i, (uint32_t)sweep, (uint32_t)((unsigned long long)sweep >> 32) ^ pv.sysmix,
          0x40000000u | (uint32_t)h};
    U4 r = philox4x32_10(cu, pv.key0, pv.key1);
    return log(u01(r.x, r.y));
}

// Working copies for the LDS swap round: L[e] = lnL at position e (through
// the map), Mp = map, and the partner JP / log-uniform LU of every walker of
// every pair (h - 1) W + i, h = 1..T-1, drawn in parallel up front (they do
// not depend on the state).  The caller synchronises afterwards.
__device__ inline void pt_swap_stage(const PTView& pv, long long sweep, double* L, int* Mp, double* LU, int* JP,
                                     int nthr) {
    const int n = pv.T * pv.W, np = (pv.T - 1) * pv.W;
    for (int e = threadIdx.x; e < n; e += nthr) {
        const int m = pv.map[e];
        Mp[e] = m;
        L[e] = pv.lnl[m];
    }
    for (int e = threadIdx.x; e < np; e += nthr) {
        const int h = e / pv.W + 1, i = e - (h - 1) * pv.W;
        JP[e] = (int)swap_partner(pv, sweep, h, i);
        LU[e] = swap_logu(pv, sweep, h, i);
    }
}

// Adjacent-temperature swaps of one ensemble, hottest pair first, by one
// workgroup of `nthr` threads.  L (likelihood by position) and Mp (the map)
// are working copies: in LDS (pre == true, then LU/JP hold the pre-drawn
// uniforms and partners of every pair) or in global memory (pre == false,
// randoms drawn inline).  On return s_red[c] = (accepted, sum |log ratio|)
// of pair (c, c+1).  Walker states never move; only map entries do.
__device__ inline void pt_swap_round(const PTView& pv, long long sweep, double* L, int* Mp, const double* LU,
                                     const int* JP, bool pre, double (*s_red)[2], int nthr) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int T = pv.T, W = pv.W;
    for (int e = tid; e < 2 * 64; e += nthr) (&s_red[0][0])[e] = 0.0;
    __syncthreads();
    for (int h = T - 1; h >= 1; --h) {
        const int c = h - 1;
        const double dbeta = pv.betas[c] - pv.betas[h];
        double nacc = 0.0, dist = 0.0;
        for (int i = tid; i < W; i += nthr) {
            const int e = (h - 1) * W + i;
            const int jp = pre ? JP[e] : (int)swap_partner(pv, sweep, h, i);
            const double lu = pre ? LU[e] : swap_logu(pv, sweep, h, i);
            const int ih = h * W + i, jc = c * W + jp;
            const double lh = L[ih], lc = L[jc];
            double lr = dbeta * (lh - lc);
            if (!(lr == lr)) lr = -__builtin_inf();
            const bool acc = lu < lr;
            if (acc) {
                L[ih] = lc;
                L[jc] = lh;
                const int mh = Mp[ih];
                Mp[ih] = Mp[jc];
                Mp[jc] = mh;
            }
            nacc += acc ? 1.0 : 0.0;
            if (lr > -__builtin_inf() && lr < __builtin_inf()) dist += fabs(lr);
        }
        if (tid - lane < W) {          // whole waves only; waves without walkers add nothing
            nacc = wave_sum(nacc);
            dist = wave_sum(dist);
            if (lane == 0) { atomicAdd(&s_red[c][0], nacc); atomicAdd(&s_red[c][1], dist); }
        }
        if (!pre) __threadfence();
        __syncthreads();
    }
}

// Vousden, Farr & Mandel (2016) ladder dynamics from this sweep's swap
// acceptance `rate` (one thread): S_i = ln(T_i - T_{i-1}) moves by
// kappa (A_{i-1} - A_i), kappa = tau / (n + tau) / nu; the coldest and
// hottest temperatures stay fixed.
__device__ inline void pt_ladder_update(const PTView& pv, long long sweep, const double* rate) {
    const int T = pv.T;
    if (pv.adapt_mode != 0 || T < 3) return;
    const double kappa = pv.adapt_tau / ((double)sweep + pv.adapt_tau) / pv.adapt_nu;
    double Tprev_old = 1.0 / pv.betas[0];
    double acc_T = Tprev_old;
    for (int i = 1; i < T - 1; ++i) {
        const double Ti_old = 1.0 / pv.betas[i];
        const double dS = kappa * (rate[i - 1] - rate[i]);
        acc_T += (Ti_old - Tprev_old) * exp(dS);
        Tprev_old = Ti_old;
        pv.betas[i] = 1.0 / acc_T;
    }
}

}  // namespace emp
