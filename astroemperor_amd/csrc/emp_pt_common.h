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

// Stretch proposal of task b of a half-step: walker k = half W/2 + (b mod
// W/2) of temperature t = b / (W/2) moves towards/away from walker j of the
// complementary half, Y = X_j + z (X_k - X_j) with z = ((a - 1) u + 1)^2 / a
// (Goodman & Weare 2010, eq. 9).
struct Proposal {
    int t, k, sk, sj;     // temperature, walker, slots of k and j
    double z, uacc;       // stretch factor, acceptance uniform
};

__device__ __forceinline__ Proposal pt_propose(const PTView& pv, long long sweep, int half, int step, int b) {
    Proposal p;
    p.t = b / pv.W2;
    p.k = half * pv.W2 + b % pv.W2;
    const uint32_t shi = (uint32_t)((unsigned long long)sweep >> 32) ^ pv.sysmix;
    U4 c1{(uint32_t)(p.t * pv.W + p.k), (uint32_t)sweep, shi, (uint32_t)(step * 2 + half)};
    const U4 r1 = philox4x32_10(c1, pv.key0, pv.key1);
    U4 c2 = c1;
    c2.w |= 0x80000000u;
    const U4 r2 = philox4x32_10(c2, pv.key0, pv.key1);
    const double uz = u01(r1.x, r1.y), uj = u01(r1.z, r1.w);
    p.uacc = u01(r2.x, r2.y);
    const double zz = (pv.a - 1.0) * uz + 1.0;
    p.z = zz * zz / pv.a;
    const int j = (1 - half) * pv.W2 + (int)(uj * (double)pv.W2);
    p.sk = pv.map[p.t * pv.W + p.k];
    p.sj = pv.map[p.t * pv.W + j];
    return p;
}

// Free parameter f of the proposal (an explicit FMA, so that every caller
// forms the same value).
__device__ __forceinline__ double pt_prop_value(const double* Xk, const double* Xj, double z, int f) {
    return __builtin_fma(z, Xk[f] - Xj[f], Xj[f]);
}

// Prior of a proposal, in the order of prior_seq, read straight from the two
// walkers' rows, by one whole wave: the lanes evaluate the terms of the
// summation sequence md.pseq in parallel, with their loads in flight
// together, then the terms are added up in sequence order (a scalar read of
// each lane) with prior_seq's -inf return at the block ends, which a ballot
// marks, so the sum is bit-identical.
__device__ inline double prior_proposal_wave(const ModelDev& md, const double* Xk, const double* Xj, double z,
                                             int lane) {
#pragma clang fp contract(off)
    double lp = 0.0;
    for (int i0 = 0; i0 < md.npseq; i0 += 64) {
        const int i = i0 + lane;
        double v = 0.0;
        int kind = 0;
        if (i < md.npseq) {
            const PseqEntry& pe = md.pseq_e[i];
            kind = pe.kind;
            if (kind == 0) {
                v = prior_term(pe.term, pt_prop_value(Xk, Xj, z, pe.fa));
            } else if (kind == 1) {
                const double s3 = pe.fa >= 0 ? pt_prop_value(Xk, Xj, z, pe.fa) : pe.va;
                const double s4 = pe.fb >= 0 ? pt_prop_value(Xk, Xj, z, pe.fb) : pe.vb;
                const double x = s3 * s3 + s4 * s4;
                v = prior_term(pe.term, x);
            }
        }
        const unsigned long long chk = __ballot(i < md.npseq && kind == 2);
        const int n = min(64, md.npseq - i0);
        for (int k = 0; k < n; ++k) {
            if ((chk >> k) & 1ull) {
                if (lp == -__builtin_inf()) return lp;
            } else {
                lp += readlane_f64(v, k);
            }
        }
    }
    return lp;
}

// Indices b < n with lp[b] > -inf, in increasing order, into list[]; every
// thread of the workgroup (nthr threads, whole waves) gets the count.
// s_wcnt: LDS scratch of one int per wave.
__device__ inline int compact_live(const double* lp, int n, int* list, int* s_wcnt, int nthr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = nthr >> 6;
    int base = 0;
    for (int c0 = 0; c0 < n; c0 += nthr) {
        const int b = c0 + (int)threadIdx.x;
        const bool f = b < n && lp[b] > -__builtin_inf();
        const unsigned long long m = __ballot(f);
        const int within = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[wave] = __popcll(m);
        __syncthreads();
        int off = base, tot = 0;
        for (int w = 0; w < nw; ++w) {
            const int c = s_wcnt[w];
            if (w < wave) off += c;
            tot += c;
        }
        if (f) list[off + within] = b;
        __syncthreads();
        base += tot;
    }
    return base;
}

// Acceptance operands of one stretch proposal that are fixed before its
// likelihood is known: the Metropolis log ratio is
//   ((D - 1) ln z + (beta lnL' - beta lnL)) + (ln pi' - ln pi),
// with beta lnL taken as 0 at beta = 0.
struct TaskOps {
    int t, k, sk;         // temperature, walker, slot of the walker
    double lp;            // ln prior of the proposal
    double lz;            // (D - 1) ln z
    double dp;            // ln pi' - ln pi
    double to;            // beta lnL of the current position (0 at beta = 0)
    double beta, logu;    // inverse temperature, log of the acceptance uniform
};

__device__ __forceinline__ TaskOps pt_task_ops(const PTView& pv, const Proposal& pr, double lp, int D) {
    TaskOps o;
    o.t = pr.t;
    o.k = pr.k;
    o.sk = pr.sk;
    o.lp = lp;
    o.beta = pv.betas[pr.t];
    o.lz = (double)(D - 1) * log(pr.z);
    o.dp = lp - pv.lnp[pr.sk];
    o.to = o.beta == 0.0 ? 0.0 : o.beta * pv.lnl[pr.sk];
    o.logu = log(pr.uacc);
    return o;
}

// Likelihood and Metropolis acceptance of one slot's task whose proposal is
// already in sl.theta (free and fixed parameters) and whose operands are in
// s_ops[slot] (LDS, read only at the acceptance so that they do not occupy
// registers through the likelihood): per-slot constants, the mixed path's
// Kepler tables, lnL, acceptance, and the accepted position written back.
// Every evaluated likelihood adds one to *n_lnl.  The sin/cos table must be
// in `tile`; with `resident` all epochs are too.  All threads of the
// workgroup call this together (it contains barriers).  POST = false
// compiles the terms added after the MA out (models without them).
template <int PREC, int WPT, bool PF = true, bool POST = true, class SL>
__device__ void pt_eval_accept(const ModelDev& md, const PTView& pv, const TaskOps* s_ops, bool live, SL& sl,
                               TileLDS& tile, ScanLDS& scan, int* s_acc, int slot, int wsub, bool resident,
                               unsigned long long* n_lnl) {
    const int lane = threadIdx.x & 63;
    const int g = wsub * 64 + lane, G = 64 * WPT;
    EpochRegs e0{};
    const bool staged = resident && md.N <= TILE;
    if (PF && !staged) epoch_prefetch(md, 0, e0);      // the first tile's loads overlap the set-up
    __syncthreads();
    if (live && wsub == 0) slot_setup(md, sl, lane);
    __syncthreads();
    if (PREC == 1 && live) {       // per-vector Kepler tables, one planet per wave
        for (int p = wsub; p < md.K; p += WPT)
            if (sl.planet[p].ktab) ktab_fill(sl.planet[p], lane, sl.ktab[p][lane]);
    }
    __syncthreads();
    double ll = block_lnl<PREC, WPT, false, false, PF, POST>(md, sl, tile, scan, live, wsub, nullptr, nullptr,
                                                            resident, nullptr, (PF && !staged) ? &e0 : nullptr);
    if (live && wsub == 0 && lane == 0) {
        const TaskOps& ops = s_ops[slot];
        int acc = 0;
        atomicAdd(n_lnl, 1ull);
        if (!(ll == ll) || ll == __builtin_inf()) ll = -__builtin_inf();
        const double tn = ops.beta == 0.0 ? 0.0 : ops.beta * ll;
        const double lnalpha = (ops.lz + (tn - ops.to)) + ops.dp;
        if (ops.logu < lnalpha) {
            pv.lnl[ops.sk] = ll;
            pv.lnp[ops.sk] = ops.lp;
            pv.accepted[ops.t * pv.W + ops.k] += 1;
            acc = 1;
        }
        s_acc[slot] = acc;
    }
    __syncthreads();
    if (live && s_acc[slot]) {
        double* Xk = pv.X + (size_t)s_ops[slot].sk * pv.D;
        for (int d = g; d < md.D; d += G) Xk[d] = sl.theta[md.free_idx[d]];
    }
}

// One slot's task of a fused stretch half-step, for a task whose proposal
// prior lp is already known (b < 0: no task for this slot in this pass):
// proposal into LDS, then pt_eval_accept (without the tile prefetch: the
// batched-systems kernel that calls this is register-bound).
template <int PREC, int WPT, bool EXT = true, class SL>
__device__ void pt_half_task(const ModelDev& md, const PTView& pv, long long sweep, int half, int step, int b,
                             double lp, SL& sl, TileLDS& tile, ScanLDS& scan, int* s_acc, TaskOps* s_ops,
                             int slot, int wsub, bool resident, unsigned long long* n_lnl) {
    const int lane = threadIdx.x & 63;
    const int g = wsub * 64 + lane, G = 64 * WPT;
    const bool live = b >= 0 && lp > -__builtin_inf();
    if (live) {
        const Proposal pr = pt_propose(pv, sweep, half, step, b);
        const double* Xk = pv.X + (size_t)pr.sk * pv.D;
        const double* Xj = pv.X + (size_t)pr.sj * pv.D;
        for (int d = g; d < md.D; d += G) sl.theta[md.free_idx[d]] = pt_prop_value(Xk, Xj, pr.z, d);
        for (int q = g; q < md.Dfull; q += G)
            if (md.inv_free[q] < 0) sl.theta[q] = md.fixed_val[q];
        if (wsub == 0 && lane == 0) s_ops[slot] = pt_task_ops(pv, pr, lp, md.D);
    }
    pt_eval_accept<PREC, WPT, false, EXT>(md, pv, s_ops, live, sl, tile, scan, s_acc, slot, wsub, resident, n_lnl);
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
// are working copies; LU/JP hold the pre-drawn log-uniforms and partners of
// every pair (h - 1) W + i, h = 1..T-1 (in LDS, or in global memory with
// `global_mem`).  Once a pair is decided its slots are reused for its
// statistics (LU: |log ratio|, 0 if infinite; JP: accepted), which are
// summed per temperature pair after the last round, so a round carries only
// the decision and the swap.  On return s_red[c] = (accepted, sum |log
// ratio|) of pair (c, c+1).  Walker states never move; only map entries do.
// bet: the ladder (LDS copy).
__device__ inline void pt_swap_round(const PTView& pv, long long sweep, double* L, int* Mp, double* LU, int* JP,
                                     bool global_mem, double (*s_red)[2], const double* bet, int nthr) {
    (void)sweep;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
    const int T = pv.T, W = pv.W;
    for (int h = T - 1; h >= 1; --h) {
        const int c = h - 1;
        const double dbeta = bet[c] - bet[h];
        for (int i = tid; i < W; i += nthr) {
            const int e = (h - 1) * W + i;
            const int jp = JP[e];
            const double lu = LU[e];
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
            LU[e] = (lr > -__builtin_inf() && lr < __builtin_inf()) ? fabs(lr) : 0.0;
            JP[e] = acc ? 1 : 0;
        }
        if (global_mem) __threadfence();
        __syncthreads();
    }
    for (int h = 1 + wave; h < T; h += nw) {      // one wave per temperature pair
        double na = 0.0, di = 0.0;
        for (int i = lane; i < W; i += 64) {
            const int e = (h - 1) * W + i;
            na += (double)JP[e];
            di += LU[e];
        }
        na = wave_sum(na);
        di = wave_sum(di);
        if (lane == 0) {
            s_red[h - 1][0] = na;
            s_red[h - 1][1] = di;
        }
    }
    __syncthreads();
}

// Spread of the log-likelihood at each temperature after the swap round, for
// the specific-heat ladder: sig[t] = sample standard deviation (n - 1) of the
// finite lnL of temperature t's W slots, 0 with fewer than two.  One wave per
// temperature, two passes (mean, then squared deviations).  The caller
// publishes sig with a barrier.
__device__ inline void pt_lnl_spread(const PTView& pv, const double* L, double* sig, int nthr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = nthr >> 6;
    const int W = pv.W;
    for (int t = wave; t < pv.T; t += nw) {
        const double* l = L + (size_t)t * W;
        double s = 0.0, c = 0.0;
        for (int i = lane; i < W; i += 64)
            if (l[i] > -__builtin_inf() && l[i] < __builtin_inf()) {
                s += l[i];
                c += 1.0;
            }
        s = wave_sum(s);
        c = wave_sum(c);
        const double mean = c > 0.0 ? s / c : 0.0;
        double q = 0.0;
        for (int i = lane; i < W; i += 64)
            if (l[i] > -__builtin_inf() && l[i] < __builtin_inf()) {
                const double d = l[i] - mean;
                q += d * d;
            }
        q = wave_sum(q);
        if (lane == 0) sig[t] = c > 1.0 ? sqrt(q / (c - 1.0)) : 0.0;
    }
}

// Ladder dynamics after a sweep.  S_i = ln(T_i - T_{i-1}) (i = 1..T-2; the
// coldest and hottest temperatures stay fixed) moves by kappa times a
// balance term, kappa = tau / (n + tau) / nu:
//   mode 0, equal swap acceptance (Vousden, Farr & Mandel 2016):
//     A_{i-1} - A_i, A_k the acceptance of pair (k, k+1) this sweep;
//   mode 1, equal swap distance: (d_i - d_{i-1}) / (d_i + d_{i-1}), d_k the
//     mean |log swap ratio| of pair (k, k+1) this sweep;
//   mode 2, specific heat (equal thermodynamic length, the metric
//     sqrt(Var_beta lnL) dbeta = sqrt(C(beta)) / beta dbeta with the heat
//     capacity C = beta^2 Var_beta lnL; Kofke 2002, Crooks 2007):
//     (l_i - l_{i-1}) / (l_i + l_{i-1}), l_k = (beta_k - beta_{k+1})
//     (sig_k + sig_{k+1}) / 2 with sig from pt_lnl_spread;
// a pair that is closer than its neighbour (higher acceptance, smaller
// distance or length) widens.  The normalised terms lie in [-1, 1] like the
// acceptance difference, and are 0 when both distances are 0 or not finite.
// The new spacings (T_i - T_{i-1}) exp(dS_i) are formed in parallel from the
// ladder `bet` (LDS copy of the current betas), then one thread adds them up
// in order and writes pv.betas and bet.  rate and smd hold T - 1 pair
// values, sig T values (read in mode 2 only).  All nthr threads call this
// (it holds a barrier); s_d is LDS scratch of 64 doubles.
__device__ inline void pt_ladder_update(const PTView& pv, long long sweep, const double* rate, const double* smd,
                                        const double* sig, double* bet, double* s_d, int nthr) {
    const int T = pv.T;
    if (pv.adapt_mode < 0 || T < 3) return;
    const double kappa = pv.adapt_tau / ((double)sweep + pv.adapt_tau) / pv.adapt_nu;
    for (int i = (int)threadIdx.x + 1; i < T - 1; i += nthr) {
        double dS;
        if (pv.adapt_mode == 0) {
            dS = kappa * (rate[i - 1] - rate[i]);
        } else {
            double lo, hi;                  // pair (i-1, i), pair (i, i+1)
            if (pv.adapt_mode == 1) {
                lo = smd[i - 1];
                hi = smd[i];
            } else {
                lo = (bet[i - 1] - bet[i]) * 0.5 * (sig[i - 1] + sig[i]);
                hi = (bet[i] - bet[i + 1]) * 0.5 * (sig[i] + sig[i + 1]);
            }
            const double s = lo + hi;
            dS = (s > 0.0 && s < __builtin_inf()) ? kappa * (hi - lo) / s : 0.0;
        }
        const double spacing = 1.0 / bet[i] - 1.0 / bet[i - 1];
        s_d[i] = spacing * exp(dS);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc_T = 1.0 / bet[0];
        for (int i = 1; i < T - 1; ++i) {
            acc_T += s_d[i];
            const double b = 1.0 / acc_T;
            bet[i] = b;
            pv.betas[i] = b;
        }
    }
}

}  // namespace emp
