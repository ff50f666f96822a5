// Parallel-tempering engine: the reddemcee sweep (emcee red/blue stretch
// moves at every temperature, adjacent-temperature swaps, ladder adaptation)
// as device-resident kernels, replacing reddemcee.PTSampler.run_mcmc
// (support/endit_reddemcee.scr:3; constructed at emp/emp.py:576-593).
//
// Layout.  Walker states live in fixed slots: X[slot][D], lnl[slot],
// lnp[slot].  map[t][i] names the slot that holds walker i of temperature t.
// A temperature swap exchanges two map entries, so no walker state moves;
// the stretch kernels and the chain record read and write through the map.
//
// One sweep = nsteps x {half 0, half 1} fused stretch half-steps, one swap
// round (with ladder update and the sweep counter), one history record.  The
// sweep counter lives on the device so a captured hipGraph of one sweep can
// be replayed unchanged; every random number is
// Philox(seed; walker, sweep, step/half, purpose), so a run is a pure
// function of (seed, initial state) and resumes bitwise.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "emp_pt_common.h"

using namespace emp;

struct PTDev {
    int T, W, W2, D;
    double a;
    uint32_t key0, key1;
    int adapt_mode;
    double adapt_tau, adapt_nu;
    int64_t max_sweeps;
    int nsteps;         // stretch steps per sweep
    int64_t max_rows;   // history rows: one per step (max_sweeps x nsteps)
    int record;
    int swap_stream;    // the swap rounds stream rows through LDS (the global-memory instance, W <= 1024)
    double* X;          // [T*W slots][D]
    double* lnl;        // [slots]
    double* lnp;        // [slots]
    int* map;           // [T][W] -> slot
    double* swap_l;     // [T][W] scratch for the global-memory swap path
    int* rec_owed;      // 1: the previous sweep's state is still to be recorded (set by the swap round)
    double* swap_lu;    // [T-1][W] log uniforms of this sweep's swap round, drawn by the last prior pass
    int* swap_jp;       // [T-1][W] partners of this sweep's swap round
    double* betas;      // [T]
    long long* accepted;  // [T][W] (logical positions)
    long long* sweep;     // [2] sweeps completed, likelihoods evaluated
    double* prop_lp;      // [T W/2] prior of each task's proposal (gradient move)
    TaskOps* rec;         // [T W/2] acceptance operands of each task's proposal (lp = -inf: outside the support)
    double* prow;         // [T W/2][D] the proposals inside the prior support
    double* h_chain;    // [rows][T][W][D], one row per step (the state after it; a sweep's last
                        // row is taken after its swap round)
    double* h_lnl;      // [rows][T][W]
    double* h_lnp;      // [rows][T][W]
    double* h_betas;    // [S][T]
    double* h_tsw;      // [S][T-1]
    double* h_smd;      // [S][T-1]
    // Gradient (Langevin) move, allocated when the move is enabled.
    double* lg_eps;     // [T] step eps of each temperature
    double* lg_g;       // [slots][D] grad lnL at each slot's position
    double* lg_l;       // [slots] lnL from the adjoint kernel at each slot's position
    double* lg_L1;      // [T][D][D] Cholesky factor L of the complementary half's covariance, L1[k D + i] = L_ik
    double* lg_L2;      // [T][D][D] the same factor, L2[k D + i] = L_ki
    int* lg_act;        // [T][D] parameter has a nonzero spread in the complementary half
    double* lg_Xp;      // [T W/2][D] proposals
    double* lg_gp;      // [T W/2][D] grad lnL at the proposals
    double* lg_lp;      // [T W/2] lnL at the proposals
    double* lg_lq;      // [T W/2] log density of the forward proposal (up to the shared constant)
    double* lg_scr;     // [T W][N] moving-average adjoint scratch (null without the moving average)
    int* lg_stale;      // [slots] 1: lg_g / lg_l are not those of the slot's position (a stretch move
                        // moved it); cleared by the adjoint launch that recomputes them
};

struct EmpPT {
    EmpModel* model;
    EmpPTDesc desc;
    PTDev dev;
    std::vector<void*> allocations;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;      // stretch sweep
    hipGraphExec_t graph_grad = nullptr;      // gradient-move sweep
    hipGraphExec_t graph_run[5] = {};         // 2, 4, 8, 16, 32 stretch sweeps in one graph
    double w_stretch = 1.0, w_grad = 0.0;     // move weights (one move per sweep)
    int wpt = 1;
    int grad_wpt = 2;            // waves per vector of the adjoint launches (one shape for both, so a
                                 // slot's gradient does not depend on which launch computed it)
    size_t swap_lds = 0;         // dynamic LDS of the swap kernel, 0 = global-memory instance
    size_t swap_lds_fit = 0;     // what the LDS instance needs when the ensemble fits, else 0
    long long host_sweeps = 0;   // mirror of the device counter
    long long hist_base = 0;     // first sweep whose history the engine holds (a restore from a
                                 // checkpoint without chain rows starts the history there)
    double last_ms_per_sweep = -1.0, last_ms_half = -1.0;
};

static __host__ __device__ inline PTView view_of(const PTDev& pd) {
    return pt_view(pd.T, pd.W, pd.D, pd.a, pd.key0, pd.key1, 0u, pd.adapt_mode, pd.adapt_tau, pd.adapt_nu, pd.X,
                   pd.lnl, pd.lnp, pd.map, pd.betas, pd.accepted);
}

// ---------------------------------------------------------------------------
// Half-step, part 1: every proposal of the moving half (one wave per walker,
// its lanes on the prior's terms), its prior, and its record: the acceptance
// operands that do not depend on the likelihood (ln prior -inf marks a
// proposal outside the prior support) and, inside the support, the proposal
// row.  The loads for the operands are issued with the rows they sit beside.
// ---------------------------------------------------------------------------
constexpr int PRIOR_BLOCK = 64;     // one wave per workgroup: the passes spread over more CUs
constexpr int PRIOR_WAVES = PRIOR_BLOCK / 64;

// Chain record of logical walker w (temperature-major) into history row s.
__device__ __forceinline__ void record_walker(const PTDev& pd, long long s, int w, int lane) {
    const size_t nw = (size_t)pd.T * pd.W;
    const int slot = pd.map[w];
    double* cx = pd.h_chain + (size_t)s * nw * pd.D + (size_t)w * pd.D;
    const double* x = pd.X + (size_t)slot * pd.D;
    for (int d = lane; d < pd.D; d += 64) cx[d] = x[d];
    if (lane == 0) {
        pd.h_lnl[(size_t)s * nw + w] = pd.lnl[slot];
        pd.h_lnp[(size_t)s * nw + w] = pd.lnp[slot];
    }
}

__global__ __launch_bounds__(PRIOR_BLOCK) void k_pt_prior(ModelDev md, PTDev pd, int half, int step, int draw_swaps,
                                                          int first) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * PRIOR_WAVES + wave;
    const PTView pv = view_of(pd);
    if (first && pd.record && (step > 0 || *pd.rec_owed) && b < pd.T * pd.W2) {
        // A step's first pass records the state before it (nothing has moved
        // since): the previous step's, or the previous sweep's after its
        // swap round.  Walker k of half 0 and its half-1 twin.
        const long long s = *pd.sweep * pd.nsteps + step - 1;
        if (s >= 0 && s < pd.max_rows) {
            const int t = b / pd.W2, k = b % pd.W2;
            record_walker(pd, s, t * pd.W + k, lane);
            record_walker(pd, s, t * pd.W + k + pd.W2, lane);
        }
    }
    if (draw_swaps) {
        // The swap round's pairings and uniforms depend on (seed, sweep, pair)
        // only; this grid has T W / 2 x 64 threads >= (T - 1) W pairs.
        const int e = blockIdx.x * PRIOR_BLOCK + threadIdx.x;
        if (e < (pd.T - 1) * pd.W) {
            const long long sweep = *pd.sweep;
            const int h = e / pd.W + 1, i = e - (h - 1) * pd.W;
            pd.swap_jp[e] = (int)swap_partner(pv, sweep, h, i);
            pd.swap_lu[e] = swap_logu(pv, sweep, h, i);
        }
    }
    if (b >= pd.T * pd.W2) return;            // whole waves
    // What does not depend on the random draws is loaded first, so that it
    // is in flight while the counter is read and Philox runs.
    const int t = b / pd.W2, k = half * pd.W2 + b % pd.W2;
    const int sk = pd.map[t * pd.W + k];
    const double beta = pd.betas[t], lnl_old = pd.lnl[sk], lnp_old = pd.lnp[sk];
    const Proposal pr = pt_propose(pv, *pd.sweep, half, step, b);
    const double* Xk = pd.X + (size_t)pr.sk * pd.D;
    const double* Xj = pd.X + (size_t)pr.sj * pd.D;
    const double lp = prior_proposal_wave(md, Xk, Xj, pr.z, lane);
    const bool live = lp > -__builtin_inf();
    if (live) {
        double* row = pd.prow + (size_t)b * pd.D;
        for (int d = lane; d < pd.D; d += 64) row[d] = pt_prop_value(Xk, Xj, pr.z, d);
    }
    if (lane == 0) {
        TaskOps o;
        o.t = pr.t;
        o.k = pr.k;
        o.sk = pr.sk;
        o.lp = lp;
        o.beta = beta;
        o.lz = (double)(pd.D - 1) * log(pr.z);
        o.dp = lp - lnp_old;
        o.to = beta == 0.0 ? 0.0 : beta * lnl_old;
        o.logu = log(pr.uacc);
        pd.rec[b] = o;
    }
}

// ---------------------------------------------------------------------------
// Half-step, part 2: likelihood and acceptance of the proposals inside the
// prior support, one slot of the workgroup per task; a workgroup whose tasks
// all lie outside leaves at once.
// ---------------------------------------------------------------------------
template <int PREC, int WPT, bool POST>
__global__ __launch_bounds__(BLOCK) void k_pt_half(ModelDev md, PTDev pd, int half, int step) {
    constexpr int SLOTS = NWAVES / WPT;
    __shared__ SlotLDS<SLOTS> s_slot[SLOTS];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    __shared__ int s_acc[SLOTS];
    __shared__ TaskOps s_ops[SLOTS];
    (void)half;
    (void)step;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave / WPT, wsub = wave % WPT;
    const int g = wsub * 64 + lane, G = 64 * WPT;
    const int b = blockIdx.x * SLOTS + slot;
    const bool live = b < pd.T * pd.W2 && pd.rec[b].lp > -__builtin_inf();
    if (SLOTS == 1) {
        if (!live) return;                   // the whole workgroup
    } else if (!__syncthreads_or(live)) {
        return;
    }
    const PTView pv = view_of(pd);
    SlotLDS<SLOTS>& sl = s_slot[slot];
    tab_load(md, s_tile);      // published by pt_eval_accept's first barrier
    if (live) {
        if (wsub == 0 && lane == 0) s_ops[slot] = pd.rec[b];
        const double* row = pd.prow + (size_t)b * pd.D;
        for (int j = g; j < md.Dfull; j += G) {
            const int fi = md.inv_free[j];
            sl.theta[j] = fi >= 0 ? row[fi] : md.fixed_val[j];
        }
    }
    pt_eval_accept<PREC, WPT, true, POST>(md, pv, s_ops, live, sl, s_tile, s_scan, s_acc, slot, wsub, false,
                              (unsigned long long*)(pd.sweep + 1));
    // an accepted stretch move leaves the slot's gradient out of date
    if (pd.lg_stale && live && wsub == 0 && lane == 0 && s_acc[slot]) pd.lg_stale[s_ops[slot].sk] = 1;
}

// ---------------------------------------------------------------------------
// Adjacent-temperature swaps (hottest pair first), swap statistics, ladder
// adaptation and the sweep counter.  One workgroup.  The pairings and
// log-uniforms were drawn by the sweep's last prior pass.  With USE_LDS the
// map, the per-position likelihoods and those randoms are staged in LDS.
// Otherwise, for W up to SWAP_BLOCK, the rows stream through LDS
// (swap_rounds_stream); for wider ensembles the round works in global memory.
// ---------------------------------------------------------------------------
constexpr int SWAP_BLOCK = 1024;
constexpr int SWAP_UNROLL = 4;

// The swap rounds of an ensemble whose positions do not fit LDS, W <=
// SWAP_BLOCK.  Round h reads only temperatures h and h - 1, and temperature
// h is final after it, so three rows of (lnL, map entry) rotate through LDS,
// row t in buffer t mod 3: the pair's two rows and the next row, staged
// while the pair is decided from loads issued a round earlier (as are the
// pair's partner and log-uniform).  One barrier per round.  Thread i owns
// position i of every row and pair i of every round; the decisions, the
// pair statistics left in swap_lu / swap_jp and the final map equal those of
// pt_swap_round.  Final rows go to the map and their likelihoods by position
// to swap_l (the specific-heat spread reads them).
__device__ inline void swap_rounds_stream(const PTDev& pd, const PTView& pv, double (*s_l)[SWAP_BLOCK],
                                          int (*s_m)[SWAP_BLOCK], double (*s_red)[2], const double* bet) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = SWAP_BLOCK >> 6;
    const int T = pd.T, W = pd.W;
    const bool own = tid < W;
    if (own) {
        for (int t = T - 1; t >= 0 && t >= T - 2; --t) {
            const int m = pd.map[(size_t)t * W + tid];
            s_m[t % 3][tid] = m;
            s_l[t % 3][tid] = pd.lnl[m];
        }
    }
    int mreg = 0, jreg = 0;
    double lreg = 0.0, ureg = 0.0;
    if (own && T >= 3) {
        mreg = pd.map[(size_t)(T - 3) * W + tid];
        lreg = pd.lnl[mreg];
    }
    if (own && T >= 2) {
        jreg = pd.swap_jp[(size_t)(T - 2) * W + tid];
        ureg = pd.swap_lu[(size_t)(T - 2) * W + tid];
    }
    __syncthreads();
    for (int h = T - 1; h >= 1; --h) {
        const int c = h - 1, bh = h % 3, bc = c % 3;
        int mnext = 0, jnext = 0;
        double lnext = 0.0, unext = 0.0;
        if (own) {
            if (h >= 2) {                 // stage row h - 2, start loading row h - 3 and pair h - 1
                s_m[(h - 2) % 3][tid] = mreg;
                s_l[(h - 2) % 3][tid] = lreg;
                jnext = pd.swap_jp[(size_t)(h - 2) * W + tid];
                unext = pd.swap_lu[(size_t)(h - 2) * W + tid];
            }
            if (h >= 3) mnext = pd.map[(size_t)(h - 3) * W + tid];
            const size_t e = (size_t)(h - 1) * W + tid;
            const int jp = jreg;
            const double lh = s_l[bh][tid], lc = s_l[bc][jp];
            double lr = (bet[c] - bet[h]) * (lh - lc);
            if (!(lr == lr)) lr = -__builtin_inf();
            const bool acc = ureg < lr;
            if (acc) {
                s_l[bh][tid] = lc;
                s_l[bc][jp] = lh;
                const int mh = s_m[bh][tid];
                s_m[bh][tid] = s_m[bc][jp];
                s_m[bc][jp] = mh;
            }
            pd.swap_lu[e] = (lr > -__builtin_inf() && lr < __builtin_inf()) ? fabs(lr) : 0.0;
            pd.swap_jp[e] = acc ? 1 : 0;
            if (h >= 3) lnext = pd.lnl[mnext];
        }
        __syncthreads();
        if (own) {                        // temperature h is final
            pd.map[(size_t)h * W + tid] = s_m[bh][tid];
            pd.swap_l[(size_t)h * W + tid] = s_l[bh][tid];
        }
        mreg = mnext;
        lreg = lnext;
        jreg = jnext;
        ureg = unext;
    }
    if (own) {
        pd.map[tid] = s_m[0][tid];
        pd.swap_l[tid] = s_l[0][tid];
    }
    __threadfence();
    __syncthreads();
    for (int h = 1 + wave; h < T; h += nw) {      // one wave per temperature pair
        double na = 0.0, di = 0.0;
        for (int i = lane; i < W; i += 64) {
            const size_t e = (size_t)(h - 1) * W + i;
            na += (double)pd.swap_jp[e];
            di += pd.swap_lu[e];
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

template <bool USE_LDS>
__global__ __launch_bounds__(SWAP_BLOCK) void k_pt_swap(PTDev pd) {
    extern __shared__ double s_dyn[];
    __shared__ double s_red[64][2];
    __shared__ double s_rate[64], s_smd[64], s_sig[64], s_bet[64], s_ld[64];
    const int tid = threadIdx.x;
    const long long sweep = *pd.sweep;
    const PTView pv = view_of(pd);
    const int T = pd.T, W = pd.W;
    const int n = T * W, np = (T - 1) * W;
    double* L = USE_LDS ? s_dyn : pd.swap_l;
    double* LU = s_dyn + n;                          // [T-1][W] log uniforms (LDS path)
    int* Mp = USE_LDS ? (int*)(s_dyn + n + np) : pd.map;
    int* JP = (int*)(s_dyn + n + np) + n;            // [T-1][W] partners (LDS path)
    for (int e = tid; e < T; e += SWAP_BLOCK) s_bet[e] = pd.betas[e];
    if (USE_LDS) {
        // Up to SWAP_UNROLL entries per thread with all independent loads in
        // flight together (map and the swap randoms, then the lnL through the
        // map), instead of one dependent pair per loop trip.
        for (int e0 = 0; e0 < n; e0 += SWAP_UNROLL * SWAP_BLOCK) {     // np < n
            int m[SWAP_UNROLL], jp[SWAP_UNROLL];
            double lu[SWAP_UNROLL], l[SWAP_UNROLL];
#pragma unroll
            for (int r = 0; r < SWAP_UNROLL; ++r) {
                const int e = e0 + r * SWAP_BLOCK + tid;
                m[r] = e < n ? pd.map[e] : 0;
                lu[r] = e < np ? pd.swap_lu[e] : 0.0;
                jp[r] = e < np ? pd.swap_jp[e] : 0;
            }
#pragma unroll
            for (int r = 0; r < SWAP_UNROLL; ++r) {
                const int e = e0 + r * SWAP_BLOCK + tid;
                l[r] = e < n ? pd.lnl[m[r]] : 0.0;
            }
#pragma unroll
            for (int r = 0; r < SWAP_UNROLL; ++r) {
                const int e = e0 + r * SWAP_BLOCK + tid;
                if (e < n) {
                    Mp[e] = m[r];
                    L[e] = l[r];
                }
                if (e < np) {
                    LU[e] = lu[r];
                    JP[e] = jp[r];
                }
            }
        }
    }
    if constexpr (!USE_LDS) {
        if (pd.swap_stream) {
            __shared__ double s_sl[3][SWAP_BLOCK];
            __shared__ int s_sm[3][SWAP_BLOCK];
            __syncthreads();
            swap_rounds_stream(pd, pv, s_sl, s_sm, s_red, s_bet);
        } else {
            for (int e = tid; e < n; e += SWAP_BLOCK) L[e] = pd.lnl[pd.map[e]];
            __threadfence();
            __syncthreads();
            pt_swap_round(pv, sweep, L, Mp, pd.swap_lu, pd.swap_jp, true, s_red, s_bet, SWAP_BLOCK);
        }
    } else {
        __syncthreads();
        pt_swap_round(pv, sweep, L, Mp, LU, JP, false, s_red, s_bet, SWAP_BLOCK);
    }
    if (tid < T - 1) {
        const double na = s_red[tid][0], di = s_red[tid][1];
        s_rate[tid] = na / W;
        s_smd[tid] = di / W;
        if (sweep < pd.max_sweeps) {
            pd.h_tsw[(size_t)sweep * (T - 1) + tid] = na / W;
            pd.h_smd[(size_t)sweep * (T - 1) + tid] = di / W;
        }
    }
    if (USE_LDS)
        for (int e = tid; e < n; e += SWAP_BLOCK) pd.map[e] = Mp[e];
    if (pd.adapt_mode == 2) pt_lnl_spread(pv, L, s_sig, SWAP_BLOCK);
    __syncthreads();
    pt_ladder_update(pv, sweep, s_rate, s_smd, s_sig, s_bet, s_ld, SWAP_BLOCK);
    __syncthreads();
    if (tid < T && sweep < pd.max_sweeps) pd.h_betas[(size_t)sweep * T + tid] = s_bet[tid];
    if (tid == 0) {
        *pd.sweep = sweep + 1;
        if (pd.record) *pd.rec_owed = 1;
    }
}

// The swap round's randoms for a sweep without stretch half-steps (the
// gradient-move sweep).
__global__ void k_pt_swap_draw(PTDev pd) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (pd.T - 1) * pd.W) return;
    const PTView pv = view_of(pd);
    const long long sweep = *pd.sweep;
    const int h = e / pd.W + 1, i = e - (h - 1) * pd.W;
    pd.swap_jp[e] = (int)swap_partner(pv, sweep, h, i);
    pd.swap_lu[e] = swap_logu(pv, sweep, h, i);
}

// History row of the state before step `step` of the current sweep (step 0:
// the previous sweep's final state).  mode 0: the end of a run, which also
// clears the debt; 1: the start of a gradient sweep, only if still owed;
// 2: before an inner step of a gradient sweep.
__global__ void k_pt_record(PTDev pd, int mode, int step) {
    const long long s = *pd.sweep * pd.nsteps + step - 1;
    if (mode == 1 && !*pd.rec_owed) return;
    if (mode == 0 && blockIdx.x == 0 && threadIdx.x == 0) *pd.rec_owed = 0;
    if (s < 0 || s >= pd.max_rows) return;
    const size_t nw = (size_t)pd.T * pd.W, nx = nw * pd.D;
    double* cx = pd.h_chain + (size_t)s * nx;
    double* cl = pd.h_lnl + (size_t)s * nw;
    double* cp = pd.h_lnp + (size_t)s * nw;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nx; e += (size_t)gridDim.x * blockDim.x) {
        const size_t w = e / pd.D, d = e - w * pd.D;
        const int slot = pd.map[w];
        cx[e] = pd.X[(size_t)slot * pd.D + d];
        if (d == 0) { cl[w] = pd.lnl[slot]; cp[w] = pd.lnp[slot]; }
    }
}

__global__ void k_mask_lnl(double* lnl, const double* lnp, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(lnp[i] > -__builtin_inf())) lnl[i] = -__builtin_inf();
}

__global__ void k_identity(int* map, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[i] = (int)i;
}

__global__ void k_fill_i32(int* a, int64_t n, int v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// Evidence ingredients from the recorded lnL history, sweeps [s0, s1), all
// walkers, one workgroup per temperature t: the mean lnL, and for t >= 1 the
// stepping stone from t to the colder t - 1,
//     log mean exp((beta_{t-1} - beta_t) lnL_t)
// (the current ladder), by a max pass and a sum pass, with the delta-method
// variance of that log (mean of exp(2 x) over the square of the mean of
// exp(x), minus one, over n; autocorrelation ignored).  Every sum runs in a
// fixed order (thread stride, then a fixed tree), so the result is
// reproducible.
// ---------------------------------------------------------------------------
constexpr int MOM_BLOCK = 256;

__device__ inline double block_tree(double v, double* s_tmp, bool take_max) {
    const int tid = threadIdx.x;
    s_tmp[tid] = v;
    __syncthreads();
    for (int o = MOM_BLOCK / 2; o > 0; o >>= 1) {
        if (tid < o) s_tmp[tid] = take_max ? fmax(s_tmp[tid], s_tmp[tid + o]) : s_tmp[tid] + s_tmp[tid + o];
        __syncthreads();
    }
    const double r = s_tmp[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(MOM_BLOCK) void k_pt_lnl_moments(PTDev pd, long long s0, long long s1, double* mean,
                                                              double* lse, double* var) {
    __shared__ double s_tmp[MOM_BLOCK];
    const int t = blockIdx.x, T = pd.T, W = pd.W;
    const long long n = (s1 - s0) * W;
    const double db = t > 0 ? pd.betas[t - 1] - pd.betas[t] : 0.0;
    double sum = 0.0, mx = -__builtin_inf();
    for (long long i = threadIdx.x; i < n; i += MOM_BLOCK) {
        const long long s = s0 + i / W;
        const double l = pd.h_lnl[((size_t)s * T + t) * W + (size_t)(i % W)];
        sum += l;
        mx = fmax(mx, db * l);
    }
    sum = block_tree(sum, s_tmp, false);
    mx = block_tree(mx, s_tmp, true);
    const bool fin = mx > -__builtin_inf() && mx < __builtin_inf();
    double se = 0.0, se2 = 0.0;
    if (t > 0 && fin)
        for (long long i = threadIdx.x; i < n; i += MOM_BLOCK) {
            const long long s = s0 + i / W;
            const double x = exp(db * pd.h_lnl[((size_t)s * T + t) * W + (size_t)(i % W)] - mx);
            se += x;
            se2 += x * x;
        }
    se = block_tree(se, s_tmp, false);
    se2 = block_tree(se2, s_tmp, false);
    if (threadIdx.x == 0) {
        const double dn = (double)n;
        mean[t] = sum / dn;
        lse[t] = t == 0 ? 0.0 : (fin ? mx + log(se / dn) : mx);
        var[t] = (t > 0 && fin) ? ((se2 / dn) / ((se / dn) * (se / dn)) - 1.0) / dn : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Gradient move: ensemble-preconditioned Metropolis-adjusted Langevin
// (Roberts & Tweedie 1996; the preconditioner is the covariance of the
// complementary half of the ensemble, Leimkuhler, Matthews & Weare 2018).
// Walker k of half h at temperature t proposes
//     y = x + L u,   u = (eps^2 / 2) beta_t L^T g(x) + eps xi,   xi ~ N(0, I),
// with L L^T the covariance of the other half's walkers at t (population
// covariance, diagonal times 1 + 1e-6; a parameter without spread there does
// not move) and g = grad lnL, and accepts with the Metropolis-Hastings ratio of
// the tempered posterior beta_t lnL + ln prior times q(x | y) / q(y | x).  The
// lnL values and gradients at both points come from the FP64 adjoint kernel.
// A non-finite gradient switches the drift off at that point; the drift stays
// a function of the point alone, so the ratio remains exact.  If the
// factorisation meets a non-positive pivot, L is the diagonal of standard
// deviations for that temperature.
//
// Random numbers: Philox counters (t W + k, sweep_lo, sweep_hi,
// 2^29 | (2 step + half) << 8 | r); r < 64 gives the normals of parameters
// 2r and 2r + 1 (Box-Muller), r = 255 the acceptance uniform.  The move of a
// sweep is chosen on the host from counter (0, sweep_lo, sweep_hi, 2^28).
// ---------------------------------------------------------------------------
constexpr uint32_t LG_TAG = 0x20000000u;
constexpr uint32_t MOVE_TAG = 0x10000000u;
constexpr int LG_BLOCK = 1024;
constexpr int LG_CHUNK = 64;
constexpr int LG_NP = MAXD * (MAXD + 1) / 2;
constexpr int LG_PAIRS = (LG_NP + LG_BLOCK - 1) / LG_BLOCK;
constexpr int LG_STAGE = (LG_CHUNK * MAXD + LG_BLOCK - 1) / LG_BLOCK;
constexpr double LG_RIDGE = 1e-6;

__device__ __forceinline__ U4 lg_draw(const PTDev& pd, long long sweep, int half, int step, int t, int k, uint32_t r) {
    const U4 c{(uint32_t)(t * pd.W + k), (uint32_t)sweep, (uint32_t)((unsigned long long)sweep >> 32),
               LG_TAG | ((uint32_t)(2 * step + half) << 8) | r};
    return philox4x32_10(c, pd.key0, pd.key1);
}

__device__ inline double lg_normal(const PTDev& pd, long long sweep, int half, int step, int t, int k, int j) {
    const U4 r = lg_draw(pd, sweep, half, step, t, k, (uint32_t)(j >> 1));
    const double rad = sqrt(-2.0 * log(u01(r.x, r.y)));
    double sn, cs;
    sincospi(2.0 * u01(r.z, r.w), &sn, &cs);
    return (j & 1) ? rad * sn : rad * cs;
}

// Value of parameter k (uniform across the wave) from the lane that holds it:
// lane k & 63, register k >> 6.
__device__ __forceinline__ double lg_bcast(const double v[2], int k) {
    const double x = k < 64 ? v[0] : v[1];
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)bits, k & 63);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), k & 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// (L^T g)_j for the two parameters j = lane, lane + 64 of this lane.
__device__ inline void lg_lt_times(const double* L2, const double g[2], int D, int lane, double v[2]) {
    v[0] = v[1] = 0.0;
    for (int k = 0; k < D; ++k) {
        const double gk = lg_bcast(g, k);
        for (int h = 0; h < 2; ++h) {
            const int j = lane + 64 * h;
            if (j < D) v[h] = __builtin_fma(L2[(size_t)k * D + j], gk, v[h]);
        }
    }
}

// Covariance of the complementary half at each temperature and its Cholesky
// factor; one workgroup of 1024 threads per temperature.  The walkers are
// staged through LDS in chunks (the next chunk's rows are loaded into
// registers while the current one is summed); the covariance and its
// factorisation stay in LDS as a packed lower triangle (row i starts at
// i (i + 1) / 2).  Each thread owns the same entries in both phases.  The
// factorisation is right-looking on the unscaled columns,
// A_ik -= A_ij A_kj / A_jj for j < k <= i, with one barrier per column; the
// columns are scaled by 1 / sqrt(A_jj) at the end.  Deterministic: every sum
// runs in a fixed order.  Block 0 of the first half-step of the sweep also
// counts the likelihoods of the current-position gradient launch.
__global__ __launch_bounds__(LG_BLOCK) void k_lg_precond(PTDev pd, int half, int step) {
    __shared__ double s_x[LG_CHUNK][MAXD];
    __shared__ double s_a[LG_NP];
    __shared__ double s_mean[MAXD], s_diag[MAXD];
    const int t = blockIdx.x, D = pd.D, W2 = pd.W2, tid = threadIdx.x;
    const int* cmap = pd.map + t * pd.W + (1 - half) * W2;
    double* A = pd.lg_L2 + (size_t)t * D * D;
    double* L1 = pd.lg_L1 + (size_t)t * D * D;
    if (t == 0 && tid == 0 && half == 0 && step == 0)
        atomicAdd((unsigned long long*)(pd.sweep + 1), (unsigned long long)pd.T * pd.W);
    const int np = D * (D + 1) / 2;
    // this thread's entries p = tid + m LG_BLOCK of the packed triangle, as i << 16 | j
    int pij[LG_PAIRS];
    {
        int i = 0, j = tid;
        while (j > i) { j -= i + 1; ++i; }
#pragma unroll
        for (int m = 0; m < LG_PAIRS; ++m) {
            pij[m] = i << 16 | j;
            j += LG_BLOCK;
            while (j > i) { j -= i + 1; ++i; }
        }
    }
    double pre[LG_STAGE];
    auto fetch = [&](int w0) {
        const int nw = min(LG_CHUNK, W2 - w0);
#pragma unroll
        for (int r = 0; r < LG_STAGE; ++r) {
            const int e = tid + r * LG_BLOCK;
            if (e < nw * D) {
                const int w = e / D, j = e - w * D;
                pre[r] = pd.X[(size_t)cmap[w0 + w] * D + j];
            }
        }
    };
    auto stage = [&](int w0, bool centre) {
        const int nw = min(LG_CHUNK, W2 - w0);
#pragma unroll
        for (int r = 0; r < LG_STAGE; ++r) {
            const int e = tid + r * LG_BLOCK;
            if (e < nw * D) {
                const int w = e / D, j = e - w * D;
                s_x[w][j] = centre ? pre[r] - s_mean[j] : pre[r];
            }
        }
    };
    // mean (thread j < D sums column j over the walkers in order)
    double sum = 0.0;
    fetch(0);
    for (int w0 = 0; w0 < W2; w0 += LG_CHUNK) {
        const int nw = min(LG_CHUNK, W2 - w0);
        __syncthreads();
        stage(w0, false);
        if (w0 + LG_CHUNK < W2) fetch(w0 + LG_CHUNK);
        __syncthreads();
        if (tid < D)
            for (int w = 0; w < nw; ++w) sum += s_x[w][tid];
    }
    if (tid < D) s_mean[tid] = sum / (double)W2;
    // lower triangle of the covariance
    double acc[LG_PAIRS];
#pragma unroll
    for (int m = 0; m < LG_PAIRS; ++m) acc[m] = 0.0;
    fetch(0);
    for (int w0 = 0; w0 < W2; w0 += LG_CHUNK) {
        const int nw = min(LG_CHUNK, W2 - w0);
        __syncthreads();
        stage(w0, true);
        if (w0 + LG_CHUNK < W2) fetch(w0 + LG_CHUNK);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < LG_PAIRS; ++m) {
            if (tid + m * LG_BLOCK < np) {
                const int i = pij[m] >> 16, j = pij[m] & 0xffff;
                double a = acc[m];
                for (int w = 0; w < nw; ++w) a = __builtin_fma(s_x[w][i], s_x[w][j], a);
                acc[m] = a;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < LG_PAIRS; ++m) {
        const int p = tid + m * LG_BLOCK;
        if (p < np) {
            const double c = acc[m] / (double)W2;
            s_a[p] = c;
            if ((pij[m] >> 16) == (pij[m] & 0xffff)) s_diag[pij[m] >> 16] = c;
        }
    }
    __syncthreads();
    // parameters without spread: row and column cleared, unit pivot, inactive
    if (tid < D) {
        const int i = tid;
        const bool act = s_diag[i] > 0.0;
        pd.lg_act[t * D + i] = act ? 1 : 0;
        const int ri = i * (i + 1) / 2;
        if (act) {
            s_a[ri + i] = s_diag[i] * (1.0 + LG_RIDGE);
        } else {
            for (int k = 0; k < i; ++k) s_a[ri + k] = 0.0;
            for (int r = i + 1; r < D; ++r) s_a[r * (r + 1) / 2 + i] = 0.0;
            s_a[ri + i] = 1.0;
        }
    }
    bool fail = false;
    for (int j = 0; j < D; ++j) {
        __syncthreads();
        const double piv = s_a[j * (j + 1) / 2 + j];    // final since step j - 1; same value in every thread
        if (!(piv > 0.0) || !(piv < __builtin_inf())) {
            fail = true;
            break;
        }
        const double inv = 1.0 / piv;
#pragma unroll
        for (int m = 0; m < LG_PAIRS; ++m) {
            const int p = tid + m * LG_BLOCK;
            const int i = pij[m] >> 16, k = pij[m] & 0xffff;
            if (p < np && k > j) {
                const double lij = s_a[i * (i + 1) / 2 + j] * inv;
                s_a[p] = __builtin_fma(-lij, s_a[k * (k + 1) / 2 + j], s_a[p]);
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < D * D; e += LG_BLOCK) {
        const int i = e / D, k = e - i * D;
        double v = 0.0;
        if (fail) {
            if (i == k) v = s_diag[i] > 0.0 ? sqrt(s_diag[i]) : 1.0;
        } else if (k <= i) {
            const double dk = sqrt(s_a[k * (k + 1) / 2 + k]);
            v = i == k ? dk : s_a[i * (i + 1) / 2 + k] / dk;
        }
        A[e] = v;
        L1[(size_t)k * D + i] = v;
    }
}

// Proposal of every walker of the moving half (one wave per walker), its
// prior, and the row handed to the adjoint kernel: the proposal when its
// prior is finite, else the current position (a valid point whose result is
// ignored).
__global__ __launch_bounds__(PRIOR_BLOCK) void k_lg_propose(ModelDev md, PTDev pd, int half, int step) {
    __shared__ double s_row[PRIOR_WAVES][MAXD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * PRIOR_WAVES + wave;
    if (b >= pd.T * pd.W2) return;            // whole waves; no workgroup barrier follows
    const int D = pd.D;
    const long long sweep = *pd.sweep;
    const int t = b / pd.W2, k = half * pd.W2 + b % pd.W2;
    const int sk = pd.map[t * pd.W + k];
    const double beta = pd.betas[t];
    const double eps = pd.lg_eps[t], e2h = 0.5 * eps * eps;
    const double* x = pd.X + (size_t)sk * D;
    const double* g = pd.lg_g + (size_t)sk * D;
    const double* L1 = pd.lg_L1 + (size_t)t * D * D;
    const double* L2 = pd.lg_L2 + (size_t)t * D * D;
    const int* act = pd.lg_act + t * D;
    double xv[2], gv[2];
    bool fin = true;
    for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        xv[h] = j < D ? x[j] : 0.0;
        gv[h] = j < D ? g[j] : 0.0;
        fin = fin && __builtin_isfinite(gv[h]);
    }
    const bool drift = __ballot(!fin) == 0ull;
    double v[2] = {0.0, 0.0};
    if (drift) lg_lt_times(L2, gv, D, lane, v);
    double u[2] = {0.0, 0.0}, lq = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        if (j < D && act[j]) {
            const double xi = lg_normal(pd, sweep, half, step, t, k, j);
            u[h] = __builtin_fma(e2h * beta, v[h], eps * xi);
            lq -= 0.5 * xi * xi;
        }
    }
    lq = wave_sum(lq);
    double dy[2] = {0.0, 0.0};
    for (int kk = 0; kk < D; ++kk) {
        const double uk = lg_bcast(u, kk);
        for (int h = 0; h < 2; ++h) {
            const int j = lane + 64 * h;
            if (j < D) dy[h] = __builtin_fma(L1[(size_t)kk * D + j], uk, dy[h]);
        }
    }
    double y[2];
    for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        y[h] = xv[h] + dy[h];
        if (j < D) s_row[wave][j] = y[h];
    }
    __threadfence_block();     // the row is read back across lanes
    const double lp = prior_proposal_wave(md, s_row[wave], s_row[wave], 0.0, lane);
    const bool live = lp > -__builtin_inf();
    double* yp = pd.lg_Xp + (size_t)b * D;
    for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        if (j < D) yp[j] = live ? y[h] : xv[h];
    }
    if (lane == 0) {
        pd.prop_lp[b] = lp;
        pd.lg_lq[b] = lq;
    }
}

// Metropolis-Hastings acceptance of the gradient move (one wave per walker).
__global__ __launch_bounds__(PRIOR_BLOCK) void k_lg_accept(PTDev pd, int half, int step) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * PRIOR_WAVES + wave;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd((unsigned long long*)(pd.sweep + 1), (unsigned long long)pd.T * pd.W2);
    if (b >= pd.T * pd.W2) return;
    const double lp = pd.prop_lp[b];
    if (!(lp > -__builtin_inf())) return;
    const int D = pd.D;
    const long long sweep = *pd.sweep;
    const int t = b / pd.W2, k = half * pd.W2 + b % pd.W2;
    const int sk = pd.map[t * pd.W + k];
    const double beta = pd.betas[t];
    const double* x = pd.X + (size_t)sk * D;
    const double* y = pd.lg_Xp + (size_t)b * D;
    const double* gy = pd.lg_gp + (size_t)b * D;
    const double* L1 = pd.lg_L1 + (size_t)t * D * D;
    const double* L2 = pd.lg_L2 + (size_t)t * D * D;
    const int* act = pd.lg_act + t * D;
    double yv[2], gv[2], dv[2];
    bool fin = true;
    for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        yv[h] = j < D ? y[j] : 0.0;
        gv[h] = j < D ? gy[j] : 0.0;
        dv[h] = j < D ? x[j] - yv[h] : 0.0;
        fin = fin && __builtin_isfinite(gv[h]);
    }
    const bool drift = __ballot(!fin) == 0ull;
    double v[2] = {0.0, 0.0};
    if (drift) lg_lt_times(L2, gv, D, lane, v);
    // r = L^-1 (x - y) by forward substitution, one column at a time
    for (int kk = 0; kk < D; ++kk) {
        const double rk = lg_bcast(dv, kk) / L1[(size_t)kk * D + kk];
        for (int h = 0; h < 2; ++h) {
            const int j = lane + 64 * h;
            if (j == kk) dv[h] = rk;
            else if (j > kk && j < D) dv[h] = __builtin_fma(-L1[(size_t)kk * D + j], rk, dv[h]);
        }
    }
    // reverse proposal: xi' = r / eps - (eps / 2) beta L^T g(y)
    const double eps = pd.lg_eps[t];
    double lq = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int j = lane + 64 * h;
        if (j < D && act[j]) {
            const double rho = dv[h] / eps - 0.5 * eps * beta * v[h];
            lq -= 0.5 * rho * rho;
        }
    }
    lq = wave_sum(lq);
    double ll = pd.lg_lp[b];
    if (!(ll == ll) || ll == __builtin_inf()) ll = -__builtin_inf();
    double lo = pd.lg_l[sk];
    if (!(lo == lo)) lo = -__builtin_inf();
    const double tn = beta == 0.0 ? 0.0 : beta * ll;
    const double to = beta == 0.0 ? 0.0 : beta * lo;
    const double lnalpha = (tn - to) + (lp - pd.lnp[sk]) + (lq - pd.lg_lq[b]);
    const U4 r = lg_draw(pd, sweep, half, step, t, k, 255u);
    if (log(u01(r.x, r.y)) < lnalpha) {
        double* xs = pd.X + (size_t)sk * D;
        double* gs = pd.lg_g + (size_t)sk * D;
        for (int h = 0; h < 2; ++h) {
            const int j = lane + 64 * h;
            if (j < D) { xs[j] = yv[h]; gs[j] = gv[h]; }
        }
        if (lane == 0) {
            pd.lnl[sk] = ll;
            pd.lnp[sk] = lp;
            pd.lg_l[sk] = ll;
            pd.accepted[t * pd.W + k] += 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
// With events, the launch itself records them (start and end of the
// kernel's execution, as a kernel trace measures it).  Models without the
// extended features (activity or cosine terms, MA orders above 1) run the
// variant that compiles them out (it keeps the one-slot variants within the
// scalar-register budget of 8 waves per SIMD).
template <int PREC, int WPT, bool POST>
static void launch_half_c(const ModelDev& md, const PTDev& pd, int half, int step, hipStream_t s, hipEvent_t e0,
                          hipEvent_t e1) {
    constexpr int SLOTS = NWAVES / WPT;
    const int64_t tasks = (int64_t)pd.T * pd.W2;
    const int64_t grid = (tasks + SLOTS - 1) / SLOTS;
    if (e0)
        hipExtLaunchKernelGGL((k_pt_half<PREC, WPT, POST>), dim3((unsigned)grid), dim3(BLOCK), 0, s, e0, e1, 0, md,
                              pd, half, step);
    else
        hipLaunchKernelGGL((k_pt_half<PREC, WPT, POST>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, pd, half,
                           step);
}

template <int PREC, int WPT>
static void launch_half_t(const ModelDev& md, const PTDev& pd, int half, int step, hipStream_t s,
                          hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (needs_ext(md)) launch_half_c<PREC, WPT, true>(md, pd, half, step, s, e0, e1);
    else launch_half_c<PREC, WPT, false>(md, pd, half, step, s, e0, e1);
}

// draw_swaps: this prior pass also draws the sweep's swap randoms (the last
// half-step of a sweep).
static int launch_prior(EmpPT* p, int half, int step, int draw_swaps, int first) {
    const int64_t tasks = (int64_t)p->dev.T * p->dev.W2;
    hipLaunchKernelGGL(k_pt_prior, dim3((unsigned)((tasks + PRIOR_WAVES - 1) / PRIOR_WAVES)), dim3(PRIOR_BLOCK), 0,
                       p->stream, p->model->dev, p->dev, half, step, draw_swaps, first);
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

static int launch_lik(EmpPT* p, int half, int step, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const ModelDev& md = p->model->dev;
    const int prec = p->desc.precision, w = p->wpt;
#define EMP_CASE(P, W) if (prec == P && w == W) { launch_half_t<P, W>(md, p->dev, half, step, p->stream, e0, e1); EMP_HIP(hipGetLastError()); return EMP_OK; }
    EMP_CASE(0, 1) EMP_CASE(0, 2) EMP_CASE(0, 4) EMP_CASE(0, 8)
    EMP_CASE(1, 1) EMP_CASE(1, 2) EMP_CASE(1, 4) EMP_CASE(1, 8)
#undef EMP_CASE
    emp_set_error("bad precision/wpt");
    return EMP_EINVAL;
}

static int launch_half(EmpPT* p, int half, int step, int draw_swaps, int first = 0) {
    const int rc = launch_prior(p, half, step, draw_swaps, first);
    return rc ? rc : launch_lik(p, half, step);
}

// Swap round (with ladder and sweep counter): the end of every sweep,
// whichever move it made.  The sweep's state is recorded by the next sweep's
// first pass, or by launch_record at the end of the run.
static int launch_tail(EmpPT* p) {
    if (p->swap_lds)
        hipLaunchKernelGGL(k_pt_swap<true>, dim3(1), dim3(SWAP_BLOCK), p->swap_lds, p->stream, p->dev);
    else
        hipLaunchKernelGGL(k_pt_swap<false>, dim3(1), dim3(SWAP_BLOCK), 0, p->stream, p->dev);
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

// A history record (k_pt_record's modes): the previous sweep's state only if
// still owed (1, the start of a gradient sweep), before an inner step of a
// gradient sweep (2), or unconditionally, clearing the debt (0, the end of a
// run).
static int launch_record(EmpPT* p, int mode, int step = 0) {
    if (!p->dev.record) return EMP_OK;
    const size_t nx = (size_t)p->dev.T * p->dev.W * p->dev.D;
    const unsigned grid = (unsigned)std::min<size_t>((nx + 255) / 256, 4096);
    hipLaunchKernelGGL(k_pt_record, dim3(grid), dim3(256), 0, p->stream, p->dev, mode, step);
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

static int launch_sweep(EmpPT* p) {
    for (int st = 0; st < p->desc.nsteps; ++st)
        for (int h = 0; h < 2; ++h) {
            int rc = launch_half(p, h, st, h == 1 && st == p->desc.nsteps - 1, h == 0);
            if (rc) return rc;
        }
    return launch_tail(p);
}

// A sweep of gradient moves: gradients at the current positions that a
// stretch move has moved since their last gradient (the other slots keep
// theirs: swaps exchange map entries, not states), then per half-step the
// preconditioner, the proposals, their gradients and the acceptance.
// Allocates nothing (the scratch is the engine's own), so it can be captured.
// ev: optional [2 nsteps][2] events recorded around each proposal adjoint.
static int launch_grad_sweep(EmpPT* p, hipEvent_t* ev = nullptr) {
    const PTDev& d = p->dev;
    const ModelDev& md = p->model->dev;
    const int64_t TW = (int64_t)d.T * d.W, tasks = (int64_t)d.T * d.W2;
    const unsigned tgrid = (unsigned)((tasks + PRIOR_WAVES - 1) / PRIOR_WAVES);
    int rc = launch_record(p, 1);
    if (rc) return rc;
    rc = emp_launch_grad_ws(p->model, d.X, TW, d.D, d.lg_l, d.lg_g, d.lg_scr, p->stream, p->desc.precision,
                            d.lg_stale, p->grad_wpt);
    if (rc) return rc;
    for (int st = 0; st < p->desc.nsteps; ++st)
        for (int h = 0; h < 2; ++h) {
            if (h == 0 && st > 0 && (rc = launch_record(p, 2, st))) return rc;
            hipLaunchKernelGGL(k_lg_precond, dim3((unsigned)d.T), dim3(LG_BLOCK), 0, p->stream, d, h, st);
            EMP_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_lg_propose, dim3(tgrid), dim3(PRIOR_BLOCK), 0, p->stream, md, d, h, st);
            EMP_HIP(hipGetLastError());
            hipEvent_t* e = ev ? ev + 2 * (2 * st + h) : nullptr;
            if (e) EMP_HIP(hipEventRecord(e[0], p->stream));
            rc = emp_launch_grad_ws(p->model, d.lg_Xp, tasks, d.D, d.lg_lp, d.lg_gp, d.lg_scr, p->stream,
                                    p->desc.precision, nullptr, p->grad_wpt);
            if (rc) return rc;
            if (e) EMP_HIP(hipEventRecord(e[1], p->stream));
            hipLaunchKernelGGL(k_lg_accept, dim3(tgrid), dim3(PRIOR_BLOCK), 0, p->stream, d, h, st);
            EMP_HIP(hipGetLastError());
        }
    if (d.T > 1) {
        const int np = (d.T - 1) * d.W;
        hipLaunchKernelGGL(k_pt_swap_draw, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, p->stream, d);
        EMP_HIP(hipGetLastError());
    }
    return launch_tail(p);
}

// The move of sweep `sweep`: 1 (gradient) with probability w_grad / (w_stretch
// + w_grad), from the engine's Philox stream; a run stays a function of the
// seed and resumes bitwise.
static int pick_move(const EmpPT* p, long long sweep) {
    if (!(p->w_grad > 0.0)) return 0;
    if (!(p->w_stretch > 0.0)) return 1;
    const U4 r = philox4x32_10(U4{0u, (uint32_t)sweep, (uint32_t)((unsigned long long)sweep >> 32), MOVE_TAG},
                               p->dev.key0, p->dev.key1);
    return u01(r.x, r.y) * (p->w_stretch + p->w_grad) < p->w_grad ? 1 : 0;
}

static void drop_graphs(EmpPT* p) {
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph_grad) (void)hipGraphExecDestroy(p->graph_grad);
    for (auto& g : p->graph_run) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    p->graph_exec = p->graph_grad = nullptr;
}

// Captures `n` consecutive sweeps of one kind into a graph.  The sweep
// counter lives on the device, so replaying a graph of n sweeps equals n
// replays of the one-sweep graph.
static int capture_sweeps(EmpPT* p, int kind, int n, hipGraphExec_t* out) {
    hipGraph_t g;
    EMP_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    int rc = EMP_OK;
    for (int i = 0; i < n && !rc; ++i) rc = kind ? launch_grad_sweep(p) : launch_sweep(p);
    hipError_t ce = hipStreamEndCapture(p->stream, &g);
    if (rc) return rc;
    EMP_HIP(ce);
    EMP_HIP(hipGraphInstantiate(out, g, nullptr, nullptr, 0));
    EMP_HIP(hipGraphDestroy(g));
    return EMP_OK;
}

template <typename T>
static int alloc(EmpPT* p, T** out, size_t n) {
    void* q = nullptr;
    EMP_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    p->allocations.push_back(q);
    *out = (T*)q;
    return EMP_OK;
}

// Every slot's gradient is recomputed by the next gradient sweep (after a
// state was loaded or restored; no-op while the gradient move was never on).
static int mark_all_stale(EmpPT* p) {
    PTDev& d = p->dev;
    if (!d.lg_stale) return EMP_OK;
    const int64_t n = (int64_t)d.T * d.W;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, d.lg_stale, n, 1);
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

// Evaluates lnp and lnl of every slot (slots hold the positions in logical
// order when the map is the identity, which is how states are loaded).
static int eval_state(EmpPT* p) {
    PTDev& d = p->dev;
    const int64_t n = (int64_t)d.T * d.W;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, d.map, n);
    EMP_HIP(hipGetLastError());
    if (int rc = mark_all_stale(p)) return rc;
    int rc = emp_launch_prior(p->model, d.X, n, d.D, d.lnp, p->stream);
    if (rc) return rc;
    rc = emp_launch_lnl(p->model, d.X, n, d.D, d.lnl, p->desc.precision, 0, p->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mask_lnl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, d.lnl, d.lnp, n);
    EMP_HIP(hipGetLastError());
    EMP_HIP(hipStreamSynchronize(p->stream));
    return EMP_OK;
}

extern "C" void emp_pt_destroy(EmpPT* p) {
    if (!p) return;
    drop_graphs(p);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (void* q : p->allocations) (void)hipFree(q);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" int emp_pt_create(EmpModel* m, const EmpPTDesc* desc, const double* betas, const double* p0, EmpPT** out) {
    if (!out) { emp_set_error("null output handle"); return EMP_EINVAL; }
    *out = nullptr;
    if (!m || !desc || !betas || !p0) { emp_set_error("null argument"); return EMP_EINVAL; }
    const EmpPTDesc& d = *desc;
    if (d.ntemps < 1 || d.ntemps > 64) { emp_set_error("ntemps must be in [1, 64]"); return EMP_EINVAL; }
    if (d.nwalkers < 2 || d.nwalkers % 2 || d.nwalkers > SWAP_BLOCK) {
        emp_set_error("nwalkers must be even and in [2, %d]", SWAP_BLOCK);
        return EMP_EINVAL;
    }
    if (d.nsteps < 1 || d.nsteps > (1 << 20)) { emp_set_error("nsteps must be in [1, 2^20]"); return EMP_EINVAL; }
    if (d.precision != EMP_FP64 && d.precision != EMP_MIXED) { emp_set_error("bad precision"); return EMP_EINVAL; }
    if (!(d.a > 1.0)) { emp_set_error("stretch scale a must be > 1"); return EMP_EINVAL; }
    if (d.adapt_mode < -1 || d.adapt_mode > 2) { emp_set_error("adapt_mode must be -1 (frozen), 0, 1 or 2"); return EMP_EUNSUP; }
    if (d.adapt_mode >= 0 && !(d.adapt_tau > 0 && d.adapt_nu > 0)) { emp_set_error("adapt_tau and adapt_nu must be > 0"); return EMP_EINVAL; }
    if (m->desc.ndim < 1) { emp_set_error("model has no free parameter"); return EMP_EINVAL; }
    for (int i = 0; i < d.ntemps; ++i)
        if (!(betas[i] >= 0.0 && betas[i] <= 1.0)) { emp_set_error("betas must lie in [0, 1]"); return EMP_EINVAL; }
    EmpPT* p = new EmpPT();
    p->model = m;
    p->desc = d;
    PTDev& v = p->dev;
    v.T = d.ntemps; v.W = d.nwalkers; v.W2 = d.nwalkers / 2; v.D = m->desc.ndim;
    v.a = d.a;
    v.key0 = (uint32_t)d.seed; v.key1 = (uint32_t)(d.seed >> 32);
    v.adapt_mode = d.adapt_mode; v.adapt_tau = d.adapt_tau; v.adapt_nu = d.adapt_nu;
    v.max_sweeps = std::max<int64_t>(d.max_sweeps, 0);
    v.nsteps = d.nsteps;
    v.max_rows = v.max_sweeps * d.nsteps;
    v.record = d.record ? 1 : 0;
    const size_t TW = (size_t)v.T * v.W, S = (size_t)v.max_sweeps, R = (size_t)v.max_rows;
    int rc = 0;
    if ((rc = alloc(p, &v.X, TW * v.D)) || (rc = alloc(p, &v.lnl, TW)) || (rc = alloc(p, &v.lnp, TW)) ||
        (rc = alloc(p, &v.map, TW)) || (rc = alloc(p, &v.swap_l, TW)) ||
        (rc = alloc(p, &v.swap_lu, (size_t)std::max(v.T - 1, 1) * v.W)) || (rc = alloc(p, &v.rec_owed, 1)) ||
        (rc = alloc(p, &v.swap_jp, (size_t)std::max(v.T - 1, 1) * v.W)) ||
        (rc = alloc(p, &v.betas, (size_t)v.T)) || (rc = alloc(p, &v.accepted, TW)) || (rc = alloc(p, &v.sweep, 2)) ||
        (rc = alloc(p, &v.h_betas, S * v.T)) || (rc = alloc(p, &v.h_tsw, S * std::max(v.T - 1, 1))) ||
        (rc = alloc(p, &v.h_smd, S * std::max(v.T - 1, 1))) || (rc = alloc(p, &v.prop_lp, TW / 2)) ||
        (rc = alloc(p, &v.rec, TW / 2)) || (rc = alloc(p, &v.prow, TW / 2 * v.D))) {
        emp_pt_destroy(p);
        return rc;
    }
    if (v.record) {
        if ((rc = alloc(p, &v.h_chain, R * TW * v.D)) || (rc = alloc(p, &v.h_lnl, R * TW)) || (rc = alloc(p, &v.h_lnp, R * TW))) {
            emp_pt_destroy(p);
            return rc;
        }
    }
    // Stage the swap round in LDS when the map, the likelihoods and the
    // pre-drawn pairings/uniforms fit (12 bytes per walker and per walker-pair,
    // inside the default 64 KiB workgroup allocation); otherwise the same
    // kernel works in global memory and draws the randoms inline.
    const size_t lds = (TW + (size_t)(v.T - 1) * v.W) * (sizeof(double) + sizeof(int));
    p->swap_lds = p->swap_lds_fit = lds <= 62 * 1024 ? lds : 0;
    v.swap_stream = (!p->swap_lds && v.W <= SWAP_BLOCK) ? 1 : 0;
#define EMP_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { int r_ = emp_hip_fail(e_, #x, __FILE__, __LINE__); emp_pt_destroy(p); return r_; } } while (0)
    EMP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    EMP_TRY(hipMemcpy(v.X, p0, TW * v.D * sizeof(double), hipMemcpyHostToDevice));
    EMP_TRY(hipMemcpy(v.betas, betas, (size_t)v.T * sizeof(double), hipMemcpyHostToDevice));
    EMP_TRY(hipMemset(v.accepted, 0, TW * sizeof(long long)));
    EMP_TRY(hipMemset(v.sweep, 0, 2 * sizeof(long long)));
    EMP_TRY(hipMemset(v.rec_owed, 0, sizeof(int)));
    EMP_TRY(hipDeviceSynchronize());     // the engine stream does not wait for the null stream
#undef EMP_TRY
    p->wpt = emp_pick_wpt(m, (int64_t)v.T * v.W2, d.precision, true);
    rc = eval_state(p);
    if (rc) { emp_pt_destroy(p); return rc; }
    *out = p;
    return EMP_OK;
}

extern "C" int emp_pt_run(EmpPT* p, int64_t nsweeps, int32_t use_graph) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    if (nsweeps < 0) { emp_set_error("nsweeps must be >= 0"); return EMP_EINVAL; }
    if (emp_fault_injected()) return EMP_EFAULT_INJECTED;
    if (p->host_sweeps + nsweeps > p->dev.max_sweeps) {
        emp_set_error("history capacity %lld sweeps exceeded (%lld done, %lld requested)",
                      (long long)p->dev.max_sweeps, p->host_sweeps, (long long)nsweeps);
        return EMP_EINVAL;
    }
    if (nsweeps == 0) return EMP_OK;
    hipEvent_t e0, e1;
    EMP_HIP(hipEventCreate(&e0));
    EMP_HIP(hipEventCreate(&e1));
    int rc = EMP_OK;
    EMP_HIP(hipEventRecord(e0, p->stream));
    for (int64_t s = 0; s < nsweeps;) {
        const int mv = pick_move(p, p->host_sweeps + s);
        if (!use_graph) {
            rc = mv ? launch_grad_sweep(p) : launch_sweep(p);
            if (rc) return rc;
            ++s;
            continue;
        }
        // A run of stretch sweeps goes out in graphs of 32, 16, ..., 2, 1
        // sweeps: one graph launch per graph instead of per sweep.
        int64_t run = 1;
        if (!mv)
            while (run < 32 && s + run < nsweeps && !pick_move(p, p->host_sweeps + s + run)) ++run;
        int lvl = -1;
        for (int l = 4; l >= 0; --l)
            if (run >= (2 << l)) { lvl = l; break; }
        hipGraphExec_t& ge = lvl >= 0 ? p->graph_run[lvl] : (mv ? p->graph_grad : p->graph_exec);
        const int n = lvl >= 0 ? (2 << lvl) : 1;
        if (!ge && (rc = capture_sweeps(p, mv, n, &ge))) return rc;
        EMP_HIP(hipGraphLaunch(ge, p->stream));
        s += n;
    }
    if ((rc = launch_record(p, 0))) return rc;    // the run's last sweep
    EMP_HIP(hipEventRecord(e1, p->stream));
    EMP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    EMP_HIP(hipEventElapsedTime(&ms, e0, e1));
    p->last_ms_per_sweep = ms / (double)nsweeps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->host_sweeps += nsweeps;
    return EMP_OK;
}

// Runs ceil(n / 2) stretch sweeps one launch at a time (no graph) and times
// the first n likelihood kernels with events recorded by their own launches:
// the measured launches are those of ordinary sweeps, which advance the chain
// and its history.
extern "C" int emp_pt_time_halfstep(EmpPT* p, int32_t n, double* ms_avg) {
    if (!p || n < 1 || !ms_avg) { emp_set_error("bad argument"); return EMP_EINVAL; }
    const int64_t ns = ((int64_t)n + 2 * p->desc.nsteps - 1) / (2 * p->desc.nsteps);
    if (p->host_sweeps + ns > p->dev.max_sweeps) {
        emp_set_error("history capacity %lld sweeps exceeded (%lld done, %lld needed to time %d half-steps)",
                      (long long)p->dev.max_sweeps, p->host_sweeps, (long long)ns, n);
        return EMP_EINVAL;
    }
    std::vector<hipEvent_t> ev(2 * (size_t)n);
    for (auto& e : ev) EMP_HIP(hipEventCreate(&e));
    int k = 0, rc = EMP_OK;
    for (int64_t sw = 0; sw < ns && !rc; ++sw) {
        for (int st = 0; st < p->desc.nsteps && !rc; ++st)
            for (int h = 0; h < 2 && !rc; ++h) {
                rc = launch_prior(p, h, st, h == 1 && st == p->desc.nsteps - 1, h == 0);
                if (rc) break;
                rc = k < n ? launch_lik(p, h, st, ev[2 * k], ev[2 * k + 1]) : launch_lik(p, h, st);
                ++k;
            }
        if (!rc) rc = launch_tail(p);
    }
    if (!rc) rc = launch_record(p, 0);
    if (rc) return rc;
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->host_sweeps += ns;
    double tot = 0.0;
    for (int i = 0; i < n; ++i) {
        float ms = 0.f;
        EMP_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        tot += ms;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_avg = tot / n;
    p->last_ms_half = *ms_avg;
    return EMP_OK;
}

// Runs ceil(n / (2 nsteps)) gradient sweeps launch by launch (no graph) and
// times the first n adjoint launches over the proposals (T W / 2 vectors
// each) with events recorded on the engine stream around each launch.  The
// sweeps are ordinary ones: they advance the chain and use history capacity.
extern "C" int emp_pt_time_gradstep(EmpPT* p, int32_t n, double* ms_avg) {
    if (!p || n < 1 || !ms_avg) { emp_set_error("bad argument"); return EMP_EINVAL; }
    if (!(p->w_grad > 0.0) || !p->dev.lg_g) { emp_set_error("the gradient move is not enabled"); return EMP_EINVAL; }
    const int per = 2 * p->desc.nsteps;
    const int64_t ns = ((int64_t)n + per - 1) / per;
    if (p->host_sweeps + ns > p->dev.max_sweeps) {
        emp_set_error("history capacity %lld sweeps exceeded (%lld done, %lld needed to time %d launches)",
                      (long long)p->dev.max_sweeps, p->host_sweeps, (long long)ns, n);
        return EMP_EINVAL;
    }
    std::vector<hipEvent_t> ev(2 * (size_t)per * ns, nullptr);
    int rc = EMP_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { rc = emp_hip_fail(hipGetLastError(), "event", __FILE__, __LINE__); break; }
    for (int64_t sw = 0; sw < ns && !rc; ++sw) rc = launch_grad_sweep(p, ev.data() + 2 * (size_t)per * sw);
    if (!rc) rc = launch_record(p, 0);
    if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "sync", __FILE__, __LINE__);
    double tot = 0.0;
    for (int i = 0; i < n && !rc; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "timing", __FILE__, __LINE__);
        tot += ms;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    p->host_sweeps += ns;
    *ms_avg = tot / n;
    return EMP_OK;
}

// Times graph-replayed stretch sweeps at 1, 2, 4 and 8 waves per walker on a
// copy of the state (positions, likelihoods, map, ladder, counters),
// restores the state after each shape and keeps the fastest.  Each shape: an
// 8-sweep graph, replayed once untimed and `trials` times timed.  History
// rows the trials write lie at or after the current sweep and are written
// again by the run that follows.
extern "C" int emp_pt_autotune(EmpPT* p, int32_t trials, int32_t* best_wpt, double* best_ms) {
    if (!p || trials < 1) { emp_set_error("bad argument"); return EMP_EINVAL; }
    PTDev& v = p->dev;
    const size_t TW = (size_t)v.T * v.W;
    const std::pair<void*, size_t> items[] = {{v.X, TW * v.D * 8}, {v.lnl, TW * 8},      {v.lnp, TW * 8},
                                              {v.accepted, TW * 8}, {v.sweep, 16},         {v.map, TW * 4},
                                              {v.betas, (size_t)v.T * 8}, {v.rec_owed, sizeof(int)}};
    std::vector<void*> copies;
    auto release = [&]() { for (void* q : copies) (void)hipFree(q); };
    EMP_HIP(hipStreamSynchronize(p->stream));
    for (const auto& it : items) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, it.second);
        if (e != hipSuccess) { release(); return emp_hip_fail(e, "hipMalloc (autotune snapshot)", __FILE__, __LINE__); }
        copies.push_back(q);
        if ((e = hipMemcpyAsync(q, it.first, it.second, hipMemcpyDeviceToDevice, p->stream)) != hipSuccess) {
            release();
            return emp_hip_fail(e, "snapshot", __FILE__, __LINE__);
        }
    }
    constexpr int NS = 8;
    const int keep = p->wpt;
    int best = keep;
    double best_t = -1.0;
    int rc = EMP_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "event", __FILE__, __LINE__);
    for (int w = 1; w <= 8 && !rc; w *= 2) {
        p->wpt = w;
        drop_graphs(p);
        hipGraphExec_t g = nullptr;
        rc = capture_sweeps(p, 0, NS, &g);
        float ms = 0.f;
        if (!rc && hipGraphLaunch(g, p->stream) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "graph", __FILE__, __LINE__);
        if (!rc && hipEventRecord(e0, p->stream) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "event", __FILE__, __LINE__);
        for (int i = 0; i < trials && !rc; ++i)
            if (hipGraphLaunch(g, p->stream) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "graph", __FILE__, __LINE__);
        if (!rc && (hipEventRecord(e1, p->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = emp_hip_fail(hipGetLastError(), "timing", __FILE__, __LINE__);
        if (g) (void)hipGraphExecDestroy(g);
        for (size_t i = 0; i < copies.size() && !rc; ++i)
            if (hipMemcpyAsync(items[i].first, copies[i], items[i].second, hipMemcpyDeviceToDevice, p->stream) != hipSuccess)
                rc = emp_hip_fail(hipGetLastError(), "restore", __FILE__, __LINE__);
        if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = emp_hip_fail(hipGetLastError(), "restore", __FILE__, __LINE__);
        const double per = (double)ms / (NS * trials);
        if (!rc && (best_t < 0 || per < best_t)) { best_t = per; best = w; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    release();
    p->wpt = rc ? keep : best;
    drop_graphs(p);
    if (rc) return rc;
    if (best_wpt) *best_wpt = best;
    if (best_ms) *best_ms = best_t;
    return EMP_OK;
}

extern "C" int emp_pt_last_timing(EmpPT* p, double* ms_per_sweep, double* ms_half) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    if (ms_per_sweep) *ms_per_sweep = p->last_ms_per_sweep;
    if (ms_half) *ms_half = p->last_ms_half;
    return EMP_OK;
}

// Current state in logical order (gathered through the map) on the host.
static int gather_state(EmpPT* p, std::vector<double>& X, std::vector<double>& l, std::vector<double>& lp) {
    const PTDev& v = p->dev;
    const size_t TW = (size_t)v.T * v.W;
    std::vector<double> Xs(TW * v.D), ls(TW), ps(TW);
    std::vector<int> mp(TW);
    EMP_HIP(hipMemcpy(Xs.data(), v.X, Xs.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(ls.data(), v.lnl, ls.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(ps.data(), v.lnp, ps.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(mp.data(), v.map, mp.size() * 4, hipMemcpyDeviceToHost));
    X.resize(TW * v.D); l.resize(TW); lp.resize(TW);
    for (size_t w = 0; w < TW; ++w) {
        const size_t s = (size_t)mp[w];
        memcpy(&X[w * v.D], &Xs[s * v.D], (size_t)v.D * 8);
        l[w] = ls[s];
        lp[w] = ps[s];
    }
    return EMP_OK;
}

extern "C" int emp_pt_get(EmpPT* p, int32_t what, void* host_out, int64_t cap, int64_t* shape) {
    if (!p || !host_out) { emp_set_error("null argument"); return EMP_EINVAL; }
    const PTDev& v = p->dev;
    EMP_HIP(hipStreamSynchronize(p->stream));
    long long done = 0;
    EMP_HIP(hipMemcpy(&done, v.sweep, sizeof(done), hipMemcpyDeviceToHost));
    const long long S = std::min<long long>(done, v.max_sweeps);
    const long long R = std::min<long long>(done * v.nsteps, v.max_rows);     // history rows (one per step)
    // history held from sweep S0 (row R0) on
    const long long S0 = std::min<long long>(p->hist_base, S), R0 = S0 * v.nsteps;
    const size_t TW = (size_t)v.T * v.W, T1 = (size_t)std::max(v.T - 1, 1);
    int64_t sh[4] = {1, 1, 1, 1};
    const void* src = nullptr;
    size_t elem = sizeof(double);
    std::vector<double> Xg, lg, pg;
    if (what == EMP_PT_STATE_X || what == EMP_PT_STATE_LNL || what == EMP_PT_STATE_LNP) {
        int rc = gather_state(p, Xg, lg, pg);
        if (rc) return rc;
    }
    bool host_src = false;
    switch (what) {
        case EMP_PT_CHAIN:
            if (!v.record) goto norec;
            sh[0] = R - R0; sh[1] = v.T; sh[2] = v.W; sh[3] = v.D; src = v.h_chain + (size_t)R0 * TW * v.D;
            break;
        case EMP_PT_LNL: if (!v.record) goto norec; sh[0] = R - R0; sh[1] = v.T; sh[2] = v.W; src = v.h_lnl + (size_t)R0 * TW; break;
        case EMP_PT_LNP: if (!v.record) goto norec; sh[0] = R - R0; sh[1] = v.T; sh[2] = v.W; src = v.h_lnp + (size_t)R0 * TW; break;
        case EMP_PT_BETAS: sh[0] = S - S0; sh[1] = v.T; src = v.h_betas + (size_t)S0 * v.T; break;
        case EMP_PT_TSW: sh[0] = S - S0; sh[1] = std::max(v.T - 1, 0); src = v.h_tsw + (size_t)S0 * T1; break;
        case EMP_PT_SMD: sh[0] = S - S0; sh[1] = std::max(v.T - 1, 0); src = v.h_smd + (size_t)S0 * T1; break;
        case EMP_PT_ACCEPTED: sh[0] = v.T; sh[1] = v.W; src = v.accepted; elem = sizeof(long long); break;
        case EMP_PT_STATE_X: sh[0] = v.T; sh[1] = v.W; sh[2] = v.D; src = Xg.data(); host_src = true; break;
        case EMP_PT_STATE_LNL: sh[0] = v.T; sh[1] = v.W; src = lg.data(); host_src = true; break;
        case EMP_PT_STATE_LNP: sh[0] = v.T; sh[1] = v.W; src = pg.data(); host_src = true; break;
        case EMP_PT_STATE_BETAS: sh[0] = v.T; src = v.betas; break;
        case EMP_PT_NSWEEPS: sh[0] = 2; src = v.sweep; elem = sizeof(long long); break;
        case EMP_PT_GRAD_STEPS:
            if (!v.lg_eps) { emp_set_error("the gradient move was never enabled"); return EMP_EINVAL; }
            sh[0] = v.T; src = v.lg_eps;
            break;
        default: emp_set_error("unknown quantity %d", what); return EMP_EINVAL;
    }
    {
        const size_t n = (size_t)sh[0] * sh[1] * sh[2] * sh[3];
        if ((int64_t)(n * elem) > cap) { emp_set_error("output buffer too small (%zu bytes needed)", n * elem); return EMP_EINVAL; }
        if (n) {
            if (host_src) memcpy(host_out, src, n * elem);
            else EMP_HIP(hipMemcpy(host_out, src, n * elem, hipMemcpyDeviceToHost));
        }
        if (shape) for (int i = 0; i < 4; ++i) shape[i] = sh[i];
        return EMP_OK;
    }
norec:
    emp_set_error("history was not recorded (record == 0)");
    return EMP_EINVAL;
}

extern "C" int emp_pt_set_swap_path(EmpPT* p, int32_t path) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    PTDev& v = p->dev;
    size_t lds = 0;
    int stream = 0;
    switch (path) {
        case EMP_SWAP_AUTO:
            lds = p->swap_lds_fit;
            stream = (!lds && v.W <= SWAP_BLOCK) ? 1 : 0;
            break;
        case EMP_SWAP_LDS:
            if (!p->swap_lds_fit) { emp_set_error("the ensemble does not fit the LDS swap round"); return EMP_EINVAL; }
            lds = p->swap_lds_fit;
            break;
        case EMP_SWAP_STREAM:
            if (v.W > SWAP_BLOCK) { emp_set_error("the streamed swap round takes at most %d walkers", SWAP_BLOCK); return EMP_EINVAL; }
            stream = 1;
            break;
        case EMP_SWAP_GLOBAL:
            break;
        default:
            emp_set_error("swap path must be EMP_SWAP_AUTO, _LDS, _STREAM or _GLOBAL");
            return EMP_EINVAL;
    }
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->swap_lds = lds;
    v.swap_stream = stream;
    drop_graphs(p);        // kernel and arguments changed: capture again on the next run
    return EMP_OK;
}

extern "C" int emp_pt_set_adaptation(EmpPT* p, int32_t mode, double tau, double nu) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    if (mode < -1 || mode > 2) { emp_set_error("adapt_mode must be -1 (frozen), 0, 1 or 2"); return EMP_EUNSUP; }
    if (mode >= 0 && !(tau > 0 && nu > 0)) { emp_set_error("adapt_tau and adapt_nu must be > 0"); return EMP_EINVAL; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->dev.adapt_mode = mode;
    p->dev.adapt_tau = tau;
    p->dev.adapt_nu = nu;
    drop_graphs(p);        // kernel arguments changed: capture again on the next run
    return EMP_OK;
}

extern "C" int emp_pt_lnl_moments(EmpPT* p, int64_t discard, double* mean, double* lse, double* var) {
    if (!p || !mean || !lse || !var) { emp_set_error("null argument"); return EMP_EINVAL; }
    const PTDev& v = p->dev;
    if (!v.record) { emp_set_error("history was not recorded (record == 0)"); return EMP_EINVAL; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    long long done = 0;
    EMP_HIP(hipMemcpy(&done, v.sweep, sizeof(done), hipMemcpyDeviceToHost));
    const long long S = std::min<long long>(done * v.nsteps, v.max_rows);      // history rows
    // rows held start at the history base; discard counts from there
    const long long R0 = std::min<long long>(p->hist_base * v.nsteps, S);
    if (discard < 0 || R0 + discard >= S) { emp_set_error("discard must lie in [0, %lld) rows", S - R0); return EMP_EINVAL; }
    discard += R0;
    double* d = nullptr;
    EMP_HIP(hipMalloc(&d, 3 * (size_t)v.T * sizeof(double)));
    hipLaunchKernelGGL(k_pt_lnl_moments, dim3((unsigned)v.T), dim3(MOM_BLOCK), 0, p->stream, v, (long long)discard, S,
                       d, d + v.T, d + 2 * v.T);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(mean, d, (size_t)v.T * sizeof(double), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(lse, d + v.T, (size_t)v.T * sizeof(double), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(var, d + 2 * v.T, (size_t)v.T * sizeof(double), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return emp_hip_fail(e, "lnl moments", __FILE__, __LINE__);
    return EMP_OK;
}

extern "C" int emp_pt_set_moves(EmpPT* p, const EmpMoveDesc* moves, int32_t n) {
    if (!p || (n > 0 && !moves)) { emp_set_error("null argument"); return EMP_EINVAL; }
    if (n < 1 || n > 2) { emp_set_error("between one and two moves"); return EMP_EINVAL; }
    double ws = 0.0, wg = 0.0, a = p->dev.a, eps = 0.0;
    bool seen[2] = {false, false};
    for (int i = 0; i < n; ++i) {
        const EmpMoveDesc& m = moves[i];
        if (m.kind != EMP_MOVE_STRETCH && m.kind != EMP_MOVE_GRADIENT) { emp_set_error("unknown move kind %d", m.kind); return EMP_EINVAL; }
        if (seen[m.kind]) { emp_set_error("move kind %d given twice", m.kind); return EMP_EINVAL; }
        seen[m.kind] = true;
        if (!(m.weight >= 0.0 && m.weight < __builtin_inf())) { emp_set_error("move weights must be finite and >= 0"); return EMP_EINVAL; }
        if (m.kind == EMP_MOVE_STRETCH) {
            if (!(m.scale > 1.0 && m.scale < __builtin_inf())) { emp_set_error("stretch scale a must be > 1"); return EMP_EINVAL; }
            ws = m.weight;
            a = m.scale;
        } else {
            if (!(m.scale > 0.0 && m.scale < __builtin_inf())) { emp_set_error("gradient step must be > 0"); return EMP_EINVAL; }
            wg = m.weight;
            eps = m.scale;
        }
    }
    if (!(ws + wg > 0.0)) { emp_set_error("move weights sum to zero"); return EMP_EINVAL; }
    PTDev& v = p->dev;
    EMP_HIP(hipStreamSynchronize(p->stream));
    if (wg > 0.0 && !v.lg_g) {
        const size_t TW = (size_t)v.T * v.W, T2 = TW / 2, D = (size_t)v.D;
        const EmpModel* m = p->model;
        int rc = 0;
        if ((rc = alloc(p, &v.lg_g, TW * D)) || (rc = alloc(p, &v.lg_l, TW)) ||
            (rc = alloc(p, &v.lg_L1, (size_t)v.T * D * D)) || (rc = alloc(p, &v.lg_L2, (size_t)v.T * D * D)) ||
            (rc = alloc(p, &v.lg_act, (size_t)v.T * D)) || (rc = alloc(p, &v.lg_Xp, T2 * D)) ||
            (rc = alloc(p, &v.lg_gp, T2 * D)) || (rc = alloc(p, &v.lg_lp, T2)) || (rc = alloc(p, &v.lg_lq, T2)) ||
            (rc = alloc(p, &v.lg_eps, (size_t)v.T)) || (rc = alloc(p, &v.lg_stale, TW)) ||
            (emp_grad_scratch_doubles(m, (int64_t)TW) && (rc = alloc(p, &v.lg_scr, emp_grad_scratch_doubles(m, (int64_t)TW))))) {
            v.lg_g = nullptr;     // the allocations stay with the engine; a later call retries
            v.lg_stale = nullptr;
            return rc;
        }
        p->grad_wpt = emp_pick_grad_wpt(m, (int64_t)T2);
        if ((rc = mark_all_stale(p))) return rc;
        EMP_HIP(hipStreamSynchronize(p->stream));
    }
    v.a = a;
    if (wg > 0.0) {
        const std::vector<double> steps((size_t)v.T, eps);
        EMP_HIP(hipMemcpy(v.lg_eps, steps.data(), steps.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    p->w_stretch = ws;
    p->w_grad = wg;
    drop_graphs(p);
    return EMP_OK;
}

extern "C" int emp_pt_set_gradient_steps(EmpPT* p, const double* eps) {
    if (!p || !eps) { emp_set_error("null argument"); return EMP_EINVAL; }
    PTDev& v = p->dev;
    if (!(p->w_grad > 0.0) || !v.lg_eps) { emp_set_error("the gradient move is not enabled (emp_pt_set_moves)"); return EMP_EINVAL; }
    for (int t = 0; t < v.T; ++t)
        if (!(eps[t] > 0.0 && eps[t] < __builtin_inf())) { emp_set_error("gradient step %d must be finite and > 0", t); return EMP_EINVAL; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    EMP_HIP(hipMemcpy(v.lg_eps, eps, (size_t)v.T * sizeof(double), hipMemcpyHostToDevice));
    return EMP_OK;
}

extern "C" int emp_pt_set_state(EmpPT* p, const double* x, const double* betas) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    PTDev& v = p->dev;
    EMP_HIP(hipStreamSynchronize(p->stream));
    if (betas) {
        for (int i = 0; i < v.T; ++i)
            if (!(betas[i] >= 0.0 && betas[i] <= 1.0)) { emp_set_error("betas must lie in [0, 1]"); return EMP_EINVAL; }
        EMP_HIP(hipMemcpy(v.betas, betas, (size_t)v.T * sizeof(double), hipMemcpyHostToDevice));
    }
    if (x) {
        EMP_HIP(hipMemcpy(v.X, x, (size_t)v.T * v.W * v.D * sizeof(double), hipMemcpyHostToDevice));
        return eval_state(p);
    }
    return EMP_OK;
}

// Checkpoint layout (version 2, little-endian native binary): magic; int32
// {version, T, W, D, nsteps, record}; int64 {sweeps, likelihoods evaluated,
// first sweep of the history held}; X, lnl, lnp in logical walker order;
// betas; accepted; then the history from that first sweep on: betas, tsw and
// smd per sweep and, when record = 1, the chain, lnl and lnp rows.
static const char CK_MAGIC[8] = {'E', 'M', 'P', 'P', 'T', 'C', 'K', '1'};
constexpr int32_t CK_VERSION = 2;
constexpr size_t CK_CHUNK = (size_t)1 << 23;     // doubles per staged block of history (64 MiB)

// Device <-> file copies of n doubles through a bounded host buffer.
static bool ck_put(FILE* f, const double* dev, size_t n, std::vector<double>& buf) {
    for (size_t o = 0; o < n; o += CK_CHUNK) {
        const size_t c = std::min(CK_CHUNK, n - o);
        buf.resize(c);
        if (hipMemcpy(buf.data(), dev + o, c * 8, hipMemcpyDeviceToHost) != hipSuccess) return false;
        if (fwrite(buf.data(), 8, c, f) != c) return false;
    }
    return true;
}

static bool ck_get(FILE* f, double* dev, size_t n, std::vector<double>& buf) {
    for (size_t o = 0; o < n; o += CK_CHUNK) {
        const size_t c = std::min(CK_CHUNK, n - o);
        buf.resize(c);
        if (fread(buf.data(), 8, c, f) != c) return false;
        if (dev && hipMemcpy(dev + o, buf.data(), c * 8, hipMemcpyHostToDevice) != hipSuccess) return false;
    }
    return true;
}

extern "C" int emp_pt_checkpoint(EmpPT* p, const char* path) {
    if (!p || !path) { emp_set_error("null argument"); return EMP_EINVAL; }
    const PTDev& v = p->dev;
    EMP_HIP(hipStreamSynchronize(p->stream));
    const size_t TW = (size_t)v.T * v.W, T1 = (size_t)std::max(v.T - 1, 1);
    std::vector<double> X, l, pr, b(v.T), buf;
    std::vector<long long> acc(TW);
    int rc = gather_state(p, X, l, pr);
    if (rc) return rc;
    long long cnt[2] = {0, 0};
    EMP_HIP(hipMemcpy(b.data(), v.betas, b.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(acc.data(), v.accepted, acc.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(cnt, v.sweep, 16, hipMemcpyDeviceToHost));
    const long long sw = cnt[0], base = std::min<long long>(p->hist_base, sw);
    const size_t ns = (size_t)(sw - base), nr = ns * (size_t)v.nsteps, r0 = (size_t)base * v.nsteps;
    FILE* f = fopen(path, "wb");
    if (!f) { emp_set_error("cannot open %s for writing", path); return EMP_EIO; }
    int32_t hdr[6] = {CK_VERSION, v.T, v.W, v.D, v.nsteps, v.record};
    long long hdr64[3] = {sw, cnt[1], base};
    bool ok = fwrite(CK_MAGIC, 1, 8, f) == 8 && fwrite(hdr, 4, 6, f) == 6 && fwrite(hdr64, 8, 3, f) == 3 &&
              fwrite(X.data(), 8, X.size(), f) == X.size() && fwrite(l.data(), 8, l.size(), f) == l.size() &&
              fwrite(pr.data(), 8, pr.size(), f) == pr.size() && fwrite(b.data(), 8, b.size(), f) == b.size() &&
              fwrite(acc.data(), 8, acc.size(), f) == acc.size();
    ok = ok && ck_put(f, v.h_betas + (size_t)base * v.T, ns * v.T, buf) &&
         ck_put(f, v.h_tsw + (size_t)base * T1, ns * T1, buf) && ck_put(f, v.h_smd + (size_t)base * T1, ns * T1, buf);
    if (v.record)
        ok = ok && ck_put(f, v.h_chain + r0 * TW * v.D, nr * TW * v.D, buf) &&
             ck_put(f, v.h_lnl + r0 * TW, nr * TW, buf) && ck_put(f, v.h_lnp + r0 * TW, nr * TW, buf);
    ok = (fclose(f) == 0) && ok;
    if (!ok) { emp_set_error("write error on %s", path); return EMP_EIO; }
    return EMP_OK;
}

extern "C" int emp_pt_restore(EmpPT* p, const char* path) {
    if (!p || !path) { emp_set_error("null argument"); return EMP_EINVAL; }
    PTDev& v = p->dev;
    FILE* f = fopen(path, "rb");
    if (!f) { emp_set_error("cannot open %s", path); return EMP_EIO; }
    char mg[8];
    int32_t hdr[6];
    long long hdr64[3];
    const size_t TW = (size_t)v.T * v.W, T1 = (size_t)std::max(v.T - 1, 1);
    std::vector<double> X(TW * v.D), l(TW), pr(TW), b(v.T), buf;
    std::vector<long long> acc(TW);
    bool ok = fread(mg, 1, 8, f) == 8 && memcmp(mg, CK_MAGIC, 8) == 0 && fread(hdr, 4, 6, f) == 6 &&
              hdr[0] == CK_VERSION && hdr[1] == v.T && hdr[2] == v.W && hdr[3] == v.D && hdr[4] == v.nsteps &&
              (hdr[5] == 0 || hdr[5] == 1) && fread(hdr64, 8, 3, f) == 3 &&
              fread(X.data(), 8, X.size(), f) == X.size() && fread(l.data(), 8, l.size(), f) == l.size() &&
              fread(pr.data(), 8, pr.size(), f) == pr.size() && fread(b.data(), 8, b.size(), f) == b.size() &&
              fread(acc.data(), 8, acc.size(), f) == acc.size();
    if (!ok) { fclose(f); emp_set_error("%s is not a checkpoint of this engine shape (T, W, D, nsteps)", path); return EMP_EIO; }
    const long long sw = hdr64[0], nlnl = hdr64[1], base = hdr64[2];
    if (sw < 0 || nlnl < 0 || base < 0 || base > sw) { fclose(f); emp_set_error("%s: bad sweep counters", path); return EMP_EIO; }
    if (sw > v.max_sweeps) {
        fclose(f);
        emp_set_error("checkpoint has %lld sweeps, capacity is %lld", sw, (long long)v.max_sweeps);
        return EMP_EINVAL;
    }
    for (int i = 0; i < v.T; ++i)
        if (!(b[i] >= 0.0 && b[i] <= 1.0)) { fclose(f); emp_set_error("%s: betas outside [0, 1]", path); return EMP_EIO; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    // A recording engine restored from a file without chain rows holds the
    // history from the restored sweep on (hist_base); otherwise it takes the
    // file's history.
    const bool file_rec = hdr[5] == 1;
    const long long nbase = (v.record && !file_rec) ? sw : base;
    const size_t ns = (size_t)(sw - base), nr = ns * (size_t)v.nsteps, r0 = (size_t)base * v.nsteps;
    const bool take = nbase == base;
    ok = ck_get(f, take ? v.h_betas + (size_t)base * v.T : nullptr, ns * v.T, buf) &&
         ck_get(f, take ? v.h_tsw + (size_t)base * T1 : nullptr, ns * T1, buf) &&
         ck_get(f, take ? v.h_smd + (size_t)base * T1 : nullptr, ns * T1, buf);
    if (ok && file_rec) {
        const bool into = v.record && take;
        ok = ck_get(f, into ? v.h_chain + r0 * TW * v.D : nullptr, nr * TW * v.D, buf) &&
             ck_get(f, into ? v.h_lnl + r0 * TW : nullptr, nr * TW, buf) &&
             ck_get(f, into ? v.h_lnp + r0 * TW : nullptr, nr * TW, buf);
    }
    fclose(f);
    if (!ok) { emp_set_error("%s: truncated history", path); return EMP_EIO; }
    const long long cnt[2] = {sw, nlnl};
    EMP_HIP(hipMemcpy(v.X, X.data(), X.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.lnl, l.data(), l.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.lnp, pr.data(), pr.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.betas, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.accepted, acc.data(), acc.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.sweep, cnt, 16, hipMemcpyHostToDevice));
    EMP_HIP(hipMemset(v.rec_owed, 0, sizeof(int)));
    const int64_t n = (int64_t)TW;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, v.map, n);
    EMP_HIP(hipGetLastError());
    if (int rc = mark_all_stale(p)) return rc;
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->host_sweeps = sw;
    p->hist_base = nbase;
    return EMP_OK;
}
