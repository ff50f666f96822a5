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

#include "emp_kernels.h"

using namespace emp;

struct PTDev {
    int T, W, W2, D;
    double a;
    uint32_t key0, key1;
    int adapt_mode;
    double adapt_tau, adapt_nu;
    int64_t max_sweeps;
    int record;
    double* X;          // [T*W slots][D]
    double* lnl;        // [slots]
    double* lnp;        // [slots]
    int* map;           // [T][W] -> slot
    double* swap_l;     // [T][W] scratch for the global-memory swap path
    double* betas;      // [T]
    long long* accepted;  // [T][W] (logical positions)
    long long* sweep;     // [2] sweeps completed, likelihoods evaluated
    double* h_chain;    // [S][T][W][D]
    double* h_lnl;      // [S][T][W]
    double* h_lnp;      // [S][T][W]
    double* h_betas;    // [S][T]
    double* h_tsw;      // [S][T-1]
    double* h_smd;      // [S][T-1]
};

struct EmpPT {
    EmpModel* model;
    EmpPTDesc desc;
    PTDev dev;
    std::vector<void*> allocations;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int wpt = 1;
    size_t swap_lds = 0;         // dynamic LDS of the swap kernel, 0 = global path
    long long host_sweeps = 0;   // mirror of the device counter
    double last_ms_per_sweep = -1.0, last_ms_half = -1.0;
};

// ---------------------------------------------------------------------------
// Fused stretch half-step: propose, prior, likelihood, accept.
// ---------------------------------------------------------------------------
template <int PREC, int WPT>
__global__ __launch_bounds__(BLOCK) void k_pt_half(ModelDev md, PTDev pd, int half, int step) {
    constexpr int SLOTS = NWAVES / WPT;
    __shared__ SlotLDS s_slot[SLOTS];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    __shared__ double s_lp[SLOTS];
    __shared__ int s_acc[SLOTS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave / WPT, wsub = wave % WPT;
    const int g = wsub * 64 + lane, G = 64 * WPT;
    const int64_t b = (int64_t)blockIdx.x * SLOTS + slot;
    const bool task = b < (int64_t)pd.T * pd.W2;
    SlotLDS& sl = s_slot[slot];
    const long long sweep = *pd.sweep;
    int t = 0, k = 0, sk = 0;
    double z = 1.0, uacc = 0.5;
    if (task) {
        t = (int)(b / pd.W2);
        k = half * pd.W2 + (int)(b % pd.W2);
        const uint32_t wid = (uint32_t)(t * pd.W + k);
        U4 c1{wid, (uint32_t)sweep, (uint32_t)((unsigned long long)sweep >> 32), (uint32_t)(step * 2 + half)};
        U4 r1 = philox4x32_10(c1, pd.key0, pd.key1);
        U4 c2 = c1;
        c2.w |= 0x80000000u;
        U4 r2 = philox4x32_10(c2, pd.key0, pd.key1);
        const double uz = u01(r1.x, r1.y), uj = u01(r1.z, r1.w);
        uacc = u01(r2.x, r2.y);
        const double zz = (pd.a - 1.0) * uz + 1.0;
        z = zz * zz / pd.a;                                  // Goodman & Weare (2010) eq. 9
        const int j = (1 - half) * pd.W2 + (int)(uj * (double)pd.W2);
        sk = pd.map[t * pd.W + k];
        const int sj = pd.map[t * pd.W + j];
        const double* Xk = pd.X + (size_t)sk * pd.D;
        const double* Xj = pd.X + (size_t)sj * pd.D;
        for (int d = g; d < md.D; d += G) sl.theta[md.free_idx[d]] = Xj[d] + z * (Xk[d] - Xj[d]);
        for (int q = g; q < md.Dfull; q += G)
            if (md.inv_free[q] < 0) sl.theta[q] = md.fixed_val[q];
    }
    tab_load(md, s_tile);
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
            if (sl.planet[p].e <= KT_EMAX) sl.ktab[p][lane] = ktab_node(sl.planet[p], s_tile.tab, lane);
    }
    __syncthreads();
    double ll = block_lnl<PREC, WPT, false>(md, sl, s_tile, s_scan, live, wsub, nullptr, nullptr);
    if (task && wsub == 0 && lane == 0) {
        int acc = 0;
        if (live) {
            atomicAdd((unsigned long long*)(pd.sweep + 1), 1ull);
            if (!(ll == ll) || ll == __builtin_inf()) ll = -__builtin_inf();
            const double beta = pd.betas[t];
            const double tn = beta == 0.0 ? 0.0 : beta * ll;
            const double to = beta == 0.0 ? 0.0 : beta * pd.lnl[sk];
            const double lnalpha = (double)(md.D - 1) * log(z) + (tn - to) + (lp - pd.lnp[sk]);
            if (log(uacc) < lnalpha) {
                pd.lnl[sk] = ll;
                pd.lnp[sk] = lp;
                pd.accepted[t * pd.W + k] += 1;
                acc = 1;
            }
        }
        s_acc[slot] = acc;
    }
    __syncthreads();
    if (task && s_acc[slot]) {
        double* Xk = pd.X + (size_t)sk * pd.D;
        for (int d = g; d < md.D; d += G) Xk[d] = sl.theta[md.free_idx[d]];
    }
}

// ---------------------------------------------------------------------------
// Adjacent-temperature swaps (hottest pair first), swap statistics, ladder
// adaptation and the sweep counter.  One workgroup.  The map and the
// per-position likelihoods are staged in LDS when they fit (USE_LDS), else
// worked on in global memory.
// ---------------------------------------------------------------------------
constexpr int SWAP_BLOCK = 1024;
constexpr int SWAP_WAVES = SWAP_BLOCK / 64;

__device__ __forceinline__ uint32_t swap_partner(const PTDev& pd, long long sweep, int h, int i) {
    return feistel_perm((uint32_t)i, (uint32_t)pd.W, pd.key0 ^ (uint32_t)sweep ^ 0x5bd1e995u,
                        pd.key1 ^ ((uint32_t)h * 0x27d4eb2fu) ^ (uint32_t)((unsigned long long)sweep >> 32));
}

__device__ __forceinline__ double swap_logu(const PTDev& pd, long long sweep, int h, int i) {
    U4 cu{(uint32_t)i, (uint32_t)sweep, (uint32_t)((unsigned long long)sweep >> 32), 0x40000000u | (uint32_t)h};
    U4 r = philox4x32_10(cu, pd.key0, pd.key1);
    return log(u01(r.x, r.y));
}

// USE_LDS: the map, the per-position likelihoods and the pairing/uniforms of
// every pair (computed in parallel up front; they do not depend on the state)
// live in LDS, so the sequential hottest-first loop is compare-and-exchange.
template <bool USE_LDS>
__global__ __launch_bounds__(SWAP_BLOCK) void k_pt_swap(PTDev pd) {
    extern __shared__ double s_dyn[];
    __shared__ double s_red[64][2];
    __shared__ double s_rate[64];
    const int tid = threadIdx.x, lane = tid & 63;
    const long long sweep = *pd.sweep;
    const int T = pd.T, W = pd.W;
    const int n = T * W, np = (T - 1) * W;
    double* L = USE_LDS ? s_dyn : pd.swap_l;
    double* LU = s_dyn + n;                          // [T-1][W] log uniforms (LDS path)
    int* Mp = USE_LDS ? (int*)(s_dyn + n + np) : pd.map;
    int* JP = (int*)(s_dyn + n + np) + n;            // [T-1][W] partners (LDS path)
    for (int e = tid; e < 128; e += SWAP_BLOCK) (&s_red[0][0])[e] = 0.0;
    for (int e = tid; e < n; e += SWAP_BLOCK) {
        const int m = pd.map[e];
        if (USE_LDS) Mp[e] = m;
        L[e] = pd.lnl[m];
    }
    if (USE_LDS) {
        for (int e = tid; e < np; e += SWAP_BLOCK) {
            const int h = e / W + 1, i = e - (h - 1) * W;
            JP[e] = (int)swap_partner(pd, sweep, h, i);
            LU[e] = swap_logu(pd, sweep, h, i);
        }
    }
    __syncthreads();
    for (int h = T - 1; h >= 1; --h) {
        const int c = h - 1;
        const double dbeta = pd.betas[c] - pd.betas[h];
        double nacc = 0.0, dist = 0.0;
        for (int i = tid; i < W; i += SWAP_BLOCK) {
            const int e = (h - 1) * W + i;
            const int jp = USE_LDS ? JP[e] : (int)swap_partner(pd, sweep, h, i);
            const double lu = USE_LDS ? LU[e] : swap_logu(pd, sweep, h, i);
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
        if (!USE_LDS) __threadfence();
        __syncthreads();
    }
    if (tid < T - 1) {
        const double na = s_red[tid][0], di = s_red[tid][1];
        s_rate[tid] = na / W;
        if (sweep < pd.max_sweeps) {
            pd.h_tsw[(size_t)sweep * (T - 1) + tid] = na / W;
            pd.h_smd[(size_t)sweep * (T - 1) + tid] = di / W;
        }
    }
    if (USE_LDS)
        for (int e = tid; e < n; e += SWAP_BLOCK) pd.map[e] = Mp[e];
    __syncthreads();
    if (tid == 0) {
        // Vousden, Farr & Mandel (2016): S_i = ln(T_i - T_{i-1}) moves by
        // kappa (A_{i-1} - A_i), kappa = tau / (n + tau) / nu; the coldest and
        // hottest temperatures are fixed.
        if (pd.adapt_mode == 0 && T >= 3) {
            const double kappa = pd.adapt_tau / ((double)sweep + pd.adapt_tau) / pd.adapt_nu;
            double Tprev_old = 1.0 / pd.betas[0];
            double acc_T = Tprev_old;
            for (int i = 1; i < T - 1; ++i) {
                const double Ti_old = 1.0 / pd.betas[i];
                const double dS = kappa * (s_rate[i - 1] - s_rate[i]);
                acc_T += (Ti_old - Tprev_old) * exp(dS);
                Tprev_old = Ti_old;
                pd.betas[i] = 1.0 / acc_T;
            }
        }
        if (sweep < pd.max_sweeps)
            for (int i = 0; i < T; ++i) pd.h_betas[(size_t)sweep * T + i] = pd.betas[i];
        *pd.sweep = sweep + 1;
    }
}

__global__ void k_pt_record(PTDev pd) {
    const long long s = *pd.sweep - 1;
    if (s < 0 || s >= pd.max_sweeps) return;
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

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
template <int PREC, int WPT>
static void launch_half_t(const ModelDev& md, const PTDev& pd, int half, int step, hipStream_t s) {
    constexpr int SLOTS = NWAVES / WPT;
    const int64_t tasks = (int64_t)pd.T * pd.W2;
    const int64_t grid = (tasks + SLOTS - 1) / SLOTS;
    hipLaunchKernelGGL((k_pt_half<PREC, WPT>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, pd, half, step);
}

static int launch_half(EmpPT* p, int half, int step) {
    const ModelDev& md = p->model->dev;
    const int prec = p->desc.precision, w = p->wpt;
#define EMP_CASE(P, W) if (prec == P && w == W) { launch_half_t<P, W>(md, p->dev, half, step, p->stream); EMP_HIP(hipGetLastError()); return EMP_OK; }
    EMP_CASE(0, 1) EMP_CASE(0, 2) EMP_CASE(0, 4) EMP_CASE(0, 8)
    EMP_CASE(1, 1) EMP_CASE(1, 2) EMP_CASE(1, 4) EMP_CASE(1, 8)
#undef EMP_CASE
    emp_set_error("bad precision/wpt");
    return EMP_EINVAL;
}

static int launch_sweep(EmpPT* p) {
    for (int st = 0; st < p->desc.nsteps; ++st)
        for (int h = 0; h < 2; ++h) {
            int rc = launch_half(p, h, st);
            if (rc) return rc;
        }
    if (p->swap_lds)
        hipLaunchKernelGGL(k_pt_swap<true>, dim3(1), dim3(SWAP_BLOCK), p->swap_lds, p->stream, p->dev);
    else
        hipLaunchKernelGGL(k_pt_swap<false>, dim3(1), dim3(SWAP_BLOCK), 0, p->stream, p->dev);
    EMP_HIP(hipGetLastError());
    if (p->dev.record) {
        const size_t nx = (size_t)p->dev.T * p->dev.W * p->dev.D;
        const unsigned grid = (unsigned)std::min<size_t>((nx + 255) / 256, 4096);
        hipLaunchKernelGGL(k_pt_record, dim3(grid), dim3(256), 0, p->stream, p->dev);
        EMP_HIP(hipGetLastError());
    }
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

// Evaluates lnp and lnl of every slot (slots hold the positions in logical
// order when the map is the identity, which is how states are loaded).
static int eval_state(EmpPT* p) {
    PTDev& d = p->dev;
    const int64_t n = (int64_t)d.T * d.W;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, d.map, n);
    EMP_HIP(hipGetLastError());
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
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
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
    if (d.nsteps < 1) { emp_set_error("nsteps must be >= 1"); return EMP_EINVAL; }
    if (d.precision != EMP_FP64 && d.precision != EMP_MIXED) { emp_set_error("bad precision"); return EMP_EINVAL; }
    if (!(d.a > 1.0)) { emp_set_error("stretch scale a must be > 1"); return EMP_EINVAL; }
    if (d.adapt_mode != -1 && d.adapt_mode != 0) { emp_set_error("adapt_mode must be -1 (frozen) or 0"); return EMP_EUNSUP; }
    if (d.adapt_mode == 0 && !(d.adapt_tau > 0 && d.adapt_nu > 0)) { emp_set_error("adapt_tau and adapt_nu must be > 0"); return EMP_EINVAL; }
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
    v.record = d.record ? 1 : 0;
    const size_t TW = (size_t)v.T * v.W, S = (size_t)v.max_sweeps;
    int rc = 0;
    if ((rc = alloc(p, &v.X, TW * v.D)) || (rc = alloc(p, &v.lnl, TW)) || (rc = alloc(p, &v.lnp, TW)) ||
        (rc = alloc(p, &v.map, TW)) || (rc = alloc(p, &v.swap_l, TW)) ||
        (rc = alloc(p, &v.betas, (size_t)v.T)) || (rc = alloc(p, &v.accepted, TW)) || (rc = alloc(p, &v.sweep, 2)) ||
        (rc = alloc(p, &v.h_betas, S * v.T)) || (rc = alloc(p, &v.h_tsw, S * std::max(v.T - 1, 1))) ||
        (rc = alloc(p, &v.h_smd, S * std::max(v.T - 1, 1)))) {
        emp_pt_destroy(p);
        return rc;
    }
    if (v.record) {
        if ((rc = alloc(p, &v.h_chain, S * TW * v.D)) || (rc = alloc(p, &v.h_lnl, S * TW)) || (rc = alloc(p, &v.h_lnp, S * TW))) {
            emp_pt_destroy(p);
            return rc;
        }
    }
    // Stage the swap round in LDS when the map, the likelihoods and the
    // pre-drawn pairings/uniforms fit (12 bytes per walker and per walker-pair,
    // inside the default 64 KiB workgroup allocation); otherwise the same
    // kernel works in global memory and draws the randoms inline.
    const size_t lds = (TW + (size_t)(v.T - 1) * v.W) * (sizeof(double) + sizeof(int));
    p->swap_lds = lds <= 62 * 1024 ? lds : 0;
#define EMP_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { int r_ = emp_hip_fail(e_, #x, __FILE__, __LINE__); emp_pt_destroy(p); return r_; } } while (0)
    EMP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    EMP_TRY(hipMemcpy(v.X, p0, TW * v.D * sizeof(double), hipMemcpyHostToDevice));
    EMP_TRY(hipMemcpy(v.betas, betas, (size_t)v.T * sizeof(double), hipMemcpyHostToDevice));
    EMP_TRY(hipMemset(v.accepted, 0, TW * sizeof(long long)));
    EMP_TRY(hipMemset(v.sweep, 0, 2 * sizeof(long long)));
#undef EMP_TRY
    p->wpt = emp_pick_wpt(m, (int64_t)v.T * v.W2, d.precision);
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
    if (use_graph) {
        if (!p->graph_exec) {
            hipGraph_t g;
            EMP_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
            rc = launch_sweep(p);
            hipError_t ce = hipStreamEndCapture(p->stream, &g);
            if (rc) return rc;
            EMP_HIP(ce);
            EMP_HIP(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
            EMP_HIP(hipGraphDestroy(g));
        }
        for (int64_t s = 0; s < nsweeps; ++s) EMP_HIP(hipGraphLaunch(p->graph_exec, p->stream));
    } else {
        for (int64_t s = 0; s < nsweeps && !rc; ++s) rc = launch_sweep(p);
        if (rc) return rc;
    }
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

// Times `n` half-step kernel launches individually with events (no graph).
// The launches advance the chain like ordinary half-steps of sweep 0's step 0.
extern "C" int emp_pt_time_halfstep(EmpPT* p, int32_t n, double* ms_avg) {
    if (!p || n < 1 || !ms_avg) { emp_set_error("bad argument"); return EMP_EINVAL; }
    std::vector<hipEvent_t> ev(2 * (size_t)n);
    for (auto& e : ev) EMP_HIP(hipEventCreate(&e));
    for (int i = 0; i < n; ++i) {
        EMP_HIP(hipEventRecord(ev[2 * i], p->stream));
        int rc = launch_half(p, i & 1, 0);
        if (rc) return rc;
        EMP_HIP(hipEventRecord(ev[2 * i + 1], p->stream));
    }
    EMP_HIP(hipStreamSynchronize(p->stream));
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
        case EMP_PT_CHAIN: if (!v.record) goto norec; sh[0] = S; sh[1] = v.T; sh[2] = v.W; sh[3] = v.D; src = v.h_chain; break;
        case EMP_PT_LNL: if (!v.record) goto norec; sh[0] = S; sh[1] = v.T; sh[2] = v.W; src = v.h_lnl; break;
        case EMP_PT_LNP: if (!v.record) goto norec; sh[0] = S; sh[1] = v.T; sh[2] = v.W; src = v.h_lnp; break;
        case EMP_PT_BETAS: sh[0] = S; sh[1] = v.T; src = v.h_betas; break;
        case EMP_PT_TSW: sh[0] = S; sh[1] = std::max(v.T - 1, 0); src = v.h_tsw; break;
        case EMP_PT_SMD: sh[0] = S; sh[1] = std::max(v.T - 1, 0); src = v.h_smd; break;
        case EMP_PT_ACCEPTED: sh[0] = v.T; sh[1] = v.W; src = v.accepted; elem = sizeof(long long); break;
        case EMP_PT_STATE_X: sh[0] = v.T; sh[1] = v.W; sh[2] = v.D; src = Xg.data(); host_src = true; break;
        case EMP_PT_STATE_LNL: sh[0] = v.T; sh[1] = v.W; src = lg.data(); host_src = true; break;
        case EMP_PT_STATE_LNP: sh[0] = v.T; sh[1] = v.W; src = pg.data(); host_src = true; break;
        case EMP_PT_STATE_BETAS: sh[0] = v.T; src = v.betas; break;
        case EMP_PT_NSWEEPS: sh[0] = 2; src = v.sweep; elem = sizeof(long long); break;
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

extern "C" int emp_pt_set_adaptation(EmpPT* p, int32_t mode, double tau, double nu) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    if (mode != -1 && mode != 0) { emp_set_error("adapt_mode must be -1 (frozen) or 0"); return EMP_EUNSUP; }
    if (mode == 0 && !(tau > 0 && nu > 0)) { emp_set_error("adapt_tau and adapt_nu must be > 0"); return EMP_EINVAL; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->dev.adapt_mode = mode;
    p->dev.adapt_tau = tau;
    p->dev.adapt_nu = nu;
    if (p->graph_exec) {   // kernel arguments changed: capture again on the next run
        EMP_HIP(hipGraphExecDestroy(p->graph_exec));
        p->graph_exec = nullptr;
    }
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

// Checkpoint layout: magic, version, T, W, D, sweeps, then X, lnl, lnp (in
// logical walker order), betas, accepted (little-endian native binary).
static const char CK_MAGIC[8] = {'E', 'M', 'P', 'P', 'T', 'C', 'K', '1'};

extern "C" int emp_pt_checkpoint(EmpPT* p, const char* path) {
    if (!p || !path) { emp_set_error("null argument"); return EMP_EINVAL; }
    const PTDev& v = p->dev;
    EMP_HIP(hipStreamSynchronize(p->stream));
    const size_t TW = (size_t)v.T * v.W;
    std::vector<double> X, l, pr, b(v.T);
    std::vector<long long> acc(TW);
    int rc = gather_state(p, X, l, pr);
    if (rc) return rc;
    long long sw = 0;
    EMP_HIP(hipMemcpy(b.data(), v.betas, b.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(acc.data(), v.accepted, acc.size() * 8, hipMemcpyDeviceToHost));
    EMP_HIP(hipMemcpy(&sw, v.sweep, 8, hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) { emp_set_error("cannot open %s for writing", path); return EMP_EIO; }
    int32_t hdr[4] = {1, v.T, v.W, v.D};
    bool ok = fwrite(CK_MAGIC, 1, 8, f) == 8 && fwrite(hdr, 4, 4, f) == 4 && fwrite(&sw, 8, 1, f) == 1 &&
              fwrite(X.data(), 8, X.size(), f) == X.size() && fwrite(l.data(), 8, l.size(), f) == l.size() &&
              fwrite(pr.data(), 8, pr.size(), f) == pr.size() && fwrite(b.data(), 8, b.size(), f) == b.size() &&
              fwrite(acc.data(), 8, acc.size(), f) == acc.size();
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
    int32_t hdr[4];
    long long sw = 0;
    const size_t TW = (size_t)v.T * v.W;
    std::vector<double> X(TW * v.D), l(TW), pr(TW), b(v.T);
    std::vector<long long> acc(TW);
    bool ok = fread(mg, 1, 8, f) == 8 && memcmp(mg, CK_MAGIC, 8) == 0 && fread(hdr, 4, 4, f) == 4 &&
              hdr[0] == 1 && hdr[1] == v.T && hdr[2] == v.W && hdr[3] == v.D && fread(&sw, 8, 1, f) == 1 &&
              fread(X.data(), 8, X.size(), f) == X.size() && fread(l.data(), 8, l.size(), f) == l.size() &&
              fread(pr.data(), 8, pr.size(), f) == pr.size() && fread(b.data(), 8, b.size(), f) == b.size() &&
              fread(acc.data(), 8, acc.size(), f) == acc.size();
    fclose(f);
    if (!ok) { emp_set_error("%s is not a checkpoint of this engine shape", path); return EMP_EIO; }
    if (sw > v.max_sweeps) { emp_set_error("checkpoint has %lld sweeps, capacity is %lld", sw, (long long)v.max_sweeps); return EMP_EINVAL; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    EMP_HIP(hipMemcpy(v.X, X.data(), X.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.lnl, l.data(), l.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.lnp, pr.data(), pr.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.betas, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.accepted, acc.data(), acc.size() * 8, hipMemcpyHostToDevice));
    EMP_HIP(hipMemcpy(v.sweep, &sw, 8, hipMemcpyHostToDevice));
    const int64_t n = (int64_t)TW;
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, v.map, n);
    EMP_HIP(hipGetLastError());
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->host_sweeps = sw;
    return EMP_OK;
}
