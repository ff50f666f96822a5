// Batched independent systems: S separate RV models, each with its own
// parallel-tempering ensemble, advanced together by ONE persistent kernel
// launch per sweep (replayed from a hipGraph).
//
// Scheduling.  A sweep of one system (nsteps x two stretch halves over its
// T x W walkers, the swap round, the ladder update, the statistics and the
// optional record) is one work item, done start to finish by one workgroup:
// the system's epochs stay resident in LDS (N <= TILE) for all its
// likelihoods, and no workgroup ever waits for another.  Workgroups pop
// items from a queue ordered longest first (cost ~ N (k c_kep + c_ep) T W),
// so the ragged tail of small systems fills in behind the big ones.  The
// last workgroup to leave resets the queue and advances the sweep counter,
// so the captured launch replays unchanged.
//
// Randomness: the counters of the single-system engine with the system index
// mixed in (system 0 draws exactly what the single-system engine draws).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "emp_pt_common.h"

using namespace emp;

constexpr int SYS_MAX_TASKS = 512;   // T W / 2 per system (per-task prior and list in LDS)

struct SysDev {
    int S, T, W, nsteps;
    double a;
    uint32_t key0, key1;
    int adapt_mode;
    double adapt_tau, adapt_nu;
    const ModelDev* md;        // [S]
    const int* order;          // [S] queue order
    const int64_t* xoff;       // [S] first element of system s in X
    double* X;                 // system after system: [T W slots][D_s]
    double* lnl;               // [S][T W] by slot
    double* lnp;               // [S][T W]
    int* map;                  // [S][T W] position -> slot
    double* betas;             // [S][T]
    long long* accepted;       // [S][T W] by position
    double* tsw_sum;           // [S][T-1] summed over sweeps
    double* smd_sum;           // [S][T-1]
    unsigned long long* ctl;   // queue head, workgroups finished, sweeps done, likelihoods evaluated
    int rec_every, rec_temps;
    int64_t max_rec;
    const int64_t* rxoff;      // [S] first element of system s in one record of h_x
    int64_t rec_x;             // elements of one record of h_x
    double* h_x;               // [R][sum_s rec_temps W D_s]
    double* h_lnl;             // [R][S][rec_temps W]
    double* h_lnp;             // [R][S][rec_temps W]
};

// The model table does not change during a launch.  Reading it through the
// constant address space lets the compiler re-load a field where it is used
// instead of holding the whole ModelDev in registers across the sweep.
typedef __attribute__((address_space(4))) const ModelDev ConstModelDev;

__device__ __forceinline__ const ModelDev& sys_model(const SysDev& sd, int s) {
    const ConstModelDev* c = (const ConstModelDev*)sd.md;
    return *(const ModelDev*)&c[s];
}

__device__ __forceinline__ PTView sys_view(const SysDev& sd, int s, int D) {
    const size_t TW = (size_t)sd.T * sd.W;
    return pt_view(sd.T, sd.W, D, sd.a, sd.key0, sd.key1, (uint32_t)s, sd.adapt_mode, sd.adapt_tau, sd.adapt_nu,
                   sd.X + sd.xoff[s], sd.lnl + s * TW, sd.lnp + s * TW, sd.map + s * TW, sd.betas + (size_t)s * sd.T,
                   sd.accepted + s * TW);
}

// Prior of every proposal of one half-step of one system into lp[ntask], in
// three workgroup-wide steps over `scratch` (bytes of LDS not otherwise in
// use): the proposals (one thread per task), every (task, term) pair of the
// prior's summation sequence md.pseq (all loads in flight together), then
// the sum of each task's terms in sequence order with prior_seq's -inf
// return (one thread per task), bit-identical to prior_seq.  Ends with a
// barrier.
__device__ inline void sys_priors(const ModelDev& md, const PTView& pv, long long sweep, int half, int step,
                                  int ntask, char* scratch, size_t bytes, double* lp, int nthr) {
#pragma clang fp contract(off)
    double* z = (double*)scratch;                      // [ntask]
    int* sk = (int*)(z + ntask);                       // [ntask]
    int* sj = sk + ntask;                              // [ntask]
    double* terms = (double*)(sj + ntask);             // 8-byte aligned: z, sk, sj take 16 ntask bytes
    const int L = md.npseq;
    const int chunk = (int)min((size_t)ntask, (bytes - (size_t)((char*)terms - scratch)) / (sizeof(double) * L));
    for (int b = threadIdx.x; b < ntask; b += nthr) {
        const Proposal pr = pt_propose(pv, sweep, half, step, b);
        z[b] = pr.z;
        sk[b] = pr.sk;
        sj[b] = pr.sj;
    }
    __syncthreads();
    for (int c0 = 0; c0 < ntask; c0 += chunk) {
        const int cn = min(chunk, ntask - c0);
        for (int e = threadIdx.x; e < cn * L; e += nthr) {
            const int bl = e / L, i = e - bl * L, b = c0 + bl;
            const int q = md.pseq[i];
            if (q == PSEQ_CHECK) continue;
            const double* Xk = pv.X + (size_t)sk[b] * pv.D;
            const double* Xj = pv.X + (size_t)sj[b] * pv.D;
            double v;
            if (q >= 0) {
                v = prior_term(md.prior[q], pt_prop_value(Xk, Xj, z[b], md.inv_free[q]));
            } else {
                const int pb = PSEQ_ECC - q;
                const int o = md.off_kep + 5 * pb + 3;
                const int f3 = md.inv_free[o], f4 = md.inv_free[o + 1];
                const double s3 = f3 >= 0 ? pt_prop_value(Xk, Xj, z[b], f3) : md.fixed_val[o];
                const double s4 = f4 >= 0 ? pt_prop_value(Xk, Xj, z[b], f4) : md.fixed_val[o + 1];
                const double x = s3 * s3 + s4 * s4;
                v = prior_term(md.ecc_prior[pb], x);
            }
            terms[e] = v;
        }
        __syncthreads();
        for (int bl = threadIdx.x; bl < cn; bl += nthr) {
            const double* tv = terms + (size_t)bl * L;
            double s = 0.0;
            for (int i = 0; i < L; ++i) {
                if (md.pseq[i] == PSEQ_CHECK) {
                    if (s == -__builtin_inf()) break;
                } else {
                    s += tv[i];
                }
            }
            lp[c0 + bl] = s;
        }
        __syncthreads();
    }
}

// One sweep of every system.  NW waves per workgroup, WPT waves per walker;
// EXT: the instance with the extended model features (needs_ext), used when
// any system of the batch has them.
template <int PREC, int WPT, int NW, bool EXT>
__global__ __launch_bounds__(NW * 64) void k_sys_sweep(SysDev sd) {
    constexpr int NT = NW * 64, SLOTS = NW / WPT;
    __shared__ SlotLDS<SLOTS> s_slot[SLOTS];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    __shared__ int s_acc[SLOTS];
    __shared__ TaskOps s_ops[SLOTS];
    __shared__ double s_bet[64], s_ld[64];
    __shared__ double s_red[64][2];
    __shared__ double s_rate[64], s_smd[64], s_sig[64];
    __shared__ double s_tlp[SYS_MAX_TASKS];    // prior of each task's proposal
    __shared__ int s_list[SYS_MAX_TASKS];      // tasks with a finite prior, in order
    __shared__ int s_wcnt[MAX_WAVES];
    __shared__ long long s_item;
    __shared__ unsigned long long s_cnt;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int slot = wave / WPT, wsub = wave % WPT;
    const long long sweep = (long long)sd.ctl[2];
    const int T = sd.T, W = sd.W, TW = T * W, ntask = T * (W / 2);
    tab_load(sys_model(sd, 0), s_tile);      // one table serves every system
    for (;;) {
        __syncthreads();             // the previous item is finished with LDS
        if (tid == 0) {
            s_item = (long long)atomicAdd(&sd.ctl[0], 1ull);
            s_cnt = 0ull;
        }
        __syncthreads();
        // uniform by construction; say so, so that the model is read with
        // scalar loads into scalar registers
        const int it = __builtin_amdgcn_readfirstlane((int)min(s_item, (long long)sd.S));
        if (it >= sd.S) break;
        const int s = __builtin_amdgcn_readfirstlane(sd.order[it]);
        const ModelDev& md = sys_model(sd, s);
        const PTView pv = sys_view(sd, s, md.D);
        const bool res = md.N <= TILE;
        if (res) tile_load(md, s_tile, 0, md.N);   // published by the first task's barrier
        for (int st = 0; st < sd.nsteps; ++st)
            for (int half = 0; half < 2; ++half) {
                // Priors of all proposals (in the idle slot memory), then only
                // the proposals inside the prior support take likelihood slots.
                sys_priors(md, pv, sweep, half, st, ntask, (char*)s_slot, sizeof(s_slot), s_tlp, NT);
                __syncthreads();
                const int nlive = compact_live(s_tlp, ntask, s_list, s_wcnt, NT);
                for (int q0 = 0; q0 < nlive; q0 += SLOTS) {
                    const int q = q0 + slot;
                    const int b = q < nlive ? s_list[q] : -1;
                    pt_half_task<PREC, WPT, EXT>(md, pv, sweep, half, st, b, b >= 0 ? s_tlp[b] : -__builtin_inf(),
                                            s_slot[slot], s_tile, s_scan, s_acc, s_ops, slot, wsub, res, &s_cnt);
                }
                __syncthreads();     // this half's writes before the other half reads them
            }
        // Swap round on LDS copies that alias the (now idle) slot scratch.
        double* L = (double*)s_slot;
        double* LU = L + TW;
        int* Mp = (int*)(LU + (T - 1) * W);
        int* JP = Mp + TW;
        pt_swap_stage(pv, sweep, L, Mp, LU, JP, NT);
        for (int e = tid; e < T; e += NT) s_bet[e] = pv.betas[e];
        __syncthreads();
        pt_swap_round(pv, sweep, L, Mp, LU, JP, false, s_red, s_bet, NT);
        for (int e = tid; e < TW; e += NT) pv.map[e] = Mp[e];
        if (tid < T - 1) {
            const double r = s_red[tid][0] / W;
            s_rate[tid] = r;
            s_smd[tid] = s_red[tid][1] / W;
            sd.tsw_sum[(size_t)s * (T - 1) + tid] += r;
            sd.smd_sum[(size_t)s * (T - 1) + tid] += s_red[tid][1] / W;
        }
        if (sd.rec_every > 0 && (sweep + 1) % sd.rec_every == 0) {
            const long long r = (sweep + 1) / sd.rec_every - 1;
            if (r < sd.max_rec) {
                const int nw = sd.rec_temps * W, D = md.D;
                double* hx = sd.h_x + r * sd.rec_x + sd.rxoff[s];
                for (int e = tid; e < nw * D; e += NT) {
                    const int w = e / D, d = e - w * D;
                    hx[e] = pv.X[(size_t)Mp[w] * D + d];
                }
                const size_t o = ((size_t)r * sd.S + s) * nw;
                for (int w = tid; w < nw; w += NT) {
                    sd.h_lnl[o + w] = pv.lnl[Mp[w]];
                    sd.h_lnp[o + w] = pv.lnp[Mp[w]];
                }
            }
        }
        if (pv.adapt_mode == 2) pt_lnl_spread(pv, L, s_sig, NT);
        __syncthreads();
        pt_ladder_update(pv, sweep, s_rate, s_smd, s_sig, s_bet, s_ld, NT);
        if (tid == 0) atomicAdd(&sd.ctl[3], s_cnt);
    }
    if (tid == 0) {
        __threadfence();
        const unsigned long long f = atomicAdd(&sd.ctl[1], 1ull);
        if (f == (unsigned long long)gridDim.x - 1) {   // every other workgroup has left
            sd.ctl[0] = 0ull;
            sd.ctl[1] = 0ull;
            sd.ctl[2] = (unsigned long long)(sweep + 1);
        }
    }
}

// Initial prior and likelihood of every walker (the state loaded into slot =
// position order), one workgroup per system; the same arithmetic as the
// batched likelihood kernel with the same waves per walker.
template <int PREC, int WPT, int NW, bool EXT>
__global__ __launch_bounds__(NW * 64) void k_sys_init(SysDev sd) {
    constexpr int NT = NW * 64, SLOTS = NW / WPT;
    __shared__ SlotLDS<SLOTS> s_slot[SLOTS];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    __shared__ double s_lp[SLOTS];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int slot = wave / WPT, wsub = wave % WPT;
    const int s = blockIdx.x;
    const ModelDev& md = sys_model(sd, s);
    const PTView pv = sys_view(sd, s, md.D);
    const int TW = sd.T * sd.W;
    SlotLDS<SLOTS>& sl = s_slot[slot];
    tab_load(md, s_tile);
    for (int e = tid; e < TW; e += NT) pv.map[e] = e;
    for (int b0 = 0; b0 < TW; b0 += SLOTS) {
        const int b = b0 + slot;
        const bool live = b < TW;
        __syncthreads();
        if (live) expand_theta(md, sl, pv.X + (size_t)b * md.D, wsub * 64 + lane, 64 * WPT);
        __syncthreads();
        if (live && wsub == 0 && lane == 0) s_lp[slot] = prior_seq(md, sl.theta);
        slot_prepare<PREC, WPT>(md, sl, s_tile, live, wsub, lane);
        double v = block_lnl<PREC, WPT, false, false, false, EXT>(md, sl, s_tile, s_scan, live, wsub, nullptr, nullptr);
        if (live && wsub == 0 && lane == 0) {
            if (!(v == v) || v == __builtin_inf()) v = -__builtin_inf();
            const double lp = s_lp[slot];
            pv.lnp[b] = lp;
            pv.lnl[b] = lp > -__builtin_inf() ? v : -__builtin_inf();
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct EmpSystems {
    EmpSystemsDesc desc;
    SysDev dev;
    std::vector<void*> allocations;
    std::vector<int> D;
    std::vector<int64_t> xoff, rxoff;
    int* d_order_lpt = nullptr;
    int* d_order_input = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    EmpSystemsTuning tune{};
    int n_cu = 256;
    bool ext = false;          // some system needs the extended instance (needs_ext)
    int64_t x_total = 0;
    double last_ms = -1.0;
};

// Static LDS of the sweep kernel's slot array (the swap round's working set
// must fit in it).
static size_t slot_bytes(int nw, int wpt) {
    const int slots = nw / wpt;
    return (size_t)slots * (slots <= 2 ? sizeof(SlotLDS<2>) : sizeof(SlotLDS<4>));
}
static size_t swap_bytes(int T, int W) { return (size_t)(2 * T - 1) * W * (sizeof(double) + sizeof(int)); }

#define EMP_SYS_CONFIGS(X) X(0, 1, 8) X(0, 1, 16) X(0, 2, 8) X(0, 2, 16) X(1, 1, 8) X(1, 1, 16) X(1, 2, 8) X(1, 2, 16)

template <int PREC, int WPT, int NW>
static const void* sweep_fn(bool ext) {
    return ext ? (const void*)&k_sys_sweep<PREC, WPT, NW, true> : (const void*)&k_sys_sweep<PREC, WPT, NW, false>;
}

static int launch_sweep(EmpSystems* p) {
    const EmpSystemsTuning& tu = p->tune;
    const int prec = p->desc.precision;
    const int grid = std::max(1, std::min(p->desc.n_systems, p->n_cu * tu.wgs_per_cu));
    p->dev.order = tu.order == 0 ? p->d_order_lpt : p->d_order_input;
#define EMP_CASE(P, WP, NW)                                                                                \
    if (prec == P && tu.waves_per_theta == WP && tu.waves_per_wg == NW) {                                 \
        if (p->ext) hipLaunchKernelGGL((k_sys_sweep<P, WP, NW, true>), dim3(grid), dim3(NW * 64), 0, p->stream, p->dev); \
        else hipLaunchKernelGGL((k_sys_sweep<P, WP, NW, false>), dim3(grid), dim3(NW * 64), 0, p->stream, p->dev); \
        EMP_HIP(hipGetLastError());                                                                       \
        return EMP_OK;                                                                                    \
    }
    EMP_SYS_CONFIGS(EMP_CASE)
#undef EMP_CASE
    emp_set_error("unsupported launch configuration (%d waves per workgroup, %d per walker)", tu.waves_per_wg,
                  tu.waves_per_theta);
    return EMP_EINVAL;
}

static int launch_init(EmpSystems* p) {
    const EmpSystemsTuning& tu = p->tune;
    const int prec = p->desc.precision;
#define EMP_CASE(P, WP, NW)                                                                                           \
    if (prec == P && tu.waves_per_theta == WP && tu.waves_per_wg == NW) {                                            \
        if (p->ext) hipLaunchKernelGGL((k_sys_init<P, WP, NW, true>), dim3(p->desc.n_systems), dim3(NW * 64), 0, p->stream, p->dev); \
        else hipLaunchKernelGGL((k_sys_init<P, WP, NW, false>), dim3(p->desc.n_systems), dim3(NW * 64), 0, p->stream, p->dev); \
        EMP_HIP(hipGetLastError());                                                                                  \
        return EMP_OK;                                                                                               \
    }
    EMP_SYS_CONFIGS(EMP_CASE)
#undef EMP_CASE
    emp_set_error("unsupported launch configuration");
    return EMP_EINVAL;
}

static int check_tuning(const EmpSystems* p, const EmpSystemsTuning& t) {
    if (!(t.waves_per_wg == 8 || t.waves_per_wg == 16) || !(t.waves_per_theta == 1 || t.waves_per_theta == 2)) {
        emp_set_error("waves_per_wg must be 8 or 16 and waves_per_theta 1 or 2");
        return EMP_EINVAL;
    }
    if (t.wgs_per_cu < 1 || t.wgs_per_cu > 4) { emp_set_error("wgs_per_cu must be in [1, 4]"); return EMP_EINVAL; }
    if (t.order != 0 && t.order != 1) { emp_set_error("order must be 0 (longest first) or 1 (input)"); return EMP_EINVAL; }
    if (swap_bytes(p->desc.ntemps, p->desc.nwalkers) > slot_bytes(t.waves_per_wg, t.waves_per_theta)) {
        emp_set_error("T x W too large for the LDS swap round of this configuration");
        return EMP_EUNSUP;
    }
    return EMP_OK;
}

template <typename T>
static int alloc(EmpSystems* p, T** out, size_t n) {
    void* q = nullptr;
    EMP_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    p->allocations.push_back(q);
    *out = (T*)q;
    return EMP_OK;
}

template <typename T>
static int upload(EmpSystems* p, const std::vector<T>& v, T** out) {
    int rc = alloc(p, out, v.size());
    if (rc) return rc;
    if (!v.empty()) EMP_HIP(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return EMP_OK;
}

extern "C" void emp_systems_destroy(EmpSystems* p) {
    if (!p) return;
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (void* q : p->allocations) (void)hipFree(q);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" int emp_systems_create(const EmpModelDesc* models, const EmpSystemsDesc* desc, const double* betas,
                                  const double* p0, EmpSystems** out) {
    if (!out) { emp_set_error("null output handle"); return EMP_EINVAL; }
    *out = nullptr;
    if (!models || !desc || !betas || !p0) { emp_set_error("null argument"); return EMP_EINVAL; }
    const EmpSystemsDesc& d = *desc;
    const int S = d.n_systems, T = d.ntemps, W = d.nwalkers;
    if (S < 1) { emp_set_error("n_systems must be >= 1"); return EMP_EINVAL; }
    if (T < 1 || T > 64) { emp_set_error("ntemps must be in [1, 64]"); return EMP_EINVAL; }
    if (W < 2 || W % 2 || W > 1024) { emp_set_error("nwalkers must be even and in [2, 1024]"); return EMP_EINVAL; }
    if (T * (W / 2) > SYS_MAX_TASKS) { emp_set_error("ntemps x nwalkers / 2 must be <= %d", SYS_MAX_TASKS); return EMP_EUNSUP; }
    if (d.nsteps < 1) { emp_set_error("nsteps must be >= 1"); return EMP_EINVAL; }
    if (d.precision != EMP_FP64 && d.precision != EMP_MIXED) { emp_set_error("bad precision"); return EMP_EINVAL; }
    if (!(d.a > 1.0)) { emp_set_error("stretch scale a must be > 1"); return EMP_EINVAL; }
    if (d.adapt_mode < -1 || d.adapt_mode > 2) { emp_set_error("adapt_mode must be -1 (frozen), 0, 1 or 2"); return EMP_EUNSUP; }
    if (d.adapt_mode >= 0 && !(d.adapt_tau > 0 && d.adapt_nu > 0)) { emp_set_error("adapt_tau and adapt_nu must be > 0"); return EMP_EINVAL; }
    if (d.record_every < 0 || d.max_records < 0) { emp_set_error("record_every and max_records must be >= 0"); return EMP_EINVAL; }
    if (d.record_every > 0 && (d.record_temps < 1 || d.record_temps > T)) { emp_set_error("record_temps must be in [1, ntemps]"); return EMP_EINVAL; }
    for (int i = 0; i < S * T; ++i)
        if (!(betas[i] >= 0.0 && betas[i] <= 1.0)) { emp_set_error("betas must lie in [0, 1] (entry %d)", i); return EMP_EINVAL; }

    // Host preparation and packing of every model.
    std::vector<ModelHost> hosts(S);
    size_t nep = 0, npar = 0, nfree = 0, nblk = 0, necc = 0, nseq = 0, nsai = 0, nsaib = 0;
    for (int s = 0; s < S; ++s) {
        int rc = emp_model_host(&models[s], &hosts[s]);
        if (rc) {
            char msg[1024];
            snprintf(msg, sizeof msg, "system %d: %s", s, emp_last_error());
            emp_set_error("%s", msg);
            return rc;
        }
        if (hosts[s].dev.D < 1) { emp_set_error("system %d has no free parameter", s); return EMP_EINVAL; }
        nep += hosts[s].t.size();
        npar += hosts[s].inv.size();
        nfree += hosts[s].free_idx.size();
        nblk += hosts[s].bstart.size();
        necc += hosts[s].ecc.size();
        nseq += hosts[s].pseq.size();
        nsai += hosts[s].sai.size();
        nsaib += hosts[s].sai_base.size();
    }
    EmpSystems* p = new EmpSystems();
    p->desc = d;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            emp_systems_destroy(p);
            emp_set_error("cannot query the HIP device");
            return EMP_EHIP;
        }
        p->n_cu = prop.multiProcessorCount;
    }
    std::vector<double> t(nep), y(nep), s2(nep), dt(nep), fixed(npar), sai(nsai);
    std::vector<int> ins(nep), inv(npar), fidx(nfree), bst(nblk), pseq(nseq), saib(nsaib);
    std::vector<EmpPriorTerm> pri(npar), ecc(necc);
    std::vector<ModelDev> mds(S);
    std::vector<size_t> o_ep(S), o_par(S), o_free(S), o_blk(S), o_ecc(S), o_seq(S), o_sai(S), o_saib(S);
    p->D.resize(S);
    p->xoff.resize(S);
    p->rxoff.resize(S);
    {
        size_t e = 0, q = 0, f = 0, b = 0, c = 0, g = 0, a = 0, ab = 0;
        int64_t xo = 0, ro = 0;
        for (int s = 0; s < S; ++s) {
            const ModelHost& h = hosts[s];
            o_ep[s] = e; o_par[s] = q; o_free[s] = f; o_blk[s] = b; o_ecc[s] = c; o_seq[s] = g;
            o_sai[s] = a; o_saib[s] = ab;
            std::copy(h.sai.begin(), h.sai.end(), sai.begin() + a);
            std::copy(h.sai_base.begin(), h.sai_base.end(), saib.begin() + ab);
            a += h.sai.size(); ab += h.sai_base.size();
            std::copy(h.t.begin(), h.t.end(), t.begin() + e);
            std::copy(h.y.begin(), h.y.end(), y.begin() + e);
            std::copy(h.s2.begin(), h.s2.end(), s2.begin() + e);
            std::copy(h.dt.begin(), h.dt.end(), dt.begin() + e);
            std::copy(h.ins.begin(), h.ins.end(), ins.begin() + e);
            std::copy(h.inv.begin(), h.inv.end(), inv.begin() + q);
            std::copy(h.fixed.begin(), h.fixed.end(), fixed.begin() + q);
            std::copy(h.prior.begin(), h.prior.end(), pri.begin() + q);
            std::copy(h.free_idx.begin(), h.free_idx.end(), fidx.begin() + f);
            std::copy(h.bstart.begin(), h.bstart.end(), bst.begin() + b);
            std::copy(h.ecc.begin(), h.ecc.end(), ecc.begin() + c);
            std::copy(h.pseq.begin(), h.pseq.end(), pseq.begin() + g);
            e += h.t.size(); q += h.inv.size(); f += h.free_idx.size(); b += h.bstart.size(); c += h.ecc.size();
            g += h.pseq.size();
            p->D[s] = h.dev.D;
            p->xoff[s] = xo;
            p->rxoff[s] = ro;
            xo += (int64_t)T * W * h.dev.D;
            ro += (int64_t)(d.record_every > 0 ? d.record_temps : 0) * W * h.dev.D;
        }
        p->x_total = xo;
        p->dev.rec_x = ro;
    }
    SysDev& v = p->dev;
    v.S = S; v.T = T; v.W = W; v.nsteps = d.nsteps; v.a = d.a;
    v.key0 = (uint32_t)d.seed; v.key1 = (uint32_t)(d.seed >> 32);
    v.adapt_mode = d.adapt_mode; v.adapt_tau = d.adapt_tau; v.adapt_nu = d.adapt_nu;
    v.rec_every = d.record_every; v.rec_temps = d.record_every > 0 ? d.record_temps : 0;
    v.max_rec = d.record_every > 0 ? d.max_records : 0;
    const size_t STW = (size_t)S * T * W;
    int rc = 0;
    double *dt_, *dy, *ds2, *ddt, *dfix, *dsai;
    int *dins, *dinv, *dfidx, *dbst, *dseq, *dsaib;
    EmpPriorTerm *dpri, *decc;
    double2* dtab;
    std::vector<double2> tab;
    emp_sincos_table(tab);
    if ((rc = upload(p, t, &dt_)) || (rc = upload(p, y, &dy)) || (rc = upload(p, s2, &ds2)) ||
        (rc = upload(p, dt, &ddt)) || (rc = upload(p, ins, &dins)) || (rc = upload(p, inv, &dinv)) ||
        (rc = upload(p, fixed, &dfix)) || (rc = upload(p, pri, &dpri)) || (rc = upload(p, fidx, &dfidx)) ||
        (rc = upload(p, bst, &dbst)) || (rc = upload(p, ecc, &decc)) || (rc = upload(p, tab, &dtab)) ||
        (rc = upload(p, pseq, &dseq)) || (rc = upload(p, sai, &dsai)) || (rc = upload(p, saib, &dsaib))) {
        emp_systems_destroy(p);
        return rc;
    }
    for (int s = 0; s < S; ++s) {
        ModelDev m = hosts[s].dev;
        m.t = dt_ + o_ep[s]; m.y = dy + o_ep[s]; m.s2 = ds2 + o_ep[s]; m.dt = ddt + o_ep[s]; m.ins = dins + o_ep[s];
        m.inv_free = dinv + o_par[s]; m.fixed_val = dfix + o_par[s]; m.prior = dpri + o_par[s];
        m.free_idx = dfidx + o_free[s]; m.block_start = dbst + o_blk[s]; m.ecc_prior = decc + o_ecc[s];
        m.sctab = dtab;
        m.pseq = dseq + o_seq[s];
        m.sai = dsai + o_sai[s];
        m.sai_base = dsaib + o_saib[s];
        if (!emp_model_dev_complete(m)) {
            emp_systems_destroy(p);
            emp_set_error("system %d: a device array of the model is missing", s);
            return EMP_EINVAL;
        }
        p->ext = p->ext || needs_ext(m);
        mds[s] = m;
    }
    // Queue orders: longest first (stable), and the input order.
    std::vector<int> lpt(S), inp(S);
    std::vector<double> cost(S);
    for (int s = 0; s < S; ++s) {
        const ModelDev& m = mds[s];
        cost[s] = (double)m.N * (m.K * 51.0 + 46.0 + (m.ma_active ? 23.0 : 0.0));
        lpt[s] = inp[s] = s;
    }
    std::stable_sort(lpt.begin(), lpt.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int64_t> xoff = p->xoff, rxoff = p->rxoff;
    ModelDev* dmd = nullptr;
    int64_t *dxoff = nullptr, *drxoff = nullptr;
    if ((rc = upload(p, mds, &dmd)) || (rc = upload(p, lpt, &p->d_order_lpt)) || (rc = upload(p, inp, &p->d_order_input)) ||
        (rc = upload(p, xoff, &dxoff)) || (rc = upload(p, rxoff, &drxoff)) ||
        (rc = alloc(p, &v.X, (size_t)p->x_total)) || (rc = alloc(p, &v.lnl, STW)) || (rc = alloc(p, &v.lnp, STW)) ||
        (rc = alloc(p, &v.map, STW)) || (rc = alloc(p, &v.betas, (size_t)S * T)) ||
        (rc = alloc(p, &v.accepted, STW)) || (rc = alloc(p, &v.tsw_sum, (size_t)S * std::max(T - 1, 1))) ||
        (rc = alloc(p, &v.smd_sum, (size_t)S * std::max(T - 1, 1))) || (rc = alloc(p, &v.ctl, 4))) {
        emp_systems_destroy(p);
        return rc;
    }
    v.md = dmd; v.xoff = dxoff; v.rxoff = drxoff; v.order = p->d_order_lpt;
    if (v.max_rec > 0) {
        const size_t nw = (size_t)v.rec_temps * W;
        if ((rc = alloc(p, &v.h_x, (size_t)v.max_rec * v.rec_x)) || (rc = alloc(p, &v.h_lnl, (size_t)v.max_rec * S * nw)) ||
            (rc = alloc(p, &v.h_lnp, (size_t)v.max_rec * S * nw))) {
            emp_systems_destroy(p);
            return rc;
        }
    }
#define EMP_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { int r_ = emp_hip_fail(e_, #x, __FILE__, __LINE__); emp_systems_destroy(p); return r_; } } while (0)
    EMP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    EMP_TRY(hipMemcpy(v.X, p0, (size_t)p->x_total * sizeof(double), hipMemcpyHostToDevice));
    EMP_TRY(hipMemcpy(v.betas, betas, (size_t)S * T * sizeof(double), hipMemcpyHostToDevice));
    EMP_TRY(hipMemset(v.accepted, 0, STW * sizeof(long long)));
    EMP_TRY(hipMemset(v.tsw_sum, 0, (size_t)S * std::max(T - 1, 1) * sizeof(double)));
    EMP_TRY(hipMemset(v.smd_sum, 0, (size_t)S * std::max(T - 1, 1) * sizeof(double)));
    EMP_TRY(hipMemset(v.ctl, 0, 4 * sizeof(unsigned long long)));
    EMP_TRY(hipDeviceSynchronize());     // the engine stream does not wait for the null stream
#undef EMP_TRY
    // Default launch: 16 waves (16 walkers) per workgroup, one workgroup per
    // compute unit, longest system first; 8-wave workgroups when the swap
    // round does not fit.
    p->tune = EmpSystemsTuning{16, 1, 1, 0, -1.0};
    if (check_tuning(p, p->tune)) {
        p->tune = EmpSystemsTuning{8, 1, 2, 0, -1.0};
        if ((rc = check_tuning(p, p->tune))) { emp_systems_destroy(p); return rc; }
    }
    if ((rc = launch_init(p))) { emp_systems_destroy(p); return rc; }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        rc = emp_hip_fail(hipGetLastError(), "initial evaluation", __FILE__, __LINE__);
        emp_systems_destroy(p);
        return rc;
    }
    *out = p;
    return EMP_OK;
}

static int run_sweeps(EmpSystems* p, int64_t nsweeps, bool use_graph, double* ms_out) {
    hipEvent_t e0, e1;
    EMP_HIP(hipEventCreate(&e0));
    EMP_HIP(hipEventCreate(&e1));
    EMP_HIP(hipEventRecord(e0, p->stream));
    if (use_graph) {
        if (!p->graph_exec) {
            hipGraph_t g;
            EMP_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_sweep(p);
            hipError_t ce = hipStreamEndCapture(p->stream, &g);
            if (rc) return rc;
            EMP_HIP(ce);
            EMP_HIP(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
            EMP_HIP(hipGraphDestroy(g));
        }
        for (int64_t s = 0; s < nsweeps; ++s) EMP_HIP(hipGraphLaunch(p->graph_exec, p->stream));
    } else {
        for (int64_t s = 0; s < nsweeps; ++s) {
            int rc = launch_sweep(p);
            if (rc) return rc;
        }
    }
    EMP_HIP(hipEventRecord(e1, p->stream));
    EMP_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    EMP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_out) *ms_out = nsweeps ? ms / (double)nsweeps : 0.0;
    return EMP_OK;
}

extern "C" int emp_systems_run(EmpSystems* p, int64_t nsweeps, int32_t use_graph) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    if (nsweeps < 0) { emp_set_error("nsweeps must be >= 0"); return EMP_EINVAL; }
    if (emp_fault_injected()) return EMP_EFAULT_INJECTED;
    if (nsweeps == 0) return EMP_OK;
    double ms = 0.0;
    int rc = run_sweeps(p, nsweeps, use_graph != 0, &ms);
    if (rc) return rc;
    p->last_ms = ms;
    return EMP_OK;
}

extern "C" int emp_systems_last_timing(EmpSystems* p, double* ms_per_sweep) {
    if (!p || !ms_per_sweep) { emp_set_error("null argument"); return EMP_EINVAL; }
    *ms_per_sweep = p->last_ms;
    return EMP_OK;
}

extern "C" int emp_systems_set_tuning(EmpSystems* p, const EmpSystemsTuning* t) {
    if (!p || !t) { emp_set_error("null argument"); return EMP_EINVAL; }
    int rc = check_tuning(p, *t);
    if (rc) return rc;
    EMP_HIP(hipStreamSynchronize(p->stream));
    p->tune = *t;
    if (p->graph_exec) {
        EMP_HIP(hipGraphExecDestroy(p->graph_exec));
        p->graph_exec = nullptr;
    }
    return EMP_OK;
}

extern "C" int emp_systems_get_tuning(EmpSystems* p, EmpSystemsTuning* t) {
    if (!p || !t) { emp_set_error("null argument"); return EMP_EINVAL; }
    *t = p->tune;
    return EMP_OK;
}

// Device-to-device copies of the mutable state, so that autotuning trials
// leave the chains exactly where they were.
struct SysSnapshot {
    std::vector<std::pair<void*, size_t>> items;   // (device buffer, bytes)
    std::vector<void*> copies;
};

static void snap_free(SysSnapshot& sn) {
    for (void* q : sn.copies) (void)hipFree(q);
    sn.copies.clear();
}

static int snap_take(EmpSystems* p, SysSnapshot& sn) {
    const SysDev& v = p->dev;
    const size_t STW = (size_t)v.S * v.T * v.W, ST1 = (size_t)v.S * std::max(v.T - 1, 1);
    sn.items = {{v.X, (size_t)p->x_total * 8}, {v.lnl, STW * 8}, {v.lnp, STW * 8}, {v.map, STW * 4},
                {v.betas, (size_t)v.S * v.T * 8}, {v.accepted, STW * 8}, {v.tsw_sum, ST1 * 8},
                {v.smd_sum, ST1 * 8}, {v.ctl, 4 * 8}};
    for (auto& it : sn.items) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, it.second);
        if (e != hipSuccess) { snap_free(sn); return emp_hip_fail(e, "hipMalloc (autotune snapshot)", __FILE__, __LINE__); }
        sn.copies.push_back(q);
        EMP_HIP(hipMemcpyAsync(q, it.first, it.second, hipMemcpyDeviceToDevice, p->stream));
    }
    EMP_HIP(hipStreamSynchronize(p->stream));
    return EMP_OK;
}

static int snap_restore(EmpSystems* p, const SysSnapshot& sn) {
    for (size_t i = 0; i < sn.items.size(); ++i)
        EMP_HIP(hipMemcpyAsync(sn.items[i].first, sn.copies[i], sn.items[i].second, hipMemcpyDeviceToDevice, p->stream));
    EMP_HIP(hipStreamSynchronize(p->stream));
    return EMP_OK;
}

extern "C" int emp_systems_autotune(EmpSystems* p, int32_t trial_sweeps, EmpSystemsTuning* best_out) {
    if (!p) { emp_set_error("null engine"); return EMP_EINVAL; }
    if (trial_sweeps < 1) { emp_set_error("trial_sweeps must be >= 1"); return EMP_EINVAL; }
    SysSnapshot sn;
    int rc = snap_take(p, sn);
    if (rc) return rc;
    const EmpSystemsTuning keep = p->tune;
    EmpSystemsTuning best = keep;
    best.ms_per_sweep = -1.0;
    const int nws[2] = {16, 8}, wps[2] = {1, 2};
    for (int order = 0; order < 2 && !rc; ++order)
        for (int a = 0; a < 2 && !rc; ++a)
            for (int b = 0; b < 2 && !rc; ++b)
                for (int wpc = 1; wpc <= 4 && !rc; ++wpc) {
                    EmpSystemsTuning c{nws[a], wps[b], wpc, order, -1.0};
                    if (check_tuning(p, c)) continue;
                    // persistent workgroups beyond what a compute unit holds would only queue
                    int fit = 0;
                    const void* fn = nullptr;
                    const int prec = p->desc.precision;
#define EMP_CASE(P, WP, NW) if (prec == P && c.waves_per_theta == WP && c.waves_per_wg == NW) fn = sweep_fn<P, WP, NW>(p->ext);
                    EMP_SYS_CONFIGS(EMP_CASE)
#undef EMP_CASE
                    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, fn, c.waves_per_wg * 64, 0) != hipSuccess)
                        fit = 1;
                    if (wpc > std::max(fit, 1)) continue;
                    p->tune = c;
                    double ms = 0.0;
                    if ((rc = run_sweeps(p, 1, false, nullptr)) || (rc = run_sweeps(p, trial_sweeps, false, &ms)) ||
                        (rc = snap_restore(p, sn)))
                        break;
                    c.ms_per_sweep = ms;
                    if (best.ms_per_sweep < 0 || ms < best.ms_per_sweep) best = c;
                }
    const int rc2 = snap_restore(p, sn);
    snap_free(sn);
    if (rc) { p->tune = keep; return rc; }
    if (rc2) return rc2;
    if (best.ms_per_sweep < 0) { emp_set_error("no launch configuration fits"); p->tune = keep; return EMP_EUNSUP; }
    p->tune = best;
    if (p->graph_exec) {
        EMP_HIP(hipGraphExecDestroy(p->graph_exec));
        p->graph_exec = nullptr;
    }
    if (best_out) *best_out = best;
    return EMP_OK;
}

// One system's current state in position order.
static int gather_one(EmpSystems* p, int s, int what, double* out) {
    const SysDev& v = p->dev;
    const size_t TW = (size_t)v.T * v.W;
    const int D = p->D[s];
    std::vector<int> mp(TW);
    EMP_HIP(hipMemcpy(mp.data(), v.map + s * TW, TW * 4, hipMemcpyDeviceToHost));
    if (what == EMP_PT_STATE_X) {
        std::vector<double> xs(TW * D);
        EMP_HIP(hipMemcpy(xs.data(), v.X + p->xoff[s], xs.size() * 8, hipMemcpyDeviceToHost));
        for (size_t w = 0; w < TW; ++w) memcpy(out + w * D, &xs[(size_t)mp[w] * D], (size_t)D * 8);
    } else {
        std::vector<double> ls(TW);
        EMP_HIP(hipMemcpy(ls.data(), (what == EMP_PT_STATE_LNL ? v.lnl : v.lnp) + s * TW, TW * 8, hipMemcpyDeviceToHost));
        for (size_t w = 0; w < TW; ++w) out[w] = ls[mp[w]];
    }
    return EMP_OK;
}

extern "C" int emp_systems_get(EmpSystems* p, int32_t what, int32_t system, void* host_out, int64_t cap,
                               int64_t* shape) {
    if (!p || !host_out) { emp_set_error("null argument"); return EMP_EINVAL; }
    const SysDev& v = p->dev;
    const int S = v.S, T = v.T, W = v.W;
    if (system < -1 || system >= S) { emp_set_error("system %d out of range [-1, %d)", system, S); return EMP_EINVAL; }
    EMP_HIP(hipStreamSynchronize(p->stream));
    unsigned long long ctl[4];
    EMP_HIP(hipMemcpy(ctl, v.ctl, sizeof ctl, hipMemcpyDeviceToHost));
    const long long sweeps = (long long)ctl[2];
    const long long R = v.rec_every > 0 ? std::min<long long>(sweeps / v.rec_every, v.max_rec) : 0;
    const int s0 = system < 0 ? 0 : system, s1 = system < 0 ? S : system + 1, ns = s1 - s0;
    const size_t TW = (size_t)T * W, nw = (size_t)v.rec_temps * W;
    int64_t sh[4] = {1, 1, 1, 1};
    size_t n = 0, elem = 8;
    auto need = [&](size_t count) -> bool {
        if ((int64_t)(count * elem) > cap) {
            emp_set_error("output buffer too small (%zu bytes needed)", count * elem);
            return false;
        }
        return true;
    };
    switch (what) {
        case EMP_PT_NSWEEPS: {
            sh[0] = 2; elem = 8; n = 2;
            if (!need(n)) return EMP_EINVAL;
            long long o[2] = {(long long)ctl[2], (long long)ctl[3]};
            memcpy(host_out, o, sizeof o);
            break;
        }
        case EMP_PT_STATE_X: {
            n = 0;
            for (int s = s0; s < s1; ++s) n += TW * p->D[s];
            if (system >= 0) { sh[0] = T; sh[1] = W; sh[2] = p->D[system]; } else { sh[0] = (int64_t)n; }
            if (!need(n)) return EMP_EINVAL;
            double* o = (double*)host_out;
            for (int s = s0; s < s1; ++s) {
                int rc = gather_one(p, s, what, o);
                if (rc) return rc;
                o += TW * p->D[s];
            }
            break;
        }
        case EMP_PT_STATE_LNL:
        case EMP_PT_STATE_LNP: {
            n = (size_t)ns * TW;
            if (system >= 0) { sh[0] = T; sh[1] = W; } else { sh[0] = S; sh[1] = T; sh[2] = W; }
            if (!need(n)) return EMP_EINVAL;
            for (int s = s0; s < s1; ++s) {
                int rc = gather_one(p, s, what, (double*)host_out + (size_t)(s - s0) * TW);
                if (rc) return rc;
            }
            break;
        }
        case EMP_PT_ACCEPTED:
        case EMP_PT_STATE_BETAS: {
            const bool acc = what == EMP_PT_ACCEPTED;
            const size_t per = acc ? TW : (size_t)T;
            n = (size_t)ns * per;
            if (system >= 0) { sh[0] = T; if (acc) sh[1] = W; }
            else { sh[0] = S; sh[1] = T; if (acc) sh[2] = W; }
            if (!need(n)) return EMP_EINVAL;
            const void* src = acc ? (const void*)(v.accepted + s0 * per) : (const void*)(v.betas + s0 * per);
            EMP_HIP(hipMemcpy(host_out, src, n * 8, hipMemcpyDeviceToHost));
            break;
        }
        case EMP_PT_TSW:
        case EMP_PT_SMD: {
            const size_t per = (size_t)std::max(T - 1, 0);
            n = (size_t)ns * per;
            if (system >= 0) { sh[0] = (int64_t)per; } else { sh[0] = S; sh[1] = (int64_t)per; }
            if (!need(n)) return EMP_EINVAL;
            if (n) {
                EMP_HIP(hipMemcpy(host_out, (what == EMP_PT_TSW ? v.tsw_sum : v.smd_sum) + s0 * per, n * 8,
                                  hipMemcpyDeviceToHost));
                double* o = (double*)host_out;
                for (size_t i = 0; i < n; ++i) o[i] = sweeps > 0 ? o[i] / (double)sweeps : 0.0;
            }
            break;
        }
        case EMP_PT_CHAIN: {
            if (system < 0) { emp_set_error("the recorded chain is returned one system at a time"); return EMP_EINVAL; }
            if (!v.max_rec) goto norec;
            const size_t per = nw * p->D[system];
            n = (size_t)R * per;
            sh[0] = R; sh[1] = v.rec_temps; sh[2] = W; sh[3] = p->D[system];
            if (!need(n)) return EMP_EINVAL;
            for (long long r = 0; r < R; ++r)
                EMP_HIP(hipMemcpy((double*)host_out + r * per, v.h_x + r * v.rec_x + p->rxoff[system], per * 8,
                                  hipMemcpyDeviceToHost));
            break;
        }
        case EMP_PT_LNL:
        case EMP_PT_LNP: {
            if (!v.max_rec) goto norec;
            const double* src = what == EMP_PT_LNL ? v.h_lnl : v.h_lnp;
            n = (size_t)R * ns * nw;
            if (system >= 0) { sh[0] = R; sh[1] = v.rec_temps; sh[2] = W; }
            else { sh[0] = R; sh[1] = S; sh[2] = v.rec_temps; sh[3] = W; }
            if (!need(n)) return EMP_EINVAL;
            for (long long r = 0; r < R; ++r)
                EMP_HIP(hipMemcpy((double*)host_out + (size_t)r * ns * nw, src + ((size_t)r * S + s0) * nw,
                                  (size_t)ns * nw * 8, hipMemcpyDeviceToHost));
            break;
        }
        default:
            emp_set_error("unknown quantity %d", what);
            return EMP_EINVAL;
    }
    if (shape) for (int i = 0; i < 4; ++i) shape[i] = sh[i];
    return EMP_OK;
norec:
    emp_set_error("nothing was recorded (record_every == 0 or max_records == 0)");
    return EMP_EINVAL;
}
