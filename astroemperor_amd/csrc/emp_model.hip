// Model handle, batched likelihood / prior / model-eval kernels and their
// C-ABI entry points (include/emp_abi.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "emp_kernels.h"

using namespace emp;

// ---------------------------------------------------------------------------
// Error plumbing.
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

void emp_set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int emp_hip_fail(hipError_t e, const char* what, const char* file, int line) {
    emp_set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    return e == hipErrorOutOfMemory ? EMP_ENOMEM : EMP_EHIP;
}

// EMP_FAULT_AT=k makes the k-th compute call of the process fail with
// EMP_EFAULT_INJECTED, so host code can test its error handling without
// touching the device.
bool emp_fault_injected() {
    static std::atomic<long> counter{0};
    static long target = [] {
        const char* s = getenv("EMP_FAULT_AT");
        return s ? atol(s) : 0L;
    }();
    if (target <= 0) return false;
    long n = ++counter;
    if (n == target) {
        emp_set_error("fault injected at call %ld (EMP_FAULT_AT)", n);
        return true;
    }
    return false;
}

extern "C" const char* emp_last_error(void) { return g_last_error.c_str(); }
extern "C" int emp_abi_version(void) { return EMP_ABI_VERSION; }

extern "C" int emp_philox4x32_10(const uint32_t* c, const uint32_t* k, uint32_t* o) {
    if (!c || !k || !o) { emp_set_error("null argument"); return EMP_EINVAL; }
    U4 r = philox4x32_10(U4{c[0], c[1], c[2], c[3]}, k[0], k[1]);
    o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w;
    return EMP_OK;
}

extern "C" int emp_device_info(int32_t* n_cu, int32_t* clock_khz, char* arch_name, int32_t arch_cap) {
    int dev = 0;
    EMP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    EMP_HIP(hipGetDeviceProperties(&p, dev));
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    if (arch_name && arch_cap > 0) {
        strncpy(arch_name, p.gcnArchName, (size_t)arch_cap - 1);
        arch_name[arch_cap - 1] = 0;
    }
    return EMP_OK;
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
template <int PREC, int WPT, bool POST>
__global__ __launch_bounds__(BLOCK) void k_lnl(ModelDev md, const double* __restrict__ theta, int64_t B,
                                               int64_t ld, double* __restrict__ out,
                                               unsigned long long* nonfinite) {
    constexpr int SLOTS = NWAVES / WPT;
    __shared__ SlotLDS<SLOTS> s_slot[SLOTS];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave / WPT, wsub = wave % WPT;
    const int64_t b = (int64_t)blockIdx.x * SLOTS + slot;
    const bool live = b < B;
    SlotLDS<SLOTS>& sl = s_slot[slot];
    if (live) expand_theta(md, sl, theta + b * ld, wsub * 64 + lane, 64 * WPT);
    tab_load(md, s_tile);
    __syncthreads();
    slot_prepare<PREC, WPT>(md, sl, s_tile, live, wsub, lane);
    double v = block_lnl<PREC, WPT, false, false, true, POST>(md, sl, s_tile, s_scan, live, wsub, nullptr, nullptr);
    if (live && wsub == 0 && lane == 0) {
        if (!(v == v) || v == __builtin_inf()) {
            v = -__builtin_inf();
            atomicAdd(nonfinite, 1ull);
        }
        out[b] = v;
    }
}

__global__ void k_prior(ModelDev md, const double* __restrict__ theta, int64_t B, int64_t ld,
                        double* __restrict__ out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double th[MAXD];
    const double* row = theta + b * ld;
    for (int j = 0; j < md.Dfull; ++j) {
        const int f = md.inv_free[j];
        th[j] = f >= 0 ? row[f] : md.fixed_val[j];
    }
    out[b] = prior_seq(md, th);
}

__global__ __launch_bounds__(BLOCK) void k_model_eval(ModelDev md, const double* __restrict__ theta,
                                                      double* out_model, double* out_err2) {
    __shared__ SlotLDS<1> s_slot[1];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    SlotLDS<1>& sl = s_slot[0];
    expand_theta(md, sl, theta, threadIdx.x, BLOCK);
    tab_load(md, s_tile);
    __syncthreads();
    slot_prepare<0, NWAVES>(md, sl, s_tile, true, wave, lane);
    block_lnl<0, NWAVES, true>(md, sl, s_tile, s_scan, true, wave, out_model, out_err2);
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
// Models without the extended features (terms added after the MA, MA orders
// above 1) run the variant that compiles them out (fewer registers).
template <int PREC, int WPT>
static void launch_lnl_t(const ModelDev& md, const double* th, int64_t B, int64_t ld, double* out,
                         unsigned long long* nf, hipStream_t s) {
    constexpr int SLOTS = NWAVES / WPT;
    const int64_t grid = (B + SLOTS - 1) / SLOTS;
    if (needs_ext(md))
        hipLaunchKernelGGL((k_lnl<PREC, WPT, true>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, th, B, ld, out, nf);
    else
        hipLaunchKernelGGL((k_lnl<PREC, WPT, false>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, th, B, ld, out,
                           nf);
}

int emp_pick_wpt(const EmpModel* m, int64_t B, int precision, bool halfstep) {
    if (m->wpt_override) return m->wpt_override;
    // Enough waves to give every SIMD (4 per CU) about 32 on the mixed path
    // (from 4 waves per vector its Kepler tables are kept per interval) and
    // 16 on the FP64 path, at least two waves per vector for series of 1024
    // epochs or more, but never more lanes per vector than a third of the
    // epochs for batched likelihoods, half for the PT half-step (whose few
    // proposals make one walker's critical path the launch's length).
    // Measured batched lnL at 1 / 2 / 4 / 8 waves per vector, mixed and FP64:
    // 7 planets x 10^4 epochs at 8192 vectors 1.48 / 1.31 / 1.28 / 1.37 ms
    // and 3.07 / 3.01 / 3.06 / 3.15 ms; 5 planets x 2000 epochs at 65,536
    // vectors 1.40 / 1.26 / 1.34 / 1.53 ms and 3.42 / 3.40 / 3.48 / 3.66 ms;
    // 3 planets x 1530 epochs at 2560 vectors 72 / 56 / 52 / 59 us and
    // 121 / 92 / 81 / 90 us.  The rule picks 4, 2, 2, 2, 4 and 4.
    const int per_simd = precision == EMP_MIXED ? 32 : 16;
    const int64_t target_waves = (int64_t)m->n_cu * 4 * per_simd;
    const int wmin = m->dev.N >= 1024 ? 2 : 1;
    const int max_lanes = halfstep ? m->dev.N / 2 : m->dev.N / 3;
    int wpt = 1;
    while (wpt < 8 && (B * wpt < target_waves || wpt < wmin) && 64 * (wpt * 2) <= max_lanes) wpt *= 2;
    return wpt;
}

int emp_launch_lnl(const EmpModel* m, const double* d_theta, int64_t B, int64_t ld, double* d_lnl,
                   int precision, int wpt, hipStream_t s) {
    if (B <= 0) return EMP_OK;
    const ModelDev& md = m->dev;
    if (wpt <= 0) wpt = emp_pick_wpt(m, B, precision);
#define EMP_CASE(P, W)                                                                 \
    if (precision == P && wpt == W) {                                                  \
        launch_lnl_t<P, W>(md, d_theta, B, ld, d_lnl, m->d_nonfinite, s);              \
        EMP_HIP(hipGetLastError());                                                    \
        return EMP_OK;                                                                 \
    }
    EMP_CASE(0, 1) EMP_CASE(0, 2) EMP_CASE(0, 4) EMP_CASE(0, 8)
    EMP_CASE(1, 1) EMP_CASE(1, 2) EMP_CASE(1, 4) EMP_CASE(1, 8)
#undef EMP_CASE
    emp_set_error("unsupported precision %d / waves per theta %d", precision, wpt);
    return EMP_EINVAL;
}

int emp_launch_prior(const EmpModel* m, const double* d_theta, int64_t B, int64_t ld, double* d_lnp,
                     hipStream_t s) {
    if (B <= 0) return EMP_OK;
    const int64_t grid = (B + 255) / 256;
    hipLaunchKernelGGL(k_prior, dim3((unsigned)grid), dim3(256), 0, s, m->dev, d_theta, B, ld, d_lnp);
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

// ---------------------------------------------------------------------------
// Model creation.
// ---------------------------------------------------------------------------
template <typename T>
static int upload(EmpModel* m, const T* host, size_t n, const T** out) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    EMP_HIP(hipMalloc(&p, bytes));
    m->allocations.push_back(p);
    if (n) EMP_HIP(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T*)p;
    return EMP_OK;
}

static int validate(const EmpModelDesc* d) {
    if (!d) { emp_set_error("null descriptor"); return EMP_EINVAL; }
    if (d->abi_version != EMP_ABI_VERSION) {
        emp_set_error("abi_version %d, library speaks %d", d->abi_version, EMP_ABI_VERSION);
        return EMP_EINVAL;
    }
    if (d->n_epochs < 1) { emp_set_error("n_epochs must be >= 1"); return EMP_EINVAL; }
    if (d->n_ins < 1 || d->n_ins > IMAX) { emp_set_error("n_ins must be in [1, %d]", IMAX); return EMP_EUNSUP; }
    if (d->n_planets < 0 || d->n_planets > KMAX) { emp_set_error("n_planets must be in [0, %d]", KMAX); return EMP_EUNSUP; }
    const int p = d->parameterisation;
    if (d->n_planets > 0 && !(p == 0 || p == 1 || p == 2 || p == 3 || p == 4 || p == 6 || p == 7)) {
        emp_set_error("parameterisation %d has no RV template", p);
        return EMP_EUNSUP;
    }
    if (d->acc_order < 0 || d->acc_order > AMAX) { emp_set_error("acc_order must be in [0, %d]", AMAX); return EMP_EUNSUP; }
    if (d->ndim_full < 1 || d->ndim_full > MAXD) { emp_set_error("ndim_full must be in [1, %d]", MAXD); return EMP_EUNSUP; }
    if (d->ndim < 0 || d->ndim > d->ndim_full) { emp_set_error("bad ndim"); return EMP_EINVAL; }
    if (!d->t || !d->y || !d->yerr || !d->ins || !d->prior || (d->ndim && !d->free_index) || !d->fixed_value) {
        emp_set_error("null array in descriptor");
        return EMP_EINVAL;
    }
    if (d->n_planets && d->off_kep + 5 * d->n_planets > d->ndim_full) { emp_set_error("Keplerian block out of range"); return EMP_EINVAL; }
    if (d->off_offset < 0 || d->off_offset + d->n_ins > d->ndim_full) { emp_set_error("offset block out of range"); return EMP_EINVAL; }
    if (d->has_jitter && (d->off_jitter < 0 || d->off_jitter + d->n_ins > d->ndim_full)) { emp_set_error("jitter block out of range"); return EMP_EINVAL; }
    if (d->acc_order && (d->off_acc < 0 || d->off_acc + d->acc_order > d->ndim_full)) { emp_set_error("acceleration block out of range"); return EMP_EINVAL; }
    if (d->ma_order > 0) {
        const int sets = d->ma_global ? 1 : d->n_ins;
        if (d->off_ma < 0 || d->off_ma + 2 * d->ma_order * sets > d->ndim_full) { emp_set_error("MA block out of range"); return EMP_EINVAL; }
    }
    if (d->off_sinusoid < -1 || (d->off_sinusoid >= 0 && d->off_sinusoid + 3 > d->ndim_full)) { emp_set_error("sinusoid block out of range"); return EMP_EINVAL; }
    if (d->off_magnetic < -1 || (d->off_magnetic >= 0 && d->off_magnetic + 5 > d->ndim_full)) { emp_set_error("magnetic-cycle block out of range"); return EMP_EINVAL; }
    if (d->n_sai < 0 || d->n_sai > SAIMAX) { emp_set_error("n_sai must be in [0, %d]", SAIMAX); return EMP_EUNSUP; }
    if ((d->n_sai > 0) != (d->off_sai >= 0) || d->off_sai < -1 || (d->off_sai >= 0 && d->off_sai + d->n_sai > d->ndim_full)) {
        emp_set_error("stellar-activity block out of range");
        return EMP_EINVAL;
    }
    if (d->n_sai > 0) {
        if (!d->sai || !d->sai_counts) { emp_set_error("null activity arrays"); return EMP_EINVAL; }
        int tot = 0;
        for (int k = 0; k < d->n_ins; ++k) {
            if (d->sai_counts[k] < 0) { emp_set_error("negative activity column count"); return EMP_EINVAL; }
            tot += d->sai_counts[k];
        }
        if (tot != d->n_sai) { emp_set_error("sai_counts sum to %d, n_sai is %d", tot, d->n_sai); return EMP_EINVAL; }
    }
    for (int i = 0; i < d->n_epochs; ++i) {
        if (d->ins[i] < 0 || d->ins[i] >= d->n_ins) { emp_set_error("instrument tag out of range at epoch %d", i); return EMP_EINVAL; }
    }
    for (int i = 1; i < d->n_epochs; ++i) {
        if (d->t[i] < d->t[i - 1]) { emp_set_error("epochs must be time-sorted (epoch %d)", i); return EMP_EINVAL; }
    }
    for (int k = 0; k < d->ndim; ++k) {
        if (d->free_index[k] < 0 || d->free_index[k] >= d->ndim_full) { emp_set_error("free_index out of range"); return EMP_EINVAL; }
    }
    if (d->n_blocks < 1 || d->n_blocks > MAXD) { emp_set_error("bad n_blocks"); return EMP_EINVAL; }
    for (int j = 0; j < d->ndim_full; ++j) {
        if (d->prior[j].block < 0 || d->prior[j].block >= d->n_blocks || (j && d->prior[j].block < d->prior[j - 1].block)) {
            emp_set_error("prior block ids must be non-decreasing in [0, n_blocks)");
            return EMP_EINVAL;
        }
    }
    return EMP_OK;
}

int emp_model_host(const EmpModelDesc* d, ModelHost* h) {
    int rc = validate(d);
    if (rc) return rc;
    const bool ma_active = d->ma_order > 0 && (d->ma_global || d->semantics == EMP_SEM_INTENDED);
    if (ma_active && d->ma_order > 4) {
        emp_set_error("moving-average order %d: the kernels implement orders 1 to 4", d->ma_order);
        return EMP_EUNSUP;
    }
    const int N = d->n_epochs;
    // Device order: per-instrument MA needs each instrument's time series
    // contiguous; otherwise keep the caller's (time) order.
    const bool by_ins = ma_active && !d->ma_global;
    h->order.resize(N);
    for (int i = 0; i < N; ++i) h->order[i] = i;
    if (by_ins) std::stable_sort(h->order.begin(), h->order.end(), [&](int a, int b) { return d->ins[a] < d->ins[b]; });
    h->t.resize(N); h->y.resize(N); h->s2.resize(N); h->dt.resize(N); h->ins.resize(N);
    for (int i = 0; i < N; ++i) {
        const int o = h->order[i];
        h->t[i] = d->t[o];
        h->y[i] = d->y[o];
        h->s2[i] = d->yerr[o] * d->yerr[o];
        h->ins[i] = d->ins[o];
    }
    for (int i = 0; i < N; ++i) {
        const bool start = i == 0 || (by_ins && h->ins[i] != h->ins[i - 1]);
        h->dt[i] = start ? -1.0 : std::fabs(h->t[i] - h->t[i - 1]);
    }
    h->inv.assign(d->ndim_full, -1);
    for (int k = 0; k < d->ndim; ++k) h->inv[d->free_index[k]] = k;
    h->free_idx.assign(d->free_index, d->free_index + d->ndim);
    h->fixed.assign(d->fixed_value, d->fixed_value + d->ndim_full);
    h->prior.assign(d->prior, d->prior + d->ndim_full);
    h->bstart.assign(d->n_blocks + 1, d->ndim_full);
    for (int b = d->n_blocks - 1; b >= 0; --b) {
        int s = h->bstart[b + 1];
        for (int j = 0; j < d->ndim_full; ++j) if (d->prior[j].block == b) { s = j; break; }
        h->bstart[b] = s;
    }
    h->bstart[0] = 0;
    h->ecc.resize(std::max(d->n_planets, 1));
    for (int k = 0; k < (int)h->ecc.size(); ++k) {
        if (d->ecc_prior && k < d->n_planets) h->ecc[k] = d->ecc_prior[k];
        else { h->ecc[k] = EmpPriorTerm{}; h->ecc[k].kind = EMP_PRIOR_NONE; }
    }
    // The prior's summation order (prior_seq): each block's free parameters,
    // the -inf check, then the block's derived eccentricity prior.
    h->pseq.clear();
    for (int b = 0; b < d->n_blocks; ++b) {
        for (int j = h->bstart[b]; j < h->bstart[b + 1]; ++j)
            if (h->inv[j] >= 0) h->pseq.push_back(j);
        h->pseq.push_back(PSEQ_CHECK);
        if (b < d->n_planets && h->ecc[b].kind != EMP_PRIOR_NONE) h->pseq.push_back(PSEQ_ECC - b);
    }
    h->pseq_e.assign(h->pseq.size(), PseqEntry{});
    for (size_t i = 0; i < h->pseq.size(); ++i) {
        const int q = h->pseq[i];
        PseqEntry& e = h->pseq_e[i];
        e.fa = e.fb = -1;
        if (q >= 0) {
            e.kind = 0;
            e.term = h->prior[q];
            e.fa = h->inv[q];
        } else if (q <= PSEQ_ECC) {
            const int b = PSEQ_ECC - q, o = d->off_kep + 5 * b + 3;
            e.kind = 1;
            e.term = h->ecc[b];
            e.fa = h->inv[o];
            e.fb = h->inv[o + 1];
            e.va = d->fixed_value[o];
            e.vb = d->fixed_value[o + 1];
        } else {
            e.kind = 2;
        }
    }
    ModelDev& md = h->dev;
    md = ModelDev{};
    md.N = N; md.I = d->n_ins; md.K = d->n_planets; md.param = d->parameterisation; md.use_c = d->use_c;
    md.acc = d->acc_order; md.has_jit = d->has_jitter;
    md.ma_q = d->ma_order; md.ma_global = d->ma_global; md.ma_active = ma_active ? 1 : 0;
    md.Dfull = d->ndim_full; md.D = d->ndim; md.nblocks = d->n_blocks;
    md.off_kep = d->off_kep; md.off_acc = d->off_acc; md.off_off = d->off_offset;
    md.off_jit = d->off_jitter; md.off_ma = d->off_ma;
    md.off_sin = d->off_sinusoid; md.off_mag = d->off_magnetic;
    md.n_cos = (d->off_sinusoid >= 0 ? 1 : 0) + (d->off_magnetic >= 0 ? 2 : 0);
    // Activity columns: each epoch keeps only its own instrument's columns
    // (the others are zero, which the caller must honour), column-major so
    // that lanes over consecutive epochs read consecutive addresses.
    md.off_sai = d->n_sai > 0 ? d->off_sai : -1;
    md.sai_w = 0;
    h->sai_base.assign(d->n_ins + 1, 0);
    if (d->n_sai > 0) {
        for (int k = 0; k < d->n_ins; ++k) {
            h->sai_base[k + 1] = h->sai_base[k] + d->sai_counts[k];
            md.sai_w = std::max(md.sai_w, (int)d->sai_counts[k]);
        }
        h->sai.assign((size_t)md.sai_w * N, 0.0);
        for (int i = 0; i < N; ++i) {
            const int o = h->order[i], k = h->ins[i];
            const double* row = d->sai + (size_t)o * d->n_sai;
            for (int c = 0; c < d->n_sai; ++c) {
                const int own = c - h->sai_base[k];
                if (own >= 0 && own < d->sai_counts[k]) h->sai[(size_t)own * N + i] = row[c];
                else if (row[c] != 0.0) {
                    emp_set_error("activity column %d is non-zero at epoch %d of instrument %d", c, o, k + 1);
                    return EMP_EINVAL;
                }
            }
        }
    }
    md.lnl_const = -0.5 * std::log(2.0 * M_PI) * (double)N;
    md.npseq = (int)h->pseq.size();
    md.t0 = d->t[0];
    // Fixed MA timescales (the usual setup): the weight exp(-dt / tau) of
    // every epoch is the same for all parameter vectors; compute it once.
    md.ma_wfixed = 0;
    if (ma_active && d->ma_order == 1) {
        bool fixed = true;
        for (int k = 0; k < d->n_ins && fixed; ++k) {
            const int j = d->off_ma + 2 * d->ma_order * (d->ma_global ? 0 : k) + 1;
            fixed = h->inv[j] < 0;
        }
        if (fixed) {
            md.ma_wfixed = 1;
            for (int i = 0; i < N; ++i) {
                const int j = d->off_ma + 2 * d->ma_order * (d->ma_global ? 0 : h->ins[i]) + 1;
                h->dt[i] = h->dt[i] < 0.0 ? 0.0 : std::exp(-h->dt[i] / d->fixed_value[j]);
            }
        }
    }
    return EMP_OK;
}

void emp_sincos_table(std::vector<double2>& tab) {
    tab.resize(SC_N + 1);
    const long double pi_l = 3.141592653589793238462643383279502884L;
    for (int k = 0; k <= SC_N; ++k) {
        const long double x = pi_l * (long double)k / (long double)SC_N;
        tab[k] = make_double2((double)sinl(x), (double)cosl(x));
    }
}

extern "C" int emp_model_create(const EmpModelDesc* d, EmpModel** out) {
    if (!out) { emp_set_error("null output handle"); return EMP_EINVAL; }
    *out = nullptr;
    ModelHost h;
    int rc = emp_model_host(d, &h);
    if (rc) return rc;
    EmpModel* m = new EmpModel();
    m->desc = *d;
    EMP_HIP(hipGetDevice(&m->device));
    {
        hipDeviceProp_t p;
        EMP_HIP(hipGetDeviceProperties(&p, m->device));
        m->n_cu = p.multiProcessorCount;
    }
    m->order = h.order;
    ModelDev& md = m->dev;
    md = h.dev;
#define EMP_UP(vec, field)                                       \
    do {                                                         \
        int r_ = upload(m, vec.data(), vec.size(), &field);      \
        if (r_) { emp_model_destroy(m); return r_; }             \
    } while (0)
    EMP_UP(h.t, md.t);
    EMP_UP(h.y, md.y);
    EMP_UP(h.s2, md.s2);
    EMP_UP(h.dt, md.dt);
    EMP_UP(h.ins, md.ins);
    EMP_UP(h.inv, md.inv_free);
    EMP_UP(h.free_idx, md.free_idx);
    EMP_UP(h.fixed, md.fixed_val);
    EMP_UP(h.prior, md.prior);
    EMP_UP(h.bstart, md.block_start);
    EMP_UP(h.ecc, md.ecc_prior);
    EMP_UP(h.pseq, md.pseq);
    EMP_UP(h.pseq_e, md.pseq_e);
    EMP_UP(h.sai, md.sai);
    EMP_UP(h.sai_base, md.sai_base);
    {
        std::vector<double2> tab;
        emp_sincos_table(tab);
        EMP_UP(tab, md.sctab);
    }
#undef EMP_UP
    if (!emp_model_dev_complete(md) || !md.pseq_e) {
        emp_model_destroy(m);
        emp_set_error("a device array of the model is missing");
        return EMP_EINVAL;
    }
    void* nf = nullptr;
    if (hipMalloc(&nf, sizeof(unsigned long long)) != hipSuccess) {
        emp_model_destroy(m);
        emp_set_error("hipMalloc failed for the non-finite counter");
        return EMP_ENOMEM;
    }
    m->allocations.push_back(nf);
    m->d_nonfinite = (unsigned long long*)nf;
    EMP_HIP(hipMemset(nf, 0, sizeof(unsigned long long)));
    m->desc.t = m->desc.y = m->desc.yerr = nullptr;
    m->desc.ins = nullptr; m->desc.free_index = nullptr; m->desc.fixed_value = nullptr;
    m->desc.prior = nullptr; m->desc.ecc_prior = nullptr;
    m->desc.sai = nullptr; m->desc.sai_counts = nullptr;
    *out = m;
    return EMP_OK;
}

extern "C" void emp_model_destroy(EmpModel* m) {
    if (!m) return;
    for (void* p : m->allocations) (void)hipFree(p);
    if (m->d_stage_in) (void)hipFree(m->d_stage_in);
    if (m->d_stage_out) (void)hipFree(m->d_stage_out);
    if (m->d_grad_scratch) (void)hipFree(m->d_grad_scratch);
    delete m;
}

static int ensure_stage(double** p, size_t* cap, size_t n) {
    if (*cap >= n) return EMP_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    EMP_HIP(hipMalloc((void**)p, n * sizeof(double)));
    *cap = n;
    return EMP_OK;
}

// Runs `fn(d_theta, d_out...)` either on caller device pointers or through the
// staging buffers for host pointers.
static int with_stage(EmpModel* m, const double* theta, int64_t B, int64_t ld, double* out, int64_t out_per,
                      double* out2, int64_t out2_per, int flags, hipStream_t s,
                      int (*fn)(EmpModel*, const double*, double*, double*, void*), void* ctx) {
    if (!(flags & EMP_HOST_POINTERS)) return fn(m, theta, out, out2, ctx);
    const size_t nin = (size_t)B * ld, nout = (size_t)B * (out_per + out2_per);
    int rc = ensure_stage(&m->d_stage_in, &m->stage_in_cap, std::max<size_t>(nin, 1));
    if (rc) return rc;
    rc = ensure_stage(&m->d_stage_out, &m->stage_out_cap, std::max<size_t>(nout, 1));
    if (rc) return rc;
    EMP_HIP(hipMemcpyAsync(m->d_stage_in, theta, nin * sizeof(double), hipMemcpyHostToDevice, s));
    double* d_out2 = out2 ? m->d_stage_out + (size_t)B * out_per : nullptr;
    rc = fn(m, m->d_stage_in, m->d_stage_out, d_out2, ctx);
    if (rc) return rc;
    EMP_HIP(hipMemcpyAsync(out, m->d_stage_out, (size_t)B * out_per * sizeof(double), hipMemcpyDeviceToHost, s));
    if (out2) EMP_HIP(hipMemcpyAsync(out2, d_out2, (size_t)B * out2_per * sizeof(double), hipMemcpyDeviceToHost, s));
    EMP_HIP(hipStreamSynchronize(s));
    return EMP_OK;
}

struct LnlCtx { int64_t B, ld; int precision; hipStream_t s; };

static int run_lnl(EmpModel* m, const double* th, double* out, double*, void* c) {
    LnlCtx* x = (LnlCtx*)c;
    return emp_launch_lnl(m, th, x->B, x->ld, out, x->precision, 0, x->s);
}

static int run_prior(EmpModel* m, const double* th, double* out, double*, void* c) {
    LnlCtx* x = (LnlCtx*)c;
    return emp_launch_prior(m, th, x->B, x->ld, out, x->s);
}

static int check_batch(EmpModel* m, const double* theta, int64_t B, int64_t ld, const void* out) {
    if (!m) { emp_set_error("null model"); return EMP_EINVAL; }
    if (B < 0 || (B > 0 && (!theta || !out))) { emp_set_error("bad batch arguments"); return EMP_EINVAL; }
    if (ld < m->desc.ndim) { emp_set_error("ld (%lld) < ndim (%d)", (long long)ld, m->desc.ndim); return EMP_EINVAL; }
    if (emp_fault_injected()) return EMP_EFAULT_INJECTED;
    return EMP_OK;
}

extern "C" int emp_lnl_batch(EmpModel* m, const double* theta, int64_t B, int64_t ld, double* lnl,
                             int32_t precision, int32_t flags, void* stream) {
    int rc = check_batch(m, theta, B, ld, lnl);
    if (rc) return rc;
    if (precision != EMP_FP64 && precision != EMP_MIXED) { emp_set_error("bad precision %d", precision); return EMP_EINVAL; }
    if (B == 0) return EMP_OK;
    LnlCtx c{B, ld, precision, (hipStream_t)stream};
    return with_stage(m, theta, B, ld, lnl, 1, nullptr, 0, flags, c.s, run_lnl, &c);
}

extern "C" int emp_lnprior_batch(EmpModel* m, const double* theta, int64_t B, int64_t ld, double* lnp,
                                 int32_t flags, void* stream) {
    int rc = check_batch(m, theta, B, ld, lnp);
    if (rc) return rc;
    if (B == 0) return EMP_OK;
    LnlCtx c{B, ld, 0, (hipStream_t)stream};
    return with_stage(m, theta, B, ld, lnp, 1, nullptr, 0, flags, c.s, run_prior, &c);
}

extern "C" int emp_model_eval(EmpModel* m, const double* theta, double* model_out, double* err2_out) {
    int rc = check_batch(m, theta, 1, m ? m->desc.ndim : 0, model_out);
    if (rc) return rc;
    if (!err2_out) { emp_set_error("null err2 output"); return EMP_EINVAL; }
    const int N = m->dev.N;
    const int D = std::max(m->desc.ndim, 1);
    rc = ensure_stage(&m->d_stage_in, &m->stage_in_cap, (size_t)D);
    if (rc) return rc;
    rc = ensure_stage(&m->d_stage_out, &m->stage_out_cap, (size_t)2 * N);
    if (rc) return rc;
    if (m->desc.ndim) EMP_HIP(hipMemcpy(m->d_stage_in, theta, (size_t)m->desc.ndim * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_model_eval, dim3(1), dim3(BLOCK), 0, 0, m->dev, m->d_stage_in, m->d_stage_out,
                       m->d_stage_out + N);
    EMP_HIP(hipGetLastError());
    std::vector<double> buf((size_t)2 * N);
    EMP_HIP(hipMemcpy(buf.data(), m->d_stage_out, (size_t)2 * N * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < N; ++i) {
        model_out[m->order[i]] = buf[i];
        err2_out[m->order[i]] = buf[N + i];
    }
    return EMP_OK;
}

extern "C" int emp_model_set_waves_per_theta(EmpModel* m, int32_t w) {
    if (!m) { emp_set_error("null model"); return EMP_EINVAL; }
    if (!(w == 0 || w == 1 || w == 2 || w == 4 || w == 8)) { emp_set_error("waves per theta must be 0, 1, 2, 4 or 8"); return EMP_EINVAL; }
    m->wpt_override = w;
    return EMP_OK;
}

extern "C" int emp_model_get_waves_per_theta(EmpModel* m, int32_t precision, int32_t* out) {
    if (!m || !out) { emp_set_error("null argument"); return EMP_EINVAL; }
    *out = emp_pick_wpt(m, 1024, precision);
    return EMP_OK;
}

extern "C" int emp_model_nonfinite_count(EmpModel* m, int64_t* out) {
    if (!m || !out) { emp_set_error("null argument"); return EMP_EINVAL; }
    unsigned long long v = 0;
    EMP_HIP(hipMemcpy(&v, m->d_nonfinite, sizeof(v), hipMemcpyDeviceToHost));
    *out = (int64_t)v;
    return EMP_OK;
}
