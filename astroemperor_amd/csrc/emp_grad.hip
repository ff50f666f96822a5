// Adjoint kernel: lnL and dlnL/dtheta for a batch of parameter vectors (FP64).
//
// The Kepler solve is differentiated implicitly: from E - e sin E = M,
//   dE/dM = 1/(1 - e cos E),  dE/de|_M = sin E/(1 - e cos E),
// so no derivative flows through the solver's iterations.  With
// D = 1 - e cos E, b = sqrt(1 - e^2), cos f = (cos E - e)/D, sin f = b sin E/D,
// the planet term mu = A (cos(f + w) + e cos w) has
//   dmu/dM = -A sin(f + w) b / D^2
//   dmu/de = -A sin(f + w) (b sin E / D^2 + sin E/(b D)) + A cos w
//   dmu/dw = -A (sin(f + w) + e sin w),   dmu/dA = cos(f + w) + e cos w,
// and dlnL/dmu_i = r_i / err2_i,  dlnL/dJ_k = J_k sum_{i in k} (r_i^2/err2_i^2 - 1/err2_i).
// The natural-parameter gradients are then mapped to the block's
// parameterisation (block_repo.py:23-31) once per vector.
#include <cmath>

#include "emp_kernels.h"

using namespace emp;

struct GradLDS {
    double g[MAXD];            // gradient w.r.t. the full vector
    double nat[KMAX][5];       // per planet: (M*tau, M, e, w, A) accumulators
};

template <int WPT>
__global__ __launch_bounds__(BLOCK) void k_lnl_grad(ModelDev md, const double* __restrict__ theta, int64_t B,
                                                    int64_t ld, double* __restrict__ out_lnl,
                                                    double* __restrict__ out_grad,
                                                    unsigned long long* nonfinite) {
    constexpr int SLOTS = NWAVES / WPT;
    __shared__ SlotLDS s_slot[SLOTS];
    __shared__ GradLDS s_grad[SLOTS];
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave / WPT, wsub = wave % WPT;
    const int g = wsub * 64 + lane, G = 64 * WPT;
    const int64_t b = (int64_t)blockIdx.x * SLOTS + slot;
    const bool live = b < B;
    SlotLDS& sl = s_slot[slot];
    GradLDS& gl = s_grad[slot];
    if (live) expand_theta(md, sl, theta + b * ld, g, G);
    for (int j = g; j < MAXD; j += G) gl.g[j] = 0.0;
    for (int j = g; j < KMAX * 5; j += G) (&gl.nat[0][0])[j] = 0.0;
    tab_load(md, s_tile);
    __syncthreads();
    slot_prepare<0, WPT>(md, sl, s_tile, live, wsub, lane);

    double acc_nat[KMAX][5];
#pragma unroll
    for (int p = 0; p < KMAX; ++p)
#pragma unroll
        for (int q = 0; q < 5; ++q) acc_nat[p][q] = 0.0;
    double acc_acc[AMAX];
#pragma unroll
    for (int a = 0; a < AMAX; ++a) acc_acc[a] = 0.0;
    double chi = 0.0, prod = 1.0;
    int pexp = 0;

    for (int tb = 0; tb < md.N; tb += TILE) {
        const int tn = min(TILE, md.N - tb);
        __syncthreads();
        tile_load(md, s_tile, tb, tn);
        __syncthreads();
        for (int cb = 0; cb < tn; cb += G) {
            const int j = cb + g;
            if (!(live && j < tn)) continue;
            const double t = s_tile.t[j];
            const int k = s_tile.ins[j];
            double mu = sl.gam[k];
            const double x = t - md.t0;
            if (md.acc) {
                double v = 0.0;
                for (int a = 0; a < md.acc; ++a) v = __builtin_fma(v, x, sl.acc[a]);
                mu += v * x;
            }
            double dM[KMAX], de[KMAX], dw[KMAX], dA[KMAX], tau[KMAX];
#pragma unroll
            for (int p = 0; p < KMAX; ++p) {
                dM[p] = de[p] = dw[p] = dA[p] = tau[p] = 0.0;
                if (p < md.K) {
                    const PlanetC& pc = sl.planet[p];
                    const double M = reduce_2pi(planet_M(pc, t));
                    double sE, cE;
                    kepler_fp64(M, pc, s_tile.tab, sE, cE);
                    const double e = pc.e;
                    const double rD = rcp64(__builtin_fma(-e, cE, 1.0));
                    const double bq = sqrt((1.0 - e) * (1.0 + e));
                    const double cosf = (cE - e) * rD, sinf = bq * sE * rD;
                    const double cfw = cosf * pc.cw - sinf * pc.sw;
                    const double sfw = sinf * pc.cw + cosf * pc.sw;
                    const double A = pc.K;
                    mu += A * (cfw + e * pc.cw);
                    dM[p] = -A * sfw * bq * rD * rD;
                    de[p] = -A * sfw * (bq * sE * rD * rD + sE * rD / bq) + A * pc.cw;
                    dw[p] = -A * (sfw + e * pc.sw);
                    dA[p] = cfw + e * pc.cw;
                    tau[p] = pc.tp_mode ? (t - pc.tp) : t;
                }
            }
            const double err2 = s_tile.s2[j] + sl.jit2[k];
            const double r = s_tile.y[j] - mu;
            const double ie = rcp64(err2);
            const double w = r * ie;
            chi = __builtin_fma(r * r, ie, chi);
            prod *= err2;
            pexp += __builtin_amdgcn_frexp_exp(prod);
            prod = __builtin_amdgcn_frexp_mant(prod);
#pragma unroll
            for (int p = 0; p < KMAX; ++p) {
                acc_nat[p][0] = __builtin_fma(w * dM[p], tau[p], acc_nat[p][0]);
                acc_nat[p][1] = __builtin_fma(w, dM[p], acc_nat[p][1]);
                acc_nat[p][2] = __builtin_fma(w, de[p], acc_nat[p][2]);
                acc_nat[p][3] = __builtin_fma(w, dw[p], acc_nat[p][3]);
                acc_nat[p][4] = __builtin_fma(w, dA[p], acc_nat[p][4]);
            }
            if (md.acc) {
                double xp = x;
#pragma unroll
                for (int a = AMAX - 1; a >= 0; --a) {
                    if (a < md.acc) {
                        acc_acc[a] = __builtin_fma(w, xp, acc_acc[a]);
                        xp *= x;
                    }
                }
            }
            atomicAdd(&gl.g[md.off_off + k], w);
            if (md.has_jit) atomicAdd(&gl.g[md.off_jit + k], sl.jit[k] * (w * w - ie));
        }
    }
    // Reduce the natural-parameter accumulators over the slot's waves.
#pragma unroll
    for (int p = 0; p < KMAX; ++p)
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            if (p < md.K) {
                const double v = wave_sum(acc_nat[p][q]);
                if (live && lane == 0) atomicAdd(&gl.nat[p][q], v);
            }
        }
#pragma unroll
    for (int a = 0; a < AMAX; ++a) {
        if (a < md.acc) {
            const double v = wave_sum(acc_acc[a]);
            if (live && lane == 0) atomicAdd(&gl.g[md.off_acc + a], v);
        }
    }
    double ldet = live ? log(prod) + (double)pexp * LN2 : 0.0;
    chi = wave_sum(chi);
    ldet = wave_sum(ldet);
    if (lane == 0) { s_scan.red[wave][0] = chi; s_scan.red[wave][1] = ldet; }
    __syncthreads();
    if (live && wsub == 0) {
        // Map natural gradients to the parameterisation (one lane per planet).
        if (lane < md.K) {
            const int p = lane;
            const PlanetC& pc = sl.planet[p];
            const double* th = sl.theta + md.off_kep + 5 * p;
            double* gp = gl.g + md.off_kep + 5 * p;
            const double per = 2.0 * PI_D / pc.freq;
            const double gP = gl.nat[p][0] * (-pc.freq / per);
            const double gph = gl.nat[p][1];
            const double ge = gl.nat[p][2], gw = gl.nat[p][3], gA = gl.nat[p][4];
            const int prm = md.param;
            const bool lnP = (prm == 2 || prm == 6 || prm == 7);
            const bool sc = (prm == 1 || prm == 2 || prm == 4 || prm == 7);
            gp[0] = lnP ? gP * per : gP;
            if (prm == 2) {
                const double As = th[1], Ac = th[2], A = pc.K;
                gp[1] = gA * 2.0 * As + gph * (Ac / A);
                gp[2] = gA * 2.0 * Ac - gph * (As / A);
            } else {
                gp[1] = gA;
                gp[2] = pc.tp_mode ? -pc.freq * gph : gph;
            }
            if (sc) {
                const double S = th[3], C = th[4], e = pc.e;
                const double thr = (prm == 2 || prm == 4) ? 1e-5 : 1e-6;
                const bool free_w = md.use_c || e >= thr;
                gp[3] = ge * 2.0 * S + (free_w ? gw * (C / e) : 0.0);
                gp[4] = ge * 2.0 * C - (free_w ? gw * (S / e) : 0.0);
            } else {
                gp[3] = ge;
                gp[4] = gw;
            }
        }
    }
    __syncthreads();
    if (live && wsub == 0) {
        double c2 = 0.0, l2 = 0.0;
        for (int w = 0; w < WPT; ++w) { c2 += s_scan.red[wave + w][0]; l2 += s_scan.red[wave + w][1]; }
        double v = -0.5 * (c2 + l2) + md.lnl_const;
        bool bad = !(v == v) || v == __builtin_inf();
        for (int d = lane; d < md.D; d += 64) {
            double gv = gl.g[md.free_idx[d]];
            if (bad) gv = 0.0;
            out_grad[b * md.D + d] = gv;
        }
        if (lane == 0) {
            if (bad) { v = -__builtin_inf(); atomicAdd(nonfinite, 1ull); }
            out_lnl[b] = v;
        }
    }
}

template <int WPT>
static void launch_grad_t(const ModelDev& md, const double* th, int64_t B, int64_t ld, double* lnl, double* grad,
                          unsigned long long* nf, hipStream_t s) {
    constexpr int SLOTS = NWAVES / WPT;
    const int64_t grid = (B + SLOTS - 1) / SLOTS;
    hipLaunchKernelGGL((k_lnl_grad<WPT>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, th, B, ld, lnl, grad, nf);
}

static int launch_grad(const EmpModel* m, const double* th, int64_t B, int64_t ld, double* lnl, double* grad,
                       hipStream_t s) {
    int wpt = emp_pick_wpt(m, B, EMP_FP64);
    if (wpt == 1) launch_grad_t<1>(m->dev, th, B, ld, lnl, grad, m->d_nonfinite, s);
    else if (wpt == 2) launch_grad_t<2>(m->dev, th, B, ld, lnl, grad, m->d_nonfinite, s);
    else if (wpt == 4) launch_grad_t<4>(m->dev, th, B, ld, lnl, grad, m->d_nonfinite, s);
    else launch_grad_t<8>(m->dev, th, B, ld, lnl, grad, m->d_nonfinite, s);
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

extern "C" int emp_lnl_grad_batch(EmpModel* m, const double* theta, int64_t B, int64_t ld, double* lnl,
                                  double* grad, int32_t flags, void* stream) {
    if (!m) { emp_set_error("null model"); return EMP_EINVAL; }
    if (B < 0 || (B > 0 && (!theta || !lnl || !grad))) { emp_set_error("bad batch arguments"); return EMP_EINVAL; }
    if (ld < m->desc.ndim) { emp_set_error("ld < ndim"); return EMP_EINVAL; }
    if (m->dev.ma_active) {
        emp_set_error("gradients with an active moving-average term are not implemented");
        return EMP_EUNSUP;
    }
    if (emp_fault_injected()) return EMP_EFAULT_INJECTED;
    if (B == 0) return EMP_OK;
    hipStream_t s = (hipStream_t)stream;
    if (!(flags & EMP_HOST_POINTERS)) return launch_grad(m, theta, B, ld, lnl, grad, s);
    const size_t nin = (size_t)B * ld, D = (size_t)m->desc.ndim;
    double *din = nullptr, *dl = nullptr, *dg = nullptr;
    EMP_HIP(hipMalloc(&din, nin * 8));
    EMP_HIP(hipMalloc(&dl, (size_t)B * 8));
    EMP_HIP(hipMalloc(&dg, std::max<size_t>((size_t)B * D, 1) * 8));
    int rc = EMP_OK;
    if (hipMemcpyAsync(din, theta, nin * 8, hipMemcpyHostToDevice, s) != hipSuccess) rc = EMP_EHIP;
    if (!rc) rc = launch_grad(m, din, B, ld, dl, dg, s);
    if (!rc && hipMemcpyAsync(lnl, dl, (size_t)B * 8, hipMemcpyDeviceToHost, s) != hipSuccess) rc = EMP_EHIP;
    if (!rc && D && hipMemcpyAsync(grad, dg, (size_t)B * D * 8, hipMemcpyDeviceToHost, s) != hipSuccess) rc = EMP_EHIP;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = EMP_EHIP;
    if (rc == EMP_EHIP) emp_set_error("HIP error in emp_lnl_grad_batch host path");
    (void)hipFree(din);
    (void)hipFree(dl);
    (void)hipFree(dg);
    return rc;
}
