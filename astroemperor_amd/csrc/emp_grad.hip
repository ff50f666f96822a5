// Adjoint kernel: lnL and dlnL/dtheta for a batch of parameter vectors (FP64).
//
// The Kepler solve is differentiated implicitly: from E - e sin E = M,
//   dE/dM = 1/(1 - e cos E),  dE/de|_M = sin E/(1 - e cos E),
// so no derivative flows through the solver's iterations.  With
// D = 1 - e cos E, b = sqrt(1 - e^2), cos f = (cos E - e)/D, sin f = b sin E/D,
// the planet term mu = A (cos(f + w) + e cos w) has
//   dmu/dM = -A sin(f + w) b / D^2
//   dmu/de = -A sin(f + w) (b sin E / D^2 + sin E/(b D)) + A cos w
//   dmu/dw = -A (sin(f + w) + e sin w),   dmu/dA = cos(f + w) + e cos w.
// Without the moving average dlnL/dmu_i = r_i / err2_i and
// dlnL/dJ_k = J_k sum_{i in k} (r_i^2/err2_i^2 - 1/err2_i).
//
// MA(1) (moav01.model:7-15; per instrument for the intended semantics):
// eps_i = r_i + a_i eps_{i-1}, a_i = -c exp(-dt_i / tau), and lnL sums
// -(eps_i^2 / err2_i + ln err2_i) / 2.  Its adjoint lambda_i = dlnL/deps_i
// obeys the reverse recurrence lambda_i = -eps_i / err2_i + a_{i+1}
// lambda_{i+1}; then dlnL/dmu_i = -lambda_i, dlnL/dc = sum lambda_i
// eps_{i-1} da_i/dc and dlnL/dtau likewise, with eps in place of r in the
// jitter term.  The kernel makes two passes: the forward likelihood, which
// stores eps_i (scratch, one row per vector), then the epochs in reverse
// (tiles, then chunks), where the planet terms and their derivatives are
// recomputed and lambda comes from a suffix scan with the carry running
// backwards.
//
// Terms added after the moving average, S_i = sum_j zeta_j a_ij (activity
// columns) + sum_c A_c cos(f_c t_i + ph_c) (sinusoid, magnetic cycle): dlnL/dS_i
// is the final residual over err2_i and does not pass through the MA adjoint:
//   dS/dzeta_j = a_ij, dS/dA = cos M, dS/dph = -A sin M, dS/df = -A sin M t,
//   df/dP = -f/P.
//
// The natural-parameter gradients are then mapped to the block's
// parameterisation (block_repo.py:23-31) once per vector.
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "emp_kernels.h"

using namespace emp;

struct GradLDS {
    double g[MAXD];            // gradient w.r.t. the full vector
    double nat[KMAX][5];       // per planet: (M*tau, M, e, w, A) sums
    double cosg[CMAX][3];      // per cosine term: (A, phase, frequency) sums
};

// Per-wave partial sums, added over a slot's waves in wave order at the end,
// so that the gradient does not depend on how the waves were scheduled.
// acc: offsets, jitters, MA coefficients and timescales (one entry per
// (set, order), order-minor; orders above 1 only in the EXT = 2 instances),
// then, with the terms added after the MA, the cosine terms and the activity
// coefficients.  nat: per planet the five natural-parameter sums, then the
// acceleration coefficients.
template <int EXT>
struct GradWave {
    static constexpr int QM = EXT == 2 ? 4 : 1;
    static constexpr int OFF = 0, JIT = IMAX, MAC = 2 * IMAX, MAT = MAC + IMAX * QM, COS = MAT + IMAX * QM;
    static constexpr int SAI = COS + (EXT >= 1 ? 3 * CMAX : 0);
    static constexpr int N = SAI + (EXT >= 1 ? SAIMAX : 0);
};
constexpr int WN_ACC = KMAX * 5, WN_N = KMAX * 5 + AMAX;
template <int EXT>
struct GradWaveLDS {
    double acc[NWAVES][GradWave<EXT>::N];
    double nat[NWAVES][WN_N];
};

// w times the derivatives of one planet's term at time t with respect to its
// natural parameters, added to a = (dmu/dM tau, dmu/dM, dmu/de, dmu/dw,
// dmu/dA).  With D = 1 - e cos E, b = sqrt(1 - e^2) (per planet, pc.bq):
// D cos f = cos E - e, D sin f = b sin E.
template <int PREC, class KTE>
__device__ __forceinline__ void planet_wgrad(const PlanetC& pc, const SinCosTab& tab, const KTE* kt, double t,
                                             double w, double (&a)[5], bool tp, int kmode) {
    double sE, cE;
    planet_sc<PREC>(pc, tab, kt, t, sE, cE, tp, kmode);
    const double e = pc.e;
    const double rD = rcp_path<PREC>(__builtin_fma(e, -cE, 1.0));
    const double u = cE - e, v = pc.bq * sE;
    const double sfw = __builtin_fma(v, pc.cw, u * pc.sw) * rD;     // sin(f + w)
    const double cfw = __builtin_fma(u, pc.cw, -(v * pc.sw)) * rD;  // cos(f + w)
    const double wA = w * pc.K;
    const double ws = wA * sfw;
    const double dM = -(ws * pc.bq) * (rD * rD);
    const double tau = tp ? t - pc.tp : t;
    a[0] = __builtin_fma(dM, tau, a[0]);
    a[1] += dM;
    a[2] = __builtin_fma(-ws, sE * rD * __builtin_fma(pc.bq, rD, pc.ibq), __builtin_fma(wA, pc.cw, a[2]));
    a[3] = __builtin_fma(-wA, __builtin_fma(e, pc.sw, sfw), a[3]);
    a[4] = __builtin_fma(w, __builtin_fma(e, pc.cw, cfw), a[4]);
}

// (x, y) exchanged between lanes 32 apart by v_permlane32_swap, which swaps
// the upper half of its first operand with the lower half of its second:
// lanes 0-31 end with (x, x of lane + 32), lanes 32-63 with (y of lane - 32,
// y), so the sum holds x summed over the pair below and y above.
__device__ __forceinline__ double swap_add32(double x, double y) {
    const long long xi = __double_as_longlong(x), yi = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xi, (unsigned)yi, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(xi >> 32), (unsigned)(yi >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) +
           __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}

// The same between lanes 16 apart within each half (v_permlane16_swap):
// x summed over the pair where lane bit 4 is 0, y where it is 1.
__device__ __forceinline__ double swap_add16(double x, double y) {
    const long long xi = __double_as_longlong(x), yi = __double_as_longlong(y);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)xi, (unsigned)yi, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(xi >> 32), (unsigned)(yi >> 32), false, false);
    return __longlong_as_double(((long long)hi[0] << 32) | lo[0]) +
           __longlong_as_double(((long long)hi[1] << 32) | lo[1]);
}

// Sum of five per-lane values over the wave, scattered: the values are
// paired across lanes 32 apart as (v0, v4), (v1, v2), (v3, v3) (permlane
// swaps: each lane keeps the member its lane bit selects, summed with its
// partner's), those across lanes 16 apart as (v0|v4, v1|v2) and (v3, v3),
// the two results across lanes 8 apart (DPP), and the remaining value is
// summed over each aligned group of eight lanes.  The totals land at lanes
// 0 (v0), 16 (v1), 48 (v2), 8 (v3) and 32 (v4); red5_index(lane) names the
// value a lane holds, -1 for lanes that hold none.  Whole wave active.
__device__ __forceinline__ double wave_reduce5(const double (&v)[5], int lane) {
    const double a0 = swap_add32(v[0], v[4]);
    const double a1 = swap_add32(v[1], v[2]);
    const double a2 = swap_add32(v[3], v[3]);
    const double b0 = swap_add16(a0, a1);
    const double b1 = swap_add16(a2, a2);
    const bool h3 = lane & 8;
    double c = (h3 ? b1 : b0) + dpp_mov_f64<0x128>(h3 ? b0 : b1);   // row_ror 8: lane ^ 8
    c += dpp_mov_f64<0x141>(c);    // row_half_mirror: lane i with 7 - i of its group of 8
    c += dpp_mov_f64<0xB1>(c);     // quad_perm [1, 0, 3, 2]
    c += dpp_mov_f64<0x4E>(c);     // quad_perm [2, 3, 0, 1]
    return c;
}

// Nibble g of 0xF2F4F130 is the value lane 8 g holds (F: none).
__device__ __forceinline__ int red5_index(int lane) {
    const int i = (int)((0xF2F4F130u >> (4 * (lane >> 3))) & 0xFu);
    return ((lane & 7) || i == 0xF) ? -1 : i;
}

// Four values: (v0, v1) and (v2, v3) across lanes 32 apart, the pair across
// lanes 16 apart, then summed over each row of 16 lanes.  The totals land at
// lanes 0 (v0), 32 (v1), 16 (v2) and 48 (v3): red4_index(lane).
__device__ __forceinline__ double wave_reduce4(const double (&v)[4]) {
    const double a0 = swap_add32(v[0], v[1]);
    const double a1 = swap_add32(v[2], v[3]);
    double c = swap_add16(a0, a1);
    c += dpp_mov_f64<0x128>(c);
    c += dpp_mov_f64<0x141>(c);
    c += dpp_mov_f64<0xB1>(c);
    c += dpp_mov_f64<0x4E>(c);
    return c;
}

__device__ __forceinline__ int red4_index(int lane) {
    return (lane & 15) ? -1 : (int)((0x3120u >> (4 * (lane >> 4))) & 0xFu);
}

// The five sums added to dst[0..4], the first n of four to dst[0..n).
__device__ __forceinline__ void wave_reduce_add5(const double (&v)[5], double* dst, int lane, int ri) {
    const double c = wave_reduce5(v, lane);
    if (ri >= 0) dst[ri] += c;
}

__device__ __forceinline__ void wave_reduce_add4(const double (&v)[4], int n, double* dst, int ri4) {
    const double c = wave_reduce4(v);
    if (ri4 >= 0 && ri4 < n) dst[ri4] += c;
}

// Adds w = dlnL/dS_n times the derivatives of the terms added after the MA
// at epoch n (time t, instrument k).
template <int EXT, class SL>
__device__ __forceinline__ void post_accumulate(const ModelDev& md, const SL& sl, const SinCosTab& tab, double t,
                                                int n, int k, double w, double* wacc) {
    using GW = GradWave<EXT>;
    if (md.sai_w) {
        const int b0 = md.sai_base[k], nc = md.sai_base[k + 1] - b0;
        for (int c = 0; c < nc; ++c) atomicAdd(&wacc[GW::SAI + b0 + c], w * md.sai[(size_t)c * md.N + n]);
    }
    double* acc = wacc + GW::COS;
    for (int c = 0; c < md.n_cos; ++c) {
        const double r = reduce_2pi(cos_phase(sl, c, t));
        double sn, cs;
        sincos_tab(tab, __builtin_fabs(r), sn, cs);
        const double dph = -sl.cosw[c][0] * (r < 0.0 ? -sn : sn);
        atomicAdd(&acc[3 * c + 0], w * cs);
        atomicAdd(&acc[3 * c + 1], w * dph);
        atomicAdd(&acc[3 * c + 2], w * dph * t);
    }
}

// MA(1) multiplier of epoch n (device order) from its instrument and its MA
// column (the fixed-timescale weight, or dt with -1 at a segment start).
template <class SL>
__device__ __forceinline__ double ma_mult(const ModelDev& md, const SL& sl, int k, double col) {
    if (md.ma_wfixed) return -(sl.ma_c[k] * col);
    return col < 0.0 ? 0.0 : -(sl.ma_c[k] * exp(-col / sl.ma_tau[k]));
}

// Adjoint of MA(q), q >= 2, for one chunk (epochs in reverse): with
// a_{n,c} the forward multipliers, lambda_n = u_n + sum_{c<q} a_{n+1+c,c}
// lambda_{n+1+c}, an affine map on the state (lambda_n, ..., lambda_{n+Q-1})
// whose companion matrix has first row (a_{n+1,0}, ..., a_{n+Q,Q-1}).  Maps
// are composed from the wave's last lane down (suffix), then across the
// slot's waves from the last one.  `carry` is the state after the chunk
// (updated to the state before it).  Reads the tile before the barrier.
template <int Q, int WPT, class SL>
__device__ __forceinline__ double lam_chunk_q(const ModelDev& md, const SL& sl, const TileLDS& tl, bool valid,
                                              int j, int tb, int tn, double t, int k, double u, int lane, int wave,
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
        const int q = md.ma_q, n = tb + j;
        const double* ph = sl.theta + md.off_ma + 2 * q * (md.ma_global ? 0 : k);
#pragma unroll
        for (int c = 0; c < Q; ++c) {
            double a = 0.0;
            const int n2 = n + 1 + c, j2 = j + 1 + c;
            if (c < q && n2 < md.N) {
                const int k2 = j2 < tn ? tl.ins[j2] : md.ins[n2];
                if (md.ma_global || k2 == k) {
                    const double t2 = j2 < tn ? tl.t[j2] : md.t[n2];
                    a = -(ph[2 * c] * exp(-__builtin_fabs(t2 - t) / ph[2 * c + 1]));
                }
            }
            m[0][c] = a;
        }
#pragma unroll
        for (int a = 1; a < Q; ++a)
#pragma unroll
            for (int b = 0; b < Q; ++b) m[a][b] = b == a - 1 ? 1.0 : 0.0;
        v[0] = u;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        double pm[Q][Q], pv[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a) {
#pragma unroll
            for (int b = 0; b < Q; ++b) pm[a][b] = __shfl_down(m[a][b], o, 64);
            pv[a] = __shfl_down(v[a], o, 64);
        }
        if (lane + o < 64) {
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
    }
    double sin_[Q];
    if (WPT == 1) {
#pragma unroll
        for (int a = 0; a < Q; ++a) sin_[a] = carry[a];
    } else {
        if (lane == 0) {
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
        for (int w = WPT - 1; w >= 0; --w) {
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
        for (int a = 0; a < Q; ++a) carry[a] = __shfl(st[a], 0, 64);
    }
    return st[0];
}

// Per-lane sums of the per-instrument gradient terms (offset, jitter, MA(1)
// coefficient and timescale) of the lane's epochs, kept in registers while
// the lane's instrument stays the same and added to the wave's LDS sums when
// it changes (epochs grouped by instrument, as the per-instrument MA orders
// them, change instrument a few times per lane in a whole pass), instead of
// four same-address LDS atomics per epoch.
struct InsAcc {
    int k = -1;
    double off = 0.0, jit = 0.0, mac = 0.0, mat = 0.0;
    template <int EXT>
    __device__ __forceinline__ void flush(double* wacc) {
        using GW = GradWave<EXT>;
        if (k < 0) return;
        atomicAdd(&wacc[GW::OFF + k], off);
        atomicAdd(&wacc[GW::JIT + k], jit);
        atomicAdd(&wacc[GW::MAC + k * GW::QM], mac);
        atomicAdd(&wacc[GW::MAT + k * GW::QM], mat);
        off = jit = mac = mat = 0.0;
    }
    template <int EXT>
    __device__ __forceinline__ void to(int kk, double* wacc) {
        if (kk != k) {
            flush<EXT>(wacc);
            k = kk;
        }
    }
};

// The MA adjoint at one chunk of a tile taken in reverse: lambda of the
// lane's epoch from eps (the forward pass's scratch row), with the carry from
// the later epochs; returns w = dlnL/dmu = -lambda (0 for an invalid lane)
// after adding the epoch's offset, jitter, post-MA and MA-parameter terms.
template <int PREC, int WPT, int EXT, class SL>
__device__ __forceinline__ double ma_chunk_w(const ModelDev& md, const SL& sl, const TileLDS& tl, ScanLDS& sc,
                                             const double* eps, double* wacc, InsAcc& ia, bool valid, int j, int tb,
                                             int tn, int lane, int wave, int wsub, int& parity, double& lam_carry,
                                             double (&lq_carry)[4]) {
    using GW = GradWave<EXT>;
    constexpr bool POST = EXT >= 1;
    const bool maq = EXT == 2 && md.ma_q > 1;
    double u = 0.0, amul = 1.0, ie = 0.0, t = 0.0, ef = 0.0;
    int k = 0;
    if (valid) {
        t = tl.t[j];
        k = tl.ins[j];
        const double err2 = tl.s2[j] + sl.jit2[k];
        ie = rcp_path<PREC>(err2);
        const double e_i = eps[tb + j];
        ef = (POST && has_post(md)) ? e_i - epoch_post(md, sl, tl.tab, t, tb + j, k) : e_i;
        u = -ef * ie;
        const int n1 = tb + j + 1;     // the next epoch's multiplier
        amul = 0.0;
        if (!maq && n1 < md.N) {
            const bool in_tile = j + 1 < tn;
            const int k1 = in_tile ? tl.ins[j + 1] : md.ins[n1];
            const double c1 = in_tile ? tl.dt[j + 1] : md.dt[n1];
            amul = ma_mult(md, sl, k1, c1);
        }
    }
    double sa = amul, sb = u;
    double lam = 0.0;
    if (maq) {
        lam = md.ma_q == 2
                  ? lam_chunk_q<2, WPT>(md, sl, tl, valid, j, tb, tn, t, k, u, lane, wave, wsub, parity, lq_carry)
                  : lam_chunk_q<4, WPT>(md, sl, tl, valid, j, tb, tn, t, k, u, lane, wave, wsub, parity, lq_carry);
    } else if (WPT == 1) {
        wave_affine_suffix_scan(sa, sb, lane);
        lam = __builtin_fma(sa, lam_carry, sb);
        lam_carry = __shfl(lam, 0, 64);
    } else {
        wave_affine_suffix_scan(sa, sb, lane);
        if (lane == 0) { sc.tot[parity][wave][0] = sa; sc.tot[parity][wave][1] = sb; }
        __syncthreads();
        const int w0 = wave - wsub;
        double pa = 1.0, pb = 0.0;
        for (int w = WPT - 1; w >= 0; --w) {
            if (w == wsub) lam = __builtin_fma(sa, __builtin_fma(pa, lam_carry, pb), sb);
            const double ta = sc.tot[parity][w0 + w][0], tbv = sc.tot[parity][w0 + w][1];
            pb = __builtin_fma(ta, pb, tbv);
            pa = ta * pa;
        }
        lam_carry = __builtin_fma(pa, lam_carry, pb);
        parity ^= 1;
    }
    if (!valid) return 0.0;
    const double w = -lam;
    ia.to<EXT>(k, wacc);
    ia.off += w;
    if (md.has_jit) ia.jit = __builtin_fma(sl.jit[k], ef * ef * ie * ie - ie, ia.jit);
    if (POST && has_post(md)) post_accumulate<EXT>(md, sl, tl.tab, t, tb + j, k, -u, wacc);
    if (maq) {
        // d a_{n,c} / d phi_c = -exp(-dt / tau_c), d a_{n,c} / d tau_c = a_{n,c} dt / tau_c^2
        const int n = tb + j, set = md.ma_global ? 0 : k;
        const double* ph = sl.theta + md.off_ma + 2 * md.ma_q * set;
        for (int c = 0; c < md.ma_q; ++c) {
            const int n1 = n - 1 - c, j1 = j - 1 - c;
            if (n1 < 0) break;
            const int k1 = j1 >= 0 ? tl.ins[j1] : md.ins[n1];
            if (!(md.ma_global || k1 == k)) break;
            const double dt = __builtin_fabs(t - (j1 >= 0 ? tl.t[j1] : md.t[n1]));
            const double tau = ph[2 * c + 1];
            const double wgt = exp(-dt / tau);
            const double le = lam * eps[n1];
            atomicAdd(&wacc[GW::MAC + set * GW::QM + c], le * -wgt);
            atomicAdd(&wacc[GW::MAT + set * GW::QM + c], le * -(ph[2 * c] * wgt * dt / (tau * tau)));
        }
    } else {
        // da_i/dc = a_i / c, da_i/dtau = a_i dt / tau^2 (a_i = 0 at a segment start)
        const double col = tl.dt[j];
        const bool start = md.ma_wfixed ? col == 0.0 : col < 0.0;
        if (!start) {
            const int n = tb + j;
            const double le = lam * eps[n - 1];
            const double wgt = md.ma_wfixed ? col : exp(-col / sl.ma_tau[k]);
            if (md.ma_global) {
                atomicAdd(&wacc[GW::MAC], le * -wgt);
                if (!md.ma_wfixed) {
                    const double tau = sl.ma_tau[k];
                    atomicAdd(&wacc[GW::MAT], le * -(sl.ma_c[k] * wgt * col / (tau * tau)));
                }
            } else {
                ia.mac = __builtin_fma(le, -wgt, ia.mac);
                if (!md.ma_wfixed) {
                    const double tau = sl.ma_tau[k];
                    ia.mat = __builtin_fma(le, -(sl.ma_c[k] * wgt * col / (tau * tau)), ia.mat);
                }
            }
        }
    }
    return w;
}

// Without the MA: the lane's epoch's residual, its share of chi^2 and of the
// log-determinant, w = dlnL/dmu = r / err2 and the epoch's offset, jitter and
// post-MA terms.
template <int PREC, int EXT, class SL>
__device__ __forceinline__ double plain_w(const ModelDev& md, const SL& sl, const TileLDS& tl, double* wacc,
                                          InsAcc& ia, int j, int tb, double& chi, double& prod, int& pexp,
                                          int kmask) {
    using GW = GradWave<EXT>;
    constexpr bool POST = EXT >= 1;
    const double t = tl.t[j];
    const int k = tl.ins[j];
    const double mu = epoch_mu<PREC>(md, sl, tl.tab, t, k, kmask);
    const double err2 = tl.s2[j] + sl.jit2[k];
    const double cs = (POST && has_post(md)) ? epoch_post(md, sl, tl.tab, t, tb + j, k) : 0.0;
    const double r = tl.y[j] - mu - cs;
    const double ie = rcp_path<PREC>(err2);
    const double w = r * ie;
    chi = __builtin_fma(r * r, ie, chi);
    prod *= err2;
    pexp += __builtin_amdgcn_frexp_exp(prod);
    prod = __builtin_amdgcn_frexp_mant(prod);
    ia.to<EXT>(k, wacc);
    ia.off += w;
    if (md.has_jit) ia.jit = __builtin_fma(sl.jit[k], w * w - ie, ia.jit);
    if (POST && has_post(md)) post_accumulate<EXT>(md, sl, tl.tab, t, tb + j, k, w, wacc);
    return w;
}

// Adds the per-wave partial sums of a slot in wave order, maps the natural
// gradients to the parameterisation and writes lnL and the free gradient of
// vector b.  All threads of the workgroup call this (it holds barriers).
template <int EXT, int WPT, class SL>
__device__ __forceinline__ void grad_finish(const ModelDev& md, const SL& sl, GradLDS& gl,
                                            const GradWaveLDS<EXT>& s_gw, bool live, int wave, int wsub, int lane,
                                            double lnl_v, int64_t b, double* __restrict__ out_lnl,
                                            double* __restrict__ out_grad, unsigned long long* nonfinite, int* stale) {
    using GW = GradWave<EXT>;
    constexpr bool POST = EXT >= 1;
    __syncthreads();
    if (live && wsub == 0) {
        const int w0 = wave;
        for (int e = lane; e < GW::N; e += 64) {
            double v = 0.0;
            for (int w = 0; w < WPT; ++w) v += s_gw.acc[w0 + w][e];
            if (POST && e >= GW::SAI) {
                if (md.sai_w && e - GW::SAI < md.sai_base[md.I]) gl.g[md.off_sai + (e - GW::SAI)] = v;
            } else if (POST && e >= GW::COS) {
                (&gl.cosg[0][0])[e - GW::COS] = v;
            } else if (e >= GW::MAC) {
                const bool tau = e >= GW::MAT;
                const int r = e - (tau ? GW::MAT : GW::MAC), set = r / GW::QM, c = r - set * GW::QM;
                if (md.ma_active && set < (md.ma_global ? 1 : md.I) && c < md.ma_q)
                    gl.g[md.off_ma + 2 * md.ma_q * set + 2 * c + (tau ? 1 : 0)] = v;
            } else if (e >= GW::JIT) {
                if (md.has_jit && e - GW::JIT < md.I) gl.g[md.off_jit + (e - GW::JIT)] = v;
            } else if (e < md.I) {
                gl.g[md.off_off + e] = v;
            }
        }
        for (int e = lane; e < WN_N; e += 64) {
            double v = 0.0;
            for (int w = 0; w < WPT; ++w) v += s_gw.nat[w0 + w][e];
            if (e < WN_ACC) {
                if (e < md.K * 5) (&gl.nat[0][0])[e] = v;
            } else if (e - WN_ACC < md.acc) {
                gl.g[md.off_acc + (e - WN_ACC)] = v;
            }
        }
    }
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
        if (POST && lane == 63) {
            for (int c = 0; c < md.n_cos; ++c) {
                const int m = md.off_sin >= 0 ? c - 1 : c;    // -1: the sinusoid
                const int o = m < 0 ? md.off_sin : md.off_mag;
                const double gP = gl.cosg[c][2] * -(sl.cosw[c][1] / sl.theta[o]);
                if (m < 0) {
                    gl.g[o] = gP;
                    gl.g[o + 1] = gl.cosg[c][0];
                    gl.g[o + 2] = gl.cosg[c][1];
                } else {
                    gl.g[o] += gP;
                    gl.g[o + 1 + m] = gl.cosg[c][0];
                    gl.g[o + 3 + m] = gl.cosg[c][1];
                }
            }
        }
    }
    __syncthreads();
    if (live && wsub == 0) {
        double v = lnl_v;
        bool bad = !(v == v) || v == __builtin_inf();
        for (int d = lane; d < md.D; d += 64) {
            double gv = gl.g[md.free_idx[d]];
            if (bad) gv = 0.0;
            out_grad[b * md.D + d] = gv;
        }
        if (lane == 0) {
            if (bad) { v = -__builtin_inf(); atomicAdd(nonfinite, 1ull); }
            out_lnl[b] = v;
            if (stale) stale[b] = 0;
        }
    }
}

// Adjoint kernel.  Per tile (forward without the MA, in reverse with it):
// stage 1 finds w = dlnL/dmu of each lane's epochs (CPT chunks of 64 WPT
// epochs, held in registers), stage 2 then takes the planets one at a time,
// adding w times each planet's derivatives over the lane's epochs and
// reducing the five sums over the wave once per planet per tile, and the
// acceleration terms likewise.  No derivative is held across planets, which
// keeps the register count near the likelihood kernel's.  With the MA a
// forward pass first computes lnL and stores eps (one scratch row per
// vector).  `stale` (optional, per vector): vectors whose flag is 0 are
// skipped; a computed vector's flag is cleared.
// EXT: 0 plain models, 1 with terms added after the MA, 2 also MA orders
// above 1.
template <int PREC, int WPT, int EXT>
__global__ __launch_bounds__(BLOCK) void k_lnl_grad(ModelDev md, const double* __restrict__ theta, int64_t B,
                                                    int64_t ld, double* __restrict__ out_lnl,
                                                    double* __restrict__ out_grad,
                                                    unsigned long long* nonfinite, double* eps_scratch, int* stale) {
    constexpr int SLOTS = NWAVES / WPT;
    constexpr int G = 64 * WPT;
    constexpr int CPT = TILE / G;             // chunks of a tile per wave
    constexpr bool POST = EXT >= 1;
    using GW = GradWave<EXT>;
    __shared__ SlotLDS<SLOTS> s_slot[SLOTS];
    __shared__ GradLDS s_grad[SLOTS];
    __shared__ GradWaveLDS<EXT> s_gw;
    __shared__ TileLDS s_tile;
    __shared__ ScanLDS s_scan;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave / WPT, wsub = wave % WPT;
    const int g = wsub * 64 + lane;
    const int64_t b = (int64_t)blockIdx.x * SLOTS + slot;
    const bool live = b < B && (!stale || stale[b] != 0);
    if (!__syncthreads_or(live)) return;       // no vector of this workgroup to compute
    SlotLDS<SLOTS>& sl = s_slot[slot];
    GradLDS& gl = s_grad[slot];
    if (live) expand_theta(md, sl, theta + b * ld, g, G);
    for (int j = g; j < MAXD; j += G) gl.g[j] = 0.0;
    for (int j = g; j < KMAX * 5; j += G) (&gl.nat[0][0])[j] = 0.0;
    double* wacc = s_gw.acc[wave];
    double* wnat = s_gw.nat[wave];
    for (int j = lane; j < GW::N; j += 64) wacc[j] = 0.0;
    for (int j = lane; j < WN_N; j += 64) wnat[j] = 0.0;
    tab_load(md, s_tile);
    __syncthreads();
    slot_prepare<PREC, WPT>(md, sl, s_tile, live, wsub, lane);
    const int kmask = wave_uniform(sl.ktab_mask);
    const bool tp = model_tp(md);

    const bool ma = md.ma_active;
    double* eps = ma ? eps_scratch + (live ? b : 0) * (int64_t)md.N : nullptr;
    double lnl_v = 0.0;
    if (ma) {
        // Pass 1: the likelihood, storing eps.
        lnl_v = block_lnl<PREC, WPT, false, true, true, POST>(md, sl, s_tile, s_scan, live, wsub, nullptr, nullptr,
                                                             false, eps);
        __syncthreads();   // eps rows written by every wave of the slot
    }
    double chi = 0.0, prod = 1.0;
    int pexp = 0;
    double lam_carry = 0.0;
    double lq_carry[4] = {0.0, 0.0, 0.0, 0.0};     // MA(q >= 2) adjoint state
    int parity = 0;
    InsAcc ia;
    const int ntile = (md.N + TILE - 1) / TILE;
    for (int i = 0; i < ntile; ++i) {
        const int it = ma ? ntile - 1 - i : i;
        const int tb = it * TILE, tn = min(TILE, md.N - tb);
        __syncthreads();
        tile_load(md, s_tile, tb, tn);
        __syncthreads();
        double wv[CPT];
        if (!ma) {
#pragma unroll
            for (int c = 0; c < CPT; ++c) {
                const int j = c * G + g;
                wv[c] = (live && j < tn) ? plain_w<PREC, EXT>(md, sl, s_tile, wacc, ia, j, tb, chi, prod, pexp, kmask)
                                         : 0.0;
            }
        } else {
#pragma unroll
            for (int c = CPT - 1; c >= 0; --c) {
                wv[c] = 0.0;
                if (c * G >= tn) continue;            // uniform over the workgroup
                const int j = c * G + g;
                wv[c] = ma_chunk_w<PREC, WPT, EXT>(md, sl, s_tile, s_scan, eps, wacc, ia, live && j < tn, j, tb, tn,
                                                   lane, wave, wsub, parity, lam_carry, lq_carry);
            }
        }
        if (live) {
            for (int p = 0; p < md.K; ++p) {
                double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                const PlanetC& pc = sl.planet[p];
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const int j = c * G + g;
                    if (j < tn) planet_wgrad<PREC>(pc, s_tile.tab, sl.ktab[p], s_tile.t[j], wv[c], a, tp, kt_mode(kmask, p));
                }
                wave_reduce_add5(a, wnat + 5 * p, lane, red5_index(lane));
            }
            if (md.acc) {
                double a[AMAX] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int c = 0; c < CPT; ++c) {
                    const int j = c * G + g;
                    if (j < tn) {
                        const double x = s_tile.t[j] - md.t0;
                        double xp = wv[c] * x;
#pragma unroll
                        for (int q = AMAX - 1; q >= 0; --q) {
                            if (q < md.acc) {
                                a[q] += xp;
                                xp *= x;
                            }
                        }
                    }
                }
                wave_reduce_add4(a, md.acc, wnat + WN_ACC, red4_index(lane));
            }
        }
    }
    ia.flush<EXT>(wacc);
    if (!ma) {
        double ldet = live ? log(prod) + (double)pexp * LN2 : 0.0;
        chi = wave_sum(chi);
        ldet = wave_sum(ldet);
        if (lane == 0) { s_scan.red[wave][0] = chi; s_scan.red[wave][1] = ldet; }
        __syncthreads();
        double c2 = 0.0, l2 = 0.0;
        const int w0 = wave - wsub;
        for (int w = 0; w < WPT; ++w) { c2 += s_scan.red[w0 + w][0]; l2 += s_scan.red[w0 + w][1]; }
        lnl_v = -0.5 * (c2 + l2) + md.lnl_const;
    }
    grad_finish<EXT, WPT>(md, sl, gl, s_gw, live, wave, wsub, lane, lnl_v, b, out_lnl, out_grad, nonfinite, stale);
}

// ---------------------------------------------------------------------------
// Single-pass adjoint for models whose moving average has order 1 or is
// absent: eight waves per vector, one vector per workgroup, one epoch per
// lane per tile.  Each planet's Kepler solve is done once per epoch; its
// (sin E, cos E) stay in registers from the model to the derivatives.
//
// The MA(1) adjoint is carried forward.  Within a tile the reverse
// recurrence lambda_i = u_i + a_{i+1} lambda_{i+1} composes (a suffix scan
// started from zero) into lambda_i = B_i + A_i Lambda, with Lambda the value
// at the next tile's first epoch, unknown while the tile is processed.  Every
// gradient sum is linear in lambda: sum_i lambda_i h_i = S_B + Lambda S_A,
// S_B = sum B_i h_i, S_A = sum A_i h_i.  With (A^n, B^n) the map of tile n as
// a whole (that of its first epoch), Lambda_{n-1} = B^n + A^n Lambda_n and
// Lambda = 0 after the last tile, so
//   sum_n (S_B^n + Lambda_n S_A^n) = sum_n (S_B^n + B^n R_n),
//   R_0 = 0, R_{n+1} = A^n R_n + S_A^n.
// Every such sum keeps a total and an R state and is complete after one pass
// over the epochs; no eps row is stored.
// ---------------------------------------------------------------------------

// Per-lane sums of the per-instrument terms (offset, jitter, MA(1)
// coefficient and timescale) of the lane's epochs, the offset and MA sums
// with their R states; added to the wave's LDS totals and R states when the
// lane's instrument changes (the MA sums are keyed by instrument, or by 0 for
// a global MA).
struct InsAcc1 {
    int k = -1, km = -1;
    double off = 0.0, offR = 0.0, jit = 0.0, mac = 0.0, macR = 0.0, mat = 0.0, matR = 0.0;
    template <int EXT>
    __device__ __forceinline__ void flush_ins(double* tot, double* rst) {
        using GW = GradWave<EXT>;
        if (k < 0) return;
        atomicAdd(&tot[GW::OFF + k], off);
        atomicAdd(&rst[GW::OFF + k], offR);
        atomicAdd(&tot[GW::JIT + k], jit);
        off = offR = jit = 0.0;
    }
    template <int EXT>
    __device__ __forceinline__ void flush_ma(double* tot, double* rst) {
        using GW = GradWave<EXT>;
        if (km < 0) return;
        atomicAdd(&tot[GW::MAC + km * GW::QM], mac);
        atomicAdd(&rst[GW::MAC + km * GW::QM], macR);
        atomicAdd(&tot[GW::MAT + km * GW::QM], mat);
        atomicAdd(&rst[GW::MAT + km * GW::QM], matR);
        mac = macR = mat = matR = 0.0;
    }
    template <int EXT>
    __device__ __forceinline__ void to(int kk, int kkm, double* tot, double* rst) {
        if (kk != k) {
            flush_ins<EXT>(tot, rst);
            k = kk;
        }
        if (kkm != km) {
            flush_ma<EXT>(tot, rst);
            km = kkm;
        }
    }
};

// The epochs of one tile of the single-pass kernel (NT of them) and the
// sin/cos table; thread j stages epoch tb + j, its loads issued a tile ahead.
template <int NT>
struct Tile1LDS {
    double t[NT], y[NT], s2[NT], dt[NT];
    int ins[NT];
    SinCosTab tab;
};

template <int NT>
__device__ __forceinline__ void tile1_prefetch(const ModelDev& md, int tb, EpochRegs& r) {
    const int n = tb + (int)threadIdx.x;
    if (n < md.N) {
        r.t = md.t[n];
        r.y = md.y[n];
        r.s2 = md.s2[n];
        r.ins = md.ins[n];
        r.dt = md.ma_active ? md.dt[n] : 0.0;
    }
}

template <int NT>
__device__ __forceinline__ void tile1_store(const ModelDev& md, Tile1LDS<NT>& tl, int tb, const EpochRegs& r) {
    const int j = threadIdx.x;
    if (tb + j < md.N) {
        tl.t[j] = r.t;
        tl.y[j] = r.y;
        tl.s2[j] = r.s2;
        tl.ins[j] = r.ins;
        tl.dt[j] = r.dt;
    }
}

template <int EXT>
struct Grad1LDS {
    double rst[NWAVES][GradWave<EXT>::N];    // R states of the per-instrument sums (per wave)
    double natR[NWAVES][WN_N];               // R states of the planet and acceleration sums (per wave)
    double suf[2][NWAVES][2];                // suffix maps of the waves (double-buffered)
};

// Single-pass adjoint (above).  EXT: 0 plain models, 1 with the terms added
// after the MA.  NW waves per workgroup (one vector), tiles of 64 NW epochs.
// TP: the model's planets are parameterised by the time of periastron
// (model_tp), a template parameter so that no planet-epoch branches on it.
// `stale` as in k_lnl_grad.
template <int PREC, int EXT, int NW, bool TP>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(EXT == 0 ? 4 : 3))) void k_lnl_grad1(ModelDev md, const double* __restrict__ theta, int64_t B,
                                                       int64_t ld, double* __restrict__ out_lnl,
                                                       double* __restrict__ out_grad, unsigned long long* nonfinite,
                                                       int* stale) {
    constexpr int WPT = NW, NT = 64 * NW, TILE1 = NT;
    constexpr bool POST = EXT >= 1;
    using GW = GradWave<EXT>;
    __shared__ SlotLDS<1> sl;
    __shared__ GradLDS gl;
    __shared__ GradWaveLDS<EXT> s_gw;
    __shared__ Grad1LDS<EXT> s_g1;
    __shared__ Tile1LDS<TILE1> s_tile;
    __shared__ ScanLDS s_scan;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = threadIdx.x;
    const int64_t b = blockIdx.x;
    if (b >= B || (stale && stale[b] == 0)) return;     // the whole workgroup: one vector
    expand_theta(md, sl, theta + b * ld, g, NT);
    for (int j = g; j < MAXD; j += NT) gl.g[j] = 0.0;
    for (int j = g; j < KMAX * 5; j += NT) (&gl.nat[0][0])[j] = 0.0;
    double* wacc = s_gw.acc[wave];
    double* wnat = s_gw.nat[wave];
    double* wrst = s_g1.rst[wave];
    double* wnatR = s_g1.natR[wave];
    for (int j = lane; j < GW::N; j += 64) wacc[j] = wrst[j] = 0.0;
    for (int j = lane; j < WN_N; j += 64) wnat[j] = wnatR[j] = 0.0;
    for (int j = g; j <= SC_N; j += NT) s_tile.tab.v[j] = md.sctab[j];
    if (wave == 0) slot_setup(md, sl, lane);
    __syncthreads();
    if (PREC == 1)      // the mixed path's per-vector Kepler tables, one planet per wave
        for (int p = wave; p < md.K; p += NW)
            if (sl.planet[p].ktab) ktab_fill(sl.planet[p], lane, sl.ktab[p][lane]);
    __syncthreads();
    const int kmask = wave_uniform(sl.ktab_mask);
    constexpr bool tp = TP;
    const int ri5 = red5_index(lane), ri4 = red4_index(lane);    // the sums a lane holds after a reduction

    const bool ma = md.ma_active;
    double chi = 0.0, prod = 1.0;
    int pexp = 0;
    double carry = 0.0;                 // eps of the last epoch so far
    int parity = 0, sparity = 0;
    InsAcc1 ia;
    EpochRegs nx{};
    tile1_prefetch<TILE1>(md, 0, nx);
    for (int tb = 0; tb < md.N; tb += TILE1) {
        const int tn = min(TILE1, md.N - tb);
        __syncthreads();                // the previous tile's readers are done
        tile1_store<TILE1>(md, s_tile, tb, nx);
        __syncthreads();
        if (tb + TILE1 < md.N) tile1_prefetch<TILE1>(md, tb + TILE1, nx);
        const int j = g;
        const bool valid = j < tn;
        // Stage 1: every planet's (sin E, cos E), the model and the residual.
        // Every lane solves the planets, a lane past the series' end at
        // t = 0: its values are finite and its weights (wB, wA) exactly zero,
        // so neither these arrays nor the derivatives below are cleared.
        double sE[KMAX], cE[KMAX];
        double t = 0.0, r = 0.0, err2 = 1.0, mu = 0.0;
        int k = 0;
        if (valid) {
            t = s_tile.t[j];
            k = s_tile.ins[j];
            mu = sl.gam[k];
            if (md.acc) {
                const double x = t - md.t0;
                double v = 0.0;
                for (int a = 0; a < md.acc; ++a) v = __builtin_fma(v, x, sl.acc[a]);
                mu += v * x;
            }
        }
#pragma unroll
        for (int p = 0; p < KMAX; ++p) {
            if (p < md.K) {
                const PlanetC& pc = sl.planet[p];
                planet_sc<PREC>(pc, s_tile.tab, sl.ktab[p], t, sE[p], cE[p], tp, kt_mode(kmask, p));
                mu += __builtin_fma(pc.c1, cE[p], pc.c2 * sE[p]) * rcp_path<PREC>(__builtin_fma(pc.e, -cE[p], 1.0));
            }
        }
        if (valid) {
            err2 = s_tile.s2[j] + sl.jit2[k];
            r = s_tile.y[j] - mu;
            ia.to<EXT>(k, md.ma_global ? 0 : k, wacc, wrst);
        }
        // The MA filter, forward (the same scan and composition as block_lnl).
        double eps = r, em1 = 0.0;      // eps of this epoch and of the previous one
        if (ma) {
            double a = 1.0, bb = 0.0;
            if (valid) {
                const double dtj = s_tile.dt[j];
                if (md.ma_wfixed) a = -(sl.ma_c[k] * dtj);
                else a = dtj < 0.0 ? 0.0 : -(sl.ma_c[k] * exp(-dtj / sl.ma_tau[k]));
                bb = r;
            }
            wave_affine_scan(a, bb, lane);
            if (lane == 63) { s_scan.tot[parity][wave][0] = a; s_scan.tot[parity][wave][1] = bb; }
            __syncthreads();
            // (pa, pb): the maps of waves 0 .. w composed
            double e0 = carry;
            if (wave == 0) eps = __builtin_fma(a, carry, bb);
            double pa = s_scan.tot[parity][0][0], pb = s_scan.tot[parity][0][1];
            for (int w = 1; w < WPT; ++w) {
                if (w == wave) {
                    e0 = __builtin_fma(pa, carry, pb);
                    eps = __builtin_fma(a, e0, bb);
                }
                const double ta = s_scan.tot[parity][w][0], tbv = s_scan.tot[parity][w][1];
                pb = __builtin_fma(ta, pb, tbv);
                pa = ta * pa;
            }
            carry = __builtin_fma(pa, carry, pb);
            parity ^= 1;
            em1 = __shfl_up(eps, 1, 64);
            if (lane == 0) em1 = e0;
        }
        // The likelihood's terms.
        double ef = 0.0, ie = 0.0;
        if (valid) {
            const double cs = (POST && has_post(md)) ? epoch_post(md, sl, s_tile.tab, t, tb + j, k) : 0.0;
            ef = eps - cs;
            ie = rcp_path<PREC>(err2);
            chi = __builtin_fma(ef * ef, ie, chi);
            prod *= err2;
            pexp += __builtin_amdgcn_frexp_exp(prod);
            prod = __builtin_amdgcn_frexp_mant(prod);
            if (md.has_jit) ia.jit = __builtin_fma(sl.jit[k], ef * ef * ie * ie - ie, ia.jit);
            if (POST && has_post(md)) post_accumulate<EXT>(md, sl, s_tile.tab, t, tb + j, k, ef * ie, wacc);
        }
        // dlnL/dmu = wB + wA Lambda; lambda = lB + lA Lambda; (tA, tB) the tile's map.
        double wB = 0.0, wA = 0.0, tA = 0.0, tB = 0.0;
        bool needA = false;
        if (!ma) {
            wB = ef * ie;
            ia.off += wB;
        } else {
            double u = 0.0, amul = 1.0;
            if (valid) {
                u = -ef * ie;
                const int n1 = tb + j + 1;     // the next epoch's multiplier
                amul = 0.0;
                if (n1 < md.N) {
                    const bool in_tile = j + 1 < tn;
                    const int k1 = in_tile ? s_tile.ins[j + 1] : md.ins[n1];
                    const double c1 = in_tile ? s_tile.dt[j + 1] : md.dt[n1];
                    amul = ma_mult(md, sl, k1, c1);
                }
            }
            double sa = amul, sb = u;
            wave_affine_suffix_scan(sa, sb, lane);
            if (lane == 0) { s_g1.suf[sparity][wave][0] = sa; s_g1.suf[sparity][wave][1] = sb; }
            __syncthreads();
            // (pa, pb): the maps of waves WPT - 1 down to w composed
            double lB = sb, lA = sa;
            double pa = s_g1.suf[sparity][WPT - 1][0], pb = s_g1.suf[sparity][WPT - 1][1];
            for (int w = WPT - 2; w >= 0; --w) {
                if (w == wave) {
                    lB = __builtin_fma(sa, pb, sb);
                    lA = sa * pa;
                }
                const double ta = s_g1.suf[sparity][w][0], tbv = s_g1.suf[sparity][w][1];
                pb = __builtin_fma(ta, pb, tbv);
                pa = ta * pa;
            }
            tA = pa;
            tB = pb;
            sparity ^= 1;
            if (valid) {
                wB = -lB;
                wA = -lA;
            }
            // A wave whose every |A_i| is below 2^-100 adds nothing measurable
            // through Lambda (the terms are under 1e-17 of any gradient the
            // data can give): it reduces the B parts alone.  With |a| <= 1/3,
            // every wave of a tile but the last is such a wave.
            needA = __any(valid && __builtin_fabs(lA) > 0x1p-100);
            // the sums this wave already holds in LDS move to the next tile
            for (int e = lane; e < GW::COS; e += 64) {
                const double R0 = wrst[e];
                wacc[e] = __builtin_fma(tB, R0, wacc[e]);
                wrst[e] = tA * R0;
            }
            // this lane's sums: the epoch's terms, then the move to the next tile
            double hc = 0.0, ht = 0.0;      // d a_j / d (c, tau) times eps_{j-1}
            if (valid) {
                const double col = s_tile.dt[j];
                const bool start = md.ma_wfixed ? col == 0.0 : col < 0.0;
                if (!start) {
                    const double wgt = md.ma_wfixed ? col : exp(-col / sl.ma_tau[k]);
                    hc = em1 * -wgt;
                    if (!md.ma_wfixed) {
                        const double tau = sl.ma_tau[k];
                        ht = em1 * -(sl.ma_c[k] * wgt * col / (tau * tau));
                    }
                }
            }
            ia.off += __builtin_fma(tB, ia.offR, wB);
            ia.offR = __builtin_fma(tA, ia.offR, wA);
            ia.mac += __builtin_fma(tB, ia.macR, lB * hc);
            ia.macR = __builtin_fma(tA, ia.macR, lA * hc);
            ia.mat += __builtin_fma(tB, ia.matR, lB * ht);
            ia.matR = __builtin_fma(tA, ia.matR, lA * ht);
        }
        // Stage 2: the planets one at a time, their derivatives from the
        // stage-1 (sin E, cos E); one wave reduction per planet per tile.
#pragma unroll
        for (int p = 0; p < KMAX; ++p) {
            if (p < md.K) {
                const PlanetC& pc = sl.planet[p];
                double f[5];    // every lane (see stage 1)
                {
                    const double e = pc.e;
                    const double rD = rcp_path<PREC>(__builtin_fma(e, -cE[p], 1.0));
                    const double uu = cE[p] - e, vv = pc.bq * sE[p];
                    const double sfw = __builtin_fma(vv, pc.cw, uu * pc.sw) * rD;    // sin(f + w)
                    const double cfw = __builtin_fma(uu, pc.cw, -(vv * pc.sw)) * rD; // cos(f + w)
                    const double Ks = pc.K * sfw;
                    const double fM = -(Ks * pc.bq) * (rD * rD);
                    f[0] = fM * (tp ? t - pc.tp : t);
                    f[1] = fM;
                    f[2] = __builtin_fma(-Ks, sE[p] * rD * __builtin_fma(pc.bq, rD, pc.ibq), pc.K * pc.cw);
                    f[3] = -pc.K * __builtin_fma(e, pc.sw, sfw);
                    f[4] = __builtin_fma(e, pc.cw, cfw);
                }
                // the B parts (one reduction for both cases), then with the MA
                // the A parts, only where some lane's A is not negligible
                const double b5[5] = {wB * f[0], wB * f[1], wB * f[2], wB * f[3], wB * f[4]};
                const double sb = wave_reduce5(b5, lane);
                double sa = 0.0;
                if (ma && needA) {
                    const double a5[5] = {wA * f[0], wA * f[1], wA * f[2], wA * f[3], wA * f[4]};
                    sa = wave_reduce5(a5, lane);
                }
                if (ri5 >= 0) {
                    const int e = 5 * p + ri5;
                    if (ma) {
                        const double R0 = wnatR[e];
                        wnat[e] += __builtin_fma(tB, R0, sb);
                        wnatR[e] = __builtin_fma(tA, R0, sa);
                    } else {
                        wnat[e] += sb;
                    }
                }
            }
        }
        if (md.acc) {
            double xq[AMAX] = {0.0, 0.0, 0.0, 0.0};
            if (valid) {
                const double x = t - md.t0;
                double xp = x;
#pragma unroll
                for (int q = AMAX - 1; q >= 0; --q) {
                    if (q < md.acc) {
                        xq[q] = xp;
                        xp *= x;
                    }
                }
            }
            const double b4[AMAX] = {wB * xq[0], wB * xq[1], wB * xq[2], wB * xq[3]};
            if (!ma) {
                wave_reduce_add4(b4, md.acc, wnat + WN_ACC, ri4);
            } else {
                const double sb = wave_reduce4(b4);
                double sa = 0.0;
                if (needA) {
                    const double a4[AMAX] = {wA * xq[0], wA * xq[1], wA * xq[2], wA * xq[3]};
                    sa = wave_reduce4(a4);
                }
                if (ri4 >= 0 && ri4 < md.acc) {
                    const int e = WN_ACC + ri4;
                    const double R0 = wnatR[e];
                    wnat[e] += __builtin_fma(tB, R0, sb);
                    wnatR[e] = __builtin_fma(tA, R0, sa);
                }
            }
        }
    }
    ia.flush_ins<EXT>(wacc, wrst);
    ia.flush_ma<EXT>(wacc, wrst);
    double ldet = log(prod) + (double)pexp * LN2;
    chi = wave_sum(chi);
    ldet = wave_sum(ldet);
    if (lane == 0) { s_scan.red[wave][0] = chi; s_scan.red[wave][1] = ldet; }
    __syncthreads();
    double c2 = 0.0, l2 = 0.0;
    for (int w = 0; w < WPT; ++w) { c2 += s_scan.red[w][0]; l2 += s_scan.red[w][1]; }
    const double lnl_v = -0.5 * (c2 + l2) + md.lnl_const;
    grad_finish<EXT, WPT>(md, sl, gl, s_gw, true, wave, wave, lane, lnl_v, b, out_lnl, out_grad, nonfinite, stale);
}

// Each model runs the instance with only the features it uses: the terms
// added after the MA cost scalar registers, MA orders above 1 vector
// registers.
template <int PREC, int WPT>
static void launch_grad_t(const ModelDev& md, const double* th, int64_t B, int64_t ld, double* lnl, double* grad,
                          unsigned long long* nf, double* scratch, int* stale, hipStream_t s) {
    constexpr int SLOTS = NWAVES / WPT;
    const int64_t grid = (B + SLOTS - 1) / SLOTS;
    if (md.ma_active && md.ma_q > 1)
        hipLaunchKernelGGL((k_lnl_grad<PREC, WPT, 2>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, th, B, ld,
                           lnl, grad, nf, scratch, stale);
    else if (has_post(md))
        hipLaunchKernelGGL((k_lnl_grad<PREC, WPT, 1>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, th, B, ld,
                           lnl, grad, nf, scratch, stale);
    else
        hipLaunchKernelGGL((k_lnl_grad<PREC, WPT, 0>), dim3((unsigned)grid), dim3(BLOCK), 0, s, md, th, B, ld,
                           lnl, grad, nf, scratch, stale);
}

// The single-pass kernel runs models whose MA order is at most 1; by
// default only those with at least GRAD1_MIN_N epochs: it keeps one epoch
// per lane, so shorter series would leave lanes idle where the two-pass
// kernel takes fewer waves per vector.  emp_model_set_grad_kernel overrides
// the choice per model.
constexpr int GRAD1_NW = 4;
constexpr int GRAD1_MIN_N = 4 * 64 * GRAD1_NW;

static bool use_grad1(const EmpModel* m) {
    const ModelDev& md = m->dev;
    if (md.ma_active && md.ma_q > 1) return false;
    if (m->grad_kernel == EMP_GRAD_KERNEL_SINGLE_PASS) return true;
    if (m->grad_kernel == EMP_GRAD_KERNEL_TWO_PASS) return false;
    return md.N >= GRAD1_MIN_N;
}

extern "C" int emp_model_set_grad_kernel(EmpModel* m, int32_t which) {
    if (!m) { emp_set_error("null model"); return EMP_EINVAL; }
    if (which != EMP_GRAD_KERNEL_AUTO && which != EMP_GRAD_KERNEL_TWO_PASS && which != EMP_GRAD_KERNEL_SINGLE_PASS) {
        emp_set_error("adjoint kernel must be EMP_GRAD_KERNEL_AUTO, _TWO_PASS or _SINGLE_PASS");
        return EMP_EINVAL;
    }
    if (which == EMP_GRAD_KERNEL_SINGLE_PASS && m->dev.ma_active && m->dev.ma_q > 1) {
        emp_set_error("the single-pass adjoint handles MA orders up to 1 (this model: %d)", m->dev.ma_q);
        return EMP_EUNSUP;
    }
    m->grad_kernel = which;
    return EMP_OK;
}

// Doubles of scratch one launch over B vectors needs: with the moving
// average in the two-pass kernel, one row of eps per vector (B x N); none
// otherwise.
size_t emp_grad_scratch_doubles(const EmpModel* m, int64_t B) {
    return (m->dev.ma_active && !use_grad1(m)) ? (size_t)B * (size_t)m->dev.N : 0;
}

// The MA adjoint's scratch, model-owned, grown (never shrunk) on demand.
static int ensure_grad_scratch(EmpModel* m, int64_t B) {
    const size_t need = emp_grad_scratch_doubles(m, B);
    if (!need) return EMP_OK;
    if (m->grad_scratch_cap >= need) return EMP_OK;
    if (m->d_grad_scratch) {
        EMP_HIP(hipDeviceSynchronize());
        (void)hipFree(m->d_grad_scratch);
        m->d_grad_scratch = nullptr;
        m->grad_scratch_cap = 0;
    }
    EMP_HIP(hipMalloc((void**)&m->d_grad_scratch, need * sizeof(double)));
    m->grad_scratch_cap = need;
    return EMP_OK;
}

// Waves per vector of the adjoint kernel for a batch of B.  Four by default:
// with one or two waves per vector a workgroup holds four or eight slots of
// LDS and the kernel's registers allow three or four waves per SIMD, with
// four both allow four to six (measured on 8192 vectors of 7 planets x 10^4
// epochs: mixed 5.1 ms at 4, 6.8 ms at 2, 6.1 ms at 8; FP64 within 4%).
// Fewer when a vector has fewer than two epochs per lane, more while the
// batch leaves SIMDs short of waves.
int emp_pick_grad_wpt(const EmpModel* m, int64_t B) {
    if (m->wpt_override) return m->wpt_override;
    const int64_t target_waves = (int64_t)m->n_cu * 4 * 6;
    int wpt = 4;
    while (wpt > 1 && 64 * wpt > m->dev.N / 2) wpt /= 2;
    while (wpt < 8 && B * wpt < target_waves && 64 * (wpt * 2) <= m->dev.N / 2) wpt *= 2;
    return wpt;
}

int emp_launch_grad_ws(const EmpModel* m, const double* th, int64_t B, int64_t ld, double* lnl, double* grad,
                       double* eps_scratch, hipStream_t s, int precision, int* stale, int wpt) {
    if (B <= 0) return EMP_OK;
    unsigned long long* nf = m->d_nonfinite;
    if (use_grad1(m)) {
        const ModelDev& md = m->dev;
        if (precision != 0 && precision != 1) { emp_set_error("bad precision"); return EMP_EINVAL; }
#define EMP_CASE1(P, E, T) hipLaunchKernelGGL((k_lnl_grad1<P, E, GRAD1_NW, T>), dim3((unsigned)B), dim3(64 * GRAD1_NW), 0, s, md, th, B, ld, lnl, grad, nf, stale)
#define EMP_CASE1T(P, E) { if (model_tp(md)) EMP_CASE1(P, E, true); else EMP_CASE1(P, E, false); }
        if (precision == 0) { if (has_post(md)) EMP_CASE1T(0, 1) else EMP_CASE1T(0, 0) }
        else { if (has_post(md)) EMP_CASE1T(1, 1) else EMP_CASE1T(1, 0) }
#undef EMP_CASE1T
#undef EMP_CASE1
        EMP_HIP(hipGetLastError());
        return EMP_OK;
    }
    if (m->dev.ma_active && !eps_scratch) { emp_set_error("moving-average gradient needs a scratch buffer"); return EMP_EINVAL; }
    if (wpt <= 0) wpt = emp_pick_grad_wpt(m, B);
#define EMP_CASE(P, W) if (precision == P && wpt == W) launch_grad_t<P, W>(m->dev, th, B, ld, lnl, grad, nf, eps_scratch, stale, s); else
    EMP_CASE(0, 1) EMP_CASE(0, 2) EMP_CASE(0, 4) EMP_CASE(0, 8)
    EMP_CASE(1, 1) EMP_CASE(1, 2) EMP_CASE(1, 4) EMP_CASE(1, 8) { emp_set_error("bad precision/wpt"); return EMP_EINVAL; }
#undef EMP_CASE
    EMP_HIP(hipGetLastError());
    return EMP_OK;
}

int emp_launch_grad(EmpModel* m, const double* th, int64_t B, int64_t ld, double* lnl, double* grad, hipStream_t s,
                    int precision) {
    int rc = ensure_grad_scratch(m, B);
    if (rc) return rc;
    return emp_launch_grad_ws(m, th, B, ld, lnl, grad, m->d_grad_scratch, s, precision);
}

extern "C" int emp_lnl_grad_batch(EmpModel* m, const double* theta, int64_t B, int64_t ld, double* lnl,
                                  double* grad, int32_t flags, void* stream) {
    if (!m) { emp_set_error("null model"); return EMP_EINVAL; }
    if (B < 0 || (B > 0 && (!theta || !lnl || !grad))) { emp_set_error("bad batch arguments"); return EMP_EINVAL; }
    if (ld < m->desc.ndim) { emp_set_error("ld < ndim"); return EMP_EINVAL; }
    if (emp_fault_injected()) return EMP_EFAULT_INJECTED;
    if (B == 0) return EMP_OK;
    hipStream_t s = (hipStream_t)stream;
    const int prec = (flags & EMP_GRAD_MIXED) ? EMP_MIXED : EMP_FP64;
    if (!(flags & EMP_HOST_POINTERS)) return emp_launch_grad(m, theta, B, ld, lnl, grad, s, prec);
    const size_t nin = (size_t)B * ld, D = (size_t)m->desc.ndim;
    double *din = nullptr, *dl = nullptr, *dg = nullptr;
    EMP_HIP(hipMalloc(&din, nin * 8));
    EMP_HIP(hipMalloc(&dl, (size_t)B * 8));
    EMP_HIP(hipMalloc(&dg, std::max<size_t>((size_t)B * D, 1) * 8));
    int rc = EMP_OK;
    if (hipMemcpyAsync(din, theta, nin * 8, hipMemcpyHostToDevice, s) != hipSuccess) rc = EMP_EHIP;
    if (!rc) rc = emp_launch_grad(m, din, B, ld, dl, dg, s, prec);
    if (!rc && hipMemcpyAsync(lnl, dl, (size_t)B * 8, hipMemcpyDeviceToHost, s) != hipSuccess) rc = EMP_EHIP;
    if (!rc && D && hipMemcpyAsync(grad, dg, (size_t)B * D * 8, hipMemcpyDeviceToHost, s) != hipSuccess) rc = EMP_EHIP;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = EMP_EHIP;
    if (rc == EMP_EHIP) emp_set_error("HIP error in emp_lnl_grad_batch host path");
    (void)hipFree(din);
    (void)hipFree(dl);
    (void)hipFree(dg);
    return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free variant of emp_lnl_grad_batch (device pointers).  The
// moving-average scratch is `scratch`, a static device array that every
// device the library runs on holds until the process ends, instead of model
// memory grown with hipMalloc.  Host code never passes the array's name to a
// kernel: it stands for the host-side shadow, not the device address.  The
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current, and is kept in scratch_dev[].
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[EMP_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_dev[MAX_DEVICES] = {};
}  // namespace

// Device address of this device's copy of `scratch` (the current device).
static int scratch_address(int dev, double** out) {
    if (dev < 0 || dev >= MAX_DEVICES) { emp_set_error("device %d outside [0, %d)", dev, MAX_DEVICES); return EMP_EINVAL; }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return emp_hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) { emp_set_error("hipGetSymbolAddress(scratch) returned a null address"); return EMP_EHIP; }
        scratch_dev[dev] = (double*)p;
    }
    *out = scratch_dev[dev];
    return EMP_OK;
}

extern "C" int emp_lnl_grad_batch_noalloc(EmpModel* m, const double* theta, int64_t B, int64_t ld, double* lnl,
                                          double* grad, int32_t flags, void* stream) {
    if (!m) { emp_set_error("null model"); return EMP_EINVAL; }
    if (B < 0 || (B > 0 && (!theta || !lnl || !grad))) { emp_set_error("bad batch arguments"); return EMP_EINVAL; }
    if (ld < m->desc.ndim) { emp_set_error("ld < ndim"); return EMP_EINVAL; }
    if (flags & EMP_HOST_POINTERS) {
        emp_set_error("the allocation-free entry point takes device pointers only (EMP_HOST_POINTERS stages through allocated memory)");
        return EMP_EINVAL;
    }
    if (flags & ~EMP_GRAD_MIXED) { emp_set_error("unknown flags 0x%x", (unsigned)flags); return EMP_EINVAL; }
    const size_t need = emp_grad_scratch_doubles(m, B);
    if (need > EMP_NOALLOC_SCRATCH_DOUBLES) {
        emp_set_error("batch of %lld vectors needs %zu doubles of scratch; the allocation-free entry point holds %zu",
                      (long long)B, need, (size_t)EMP_NOALLOC_SCRATCH_DOUBLES);
        return EMP_EINVAL;
    }
    if (emp_fault_injected()) return EMP_EFAULT_INJECTED;
    if (B == 0) return EMP_OK;
    int dev = -1;
    EMP_HIP(hipGetDevice(&dev));
    if (dev != m->device) { emp_set_error("model lives on device %d, device %d is current", m->device, dev); return EMP_EINVAL; }
    double* ws = nullptr;
    if (need) {
        const int rc = scratch_address(dev, &ws);
        if (rc) return rc;
    }
    const int prec = (flags & EMP_GRAD_MIXED) ? EMP_MIXED : EMP_FP64;
    return emp_launch_grad_ws(m, theta, B, ld, lnl, grad, ws, (hipStream_t)stream, prec);
}
