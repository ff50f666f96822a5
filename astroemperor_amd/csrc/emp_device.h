// Device-side numerics for the Keplerian RV likelihood on gfx950 (CDNA4).
//
// Everything here is wave64 code: one parameter vector (theta) is owned by a
// group of 1..8 wavefronts, and the epochs are strided over the group's
// lanes.  Per-theta quantities are computed redundantly by every lane and
// kept in LDS; per-epoch quantities live in VGPRs.
//
// Kepler solve.  The reference calls kepler.solve (kepler.py 0.0.7) inside
// support/models/kep0*.model.  Here the root of E - e sin E = M is found
// with Markley's non-iterative method (F. L. Markley, "Kepler Equation
// Solver", Celest. Mech. Dyn. Astron. 63, 101-111, 1995): a cubic starter
// with ~4e-4 rad worst-case error followed by one fifth-order Householder
// correction, which reaches machine precision for 0 <= e < 1.
//
// Projection.  The templates compute f = 2 atan(sqrt((1+e)/(1-e)) tan(E/2))
// and A (cos(f + w) + e cos w).  With cos f = (cos E - e)/(1 - e cos E) and
// sin f = sqrt(1-e^2) sin E/(1 - e cos E) this is the same function
//     mu = [A (1 - e^2) cos w cos E - A sqrt(1 - e^2) sin w sin E] / (1 - e cos E),
// which needs one reciprocal per planet-epoch and no atan/tan.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emp {

constexpr double PI_D      = 3.141592653589793116;        // nearest double to pi
constexpr double TWO_PI_HI = 6.283185307179586232;        // nearest double to 2 pi
constexpr double TWO_PI_LO = 2.4492935982947064e-16;      // 2 pi - TWO_PI_HI
constexpr double INV_TWO_PI = 0.15915494309189535;
// Markley's alpha(M, e) = (3 pi^2 + 1.6 pi (pi - |M|)/(1 + e)) / (pi^2 - 6)
constexpr double MK_A0 = 3.0 * PI_D * PI_D / (PI_D * PI_D - 6.0);
constexpr double MK_A1 = 1.6 * PI_D / (PI_D * PI_D - 6.0);

// ----------------------------------------------------------------------------
// Wave helpers (wave64).
// ----------------------------------------------------------------------------
// Cross-lane moves by DPP (register to register, no LDS round trip).
// update_dpp keeps `old` in lanes whose source is outside the row or whose
// row is masked off.  Controls: row_shr:n = 0x110 + n (within rows of 16
// lanes), row_bcast:15 = 0x142 (lane 15 of each row to the next row),
// row_bcast:31 = 0x143 (lane 31 to rows 2 and 3).  Callers run with every
// lane of the wave active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double old, double src) {
    const long long s = __double_as_longlong(src), o = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The same for controls under which every lane reads a lane of its own row
// (row_ror, row_half_mirror = 0x141, quad_perm) with every row enabled: no
// lane keeps an old value, so none is passed (and no register is cleared
// for one).
template <int CTRL>
__device__ __forceinline__ double dpp_mov_f64(double src) {
    const long long s = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_mov_dpp((int)s, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(s >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// row_shr:n (0x110 + n) or row_shl:n (0x100 + n) with every row enabled
// and zero in the lanes that have no source (the first, or last, n of each
// row): update_dpp with `old` = 0.0, without a register cleared to hold the
// 0.0.
template <int CTRL>
__device__ __forceinline__ double dpp_shr_zero_f64(double src) {
    static_assert((CTRL > 0x110 && CTRL < 0x120) || (CTRL > 0x100 && CTRL < 0x110), "row_shr or row_shl only");
    const long long s = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_mov_dpp((int)s, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(s >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// row_shr:n or row_shl:n with every row enabled and 1.0 in the lanes that
// have no source: the low word of 1.0 is zero (bound_ctrl zero fill), only
// the high word needs an old value.
template <int CTRL>
__device__ __forceinline__ double dpp_shr_one_f64(double src) {
    static_assert((CTRL > 0x110 && CTRL < 0x120) || (CTRL > 0x100 && CTRL < 0x110), "row_shr or row_shl only");
    const long long s = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_mov_dpp((int)s, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0x3ff00000, (int)(s >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Value of lane k (uniform k) in every lane.
__device__ __forceinline__ double readlane_f64(double v, int k) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, k);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the wave (inclusive prefix sum by DPP, then lane 63's total),
// the same value in every lane.  Whole wave active.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0x111, 0xf>(0.0, v);
    v += dpp_f64<0x112, 0xf>(0.0, v);
    v += dpp_f64<0x114, 0xf>(0.0, v);
    v += dpp_f64<0x118, 0xf>(0.0, v);
    v += dpp_f64<0x142, 0xa>(0.0, v);
    v += dpp_f64<0x143, 0xc>(0.0, v);
    return readlane_f64(v, 63);
}

__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// FP64 reciprocal for arguments in a normal range: hardware estimate plus two
// Newton-Raphson refinements (each doubles the correct bits).
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// One refinement only (mixed path): the hardware estimate's error squared,
// below 1e-13 relative.
__device__ __forceinline__ double rcp64_1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// Reduce an FP64 angle to [-pi, pi] with a two-term (Cody-Waite) 2 pi.
__device__ __forceinline__ double reduce_2pi(double M) {
    double n = __builtin_rint(M * INV_TWO_PI);
    double r = __builtin_fma(-n, TWO_PI_HI, M);
    return __builtin_fma(-n, TWO_PI_LO, r);
}

// ----------------------------------------------------------------------------
// Per-planet constants, computed once per theta (per planet) and kept in LDS.
// ----------------------------------------------------------------------------
struct PlanetC {
    double freq;     // 2 pi / P
    double phase;    // M = freq * t + phase         (tp_mode == 0)
    double tp;       // M = freq * (t - tp)          (tp_mode == 1)
    double e;        // eccentricity
    double c1;       // A (1 - e^2) cos w
    double c2;       // -A sqrt(1 - e^2) sin w
    double mk_a1;    // MK_A1 / (1 + e)
    double mk_d0;    // 3 (1 - e)
    double K;        // amplitude (for gradients)
    double cw, sw;   // cos w, sin w (for gradients)
    double bq, ibq;  // sqrt(1 - e^2) and its reciprocal (for gradients)
    int tp_mode;
    int bad;         // NaN-producing parameter set (reference would give NaN)
    int ktab;        // mixed path: this planet's per-vector Kepler table is built and used
};

// Builds PlanetC from the 5 block parameters, following the reference's
// per-parameterisation template (support/models/kep0p.model, Appendix
// "derived" column): P from lnP for 2/6/7, K = As^2+Ac^2 and
// phase = arccos(Ac/sqrt K) for 2, e = S^2+C^2 and w = arccos(C/sqrt e)
// (2 pi - w when S < 0) for 1/2/4/7 with w forced to 0 below 1e-6 (1/7) or
// 1e-5 (2/4) unless `use_c` (clib/ckep0*.model has no guard).
__device__ inline PlanetC planet_setup(const double* th, int param, int use_c) {
    PlanetC pc;
    double per, A, e, cw = 1.0, sw = 0.0;
    pc.tp_mode = 0;
    pc.phase = 0.0;
    pc.tp = 0.0;
    pc.bad = 0;
    pc.ktab = 0;
    bool sc = (param == 1 || param == 2 || param == 4 || param == 7);
    if (param == 2 || param == 6 || param == 7) per = exp(th[0]); else per = th[0];
    if (param == 2) {
        A = th[1] * th[1] + th[2] * th[2];
        double ph = acos(th[2] / sqrt(A));
        if (th[1] < 0) ph = 2.0 * PI_D - ph;
        pc.phase = ph;
    } else {
        A = th[1];
        if (param == 3 || param == 4) { pc.tp_mode = 1; pc.tp = th[2]; }
        else pc.phase = th[2];
    }
    if (sc) {
        e = th[3] * th[3] + th[4] * th[4];
        double thr = (param == 2 || param == 4) ? 1e-5 : 1e-6;
        if (!use_c && e < thr) {
            cw = 1.0; sw = 0.0;
        } else {
            double se = sqrt(e);
            cw = th[4] / se;
            sw = th[3] / se;
            if (!(cw >= -1.0 && cw <= 1.0)) pc.bad = 1;   // arccos domain error -> NaN
        }
    } else {
        e = th[3];
        sincos(th[4], &sw, &cw);
    }
    if (!(e >= 0.0 && e < 1.0)) pc.bad = 1;
    double q = sqrt((1.0 - e) * (1.0 + e));
    pc.freq = 2.0 * PI_D / per;
    pc.e = e;
    pc.c1 = A * (1.0 - e) * (1.0 + e) * cw;
    pc.c2 = -A * q * sw;
    pc.mk_a1 = MK_A1 / (1.0 + e);
    pc.mk_d0 = 3.0 * (1.0 - e);
    pc.K = A;
    pc.cw = cw;
    pc.sw = sw;
    pc.bq = q;
    pc.ibq = 1.0 / q;
    if (pc.bad) {   // the reference evaluates to NaN here; make that explicit
        pc.c1 = __builtin_nan("");
        pc.c2 = __builtin_nan("");
    }
    return pc;
}

// tp: the model's planets are parameterised by the time of periastron
// (model_tp), a kernel-argument scalar, so the choice is a scalar branch.
__device__ __forceinline__ double planet_M(const PlanetC& pc, double t, bool tp) {
    return tp ? (t - pc.tp) * pc.freq : pc.freq * t + pc.phase;
}

// ----------------------------------------------------------------------------
// FP64 sin/cos on [0, pi] from a table of (sin, cos) at k pi / SC_N held in
// LDS: x = k h + d with |d| <= h/2 = pi/512, and
//   sin x = s_k cos d + c_k sin d,  cos x = c_k cos d - s_k sin d,
//   sin d = d - d^3/6 + d^5/120,    cos d = 1 - d^2/2 + d^4/24 - d^6/720
// (series truncation below 1e-19).  The table is filled on the host in
// extended precision.  Callers clamp the argument into [0, pi].
// ----------------------------------------------------------------------------
constexpr int SC_N = 256;
constexpr double SC_H = PI_D / SC_N;                        // exact: power-of-two scaling
constexpr double SC_H_LO = 1.2246467991473532e-16 / SC_N;  // (pi - PI_D) / SC_N
constexpr double SC_INV_H = SC_N / PI_D;

struct SinCosTab { double2 v[SC_N + 1]; };

// floor(v + 1/2) as an integer in one instruction (v_cvt_rpi_i32_f32; NaN
// gives 0).
__device__ __forceinline__ int round_half_up(float v) {
    int k;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(k) : "v"(v));
    return k;
}

// The index is x / h rounded half up from x's FP32 value (a node either
// side of a tie, |d| within h/2 + 2e-7 rad, where the series is exact).
__device__ __forceinline__ void sincos_tab(const SinCosTab& tb, double x, double& s, double& c) {
    const int k = round_half_up((float)x * (float)SC_INV_H);
    const double kd = (double)k;
    const double d = __builtin_fma(-kd, SC_H_LO, __builtin_fma(-kd, SC_H, x));
    const double2 sc = tb.v[k];
    const double d2 = d * d;
    const double sd = __builtin_fma(d * d2, __builtin_fma(d2, 1.0 / 120.0, -1.0 / 6.0), d);
    const double cd = __builtin_fma(d2, __builtin_fma(d2, __builtin_fma(d2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
    s = __builtin_fma(sc.x, cd, sc.y * sd);
    c = __builtin_fma(sc.y, cd, -sc.x * sd);
}

// Mixed path: sin/cos of E = (double)Ef for an FP32 estimate Ef in
// [-h/2, pi + h/2].  The table index is Ef / h rounded half up (|d| <= h/2
// either way at a tie); d = E - k h in one FMA (the double h = pi / SC_N is
// within 4.8e-19 of pi / SC_N, so d within k 4.8e-19 <= 1.2e-16, largest
// near pi); sin d to d^3 (d^5 / 120 < 7e-14 dropped) and cos d to d^4
// (d^6 / 720 < 8e-17 dropped), or, with D5, both to the sincos_tab degree
// (d^5 and d^6).
template <bool D5 = false>
__device__ __forceinline__ void sincos_tab_f(const SinCosTab& tb, float Ef, double E, double& s, double& c) {
    const int k = round_half_up(Ef * (float)SC_INV_H);
    const double d = __builtin_fma(-(double)k, SC_H, E);
    const double2 sc = tb.v[k];
    const double d2 = d * d;
    const double sd = D5 ? __builtin_fma(d * d2, __builtin_fma(d2, 1.0 / 120.0, -1.0 / 6.0), d)
                         : __builtin_fma(d * d2, -1.0 / 6.0, d);
    const double cd = D5 ? __builtin_fma(d2, __builtin_fma(d2, __builtin_fma(d2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0)
                         : __builtin_fma(d2, __builtin_fma(d2, 1.0 / 24.0, -0.5), 1.0);
    s = __builtin_fma(sc.x, cd, sc.y * sd);
    c = __builtin_fma(sc.y, cd, -sc.x * sd);
}

// x^(2/3) for x >= 0 to FP32 accuracy (hardware log2 / exp2, ~3e-7
// relative): the Markley starter's cube root (see kepler_fp64_E).
__device__ __forceinline__ double pow23_start(double x) {
    const float xf = (float)x;
    return xf > 0.0f ? (double)__builtin_amdgcn_exp2f(0.6666667f * __builtin_amdgcn_logf(xf)) : 0.0;
}

// ----------------------------------------------------------------------------
// FP64 Kepler solve (Markley 1995).  Input: M reduced to [-pi, pi].
// Output: sin E, cos E (and E for M >= 0 through kepler_fp64_E).
// ----------------------------------------------------------------------------
__device__ __forceinline__ void kepler_fp64_E(double M, const PlanetC& pc, const SinCosTab& tb, double& E,
                                              double& sE, double& cE) {
    const double e = pc.e;
    const double Ma = __builtin_fabs(M);
    const double ome = 1.0 - e;
    const double alpha = __builtin_fma(pc.mk_a1, PI_D - Ma, MK_A0);
    const double d = __builtin_fma(alpha, e, pc.mk_d0);
    const double ad = alpha * d;
    const double M2 = Ma * Ma;
    const double q = __builtin_fma(2.0 * ad, ome, -M2);
    const double r = __builtin_fma(3.0 * ad * (d - ome), Ma, M2 * Ma);
    const double q2 = q * q;
    // The starter's arithmetic is FP64 (q and r cancel near periastron at
    // high e); its square root, cube root and reciprocals are the hardware's
    // FP32-accurate ones (about 3e-7 relative): the start is up to 4e-4 rad
    // from the root in any case, and the fifth-order correction below
    // leaves (4e-4)^5 of that.
    const double w = pow23_start(__builtin_fabs(r) + (double)__builtin_amdgcn_sqrtf((float)__builtin_fma(q2, q, r * r)));
    double E1 = __builtin_fma(2.0 * r * w, __builtin_amdgcn_rcp(__builtin_fma(w, w + q, q2)), Ma) * __builtin_amdgcn_rcp(d);
    E1 = __builtin_fmin(__builtin_fmax(E1, 0.0), PI_D);
    double s, c;
    sincos_tab(tb, E1, s, c);
    // Fifth-order Householder correction (Markley eqs. 20-22).
    const double f0 = E1 - __builtin_fma(e, s, Ma);
    const double f1 = __builtin_fma(-e, c, 1.0);
    const double f2 = e * s;
    const double f3 = 1.0 - f1;
    // The corrections are below 4e-4 rad: reciprocals to 2e-15 (one
    // refinement) leave them exact to 1e-18.
    // dn only seeds d3's denominator (0.5 dn f2): the hardware reciprocal's
    // 4.6e-8 relative error reaches d5 below 1e-18 rad, up to e = 0.99.
    const double dn = -f0 * __builtin_amdgcn_rcp(f1);
    const double d3 = -f0 * rcp64_1(__builtin_fma(0.5 * dn, f2, f1));
    const double d4 = -f0 * rcp64_1(f1 + 0.5 * d3 * f2 + d3 * d3 * f3 * (1.0 / 6.0));
    const double d5 = -f0 * rcp64_1(f1 + 0.5 * d4 * f2 + d4 * d4 * f3 * (1.0 / 6.0)
                                  - d4 * d4 * d4 * f2 * (1.0 / 24.0));
    // Rotate (s, c) by d5; |d5| < 1e-3 so three Taylor terms are exact to 1e-19.
    const double dd = d5 * d5;
    const double sd = d5 * __builtin_fma(-dd, 1.0 / 6.0, 1.0);
    const double cd = __builtin_fma(dd, __builtin_fma(dd, 1.0 / 24.0, -0.5), 1.0);
    sE = __builtin_fma(s, cd, c * sd);
    cE = __builtin_fma(c, cd, -s * sd);
    E = E1 + d5;
    if (M < 0.0) sE = -sE;
}

__device__ __forceinline__ void kepler_fp64(double M, const PlanetC& pc, const SinCosTab& tb, double& sE,
                                            double& cE) {
    double E;
    kepler_fp64_E(M, pc, tb, E, sE, cE);
}

// ----------------------------------------------------------------------------
// Per-vector Kepler tables (mixed path).  For one parameter vector the
// eccentricity of a planet is fixed, so E(M) on [0, pi] is tabulated once at
// KT_N nodes M_j = j h (h = pi / (KT_N - 1)) as (E_j, h dE/dM_j) in FP32;
// each epoch then takes the cubic Hermite interpolant between its two nodes
// and one FP64 step.  The table is kept in one of two forms: the nodes
// (float2, 512 B per planet), or the interpolant's coefficients on each of
// the KT_N - 1 intervals and, as interval KT_N - 1, its linear extension
// past pi, where rounding can put |M| (float4, 1 KiB per planet, five fewer
// operations per epoch).  The FP64 step is: a
// Newton step up to e = KT_ENEWTON (start error <= 6.1e-7 rad, after the
// step <= 6.1e-14), a Halley step up to e = KT_EH64 (start error <= 3.5e-5
// rad, after the step <= 1.5e-14; a Newton step would leave 4.7e-10).
// Above KT_EH64 the start error grows fast (1.4e-4 at e = 0.8, 7e-3 at
// 0.95): up to KT_EMAX an FP32 Halley step comes first (then <= 7.3e-7 rad,
// <= 1.3e-15 after the FP64 Halley step).  Above KT_EMAX (start error
// 1.3e-2 at e = 0.96, 4.6e-2 at 0.97) the FP32 Markley starter (with the
// FP32 Halley step, error <= 5e-7 rad to e = 0.99) and one FP64 Halley step
// are used instead.  These bounds are the maxima over 2e6 mean anomalies on
// [0, pi] of an FP32 emulation of the interpolant and the FP32 step.
// ----------------------------------------------------------------------------
constexpr int KT_N = 64;
constexpr double KT_H = PI_D / (KT_N - 1);
constexpr double KT_ENEWTON = 0.5;
constexpr double KT_EH64 = 0.75;
constexpr double KT_EMAX = 0.95;
// A planet's Kepler path on the mixed path, as flags: the start from the
// table (else the FP32 Markley starter), an FP32 Halley step after it, and
// an FP64 Halley step (else an FP64 Newton step).  Markley: H32 | H64;
// table: up to KT_ENEWTON TABLE, up to KT_EH64 TABLE | H64, up to KT_EMAX
// TABLE | H32 | H64.
constexpr int KT_F_TABLE = 1, KT_F_H32 = 2, KT_F_H64 = 4;
__device__ __forceinline__ int kt_flags(bool table, double e) {
    if (!table) return KT_F_H32 | KT_F_H64;
    return KT_F_TABLE | (e > KT_EH64 ? KT_F_H32 : 0) | (e > KT_ENEWTON ? KT_F_H64 : 0);
}
// Planet p's flags from SlotLDS::ktab_mask (four bits per planet).
__device__ __forceinline__ int kt_mode(int kmask, int p) { return (int)(((unsigned)kmask >> (4 * p)) & 7u); }
// Building a table costs about as much as 64 FP64 planet-epochs, and saves
// the FP32 Markley starter's ~40 extra operations per epoch: it pays from
// about 256 epochs on.
constexpr int KT_MIN_EPOCHS = 256;

// The cubic Hermite interpolant on [M_j, M_j+1] in Horner form: with
// d = E_j+1 - E_j and the scaled slopes s0 = h E'_j, s1 = h E'_j+1,
//   E(x) = E_j + x (s0 + x (3d - 2 s0 - s1 + x (s0 + s1 - 2d))),
// x = M / h - j in [0, 1].  Node form: kt[j] = (E_j, s_j).
__device__ __forceinline__ float ktab_start(const float2* kt, float Mf) {
    constexpr float SCALE = (float)((KT_N - 1) / PI_D);
    const int j = min((int)(Mf * SCALE), KT_N - 2);
    const float x = __builtin_fmaf(Mf, SCALE, -(float)j);
    const float2 a = kt[j], b = kt[j + 1];
    const float d = b.x - a.x, ss = a.y + b.y;
    const float c3 = __builtin_fmaf(-2.0f, d, ss);              // s0 + s1 - 2d
    const float c2 = __builtin_fmaf(3.0f, d, -(a.y + ss));      // 3d - 2 s0 - s1
    return __builtin_fmaf(x, __builtin_fmaf(x, __builtin_fmaf(x, c3, c2), a.y), a.x);
}

// Interval form: kt[j] = (E_j, s0, 3d - 2 s0 - s1, s0 + s1 - 2d), the same
// FP32 values the node form computes per epoch; x is the fraction of the
// rounded M / h (one instruction; the node form keeps M / h - j unrounded,
// a difference of at most 9.5e-8 rad in M, below the 1.2e-7 of M's own
// rounding to FP32).  j is clamped to the table, which only matters for a
// mean anomaly that is not finite (x is then not M / h - j: the start is
// wrong, but stays a finite value in range, or NaN).
__device__ __forceinline__ float ktab_start(const float4* kt, float Mf) {
    constexpr float SCALE = (float)((KT_N - 1) / PI_D);
    const float u = Mf * SCALE;
    const int j = min((int)u, KT_N - 1);
    const float x = __builtin_amdgcn_fractf(u);
    const float4 c = kt[j];
    return __builtin_fmaf(x, __builtin_fmaf(x, __builtin_fmaf(x, c.w, c.z), c.y), c.x);
}

// ----------------------------------------------------------------------------
// Mixed-precision Kepler solve.  FP32 buys the iteration count: the starter
// comes from the per-vector FP32 table (e <= KT_EMAX) or, above it, from
// Markley's starter in FP32 (hardware log2/exp2/rcp/sqrt); the Markley start
// and the table start above KT_EH64 take one FP32 Halley step with the
// hardware sin/cos.  One FP64 step from that root and the FP64 projection
// restore the accuracy the likelihood needs.
// ----------------------------------------------------------------------------
__device__ __forceinline__ float kepler_fp32_markley(float Mf, const PlanetC& pc) {
    const float ef = (float)pc.e;
    const float omef = 1.0f - ef;
    const float alpha = __builtin_fmaf((float)pc.mk_a1, (float)PI_D - Mf, (float)MK_A0);
    const float d = __builtin_fmaf(alpha, ef, (float)pc.mk_d0);
    const float ad = alpha * d;
    const float M2 = Mf * Mf;
    const float q = __builtin_fmaf(2.0f * ad, omef, -M2);
    const float r = __builtin_fmaf(3.0f * ad * (d - omef), Mf, M2 * Mf);
    const float q2 = q * q;
    const float x = __builtin_fabsf(r) + __builtin_amdgcn_sqrtf(__builtin_fmaf(q2, q, r * r));
    // w = x^(2/3) = 2^(2/3 log2 x)
    const float w = __builtin_amdgcn_exp2f(0.66666667f * __builtin_amdgcn_logf(x));
    return (2.0f * r * w * __builtin_amdgcn_rcpf(__builtin_fmaf(w, w + q, q2)) + Mf) * __builtin_amdgcn_rcpf(d);
}

// FP32 Halley step with hardware sin/cos (argument in revolutions).
__device__ __forceinline__ float kepler_fp32_halley(float E, float Mf, const PlanetC& pc) {
    const float ef = (float)pc.e;
    const float rev = E * 0.15915494f;
    const float s = __builtin_amdgcn_sinf(rev), c = __builtin_amdgcn_cosf(rev);
    const float f0 = E - __builtin_fmaf(ef, s, Mf);
    const float f1 = __builtin_fmaf(-ef, c, 1.0f);
    const float dn = -f0 * __builtin_amdgcn_rcpf(f1);
    return E - f0 * __builtin_amdgcn_rcpf(__builtin_fmaf(0.5f * dn, ef * s, f1));
}

__device__ __forceinline__ float kepler_fp32_start(float Mf, const PlanetC& pc) {
    return kepler_fp32_halley(kepler_fp32_markley(Mf, pc), Mf, pc);
}

// Node `lane` of planet pc's table, solved in FP32: the nodes are stored in
// FP32, and FP32 roots leave the interpolant's error where exact nodes put
// it (<= 6.2e-7 rad for e <= 0.5).
__device__ __forceinline__ float2 ktab_node(const PlanetC& pc, int lane) {
    const float E = kepler_fp32_start((float)(lane * KT_H), pc);
    const float c = __builtin_amdgcn_cosf(E * 0.15915494f);
    return make_float2(E, (float)KT_H * __builtin_amdgcn_rcpf(__builtin_fmaf(-(float)pc.e, c, 1.0f)));
}

// Entry `lane` of planet pc's table in either form (see ktab_start): node
// lane, or interval lane from nodes lane and lane + 1, the last lane holding
// the linear extension past pi.  Called by a whole wave (64 lanes, 64 nodes).
__device__ __forceinline__ void ktab_fill(const PlanetC& pc, int lane, float2& out) { out = ktab_node(pc, lane); }
__device__ __forceinline__ void ktab_fill(const PlanetC& pc, int lane, float4& out) {
    const float2 a = ktab_node(pc, lane);
    const float bx = __shfl_down(a.x, 1), by = __shfl_down(a.y, 1);
    if (lane == KT_N - 1) {
        out = make_float4(a.x, a.y, 0.0f, 0.0f);
    } else {
        const float d = bx - a.x, ss = a.y + by;
        out = make_float4(a.x, a.y, __builtin_fmaf(3.0f, d, -(a.y + ss)), __builtin_fmaf(-2.0f, d, ss));
    }
}

// One FP64 step from the FP32 root Ef (Ma = |M| in [0, pi]): Newton from the
// table start at e <= KT_ENEWTON (<= 6.1e-7 rad, so sin E, cos E follow
// from a first-order rotation, neglected d^2 / 2 < 2e-13), otherwise Halley
// with a third-order rotation (from the table start, after the FP32 Halley
// step above KT_EH64, or from the FP32 Markley start and Halley step).
// Ef is within 3.5e-5 rad of the root in [0, pi] when it comes from the
// table, clamped into [0, (float)pi] when it comes from the Markley start,
// or NaN; sincos_tab_f takes it as it is (its index stays in [0, SC_N] for
// any start within h/2 = 6.1e-3 rad of [0, pi], and is 0 for NaN), and the
// step converges from either side of the root.
template <bool HALLEY>
__device__ __forceinline__ void kepler_fp64_step(double Ma, float Ef, double e, const SinCosTab& tb, double& sE,
                                                 double& cE) {
    const double E0 = (double)Ef;
    double s, c;
    // Near periastron at high e the step amplifies an error in sin E0 by
    // 1 / (1 - e cos E0) (up to 20 at e = 0.95): the Halley step takes sin d
    // to d^5; the error of d is largest near pi, where it is not amplified.
    sincos_tab_f<HALLEY>(tb, Ef, E0, s, c);
    // 1 - e c as fma(e, -c, 1) (the same value): the negation then folds into
    // the operand of each use instead of a negated e kept between blocks.
    const double f0 = E0 - __builtin_fma(e, s, Ma);
    const double f1 = __builtin_fma(e, -c, 1.0);
    // The Newton step follows the table start at e <= KT_ENEWTON, at most
    // 6.1e-7 rad from the root: the hardware reciprocal's 4.6e-8 relative
    // error then moves the step by under 3e-14 rad, and its refinement is
    // left out.  The Halley step may start 1e-4 rad off (the Markley start
    // near periastron at high e) and keeps it.
    const double r1 = HALLEY ? rcp64_1(f1) : __builtin_amdgcn_rcp(f1);
    double dl = -f0 * r1;
    if (HALLEY) {
        // The FP32 Markley start (hardware sin/cos) is off by up to ~1e-4 rad
        // near periastron at e -> 1: rotate to third order (d^4 / 24 < 5e-18).
        dl *= __builtin_fma(-0.5 * dl * e * s, r1, 1.0);
        const double dd = dl * dl;
        const double sd = dl * __builtin_fma(dd, -1.0 / 6.0, 1.0);
        const double cd = __builtin_fma(dd, -0.5, 1.0);
        sE = __builtin_fma(c, sd, s * cd);
        cE = __builtin_fma(-s, sd, c * cd);
    } else {
        sE = __builtin_fma(c, dl, s);
        cE = __builtin_fma(-s, dl, c);
    }
}

// (sin E, cos E) of the planet at time t.  PREC 0: FP64 Markley; PREC 1:
// mixed, on the path kmode (kt_mode of the slot's mask, a wave-uniform
// scalar so that each choice is a scalar branch), with `kt` the planet's
// per-vector table (either form).  tp as in planet_M.
template <int PREC, class KTE>
__device__ __forceinline__ void planet_sc(const PlanetC& pc, const SinCosTab& tb, const KTE* kt, double t,
                                          double& sE, double& cE, bool tp, int kmode) {
    // Both paths reduce the same FP64 mean anomaly: the mixed path's M is
    // rounded exactly as the FP64 path's, so its error near periastron at
    // high e (amplified by dE/dM = 1 / (1 - e cos E)) stays the FP64 path's.
    const double M = reduce_2pi(planet_M(pc, t, tp));
    if (PREC == 0) {
        kepler_fp64(M, pc, tb, sE, cE);
        return;
    }
    const double Ma = __builtin_fabs(M);
    const float Mf = (float)Ma;
    float Ef;
    if (kmode & KT_F_TABLE) {
        Ef = ktab_start(kt, Mf);
        if (kmode & KT_F_H32) Ef = kepler_fp32_halley(Ef, Mf, pc);
    } else {
        // the Markley start always takes the FP32 Halley step (kt_flags);
        // near periastron at e -> 1 it may leave [0, pi], so it is clamped
        Ef = __builtin_amdgcn_fmed3f(kepler_fp32_start(Mf, pc), 0.0f, (float)PI_D);
    }
    if (kmode & KT_F_H64) kepler_fp64_step<true>(Ma, Ef, pc.e, tb, sE, cE);
    else kepler_fp64_step<false>(Ma, Ef, pc.e, tb, sE, cE);
    // E(-M) = -E(M), and sin E >= 0 on [0, pi]: the sign of M, one bit insert
    // (a rounding-level negative sE at E ~ 0 or pi comes back positive).
    sE = __builtin_copysign(sE, M);
}

// Reciprocal of the path: two Newton steps (FP64) or one (mixed).
template <int PREC>
__device__ __forceinline__ double rcp_path(double x) {
    return PREC == 0 ? rcp64(x) : rcp64_1(x);
}

// Planet term at time t (see planet_sc).
template <int PREC, class KTE>
__device__ __forceinline__ double planet_mu(const PlanetC& pc, const SinCosTab& tb, const KTE* kt, double t,
                                           bool tp, int kmode) {
    double sE, cE;
    planet_sc<PREC>(pc, tb, kt, t, sE, cE, tp, kmode);
    const double den = __builtin_fma(pc.e, -cE, 1.0);
    return __builtin_fma(pc.c1, cE, pc.c2 * sE) * rcp_path<PREC>(den);
}

// ----------------------------------------------------------------------------
// Counter-based RNG: Philox-4x32-10 (Salmon, Moraes, Dror & Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// ----------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c.x;
        uint64_t p1 = (uint64_t)M1 * c.z;
        U4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Uniform double in (0, 1) from two 32-bit words (53 random bits).
__host__ __device__ inline double u01(uint32_t hi, uint32_t lo) {
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)v + 0.5) * 1.1102230246251565e-16;   // 2^-53
}

// Keyed bijection of [0, n) (4-round Feistel on the enclosing power of four,
// cycle-walking back into range).  Used to pair walkers between adjacent
// temperatures independently of their states.
__host__ __device__ inline uint32_t feistel_perm(uint32_t i, uint32_t n, uint32_t key0, uint32_t key1) {
    uint32_t bits = 2;
    while ((1u << bits) < n) bits += 2;
    const uint32_t half = bits / 2, mask = (1u << half) - 1u;
    uint32_t x = i;
    do {
        uint32_t L = x >> half, R = x & mask;
        for (uint32_t r = 0; r < 4; ++r) {
            uint32_t h = R * 0x9E3779B1u ^ (key0 + r * 0x85EBCA77u);
            h ^= h >> 15; h *= 0x2C1B3C6Du ^ key1; h ^= h >> 12;
            uint32_t nL = R;
            R = (L ^ h) & mask;
            L = nL;
        }
        x = (L << half) | R;
    } while (x >= n);
    return x;
}

}  // namespace emp
