// Internal declarations shared by the HIP translation units of libemp_mi355x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <vector>

#include "../../include/emp_abi.h"
#include "emp_device.h"

namespace emp {

constexpr int MAXD = 128;        // full parameter vector length supported by the kernels
constexpr int KMAX = 8;          // Keplerians per model
constexpr int IMAX = 16;         // instruments per model
constexpr int AMAX = 4;          // acceleration order
constexpr int CMAX = 3;          // cosine terms per model (a sinusoid, a magnetic cycle's two)
constexpr int SAIMAX = 32;       // stellar-activity columns per model
constexpr int BLOCK = 512;       // threads per workgroup (8 waves)
constexpr int NWAVES = BLOCK / 64;
constexpr int MAX_WAVES = 16;    // largest workgroup any kernel launches (1024 threads)
constexpr int TILE = 512;        // epochs staged in LDS per tile
// Prior sequence entries: j >= 0 the prior of (free) full parameter j;
// PSEQ_ECC - b the derived eccentricity prior of Keplerian block b;
// PSEQ_CHECK: return -inf if the sum is -inf (the end of a block's own terms).
constexpr int PSEQ_CHECK = -1;
constexpr int PSEQ_ECC = -2;

// One entry of the prior's summation sequence, resolved on the host: the
// term and where its argument comes from (free indices, or fixed values for
// the S, C of a derived eccentricity), so that a lane goes straight from its
// entry to the walker's row.
struct PseqEntry {
    EmpPriorTerm term;
    double va, vb;        // fixed values of S and C (eccentricity entries whose S or C is fixed)
    int kind;             // 0: parameter prior (fa), 1: eccentricity S^2 + C^2 (fa, fb), 2: -inf check
    int fa, fb;           // free indices, -1 when fixed
    int pad;
};

// Everything a kernel needs to know about a model, passed by value.
struct ModelDev {
    int N, I, K, param, use_c, acc, has_jit;
    int ma_q, ma_global, ma_active;
    int ma_wfixed;        // every MA timescale is fixed: dt holds the weights exp(-dt / tau)
    int Dfull, D, nblocks;
    int off_kep, off_acc, off_off, off_jit, off_ma;
    int off_sin, off_mag; // Sinusoid / MagneticCycle block starts, -1 when absent
    int n_cos;            // cosine terms: 1 per sinusoid, 2 per magnetic cycle
    int off_sai;          // StellarActivity block start, -1 when absent
    int sai_w;            // most activity columns of one instrument (0: no activity terms)
    const double* sai;    // [sai_w][N] column c of each epoch's own instrument, device order
    const int* sai_base;  // [I + 1] first coefficient (within the block) of each instrument
    const double* t;      // [N] device order
    const double* y;      // [N]
    const double* s2;     // [N] yerr^2
    const double* dt;     // [N] t_i - t_{i-1} within an MA segment, -1 at a segment start;
                          //     with ma_wfixed exp(-dt / tau), 0 at a segment start
    const int* ins;       // [N]
    const int* inv_free;  // [Dfull] free index of each full parameter, -1 if fixed
    const int* free_idx;  // [D]
    const double* fixed_val;          // [Dfull]
    const EmpPriorTerm* prior;        // [Dfull]
    const int* block_start;           // [nblocks + 1]
    const EmpPriorTerm* ecc_prior;    // [K]
    const int* pseq;                  // [npseq] prior terms in summation order (see PSEQ_*)
    const PseqEntry* pseq_e;          // [npseq] the same, resolved (single-model engines; null in batches)
    int npseq;
    const double2* sctab;             // [SC_N + 1] (sin, cos) of k pi / SC_N
    double lnl_const;                 // -0.5 ln(2 pi) N
    double t0;                        // X_[0] (acceleration origin)
};

// Per-slot LDS scratch of the likelihood kernels.  KTE: the entry of the
// mixed path's Kepler tables (ktab_start), float4 intervals or float2 nodes.
template <class KTE>
struct SlotLDSt {
    double theta[MAXD];
    PlanetC planet[KMAX];
    double gam[IMAX], jit2[IMAX], ma_c[IMAX], ma_tau[IMAX], jit[IMAX];
    double acc[AMAX];
    double cosw[CMAX][3];      // per cosine term: (amplitude, angular frequency, phase)
    KTE ktab[KMAX][KT_N];      // mixed path: per planet, the Kepler table
    int ktab_mask;             // bits 4 p .. 4 p + 2: planet p's Kepler path flags (kt_flags)
};

// The slot of a workgroup that holds `SLOTS` of them: the interval tables
// (8 KiB per slot) where there are at most two, the node tables (4 KiB)
// where more would cost occupancy or not fit (16 slots: 175 KiB).
template <int SLOTS>
using SlotLDS = SlotLDSt<std::conditional_t<(SLOTS <= 2), float4, float2>>;

}  // namespace emp

struct EmpModel {
    EmpModelDesc desc;        // scalars copied from the caller (pointers cleared)
    emp::ModelDev dev;
    std::vector<void*> allocations;
    std::vector<int> order;   // device position -> caller (time-sorted) position
    unsigned long long* d_nonfinite = nullptr;
    double* d_stage_in = nullptr;     // staging for EMP_HOST_POINTERS
    double* d_stage_out = nullptr;
    size_t stage_in_cap = 0, stage_out_cap = 0;
    int wpt_override = 0;
    int grad_kernel = 0;      // EMP_GRAD_KERNEL_*
    int device = 0;
    int n_cu = 256;
    double* d_grad_scratch = nullptr;   // MA adjoint: eps rows [B][N]
    size_t grad_scratch_cap = 0;
};

// Host-side preparation of one model: validated, epochs in device order
// (grouped by instrument when the per-instrument moving average is active),
// MA time steps, free/fixed maps, prior block starts.  dev holds the scalars;
// its pointers are left null for the caller to place the arrays.
struct ModelHost {
    emp::ModelDev dev{};
    std::vector<int> order;   // device position -> caller position
    std::vector<double> t, y, s2, dt, fixed, sai;
    std::vector<int> ins, inv, free_idx, bstart, pseq, sai_base;
    std::vector<emp::PseqEntry> pseq_e;
    std::vector<EmpPriorTerm> prior, ecc;
};
int emp_model_host(const EmpModelDesc* d, ModelHost* h);
// Every device array a likelihood kernel may read is set (pseq_e is only
// used by the single-model engines and may be null).
inline bool emp_model_dev_complete(const emp::ModelDev& m) {
    return m.t && m.y && m.s2 && m.dt && m.ins && m.inv_free && m.free_idx && m.fixed_val && m.prior &&
           m.block_start && m.ecc_prior && m.pseq && m.sctab && m.sai && m.sai_base;
}
// (sin, cos) of k pi / SC_N, k = 0..SC_N, rounded from extended precision.
void emp_sincos_table(std::vector<double2>& tab);

// Error handling.
void emp_set_error(const char* fmt, ...);
int emp_hip_fail(hipError_t e, const char* what, const char* file, int line);
bool emp_fault_injected();
#define EMP_HIP(x)                                                              \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) return emp_hip_fail(e_, #x, __FILE__, __LINE__);  \
    } while (0)

// Likelihood launchers (emp_model.hip), used by the PT engine as well.
// Waves per vector for B vectors: batched likelihoods, or (halfstep) the
// PT half-step's proposals.
int emp_pick_wpt(const EmpModel* m, int64_t B, int precision, bool halfstep = false);
int emp_launch_lnl(const EmpModel* m, const double* d_theta, int64_t B, int64_t ld,
                   double* d_lnl, int precision, int wpt, hipStream_t s);
int emp_launch_prior(const EmpModel* m, const double* d_theta, int64_t B, int64_t ld,
                     double* d_lnp, hipStream_t s);
// lnL and its gradient (emp_grad.hip); grows the model's MA scratch if needed.
int emp_launch_grad(EmpModel* m, const double* d_theta, int64_t B, int64_t ld, double* d_lnl,
                    double* d_grad, hipStream_t s, int precision = EMP_FP64);
// The same with a caller-owned scratch of B x N doubles (used only when the
// moving average is active); allocates nothing, so it may be captured.
// stale (optional, per vector): only vectors flagged 1 are computed, and their
// flags are cleared.  wpt: waves per vector (0: emp_pick_grad_wpt).
int emp_launch_grad_ws(const EmpModel* m, const double* d_theta, int64_t B, int64_t ld, double* d_lnl,
                       double* d_grad, double* eps_scratch, hipStream_t s, int precision = EMP_FP64,
                       int* stale = nullptr, int wpt = 0);
int emp_pick_grad_wpt(const EmpModel* m, int64_t B);
// Doubles of scratch one adjoint launch over B vectors needs.
size_t emp_grad_scratch_doubles(const EmpModel* m, int64_t B);
