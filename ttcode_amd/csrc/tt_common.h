// tt_common.h -- shared definitions of the gfx950 implementation of the ttcode
// hot path (libttgpu.so).
//
// Device layout.  Every 3-D field lives in HBM over the full ghosted box
// (-1:nx+2, -1:ny+2, -1:nz+2) in the reference's own column-major order (r
// fastest), so the wire format of tt_upload/tt_download is a sub-box copy.
// Fields whose declared bounds are smaller (ppmod.f) use a part of the box.
// The SOR working set has a separate colour-split layout (tt_sor.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>

#include "../../include/ttgpu.h"
#include "../../include/ttgpu_fields.h"

namespace tt {

// ppmod.f:1071-1093: the reference stores its constants from REAL literals.
constexpr double FD = (double)2.99792458e10f;
constexpr double ARAD = (double)7.56596e-15f;
constexpr double RGAS = (double)8.315e7f;
constexpr double SBCONST = (double)5.6705e-5f;
constexpr double PI = 3.141592653589793;

// Box geometry passed by value to kernels.
struct Geo {
    int nx, ny, nz;
    int gx, gy, gz;
    long long n3;
    // 32-bit arithmetic: tt_init rejects boxes of 2^31 cells or more
    __host__ __device__ inline long long idx(int i, int j, int k) const {
        return (long long)((i + 1) + gx * ((j + 1) + gy * (k + 1)));
    }
    __host__ __device__ inline int idx2(int j, int k) const { return (j + 1) + gy * (k + 1); }
};

// 1-D metric arrays: one buffer, metric m at base[m*len + i + 2].
// The table is written only by host copies, never while a kernel runs, so it
// is read through the constant address space: a metric whose index is
// uniform across the wavefront (a plane's k, a block's row) becomes a scalar
// load into scalar registers, which no later store in the kernel can alias.
using cdouble_p = const __attribute__((address_space(4))) double*;
struct Met {
    const double* base;
    int len;
    __device__ inline double operator()(int m, int i) const {
        return ((cdouble_p)base)[(long long)m * len + i + 2];
    }
};

// div(v) of a cell in the reference's operation order (the p dV work,
// ppfor.f:190-196, and VISC's compression term, ppvisc.f:36-39)
__device__ inline double divv(const Geo& g, const Met& m, const double* V, const double* G, const double* H, int i,
                              int j, int k) {
    return (m(TT_M_SURXA, i + 1) * V[g.idx(i + 1, j, k)] - m(TT_M_SURXA, i) * V[g.idx(i, j, k)]) / m(TT_M_VOLXB, i) +
           (m(TT_M_SURYA, j + 1) * G[g.idx(i, j + 1, k)] - m(TT_M_SURYA, j) * G[g.idx(i, j, k)]) / m(TT_M_VOLYB, j) +
           (m(TT_M_SURZA, k + 1) * H[g.idx(i, j, k + 1)] - m(TT_M_SURZA, k) * H[g.idx(i, j, k)]) / m(TT_M_VOLZB, k);
}

// Streaming store for a kernel's output arrays: each element is written once
// and no later access of the same kernel reads it; the next kernel reads the
// array from memory (36 MB fields do not stay in a 4 MB L2).  The
// nontemporal form skips the L2's write allocation, which otherwise bounds
// the many-output stencil kernels (the three face families: 235 -> 163 us at
// 128x128x256, bit-identical).
template <typename T>
__device__ inline void st_nt(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}

// Device-side control block: reductions, convergence state, error flags.
struct Ctrl {
    // SOR
    int done;          // 0 running, 1 converged, 2 itmax reached
    int it;            // completed iterations
    int itmax, itmin;
    unsigned int counter;   // last-block-done counter
    unsigned int counter2;
    double rhsnorm, eps, rnorm;
    // Newton
    double maxdq, maxde;
    int bad_q;         // delta_Q positivity failure
    int bad_bound;     // BOUNDXMAX(2) upperbnd above the domain
    int bad_state;     // SCHRITT positivity failure
    int bad_kappa;     // table opacities: T <= 0 or an undefined pressure
    int spec_hit;      // a deferred (speculative) test ended the solve
    double dtmin;
    double scratch[8];
    double av_min[3];  // NR_ARTIFICIAL_VISC: min dv/dx of the x, y and z passes
    double bot_dtdr;   // BOUNDXMIN(9) at III <= 0: the gradient of cell (1,1,1)
    // stagnation watch of inner (FP32) solves
    double best;       // best L1 ratio seen
    int since;         // iterations since it last improved by 1%
    int stag;          // window (0: off)
    // the FP64 defect correction's outer loop, decided on the device
    double rhs64;      // sum |RHS| of the FP64 system
    double rn64;       // FP64 L1 residual after the latest correction
    int outer;         // corrections started
    int outer_done;    // 0 running, 1 converged, 2 no progress, 3 a cap reached
    int total;         // inner iterations of the finished corrections
    int pad_c;
    int hist[16];      // inner iterations of each of the first 16 corrections (0: not run)
    // the Newton loop's decision after an update, on the device (k_newton_decide)
    int newton_fin;    // 1: the loop ends here (converged, the iteration cap, or a fatal flag)
    int pad_n;
};

// Library state (one device, one stream).
struct Lib {
    bool ready = false;
    tt_params p{};
    Geo g{};
    hipStream_t s = nullptr;
    double* f[TT_FIELD_MAX] = {};   // ghosted device arrays, by field id
    int32_t* ub = nullptr;          // upperbnd (-1:ny+2,-1:nz+2)
    double* mustar = nullptr;
    double* sincoef = nullptr;      // per (j,k) column: SBCONST Tinc^4 / mu* of the grey irradiation
    double* scza = nullptr;         // sin and cos of XZA(k), k = -2 .. (metric slots), from the device's functions
    double* met = nullptr;          // device metric buffer
    double* hmet = nullptr;         // host copy
    int metlen = 0;
    // temporaries
    double *ED = nullptr, *VD = nullptr, *GDN = nullptr, *HDN = nullptr;
    double *FLMX = nullptr, *FLMY = nullptr, *FLMZ = nullptr;
    double *TOLD = nullptr, *VH = nullptr, *GH = nullptr, *HH = nullptr;
    double *DFQ = nullptr;
    double *DE1 = nullptr;          // first Newton correction of the last step (warm start)
    double* ftab = nullptr;         // IKAPTYP 4/5 tables on the device and 10^ of them (6 x 42 x 18)
    int ftab_version = -1;          // version of the host tables ftab holds
    double* btab = nullptr;         // IKAPTYP 9-14 tables and log10 T axis, then 10^ of them (2 x (3 x 50 x 50 + 50))
    int btab_version = -1;
    double* ctab = nullptr;         // IKAPTYP 6/7 column table, then 10^ of it (2 x 8 x 42 x 18)
    int ctab_version = -1;
    double* wtab = nullptr;         // IKAPTYP 20-25 irradiation bins (50 x 50 x 30, log10)
    int wtab_version = -1;
    int have_de1 = 0;
    double* de1_out = nullptr;      // set: the next energy solve also stores its correction there (DE1)
    double* partial = nullptr;      // block partial sums (reductions)
    int npartial = 0;
    Ctrl* ctrl = nullptr;           // device
    Ctrl* hctrl = nullptr;          // pinned host mirror
    Ctrl* hctrl_d = nullptr;        // its device-side address (kernels write it directly)
    tt_solver_opts opts{};
    tt_limits lim{};                // clamps applied by tt_step
    // learned guess: consecutive rejections, and first solves since the
    // network last ran
    int learned_rejects = 0, learned_idle = 0;
    tt_viscosity visc{};            // ETACALC/VISC in FORGAS3D (off: ivis 0)
    int av_valid = 0;               // ctrl->av_min holds a NR_ARTIFICIAL_VISC result
    // guess 2 without the network: moving averages of the first solve's SOR
    // phase time (ms: the guess's check and the solve) from the cold [0] and
    // the warm [1] start (0: none yet), and the count of such solves (every
    // 16th tries the start not preferred)
    double start_avg[2] = {0.0, 0.0};
    int start_picks = 0;
    std::string err;
    // sweep timing
    double line_om_auto = 0.0;  // the line sweeps' omega a search chose (tt_solver_opts.line_omega 0)
    int line_om_trials = 0;     // that search's trial solves
    bool timing = false;
    int timing_period = 1;
    int64_t timing_count = 0;
    int64_t n64 = 0, n32 = 0;
    double ms64 = 0.0, ms32 = 0.0;
    // speculative convergence tests of the FP64 SOR solve
    double spec_margin = 2.5;
    int64_t spec_batches = 0, spec_reruns = 0;
    int energy_iter = 0;
    int iii = 0;
    int have_prev_guess = 0;
    // upperbnd holds the search result of the last BOUNDS and RH has changed
    // since only through that BOUNDS (cleared by every other writer of RH or
    // upperbnd): the next BOUNDS may take its search from those results
    bool ub_repeat = false;
    // GRAVY / GRAVZ hold +0.0 in every cell (as tt_init leaves them; cleared
    // by an upload of anything else): the force kernel then adds the
    // constant instead of reading the array, the same arithmetic
    bool ccy_zero = true;  // CENTCORY is +0.0 on the interior (zeroed at tt_init, not uploaded since)
    bool gy_zero = true, gz_zero = true;
    // non-null: the kernels the *_impl helpers launch run only where this
    // device flag is non-zero (work queued ahead of a control-block read,
    // which returns at once when the flag was not set in stream order)
    const int* run_if = nullptr;
    // the work after a line-sweep energy solve, queued ahead of its control-
    // block read under run_if = outer_done (set by energy_impl around the
    // solve; sor_solve_impl then leaves storing the result to it)
    std::function<int()> sor_tail;
    // tt_step's work after COUPLED_ENERGY_EQS (BOUNDS, BOUNDS, PRESSG),
    // queued by energy_impl ahead of its last read under newton_fin; set by
    // tt_step, and trailing_done reports whether that read found it done
    std::function<int()> step_trailing;
    bool trailing_done = false;
    int64_t dev_bytes = 0;          // device memory held by the library
    hipEvent_t step_ev[4] = {};     // tt_step's phase boundaries (created by tt_init)
    hipEvent_t en_ev[4] = {};       // a Newton iteration's build / SOR / update boundaries
};

extern Lib L;

int fail(int code, const std::string& msg);

#define TT_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) return tt::fail(10, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define TT_LAUNCH()                                                                   \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess) return tt::fail(11, std::string("launch: ") + hipGetErrorString(e_)); \
    } while (0)

inline Met met() { return Met{L.met, L.metlen}; }

// hipMalloc that counts what the library holds (freed all at once by tt_finalize)
template <typename T>
inline hipError_t dev_alloc(T** p, size_t bytes) {
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e == hipSuccess) L.dev_bytes += (int64_t)bytes;
    return e;
}
inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Device-to-device copy and zero fill on the library's stream, as kernels
// (sizes in bytes, multiples of 4)
int dev_copy(void* dst, const void* src, size_t bytes);
int dev_zero(void* dst, size_t bytes);

// Fortran order helpers for the full box
__device__ inline void box_ijk(const Geo& g, long long t, int& i, int& j, int& k) {
    const int u = (int)t;  // boxes stay below 2^31 cells
    const int r = u / g.gx;
    i = u - r * g.gx - 1;
    k = r / g.gy;
    j = r - k * g.gy - 1;
    k -= 1;
}

// Sub-box decomposition: t over ni*nj*nk with origin (i0,j0,k0)
__device__ inline void sub_ijk(long long t, int ni, int nj, int i0, int j0, int k0, int& i, int& j, int& k) {
    const int u = (int)t;
    const int r = u / ni;
    i = u - r * ni + i0;
    const int q = r / nj;
    j = r - q * nj + j0;
    k = q + k0;
}

// Wavefront + block reductions (64-wide wavefronts)
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// Reduction over the 256 threads of a block, in a fixed order (wavefront
// butterflies, then the four wavefront results in order); op 0 sum, 1 max,
// 2 min.  The result is valid in thread 0.
__device__ inline double block_reduce256(double v, double* sh, int op) {
    for (int o = 32; o > 0; o >>= 1) {
        double u = __shfl_xor(v, o, 64);
        v = op == 0 ? v + u : (op == 1 ? fmax(v, u) : fmin(v, u));
    }
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        s = sh[0];
        for (int q = 1; q < 4; ++q) s = op == 0 ? s + sh[q] : (op == 1 ? fmax(s, sh[q]) : fmin(s, sh[q]));
    }
    __syncthreads();
    return s;
}
__device__ inline double block_sum256(double v, double* sh) { return block_reduce256(v, sh, 0); }

// Fold nvec (<= 3) vectors of nb block partials, part[v*nb + b], into the
// device scalars out[v] with one block launched after the producing kernel
// (the kernel boundary orders the partials before the fold reads them).
int fold_partials(const double* part, int nb, int nvec, int op, double* out0, double* out1 = nullptr,
                  double* out2 = nullptr);

// A kernel launched under L.run_if returns at once unless the flag is set.
__device__ inline bool skip_run(const int* run) { return run && *run == 0; }

// Internal entry points shared between translation units
int bounds_impl();
int pressg_impl();
int qdens_impl();
int sound_impl();
int planetshift_impl(double* a, int f);
int sor_solve_impl(int fill_from_evals, tt_sor_stats* st, int fieldmode);
int sor_store_result(int mode);   // the solution into delta_ER (mode 0) or T (1), phi ghosts wrapped
int sor_adjust_impl(int fill_from_evals, double* omega_out, int32_t* iters);
int limits_impl(int stage, double delt);
int sor_pack_impl();
// The SOR system lives in two forms: the module arrays SOR_A..G, RHS_ER
// (wire layout) and the colour-split copy the sweeps read.  A kernel that
// writes the colour-split copy directly reports it; the module arrays are
// then rebuilt from it only when the host downloads one of them.
struct PackedSys {
    double* c64[2];
    float* c32[2];
    int hx;
};
PackedSys sor_packed();
void sor_sys_packed_written(bool with32);
void sor_sys_module_written();
int sor_sys_before_download(int field);
int diffusion_impl(double delt, double xnue, double pr, int* iters);
int energy_impl(double delt, int iii, tt_energy_stats* st);
// EVAL1..6 / EVAL_1_6 are written to their module arrays on demand
// (tt_energy.hip): evals_materialize() writes them if they are pending;
// evals_before_access(field) does so before a host transfer of a field they
// derive from or are; evals_reset() drops the pending state (tt_init)
int evals_materialize();
int scza_impl();                  // L.scza from the metric table (tt_state.hip)
int evals_before_access(int field);
void evals_reset();
int burrows_cell111_impl();        // BurrowsOPC at cell (1,1,1) (BOUNDXMIN(9), III <= 0)
int advect_impl(double delt);
int forgas_impl(double delt);
int visc_impl();
void visc_free();
int artvis_impl(double dt);       // NR_ARTIFICIAL_VISC, ppfor.f:1032-1230
double dt_artvis();               // dt_artvis after the last artvis_impl (host; syncs)
int sphericalvellimit_impl(int ivel, double velmax, double hydrotime);
int sync_check();
int check_flags();                // sync_check's tests on the host copy already read
// The control block into its pinned host mirror (a one-wavefront kernel that
// writes the mapped host memory, then a stream sync): no copy engine or blit
// kernel per read.
int ctrl_to_host();

}  // namespace tt
