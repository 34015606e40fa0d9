// tt_sor.hip -- red-black SOR of ttcode (ppsor.f, ppdif.f) on gfx950.
//
// SOLVE_WITH_SOR (ppsor.f:2-187) sweeps red cells ((i+j+k) even) then black
// cells, refreshing the periodic phi ghosts after each half sweep, and stops
// when the L1 sum of the residuals taken during the sweep, divided by
// sum|RHS|, is <= EPSMAT (or at ITMAX).  Red cells depend only on black cells
// and vice versa (ny even keeps the colouring consistent across the phi
// wrap), so each half sweep is one fully parallel kernel whose result equals
// the reference's sequential sweep; only the summation order of the L1 norm
// differs.
//
// Colour-split layout.  For colour c and row (j,k) the cells of that colour
// are i = 2h + 1 + s, h = 0..hx-1 (hx = nx/2), s = (j+k+c+1) & 1.  The seven
// coefficients and the RHS of a colour are stored row-blocked: row r = (j-1)
// + ny (k-1) holds the eight vectors A..G, RHS of length hx one after the
// other, so a wavefront reads one contiguous 8*hx block per row.  The unknowns
// of a colour are stored in rows of px = hx + 4 slots, slot h at row + 2 + h
// (even, so FP32 pairs load as one 8-byte word when hx is even); slots h = -1
// and h = hx hold the fixed r ghost values and one extra row at k = 0 and
// k = nz+1 the theta ghosts (zero for delta_ER).  The r neighbours of a cell
// of colour c are slots h+s-1 and h+s of the other colour's row, its phi and
// theta neighbours slot h of the neighbouring rows.
//
// Convergence.  Each half sweep writes one partial |resid| sum per block; a
// one-block finish kernel after the black sweep sums them in a fixed order,
// advances the iteration counter and sets ctrl->done, which every later
// sweep reads first and returns on.  The kernel boundary orders the partial
// sums before the finish kernel reads them (a per-block device-scope fence
// would write the XCD's L2 back once per block).  The host queues `batch`
// iterations between reads of the control block, so the iteration at which
// the solve stops is exactly the first one that meets the tolerance.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <utility>
#include <vector>

#include <hip/hip_ext.h>

#include "tt_common.h"

namespace tt {

struct CS {
    int hx, ny, nz, px;  // px = hx + 4 slots per row of unknowns, slot h at 2 + h
    long long ncell;  // cells per colour
};

static CS g_cs;
// Coefficients + RHS per colour, row-blocked: element (q, r, h) of the
// vectors A,B,C,D,E,F,G,RHS (q = 0..7) at (8 r + q) hx + h.
static double* c64[2];
static float* c32[2];
static double* x64[2];  // padded unknowns
static float* x32[2];
static int g_nb = 0;           // sweep blocks
static int g_chunk = 0;
static std::vector<hipEvent_t> ev_pool, ev_pend;
static std::vector<int> ev_kind;
static std::vector<int> ev_iter;  // SOR iteration (0-based) of each timed launch
static std::vector<long long> ev_bid;  // batch of each timed launch
// correction of each timed launch in a queued line defect correction (-1:
// another solve), and the correction being queued
static std::vector<int> ev_outer;
static int g_outer_tag = -1;
// Two pinned copies of the control block, so that the host reads batch k's
// convergence state while batch k+1 already runs on the device.
static Ctrl* g_hc[2] = {nullptr, nullptr};
static Ctrl* g_hcd[2] = {nullptr, nullptr};  // their device-side addresses
static hipEvent_t g_cev[2] = {nullptr, nullptr};
static long long g_cbid[2] = {0, 0};
static long long g_bid = 0;  // batches enqueued so far

static long long padded_len() { return (long long)g_cs.px * g_cs.ny * (g_cs.nz + 2); }

// Radial-line inner solves (tt_solver_opts.inner = 1).  Line (j,k) is the
// whole radial column i = 1..nx; lines are coloured by (j+k) parity, so a
// line's phi and theta neighbours are lines of the other colour and the
// lines of one colour are solved independently.  Row line_row(j-1, k-1) of
// g_lco (below) holds eight FP32 vectors of nx values: the couplings D, E (phi
// -/+), F, G (theta -/+), then the LU factors of the line's tridiagonal
// part, m (the multiplier of row i-1) and u (the pivot), the super-diagonal
// C, and the right-hand side.  The unknowns are rows of nx values over
// (j, k = 0..nz+1); the theta ghost rows k = 0, nz+1 and the r ghosts stay
// zero (the correction's boundary values).
struct LS {
    int nx, ny, nz;
};
static LS g_ls;

// Row of line (jj, kk) (0-based j-1, k-1) in the line system: the lines of
// each colour (j+k parity) are stored together, colour 0 first, in (k, j)
// order within it, so that a half sweep streams one contiguous block.
__host__ __device__ inline long long line_row(int jj, int kk, int ny, int nz) {
    const int hy = ny >> 1;
    return (long long)((jj + kk) & 1) * hy * nz + (long long)kk * hy + (jj >> 1);
}
static float* g_lco = nullptr;
static float* g_lx = nullptr;
static int g_lchunk = 0;  // lines of a colour per logical block
static int g_lnb = 0;     // line-sweep blocks
// lines per wavefront in a line half sweep (2, 4 and 8 measure the same,
// 15.2-15.3 us at 128x256x128)
constexpr int LINE_LPW = 4;
static int g_pn = 0;      // partials per colour of the sweeps being queued (g_nb or g_lnb)
static int g_line = 0;    // FP32 sweeps launch the line kernel
static bool g_line_ok = false;  // g_lco's factors match the FP64 system
static long long line_x_len() { return (long long)g_ls.nx * g_ls.ny * (g_ls.nz + 2); }

void sor_free() {
    for (int c = 0; c < 2; ++c) {
        if (c64[c]) hipFree(c64[c]);
        if (c32[c]) hipFree(c32[c]);
        c64[c] = nullptr;
        c32[c] = nullptr;
        if (x64[c]) hipFree(x64[c]);
        if (x32[c]) hipFree(x32[c]);
        x64[c] = nullptr;
        x32[c] = nullptr;
    }
    if (g_lco) hipFree(g_lco);
    if (g_lx) hipFree(g_lx);
    g_lco = nullptr;
    g_lx = nullptr;
    g_line_ok = false;
    for (auto e : ev_pool) hipEventDestroy(e);
    for (auto e : ev_pend) hipEventDestroy(e);
    ev_pool.clear();
    ev_pend.clear();
    ev_kind.clear();
    ev_iter.clear();
    ev_bid.clear();
    ev_outer.clear();
    for (int q = 0; q < 2; ++q) {
        if (g_hc[q]) hipHostFree(g_hc[q]);
        if (g_cev[q]) hipEventDestroy(g_cev[q]);
        g_hc[q] = nullptr;
        g_hcd[q] = nullptr;
        g_cev[q] = nullptr;
    }
}

int sor_alloc() {
    g_cs.hx = L.g.nx / 2;
    g_cs.ny = L.g.ny;
    g_cs.nz = L.g.nz;
    g_cs.px = g_cs.hx + 4;
    g_cs.ncell = (long long)g_cs.hx * g_cs.ny * g_cs.nz;
    for (int c = 0; c < 2; ++c) {
        TT_HIP(dev_alloc(&c64[c], 8 * g_cs.ncell * sizeof(double)));
        TT_HIP(dev_alloc(&c32[c], 8 * g_cs.ncell * sizeof(float)));
        TT_HIP(dev_alloc(&x64[c], padded_len() * sizeof(double)));
        TT_HIP(dev_alloc(&x32[c], padded_len() * sizeof(float)));
        TT_HIP(hipMemsetAsync(x64[c], 0, padded_len() * sizeof(double), L.s));
        TT_HIP(hipMemsetAsync(x32[c], 0, padded_len() * sizeof(float), L.s));
    }
    for (int q = 0; q < 2; ++q) {
        TT_HIP(hipHostMalloc((void**)&g_hc[q], sizeof(Ctrl), hipHostMallocDefault));
        TT_HIP(hipHostGetDevicePointer((void**)&g_hcd[q], g_hc[q], 0));
        TT_HIP(hipEventCreateWithFlags(&g_cev[q], hipEventDisableTiming));
    }
    // sweep blocks: about 4 rows per wavefront, a multiple of 8 (XCDs)
    int nrows = g_cs.ny * g_cs.nz;
    int nb = (nrows + 15) / 16;
    if (nb > 2048) nb = 2048;
    nb = ((nb + 7) / 8) * 8;
    if (nb < 8) nb = 8;
    // partials: two iterations (by parity) x two colours x nb, in the first
    // half of the partial buffer; the fold reads at most 16 per thread of 256
    if (4 * nb > L.npartial || 2 * nb > 16 * 256) return fail(1, "SOR: partial buffer too small");
    g_nb = nb;
    g_pn = nb;
    g_chunk = (nrows + nb - 1) / nb;  // rows per block
    return 0;
}

// The line system and unknowns, on the first line solve (134 + 17 MB at
// 128 x 256 x 128), kept to the end.
static int line_alloc() {
    if (g_lco) return 0;
    if (L.g.nx > 512) return fail(1, "line sweeps: nx must be <= 512 (eight cells per lane)");
    g_ls = LS{L.g.nx, L.g.ny, L.g.nz};
    const long long ncell = (long long)g_ls.nx * g_ls.ny * g_ls.nz;
    TT_HIP(dev_alloc(&g_lco, 8 * ncell * sizeof(float)));
    TT_HIP(dev_alloc(&g_lx, line_x_len() * sizeof(float)));
    TT_HIP(hipMemsetAsync(g_lx, 0, line_x_len() * sizeof(float), L.s));
    const int nl = (g_ls.ny / 2) * g_ls.nz;  // lines of a colour
    g_lchunk = 4 * LINE_LPW;
    g_lnb = (nl + g_lchunk - 1) / g_lchunk;
    if (g_lnb > g_nb) {  // the partial slots hold g_nb per colour
        g_lnb = g_nb;
        g_lchunk = (nl + g_nb - 1) / g_nb;
    }
    g_line_ok = false;
    return 0;
}

// logical block id: blocks b, b+8, b+16, ... run on the same XCD; give each
// XCD one contiguous slab of the domain so phi/theta neighbour rows share its L2.
__device__ inline int xcd_block(int nb) {
    int b = blockIdx.x;
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Finish an iteration with one block of 256 threads: sum the 2*nb partials
// of the iteration in a fixed order (thread t adds part[t + 256 u], u
// ascending, then the wavefront and block sums) and update the control
// block.  step is the number of iterations since the previous test (1 on
// the reference path, which tests every iteration).  Every load is issued
// before the first add, so the block waits for one memory round trip.
__device__ inline void finish_block(Ctrl* ctrl, const double* __restrict__ part, int nparts, int step, double* sh,
                                    int deferred, Ctrl* snap = nullptr) {
    const int tid = threadIdx.x;
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = tid + 256 * u;
        v[u] = q < nparts ? part[q] : 0.0;
    }
    const int done = ctrl->done;
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
    s = block_sum256(s, sh);
    if (tid == 0 && !done) {
        int it = ctrl->it + step;
        ctrl->it = it;
        ctrl->rnorm = s;
        double ratio = s / ctrl->rhsnorm;
        int d = 0;
        if (it >= ctrl->itmax) d = 2;
        else if (!(ratio > ctrl->eps || it < ctrl->itmin)) d = 1;
        else if (ctrl->stag > 0) {  // inner solves: stop when the ratio no longer falls
            if (ratio < 0.99 * ctrl->best) {
                ctrl->best = ratio;
                ctrl->since = 0;
            } else {
                ctrl->since += step;
                if (ctrl->since >= ctrl->stag) d = 3;
            }
        }
        if (d) {
            ctrl->done = d;
            if (deferred) ctrl->spec_hit = 1;
        }
    }
    // the block as the host reads it after the batch (pinned host memory)
    if (tid == 0 && snap) *snap = *ctrl;
}

__global__ __launch_bounds__(256) void k_finish(Ctrl* ctrl, const double* __restrict__ part, int nparts, int step,
                                                Ctrl* snap) {
    __shared__ double sh[4];
    finish_block(ctrl, part, nparts, step, sh, 0, snap);
}

// Arguments of the convergence test a red sweep carries for the iteration
// before it (fin != 0): one extra block, launched after the nb sweep blocks,
// runs finish_block on that iteration's partials while the sweep proceeds.
struct FinArgs {
    const double* part;
    int nparts, step, fin;
};

// One half sweep.  Rows (j,k) of the colour are split into contiguous
// ranges, one per logical block; the four wavefronts of a block take every
// fourth row of its range and the 64 lanes of a wavefront walk h.  Index
// arithmetic is 32-bit and per row.  With rev set the logical blocks run in
// reverse order: alternating the direction from one iteration to the next
// makes a sweep start on the rows the previous sweep touched last, which are
// still in the memory-side (Infinity) cache.
//
// The colour's unknowns are updated in place in xc; xo holds the other
// colour's.  Block nb (present when fa.fin is set) runs the previous
// iteration's convergence test instead of sweeping rows; a solve that test
// ends has then had one more red half sweep, which inner solves of the
// defect correction allow and the FP64 solve undoes by running again.
template <typename T>
__global__ __launch_bounds__(256) void k_sweep(int c, const T* __restrict__ coef, T* __restrict__ xc,
                                               const T* __restrict__ xo, CS cs, T omega, double* part, int nb,
                                               int rows_per_block, Ctrl* ctrl, int rev, FinArgs fa) {
    __shared__ double sh[4];
    if ((int)blockIdx.x == nb) {
        finish_block(ctrl, fa.part, fa.nparts, fa.step, sh, 1);
        return;
    }
    // The test block may set done while this sweep runs, so wavefronts of a
    // block can read different values: none returns early (all reach the
    // block sum's barriers); a sweep the test stops is discarded by the host.
    const int done = ctrl->done;
    int lb = xcd_block(nb);
    if (rev) lb = nb - 1 - lb;
    // the wavefront index, uniform: the row arithmetic runs on the scalar unit
    const int tx = threadIdx.x & 63, ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int hx = cs.hx, ny = cs.ny, px = cs.px;
    const int nrows = ny * cs.nz;
    const int r0 = lb * rows_per_block;
    const int r1 = min(r0 + rows_per_block, nrows);
    double acc = 0.0;
    for (int r = done ? r1 : r0 + ty; r < r1; r += 4) {
        const int kk = r / ny;
        const int jj = r - kk * ny;
        const int s = (jj + kk + c + 1) & 1;  // (j+k+c+1)&1 with j=jj+1, k=kk+1
        const int jm = jj == 0 ? ny - 1 : jj - 1;
        const int jp = jj == ny - 1 ? 0 : jj + 1;
        const int row = px * (jj + ny * (kk + 1));
        const T* __restrict__ xr = xo + row + 1 + s;  // slot h+s-1
        const T* __restrict__ xrjm = xo + px * (jm + ny * (kk + 1)) + 2;
        const T* __restrict__ xrjp = xo + px * (jp + ny * (kk + 1)) + 2;
        const T* __restrict__ xrkm = xo + px * (jj + ny * kk) + 2;
        const T* __restrict__ xrkp = xo + px * (jj + ny * (kk + 2)) + 2;
        T* __restrict__ xw = xc + row + 2;
        const T* __restrict__ cr = coef + (long long)r * 8 * hx;
        for (int h = tx; h < hx; h += 64) {
            // all fourteen loads first, so they are in flight together
            T q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = cr[u * hx + h];
            const T xs = xw[h];
            const T x1 = xr[h], x2 = xr[h + 1], x3 = xrjm[h], x4 = xrjp[h], x5 = xrkm[h], x6 = xrkp[h];
            const T resid = q[0] * xs + q[1] * x1 + q[2] * x2 + q[3] * x3 + q[4] * x4 + q[5] * x5 + q[6] * x6 - q[7];
            acc += fabs((double)resid);
            xw[h] = xs - omega * resid / q[0];
        }
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0 && !done) part[c * nb + lb] = bs;
}

// FP32 half sweep with two cells per lane (hx even): half a wavefront per
// row, so the eight coefficient loads and the six unknown loads of a lane are
// 8-byte words and a wavefront still moves 512 bytes per load.  Same
// arithmetic per cell as k_sweep<float>.
__global__ __launch_bounds__(256) void k_sweep_f32x2(int c, const float* __restrict__ coef, float* __restrict__ xc,
                                                     const float* __restrict__ xo, CS cs, float omega,
                                                     double* part, int nb, int rows_per_block, Ctrl* ctrl,
                                                     int rev, FinArgs fa) {
    __shared__ double sh[4];
    if ((int)blockIdx.x == nb) {
        finish_block(ctrl, fa.part, fa.nparts, fa.step, sh, 1);
        return;
    }
    // The test block may set done while this sweep runs, so wavefronts of a
    // block can read different values: none returns early (all reach the
    // block sum's barriers); a sweep the test stops is discarded by the host.
    const int done = ctrl->done;
    int lb = xcd_block(nb);
    if (rev) lb = nb - 1 - lb;
    // the wavefront's row pair starts at a uniform row re (scalar arithmetic);
    // the upper half-wavefront takes re + 1
    const int lane = threadIdx.x & 63, ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int half = lane >> 5, tp = lane & 31;
    const int hx = cs.hx, hp = hx >> 1, ny = cs.ny, px = cs.px;
    const int nrows = ny * cs.nz;
    const int r0 = lb * rows_per_block;
    const int r1 = min(r0 + rows_per_block, nrows);
    double acc = 0.0;
    for (int re = done ? r1 : r0 + 2 * ty; re < r1; re += 8) {
        const int r = re + half;
        if (r >= r1) continue;
        const int ke = re / ny, je = re - ke * ny;
        const bool wrap = je + half == ny;
        const int kk = wrap ? ke + 1 : ke;
        const int jj = wrap ? 0 : je + half;
        const int s = (jj + kk + c + 1) & 1;
        const int jm = jj == 0 ? ny - 1 : jj - 1;
        const int jp = jj == ny - 1 ? 0 : jj + 1;
        const int row = px * (jj + ny * (kk + 1));
        const float* __restrict__ xrjm = xo + px * (jm + ny * (kk + 1)) + 2;
        const float* __restrict__ xrjp = xo + px * (jp + ny * (kk + 1)) + 2;
        const float* __restrict__ xrkm = xo + px * (jj + ny * kk) + 2;
        const float* __restrict__ xrkp = xo + px * (jj + ny * (kk + 2)) + 2;
        const float* __restrict__ xrr = xo + row + 2 * s;  // aligned window over slots h+s-1 .. h+s+1
        float* __restrict__ xw = xc + row + 2;
        const float* __restrict__ cr = coef + (long long)r * 8 * hx;
        for (int pp = tp; pp < hp; pp += 32) {
            const int h = 2 * pp;
            float2 q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = *(const float2*)(cr + u * hx + h);
            const float2 xs = *(const float2*)(xw + h);
            const float2 w0 = *(const float2*)(xrr + h), w1 = *(const float2*)(xrr + h + 2);
            const float2 a3 = *(const float2*)(xrjm + h), a4 = *(const float2*)(xrjp + h);
            const float2 a5 = *(const float2*)(xrkm + h), a6 = *(const float2*)(xrkp + h);
            // slots h+s-1, h+s, h+s+1 of the other colour
            const float v0 = s ? w0.x : w0.y, v1 = s ? w0.y : w1.x, v2 = s ? w1.x : w1.y;
            const float ra = q[0].x * xs.x + q[1].x * v0 + q[2].x * v1 + q[3].x * a3.x + q[4].x * a4.x +
                             q[5].x * a5.x + q[6].x * a6.x - q[7].x;
            const float rb = q[0].y * xs.y + q[1].y * v1 + q[2].y * v2 + q[3].y * a3.y + q[4].y * a4.y +
                             q[5].y * a5.y + q[6].y * a6.y - q[7].y;
            acc += fabs((double)ra);
            acc += fabs((double)rb);
            float2 o;
            o.x = xs.x - omega * ra / q[0].x;
            o.y = xs.y - omega * rb / q[0].y;
            *(float2*)(xw + h) = o;
        }
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0 && !done) part[c * nb + lb] = bs;
}

// ------------------------------------------------------ radial-line sweeps
// Lane moves on the vector unit (DPP), for the scans of the line sweep.
// dpp<CTRL>(old, v): v from the lane CTRL selects; lanes whose source lies
// outside the wavefront or the 16-lane row keep old.  Controls: 0x111 + n - 1
// row_shr:n (lane l - n of the row), 0x101 + n - 1 row_shl:n (lane l + n),
// 0x138 wave_shr:1 (lane l - 1), 0x130 wave_shl:1 (lane l + 1).
template <int CTRL>
__device__ inline float dpp(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xf, 0xf, false));
}
__device__ inline float rdlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// Inclusive scan of affine maps z -> a z + b over the lanes in ascending
// order (lane l ends with the composition of lanes 0..l, lane 0 applied
// first): within each 16-lane row by four shifted compositions, then the
// totals of the rows before, read from lanes 15, 31, 47.
__device__ inline void scan_affine_up(float& a, float& b, int lane) {
#define TT_UP(N)                                                                                    \
    {                                                                                               \
        const float ua = dpp<0x110 + N>(1.f, a), ub = dpp<0x110 + N>(0.f, b);                       \
        b = a * ub + b;                                                                             \
        a = a * ua;                                                                                 \
    }
    TT_UP(1) TT_UP(2) TT_UP(4) TT_UP(8)
#undef TT_UP
    const float a0 = rdlane(a, 15), b0 = rdlane(b, 15), a1 = rdlane(a, 31), b1 = rdlane(b, 31);
    const float a2 = rdlane(a, 47), b2 = rdlane(b, 47);
    // rows 0..r-1 composed: r = 1: T0; r = 2: T1 T0; r = 3: T2 T1 T0
    const float a01 = a1 * a0, b01 = a1 * b0 + b1;
    const float a012 = a2 * a01, b012 = a2 * b01 + b2;
    const int row = lane >> 4;
    const float ca = row == 0 ? 1.f : row == 1 ? a0 : row == 2 ? a01 : a012;
    const float cb = row == 0 ? 0.f : row == 1 ? b0 : row == 2 ? b01 : b012;
    b = a * cb + b;
    a = a * ca;
}

// The same in descending order (lane l ends with lanes l..63, lane 63
// applied first); row totals from lanes 48, 32, 16.
__device__ inline void scan_affine_down(float& a, float& b, int lane) {
#define TT_DN(N)                                                                                    \
    {                                                                                               \
        const float va = dpp<0x100 + N>(1.f, a), vb = dpp<0x100 + N>(0.f, b);                       \
        b = a * vb + b;                                                                             \
        a = a * va;                                                                                 \
    }
    TT_DN(1) TT_DN(2) TT_DN(4) TT_DN(8)
#undef TT_DN
    const float a3 = rdlane(a, 48), b3 = rdlane(b, 48), a2 = rdlane(a, 32), b2 = rdlane(b, 32);
    const float a1 = rdlane(a, 16), b1 = rdlane(b, 16);
    // rows r+1..3 composed: r = 2: T3; r = 1: T2 T3; r = 0: T1 T2 T3
    const float a23 = a2 * a3, b23 = a2 * b3 + b2;
    const float a123 = a1 * a23, b123 = a1 * b23 + b1;
    const int row = lane >> 4;
    const float ca = row == 3 ? 1.f : row == 2 ? a3 : row == 1 ? a23 : a123;
    const float cb = row == 3 ? 0.f : row == 2 ? b3 : row == 1 ? b23 : b123;
    b = a * cb + b;
    a = a * ca;
}

// Q cells of a lane from p (FULL: all 64 Q cells of the line exist, so the
// lane's cells load as 8-byte pairs where Q is even); cells past nx read fill.
template <int Q, bool FULL>
__device__ inline void load_cells(float (&v)[Q], const float* __restrict__ p, int i0, int nx, float fill) {
    if (FULL && Q % 2 == 0) {
#pragma unroll
        for (int t = 0; t < Q; t += 2) {
            const float2 w = *(const float2*)(p + t);
            v[t] = w.x;
            v[t + 1] = w.y;
        }
    } else {
#pragma unroll
        for (int t = 0; t < Q; ++t) v[t] = (FULL || i0 + t < nx) ? p[t] : fill;
    }
}

// One half sweep of red-black line SOR: every line of colour c (lines
// (j,k) with (j+k) parity c) is solved exactly for the current values of its
// four neighbour lines,
//   T y = f,  f = b - D x(j-1) - E x(j+1) - F x(k-1) - G x(k+1),
// with T the line's tridiagonal part (B, A, C) in its LU form (m, u, C):
// z_i = f_i - m_i z_{i-1}, y_i = (z_i - C_i y_{i+1}) / u_i; then
// x <- x + omega (y - x).  One wavefront per line: lane l holds the Q cells
// l Q .. l Q + Q - 1 (0-based i - 1), so both recurrences are affine maps a
// lane composes over its own cells and the wavefront combines with a scan
// (forward over lanes, then backward).  The residual of the line before its
// update, f - L U x, is summed in |.| for the convergence test, as the point
// sweep sums its residuals.  Cells past nx act as identity rows with zero
// right-hand side.  Rows, blocks, the test block and the partials are laid
// out as in k_sweep.
// The operands of one line for a lane: eight coefficient vectors, the own
// unknowns and the four neighbour lines' (Q cells each).
template <int Q>
struct LineOps {
    float cD[Q], cE[Q], cF[Q], cG[Q], cM[Q], cU[Q], cC[Q], cB[Q], X[Q], n1[Q], n2[Q], n3[Q], n4[Q];
};

template <int Q, bool FULL>
__device__ inline void line_load(LineOps<Q>& o, const float* __restrict__ coef, const float* __restrict__ x, LS ls,
                                 int c, int q, int i0) {
    const int nx = ls.nx, ny = ls.ny, hy = ny >> 1;
    const int kk = q / hy;
    const int jj = 2 * (q - kk * hy) + ((kk + c) & 1);  // (jj + kk) & 1 == c
    const int jm = jj == 0 ? ny - 1 : jj - 1;
    const int jp = jj == ny - 1 ? 0 : jj + 1;
    const float* __restrict__ cr = coef + ((long long)c * hy * ls.nz + q) * 8 * nx + i0;  // line_row(jj, kk)
    load_cells<Q, FULL>(o.cD, cr, i0, nx, 0.f);
    load_cells<Q, FULL>(o.cE, cr + nx, i0, nx, 0.f);
    load_cells<Q, FULL>(o.cF, cr + 2 * nx, i0, nx, 0.f);
    load_cells<Q, FULL>(o.cG, cr + 3 * nx, i0, nx, 0.f);
    load_cells<Q, FULL>(o.cM, cr + 4 * nx, i0, nx, 0.f);
    load_cells<Q, FULL>(o.cU, cr + 5 * nx, i0, nx, 1.f);
    load_cells<Q, FULL>(o.cC, cr + 6 * nx, i0, nx, 0.f);
    load_cells<Q, FULL>(o.cB, cr + 7 * nx, i0, nx, 0.f);
    load_cells<Q, FULL>(o.X, x + (long long)(jj + ny * (kk + 1)) * nx + i0, i0, nx, 0.f);
    load_cells<Q, FULL>(o.n1, x + (long long)(jm + ny * (kk + 1)) * nx + i0, i0, nx, 0.f);
    load_cells<Q, FULL>(o.n2, x + (long long)(jp + ny * (kk + 1)) * nx + i0, i0, nx, 0.f);
    load_cells<Q, FULL>(o.n3, x + (long long)(jj + ny * kk) * nx + i0, i0, nx, 0.f);
    load_cells<Q, FULL>(o.n4, x + (long long)(jj + ny * (kk + 2)) * nx + i0, i0, nx, 0.f);
}

// One line's solve for a lane (see k_line_sweep): returns the lane's part
// of the residual sum and leaves the new unknowns in y.
template <int Q>
__device__ inline float line_solve(const LineOps<Q>& o, float omega, int lane, float (&y)[Q]) {
    float f[Q], g[Q];
#pragma unroll
    for (int t = 0; t < Q; ++t)
        f[t] = o.cB[t] - o.cD[t] * o.n1[t] - o.cE[t] * o.n2[t] - o.cF[t] * o.n3[t] - o.cG[t] * o.n4[t];
    // residual before the update: f - L (U x), (U x)_i = u_i x_i + C_i x_{i+1},
    // (L w)_i = w_i + m_i w_{i-1}; the r ghosts are zero
    const float xnext = dpp<0x130>(0.f, o.X[0]);
#pragma unroll
    for (int t = 0; t < Q; ++t) g[t] = o.cU[t] * o.X[t] + o.cC[t] * (t + 1 < Q ? o.X[t + 1] : xnext);
    const float gprev = dpp<0x138>(0.f, g[Q - 1]);
    float rs = 0.f;
#pragma unroll
    for (int t = 0; t < Q; ++t) rs += fabsf(f[t] - (g[t] + o.cM[t] * (t ? g[t - 1] : gprev)));
    // forward substitution z_i = f_i - m_i z_{i-1}: the lane's map
    // z_out = pa z_in + pb, composed with the lanes before it
    float pa = 1.f, pb = 0.f;
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        pb = f[t] - o.cM[t] * pb;
        pa = -o.cM[t] * pa;
    }
    scan_affine_up(pa, pb, lane);
    float zin = dpp<0x138>(0.f, pb);  // z before this lane's first cell (0 before cell 1)
    float ca[Q], cb[Q];
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        zin = f[t] - o.cM[t] * zin;
        const float w = __builtin_amdgcn_rcpf(o.cU[t]);
        ca[t] = -o.cC[t] * w;
        cb[t] = zin * w;
    }
    // back substitution y_i = ca_i y_{i+1} + cb_i: the lane's map
    // y_first = qa y_in + qb, composed with the lanes after it
    float qa = 1.f, qb = 0.f;
#pragma unroll
    for (int t = Q - 1; t >= 0; --t) {
        qb = ca[t] * qb + cb[t];
        qa = ca[t] * qa;
    }
    scan_affine_down(qa, qb, lane);
    float yin = dpp<0x130>(0.f, qb);  // y after this lane's last cell (0 after cell nx)
#pragma unroll
    for (int t = Q - 1; t >= 0; --t) {
        yin = ca[t] * yin + cb[t];
        y[t] = o.X[t] + omega * (yin - o.X[t]);
    }
    return rs;
}

template <int Q, bool FULL>
__device__ inline void line_store(float* __restrict__ xw, const float (&y)[Q], int i0, int nx) {
    if (FULL && Q % 2 == 0) {
#pragma unroll
        for (int t = 0; t < Q; t += 2) *(float2*)(xw + t) = make_float2(y[t], y[t + 1]);
    } else {
#pragma unroll
        for (int t = 0; t < Q; ++t)
            if (FULL || i0 + t < nx) xw[t] = y[t];
    }
}

// One half sweep of red-black line SOR: every line of colour c (lines
// (j,k) with (j+k) parity c) is solved exactly for the current values of its
// four neighbour lines,
//   T y = f,  f = b - D x(j-1) - E x(j+1) - F x(k-1) - G x(k+1),
// with T the line's tridiagonal part (B, A, C) in its LU form (m, u, C):
// z_i = f_i - m_i z_{i-1}, y_i = (z_i - C_i y_{i+1}) / u_i; then
// x <- x + omega (y - x).  One wavefront per line: lane l holds the Q cells
// l Q .. l Q + Q - 1 (0-based i - 1), so both recurrences are affine maps a
// lane composes over its own cells and the wavefront combines with a scan
// (forward over lanes, then backward).  The residual of the line before its
// update, f - L U x, is summed in |.| for the convergence test, as the point
// sweep sums its residuals.  Cells past nx act as identity rows with zero
// right-hand side.  A wavefront solves two of its lines at once (lines of
// one colour never read each other), so that their two dependent chains of
// lane moves overlap.  Rows, blocks, the test block and the partials are
// laid out as in k_sweep.
template <int Q, bool FULL>
__global__ __launch_bounds__(256) void k_line_sweep(int c, const float* __restrict__ coef, float* __restrict__ x, LS ls,
                                                    float omega, double* part, int nb, int lines_per_block, Ctrl* ctrl,
                                                    int rev, FinArgs fa) {
    __shared__ double sh[4];
    if ((int)blockIdx.x == nb) {
        finish_block(ctrl, fa.part, fa.nparts, fa.step, sh, 1);
        return;
    }
    const int done = ctrl->done;
    int lb = xcd_block(nb);
    if (rev) lb = nb - 1 - lb;
    // the wavefront index, uniform: the line arithmetic and the loop run on
    // the scalar unit
    const int lane = threadIdx.x & 63, ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nx = ls.nx, ny = ls.ny, hy = ny >> 1;
    const int nl = hy * ls.nz;
    const int q0 = lb * lines_per_block;
    const int q1 = min(q0 + lines_per_block, nl);
    const int i0 = lane * Q;
    double acc = 0.0;
    for (int q = done ? q1 : q0 + ty; q < q1; q += 8) {
        const bool two = q + 4 < q1;  // wavefront-uniform
        const int qb = two ? q + 4 : q;
        LineOps<Q> oa, ob;
        line_load<Q, FULL>(oa, coef, x, ls, c, q, i0);
        line_load<Q, FULL>(ob, coef, x, ls, c, qb, i0);
        float ya[Q], yb[Q];
        const float ra = line_solve<Q>(oa, omega, lane, ya);
        const float rb = line_solve<Q>(ob, omega, lane, yb);
        const int ka = q / hy, ja = 2 * (q - ka * hy) + ((ka + c) & 1);
        line_store<Q, FULL>(x + (long long)(ja + ny * (ka + 1)) * nx + i0, ya, i0, nx);
        acc += (double)ra;
        if (two) {
            const int kb = qb / hy, jb = 2 * (qb - kb * hy) + ((kb + c) & 1);
            line_store<Q, FULL>(x + (long long)(jb + ny * (kb + 1)) * nx + i0, yb, i0, nx);
            acc += (double)rb;
        }
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0 && !done) part[c * nb + lb] = bs;
}

// FP64 lane moves and reads (two 32-bit halves each).
template <int CTRL>
__device__ inline double dppd(double old, double v) {
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double rdlaned(double v, int l) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, l), hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 2x2 matrices [[a, b], [c, d]] for the pivot recurrence, kept scaled by a
// power of two (exact) so that long products neither overflow nor underflow.
struct M2 {
    double a, b, c, d;
};
__device__ inline M2 mmul(const M2& x, const M2& y) {
    return M2{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
__device__ inline M2 mnorm(const M2& x) {
    const double s = fmax(fmax(fabs(x.a), fabs(x.b)), fmax(fabs(x.c), fabs(x.d)));
    const int e = ilogb(s);
    return M2{ldexp(x.a, -e), ldexp(x.b, -e), ldexp(x.c, -e), ldexp(x.d, -e)};
}
template <int CTRL>
__device__ inline M2 mdpp(const M2& x) {  // identity where the source lane is outside
    return M2{dppd<CTRL>(1.0, x.a), dppd<CTRL>(0.0, x.b), dppd<CTRL>(0.0, x.c), dppd<CTRL>(1.0, x.d)};
}
__device__ inline M2 mread(const M2& x, int l) {
    return M2{rdlaned(x.a, l), rdlaned(x.b, l), rdlaned(x.c, l), rdlaned(x.d, l)};
}

// The line system from the FP64 colour-split one, one wavefront per line
// (lane l holds cells l Q .. l Q + Q - 1): the couplings D..G, and the LU
// factors of the tridiagonal part, u_1 = A_1, m_i = B_i / u_{i-1},
// u_i = A_i - m_i C_{i-1}, all stored FP32.  B_1 and C_nx couple to the r
// ghosts, whose corrections are zero, and drop out.  The pivots are ratios
// of leading principal minors, u_i = p_i / p_{i-1}, with
// (p_i, p_{i-1}) = M_i (p_{i-1}, p_{i-2}), M_i = [[A_i, -B_i C_{i-1}], [1, 0]],
// so each is read from a prefix product of the M_i: a lane multiplies its
// own cells', the wavefront scans the lanes' products (four shifted
// multiplications in each 16-lane row, then the row carries), and the lane
// steps through its cells from the product before them.  In FP64 the
// pivots agree with the sequential recurrence to about 1e-14 relative.
template <int Q, bool FULL>
__global__ __launch_bounds__(256) void k_line_factor(CS cs, LS ls, const double* __restrict__ C0,
                                                     const double* __restrict__ C1, float* __restrict__ lco) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // line (jj, kk), r = jj + ny kk
    if (r >= ls.ny * ls.nz) return;                      // whole wavefronts leave together
    const int kk = r / ls.ny, jj = r - kk * ls.ny;
    const int hx = cs.hx, nx = ls.nx;
    const int i0 = lane * Q;
    const double* __restrict__ b0 = C0 + (long long)r * 8 * hx;
    const double* __restrict__ b1 = C1 + (long long)r * 8 * hx;
    double va[Q], vb[Q], vc[Q], cpl[7][Q];
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        const int idx = i0 + t;  // cell i = idx + 1, colour (i + j + k) & 1
        const bool in = FULL || idx < nx;
        const double* __restrict__ bb = (((idx + 1 + jj + kk) & 1) ? b1 : b0) + (idx >> 1);
        va[t] = in ? bb[0] : 1.0;
        vb[t] = in ? bb[hx] : 0.0;
        vc[t] = in ? bb[2 * hx] : 0.0;
#pragma unroll
        for (int v = 3; v < 7; ++v) cpl[v][t] = in ? bb[v * hx] : 0.0;
    }
    // C of the cell before this lane's first
    const double cbefore = dppd<0x138>(0.0, vc[Q - 1]);
    M2 P{1.0, 0.0, 0.0, 1.0};
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        const double cprev = t ? vc[t - 1] : cbefore;
        const M2 m{va[t], i0 + t == 0 ? 0.0 : -vb[t] * cprev, 1.0, 0.0};
        P = mnorm(mmul(m, P));
    }
    // inclusive scan: lane l gets M(lane l) ... M(lane 0)
#define TT_MS(N) P = mnorm(mmul(P, mdpp<0x110 + N>(P)));
    TT_MS(1) TT_MS(2) TT_MS(4) TT_MS(8)
#undef TT_MS
    {
        const M2 t0 = mread(P, 15), t1 = mread(P, 31), t2 = mread(P, 47);
        const M2 t10 = mnorm(mmul(t1, t0));
        const M2 t210 = mnorm(mmul(t2, t10));
        const int row = lane >> 4;
        const M2 cy = row == 0 ? M2{1.0, 0.0, 0.0, 1.0} : row == 1 ? t0 : row == 2 ? t10 : t210;
        P = mnorm(mmul(P, cy));
    }
    // the product over the cells before this lane's first (identity on lane 0)
    M2 cur = mdpp<0x138>(P);
    double uprev = cur.a / cur.c;  // u of the cell before (unused for cell 1)
    float* __restrict__ o = lco + line_row(jj, kk, ls.ny, ls.nz) * 8 * nx + i0;
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        const double cprev = t ? vc[t - 1] : cbefore;
        const M2 m{va[t], i0 + t == 0 ? 0.0 : -vb[t] * cprev, 1.0, 0.0};
        cur = mnorm(mmul(m, cur));
        const double u = cur.a / cur.c;
        const double mm = i0 + t == 0 ? 0.0 : vb[t] / uprev;
        uprev = u;
        if (FULL || i0 + t < nx) {
            o[t] = (float)cpl[3][t];
            o[nx + t] = (float)cpl[4][t];
            o[2 * nx + t] = (float)cpl[5][t];
            o[3 * nx + t] = (float)cpl[6][t];
            o[4 * nx + t] = (float)mm;
            o[5 * nx + t] = (float)u;
            o[6 * nx + t] = (float)vc[t];
        }
    }
}

template <int Q>
static void line_factor_q(int nl) {
    const int grid = (nl + 3) / 4;
    if (g_ls.nx == 64 * Q)
        k_line_factor<Q, true><<<grid, 256, 0, L.s>>>(g_cs, g_ls, c64[0], c64[1], g_lco);
    else
        k_line_factor<Q, false><<<grid, 256, 0, L.s>>>(g_cs, g_ls, c64[0], c64[1], g_lco);
}

// ---------------------------------------------------------- matrix fill
// FILL_SOR_MATRIX (ppsor.f:189-277) + B_VECTOR_BOUNDARIES (ppsor.f:280-329)
// for one rank without poles: identity rows at i >= upperbnd, i = nx, k = 1,
// k = nz; couplings cut next to them.  Writes the ghosted SOR_A..G, RHS_ER.
__global__ void k_fill_matrix(Geo g, const int32_t* __restrict__ ub, const double* __restrict__ e1,
                              const double* __restrict__ e2, const double* __restrict__ e3,
                              const double* __restrict__ e4, const double* __restrict__ e5,
                              const double* __restrict__ e6, const double* __restrict__ e7,
                              const double* __restrict__ e8, const double* __restrict__ e16, double* SA,
                              double* SB, double* SC, double* SD, double* SE, double* SF, double* SG,
                              double* RHS) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long n = g.idx(i, j, k);
    double E8 = e8[n];
    double a = e7[n] + E8 * e16[n], b = -E8 * e2[n], c = -E8 * e1[n], d = -E8 * e4[n], e = -E8 * e3[n],
           f = -E8 * e6[n], gg = -E8 * e5[n];
    int u = ub[g.idx2(j, k)];
    bool ident = (i >= u) || (i == g.nx) || (k == 1) || (k == g.nz);
    if (ident) { a = 1.0; b = c = d = e = f = gg = 0.0; }
    if (i == u - 1 || i == g.nx - 1) c = 0.0;
    if (k == 2) f = 0.0;
    if (k == g.nz - 1) gg = 0.0;
    SA[n] = a; SB[n] = b; SC[n] = c; SD[n] = d; SE[n] = e; SF[n] = f; SG[n] = gg;
    if (ident) RHS[n] = 0.0;
}

// ghosted SOR_A..G, RHS -> colour-split row-blocked FP64, plus per-block
// sums of |RHS|.  One thread per cell pair (2p+1, 2p+2) of a row: the two
// cells are adjacent in the ghosted arrays and land on slot p of the two
// colours (the odd cell has s = 0).
__global__ __launch_bounds__(256) void k_pack(Geo g, CS cs, const double* __restrict__ SA,
                                              const double* __restrict__ SB, const double* __restrict__ SC,
                                              const double* __restrict__ SD, const double* __restrict__ SE,
                                              const double* __restrict__ SF, const double* __restrict__ SG,
                                              const double* __restrict__ RHS, double* C0, double* C1, double* part) {
    __shared__ double sh[4];
    const int hx = cs.hx;
    const long long tot = cs.ncell;  // pairs = cells per colour
    double acc = 0.0;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < tot; t += (long long)gridDim.x * 256) {
        const int e = (int)t;
        const int r = e / hx;
        const int p = e - r * hx;
        const int kk = r / cs.ny;
        const int j = r - kk * cs.ny + 1, k = kk + 1;
        const int codd = (j + k + 1) & 1;  // colour of cell i = 2p+1
        const long long n = g.idx(2 * p + 1, j, k);
        double* co = (codd ? C1 : C0) + (long long)r * 8 * hx + p;
        double* ce = (codd ? C0 : C1) + (long long)r * 8 * hx + p;
        const double* src[8] = {SA, SB, SC, SD, SE, SF, SG, RHS};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double vo = src[q][n], ve = src[q][n + 1];
            co[q * hx] = vo;
            ce[q * hx] = ve;
            if (q == 7) acc += fabs(vo) + fabs(ve);
        }
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// padded unknowns of colour c from a ghosted field (every slot, pads included)
__global__ void k_load_x(Geo g, CS cs, int c, const double* __restrict__ fld, double* x, int zero_ghosts) {
    long long n = (long long)cs.px * cs.ny * (cs.nz + 2);
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int tt = (int)t;
    int r = tt / cs.px;
    int hp = tt - r * cs.px;
    int kq = r / cs.ny;
    int j = r - kq * cs.ny + 1, k = kq;
    int h = hp - 2;
    int s = (j + k + c + 1) & 1;
    int i = 2 * h + 1 + s;
    bool interior = (h >= 0 && h < cs.hx && k >= 1 && k <= cs.nz);
    double v = 0.0;
    if (h >= -1 && h <= cs.hx && (interior || !zero_ghosts)) {
        if (i >= -1 && i <= g.nx + 2) v = fld[g.idx(i, j, k)];
    }
    x[t] = v;
}

// unknowns of both colours -> interior of a ghosted field
__global__ void k_store_x(Geo g, CS cs, const double* __restrict__ x0, const double* __restrict__ x1, double* fld) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * cs.ncell) return;
    int c = t >= cs.ncell;
    int e = (int)(t - c * cs.ncell);
    int r = e / cs.hx;
    int h = e - r * cs.hx;
    int kq = r / cs.ny;
    int j = r - kq * cs.ny + 1, k = kq + 1;
    int s = (j + k + c + 1) & 1;
    int i = 2 * h + 1 + s;
    const double* x = c ? x1 : x0;
    fld[g.idx(i, j, k)] = x[(long long)cs.px * ((j - 1) + (long long)cs.ny * k) + h + 2];
}

// unknowns of both colours -> a whole ghosted field: the interior from x,
// every other cell 0 (a zero fill and k_store_x in one pass)
// (fld2, when set, receives the same values: the first Newton correction
// kept for the next step's warm start, whose phi ghosts no reader uses)
__global__ void k_store_x_box(Geo g, CS cs, const double* __restrict__ x0, const double* __restrict__ x1,
                              double* fld, double* fld2, const int* run) {
    if (skip_run(run)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    double v = 0.0;
    if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) {
        const double* x = ((i + j + k) & 1) ? x1 : x0;
        v = x[(long long)cs.px * ((j - 1) + (long long)cs.ny * k) + ((i - 1) >> 1) + 2];
    }
    st_nt(&fld[t], v);
    if (fld2) st_nt(&fld2[t], v);
}

// PLANETSHIFT for one rank (ppshift.f:1009-1214): periodic phi ghosts for
// k = 1..nz, i = low..nx+2; skipm1 for arrays without a j = -1 plane.
__global__ void k_planetshift(double* a, Geo g, int low, int skipm1, const int* run) {
    if (skip_run(run)) return;
    int ni = g.nx + 2 - low + 1;
    long long n = (long long)ni * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % ni) + low, k = (int)(t / ni) + 1;
    a[g.idx(i, 0, k)] = a[g.idx(i, g.ny, k)];
    if (!skipm1) a[g.idx(i, -1, k)] = a[g.idx(i, g.ny - 1, k)];
    a[g.idx(i, g.ny + 1, k)] = a[g.idx(i, 1, k)];
    a[g.idx(i, g.ny + 2, k)] = a[g.idx(i, 2, k)];
}

int planetshift_impl(double* a, int f) {
    int low = (f == 3) ? 0 : -1;
    int skip = (f == 4 || f == 11) ? 1 : 0;
    long long n = (long long)(L.g.nx + 3 - low) * L.g.nz;
    k_planetshift<<<nblk(n, 256), 256, 0, L.s>>>(a, L.g, low, skip, L.run_if);
    TT_LAUNCH();
    return 0;
}

// FP64 true residual r = RHS - A x for both colours (defect correction and
// the guess check); per-block sums of |r| (folded into ctrl->scratch[0]);
// optionally the residual in FP32 into the RHS slot of the FP32 operator.
// With guard set it does nothing once the outer loop has ended.
__global__ __launch_bounds__(256) void k_resid64(CS cs, const double* __restrict__ C0, const double* __restrict__ C1,
                                                  const double* x0, const double* x1, float* r32_0, float* r32_1,
                                                  float* rline, int nx, double* part, int chunk,
                                                  const Ctrl* guard) {
    __shared__ double sh[4];
    if (guard && guard->outer_done) return;  // a queued correction after the loop has ended
    const int lb = blockIdx.x;
    long long start = (long long)lb * chunk, end = start + chunk;
    if (end > 2 * cs.ncell) end = 2 * cs.ncell;
    double acc = 0.0;
    const int hx = cs.hx, ny = cs.ny, px = cs.px;
    for (long long t = start + threadIdx.x; t < end; t += 256) {
        int c = t >= cs.ncell;
        int e = (int)(t - c * cs.ncell);
        int rr = e / hx;
        int h = e - rr * hx;
        int kk = rr / ny;
        int jj = rr - kk * ny;
        int s = (jj + kk + c + 1) & 1;
        const double* xc = c ? x1 : x0;
        const double* xo = c ? x0 : x1;
        const double* q = (c ? C1 : C0) + (long long)rr * 8 * hx + h;
        long long row = (long long)px * (jj + (long long)ny * (kk + 1));
        int jm = jj == 0 ? ny - 1 : jj - 1, jp = jj == ny - 1 ? 0 : jj + 1;
        double ax = q[0] * xc[row + h + 2] + q[hx] * xo[row + h + s + 1] + q[2 * hx] * xo[row + h + s + 2] +
                    q[3 * hx] * xo[(long long)px * (jm + (long long)ny * (kk + 1)) + h + 2] +
                    q[4 * hx] * xo[(long long)px * (jp + (long long)ny * (kk + 1)) + h + 2] +
                    q[5 * hx] * xo[(long long)px * (jj + (long long)ny * kk) + h + 2] +
                    q[6 * hx] * xo[(long long)px * (jj + (long long)ny * (kk + 2)) + h + 2];
        double res = q[7 * hx] - ax;
        acc += fabs(res);
        if (r32_0) (c ? r32_1 : r32_0)[(long long)rr * 8 * hx + 7 * hx + h] = (float)res;
        if (rline) rline[(line_row(jj, kk, ny, cs.nz) * 8 + 7) * nx + 2 * h + s] = (float)res;  // cell i = 2h + 1 + s
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[lb] = bs;
}

// dot products of the guess check: per block sum RHS * (A x) and (A x)^2
// over both colours (A x = RHS - r)
__global__ __launch_bounds__(256) void k_guess_dots(CS cs, const double* __restrict__ C0,
                                                     const double* __restrict__ C1, const double* x0,
                                                     const double* x1, double* part, int chunk, int nb) {
    __shared__ double sh[4];
    const int lb = blockIdx.x;
    long long start = (long long)lb * chunk, end = start + chunk;
    if (end > 2 * cs.ncell) end = 2 * cs.ncell;
    double s1 = 0.0, s2 = 0.0;
    const int hx = cs.hx, ny = cs.ny, px = cs.px;
    for (long long t = start + threadIdx.x; t < end; t += 256) {
        int c = t >= cs.ncell;
        int e = (int)(t - c * cs.ncell);
        int rr = e / hx;
        int h = e - rr * hx;
        int kk = rr / ny;
        int jj = rr - kk * ny;
        int s = (jj + kk + c + 1) & 1;
        const double* xc = c ? x1 : x0;
        const double* xo = c ? x0 : x1;
        const double* q = (c ? C1 : C0) + (long long)rr * 8 * hx + h;
        long long row = (long long)px * (jj + (long long)ny * (kk + 1));
        int jm = jj == 0 ? ny - 1 : jj - 1, jp = jj == ny - 1 ? 0 : jj + 1;
        double ax = q[0] * xc[row + h + 2] + q[hx] * xo[row + h + s + 1] + q[2 * hx] * xo[row + h + s + 2] +
                    q[3 * hx] * xo[(long long)px * (jm + (long long)ny * (kk + 1)) + h + 2] +
                    q[4 * hx] * xo[(long long)px * (jp + (long long)ny * (kk + 1)) + h + 2] +
                    q[5 * hx] * xo[(long long)px * (jj + (long long)ny * kk) + h + 2] +
                    q[6 * hx] * xo[(long long)px * (jj + (long long)ny * (kk + 2)) + h + 2];
        s1 += q[7 * hx] * ax;
        s2 += ax * ax;
    }
    double b1 = block_sum256(s1, sh);
    double b2 = block_sum256(s2, sh);
    if (threadIdx.x == 0) {
        part[lb] = b1;
        part[nb + lb] = b2;
    }
}

__global__ void k_to_f32(const double* __restrict__ a, float* b, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (float)a[t];
}

// x64 (padded) += x32 (padded) over the interior slots
__global__ void k_add_corr(CS cs, double* x, const float* __restrict__ c) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cs.ncell) return;
    int tt = (int)t;
    int r = tt / cs.hx;
    int h = tt - r * cs.hx;
    int kk = r / cs.ny;
    int jj = r - kk * cs.ny;
    long long p = (long long)cs.px * (jj + (long long)cs.ny * (kk + 1)) + h + 2;
    x[p] = x[p] + (double)c[p];
}

// The residual of a zero start, RHS itself, in FP32 into the point system's
// RHS slots (r32_0/1) or the line system's (rline).
__global__ void k_rhs_f32(CS cs, const double* __restrict__ C0, const double* __restrict__ C1, float* r32_0,
                          float* r32_1, float* rline, int nx) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * cs.ncell) return;
    const int c = t >= cs.ncell;
    const int e = (int)(t - c * cs.ncell);
    const int hx = cs.hx;
    const int rr = e / hx, h = e - rr * hx;
    const int kk = rr / cs.ny, jj = rr - kk * cs.ny;
    const long long o = (long long)rr * 8 * hx + 7 * hx + h;
    const double v = (c ? C1 : C0)[o];
    if (r32_0) (c ? r32_1 : r32_0)[o] = (float)v;
    if (rline) rline[(line_row(jj, kk, cs.ny, cs.nz) * 8 + 7) * nx + 2 * h + ((jj + kk + c + 1) & 1)] = (float)v;
}

// x64 (padded, both colours) += the line unknowns over the interior cells,
// which are then zeroed: the next correction's inner solve starts from 0.
// Nothing once the outer loop has ended.
__global__ void k_add_corr_line(CS cs, int nx, double* x0, double* x1, float* xl, const Ctrl* guard) {
    if (guard->outer_done) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * cs.ncell) return;
    const int cc = t >= cs.ncell;
    double* x = cc ? x1 : x0;
    int tt = (int)(t - cc * cs.ncell);
    int r = tt / cs.hx;
    int h = tt - r * cs.hx;
    int kk = r / cs.ny;
    int jj = r - kk * cs.ny;
    int s = (jj + kk + cc + 1) & 1;
    long long row = jj + (long long)cs.ny * (kk + 1);
    long long p = (long long)cs.px * row + h + 2;
    const long long q = row * nx + 2 * h + s;
    x[p] = x[p] + (double)xl[q];
    xl[q] = 0.0f;
}

__global__ void k_scale(double* a, double s, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = a[t] * s;
}

// One step of the defect correction's outer loop, on the device, in the
// stream after the FP64 residual of the latest correction (folded into
// scratch[0]; first = 1: the zero start, whose residual is RHS with norm
// rhsnorm; first = 2: a guess, residual in scratch[0]).  It stops the loop
// when the FP64 L1 ratio meets eps64, when a correction made no progress or
// at a cap, and otherwise sets up the next inner solve: its tolerance is
// what takes the FP64 residual below the target with a factor 2 margin,
// max(inner_tol, min(0.5 eps64 sum|RHS| / r, 0.5)).  When the loop stops,
// done is set, so the inner solve's sweeps the host has queued return at once.
__device__ inline void outer_step_body(Ctrl* c, int first, double eps64, double inner_tol, int inner_max, int stag,
                                       int max_outer, int itmax, int restore) {
    if (first) {
        c->rhs64 = c->rhsnorm;
        c->outer = 0;
        c->total = 0;
        c->outer_done = 0;
        for (int q = 0; q < 16; ++q) c->hist[q] = 0;
    } else {
        c->total += c->it;  // the correction just finished
        if (c->outer - 1 < 16) c->hist[c->outer - 1] = c->it;
    }
    const double rn = first == 1 ? c->rhs64 : c->scratch[0];
    int stop = 0;
    if (rn / c->rhs64 <= eps64) stop = 1;
    else if (!first && !(rn < c->rn64)) stop = 2;
    else if (c->outer >= max_outer || c->total >= itmax) stop = 3;
    c->rn64 = rn;
    if (stop) {
        c->outer_done = stop;
        c->done = 1;
        if (restore) {  // the FP64 solve's fields, as the reference's solve leaves them
            c->rhsnorm = c->rhs64;
            c->it = c->total;
            c->done = (rn / c->rhs64 <= eps64) ? 1 : 2;
        }
        return;
    }
    c->outer++;
    c->done = 0;
    c->it = 0;
    c->counter = 0;
    c->counter2 = 0;
    c->itmin = 1;
    c->itmax = inner_max;
    c->best = 1e300;
    c->since = 0;
    c->stag = stag;
    c->spec_hit = 0;
    const double need = 0.5 * eps64 * c->rhs64 / rn;
    c->eps = need > inner_tol ? (need < 0.5 ? need : 0.5) : inner_tol;
    c->rhsnorm = rn;
}

__global__ void k_outer_step(Ctrl* c, int first, double eps64, double inner_tol, int inner_max, int stag,
                             int max_outer, int itmax, int restore) {
    if (threadIdx.x != 0) return;
    outer_step_body(c, first, eps64, inner_tol, inner_max, stag, max_outer, itmax, restore);
}

// The fold of k_resid64's block sums into scratch[0] (k_fold's order, so
// the same bits) and the outer step after it, in one launch; nothing once
// the loop has ended, so corrections queued beyond its end cost a launch.
__global__ __launch_bounds__(256) void k_outer_fold(Ctrl* c, const double* __restrict__ part, int nb, double eps64,
                                                    double inner_tol, int inner_max, int stag, int max_outer,
                                                    int itmax) {
    __shared__ double sh[4];
    if (c->outer_done) return;
    double s = 0.0;
    for (int q = threadIdx.x; q < nb; q += 256) s = s + part[q];
    s = block_reduce256(s, sh, 0);
    if (threadIdx.x != 0) return;
    c->scratch[0] = s;
    outer_step_body(c, 0, eps64, inner_tol, inner_max, stag, max_outer, itmax, 1);
}

// ------------------------------------------------------------- host side
static hipEvent_t ev_get() {
    hipEvent_t e;
    if (!ev_pool.empty()) {
        e = ev_pool.back();
        ev_pool.pop_back();
    } else {
        hipEventCreate(&e);
    }
    return e;
}

// Fold the pending event pairs of batches <= upto into the totals; st is
// the control block after those batches.  Launches queued after the solve
// had finished (iteration >= the completed count once done is set) return
// at once and are left out.
static void ev_drain(const Ctrl* st, long long upto) {
    const bool fin = st && st->done != 0;
    const int itc = st ? st->it : 0;
    size_t keep = 0;
    for (size_t u = 0; u < ev_kind.size(); ++u) {
        hipEvent_t a = ev_pend[2 * u], b = ev_pend[2 * u + 1];
        if (ev_bid[u] > upto) {
            ev_pend[2 * keep] = a;
            ev_pend[2 * keep + 1] = b;
            ev_kind[keep] = ev_kind[u];
            ev_iter[keep] = ev_iter[u];
            ev_bid[keep] = ev_bid[u];
            ev_outer[keep] = ev_outer[u];
            ++keep;
            continue;
        }
        // a queued correction's sweep ran if its iteration is below that
        // correction's count (corrections queued after the loop ended have 0)
        const int o = ev_outer[u];
        const bool ran = o >= 0 ? (st && o < 16 && ev_iter[u] < st->hist[o]) : !(fin && ev_iter[u] >= itc);
        if (ran) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, a, b);
            if (ev_kind[u] == 64) { L.ms64 += ms; L.n64++; }
            else { L.ms32 += ms; L.n32++; }
        }
        ev_pool.push_back(a);
        ev_pool.push_back(b);
    }
    ev_pend.resize(2 * keep);
    ev_kind.resize(keep);
    ev_iter.resize(keep);
    ev_bid.resize(keep);
    ev_outer.resize(keep);
}

template <int Q, bool FULL>
static void line_go(bool timed, hipEvent_t e0, hipEvent_t e1, int grid, int c, float omega, int rev, double* part,
                    FinArgs fa) {
    if (!timed)
        k_line_sweep<Q, FULL><<<grid, 256, 0, L.s>>>(c, g_lco, g_lx, g_ls, omega, part, g_lnb, g_lchunk, L.ctrl, rev,
                                                     fa);
    else
        hipExtLaunchKernelGGL(k_line_sweep<Q, FULL>, dim3(grid), dim3(256), 0, L.s, e0, e1, 0, c,
                              (const float*)g_lco, g_lx, g_ls, omega, part, g_lnb, g_lchunk, L.ctrl, rev, fa);
}

template <int Q>
static void line_go_q(bool timed, hipEvent_t e0, hipEvent_t e1, int grid, int c, float omega, int rev, double* part,
                      FinArgs fa) {
    if (g_ls.nx == 64 * Q) line_go<Q, true>(timed, e0, e1, grid, c, omega, rev, part, fa);
    else line_go<Q, false>(timed, e0, e1, grid, c, omega, rev, part, fa);
}

// A half sweep of the line solver (same partials, test block and timing as
// the point sweeps); Q = cells per lane, nx <= 64 Q.
static int launch_line(int c, double omega, int rev, int iter, double* part, FinArgs fa, int grid, bool timed) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {
        e0 = ev_get();
        e1 = ev_get();
    }
    const float w = (float)omega;
    const int q = (g_ls.nx + 63) / 64;
    if (q <= 1) line_go_q<1>(timed, e0, e1, grid, c, w, rev, part, fa);
    else if (q == 2) line_go_q<2>(timed, e0, e1, grid, c, w, rev, part, fa);
    else if (q == 3) line_go_q<3>(timed, e0, e1, grid, c, w, rev, part, fa);
    else if (q == 4) line_go_q<4>(timed, e0, e1, grid, c, w, rev, part, fa);
    else if (q <= 6) line_go_q<6>(timed, e0, e1, grid, c, w, rev, part, fa);
    else line_go_q<8>(timed, e0, e1, grid, c, w, rev, part, fa);
    TT_LAUNCH();
    if (timed) {
        ev_pend.push_back(e0);
        ev_pend.push_back(e1);
        ev_kind.push_back(32);
        ev_iter.push_back(iter);
        ev_bid.push_back(g_bid);
        ev_outer.push_back(g_outer_tag);
    }
    return 0;
}

// Half sweep c of iteration iter (0-based within the solve), in place.  The
// partials of iteration iter go to slot iter % 2 (two colours x nb each).
// fin: the red sweep also runs the convergence test of iteration iter - 1.
template <typename T>
static int launch_sweep(int c, T* coef0, T* coef1, T* xa, T* xb, double omega, int rev, int iter, int fin,
                        int step) {
    T* q = c ? coef1 : coef0;
    T* xc = c ? xb : xa;
    const T* xo = c ? xa : xb;
    double* part = L.partial + (iter & 1) * 2 * g_nb;
    FinArgs fa{L.partial + ((iter + 1) & 1) * 2 * g_nb, 2 * g_pn, step, fin};
    const int grid = (g_line && sizeof(T) == 4 ? g_lnb : g_nb) + (fin ? 1 : 0);
    const bool timed = L.timing && (L.timing_count++ % L.timing_period) == 0;
    const bool pairs = sizeof(T) == 4 && g_cs.hx % 2 == 0;
    if (sizeof(T) == 4 && g_line) return launch_line(c, omega, rev, iter, part, fa, grid, timed);
    if (!timed) {
        // the plain launch path: the host's cost per launch bounds how fast
        // queued sweeps that return at once (after convergence) drain
        if (pairs)
            k_sweep_f32x2<<<grid, 256, 0, L.s>>>(c, (const float*)q, (float*)xc, (const float*)xo, g_cs, (float)omega,
                                                 part, g_nb, g_chunk, L.ctrl, rev, fa);
        else
            k_sweep<T><<<grid, 256, 0, L.s>>>(c, (const T*)q, xc, xo, g_cs, (T)omega, part, g_nb, g_chunk, L.ctrl, rev,
                                              fa);
        TT_LAUNCH();
        return 0;
    }
    // timed launches: the dispatch packet itself records the start and stop
    // events, so the interval is the kernel's own execution
    hipEvent_t e0 = ev_get(), e1 = ev_get();
    if (pairs) {
        hipExtLaunchKernelGGL(k_sweep_f32x2, dim3(grid), dim3(256), 0, L.s, e0, e1, 0, c, (const float*)q,
                              (float*)xc, (const float*)xo, g_cs, (float)omega, part, g_nb, g_chunk, L.ctrl, rev, fa);
    } else {
        hipExtLaunchKernelGGL(k_sweep<T>, dim3(grid), dim3(256), 0, L.s, e0, e1, 0, c, (const T*)q, xc, xo, g_cs,
                              (T)omega, part, g_nb, g_chunk, L.ctrl, rev, fa);
    }
    TT_LAUNCH();
    ev_pend.push_back(e0);
    ev_pend.push_back(e1);
    ev_kind.push_back(sizeof(T) == 8 ? 64 : 32);
    ev_iter.push_back(iter);
    ev_bid.push_back(g_bid);
    ev_outer.push_back(g_outer_tag);
    return 0;
}

static int read_ctrl() {
    int rc = ctrl_to_host();
    if (rc) return rc;
    if (L.timing) ev_drain(L.hctrl, g_bid);
    return 0;
}

// The solve fields of the control block, set in stream order by a kernel
// (the other fields -- flags of earlier kernels not yet read -- stay as the
// device has them); the host copy gets the same values.
__global__ void k_ctrl_solve(Ctrl* c, int itmax, int itmin, double eps) {
    if (threadIdx.x != 0) return;
    c->done = 0;
    c->it = 0;
    c->itmax = itmax;
    c->itmin = itmin;
    c->eps = eps;
    c->counter = 0;
    c->counter2 = 0;
    c->stag = 0;
    c->spec_hit = 0;
}

static int set_ctrl_solve(int itmax, int itmin, double eps) {
    k_ctrl_solve<<<1, 64, 0, L.s>>>(L.ctrl, itmax, itmin, eps);
    TT_LAUNCH();
    Ctrl& c = *L.hctrl;
    c.done = 0;
    c.it = 0;
    c.itmax = itmax;
    c.itmin = itmin;
    c.eps = eps;
    c.counter = 0;
    c.counter2 = 0;
    c.stag = 0;
    c.spec_hit = 0;
    return 0;
}

// The whole control block from a host value, in stream order (the value is
// a kernel argument, captured at launch).  Only for a host copy read after
// the last kernel that sets a device flag.
__global__ void k_ctrl_put(Ctrl* dst, Ctrl v) {
    if (threadIdx.x == 0) *dst = v;
}

static int ctrl_put() {
    k_ctrl_put<<<1, 64, 0, L.s>>>(L.ctrl, *L.hctrl);
    TT_LAUNCH();
    return 0;
}

// SOR iterations on (coef, x) until ctrl->done; returns completed
// iterations.  The convergence test runs every check_every-th iteration:
// with deferred set (inner solves) inside the next red sweep, except for the
// last iteration of a batch, whose k_finish keeps the control block the host
// copies current.  With deferred unset (the reference's solve, which must
// stop exactly where the test is met) in a k_finish of its own, or, when
// speculate is set and the latest ratios say the solve cannot converge
// within the batch being queued, deferred as for inner solves; if such a
// test ends the solve after all, SOR_SPEC_MISS is returned and the caller
// reruns the solve without speculation.
static const int SOR_SPEC_MISS = -100;

// The host's view of a solve in progress: the two latest (iteration, log L1
// ratio) states it has read.
struct SolveView {
    int nknown = 0, it_a = 0, it_b = 0;
    double lr_a = 0.0, lr_b = 0.0;
    double leps = 0.0;  // log of the solve's tolerance, from the snapshots (the device may set it)
    // log-ratio fall per iteration over the two reads (0: unknown)
    double rate() const { return nknown >= 2 && it_b > it_a ? (lr_a - lr_b) / (double)(it_b - it_a) : 0.0; }
};

// Wait for the control-block copy of batch q; *done is set when the solve
// has finished (L.hctrl then holds the final state).
static int take_batch(int q, SolveView& v, bool* done) {
    TT_HIP(hipEventSynchronize(g_cev[q]));
    if (L.timing) ev_drain(g_hc[q], g_cbid[q]);
    const Ctrl* h = g_hc[q];
    *done = h->done != 0;
    if (*done) {
        *L.hctrl = *h;
        return 0;
    }
    v.leps = log(h->eps);
    if (h->it > v.it_b && h->rnorm > 0.0 && h->rhsnorm > 0.0) {
        v.it_a = v.it_b;
        v.lr_a = v.lr_b;
        v.it_b = h->it;
        v.lr_b = log(h->rnorm / h->rhsnorm);
        ++v.nknown;
    }
    return 0;
}

template <typename T>
static int run_sweeps(T* coef0, T* coef1, T* xa, T* xb, double omega, int* iters, double* rnorm,
                      int check_every, bool deferred, bool speculate = false, int max_batch = 0) {
    int batch = L.opts.batch > 0 ? L.opts.batch : 16;
    batch = ((batch + 3) / 4) * 4;  // a multiple of 4: directions alternate and checks stay aligned
    // (even suffices where every iteration is tested)
    if (max_batch > 0 && batch > max_batch) batch = max_batch;
    long long launched = 0;
    int slot = 0, pend = -1, last = 0;
    SolveView view;
    view.leps = log(L.hctrl->eps);
    bool done = false;
    for (;;) {
        // the control block's iteration cap ends the loop; a solve that has
        // run past it without the device reporting completion is an error
        if (launched > (long long)L.hctrl->itmax + 3 * batch)
            return fail(12, "SOR: device did not report completion");
        // Near the end, sweeps queued after the converging iteration return
        // at once but still cost a launch each.  From the observed rate: if
        // the solve should end within the queued work, read the batch in
        // flight before queuing more; if it should end within less than a
        // batch, queue a shorter one (a multiple of 4, with 25% + 2 spare).
        const double rate = view.rate();
        int bsize = batch;
        if (rate > 0.0) {
            const double remain = (double)view.it_b + (view.lr_b - view.leps) / rate - (double)launched;
            if (remain <= 0.0 && pend >= 0) {
                int rc = take_batch(pend, view, &done);
                if (rc) return rc;
                if (done) break;
                pend = -1;
                continue;
            }
            if (remain < batch) {
                const int want = (int)ceil(1.25 * remain) + 2;
                bsize = ((want + 3) / 4) * 4;
                if (bsize < 4) bsize = 4;
                if (bsize > batch) bsize = batch;
            }
        }
        launched += bsize;
        ++g_bid;
        bool defer = deferred;
        if (!deferred && speculate && rate != 0.0) {
            // iterations from the newer read to the end of this batch
            const double ahead = (double)(launched - view.it_b);
            const double m = L.spec_margin;
            defer = m == 0.0 || (m > 0.0 && rate > 0.0 && view.lr_b - view.leps > m * rate * ahead);
            if (defer) ++L.spec_batches;
        }
        for (int b = 0; b < bsize; ++b) {
            const int rev = (b & 1);  // bsize is even: the direction alternates across batches too
            const int iter = (int)(launched - bsize) + b;
            // the previous iteration was a test point inside this batch
            const int fin = (defer && b > 0 && b % check_every == 0) ? 1 : 0;
            int rc = launch_sweep<T>(0, coef0, coef1, xa, xb, omega, rev, iter, fin, check_every);
            if (rc) return rc;
            rc = launch_sweep<T>(1, coef0, coef1, xa, xb, omega, rev, iter, 0, check_every);
            if (rc) return rc;
            if (defer ? b == bsize - 1 : (b + 1) % check_every == 0) {
                // the batch's last test also leaves the control block in the
                // host's pinned copy for this batch (bsize is a multiple of 4)
                k_finish<<<1, 256, 0, L.s>>>(L.ctrl, L.partial + (iter & 1) * 2 * g_nb, 2 * g_pn, check_every,
                                             b == bsize - 1 ? g_hcd[slot] : nullptr);
                TT_LAUNCH();
            }
        }
        TT_HIP(hipEventRecord(g_cev[slot], L.s));
        g_cbid[slot] = g_bid;
        last = slot;
        // the previous batch's state: this batch is already queued behind it,
        // and returns at once if the previous one finished the solve
        if (pend >= 0) {
            int rc = take_batch(pend, view, &done);
            if (rc) return rc;
            if (done) break;
        }
        pend = slot;
        slot ^= 1;
    }
    if (L.timing) {
        TT_HIP(hipEventSynchronize(g_cev[last]));
        ev_drain(L.hctrl, g_cbid[last]);
    }
    *iters = L.hctrl->it;
    *rnorm = L.hctrl->rnorm;
    if (!deferred && L.hctrl->spec_hit) return SOR_SPEC_MISS;
    return 0;
}

// Which form of the system is current: 0 the module arrays (the packed copy
// must be rebuilt), 1 the packed copy (the module arrays are stale), 2 both.
// g_c32_ok: the FP32 copy's A..G match the FP64 one (its RHS slot is
// rewritten with the residual before every inner solve).
static int g_sys = 0;
static bool g_c32_ok = false;

PackedSys sor_packed() { return PackedSys{{c64[0], c64[1]}, {c32[0], c32[1]}, g_cs.hx}; }
void sor_sys_packed_written(bool with32) {
    g_sys = 1;
    g_c32_ok = with32;
    g_line_ok = false;
}
void sor_sys_module_written() {
    g_sys = 0;
    g_c32_ok = false;
    g_line_ok = false;
}

// sum |RHS| over the packed copy, pair by pair in k_pack's order (same
// blocks, same per-thread sums), so the norm is the same to the bit
__global__ __launch_bounds__(256) void k_rhsnorm(CS cs, const double* __restrict__ C0, const double* __restrict__ C1,
                                                 double* part) {
    __shared__ double sh[4];
    const int hx = cs.hx;
    const long long tot = cs.ncell;
    double acc = 0.0;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < tot; t += (long long)gridDim.x * 256) {
        const int e = (int)t;
        const int r = e / hx;
        const int p = e - r * hx;
        const int kk = r / cs.ny;
        const int j = r - kk * cs.ny + 1, k = kk + 1;
        const int codd = (j + k + 1) & 1;
        const long long o = ((long long)r * 8 + 7) * hx + p;
        const double vo = (codd ? C1 : C0)[o], ve = (codd ? C0 : C1)[o];
        acc += fabs(vo) + fabs(ve);
    }
    double bs = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// the packed copy back into the module arrays (interior cells)
__global__ void k_unpack(Geo g, CS cs, const double* __restrict__ C0, const double* __restrict__ C1, double* SA,
                         double* SB, double* SC, double* SD, double* SE, double* SF, double* SG, double* RHS) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    const int r = (j - 1) + cs.ny * (k - 1);
    const double* C = ((i + j + k) & 1) ? C1 : C0;
    const long long o = (long long)r * 8 * cs.hx + (i - 1) / 2;
    const long long n = g.idx(i, j, k);
    double* dst[8] = {SA, SB, SC, SD, SE, SF, SG, RHS};
#pragma unroll
    for (int q = 0; q < 8; ++q) dst[q][n] = C[o + (long long)q * cs.hx];
}

int sor_sys_before_download(int field) {
    const bool sys = (field >= TT_SOR_A && field <= TT_SOR_G) || field == TT_RHS_ER;
    if (!sys || g_sys != 1) return 0;
    double** f = L.f;
    long long n = (long long)L.g.nx * L.g.ny * L.g.nz;
    k_unpack<<<nblk(n, 256), 256, 0, L.s>>>(L.g, g_cs, c64[0], c64[1], f[TT_SOR_A], f[TT_SOR_B], f[TT_SOR_C],
                                            f[TT_SOR_D], f[TT_SOR_E], f[TT_SOR_F], f[TT_SOR_G], f[TT_RHS_ER]);
    TT_LAUNCH();
    g_sys = 2;
    return 0;
}

// The packed system and sum |RHS| for a solve: packed from the module
// arrays when those are newer, else the norm of the packed copy.
int sor_pack_impl() {
    double** f = L.f;
    int nb = (int)((g_cs.ncell + 255) / 256);
    if (nb > L.npartial) nb = L.npartial;
    double* part = L.partial + L.npartial;  // second half: the sweeps use the first
    if (g_sys == 0) {
        k_pack<<<nb, 256, 0, L.s>>>(L.g, g_cs, f[TT_SOR_A], f[TT_SOR_B], f[TT_SOR_C], f[TT_SOR_D], f[TT_SOR_E],
                                    f[TT_SOR_F], f[TT_SOR_G], f[TT_RHS_ER], c64[0], c64[1], part);
        g_sys = 2;
        g_c32_ok = false;
        g_line_ok = false;
    } else {
        k_rhsnorm<<<nb, 256, 0, L.s>>>(g_cs, c64[0], c64[1], part);
    }
    TT_LAUNCH();
    return fold_partials(part, nb, 1, 0, &L.ctrl->rhsnorm);
}

// to: 0 the norm only, 1 also the FP32 residual into the point system's RHS
// slots, 2 into the line system's
static int resid64(double* out_norm, int to) {
    int nb = L.npartial;
    long long tot = 2 * g_cs.ncell;
    int chunk = (int)((tot + nb - 1) / nb);
    chunk = ((chunk + 255) / 256) * 256;
    nb = (int)((tot + chunk - 1) / chunk);
    double* part = L.partial + L.npartial;
    k_resid64<<<nb, 256, 0, L.s>>>(g_cs, c64[0], c64[1], x64[0], x64[1], to == 1 ? c32[0] : nullptr,
                                   to == 1 ? c32[1] : nullptr, to == 2 ? g_lco : nullptr, g_ls.nx, part, chunk,
                                   nullptr);
    TT_LAUNCH();
    int rc = fold_partials(part, nb, 1, 0, &L.ctrl->scratch[0]);
    if (rc) return rc;
    if (!out_norm) return 0;  // stays on the device, for k_outer_step
    rc = read_ctrl();
    if (rc) return rc;
    *out_norm = L.hctrl->scratch[0];
    return 0;
}

// Inner iterations of each correction in the previous line solve (0: none
// yet), the host's forecast for the next solve.
static int g_hist_prev[16];
static int g_nouter_prev = 0;

// The line-sweep defect correction with no host read inside it.  The host
// queues whole corrections -- the inner solve's iterations, the test of its
// last one, the correction added (which zeroes the line unknowns for the
// next inner solve), the FP64 residual, and its fold with the outer step --
// and reads the control block once after each round of them.  An inner
// solve gets the iterations the same correction took in the previous solve
// plus one (6 for corrections it did not reach): its sweeps return at once
// after its own test ends it, and if it is not over by then its correction
// is added as it stands and the device's outer step judges the result as
// for any other.  The first round queues as many corrections as the
// previous solve took (a correction queued after the loop has ended costs a
// launch per kernel, about as much as one more read), later rounds one.
// The line sweeps' over-relaxation: tt_solver_opts.line_omega where it is
// set, else the value the search below chose (line_om_auto)
static double line_om() { return L.opts.line_omega > 0.0 ? L.opts.line_omega : L.line_om_auto; }

static int line_chain(int stag, int max_outer) {
    int nb = L.npartial;
    const long long tot = 2 * g_cs.ncell;
    int chunk = (int)((tot + nb - 1) / nb);
    chunk = ((chunk + 255) / 256) * 256;
    nb = (int)((tot + chunk - 1) / chunk);
    double* part = L.partial + L.npartial;
    if (int e = dev_zero(g_lx, line_x_len() * sizeof(float))) return e;
    g_line = 1;
    g_pn = g_lnb;
    int rc = 0, o = 0;
    for (int round = 0; rc == 0; ++round) {
        const int nq = round == 0 ? (g_nouter_prev > 0 ? g_nouter_prev : 2) : 1;
        ++g_bid;
        for (int q = 0; q < nq && rc == 0; ++q, ++o) {
            const int ni = (o < 16 && g_hist_prev[o] > 0) ? g_hist_prev[o] + 1 : 6;
            g_outer_tag = o;
            for (int b = 0; b < ni && rc == 0; ++b) {
                // every iteration tested, inside the next red sweep
                rc = launch_sweep<float>(0, c32[0], c32[1], x32[0], x32[1], line_om(), b & 1, b, b > 0 ? 1 : 0, 1);
                if (rc == 0) rc = launch_sweep<float>(1, c32[0], c32[1], x32[0], x32[1], line_om(), b & 1, b, 0, 1);
            }
            if (rc) break;
            k_finish<<<1, 256, 0, L.s>>>(L.ctrl, L.partial + ((ni - 1) & 1) * 2 * g_nb, 2 * g_pn, 1, nullptr);
            TT_LAUNCH();
            k_add_corr_line<<<nblk(2 * g_cs.ncell, 256), 256, 0, L.s>>>(g_cs, g_ls.nx, x64[0], x64[1], g_lx, L.ctrl);
            TT_LAUNCH();
            k_resid64<<<nb, 256, 0, L.s>>>(g_cs, c64[0], c64[1], x64[0], x64[1], nullptr, nullptr, g_lco, g_ls.nx,
                                           part, chunk, L.ctrl);
            TT_LAUNCH();
            k_outer_fold<<<1, 256, 0, L.s>>>(L.ctrl, part, nb, L.p.epsmat, L.opts.inner_tol, L.opts.inner_max, stag,
                                             max_outer, L.p.itmax);
            TT_LAUNCH();
        }
        g_outer_tag = -1;
        if (rc) break;
        if (L.sor_tail) {  // what follows the solve, on the device once it has ended
            L.run_if = &L.ctrl->outer_done;
            rc = L.sor_tail();
            L.run_if = nullptr;
            if (rc) break;
        }
        rc = ctrl_to_host();
        if (rc) break;
        if (L.timing) ev_drain(L.hctrl, g_bid);
        if (L.hctrl->outer_done) break;
        if (o > max_outer + 4) rc = fail(12, "line defect correction: the device did not end the outer loop");
    }
    g_outer_tag = -1;
    g_line = 0;
    g_pn = g_nb;
    if (rc) return rc;
    g_nouter_prev = L.hctrl->outer;
    for (int q = 0; q < 16; ++q) g_hist_prev[q] = L.hctrl->hist[q];
    return 0;
}

// the line system of the FP64 system on the device: allocated on first use,
// factored once per system
static int line_prepare() {
    if (int rc = line_alloc()) return rc;
    if (!g_line_ok) {
        const int nl = g_ls.ny * g_ls.nz;
        const int q = (g_ls.nx + 63) / 64;
        if (q <= 1) line_factor_q<1>(nl);
        else if (q == 2) line_factor_q<2>(nl);
        else if (q == 3) line_factor_q<3>(nl);
        else if (q == 4) line_factor_q<4>(nl);
        else if (q <= 6) line_factor_q<6>(nl);
        else line_factor_q<8>(nl);
        TT_LAUNCH();
        g_line_ok = true;
    }
    return 0;
}

// The over-relaxation of the line sweeps, where tt_solver_opts.line_omega is
// 0: chosen on the first system a line solve meets after tt_init or
// tt_set_solver, as ADJUST_SOR_PARAM (ppsor.f:332-354) chooses SORPARAM for
// the point sweeps: cold-start solves of that system, each the whole defect
// correction to the FP64 target, with omega = 1.0, 1.1, ... (until the count
// exceeds 1.5 times the best, past it), then in steps of 0.02 from 0.08 below
// the fastest coarse value to 0.08 above (below the lowest and above the
// highest where several tie); the fewest line iterations win, the middle
// value of a tie (where ITMAX ends every trial first, the smallest FP64
// residual ratio reached).  The chosen omega's corrections then keep the budget its
// trial needed, as a solve keeps the previous solve's.  The best omega grows with the weight of the phi and theta
// couplings: 1.13-1.16 at 128x256x128, about 1.55 at 192x512x256.  Nothing
// here writes delta_ER; the unknowns are the caller's to set afterwards.
static int line_omega_search() {
    if (int rc = line_prepare()) return rc;
    const bool timing = L.timing;
    std::function<int()> tail;
    tail.swap(L.sor_tail);  // what follows the solve waits for the solve itself
    L.timing = false;
    int best_no = 0, best_hist[16] = {0};
    struct Trial {
        double om;
        int n;         // line iterations to the FP64 target; INT_MAX where a cap ended it first
        double ratio;  // the FP64 residual ratio it reached
    };
    std::vector<Trial> tried;
    auto better = [](const Trial& a, const Trial& b) { return a.n != b.n ? a.n < b.n : a.ratio < b.ratio; };
    auto best = [&]() -> const Trial& {
        size_t k = 0;
        for (size_t q = 1; q < tried.size(); ++q)
            if (better(tried[q], tried[k])) k = q;
        return tried[k];
    };
    long long pl = padded_len();
    auto trial = [&](double om) -> int {
        for (auto& t : tried)
            if (fabs(t.om - om) < 1e-9) return 0;
        int rc = set_ctrl_solve(L.p.itmax, 1, L.p.epsmat);
        if (!rc) rc = dev_zero(x64[0], pl * sizeof(double));
        if (!rc) rc = dev_zero(x64[1], pl * sizeof(double));
        if (rc) return rc;
        k_rhs_f32<<<nblk(2 * g_cs.ncell, 256), 256, 0, L.s>>>(g_cs, c64[0], c64[1], nullptr, nullptr, g_lco, g_ls.nx);
        TT_LAUNCH();
        k_outer_step<<<1, 64, 0, L.s>>>(L.ctrl, 1, L.p.epsmat, L.opts.inner_tol, L.opts.inner_max, 8, 200, L.p.itmax,
                                        1);
        TT_LAUNCH();
        L.line_om_auto = om;
        // each correction may run until its own test ends it (64 line
        // iterations queued per correction, two corrections per round), so
        // the count is omega's and not the previous solve's budget
        g_nouter_prev = 2;
        for (int q = 0; q < 16; ++q) g_hist_prev[q] = 63;
        if ((rc = line_chain(8, 200))) return rc;
        const Ctrl& c = *L.hctrl;
        Trial t{om, c.done == 1 ? c.total : INT_MAX, c.rhs64 > 0.0 ? c.rn64 / c.rhs64 : 0.0};
        if (tried.empty() || better(t, best())) {
            best_no = g_nouter_prev;
            for (int q = 0; q < 16; ++q) best_hist[q] = g_hist_prev[q];
        }
        tried.push_back(t);
        return 0;
    };
    int rc = 0;
    for (int c = 0; c <= 9 && rc == 0; ++c) {
        const double om = 1.0 + 0.1 * c;
        if ((rc = trial(om))) break;
        const Trial& b = best();
        if (om > b.om + 0.15 && (tried.back().n == INT_MAX || (double)tried.back().n > 1.5 * b.n)) break;
    }
    if (!rc) {
        const Trial b = best();
        double lo = b.om, hi = b.om;
        for (auto& t : tried)
            if (b.n != INT_MAX && t.n == b.n) lo = std::min(lo, t.om), hi = std::max(hi, t.om);
        for (double om = std::max(1.0, lo - 0.08); om <= std::min(1.95, hi + 0.08) + 1e-9 && rc == 0; om += 0.02)
            rc = trial(std::round(om * 100.0) / 100.0);
    }
    if (!rc) {
        // the fewest iterations, the middle omega of a tie; where a cap
        // (ITMAX) ended every trial, the smallest residual ratio
        const Trial b = best();
        std::vector<double> ties;
        for (auto& t : tried)
            if (b.n != INT_MAX ? t.n == b.n : t.om == b.om) ties.push_back(t.om);
        std::sort(ties.begin(), ties.end());
        L.line_om_auto = ties[ties.size() / 2];
        L.line_om_trials = (int)tried.size();
        g_nouter_prev = best_no;  // the budget of the chosen omega's corrections
        for (int q = 0; q < 16; ++q) g_hist_prev[q] = best_hist[q];
    }
    if (rc) L.line_om_auto = 0.0;
    L.timing = timing;
    tail.swap(L.sor_tail);
    if (rc) return rc;
    return set_ctrl_solve(L.p.itmax, 1, L.p.epsmat);
}

// mode 0: delta_ER (x from 0 or a guess held in x64, zero ghosts, ITMIN=1,
// ITMAX from the parameters); mode 1: DIFFUSION of T (x and ghosts from T,
// ITMIN=5, SORITMAX=500, ppdif.f:14-15).
int sor_solve_impl(int fill_from_evals, tt_sor_stats* st, int mode) {
    Geo& g = L.g;
    double** f = L.f;
    tt_sor_stats s{};
    if (fill_from_evals) {
        if (int e = evals_materialize()) return e;
        long long n = (long long)g.nx * g.ny * g.nz;
        k_fill_matrix<<<nblk(n, 256), 256, 0, L.s>>>(g, L.ub, f[TT_EVAL1], f[TT_EVAL2], f[TT_EVAL3], f[TT_EVAL4],
                                                     f[TT_EVAL5], f[TT_EVAL6], f[TT_EVAL7], f[TT_EVAL8],
                                                     f[TT_EVAL_1_6], f[TT_SOR_A], f[TT_SOR_B], f[TT_SOR_C],
                                                     f[TT_SOR_D], f[TT_SOR_E], f[TT_SOR_F], f[TT_SOR_G],
                                                     f[TT_RHS_ER]);
        TT_LAUNCH();
        sor_sys_module_written();
    }
    int rc = set_ctrl_solve(mode == 1 ? 500 : L.p.itmax, mode == 1 ? 5 : 1, L.p.epsmat);
    if (rc) return rc;
    rc = sor_pack_impl();
    if (rc) return rc;
    if (mode == 0 && L.opts.precision == 1 && L.opts.inner == 1 && !(line_om() > 0.0))
        if ((rc = line_omega_search())) return rc;
    long long pl = padded_len();
    double* target = mode == 1 ? f[TT_T] : f[TT_DELTA_ER];
    int prec = mode == 1 ? 0 : L.opts.precision;
    // initial unknowns
    double alpha = 0.0;
    // the start values again, for a rerun (their sources are unchanged)
    auto reload = [&]() -> int {
        for (int c = 0; c < 2; ++c) {
            if (mode == 1 || s.guess_used) {
                k_load_x<<<nblk(pl, 256), 256, 0, L.s>>>(g, g_cs, c, mode == 1 ? f[TT_T] : f[TT_DELTA_ER], x64[c],
                                                         mode == 1 ? 0 : 1);
                TT_LAUNCH();
                if (mode == 0) {
                    k_scale<<<nblk(pl, 256), 256, 0, L.s>>>(x64[c], alpha, pl);
                    TT_LAUNCH();
                }
            } else if (int e = dev_zero(x64[c], pl * sizeof(double))) {
                return e;
            }
        }
        return 0;
    };
    if (mode == 1) {
        for (int c = 0; c < 2; ++c) {
            k_load_x<<<nblk(pl, 256), 256, 0, L.s>>>(g, g_cs, c, f[TT_T], x64[c], 0);
            TT_LAUNCH();
        }
    } else {
        // cold start: zero; a guess (warm/learned) is staged in DELTA_ER by
        // the caller and accepted only if its residual is not larger
        bool use_guess = L.opts.guess != 0 && L.have_prev_guess != 0;
        if (use_guess)
            for (int c = 0; c < 2; ++c) {
                k_load_x<<<nblk(pl, 256), 256, 0, L.s>>>(g, g_cs, c, f[TT_DELTA_ER], x64[c], 1);
                TT_LAUNCH();
            }
        s.guess_used = 0;
        if (use_guess) {
            rc = read_ctrl();
            if (rc) return rc;
            s.r0_cold = L.hctrl->rhsnorm;
            // scale the guess by the factor that minimises its residual's
            // L2 norm, alpha = <b, A x> / <A x, A x>: a guess with the right
            // shape but the wrong amplitude, or rough error that A amplifies,
            // is pulled towards the cold start instead of being lost
            // two partial vectors (nb each) in the second half of the buffer
            int nb = L.npartial / 2;
            long long tot = 2 * g_cs.ncell;
            int chunk = (int)((tot + nb - 1) / nb);
            chunk = ((chunk + 255) / 256) * 256;
            nb = (int)((tot + chunk - 1) / chunk);
            double* part = L.partial + L.npartial;
            k_guess_dots<<<nb, 256, 0, L.s>>>(g_cs, c64[0], c64[1], x64[0], x64[1], part, chunk, nb);
            TT_LAUNCH();
            rc = fold_partials(part, nb, 2, 0, &L.ctrl->scratch[3], &L.ctrl->scratch[4]);
            if (rc) return rc;
            rc = read_ctrl();
            if (rc) return rc;
            double s1 = L.hctrl->scratch[3], s2 = L.hctrl->scratch[4];
            alpha = (s2 > 0.0 && s1 > 0.0) ? s1 / s2 : 0.0;
            s.guess_scale = alpha;
            if (alpha > 0.0) {
                for (int c = 0; c < 2; ++c) {
                    k_scale<<<nblk(pl, 256), 256, 0, L.s>>>(x64[c], alpha, pl);
                    TT_LAUNCH();
                }
                double r0;
                rc = resid64(&r0, 0);
                if (rc) return rc;
                s.r0_guess = r0;
                if (r0 <= s.r0_cold) s.guess_used = L.have_prev_guess;  // 1 warm, 2 learned
            }
        }
        if (!s.guess_used) {
            if ((rc = dev_zero(x64[0], pl * sizeof(double))) || (rc = dev_zero(x64[1], pl * sizeof(double)))) return rc;
        }
    }
    int iters = 0;
    double rnorm = 0.0;
    if (prec == 0) {
        rc = run_sweeps<double>(c64[0], c64[1], x64[0], x64[1], L.p.sorparam, &iters, &rnorm, 1, false,
                                L.spec_margin >= 0.0);
        if (rc == SOR_SPEC_MISS) {
            // a speculative test ended the solve one red half sweep late:
            // run it again from the same start with every test in a kernel
            // of its own
            ++L.spec_reruns;
            if ((rc = reload())) return rc;
            if ((rc = read_ctrl())) return rc;
            L.hctrl->done = 0;
            L.hctrl->it = 0;
            L.hctrl->spec_hit = 0;
            L.hctrl->rnorm = 0.0;
            if ((rc = ctrl_put())) return rc;
            rc = run_sweeps<double>(c64[0], c64[1], x64[0], x64[1], L.p.sorparam, &iters, &rnorm, 1, false, false);
        }
        if (rc) return rc;
        s.sor_iter = iters;
        s.rnorm = rnorm;
    } else {
        // FP32 SOR inside FP64 defect correction
        const bool line = L.opts.inner == 1;
        if (line) {
            if ((rc = line_prepare())) return rc;
        } else if (!g_c32_ok) {
            for (int c = 0; c < 2; ++c) {
                k_to_f32<<<nblk(8 * g_cs.ncell, 256), 256, 0, L.s>>>(c64[c], c32[c], 8 * g_cs.ncell);
                TT_LAUNCH();
            }
            g_c32_ok = true;
        }
        const int to = line ? 2 : 1;
        if (s.guess_used) {
            rc = resid64(nullptr, to);  // the guess's residual, folded into scratch[0]
            if (rc) return rc;
        } else {
            // zero start: the residual is RHS, its norm sum|RHS|
            k_rhs_f32<<<nblk(2 * g_cs.ncell, 256), 256, 0, L.s>>>(g_cs, c64[0], c64[1], line ? nullptr : c32[0],
                                                                  line ? nullptr : c32[1], line ? g_lco : nullptr,
                                                                  g_ls.nx);
            TT_LAUNCH();
        }
        // The outer loop's decisions run on the device (k_outer_step): the
        // host queues a correction's kernels and the next inner solve without
        // reading the FP64 residual first, and learns that the loop has
        // ended from the inner solve's control-block snapshots (point
        // sweeps) or from one read after all the corrections it queued
        // (line sweeps, line_chain).  FP32 sweeps stop reducing the residual
        // at their rounding floor: an inner solve hands over when its ratio
        // has not fallen by 1% over 256 iterations (8 for line sweeps, which
        // gain a decade in a few).
        const int stag = line ? 8 : 256, max_outer = 200;
        k_outer_step<<<1, 64, 0, L.s>>>(L.ctrl, s.guess_used ? 2 : 1, L.p.epsmat, L.opts.inner_tol, L.opts.inner_max,
                                        stag, max_outer, L.p.itmax, line ? 1 : 0);
        TT_LAUNCH();
        if (line) {
            rc = line_chain(stag, max_outer);
            if (rc) return rc;
        } else {
            for (;;) {
                // inner: A c = r in FP32 from c = 0, L1 during-sweep ratio vs sum|r|
                if ((rc = dev_zero(x32[0], pl * sizeof(float))) || (rc = dev_zero(x32[1], pl * sizeof(float)))) return rc;
                L.hctrl->itmax = L.opts.inner_max;  // the host's guard in run_sweeps
                int it32 = 0;
                double rn32 = 0;
                rc = run_sweeps<float>(c32[0], c32[1], x32[0], x32[1], L.p.sorparam, &it32, &rn32, 4, true);
                if (rc) return rc;
                if (L.hctrl->outer_done) break;  // the loop had ended: these sweeps returned at once
                for (int cc = 0; cc < 2; ++cc) {
                    k_add_corr<<<nblk(g_cs.ncell, 256), 256, 0, L.s>>>(g_cs, x64[cc], x32[cc]);
                    TT_LAUNCH();
                }
                rc = resid64(nullptr, to);
                if (rc) return rc;
                k_outer_step<<<1, 64, 0, L.s>>>(L.ctrl, 0, L.p.epsmat, L.opts.inner_tol, L.opts.inner_max, stag,
                                                max_outer, L.p.itmax, 0);
                TT_LAUNCH();
            }
        }
        const double rhs_norm = L.hctrl->rhs64, rn = L.hctrl->rn64;
        const int total32 = L.hctrl->total;
        s.sor_iter = total32;
        s.fp32_sweeps = total32;
        s.outer_iter = L.hctrl->outer;
        s.rnorm = rn;
        // the FP64 control block's solve fields: the line path's outer step
        // restored them on the device when it ended the loop (the snapshot
        // holds them); the point path restores them here
        if (!line) {
            L.hctrl->rhsnorm = rhs_norm;
            L.hctrl->it = total32;
            L.hctrl->done = (rn / rhs_norm <= L.p.epsmat) ? 1 : 2;
            if ((rc = ctrl_put())) return rc;
        }
    }
    s.rhsnorm = L.hctrl->rhsnorm;
    if (mode == 0 && !s.guess_used && s.r0_cold == 0.0) s.r0_cold = s.rhsnorm;  // zero start: residual = RHS
    // unknowns back into the ghosted field (delta_ER: zero outside the
    // interior), then the phi wrap; after line sweeps with a tail, the tail
    // has queued that ahead of the solve's read
    const bool tail = prec == 1 && L.opts.inner == 1 && L.sor_tail;
    if (!tail && (rc = sor_store_result(mode))) return rc;
    if (st) *st = s;
    if (L.hctrl->done == 2) {
        fail(-1, mode == 1 ? "Max iteration reached: DIF SOR matrix solver" : "MAX COUPLED SOR ITERATION");
        return -1;
    }
    return 0;
}

// scale the last delta_ER in place (warm start: delta_ER <- s * delta_ER)
int sor_store_result(int mode) {
    Geo& g = L.g;
    double* target = mode == 1 ? L.f[TT_T] : L.f[TT_DELTA_ER];
    if (mode == 0)
        k_store_x_box<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, g_cs, x64[0], x64[1], target, L.de1_out, L.run_if);
    else
        k_store_x<<<nblk(2 * g_cs.ncell, 256), 256, 0, L.s>>>(g, g_cs, x64[0], x64[1], target);
    TT_LAUNCH();
    return planetshift_impl(target, mode == 1 ? 10 : 8);
}

int sor_scale_guess(double sc) {
    k_scale<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.f[TT_DELTA_ER], sc, L.g.n3);
    TT_LAUNCH();
    return 0;
}

// ADJUST_SOR_PARAM, ppsor.f:332-354: fifty cold-start solves of the same
// system with SORPARAM from 1.01 in steps of 0.02 (both REAL literals, so
// the float values widened, the sum formed in double as SORPARAM is
// double), keeping the first omega with the fewest iterations (minloc).
int sor_adjust_impl(int fill_from_evals, double* omega_out, int32_t* iters) {
    const int ntry = 50;
    int32_t cnt[ntry];
    double om[ntry];
    const double omega0 = L.p.sorparam;
    const int guess = L.opts.guess, have_guess = L.have_prev_guess;
    L.opts.guess = 0;  // SOLVE_WITH_SOR starts from delta_ER = 0 (ppsor.f:29)
    L.have_prev_guess = 0;
    double w = (double)1.01f;
    int rc = 0;
    for (int t = 0; t < ntry; ++t) {
        L.p.sorparam = w;
        tt_sor_stats s{};
        rc = sor_solve_impl(t == 0 ? fill_from_evals : 0, &s, 0);
        if (rc > 0) break;
        rc = 0;  // an iteration cap only counts ITMAX iterations
        cnt[t] = s.sor_iter;
        om[t] = w;
        w = w + (double)0.02f;
    }
    L.opts.guess = guess;
    L.have_prev_guess = have_guess;
    if (rc > 0) {
        L.p.sorparam = omega0;
        return rc;
    }
    int best = 0;
    for (int t = 1; t < ntry; ++t)
        if (cnt[t] < cnt[best]) best = t;
    L.p.sorparam = om[best];
    if (omega_out) *omega_out = om[best];
    if (iters) memcpy(iters, cnt, sizeof cnt);
    return 0;
}

}  // namespace tt

using namespace tt;

extern "C" {

int tt_solve_sor(int fill_matrix, tt_sor_stats* st) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = sor_solve_impl(fill_matrix, st, 0);
    int rc2 = sync_check();
    return rc2 ? rc2 : rc;
}

int tt_adjust_sor_param(int fill_matrix, double* omega, int32_t* iters) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = sor_adjust_impl(fill_matrix, omega, iters);
    int rc2 = sync_check();
    return rc2 ? rc2 : rc;
}

int tt_set_sorparam(double omega) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!(omega > 0.0 && omega < 2.0)) return fail(1, "SORPARAM must lie in (0, 2)");
    L.p.sorparam = omega;
    return 0;
}

double tt_get_sorparam(void) { return L.ready ? L.p.sorparam : 0.0; }

int tt_get_line_omega(double* omega, int32_t* trials) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (omega) *omega = line_om();
    if (trials) *trials = L.opts.line_omega > 0.0 ? 0 : L.line_om_trials;
    return 0;
}

int tt_sweep_timing(int period) {
    if (L.s) hipStreamSynchronize(L.s);
    ev_drain(L.hctrl, g_bid);
    L.timing = period > 0;
    L.timing_period = period > 0 ? period : 1;
    L.timing_count = 0;
    L.n64 = L.n32 = 0;
    L.ms64 = L.ms32 = 0.0;
    return 0;
}

int tt_sor_speculation(double margin) {
    L.spec_margin = margin;
    L.spec_batches = 0;
    L.spec_reruns = 0;
    return 0;
}

int tt_sor_speculation_get(int64_t* spec_batches, int64_t* reruns) {
    if (spec_batches) *spec_batches = L.spec_batches;
    if (reruns) *reruns = L.spec_reruns;
    return 0;
}

int tt_sweep_timing_get(int64_t* n64, double* ms64, int64_t* n32, double* ms32) {
    if (L.s) hipStreamSynchronize(L.s);
    ev_drain(L.hctrl, g_bid);
    if (n64) *n64 = L.n64;
    if (ms64) *ms64 = L.ms64;
    if (n32) *n32 = L.n32;
    if (ms32) *ms32 = L.ms32;
    return 0;
}

}  // extern "C"
