// tt_cnn.hip -- learned initial guess for the FLD solve (new: no counterpart
// in the reference).  A 3-D convolutional network maps five per-cell
// features (log E, log T, log rho, log kappa_R, asinh S) to the first Newton
// correction delta_ER of COUPLED_ENERGY_EQS as a rate, delta_ER / (E dt)
// (the rate form carries over between grids whose CFL time steps differ).
//
//   5 -> 16 -> 16 -> 16 -> 16 -> 1 channels, 3x3x3 "same" convolutions (zero
//   padding in r and theta, periodic in phi), GELU (tanh form) after each
//   hidden layer.  Weights W[tap][ci][co], tap = (dk+1)*9 + (dj+1)*3 + (di+1).
//
// Arithmetic: bf16 weights and activations, fp32 accumulation.  The hidden
// layers are implicit GEMMs on the matrix cores (v_mfma_f32_16x16x32_bf16):
// a wavefront computes 16 consecutive cells of one r row (M) x 16 output
// channels (N); K runs over taps x input channels, 32 per instruction (two
// taps of 16 channels, or four taps of the 8-channel padded input).  Operand
// fragments: lane l holds A[l%16][8(l/16)..+7] (8 channels of one tap of one
// cell: one 16-byte load from the channels-last activations) and
// B[8(l/16)..+7][l%16] (weights, pre-packed in that order), and receives
// D[4(l/16)+i][l%16], i = 0..3.  The input gradient of a hidden layer is the
// same GEMM over the output gradient with mirrored taps and transposed
// weights; the weight gradient is a GEMM with K over cells.  The last layer
// (16 -> 1) runs on the vector units.
//
// Training data come from the library's own FP64 solves: with collection on,
// the first Newton iteration of every COUPLED_ENERGY_EQS call appends its
// features and its delta_ER / (E dt) to a device-resident set; tt_cnn_train runs
// Adam over that set with the kernels below.  At inference the network's
// output, times E dt, is staged as the SOR initial guess; the SOR solve
// scales it to its least-residual multiple and keeps it only if that
// residual is not above the cold start's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "tt_common.h"

namespace tt {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef uint16_t bf16_t;

constexpr int NL = 5;
constexpr int NF = 5;   // features
constexpr int C0P = 8;  // padded input channels
__host__ __device__ constexpr int ch(int l) { return l == 0 ? 5 : (l == NL ? 1 : 16); }
__host__ __device__ constexpr int woff(int l) {
    int o = 0;
    for (int m = 0; m < l; ++m) o += 27 * ch(m) * ch(m + 1);
    return o;
}
__host__ __device__ constexpr int boff(int l) {
    int o = 0;
    for (int m = 0; m < l; ++m) o += ch(m + 1);
    return o;
}
constexpr int NW = woff(NL);  // 23,328
constexpr int NB = boff(NL);  // 65

__host__ __device__ inline bf16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

__device__ inline float gelu(float u) {
    const float c = 0.7978845608028654f;
    return 0.5f * u * (1.0f + tanhf(c * (u + 0.044715f * u * u * u)));
}
__device__ inline float gelu_d(float u) {
    const float c = 0.7978845608028654f;
    float t = tanhf(c * (u + 0.044715f * u * u * u));
    return 0.5f * (1.0f + t) + 0.5f * u * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * u * u);
}

struct Grid3 {
    int nx, ny, nz;
    __device__ inline int cell(int i, int j, int k) const { return i + nx * (j + ny * k); }
    // neighbour of cell (i,j,k) through tap t: -1 outside (r, theta), phi wraps
    __device__ inline int nbr(int i, int j, int k, int t) const {
        int dk = t / 9 - 1, dj = (t / 3) % 3 - 1, di = t % 3 - 1;
        int ii = i + di, kk = k + dk;
        if (ii < 0 || ii >= nx || kk < 0 || kk >= nz) return -1;
        int jj = j + dj;
        jj = jj < 0 ? jj + ny : (jj >= ny ? jj - ny : jj);
        return cell(ii, jj, kk);
    }
};

// ------------------------------------------------------------ model state
struct Model {
    bool have_w = false, have_norm = false;
    float *w = nullptr, *b = nullptr;     // fp32 master parameters
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
    float *gw = nullptr, *gb = nullptr;   // gradient accumulators
    long long adam_t = 0;
    bf16_t* frag = nullptr;               // packed bf16 fragments, see k_pack_frags
    float* w4 = nullptr;                  // last layer weights rounded to bf16, [27][16]
    float norm[12] = {0};                 // mean[6], std[6]: 5 features + target
    float* dnorm = nullptr;
    // workspace for one grid
    long long wcells = 0;
    bf16_t* act[NL] = {};                 // act[0]: [cells][8]; act[1..4]: [cells][16]
    float* pre[NL - 1] = {};              // pre-activations of layers 0..3
    bf16_t *gu = nullptr, *gu2 = nullptr; // dU ping-pong [cells][16]
    float *y = nullptr, *tg = nullptr, *gy = nullptr;
    float* part = nullptr;                // reduction partials
    long long npart = 0;
    // dataset
    int ds_cap = 0, ds_n = 0;
    long long ds_cells = 0;
    int ds_nx = 0, ds_ny = 0, ds_nz = 0;
    float *ds_feat = nullptr, *ds_tgt = nullptr;
    bool collect = false;
    int64_t bytes = 0;
};
static Model M;

// fragment offsets (in bf16 elements): layer 0 fwd 7 q, hidden fwd 3 x 14 q,
// hidden input-gradient 3 x 14 q; 64 lanes x 8 values per q
constexpr int FQ0 = 7, FQH = 14;
constexpr long long FRAG_L0 = 0;
__host__ __device__ constexpr long long frag_fwd(int l) { return (FQ0 + (long long)(l - 1) * FQH) * 512; }
__host__ __device__ constexpr long long frag_bwd(int l) { return (FQ0 + 3LL * FQH + (long long)(l - 1) * FQH) * 512; }
constexpr long long FRAG_N = (FQ0 + 6LL * FQH) * 512;

template <typename T>
static hipError_t cnn_alloc(T** p, size_t bytes) {
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e == hipSuccess) M.bytes += (int64_t)bytes;
    return e;
}
#define FREE0(p) do { if (p) hipFree(p); p = nullptr; } while (0)

// pack the fp32 master weights into bf16 MFMA operand order, and the last
// layer's weights rounded to bf16
__global__ void k_pack_frags(const float* __restrict__ w, bf16_t* frag, float* w4) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < FRAG_N) {
        int e = (int)(t & 7), lane = (int)((t >> 3) & 63);
        long long q = t >> 9;  // global q index
        int kg = lane >> 4, n = lane & 15;
        float v = 0.f;
        if (q < FQ0) {  // layer 0: four taps of 8 (5 real) channels per instruction
            int tap = 4 * (int)q + kg, ci = e;
            if (tap < 27 && ci < 5) v = w[woff(0) + (tap * 5 + ci) * 16 + n];
        } else if (q < FQ0 + 3 * FQH) {  // hidden forward: two taps of 16 channels
            int l = 1 + (int)((q - FQ0) / FQH), qq = (int)((q - FQ0) % FQH);
            int tap = 2 * qq + (kg >> 1), ci = 8 * (kg & 1) + e;
            if (tap < 27) v = w[woff(l) + (tap * 16 + ci) * 16 + n];
        } else {  // hidden input gradient: mirrored taps, transposed W
            int l = 1 + (int)((q - FQ0 - 3 * FQH) / FQH), qq = (int)((q - FQ0 - 3 * FQH) % FQH);
            int tp = 2 * qq + (kg >> 1), co = 8 * (kg & 1) + e;
            if (tp < 27) v = w[woff(l) + ((26 - tp) * 16 + n) * 16 + co];
        }
        frag[t] = f2bf(v);
    }
    if (t < 27 * 16) w4[t] = bf2f(f2bf(w[woff(4) + t]));
}

__device__ inline bf16x8 ld8(const bf16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }

// Hidden-width convolution on the matrix cores.  CIN 8 (layer 0 input) or 16.
// A block computes a tile of 16 (r) x 4 (phi) x 4 (theta) cells: its input
// with a one-cell halo (18 x 6 x 6 cells, zero outside r and theta, phi
// wrapped) is staged in LDS once, and every A fragment (one tap, 8 channels
// of one cell per lane) is a 16-byte LDS read instead of a gather from
// global memory.  Each wavefront owns 4 rows of 16 cells (4 accumulation
// chains); each weight fragment is loaded once for the 4.  K runs over the
// taps in the same order as the weight fragments.
// mode 0: out = bf16(gelu(acc + bias)), pre = acc + bias (when pre != null)
// mode 1: out = bf16(acc * gelu'(gpre))  (input gradient of the layer, times
//         the previous layer's gelu derivative)
constexpr int TPW = 4;
constexpr int CT_I = 16, CT_J = 4, CT_K = 4;
constexpr int CH_I = CT_I + 2, CH_J = CT_J + 2, CH_K = CT_K + 2, CH_N = CH_I * CH_J * CH_K;
__host__ __device__ inline unsigned conv_blocks(const Grid3& g) {
    return (unsigned)((g.nx / CT_I) * ((g.ny + CT_J - 1) / CT_J) * ((g.nz + CT_K - 1) / CT_K));
}

template <int CIN>
__global__ __launch_bounds__(256) void k_conv_mfma(const bf16_t* __restrict__ x, const bf16_t* __restrict__ frag,
                                                   const float* __restrict__ bias, Grid3 g, int mode,
                                                   bf16_t* __restrict__ out, float* __restrict__ pre,
                                                   const float* __restrict__ gpre) {
    constexpr int V = CIN / 8;  // 16-byte words per cell
    constexpr int NQ = CIN == 8 ? FQ0 : FQH;
    __shared__ uint4 tile[CH_N * V];
    const int nti = g.nx / CT_I, ntj = (g.ny + CT_J - 1) / CT_J;
    int b = blockIdx.x;
    const int bi = b % nti;
    b /= nti;
    const int bj = b % ntj, bk = b / ntj;
    const int i0 = bi * CT_I, j0 = bj * CT_J, k0 = bk * CT_K;
    const uint4 zero = {0u, 0u, 0u, 0u};
    for (int e = threadIdx.x; e < CH_N * V; e += 256) {
        const int cell = e / V, v = e - cell * V;
        const int ii = cell % CH_I, rr = cell / CH_I, jj = rr % CH_J, kk = rr / CH_J;
        const int i = i0 - 1 + ii, k = k0 - 1 + kk;
        int j = (j0 - 1 + jj) % g.ny;
        if (j < 0) j += g.ny;
        uint4 val = zero;
        if (i >= 0 && i < g.nx && k >= 0 && k < g.nz) val = *(const uint4*)(x + (long long)g.cell(i, j, k) * CIN + 8 * v);
        tile[e] = val;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    int jr[TPW], kr[TPW];
    bool tv[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int row = w * TPW + u;  // 16 rows: (phi, theta) = (row % 4, row / 4)
        jr[u] = row % CT_J;
        kr[u] = row / CT_J;
        tv[u] = j0 + jr[u] < g.ny && k0 + kr[u] < g.nz;
    }
    f32x4 acc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) acc[u] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < NQ; ++q) {
        const int tap = CIN == 8 ? 4 * q + kg : 2 * q + (kg >> 1);
        const int c0 = CIN == 8 ? 0 : (kg & 1);  // 16-byte word within the cell
        const bf16x8 bw = ld8(frag + ((long long)q * 64 + lane) * 8);
        const int dk = tap / 9, dj = (tap / 3) % 3, di = tap % 3;  // offsets + 1
        uint4 av[TPW];
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const int lc = ((kr[u] + dk) * CH_J + (jr[u] + dj)) * CH_I + (m + di);
            av[u] = tap < 27 ? tile[lc * V + c0] : zero;
        }
#pragma unroll
        for (int u = 0; u < TPW; ++u)
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[u]), bw, acc[u], 0, 0, 0);
    }
    const int n = lane & 15;
    const float bn = mode == 0 ? bias[n] : 0.f;
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        if (!tv[u]) continue;
        const int base = g.cell(i0, j0 + jr[u], k0 + kr[u]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long long c = (long long)(base + 4 * kg + r) * 16 + n;
            if (mode == 0) {
                float v = acc[u][r] + bn;
                if (pre) pre[c] = v;
                out[c] = f2bf(gelu(v));
            } else {
                out[c] = f2bf(acc[u][r] * gelu_d(gpre[c]));
            }
        }
    }
}

// last layer forward: y = b4 + sum over taps of act4[nbr] . w4[tap]
__global__ void k_conv_out(const bf16_t* __restrict__ x, const float* __restrict__ w4, const float* __restrict__ b,
                           Grid3 g, long long n, float* __restrict__ y) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int c = (int)t, i = c % g.nx, r = c / g.nx, j = r % g.ny, k = r / g.ny;
    float acc = b[boff(4)];
    for (int tap = 0; tap < 27; ++tap) {
        int nb = g.nbr(i, j, k, tap);
        if (nb < 0) continue;
        const uint4* p = (const uint4*)(x + (long long)nb * 16);
        uint4 u0 = p[0], u1 = p[1];
        uint32_t wds[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            acc += bf2f((bf16_t)(wds[q] & 0xFFFFu)) * w4[tap * 16 + 2 * q];
            acc += bf2f((bf16_t)(wds[q] >> 16)) * w4[tap * 16 + 2 * q + 1];
        }
    }
    y[t] = acc;
}

// loss = mean (y - tg)^2: gy = 2 (y - tg) / n, per-block sums of squares
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ tg, long long n,
                                              float* __restrict__ gy, double* part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
        float r = y[t] - tg[t];
        gy[t] = 2.0f * r / (float)n;
        acc += (double)r * r;
    }
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// last layer weight gradient: part[(chunk*27 + tap)*16 + ci] = sum over the
// chunk's cells c of act4[nbr(c,tap)][ci] * gy[c]; grid (chunks, 27)
__global__ __launch_bounds__(256) void k_wgrad_out(const bf16_t* __restrict__ x, const float* __restrict__ gy, Grid3 g,
                                                   long long n, int chunk, float* __restrict__ part) {
    __shared__ float sh[16][257];
    const int tap = blockIdx.y;
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    long long c0 = (long long)blockIdx.x * chunk, c1 = c0 + chunk < n ? c0 + chunk : n;
    for (long long t = c0 + threadIdx.x; t < c1; t += 256) {
        int c = (int)t, i = c % g.nx, r = c / g.nx, j = r % g.ny, k = r / g.ny;
        int nb = g.nbr(i, j, k, tap);
        if (nb < 0) continue;
        float gv = gy[t];
        const bf16_t* p = x + (long long)nb * 16;
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] += bf2f(p[q]) * gv;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    if (threadIdx.x < 16) {
        float s = 0.f;
        for (int u = 0; u < 256; ++u) s += sh[threadIdx.x][u];
        part[((long long)blockIdx.x * 27 + tap) * 16 + threadIdx.x] = s;
    }
}

// last layer input gradient times gelu'(pre3): gu[s][ci] =
// bf16(gelu'(pre3[s][ci]) * sum over taps of w4[tap][ci] * gy[s - off(tap)])
__global__ void k_bwd_out(const float* __restrict__ gy, const float* __restrict__ w4, const float* __restrict__ pre3,
                          Grid3 g, long long n, bf16_t* __restrict__ gu) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int c = (int)t, i = c % g.nx, r = c / g.nx, j = r % g.ny, k = r / g.ny;
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (int tap = 0; tap < 27; ++tap) {
        int nb = g.nbr(i, j, k, 26 - tap);  // the cell whose tap lands on this one
        if (nb < 0) continue;
        float gv = gy[nb];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] += w4[tap * 16 + q] * gv;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) gu[t * 16 + q] = f2bf(acc[q] * gelu_d(pre3[t * 16 + q]));
}

// hidden weight gradient on the matrix cores: dW[tap][ci][co] = sum over
// cells c of x[nbr(c,tap)][ci] * gu[c][co].  A = x^T (M = ci), B = gu
// (N = co), K = 32 cells per instruction.  grid (chunks, 27); the four
// wavefronts of a block split the chunk and are summed through LDS.
template <int CIN>
__global__ __launch_bounds__(256) void k_wgrad_mfma(const bf16_t* __restrict__ x, const bf16_t* __restrict__ gu,
                                                    Grid3 g, long long n, int chunk, float* __restrict__ part) {
    __shared__ float sh[4][256];
    const int tap = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    long long c0 = (long long)blockIdx.x * chunk, c1 = c0 + chunk < n ? c0 + chunk : n;
    for (long long base = c0 + 32 * w; base < c1; base += 128) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            long long t = base + 8 * kg + e;
            float av = 0.f, bv = 0.f;
            if (t < c1) {
                int c = (int)t, i = c % g.nx, r = c / g.nx, j = r % g.ny, k = r / g.ny;
                int nb = g.nbr(i, j, k, tap);
                if (nb >= 0 && m < CIN) av = bf2f(x[(long long)nb * CIN + m]);
                bv = bf2f(gu[t * 16 + m]);
            }
            a[e] = (__bf16)av;
            b[e] = (__bf16)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) sh[w][(4 * kg + r) * 16 + m] = acc[r];
    __syncthreads();
    const int q = threadIdx.x;  // ci*16 + co
    part[((long long)blockIdx.x * 27 + tap) * 256 + q] = sh[0][q] + sh[1][q] + sh[2][q] + sh[3][q];
}

// per-block column sums of a [cells][16] bf16 array (bias gradients)
__global__ __launch_bounds__(256) void k_colsum16(const bf16_t* __restrict__ gu, long long n, float* __restrict__ part) {
    __shared__ float sh[16][17];
    const int q = threadIdx.x & 15, r = threadIdx.x >> 4;
    float acc = 0.f;
    for (long long t = (long long)blockIdx.x * 16 + r; t < n; t += (long long)gridDim.x * 16) acc += bf2f(gu[t * 16 + q]);
    sh[r][q] = acc;
    __syncthreads();
    if (threadIdx.x < 16) {
        float s = 0.f;
        for (int u = 0; u < 16; ++u) s += sh[u][threadIdx.x];
        part[(long long)blockIdx.x * 16 + threadIdx.x] = s;
    }
}

// dst[idx] += sum over chunks of part[chunk][idx]; idx over 27 x cis x cos
// mapped from the partial layout 27 x 16 x cos (layer 0 has cis = 5)
__global__ void k_fold_wgrad(const float* __restrict__ part, int nchunks, int cis, int cos_, int src_ci,
                             float* __restrict__ dst) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 27 * cis * cos_;
    if (t >= total) return;
    int co = t % cos_, r = t / cos_, ci = r % cis, tap = r / cis;
    int src = (tap * src_ci + ci) * cos_ + co;
    float s = 0.f;
    for (int q = 0; q < nchunks; ++q) s += part[(long long)q * 27 * src_ci * cos_ + src];
    dst[t] += s;
}

__global__ void k_fold_vec(const float* __restrict__ part, int nchunks, int w, float* __restrict__ dst) {
    int t = threadIdx.x;
    if (t >= w) return;
    float s = 0.f;
    for (int q = 0; q < nchunks; ++q) s += part[(long long)q * w + t];
    dst[t] += s;
}

__global__ void k_sum_gy(const float* __restrict__ gy, long long n, float* part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) acc += gy[t];
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = (float)acc;
}

__global__ void k_adam(float* w, const float* __restrict__ g, float* m, float* v, int n, float lr, float scale,
                       float b1, float b2, float eps, float c1, float c2) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float gg = g[t] * scale;
    float mm = b1 * m[t] + (1.f - b1) * gg;
    float vv = b2 * v[t] + (1.f - b2) * gg * gg;
    m[t] = mm;
    v[t] = vv;
    w[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---------------------------------------------------------------- features
// raw features of interior cell (i,j,k) of the library grid, in the cell
// order (i fastest) of the network: log E, log T, log rho, log kappa_R,
// asinh S; target delta_ER / (E dt)
__device__ inline void raw_features(const Geo& g, int c, const double* ER, const double* T, const double* RH,
                                    const double* KR, const double* SIN, float f[NF], int& i, int& j, int& k) {
    int nx = g.nx, ny = g.ny;
    i = c % nx + 1;
    int r = c / nx;
    j = r % ny + 1;
    k = r / ny + 1;
    long long n = g.idx(i, j, k);
    f[0] = (float)log(ER[n]);
    f[1] = (float)log(T[n]);
    f[2] = (float)log(RH[n]);
    f[3] = (float)log(KR[n]);
    f[4] = (float)asinh(SIN[n]);
}

__global__ void k_collect(Geo g, const double* __restrict__ ER, const double* __restrict__ T,
                          const double* __restrict__ RH, const double* __restrict__ KR,
                          const double* __restrict__ SIN, const double* __restrict__ DER, double delt,
                          long long n, float* __restrict__ feat, float* __restrict__ tgt) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float f[NF];
    int i, j, k;
    raw_features(g, (int)t, ER, T, RH, KR, SIN, f, i, j, k);
    for (int q = 0; q < NF; ++q) feat[t * NF + q] = f[q];
    long long c = g.idx(i, j, k);
    tgt[t] = (float)(DER[c] / (ER[c] * delt));
}

// normalised bf16 network input [cells][8] from raw features
__global__ void k_norm_in(const float* __restrict__ feat, const float* __restrict__ nrm, long long n,
                          bf16_t* __restrict__ x) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (int q = 0; q < C0P; ++q)
        x[t * C0P + q] = q < NF ? f2bf((feat[t * NF + q] - nrm[q]) / nrm[6 + q]) : (bf16_t)0;
}

__global__ void k_norm_tgt(const float* __restrict__ tgt, const float* __restrict__ nrm, long long n, float* out) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (tgt[t] - nrm[5]) / nrm[11];
}

// inference input straight from the fields
__global__ void k_infer_in(Geo g, const double* __restrict__ ER, const double* __restrict__ T,
                           const double* __restrict__ RH, const double* __restrict__ KR,
                           const double* __restrict__ SIN, const float* __restrict__ nrm, long long n,
                           bf16_t* __restrict__ x) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float f[NF];
    int i, j, k;
    raw_features(g, (int)t, ER, T, RH, KR, SIN, f, i, j, k);
    for (int q = 0; q < C0P; ++q) x[t * C0P + q] = q < NF ? f2bf((f[q] - nrm[q]) / nrm[6 + q]) : (bf16_t)0;
}

// guess delta_ER = E dt (mean + std y) on the solved cells, 0 on the identity
// rows of the SOR matrix (i >= upperbnd, i = nx, k = 1, k = nz) and ghosts
__global__ void k_guess_out(Geo g, const float* __restrict__ y, const float* __restrict__ nrm,
                            const double* __restrict__ ER, const int32_t* __restrict__ ub, double delt,
                            double* DER) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    double v = 0.0;
    if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) {
        int u = ub[g.idx2(j, k)];
        bool ident = (i >= u) || (i == g.nx) || (k == 1) || (k == g.nz);
        if (!ident) {
            int c = (i - 1) + g.nx * ((j - 1) + g.ny * (k - 1));
            v = ER[t] * delt * ((double)nrm[5] + (double)nrm[11] * (double)y[c]);
        }
    }
    DER[t] = v;
}

// per-block sums and sums of squares of the 5 features and the target
__global__ __launch_bounds__(256) void k_moments(const float* __restrict__ feat, const float* __restrict__ tgt,
                                                 long long n, double* part) {
    __shared__ double sh[4];
    for (int q = 0; q < 2 * (NF + 1); ++q) {
        int ch_ = q % (NF + 1);
        bool sq = q >= NF + 1;
        double acc = 0.0;
        for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
            double v = ch_ < NF ? feat[t * NF + ch_] : tgt[t];
            acc += sq ? v * v : v;
        }
        acc = block_sum256(acc, sh);
        if (threadIdx.x == 0) part[(long long)blockIdx.x * 2 * (NF + 1) + q] = acc;
    }
}

// ------------------------------------------------------------ host side
static int ensure_model() {
    if (M.w) return 0;
    TT_HIP(cnn_alloc(&M.w, NW * 4));
    TT_HIP(cnn_alloc(&M.b, NB * 4));
    TT_HIP(cnn_alloc(&M.mw, NW * 4));
    TT_HIP(cnn_alloc(&M.vw, NW * 4));
    TT_HIP(cnn_alloc(&M.mb, NB * 4));
    TT_HIP(cnn_alloc(&M.vb, NB * 4));
    TT_HIP(cnn_alloc(&M.gw, NW * 4));
    TT_HIP(cnn_alloc(&M.gb, NB * 4));
    TT_HIP(cnn_alloc(&M.frag, FRAG_N * 2));
    TT_HIP(cnn_alloc(&M.w4, 27 * 16 * 4));
    TT_HIP(cnn_alloc(&M.dnorm, 12 * 4));
    return 0;
}

static int ensure_work(long long cells) {
    if (M.wcells >= cells) return 0;
    for (int l = 0; l < NL; ++l) FREE0(M.act[l]);
    for (int l = 0; l < NL - 1; ++l) FREE0(M.pre[l]);
    FREE0(M.gu); FREE0(M.gu2); FREE0(M.y); FREE0(M.tg); FREE0(M.gy); FREE0(M.part);
    TT_HIP(cnn_alloc(&M.act[0], cells * C0P * 2));
    for (int l = 1; l < NL; ++l) TT_HIP(cnn_alloc(&M.act[l], cells * 16 * 2));
    for (int l = 0; l < NL - 1; ++l) TT_HIP(cnn_alloc(&M.pre[l], cells * 16 * 4));
    TT_HIP(cnn_alloc(&M.gu, cells * 16 * 2));
    TT_HIP(cnn_alloc(&M.gu2, cells * 16 * 2));
    TT_HIP(cnn_alloc(&M.y, cells * 4));
    TT_HIP(cnn_alloc(&M.tg, cells * 4));
    TT_HIP(cnn_alloc(&M.gy, cells * 4));
    // weight-gradient partials: chunks of 2048 cells x 27 taps x 256
    M.npart = ((cells + 2047) / 2048) * 27 * 256 + 4096;
    TT_HIP(cnn_alloc(&M.part, M.npart * 4));
    M.wcells = cells;
    return 0;
}

static int repack() {
    k_pack_frags<<<nblk(FRAG_N, 256), 256, 0, L.s>>>(M.w, M.frag, M.w4);
    TT_LAUNCH();
    return 0;
}

// forward over a grid whose bf16 input is already in act[0]; train: keep
// the pre-activations
static int forward(Grid3 g, bool train) {
    long long n = (long long)g.nx * g.ny * g.nz;
    const unsigned nb = conv_blocks(g);
    k_conv_mfma<8><<<nb, 256, 0, L.s>>>(M.act[0], M.frag + FRAG_L0, M.b + boff(0), g, 0, M.act[1],
                                        train ? M.pre[0] : nullptr, nullptr);
    TT_LAUNCH();
    for (int l = 1; l < NL - 1; ++l) {
        k_conv_mfma<16><<<nb, 256, 0, L.s>>>(M.act[l], M.frag + frag_fwd(l), M.b + boff(l), g, 0, M.act[l + 1],
                                             train ? M.pre[l] : nullptr, nullptr);
        TT_LAUNCH();
    }
    k_conv_out<<<nblk(n, 256), 256, 0, L.s>>>(M.act[NL - 1], M.w4, M.b, g, n, M.y);
    TT_LAUNCH();
    return 0;
}

// loss and gradients (accumulated into gw, gb) for the target in tg; the
// loss of this sample is added to *loss_dev (per-block partials folded)
static int backward(Grid3 g, double* loss_part, int* loss_nb) {
    long long n = (long long)g.nx * g.ny * g.nz;
    int lnb = (int)std::min<long long>(1024, (n + 255) / 256);
    k_loss<<<lnb, 256, 0, L.s>>>(M.y, M.tg, n, M.gy, loss_part);
    TT_LAUNCH();
    *loss_nb = lnb;
    const int chunk = 2048;
    int nch = (int)((n + chunk - 1) / chunk);
    // last layer
    k_wgrad_out<<<dim3(nch, 27), 256, 0, L.s>>>(M.act[4], M.gy, g, n, chunk, M.part);
    TT_LAUNCH();
    k_fold_wgrad<<<nblk(27 * 16, 256), 256, 0, L.s>>>(M.part, nch, 16, 1, 16, M.gw + woff(4));
    TT_LAUNCH();
    int snb = (int)std::min<long long>(1024, (n + 255) / 256);
    k_sum_gy<<<snb, 256, 0, L.s>>>(M.gy, n, M.part);
    TT_LAUNCH();
    k_fold_vec<<<1, 64, 0, L.s>>>(M.part, snb, 1, M.gb + boff(4));
    TT_LAUNCH();
    k_bwd_out<<<nblk(n, 256), 256, 0, L.s>>>(M.gy, M.w4, M.pre[3], g, n, M.gu);
    TT_LAUNCH();
    const unsigned tnb = conv_blocks(g);
    bf16_t* gu = M.gu;
    bf16_t* gn = M.gu2;
    for (int l = NL - 2; l >= 0; --l) {  // hidden layers 3..0: gu = dL/d(pre-activation of layer l)
        if (l == 0) k_wgrad_mfma<8><<<dim3(nch, 27), 256, 0, L.s>>>(M.act[0], gu, g, n, chunk, M.part);
        else k_wgrad_mfma<16><<<dim3(nch, 27), 256, 0, L.s>>>(M.act[l], gu, g, n, chunk, M.part);
        TT_LAUNCH();
        int cis = l == 0 ? 5 : 16;
        k_fold_wgrad<<<nblk(27 * cis * 16, 256), 256, 0, L.s>>>(M.part, nch, cis, 16, 16, M.gw + woff(l));
        TT_LAUNCH();
        int cnb = (int)std::min<long long>(1024, (n + 15) / 16);
        k_colsum16<<<cnb, 256, 0, L.s>>>(gu, n, M.part);
        TT_LAUNCH();
        k_fold_vec<<<1, 64, 0, L.s>>>(M.part, cnb, 16, M.gb + boff(l));
        TT_LAUNCH();
        if (l > 0) {
            k_conv_mfma<16><<<tnb, 256, 0, L.s>>>(gu, M.frag + frag_bwd(l), nullptr, g, 1, gn, nullptr,
                                                  M.pre[l - 1]);
            TT_LAUNCH();
            std::swap(gu, gn);
        }
    }
    return 0;
}

static int check_grid(int nx, int ny, int nz) {
    if (nx % 16) return fail(1, "learned guess: n_r must be a multiple of 16");
    if (nx < 1 || ny < 1 || nz < 1) return fail(1, "learned guess: empty grid");
    return 0;
}

bool cnn_ready() { return M.have_w && M.have_norm; }
bool cnn_collecting() { return M.collect != 0; }

void cnn_free() {}  // the model and the training set outlive tt_init/tt_finalize

static void cnn_release() {
    FREE0(M.w); FREE0(M.b); FREE0(M.mw); FREE0(M.vw); FREE0(M.mb); FREE0(M.vb); FREE0(M.gw); FREE0(M.gb);
    FREE0(M.frag); FREE0(M.w4); FREE0(M.dnorm);
    for (int l = 0; l < NL; ++l) FREE0(M.act[l]);
    for (int l = 0; l < NL - 1; ++l) FREE0(M.pre[l]);
    FREE0(M.gu); FREE0(M.gu2); FREE0(M.y); FREE0(M.tg); FREE0(M.gy); FREE0(M.part);
    FREE0(M.ds_feat); FREE0(M.ds_tgt);
    M = Model();
}

int64_t cnn_bytes() { return M.bytes; }

// learned guess for the current Newton system, staged in DELTA_ER
int cnn_predict_guess(double delt) {
    Geo& g = L.g;
    int rc = check_grid(g.nx, g.ny, g.nz);
    if (rc) return rc;
    long long n = (long long)g.nx * g.ny * g.nz;
    if ((rc = ensure_work(n))) return rc;
    k_infer_in<<<nblk(n, 256), 256, 0, L.s>>>(g, L.f[TT_ER], L.f[TT_T], L.f[TT_RH], L.f[TT_XKAPR],
                                             L.f[TT_STELLARINPUT], M.dnorm, n, M.act[0]);
    TT_LAUNCH();
    if ((rc = forward(Grid3{g.nx, g.ny, g.nz}, false))) return rc;
    k_guess_out<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, M.y, M.dnorm, L.f[TT_ER], L.ub, delt, L.f[TT_DELTA_ER]);
    TT_LAUNCH();
    return 0;
}

// append the features of the current Newton system and its solution
// (delta_ER after the first SOR solve) to the training set
int cnn_collect_sample(double delt) {
    if (!M.collect) return 0;
    Geo& g = L.g;
    long long n = (long long)g.nx * g.ny * g.nz;
    if (M.ds_n >= M.ds_cap) return 0;  // full: further samples are dropped
    if (M.ds_n > 0 && (g.nx != M.ds_nx || g.ny != M.ds_ny || g.nz != M.ds_nz))
        return fail(1, "learned guess: every training sample must have the grid of the first");
    if (n != M.ds_cells) return fail(1, "learned guess: training set reserved for another grid size");
    M.ds_nx = g.nx; M.ds_ny = g.ny; M.ds_nz = g.nz;
    k_collect<<<nblk(n, 256), 256, 0, L.s>>>(g, L.f[TT_ER], L.f[TT_T], L.f[TT_RH], L.f[TT_XKAPR],
                                             L.f[TT_STELLARINPUT], L.f[TT_DELTA_ER], delt, n,
                                             M.ds_feat + (long long)M.ds_n * n * NF, M.ds_tgt + (long long)M.ds_n * n);
    TT_LAUNCH();
    M.ds_n++;
    return 0;
}

}  // namespace tt

using namespace tt;

extern "C" {

int64_t tt_cnn_num_weights(void) { return NW; }
int64_t tt_cnn_num_biases(void) { return NB; }

int tt_cnn_set_weights(const float* w, int64_t nw, const float* b, int64_t nb) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (nw != NW || nb != NB || !w || !b) return fail(1, "learned guess: wrong parameter count");
    int rc = ensure_model();
    if (rc) return rc;
    TT_HIP(hipMemcpyAsync(M.w, w, NW * 4, hipMemcpyHostToDevice, L.s));
    TT_HIP(hipMemcpyAsync(M.b, b, NB * 4, hipMemcpyHostToDevice, L.s));
    TT_HIP(hipMemsetAsync(M.mw, 0, NW * 4, L.s));
    TT_HIP(hipMemsetAsync(M.vw, 0, NW * 4, L.s));
    TT_HIP(hipMemsetAsync(M.mb, 0, NB * 4, L.s));
    TT_HIP(hipMemsetAsync(M.vb, 0, NB * 4, L.s));
    M.adam_t = 0;
    L.learned_rejects = 0;  // a new model gets its chances again
    L.learned_idle = 0;
    if ((rc = repack())) return rc;
    TT_HIP(hipStreamSynchronize(L.s));
    M.have_w = true;
    return 0;
}

int tt_cnn_get_weights(float* w, int64_t nw, float* b, int64_t nb) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!M.have_w) return fail(1, "learned guess: no weights set");
    if (nw != NW || nb != NB || !w || !b) return fail(1, "learned guess: wrong parameter count");
    TT_HIP(hipMemcpyAsync(w, M.w, NW * 4, hipMemcpyDeviceToHost, L.s));
    TT_HIP(hipMemcpyAsync(b, M.b, NB * 4, hipMemcpyDeviceToHost, L.s));
    TT_HIP(hipStreamSynchronize(L.s));
    return 0;
}

int tt_cnn_set_norm(const float* norm12) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = ensure_model();
    if (rc) return rc;
    for (int q = 0; q < 6; ++q)
        if (!(norm12[6 + q] > 0.f)) return fail(1, "learned guess: normalisation scales must be positive");
    memcpy(M.norm, norm12, sizeof(M.norm));
    TT_HIP(hipMemcpyAsync(M.dnorm, M.norm, sizeof(M.norm), hipMemcpyHostToDevice, L.s));
    TT_HIP(hipStreamSynchronize(L.s));
    M.have_norm = true;
    return 0;
}

int tt_cnn_get_norm(float* norm12) {
    if (!M.have_norm) return fail(1, "learned guess: no normalisation set");
    memcpy(norm12, M.norm, sizeof(M.norm));
    return 0;
}

int tt_cnn_eval_host(int nx, int ny, int nz, const float* x0, const float* target, float* y, float* loss, float* gw,
                     float* gb) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!M.have_w) return fail(1, "learned guess: no weights set");
    int rc = check_grid(nx, ny, nz);
    if (rc) return rc;
    long long n = (long long)nx * ny * nz;
    if ((rc = ensure_work(n))) return rc;
    // x0 [cells][5] normalised features -> bf16 [cells][8] (host staging)
    std::vector<bf16_t> xb((size_t)n * C0P, 0);
    for (long long t = 0; t < n; ++t)
        for (int q = 0; q < NF; ++q) xb[t * C0P + q] = f2bf(x0[t * NF + q]);
    TT_HIP(hipMemcpyAsync(M.act[0], xb.data(), n * C0P * 2, hipMemcpyHostToDevice, L.s));
    Grid3 g{nx, ny, nz};
    const bool grad = target != nullptr;
    if ((rc = forward(g, grad))) return rc;
    if (grad) {
        TT_HIP(hipMemcpyAsync(M.tg, target, n * 4, hipMemcpyHostToDevice, L.s));
        TT_HIP(hipMemsetAsync(M.gw, 0, NW * 4, L.s));
        TT_HIP(hipMemsetAsync(M.gb, 0, NB * 4, L.s));
        double* lp = (double*)(M.part + M.npart - 2048);
        int lnb = 0;
        if ((rc = backward(g, lp, &lnb))) return rc;
        std::vector<double> hl(lnb);
        TT_HIP(hipMemcpyAsync(hl.data(), lp, lnb * 8, hipMemcpyDeviceToHost, L.s));
        if (gw) TT_HIP(hipMemcpyAsync(gw, M.gw, NW * 4, hipMemcpyDeviceToHost, L.s));
        if (gb) TT_HIP(hipMemcpyAsync(gb, M.gb, NB * 4, hipMemcpyDeviceToHost, L.s));
        TT_HIP(hipStreamSynchronize(L.s));
        double s = 0.0;
        for (double v : hl) s += v;
        if (loss) *loss = (float)(s / (double)n);
    }
    if (y) TT_HIP(hipMemcpyAsync(y, M.y, n * 4, hipMemcpyDeviceToHost, L.s));
    TT_HIP(hipStreamSynchronize(L.s));
    return 0;
}

int tt_cnn_reserve(int max_samples, int nx, int ny, int nz) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = check_grid(nx, ny, nz);
    if (rc) return rc;
    FREE0(M.ds_feat);
    FREE0(M.ds_tgt);
    long long n = (long long)nx * ny * nz;
    TT_HIP(cnn_alloc(&M.ds_feat, (size_t)max_samples * n * NF * 4));
    TT_HIP(cnn_alloc(&M.ds_tgt, (size_t)max_samples * n * 4));
    M.ds_cap = max_samples;
    M.ds_n = 0;
    M.ds_cells = n;
    M.ds_nx = nx; M.ds_ny = ny; M.ds_nz = nz;
    return 0;
}

int tt_cnn_collect(int enable) {
    if (enable && !M.ds_feat) return fail(1, "learned guess: tt_cnn_reserve first");
    M.collect = enable != 0;
    return 0;
}

int tt_cnn_dataset_size(void) { return M.ds_n; }

int tt_cnn_fit_norm(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (M.ds_n < 1) return fail(1, "learned guess: empty training set");
    int rc = ensure_model();
    if (rc) return rc;
    long long n = (long long)M.ds_n * M.ds_cells;
    const int nb = 1024;
    double* part;
    TT_HIP(hipMalloc(&part, (size_t)nb * 12 * 8));
    k_moments<<<nb, 256, 0, L.s>>>(M.ds_feat, M.ds_tgt, n, part);
    hipError_t e = hipGetLastError();
    std::vector<double> h((size_t)nb * 12);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, L.s);
    if (e == hipSuccess) e = hipStreamSynchronize(L.s);
    hipFree(part);
    if (e != hipSuccess) return fail(10, std::string("moments: ") + hipGetErrorString(e));
    float nrm[12];
    for (int q = 0; q < 6; ++q) {
        double s = 0, s2 = 0;
        for (int b = 0; b < nb; ++b) { s += h[(size_t)b * 12 + q]; s2 += h[(size_t)b * 12 + 6 + q]; }
        double mean = s / (double)n, var = s2 / (double)n - mean * mean;
        nrm[q] = (float)mean;
        nrm[6 + q] = (float)std::sqrt(var > 1e-30 ? var : 1e-30);
    }
    return tt_cnn_set_norm(nrm);
}

int tt_cnn_train(int epochs, int batch, float lr, uint64_t seed, float* epoch_loss) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!M.have_w || !M.have_norm) return fail(1, "learned guess: set weights and normalisation first");
    if (M.ds_n < 1) return fail(1, "learned guess: empty training set");
    if (batch < 1) batch = 1;
    Grid3 g{M.ds_nx, M.ds_ny, M.ds_nz};
    long long n = M.ds_cells;
    int rc = ensure_work(n);
    if (rc) return rc;
    std::vector<int> order(M.ds_n);
    for (int q = 0; q < M.ds_n; ++q) order[q] = q;
    std::mt19937_64 rng(seed);
    double* lp;
    const int lmax = 1024;
    TT_HIP(hipMalloc(&lp, (size_t)M.ds_n * lmax * 8));
    std::vector<int> lnbs(M.ds_n);
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        for (int s0 = 0; s0 < M.ds_n; s0 += batch) {
            int s1 = std::min(M.ds_n, s0 + batch);
            TT_HIP(hipMemsetAsync(M.gw, 0, NW * 4, L.s));
            TT_HIP(hipMemsetAsync(M.gb, 0, NB * 4, L.s));
            for (int s = s0; s < s1; ++s) {
                int id = order[s];
                k_norm_in<<<nblk(n, 256), 256, 0, L.s>>>(M.ds_feat + (long long)id * n * NF, M.dnorm, n, M.act[0]);
                TT_LAUNCH();
                k_norm_tgt<<<nblk(n, 256), 256, 0, L.s>>>(M.ds_tgt + (long long)id * n, M.dnorm, n, M.tg);
                TT_LAUNCH();
                if ((rc = forward(g, true))) return rc;
                if ((rc = backward(g, lp + (long long)s * lmax, &lnbs[s]))) return rc;
            }
            M.adam_t++;
            const float b1 = 0.9f, b2 = 0.999f;
            float c1 = 1.f - std::pow(b1, (float)M.adam_t), c2 = 1.f - std::pow(b2, (float)M.adam_t);
            float scale = 1.f / (float)(s1 - s0);
            k_adam<<<nblk(NW, 256), 256, 0, L.s>>>(M.w, M.gw, M.mw, M.vw, NW, lr, scale, b1, b2, 1e-8f, c1, c2);
            TT_LAUNCH();
            k_adam<<<1, 128, 0, L.s>>>(M.b, M.gb, M.mb, M.vb, NB, lr, scale, b1, b2, 1e-8f, c1, c2);
            TT_LAUNCH();
            if ((rc = repack())) return rc;
        }
        std::vector<double> h((size_t)M.ds_n * lmax);
        TT_HIP(hipMemcpyAsync(h.data(), lp, h.size() * 8, hipMemcpyDeviceToHost, L.s));
        TT_HIP(hipStreamSynchronize(L.s));
        double tot = 0.0;
        for (int s = 0; s < M.ds_n; ++s)
            for (int b = 0; b < lnbs[s]; ++b) tot += h[(size_t)s * lmax + b];
        if (epoch_loss) epoch_loss[ep] = (float)(tot / ((double)M.ds_n * (double)n));
    }
    hipFree(lp);
    return 0;
}

int tt_cnn_release(void) {
    if (L.s) hipStreamSynchronize(L.s);
    cnn_release();
    return 0;
}

}  // extern "C"
