// tt_noalloc.hip -- SCHRITT's Courant bound on caller-owned device arrays,
// allocating nothing (tt_schritt_noalloc, include/ttgpu.h).
//
// The block partials of the reduction live in `scratch`, a file-scope
// __device__ array of fixed size that each GPU holds from the library's load
// to the end of the process.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): at the first
// call on a device, with that device current, hipGetSymbolAddress gives that
// device's copy, kept per device in scratch_ptr[] and passed to the kernels.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two host threads) race.
#include <cmath>

#include "tt_common.h"

namespace tt {

constexpr int NA_BLOCKS = 1024;  // reduction blocks: the partials are 4 x 1024 doubles
constexpr int NA_MAXDEV = 64;

static __device__ __align__(256) double scratch[4 * NA_BLOCKS];
static double* scratch_ptr[NA_MAXDEV] = {};

// the six metric arrays SCHRITT reads, each from its declared lower bound
struct NaMet {
    const double *dxa, *dya, *dza;          // 1:n+2
    const double *geoxg, *geoxh, *geozg;    // -1:nx+2, -1:nx+2, -1:nz+2
};

// One pass over the ghosted box: minima of RH and T over the box, whether an
// interior cell has RH, T or PG <= 0 (ppsub.f:26-46), and the Courant bound
// over the interior (ppsub.f:49-78) with CS = sqrt(gamma PG / RH), SOUND's
// ICSTYP 0 (ppphy.f:92-98), formed in place.  Block minima into part[q*nb+b].
__global__ __launch_bounds__(256) void k_schritt_na(Geo g, NaMet m, const double* __restrict__ RH,
                                                     const double* __restrict__ T, const double* __restrict__ PG,
                                                     const double* __restrict__ V, const double* __restrict__ G,
                                                     const double* __restrict__ H, double gamma, double dtmax,
                                                     double* part, int nb) {
    __shared__ double sh[4][4];
    double mdt = dtmax, mrh = 1e308, mt = 1e308, bad = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.n3;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        box_ijk(g, t, i, j, k);
        const double rh = RH[t], tt = T[t];
        mrh = fmin(mrh, rh);
        mt = fmin(mt, tt);
        if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) {
            const double pg = PG[t];
            if (rh <= 0.0 || tt <= 0.0 || pg <= 0.0) bad = 1.0;
            const double v1 = fabs(V[g.idx(i + 1, j, k)]), v2 = fabs(G[g.idx(i, j + 1, k)]);
            const double v3 = fabs(H[g.idx(i, j, k + 1)]);
            const double cs = sqrt(gamma * pg / rh);
            const double z1 = m.dxa[i] / (v1 + cs);                                  // DXA(i+1)
            const double z2 = m.dya[j] / (v2 + cs / m.geoxg[i + 1] / m.geozg[k + 1]);  // DYA(j+1), GEOXG(i), GEOZG(k)
            const double z3 = m.dza[k] / (v3 + cs / m.geoxh[i + 1]);                  // DZA(k+1), GEOXH(i)
            mdt = fmin(mdt, fmin(z1, fmin(z2, z3)));
        }
    }
    const double a = wave_min(mdt), b = wave_min(mrh), c = wave_min(mt), d = wave_max(bad);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; sh[3][w] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = fmin(fmin(sh[0][0], sh[0][1]), fmin(sh[0][2], sh[0][3]));
        part[nb + blockIdx.x] = fmin(fmin(sh[1][0], sh[1][1]), fmin(sh[1][2], sh[1][3]));
        part[2 * nb + blockIdx.x] = fmin(fmin(sh[2][0], sh[2][1]), fmin(sh[2][2], sh[2][3]));
        part[3 * nb + blockIdx.x] = fmax(fmax(sh[3][0], sh[3][1]), fmax(sh[3][2], sh[3][3]));
    }
}

// One block folds the partials (a minimum does not depend on the order) and
// writes out[0] = FDELT min(min Courant, DTMAX) (ppsub.f:129-131) and
// out[1] = 1 where the reference stops on a non-positive RH, T or PG (its
// box minima of RH or T <= 0 and an interior cell <= 0), 2 where the bound
// is <= 0 (ppsub.f:71-76), else 0.
__global__ __launch_bounds__(256) void k_schritt_na_fold(const double* __restrict__ part, int nb, double fdelt,
                                                          double dtmax, double* out) {
    __shared__ double sh[4][4];
    double a = 1e308, b = 1e308, c = 1e308, d = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) {
        a = fmin(a, part[q]);
        b = fmin(b, part[nb + q]);
        c = fmin(c, part[2 * nb + q]);
        d = fmax(d, part[3 * nb + q]);
    }
    a = wave_min(a); b = wave_min(b); c = wave_min(c); d = wave_max(d);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; sh[3][w] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ldt = sh[0][0], mrh = sh[1][0], mt = sh[2][0], bad = sh[3][0];
        for (int q = 1; q < 4; ++q) {
            ldt = fmin(ldt, sh[0][q]);
            mrh = fmin(mrh, sh[1][q]);
            mt = fmin(mt, sh[2][q]);
            bad = fmax(bad, sh[3][q]);
        }
        double status = 0.0;
        if ((mrh <= 0.0 || mt <= 0.0) && bad > 0.0) status = 1.0;
        else if (ldt <= 0.0) status = 2.0;
        out[0] = (ldt < dtmax ? ldt : dtmax) * fdelt;
        out[1] = status;
    }
}

}  // namespace tt

using namespace tt;

extern "C" int tt_schritt_noalloc(int nx, int ny, int nz, const double* rh, const double* t, const double* pg,
                                  const double* v, const double* g, const double* h, const double* dxa,
                                  const double* dya, const double* dza, const double* geoxg, const double* geoxh,
                                  const double* geozg, double gamma, double fdelt, double dtmax, double* out,
                                  void* stream) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(1, "tt_schritt_noalloc: nx, ny, nz must be >= 1");
    const long long n3 = (long long)(nx + 4) * (ny + 4) * (nz + 4);
    if (n3 >= (1LL << 31))
        return fail(1, "tt_schritt_noalloc: the ghosted box must hold fewer than 2^31 cells (32-bit indexing)");
    if (!rh || !t || !pg || !v || !g || !h || !dxa || !dya || !dza || !geoxg || !geoxh || !geozg || !out)
        return fail(1, "tt_schritt_noalloc: null pointer");
    int dev = 0;
    TT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NA_MAXDEV) return fail(1, "tt_schritt_noalloc: device index beyond the scratch table");
    if (!scratch_ptr[dev]) {  // this device's copy of `scratch`, once per device, with it current
        void* p = nullptr;
        TT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_ptr[dev] = (double*)p;
    }
    Geo geo{};
    geo.nx = nx; geo.ny = ny; geo.nz = nz;
    geo.gx = nx + 4; geo.gy = ny + 4; geo.gz = nz + 4;
    geo.n3 = n3;
    NaMet m{dxa, dya, dza, geoxg, geoxh, geozg};
    hipStream_t s = (hipStream_t)stream;
    double* part = scratch_ptr[dev];
    k_schritt_na<<<NA_BLOCKS, 256, 0, s>>>(geo, m, rh, t, pg, v, g, h, gamma, dtmax, part, NA_BLOCKS);
    TT_LAUNCH();
    k_schritt_na_fold<<<1, 256, 0, s>>>(part, NA_BLOCKS, fdelt, dtmax, out);
    TT_LAUNCH();
    return 0;
}
