// tt_phys.hip -- BOUNDS (ppbou.f), PRESSG/QDENS/SOUND (ppphy.f), SCHRITT
// (ppsub.f) and DIFFUSION (ppdif.f) on gfx950.
//
// BOUNDS for MODTYP=2 on one rank (ppbou.f:1-92) is: PLANETSHIFT of
// T,RH,V,G,H; BOUNDPLANET (upperbnd search + PASSUPPERBND, BOUNDXMIN(bcxmin),
// BOUNDXMAX(2), BOUNDZMIN(0), BOUNDZMAX(0)); PLANETSHIFT again.  The kernels
// keep that order: the r boundaries of one (j,k) column touch only that
// column, the theta boundaries of one (i,j) line only that line, so one thread
// per column / line reproduces the reference's loops exactly.
#include <cmath>

#include "tt_common.h"

namespace tt {

int planetshift_impl(double* a, int f);

// PLANETSHIFT of T, RH, V, G, H (phi wrap of the two ghost planes), one
// thread per (i, k) line; b is the block's index within this work.  The
// ghost planes j = -1, 0, ny+1, ny+2 and their sources j = ny-1, ny, 1, 2 are
// disjoint (tt_init takes ny >= 4), so all the line's sources are read
// first, in one batch, and then written.
__device__ inline void shift5_body(const Geo& g, double* T, double* RH, double* V, double* G, double* H, int b) {
    int ni = g.nx + 4;
    long long n = (long long)ni * g.nz;
    long long t = (long long)b * 256 + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % ni) - 1, k = (int)(t / ni) + 1;
    const int ny = g.ny;
    double* f3[3] = {T, RH, H};
    const bool hv = i >= 0;  // V(0:nx+2,...)
    double s3[3][4], sv[4], sg[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double* a = f3[q];
        s3[q][0] = a[g.idx(i, ny, k)];
        s3[q][1] = a[g.idx(i, ny - 1, k)];
        s3[q][2] = a[g.idx(i, 1, k)];
        s3[q][3] = a[g.idx(i, 2, k)];
    }
    if (hv) {
        sv[0] = V[g.idx(i, ny, k)];
        sv[1] = V[g.idx(i, ny - 1, k)];
        sv[2] = V[g.idx(i, 1, k)];
        sv[3] = V[g.idx(i, 2, k)];
    }
    sg[0] = G[g.idx(i, ny, k)];
    sg[1] = G[g.idx(i, 1, k)];
    sg[2] = G[g.idx(i, 2, k)];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double* a = f3[q];
        a[g.idx(i, 0, k)] = s3[q][0];
        a[g.idx(i, -1, k)] = s3[q][1];
        a[g.idx(i, ny + 1, k)] = s3[q][2];
        a[g.idx(i, ny + 2, k)] = s3[q][3];
    }
    if (hv) {
        V[g.idx(i, 0, k)] = sv[0];
        V[g.idx(i, -1, k)] = sv[1];
        V[g.idx(i, ny + 1, k)] = sv[2];
        V[g.idx(i, ny + 2, k)] = sv[3];
    }
    // G(:,0:ny+2,:) has no j = -1 plane
    G[g.idx(i, 0, k)] = sg[0];
    G[g.idx(i, ny + 1, k)] = sg[1];
    G[g.idx(i, ny + 2, k)] = sg[2];
}

__global__ __launch_bounds__(256) void k_shift5(Geo g, double* T, double* RH, double* V, double* G, double* H,
                                                const int* run) {
    if (skip_run(run)) return;
    shift5_body(g, T, RH, V, G, H, blockIdx.x);
}

// upperbnd (ppbou.f:124-138) for every column incl. ghosts, which PASSUPPERBND
// (ppbou.f:522-682) fills from the phi-wrapped column and the nearest theta
// row when the theta/corner partners are MPI_PROC_NULL.  One wavefront per
// column: the lanes read 64 consecutive r cells (one coalesced load) and a
// ballot finds the first with rho below the threshold.
__device__ inline void upperbnd_body(const Geo& g, const double* __restrict__ RH, int32_t* ub, int b) {
    int nj = g.ny + 4, nk = g.nz + 4;
    int col = b * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (col >= nj * nk) return;
    int j = col % nj - 1, k = col / nj - 1;
    int jw = j <= 0 ? j + g.ny : (j > g.ny ? j - g.ny : j);
    int kc = k < 1 ? 1 : (k > g.nz ? g.nz : k);
    const double thr = (double)1.0e-9f;
    int u = g.nx - 1;
    for (int base = 1; base <= g.nx; base += 64) {
        int i = base + lane;
        bool low = i <= g.nx && RH[g.idx(i, jw, kc)] < thr;
        unsigned long long m = __ballot(low);
        if (m) {
            u = base + __ffsll((long long)m) - 1;
            break;
        }
    }
    if (lane == 0) ub[g.idx2(j, k)] = u;
}

// upperbnd again, right after a BOUNDS that searched it, with RH changed
// since only by that BOUNDS.  For an interior column (j, k), 2 <= k <= nz-1,
// that BOUNDS left RH(1..u-1) as they were, all at or above the threshold
// (u is the first cell below it, or nx-1 where none is), kept RH(u) unless
// u = nx, and set RH to 1e-10 above u and at nx when u = nx (BOUNDXMAX).  So
// the search now finds u if RH(u) is below the threshold and nx otherwise
// (u = nx-1 with no cell below: RH(nx) is now 1e-10); reading RH(u) is
// enough.  Columns k = 1 and nz hold copies of k = 2 and nz-1 (BOUNDZMIN/
// ZMAX copy all of them), and the ghost columns search the interior columns
// they wrap to, so each column takes the result of interior column (jw, ks).
__device__ inline void upperbnd_repeat_body(const Geo& g, const double* __restrict__ RH, int32_t* ub, int t) {
    int nj = g.ny + 4, nk = g.nz + 4;
    if (t >= nj * nk) return;
    int j = t % nj - 1, k = t / nj - 1;
    int jw = j <= 0 ? j + g.ny : (j > g.ny ? j - g.ny : j);
    int ks = k < 2 ? 2 : (k > g.nz - 1 ? g.nz - 1 : k);
    const double thr = (double)1.0e-9f;
    const int up = ub[g.idx2(jw, ks)];
    // the source column's own thread may have stored its result already:
    // this map is idempotent (u -> u where RH(u) is below, nx -> nx), so
    // either value gives the same
    const int u = RH[g.idx(up, jw, ks)] < thr ? up : g.nx;
    ub[g.idx2(j, k)] = u;
}

// BOUNDS' first PLANETSHIFT and upperbnd in one launch: upperbnd reads only
// the interior columns it wraps to, which the shift does not write, and the
// shift does not read upperbnd.  Blocks [0, nshift) shift, the rest search.
__global__ __launch_bounds__(256) void k_shift5_upperbnd(Geo g, double* T, double* RH, double* V, double* G,
                                                         double* H, int32_t* ub, int nshift, int repeat,
                                                         const int* run) {
    if (skip_run(run)) return;
    if ((int)blockIdx.x < nshift) shift5_body(g, T, RH, V, G, H, blockIdx.x);
    else if (repeat) upperbnd_repeat_body(g, RH, ub, (blockIdx.x - nshift) * 256 + threadIdx.x);
    else upperbnd_body(g, RH, ub, blockIdx.x - nshift);
}

// BOUNDXMIN(12) (ppbou.f:1118-1139) or (9) (:1029-1066), then BOUNDXMAX(2)
// (ppbou.f:1269-1395), one thread per (j,k) column.  The column's reads are
// issued first, in two batches (the bottom cells and upperbnd, then the cells
// at upperbnd), and the writes follow in the reference's order.  No read
// sees a cell this column writes earlier in that order, except T(u-1) with
// u = 1, which is the T(0) BOUNDXMIN has just set and is taken from it.
__global__ void k_xbc(Geo g, Met m, const int32_t* __restrict__ ub, double* T, double* RH, double* V, double* G,
                      double* H, const double* __restrict__ ER, const double* __restrict__ SIN,
                      const double* __restrict__ XKP, const double* __restrict__ XKR, int bcxmin, double d_t,
                      double fbottom, int ibdrad, int iii, double tirr, Ctrl* ctrl, const int* run) {
    if (skip_run(run)) return;
    int nj = g.ny + 4, nk = g.nz + 4;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nj * nk) return;
    int j = t % nj - 1, k = t / nj - 1;
    const long long c1 = g.idx(1, j, k);
    const bool inner = j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
    // batch 1
    const int u = ub[g.idx2(j, k)];
    const double t1 = T[c1];
    const double g1 = j >= 0 ? G[c1] : 0.0, h1 = k >= 0 ? H[c1] : 0.0;
    const bool flux9 = bcxmin != 12 && iii > 0;
    const double xkr1 = flux9 ? XKR[c1] : 0.0, rh1 = flux9 ? RH[c1] : 0.0;
    // batch 2: the cells at upperbnd (u is within 1..nx+1 here; bad columns stop below)
    const int uc = u < 1 ? 1 : (u > g.nx + 2 ? g.nx + 2 : u);
    const long long q = g.idx(uc - 1, j, k);
    double tum1 = T[q];
    const double gu = j >= 0 ? G[g.idx(uc, j, k)] : 0.0, hu = k >= 0 ? H[g.idx(uc, j, k)] : 0.0;
    const bool rad9 = inner && ibdrad == 9 && iii > 1;
    const double eq = rad9 ? ER[q] : 0.0, sq = rad9 ? SIN[q] : 0.0, rq = rad9 ? RH[q] : 0.0, kq = rad9 ? XKP[q] : 0.0;
    // BOUNDXMIN
    double t0;
    if (bcxmin == 12) {  // fixed dT/dr = D_T
        t0 = t1 + d_t * (m(TT_M_XXB, 0) - m(TT_M_XXB, 1));
        T[g.idx(0, j, k)] = t0;
        T[g.idx(-1, j, k)] = t1 + d_t * (m(TT_M_XXB, -1) - m(TT_M_XXB, 1));
    } else {  // 9: the radiative-diffusion gradient that carries FBottom
        double dtdr;
        if (iii > 0) {
            dtdr = -3.0 * fbottom * xkr1 * rh1 / (4.0 * ARAD * FD * (t1 * t1 * t1));  // T**3.d0 as T*T*T
        } else {
            dtdr = ctrl->bot_dtdr;  // cell (1,1,1)'s, from BurrowsOPC (k_bur_cell111)
        }
        t0 = t1 - dtdr * (m(TT_M_XXB, 1) - m(TT_M_XXB, 0));
        T[g.idx(0, j, k)] = t0;
        T[g.idx(-1, j, k)] = t1 - dtdr * (m(TT_M_XXB, 1) - m(TT_M_XXB, -1));
    }
    if (uc == 1) tum1 = t0;  // T(u-1) = the T(0) just set
    V[c1] = 0.0;
    V[g.idx(0, j, k)] = 0.0;
    if (j >= 0) {
        G[g.idx(0, j, k)] = g1;
        G[g.idx(-1, j, k)] = g1;
    }
    if (k >= 0) {
        H[g.idx(0, j, k)] = h1;
        H[g.idx(-1, j, k)] = h1;
    }
    // BOUNDXMAX(2): temperature, density and radial velocity above upperbnd
    if (inner) {
        if (u >= g.nx + 1) {
            ctrl->bad_bound = 1;
            return;
        }
        double TJ;
        if (ibdrad == 3) {
            TJ = tirr * pow(cos(m(TT_M_XYB, j)) * cos(m(TT_M_XZB, k)), 0.25);
            TJ = TJ > 100.0 ? TJ : 100.0;
        } else if (ibdrad == 4) {
            TJ = tirr;
        } else if (rad9) {
            TJ = pow((eq / ARAD) + (sq / (4.0 * SBCONST * rq * kq)), 0.25);
        } else {  // 7, or 9 at III <= 1
            TJ = tum1;
        }
        T[g.idx(u, j, k)] = TJ;
        const double rh_top = (double)1.0e-10f;
        for (int i = u + 1; i <= g.nx + 2; ++i) {
            T[g.idx(i, j, k)] = TJ;
            RH[g.idx(i, j, k)] = rh_top;
            V[g.idx(i, j, k)] = 0.0;
        }
        if (u == g.nx) {
            V[g.idx(g.nx, j, k)] = 0.0;
            RH[g.idx(g.nx, j, k)] = rh_top;
            T[g.idx(g.nx, j, k)] = TJ;
        }
    }
    // stress free above upperbnd
    if (j >= 0)
        for (int i = u + 1; i <= g.nx + 2; ++i) G[g.idx(i, j, k)] = gu;
    if (k >= 0)
        for (int i = u + 1; i <= g.nx + 2; ++i) H[g.idx(i, j, k)] = hu;
}

// BOUNDZMIN(0) (ppbou.f:1833-1878) then BOUNDZMAX(0) (ppbou.f:1946-1988),
// one thread per (i, j) line.  Each written plane takes its value from k = 2
// or nz-1 or from a plane written just before (k = 1, nz), and neither half
// writes the other's source plane (tt_init takes nz >= 4), so the line reads
// its two source planes first, in one batch, and forms every written value
// from them in the reference's order.
__global__ void k_zbc(Geo g, Met m, double* T, double* RH, double* V, double* G, double* H, const int* run) {
    if (skip_run(run)) return;
    int ni = g.nx + 4, nj = g.ny + 4;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj) return;
    int i = t % ni - 1, j = t / ni - 1;
    const int nz = g.nz;
    const bool hv = i >= 0, hg = j >= 0;
    const double rh2 = RH[g.idx(i, j, 2)], t2 = T[g.idx(i, j, 2)], g2 = hg ? G[g.idx(i, j, 2)] : 0.0;
    const double rhn = RH[g.idx(i, j, nz - 1)], tn = T[g.idx(i, j, nz - 1)], gn = hg ? G[g.idx(i, j, nz - 1)] : 0.0;
    RH[g.idx(i, j, 1)] = rh2;
    RH[g.idx(i, j, 0)] = rh2;
    RH[g.idx(i, j, -1)] = rh2;
    T[g.idx(i, j, 1)] = t2;
    T[g.idx(i, j, 0)] = t2;
    T[g.idx(i, j, -1)] = t2;
    H[g.idx(i, j, 2)] = 0.0;
    H[g.idx(i, j, 1)] = 0.0;
    H[g.idx(i, j, 0)] = 0.0;
    if (hv) {
        V[g.idx(i, j, 1)] = 0.0;
        V[g.idx(i, j, 0)] = 0.0;
        V[g.idx(i, j, -1)] = 0.0;
    }
    if (hg) {
        const double g1 = g2 * m(TT_M_SURZB, 2) / m(TT_M_SURZB, 1);
        G[g.idx(i, j, 1)] = g1;
        G[g.idx(i, j, 0)] = g1 * m(TT_M_SURZB, 1) / m(TT_M_SURZB, 0);
        G[g.idx(i, j, -1)] = g1 * m(TT_M_SURZB, 1) / cos(m(TT_M_XZB, -1));
    }
    RH[g.idx(i, j, nz)] = rhn;
    RH[g.idx(i, j, nz + 1)] = rhn;
    RH[g.idx(i, j, nz + 2)] = rhn;
    T[g.idx(i, j, nz)] = tn;
    T[g.idx(i, j, nz + 1)] = tn;
    T[g.idx(i, j, nz + 2)] = tn;
    H[g.idx(i, j, nz)] = 0.0;
    H[g.idx(i, j, nz + 1)] = 0.0;
    H[g.idx(i, j, nz + 2)] = 0.0;
    if (hv) {
        V[g.idx(i, j, nz)] = 0.0;
        V[g.idx(i, j, nz + 1)] = 0.0;
        V[g.idx(i, j, nz + 2)] = 0.0;
    }
    if (hg) {
        const double gz = gn * m(TT_M_SURZB, nz - 1) / m(TT_M_SURZB, nz);
        G[g.idx(i, j, nz)] = gz;
        G[g.idx(i, j, nz + 1)] = gz * m(TT_M_SURZB, nz) / m(TT_M_SURZB, nz + 1);
        G[g.idx(i, j, nz + 2)] = gz * m(TT_M_SURZB, nz) / m(TT_M_SURZB, nz + 2);
    }
}

int bounds_impl() {
    Geo& g = L.g;
    double** f = L.f;
    if (L.p.bcxmin != 12 && L.p.bcxmin != 9) return fail(1, "BOUNDXMIN: only types 9 and 12 are supported");
    if (L.p.bcxmin == 9 && L.visc.ieta == 5) return fail(1, "BOUNDXMIN(9) with IETA 5 is not supported");
    if (L.p.ibdrad != 3 && L.p.ibdrad != 4 && L.p.ibdrad != 7 && L.p.ibdrad != 9)
        return fail(1, "BOUNDXMAX(2): IBDRAD must be 3, 4, 7 or 9");
    long long ns = (long long)(g.nx + 4) * g.nz;
    int ncol = (g.ny + 4) * (g.nz + 4), nline = (g.nx + 4) * (g.ny + 4);
    const unsigned nshift = nblk(ns, 256);
    const int rep = L.ub_repeat ? 1 : 0;
    k_shift5_upperbnd<<<nshift + (rep ? nblk(ncol, 256) : nblk(ncol, 4)), 256, 0, L.s>>>(
        g, f[TT_T], f[TT_RH], f[TT_V], f[TT_G], f[TT_H], L.ub, (int)nshift, rep, L.run_if);
    TT_LAUNCH();
    if (L.p.bcxmin == 9 && L.iii <= 0) {
        if (L.run_if) return fail(1, "BOUNDS: BurrowsOPC at III <= 0 cannot be queued ahead");
        int rc = burrows_cell111_impl();
        if (rc) return rc;
    }
    k_xbc<<<nblk(ncol, 128), 128, 0, L.s>>>(g, met(), L.ub, f[TT_T], f[TT_RH], f[TT_V], f[TT_G], f[TT_H],
                                            f[TT_ER], f[TT_STELLARINPUT], f[TT_XKAPP], f[TT_XKAPR], L.p.bcxmin,
                                            L.p.d_t, L.p.fbottom, L.p.ibdrad, L.iii, L.p.tirr, L.ctrl, L.run_if);
    TT_LAUNCH();
    k_zbc<<<nblk(nline, 256), 256, 0, L.s>>>(g, met(), f[TT_T], f[TT_RH], f[TT_V], f[TT_G], f[TT_H], L.run_if);
    TT_LAUNCH();
    k_shift5<<<nblk(ns, 256), 256, 0, L.s>>>(g, f[TT_T], f[TT_RH], f[TT_V], f[TT_G], f[TT_H], L.run_if);
    TT_LAUNCH();
    L.ub_repeat = true;
    return 0;
}

__global__ void k_pressg(double* PG, const double* __restrict__ T, const double* __restrict__ RH, double mu,
                         long long n, const int* run) {
    if (skip_run(run)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) PG[t] = RGAS * T[t] * RH[t] / mu;
}

__global__ void k_sound(double* CS, const double* __restrict__ PG, const double* __restrict__ RH, double gam,
                        long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) CS[t] = sqrt(gam * PG[t] / RH[t]);
}

// QDENS (ppphy.f:47-76): face densities over their declared ranges
__global__ void k_qdens(Geo g, Met m, const double* __restrict__ RH, double* QX, double* QY, double* QZ) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    double r = RH[t];
    if (i >= 0) st_nt(&QX[t], m(TT_M_DDX1, i) * r + m(TT_M_DDX0, i) * RH[g.idx(i - 1, j, k)]);
    if (j >= 0) st_nt(&QY[t], m(TT_M_DDY1, j) * r + m(TT_M_DDY0, j) * RH[g.idx(i, j - 1, k)]);
    if (k >= 0) st_nt(&QZ[t], m(TT_M_DDZ1, k) * r + m(TT_M_DDZ0, k) * RH[g.idx(i, j, k - 1)]);
}

int pressg_impl() {
    k_pressg<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.f[TT_PG], L.f[TT_T], L.f[TT_RH], L.p.mu_gas, L.g.n3, L.run_if);
    TT_LAUNCH();
    return 0;
}

int qdens_impl() {
    k_qdens<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.g, met(), L.f[TT_RH], L.f[TT_RHQX], L.f[TT_RHQY], L.f[TT_RHQZ]);
    TT_LAUNCH();
    return 0;
}

int sound_impl() {
    k_sound<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.f[TT_CS], L.f[TT_PG], L.f[TT_RH], L.p.gamma, L.g.n3);
    TT_LAUNCH();
    return 0;
}

// ---------------------------------------------------------------- SCHRITT
// ppsub.f:1-137: minima of RH and T over the whole box, then the Courant
// limit over the interior (block minima, last block folds them).
__global__ __launch_bounds__(256) void k_schritt(Geo g, Met m, const double* __restrict__ RH,
                                                  const double* __restrict__ T, const double* __restrict__ PG,
                                                  const double* __restrict__ V, const double* __restrict__ G,
                                                  const double* __restrict__ H, const double* __restrict__ CS,
                                                  const double* __restrict__ ETA, double dtmax, double* part, int nb,
                                                  Ctrl* ctrl) {
    __shared__ double sh[3][4];
    double mdt = dtmax, mrh = 1e308, mt = 1e308;
    int badint = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < g.n3;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        box_ijk(g, t, i, j, k);
        mrh = fmin(mrh, RH[t]);
        mt = fmin(mt, T[t]);
        if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) {
            if (RH[t] <= 0.0 || T[t] <= 0.0 || PG[t] <= 0.0) badint = 1;
            double v1 = fabs(V[g.idx(i + 1, j, k)]), v2 = fabs(G[g.idx(i, j + 1, k)]), v3 = fabs(H[g.idx(i, j, k + 1)]);
            double cs = CS[t];
            double z1 = m(TT_M_DXA, i + 1) / (v1 + cs);
            double z2 = m(TT_M_DYA, j + 1) / (v2 + cs / m(TT_M_GEOXG, i) / m(TT_M_GEOZG, k));
            double z3 = m(TT_M_DZA, k + 1) / (v3 + cs / m(TT_M_GEOXH, i));
            mdt = fmin(mdt, fmin(z1, fmin(z2, z3)));
            if (ETA) {  // the viscous condition, IVIS > 0 (ppsub.f:97-119)
                const double xnu = ETA[t] / RH[t];
                const double zcf = (1.0 / 8.0) / xnu /
                                   (pow(m(TT_M_DXA, i + 1), -2.0) +
                                    pow(m(TT_M_DYA, j + 1) * m(TT_M_GEOXG, i) * m(TT_M_GEOZG, k), -2.0) +
                                    pow(m(TT_M_DZA, k + 1) * m(TT_M_GEOXH, i), -2.0));
                mdt = fmin(mdt, zcf);
            }
        }
    }
    if (badint) atomicOr(&ctrl->bad_state, 2);
    double a = wave_min(mdt), b = wave_min(mrh), c = wave_min(mt);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q)
            part[q * nb + blockIdx.x] = fmin(fmin(sh[q][0], sh[q][1]), fmin(sh[q][2], sh[q][3]));
    }
}

// Volume integrals over the interior cells, dV = volxb(i) volyb(j) volzb(k):
// thermal energy rho c_v T, radiation energy E and the irradiation heating
// STELLARINPUT (per-block partials, folded by k_fold)
__global__ __launch_bounds__(256) void k_energy_totals(Geo g, Met m, const double* __restrict__ RH,
                                                        const double* __restrict__ CV, const double* __restrict__ T,
                                                        const double* __restrict__ ER,
                                                        const double* __restrict__ SIN, double* part, int nb) {
    __shared__ double sh[4];
    double a = 0.0, b = 0.0, c = 0.0;
    long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) {
        int i, j, k;
        sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
        long long q = g.idx(i, j, k);
        double dv = m(TT_M_VOLXB, i) * m(TT_M_VOLYB, j) * m(TT_M_VOLZB, k);
        a += RH[q] * CV[q] * T[q] * dv;
        b += ER[q] * dv;
        c += SIN[q] * dv;
    }
    a = block_sum256(a, sh);
    b = block_sum256(b, sh);
    c = block_sum256(c, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[nb + blockIdx.x] = b;
        part[2 * nb + blockIdx.x] = c;
    }
}

// -------------------------------------------------------------- DIFFUSION
// DIFFUSION_COEF + FILL_DVALUES + DIFFUSION_MAT_ELEM (ppdif.f:169-308):
// constant diffusivity c = xnue/Pr, SOR matrix with the zero-gradient top
// row; RHS = the field.
__global__ void k_dif_matrix(Geo g, Met m, double delt, double c, const double* __restrict__ fld, double* SA,
                             double* SB, double* SC, double* SD, double* SE, double* SF, double* SG, double* RHS) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long n = g.idx(i, j, k);
    double rx0 = c / m(TT_M_DXB, i), rx1 = c / m(TT_M_DXB, i + 1);
    double ry0 = c / (m(TT_M_DYB, j) * m(TT_M_GEOXG, i) * m(TT_M_GEOZG, k));
    double ry1 = c / (m(TT_M_DYB, j + 1) * m(TT_M_GEOXG, i) * m(TT_M_GEOZG, k));
    double rz0 = c / (m(TT_M_DZB, k) * m(TT_M_GEOXH, i)), rz1 = c / (m(TT_M_DZB, k + 1) * m(TT_M_GEOXH, i));
    double sxb = m(TT_M_SURXB, i), dxa = m(TT_M_DXA, i);
    double gx = m(TT_M_GEOXG, i), szb = m(TT_M_SURZB, k), dya = m(TT_M_DYA, j), dza = m(TT_M_DZA, k);
    double d1 = delt * m(TT_M_SURXA, i + 1) * rx1 / (sxb * dxa);
    double d2 = delt * m(TT_M_SURXA, i) * rx0 / (sxb * dxa);
    double d3 = delt * ry1 / (gx * szb * dya);
    double d4 = delt * ry0 / (gx * szb * dya);
    double d5 = delt * rz1 * m(TT_M_SURZA, k + 1) / (gx * szb * dza);
    double d6 = delt * rz0 * m(TT_M_SURZA, k) / (gx * szb * dza);
    double d16 = d1 + d2 + d3 + d4 + d5 + d6;
    if (i == g.nx) {
        SA[n] = 1.0 + d2; SB[n] = -d2;
        SC[n] = 0.0; SD[n] = 0.0; SE[n] = 0.0; SF[n] = 0.0; SG[n] = 0.0;
    } else {
        SA[n] = 1.0 + d16; SB[n] = -d2; SC[n] = -d1; SD[n] = -d4; SE[n] = -d3; SF[n] = -d6; SG[n] = -d5;
    }
    RHS[n] = fld[n];
}

int sync_check() {
    int rc0 = ctrl_to_host();
    if (rc0) return rc0;
    return check_flags();
}

int check_flags() {
    if (L.hctrl->bad_bound) return fail(3, "error in boundxmax: upperbnd above the domain");
    if (L.hctrl->bad_q) return fail(2, "ERROR IN DELTA_Q: non-positive value of Q");
    if (L.hctrl->bad_kappa)
        return fail(2, L.p.ikaptyp >= 4 && L.p.ikaptyp <= 7
                           ? "NEG. T. IN FREEDMAN OPC (or an undefined pressure)"
                           : "NEG. T. IN BURROWS OPC (or a density outside the table)");
    return 0;
}

int diffusion_impl(double delt, double xnue, double pr, int* iters) {
    Geo& g = L.g;
    double** f = L.f;
    long long n = (long long)g.nx * g.ny * g.nz;
    k_dif_matrix<<<nblk(n, 256), 256, 0, L.s>>>(g, met(), delt, xnue / pr, f[TT_T], f[TT_SOR_A], f[TT_SOR_B],
                                                f[TT_SOR_C], f[TT_SOR_D], f[TT_SOR_E], f[TT_SOR_F], f[TT_SOR_G],
                                                f[TT_RHS_ER]);
    TT_LAUNCH();
    sor_sys_module_written();
    tt_solver_opts keep = L.opts;
    L.opts.precision = 0;
    tt_sor_stats st{};
    int rc = sor_solve_impl(0, &st, 1);
    L.opts = keep;
    if (iters) *iters = st.sor_iter;
    return rc;
}

// VARLIMIT (ppsub.f:140-154) over a declared sub-box of a ghosted field
__global__ void k_clamp(Geo g, double* a, int lo0, int lo1, int lo2, int n0, int n1, long long n, double vmin,
                        double vmax) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % n0) + lo0;
    long long r = t / n0;
    int j = (int)(r % n1) + lo1, k = (int)(r / n1) + lo2;
    long long c = g.idx(i, j, k);
    double x = a[c];
    x = x < vmax ? x : vmax;
    x = x > vmin ? x : vmin;
    a[c] = x;
}

// the G and H loops of SPHERICALVELLIMIT (ppsub.f:364-384), no poles: over
// the whole box, G for j /= -1 and H for k /= -1 (their declared ranges)
__global__ void k_vel_sph(Geo g, Met m, double* G, double* H, double vmin, double vmax) {
    long long n = (long long)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx + 4, g.ny + 4, -1, -1, -1, i, j, k);
    long long c = g.idx(i, j, k);
    double xxb = m(TT_M_XXB, i);
    if (j != -1) {
        double x = G[c];
        double top = vmax / (xxb * m(TT_M_GEOZG, k));  // geozg = cos(xzb), ppinit.f:1396
        x = x < top ? x : top;
        double bot = vmin / (xxb * m(TT_M_GEOZG, k));
        x = x > bot ? x : bot;
        G[c] = x;
    }
    if (k != -1) {
        double x = H[c];
        double top = vmax / xxb;
        x = x < top ? x : top;
        double bot = vmin / xxb;
        x = x > bot ? x : bot;
        H[c] = x;
    }
}

static int varlimit_impl(int f, double vmin, double vmax) {
    if (f != TT_T && f != TT_RH && f != TT_V && f != TT_G && f != TT_H)
        return fail(1, "VARLIMIT: field must be T, RH, V, G or H");
    Geo& g = L.g;
    int32_t lo[3], hi[3], rank = 0, isint = 0;
    tt_field_bounds(f, g.nx, g.ny, g.nz, lo, hi, &rank, &isint);
    int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1;
    long long n = (long long)n0 * n1 * (hi[2] - lo[2] + 1);
    k_clamp<<<nblk(n, 256), 256, 0, L.s>>>(g, L.f[f], lo[0], lo[1], lo[2], n0, n1, n, vmin, vmax);
    TT_LAUNCH();
    if (f == TT_RH) L.ub_repeat = false;
    return 0;
}

// SPHERICALVELLIMIT's ceiling (ppsub.f:338-363): DAY = 8.64E4 is a REAL
// constant (ppmod.f:1083, exact in float); TIME_START = VELDAMP_START = 0
int sphericalvellimit_impl(int ivel, double velmax, double hydrotime) {
    double vmin, vmax;
    if (ivel == 1 || ivel == 3) {
        vmin = -velmax;
        vmax = velmax;
    } else if (ivel == 2 || ivel == 4) {
        const double day = (double)8.64e4f, time_start = 0.0 * day, damp_time = 25.0 * day, start = 0.0;
        const double pi = 3.141592653589793;
        if (hydrotime >= time_start && hydrotime <= time_start + damp_time) {
            vmax = (velmax - start) * pow(sin((pi / 2.0) * (hydrotime - time_start) / damp_time), (double)4.0f) +
                   start;
            vmin = -vmax;
        } else if (hydrotime > time_start + damp_time) {
            vmax = velmax;
            vmin = -velmax;
        } else {
            vmax = start;
            vmin = -start;
        }
    } else {
        return fail(1, "SPHERICALVELLIMIT: ivel must be 1, 2, 3 or 4");
    }
    if (ivel == 1 || ivel == 2) {
        Geo& g = L.g;
        long long n = (long long)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
        k_vel_sph<<<nblk(n, 256), 256, 0, L.s>>>(g, met(), L.f[TT_G], L.f[TT_H], vmin, vmax);
        TT_LAUNCH();
    }
    return varlimit_impl(TT_V, vmin, vmax);
}

// the clamps of pp.f's step loop: stage 0 before ADVECT3D (velocity,
// :543-557; temperature, :559-562; then the time advance, :637-639),
// stage 1 after ADVECT3D (:673-678), stage 2 after FORGAS3D (:708-710)
int limits_impl(int stage, double delt) {
    const tt_limits& l = L.lim;
    int rc = 0;
    if (stage == 0) {
        if (l.ivel != 0) {
            if ((rc = varlimit_impl(TT_V, -l.velmax, l.velmax))) return rc;
            if (l.ivel < 3 && (rc = sphericalvellimit_impl(l.ivel, l.velmax, l.hydrotime))) return rc;
        }
        if (l.itemp != 0 && (rc = varlimit_impl(TT_T, l.tempmin, l.tempmax))) return rc;
        L.lim.hydrotime = l.hydrotime + delt;
    } else if (stage == 1) {
        if (l.irho != 0 && (rc = varlimit_impl(TT_RH, l.rhmin, 1.0e10))) return rc;
        if (l.itemp != 0 && (rc = varlimit_impl(TT_T, l.tempmin, l.tempmax))) return rc;
    } else {
        if (l.itemp != 0 && (rc = varlimit_impl(TT_T, l.tempmin, l.tempmax))) return rc;
    }
    return 0;
}

}  // namespace tt

using namespace tt;

extern "C" {

int tt_set_limits(const tt_limits* l) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!l) return fail(1, "null limits");
    if (l->ivel < 0 || l->ivel > 4) return fail(1, "IVEL must be 0, 1, 2, 3 or 4");
    L.lim = *l;
    return 0;
}

int tt_get_limits(tt_limits* l) {
    if (!l) return fail(1, "null limits");
    *l = L.lim;
    return 0;
}

int tt_varlimit(int field, double varmin, double varmax) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = varlimit_impl(field, varmin, varmax);
    return rc ? rc : sync_check();
}

int tt_sphericalvellimit(int ivel, double velmax, double hydrotime) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = sphericalvellimit_impl(ivel, velmax, hydrotime);
    return rc ? rc : sync_check();
}

int tt_bounds(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = bounds_impl();
    if (rc) return rc;
    return sync_check();
}

int tt_pressg(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = pressg_impl();
    return rc ? rc : sync_check();
}

int tt_qdens(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = qdens_impl();
    return rc ? rc : sync_check();
}

int tt_sound(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = sound_impl();
    return rc ? rc : sync_check();
}

int tt_schritt(double* delt) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!delt) return fail(1, "null delt");
    Geo& g = L.g;
    if (L.p.fdelt == 0.0) {
        *delt = L.p.dtmax;
        return 0;
    }
    int rc = sound_impl();
    if (rc) return rc;
    int nb = 1024;
    if (3 * nb > 2 * L.npartial) nb = 2 * L.npartial / 3;
    k_schritt<<<nb, 256, 0, L.s>>>(g, met(), L.f[TT_RH], L.f[TT_T], L.f[TT_PG], L.f[TT_V], L.f[TT_G], L.f[TT_H],
                                   L.f[TT_CS], L.visc.ivis > 0 ? L.f[TT_ETA] : nullptr, L.p.dtmax, L.partial, nb,
                                   L.ctrl);
    TT_LAUNCH();
    rc = fold_partials(L.partial, nb, 1, 2, &L.ctrl->dtmin);
    if (rc) return rc;
    rc = fold_partials(L.partial + nb, nb, 2, 2, &L.ctrl->scratch[1], &L.ctrl->scratch[2]);
    if (rc) return rc;
    rc = sync_check();
    if (rc) return rc;
    if (L.hctrl->scratch[1] <= 0.0 || L.hctrl->scratch[2] <= 0.0) {
        if (L.hctrl->bad_state & 2) {
            L.hctrl->bad_state = 0;
            hipMemcpyAsync(L.ctrl, L.hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, L.s);
            hipStreamSynchronize(L.s);
            return fail(2, "non-positive value of rh/pg/t");
        }
    }
    if (L.hctrl->bad_state) {
        L.hctrl->bad_state = 0;
        hipMemcpyAsync(L.ctrl, L.hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, L.s);
        hipStreamSynchronize(L.s);
    }
    double ldt = L.hctrl->dtmin;
    if (ldt <= 0.0) return fail(2, "local dt negative");
    if (L.visc.iartvis > 0) {  // ppsub.f:127-128
        const double da = dt_artvis();
        ldt = da < ldt ? da : ldt;
    }
    ldt = (ldt < L.p.dtmax ? ldt : L.p.dtmax) * L.p.fdelt;
    *delt = ldt;
    return 0;
}

int tt_diffusion(double delt, double xnue, double pr_num, int* iters) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = diffusion_impl(delt, xnue, pr_num, iters);
    int rc2 = sync_check();
    return rc2 ? rc2 : rc;
}

}  // extern "C"

extern "C" int tt_energy_totals(double* out3) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!out3) return fail(1, "null output");
    const int nb = 1024;
    if (3 * nb > 2 * L.npartial) return fail(1, "energy totals: partial buffer too small");
    k_energy_totals<<<nb, 256, 0, L.s>>>(L.g, met(), L.f[TT_RH], L.f[TT_CV], L.f[TT_T], L.f[TT_ER],
                                         L.f[TT_STELLARINPUT], L.partial, nb);
    TT_LAUNCH();
    int rc = fold_partials(L.partial, nb, 3, 0, &L.ctrl->scratch[5], &L.ctrl->scratch[6], &L.ctrl->scratch[7]);
    if (rc) return rc;
    rc = ctrl_to_host();
    if (rc) return rc;
    out3[0] = L.hctrl->scratch[5];
    out3[1] = L.hctrl->scratch[6];
    out3[2] = L.hctrl->scratch[7];
    return 0;
}
