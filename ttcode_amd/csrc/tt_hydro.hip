// tt_hydro.hip -- ADVECT3D (ppadv.f:525-1577) and FORGAS3D (ppfor.f:142-397)
// on gfx950, plus the step driver.
//
// The reference advects with 1-D line sweeps that fill per-line scratch
// (S, IND, GRAD) and then the face fluxes.  Every face flux depends only on
// the cells within two of it along its line, so here each thread evaluates
// the fluxes through its own cell's faces directly from the state (the two
// faces shared with a neighbour are evaluated by both; the arithmetic is the
// reference's, so both copies are equal).  Mass fluxes, which the momentum
// sweeps reuse through face averages, are stored once.
#include <cmath>

#include "tt_common.h"

namespace tt {

#define IX(i, j, k) g.idx(i, j, k)

// van Leer slope (e.g. ppadv.f:568-575).  order 0: (dlo*whi + dhi*wlo) as in
// the density and energy sweeps; order 1: (dhi*wlo + dlo*whi) as in the
// momentum sweeps.  w_lo/w_hi are the spacing at the cell and the next one.
__device__ inline double vl_slope(double qm, double q0, double qp, double wlo, double whi, int order) {
    double g1 = (q0 - qm) * (qp - q0);
    if (!(g1 > 0.0)) return 0.0;
    double den = order == 0 ? ((q0 - qm) * whi + (qp - q0) * wlo) : ((qp - q0) * wlo + (q0 - qm) * whi);
    return 2.0 * g1 / den;
}

struct Box3 {
    const double* a;
    Geo g;
    __device__ double operator()(int i, int j, int k) const { return a[g.idx(i, j, k)]; }
};

// ---- mass fluxes (ppadv.f:559-644) and their phi plane exchange (:654-711)
__global__ void k_adv_mass(Geo g, Met m, double dt, const double* __restrict__ RH, const double* __restrict__ V,
                           const double* __restrict__ G, const double* __restrict__ H, double* FLMX, double* FLMY,
                           double* FLMZ) {
    long long n = (long long)(g.nx + 1) * (g.ny + 2) * (g.nz + 1);
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx + 1, g.ny + 2, 1, 0, 1, i, j, k);
    const double half = 0.5;
    Box3 R{RH, g};
    // x faces: i = 1..nx+1, j = 0..ny (j = 0 is the periodic copy of j = ny)
    if (k <= g.nz) {
        int jj = j == 0 ? g.ny : j;
        if (jj <= g.ny) {
            double v = V[IX(i, jj, k)];
            double S = v * dt;
            int ind = v > 0.0 ? i - 1 : i;
            double gr = vl_slope(R(ind - 1, jj, k), R(ind, jj, k), R(ind + 1, jj, k), m(TT_M_DXB, ind), m(TT_M_DXB, ind + 1), 0);
            FLMX[IX(i, j, k)] = (m(TT_M_SURXA, i) - m(TT_M_CCXA, i) * S) * S *
                                (R(ind, jj, k) + (m(TT_M_XXA, i) - m(TT_M_XXB, ind) - half * S) * gr);
        }
    }
    // y faces: i = 1..nx, j = 1..ny+1, k = 1..nz
    if (i <= g.nx && j >= 1 && k <= g.nz) {
        double gv = G[IX(i, j, k)];
        double S = gv * dt;
        int ind = gv > 0.0 ? j - 1 : j;
        double gr = vl_slope(R(i, ind - 1, k), R(i, ind, k), R(i, ind + 1, k), m(TT_M_DYB, ind), m(TT_M_DYB, ind + 1), 0);
        FLMY[IX(i, j, k)] = S * (m(TT_M_SURYA, j) + m(TT_M_CCYA, j) * S) *
                            (R(i, ind, k) + (m(TT_M_XYA, j) - m(TT_M_XYB, ind) - half * S) * gr);
    }
    // z faces: i = 1..nx, j = 0..ny (j = 0 periodic copy), k = 1..nz+1
    if (i <= g.nx && j <= g.ny) {
        int jj = j == 0 ? g.ny : j;
        double hv = H[IX(i, jj, k)];
        double S = hv * dt;
        int ind = hv > 0.0 ? k - 1 : k;
        double gr = vl_slope(R(i, jj, ind - 1), R(i, jj, ind), R(i, jj, ind + 1), m(TT_M_DZB, ind), m(TT_M_DZB, ind + 1), 0);
        FLMZ[IX(i, j, k)] = (m(TT_M_SURZA, k) + m(TT_M_CCZA, k) * S) * S *
                            (R(i, jj, ind) + (m(TT_M_XZA, k) - m(TT_M_XZB, ind) - half * S) * gr);
    }
}

// ---- thermal energy ED = T*RH*CV (ppadv.f:741-891).  The product is formed
// once per cell of the box (k_ed0, the reference's expression), so the
// sweep reads one value per stencil point instead of three.
__global__ void k_ed0(const double* __restrict__ T, const double* __restrict__ RH, const double* __restrict__ CV,
                      double* E0, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) E0[t] = T[t] * RH[t] * CV[t];
}

__global__ void k_adv_energy(Geo g, Met m, double dt, const double* __restrict__ E0, const double* __restrict__ V,
                             const double* __restrict__ G, const double* __restrict__ H, double* ED) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    const double half = 0.5;
    auto E = [&](int a, int b, int c) { return E0[g.idx(a, b, c)]; };
    double fx[2], fy[2], fz[2];
    for (int q = 0; q < 2; ++q) {
        int I = i + q;
        double S = V[IX(I, j, k)] * dt;
        int ind = V[IX(I, j, k)] > 0.0 ? I - 1 : I;
        double gr = vl_slope(E(ind - 1, j, k), E(ind, j, k), E(ind + 1, j, k), m(TT_M_DXB, ind), m(TT_M_DXB, ind + 1), 0);
        fx[q] = (m(TT_M_SURXA, I) - m(TT_M_CCXA, I) * S) * S * (E(ind, j, k) + (m(TT_M_XXA, I) - m(TT_M_XXB, ind) - half * S) * gr);
        int J = j + q;
        S = G[IX(i, J, k)] * dt;
        ind = G[IX(i, J, k)] > 0.0 ? J - 1 : J;
        gr = vl_slope(E(i, ind - 1, k), E(i, ind, k), E(i, ind + 1, k), m(TT_M_DYB, ind), m(TT_M_DYB, ind + 1), 0);
        fy[q] = (m(TT_M_SURYA, J) - m(TT_M_CCYA, J) * S) * S * (E(i, ind, k) + (m(TT_M_XYA, J) - m(TT_M_XYB, ind) - half * S) * gr);
        int K = k + q;
        S = H[IX(i, j, K)] * dt;
        ind = H[IX(i, j, K)] > 0.0 ? K - 1 : K;
        gr = vl_slope(E(i, j, ind - 1), E(i, j, ind), E(i, j, ind + 1), m(TT_M_DZB, ind), m(TT_M_DZB, ind + 1), 0);
        fz[q] = (m(TT_M_SURZA, K) - m(TT_M_CCZA, K) * S) * S * (E(i, j, ind) + (m(TT_M_XZA, K) - m(TT_M_XZB, ind) - half * S) * gr);
    }
    ED[IX(i, j, k)] = E(i, j, k) - (fx[1] - fx[0]) / m(TT_M_VOLXB, i) - (fy[1] - fy[0]) / m(TT_M_VOLYB, j) -
                      (fz[1] - fz[0]) / m(TT_M_VOLZB, k);
}

// ---- radial momentum (ppadv.f:896-1073); cells i = 2..nx
__global__ void k_adv_v(Geo g, Met m, double dt, const double* __restrict__ V, const double* __restrict__ G,
                        const double* __restrict__ H, const double* __restrict__ FLMX, const double* __restrict__ FLMY,
                        const double* __restrict__ FLMZ, const double* __restrict__ RHQX, double* VD) {
    long long n = (long long)(g.nx - 1) * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx - 1, g.ny, 2, 1, 1, i, j, k);
    const double half = 0.5;
    Box3 Vb{V, g};
    double fx[2], fy[2], fz[2];
    for (int q = 0; q < 2; ++q) {
        int I = i - 1 + q;  // FLX(I) for I = i-1, i
        double S = (Vb(I, j, k) + Vb(I + 1, j, k)) * dt / 2.0;
        int ind = S > 0.0 ? I : I + 1;
        double gr = vl_slope(Vb(ind - 1, j, k), Vb(ind, j, k), Vb(ind + 1, j, k), m(TT_M_DXA, ind), m(TT_M_DXA, ind + 1), 1);
        fx[q] = half * (FLMX[IX(I, j, k)] + FLMX[IX(I + 1, j, k)]) *
                (Vb(ind, j, k) + (m(TT_M_XXB, I) - m(TT_M_XXA, ind) - half * S) * gr);
        int J = j + q;
        S = (m(TT_M_DDX1, i) * G[IX(i, J, k)] + m(TT_M_DDX0, i) * G[IX(i - 1, J, k)]) * dt;
        ind = S > 0.0 ? J - 1 : J;
        gr = vl_slope(Vb(i, ind - 1, k), Vb(i, ind, k), Vb(i, ind + 1, k), m(TT_M_DYB, ind), m(TT_M_DYB, ind + 1), 1);
        fy[q] = (m(TT_M_DDX1, i) * FLMY[IX(i, J, k)] + m(TT_M_DDX0, i) * FLMY[IX(i - 1, J, k)]) *
                (Vb(i, ind, k) + (m(TT_M_XYA, J) - m(TT_M_XYB, ind) - half * S) * gr);
        int K = k + q;
        S = (m(TT_M_DDX1, i) * H[IX(i, j, K)] + m(TT_M_DDX0, i) * H[IX(i - 1, j, K)]) * dt;
        ind = S > 0.0 ? K - 1 : K;
        gr = vl_slope(Vb(i, j, ind - 1), Vb(i, j, ind), Vb(i, j, ind + 1), m(TT_M_DZB, ind), m(TT_M_DZB, ind + 1), 1);
        fz[q] = (m(TT_M_DDX1, i) * FLMZ[IX(i, j, K)] + m(TT_M_DDX0, i) * FLMZ[IX(i - 1, j, K)]) *
                (Vb(i, j, ind) + (m(TT_M_XZA, K) - m(TT_M_XZB, ind) - half * S) * gr);
    }
    VD[IX(i, j, k)] = Vb(i, j, k) * RHQX[IX(i, j, k)] - (fx[1] - fx[0]) / m(TT_M_VOLXA, i) -
                      (fy[1] - fy[0]) / m(TT_M_VOLYB, j) - (fz[1] - fz[0]) / m(TT_M_VOLZB, k);
}

// GD = (G + OMROT) r^2 cos^2 (ppadv.f:1078-1084); the phi ghosts equal the
// formula at the wrapped column because G's ghosts were wrapped by BOUNDS.
__global__ void k_gd_init(Geo g, Met m, double om, const double* __restrict__ G, double* GD) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    if (j < 0) return;
    double gx = m(TT_M_GEOXG, i), gz = m(TT_M_GEOZG, k);
    GD[t] = (G[t] + om) * (gx * gx) * (gz * gz);
}

// ---- azimuthal angular momentum (ppadv.f:1096-1266)
__global__ void k_adv_g(Geo g, Met m, double dt, const double* __restrict__ V, const double* __restrict__ G,
                        const double* __restrict__ H, const double* __restrict__ GD, const double* __restrict__ FLMX,
                        const double* __restrict__ FLMY, const double* __restrict__ FLMZ,
                        const double* __restrict__ RHQY, double* GDN) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    const double half = 0.5;
    Box3 D{GD, g};
    double fx[2], fy[2], fz[2];
    for (int q = 0; q < 2; ++q) {
        int I = i + q;
        double S = (m(TT_M_DDY1, j) * V[IX(I, j, k)] + m(TT_M_DDY0, j) * V[IX(I, j - 1, k)]) * dt;
        int ind = S > 0.0 ? I - 1 : I;
        double gr = vl_slope(D(ind - 1, j, k), D(ind, j, k), D(ind + 1, j, k), m(TT_M_DXB, ind), m(TT_M_DXB, ind + 1), 1);
        fx[q] = (m(TT_M_DDY1, j) * FLMX[IX(I, j, k)] + m(TT_M_DDY0, j) * FLMX[IX(I, j - 1, k)]) *
                (D(ind, j, k) + (m(TT_M_XXA, I) - m(TT_M_XXB, ind) - half * S) * gr);
        // FLY(J) for J = j-1, j; FLY(0) is the periodic copy of FLY(ny)
        int J = j - 1 + q;
        if (J == 0) J = g.ny;
        S = (m(TT_M_DCY0, J + 1) * G[IX(i, J, k)] + m(TT_M_DCY1, J + 1) * G[IX(i, J + 1, k)]) * dt;
        ind = S > 0.0 ? J : J + 1;
        gr = vl_slope(D(i, ind - 1, k), D(i, ind, k), D(i, ind + 1, k), m(TT_M_DYA, ind), m(TT_M_DYA, ind + 1), 1);
        fy[q] = (m(TT_M_DCY0, J + 1) * FLMY[IX(i, J, k)] + m(TT_M_DCY1, J + 1) * FLMY[IX(i, J + 1, k)]) *
                (D(i, ind, k) + (m(TT_M_XYB, J) - m(TT_M_XYA, ind) - half * S) * gr);
        int K = k + q;
        S = (m(TT_M_DDY1, j) * H[IX(i, j, K)] + m(TT_M_DDY0, j) * H[IX(i, j - 1, K)]) * dt;
        ind = S > 0.0 ? K - 1 : K;
        gr = vl_slope(D(i, j, ind - 1), D(i, j, ind), D(i, j, ind + 1), m(TT_M_DZB, ind), m(TT_M_DZB, ind + 1), 1);
        fz[q] = (m(TT_M_DDY1, j) * FLMZ[IX(i, j, K)] + m(TT_M_DDY0, j) * FLMZ[IX(i, j - 1, K)]) *
                (D(i, j, ind) + (m(TT_M_XZA, K) - m(TT_M_XZB, ind) - half * S) * gr);
    }
    GDN[IX(i, j, k)] = D(i, j, k) * RHQY[IX(i, j, k)] - (fx[1] - fx[0]) / m(TT_M_VOLXB, i) -
                       (fy[1] - fy[0]) / m(TT_M_VOLYA, j) - (fz[1] - fz[0]) / m(TT_M_VOLZB, k);
}

__global__ void k_hd_init(Geo g, Met m, const double* __restrict__ H, double* HD) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    if (k < 0) return;
    double gh = m(TT_M_GEOXH, i);
    HD[t] = H[t] * (gh * gh);
}

// ---- meridional angular momentum (ppadv.f:1278-1461); cells k = 2..nz
__global__ void k_adv_h(Geo g, Met m, double dt, const double* __restrict__ V, const double* __restrict__ G,
                        const double* __restrict__ H, const double* __restrict__ HD, const double* __restrict__ FLMX,
                        const double* __restrict__ FLMY, const double* __restrict__ FLMZ,
                        const double* __restrict__ RHQZ, double* HDN) {
    long long n = (long long)g.nx * g.ny * (g.nz - 1);
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 2, i, j, k);
    const double half = 0.5;
    Box3 D{HD, g};
    double fx[2], fy[2], fz[2];
    for (int q = 0; q < 2; ++q) {
        int I = i + q;
        double S = (m(TT_M_DDZ1, k) * V[IX(I, j, k)] + m(TT_M_DDZ0, k) * V[IX(I, j, k - 1)]) * dt;
        int ind = S > 0.0 ? I - 1 : I;
        double gr = vl_slope(D(ind - 1, j, k), D(ind, j, k), D(ind + 1, j, k), m(TT_M_DXB, ind), m(TT_M_DXB, ind + 1), 1);
        fx[q] = (m(TT_M_DDZ1, k) * FLMX[IX(I, j, k)] + m(TT_M_DDZ0, k) * FLMX[IX(I, j, k - 1)]) *
                (D(ind, j, k) + (m(TT_M_XXA, I) - m(TT_M_XXB, ind) - half * S) * gr);
        int J = j + q;
        S = (m(TT_M_DDZ1, k) * G[IX(i, J, k)] + m(TT_M_DDZ0, k) * G[IX(i, J, k - 1)]) * dt;
        ind = S > 0.0 ? J - 1 : J;
        gr = vl_slope(D(i, ind - 1, k), D(i, ind, k), D(i, ind + 1, k), m(TT_M_DYB, ind), m(TT_M_DYB, ind + 1), 1);
        fy[q] = (m(TT_M_DDZ1, k) * FLMY[IX(i, J, k)] + m(TT_M_DDZ0, k) * FLMY[IX(i, J, k - 1)]) *
                (D(i, ind, k) + (m(TT_M_XYA, J) - m(TT_M_XYB, ind) - half * S) * gr);
        int K = k - 1 + q;  // FLZ(K) for K = k-1, k
        S = half * (H[IX(i, j, K)] + H[IX(i, j, K + 1)]) * dt;
        ind = S > 0.0 ? K : K + 1;
        gr = vl_slope(D(i, j, ind - 1), D(i, j, ind), D(i, j, ind + 1), m(TT_M_DZA, ind), m(TT_M_DZA, ind + 1), 1);
        fz[q] = half * (FLMZ[IX(i, j, K)] + FLMZ[IX(i, j, K + 1)]) *
                (D(i, j, ind) + (m(TT_M_XZB, K) - m(TT_M_XZA, ind) - half * S) * gr);
    }
    HDN[IX(i, j, k)] = D(i, j, k) * RHQZ[IX(i, j, k)] - (fx[1] - fx[0]) / m(TT_M_VOLXB, i) -
                       (fy[1] - fy[0]) / m(TT_M_VOLYB, j) - (fz[1] - fz[0]) / m(TT_M_VOLZA, k);
}

__global__ void k_adv_rho(Geo g, Met m, double* RH, const double* __restrict__ FLMX, const double* __restrict__ FLMY,
                          const double* __restrict__ FLMZ) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long c = IX(i, j, k);
    RH[c] = RH[c] - (FLMX[IX(i + 1, j, k)] - FLMX[c]) / m(TT_M_VOLXB, i) -
            (FLMY[IX(i, j + 1, k)] - FLMY[c]) / m(TT_M_VOLYB, j) - (FLMZ[IX(i, j, k + 1)] - FLMZ[c]) / m(TT_M_VOLZB, k);
}

// new V, G, H, T (ppadv.f:1498-1575); GD keeps the updated momentum density
__global__ void k_adv_final(Geo g, Met m, double om, double* V, double* G, double* H, double* T, double* GD,
                            const double* __restrict__ VD, const double* __restrict__ GDN,
                            const double* __restrict__ HDN, const double* __restrict__ ED,
                            const double* __restrict__ RH, const double* __restrict__ CV,
                            const double* __restrict__ RHQX, const double* __restrict__ RHQY,
                            const double* __restrict__ RHQZ) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long c = IX(i, j, k);
    if (i >= 2) V[c] = VD[c] / RHQX[c];
    double gx = m(TT_M_GEOXG, i), gz = m(TT_M_GEOZG, k), gh = m(TT_M_GEOXH, i);
    GD[c] = GDN[c];
    G[c] = GDN[c] / (RHQY[c] * (gx * gx) * (gz * gz)) - om;
    if (k >= 2) H[c] = HDN[c] / (RHQZ[c] * (gh * gh));
    T[c] = ED[c] / (RH[c] * CV[c]);
}

int advect_impl(double dt) {
    Geo& g = L.g;
    double** f = L.f;
    int rc;
    if ((rc = bounds_impl())) return rc;
    long long nm = (long long)(g.nx + 1) * (g.ny + 2) * (g.nz + 1);
    k_adv_mass<<<nblk(nm, 256), 256, 0, L.s>>>(g, met(), dt, f[TT_RH], f[TT_V], f[TT_G], f[TT_H], L.FLMX, L.FLMY, L.FLMZ);
    TT_LAUNCH();
    if ((rc = qdens_impl())) return rc;
    long long nc = (long long)g.nx * g.ny * g.nz;
    k_ed0<<<nblk(g.n3, 256), 256, 0, L.s>>>(f[TT_T], f[TT_RH], f[TT_CV], L.TOLD, g.n3);
    TT_LAUNCH();
    k_adv_energy<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, L.TOLD, f[TT_V], f[TT_G], f[TT_H], L.ED);
    TT_LAUNCH();
    long long nv = (long long)(g.nx - 1) * g.ny * g.nz;
    k_adv_v<<<nblk(nv, 256), 256, 0, L.s>>>(g, met(), dt, f[TT_V], f[TT_G], f[TT_H], L.FLMX, L.FLMY, L.FLMZ, f[TT_RHQX], L.VD);
    TT_LAUNCH();
    k_gd_init<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, met(), L.p.omrot, f[TT_G], f[TT_GD]);
    TT_LAUNCH();
    k_adv_g<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, f[TT_V], f[TT_G], f[TT_H], f[TT_GD], L.FLMX, L.FLMY, L.FLMZ,
                                            f[TT_RHQY], L.GDN);
    TT_LAUNCH();
    k_hd_init<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, met(), f[TT_H], L.HH);
    TT_LAUNCH();
    long long nh = (long long)g.nx * g.ny * (g.nz - 1);
    k_adv_h<<<nblk(nh, 256), 256, 0, L.s>>>(g, met(), dt, f[TT_V], f[TT_G], f[TT_H], L.HH, L.FLMX, L.FLMY, L.FLMZ,
                                            f[TT_RHQZ], L.HDN);
    TT_LAUNCH();
    k_adv_rho<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), f[TT_RH], L.FLMX, L.FLMY, L.FLMZ);
    TT_LAUNCH();
    if ((rc = bounds_impl())) return rc;
    if ((rc = qdens_impl())) return rc;
    k_adv_final<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), L.p.omrot, f[TT_V], f[TT_G], f[TT_H], f[TT_T], f[TT_GD], L.VD,
                                                L.GDN, L.HDN, L.ED, f[TT_RH], f[TT_CV], f[TT_RHQX], f[TT_RHQY],
                                                f[TT_RHQZ]);
    TT_LAUNCH();
    return 0;
}

// ------------------------------------------------------------- FORGAS3D
// CENTRIFCORIOLIS, NCOSYS=2 (ppfor.f:435-486)
__global__ void k_centcor(Geo g, Met m, double om, const double* __restrict__ G, const double* __restrict__ H,
                          double* CCX, double* CCY, double* CCZ) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    double G1 = m(TT_M_DCY0, j + 1) * G[IX(i - 1, j, k)] + m(TT_M_DCY1, j + 1) * G[IX(i - 1, j + 1, k)] + om;
    double G2 = m(TT_M_DCY0, j + 1) * G[IX(i, j, k)] + m(TT_M_DCY1, j + 1) * G[IX(i, j + 1, k)] + om;
    double GQk = m(TT_M_DDX0KEP, i) * G1 + m(TT_M_DDX1KEP, i) * G2;
    double H1 = m(TT_M_DDZ0, k + 1) * H[IX(i - 1, j, k)] + m(TT_M_DDZ1, k + 1) * H[IX(i - 1, j, k + 1)];
    double H2 = m(TT_M_DDZ0, k + 1) * H[IX(i, j, k)] + m(TT_M_DDZ1, k + 1) * H[IX(i, j, k + 1)];
    double HQ = m(TT_M_DDX0, i) * H1 + m(TT_M_DDX1, i) * H2;
    double xa = m(TT_M_XXA, i);
    double tq = GQk / xa * m(TT_M_GEOZG, k);
    long long c = IX(i, j, k);
    CCX[c] = tq * tq + xa * (HQ * HQ);
    CCY[c] = 0.0;
    G1 = m(TT_M_DCY0, j + 1) * G[IX(i, j, k - 1)] + m(TT_M_DCY1, j + 1) * G[IX(i, j + 1, k - 1)];
    G2 = m(TT_M_DCY0, j + 1) * G[IX(i, j, k)] + m(TT_M_DCY1, j + 1) * G[IX(i, j + 1, k)];
    double GQ = m(TT_M_DDZ1, k) * G2 + m(TT_M_DDZ0, k) * G1 + om;
    double za = m(TT_M_XZA, k);
    CCZ[c] = -m(TT_M_XXB, i) * GQ * GQ * sin(za) * cos(za);
}

__device__ inline double divv(const Geo& g, const Met& m, const double* V, const double* G, const double* H, int i,
                              int j, int k) {
    return (m(TT_M_SURXA, i + 1) * V[g.idx(i + 1, j, k)] - m(TT_M_SURXA, i) * V[g.idx(i, j, k)]) / m(TT_M_VOLXB, i) +
           (m(TT_M_SURYA, j + 1) * G[g.idx(i, j + 1, k)] - m(TT_M_SURYA, j) * G[g.idx(i, j, k)]) / m(TT_M_VOLYB, j) +
           (m(TT_M_SURZA, k + 1) * H[g.idx(i, j, k + 1)] - m(TT_M_SURZA, k) * H[g.idx(i, j, k)]) / m(TT_M_VOLZB, k);
}

// p dV work: T -= fac*dt PG/RH/CV div(v) over cells lo..hi in each axis
__global__ void k_pdv(Geo g, Met m, double dtf, int lo, int hiadd, double* T, const double* __restrict__ PG,
                      const double* __restrict__ RH, const double* __restrict__ CV, const double* __restrict__ V,
                      const double* __restrict__ G, const double* __restrict__ H) {
    int ni = g.nx + hiadd - lo + 1, nj = g.ny + hiadd - lo + 1, nk = g.nz + hiadd - lo + 1;
    long long n = (long long)ni * nj * nk;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, ni, nj, lo, lo, lo, i, j, k);
    long long c = IX(i, j, k);
    T[c] = T[c] - dtf * PG[c] / RH[c] / CV[c] * divv(g, m, V, G, H, i, j, k);
}

// VH, GH, HH (ppfor.f:217-278); DIVT* are zero with IVIS=0
__global__ void k_force_vh(Geo g, Met m, double dt, const double* __restrict__ V, const double* __restrict__ G,
                           const double* __restrict__ H, const double* __restrict__ PG, const double* __restrict__ QX,
                           const double* __restrict__ QY, const double* __restrict__ QZ, const double* __restrict__ CCX,
                           const double* __restrict__ CCY, const double* __restrict__ CCZ,
                           const double* __restrict__ GX, const double* __restrict__ GY, const double* __restrict__ GZ,
                           double* VH, double* GH, double* HH) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long c = IX(i, j, k);
    double gx = m(TT_M_GEOXG, i), gz = m(TT_M_GEOZG, k), gh = m(TT_M_GEOXH, i);
    VH[c] = V[c] + dt * ((-(PG[c] - PG[IX(i - 1, j, k)]) / m(TT_M_DXB, i) + 0.0) / QX[c] + CCX[c] + GX[c]);
    GH[c] = G[c] + dt * (((-(PG[c] - PG[IX(i, j - 1, k)]) / m(TT_M_DYB, j) / gx / gz + 0.0) / QY[c] + CCY[c] + GY[c]) / gx / gz);
    HH[c] = H[c] + dt * (((-(PG[c] - PG[IX(i, j, k - 1)]) / m(TT_M_DZB, k) / gh + 0.0) / QZ[c] + CCZ[c] + GZ[c]) / gh);
}

// mode 0: V = (V+VH)/2 etc.; mode 1: V = VH etc.
__global__ void k_force_set(Geo g, int mode, double* V, double* G, double* H, const double* __restrict__ VH,
                            const double* __restrict__ GH, const double* __restrict__ HH) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long c = IX(i, j, k);
    if (mode == 0) {
        V[c] = (V[c] + VH[c]) / 2.0;
        G[c] = (G[c] + GH[c]) / 2.0;
        H[c] = (H[c] + HH[c]) / 2.0;
    } else {
        V[c] = VH[c];
        G[c] = GH[c];
        H[c] = HH[c];
    }
}

int forgas_impl(double dt) {
    Geo& g = L.g;
    double** f = L.f;
    int rc;
    if ((rc = bounds_impl())) return rc;
    if ((rc = pressg_impl())) return rc;
    long long nc = (long long)g.nx * g.ny * g.nz;
    k_centcor<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), L.p.omrot, f[TT_G], f[TT_H], f[TT_CENTCORX], f[TT_CENTCORY],
                                              f[TT_CENTCORZ]);
    TT_LAUNCH();
    TT_HIP(hipMemcpyAsync(L.TOLD, f[TT_T], (size_t)g.n3 * sizeof(double), hipMemcpyDeviceToDevice, L.s));
    if (L.p.isoth == 0) {
        long long nh = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
        k_pdv<<<nblk(nh, 256), 256, 0, L.s>>>(g, met(), dt / 2.0, 0, 1, f[TT_T], f[TT_PG], f[TT_RH], f[TT_CV], f[TT_V],
                                              f[TT_G], f[TT_H]);
        TT_LAUNCH();
        if ((rc = bounds_impl())) return rc;
        if ((rc = pressg_impl())) return rc;
    }
    k_force_vh<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, f[TT_V], f[TT_G], f[TT_H], f[TT_PG], f[TT_RHQX], f[TT_RHQY],
                                               f[TT_RHQZ], f[TT_CENTCORX], f[TT_CENTCORY], f[TT_CENTCORZ], f[TT_GRAVX],
                                               f[TT_GRAVY], f[TT_GRAVZ], L.VH, L.GH, L.HH);
    TT_LAUNCH();
    TT_HIP(hipMemcpyAsync(f[TT_T], L.TOLD, (size_t)g.n3 * sizeof(double), hipMemcpyDeviceToDevice, L.s));
    if (L.p.isoth == 0) {
        k_force_set<<<nblk(nc, 256), 256, 0, L.s>>>(g, 0, f[TT_V], f[TT_G], f[TT_H], L.VH, L.GH, L.HH);
        TT_LAUNCH();
        if ((rc = bounds_impl())) return rc;
        k_pdv<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, 1, 0, f[TT_T], f[TT_PG], f[TT_RH], f[TT_CV], f[TT_V], f[TT_G],
                                              f[TT_H]);
        TT_LAUNCH();
    }
    k_force_set<<<nblk(nc, 256), 256, 0, L.s>>>(g, 1, f[TT_V], f[TT_G], f[TT_H], L.VH, L.GH, L.HH);
    TT_LAUNCH();
    return bounds_impl();
}

}  // namespace tt

using namespace tt;

extern "C" {

int tt_advect3d(double delt) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = advect_impl(delt);
    return rc ? rc : sync_check();
}

int tt_forgas3d(double delt) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = forgas_impl(delt);
    return rc ? rc : sync_check();
}

int tt_step(double delt, int iii, tt_step_stats* st) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    tt_step_stats s{};
    hipEvent_t e[4];
    for (auto& x : e) TT_HIP(hipEventCreate(&x));
    L.iii = iii;
    TT_HIP(hipEventRecord(e[0], L.s));
    int rc = limits_impl(0, delt);
    if (rc) return rc;
    rc = advect_impl(delt);
    if (rc) return rc;
    if ((rc = limits_impl(1, delt))) return rc;
    TT_HIP(hipEventRecord(e[1], L.s));
    rc = forgas_impl(delt);
    if (rc) return rc;
    if ((rc = limits_impl(2, delt))) return rc;
    TT_HIP(hipEventRecord(e[2], L.s));
    int erc = energy_impl(delt, iii, &s.energy);
    if (erc > 0) return erc;
    if ((rc = bounds_impl())) return rc;
    if ((rc = bounds_impl())) return rc;
    if ((rc = pressg_impl())) return rc;
    TT_HIP(hipEventRecord(e[3], L.s));
    rc = sync_check();
    if (rc) return rc;
    float a = 0, b = 0, c = 0;
    hipEventElapsedTime(&a, e[0], e[1]);
    hipEventElapsedTime(&b, e[1], e[2]);
    hipEventElapsedTime(&c, e[2], e[3]);
    s.ms_advect = a;
    s.ms_force = b;
    s.ms_energy = c;
    s.ms_total = (double)a + b + c;
    for (auto& x : e) hipEventDestroy(x);
    if (st) *st = s;
    return erc;
}

}  // extern "C"
