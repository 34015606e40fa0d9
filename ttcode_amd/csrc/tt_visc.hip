// tt_visc.hip -- viscosity in FORGAS3D (ppfor.f:170-174): ETACALC
// (ppvisc.f:375-484) and VISC (ppvisc.f:2-373) for IVISTYP 1 in spherical
// coordinates without poles.
//
// VISC forms the viscous stress tensor from the velocities: the diagonal
// components TXX, TYY, TZZ and the compression term DIVV at cell centres,
// the shear components TXY, TXZ, TYZ at cell edges; then eta-weighted edge
// averages at the centres (TXYAV...), the dissipation function DISFN, the
// edge components scaled by the edge viscosity, and the divergence of the
// tensor on the faces, divtx/y/z.  Four passes, each one thread per cell over
// the index range the reference's loop covers, with the intermediate tensor
// components in scratch arrays: a pass reads the neighbours of the one
// before.  Every expression keeps the reference's operation order.
#include <cmath>
#include <vector>

#include "tt_common.h"

namespace tt {

#define IX(i, j, k) g.idx(i, j, k)

namespace {

// scratch: DIVV, TXX, TYY, TZZ, TXY, TXZ, TYZ, TXYAV, TYZAV, TXZAV
enum { S_DIVV, S_TXX, S_TYY, S_TZZ, S_TXY, S_TXZ, S_TYZ, S_AXY, S_AYZ, S_AXZ, S_N };
double* g_scr[S_N] = {};
double* g_nu = nullptr;  // IETA 3: the radial profile of the kinematic viscosity
double g_nu_xnue = 0.0;

struct VArgs {
    const double *V, *G, *H, *ETA;
    double* s[S_N];
    double *DISFN, *DTX, *DTY, *DTZ;
};

// ETACALC: IETA 0 and 5, eta = RH XNUE; IETA 4, eta = RH min(CS mu_gas c /
// RH, XNUE) with c = 1.67 10^-8 (the host forms it as the reference does)
// IETA 2: eta = RH XNUE_VAL (the host forms the ramped value); IETA 3: eta =
// RH nu(i) with the radial profile nu(i) from the host (nu_i, index i + 1)
__global__ void k_eta(Geo g, int ieta, double xnue, double mu, double c, const double* __restrict__ nu_i,
                      const double* __restrict__ RH, const double* __restrict__ CS, double* ETA) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    if (ieta == 4) {
        double e = CS[t] * mu * c / RH[t];
        e = e < xnue ? e : xnue;
        ETA[t] = e * RH[t];
    } else if (nu_i) {
        int i, j, k;
        box_ijk(g, t, i, j, k);
        ETA[t] = RH[t] * nu_i[i + 1];
    } else {
        ETA[t] = RH[t] * xnue;
    }
}

// pass 1 over (0:nx+2)^3: the edge components (ppvisc.f:95-118) everywhere,
// the centre components and DIVV (:30-42, :75-88) up to n+1
__global__ void k_visc_a(Geo g, Met m, VArgs a) {
    const int ni = g.nx + 3, nj = g.ny + 3, nk = g.nz + 3;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    int i, j, k;
    sub_ijk(t, ni, nj, 0, 0, 0, i, j, k);
    const long long c = IX(i, j, k);
    const double *V = a.V, *G = a.G, *H = a.H;
    const double alprat = 1.0;
    const double gxga = m(TT_M_GEOXGA, i), gzga = m(TT_M_GEOZGA, k), gxha = m(TT_M_GEOXHA, i);
    const double gxg = m(TT_M_GEOXG, i), gxh = m(TT_M_GEOXH, i);
    a.s[S_TXY][c] = 0.5 * ((V[c] - V[IX(i, j - 1, k)]) / (m(TT_M_XYB, j) - m(TT_M_XYB, j - 1)) / gxga / gzga +
                           (G[c] - G[IX(i - 1, j, k)]) / (m(TT_M_XXB, i) - m(TT_M_XXB, i - 1)) * gxga * gzga / alprat);
    double txz = 0.5 * ((V[c] - V[IX(i, j, k - 1)]) / (m(TT_M_XZB, k) - m(TT_M_XZB, k - 1)) / gxha +
                        (H[c] - H[IX(i - 1, j, k)]) / (m(TT_M_XXB, i) - m(TT_M_XXB, i - 1)) * gxha);
    if (gxha < (double)1.e-10f) txz = 0.0;
    a.s[S_TXZ][c] = txz;
    a.s[S_TYZ][c] = 0.5 * ((H[c] - H[IX(i, j - 1, k)]) / (m(TT_M_XYB, j) - m(TT_M_XYB, j - 1)) / gxg / gzga * gxh +
                           (G[c] - G[IX(i, j, k - 1)]) / (m(TT_M_XZB, k) - m(TT_M_XZB, k - 1)) * gxg * gzga / gxh /
                               alprat);
    if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
    const double eta = a.ETA[c];
    const double xxb = m(TT_M_XXB, i);
    a.s[S_TXX][c] = 2.0 * eta * (V[IX(i + 1, j, k)] - V[c]) / (m(TT_M_XXA, i + 1) - m(TT_M_XXA, i));
    a.s[S_DIVV][c] = 2.0 / 3.0 * eta * divv(g, m, V, G, H, i, j, k);
    a.s[S_TYY][c] = 2.0 * eta *
                    ((G[IX(i, j + 1, k)] - G[c]) / (m(TT_M_XYA, j + 1) - m(TT_M_XYA, j)) +
                     (V[IX(i + 1, j, k)] + V[c]) / 2.0 / xxb + (H[IX(i, j, k + 1)] + H[c]) / 2.0 * (-tan(m(TT_M_XZB, k))));
    a.s[S_TZZ][c] = 2.0 * eta *
                    ((H[IX(i, j, k + 1)] - H[c]) / (m(TT_M_XZA, k + 1) - m(TT_M_XZA, k)) +
                     (V[IX(i + 1, j, k)] + V[c]) / 2.0 / xxb);
}

// pass 2 over (0:nx+1)^3: the eta-weighted averages of the unscaled edge
// components (:127-139) and DISFN (:155-167)
__global__ void k_visc_b(Geo g, VArgs a, int idiss) {
    const int ni = g.nx + 2, nj = g.ny + 2, nk = g.nz + 2;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    int i, j, k;
    sub_ijk(t, ni, nj, 0, 0, 0, i, j, k);
    const long long c = IX(i, j, k);
    const double *TXY = a.s[S_TXY], *TXZ = a.s[S_TXZ], *TYZ = a.s[S_TYZ];
    const double eta = a.ETA[c];
    const double axy = eta * (TXY[c] + TXY[IX(i + 1, j, k)] + TXY[IX(i, j + 1, k)] + TXY[IX(i + 1, j + 1, k)]) / 2.0;
    const double ayz = eta * (TYZ[c] + TYZ[IX(i, j + 1, k)] + TYZ[IX(i, j, k + 1)] + TYZ[IX(i, j + 1, k + 1)]) / 2.0;
    const double axz = eta * (TXZ[c] + TXZ[IX(i + 1, j, k)] + TXZ[IX(i, j, k + 1)] + TXZ[IX(i + 1, j, k + 1)]) / 2.0;
    a.s[S_AXY][c] = axy;
    a.s[S_AYZ][c] = ayz;
    a.s[S_AXZ][c] = axz;
    if (idiss) {
        const double alprat = 1.0;
        const double dv = a.s[S_DIVV][c], txx = a.s[S_TXX][c], tyy = a.s[S_TYY][c], tzz = a.s[S_TZZ][c];
        a.DISFN[c] = (3.0 * dv * dv + txx * txx + 2.0 * axy * axy * alprat + tyy * tyy + 2.0 * ayz * ayz * alprat +
                      tzz * tzz + 2.0 * axz * axz) /
                     2.0 / eta;
    }
}

// pass 3 over (0:nx+2)^3: the edge components times their edge viscosity
// (:141-152), in place (each cell reads only itself and eta)
__global__ void k_visc_c(Geo g, VArgs a) {
    const int ni = g.nx + 3, nj = g.ny + 3, nk = g.nz + 3;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    int i, j, k;
    sub_ijk(t, ni, nj, 0, 0, 0, i, j, k);
    const long long c = IX(i, j, k);
    const double* E = a.ETA;
    a.s[S_TXY][c] = a.s[S_TXY][c] / 2.0 * (E[c] + E[IX(i - 1, j, k)] + E[IX(i, j - 1, k)] + E[IX(i - 1, j - 1, k)]);
    a.s[S_TXZ][c] = a.s[S_TXZ][c] / 2.0 * (E[c] + E[IX(i - 1, j, k)] + E[IX(i, j, k - 1)] + E[IX(i - 1, j, k - 1)]);
    a.s[S_TYZ][c] = a.s[S_TYZ][c] / 2.0 * (E[c] + E[IX(i, j, k - 1)] + E[IX(i, j - 1, k)] + E[IX(i, j - 1, k - 1)]);
}

// pass 4 over (0:nx+1)^3: divtx for i >= 1, divty for j >= 1, divtz for
// k >= 1 (ppvisc.f:272-321, NCOSYS=2)
__global__ void k_visc_d(Geo g, Met m, VArgs a) {
    const int ni = g.nx + 2, nj = g.ny + 2, nk = g.nz + 2;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    int i, j, k;
    sub_ijk(t, ni, nj, 0, 0, 0, i, j, k);
    const long long c = IX(i, j, k);
    const double *DIVV = a.s[S_DIVV], *TXX = a.s[S_TXX], *TYY = a.s[S_TYY], *TZZ = a.s[S_TZZ];
    const double *TXY = a.s[S_TXY], *TXZ = a.s[S_TXZ], *TYZ = a.s[S_TYZ];
    const double *AXY = a.s[S_AXY], *AYZ = a.s[S_AYZ], *AXZ = a.s[S_AXZ];
    if (i >= 1) {
        const long long w = IX(i - 1, j, k);
        const double vxa = m(TT_M_VOLXA, i), sxb0 = m(TT_M_SURXB, i - 1), sxb1 = m(TT_M_SURXB, i);
        a.DTX[c] = (sxb0 * DIVV[w] - sxb1 * DIVV[c]) / vxa + (sxb1 * TXX[c] - sxb0 * TXX[w]) / vxa +
                   (m(TT_M_SURYA, j + 1) * TXY[IX(i, j + 1, k)] - m(TT_M_SURYA, j) * TXY[c]) / m(TT_M_VOLYB, j) /
                       m(TT_M_GEOXGA, i) / m(TT_M_GEOZG, k) +
                   (m(TT_M_SURZA, k + 1) * TXZ[IX(i, j, k + 1)] - m(TT_M_SURZA, k) * TXZ[c]) / m(TT_M_VOLZB, k) /
                       m(TT_M_GEOXHA, i) -
                   (TYY[c] + TYY[w] + TZZ[c] + TZZ[w] - (DIVV[c] + DIVV[w]) * 2.0) / 2.0 / m(TT_M_XXA, i);
    }
    if (j >= 1) {
        const long long s = IX(i, j - 1, k);
        const double vya = m(TT_M_VOLYA, j), syb0 = m(TT_M_SURYB, j - 1), syb1 = m(TT_M_SURYB, j);
        const double gxg = m(TT_M_GEOXG, i), gzg = m(TT_M_GEOZG, k);
        a.DTY[c] = (syb0 * DIVV[s] - syb1 * DIVV[c]) / vya / gxg / gzg +
                   (m(TT_M_SURXA, i + 1) * TXY[IX(i + 1, j, k)] - m(TT_M_SURXA, i) * TXY[c]) / m(TT_M_VOLXB, i) +
                   (syb1 * TYY[c] - syb0 * TYY[s]) / vya / gxg / gzg +
                   (m(TT_M_SURZA, k + 1) * TYZ[IX(i, j, k + 1)] - m(TT_M_SURZA, k) * TYZ[c]) / m(TT_M_VOLZB, k) /
                       m(TT_M_GEOXH, i) +
                   ((AYZ[c] + AYZ[s]) * (-tan(m(TT_M_XZB, k))) + (AXY[c] + AXY[s])) / 2.0 / m(TT_M_XXB, i);
    }
    if (k >= 1) {
        const long long b = IX(i, j, k - 1);
        const double vza = m(TT_M_VOLZA, k), szb0 = m(TT_M_SURZB, k - 1), szb1 = m(TT_M_SURZB, k);
        const double gxh = m(TT_M_GEOXH, i);
        a.DTZ[c] = (szb0 * DIVV[b] - szb1 * DIVV[c]) / vza / gxh +
                   (m(TT_M_SURXA, i + 1) * TXZ[IX(i + 1, j, k)] - m(TT_M_SURXA, i) * TXZ[c]) / m(TT_M_VOLXB, i) +
                   (m(TT_M_SURYA, j + 1) * TYZ[IX(i, j + 1, k)] - m(TT_M_SURYA, j) * TYZ[c]) / m(TT_M_VOLYB, j) /
                       m(TT_M_GEOXG, i) / m(TT_M_GEOZGA, k) +
                   (szb1 * TZZ[c] - szb0 * TZZ[b]) / vza / gxh +
                   ((TYY[c] + TYY[b] - (DIVV[c] + DIVV[b])) * tan(m(TT_M_XZA, k)) + AXZ[c] + AXZ[b]) / 2.0 /
                       m(TT_M_XXB, i);
    }
}

// ---------------------------------------------------------------------------
// NR_ARTIFICIAL_VISC (ppfor.f:1032-1230), von Neumann-Richtmyer, one rank
// without poles (z range 1..nz).  Per direction the reference sweeps the
// cells i = 1..upperbnd-2 of each column in order, forming the compression
// flux Q = 3 delt rho min(dv, 0)^2 on the cell and updating the velocity by
// the flux difference and T by the flux's work.  Each Q depends on the
// velocities before that direction's update only, so a direction is two
// parallel passes: the fluxes, then the updates; the passes are fused
// across directions (X update + Y fluxes, Y update + Z fluxes) as they touch
// disjoint arrays.  The lower neighbour's flux a Y or Z update reads is the
// one this call wrote, or, where upperbnd of that neighbour column is lower,
// what an earlier call left: QviscY and QviscZ persist, as the reference's
// module arrays do.  Each flux pass also takes the minimum of dv/dx over the
// fluxes it forms, the time-step bound (block partials, folded to
// ctrl->av_min).
enum { A_QX, A_DX, A_DY, A_DZ, A_QY, A_QZ, A_N };  // A_QY, A_QZ persist
double* g_avq[A_N] = {};

struct AArgs {
    double *V, *G, *H, *T;
    const double *RH, *CV, *RHQX, *RHQY, *RHQZ;
    const int32_t* ub;
    double* q[A_N];
    double coef;  // artvisc_coeff delt
    double* part;
    int nb;
};

__device__ inline void block_min_out(double mn, double* out) {
    __shared__ double sh[4];
    double a = wave_min(mn);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
}

// fluxes of one direction at cell c from the velocity difference d (the
// expression order of ppfor.f:1074-1075)
__device__ inline double nr_flux(double coef, double rh, double& d) {
    if (d > 0.0) d = 0.0;
    return coef * rh * (d * d);
}

// pass 1: X fluxes (ppfor.f:1066-1099), Q(0) for every column
__global__ __launch_bounds__(256) void k_av_x(Geo g, Met m, AArgs a) {
    double mn = 0.0;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
        if (i == 1) {
            const long long c0 = IX(0, j, k);
            double d = a.V[IX(1, j, k)] - a.V[c0];
            a.q[A_QX][c0] = nr_flux(a.coef, a.RH[c0], d);
            mn = fmin(mn, d / m(TT_M_DXB, 0));
        }
        if (i > a.ub[g.idx2(j, k)] - 2) continue;
        const long long c = IX(i, j, k);
        double d = a.V[IX(i + 1, j, k)] - a.V[c];
        a.q[A_QX][c] = nr_flux(a.coef, a.RH[c], d);
        a.q[A_DX][c] = d;
        mn = fmin(mn, d / m(TT_M_DXB, i));
    }
    block_min_out(mn, a.part + blockIdx.x);
}

// pass 2: X updates (ppfor.f:1076-1079), then Y fluxes (:1100-1142), Q(j=0)
// from the j = 1 cells
__global__ __launch_bounds__(256) void k_av_xy(Geo g, Met m, AArgs a) {
    double mn = 0.0;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
        if (i > a.ub[g.idx2(j, k)] - 2) continue;
        const long long c = IX(i, j, k);
        const double rh = a.RH[c], cv = a.CV[c];
        const double qx = a.q[A_QX][c];
        a.V[c] = a.V[c] - (qx - a.q[A_QX][IX(i - 1, j, k)]) / (m(TT_M_DXB, i) * a.RHQX[c]);
        a.T[c] = a.T[c] - qx * a.q[A_DX][c] / (m(TT_M_DXA, i) * rh * cv);
        if (j == 1) {
            const long long c0 = IX(i, 0, k);
            double d = a.G[c] - a.G[c0];
            a.q[A_QY][c0] = nr_flux(a.coef, a.RH[c0], d);
            mn = fmin(mn, d / m(TT_M_DYB, 0));
        }
        double d = a.G[IX(i, j + 1, k)] - a.G[c];
        a.q[A_QY][c] = nr_flux(a.coef, rh, d);
        a.q[A_DY][c] = d;
        mn = fmin(mn, d / m(TT_M_DYB, j));
    }
    block_min_out(mn, a.part + a.nb + blockIdx.x);
}

// pass 3: Y updates (ppfor.f:1124-1127), then Z fluxes (:1154-1197), Q(k=0)
// from the k = 1 cells
__global__ __launch_bounds__(256) void k_av_yz(Geo g, Met m, AArgs a) {
    double mn = 0.0;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
        if (i > a.ub[g.idx2(j, k)] - 2) continue;
        const long long c = IX(i, j, k);
        const double rh = a.RH[c], cv = a.CV[c];
        const double qy = a.q[A_QY][c];
        a.G[c] = a.G[c] - (qy - a.q[A_QY][IX(i, j - 1, k)]) / (m(TT_M_DYB, j) * a.RHQY[c]);
        a.T[c] = a.T[c] - qy * a.q[A_DY][c] / (m(TT_M_DYA, j) * rh * cv);
        if (k == 1) {
            const long long c0 = IX(i, j, 0);
            double d = a.H[c] - a.H[c0];
            a.q[A_QZ][c0] = nr_flux(a.coef, a.RH[c0], d);
            mn = fmin(mn, d / m(TT_M_DZB, 0));
        }
        double d = a.H[IX(i, j, k + 1)] - a.H[c];
        a.q[A_QZ][c] = nr_flux(a.coef, rh, d);
        a.q[A_DZ][c] = d;
        mn = fmin(mn, d / m(TT_M_DZB, k));
    }
    block_min_out(mn, a.part + 2 * a.nb + blockIdx.x);
}

// pass 4: Z updates (ppfor.f:1176-1179)
__global__ __launch_bounds__(256) void k_av_z(Geo g, Met m, AArgs a) {
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
        if (i > a.ub[g.idx2(j, k)] - 2) continue;
        const long long c = IX(i, j, k);
        const double qz = a.q[A_QZ][c];
        a.H[c] = a.H[c] - (qz - a.q[A_QZ][IX(i, j, k - 1)]) / (m(TT_M_DZB, k) * a.RHQZ[c]);
        a.T[c] = a.T[c] - qz * a.q[A_DZ][c] / (m(TT_M_DZA, k) * a.RH[c] * a.CV[c]);
    }
}

}  // namespace

void visc_free() {
    for (auto& p : g_scr) {
        if (p) hipFree(p);
        p = nullptr;
    }
    for (auto& p : g_avq) {
        if (p) hipFree(p);
        p = nullptr;
    }
    if (g_nu) hipFree(g_nu);
    g_nu = nullptr;
    L.av_valid = 0;
}

// ETACALC + VISC on the device state (after PRESSG, as FORGAS3D calls them)
int visc_impl() {
    Geo& g = L.g;
    double** f = L.f;
    const tt_viscosity& v = L.visc;
    if (v.ieta == 4) {
        int rc = sound_impl();  // ETACALC calls SOUND (ppvisc.f:468)
        if (rc) return rc;
    }
    const double c = (double)1.67f * pow(10.0, -8.0);  // 1.67*10.d0**(-8.d0), 1.67 a REAL literal
    double xnue = v.xnue;
    const double* nu_i = nullptr;
    if (v.ieta == 2) {  // ramp of the kinematic viscosity in time (ppvisc.f:402-420)
        const double day = (double)8.64e4f, xs = 9.0;
        const double t0 = (double)282.5f * day, t1 = t0 + 10.0 * day, ht = L.lim.hydrotime;
        if (ht < t0) xnue = pow(10.0, xs);
        else if (ht >= t0 && ht < t1) xnue = pow(10.0, (log10(v.xnue) - xs) / (t1 - t0) * (ht - t0) + xs);
        else xnue = v.xnue;
    } else if (v.ieta == 3) {  // raised at the bottom boundary (ppvisc.f:427-441)
        if (!g_nu || g_nu_xnue != v.xnue) {
            std::vector<double> h((size_t)g.nx + 4);
            const double extra = 4.0, ms = extra / (1.0 - 1.0 * 15);
            for (int i = -1; i <= g.nx + 2; ++i) {
                double nu = v.xnue;
                if (i >= 1 && i <= 15) nu = pow(10.0, ms * (double)(i - 15) + log10(v.xnue));
                if (i <= 0) nu = pow(10.0, extra + log10(v.xnue));
                h[i + 1] = nu;
            }
            if (!g_nu) TT_HIP(hipMalloc(&g_nu, h.size() * sizeof(double)));
            TT_HIP(hipMemcpy(g_nu, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
            g_nu_xnue = v.xnue;
        }
        nu_i = g_nu;
    }
    k_eta<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, v.ieta, xnue, L.p.mu_gas, c, nu_i, f[TT_RH], f[TT_CS], f[TT_ETA]);
    TT_LAUNCH();
    if (v.ivistyp == 0) {  // divtx/y/z = 0 (ppvisc.f:23-27)
        for (int q : {TT_DIVTX, TT_DIVTY, TT_DIVTZ}) TT_HIP(hipMemsetAsync(f[q], 0, (size_t)g.n3 * sizeof(double), L.s));
        return 0;
    }
    if (!g_scr[0]) {
        for (auto& p : g_scr) TT_HIP(hipMalloc(&p, (size_t)g.n3 * sizeof(double)));
    }
    VArgs a{f[TT_V], f[TT_G], f[TT_H], f[TT_ETA], {}, f[TT_DISFN], f[TT_DIVTX], f[TT_DIVTY], f[TT_DIVTZ]};
    for (int q = 0; q < S_N; ++q) a.s[q] = g_scr[q];
    const long long n2 = (long long)(g.nx + 3) * (g.ny + 3) * (g.nz + 3);
    const long long n1 = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    k_visc_a<<<nblk(n2, 256), 256, 0, L.s>>>(g, met(), a);
    TT_LAUNCH();
    k_visc_b<<<nblk(n1, 256), 256, 0, L.s>>>(g, a, v.idiss != 0);
    TT_LAUNCH();
    k_visc_c<<<nblk(n2, 256), 256, 0, L.s>>>(g, a);
    TT_LAUNCH();
    k_visc_d<<<nblk(n1, 256), 256, 0, L.s>>>(g, met(), a);
    TT_LAUNCH();
    return 0;
}

// NR_ARTIFICIAL_VISC on the device state (ppfor.f:1032-1230); FORGAS3D
// calls BOUNDS before and after it
int artvis_impl(double dt) {
    Geo& g = L.g;
    double** f = L.f;
    const size_t bytes = (size_t)g.n3 * sizeof(double);
    if (!g_avq[0]) {
        for (auto& p : g_avq) TT_HIP(hipMalloc(&p, bytes));
        // QviscY/QviscZ start at zero: entries a call does not write are read
        // by the next one where upperbnd varies
        TT_HIP(hipMemsetAsync(g_avq[A_QY], 0, bytes, L.s));
        TT_HIP(hipMemsetAsync(g_avq[A_QZ], 0, bytes, L.s));
    }
    int nb = 1024;
    if (3 * nb > 2 * L.npartial) nb = 2 * L.npartial / 3;
    const double artvisc_coeff = 3.0;  // artvis_var_init, ppmod.f:392
    AArgs a{f[TT_V], f[TT_G], f[TT_H], f[TT_T], f[TT_RH], f[TT_CV], f[TT_RHQX], f[TT_RHQY], f[TT_RHQZ], L.ub, {},
            artvisc_coeff * dt, L.partial, nb};
    for (int q = 0; q < A_N; ++q) a.q[q] = g_avq[q];
    const Met mt = met();
    k_av_x<<<nb, 256, 0, L.s>>>(g, mt, a);
    TT_LAUNCH();
    k_av_xy<<<nb, 256, 0, L.s>>>(g, mt, a);
    TT_LAUNCH();
    k_av_yz<<<nb, 256, 0, L.s>>>(g, mt, a);
    TT_LAUNCH();
    k_av_z<<<nb, 256, 0, L.s>>>(g, mt, a);
    TT_LAUNCH();
    int rc = fold_partials(L.partial, nb, 3, 2, &L.ctrl->av_min[0], &L.ctrl->av_min[1], &L.ctrl->av_min[2]);
    if (rc) return rc;
    L.av_valid = 1;
    return 0;
}

// dt_artvis (ppfor.f:1211-1212) from the minima of the last artvis_impl,
// read from the host mirror of the control block after a sync_check; DTMAX
// before the first call (pp.f:439)
double dt_artvis() {
    if (!L.av_valid) return L.p.dtmax;
    const double* m = L.hctrl->av_min;
    double mn = m[0];
    if (m[1] < mn) mn = m[1];
    if (m[2] < mn) mn = m[2];
    const double tiny = (double)1.0e-30f;  // TINY, a REAL literal (ppmod.f:1092)
    const double d = 1.0 / fabs(8.0 * 3.0 * (mn - tiny));
    return d < L.p.dtmax ? d : L.p.dtmax;
}

}  // namespace tt

using namespace tt;

extern "C" int tt_set_viscosity(const tt_viscosity* v) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!v) return fail(1, "null viscosity options");
    if (v->ivis > 0) {
        if (v->ivistyp != 0 && v->ivistyp != 1) return fail(1, "VISC: IVISTYP must be 0 or 1");
        if (v->ieta < 0 || v->ieta == 1 || v->ieta > 5) return fail(1, "ETACALC: IETA must be 0, 2, 3, 4 or 5");
        if (v->ieta == 3 && L.g.nx < 13) return fail(1, "ETACALC: IETA 3 sets cells 1..15 and needs nx >= 13");
        if (!(v->xnue > 0.0)) return fail(1, "ETACALC: XNUE must be positive");
    }
    if (v->iartvis > 1) return fail(1, "unknown IARTVIS: only 1 (NR_ARTIFICIAL_VISC) is defined");
    if (v->iartvis > 0 && L.visc.iartvis == 0) L.av_valid = 0;  // dt_artvis = DTMAX (pp.f:439)
    L.visc = *v;
    return 0;
}

extern "C" int tt_get_dt_artvis(double* dt) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!dt) return fail(1, "null dt");
    int rc = sync_check();
    if (rc) return rc;
    *dt = dt_artvis();
    return 0;
}

extern "C" int tt_get_viscosity(tt_viscosity* v) {
    if (!v) return fail(1, "null viscosity options");
    *v = L.visc;
    return 0;
}
