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

#include "tt_common.h"

namespace tt {

#define IX(i, j, k) g.idx(i, j, k)

namespace {

// scratch: DIVV, TXX, TYY, TZZ, TXY, TXZ, TYZ, TXYAV, TYZAV, TXZAV
enum { S_DIVV, S_TXX, S_TYY, S_TZZ, S_TXY, S_TXZ, S_TYZ, S_AXY, S_AYZ, S_AXZ, S_N };
double* g_scr[S_N] = {};

struct VArgs {
    const double *V, *G, *H, *ETA;
    double* s[S_N];
    double *DISFN, *DTX, *DTY, *DTZ;
};

// ETACALC: IETA 0 and 5, eta = RH XNUE; IETA 4, eta = RH min(CS mu_gas c /
// RH, XNUE) with c = 1.67 10^-8 (the host forms it as the reference does)
__global__ void k_eta(long long n, int ieta, double xnue, double mu, double c, const double* __restrict__ RH,
                      const double* __restrict__ CS, double* ETA) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (ieta == 4) {
        double e = CS[t] * mu * c / RH[t];
        e = e < xnue ? e : xnue;
        ETA[t] = e * RH[t];
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

}  // namespace

void visc_free() {
    for (auto& p : g_scr) {
        if (p) hipFree(p);
        p = nullptr;
    }
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
    k_eta<<<nblk(g.n3, 256), 256, 0, L.s>>>(g.n3, v.ieta, v.xnue, L.p.mu_gas, c, f[TT_RH], f[TT_CS], f[TT_ETA]);
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

}  // namespace tt

using namespace tt;

extern "C" int tt_set_viscosity(const tt_viscosity* v) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    if (!v) return fail(1, "null viscosity options");
    if (v->ivis > 0) {
        if (v->ivistyp != 0 && v->ivistyp != 1) return fail(1, "VISC: IVISTYP must be 0 or 1");
        if (v->ieta != 0 && v->ieta != 4 && v->ieta != 5) return fail(1, "ETACALC: IETA must be 0, 4 or 5");
        if (!(v->xnue > 0.0)) return fail(1, "ETACALC: XNUE must be positive");
    }
    L.visc = *v;
    return 0;
}

extern "C" int tt_get_viscosity(tt_viscosity* v) {
    if (!v) return fail(1, "null viscosity options");
    *v = L.visc;
    return 0;
}
