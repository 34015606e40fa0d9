// tt_hydro.hip -- ADVECT3D (ppadv.f:525-1577) and FORGAS3D (ppfor.f:142-397)
// on gfx950, plus the step driver.
//
// The reference advects with 1-D line sweeps that fill per-line scratch
// (S, IND, GRAD) and then the face fluxes.  Every face flux depends only on
// the cells within two of it along its line, so here the fluxes are formed
// per face directly from the state, each face once: the mass fluxes by
// k_adv_mass (stored, the momentum sweeps reuse them through face
// averages), the transported quantities by k_adv_tiled, which shares each
// face between its two cells through LDS.  The arithmetic is the
// reference's, face by face.
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

// The upwind cell of a face depends on the sign of the face velocity, and
// the van Leer slope reads the cells on either side of it.  Loading only
// after the sign is known chains two memory round trips; each face function
// below loads the whole window (the four values straddling the face and the
// spacings and coordinates of both candidate cells) at once and then
// selects, with the reference's arithmetic on the selected values.
// up != 0 selects the low-side candidate.
__device__ inline double vl_face(bool up, double q_2, double q_1, double q0, double q1, double w_1, double w0,
                                 double w1, int order, double* qup) {
    // candidates: low side (q_2, q_1, q0; spacings w_1, w0), high side
    // (q_1, q0, q1; spacings w0, w1)
    const double qm = up ? q_2 : q_1, qc = up ? q_1 : q0, qp = up ? q0 : q1;
    const double wlo = up ? w_1 : w0, whi = up ? w0 : w1;
    *qup = qc;
    return vl_slope(qm, qc, qp, wlo, whi, order);
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

__global__ void k_hd_init(Geo g, Met m, const double* __restrict__ H, double* HD) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    if (k < 0) return;
    double gh = m(TT_M_GEOXH, i);
    HD[t] = H[t] * (gh * gh);
}

// ---- face-shared sweeps.  The four transport updates above evaluate, per
// cell, the fluxes through all six of its faces, so every face flux is
// formed twice.  k_adv_tiled forms each face flux once: a block takes a tile
// of 32 (i) x 8 (j) cells and marches through 2 planes of k (longer marches
// share more z faces but leave too few blocks to hide the loads' latency:
// 2 planes measured best at 128x128x256).  Per
// plane each cell lane forms the x face on its low side and the y face on its
// low side into LDS (double-buffered, one barrier per plane) and the z face
// on its high side in a register, which is the low z face of the next plane;
// a fifth wavefront forms the tile's extra high x faces (8) and high y faces
// (32).  The flux expressions are the reference's, face by face, so the
// results equal the per-cell form bit for bit.
struct AdvArgs {
    const double *A;                 // advected quantity: rho c_v T, V, GD or HD
    const double *V, *G, *H;         // velocities
    const double *FX, *FY, *FZ;      // mass fluxes FLMX, FLMY, FLMZ
    const double *RHQ;               // face density of the momentum (unused for energy)
    double* out;
    double dt;
    double *cf1, *cf2, *cf3;         // energy only: conv_flux, the low-face fluxes (ppadv.f:857-865)
};

constexpr int ADV_TI = 32, ADV_TJ = 8, ADV_KC = 2;

// KIND 0: thermal energy (ppadv.f:741-891), cell faces I = i, i+1 etc.; the
// upwind cell is I-1 for a positive velocity, else I
struct AdvE {
    static constexpr int ox = 0, oy = 0, oz = 0;
    __device__ static int ilo(const Geo&) { return 1; }
    __device__ static int ihi(const Geo& g) { return g.nx; }
    __device__ static int jlo(const Geo&) { return 1; }
    __device__ static int jhi(const Geo& g) { return g.ny; }
    __device__ static int klo(const Geo&) { return 1; }
    __device__ static int khi(const Geo& g) { return g.nz; }
    __device__ static double fx(const Geo& g, const Met& m, const AdvArgs& a, int I, int j, int k) {
        const double half = 0.5;
        Box3 E{a.A, g};
        const double v = a.V[IX(I, j, k)];
        const double e_2 = E(I - 2, j, k), e_1 = E(I - 1, j, k), e0 = E(I, j, k), e1 = E(I + 1, j, k);
        const double d_1 = m(TT_M_DXB, I - 1), d0 = m(TT_M_DXB, I), d1 = m(TT_M_DXB, I + 1);
        const double x_1 = m(TT_M_XXB, I - 1), x0 = m(TT_M_XXB, I);
        const double sa = m(TT_M_SURXA, I), ca = m(TT_M_CCXA, I), xa = m(TT_M_XXA, I);
        double S = v * a.dt;
        const bool up = v > 0.0;
        double eu;
        double gr = vl_face(up, e_2, e_1, e0, e1, d_1, d0, d1, 0, &eu);
        return (sa - ca * S) * S * (eu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static double fy(const Geo& g, const Met& m, const AdvArgs& a, int i, int J, int k) {
        const double half = 0.5;
        Box3 E{a.A, g};
        const double v = a.G[IX(i, J, k)];
        const double e_2 = E(i, J - 2, k), e_1 = E(i, J - 1, k), e0 = E(i, J, k), e1 = E(i, J + 1, k);
        const double d_1 = m(TT_M_DYB, J - 1), d0 = m(TT_M_DYB, J), d1 = m(TT_M_DYB, J + 1);
        const double x_1 = m(TT_M_XYB, J - 1), x0 = m(TT_M_XYB, J);
        const double sa = m(TT_M_SURYA, J), ca = m(TT_M_CCYA, J), xa = m(TT_M_XYA, J);
        double S = v * a.dt;
        const bool up = v > 0.0;
        double eu;
        double gr = vl_face(up, e_2, e_1, e0, e1, d_1, d0, d1, 0, &eu);
        return (sa - ca * S) * S * (eu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static double fz(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int K) {
        const double half = 0.5;
        Box3 E{a.A, g};
        const double v = a.H[IX(i, j, K)];
        const double e_2 = E(i, j, K - 2), e_1 = E(i, j, K - 1), e0 = E(i, j, K), e1 = E(i, j, K + 1);
        const double d_1 = m(TT_M_DZB, K - 1), d0 = m(TT_M_DZB, K), d1 = m(TT_M_DZB, K + 1);
        const double x_1 = m(TT_M_XZB, K - 1), x0 = m(TT_M_XZB, K);
        const double sa = m(TT_M_SURZA, K), ca = m(TT_M_CCZA, K), xa = m(TT_M_XZA, K);
        double S = v * a.dt;
        const bool up = v > 0.0;
        double eu;
        double gr = vl_face(up, e_2, e_1, e0, e1, d_1, d0, d1, 0, &eu);
        return (sa - ca * S) * S * (eu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static void finish(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int k, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        a.out[IX(i, j, k)] = a.A[IX(i, j, k)] - (fx1 - fx0) / m(TT_M_VOLXB, i) - (fy1 - fy0) / m(TT_M_VOLYB, j) -
                             (fz1 - fz0) / m(TT_M_VOLZB, k);
        a.cf1[IX(i, j, k)] = fx0;
        a.cf2[IX(i, j, k)] = fy0;
        a.cf3[IX(i, j, k)] = fz0;
    }
};

// KIND 1: radial momentum (ppadv.f:896-1073), cells i = 2..nx, x faces
// FLX(I) at I = i-1, i (upwind cell I for a positive S, else I+1)
struct AdvV {
    static constexpr int ox = -1, oy = 0, oz = 0;
    __device__ static int ilo(const Geo&) { return 2; }
    __device__ static int ihi(const Geo& g) { return g.nx; }
    __device__ static int jlo(const Geo&) { return 1; }
    __device__ static int jhi(const Geo& g) { return g.ny; }
    __device__ static int klo(const Geo&) { return 1; }
    __device__ static int khi(const Geo& g) { return g.nz; }
    __device__ static double fx(const Geo& g, const Met& m, const AdvArgs& a, int I, int j, int k) {
        const double half = 0.5;
        Box3 Vb{a.V, g};
        const double v_1 = Vb(I - 1, j, k), v0 = Vb(I, j, k), v1 = Vb(I + 1, j, k), v2 = Vb(I + 2, j, k);
        const double f0 = a.FX[IX(I, j, k)], f1 = a.FX[IX(I + 1, j, k)];
        const double d0 = m(TT_M_DXA, I), d1 = m(TT_M_DXA, I + 1), d2 = m(TT_M_DXA, I + 2);
        const double x0 = m(TT_M_XXA, I), x1 = m(TT_M_XXA, I + 1), xb = m(TT_M_XXB, I);
        double S = (v0 + v1) * a.dt / 2.0;
        const bool up = S > 0.0;
        double vu;
        double gr = vl_face(up, v_1, v0, v1, v2, d0, d1, d2, 1, &vu);
        return half * (f0 + f1) * (vu + (xb - (up ? x0 : x1) - half * S) * gr);
    }
    __device__ static double fy(const Geo& g, const Met& m, const AdvArgs& a, int i, int J, int k) {
        const double half = 0.5;
        Box3 Vb{a.V, g};
        const double p1 = m(TT_M_DDX1, i), p0 = m(TT_M_DDX0, i);
        const double g1 = a.G[IX(i, J, k)], g0 = a.G[IX(i - 1, J, k)];
        const double f1 = a.FY[IX(i, J, k)], f0 = a.FY[IX(i - 1, J, k)];
        const double v_2 = Vb(i, J - 2, k), v_1 = Vb(i, J - 1, k), v0 = Vb(i, J, k), v1 = Vb(i, J + 1, k);
        const double d_1 = m(TT_M_DYB, J - 1), d0 = m(TT_M_DYB, J), d1 = m(TT_M_DYB, J + 1);
        const double x_1 = m(TT_M_XYB, J - 1), x0 = m(TT_M_XYB, J), xa = m(TT_M_XYA, J);
        double S = (p1 * g1 + p0 * g0) * a.dt;
        const bool up = S > 0.0;
        double vu;
        double gr = vl_face(up, v_2, v_1, v0, v1, d_1, d0, d1, 1, &vu);
        return (p1 * f1 + p0 * f0) * (vu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static double fz(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int K) {
        const double half = 0.5;
        Box3 Vb{a.V, g};
        const double p1 = m(TT_M_DDX1, i), p0 = m(TT_M_DDX0, i);
        const double h1 = a.H[IX(i, j, K)], h0 = a.H[IX(i - 1, j, K)];
        const double f1 = a.FZ[IX(i, j, K)], f0 = a.FZ[IX(i - 1, j, K)];
        const double v_2 = Vb(i, j, K - 2), v_1 = Vb(i, j, K - 1), v0 = Vb(i, j, K), v1 = Vb(i, j, K + 1);
        const double d_1 = m(TT_M_DZB, K - 1), d0 = m(TT_M_DZB, K), d1 = m(TT_M_DZB, K + 1);
        const double x_1 = m(TT_M_XZB, K - 1), x0 = m(TT_M_XZB, K), xa = m(TT_M_XZA, K);
        double S = (p1 * h1 + p0 * h0) * a.dt;
        const bool up = S > 0.0;
        double vu;
        double gr = vl_face(up, v_2, v_1, v0, v1, d_1, d0, d1, 1, &vu);
        return (p1 * f1 + p0 * f0) * (vu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static void finish(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int k, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        a.out[IX(i, j, k)] = a.V[IX(i, j, k)] * a.RHQ[IX(i, j, k)] - (fx1 - fx0) / m(TT_M_VOLXA, i) -
                             (fy1 - fy0) / m(TT_M_VOLYB, j) - (fz1 - fz0) / m(TT_M_VOLZB, k);
    }
};

// KIND 2: azimuthal angular momentum (ppadv.f:1096-1266), y faces FLY(J) at
// J = j-1, j (upwind J for a positive S, else J+1); FLY(0) is the periodic
// copy of FLY(ny)
struct AdvG {
    static constexpr int ox = 0, oy = -1, oz = 0;
    __device__ static int ilo(const Geo&) { return 1; }
    __device__ static int ihi(const Geo& g) { return g.nx; }
    __device__ static int jlo(const Geo&) { return 1; }
    __device__ static int jhi(const Geo& g) { return g.ny; }
    __device__ static int klo(const Geo&) { return 1; }
    __device__ static int khi(const Geo& g) { return g.nz; }
    __device__ static double fx(const Geo& g, const Met& m, const AdvArgs& a, int I, int j, int k) {
        const double half = 0.5;
        Box3 D{a.A, g};
        const double p1 = m(TT_M_DDY1, j), p0 = m(TT_M_DDY0, j);
        const double u1 = a.V[IX(I, j, k)], u0 = a.V[IX(I, j - 1, k)];
        const double f1 = a.FX[IX(I, j, k)], f0 = a.FX[IX(I, j - 1, k)];
        const double q_2 = D(I - 2, j, k), q_1 = D(I - 1, j, k), q0 = D(I, j, k), q1 = D(I + 1, j, k);
        const double d_1 = m(TT_M_DXB, I - 1), d0 = m(TT_M_DXB, I), d1 = m(TT_M_DXB, I + 1);
        const double x_1 = m(TT_M_XXB, I - 1), x0 = m(TT_M_XXB, I), xa = m(TT_M_XXA, I);
        double S = (p1 * u1 + p0 * u0) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_2, q_1, q0, q1, d_1, d0, d1, 1, &qu);
        return (p1 * f1 + p0 * f0) * (qu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static double fy(const Geo& g, const Met& m, const AdvArgs& a, int i, int J, int k) {
        const double half = 0.5;
        Box3 D{a.A, g};
        if (J == 0) J = g.ny;
        const double c0 = m(TT_M_DCY0, J + 1), c1 = m(TT_M_DCY1, J + 1);
        const double g0 = a.G[IX(i, J, k)], g1 = a.G[IX(i, J + 1, k)];
        const double f0 = a.FY[IX(i, J, k)], f1 = a.FY[IX(i, J + 1, k)];
        const double q_1 = D(i, J - 1, k), q0 = D(i, J, k), q1 = D(i, J + 1, k), q2 = D(i, J + 2, k);
        const double d0 = m(TT_M_DYA, J), d1 = m(TT_M_DYA, J + 1), d2 = m(TT_M_DYA, J + 2);
        const double y0 = m(TT_M_XYA, J), y1 = m(TT_M_XYA, J + 1), yb = m(TT_M_XYB, J);
        double S = (c0 * g0 + c1 * g1) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_1, q0, q1, q2, d0, d1, d2, 1, &qu);
        return (c0 * f0 + c1 * f1) * (qu + (yb - (up ? y0 : y1) - half * S) * gr);
    }
    __device__ static double fz(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int K) {
        const double half = 0.5;
        Box3 D{a.A, g};
        const double p1 = m(TT_M_DDY1, j), p0 = m(TT_M_DDY0, j);
        const double h1 = a.H[IX(i, j, K)], h0 = a.H[IX(i, j - 1, K)];
        const double f1 = a.FZ[IX(i, j, K)], f0 = a.FZ[IX(i, j - 1, K)];
        const double q_2 = D(i, j, K - 2), q_1 = D(i, j, K - 1), q0 = D(i, j, K), q1 = D(i, j, K + 1);
        const double d_1 = m(TT_M_DZB, K - 1), d0 = m(TT_M_DZB, K), d1 = m(TT_M_DZB, K + 1);
        const double x_1 = m(TT_M_XZB, K - 1), x0 = m(TT_M_XZB, K), xa = m(TT_M_XZA, K);
        double S = (p1 * h1 + p0 * h0) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_2, q_1, q0, q1, d_1, d0, d1, 1, &qu);
        return (p1 * f1 + p0 * f0) * (qu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static void finish(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int k, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        a.out[IX(i, j, k)] = a.A[IX(i, j, k)] * a.RHQ[IX(i, j, k)] - (fx1 - fx0) / m(TT_M_VOLXB, i) -
                             (fy1 - fy0) / m(TT_M_VOLYA, j) - (fz1 - fz0) / m(TT_M_VOLZB, k);
    }
};

// KIND 3: meridional angular momentum (ppadv.f:1278-1461), cells k = 2..nz,
// z faces FLZ(K) at K = k-1, k (upwind K for a positive S, else K+1)
struct AdvH {
    static constexpr int ox = 0, oy = 0, oz = -1;
    __device__ static int ilo(const Geo&) { return 1; }
    __device__ static int ihi(const Geo& g) { return g.nx; }
    __device__ static int jlo(const Geo&) { return 1; }
    __device__ static int jhi(const Geo& g) { return g.ny; }
    __device__ static int klo(const Geo&) { return 2; }
    __device__ static int khi(const Geo& g) { return g.nz; }
    __device__ static double fx(const Geo& g, const Met& m, const AdvArgs& a, int I, int j, int k) {
        const double half = 0.5;
        Box3 D{a.A, g};
        const double p1 = m(TT_M_DDZ1, k), p0 = m(TT_M_DDZ0, k);
        const double u1 = a.V[IX(I, j, k)], u0 = a.V[IX(I, j, k - 1)];
        const double f1 = a.FX[IX(I, j, k)], f0 = a.FX[IX(I, j, k - 1)];
        const double q_2 = D(I - 2, j, k), q_1 = D(I - 1, j, k), q0 = D(I, j, k), q1 = D(I + 1, j, k);
        const double d_1 = m(TT_M_DXB, I - 1), d0 = m(TT_M_DXB, I), d1 = m(TT_M_DXB, I + 1);
        const double x_1 = m(TT_M_XXB, I - 1), x0 = m(TT_M_XXB, I), xa = m(TT_M_XXA, I);
        double S = (p1 * u1 + p0 * u0) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_2, q_1, q0, q1, d_1, d0, d1, 1, &qu);
        return (p1 * f1 + p0 * f0) * (qu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static double fy(const Geo& g, const Met& m, const AdvArgs& a, int i, int J, int k) {
        const double half = 0.5;
        Box3 D{a.A, g};
        const double p1 = m(TT_M_DDZ1, k), p0 = m(TT_M_DDZ0, k);
        const double g1 = a.G[IX(i, J, k)], g0 = a.G[IX(i, J, k - 1)];
        const double f1 = a.FY[IX(i, J, k)], f0 = a.FY[IX(i, J, k - 1)];
        const double q_2 = D(i, J - 2, k), q_1 = D(i, J - 1, k), q0 = D(i, J, k), q1 = D(i, J + 1, k);
        const double d_1 = m(TT_M_DYB, J - 1), d0 = m(TT_M_DYB, J), d1 = m(TT_M_DYB, J + 1);
        const double x_1 = m(TT_M_XYB, J - 1), x0 = m(TT_M_XYB, J), xa = m(TT_M_XYA, J);
        double S = (p1 * g1 + p0 * g0) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_2, q_1, q0, q1, d_1, d0, d1, 1, &qu);
        return (p1 * f1 + p0 * f0) * (qu + (xa - (up ? x_1 : x0) - half * S) * gr);
    }
    __device__ static double fz(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int K) {
        const double half = 0.5;
        Box3 D{a.A, g};
        const double h0 = a.H[IX(i, j, K)], h1 = a.H[IX(i, j, K + 1)];
        const double f0 = a.FZ[IX(i, j, K)], f1 = a.FZ[IX(i, j, K + 1)];
        const double q_1 = D(i, j, K - 1), q0 = D(i, j, K), q1 = D(i, j, K + 1), q2 = D(i, j, K + 2);
        const double d0 = m(TT_M_DZA, K), d1 = m(TT_M_DZA, K + 1), d2 = m(TT_M_DZA, K + 2);
        const double z0 = m(TT_M_XZA, K), z1 = m(TT_M_XZA, K + 1), zb = m(TT_M_XZB, K);
        double S = half * (h0 + h1) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_1, q0, q1, q2, d0, d1, d2, 1, &qu);
        return half * (f0 + f1) * (qu + (zb - (up ? z0 : z1) - half * S) * gr);
    }
    __device__ static void finish(const Geo& g, const Met& m, const AdvArgs& a, int i, int j, int k, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        a.out[IX(i, j, k)] = a.A[IX(i, j, k)] * a.RHQ[IX(i, j, k)] - (fx1 - fx0) / m(TT_M_VOLXB, i) -
                             (fy1 - fy0) / m(TT_M_VOLYB, j) - (fz1 - fz0) / m(TT_M_VOLZA, k);
    }
};

template <class K>
__global__ __launch_bounds__(320) void k_adv_tiled(Geo g, Met m, AdvArgs a) {
    __shared__ double sx[2][ADV_TJ][ADV_TI + 1];  // x faces of the tile rows, plus the high one
    __shared__ double sy[2][ADV_TJ + 1][ADV_TI];  // y faces of the tile columns, plus the high row
    const int ilo = K::ilo(g), ihi = K::ihi(g), jlo = K::jlo(g), jhi = K::jhi(g);
    const int klo = K::klo(g), khi = K::khi(g);
    const int nti = (ihi - ilo + ADV_TI) / ADV_TI, ntj = (jhi - jlo + ADV_TJ) / ADV_TJ;
    int b = blockIdx.x;
    const int ti = b % nti;
    b /= nti;
    const int tj = b % ntj, kc = b / ntj;
    const int ci0 = ilo + ti * ADV_TI, cj0 = jlo + tj * ADV_TJ;
    const int k0 = klo + kc * ADV_KC;
    const int k1 = min(khi, k0 + ADV_KC - 1);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const bool helper = w == 4;
    const int r = 2 * w + (l >> 5), c = l & 31;
    const int i = ci0 + c, j = cj0 + r;
    const bool vi = i <= ihi, vj = j <= jhi;
    const bool cell = !helper && vi && vj;
    // the highest face of each family
    const int Ihi = ihi + K::ox + 1, Jhi = jhi + K::oy + 1;
    double zlo = 0.0;
    int p = 0;
    for (int k = k0; k <= k1; ++k, p ^= 1) {
        double zhi = 0.0;
        if (!helper) {
            // the low x face of the lane's cell is also the high face of the
            // cell before it, and of the last valid cell when this one is
            // past the end of the row
            if (vj && i + K::ox <= Ihi) sx[p][r][c] = K::fx(g, m, a, i + K::ox, j, k);
            if (vi && j + K::oy <= Jhi) sy[p][r][c] = K::fy(g, m, a, i, j + K::oy, k);
            if (cell) {
                zhi = K::fz(g, m, a, i, j, k + K::oz + 1);
                if (k == k0) zlo = K::fz(g, m, a, i, j, k + K::oz);
            }
        } else if (l < ADV_TJ) {
            const int jr = cj0 + l, I = ci0 + ADV_TI + K::ox;
            if (jr <= jhi && I <= Ihi) sx[p][l][ADV_TI] = K::fx(g, m, a, I, jr, k);
        } else if (l < ADV_TJ + ADV_TI) {
            const int ic = ci0 + (l - ADV_TJ), J = cj0 + ADV_TJ + K::oy;
            if (ic <= ihi && J <= Jhi) sy[p][ADV_TJ][l - ADV_TJ] = K::fy(g, m, a, ic, J, k);
        }
        __syncthreads();
        if (cell) {
            K::finish(g, m, a, i, j, k, sx[p][r][c], sx[p][r][c + 1], sy[p][r][c], sy[p][r + 1][c], zlo, zhi);
            zlo = zhi;
        }
    }
}

template <class K>
static int launch_adv_tiled(const AdvArgs& a) {
    Geo& g = L.g;
    // host copies of the cell ranges (the device functions above)
    const int ni = K::ox == -1 ? g.nx - 1 : g.nx, nj = g.ny, nk = K::oz == -1 ? g.nz - 1 : g.nz;
    const long long nb = (long long)((ni + ADV_TI - 1) / ADV_TI) * ((nj + ADV_TJ - 1) / ADV_TJ) *
                         ((nk + ADV_KC - 1) / ADV_KC);
    k_adv_tiled<K><<<(unsigned)nb, 320, 0, L.s>>>(g, met(), a);
    TT_LAUNCH();
    return 0;
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
    AdvArgs aa{L.TOLD, f[TT_V], f[TT_G], f[TT_H], L.FLMX, L.FLMY, L.FLMZ, nullptr, L.ED, dt,
               f[TT_CONV_FLUX1], f[TT_CONV_FLUX2], f[TT_CONV_FLUX3]};
    if ((rc = launch_adv_tiled<AdvE>(aa))) return rc;
    aa.A = f[TT_V]; aa.RHQ = f[TT_RHQX]; aa.out = L.VD;
    if ((rc = launch_adv_tiled<AdvV>(aa))) return rc;
    k_gd_init<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, met(), L.p.omrot, f[TT_G], f[TT_GD]);
    TT_LAUNCH();
    aa.A = f[TT_GD]; aa.RHQ = f[TT_RHQY]; aa.out = L.GDN;
    if ((rc = launch_adv_tiled<AdvG>(aa))) return rc;
    k_hd_init<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, met(), f[TT_H], L.HH);
    TT_LAUNCH();
    aa.A = L.HH; aa.RHQ = f[TT_RHQZ]; aa.out = L.HDN;
    if ((rc = launch_adv_tiled<AdvH>(aa))) return rc;
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

// p dV work: T -= fac*dt PG/RH/CV div(v) over cells lo..hi in each axis;
// with DISFN (dissipative heating, IDISS) then T += fac*dt DISFN/CV/RH on the
// interior, as the reference's following loop does cell by cell
// (ppfor.f:202-210,335-343)
__global__ void k_pdv(Geo g, Met m, double dtf, int lo, int hiadd, double* T, const double* __restrict__ PG,
                      const double* __restrict__ RH, const double* __restrict__ CV, const double* __restrict__ V,
                      const double* __restrict__ G, const double* __restrict__ H, const double* __restrict__ DISFN) {
    int ni = g.nx + hiadd - lo + 1, nj = g.ny + hiadd - lo + 1, nk = g.nz + hiadd - lo + 1;
    long long n = (long long)ni * nj * nk;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, ni, nj, lo, lo, lo, i, j, k);
    long long c = IX(i, j, k);
    double tn = T[c] - dtf * PG[c] / RH[c] / CV[c] * divv(g, m, V, G, H, i, j, k);
    if (DISFN && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz)
        tn = tn + dtf * DISFN[c] / CV[c] / RH[c];
    T[c] = tn;
}

// VH, GH, HH (ppfor.f:217-278); DX/DY/DZ: divtx/y/z, null without viscosity
// (the reference's zero arrays)
__global__ void k_force_vh(Geo g, Met m, double dt, const double* __restrict__ V, const double* __restrict__ G,
                           const double* __restrict__ H, const double* __restrict__ PG, const double* __restrict__ QX,
                           const double* __restrict__ QY, const double* __restrict__ QZ, const double* __restrict__ CCX,
                           const double* __restrict__ CCY, const double* __restrict__ CCZ,
                           const double* __restrict__ GX, const double* __restrict__ GY, const double* __restrict__ GZ,
                           const double* __restrict__ DX, const double* __restrict__ DY, const double* __restrict__ DZ,
                           double* VH, double* GH, double* HH) {
    long long n = (long long)g.nx * g.ny * g.nz;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long c = IX(i, j, k);
    double gx = m(TT_M_GEOXG, i), gz = m(TT_M_GEOZG, k), gh = m(TT_M_GEOXH, i);
    const double dx = DX ? DX[c] : 0.0, dy = DY ? DY[c] : 0.0, dz = DZ ? DZ[c] : 0.0;
    VH[c] = V[c] + dt * ((-(PG[c] - PG[IX(i - 1, j, k)]) / m(TT_M_DXB, i) + dx) / QX[c] + CCX[c] + GX[c]);
    GH[c] = G[c] + dt * (((-(PG[c] - PG[IX(i, j - 1, k)]) / m(TT_M_DYB, j) / gx / gz + dy) / QY[c] + CCY[c] + GY[c]) / gx / gz);
    HH[c] = H[c] + dt * (((-(PG[c] - PG[IX(i, j, k - 1)]) / m(TT_M_DZB, k) / gh + dz) / QZ[c] + CCZ[c] + GZ[c]) / gh);
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
    const bool vis = L.visc.ivis > 0;
    if (vis && (rc = visc_impl())) return rc;  // ETACALC + VISC, ppfor.f:170-174
    const double* disfn = vis && L.visc.idiss != 0 ? f[TT_DISFN] : nullptr;
    long long nc = (long long)g.nx * g.ny * g.nz;
    k_centcor<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), L.p.omrot, f[TT_G], f[TT_H], f[TT_CENTCORX], f[TT_CENTCORY],
                                              f[TT_CENTCORZ]);
    TT_LAUNCH();
    TT_HIP(hipMemcpyAsync(L.TOLD, f[TT_T], (size_t)g.n3 * sizeof(double), hipMemcpyDeviceToDevice, L.s));
    if (L.p.isoth == 0) {
        long long nh = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
        k_pdv<<<nblk(nh, 256), 256, 0, L.s>>>(g, met(), dt / 2.0, 0, 1, f[TT_T], f[TT_PG], f[TT_RH], f[TT_CV], f[TT_V],
                                              f[TT_G], f[TT_H], disfn);
        TT_LAUNCH();
        if ((rc = bounds_impl())) return rc;
        if ((rc = pressg_impl())) return rc;
    }
    k_force_vh<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, f[TT_V], f[TT_G], f[TT_H], f[TT_PG], f[TT_RHQX], f[TT_RHQY],
                                               f[TT_RHQZ], f[TT_CENTCORX], f[TT_CENTCORY], f[TT_CENTCORZ], f[TT_GRAVX],
                                               f[TT_GRAVY], f[TT_GRAVZ], vis ? f[TT_DIVTX] : nullptr,
                                               vis ? f[TT_DIVTY] : nullptr, vis ? f[TT_DIVTZ] : nullptr, L.VH, L.GH,
                                               L.HH);
    TT_LAUNCH();
    TT_HIP(hipMemcpyAsync(f[TT_T], L.TOLD, (size_t)g.n3 * sizeof(double), hipMemcpyDeviceToDevice, L.s));
    if (L.p.isoth == 0) {
        k_force_set<<<nblk(nc, 256), 256, 0, L.s>>>(g, 0, f[TT_V], f[TT_G], f[TT_H], L.VH, L.GH, L.HH);
        TT_LAUNCH();
        if ((rc = bounds_impl())) return rc;
        k_pdv<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, 1, 0, f[TT_T], f[TT_PG], f[TT_RH], f[TT_CV], f[TT_V], f[TT_G],
                                              f[TT_H], disfn);
        TT_LAUNCH();
    }
    k_force_set<<<nblk(nc, 256), 256, 0, L.s>>>(g, 1, f[TT_V], f[TT_G], f[TT_H], L.VH, L.GH, L.HH);
    TT_LAUNCH();
    // the velocity ceiling (ppfor.f:361-371: SPHERICALVELLIMIT for NCOSYS=2)
    if (L.lim.ivel != 0 && (rc = sphericalvellimit_impl(L.lim.ivel, L.lim.velmax, L.lim.hydrotime))) return rc;
    if (L.visc.iartvis > 0) {  // ppfor.f:374-395
        if ((rc = bounds_impl())) return rc;
        if ((rc = artvis_impl(dt))) return rc;
    }
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

int tt_set_step(int iii) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    L.iii = iii;
    return 0;
}

int tt_step(double delt, int iii, tt_step_stats* st) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    tt_step_stats s{};
    hipEvent_t* e = L.step_ev;  // created once by tt_init, freed by tt_finalize
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
    TT_HIP(hipEventElapsedTime(&a, e[0], e[1]));
    TT_HIP(hipEventElapsedTime(&b, e[1], e[2]));
    TT_HIP(hipEventElapsedTime(&c, e[2], e[3]));
    s.ms_advect = a;
    s.ms_force = b;
    s.ms_energy = c;
    s.ms_total = (double)a + b + c;
    if (st) *st = s;
    return erc;
}

}  // extern "C"
