// tt_hydro.hip -- ADVECT3D (ppadv.f:525-1577) and FORGAS3D (ppfor.f:142-397)
// on gfx950, plus the step driver.
//
// The reference advects with 1-D line sweeps that fill per-line scratch
// (S, IND, GRAD) and then the face fluxes.  Every face flux depends only on
// the cells within two of it along its line, so here the fluxes are formed
// per face directly from the state, each face once: the mass fluxes by
// k_adv_mass (stored, the momentum sweeps reuse them through face
// averages), the transported quantities by k_adv_t, which shares each
// face between its two cells through LDS.  The arithmetic is the
// reference's, face by face.
#include <cmath>
#include <utility>

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

// ---- mass fluxes (ppadv.f:559-644) and their phi plane exchange (:654-711),
// one thread per cell of the box.  The same pass forms the transported
// quantities the later sweeps read over the whole box: the thermal energy
// E0 = T*RH*CV (ppadv.f:741-891), GD = (G + OMROT) r^2 cos^2 (:1078-1084;
// the phi ghosts equal the formula at the wrapped column because G's ghosts
// were wrapped by BOUNDS) and HD = H r^2 (:1278-1283), as the reference's
// expressions.
struct MassArgs {
    const double *RH, *V, *G, *H, *T, *CV;
    double *FLMX, *FLMY, *FLMZ, *E0, *GD, *HD;
    double dt, om;
};
__global__ void k_adv_mass(Geo g, Met m, MassArgs a) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    {
        const double r = a.RH[t];
        st_nt(&a.E0[t], a.T[t] * r * a.CV[t]);
        if (j >= 0) {
            const double gx = m(TT_M_GEOXG, i), gz = m(TT_M_GEOZG, k);
            st_nt(&a.GD[t], (a.G[t] + a.om) * (gx * gx) * (gz * gz));
        }
        if (k >= 0) {
            const double gh = m(TT_M_GEOXH, i);
            st_nt(&a.HD[t], a.H[t] * (gh * gh));
        }
    }
    if (i < 1 || i > g.nx + 1 || j < 0 || j > g.ny + 1 || k < 1 || k > g.nz + 1) return;
    const double dt = a.dt;
    const double* __restrict__ RH = a.RH;
    const double* __restrict__ V = a.V;
    const double* __restrict__ G = a.G;
    const double* __restrict__ H = a.H;
    double* FLMX = a.FLMX;
    double* FLMY = a.FLMY;
    double* FLMZ = a.FLMZ;
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
            st_nt(&FLMX[IX(i, j, k)], (m(TT_M_SURXA, i) - m(TT_M_CCXA, i) * S) * S *
                                          (R(ind, jj, k) + (m(TT_M_XXA, i) - m(TT_M_XXB, ind) - half * S) * gr));
        }
    }
    // y faces: i = 1..nx, j = 1..ny+1, k = 1..nz
    if (i <= g.nx && j >= 1 && k <= g.nz) {
        double gv = G[IX(i, j, k)];
        double S = gv * dt;
        int ind = gv > 0.0 ? j - 1 : j;
        double gr = vl_slope(R(i, ind - 1, k), R(i, ind, k), R(i, ind + 1, k), m(TT_M_DYB, ind), m(TT_M_DYB, ind + 1), 0);
        st_nt(&FLMY[IX(i, j, k)], S * (m(TT_M_SURYA, j) + m(TT_M_CCYA, j) * S) *
                                      (R(i, ind, k) + (m(TT_M_XYA, j) - m(TT_M_XYB, ind) - half * S) * gr));
    }
    // z faces: i = 1..nx, j = 0..ny (j = 0 periodic copy), k = 1..nz+1
    if (i <= g.nx && j <= g.ny) {
        int jj = j == 0 ? g.ny : j;
        double hv = H[IX(i, jj, k)];
        double S = hv * dt;
        int ind = hv > 0.0 ? k - 1 : k;
        double gr = vl_slope(R(i, jj, ind - 1), R(i, jj, ind), R(i, jj, ind + 1), m(TT_M_DZB, ind), m(TT_M_DZB, ind + 1), 0);
        st_nt(&FLMZ[IX(i, j, k)], (m(TT_M_SURZA, k) + m(TT_M_CCZA, k) * S) * S *
                                      (R(i, jj, ind) + (m(TT_M_XZA, k) - m(TT_M_XZB, ind) - half * S) * gr));
    }
}

// ---- face-shared sweeps.  The four transport updates evaluate, per cell,
// the fluxes through all six of its faces, so every face flux would be
// formed twice.  k_adv_t forms each face flux once: a block takes a tile of
// 32 (i) x 8 (j) cells and marches through KC planes of k.  Per plane each
// cell lane forms the x face on its low side and the y face on its low side
// into LDS (double-buffered, one barrier per plane) and the z face on its
// high side in a register, which is the low z face of the next plane; a
// fifth wavefront forms the tile's extra high x faces (8) and high y faces
// (32).  Addressing and metrics are arranged for few instructions per face:
// one 32-bit index per cell and plane with fixed offsets along the lines
// (+-1 in i, +-gx in j, +-gx*gy in k); the 1-D metrics a face or cell needs
// are loaded once per lane (x faces and i, y faces and j: before the march)
// or once per plane (z faces and k).  The flux expressions are the
// reference's, face by face, so the results equal the per-cell form bit for
// bit.
struct AdvArgs {
    const double *A;                 // advected quantity: rho c_v T, V, GD or HD
    const double *V, *G, *H;         // velocities
    const double *FX, *FY, *FZ;      // mass fluxes FLMX, FLMY, FLMZ
    const double *RHQ;               // face density of the momentum (unused for energy)
    double* out;
    double dt;
    double *cf1, *cf2, *cf3;         // energy only: conv_flux, the low-face fluxes (ppadv.f:857-865)
};

constexpr int ADV_TI = 32, ADV_TJ = 8;

// Offsets of a face's neighbours: one step along i, j, k.
struct Step {
    int y, z;  // gx, gx*gy (x is 1)
};

// The metrics of one face along its line: the spacings of the cells around
// it (d0..d2 or d_1..d1), two cell coordinates and a face coordinate.
struct FaceMet {
    double d0, d1, d2, x0, x1, xa;
};
// the metrics of a face whose low candidate cell is I-1 (centred quantities:
// DXB, XXB spacings; d_1, d0, d1 in d0..d2, x_1, x0 in x0, x1)
__device__ inline FaceMet fm_b(const Met& m, int D, int X, int XA, int I) {
    return FaceMet{m(D, I - 1), m(D, I), m(D, I + 1), m(X, I - 1), m(X, I), m(XA, I)};
}
// a face of a staggered quantity whose candidate cells are I and I+1 (DXA,
// XXA spacings: d0..d2 at I..I+2, x0, x1 at I, I+1, xa the cell's XXB(I))
__device__ inline FaceMet fm_a(const Met& m, int D, int X, int XB, int I) {
    return FaceMet{m(D, I), m(D, I + 1), m(D, I + 2), m(X, I), m(X, I + 1), m(XB, I)};
}

// KIND 0: thermal energy (ppadv.f:741-891), faces I = i, i+1 etc.; the
// upwind cell is I-1 for a positive velocity, else I.  All three face
// families have the same form.
struct AdvE {
    static constexpr int ox = 0, oy = 0, oz = 0;
    static constexpr int ilo = 1, jlo = 1, klo = 1;
    struct XF {
        FaceMet f;
        double sa, ca;
    };
    __device__ static XF xmet(const Met& m, int I, int) {
        return XF{fm_b(m, TT_M_DXB, TT_M_XXB, TT_M_XXA, I), m(TT_M_SURXA, I), m(TT_M_CCXA, I)};
    }
    __device__ static XF ymet(const Met& m, int J, int) {
        return XF{fm_b(m, TT_M_DYB, TT_M_XYB, TT_M_XYA, J), m(TT_M_SURYA, J), m(TT_M_CCYA, J)};
    }
    __device__ static XF zmet(const Met& m, int K) {
        return XF{fm_b(m, TT_M_DZB, TT_M_XZB, TT_M_XZA, K), m(TT_M_SURZA, K), m(TT_M_CCZA, K)};
    }
    struct PM {
        double volz;
    };
    __device__ static PM pmet(const Met& m, int k) { return PM{m(TT_M_VOLZB, k)}; }
    __device__ static double vol_x(const Met& m, int i) { return m(TT_M_VOLXB, i); }
    __device__ static double vol_y(const Met& m, int j) { return m(TT_M_VOLYB, j); }
    // face flux along a line with step s through the face's own cell n
    __device__ static double face(const AdvArgs& a, const double* __restrict__ vel, int n, int s, const XF& x) {
        const double half = 0.5;
        const double* __restrict__ E = a.A;
        const double v = vel[n];
        const double e_2 = E[n - 2 * s], e_1 = E[n - s], e0 = E[n], e1 = E[n + s];
        double S = v * a.dt;
        const bool up = v > 0.0;
        double eu;
        double gr = vl_face(up, e_2, e_1, e0, e1, x.f.d0, x.f.d1, x.f.d2, 0, &eu);
        return (x.sa - x.ca * S) * S * (eu + (x.f.xa - (up ? x.f.x0 : x.f.x1) - half * S) * gr);
    }
    __device__ static double fx(const AdvArgs& a, int n, Step, const XF& x, const PM&) { return face(a, a.V, n, 1, x); }
    __device__ static double fy(const AdvArgs& a, int n, Step s, const XF& y, const PM&) { return face(a, a.G, n, s.y, y); }
    __device__ static double fz(const AdvArgs& a, int n, Step s, const XF& z, int, double, double) {
        return face(a, a.H, n, s.z, z);
    }
    // the cell's own transported amount the update starts from
    __device__ static double own(const AdvArgs& a, int n) { return a.A[n]; }
    __device__ static void finish(const AdvArgs& a, int n, double q0, double vx, double vy, const PM& pm, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        st_nt(&a.out[n], q0 - (fx1 - fx0) / vx - (fy1 - fy0) / vy - (fz1 - fz0) / pm.volz);
        st_nt(&a.cf1[n], fx0);
        st_nt(&a.cf2[n], fy0);
        st_nt(&a.cf3[n], fz0);
    }
};

// KIND 1: radial momentum (ppadv.f:896-1073), cells i = 2..nx, x faces
// FLX(I) at I = i-1, i (upwind cell I for a positive S, else I+1)
struct AdvV {
    static constexpr int ox = -1, oy = 0, oz = 0;
    static constexpr int ilo = 2, jlo = 1, klo = 1;
    struct XF {
        FaceMet f;
        double p1, p0;  // DDX1, DDX0 of the face's cell column (y and z faces)
    };
    __device__ static XF xmet(const Met& m, int I, int) {
        return XF{fm_a(m, TT_M_DXA, TT_M_XXA, TT_M_XXB, I), 0.0, 0.0};
    }
    __device__ static XF ymet(const Met& m, int J, int i) {
        return XF{fm_b(m, TT_M_DYB, TT_M_XYB, TT_M_XYA, J), m(TT_M_DDX1, i), m(TT_M_DDX0, i)};
    }
    __device__ static XF zmet(const Met& m, int K) { return XF{fm_b(m, TT_M_DZB, TT_M_XZB, TT_M_XZA, K), 0.0, 0.0}; }
    struct PM {
        double volz;
    };
    __device__ static PM pmet(const Met& m, int k) { return PM{m(TT_M_VOLZB, k)}; }
    __device__ static double vol_x(const Met& m, int i) { return m(TT_M_VOLXA, i); }
    __device__ static double vol_y(const Met& m, int j) { return m(TT_M_VOLYB, j); }
    __device__ static double fx(const AdvArgs& a, int n, Step, const XF& x, const PM&) {
        const double half = 0.5;
        const double* __restrict__ Vb = a.V;
        const double v_1 = Vb[n - 1], v0 = Vb[n], v1 = Vb[n + 1], v2 = Vb[n + 2];
        const double f0 = a.FX[n], f1 = a.FX[n + 1];
        double S = (v0 + v1) * a.dt / 2.0;
        const bool up = S > 0.0;
        double vu;
        double gr = vl_face(up, v_1, v0, v1, v2, x.f.d0, x.f.d1, x.f.d2, 1, &vu);
        return half * (f0 + f1) * (vu + (x.f.xa - (up ? x.f.x0 : x.f.x1) - half * S) * gr);
    }
    // a y or z face at n along step s, with the column factors of the cell
    __device__ static double fyz(const AdvArgs& a, const double* __restrict__ vel, const double* __restrict__ fl,
                                 int n, int s, const FaceMet& f, double p1, double p0) {
        const double half = 0.5;
        const double* __restrict__ Vb = a.V;
        const double g1 = vel[n], g0 = vel[n - 1];
        const double f1 = fl[n], f0 = fl[n - 1];
        const double v_2 = Vb[n - 2 * s], v_1 = Vb[n - s], v0 = Vb[n], v1 = Vb[n + s];
        double S = (p1 * g1 + p0 * g0) * a.dt;
        const bool up = S > 0.0;
        double vu;
        double gr = vl_face(up, v_2, v_1, v0, v1, f.d0, f.d1, f.d2, 1, &vu);
        return (p1 * f1 + p0 * f0) * (vu + (f.xa - (up ? f.x0 : f.x1) - half * S) * gr);
    }
    __device__ static double fy(const AdvArgs& a, int n, Step s, const XF& y, const PM&) {
        return fyz(a, a.G, a.FY, n, s.y, y.f, y.p1, y.p0);
    }
    __device__ static double fz(const AdvArgs& a, int n, Step s, const XF& z, int, double p1, double p0) {
        return fyz(a, a.H, a.FZ, n, s.z, z.f, p1, p0);
    }
    __device__ static double own(const AdvArgs& a, int n) { return a.V[n] * a.RHQ[n]; }
    __device__ static void finish(const AdvArgs& a, int n, double q0, double vx, double vy, const PM& pm, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        st_nt(&a.out[n], q0 - (fx1 - fx0) / vx - (fy1 - fy0) / vy - (fz1 - fz0) / pm.volz);
    }
};

// KIND 2: azimuthal angular momentum (ppadv.f:1096-1266), y faces FLY(J) at
// J = j-1, j (upwind J for a positive S, else J+1); FLY(0) is the periodic
// copy of FLY(ny)
struct AdvG {
    static constexpr int ox = 0, oy = -1, oz = 0;
    static constexpr int ilo = 1, jlo = 1, klo = 1;
    struct XF {
        FaceMet f;
        double p1, p0;  // x and z faces: DDY1, DDY0 of the row; y faces: DCY0, DCY1 of J+1
    };
    __device__ static XF xmet(const Met& m, int I, int j) {
        return XF{fm_b(m, TT_M_DXB, TT_M_XXB, TT_M_XXA, I), m(TT_M_DDY1, j), m(TT_M_DDY0, j)};
    }
    __device__ static XF ymet(const Met& m, int J, int) {
        return XF{fm_a(m, TT_M_DYA, TT_M_XYA, TT_M_XYB, J), m(TT_M_DCY0, J + 1), m(TT_M_DCY1, J + 1)};
    }
    __device__ static XF zmet(const Met& m, int K) { return XF{fm_b(m, TT_M_DZB, TT_M_XZB, TT_M_XZA, K), 0.0, 0.0}; }
    struct PM {
        double volz;
    };
    __device__ static PM pmet(const Met& m, int k) { return PM{m(TT_M_VOLZB, k)}; }
    __device__ static double vol_x(const Met& m, int i) { return m(TT_M_VOLXB, i); }
    __device__ static double vol_y(const Met& m, int j) { return m(TT_M_VOLYA, j); }
    // x or z face at n along step s (s1: the step to j-1)
    __device__ static double fxz(const AdvArgs& a, const double* __restrict__ vel, const double* __restrict__ fl,
                                 int n, int s, int s1, const FaceMet& f, double p1, double p0) {
        const double half = 0.5;
        const double* __restrict__ D = a.A;
        const double u1 = vel[n], u0 = vel[n - s1];
        const double f1 = fl[n], f0 = fl[n - s1];
        const double q_2 = D[n - 2 * s], q_1 = D[n - s], q0 = D[n], q1 = D[n + s];
        double S = (p1 * u1 + p0 * u0) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_2, q_1, q0, q1, f.d0, f.d1, f.d2, 1, &qu);
        return (p1 * f1 + p0 * f0) * (qu + (f.xa - (up ? f.x0 : f.x1) - half * S) * gr);
    }
    __device__ static double fx(const AdvArgs& a, int n, Step s, const XF& x, const PM&) {
        return fxz(a, a.V, a.FX, n, 1, s.y, x.f, x.p1, x.p0);
    }
    __device__ static double fy(const AdvArgs& a, int n, Step s, const XF& y, const PM&) {
        const double half = 0.5;
        const double* __restrict__ D = a.A;
        const int sy = s.y;
        const double c0 = y.p1, c1 = y.p0;
        const double g0 = a.G[n], g1 = a.G[n + sy];
        const double f0 = a.FY[n], f1 = a.FY[n + sy];
        const double q_1 = D[n - sy], q0 = D[n], q1 = D[n + sy], q2 = D[n + 2 * sy];
        double S = (c0 * g0 + c1 * g1) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_1, q0, q1, q2, y.f.d0, y.f.d1, y.f.d2, 1, &qu);
        return (c0 * f0 + c1 * f1) * (qu + (y.f.xa - (up ? y.f.x0 : y.f.x1) - half * S) * gr);
    }
    __device__ static double fz(const AdvArgs& a, int n, Step s, const XF& z, int, double p1, double p0) {
        return fxz(a, a.H, a.FZ, n, s.z, s.y, z.f, p1, p0);
    }
    __device__ static double own(const AdvArgs& a, int n) { return a.A[n] * a.RHQ[n]; }
    __device__ static void finish(const AdvArgs& a, int n, double q0, double vx, double vy, const PM& pm, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        st_nt(&a.out[n], q0 - (fx1 - fx0) / vx - (fy1 - fy0) / vy - (fz1 - fz0) / pm.volz);
    }
};

// KIND 3: meridional angular momentum (ppadv.f:1278-1461), cells k = 2..nz,
// z faces FLZ(K) at K = k-1, k (upwind K for a positive S, else K+1)
struct AdvH {
    static constexpr int ox = 0, oy = 0, oz = -1;
    static constexpr int ilo = 1, jlo = 1, klo = 2;
    struct XF {
        FaceMet f;
        double p1, p0;  // unused: the x and y faces take DDZ1, DDZ0 of the plane
    };
    __device__ static XF xmet(const Met& m, int I, int) {
        return XF{fm_b(m, TT_M_DXB, TT_M_XXB, TT_M_XXA, I), 0.0, 0.0};
    }
    __device__ static XF ymet(const Met& m, int J, int) {
        return XF{fm_b(m, TT_M_DYB, TT_M_XYB, TT_M_XYA, J), 0.0, 0.0};
    }
    __device__ static XF zmet(const Met& m, int K) { return XF{fm_a(m, TT_M_DZA, TT_M_XZA, TT_M_XZB, K), 0.0, 0.0}; }
    struct PM {
        double volz, p1, p0;
    };
    __device__ static PM pmet(const Met& m, int k) { return PM{m(TT_M_VOLZA, k), m(TT_M_DDZ1, k), m(TT_M_DDZ0, k)}; }
    __device__ static double vol_x(const Met& m, int i) { return m(TT_M_VOLXB, i); }
    __device__ static double vol_y(const Met& m, int j) { return m(TT_M_VOLYB, j); }
    // x or y face at n along step s, with the plane factors (k-1: step sz)
    __device__ static double fxy(const AdvArgs& a, const double* __restrict__ vel, const double* __restrict__ fl,
                                 int n, int s, int sz, const FaceMet& f, double p1, double p0) {
        const double half = 0.5;
        const double* __restrict__ D = a.A;
        const double u1 = vel[n], u0 = vel[n - sz];
        const double f1 = fl[n], f0 = fl[n - sz];
        const double q_2 = D[n - 2 * s], q_1 = D[n - s], q0 = D[n], q1 = D[n + s];
        double S = (p1 * u1 + p0 * u0) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_2, q_1, q0, q1, f.d0, f.d1, f.d2, 1, &qu);
        return (p1 * f1 + p0 * f0) * (qu + (f.xa - (up ? f.x0 : f.x1) - half * S) * gr);
    }
    __device__ static double fx(const AdvArgs& a, int n, Step s, const XF& x, const PM& pm) {
        return fxy(a, a.V, a.FX, n, 1, s.z, x.f, pm.p1, pm.p0);
    }
    __device__ static double fy(const AdvArgs& a, int n, Step s, const XF& y, const PM& pm) {
        return fxy(a, a.G, a.FY, n, s.y, s.z, y.f, pm.p1, pm.p0);
    }
    __device__ static double fz(const AdvArgs& a, int n, Step s, const XF& z, int, double, double) {
        const double half = 0.5;
        const double* __restrict__ D = a.A;
        const int sz = s.z;
        const double h0 = a.H[n], h1 = a.H[n + sz];
        const double f0 = a.FZ[n], f1 = a.FZ[n + sz];
        const double q_1 = D[n - sz], q0 = D[n], q1 = D[n + sz], q2 = D[n + 2 * sz];
        double S = half * (h0 + h1) * a.dt;
        const bool up = S > 0.0;
        double qu;
        double gr = vl_face(up, q_1, q0, q1, q2, z.f.d0, z.f.d1, z.f.d2, 1, &qu);
        return half * (f0 + f1) * (qu + (z.f.xa - (up ? z.f.x0 : z.f.x1) - half * S) * gr);
    }
    __device__ static double own(const AdvArgs& a, int n) { return a.A[n] * a.RHQ[n]; }
    __device__ static void finish(const AdvArgs& a, int n, double q0, double vx, double vy, const PM& pm, double fx0,
                                  double fx1, double fy0, double fy1, double fz0, double fz1) {
        st_nt(&a.out[n], q0 - (fx1 - fx0) / vx - (fy1 - fy0) / vy - (fz1 - fz0) / pm.volz);
    }
};

// the z face's column factors, by kind (V: DDX of the cell's i; G: DDY of
// its j; E, H: none)
template <class K> __device__ inline void zcol(const Met&, int, int, double& p1, double& p0) { p1 = p0 = 0.0; }
template <> __device__ inline void zcol<AdvV>(const Met& m, int i, int, double& p1, double& p0) {
    p1 = m(TT_M_DDX1, i);
    p0 = m(TT_M_DDX0, i);
}
template <> __device__ inline void zcol<AdvG>(const Met& m, int, int j, double& p1, double& p0) {
    p1 = m(TT_M_DDY1, j);
    p0 = m(TT_M_DDY0, j);
}

template <class K, int KC>
__global__ __launch_bounds__(320) void k_adv_t(Geo g, Met m, AdvArgs a) {
    __shared__ double sx[2][ADV_TJ][ADV_TI + 1];  // x faces of the tile rows, plus the high one
    __shared__ double sy[2][ADV_TJ + 1][ADV_TI];  // y faces of the tile columns, plus the high row
    const int ilo = K::ilo, ihi = g.nx, jlo = K::jlo, jhi = g.ny, klo = K::klo, khi = g.nz;
    const int nti = (ihi - ilo + ADV_TI) / ADV_TI, ntj = (jhi - jlo + ADV_TJ) / ADV_TJ;
    const int nkc = (khi - klo + KC) / KC;
    // Hardware blocks b, b+8, ... run on one XCD; each XCD takes a contiguous
    // range of logical blocks, ordered with the tiles of a k chunk fastest,
    // so the blocks resident on one XCD at a time cover whole planes of a few
    // consecutive k chunks: a tile's x and y halos (its neighbour tiles) and
    // its z halo (the next chunk) are then read through that XCD's L2 rather
    // than fetched again (at 128x128x256 the fetched bytes fell from up to
    // 2.3x to 1.0-1.1x of the arrays' size).
    const int nb = nti * ntj * nkc;
    int b = (nb % 8 == 0) ? (int)(blockIdx.x % 8) * (nb / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    const int ti = b % nti;
    b /= nti;
    const int tj = b % ntj, kc = b / ntj;
    const int ci0 = ilo + ti * ADV_TI, cj0 = jlo + tj * ADV_TJ;
    const int k0 = klo + kc * KC;
    const int k1 = min(khi, k0 + KC - 1);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const bool helper = w == 4;
    const int r = 2 * w + (l >> 5), c = l & 31;
    // the highest face of each family
    const int Ihi = ihi + K::ox + 1, Jhi = jhi + K::oy + 1;
    const Step st{g.gx, g.gx * g.gy};
    // the lane's cell (i, j) and faces: cell lanes own a cell; the helper's
    // lanes 0..7 the tile's high x faces (row cj0 + l), lanes 8..39 its high
    // y faces (column ci0 + l - 8)
    int i, j;
    bool dox, doy, cell;
    if (!helper) {
        i = ci0 + c;
        j = cj0 + r;
        dox = j <= jhi && i + K::ox <= Ihi;
        doy = i <= ihi && j + K::oy <= Jhi;
        cell = i <= ihi && j <= jhi;
    } else if (l < ADV_TJ) {
        i = ci0 + ADV_TI;
        j = cj0 + l;
        dox = j <= jhi && i + K::ox <= Ihi;
        doy = cell = false;
    } else if (l < ADV_TJ + ADV_TI) {
        i = ci0 + (l - ADV_TJ);
        j = cj0 + ADV_TJ;
        doy = i <= ihi && j + K::oy <= Jhi;
        dox = cell = false;
    } else {
        i = ci0;
        j = cj0;
        dox = doy = cell = false;
    }
    // faces and metrics of lanes without work are clamped into the box
    const int ic = min(i, ihi + 1), jc = min(j, jhi + 1);
    int Jf = jc + K::oy;
    if (Jf == 0) Jf = g.ny;  // FLY(0) is the periodic copy of FLY(ny) (AdvG)
    const typename K::XF xm = K::xmet(m, ic + K::ox, jc);
    const typename K::XF ym = K::ymet(m, Jf, ic);
    const double volx = K::vol_x(m, min(i, ihi)), voly = K::vol_y(m, min(j, jhi));
    double zp1, zp0;
    zcol<K>(m, ic, jc, zp1, zp0);
    int n = (int)g.idx(ic, jc, k0);
    const int ny_wrap = (jc + K::oy == 0) ? g.ny * st.y : 0;  // the y face's index shift for the wrap
    double zlo = 0.0;
    if (cell) zlo = K::fz(a, n + K::oz * st.z, st, K::zmet(m, k0 + K::oz), k0 + K::oz, zp1, zp0);
    int p = 0;
    for (int k = k0; k <= k1; ++k, p ^= 1, n += st.z) {
        const typename K::PM pm = K::pmet(m, k);
        double zhi = 0.0;
        if (dox) {
            const double fx = K::fx(a, n + K::ox, st, xm, pm);
            if (!helper) sx[p][r][c] = fx;
            else sx[p][l][ADV_TI] = fx;
        }
        if (doy) {
            const double fy = K::fy(a, n + K::oy * st.y + ny_wrap, st, ym, pm);
            if (!helper) sy[p][r][c] = fy;
            else sy[p][ADV_TJ][l - ADV_TJ] = fy;
        }
        if (cell) zhi = K::fz(a, n + (K::oz + 1) * st.z, st, K::zmet(m, k + K::oz + 1), k + K::oz + 1, zp1, zp0);
        // the cell's own amount is read before the barrier, with the faces'
        // loads, so the update after it waits on LDS only
        const double q0 = cell ? K::own(a, n) : 0.0;
        __syncthreads();
        if (cell) {
            K::finish(a, n, q0, volx, voly, pm, sx[p][r][c], sx[p][r][c + 1], sy[p][r][c], sy[p][r + 1][c], zlo, zhi);
            zlo = zhi;
        }
    }
}

// planes per block: at 128x128x256 the four kernels take 561, 535 and 544 us
// together with 2, 4 and 8
constexpr int ADV_KC = 4;

template <class K>
static int launch_adv(const AdvArgs& a) {
    Geo& g = L.g;
    const int ni = g.nx - K::ilo + 1, nj = g.ny - K::jlo + 1, nk = g.nz - K::klo + 1;
    const long long nb = (long long)((ni + ADV_TI - 1) / ADV_TI) * ((nj + ADV_TJ - 1) / ADV_TJ) *
                         ((nk + ADV_KC - 1) / ADV_KC);
    k_adv_t<K, ADV_KC><<<(unsigned)nb, 320, 0, L.s>>>(g, met(), a);
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
    st_nt(&RH[c], RH[c] - (FLMX[IX(i + 1, j, k)] - FLMX[c]) / m(TT_M_VOLXB, i) -
                      (FLMY[IX(i, j + 1, k)] - FLMY[c]) / m(TT_M_VOLYB, j) -
                      (FLMZ[IX(i, j, k + 1)] - FLMZ[c]) / m(TT_M_VOLZB, k));
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
    if (i >= 2) st_nt(&V[c], VD[c] / RHQX[c]);
    double gx = m(TT_M_GEOXG, i), gz = m(TT_M_GEOZG, k), gh = m(TT_M_GEOXH, i);
    st_nt(&GD[c], GDN[c]);
    st_nt(&G[c], GDN[c] / (RHQY[c] * (gx * gx) * (gz * gz)) - om);
    if (k >= 2) st_nt(&H[c], HDN[c] / (RHQZ[c] * (gh * gh)));
    st_nt(&T[c], ED[c] / (RH[c] * CV[c]));
}

int advect_impl(double dt) {
    Geo& g = L.g;
    double** f = L.f;
    int rc;
    if ((rc = bounds_impl())) return rc;
    // mass fluxes, and E0 (into TOLD), GD and HD (into HH) for the sweeps
    MassArgs ma{f[TT_RH], f[TT_V], f[TT_G], f[TT_H], f[TT_T], f[TT_CV], L.FLMX, L.FLMY, L.FLMZ, L.TOLD, f[TT_GD],
                L.HH, dt, L.p.omrot};
    k_adv_mass<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, met(), ma);
    TT_LAUNCH();
    if ((rc = qdens_impl())) return rc;
    long long nc = (long long)g.nx * g.ny * g.nz;
    AdvArgs aa{L.TOLD, f[TT_V], f[TT_G], f[TT_H], L.FLMX, L.FLMY, L.FLMZ, nullptr, L.ED, dt,
               f[TT_CONV_FLUX1], f[TT_CONV_FLUX2], f[TT_CONV_FLUX3]};
    if ((rc = launch_adv<AdvE>(aa))) return rc;
    aa.A = f[TT_V]; aa.RHQ = f[TT_RHQX]; aa.out = L.VD;
    if ((rc = launch_adv<AdvV>(aa))) return rc;
    aa.A = f[TT_GD]; aa.RHQ = f[TT_RHQY]; aa.out = L.GDN;
    if ((rc = launch_adv<AdvG>(aa))) return rc;
    aa.A = L.HH; aa.RHQ = f[TT_RHQZ]; aa.out = L.HDN;
    if ((rc = launch_adv<AdvH>(aa))) return rc;
    k_adv_rho<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), f[TT_RH], L.FLMX, L.FLMY, L.FLMZ);
    TT_LAUNCH();
    L.ub_repeat = false;
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
// sin and cos of XZA(k) come from SCZ (k_scza: the same device functions,
// once per theta face instead of once per cell)
__global__ void k_centcor(Geo g, Met m, double om, const double* __restrict__ G, const double* __restrict__ H,
                          const double* __restrict__ SCZ, double* CCX, double* CCY, double* CCZ) {
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
    st_nt(&CCX[c], tq * tq + xa * (HQ * HQ));
    if (CCY) st_nt(&CCY[c], 0.0);  // null: CCY holds these zeros already
    G1 = m(TT_M_DCY0, j + 1) * G[IX(i, j, k - 1)] + m(TT_M_DCY1, j + 1) * G[IX(i, j + 1, k - 1)];
    G2 = m(TT_M_DCY0, j + 1) * G[IX(i, j, k)] + m(TT_M_DCY1, j + 1) * G[IX(i, j + 1, k)];
    double GQ = m(TT_M_DDZ1, k) * G2 + m(TT_M_DDZ0, k) * G1 + om;
    st_nt(&CCZ[c], -m(TT_M_XXB, i) * GQ * GQ * SCZ[2 * (k + 2)] * SCZ[2 * (k + 2) + 1]);
}

// p dV work of cell (i, j, k): T - fac*dt PG/RH/CV div(v); with DISFN
// (dissipative heating, IDISS) then + fac*dt DISFN/CV/RH on the interior, as
// the reference's following loop does cell by cell (ppfor.f:190-210,335-343)
__device__ inline double pdv_cell(const Geo& g, const Met& m, double dtf, double t, long long c, int i, int j, int k,
                                  const double* __restrict__ PG, const double* __restrict__ RH,
                                  const double* __restrict__ CV, const double* __restrict__ V,
                                  const double* __restrict__ G, const double* __restrict__ H,
                                  const double* __restrict__ DISFN) {
    double tn = t - dtf * PG[c] / RH[c] / CV[c] * divv(g, m, V, G, H, i, j, k);
    if (DISFN && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz)
        tn = tn + dtf * DISFN[c] / CV[c] / RH[c];
    return tn;
}

// p dV work in place over cells lo..hi in each axis (hi = n + hiadd)
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
    st_nt(&T[c], pdv_cell(g, m, dtf, T[c], c, i, j, k, PG, RH, CV, V, G, H, DISFN));
}

// The same over cells lo..hi, written out of place: TN over the whole box
// holds T with the work done there and T's own value elsewhere, so TN is
// exactly what the in-place form leaves in T while T keeps its old values.
__global__ void k_pdv_box(Geo g, Met m, double dtf, int lo, int hiadd, const double* __restrict__ T,
                          double* __restrict__ TN, const double* __restrict__ PG, const double* __restrict__ RH,
                          const double* __restrict__ CV, const double* __restrict__ V, const double* __restrict__ G,
                          const double* __restrict__ H, const double* __restrict__ DISFN) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    const bool in = i >= lo && i <= g.nx + hiadd && j >= lo && j <= g.ny + hiadd && k >= lo && k <= g.nz + hiadd;
    st_nt(&TN[t], in ? pdv_cell(g, m, dtf, T[t], t, i, j, k, PG, RH, CV, V, G, H, DISFN) : T[t]);
}

// VH, GH, HH (ppfor.f:217-278); DX/DY/DZ: divtx/y/z, null without viscosity
// (the reference's zero arrays); CCY (CENTRIFCORIOLIS sets it to zero), GY
// and GZ null where they hold +0.0 throughout, which is then added as a
// constant: the same operations on the same values.  Each cell reads only
// its own V, G, H, so
// the kernel also applies the update that follows in place: avg != 0 sets
// V = (V+VH)/2 etc. and stores VH, GH, HH for the second half (ppfor.f:
// 322-328); avg == 0 sets V = VH etc. (the isothermal case, :346-352).
__global__ void k_force_vh(Geo g, Met m, double dt, int avg, double* V, double* G, double* H,
                           const double* __restrict__ PG, const double* __restrict__ QX,
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
    const double v = V[c], gv = G[c], hv = H[c];
    const double vh = v + dt * ((-(PG[c] - PG[IX(i - 1, j, k)]) / m(TT_M_DXB, i) + dx) / QX[c] + CCX[c] + GX[c]);
    const double ccy = CCY ? CCY[c] : 0.0, gyc = GY ? GY[c] : 0.0, gzc = GZ ? GZ[c] : 0.0;
    const double gvh =
        gv + dt * (((-(PG[c] - PG[IX(i, j - 1, k)]) / m(TT_M_DYB, j) / gx / gz + dy) / QY[c] + ccy + gyc) / gx / gz);
    const double hvh = hv + dt * (((-(PG[c] - PG[IX(i, j, k - 1)]) / m(TT_M_DZB, k) / gh + dz) / QZ[c] + CCZ[c] + gzc) / gh);
    if (avg) {
        st_nt(&VH[c], vh);
        st_nt(&GH[c], gvh);
        st_nt(&HH[c], hvh);
        st_nt(&V[c], (v + vh) / 2.0);
        st_nt(&G[c], (gv + gvh) / 2.0);
        st_nt(&H[c], (hv + hvh) / 2.0);
    } else {
        st_nt(&V[c], vh);
        st_nt(&G[c], gvh);
        st_nt(&H[c], hvh);
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
    k_centcor<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), L.p.omrot, f[TT_G], f[TT_H], L.scza, f[TT_CENTCORX],
                                              L.ccy_zero ? nullptr : f[TT_CENTCORY], f[TT_CENTCORZ]);
    TT_LAUNCH();
    L.ccy_zero = true;
    const bool adiab = L.p.isoth == 0;
    if (adiab) {
        // TOLD = T, then T at t(n+1/2) by the p dV work: formed out of place
        // into TOLD's buffer, after which the two buffers trade places, so T
        // holds the half-step temperature and TOLD the saved one without a
        // copy; they trade back where the reference sets T = TOLD
        k_pdv_box<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, met(), dt / 2.0, 0, 1, f[TT_T], L.TOLD, f[TT_PG], f[TT_RH],
                                                    f[TT_CV], f[TT_V], f[TT_G], f[TT_H], disfn);
        TT_LAUNCH();
        std::swap(f[TT_T], L.TOLD);
        if ((rc = bounds_impl()) || (rc = pressg_impl())) {
            std::swap(f[TT_T], L.TOLD);
            return rc;
        }
    }
    // VH, GH, HH, and V = (V+VH)/2 etc. (adiabatic) or V = VH etc.
    k_force_vh<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, adiab ? 1 : 0, f[TT_V], f[TT_G], f[TT_H], f[TT_PG],
                                               f[TT_RHQX], f[TT_RHQY], f[TT_RHQZ], f[TT_CENTCORX], nullptr,
                                               f[TT_CENTCORZ], f[TT_GRAVX], L.gy_zero ? nullptr : f[TT_GRAVY],
                                               L.gz_zero ? nullptr : f[TT_GRAVZ],
                                               vis ? f[TT_DIVTX] : nullptr, vis ? f[TT_DIVTY] : nullptr,
                                               vis ? f[TT_DIVTZ] : nullptr, L.VH, L.GH, L.HH);
    TT_LAUNCH();
    if (adiab) {
        std::swap(f[TT_T], L.TOLD);  // T = TOLD
        // the velocity ceiling on the intermediate velocities (ppfor.f:303-315)
        if (L.lim.ivel != 0 && (rc = sphericalvellimit_impl(L.lim.ivel, L.lim.velmax, L.lim.hydrotime))) return rc;
        if ((rc = bounds_impl())) return rc;
        k_pdv<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), dt, 1, 0, f[TT_T], f[TT_PG], f[TT_RH], f[TT_CV], f[TT_V], f[TT_G],
                                              f[TT_H], disfn);
        TT_LAUNCH();
        // V = VH etc. on the interior: the buffers trade places.  The ghost
        // cells V, G and H then hold are not read before the BOUNDS below,
        // which rewrites every one of them from interior values and
        // constants; the velocity ceiling, if on, clamps cell by cell in
        // between.
        std::swap(f[TT_V], L.VH);
        std::swap(f[TT_G], L.GH);
        std::swap(f[TT_H], L.HH);
    }
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
    // BOUNDS, BOUNDS, PRESSG after COUPLED_ENERGY_EQS; the energy solve may
    // queue them ahead of its last control-block read (then that read has
    // checked their flags)
    auto trailing = [&]() -> int {
        int r;
        if ((r = bounds_impl()) || (r = bounds_impl()) || (r = pressg_impl())) return r;
        TT_HIP(hipEventRecord(e[3], L.s));
        return 0;
    };
    const bool queue_ahead = !(L.p.bcxmin == 9 && iii <= 0);  // BurrowsOPC at III <= 0 is not queued ahead
    if (queue_ahead) L.step_trailing = trailing;
    int erc = energy_impl(delt, iii, &s.energy);
    L.step_trailing = nullptr;
    if (erc > 0) return erc;
    if (!L.trailing_done) {
        if ((rc = trailing())) return rc;
        if ((rc = sync_check())) return rc;
    }
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
