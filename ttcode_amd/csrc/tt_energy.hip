// tt_energy.hip -- COUPLED_ENERGY_EQS (ppenergy.f:2-791) on gfx950: the
// backward-Euler gas/radiation energy Newton loop with the flux-limited
// diffusion build (KAPPA, FLIM, DIFF, CALCULATEFLUX, FLUX_BC,
// CALCULATEDIVF_WITHF, FILL_EVALUES, FILL_SOR_MATRIX) and one red-black SOR
// solve per Newton iteration.
//
// Per Newton iteration: the phi wrap of ER, one kernel per face direction
// (limiter, diffusion coefficient, flux and its boundary values), one
// per-cell kernel (divergence, EVAL1..8, the pointwise Newton terms, the SOR
// matrix and RHS with their boundary rows), the SOR solve, one per-cell kernel
// for delta_Q and the damped update with the two max reductions, then the
// Q/ER boundary blocks.
#include <cmath>
#include <cstring>
#include <vector>

#include "tt_common.h"

namespace tt {

int sor_scale_guess(double sc);
int cnn_predict_guess(double delt);
bool cnn_ready();
int cnn_collect_sample(double delt);
bool cnn_collecting();

#define IX(i, j, k) g.idx(i, j, k)

__global__ void k_q_init(double* Q, const double* __restrict__ RH, const double* __restrict__ CV,
                         const double* __restrict__ T, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) st_nt(&Q[t], RH[t] * CV[t] * T[t]);
}

// x**y for the power law's exponents: 1/2, 1, 2 and 3 by sqrt and products
// (correctly rounded, or at most two roundings, where the general pow of the
// device library is within one unit in the last place), anything else by
// pow.  y is a kernel argument, so the choice is uniform.
__device__ inline double pow_exp(double x, double y) {
    if (y == 0.5) return sqrt(x);
    if (y == 1.0) return x;
    if (y == 2.0) return x * x;
    if (y == 3.0) return x * x * x;
    return pow(x, y);
}

// KAPPA (pprphy.f:1-360): IKAPTYP=1 constant; IKAPTYP=26 (new) power law
// evaluated up to upperbnd, cells above take the upperbnd value; SUBSIG.
// Q (non-null) receives COUPLED_ENERGY_EQS's Q = RH*CV*T (ppenergy.f:63-69)
// in the same pass.
__global__ void k_kappa(Geo g, tt_params p, const int32_t* __restrict__ ub, const double* __restrict__ RH,
                        const double* __restrict__ T, double* XKR, double* XKP, double* XKA, double* DKP,
                        double* DKA, double* SIG, double sig, double* Q, const double* __restrict__ CV) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    if (Q) st_nt(&Q[t], RH[t] * CV[t] * T[t]);
    st_nt(&SIG[t], sig);
    if (p.ikaptyp == 1) {
        st_nt(&XKR[t], p.zkapmult); st_nt(&XKP[t], p.zkapmult); st_nt(&XKA[t], p.zkapmult);
        st_nt(&DKP[t], 0.0); st_nt(&DKA[t], 0.0);
        return;
    }
    int i, j, k;
    box_ijk(g, t, i, j, k);
    int u = ub[g.idx2(j, k)];
    int ii = i <= u ? i : u;
    long long n = g.idx(ii, j, k);
    double rr = RH[n] / p.kap_rho0, tt = T[n] / p.kap_t0;
    const double pa = pow_exp(rr, p.kap_a);
    double kap = p.kap0 * pa * pow_exp(tt, p.kap_b);
    double dk = p.dkap0 * pa * pow_exp(tt, p.kap_b - 1.0);
    st_nt(&XKR[t], kap); st_nt(&XKP[t], kap); st_nt(&XKA[t], kap);
    st_nt(&DKP[t], dk); st_nt(&DKA[t], dk);
}

// ---- IKAPTYP 4/5: table opacities
// Axes of the tables: the reference's DATA statements (ppmod.f:690-699),
// temperatures in K, pressures in mbar.
__constant__ double c_ft[TT_FREEDMAN_NT] = {
    75.0, 100.0, 110.0, 120.0, 130.0, 140.0, 150.0, 160.0, 170.0, 180.0, 190.0, 200.0, 210.0, 220.0,
    230.0, 240.0, 250.0, 275.0, 300.0, 400.0, 500.0, 575.0, 650.0, 725.0, 800.0, 900.0, 1000.0, 1100.0,
    1200.0, 1300.0, 1400.0, 1500.0, 1600.0, 1700.0, 1800.0, 1900.0, 2000.0, 2300.0, 2600.0, 3000.0,
    3500.0, 4000.0};
__constant__ double c_fp[TT_FREEDMAN_NP] = {0.001, 0.003, 0.01, 0.03, 0.1, 0.3, 1.0, 3.0, 10.0,
                                            30.0, 100.0, 300.0, 1000.0, 3000.0, 10000.0, 30000.0,
                                            100000.0, 300000.0};

// KAPPA, IKAPTYP 4/5 (pprphy.f:84-120, MODTYP=2): FreedmanOPC (pprphy.f:800-
// 942) for cells up to upperbnd, whose values the cells above take, over the
// whole ghosted box; SUBSIG.  One thread per cell evaluates at
// (min(i, upperbnd), j, k).  FreedmanOPC: (T, P) clamped into the table, the
// first node above each, bilinear weights as the reference forms them, the
// log10 interpolants raised to 10^x, and dk/dT as the one-sided difference
// of the upper T node's column.  tab: Planck, Rosseland, absorption, then
// 10^ of each entry (formed on the host), which the difference reads.  Compute-bound:
// three 10^x per cell (the four of the derivative come from the 10^ table).
//
// IKAPTYP 6/7 (pprphy.f:123-175, columnOPC pprphy.f:1543-1745) with col set:
// the absorption mean from the column-depth table instead, interpolated in
// (T, P) on the two column levels around the cut-off log10(2 mu m_p / (3 rho
// (r(upperbnd) - r + 1))) (the reference's -23.9526937 + log10 mu - log10
// rho - log10(dr + 1), clamped into -26 ... -19), with the level weights as
// the reference forms them (its CS0 takes the pressure node FreedmanPVAL(IC1),
// kept for parity); dkA/dT from the upper level.  col: the 42 x 18 x 8 table
// then 10^ of it.
__global__ void k_kappa_tab(Geo g, Met m, const int32_t* __restrict__ ub, const double* __restrict__ T,
                            const double* __restrict__ PG, const double* __restrict__ RH,
                            const double* __restrict__ tab, const double* __restrict__ col, double logmu,
                            double* XKR, double* XKP, double* XKA, double* DKP, double* DKA, double* SIG,
                            double sig, Ctrl* ctrl) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    SIG[t] = sig;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    const int u = ub[g.idx2(j, k)];
    const int ie = i <= u ? i : u;
    const long long n = g.idx(ie, j, k);
    const double t_in = T[n];
    double pg1 = PG[n] * (0.001);
    if (!(t_in > 0.0) || pg1 != pg1) {  // the reference stops (pprphy.f:825-828,841-862)
        ctrl->bad_kappa = 1;
        return;
    }
    double tm1 = t_in;
    if (tm1 < 75.0) tm1 = 75.001;
    if (tm1 > 4000.0) tm1 = 3999.9;
    if (pg1 < 0.001) pg1 = 0.001001;
    if (pg1 > 300000.0) pg1 = 299999.9;
    int it = 1;  // the clamps put the first node above T in 1..41
    while (it < TT_FREEDMAN_NT - 1 && !(c_ft[it] > tm1)) ++it;
    const int it1 = it - 1;
    double ws1 = c_ft[it1] - tm1;
    double ws = c_ft[it1] - c_ft[it];
    ws1 = ws1 / ws;
    ws = 1.0 - ws1;
    int iw = 1;
    while (iw < TT_FREEDMAN_NP - 1 && !(c_fp[iw] > pg1)) ++iw;
    const int iw1 = iw - 1;
    double ws2 = pg1 - c_fp[iw1];
    double ws3 = c_fp[iw] - c_fp[iw1];
    ws2 = ws2 / ws3;
    ws3 = 1.0 - ws2;
    const int TN = TT_FREEDMAN_NT * TT_FREEDMAN_NP;
    double out[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double* tq = tab + q * TN;
        const double a00 = tq[it + TT_FREEDMAN_NT * iw], a10 = tq[it + TT_FREEDMAN_NT * iw1];
        const double a11 = tq[it1 + TT_FREEDMAN_NT * iw1], a01 = tq[it1 + TT_FREEDMAN_NT * iw];
        const double w4 = ws2 * (ws1 * a00 + ws * a01);
        const double w5 = ws3 * (ws1 * a10 + ws * a11);
        out[q] = exp10(w4 + w5);  // 10.d0**x; within a few ulp of pow(10, x), at half its cost
    }
    XKP[t] = out[0];
    XKR[t] = out[1];
    const double dt = c_ft[it] - c_ft[it1];
    const double* p10 = tab + 3 * TN;  // 10^ of the Planck table
    DKP[t] = (p10[it + TT_FREEDMAN_NT * iw] - p10[it1 + TT_FREEDMAN_NT * iw]) / dt;
    if (!col) {
        const double* a10 = tab + 5 * TN;  // and of the absorption table
        XKA[t] = out[2];
        DKA[t] = (a10[it + TT_FREEDMAN_NT * iw] - a10[it1 + TT_FREEDMAN_NT * iw]) / dt;
        return;
    }
    const double dr = m(TT_M_XXB, u) - m(TT_M_XXB, ie);
    double kc = (double)-23.9526937f + logmu - log10(RH[n]) - log10(dr + 1.0);
    if (kc < -26.0) kc = -25.9999;
    if (kc > -19.0) kc = -19.0001;
    if (kc != kc) {  // rho < 0: no level above, the reference stops (pprphy.f:1639-1651)
        ctrl->bad_kappa = 1;
        return;
    }
    int ic = 1;  // levels -26 ... -19: the first above kc is 1..7
    while (ic < 7 && !(-26.0 + ic > kc)) ++ic;
    const int ic1 = ic - 1;
    double cs1 = kc - (-26.0 + ic1);
    double cs0 = (-26.0 + ic) - c_fp[ic1];
    cs1 = cs1 / cs0;
    cs0 = 1.0 - cs1;
    const int CN = TT_FREEDMAN_NT * TT_FREEDMAN_NP;
    const double* c0 = col + ic * CN;
    const double* c1 = col + ic1 * CN;
    const double a000 = c0[it + TT_FREEDMAN_NT * iw], a010 = c0[it + TT_FREEDMAN_NT * iw1];
    const double a011 = c0[it1 + TT_FREEDMAN_NT * iw1], a001 = c0[it1 + TT_FREEDMAN_NT * iw];
    const double a_0 = cs1 * (ws2 * (ws1 * a000 + ws * a001) + ws3 * (ws1 * a010 + ws * a011));
    const double a100 = c1[it + TT_FREEDMAN_NT * iw], a110 = c1[it + TT_FREEDMAN_NT * iw1];
    const double a111 = c1[it1 + TT_FREEDMAN_NT * iw1], a101 = c1[it1 + TT_FREEDMAN_NT * iw];
    const double a_1 = cs0 * (ws2 * (ws1 * a100 + ws * a101) + ws3 * (ws1 * a110 + ws * a111));
    XKA[t] = exp10(a_0 + a_1);
    const double* c10 = col + 8 * CN + ic * CN;  // 10^ of level ic
    DKA[t] = (c10[it + TT_FREEDMAN_NT * iw] - c10[it1 + TT_FREEDMAN_NT * iw]) / dt;
}

// ---- IKAPTYP 8: H- bound-free absorption (pprphy.f:176-198).  The
// reference's j loop starts at 1: cells at j < 1 keep their xkapR and dk/dT
// from before, and kappa_P = kappa_A = kappa_R everywhere after it.  Cells
// above upperbnd take the upperbnd values.  0.00342 is a REAL constant.
__global__ void k_kappa_hminus(Geo g, const int32_t* __restrict__ ub, const double* __restrict__ RH,
                               const double* __restrict__ T, double* XKR, double* XKP, double* XKA, double* DKP,
                               double* DKA, double* SIG, double sig) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    SIG[t] = sig;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    double kr = XKR[t];
    if (j >= 1) {
        const int u = ub[g.idx2(j, k)];
        const long long n = g.idx(i <= u ? i : u, j, k);
        const double r5 = pow(RH[n], 0.5), tt = T[n] / 2500.0;
        kr = 0.95 * r5 * pow(tt, 9.0);
        const double dk = (double)0.00342f * r5 * pow(tt, 8.0);
        XKR[t] = kr;
        DKP[t] = dk;
        DKA[t] = dk;
    }
    XKP[t] = kr;
    XKA[t] = kr;
}

// ---- IKAPTYP 9-12, 14: Burrows tables
// Axes (ppmod.f:724-745): log10 T and log10 rho, REAL constants of DATA
// statements into double arrays, so each is the float value widened.
#define TT_BURROWS_LOGT \
    1.69896f, 1.73978f, 1.78061f, 1.82143f, 1.86225f, 1.90308f, 1.94390f, 1.98473f, 2.02555f, 2.06637f, \
    2.10720f, 2.14802f, 2.18884f, 2.22967f, 2.27049f, 2.31132f, 2.35214f, 2.39296f, 2.43379f, 2.47461f, \
    2.51543f, 2.55626f, 2.59708f, 2.63790f, 2.67873f, 2.71955f, 2.76038f, 2.80120f, 2.84202f, 2.88285f, \
    2.92367f, 2.96406f, 3.00488f, 3.04571f, 3.08653f, 3.12735f, 3.16818f, 3.20900f, 3.24983f, 3.29065f, \
    3.33147f, 3.37230f, 3.41312f, 3.45394f, 3.49477f, 3.53559f, 3.57642f, 3.61724f, 3.65806f, 3.69889f
#define TT_BURROWS_LOGRH \
    -11.9996f, -11.7954f, -11.5913f, -11.3872f, -11.1831f, -10.9790f, -10.7748f, -10.5707f, -10.3666f, \
    -10.1625f, -9.95837f, -9.75425f, -9.55014f, -9.34602f, -9.14190f, -8.93778f, -8.73366f, -8.52954f, \
    -8.32543f, -8.12131f, -7.91719f, -7.71307f, -7.50895f, -7.30483f, -7.10071f, -6.89660f, -6.69248f, \
    -6.48836f, -6.28424f, -6.08012f, -5.87600f, -5.67189f, -5.46777f, -5.26365f, -5.05953f, -4.85541f, \
    -4.65129f, -4.44718f, -4.24479f, -4.04068f, -3.83656f, -3.63244f, -3.42875f, -3.22464f, -3.02052f, \
    -2.81640f, -2.61228f, -2.40816f, -2.20404f, -1.99993f
__constant__ double c_bt[TT_BURROWS_NT] = {TT_BURROWS_LOGT};
__constant__ double c_brh[TT_BURROWS_NR] = {TT_BURROWS_LOGRH};
static const double h_bt[TT_BURROWS_NT] = {TT_BURROWS_LOGT};

// index of the first of the 50 strictly increasing nodes ax[] above x, given
// x inside (ax[0], ax[49]) and a guess g of it - 1
__device__ inline int first_above(const double* ax, double x, double g) {
    int q = (int)g + 1;
    q = q < 1 ? 1 : (q > 49 ? 49 : q);
    while (q > 1 && ax[q - 1] > x) --q;
    while (q < 49 && !(ax[q] > x)) ++q;
    return q;
}

// KAPPA, IKAPTYP 9-12 and 14 (pprphy.f:199-234,242-274, MODTYP=2): BurrowsOPC
// (pprphy.f:1005-1225) for cells up to upperbnd, whose values the cells
// above take; SUBSIG.  One thread per cell evaluates at (min(i, upperbnd),
// j, k): log10 T and log10 rho clamped into the table (the clamp bounds are
// REAL constants too), the first node above each, the reference's bilinear
// weights, 10^ of the interpolants, and dk/dT as the one-sided difference of
// 10^ of the upper T node's column over 10^ of the two T nodes.  tab: the
// Planck, Rosseland and absorption tables and the log10 T axis, then 10^ of
// all of them (formed on the host).
// BurrowsOPC (pprphy.f:1005-1225, its grey part) of one cell: log10 T and
// log10 rho clamped into the table (the clamp bounds are REAL constants too),
// the first node above each, the reference's bilinear weights, 10^ of the
// interpolants, and dk/dT as the one-sided difference of 10^ of the upper T
// node's column over 10^ of the two T nodes.  tab: the Planck, Rosseland and
// absorption tables and the log10 T axis, then 10^ of all of them (formed on
// the host).  out: kP, kR, kA, dkP/dT, dkA/dT.  Returns false where the
// reference stops: T <= 0 (pprphy.f:1039-1045) or rho < 0 (:1091-1097).
__device__ inline bool burrows_cell(double t_in, double rh_in, const double* __restrict__ tab, double out[5]) {
    if (!(t_in > 0.0)) return false;
    double rh1 = log10(rh_in);
    double tm1 = log10(t_in);
    if (tm1 < (double)1.69896f) tm1 = (double)1.69897f;
    if (tm1 > (double)3.69889f) tm1 = (double)3.69888f;
    if (rh1 < (double)-11.9996f) rh1 = (double)-11.9995f;
    if (rh1 > (double)-1.99993f) rh1 = (double)-1.99994f;
    if (rh1 != rh1) return false;  // rho < 0: no node above
    // the first node above each value (the clamps put it in 1..49): the axes
    // are nearly uniform, so a guess from the mean spacing, then the step to
    // the node the reference's scan finds (the axes increase strictly)
    const int it = first_above(c_bt, tm1, (tm1 - c_bt[0]) * (49.0 / (c_bt[49] - c_bt[0])));
    const int it1 = it - 1;
    double ws1 = c_bt[it1] - tm1;
    double ws = c_bt[it1] - c_bt[it];
    ws1 = ws1 / ws;
    ws = 1.0 - ws1;
    const int iw = first_above(c_brh, rh1, (rh1 - c_brh[0]) * (49.0 / (c_brh[49] - c_brh[0])));
    const int iw1 = iw - 1;
    double ws2 = rh1 - c_brh[iw1];
    double ws3 = c_brh[iw] - c_brh[iw1];
    ws2 = ws2 / ws3;
    ws3 = 1.0 - ws2;
    const int TN = TT_BURROWS_NT * TT_BURROWS_NR;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double* tq = tab + q * TN;
        const double a00 = tq[it + TT_BURROWS_NT * iw], a10 = tq[it + TT_BURROWS_NT * iw1];
        const double a11 = tq[it1 + TT_BURROWS_NT * iw1], a01 = tq[it1 + TT_BURROWS_NT * iw];
        const double w4 = ws2 * (ws1 * a00 + ws * a01);
        const double w5 = ws3 * (ws1 * a10 + ws * a11);
        out[q] = exp10(w4 + w5);  // 10.d0**x, within a few ulp
    }
    const double* p10 = tab + 3 * TN + TT_BURROWS_NT;  // 10^ of the tables and of the T axis
    const double* t10 = p10 + 3 * TN;
    const double dt = t10[it] - t10[it1];
    out[3] = (p10[it + TT_BURROWS_NT * iw] - p10[it1 + TT_BURROWS_NT * iw]) / dt;
    out[4] = (p10[2 * TN + it + TT_BURROWS_NT * iw] - p10[2 * TN + it1 + TT_BURROWS_NT * iw]) / dt;
    return true;
}

// KAPPA, IKAPTYP 9-12 and 14 (pprphy.f:199-234,242-274, MODTYP=2): BurrowsOPC
// for cells up to upperbnd, whose values the cells above take; SUBSIG.  One
// thread per cell evaluates at (min(i, upperbnd), j, k).  IKAPTYP 20-25
// (below != 0, pprphy.f:321-349): the cells from upperbnd up take
// upperbnd-1's values.
__global__ void k_kappa_bur(Geo g, const int32_t* __restrict__ ub, const double* __restrict__ T,
                            const double* __restrict__ RH, const double* __restrict__ tab, double* XKR,
                            double* XKP, double* XKA, double* DKP, double* DKA, double* SIG, double sig,
                            int below, Ctrl* ctrl) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    SIG[t] = sig;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    const int u = ub[g.idx2(j, k)];
    const int src = below ? (i < u ? i : u - 1) : (i <= u ? i : u);
    const long long n = g.idx(src, j, k);
    double o[5];
    if (!burrows_cell(T[n], RH[n], tab, o)) {
        ctrl->bad_kappa = 1;
        return;
    }
    XKP[t] = o[0];
    XKR[t] = o[1];
    XKA[t] = o[2];
    DKP[t] = o[3];
    DKA[t] = o[4];
}

// BOUNDXMIN(9) at III <= 0 (ppbou.f:1036-1046): BurrowsOPC at cell (1,1,1),
// written there (the reference passes those array elements)
// written there (the reference passes those array elements), and the
// bottom gradient -3 FBottom kappa_R rho / (4 a c T^3) every column then
// takes, into ctrl->bot_dtdr (T**3.d0 as T*T*T, as the reference's compiler
// evaluates it)
__global__ void k_bur_cell111(Geo g, const double* __restrict__ T, const double* __restrict__ RH,
                              const double* __restrict__ tab, double* XKR, double* XKP, double* XKA, double* DKP,
                              double* DKA, double fbottom, Ctrl* ctrl) {
    if (threadIdx.x != 0) return;
    const long long n = g.idx(1, 1, 1);
    double o[5];
    if (!burrows_cell(T[n], RH[n], tab, o)) {
        ctrl->bad_kappa = 1;
        return;
    }
    XKP[n] = o[0];
    XKR[n] = o[1];
    XKA[n] = o[2];
    DKP[n] = o[3];
    DKA[n] = o[4];
    const double t1 = T[n];
    ctrl->bot_dtdr = -3.0 * fbottom * o[1] * RH[n] / (4.0 * ARAD * FD * (t1 * t1 * t1));
}


// The grey irradiation, one 64-lane wavefront per (j, k) column (ghost
// columns included): CALCSTELLARTAU (pprad.f:191-221), STELLAR_DTAU
// (:223-242) and STELLARINPUT_HEATING's grey branch (:490-502).
//  1. lanes over the cells load xkapA and RH (coalesced along r) and form
//     (xkapA RH) DR into LDS;
//  2. one lane sums them from upperbnd down in the reference's order, tau =
//     tau + (xkapA RH) DR, sixteen LDS values per batch so the reads overlap;
//     above upperbnd tau is zeroed when upperbnd < nx, else STELLARTAU(nx+1)
//     keeps its value, as the reference's zeroing covers upperbnd+1..nx+1;
//  3. lanes over the cells write STELLARTAU, DTAUS = (tau(i-1) - tau(i+1)) /
//     2 with DTAUS(1) = DTAUS(2) and DTAUS(nx+1) = DTAUS(nx), and
//     STELLARINPUT = SBCONST (Tinc^4 / mu*) exp(-tau / mu*) DTAUS / dxb for
//     i = 1..nx (the reference's i = nx+1, nx+2 land in the next column and
//     are overwritten there); the column's SBCONST (Tinc^4 / mu*) comes from
//     k_sin_coef.
// CALC_TINCIDENT (F_INCIDENT 2, pprad.f:321-330) and the column factor
// SBCONST (Tinc^4 / mu*) of the grey STELLARINPUT, once per (j, k) column
__global__ void k_sin_coef(Geo g, Met m, double tirr, const double* __restrict__ MU, double* COEF) {
    const int nj = g.ny + 4, ncol = nj * (g.nz + 4);
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    const int j = col % nj - 1, k = col / nj - 1;
    const double angle = cos(m(TT_M_XYB, j)) * cos(m(TT_M_XZB, k));
    double tinc = 0.0;
    if (angle >= 0.0) {
        tinc = tirr * pow(angle, 0.25);
        tinc = tinc > 0.0 ? tinc : 0.0;
    }
    COEF[col] = SBCONST * (pow(tinc, 4.0) / MU[g.idx2(j, k)]);
}

__global__ __launch_bounds__(64) void k_stau_sin(Geo g, Met m, const int32_t* __restrict__ ub,
                                                 const double* __restrict__ XKA, const double* __restrict__ RH,
                                                 const double* __restrict__ MU, const double* __restrict__ COEF,
                                                 double* STAU, double* DTAUS, double* SIN) {
    extern __shared__ double tau[];  // tau[i], i = 0..nx+1
    const int nj = g.ny + 4;
    const int j = (int)(blockIdx.x % nj) - 1, k = (int)(blockIdx.x / nj) - 1;
    const int lane = threadIdx.x;
    const int u = ub[g.idx2(j, k)];
    const double dr = m(TT_M_XXB, 2) - m(TT_M_XXB, 1);
    for (int i = lane; i <= g.nx + 1; i += 64) {
        double v = 0.0;
        if (i >= 1 && i <= u) v = XKA[IX(i, j, k)] * RH[IX(i, j, k)] * dr;
        else if (i == g.nx + 1 && u >= g.nx) v = STAU[IX(i, j, k)];
        tau[i] = v;
    }
    __syncthreads();
    if (lane == 0) {
        double tot = 0.0;
        for (int i0 = u; i0 >= 1; i0 -= 16) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = i0 - q >= 1 ? tau[i0 - q] : 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                if (i0 - q >= 1) {
                    tot = tot + v[q];
                    tau[i0 - q] = tot;
                }
            }
        }
    }
    __syncthreads();
    const double mu = MU[g.idx2(j, k)];
    const double coef = COEF[blockIdx.x];
    for (int i = 1 + lane; i <= g.nx + 1; i += 64) {
        if (i <= u || u < g.nx) STAU[IX(i, j, k)] = tau[i];
        const int ic = i == 1 ? 2 : (i == g.nx + 1 ? g.nx : i);
        const double dt = (tau[ic - 1] - tau[ic + 1]) / 2.0;
        DTAUS[IX(i, j, k)] = dt;
        if (i <= g.nx) SIN[IX(i, j, k)] = coef * (exp(-tau[i] / mu)) * dt / m(TT_M_DXB, i);
    }
}

// The same three steps for C columns per 64-lane wavefront (consecutive
// (j, k) columns), which spreads the serial scan's instructions over C
// columns: step 1 loads the C columns' cells, NCH = ceil((nx + 2) / 64)
// per lane and column, in one batch (the addresses do not depend on
// upperbnd; the selection does), lanes 0..C-1 then scan one column each, in
// the reference's order, and step 3 writes the columns one after another.
// LDS rows are nx + 3 doubles apart, so the scanning lanes fall in distinct
// banks.  The last block's missing columns repeat the last column and are
// not written.
template <int C, int NCH>
__global__ __launch_bounds__(64) void k_stau_sin_c(Geo g, Met m, const int32_t* __restrict__ ub,
                                                   const double* __restrict__ XKA, const double* __restrict__ RH,
                                                   const double* __restrict__ MU, const double* __restrict__ COEF,
                                                   double* STAU, double* DTAUS, double* SIN) {
    extern __shared__ double tau[];  // tau[c * ld + i], i = 0..nx+1
    const int nj = g.ny + 4, ncol = nj * (g.nz + 4), ld = g.nx + 3;
    const int lane = threadIdx.x, c0 = blockIdx.x * C;
    const double dr = m(TT_M_XXB, 2) - m(TT_M_XXB, 1);
    int uc[C];
    double va[C][NCH], vb[C][NCH];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int col = min(c0 + c, ncol - 1);
        const int j = col % nj - 1, k = col / nj - 1;
        uc[c] = ub[g.idx2(j, k)];
#pragma unroll
        for (int h = 0; h < NCH; ++h) {
            const int i = min(lane + 64 * h, g.nx + 1);
            const long long n = IX(i, j, k);
            va[c][h] = (i == g.nx + 1 ? STAU : XKA)[n];
            vb[c][h] = RH[n];
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int u = uc[c];
#pragma unroll
        for (int h = 0; h < NCH; ++h) {
            const int i = lane + 64 * h;
            double v = 0.0;
            if (i >= 1 && i <= u) v = va[c][h] * vb[c][h] * dr;
            else if (i == g.nx + 1 && u >= g.nx) v = va[c][h];
            if (i <= g.nx + 1) tau[c * ld + i] = v;
        }
    }
    __syncthreads();
    if (lane < C) {
        const int col = min(c0 + lane, ncol - 1);
        const int u = ub[g.idx2(col % nj - 1, col / nj - 1)];
        double* t = tau + lane * ld;
        double tot = 0.0;
        for (int i0 = u; i0 >= 1; i0 -= 16) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = i0 - q >= 1 ? t[i0 - q] : 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                if (i0 - q >= 1) {
                    tot = tot + v[q];
                    t[i0 - q] = tot;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int col = c0 + c;
        if (col >= ncol) break;
        const int j = col % nj - 1, k = col / nj - 1;
        const int u = uc[c];
        const double* t = tau + c * ld;
        const double mu = MU[g.idx2(j, k)];
        const double coef = COEF[col];
        for (int i = 1 + lane; i <= g.nx + 1; i += 64) {
            if (i <= u || u < g.nx) STAU[IX(i, j, k)] = t[i];
            const int ic = i == 1 ? 2 : (i == g.nx + 1 ? g.nx : i);
            const double dt = (t[ic - 1] - t[ic + 1]) / 2.0;
            DTAUS[IX(i, j, k)] = dt;
            if (i <= g.nx) SIN[IX(i, j, k)] = coef * (exp(-tau[c * ld + i] / mu)) * dt / m(TT_M_DXB, i);
        }
    }
}

// The wavelength-binned STELLARINPUT_HEATING (pprad.f:429-447) for IKAPTYP
// 20, 21, 23 and 24, one 64-lane wavefront per (j, k) column:
//  1. lanes over the cells: each cell's 30 bins of xkapW (BurrowsOPC's
//     bilinear interpolant of the irradiation table, pprphy.f:1138-1151, at
//     the cell, or at upperbnd-1 for the cell upperbnd, as KAPPA copies it)
//     times RH, into LDS, tw[l][i];
//  2. lanes over the bins: the reverse scan from upperbnd, tau = tau +
//     (xkapW RH) DR, in the reference's order (WAVELENGTH_CALCSTELLARTAU,
//     pprad.f:519-547), in place; tau = 0 above upperbnd;
//  3. lanes over the cells: dtau = (tau(i-1) - tau(i+1)) / 2, dtau(1) =
//     dtau(2) (WAVELENGTH_STELLAR_DTAU), and the sum over the bins in order
//     of BSTAR_BIN exp(-tau/mu*) dtau, times the dilution, over dxb.
// The reference zeroes STELLARINPUT once and sums into it over i = 1..nx+2,
// while the array is declared 1:nx and tau, dtau 1:nx+1: the column before
// (in Fortran order) leaves at its i = nx+2 a sum of this column's tau(1),
// dtau(1) over its own mu* (times the dilution over dxb(nx+2)) in this
// column's STELLARINPUT(2), to which the own terms are added (its i = nx+1
// adds zeros: dtau(nx+1) = 0).  Reproduced as `carry`.
struct WaveArgs {
    double bstar[TT_NWAVE];
    double dil;
};

__global__ __launch_bounds__(64) void k_sin_wave(Geo g, Met m, const int32_t* __restrict__ ub,
                                                 const double* __restrict__ T, const double* __restrict__ RH,
                                                 const double* __restrict__ wtab, const double* __restrict__ MU,
                                                 WaveArgs w, double* SIN, Ctrl* ctrl) {
    extern __shared__ double tw[];  // tw[l * nt + i], i = 0..nx+1
    const int nt = g.nx + 2;
    const int nj = g.ny + 4;
    const int j = (int)(blockIdx.x % nj) - 1, k = (int)(blockIdx.x / nj) - 1;
    const int lane = threadIdx.x;
    const int u = ub[g.idx2(j, k)];
    const double dr = m(TT_M_XXB, 2) - m(TT_M_XXB, 1);
    const int TN = TT_BURROWS_NT * TT_BURROWS_NR;
    for (int i = lane; i <= g.nx + 1; i += 64) {
        if (i >= 1 && i <= u) {
            const int c = i < u ? i : u - 1;
            const double t_in = T[IX(c, j, k)], rh_in = RH[IX(c, j, k)];
            const double rh = RH[IX(i, j, k)];
            bool ok = t_in > 0.0;
            double rh1 = log10(rh_in), tm1 = log10(t_in);
            if (tm1 < (double)1.69896f) tm1 = (double)1.69897f;
            if (tm1 > (double)3.69889f) tm1 = (double)3.69888f;
            if (rh1 < (double)-11.9996f) rh1 = (double)-11.9995f;
            if (rh1 > (double)-1.99993f) rh1 = (double)-1.99994f;
            if (rh1 != rh1) ok = false;
            if (!ok) {
                ctrl->bad_kappa = 1;
                for (int l = 0; l < TT_NWAVE; ++l) tw[l * nt + i] = 0.0;
                continue;
            }
            const int it = first_above(c_bt, tm1, (tm1 - c_bt[0]) * (49.0 / (c_bt[49] - c_bt[0])));
            const int it1 = it - 1;
            double ws1 = c_bt[it1] - tm1;
            double ws = c_bt[it1] - c_bt[it];
            ws1 = ws1 / ws;
            ws = 1.0 - ws1;
            const int iw = first_above(c_brh, rh1, (rh1 - c_brh[0]) * (49.0 / (c_brh[49] - c_brh[0])));
            const int iw1 = iw - 1;
            double ws2 = rh1 - c_brh[iw1];
            double ws3 = c_brh[iw] - c_brh[iw1];
            ws2 = ws2 / ws3;
            ws3 = 1.0 - ws2;
            for (int l = 0; l < TT_NWAVE; ++l) {
                const double* tq = wtab + l * TN;
                const double a00 = tq[it + TT_BURROWS_NT * iw], a10 = tq[it + TT_BURROWS_NT * iw1];
                const double a11 = tq[it1 + TT_BURROWS_NT * iw1], a01 = tq[it1 + TT_BURROWS_NT * iw];
                const double w4 = ws2 * (ws1 * a00 + ws * a01);
                const double w5 = ws3 * (ws1 * a10 + ws * a11);
                tw[l * nt + i] = exp10(w4 + w5) * rh;
            }
        } else {
            for (int l = 0; l < TT_NWAVE; ++l) tw[l * nt + i] = 0.0;
        }
    }
    __syncthreads();
    if (lane < TT_NWAVE) {
        double* t = tw + lane * nt;
        double tot = 0.0;
        for (int i = u; i >= 1; --i) {
            tot = tot + (t[i] * dr);
            t[i] = tot;
        }
    }
    __syncthreads();
    const double mu = MU[g.idx2(j, k)];
    for (int i = 1 + lane; i <= g.nx; i += 64) {
        double s = 0.0;
        if (i == 2 && (j > -1 || k > -1)) {
            const double mup = j > -1 ? MU[g.idx2(j - 1, k)] : MU[g.idx2(g.ny + 2, k - 1)];
            double carry = 0.0;
            for (int l = 0; l < TT_NWAVE; ++l) {
                const double* t = tw + l * nt;
                const double d = (t[1] - t[3]) / 2.0;  // dtau(1) = dtau(2)
                carry = carry + w.bstar[l] * (exp(-t[1] / mup)) * d;
            }
            s = carry * w.dil / m(TT_M_DXB, g.nx + 2);
        }
        const int ic = i == 1 ? 2 : i;
        for (int l = 0; l < TT_NWAVE; ++l) {
            const double* t = tw + l * nt;
            const double d = (t[ic - 1] - t[ic + 1]) / 2.0;
            s = s + w.bstar[l] * (exp(-t[i] / mu)) * d;
        }
        SIN[IX(i, j, k)] = s * w.dil / m(TT_M_DXB, i);
    }
}

__global__ void k_fill_box(double* a, double v, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// Levermore-Pomraning limiter (pprphy.f:410-522)
__device__ inline double lp(double rr) { return (2.0 + rr) / (6.0 + 3.0 * rr + rr * rr); }

// The three face families: FLIM (x loop, y and z with FLIM's stale
// xmax_index, pprphy.f:410-522), DIFF (pprad.f:74-156), CALCULATEFLUX and
// FLUX_BC (ppenergy.f:880-982).
struct FaceArgs {
    const double *ER, *XKR, *SIG;
    const double *RHQ[3];
    double *FLIM[3], *DIFR[3], *FLUX[3];
    const int32_t* ub;
    int iflimtyp;
    double fbottom;
};

// All three families in one pass over (i, j, k) in [1, nx+1] x [1, ny+1] x
// [1, nz+1]: each thread forms the faces on the low side of its cell that
// exist there.
//
// Every limiter of the thread is evaluated at one cell, s = min(i, upperbnd-1):
// FLIMX takes the limiter of face min(i, upperbnd-1); FLIMY/FLIMZ copy the one
// of upperbnd-1 above the upper boundary, form their own below xmax_index and
// keep an earlier call's value (FLIM read back) for s beyond xmax_index
// (pprphy.f:464).  So the thread loads the union of the three limiters' E
// stencils around s (16 cells), and the cells at i that DIFF and the fluxes
// read, unconditionally and before any arithmetic; the expressions are those
// of the reference, operation for operation.  A stale limiter read is of a
// cell no thread changes in this call.  Indices stay inside the two ghost
// layers: s >= 0 since upperbnd >= 1, and i, j, k reach nx+1, ny+1, nz+1.
__global__ __launch_bounds__(256) void k_faces(Geo g, Met m, FaceArgs a) {
    long long n = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, g.nx + 1, g.ny + 1, 1, 1, 1, i, j, k);
    const bool hx = j <= g.ny && k <= g.nz, hy = i <= g.nx && k <= g.nz, hz = i <= g.nx && j <= g.ny;
    const int u = a.ub[g.idx2(j, k)];
    const int s = i <= u - 1 ? i : u - 1;
    const long long sy = g.gx, sz = (long long)g.gx * g.gy;
    const long long cs = IX(s, j, k), c = IX(i, j, k);
    const double* __restrict__ ER = a.ER;
    const double* __restrict__ XKR = a.XKR;
    const double* __restrict__ SIG = a.SIG;
    // E around s: eABC = E(s+A, j+B, k+C), m = -1, p = +1
    const double e000 = ER[cs], em00 = ER[cs - 1], ep00 = ER[cs + 1];
    const double e0p0 = ER[cs + sy], e0m0 = ER[cs - sy], emp0 = ER[cs - 1 + sy], emm0 = ER[cs - 1 - sy],
                 epm0 = ER[cs + 1 - sy];
    const double e00p = ER[cs + sz], e00m = ER[cs - sz], em0p = ER[cs - 1 + sz], em0m = ER[cs - 1 - sz],
                 ep0m = ER[cs + 1 - sz];
    const double e0mp = ER[cs - sy + sz], e0mm = ER[cs - sy - sz], e0pm = ER[cs + sy - sz];
    const double ks = XKR[cs], ksx = XKR[cs - 1], ksy = XKR[cs - sy], ksz = XKR[cs - sz], ss = SIG[cs];
    const double qsx = a.RHQ[0][cs], qsy = a.RHQ[1][cs], qsz = a.RHQ[2][cs];
    // the cells at i
    const double ei = ER[c], eix = ER[c - 1], eiy = ER[c - sy], eiz = ER[c - sz];
    const double ki = XKR[c], kix = XKR[c - 1], kiy = XKR[c - sy], kiz = XKR[c - sz];
    const double si = SIG[c], six = SIG[c - 1], siy = SIG[c - sy], siz = SIG[c - sz];
    const double qix = a.RHQ[0][c], qiy = a.RHQ[1][c], qiz = a.RHQ[2][c];
    const int xmaxs = a.ub[g.idx2(g.ny, g.nz)] - 1;

    double flx = 1.0 / 3.0, fly = 1.0 / 3.0, flz = 1.0 / 3.0;
    if (a.iflimtyp != 1) {
        {  // FLIMX at (s, j, k)
            double gx = (e000 - em00) / m(TT_M_DXB, s);
            double gy = (m(TT_M_DDX1, s) * (e0p0 - e0m0) + m(TT_M_DDX0, s) * (emp0 - emm0)) /
                        (m(TT_M_GEOXGA, s) * m(TT_M_GEOZG, k) * (m(TT_M_XYB, j + 1) - m(TT_M_XYB, j - 1)));
            double gz = (m(TT_M_DDX1, s) * (e00p - e00m) + m(TT_M_DDX0, s) * (em0p - em0m)) /
                        (m(TT_M_GEOXHA, s) * (m(TT_M_XZB, k + 1) - m(TT_M_XZB, k - 1)));
            double grad = sqrt(gx * gx + gy * gy + gz * gz) + 1.0e-50;
            double rr = 2.0 * grad /
                        (qsx * (m(TT_M_DDX1, s) * (ks + ss) + m(TT_M_DDX0, s) * (ksx + ss)) * (e000 + em00));
            flx = lp(rr);
        }
        if (s <= xmaxs) {
            {  // FLIMY at (s, j, k)
                double gx = (m(TT_M_DDY1, j) * (ep00 - em00) + m(TT_M_DDY0, j) * (epm0 - emm0)) /
                            (m(TT_M_XXB, s + 1) - m(TT_M_XXB, s - 1));
                double gy = (e000 - e0m0) / (m(TT_M_GEOXG, s) * m(TT_M_GEOZG, k) * m(TT_M_DYB, j));
                double gz = (m(TT_M_DDY1, j) * (e00p - e00m) + m(TT_M_DDY0, j) * (e0mp - e0mm)) /
                            ((m(TT_M_XZB, k + 1) - m(TT_M_XZB, k - 1)) * m(TT_M_GEOXH, s));
                double grad = sqrt(gx * gx + gy * gy + gz * gz) + 1.0e-50;
                double rr = 2.0 * grad /
                            (qsy * (m(TT_M_DDY1, j) * (ks + ss) + m(TT_M_DDY0, j) * (ksy + ss)) * (e000 + e0m0));
                fly = lp(rr);
            }
            {  // FLIMZ at (s, j, k)
                double gx = (m(TT_M_DDZ1, k) * (ep00 - em00) + m(TT_M_DDZ0, k) * (ep0m - em0m)) /
                            (m(TT_M_XXB, s + 1) - m(TT_M_XXB, s - 1));
                double gy = (m(TT_M_DDZ1, k) * (e0p0 - e0m0) + m(TT_M_DDZ0, k) * (e0pm - e0mm)) /
                            ((m(TT_M_XYB, j + 1) - m(TT_M_XYB, j - 1)) * m(TT_M_GEOXG, s) * m(TT_M_GEOZGA, k));
                double gz = (e000 - e00m) / (m(TT_M_GEOXH, s) * m(TT_M_DZB, k));
                double grad = sqrt(gx * gx + gy * gy + gz * gz) + 1.0e-50;
                double rr = 2.0 * grad /
                            (qsz * (m(TT_M_DDZ1, k) * (ks + ss) + m(TT_M_DDZ0, k) * (ksz + ss)) * (e000 + e00m));
                flz = lp(rr);
            }
        } else {
            fly = a.FLIM[1][cs];
            flz = a.FLIM[2][cs];
        }
    }
    // DIFF and CALCULATEFLUX + FLUX_BC at the faces of (i, j, k)
    const double dx = FD * flx / (m(TT_M_DXB, i) * qix * (m(TT_M_DDX1, i) * (ki + si) + m(TT_M_DDX0, i) * (kix + six)));
    double fx = -dx * (ei - eix);
    if (i >= u) fx = FD * e000 / 2.0;  // s = upperbnd-1 here
    if (i == 1) fx = a.fbottom;
    const double dy = FD * fly /
                      (m(TT_M_DYB, j) * m(TT_M_GEOXG, i) * m(TT_M_GEOZG, k) * qiy *
                       (m(TT_M_DDY1, j) * (ki + si) + m(TT_M_DDY0, j) * (kiy + siy)));
    double fy = -dy * (ei - eiy);
    if (j <= g.ny && i >= u) fy = 0.0;
    const double dz = FD * flz /
                      (m(TT_M_DZB, k) * m(TT_M_GEOXH, i) * qiz *
                       (m(TT_M_DDZ1, k) * (ki + si) + m(TT_M_DDZ0, k) * (kiz + siz)));
    double fz = -dz * (ei - eiz);
    if (k <= g.nz && i >= u) fz = 0.0;
    if (hx) { st_nt(&a.FLIM[0][c], flx); st_nt(&a.DIFR[0][c], dx); st_nt(&a.FLUX[0][c], fx); }
    if (hy) { st_nt(&a.FLIM[1][c], fly); st_nt(&a.DIFR[1][c], dy); st_nt(&a.FLUX[1][c], fy); }
    if (hz) { st_nt(&a.FLIM[2][c], flz); st_nt(&a.DIFR[2][c], dz); st_nt(&a.FLUX[2][c], fz); }
}

// The CALCULATEFLUX + FLUX_BC pair before the Newton loop (ppenergy.f:78-79)
// feeds only the top-boundary blocks, which read flux_x at i >= upperbnd and,
// through the i = nx+2 alias, at i = 1: faces FLUX_BC fixes to c E(upperbnd-1)/2
// and Fbottom (ppenergy.f:947,966-967,976).  Those values are written here,
// per interior column, with the expressions face_x_eval uses.  Every other
// flux, limiter and diffusion coefficient that call forms is overwritten by
// the Newton loop's first call before it is read; the ghost columns the alias
// reaches are rewritten by the theta blocks and the phi wrap.
__global__ void k_flux_x_top(Geo g, const double* __restrict__ ER, const int32_t* __restrict__ ub, double* FX,
                             double fbottom) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ny * g.nz) return;
    int j = t % g.ny + 1, k = t / g.ny + 1;
    int u = ub[g.idx2(j, k)];
    double top = FD * ER[IX(u - 1, j, k)] / 2.0;
    for (int i = u > 1 ? u : 2; i <= g.nx + 1; ++i) FX[IX(i, j, k)] = top;
    FX[IX(1, j, k)] = fbottom;
}

// FILL_EVALUES' six face terms and their sum (ppenergy.f:984-1030), from the
// diffusion coefficients DIFR and the metrics alone
struct Evals {
    double e1, e2, e3, e4, e5, e6, e16;
};
__device__ inline Evals fill_evalues(const Geo& g, const Met& m, const double* __restrict__ DX,
                                     const double* __restrict__ DY, const double* __restrict__ DZ, int i, int j,
                                     int k) {
    const long long c = IX(i, j, k);
    const double sxb = m(TT_M_SURXB, i), dxa = m(TT_M_DXA, i), xxb = m(TT_M_XXB, i), szb = m(TT_M_SURZB, k);
    const double dya = m(TT_M_DYA, j), dza = m(TT_M_DZA, k), sxa0 = m(TT_M_SURXA, i), sxa1 = m(TT_M_SURXA, i + 1);
    const double sza0 = m(TT_M_SURZA, k), sza1 = m(TT_M_SURZA, k + 1);
    Evals e;
    e.e1 = -sxa1 * DX[IX(i + 1, j, k)] / (sxb * dxa);
    e.e2 = -sxa0 * DX[c] / (sxb * dxa);
    e.e3 = -DY[IX(i, j + 1, k)] / (xxb * szb * dya);
    e.e4 = -DY[c] / (xxb * szb * dya);
    e.e5 = -DZ[IX(i, j, k + 1)] * sza1 / (xxb * szb * dza);
    e.e6 = -DZ[c] * sza0 / (xxb * szb * dza);
    e.e16 = e.e1 + e.e2 + e.e3 + e.e4 + e.e5 + e.e6;
    return e;
}

// EVAL1..6 and EVAL_1_6 into their module arrays over the interior, from the
// same expressions k_newton_cell evaluates
__global__ void k_evals(Geo g, Met m, const double* __restrict__ DX, const double* __restrict__ DY,
                        const double* __restrict__ DZ, double* E1, double* E2, double* E3, double* E4, double* E5,
                        double* E6, double* E16) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    const long long c = IX(i, j, k);
    const Evals e = fill_evalues(g, m, DX, DY, DZ, i, j, k);
    st_nt(&E1[c], e.e1); st_nt(&E2[c], e.e2); st_nt(&E3[c], e.e3); st_nt(&E4[c], e.e4);
    st_nt(&E5[c], e.e5); st_nt(&E6[c], e.e6); st_nt(&E16[c], e.e16);
}

struct CellArgs {
    const double *FX, *FY, *FZ, *DX, *DY, *DZ, *Q, *QOLD, *ER, *EROLD, *RH, *CV, *XKP, *XKA, *DKP, *DKA, *SIN;
    double *E1, *E2, *E3, *E4, *E5, *E6, *E7, *E8, *E16, *ETA, *FQ, *FR, *RHS, *DFQ;
    double *P0, *P1;  // the packed FP64 system, colours 0 and 1
    float *Q0, *Q1;   // its FP32 copy (null: not written)
    int hx;
    const int32_t* ub;
    int f_incident;
    double delt;
};

// CALCULATEDIVF_WITHF (:825-878) + FILL_EVALUES (:984-1030) + the Newton
// pointwise loop (:286-345) + FILL_EVALUES + FILL_SOR_MATRIX and
// B_VECTOR_BOUNDARIES (ppsor.f:189-329), per interior cell.
__global__ void k_newton_cell(Geo g, Met m, CellArgs a) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.nx * g.ny * g.nz) return;
    int i, j, k;
    sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
    long long c = IX(i, j, k);
    const double theta = 1.0, delt = a.delt;
    double sxb = m(TT_M_SURXB, i), dxa = m(TT_M_DXA, i), xxb = m(TT_M_XXB, i), szb = m(TT_M_SURZB, k);
    double dya = m(TT_M_DYA, j), dza = m(TT_M_DZA, k), sxa0 = m(TT_M_SURXA, i), sxa1 = m(TT_M_SURXA, i + 1);
    double sza0 = m(TT_M_SURZA, k), sza1 = m(TT_M_SURZA, k + 1);
    double divf = ((sxa1 / (sxb * dxa)) * a.FX[IX(i + 1, j, k)] - (sxa0 / (sxb * dxa)) * a.FX[c]) +
                  ((1.0 / (xxb * szb * dya)) * a.FY[IX(i, j + 1, k)] - (1.0 / (xxb * szb * dya)) * a.FY[c]) +
                  ((sza1 / (xxb * szb * dza)) * a.FZ[IX(i, j, k + 1)] - (sza0 / (xxb * szb * dza)) * a.FZ[c]);
    const Evals ev = fill_evalues(g, m, a.DX, a.DY, a.DZ, i, j, k);
    const double e1 = ev.e1, e2 = ev.e2, e3 = ev.e3, e4 = ev.e4, e5 = ev.e5, e6 = ev.e6, e16 = ev.e16;
    double rh = a.RH[c], cv = a.CV[c], q = a.Q[c], er = a.ER[c], kp = a.XKP[c];
    double B = ARAD * pow(q / (rh * cv), 4.0);
    double fq, heat = 0.0;
    // (1-theta)*(divF_old - divF) vanishes for theta = 1
    if (a.f_incident > 1) {
        heat = a.SIN[c] / (rh * a.XKA[c]);
        fq = q - a.QOLD[c] + delt * ((1.0 - theta) * 0.0 + (rh * kp * FD * (B - er)) - a.SIN[c]);
    } else {
        fq = q - a.QOLD[c] + delt * ((1.0 - theta) * 0.0 + rh * kp * FD * (B - er));
    }
    double fr = er - a.EROLD[c] + delt * (divf - rh * kp * FD * (B - er));
    double t3 = pow(q / (rh * cv), 3.0);
    double dfq;
    if (a.f_incident > 1)
        dfq = 1.0 + (delt / cv) * (kp * 4.0 * ARAD * FD * t3 + (FD * (B - er) * a.DKP[c]) - heat * a.DKA[c]);
    else
        dfq = 1.0 + (delt / cv) * (kp * 4.0 * ARAD * FD * t3 + FD * (B - er) * a.DKP[c]);
    double dfr = -(delt / cv) * (kp * 4.0 * ARAD * FD * t3 + FD * (B - er) * a.DKP[c]);
    double eta = dfq != 0.0 ? -dfr / dfq : 0.0;
    double rhs = -fr - eta * fq;
    double e7 = 1.0 + (1.0 - eta) * delt * rh * kp * FD;
    double e8 = (eta * (1.0 - theta) - 1.0) * delt;
    // EVAL1..6 and EVAL_1_6 are left to k_evals (they depend on DIFR only)
    st_nt(&a.E7[c], e7); st_nt(&a.E8[c], e8); st_nt(&a.ETA[c], eta);
    st_nt(&a.FQ[c], fq); st_nt(&a.FR[c], fr); st_nt(&a.DFQ[c], dfq);
    (void)a.RHS;
    // FILL_SOR_MATRIX + B_VECTOR_BOUNDARIES
    double sa = e7 + e8 * e16, sb = -e8 * e2, sc = -e8 * e1, sd = -e8 * e4, se = -e8 * e3, sf = -e8 * e6,
           sg = -e8 * e5;
    int u = a.ub[g.idx2(j, k)];
    bool ident = (i >= u) || (i == g.nx) || (k == 1) || (k == g.nz);
    if (ident) { sa = 1.0; sb = sc = sd = se = sf = sg = 0.0; rhs = 0.0; }
    if (i == u - 1 || i == g.nx - 1) sc = 0.0;
    if (k == 2) sf = 0.0;
    if (k == g.nz - 1) sg = 0.0;
    // straight into the colour-split row-blocked system the sweeps read:
    // colour (i+j+k)&1, row (j-1) + ny (k-1), slot (i-1)/2, vectors A..G, RHS
    const int col = (i + j + k) & 1;
    const long long o = (long long)((j - 1) + g.ny * (k - 1)) * 8 * a.hx + (i - 1) / 2;
    const double v[8] = {sa, sb, sc, sd, se, sf, sg, rhs};
    double* P = col ? a.P1 : a.P0;
#pragma unroll
    for (int q = 0; q < 8; ++q) P[o + (long long)q * a.hx] = v[q];
    if (a.Q0) {
        float* Q = col ? a.Q1 : a.Q0;
#pragma unroll
        for (int q = 0; q < 8; ++q) Q[o + (long long)q * a.hx] = (float)v[q];
    }
}

struct UpdArgs {
    const double *DER, *DFQ, *FQ, *RH, *XKP;
    double *Q, *ER;
    const int32_t* ub;
    double delt, conv;
    int imin;  // first cell of the update: 1 (BCXMIN 12), 2 (else; ppenergy.f:390-396)
};

// delta_Q from the delta_ER stencil (ppenergy.f:379-408), the positivity
// test, the two max reductions and the damped update (ppenergy.f:551-573).
// The stencil's seven terms are each multiplied by Q1 = -delt*(1 - theta),
// and theta is fixed at 1 (ppenergy.f:60): each term is a signed zero, and a
// signed zero added to Q2*d - FQ changes neither the result nor, where that
// is zero, anything computed from it (q + conv*dq, |dq/q|).  So the kernel
// forms dq = (Q2*d - FQ)/DFQ and reads neither EVAL1..6, EVAL_1_6 nor the
// neighbouring delta_ER values (as the reference does for finite ones).
__global__ __launch_bounds__(256) void k_dq_update(Geo g, UpdArgs a, double* part, int nb, Ctrl* ctrl,
                                                    const int* run) {
    __shared__ double sh[2][4];
    if (skip_run(run)) return;
    double mq = 0.0, me = 0.0;
    int bad = 0;
    long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        sub_ijk(t, g.nx, g.ny, 1, 1, 1, i, j, k);
        int u = a.ub[g.idx2(j, k)];
        if (i > u - 1 || i < a.imin) continue;
        long long c = IX(i, j, k);
        double Q2 = a.delt * a.RH[c] * a.XKP[c] * FD;
        double d = a.DER[c];
        double dq = (1.0 / a.DFQ[c]) * (Q2 * d - a.FQ[c]);
        double q = a.Q[c], er = a.ER[c];
        if ((q + a.conv * dq) <= 0.0) bad = 1;
        mq = fmax(mq, fabs(dq / q));
        me = fmax(me, fabs(d / er));
        st_nt(&a.ER[c], er + a.conv * d);
        st_nt(&a.Q[c], q + a.conv * dq);
    }
    if (bad) ctrl->bad_q = 1;
    double x = wave_max(mq), y = wave_max(me);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = x; sh[1][w] = y; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
        part[nb + blockIdx.x] = fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3]));
    }
}

// flux_x as the top-boundary blocks read it: declared (1:nx+1,1:ny,1:nz) but
// indexed over ghost j,k and i = nx+2; the read lands on the element with the
// same column-major offset (0 outside the array: those ghost values are
// overwritten by the theta blocks and the phi wrap).
__device__ inline double flux_x_alias(const Geo& g, const double* FX, int i, int j, int k) {
    long long sx = g.nx + 1, sy = g.ny;
    long long off = (long long)(i - 1) + sx * ((long long)(j - 1) + sy * (long long)(k - 1));
    if (off < 0 || off >= sx * sy * g.nz) return 0.0;
    int ii = (int)(off % sx) + 1;
    long long r = off / sx;
    int jj = (int)(r % sy) + 1, kk = (int)(r / sy) + 1;
    return FX[IX(ii, jj, kk)];
}

// Q/ER r boundaries per (j,k) column: upper (ppenergy.f:80-104, :681-723) and
// the bottom, BCXMIN=12 (:184-195, :640-651) or 9 (:150-167, :621-638).
// order 0: top first, 1: bottom first.
__global__ void k_energy_bc_x(Geo g, Met m, const int32_t* __restrict__ ub, double* Q, double* ER,
                              const double* __restrict__ RH, const double* __restrict__ CV,
                              const double* __restrict__ T, const double* __restrict__ FX,
                              const double* __restrict__ XKR, int bcxmin, double d_t, double fbottom, int order,
                              const int* run) {
    if (skip_run(run)) return;
    int nj = g.ny + 4, nk = g.nz + 4;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nj * nk) return;
    int j = t % nj - 1, k = t / nj - 1;
    for (int pass = 0; pass < 2; ++pass) {
        bool top = (pass == 0) == (order == 0);
        if (top) {
            int u = ub[g.idx2(j, k)];
            double qup = RH[IX(u - 1, j, k)] * CV[IX(u - 1, j, k)] * T[IX(u, j, k)];
            for (int i = u; i <= g.nx + 2; ++i) {
                Q[IX(i, j, k)] = qup;
                ER[IX(i, j, k)] = fabs(2.0 * flux_x_alias(g, FX, i, j, k) / FD);
            }
        } else if (bcxmin == 9) {
            // fixed flux, free T and Q: dQ/dr of the radiative diffusion that
            // carries FBottom (T**3.d0 as T*T*T), optically thick ER = a T^4
            const long long c1 = IX(1, j, k), c0 = IX(0, j, k), cm = IX(-1, j, k);
            const double rh = RH[c1], cv = CV[c1], t1 = T[c1];
            const double dq = -rh * cv * 3.0 * fbottom * XKR[c1] * rh / (4.0 * ARAD * FD * (t1 * t1 * t1));
            const double q1 = Q[c1];
            const double q0 = q1 - dq * (m(TT_M_XXB, 1) - m(TT_M_XXB, 0));
            const double qm = q1 - dq * (m(TT_M_XXB, 1) - m(TT_M_XXB, -1));
            Q[c0] = q0;
            Q[cm] = qm;
            ER[c1] = ARAD * pow(q1 / (rh * cv), 4.0);
            ER[c0] = ARAD * pow(q0 / (RH[c0] * CV[c0]), 4.0);
            ER[cm] = ARAD * pow(qm / (RH[cm] * CV[cm]), 4.0);
        } else {
            double dq = CV[IX(1, j, k)] * RH[IX(1, j, k)] * d_t;
            double q1 = Q[IX(1, j, k)];
            double q0 = q1 + dq * (m(TT_M_XXB, 0) - m(TT_M_XXB, 1));
            double qm = q1 + dq * (m(TT_M_XXB, -1) - m(TT_M_XXB, 1));
            Q[IX(0, j, k)] = q0;
            Q[IX(-1, j, k)] = qm;
            double den = RH[IX(1, j, k)] * CV[IX(1, j, k)];
            ER[IX(0, j, k)] = ARAD * pow(q0 / den, 4.0);
            ER[IX(-1, j, k)] = ARAD * pow(qm / den, 4.0);
        }
    }
}

// theta edges: zero gradient (ppenergy.f:204-227, :727-757) per (i,j)
__global__ void k_energy_bc_z(Geo g, double* Q, double* ER, const int* run) {
    if (skip_run(run)) return;
    int ni = g.nx + 4, nj = g.ny + 4;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj) return;
    int i = t % ni - 1, j = t / ni - 1;
    const int nz = g.nz;
    double q = Q[IX(i, j, 2)], e = ER[IX(i, j, 2)];
    Q[IX(i, j, 1)] = q; Q[IX(i, j, 0)] = q; Q[IX(i, j, -1)] = q;
    ER[IX(i, j, 1)] = e; ER[IX(i, j, 0)] = e; ER[IX(i, j, -1)] = e;
    q = Q[IX(i, j, nz - 1)];
    e = ER[IX(i, j, nz - 1)];
    Q[IX(i, j, nz)] = q; Q[IX(i, j, nz + 1)] = q; Q[IX(i, j, nz + 2)] = q;
    ER[IX(i, j, nz)] = e; ER[IX(i, j, nz + 1)] = e; ER[IX(i, j, nz + 2)] = e;
}

__global__ void k_energy_final(double* DER, double* T, const double* __restrict__ ER, const double* __restrict__ EROLD,
                               const double* __restrict__ Q, const double* __restrict__ RH,
                               const double* __restrict__ CV, long long n, const int* run) {
    if (skip_run(run)) return;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    st_nt(&DER[t], ER[t] - EROLD[t]);
    st_nt(&T[t], Q[t] / (RH[t] * CV[t]));
}

// IKAPTYP 4/5 tables on the host (log10 Planck, Rosseland, absorption;
// temperature fastest), kept across tt_init/tt_finalize like the reference's
// opc_init module; KAPPA copies them to the device when they have changed.
static double g_ftab[3 * TT_FREEDMAN_NT * TT_FREEDMAN_NP];
static int g_ftab_version = 0;  // 0: never set
// IKAPTYP 9-14 tables likewise, followed by the log10 T axis
static double g_btab[3 * TT_BURROWS_NT * TT_BURROWS_NR + TT_BURROWS_NT];
static int g_btab_version = 0;
// IKAPTYP 20-25 irradiation bins (log10, T fastest, then rho, then bin), the
// bins' stellar flux and the dilution
static double g_wtab[TT_NWAVE * TT_BURROWS_NT * TT_BURROWS_NR];
static double g_bstar[TT_NWAVE];
static double g_wdil = 0.0;
static int g_wtab_version = 0;
// IKAPTYP 6/7 column-depth absorption table (42 x 18 x 8)
static double g_ctab[8 * TT_FREEDMAN_NT * TT_FREEDMAN_NP];
static int g_ctab_version = 0;

// a table followed by 10^ of each entry, the powers from the host's libm
// (the reference's 10.d0**x): the derivatives difference neighbouring 10^
// values, which can cancel, so they are formed from the same powers the
// reference forms
static std::vector<double> with_pow10(const double* a, int n) {
    std::vector<double> t(2 * (size_t)n);
    for (int q = 0; q < n; ++q) {
        t[q] = a[q];
        t[n + q] = std::pow(10.0, a[q]);
    }
    return t;
}

__global__ void k_clear_bad_kappa(Ctrl* c) {
    if (threadIdx.x == 0) c->bad_kappa = 0;
}

// IKAPTYP 4-7 and 9-12, 14 interpolate tables, whose kernels flag states the
// reference stops on (ctrl->bad_kappa)
static bool table_kappa() {
    const int t = L.p.ikaptyp;
    return (t >= 4 && t <= 7) || (t >= 9 && t <= 12) || t == 14 || (t >= 20 && t <= 25);
}

// the Burrows tables on the device, refreshed when the host's have changed
static int btab_sync() {
    if (L.btab_version != g_btab_version) {
        const std::vector<double> t = with_pow10(g_btab, 3 * TT_BURROWS_NT * TT_BURROWS_NR + TT_BURROWS_NT);
        TT_HIP(hipMemcpyAsync(L.btab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, L.s));
        TT_HIP(hipStreamSynchronize(L.s));  // t is released after this call
        L.btab_version = g_btab_version;
    }
    return 0;
}

int burrows_cell111_impl() {
    if (g_btab_version == 0)
        return fail(1, "BOUNDXMIN(9) at III <= 0 evaluates BurrowsOPC: call tt_set_burrows_table first");
    int rc = btab_sync();
    if (rc) return rc;
    double** f = L.f;
    k_clear_bad_kappa<<<1, 64, 0, L.s>>>(L.ctrl);
    TT_LAUNCH();
    k_bur_cell111<<<1, 64, 0, L.s>>>(L.g, f[TT_T], f[TT_RH], L.btab, f[TT_XKAPR], f[TT_XKAPP], f[TT_XKAPA],
                                     f[TT_DXKAPP_DT], f[TT_DXKAPA_DT], L.p.fbottom, L.ctrl);
    TT_LAUNCH();
    return 0;
}

// Q: also form Q = RH*CV*T where the opacity kernel passes over every cell
// anyway (the power law and the constant); *q_done tells the caller
static int kappa_impl(double* Q = nullptr, bool* q_done = nullptr) {
    if (q_done) *q_done = false;
    if (table_kappa()) {
        k_clear_bad_kappa<<<1, 64, 0, L.s>>>(L.ctrl);
        TT_LAUNCH();
    }
    double sig;
    if (L.p.isigtyp == 0) sig = 0.0;
    else if (L.p.isigtyp == 1) sig = 0.4;
    else return fail(1, "SUBSIG: ISIGTYP must be 0 or 1");
    double** f = L.f;
    if (L.p.ikaptyp >= 4 && L.p.ikaptyp <= 7) {
        if (g_ftab_version == 0) return fail(1, "KAPPA: IKAPTYP 4-7 need tt_set_opacity_table first");
        const bool colmode = L.p.ikaptyp >= 6;
        if (colmode && g_ctab_version == 0) return fail(1, "KAPPA: IKAPTYP 6 and 7 need tt_set_column_table first");
        if (colmode && L.ctab_version != g_ctab_version) {
            const std::vector<double> t = with_pow10(g_ctab, 8 * TT_FREEDMAN_NT * TT_FREEDMAN_NP);
            TT_HIP(hipMemcpyAsync(L.ctab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, L.s));
            TT_HIP(hipStreamSynchronize(L.s));
            L.ctab_version = g_ctab_version;
        }
        if (L.ftab_version != g_ftab_version) {
            const std::vector<double> t = with_pow10(g_ftab, 3 * TT_FREEDMAN_NT * TT_FREEDMAN_NP);
            TT_HIP(hipMemcpyAsync(L.ftab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, L.s));
            TT_HIP(hipStreamSynchronize(L.s));  // t is released after this call
            L.ftab_version = g_ftab_version;
        }
        int rc = pressg_impl();  // pprphy.f:88
        if (rc) return rc;
        // XMUE = mu_gas (GASINIT, ppinit.f:1414)
        k_kappa_tab<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.g, met(), L.ub, f[TT_T], f[TT_PG], f[TT_RH], L.ftab,
                                                       colmode ? L.ctab : nullptr, std::log10(L.p.mu_gas),
                                                       f[TT_XKAPR], f[TT_XKAPP], f[TT_XKAPA], f[TT_DXKAPP_DT],
                                                       f[TT_DXKAPA_DT], f[TT_SIG], sig, L.ctrl);
        TT_LAUNCH();
        return 0;
    }
    const bool wave = L.p.ikaptyp >= 20 && L.p.ikaptyp <= 25;
    if ((L.p.ikaptyp >= 9 && L.p.ikaptyp <= 12) || L.p.ikaptyp == 14 || wave) {
        if (g_btab_version == 0) return fail(1, "KAPPA: IKAPTYP 9-14 and 20-25 need tt_set_burrows_table first");
        int rc = btab_sync();
        if (rc) return rc;
        k_kappa_bur<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.g, L.ub, f[TT_T], f[TT_RH], L.btab, f[TT_XKAPR],
                                                       f[TT_XKAPP], f[TT_XKAPA], f[TT_DXKAPP_DT],
                                                       f[TT_DXKAPA_DT], f[TT_SIG], sig, wave ? 1 : 0, L.ctrl);
        TT_LAUNCH();
        return 0;
    }
    if (L.p.ikaptyp == 8) {
        k_kappa_hminus<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.g, L.ub, f[TT_RH], f[TT_T], f[TT_XKAPR], f[TT_XKAPP],
                                                           f[TT_XKAPA], f[TT_DXKAPP_DT], f[TT_DXKAPA_DT],
                                                           f[TT_SIG], sig);
        TT_LAUNCH();
        return 0;
    }
    if (L.p.ikaptyp == 13) {  // constants (pprphy.f:235-241); gamma_OPC**4.d0 from the host's libm
        const double xa = std::pow(L.p.gamma_opc, 4.0) * L.p.zkapmult;
        const double v[6] = {L.p.zkapmult, L.p.zkapmult, xa, 0.0, 0.0, sig};
        const int id[6] = {TT_XKAPR, TT_XKAPP, TT_XKAPA, TT_DXKAPP_DT, TT_DXKAPA_DT, TT_SIG};
        for (int q = 0; q < 6; ++q) {
            k_fill_box<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(f[id[q]], v[q], L.g.n3);
            TT_LAUNCH();
        }
        return 0;
    }
    if (L.p.ikaptyp != 1 && L.p.ikaptyp != 26) return fail(1, "KAPPA: IKAPTYP must be 1, 4-14 or 26");
    k_kappa<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(L.g, L.p, L.ub, f[TT_RH], f[TT_T], f[TT_XKAPR], f[TT_XKAPP],
                                                f[TT_XKAPA], f[TT_DXKAPP_DT], f[TT_DXKAPA_DT], f[TT_SIG], sig, Q,
                                                f[TT_CV]);
    TT_LAUNCH();
    if (q_done) *q_done = Q != nullptr;
    return 0;
}

#ifndef STAU_COLS
#define STAU_COLS 4
#endif

static int stellar_impl() {
    double** f = L.f;
    if (L.p.f_incident == 0) {
        k_fill_box<<<nblk(L.g.n3, 256), 256, 0, L.s>>>(f[TT_STELLARINPUT], 0.0, L.g.n3);
        TT_LAUNCH();
        return 0;
    }
    if (L.p.f_incident != 2) return fail(1, "STELLARINPUT_HEATING: F_INCIDENT must be 0 or 2");
    int ncol = (L.g.ny + 4) * (L.g.nz + 4);
    const int ik = L.p.ikaptyp;
    if (ik == 20 || ik == 21 || ik == 23 || ik == 24) {  // the wavelength-binned branch (pprad.f:429-447)
        if (g_wtab_version == 0)
            return fail(1, "STELLARINPUT_HEATING: IKAPTYP 20-24 need tt_set_burrows_wave_table first");
        if (L.wtab_version != g_wtab_version) {
            TT_HIP(hipMemcpyAsync(L.wtab, g_wtab, sizeof g_wtab, hipMemcpyHostToDevice, L.s));
            TT_HIP(hipStreamSynchronize(L.s));
            L.wtab_version = g_wtab_version;
        }
        const size_t lds = (size_t)TT_NWAVE * (L.g.nx + 2) * sizeof(double);
        if (lds > 64 * 1024) return fail(1, "STELLARINPUT_HEATING: the wavelength branch needs nx <= 270");
        WaveArgs w;
        for (int l = 0; l < TT_NWAVE; ++l) w.bstar[l] = g_bstar[l];
        w.dil = g_wdil;
        k_sin_wave<<<ncol, 64, lds, L.s>>>(L.g, met(), L.ub, f[TT_T], f[TT_RH], L.wtab, L.mustar, w,
                                          f[TT_STELLARINPUT], L.ctrl);
        TT_LAUNCH();
        return 0;
    }
    const size_t lds = (size_t)(L.g.nx + 2) * sizeof(double);
    if (lds > 64 * 1024) return fail(1, "STELLARINPUT_HEATING: nx too large for the column kernel");
    double* coef = L.sincoef;  // mu_star may have been uploaded since the last call: formed each call
    k_sin_coef<<<nblk(ncol, 256), 256, 0, L.s>>>(L.g, met(), L.p.tirr, L.mustar, coef);
    TT_LAUNCH();
    constexpr int SC = STAU_COLS;
    const int nch = (L.g.nx + 2 + 63) / 64;
    const size_t ldsc = (size_t)SC * (L.g.nx + 3) * sizeof(double);
    const unsigned nbc = (unsigned)((ncol + SC - 1) / SC);
#define TT_STAU_C(N)                                                                                         \
    case N:                                                                                                  \
        k_stau_sin_c<SC, N><<<nbc, 64, ldsc, L.s>>>(L.g, met(), L.ub, f[TT_XKAPA], f[TT_RH], L.mustar, coef, \
                                                   f[TT_STELLARTAU], f[TT_DTAUS], f[TT_STELLARINPUT]);      \
        break;
    switch (nch) {
        TT_STAU_C(1) TT_STAU_C(2) TT_STAU_C(3) TT_STAU_C(4) TT_STAU_C(5) TT_STAU_C(6) TT_STAU_C(7) TT_STAU_C(8)
        default:  // nx > 510: one column per wavefront
            k_stau_sin<<<ncol, 64, lds, L.s>>>(L.g, met(), L.ub, f[TT_XKAPA], f[TT_RH], L.mustar, coef,
                                               f[TT_STELLARTAU], f[TT_DTAUS], f[TT_STELLARINPUT]);
    }
#undef TT_STAU_C
    TT_LAUNCH();
    return 0;
}

// EVAL1..6 and EVAL_1_6 after a Newton iteration: k_newton_cell forms them
// in registers and does not store them (seven arrays fewer to write, about a
// quarter of its time).  They depend on DIFR and the metrics only, so their
// module arrays are written from those (k_evals) when something reads them
// (a download, a solve that fills its matrix from them) or before DIFR, an
// EVAL array or a metric is overwritten from the host.
static bool g_evals_lazy = false;

int evals_materialize() {
    if (!g_evals_lazy) return 0;
    double** f = L.f;
    const long long n = (long long)L.g.nx * L.g.ny * L.g.nz;
    k_evals<<<nblk(n, 256), 256, 0, L.s>>>(L.g, met(), f[TT_DIFRX], f[TT_DIFRY], f[TT_DIFRZ], f[TT_EVAL1], f[TT_EVAL2],
                                            f[TT_EVAL3], f[TT_EVAL4], f[TT_EVAL5], f[TT_EVAL6], f[TT_EVAL_1_6]);
    TT_LAUNCH();
    g_evals_lazy = false;
    return 0;
}

int evals_before_access(int field) {
    if (field == TT_DIFRX || field == TT_DIFRY || field == TT_DIFRZ || (field >= TT_EVAL1 && field <= TT_EVAL6) ||
        field == TT_EVAL_1_6)
        return evals_materialize();
    return 0;
}

void evals_reset() { g_evals_lazy = false; }

static int flux_impl() {
    if (L.p.iflimtyp != 1 && L.p.iflimtyp != 2) return fail(1, "FLIM: IFLIMTYP must be 1 or 2");
    Geo& g = L.g;
    double** f = L.f;
    FaceArgs a{f[TT_ER], f[TT_XKAPR], f[TT_SIG], {f[TT_RHQX], f[TT_RHQY], f[TT_RHQZ]},
               {f[TT_FLIMX], f[TT_FLIMY], f[TT_FLIMZ]}, {f[TT_DIFRX], f[TT_DIFRY], f[TT_DIFRZ]},
               {f[TT_FLUX_X], f[TT_FLUX_Y], f[TT_FLUX_Z]}, L.ub, L.p.iflimtyp, L.p.fbottom};
    long long n = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    k_faces<<<nblk(n, 256), 256, 0, L.s>>>(g, met(), a);
    TT_LAUNCH();
    return 0;
}

static int energy_bcs(int order) {
    Geo& g = L.g;
    double** f = L.f;
    int ncol = (g.ny + 4) * (g.nz + 4), nline = (g.nx + 4) * (g.ny + 4);
    k_energy_bc_x<<<nblk(ncol, 128), 128, 0, L.s>>>(g, met(), L.ub, f[TT_Q], f[TT_ER], f[TT_RH], f[TT_CV], f[TT_T],
                                                    f[TT_FLUX_X], f[TT_XKAPR], L.p.bcxmin, L.p.d_t, L.p.fbottom,
                                                    order, L.run_if);
    TT_LAUNCH();
    k_energy_bc_z<<<nblk(nline, 256), 256, 0, L.s>>>(g, f[TT_Q], f[TT_ER], L.run_if);
    TT_LAUNCH();
    return 0;
}

// The Newton loop's test after an update, on the device: the loop ends when
// max|dQ/Q| is not above delta_max or at the iteration cap (ppenergy.f:
// 750-791), and not where a fatal flag is set (the host stops there).
// newton_fin is written whatever the run flag says, so it never holds an
// earlier solve's value.
__global__ void k_newton_decide(Ctrl* c, double delta_max, int iter, int itmax, const int* run) {
    if (threadIdx.x) return;
    if (run && *run == 0) {
        c->newton_fin = 0;
        return;
    }
    const bool fatal = c->bad_q || c->bad_bound || c->bad_kappa;
    c->newton_fin = !fatal && (iter == itmax || !(c->maxdq > delta_max)) ? 1 : 0;
}

int energy_impl(double delt, int iii, tt_energy_stats* st) {
    Geo& g = L.g;
    double** f = L.f;
    const double delta_max = 0.005, conv_parm = (double)0.1f;  // ppenergy.f:42,49 (0.1 is a REAL literal)
    tt_energy_stats es{};
    if (L.p.bcxmin != 12 && L.p.bcxmin != 9) return fail(1, "COUPLED_ENERGY_EQS: BCXMIN must be 9 or 12");
    if (L.p.bcxmin == 9 && L.visc.ieta == 5)
        return fail(1, "COUPLED_ENERGY_EQS: BCXMIN 9 with IETA 5 is not supported");
    L.iii = iii;
    int rc;
    L.energy_iter = 0;
    if ((rc = dev_zero(f[TT_DELTA_ER], (size_t)g.n3 * sizeof(double)))) return rc;
    if ((rc = bounds_impl())) return rc;
    // Q = RH*CV*T, and KAPPA, SUBSIG, QDENS: T and RH stay fixed through the
    // Newton loop, so the opacities FLIM recomputes while energy_iter <= 1
    // are these.  The opacity pass forms Q too where it covers every cell.
    bool q_done = false;
    if ((rc = kappa_impl(f[TT_Q], &q_done))) return rc;
    if (!q_done) {
        k_q_init<<<nblk(g.n3, 256), 256, 0, L.s>>>(f[TT_Q], f[TT_RH], f[TT_CV], f[TT_T], g.n3);
        TT_LAUNCH();
    }
    // a state outside the opacity tables stops here, before the Newton loop
    // runs on partly stale opacities (pprphy.f:825-828,1040-1105 stop at once)
    if (table_kappa() && (rc = sync_check())) return rc;
    if ((rc = qdens_impl())) return rc;
    if (L.p.f_incident > 1 && (rc = stellar_impl())) return rc;
    if (L.p.iflimtyp != 1 && L.p.iflimtyp != 2) return fail(1, "FLIM: IFLIMTYP must be 1 or 2");
    k_flux_x_top<<<nblk((long long)g.ny * g.nz, 256), 256, 0, L.s>>>(g, f[TT_ER], L.ub, f[TT_FLUX_X], L.p.fbottom);
    TT_LAUNCH();
    if ((rc = energy_bcs(0))) return rc;
    if ((rc = planetshift_impl(f[TT_ER], 6))) return rc;
    if ((rc = dev_copy(f[TT_ER_OLD], f[TT_ER], (size_t)g.n3 * sizeof(double)))) return rc;
    if ((rc = dev_copy(L.VD, f[TT_Q], (size_t)g.n3 * sizeof(double)))) return rc;  // Q_old
    int srcode = 0;
    bool final_done = false;   // k_energy_final has run (queued ahead, k_newton_decide set newton_fin)
    L.trailing_done = false;
    auto final_impl = [&]() -> int {
        k_energy_final<<<nblk(g.n3, 256), 256, 0, L.s>>>(f[TT_DELTA_ER], f[TT_T], f[TT_ER], f[TT_ER_OLD], f[TT_Q],
                                                         f[TT_RH], f[TT_CV], g.n3, L.run_if);
        TT_LAUNCH();
        return 0;
    };
    int first_choice = 0;      // the first solve's start: 0 cold, 1 warm, 2 learned
    hipEvent_t* ev = L.en_ev;  // phase boundaries of one Newton iteration
    for (;;) {
        L.energy_iter++;
        TT_HIP(hipEventRecord(ev[0], L.s));
        if ((rc = planetshift_impl(f[TT_ER], 6))) return rc;
        if ((rc = flux_impl())) return rc;
        CellArgs ca{f[TT_FLUX_X], f[TT_FLUX_Y], f[TT_FLUX_Z], f[TT_DIFRX], f[TT_DIFRY], f[TT_DIFRZ], f[TT_Q], L.VD,
                    f[TT_ER], f[TT_ER_OLD], f[TT_RH], f[TT_CV], f[TT_XKAPP], f[TT_XKAPA], f[TT_DXKAPP_DT],
                    f[TT_DXKAPA_DT], f[TT_STELLARINPUT], f[TT_EVAL1], f[TT_EVAL2], f[TT_EVAL3], f[TT_EVAL4],
                    f[TT_EVAL5], f[TT_EVAL6], f[TT_EVAL7], f[TT_EVAL8], f[TT_EVAL_1_6], f[TT_JACOB_ETA], f[TT_FQ],
                    f[TT_FR], f[TT_RHS_ER], L.DFQ, nullptr, nullptr, nullptr, nullptr, 0, L.ub, L.p.f_incident,
                    delt};
        const PackedSys ps = sor_packed();
        const bool with32 = L.opts.precision == 1 && L.opts.inner == 0;  // line sweeps build their own FP32 system
        ca.P0 = ps.c64[0]; ca.P1 = ps.c64[1]; ca.hx = ps.hx;
        if (with32) { ca.Q0 = ps.c32[0]; ca.Q1 = ps.c32[1]; }
        long long nc = (long long)g.nx * g.ny * g.nz;
        k_newton_cell<<<nblk(nc, 256), 256, 0, L.s>>>(g, met(), ca);
        TT_LAUNCH();
        g_evals_lazy = true;
        sor_sys_packed_written(with32);
        TT_HIP(hipEventRecord(ev[1], L.s));
        // ADJUST_SOR_PARAM at the first Newton iteration of step 1 and of
        // every 10000th step (ppenergy.f:372-374), when switched on
        if (L.opts.adjust_omega && L.energy_iter == 1 && (iii % 10000 == 0 || iii == 1)) {
            if ((rc = sor_adjust_impl(0, nullptr, nullptr))) return rc;
        }
        // initial guess (new).  Later Newton iterations: the previous
        // correction scaled by warm_scale.  First Newton iteration: the
        // first correction of the previous step (guess 1), or the learned
        // prediction (guess 2, when a network is loaded).  When guess 2 does
        // not run the network, the first solve starts warm or cold, whichever
        // has taken less time on recent first solves (the SOR phase: the
        // guess's check and the solve; with line sweeps both starts take the
        // same iterations and the check is the difference); every 16th such
        // solve takes the other, so its average stays current.  The SOR solve
        // keeps a guess only if its residual is not above the cold start's.
        L.have_prev_guess = 0;
        if (L.opts.guess != 0) {
            if (L.energy_iter > 1) {
                if ((rc = sor_scale_guess(L.opts.warm_scale))) return rc;
                L.have_prev_guess = 1;
            } else {
                // the network's prediction, unless it was rejected on the
                // last 3 first solves: then it runs again on every 16th only
                // (inference costs about as much as 40 FP32 sweeps)
                const bool learned = L.opts.guess == 2 && cnn_ready();
                if (learned && (L.learned_rejects < 3 || L.learned_idle >= 15)) {
                    if ((rc = cnn_predict_guess(delt))) return rc;
                    L.have_prev_guess = 2;
                    L.learned_idle = 0;
                } else {
                    if (learned) {
                        es.learned_skipped++;
                        L.learned_idle++;
                    }
                    bool warm = true;
                    if (L.opts.guess == 2) {
                        const double ac = L.start_avg[0], aw = L.start_avg[1];
                        const bool prefer_warm = aw == 0.0 || (ac > 0.0 && aw <= ac);  // untried: try it
                        warm = (++L.start_picks % 16 == 0) ? !prefer_warm : prefer_warm;
                    }
                    if (warm && L.have_de1) {
                        if ((rc = dev_copy(f[TT_DELTA_ER], L.DE1, (size_t)g.n3 * sizeof(double)))) return rc;
                        L.have_prev_guess = 1;
                    }
                }
            }
        }
        if (L.energy_iter == 1) first_choice = L.have_prev_guess;
        tt_sor_stats ss{};
        // delta_Q, the damped update and the energy blocks after the solve
        auto update = [&]() -> int {
            TT_HIP(hipEventRecord(ev[2], L.s));
            // 4 blocks per CU keep the grid-stride loop's loads in flight (max
            // reductions: the result does not depend on the block count)
            const long long ncell = (long long)g.nx * g.ny * g.nz;
            int nb = (int)std::min<long long>(4096, (ncell + 255) / 256);
            UpdArgs ua{f[TT_DELTA_ER], L.DFQ, f[TT_FQ], f[TT_RH], f[TT_XKAPP], f[TT_Q], f[TT_ER], L.ub, delt, conv_parm,
                       L.p.bcxmin == 12 ? 1 : 2};
            k_dq_update<<<nb, 256, 0, L.s>>>(g, ua, L.partial, nb, L.ctrl, L.run_if);
            TT_LAUNCH();
            int r = fold_partials(L.partial, nb, 2, 1, &L.ctrl->maxdq, &L.ctrl->maxde);
            if (r) return r;
            if ((r = energy_bcs(1))) return r;
            TT_HIP(hipEventRecord(ev[3], L.s));
            return 0;
        };
        // With line sweeps the solve's one control-block read comes after
        // everything that follows it in the common case, queued ahead on
        // device flags: the result's store, the update, the Newton test
        // (k_newton_decide) and, where that test ends the loop, the final
        // T and delta_ER and tt_step's trailing BOUNDS/PRESSG.  Every kernel
        // of it returns at once where its flag is not set in stream order
        // (the solve not ended, the loop going on), and the host then queues
        // what is still to do as before.
        const bool tail = L.opts.precision == 1 && L.opts.inner == 1 && !cnn_collecting();
        if (tail)
            L.sor_tail = [&]() -> int {
                int r;
                if ((r = sor_store_result(0))) return r;
                if ((r = update())) return r;
                k_newton_decide<<<1, 64, 0, L.s>>>(L.ctrl, delta_max, L.energy_iter, L.p.itmax, L.run_if);
                TT_LAUNCH();
                const int* keep = L.run_if;
                L.run_if = &L.ctrl->newton_fin;
                r = final_impl();
                if (!r && L.step_trailing) r = L.step_trailing();
                L.run_if = keep;
                return r;
            };
        L.de1_out = L.energy_iter == 1 ? L.DE1 : nullptr;  // the first correction, for the next warm start
        int src = sor_solve_impl(0, &ss, 0);
        L.de1_out = nullptr;
        L.sor_tail = nullptr;
        if (src > 0) return src;
        if (src < 0) srcode = src;
        if (L.energy_iter == 1) {
            L.have_de1 = 1;
            if ((rc = cnn_collect_sample(delt))) return rc;  // training set of the learned guess, when collecting
        }
        es.sor_iter_total += ss.sor_iter;
        es.sor_iter_last = ss.sor_iter;
        es.fp32_sweeps += ss.fp32_sweeps;
        if (L.energy_iter == 1) {
            es.sor_iter_first = ss.sor_iter;
            es.first_start = ss.guess_used;
            es.guess_scale_first = ss.guess_scale;
            es.guess_r0_first = (L.have_prev_guess && ss.r0_cold > 0.0) ? ss.r0_guess / ss.r0_cold : 0.0;
        }
        if (L.have_prev_guess && ss.guess_used == 0) es.fallbacks++;
        if (ss.guess_used == 2) es.learned_used++;
        if (L.have_prev_guess == 2) L.learned_rejects = ss.guess_used == 2 ? 0 : L.learned_rejects + 1;
        if (tail) {
            // the solve's read came after the update (and the rest): its
            // flags are in the host copy already
            if ((rc = check_flags())) return rc;
            final_done = L.hctrl->newton_fin != 0;
        } else {
            if ((rc = update())) return rc;
            if ((rc = sync_check())) return rc;
        }
        float t01 = 0, t12 = 0, t23 = 0;
        TT_HIP(hipEventElapsedTime(&t01, ev[0], ev[1]));
        TT_HIP(hipEventElapsedTime(&t12, ev[1], ev[2]));
        TT_HIP(hipEventElapsedTime(&t23, ev[2], ev[3]));
        es.ms_build += t01;
        es.ms_sor += t12;
        es.ms_update += t23;
        if (L.energy_iter == 1) {
            es.first_choice = first_choice;
            es.ms_sor_first = t12;
            if (L.opts.guess == 2 && first_choice < 2) {  // the start the first solve was given
                double& a = L.start_avg[first_choice];
                a = a == 0.0 ? (double)t12 : 0.75 * a + 0.25 * (double)t12;
            }
        }
        es.max_dq = L.hctrl->maxdq;
        es.max_de = L.hctrl->maxde;
        es.energy_iter = L.energy_iter;
        if (L.energy_iter == L.p.itmax) {
            srcode = fail(-2, "Max iteration reached in coupled_energy_eqs");
            break;
        }
        if (!(es.max_dq > delta_max)) break;
    }
    // the final T and delta_ER, unless queued ahead and run already; then
    // tt_step's trailing work has run too, where it gave some
    if (!final_done && (rc = final_impl())) return rc;
    L.trailing_done = final_done && L.step_trailing;
    if (st) *st = es;
    return srcode;
}

}  // namespace tt

using namespace tt;

extern "C" int tt_coupled_energy(double delt, int iii, tt_energy_stats* st) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = energy_impl(delt, iii, st);
    if (rc > 0) return rc;
    int rc2 = sync_check();
    return rc2 ? rc2 : rc;
}

extern "C" int tt_set_opacity_table(const double* kap_planck, const double* kap_ross, const double* kap_abs) {
    if (!kap_planck || !kap_ross || !kap_abs) return fail(1, "tt_set_opacity_table: null table");
    const int n = TT_FREEDMAN_NT * TT_FREEDMAN_NP;
    memcpy(g_ftab, kap_planck, n * sizeof(double));
    memcpy(g_ftab + n, kap_ross, n * sizeof(double));
    memcpy(g_ftab + 2 * n, kap_abs, n * sizeof(double));
    ++g_ftab_version;
    return 0;
}

extern "C" int tt_set_burrows_table(const double* kap_planck, const double* kap_ross, const double* kap_abs) {
    if (!kap_planck || !kap_ross || !kap_abs) return fail(1, "tt_set_burrows_table: null table");
    const int n = TT_BURROWS_NT * TT_BURROWS_NR;
    memcpy(g_btab, kap_planck, n * sizeof(double));
    memcpy(g_btab + n, kap_ross, n * sizeof(double));
    memcpy(g_btab + 2 * n, kap_abs, n * sizeof(double));
    memcpy(g_btab + 3 * n, h_bt, sizeof h_bt);
    ++g_btab_version;
    return 0;
}

extern "C" int tt_set_burrows_wave_table(const double* kap_wave, const double* bstar, double dilution) {
    if (!kap_wave || !bstar) return fail(1, "tt_set_burrows_wave_table: null table");
    memcpy(g_wtab, kap_wave, sizeof g_wtab);
    memcpy(g_bstar, bstar, sizeof g_bstar);
    g_wdil = dilution;
    ++g_wtab_version;
    return 0;
}

extern "C" int tt_set_column_table(const double* kap_abs) {
    if (!kap_abs) return fail(1, "tt_set_column_table: null table");
    memcpy(g_ctab, kap_abs, sizeof g_ctab);
    ++g_ctab_version;
    return 0;
}

extern "C" int tt_kappa(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int rc = kappa_impl();
    if (rc) return rc;
    return sync_check();
}
