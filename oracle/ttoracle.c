/*
 * ttoracle.c -- CPU FP64 restatement of ttcode's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library
 * (tests, __graft_entry__.smoke) and the timed CPU baseline of bench.py.  It
 * is never linked into, loaded by or called from the product path.
 *
 * It restates, routine by routine, the reference Fortran (file:line cited at
 * each function) for one MPI rank holding the whole domain (proc_dims =
 * (1,1,1): phi periodic through self-exchange, every r and theta neighbour
 * MPI_PROC_NULL, ppshift.f:694-819).  Loop order and the order of every
 * floating-point operation follow the reference so that, compiled without FMA
 * contraction, results agree bit for bit with a flang build of the reference
 * on the same inputs (checked by tests/test_oracle_reference.py against the
 * fixtures in tests/golden/).
 *
 * Physical constants keep the values the reference actually uses: ppmod.f
 * declares them DOUBLE PRECISION but writes REAL literals (e.g. "FD =
 * 2.99792458 E 10", ppmod.f:1081), so the stored value is the nearest float.
 *
 * Storage: every 3-D array is held over the full ghosted box
 * (-1:n+2)^3 regardless of its declared bounds (a superset of them); the wire
 * format of tor_upload/tor_download is the declared bounds of
 * include/ttgpu_fields.h, as for the device library.
 *
 * OpenMP: loops whose result does not depend on order are parallel; the SOR
 * sweeps and every reduction are sequential when tor_set_threads(1) (the
 * checker setting) and parallel otherwise (the CPU-baseline setting).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ttgpu.h"
#include "../include/ttgpu_fields.h"

/* ppmod.f:1071-1093 (REAL literals stored in DOUBLE PRECISION parameters) */
static const double PI = 3.141592653589793;
#define FD ((double)2.99792458e10f)
#define ARAD ((double)7.56596e-15f)
#define RGAS ((double)8.315e7f)
#define SBCONST ((double)5.6705e-5f)

typedef struct {
    tt_params p;
    int nx, ny, nz, gx, gy, gz;
    size_t n3, n2;
    int nthreads;
    /* 3-D fields, ghosted (-1:n+2)^3 */
    double *T, *RH, *V, *G, *H, *ER, *ER_OLD, *DER, *GD, *PG, *CS, *CV;
    double *RHQX, *RHQY, *RHQZ, *XKR, *XKP, *XKA, *DKP, *DKA, *SIG;
    double *GRAVX, *GRAVY, *GRAVZ, *SIN, *STAU, *DTAUS;
    double *RHS, *SA, *SB, *SC, *SD, *SE, *SF, *SG;
    double *EV1, *EV2, *EV3, *EV4, *EV5, *EV6, *EV7, *EV8, *EV16, *JETA, *FQ, *FR;
    double *FLIMX, *FLIMY, *FLIMZ, *DIFRX, *DIFRY, *DIFRZ, *FLUXX, *FLUXY, *FLUXZ;
    double *CCX, *CCY, *CCZ, *Q;
    double *ETA, *DIVTX, *DIVTY, *DIVTZ, *DISFN;
    double *CF1, *CF2, *CF3;  /* conv_flux(1..3,:,:,:) */
    /* 2-D */
    int32_t* UB;
    double* MUSTAR;
    /* 1-D metric, index i stored at [i+2] */
    double* m[TT_M_COUNT];
    int energy_iter;
    char err[256];
} oracle_t;

static oracle_t O;

#define I3(i, j, k) ((((size_t)((k) + 1)) * (size_t)O.gy + (size_t)((j) + 1)) * (size_t)O.gx + (size_t)((i) + 1))
#define A(a, i, j, k) (a)[I3(i, j, k)]
#define I2(j, k) (((size_t)((k) + 1)) * (size_t)O.gy + (size_t)((j) + 1))
#define UBND(j, k) O.UB[I2(j, k)]
#define M(id, i) (O.m[id][(i) + 2])

#define XXA(i) M(TT_M_XXA, i)
#define XXB(i) M(TT_M_XXB, i)
#define DXA(i) M(TT_M_DXA, i)
#define DXB(i) M(TT_M_DXB, i)
#define XYA(i) M(TT_M_XYA, i)
#define XYB(i) M(TT_M_XYB, i)
#define DYA(i) M(TT_M_DYA, i)
#define DYB(i) M(TT_M_DYB, i)
#define XZA(i) M(TT_M_XZA, i)
#define XZB(i) M(TT_M_XZB, i)
#define DZA(i) M(TT_M_DZA, i)
#define DZB(i) M(TT_M_DZB, i)
#define DDX0(i) M(TT_M_DDX0, i)
#define DDX1(i) M(TT_M_DDX1, i)
#define DDY0(i) M(TT_M_DDY0, i)
#define DDY1(i) M(TT_M_DDY1, i)
#define DDZ0(i) M(TT_M_DDZ0, i)
#define DDZ1(i) M(TT_M_DDZ1, i)
#define DCY0(i) M(TT_M_DCY0, i)
#define DCY1(i) M(TT_M_DCY1, i)
#define DDX0KEP(i) M(TT_M_DDX0KEP, i)
#define DDX1KEP(i) M(TT_M_DDX1KEP, i)
#define SURXA(i) M(TT_M_SURXA, i)
#define SURXB(i) M(TT_M_SURXB, i)
#define VOLXA(i) M(TT_M_VOLXA, i)
#define VOLXB(i) M(TT_M_VOLXB, i)
#define CCXA(i) M(TT_M_CCXA, i)
#define SURYA(i) M(TT_M_SURYA, i)
#define SURYB(i) M(TT_M_SURYB, i)
#define VOLYA(i) M(TT_M_VOLYA, i)
#define VOLYB(i) M(TT_M_VOLYB, i)
#define CCYA(i) M(TT_M_CCYA, i)
#define SURZA(i) M(TT_M_SURZA, i)
#define SURZB(i) M(TT_M_SURZB, i)
#define VOLZA(i) M(TT_M_VOLZA, i)
#define VOLZB(i) M(TT_M_VOLZB, i)
#define CCZA(i) M(TT_M_CCZA, i)
#define GEOXGA(i) M(TT_M_GEOXGA, i)
#define GEOXHA(i) M(TT_M_GEOXHA, i)
#define GEOXG(i) M(TT_M_GEOXG, i)
#define GEOXH(i) M(TT_M_GEOXH, i)
#define GEOZG(i) M(TT_M_GEOZG, i)
#define GEOZGA(i) M(TT_M_GEOZGA, i)

#define NX O.nx
#define NY O.ny
#define NZ O.nz

static int seterr(int code, const char* msg)
{
    snprintf(O.err, sizeof O.err, "%s", msg);
    return code;
}

const char* tor_last_error(void) { return O.err; }

void tor_set_threads(int n)
{
    O.nthreads = n < 1 ? 1 : n;
#ifdef _OPENMP
    omp_set_num_threads(O.nthreads);
#endif
}

static int g_iii; /* III of the step loop, as BOUNDS reads it (set by tor_step, tor_coupled_energy, tor_set_iii) */

static double* field_ptr(int f)
{
    switch (f) {
    case TT_T: return O.T;
    case TT_RH: return O.RH;
    case TT_V: return O.V;
    case TT_G: return O.G;
    case TT_H: return O.H;
    case TT_ER: return O.ER;
    case TT_DELTA_ER: return O.DER;
    case TT_GD: return O.GD;
    case TT_PG: return O.PG;
    case TT_CS: return O.CS;
    case TT_CV: return O.CV;
    case TT_RHQX: return O.RHQX;
    case TT_RHQY: return O.RHQY;
    case TT_RHQZ: return O.RHQZ;
    case TT_XKAPR: return O.XKR;
    case TT_XKAPP: return O.XKP;
    case TT_XKAPA: return O.XKA;
    case TT_DXKAPP_DT: return O.DKP;
    case TT_DXKAPA_DT: return O.DKA;
    case TT_SIG: return O.SIG;
    case TT_GRAVX: return O.GRAVX;
    case TT_GRAVY: return O.GRAVY;
    case TT_GRAVZ: return O.GRAVZ;
    case TT_STELLARINPUT: return O.SIN;
    case TT_STELLARTAU: return O.STAU;
    case TT_DTAUS: return O.DTAUS;
    case TT_ER_OLD: return O.ER_OLD;
    case TT_RHS_ER: return O.RHS;
    case TT_SOR_A: return O.SA;
    case TT_SOR_B: return O.SB;
    case TT_SOR_C: return O.SC;
    case TT_SOR_D: return O.SD;
    case TT_SOR_E: return O.SE;
    case TT_SOR_F: return O.SF;
    case TT_SOR_G: return O.SG;
    case TT_EVAL1: return O.EV1;
    case TT_EVAL2: return O.EV2;
    case TT_EVAL3: return O.EV3;
    case TT_EVAL4: return O.EV4;
    case TT_EVAL5: return O.EV5;
    case TT_EVAL6: return O.EV6;
    case TT_EVAL7: return O.EV7;
    case TT_EVAL8: return O.EV8;
    case TT_EVAL_1_6: return O.EV16;
    case TT_JACOB_ETA: return O.JETA;
    case TT_FQ: return O.FQ;
    case TT_FR: return O.FR;
    case TT_FLIMX: return O.FLIMX;
    case TT_FLIMY: return O.FLIMY;
    case TT_FLIMZ: return O.FLIMZ;
    case TT_DIFRX: return O.DIFRX;
    case TT_DIFRY: return O.DIFRY;
    case TT_DIFRZ: return O.DIFRZ;
    case TT_FLUX_X: return O.FLUXX;
    case TT_FLUX_Y: return O.FLUXY;
    case TT_FLUX_Z: return O.FLUXZ;
    case TT_CENTCORX: return O.CCX;
    case TT_ETA: return O.ETA;
    case TT_DIVTX: return O.DIVTX;
    case TT_DIVTY: return O.DIVTY;
    case TT_DIVTZ: return O.DIVTZ;
    case TT_DISFN: return O.DISFN;
    case TT_CONV_FLUX1: return O.CF1;
    case TT_CONV_FLUX2: return O.CF2;
    case TT_CONV_FLUX3: return O.CF3;
    case TT_CENTCORY: return O.CCY;
    case TT_CENTCORZ: return O.CCZ;
    case TT_Q: return O.Q;
    case TT_MU_STAR: return O.MUSTAR;
    default: return NULL;
    }
}

int tor_field_info(int f, int32_t lo[3], int32_t hi[3], int32_t* rank, int32_t* is_int)
{
    return tt_field_bounds(f, NX, NY, NZ, lo, hi, rank, is_int);
}

int tor_metric_info(int m, int32_t* lo, int32_t* hi) { return tt_metric_bounds(m, NX, NY, NZ, lo, hi); }

/* Copy between the ghosted internal box and the declared-bounds wire format. */
static int xfer(int f, void* host, int to_host)
{
    int32_t lo[3], hi[3], rank, is_int;
    if (tt_field_bounds(f, NX, NY, NZ, lo, hi, &rank, &is_int)) return seterr(1, "unknown field");
    if (rank == 2) {
        size_t n = 0;
        for (int k = lo[1]; k <= hi[1]; ++k)
            for (int j = lo[0]; j <= hi[0]; ++j, ++n) {
                if (is_int) {
                    int32_t* h = (int32_t*)host;
                    if (to_host) h[n] = UBND(j, k); else UBND(j, k) = h[n];
                } else {
                    double* h = (double*)host;
                    if (to_host) h[n] = O.MUSTAR[I2(j, k)]; else O.MUSTAR[I2(j, k)] = h[n];
                }
            }
        return 0;
    }
    double* a = field_ptr(f);
    double* h = (double*)host;
    if (!a) return seterr(1, "field not held by the oracle");
    size_t n = 0;
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i, ++n) {
                if (to_host) h[n] = A(a, i, j, k); else A(a, i, j, k) = h[n];
            }
    return 0;
}

int tor_upload(int f, const void* host) { return xfer(f, (void*)host, 0); }
int tor_download(int f, void* host) { return xfer(f, host, 1); }

int tor_upload_metric(int m, const double* host)
{
    int32_t lo, hi;
    if (tt_metric_bounds(m, NX, NY, NZ, &lo, &hi)) return seterr(1, "unknown metric");
    for (int i = lo; i <= hi; ++i) M(m, i) = host[i - lo];
    return 0;
}

int tor_download_metric(int m, double* host)
{
    int32_t lo, hi;
    if (tt_metric_bounds(m, NX, NY, NZ, &lo, &hi)) return seterr(1, "unknown metric");
    for (int i = lo; i <= hi; ++i) host[i - lo] = M(m, i);
    return 0;
}

/* ---------------------------------------------------------------- GITTER
 * ppinit.f:1028-1406, NCOSYS=2, uniform grids (DXMIN=DYMIN=DZMIN=0),
 * proc_coords = 0, proc_dims = 1. */
static void gitter(void)
{
    const tt_params* p = &O.p;
    double YMAX1 = p->ymax * 2.0 * PI, YMIN1 = p->ymin * 2.0 * PI;
    double ZMAX1 = (p->zmax / 2.0) * PI, ZMIN1 = (p->zmin / 2.0) * PI;
    double DELX, DELY, DELZ, DRAN3, XXAN3;
    int i;
    /* x */
    DELX = (p->xmax - p->xmin) / (double)NX;
    for (i = 1; i <= NX + 2; ++i) DXA(i) = DELX;
    XXA(1) = p->xmin + (0.0 * (p->xmax - p->xmin) / 1.0);
    XXA(0) = XXA(1) - DXA(1);
    for (i = 2; i <= NX + 2; ++i) XXA(i) = XXA(i - 1) + DXA(i);
    for (i = 0; i <= NX + 1; ++i) XXB(i) = (XXA(i) + XXA(i + 1)) / 2.0;
    DRAN3 = DXA(NX + 2) * 1.0;
    XXAN3 = XXA(NX + 2) + DRAN3;
    XXB(NX + 2) = (XXA(NX + 2) + XXAN3) / 2.0;
    DRAN3 = DXA(1) / 1.0;
    XXAN3 = XXA(0) - DRAN3;
    XXB(-1) = (XXA(0) + XXAN3) / 2.0;
    /* y */
    DELY = (YMAX1 - YMIN1) / (double)NY;
    for (i = 1; i <= NY + 2; ++i) DYA(i) = DELY;
    XYA(1) = YMIN1 + (0.0 * (YMAX1 - YMIN1) / 1.0);
    XYA(0) = XYA(1) - DYA(1);
    for (i = 2; i <= NY + 2; ++i) XYA(i) = XYA(i - 1) + DYA(i);
    for (i = 0; i <= NY + 1; ++i) XYB(i) = (XYA(i) + XYA(i + 1)) / 2.0;
    DRAN3 = DYA(NY + 2) * 1.0;
    XXAN3 = XYA(NY + 2) + DRAN3;
    XYB(NY + 2) = (XYA(NY + 2) + XXAN3) / 2.0;
    DRAN3 = DYA(1) / 1.0;
    XXAN3 = XYA(0) - DRAN3;
    XYB(-1) = (XYA(0) + XXAN3) / 2.0;
    /* z */
    DELZ = (ZMAX1 - ZMIN1) / (double)NZ;
    for (i = 1; i <= NZ + 2; ++i) DZA(i) = DELZ;
    XZA(1) = ZMIN1 + (0.0 * (ZMAX1 - ZMIN1) / 1.0);
    XZA(0) = XZA(1) - DZA(1);
    for (i = 2; i <= NZ + 2; ++i) XZA(i) = XZA(i - 1) + DZA(i);
    for (i = 0; i <= NZ + 1; ++i) XZB(i) = (XZA(i) + XZA(i + 1)) / 2.0;
    DRAN3 = DZA(NZ + 2) * 1.0;
    XXAN3 = XZA(NZ + 2) + DRAN3;
    XZB(NZ + 2) = (XZA(NZ + 2) + XXAN3) / 2.0;
    DRAN3 = DZA(1) / 1.0;
    XXAN3 = XZA(0) - DRAN3;
    XZB(-1) = (XZA(0) + XXAN3) / 2.0;
    for (i = 0; i <= NX + 2; ++i) DXB(i) = XXB(i) - XXB(i - 1);
    for (i = 0; i <= NY + 2; ++i) DYB(i) = XYB(i) - XYB(i - 1);
    for (i = 0; i <= NZ + 2; ++i) DZB(i) = XZB(i) - XZB(i - 1);
    for (i = 0; i <= NX + 2; ++i) { DDX0(i) = 0.5; DDX1(i) = 0.5; }
    for (i = 0; i <= NY + 2; ++i) { DDY0(i) = 0.5; DDY1(i) = 0.5; }
    for (i = 1; i <= NY + 2; ++i) { DCY0(i) = 0.5; DCY1(i) = 0.5; }
    for (i = 0; i <= NZ + 2; ++i) { DDZ0(i) = 0.5; DDZ1(i) = 0.5; }
    for (i = 1; i <= NX + 2; ++i) {
        DDX0KEP(i) = DDX0(i) * pow(XXB(i - 1), 1.5);
        DDX1KEP(i) = DDX1(i) * pow(XXB(i), 1.5);
    }
    double xxa1 = XXA(1);
    if (xxa1 == 0.0) xxa1 = (double)1.e-20f;
    for (i = 0; i <= NX + 1; ++i)
        VOLXB(i) = (XXA(i + 1) * XXA(i + 1) * XXA(i + 1) - XXA(i) * XXA(i) * XXA(i)) / 3.0;
    for (i = 0; i <= NX + 2; ++i) {
        SURXA(i) = XXA(i) * XXA(i);
        CCXA(i) = XXA(i);
        VOLXA(i) = (XXB(i) * XXB(i) * XXB(i) - XXB(i - 1) * XXB(i - 1) * XXB(i - 1)) / 3.0;
        M(TT_M_SURXB, i) = XXB(i) * XXB(i);
        M(TT_M_CCXB, i) = XXB(i);
        GEOXGA(i) = XXA(i);
        GEOXHA(i) = XXA(i);
    }
    GEOXGA(1) = xxa1;
    GEOXHA(1) = xxa1;
    for (i = -1; i <= NX + 2; ++i) { GEOXG(i) = XXB(i); GEOXH(i) = XXB(i); }
    for (i = 0; i <= NY + 1; ++i) VOLYB(i) = XYA(i + 1) - XYA(i);
    for (i = 0; i <= NY + 2; ++i) {
        SURYA(i) = 1.0;
        CCYA(i) = 0.0;
        VOLYA(i) = XYB(i) - XYB(i - 1);
        M(TT_M_SURYB, i) = 1.0;
        M(TT_M_CCYB, i) = 0.0;
    }
    for (i = 0; i <= NZ + 1; ++i) VOLZB(i) = sin(XZA(i + 1)) - sin(XZA(i));
    for (i = 0; i <= NZ + 2; ++i) {
        SURZA(i) = cos(XZA(i));
        CCZA(i) = sin(XZA(i)) / 2.0;
        VOLZA(i) = fabs(sin(XZB(i)) - sin(XZB(i - 1)));
        SURZB(i) = cos(XZB(i));
        M(TT_M_CCZB, i) = sin(XZB(i)) / 2.0;
        GEOZGA(i) = cos(XZA(i));
    }
    for (i = -1; i <= NZ + 2; ++i) GEOZG(i) = cos(XZB(i));
    if (GEOZGA(NZ + 1) == 0.0) GEOZGA(NZ + 1) = (double)1.e-10f;
    if (GEOZGA(1) == 0.0) GEOZGA(1) = (double)1.e-10f;
}

/* CALCMUSTAR, ppsub.f:156-179 */
static void calcmustar(void)
{
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            if (O.p.f_incident == 2 || O.p.f_incident == 3 || O.p.f_incident == 8) {
                double a = cos(XYB(j)) * cos(XZB(k));
                double lim = (double)1.e-7f;
                O.MUSTAR[I2(j, k)] = a > lim ? a : lim;
            } else {
                O.MUSTAR[I2(j, k)] = 1.0;
            }
        }
}

static double* alloc3(void) { return (double*)calloc(O.n3, sizeof(double)); }

void tor_free(void)
{
    double** f3[] = {&O.T, &O.RH, &O.V, &O.G, &O.H, &O.ER, &O.ER_OLD, &O.DER, &O.GD, &O.PG, &O.CS,
                     &O.CV, &O.RHQX, &O.RHQY, &O.RHQZ, &O.XKR, &O.XKP, &O.XKA, &O.DKP, &O.DKA,
                     &O.SIG, &O.GRAVX, &O.GRAVY, &O.GRAVZ, &O.SIN, &O.STAU, &O.DTAUS, &O.RHS,
                     &O.SA, &O.SB, &O.SC, &O.SD, &O.SE, &O.SF, &O.SG, &O.EV1, &O.EV2, &O.EV3,
                     &O.EV4, &O.EV5, &O.EV6, &O.EV7, &O.EV8, &O.EV16, &O.JETA, &O.FQ, &O.FR,
                     &O.FLIMX, &O.FLIMY, &O.FLIMZ, &O.DIFRX, &O.DIFRY, &O.DIFRZ, &O.FLUXX,
                     &O.FLUXY, &O.FLUXZ, &O.CCX, &O.CCY, &O.CCZ, &O.Q, &O.ETA, &O.DIVTX,
                     &O.DIVTY, &O.DIVTZ, &O.DISFN, &O.CF1, &O.CF2, &O.CF3};
    for (size_t i = 0; i < sizeof f3 / sizeof f3[0]; ++i) { free(*f3[i]); *f3[i] = NULL; }
    free(O.UB); O.UB = NULL;
    free(O.MUSTAR); O.MUSTAR = NULL;
    for (int m = 0; m < TT_M_COUNT; ++m) { free(O.m[m]); O.m[m] = NULL; }
}

static int g_adjust; /* ADJUST_SOR_PARAM inside COUPLED_ENERGY_EQS (off by default) */
static tt_limits g_lim; /* clamps of the step loop (off by default) */
static tt_viscosity g_visc; /* viscosity in FORGAS3D (off by default) */

/* NR_ARTIFICIAL_VISC's module state (artvis_var_init, ppmod.f:388-395): the
 * compression fluxes, allocated at first use and zeroed (the reference leaves
 * them undefined; entries one call does not write are read by the next where
 * upperbnd varies), and dt_artvis, DTMAX until the first call (pp.f:439) */
static double *g_qvx, *g_qvy, *g_qvz;
static double g_dt_artvis;

int tor_set_viscosity(const tt_viscosity* v)
{
    if (v->ivis > 0) {
        if (v->ivistyp != 0 && v->ivistyp != 1) return seterr(1, "VISC: IVISTYP must be 0 or 1");
        if (v->ieta < 0 || v->ieta == 1 || v->ieta > 5) return seterr(1, "ETACALC: IETA must be 0, 2, 3, 4 or 5");
    }
    if (v->iartvis > 1) return seterr(1, "unknown IARTVIS"); /* ppfor.f:388-390 */
    if (v->iartvis > 0 && g_visc.iartvis <= 0) g_dt_artvis = O.p.dtmax;
    g_visc = *v;
    return 0;
}

int tor_get_dt_artvis(double* dt)
{
    *dt = g_visc.iartvis > 0 ? g_dt_artvis : O.p.dtmax;
    return 0;
}

int tor_get_viscosity(tt_viscosity* v)
{
    *v = g_visc;
    return 0;
}

int tor_init(const tt_params* p)
{
    tor_free();
    g_adjust = 0;
    memset(&g_lim, 0, sizeof g_lim);
    memset(&g_visc, 0, sizeof g_visc);
    g_iii = 0;
    free(g_qvx); free(g_qvy); free(g_qvz);
    g_qvx = g_qvy = g_qvz = NULL;
    int nt = O.nthreads ? O.nthreads : 1;
    memset(&O, 0, sizeof O);
    O.nthreads = nt;
    O.p = *p;
    O.nx = p->nx; O.ny = p->ny; O.nz = p->nz;
    if (O.nx < 4 || O.ny < 4 || O.nz < 4) return seterr(1, "nx, ny, nz must be >= 4");
    O.gx = O.nx + 4; O.gy = O.ny + 4; O.gz = O.nz + 4;
    O.n3 = (size_t)O.gx * O.gy * O.gz;
    O.n2 = (size_t)O.gy * O.gz;
    double** f3[] = {&O.T, &O.RH, &O.V, &O.G, &O.H, &O.ER, &O.ER_OLD, &O.DER, &O.GD, &O.PG, &O.CS,
                     &O.CV, &O.RHQX, &O.RHQY, &O.RHQZ, &O.XKR, &O.XKP, &O.XKA, &O.DKP, &O.DKA,
                     &O.SIG, &O.GRAVX, &O.GRAVY, &O.GRAVZ, &O.SIN, &O.STAU, &O.DTAUS, &O.RHS,
                     &O.SA, &O.SB, &O.SC, &O.SD, &O.SE, &O.SF, &O.SG, &O.EV1, &O.EV2, &O.EV3,
                     &O.EV4, &O.EV5, &O.EV6, &O.EV7, &O.EV8, &O.EV16, &O.JETA, &O.FQ, &O.FR,
                     &O.FLIMX, &O.FLIMY, &O.FLIMZ, &O.DIFRX, &O.DIFRY, &O.DIFRZ, &O.FLUXX,
                     &O.FLUXY, &O.FLUXZ, &O.CCX, &O.CCY, &O.CCZ, &O.Q, &O.ETA, &O.DIVTX,
                     &O.DIVTY, &O.DIVTZ, &O.DISFN, &O.CF1, &O.CF2, &O.CF3};
    for (size_t i = 0; i < sizeof f3 / sizeof f3[0]; ++i) {
        *f3[i] = alloc3();
        if (!*f3[i]) return seterr(1, "out of memory");
    }
    O.UB = (int32_t*)calloc(O.n2, sizeof(int32_t));
    O.MUSTAR = (double*)calloc(O.n2, sizeof(double));
    int nmax = O.nx > O.ny ? O.nx : O.ny;
    if (O.nz > nmax) nmax = O.nz;
    for (int m = 0; m < TT_M_COUNT; ++m) O.m[m] = (double*)calloc((size_t)nmax + 6, sizeof(double));
    gitter();
    calcmustar();
    /* GASINIT, ppinit.f:1414-1415 */
    double cv = RGAS / (p->mu_gas * (p->gamma - 1.0));
    for (size_t n = 0; n < O.n3; ++n) O.CV[n] = cv;
    /* FGRAV, IGRAV 4 with NCOSYS 2 (ppfor.f:753-766): the point mass
       MPLANET = MSOL MSTAR2 (GRAV, MSOL REAL literals, ppmod.f:1075,1079) */
    if (p->igrav == 4) {
        const double grav = (double)6.672e-8f, msol = (double)1.989e33f;
        const double mplanet = msol * p->mstar2;
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j)
                for (int i = -1; i <= NX + 2; ++i) {
                    const double x = i == -1 ? XXA(0) - DXA(1) : XXA(i); /* xxa(-1) is not defined */
                    A(O.GRAVX, i, j, k) = -grav * mplanet / (x * x);
                }
    }
    return 0;
}

/* --------------------------------------------------------------- ppphy.f */
int tor_pressg(void) /* ppphy.f:1-45, IGASTYP=1 */
{
    const double mu = O.p.mu_gas;
#pragma omp parallel for
    for (long n = 0; n < (long)O.n3; ++n) O.PG[n] = RGAS * O.T[n] * O.RH[n] / mu;
    return 0;
}

int tor_qdens(void) /* ppphy.f:47-76 */
{
#pragma omp parallel for
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j)
            for (int i = 0; i <= NX + 2; ++i)
                A(O.RHQX, i, j, k) = DDX1(i) * A(O.RH, i, j, k) + DDX0(i) * A(O.RH, i - 1, j, k);
#pragma omp parallel for
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = 0; j <= NY + 2; ++j)
            for (int i = -1; i <= NX + 2; ++i)
                A(O.RHQY, i, j, k) = DDY1(j) * A(O.RH, i, j, k) + DDY0(j) * A(O.RH, i, j - 1, k);
#pragma omp parallel for
    for (int k = 0; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j)
            for (int i = -1; i <= NX + 2; ++i)
                A(O.RHQZ, i, j, k) = DDZ1(k) * A(O.RH, i, j, k) + DDZ0(k) * A(O.RH, i, j, k - 1);
    return 0;
}

int tor_sound(void) /* ppphy.f:79-123, ICSTYP=0 */
{
    const double g = O.p.gamma;
#pragma omp parallel for
    for (long n = 0; n < (long)O.n3; ++n) O.CS[n] = sqrt(g * O.PG[n] / O.RH[n]);
    return 0;
}

/* ------------------------------------------------------------- ppshift.f
 * PLANETSHIFT (ppshift.f:1009-1214) with proc_dims = (1,1,1): the phi
 * neighbours are this rank (periodic), every neighbour that also steps in
 * theta is MPI_PROC_NULL, so only the "from lower" and "from upper" slabs
 * move: j=0,-1 <- ny,ny-1 and j=ny+1,ny+2 <- 1,2, for k=1..nz and
 * i=low..nx+2.  f = 3 (V) starts at i=0; f = 4, 11 (G, GD) have no j=-1. */
static void planetshift(double* a, int f)
{
    int low = (f == 3) ? 0 : -1;
    for (int k = 1; k <= NZ; ++k)
        for (int i = low; i <= NX + 2; ++i) A(a, i, 0, k) = A(a, i, NY, k);
    if (f != 4 && f != 11)
        for (int k = 1; k <= NZ; ++k)
            for (int i = low; i <= NX + 2; ++i) A(a, i, -1, k) = A(a, i, NY - 1, k);
    for (int k = 1; k <= NZ; ++k)
        for (int i = low; i <= NX + 2; ++i) A(a, i, NY + 1, k) = A(a, i, 1, k);
    for (int k = 1; k <= NZ; ++k)
        for (int i = low; i <= NX + 2; ++i) A(a, i, NY + 2, k) = A(a, i, 2, k);
}

/* PASSUPPERBND, ppbou.f:522-682 for one rank: phi slabs wrap, theta and
 * corner partners are MPI_PROC_NULL and take the k=1 / k=nz values. */
static void passupperbnd(void)
{
    for (int k = 1; k <= NZ; ++k) {
        UBND(0, k) = UBND(NY, k);
        UBND(-1, k) = UBND(NY - 1, k);
        UBND(NY + 1, k) = UBND(1, k);
        UBND(NY + 2, k) = UBND(2, k);
    }
    /* lowerbelow NULL */
    UBND(0, 0) = UBND(0, 1); UBND(0, -1) = UBND(0, 1);
    UBND(-1, 0) = UBND(-1, 1); UBND(-1, -1) = UBND(-1, 1);
    /* lowerabove NULL */
    UBND(0, NZ + 1) = UBND(0, NZ); UBND(0, NZ + 2) = UBND(0, NZ);
    UBND(-1, NZ + 1) = UBND(-1, NZ); UBND(-1, NZ + 2) = UBND(-1, NZ);
    /* above NULL */
    for (int j = 1; j <= NY; ++j) { UBND(j, NZ + 1) = UBND(j, NZ); UBND(j, NZ + 2) = UBND(j, NZ); }
    /* upperabove NULL */
    UBND(NY + 1, NZ + 1) = UBND(NY + 1, NZ); UBND(NY + 1, NZ + 2) = UBND(NY + 1, NZ);
    UBND(NY + 2, NZ + 1) = UBND(NY + 2, NZ); UBND(NY + 2, NZ + 2) = UBND(NY + 2, NZ);
    /* upperbelow NULL */
    UBND(NY + 1, 0) = UBND(NY + 1, 1); UBND(NY + 1, -1) = UBND(NY + 1, 1);
    UBND(NY + 2, 0) = UBND(NY + 2, 1); UBND(NY + 2, -1) = UBND(NY + 2, 1);
    /* below NULL */
    for (int j = 1; j <= NY; ++j) { UBND(j, 0) = UBND(j, 1); UBND(j, -1) = UBND(j, 1); }
}

/* --------------------------------------------------------------- ppbou.f */
static int burrows_opc(double t_in, double rh_in, double* bkp, double* bkr, double* bka, double* dkp,
                       double* dka);
static int g_btab_set;

/* ppbou.f:734-1148: CASE(12) :1118-1139 (T ghosts from the fixed gradient
   D_T) and CASE(9) :1029-1066 (the gradient that carries FBottom through the
   bottom by radiative diffusion); then V = 0 and stress-free G, H for both */
static int boundxmin(int type, int iii)
{
    if (type != 12 && type != 9) return seterr(1, "BOUNDXMIN: only types 9 and 12 are supported");
    if (type == 9 && g_visc.ieta == 5) return seterr(1, "BOUNDXMIN(9) with IETA 5 is not supported");
    const double D_T = O.p.d_t, fb = O.p.fbottom;
    if (type == 9 && iii <= 0) {
        /* III = 0 (set-up): BurrowsOPC at cell (1,1,1) (writing its opacities
           there), whose gradient every column then takes */
        if (!g_btab_set) return seterr(1, "BOUNDXMIN(9) at III <= 0 evaluates BurrowsOPC: set a Burrows table");
        if (burrows_opc(A(O.T, 1, 1, 1), A(O.RH, 1, 1, 1), &A(O.XKP, 1, 1, 1), &A(O.XKR, 1, 1, 1),
                        &A(O.XKA, 1, 1, 1), &A(O.DKP, 1, 1, 1), &A(O.DKA, 1, 1, 1)))
            return seterr(2, "NEG. T. IN BURROWS OPC (or a density outside the table)");
    }
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            if (type == 12) {
                A(O.T, 0, j, k) = A(O.T, 1, j, k) + D_T * (XXB(0) - XXB(1));
                A(O.T, -1, j, k) = A(O.T, 1, j, k) + D_T * (XXB(-1) - XXB(1));
            } else {
                /* T**3.d0: the reference's compiler multiplies, T*T*T (its
                   fixtures match that, not pow) */
                const int jj = iii > 0 ? j : 1, kk = iii > 0 ? k : 1;
                const double dtdr = -3.0 * fb * A(O.XKR, 1, jj, kk) * A(O.RH, 1, jj, kk) /
                                    (4.0 * ARAD * FD * (A(O.T, 1, jj, kk) * A(O.T, 1, jj, kk) * A(O.T, 1, jj, kk)));
                A(O.T, 0, j, k) = A(O.T, 1, j, k) - dtdr * (XXB(1) - XXB(0));
                A(O.T, -1, j, k) = A(O.T, 1, j, k) - dtdr * (XXB(1) - XXB(-1));
            }
            A(O.V, 1, j, k) = 0.0;
            A(O.V, 0, j, k) = 0.0;
        }
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = 0; j <= NY + 2; ++j) {
            A(O.G, 0, j, k) = A(O.G, 1, j, k);
            A(O.G, -1, j, k) = A(O.G, 1, j, k);
        }
    for (int k = 0; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            A(O.H, 0, j, k) = A(O.H, 1, j, k);
            A(O.H, -1, j, k) = A(O.H, 1, j, k);
        }
    return 0;
}

static int boundxmax2(int iii) /* ppbou.f:1269-1395, CASE(2) */
{
    const int ver = O.p.ibdrad;
    const double rh_top = (double)1.0e-10f;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            int ub = UBND(j, k);
            double TJ;
            if (ub >= NX + 1) return seterr(3, "error in boundxmax: upperbnd above the domain");
            if (ver == 3) {
                TJ = O.p.tirr * pow(cos(XYB(j)) * cos(XZB(k)), 0.25);
                TJ = TJ > 100.0 ? TJ : 100.0;
            } else if (ver == 4) {
                TJ = O.p.tirr;
            } else if (ver == 7) {
                TJ = A(O.T, ub - 1, j, k);
            } else if (ver == 9) {
                if (iii <= 1)
                    TJ = A(O.T, ub - 1, j, k);
                else
                    TJ = pow((A(O.ER, ub - 1, j, k) / ARAD) +
                                 (A(O.SIN, ub - 1, j, k) /
                                  (4.0 * SBCONST * A(O.RH, ub - 1, j, k) * A(O.XKP, ub - 1, j, k))),
                             0.25);
            } else {
                return seterr(1, "BOUNDXMAX(2): IBDRAD must be 3, 4, 7 or 9");
            }
            A(O.T, ub, j, k) = TJ;
            for (int i = ub + 1; i <= NX + 2; ++i) {
                A(O.T, i, j, k) = TJ;
                A(O.RH, i, j, k) = rh_top;
                A(O.V, i, j, k) = 0.0;
            }
            if (ub == NX) {
                A(O.V, NX, j, k) = 0.0;
                A(O.RH, NX, j, k) = rh_top;
                A(O.T, NX, j, k) = TJ;
            }
        }
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = 0; j <= NY + 2; ++j) {
            int ub = UBND(j, k);
            double g = A(O.G, ub, j, k);
            for (int i = ub + 1; i <= NX + 2; ++i) A(O.G, i, j, k) = g;
        }
    for (int k = 0; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            int ub = UBND(j, k);
            double h = A(O.H, ub, j, k);
            for (int i = ub + 1; i <= NX + 2; ++i) A(O.H, i, j, k) = h;
        }
    return 0;
}

static void boundzmin0(void) /* ppbou.f:1823-1878 */
{
    for (int i = -1; i <= NX + 2; ++i)
        for (int j = -1; j <= NY + 2; ++j) {
            A(O.RH, i, j, 1) = A(O.RH, i, j, 2);
            A(O.RH, i, j, 0) = A(O.RH, i, j, 1);
            A(O.RH, i, j, -1) = A(O.RH, i, j, 1);
            A(O.T, i, j, 1) = A(O.T, i, j, 2);
            A(O.T, i, j, 0) = A(O.T, i, j, 1);
            A(O.T, i, j, -1) = A(O.T, i, j, 1);
            A(O.H, i, j, 2) = 0.0;
            A(O.H, i, j, 1) = 0.0;
            A(O.H, i, j, 0) = 0.0;
        }
    for (int i = 0; i <= NX + 2; ++i)
        for (int j = -1; j <= NY + 2; ++j) {
            A(O.V, i, j, 1) = 0.0;
            A(O.V, i, j, 0) = A(O.V, i, j, 1);
            A(O.V, i, j, -1) = A(O.V, i, j, 1);
        }
    for (int i = -1; i <= NX + 2; ++i)
        for (int j = 0; j <= NY + 2; ++j) {
            A(O.G, i, j, 1) = A(O.G, i, j, 2) * SURZB(2) / SURZB(1);
            A(O.G, i, j, 0) = A(O.G, i, j, 1) * SURZB(1) / SURZB(0);
            A(O.G, i, j, -1) = A(O.G, i, j, 1) * SURZB(1) / cos(XZB(-1));
        }
}

static void boundzmax0(void) /* ppbou.f:1936-1988 */
{
    for (int i = -1; i <= NX + 2; ++i)
        for (int j = -1; j <= NY + 2; ++j) {
            A(O.RH, i, j, NZ) = A(O.RH, i, j, NZ - 1);
            A(O.RH, i, j, NZ + 1) = A(O.RH, i, j, NZ);
            A(O.RH, i, j, NZ + 2) = A(O.RH, i, j, NZ);
            A(O.T, i, j, NZ) = A(O.T, i, j, NZ - 1);
            A(O.T, i, j, NZ + 1) = A(O.T, i, j, NZ);
            A(O.T, i, j, NZ + 2) = A(O.T, i, j, NZ);
            A(O.H, i, j, NZ) = 0.0;
            A(O.H, i, j, NZ + 1) = 0.0;
            A(O.H, i, j, NZ + 2) = 0.0;
        }
    for (int i = 0; i <= NX + 2; ++i)
        for (int j = -1; j <= NY + 2; ++j) {
            A(O.V, i, j, NZ) = 0.0;
            A(O.V, i, j, NZ + 1) = A(O.V, i, j, NZ);
            A(O.V, i, j, NZ + 2) = A(O.V, i, j, NZ);
        }
    for (int i = -1; i <= NX + 2; ++i)
        for (int j = 0; j <= NY + 2; ++j) {
            A(O.G, i, j, NZ) = A(O.G, i, j, NZ - 1) * SURZB(NZ - 1) / SURZB(NZ);
            A(O.G, i, j, NZ + 1) = A(O.G, i, j, NZ) * SURZB(NZ) / SURZB(NZ + 1);
            A(O.G, i, j, NZ + 2) = A(O.G, i, j, NZ) * SURZB(NZ) / SURZB(NZ + 2);
        }
}


static int boundplanet(void) /* ppbou.f:94-249, MODVER 0/2/3/4/5 branch */
{
    const double thr = (double)1.0e-9f;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            int ub = NX - 1;
            for (int i = 1; i <= NX; ++i)
                if (A(O.RH, i, j, k) < thr) { ub = i; break; }
            UBND(j, k) = ub;
        }
    passupperbnd();
    int rc = boundxmin(O.p.bcxmin, g_iii);
    if (rc) return rc;
    rc = boundxmax2(g_iii);
    if (rc) return rc;
    boundzmin0();
    boundzmax0();
    return 0;
}

int tor_bounds(void) /* BOUNDS, ppbou.f:1-92, MODTYP=2 */
{
    double* f[5] = {O.T, O.RH, O.V, O.G, O.H};
    for (int n = 0; n < 5; ++n) planetshift(f[n], n + 1);
    int rc = boundplanet();
    if (rc) return rc;
    for (int n = 0; n < 5; ++n) planetshift(f[n], n + 1);
    return 0;
}

/* -------------------------------------------------------------- pprphy.f */
/* Axes of the IKAPTYP 4/5 tables: the reference's DATA statements
 * (ppmod.f:690-699); temperatures in K, pressures in mbar. */
static const double FREEDMAN_T[42] = {
    75.0, 100.0, 110.0, 120.0, 130.0, 140.0, 150.0, 160.0, 170.0, 180.0, 190.0, 200.0, 210.0, 220.0,
    230.0, 240.0, 250.0, 275.0, 300.0, 400.0, 500.0, 575.0, 650.0, 725.0, 800.0, 900.0, 1000.0, 1100.0,
    1200.0, 1300.0, 1400.0, 1500.0, 1600.0, 1700.0, 1800.0, 1900.0, 2000.0, 2300.0, 2600.0, 3000.0,
    3500.0, 4000.0};
static const double FREEDMAN_P[18] = {0.001, 0.003, 0.01, 0.03, 0.1, 0.3, 1.0, 3.0, 10.0,
                                      30.0, 100.0, 300.0, 1000.0, 3000.0, 10000.0, 30000.0,
                                      100000.0, 300000.0};
/* log10 Planck, Rosseland and absorption tables, temperature fastest */
static double g_ftab[3][42 * 18];
static int g_ftab_set;

int tor_set_opacity_table(const double* kp, const double* kr, const double* ka)
{
    memcpy(g_ftab[0], kp, sizeof g_ftab[0]);
    memcpy(g_ftab[1], kr, sizeof g_ftab[1]);
    memcpy(g_ftab[2], ka, sizeof g_ftab[2]);
    g_ftab_set = 1;
    return 0;
}

/* FreedmanOPC, pprphy.f:800-942: bilinear interpolation of the log10 tables
 * in (T, P), inside the table's clamps, then 10^x; the temperature
 * derivatives are the reference's one-sided difference ("CLUDGE FOR NOW").
 * T <= 0 is the caller's error. */
static void freedman_opc(double t_in, double pg_in, double* bkp, double* bkr, double* bka, double* dkp,
                         double* dka)
{
    double pg1 = pg_in * (0.001);
    double tm1 = t_in;
    if (tm1 < 75.0) tm1 = 75.001;
    if (tm1 > 4000.0) tm1 = 3999.9;
    if (pg1 < 0.001) pg1 = 0.001001;
    if (pg1 > 300000.0) pg1 = 299999.9;
    int it = 0; /* first node above T; the clamps keep it in 1..41 */
    while (!(FREEDMAN_T[it] > tm1)) ++it;
    const int it1 = it - 1;
    double ws1 = FREEDMAN_T[it1] - tm1;
    double ws = FREEDMAN_T[it1] - FREEDMAN_T[it];
    ws1 = ws1 / ws;
    ws = 1.0 - ws1;
    int iw = 0;
    while (!(FREEDMAN_P[iw] > pg1)) ++iw;
    const int iw1 = iw - 1;
    double ws2 = pg1 - FREEDMAN_P[iw1];
    double ws3 = FREEDMAN_P[iw] - FREEDMAN_P[iw1];
    ws2 = ws2 / ws3;
    ws3 = 1.0 - ws2;
    double out[3], k00[3], k01[3];
    for (int q = 0; q < 3; ++q) {
        const double* t = g_ftab[q];
        const double a00 = t[it + 42 * iw], a10 = t[it + 42 * iw1], a11 = t[it1 + 42 * iw1],
                     a01 = t[it1 + 42 * iw];
        const double w4 = ws2 * (ws1 * a00 + ws * a01);
        const double w5 = ws3 * (ws1 * a10 + ws * a11);
        out[q] = pow(10.0, w4 + w5);
        k00[q] = a00;
        k01[q] = a01;
    }
    *bkp = out[0];
    *bkr = out[1];
    *bka = out[2];
    *dkp = (pow(10.0, k00[0]) - pow(10.0, k01[0])) / (FREEDMAN_T[it] - FREEDMAN_T[it1]);
    *dka = (pow(10.0, k00[2]) - pow(10.0, k01[2])) / (FREEDMAN_T[it] - FREEDMAN_T[it1]);
}

/* Axes of the IKAPTYP 9-14 tables (ppmod.f:724-745): log10 T and log10 rho,
 * REAL constants of DATA statements stored in double arrays (float values,
 * widened). */
static const double BURROWS_T[50] = {
    1.69896f, 1.73978f, 1.78061f, 1.82143f, 1.86225f, 1.90308f, 1.94390f, 1.98473f, 2.02555f, 2.06637f,
    2.10720f, 2.14802f, 2.18884f, 2.22967f, 2.27049f, 2.31132f, 2.35214f, 2.39296f, 2.43379f, 2.47461f,
    2.51543f, 2.55626f, 2.59708f, 2.63790f, 2.67873f, 2.71955f, 2.76038f, 2.80120f, 2.84202f, 2.88285f,
    2.92367f, 2.96406f, 3.00488f, 3.04571f, 3.08653f, 3.12735f, 3.16818f, 3.20900f, 3.24983f, 3.29065f,
    3.33147f, 3.37230f, 3.41312f, 3.45394f, 3.49477f, 3.53559f, 3.57642f, 3.61724f, 3.65806f, 3.69889f};
static const double BURROWS_RH[50] = {
    -11.9996f, -11.7954f, -11.5913f, -11.3872f, -11.1831f, -10.9790f, -10.7748f, -10.5707f, -10.3666f,
    -10.1625f, -9.95837f, -9.75425f, -9.55014f, -9.34602f, -9.14190f, -8.93778f, -8.73366f, -8.52954f,
    -8.32543f, -8.12131f, -7.91719f, -7.71307f, -7.50895f, -7.30483f, -7.10071f, -6.89660f, -6.69248f,
    -6.48836f, -6.28424f, -6.08012f, -5.87600f, -5.67189f, -5.46777f, -5.26365f, -5.05953f, -4.85541f,
    -4.65129f, -4.44718f, -4.24479f, -4.04068f, -3.83656f, -3.63244f, -3.42875f, -3.22464f, -3.02052f,
    -2.81640f, -2.61228f, -2.40816f, -2.20404f, -1.99993f};
static double g_btab[3][50 * 50]; /* log10 Planck, Rosseland, absorption; T fastest */
static int g_btab_set;

int tor_set_burrows_table(const double* kp, const double* kr, const double* ka)
{
    memcpy(g_btab[0], kp, sizeof g_btab[0]);
    memcpy(g_btab[1], kr, sizeof g_btab[1]);
    memcpy(g_btab[2], ka, sizeof g_btab[2]);
    g_btab_set = 1;
    return 0;
}

/* BurrowsOPC, pprphy.f:1005-1225 (the grey part: the wavelength bins of
 * IKAPTYP 20-25 are not evaluated): log10 T and log10 rho clamped into the
 * table, the first node above each, bilinear interpolation of the log
 * tables, 10^x; dk/dT the one-sided difference over 10^ of the T nodes.
 * Returns 1 where the reference stops (T <= 0, or a NaN log10 rho, which
 * finds no node above it). */
/* BurrowsOPC's table nodes and bilinear weights (pprphy.f:1040-1105) */
typedef struct { int it, it1, iw, iw1; double ws1, ws, ws2, ws3; } bur_nodes;

static int burrows_nodes(double t_in, double rh_in, bur_nodes* b)
{
    if (t_in <= 0.0) return 1;
    double rh1 = log10(rh_in);
    double tm1 = log10(t_in);
    if (tm1 < (double)1.69896f) tm1 = (double)1.69897f;
    if (tm1 > (double)3.69889f) tm1 = (double)3.69888f;
    if (rh1 < (double)-11.9996f) rh1 = (double)-11.9995f;
    if (rh1 > (double)-1.99993f) rh1 = (double)-1.99994f;
    int it = 0;
    while (it < 50 && !(BURROWS_T[it] > tm1)) ++it;
    if (it == 50 || it == 0) return 1;
    b->it = it;
    b->it1 = it - 1;
    b->ws1 = BURROWS_T[b->it1] - tm1;
    b->ws = BURROWS_T[b->it1] - BURROWS_T[it];
    b->ws1 = b->ws1 / b->ws;
    b->ws = 1.0 - b->ws1;
    int iw = 0;
    while (iw < 50 && !(BURROWS_RH[iw] > rh1)) ++iw;
    if (iw == 50 || iw == 0) return 1;
    b->iw = iw;
    b->iw1 = iw - 1;
    b->ws2 = rh1 - BURROWS_RH[b->iw1];
    b->ws3 = BURROWS_RH[iw] - BURROWS_RH[b->iw1];
    b->ws2 = b->ws2 / b->ws3;
    b->ws3 = 1.0 - b->ws2;
    return 0;
}

/* IKAPTYP 20-25: log10 of the 30 wavelength bins' absorption opacity of the
   stellar irradiation, BurrowsXKAP_waveopac(50,50,30) (T fastest, then rho,
   then bin), the bins' stellar flux BSTAR_BIN and the dilution
   (RSTAR RSOL / semi)^2 */
static double g_wtab[30 * 2500], g_bstar[30], g_wdil;
static int g_wtab_set;

int tor_set_burrows_wave_table(const double* kw, const double* bstar, double dil)
{
    memcpy(g_wtab, kw, sizeof g_wtab);
    memcpy(g_bstar, bstar, sizeof g_bstar);
    g_wdil = dil;
    g_wtab_set = 1;
    return 0;
}

/* BKAP_W of BurrowsOPC (pprphy.f:1138-1151): the bins' 10^ of the bilinear
   interpolant */
static int burrows_wave(double t_in, double rh_in, double kw[30])
{
    bur_nodes b;
    if (burrows_nodes(t_in, rh_in, &b)) return 1;
    for (int l = 0; l < 30; ++l) {
        const double* t = g_wtab + 2500 * l;
        const double w00 = t[b.it + 50 * b.iw], w10 = t[b.it + 50 * b.iw1], w11 = t[b.it1 + 50 * b.iw1],
                     w01 = t[b.it1 + 50 * b.iw];
        const double w4 = b.ws2 * (b.ws1 * w00 + b.ws * w01);
        const double w5 = b.ws3 * (b.ws1 * w10 + b.ws * w11);
        kw[l] = pow(10.0, w4 + w5);
    }
    return 0;
}

static int burrows_opc(double t_in, double rh_in, double* bkp, double* bkr, double* bka, double* dkp,
                       double* dka)
{
    if (t_in <= 0.0) return 1;
    double rh1 = log10(rh_in);
    double tm1 = log10(t_in);
    if (tm1 < (double)1.69896f) tm1 = (double)1.69897f;
    if (tm1 > (double)3.69889f) tm1 = (double)3.69888f;
    if (rh1 < (double)-11.9996f) rh1 = (double)-11.9995f;
    if (rh1 > (double)-1.99993f) rh1 = (double)-1.99994f;
    int it = 0;
    while (it < 50 && !(BURROWS_T[it] > tm1)) ++it;
    if (it == 50 || it == 0) return 1;
    const int it1 = it - 1;
    double ws1 = BURROWS_T[it1] - tm1;
    double ws = BURROWS_T[it1] - BURROWS_T[it];
    ws1 = ws1 / ws;
    ws = 1.0 - ws1;
    int iw = 0;
    while (iw < 50 && !(BURROWS_RH[iw] > rh1)) ++iw;
    if (iw == 50 || iw == 0) return 1;
    const int iw1 = iw - 1;
    double ws2 = rh1 - BURROWS_RH[iw1];
    double ws3 = BURROWS_RH[iw] - BURROWS_RH[iw1];
    ws2 = ws2 / ws3;
    ws3 = 1.0 - ws2;
    double out[3], k00[3], k01[3];
    for (int q = 0; q < 3; ++q) {
        const double* t = g_btab[q];
        const double a00 = t[it + 50 * iw], a10 = t[it + 50 * iw1], a11 = t[it1 + 50 * iw1], a01 = t[it1 + 50 * iw];
        const double w4 = ws2 * (ws1 * a00 + ws * a01);
        const double w5 = ws3 * (ws1 * a10 + ws * a11);
        out[q] = pow(10.0, w4 + w5);
        k00[q] = a00;
        k01[q] = a01;
    }
    *bkp = out[0];
    *bkr = out[1];
    *bka = out[2];
    const double dt = pow(10.0, BURROWS_T[it]) - pow(10.0, BURROWS_T[it1]);
    *dkp = (pow(10.0, k00[0]) - pow(10.0, k01[0])) / dt;
    *dka = (pow(10.0, k00[2]) - pow(10.0, k01[2])) / dt;
    return 0;
}

static double g_ctab[8 * 42 * 18]; /* IKAPTYP 6/7 column table, T fastest, then P, then level */
static int g_ctab_set;

int tor_set_column_table(const double* ka)
{
    memcpy(g_ctab, ka, sizeof g_ctab);
    g_ctab_set = 1;
    return 0;
}

/* columnOPC's absorption mean, pprphy.f:1543-1745, after freedman_opc has
 * set the Planck and Rosseland means and dkP/dT (the reference calls
 * FreedmanOPC first and columnOPC then sets all five to the same P, R
 * values): the column cut-off, its level weights (CS0 with the pressure
 * node FreedmanPVAL(IC1), as the reference has it) and dkA/dT from the upper
 * level.  Returns 1 where the reference stops. */
static int column_opc(double t_in, double pg_in, double rh_in, double mu_in, double dr, double* bka, double* dka)
{
    double pg1 = pg_in * (0.001);
    double tm1 = t_in;
    double kc = (double)-23.9526937f + log10(mu_in) - log10(rh_in) - log10(dr + 1.0);
    if (kc < -26.0) kc = -25.9999;
    if (kc > -19.0) kc = -19.0001;
    if (tm1 < 75.0) tm1 = 75.001;
    if (tm1 > 4000.0) tm1 = 3999.9;
    if (pg1 < 0.001) pg1 = 0.001001;
    if (pg1 > 300000.0) pg1 = 299999.9;
    int it = 0;
    while (!(FREEDMAN_T[it] > tm1)) ++it;
    const int it1 = it - 1;
    double ws1 = FREEDMAN_T[it1] - tm1;
    double ws = FREEDMAN_T[it1] - FREEDMAN_T[it];
    ws1 = ws1 / ws;
    ws = 1.0 - ws1;
    int iw = 0;
    while (!(FREEDMAN_P[iw] > pg1)) ++iw;
    const int iw1 = iw - 1;
    double ws2 = pg1 - FREEDMAN_P[iw1];
    double ws3 = FREEDMAN_P[iw] - FREEDMAN_P[iw1];
    ws2 = ws2 / ws3;
    ws3 = 1.0 - ws2;
    static const double CVAL[8] = {-26.0, -25.0, -24.0, -23.0, -22.0, -21.0, -20.0, -19.0};
    int ic = 0;
    while (ic < 8 && !(CVAL[ic] > kc)) ++ic;
    if (ic == 8 || ic == 0) return 1;
    const int ic1 = ic - 1;
    double cs1 = kc - CVAL[ic1];
    double cs0 = CVAL[ic] - FREEDMAN_P[ic1];
    cs1 = cs1 / cs0;
    cs0 = 1.0 - cs1;
#define CA(a, b, c) g_ctab[(a) + 42 * ((b) + 18 * (c))]
    const double a000 = CA(it, iw, ic), a010 = CA(it, iw1, ic), a011 = CA(it1, iw1, ic), a001 = CA(it1, iw, ic);
    double w4 = ws2 * (ws1 * a000 + ws * a001);
    double w5 = ws3 * (ws1 * a010 + ws * a011);
    const double a_0 = cs1 * (w4 + w5);
    const double a100 = CA(it, iw, ic1), a110 = CA(it, iw1, ic1), a111 = CA(it1, iw1, ic1), a101 = CA(it1, iw, ic1);
    w4 = ws2 * (ws1 * a100 + ws * a101);
    w5 = ws3 * (ws1 * a110 + ws * a111);
    const double a_1 = cs0 * (w4 + w5);
#undef CA
    *bka = pow(10.0, a_0 + a_1);
    *dka = (pow(10.0, a000) - pow(10.0, a001)) / (FREEDMAN_T[it] - FREEDMAN_T[it1]);
    return 0;
}

static int kappa(void) /* KAPPA, pprphy.f:1-360; case 26 is new */
{
    const tt_params* p = &O.p;
    if (p->ikaptyp >= 4 && p->ikaptyp <= 7) { /* pprphy.f:84-175, MODTYP=2 */
        if (!g_ftab_set) return seterr(1, "KAPPA: IKAPTYP 4-7 need an opacity table");
        const int colm = p->ikaptyp >= 6;
        if (colm && !g_ctab_set) return seterr(1, "KAPPA: IKAPTYP 6/7 need a column table");
        tor_pressg(); /* pprphy.f:88 */
        int bad = 0;
#pragma omp parallel for reduction(| : bad)
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j) {
                int ub = UBND(j, k);
                for (int i = -1; i <= ub; ++i) {
                    if (A(O.T, i, j, k) <= 0.0) { bad = 1; continue; }
                    freedman_opc(A(O.T, i, j, k), A(O.PG, i, j, k), &A(O.XKP, i, j, k), &A(O.XKR, i, j, k),
                                 &A(O.XKA, i, j, k), &A(O.DKP, i, j, k), &A(O.DKA, i, j, k));
                    if (colm) /* XMUE = mu_gas (GASINIT, ppinit.f:1414) */
                        bad |= column_opc(A(O.T, i, j, k), A(O.PG, i, j, k), A(O.RH, i, j, k), p->mu_gas,
                                          XXB(ub) - XXB(i), &A(O.XKA, i, j, k), &A(O.DKA, i, j, k));
                }
                for (int i = ub + 1; i <= NX + 2; ++i) {
                    A(O.XKR, i, j, k) = A(O.XKR, ub, j, k);
                    A(O.XKP, i, j, k) = A(O.XKP, ub, j, k);
                    A(O.XKA, i, j, k) = A(O.XKA, ub, j, k);
                    A(O.DKP, i, j, k) = A(O.DKP, ub, j, k);
                    A(O.DKA, i, j, k) = A(O.DKA, ub, j, k);
                }
            }
        if (bad) return seterr(2, "NEG. T. IN FREEDMAN OPC");
        return 0;
    }
    if ((p->ikaptyp >= 9 && p->ikaptyp <= 12) || p->ikaptyp == 14) { /* pprphy.f:199-234,242-274 */
        if (!g_btab_set) return seterr(1, "KAPPA: IKAPTYP 9-12 and 14 need an opacity table");
        int bad = 0;
#pragma omp parallel for reduction(| : bad)
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j) {
                int ub = UBND(j, k);
                for (int i = -1; i <= ub; ++i)
                    bad |= burrows_opc(A(O.T, i, j, k), A(O.RH, i, j, k), &A(O.XKP, i, j, k), &A(O.XKR, i, j, k),
                                       &A(O.XKA, i, j, k), &A(O.DKP, i, j, k), &A(O.DKA, i, j, k));
                for (int i = ub + 1; i <= NX + 2; ++i) {
                    A(O.XKR, i, j, k) = A(O.XKR, ub, j, k);
                    A(O.XKP, i, j, k) = A(O.XKP, ub, j, k);
                    A(O.XKA, i, j, k) = A(O.XKA, ub, j, k);
                    A(O.DKP, i, j, k) = A(O.DKP, ub, j, k);
                    A(O.DKA, i, j, k) = A(O.DKA, ub, j, k);
                }
            }
        if (bad) return seterr(2, "NEG. T. IN BURROWS OPC (or a density outside the table)");
        return 0;
    }
    if (p->ikaptyp >= 20 && p->ikaptyp <= 25) { /* pprphy.f:321-349: the cells from upperbnd up take
                                                   upperbnd-1's values (the wavelength bins are formed where
                                                   STELLARINPUT_HEATING uses them) */
        if (!g_btab_set) return seterr(1, "KAPPA: IKAPTYP 20-25 need an opacity table");
        int bad = 0;
#pragma omp parallel for reduction(| : bad)
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j) {
                int ub = UBND(j, k);
                for (int i = -1; i <= ub; ++i)
                    bad |= burrows_opc(A(O.T, i, j, k), A(O.RH, i, j, k), &A(O.XKP, i, j, k), &A(O.XKR, i, j, k),
                                       &A(O.XKA, i, j, k), &A(O.DKP, i, j, k), &A(O.DKA, i, j, k));
                for (int i = ub; i <= NX + 2; ++i) {
                    A(O.XKR, i, j, k) = A(O.XKR, ub - 1, j, k);
                    A(O.XKP, i, j, k) = A(O.XKP, ub - 1, j, k);
                    A(O.XKA, i, j, k) = A(O.XKA, ub - 1, j, k);
                    A(O.DKP, i, j, k) = A(O.DKP, ub - 1, j, k);
                    A(O.DKA, i, j, k) = A(O.DKA, ub - 1, j, k);
                }
            }
        if (bad) return seterr(2, "NEG. T. IN BURROWS OPC (or a density outside the table)");
        return 0;
    }
    if (p->ikaptyp == 1) { /* pprphy.f:41-47 */
        for (size_t n = 0; n < O.n3; ++n) {
            O.XKR[n] = p->zkapmult; O.XKP[n] = p->zkapmult; O.XKA[n] = p->zkapmult;
            O.DKP[n] = 0.0; O.DKA[n] = 0.0;
        }
        return 0;
    }
    if (p->ikaptyp == 8) { /* H- bound-free, pprphy.f:176-198: j from 1; 0.00342 a REAL */
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = 1; j <= NY + 2; ++j) {
                int ub = UBND(j, k);
                for (int i = -1; i <= ub; ++i) {
                    A(O.XKR, i, j, k) = 0.95 * pow(A(O.RH, i, j, k), 0.5) * pow(A(O.T, i, j, k) / 2500.0, 9.0);
                    A(O.DKP, i, j, k) = (double)0.00342f * pow(A(O.RH, i, j, k), 0.5) * pow(A(O.T, i, j, k) / 2500.0, 8.0);
                    A(O.DKA, i, j, k) = (double)0.00342f * pow(A(O.RH, i, j, k), 0.5) * pow(A(O.T, i, j, k) / 2500.0, 8.0);
                }
                for (int i = ub + 1; i <= NX + 2; ++i) {
                    A(O.XKR, i, j, k) = A(O.XKR, ub, j, k);
                    A(O.DKP, i, j, k) = A(O.DKP, ub, j, k);
                    A(O.DKA, i, j, k) = A(O.DKA, ub, j, k);
                }
            }
        for (size_t n = 0; n < O.n3; ++n) { O.XKP[n] = O.XKR[n]; O.XKA[n] = O.XKR[n]; }
        return 0;
    }
    if (p->ikaptyp == 13) { /* pprphy.f:235-241 */
        const double xa = pow(p->gamma_opc, 4.0) * p->zkapmult;
        for (size_t n = 0; n < O.n3; ++n) {
            O.XKR[n] = p->zkapmult; O.XKP[n] = p->zkapmult; O.XKA[n] = xa;
            O.DKP[n] = 0.0; O.DKA[n] = 0.0;
        }
        return 0;
    }
    if (p->ikaptyp != 26) return seterr(1, "KAPPA: IKAPTYP must be 1, 4-14, 20-25 or 26");
    /* Power law kappa_R = kappa_P = kappa_A = kap0 (rho/rho0)^a (T/T0)^b in the
     * loop structure of the IKAPTYP=8 power law (pprphy.f:176-198): cells up to
     * upperbnd are evaluated, cells above copy the upperbnd value.  Unlike
     * case 8 (whose j loop starts at 1) all ghost columns are filled. */
#pragma omp parallel for
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            int ub = UBND(j, k);
            for (int i = -1; i <= ub; ++i) {
                double rr = A(O.RH, i, j, k) / p->kap_rho0;
                double tt = A(O.T, i, j, k) / p->kap_t0;
                A(O.XKR, i, j, k) = p->kap0 * pow(rr, p->kap_a) * pow(tt, p->kap_b);
                A(O.DKP, i, j, k) = p->dkap0 * pow(rr, p->kap_a) * pow(tt, p->kap_b - 1.0);
                A(O.DKA, i, j, k) = p->dkap0 * pow(rr, p->kap_a) * pow(tt, p->kap_b - 1.0);
            }
            for (int i = ub + 1; i <= NX + 2; ++i) {
                A(O.XKR, i, j, k) = A(O.XKR, ub, j, k);
                A(O.DKP, i, j, k) = A(O.DKP, ub, j, k);
                A(O.DKA, i, j, k) = A(O.DKA, ub, j, k);
            }
        }
    for (size_t n = 0; n < O.n3; ++n) { O.XKP[n] = O.XKR[n]; O.XKA[n] = O.XKR[n]; }
    return 0;
}

static int subsig(void) /* SUBSIG, pprphy.f:567-583 */
{
    double s;
    if (O.p.isigtyp == 0) s = 0.0;
    else if (O.p.isigtyp == 1) s = 0.4;
    else return seterr(1, "SUBSIG: ISIGTYP must be 0 or 1");
    for (size_t n = 0; n < O.n3; ++n) O.SIG[n] = s;
    return 0;
}

int tor_kappa(void) /* KAPPA + SUBSIG as FLIM calls them (pprphy.f:362-380) */
{
    int rc = kappa();
    return rc ? rc : subsig();
}

static int flim(void) /* FLIM, pprphy.f:362-565; IRAD=3 */
{
    const double xdmin = 1.0e-50;
    if (O.energy_iter <= 1) {
        int rc = kappa();
        if (rc) return rc;
        rc = subsig();
        if (rc) return rc;
        tor_qdens();
    }
    if (O.p.iflimtyp == 1) {
        for (size_t n = 0; n < O.n3; ++n) {
            O.FLIMX[n] = 1.0 / 3.0; O.FLIMY[n] = 1.0 / 3.0; O.FLIMZ[n] = 1.0 / 3.0;
        }
        return 0;
    }
    if (O.p.iflimtyp != 2) return seterr(1, "FLIM: IFLIMTYP must be 1 or 2");
    int xmax_index = 0;
    /* x faces; xmax_index keeps its last value (ub(ny,nz)-1) for y and z,
     * as the reference does (pprphy.f:464,495) */
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            xmax_index = UBND(j, k) - 1;
            for (int i = 1; i <= xmax_index; ++i) {
                double gx = (A(O.ER, i, j, k) - A(O.ER, i - 1, j, k)) / DXB(i);
                double gy = (DDX1(i) * (A(O.ER, i, j + 1, k) - A(O.ER, i, j - 1, k)) +
                             DDX0(i) * (A(O.ER, i - 1, j + 1, k) - A(O.ER, i - 1, j - 1, k))) /
                            (GEOXGA(i) * GEOZG(k) * (XYB(j + 1) - XYB(j - 1)));
                double gz = (DDX1(i) * (A(O.ER, i, j, k + 1) - A(O.ER, i, j, k - 1)) +
                             DDX0(i) * (A(O.ER, i - 1, j, k + 1) - A(O.ER, i - 1, j, k - 1))) /
                            (GEOXHA(i) * (XZB(k + 1) - XZB(k - 1)));
                double grad = sqrt(gx * gx + gy * gy + gz * gz) + xdmin;
                double rr = 2.0 * grad /
                            (A(O.RHQX, i, j, k) *
                             (DDX1(i) * (A(O.XKR, i, j, k) + A(O.SIG, i, j, k)) +
                              DDX0(i) * (A(O.XKR, i - 1, j, k) + A(O.SIG, i, j, k))) *
                             (A(O.ER, i, j, k) + A(O.ER, i - 1, j, k)));
                A(O.FLIMX, i, j, k) = (2.0 + rr) / (6.0 + 3.0 * rr + rr * rr);
            }
            for (int i = UBND(j, k); i <= NX + 1; ++i)
                A(O.FLIMX, i, j, k) = A(O.FLIMX, UBND(j, k) - 1, j, k);
        }
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY + 1; ++j) {
            for (int i = 1; i <= xmax_index; ++i) {
                double gx = (DDY1(j) * (A(O.ER, i + 1, j, k) - A(O.ER, i - 1, j, k)) +
                             DDY0(j) * (A(O.ER, i + 1, j - 1, k) - A(O.ER, i - 1, j - 1, k))) /
                            (XXB(i + 1) - XXB(i - 1));
                double gy = (A(O.ER, i, j, k) - A(O.ER, i, j - 1, k)) / (GEOXG(i) * GEOZG(k) * DYB(j));
                double gz = (DDY1(j) * (A(O.ER, i, j, k + 1) - A(O.ER, i, j, k - 1)) +
                             DDY0(j) * (A(O.ER, i, j - 1, k + 1) - A(O.ER, i, j - 1, k - 1))) /
                            ((XZB(k + 1) - XZB(k - 1)) * GEOXH(i));
                double grad = sqrt(gx * gx + gy * gy + gz * gz) + xdmin;
                double rr = 2.0 * grad /
                            (A(O.RHQY, i, j, k) *
                             (DDY1(j) * (A(O.XKR, i, j, k) + A(O.SIG, i, j, k)) +
                              DDY0(j) * (A(O.XKR, i, j - 1, k) + A(O.SIG, i, j, k))) *
                             (A(O.ER, i, j, k) + A(O.ER, i, j - 1, k)));
                A(O.FLIMY, i, j, k) = (2.0 + rr) / (6.0 + 3.0 * rr + rr * rr);
            }
            for (int i = UBND(j, k); i <= NX; ++i)
                A(O.FLIMY, i, j, k) = A(O.FLIMY, UBND(j, k) - 1, j, k);
        }
    for (int k = 1; k <= NZ + 1; ++k)
        for (int j = 1; j <= NY; ++j) {
            for (int i = 1; i <= xmax_index; ++i) {
                double gx = (DDZ1(k) * (A(O.ER, i + 1, j, k) - A(O.ER, i - 1, j, k)) +
                             DDZ0(k) * (A(O.ER, i + 1, j, k - 1) - A(O.ER, i - 1, j, k - 1))) /
                            (XXB(i + 1) - XXB(i - 1));
                double gy = (DDZ1(k) * (A(O.ER, i, j + 1, k) - A(O.ER, i, j - 1, k)) +
                             DDZ0(k) * (A(O.ER, i, j + 1, k - 1) - A(O.ER, i, j - 1, k - 1))) /
                            ((XYB(j + 1) - XYB(j - 1)) * GEOXG(i) * GEOZGA(k));
                double gz = (A(O.ER, i, j, k) - A(O.ER, i, j, k - 1)) / (GEOXH(i) * DZB(k));
                double grad = sqrt(gx * gx + gy * gy + gz * gz) + xdmin;
                double rr = 2.0 * grad /
                            (A(O.RHQZ, i, j, k) *
                             (DDZ1(k) * (A(O.XKR, i, j, k) + A(O.SIG, i, j, k)) +
                              DDZ0(k) * (A(O.XKR, i, j, k - 1) + A(O.SIG, i, j, k))) *
                             (A(O.ER, i, j, k) + A(O.ER, i, j, k - 1)));
                A(O.FLIMZ, i, j, k) = (2.0 + rr) / (6.0 + 3.0 * rr + rr * rr);
            }
            for (int i = UBND(j, k); i <= NX; ++i)
                A(O.FLIMZ, i, j, k) = A(O.FLIMZ, UBND(j, k) - 1, j, k);
        }
    return 0;
}

/* --------------------------------------------------------------- pprad.f */
static int diff(void) /* DIFF, pprad.f:74-156 (IETA != 5) */
{
    int rc = flim();
    if (rc) return rc;
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX + 1; ++i)
                A(O.DIFRX, i, j, k) =
                    FD * A(O.FLIMX, i, j, k) /
                    (DXB(i) * A(O.RHQX, i, j, k) *
                     (DDX1(i) * (A(O.XKR, i, j, k) + A(O.SIG, i, j, k)) +
                      DDX0(i) * (A(O.XKR, i - 1, j, k) + A(O.SIG, i - 1, j, k))));
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY + 1; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.DIFRY, i, j, k) =
                    FD * A(O.FLIMY, i, j, k) /
                    (DYB(j) * GEOXG(i) * GEOZG(k) * A(O.RHQY, i, j, k) *
                     (DDY1(j) * (A(O.XKR, i, j, k) + A(O.SIG, i, j, k)) +
                      DDY0(j) * (A(O.XKR, i, j - 1, k) + A(O.SIG, i, j - 1, k))));
#pragma omp parallel for
    for (int k = 1; k <= NZ + 1; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.DIFRZ, i, j, k) =
                    FD * A(O.FLIMZ, i, j, k) /
                    (DZB(k) * GEOXH(i) * A(O.RHQZ, i, j, k) *
                     (DDZ1(k) * (A(O.XKR, i, j, k) + A(O.SIG, i, j, k)) +
                      DDZ0(k) * (A(O.XKR, i, j, k - 1) + A(O.SIG, i, j, k - 1))));
    return 0;
}

static void stellar_tau(void) /* CALCSTELLARTAU + STELLAR_DTAU, pprad.f:191-242 */
{
    const double dr = XXB(2) - XXB(1);
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            int ub = UBND(j, k);
            if (ub < NX)
                for (int i = ub + 1; i <= NX + 1; ++i) A(O.STAU, i, j, k) = 0.0;
            double tot = 0.0;
            for (int i = ub; i >= 1; --i) {
                tot = tot + (A(O.XKA, i, j, k) * A(O.RH, i, j, k) * dr);
                A(O.STAU, i, j, k) = tot;
            }
        }
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            for (int i = 2; i <= NX; ++i)
                A(O.DTAUS, i, j, k) = (A(O.STAU, i - 1, j, k) - A(O.STAU, i + 1, j, k)) / 2.0;
            A(O.DTAUS, 1, j, k) = A(O.DTAUS, 2, j, k);
            A(O.DTAUS, NX + 1, j, k) = A(O.DTAUS, NX, j, k);
        }
}

static double calc_tincident(int j, int k) /* CALC_TINCIDENT, pprad.f:296-405, F_INCIDENT 2:3 */
{
    double angle = cos(XYB(j)) * cos(XZB(k));
    if (angle >= 0.0) {
        double t = O.p.tirr * pow(angle, 0.25);
        return t > 0.0 ? t : 0.0;
    }
    return 0.0;
}

/* The wavelength-binned branch of STELLARINPUT_HEATING (pprad.f:429-447) for
 * IKAPTYP 20, 21, 23 and 24: per bin, the optical depth by a reverse scan of
 * xkapW rho DR from upperbnd (WAVELENGTH_CALCSTELLARTAU, pprad.f:519-547;
 * xkapW at upperbnd is upperbnd-1's, as KAPPA copies it), its central
 * difference (WAVELENGTH_STELLAR_DTAU, :549-569: dtau(1) = dtau(2),
 * dtau(nx+1) = 0), and STELLARINPUT = sum over the bins in order of
 * BSTAR_BIN exp(-tau/mu*) dtau, times the dilution, over dxb.
 * The reference's i loop runs to nx+2 over arrays declared 1:nx (STELLARINPUT)
 * and 1:nx+1 (tau, dtau), and STELLARINPUT is zeroed once and then summed
 * into, so the column before in Fortran order (j-1, or ny+2 of k-1) leaves
 * at i = nx+2 a sum over its own mu* of this column's tau(1), dtau(1) (times
 * the dilution over dxb(nx+2)) in this column's STELLARINPUT(2), to which
 * this column's own terms are then added; at i = nx+1 it adds zeros
 * (dtau(nx+1) = 0).  The first column of the box gets nothing. */
static int stellarinput_wave(void)
{
    if (!g_wtab_set) return seterr(1, "STELLARINPUT_HEATING: IKAPTYP 20-24 need the wavelength table");
    const double dr = XXB(2) - XXB(1);
    const int nt = NX + 2; /* tau(1..nx+1), tau[l * nt + i] */
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int k = -1; k <= NZ + 2; ++k) {
        double* tau = (double*)calloc((size_t)30 * nt, sizeof(double));
        double* dtau = (double*)calloc((size_t)30 * nt, sizeof(double));
        for (int j = -1; j <= NY + 2; ++j) {
            const int ub = UBND(j, k);
            memset(tau, 0, sizeof(double) * 30 * nt);
            double tot[30], kw[30];
            for (int l = 0; l < 30; ++l) tot[l] = 0.0;
            for (int i = ub; i >= 1; --i) {
                const int c = i < ub ? i : ub - 1;
                if (burrows_wave(A(O.T, c, j, k), A(O.RH, c, j, k), kw)) { bad = 1; continue; }
                for (int l = 0; l < 30; ++l) {
                    tot[l] = tot[l] + (kw[l] * A(O.RH, i, j, k) * dr);
                    tau[l * nt + i] = tot[l];
                }
            }
            for (int l = 0; l < 30; ++l) {
                double* t = tau + l * nt;
                double* d = dtau + l * nt;
                for (int i = 2; i <= NX; ++i) d[i] = (t[i - 1] - t[i + 1]) / 2.0;
                d[1] = d[2];
                d[NX + 1] = 0.0;
            }
            const double mu = O.MUSTAR[I2(j, k)];
            double carry = 0.0; /* the previous column's i = nx+2 sum */
            if (j > -1 || k > -1) {
                const double mup = j > -1 ? O.MUSTAR[I2(j - 1, k)] : O.MUSTAR[I2(NY + 2, k - 1)];
                for (int l = 0; l < 30; ++l)
                    carry = carry + g_bstar[l] * (exp(-tau[l * nt + 1] / mup)) * dtau[l * nt + 1];
                carry = carry * g_wdil / DXB(NX + 2);
            }
            for (int i = 1; i <= NX; ++i) {
                double s = i == 2 ? carry : 0.0;
                for (int l = 0; l < 30; ++l)
                    s = s + g_bstar[l] * (exp(-tau[l * nt + i] / mu)) * dtau[l * nt + i];
                A(O.SIN, i, j, k) = s * g_wdil / DXB(i);
            }
        }
        free(tau);
        free(dtau);
    }
    if (bad) return seterr(2, "NEG. T. IN BURROWS OPC (or a density outside the table)");
    return 0;
}

static int stellarinput_heating(void) /* STELLARINPUT_HEATING, pprad.f:408-516, grey */
{
    if (O.p.f_incident == 0) {
        for (size_t n = 0; n < O.n3; ++n) O.SIN[n] = 0.0;
        return 0;
    }
    if (O.p.f_incident != 2) return seterr(1, "STELLARINPUT_HEATING: F_INCIDENT must be 0 or 2");
    const int ik = O.p.ikaptyp;
    if (ik == 20 || ik == 21 || ik == 23 || ik == 24) return stellarinput_wave();
    stellar_tau();
    /* The reference's i loop runs to nx+2 over an array declared 1:nx; the
     * two extra writes per column land on (1:2, j+1, k) and are overwritten
     * by the next column, so the result is that of i = 1..nx. */
#pragma omp parallel for
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            double tinc = calc_tincident(j, k);
            double mu = O.MUSTAR[I2(j, k)];
            for (int i = 1; i <= NX; ++i)
                A(O.SIN, i, j, k) = SBCONST * (pow(tinc, 4.0) / mu) *
                                    (exp(-A(O.STAU, i, j, k) / mu)) * A(O.DTAUS, i, j, k) / DXB(i);
        }
    return 0;
}

/* ------------------------------------------------------------ ppenergy.f */
static int calculateflux(void) /* CALCULATEFLUX, ppenergy.f:880-920 */
{
    int rc = diff();
    if (rc) return rc;
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX + 1; ++i)
                A(O.FLUXX, i, j, k) = -A(O.DIFRX, i, j, k) * (A(O.ER, i, j, k) - A(O.ER, i - 1, j, k));
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY + 1; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.FLUXY, i, j, k) = -A(O.DIFRY, i, j, k) * (A(O.ER, i, j, k) - A(O.ER, i, j - 1, k));
#pragma omp parallel for
    for (int k = 1; k <= NZ + 1; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.FLUXZ, i, j, k) = -A(O.DIFRZ, i, j, k) * (A(O.ER, i, j, k) - A(O.ER, i, j, k - 1));
    return 0;
}

static void flux_bc(void) /* FLUX_BC, ppenergy.f:922-982 */
{
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            int ub = UBND(j, k);
            for (int i = ub; i <= NX + 1; ++i) {
                A(O.FLUXX, i, j, k) = FD * A(O.ER, ub - 1, j, k) / 2.0;
                if (i != NX + 1) {
                    A(O.FLUXY, i, j, k) = 0.0;
                    A(O.FLUXZ, i, j, k) = 0.0;
                }
            }
            A(O.FLUXX, 1, j, k) = O.p.fbottom;
        }
}

/* flux_x(i,j,k) as the reference reads it in the top boundary blocks
 * (ppenergy.f:94-101,712-722): the array is declared (1:nx+1,1:ny,1:nz) but
 * indexed over ghost j, k and i = nx+2, so a read lands on the element with
 * the same column-major offset.  Offsets outside the array read memory that
 * is not part of it; those values only reach ghost cells that the theta
 * boundary blocks and PLANETSHIFT(6) overwrite, and 0 is used here. */
static double flux_x_alias(int i, int j, int k)
{
    long sx = NX + 1, sy = NY;
    long off = (long)(i - 1) + sx * ((long)(j - 1) + sy * (long)(k - 1));
    if (off < 0 || off >= sx * sy * (long)NZ) return 0.0;
    int ii = (int)(off % sx) + 1;
    long r = off / sx;
    int jj = (int)(r % sy) + 1;
    int kk = (int)(r / sy) + 1;
    return A(O.FLUXX, ii, jj, kk);
}

static void fill_evalues(double theta, double delt) /* FILL_EVALUES, ppenergy.f:984-1030 */
{
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                double e1 = -SURXA(i + 1) * A(O.DIFRX, i + 1, j, k) / (SURXB(i) * DXA(i));
                double e2 = -SURXA(i) * A(O.DIFRX, i, j, k) / (SURXB(i) * DXA(i));
                double e3 = -A(O.DIFRY, i, j + 1, k) / (XXB(i) * SURZB(k) * DYA(j));
                double e4 = -A(O.DIFRY, i, j, k) / (XXB(i) * SURZB(k) * DYA(j));
                double e5 = -A(O.DIFRZ, i, j, k + 1) * SURZA(k + 1) / (XXB(i) * SURZB(k) * DZA(k));
                double e6 = -A(O.DIFRZ, i, j, k) * SURZA(k) / (XXB(i) * SURZB(k) * DZA(k));
                double eta = A(O.JETA, i, j, k);
                A(O.EV1, i, j, k) = e1; A(O.EV2, i, j, k) = e2; A(O.EV3, i, j, k) = e3;
                A(O.EV4, i, j, k) = e4; A(O.EV5, i, j, k) = e5; A(O.EV6, i, j, k) = e6;
                A(O.EV7, i, j, k) = 1.0 + (1.0 - eta) * delt * A(O.RH, i, j, k) * A(O.XKP, i, j, k) * FD;
                A(O.EV8, i, j, k) = (eta * (1.0 - theta) - 1.0) * delt;
                A(O.EV16, i, j, k) = e1 + e2 + e3 + e4 + e5 + e6;
            }
}

static void calculatedivf_withf(double* divf, double theta, double delt) /* ppenergy.f:825-878 */
{
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(divf, i, j, k) =
                    ((SURXA(i + 1) / (SURXB(i) * DXA(i))) * A(O.FLUXX, i + 1, j, k) -
                     (SURXA(i) / (SURXB(i) * DXA(i))) * A(O.FLUXX, i, j, k)) +
                    ((1.0 / (XXB(i) * SURZB(k) * DYA(j))) * A(O.FLUXY, i, j + 1, k) -
                     (1.0 / (XXB(i) * SURZB(k) * DYA(j))) * A(O.FLUXY, i, j, k)) +
                    ((SURZA(k + 1) / (XXB(i) * SURZB(k) * DZA(k))) * A(O.FLUXZ, i, j, k + 1) -
                     (SURZA(k) / (XXB(i) * SURZB(k) * DZA(k))) * A(O.FLUXZ, i, j, k));
    fill_evalues(theta, delt);
}

/* ---------------------------------------------------------------- ppsor.f */
static void fill_sor_matrix(void) /* FILL_SOR_MATRIX, ppsor.f:189-277 (no poles) */
{
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                size_t n = I3(i, j, k);
                double e8 = O.EV8[n];
                O.SA[n] = O.EV7[n] + e8 * O.EV16[n];
                O.SB[n] = -e8 * O.EV2[n];
                O.SC[n] = -e8 * O.EV1[n];
                O.SD[n] = -e8 * O.EV4[n];
                O.SE[n] = -e8 * O.EV3[n];
                O.SF[n] = -e8 * O.EV6[n];
                O.SG[n] = -e8 * O.EV5[n];
                int ub = UBND(j, k);
                if (i >= ub) {
                    O.SA[n] = 1.0; O.SB[n] = 0.0; O.SC[n] = 0.0; O.SD[n] = 0.0;
                    O.SE[n] = 0.0; O.SF[n] = 0.0; O.SG[n] = 0.0;
                }
                if (i == ub - 1) O.SC[n] = 0.0;
                if (i == NX || k == 1 || k == NZ) {
                    O.SA[n] = 1.0; O.SB[n] = 0.0; O.SC[n] = 0.0; O.SD[n] = 0.0;
                    O.SE[n] = 0.0; O.SF[n] = 0.0; O.SG[n] = 0.0;
                }
                if (i == NX - 1) O.SC[n] = 0.0;
                if (k == 2) O.SF[n] = 0.0;
                if (k == NZ - 1) O.SG[n] = 0.0;
            }
}

static void b_vector_boundaries(void) /* B_VECTOR_BOUNDARIES, ppsor.f:280-329 */
{
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                if (i >= UBND(j, k) || i == NX || k == 1 || k == NZ) A(O.RHS, i, j, k) = 0.0;
            }
}

/* One red-black SOR iteration (ppsor.f:56-150) on x with fixed ghosts except
 * the periodic phi wrap (PLANETSHIFT after each half sweep).  Returns the L1
 * during-sweep norm.  Sequential in the reference order when nthreads == 1. */
static double sor_iteration(double* x, const double* rhs, double omega, int f)
{
    double rnorm = 0.0;
    for (int color = 0; color < 2; ++color) {
        if (O.nthreads <= 1) {
            for (int k = 1; k <= NZ; ++k)
                for (int j = 1; j <= NY; ++j) {
                    int imin = color == 0 ? 2 - ((j + k) % 2) : 2 - ((j + k + 1) % 2);
                    for (int i = imin; i <= NX; i += 2) {
                        size_t n = I3(i, j, k);
                        double resid = O.SA[n] * x[n] + O.SB[n] * A(x, i - 1, j, k) +
                                       O.SC[n] * A(x, i + 1, j, k) + O.SD[n] * A(x, i, j - 1, k) +
                                       O.SE[n] * A(x, i, j + 1, k) + O.SF[n] * A(x, i, j, k - 1) +
                                       O.SG[n] * A(x, i, j, k + 1) - rhs[n];
                        rnorm = rnorm + fabs(resid);
                        x[n] = x[n] - omega * resid / O.SA[n];
                    }
                }
        } else {
            double part = 0.0;
#pragma omp parallel for reduction(+ : part) schedule(static)
            for (int k = 1; k <= NZ; ++k)
                for (int j = 1; j <= NY; ++j) {
                    int imin = color == 0 ? 2 - ((j + k) % 2) : 2 - ((j + k + 1) % 2);
                    for (int i = imin; i <= NX; i += 2) {
                        size_t n = I3(i, j, k);
                        double resid = O.SA[n] * x[n] + O.SB[n] * A(x, i - 1, j, k) +
                                       O.SC[n] * A(x, i + 1, j, k) + O.SD[n] * A(x, i, j - 1, k) +
                                       O.SE[n] * A(x, i, j + 1, k) + O.SF[n] * A(x, i, j, k - 1) +
                                       O.SG[n] * A(x, i, j, k + 1) - rhs[n];
                        part += fabs(resid);
                        x[n] = x[n] - omega * resid / O.SA[n];
                    }
                }
            rnorm += part;
        }
        planetshift(x, f);
    }
    return rnorm;
}

static double l1_interior(const double* a)
{
    double s = 0.0;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) s = s + fabs(A(a, i, j, k));
    return s;
}

/* SOLVE_WITH_SOR, ppsor.f:2-187 (IRAD=3). */
static int solve_with_sor(int fill, tt_sor_stats* st)
{
    if (fill) {
        fill_sor_matrix();
        b_vector_boundaries();
    }
    memset(O.DER, 0, O.n3 * sizeof(double));
    double rhsnorm = l1_interior(O.RHS);
    int it = 0;
    double rnorm = 0.0;
    int rc = 0;
    for (;;) {
        ++it;
        rnorm = sor_iteration(O.DER, O.RHS, O.p.sorparam, 8);
        if (it == O.p.itmax) { rc = -1; seterr(-1, "MAX COUPLED SOR ITERATION"); break; }
        if (!(rnorm / rhsnorm > O.p.epsmat || it < 1)) break;
    }
    if (st) {
        memset(st, 0, sizeof *st);
        st->sor_iter = it;
        st->rhsnorm = rhsnorm;
        st->rnorm = rnorm;
    }
    return rc;
}

int tor_solve_sor(int fill, tt_sor_stats* st) { return solve_with_sor(fill, st); }

/* ADJUST_SOR_PARAM, ppsor.f:332-354: SORPARAM = 1.01 then += 0.02, fifty
 * times (REAL literals: float values widened to double); every trial is a
 * full SOLVE_WITH_SOR, which refills the matrix from the same EVALs; keep the
 * first omega of minimal SOR_ITER (minloc). */
int tor_adjust_sor_param(double* omega, int32_t* iters)
{
    int32_t cnt[50];
    double om[50];
    O.p.sorparam = (double)1.01f;
    for (int t = 0; t < 50; ++t) {
        tt_sor_stats ss;
        solve_with_sor(1, &ss);
        cnt[t] = ss.sor_iter;
        om[t] = O.p.sorparam;
        O.p.sorparam = O.p.sorparam + (double)0.02f;
    }
    int best = 0;
    for (int t = 1; t < 50; ++t)
        if (cnt[t] < cnt[best]) best = t;
    O.p.sorparam = om[best];
    if (omega) *omega = om[best];
    if (iters) memcpy(iters, cnt, sizeof cnt);
    return 0;
}

void tor_set_adjust(int on) { g_adjust = on; }
double tor_get_sorparam(void) { return O.p.sorparam; }

/* COUPLED_ENERGY_EQS, ppenergy.f:2-791 (IRAD=3, theta=1, conv_parm=0.1,
 * delta_max=0.005, NCOSYS=2, BCXMIN=12, no poles). */
static void energy_bottom_bc(void) /* ppenergy.f:184-195 / :640-651 (BCXMIN=12); :150-167 / :621-638 (9) */
{
    if (O.p.bcxmin == 9) {
        /* fixed flux, free T and Q: the radiative-diffusion gradient of Q that
           carries FBottom, and the optically thick ER = a T^4 at i = 1, 0, -1 */
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j) {
                const double rh = A(O.RH, 1, j, k), cv = A(O.CV, 1, j, k); /* T**3.d0 as T*T*T */
                const double dqdr = -rh * cv * 3.0 * O.p.fbottom * A(O.XKR, 1, j, k) * rh /
                                    (4.0 * ARAD * FD * (A(O.T, 1, j, k) * A(O.T, 1, j, k) * A(O.T, 1, j, k)));
                A(O.Q, 0, j, k) = A(O.Q, 1, j, k) - dqdr * (XXB(1) - XXB(0));
                A(O.Q, -1, j, k) = A(O.Q, 1, j, k) - dqdr * (XXB(1) - XXB(-1));
                for (int i = 1; i >= -1; --i)
                    A(O.ER, i, j, k) = ARAD * pow(A(O.Q, i, j, k) / (A(O.RH, i, j, k) * A(O.CV, i, j, k)), 4.0);
            }
        return;
    }
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            double dqdr = A(O.CV, 1, j, k) * A(O.RH, 1, j, k) * O.p.d_t;
            A(O.Q, 0, j, k) = A(O.Q, 1, j, k) + dqdr * (XXB(0) - XXB(1));
            A(O.Q, -1, j, k) = A(O.Q, 1, j, k) + dqdr * (XXB(-1) - XXB(1));
            A(O.ER, 0, j, k) = ARAD * pow(A(O.Q, 0, j, k) / (A(O.RH, 1, j, k) * A(O.CV, 1, j, k)), 4.0);
            A(O.ER, -1, j, k) = ARAD * pow(A(O.Q, -1, j, k) / (A(O.RH, 1, j, k) * A(O.CV, 1, j, k)), 4.0);
        }
}

static void energy_top_bc(void) /* ppenergy.f:80-104 / :681-723 */
{
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j) {
            int ub = UBND(j, k);
            double qup = A(O.RH, ub - 1, j, k) * A(O.CV, ub - 1, j, k) * A(O.T, ub, j, k);
            for (int i = ub; i <= NX + 2; ++i) {
                A(O.Q, i, j, k) = qup;
                A(O.ER, i, j, k) = fabs(2.0 * flux_x_alias(i, j, k) / FD);
            }
        }
}

static void energy_z_bc(void) /* ppenergy.f:204-227 / :727-757 */
{
    for (int j = -1; j <= NY + 2; ++j)
        for (int i = -1; i <= NX + 2; ++i) {
            A(O.Q, i, j, 1) = A(O.Q, i, j, 2);
            A(O.ER, i, j, 1) = A(O.ER, i, j, 2);
            A(O.Q, i, j, 0) = A(O.Q, i, j, 1);
            A(O.Q, i, j, -1) = A(O.Q, i, j, 1);
            A(O.ER, i, j, 0) = A(O.ER, i, j, 1);
            A(O.ER, i, j, -1) = A(O.ER, i, j, 1);
        }
    for (int j = -1; j <= NY + 2; ++j)
        for (int i = -1; i <= NX + 2; ++i) {
            A(O.Q, i, j, NZ) = A(O.Q, i, j, NZ - 1);
            A(O.ER, i, j, NZ) = A(O.ER, i, j, NZ - 1);
            A(O.Q, i, j, NZ + 1) = A(O.Q, i, j, NZ);
            A(O.Q, i, j, NZ + 2) = A(O.Q, i, j, NZ);
            A(O.ER, i, j, NZ + 1) = A(O.ER, i, j, NZ);
            A(O.ER, i, j, NZ + 2) = A(O.ER, i, j, NZ);
        }
}

int tor_coupled_energy(double delt, int iii, tt_energy_stats* st)
{
    /* conv_parm = 0.1 is a REAL literal (ppenergy.f:49): the damping in use
     * is the nearest float, 0.100000001490116 */
    const double delta_max = 0.005, conv_parm = (double)0.1f, theta = 1.0;
    int rc;
    if (O.p.bcxmin != 12 && O.p.bcxmin != 9) return seterr(1, "COUPLED_ENERGY_EQS: BCXMIN must be 9 or 12");
    if (O.p.bcxmin == 9 && g_visc.ieta == 5) return seterr(1, "COUPLED_ENERGY_EQS: BCXMIN 9 with IETA 5 is not supported");
    g_iii = iii;
    double* dQ = alloc3();
    double* divF = alloc3();
    double* divF_old = alloc3();
    double* dFQ = alloc3();
    double* dFR = alloc3();
    double* B = alloc3();
    double* Q_old = alloc3();
    if (st) memset(st, 0, sizeof *st);
    memset(O.DER, 0, O.n3 * sizeof(double));
    O.energy_iter = 0;
    rc = tor_bounds();
    if (rc) goto done;
    for (size_t n = 0; n < O.n3; ++n) O.Q[n] = O.RH[n] * O.CV[n] * O.T[n];
    if (O.p.f_incident > 1) {
        rc = kappa(); if (rc) goto done;
        rc = stellarinput_heating(); if (rc) goto done;
    }
    rc = calculateflux(); if (rc) goto done;
    flux_bc();
    energy_top_bc();
    energy_bottom_bc();
    energy_z_bc();
    planetshift(O.ER, 6);
    memcpy(O.ER_OLD, O.ER, O.n3 * sizeof(double));
    memcpy(Q_old, O.Q, O.n3 * sizeof(double));
    rc = calculateflux(); if (rc) goto done;
    flux_bc();
    calculatedivf_withf(divF_old, theta, delt);
    for (;;) {
        O.energy_iter++;
        planetshift(O.ER, 6);
        rc = calculateflux(); if (rc) goto done;
        flux_bc();
        calculatedivf_withf(divF, theta, delt);
#pragma omp parallel for
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i) {
                    size_t n = I3(i, j, k);
                    double rh = O.RH[n], cv = O.CV[n], q = O.Q[n], er = O.ER[n], kp = O.XKP[n];
                    B[n] = ARAD * pow(q / (rh * cv), 4.0);
                    double heat = 0.0;
                    if (O.p.f_incident > 1) {
                        heat = O.SIN[n] / (rh * O.XKA[n]);
                        O.FQ[n] = q - Q_old[n] +
                                  delt * ((1.0 - theta) * (divF_old[n] - divF[n]) +
                                          (rh * kp * FD * (B[n] - er)) - O.SIN[n]);
                    } else {
                        O.FQ[n] = q - Q_old[n] +
                                  delt * ((1.0 - theta) * (divF_old[n] - divF[n]) +
                                          rh * kp * FD * (B[n] - er));
                    }
                    O.FR[n] = er - O.ER_OLD[n] + delt * (divF[n] - rh * kp * FD * (B[n] - er));
                    double t3 = pow(q / (rh * cv), 3.0);
                    if (O.p.f_incident > 1)
                        dFQ[n] = 1.0 + (delt / cv) * (kp * 4.0 * ARAD * FD * t3 +
                                                      (FD * (B[n] - er) * O.DKP[n]) - heat * O.DKA[n]);
                    else
                        dFQ[n] = 1.0 + (delt / cv) * (kp * 4.0 * ARAD * FD * t3 + FD * (B[n] - er) * O.DKP[n]);
                    dFR[n] = -(delt / cv) * (kp * 4.0 * ARAD * FD * t3 + FD * (B[n] - er) * O.DKP[n]);
                    if (dFQ[n] != 0.0) O.JETA[n] = -dFR[n] / dFQ[n];
                    else O.JETA[n] = 0.0;
                    O.RHS[n] = -O.FR[n] - O.JETA[n] * O.FQ[n];
                }
        fill_evalues(theta, delt);
        /* ppenergy.f:372-374 */
        if (g_adjust && O.energy_iter == 1 && (iii % 10000 == 0 || iii == 1)) tor_adjust_sor_param(NULL, NULL);
        tt_sor_stats ss;
        int src = solve_with_sor(1, &ss);
        if (st) {
            st->sor_iter_total += ss.sor_iter;
            st->sor_iter_last = ss.sor_iter;
            if (O.energy_iter == 1) st->sor_iter_first = ss.sor_iter;
        }
        (void)src;
        double locmq = 0.0, locme = 0.0;
        int bad = 0;
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j) {
                int xmax = UBND(j, k) - 1, xmin = O.p.bcxmin == 12 ? 1 : 2; /* ppenergy.f:390-396 */
                for (int i = xmin; i <= xmax; ++i) {
                    size_t n = I3(i, j, k);
                    double Q1 = -delt * (1.0 - theta);
                    double Q2 = delt * O.RH[n] * O.XKP[n] * FD;
                    double d = O.DER[n];
                    dQ[n] = (1.0 / dFQ[n]) *
                            (Q1 * O.EV16[n] * d - Q1 * O.EV1[n] * A(O.DER, i + 1, j, k) -
                             Q1 * O.EV2[n] * A(O.DER, i - 1, j, k) - Q1 * O.EV3[n] * A(O.DER, i, j + 1, k) -
                             Q1 * O.EV4[n] * A(O.DER, i, j - 1, k) - Q1 * O.EV5[n] * A(O.DER, i, j, k + 1) -
                             Q1 * O.EV6[n] * A(O.DER, i, j, k - 1) + Q2 * d - O.FQ[n]);
                    if ((O.Q[n] + conv_parm * dQ[n]) <= 0.0) bad = 1;
                    double a = fabs(dQ[n] / O.Q[n]);
                    double b = fabs(d / O.ER[n]);
                    locmq = locmq > a ? locmq : a;
                    locme = locme > b ? locme : b;
                }
            }
        if (bad) { rc = seterr(2, "ERROR IN DELTA_Q: non-positive value of Q"); goto done; }
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j) {
                int xmax = UBND(j, k) - 1, xmin = O.p.bcxmin == 12 ? 1 : 2; /* ppenergy.f:390-396 */
                for (int i = xmin; i <= xmax; ++i) {
                    size_t n = I3(i, j, k);
                    O.ER[n] = O.ER[n] + conv_parm * O.DER[n];
                    O.Q[n] = O.Q[n] + conv_parm * dQ[n];
                }
            }
        energy_bottom_bc();
        energy_top_bc();
        energy_z_bc();
        if (st) { st->max_dq = locmq; st->max_de = locme; st->energy_iter = O.energy_iter; }
        if (O.energy_iter == O.p.itmax) { rc = seterr(-2, "Max iteration reached in coupled_energy_eqs"); break; }
        if (!(locmq > delta_max)) break;
    }
    for (size_t n = 0; n < O.n3; ++n) {
        O.DER[n] = O.ER[n] - O.ER_OLD[n];
        O.T[n] = O.Q[n] / (O.RH[n] * O.CV[n]);
    }
done:
    free(Q_old);
    free(dQ); free(divF); free(divF_old); free(dFQ); free(dFR); free(B);
    return rc;
}

/* --------------------------------------------------------------- ppadv.f */
/* van Leer slope of the reference's line sweeps: grad = 2 d1 d2 /
 * (d1*w_hi + d2*w_lo) when d1*d2 > 0, else 0 (e.g. ppadv.f:568-575). */
int tor_advect3d(double delt) /* ADVECT3D, ppadv.f:525-1577 (IRHT=1, ISOTH=0, IETOT=0) */
{
    const double half = 0.5, two = 2.0;
    int rc = tor_bounds();
    if (rc) return rc;
    double* FLMX = alloc3(); double* FLMY = alloc3(); double* FLMZ = alloc3();
    double* FLX = alloc3(); double* FLY = alloc3(); double* FLZ = alloc3();
    double* ED = alloc3(); double* VD = alloc3(); double* HD = alloc3();
    int ntop = (NX > NY ? NX : NY); if (NZ > ntop) ntop = NZ; ntop += 2;
#define LINEBUF double S[ntop + 4], GRAD[ntop + 4], GRAD1[ntop + 4]; int IND[ntop + 4];
    /* ---- mass fluxes ---- */
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            LINEBUF
            for (int i = 1; i <= NX + 1; ++i) {
                IND[i] = i; S[i] = A(O.V, i, j, k) * delt;
                if (A(O.V, i, j, k) > 0.0) IND[i] = i - 1;
            }
            for (int i = 0; i <= NX + 1; ++i) {
                GRAD1[i] = (A(O.RH, i, j, k) - A(O.RH, i - 1, j, k)) * (A(O.RH, i + 1, j, k) - A(O.RH, i, j, k));
                GRAD[i] = 0.0;
                if (GRAD1[i] > 0.0)
                    GRAD[i] = two * GRAD1[i] / ((A(O.RH, i, j, k) - A(O.RH, i - 1, j, k)) * DXB(i + 1) +
                                                (A(O.RH, i + 1, j, k) - A(O.RH, i, j, k)) * DXB(i));
            }
            for (int i = 1; i <= NX + 1; ++i)
                A(FLMX, i, j, k) = (SURXA(i) - CCXA(i) * S[i]) * S[i] *
                                   (A(O.RH, IND[i], j, k) + (XXA(i) - XXB(IND[i]) - half * S[i]) * GRAD[IND[i]]);
        }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int j = 1; j <= NY + 1; ++j) {
                S[j] = A(O.G, i, j, k) * delt; IND[j] = j;
                if (A(O.G, i, j, k) > 0.0) IND[j] = j - 1;
            }
            for (int j = 0; j <= NY + 1; ++j) {
                GRAD1[j] = (A(O.RH, i, j, k) - A(O.RH, i, j - 1, k)) * (A(O.RH, i, j + 1, k) - A(O.RH, i, j, k));
                GRAD[j] = 0.0;
                if (GRAD1[j] > 0.0)
                    GRAD[j] = two * GRAD1[j] / ((A(O.RH, i, j, k) - A(O.RH, i, j - 1, k)) * DYB(j + 1) +
                                                (A(O.RH, i, j + 1, k) - A(O.RH, i, j, k)) * DYB(j));
            }
            for (int j = 1; j <= NY + 1; ++j)
                A(FLMY, i, j, k) = S[j] * (SURYA(j) + CCYA(j) * S[j]) *
                                   (A(O.RH, i, IND[j], k) + (XYA(j) - XYB(IND[j]) - half * S[j]) * GRAD[IND[j]]);
        }
#pragma omp parallel for
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int k = 1; k <= NZ + 1; ++k) {
                IND[k] = k; S[k] = A(O.H, i, j, k) * delt;
                if (A(O.H, i, j, k) > 0.0) IND[k] = k - 1;
            }
            for (int k = 0; k <= NZ + 1; ++k) {
                GRAD1[k] = (A(O.RH, i, j, k) - A(O.RH, i, j, k - 1)) * (A(O.RH, i, j, k + 1) - A(O.RH, i, j, k));
                GRAD[k] = 0.0;
                if (GRAD1[k] > 0.0)
                    GRAD[k] = two * GRAD1[k] / ((A(O.RH, i, j, k) - A(O.RH, i, j, k - 1)) * DZB(k + 1) +
                                                (A(O.RH, i, j, k + 1) - A(O.RH, i, j, k)) * DZB(k));
            }
            for (int k = 1; k <= NZ + 1; ++k)
                A(FLMZ, i, j, k) = (SURZA(k) + CCZA(k) * S[k]) * S[k] *
                                   (A(O.RH, i, j, IND[k]) + (XZA(k) - XZB(IND[k]) - half * S[k]) * GRAD[IND[k]]);
        }
    /* flux-plane exchanges, ppadv.f:654-721: only the periodic phi ones move */
    for (int k = 1; k <= NZ; ++k)
        for (int i = 1; i <= NX + 1; ++i) A(FLMX, i, 0, k) = A(FLMX, i, NY, k);
    for (int k = 1; k <= NZ + 1; ++k)
        for (int i = 1; i <= NX; ++i) A(FLMZ, i, 0, k) = A(FLMZ, i, NY, k);
    tor_qdens();
    /* ---- thermal energy ED = T*RH*CV ---- */
    for (size_t n = 0; n < O.n3; ++n) ED[n] = O.T[n] * O.RH[n] * O.CV[n];
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            LINEBUF
            for (int i = 1; i <= NX + 1; ++i) {
                IND[i] = i; S[i] = A(O.V, i, j, k) * delt;
                if (A(O.V, i, j, k) > 0.0) IND[i] = i - 1;
            }
            for (int i = 0; i <= NX + 1; ++i) {
                GRAD1[i] = (A(ED, i, j, k) - A(ED, i - 1, j, k)) * (A(ED, i + 1, j, k) - A(ED, i, j, k));
                GRAD[i] = 0.0;
                if (GRAD1[i] > 0.0)
                    GRAD[i] = two * GRAD1[i] / ((A(ED, i, j, k) - A(ED, i - 1, j, k)) * DXB(i + 1) +
                                                (A(ED, i + 1, j, k) - A(ED, i, j, k)) * DXB(i));
            }
            for (int i = 1; i <= NX + 1; ++i)
                A(FLX, i, j, k) = (SURXA(i) - CCXA(i) * S[i]) * S[i] *
                                  (A(ED, IND[i], j, k) + (XXA(i) - XXB(IND[i]) - half * S[i]) * GRAD[IND[i]]);
        }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int j = 1; j <= NY + 1; ++j) {
                IND[j] = j; S[j] = A(O.G, i, j, k) * delt;
                if (A(O.G, i, j, k) > 0.0) IND[j] = j - 1;
            }
            for (int j = 0; j <= NY + 1; ++j) {
                GRAD1[j] = (A(ED, i, j, k) - A(ED, i, j - 1, k)) * (A(ED, i, j + 1, k) - A(ED, i, j, k));
                GRAD[j] = 0.0;
                if (GRAD1[j] > 0.0)
                    GRAD[j] = two * GRAD1[j] / ((A(ED, i, j, k) - A(ED, i, j - 1, k)) * DYB(j + 1) +
                                                (A(ED, i, j + 1, k) - A(ED, i, j, k)) * DYB(j));
            }
            for (int j = 1; j <= NY + 1; ++j)
                A(FLY, i, j, k) = (SURYA(j) - CCYA(j) * S[j]) * S[j] *
                                  (A(ED, i, IND[j], k) + (XYA(j) - XYB(IND[j]) - half * S[j]) * GRAD[IND[j]]);
        }
#pragma omp parallel for
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int k = 1; k <= NZ + 1; ++k) {
                IND[k] = k; S[k] = A(O.H, i, j, k) * delt;
                if (A(O.H, i, j, k) > 0.0) IND[k] = k - 1;
            }
            for (int k = 0; k <= NZ + 1; ++k) {
                GRAD1[k] = (A(ED, i, j, k) - A(ED, i, j, k - 1)) * (A(ED, i, j, k + 1) - A(ED, i, j, k));
                GRAD[k] = 0.0;
                if (GRAD1[k] > 0.0)
                    GRAD[k] = two * GRAD1[k] / ((A(ED, i, j, k) - A(ED, i, j, k - 1)) * DZB(k + 1) +
                                                (A(ED, i, j, k + 1) - A(ED, i, j, k)) * DZB(k));
            }
            for (int k = 1; k <= NZ + 1; ++k)
                A(FLZ, i, j, k) = (SURZA(k) - CCZA(k) * S[k]) * S[k] *
                                  (A(ED, i, j, IND[k]) + (XZA(k) - XZB(IND[k]) - half * S[k]) * GRAD[IND[k]]);
        }
    /* the thermal advection flux kept (ppadv.f:857-865); PRINTOUT writes
       cells 1..n of each component */
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                A(O.CF1, i, j, k) = A(FLX, i, j, k);
                A(O.CF2, i, j, k) = A(FLY, i, j, k);
                A(O.CF3, i, j, k) = A(FLZ, i, j, k);
            }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(ED, i, j, k) = A(ED, i, j, k) - (A(FLX, i + 1, j, k) - A(FLX, i, j, k)) / VOLXB(i) -
                                 (A(FLY, i, j + 1, k) - A(FLY, i, j, k)) / VOLYB(j) -
                                 (A(FLZ, i, j, k + 1) - A(FLZ, i, j, k)) / VOLZB(k);
    /* ---- radial momentum V ---- */
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            LINEBUF
            for (int i = 1; i <= NX; ++i) {
                S[i] = (A(O.V, i, j, k) + A(O.V, i + 1, j, k)) * delt / 2.0;
                IND[i] = i + 1;
                if (S[i] > 0.0) IND[i] = i;
            }
            for (int i = 1; i <= NX + 1; ++i) {
                GRAD1[i] = (A(O.V, i, j, k) - A(O.V, i - 1, j, k)) * (A(O.V, i + 1, j, k) - A(O.V, i, j, k));
                GRAD[i] = 0.0;
                if (GRAD1[i] > 0.0)
                    GRAD[i] = two * GRAD1[i] / ((A(O.V, i + 1, j, k) - A(O.V, i, j, k)) * DXA(i) +
                                                (A(O.V, i, j, k) - A(O.V, i - 1, j, k)) * DXA(i + 1));
            }
            for (int i = 1; i <= NX; ++i)
                A(FLX, i, j, k) = half * (A(FLMX, i, j, k) + A(FLMX, i + 1, j, k)) *
                                  (A(O.V, IND[i], j, k) + (XXB(i) - XXA(IND[i]) - half * S[i]) * GRAD[IND[i]]);
        }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int i = 2; i <= NX; ++i) {
            LINEBUF
            for (int j = 1; j <= NY + 1; ++j) {
                S[j] = (DDX1(i) * A(O.G, i, j, k) + DDX0(i) * A(O.G, i - 1, j, k)) * delt;
                IND[j] = j;
                if (S[j] > 0.0) IND[j] = j - 1;
            }
            for (int j = 0; j <= NY + 1; ++j) {
                GRAD1[j] = (A(O.V, i, j, k) - A(O.V, i, j - 1, k)) * (A(O.V, i, j + 1, k) - A(O.V, i, j, k));
                GRAD[j] = 0.0;
                if (GRAD1[j] > 0.0)
                    GRAD[j] = two * GRAD1[j] / ((A(O.V, i, j + 1, k) - A(O.V, i, j, k)) * DYB(j) +
                                                (A(O.V, i, j, k) - A(O.V, i, j - 1, k)) * DYB(j + 1));
            }
            for (int j = 1; j <= NY + 1; ++j)
                A(FLY, i, j, k) = (DDX1(i) * A(FLMY, i, j, k) + DDX0(i) * A(FLMY, i - 1, j, k)) *
                                  (A(O.V, i, IND[j], k) + (XYA(j) - XYB(IND[j]) - half * S[j]) * GRAD[IND[j]]);
        }
#pragma omp parallel for
    for (int j = 1; j <= NY; ++j)
        for (int i = 2; i <= NX; ++i) {
            LINEBUF
            for (int k = 1; k <= NZ + 1; ++k) {
                S[k] = (DDX1(i) * A(O.H, i, j, k) + DDX0(i) * A(O.H, i - 1, j, k)) * delt;
                IND[k] = k;
                if (S[k] > 0.0) IND[k] = k - 1;
            }
            for (int k = 0; k <= NZ + 1; ++k) {
                GRAD1[k] = (A(O.V, i, j, k) - A(O.V, i, j, k - 1)) * (A(O.V, i, j, k + 1) - A(O.V, i, j, k));
                GRAD[k] = 0.0;
                if (GRAD1[k] > 0.0)
                    GRAD[k] = two * GRAD1[k] / ((A(O.V, i, j, k + 1) - A(O.V, i, j, k)) * DZB(k) +
                                                (A(O.V, i, j, k) - A(O.V, i, j, k - 1)) * DZB(k + 1));
            }
            for (int k = 1; k <= NZ + 1; ++k)
                A(FLZ, i, j, k) = (DDX1(i) * A(FLMZ, i, j, k) + DDX0(i) * A(FLMZ, i - 1, j, k)) *
                                  (A(O.V, i, j, IND[k]) + (XZA(k) - XZB(IND[k]) - half * S[k]) * GRAD[IND[k]]);
        }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 2; i <= NX; ++i)
                A(VD, i, j, k) = A(O.V, i, j, k) * A(O.RHQX, i, j, k) -
                                 (A(FLX, i, j, k) - A(FLX, i - 1, j, k)) / VOLXA(i) -
                                 (A(FLY, i, j + 1, k) - A(FLY, i, j, k)) / VOLYB(j) -
                                 (A(FLZ, i, j, k + 1) - A(FLZ, i, j, k)) / VOLZB(k);
    /* ---- azimuthal angular momentum GD ---- */
#pragma omp parallel for
    for (int k = -1; k <= NZ + 2; ++k)
        for (int j = 0; j <= NY + 2; ++j)
            for (int i = -1; i <= NX + 2; ++i)
                A(O.GD, i, j, k) = (A(O.G, i, j, k) + O.p.omrot) * (GEOXG(i) * GEOXG(i)) * (GEOZG(k) * GEOZG(k));
    planetshift(O.GD, 11);
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            LINEBUF
            for (int i = 1; i <= NX + 1; ++i) {
                S[i] = (DDY1(j) * A(O.V, i, j, k) + DDY0(j) * A(O.V, i, j - 1, k)) * delt;
                IND[i] = i;
                if (S[i] > 0.0) IND[i] = i - 1;
            }
            for (int i = 0; i <= NX + 1; ++i) {
                GRAD1[i] = (A(O.GD, i, j, k) - A(O.GD, i - 1, j, k)) * (A(O.GD, i + 1, j, k) - A(O.GD, i, j, k));
                GRAD[i] = 0.0;
                if (GRAD1[i] > 0.0)
                    GRAD[i] = two * GRAD1[i] / ((A(O.GD, i + 1, j, k) - A(O.GD, i, j, k)) * DXB(i) +
                                                (A(O.GD, i, j, k) - A(O.GD, i - 1, j, k)) * DXB(i + 1));
            }
            for (int i = 1; i <= NX + 1; ++i)
                A(FLX, i, j, k) = (DDY1(j) * A(FLMX, i, j, k) + DDY0(j) * A(FLMX, i, j - 1, k)) *
                                  (A(O.GD, IND[i], j, k) + (XXA(i) - XXB(IND[i]) - half * S[i]) * GRAD[IND[i]]);
        }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int j = 1; j <= NY; ++j) {
                S[j] = (DCY0(j + 1) * A(O.G, i, j, k) + DCY1(j + 1) * A(O.G, i, j + 1, k)) * delt;
                IND[j] = j + 1;
                if (S[j] > 0.0) IND[j] = j;
            }
            for (int j = 1; j <= NY + 1; ++j) {
                GRAD1[j] = (A(O.GD, i, j, k) - A(O.GD, i, j - 1, k)) * (A(O.GD, i, j + 1, k) - A(O.GD, i, j, k));
                GRAD[j] = 0.0;
                if (GRAD1[j] > 0.0)
                    GRAD[j] = two * GRAD1[j] / ((A(O.GD, i, j + 1, k) - A(O.GD, i, j, k)) * DYA(j) +
                                                (A(O.GD, i, j, k) - A(O.GD, i, j - 1, k)) * DYA(j + 1));
            }
            for (int j = 1; j <= NY; ++j)
                A(FLY, i, j, k) = (DCY0(j + 1) * A(FLMY, i, j, k) + DCY1(j + 1) * A(FLMY, i, j + 1, k)) *
                                  (A(O.GD, i, IND[j], k) + (XYB(j) - XYA(IND[j]) - half * S[j]) * GRAD[IND[j]]);
        }
#pragma omp parallel for
    for (int i = 1; i <= NX; ++i)
        for (int j = 1; j <= NY; ++j) {
            LINEBUF
            for (int k = 1; k <= NZ + 1; ++k) {
                S[k] = (DDY1(j) * A(O.H, i, j, k) + DDY0(j) * A(O.H, i, j - 1, k)) * delt;
                IND[k] = k;
                if (S[k] > 0.0) IND[k] = k - 1;
            }
            for (int k = 0; k <= NZ + 1; ++k) {
                GRAD1[k] = (A(O.GD, i, j, k) - A(O.GD, i, j, k - 1)) * (A(O.GD, i, j, k + 1) - A(O.GD, i, j, k));
                GRAD[k] = 0.0;
                if (GRAD1[k] > 0.0)
                    GRAD[k] = two * GRAD1[k] / ((A(O.GD, i, j, k + 1) - A(O.GD, i, j, k)) * DZB(k) +
                                                (A(O.GD, i, j, k) - A(O.GD, i, j, k - 1)) * DZB(k + 1));
            }
            for (int k = 1; k <= NZ + 1; ++k)
                A(FLZ, i, j, k) = (DDY1(j) * A(FLMZ, i, j, k) + DDY0(j) * A(FLMZ, i, j - 1, k)) *
                                  (A(O.GD, i, j, IND[k]) + (XZA(k) - XZB(IND[k]) - half * S[k]) * GRAD[IND[k]]);
        }
    for (int k = 1; k <= NZ; ++k)
        for (int i = 1; i <= NX; ++i) A(FLY, i, 0, k) = A(FLY, i, NY, k);
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.GD, i, j, k) = A(O.GD, i, j, k) * A(O.RHQY, i, j, k) -
                                   (A(FLX, i + 1, j, k) - A(FLX, i, j, k)) / VOLXB(i) -
                                   (A(FLY, i, j, k) - A(FLY, i, j - 1, k)) / VOLYA(j) -
                                   (A(FLZ, i, j, k + 1) - A(FLZ, i, j, k)) / VOLZB(k);
    /* ---- meridional angular momentum HD ---- */
#pragma omp parallel for
    for (int k = 0; k <= NZ + 2; ++k)
        for (int j = -1; j <= NY + 2; ++j)
            for (int i = -1; i <= NX + 2; ++i) A(HD, i, j, k) = A(O.H, i, j, k) * (GEOXH(i) * GEOXH(i));
#pragma omp parallel for
    for (int k = 2; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            LINEBUF
            for (int i = 1; i <= NX + 1; ++i) {
                S[i] = (DDZ1(k) * A(O.V, i, j, k) + DDZ0(k) * A(O.V, i, j, k - 1)) * delt;
                IND[i] = i;
                if (S[i] > 0.0) IND[i] = i - 1;
            }
            for (int i = 0; i <= NX + 1; ++i) {
                GRAD1[i] = (A(HD, i, j, k) - A(HD, i - 1, j, k)) * (A(HD, i + 1, j, k) - A(HD, i, j, k));
                GRAD[i] = 0.0;
                if (GRAD1[i] > 0.0)
                    GRAD[i] = two * GRAD1[i] / ((A(HD, i + 1, j, k) - A(HD, i, j, k)) * DXB(i) +
                                                (A(HD, i, j, k) - A(HD, i - 1, j, k)) * DXB(i + 1));
            }
            for (int i = 1; i <= NX + 1; ++i)
                A(FLX, i, j, k) = (DDZ1(k) * A(FLMX, i, j, k) + DDZ0(k) * A(FLMX, i, j, k - 1)) *
                                  (A(HD, IND[i], j, k) + (XXA(i) - XXB(IND[i]) - half * S[i]) * GRAD[IND[i]]);
        }
#pragma omp parallel for
    for (int k = 2; k <= NZ; ++k)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int j = 1; j <= NY + 1; ++j) {
                S[j] = (DDZ1(k) * A(O.G, i, j, k) + DDZ0(k) * A(O.G, i, j, k - 1)) * delt;
                IND[j] = j;
                if (S[j] > 0.0) IND[j] = j - 1;
            }
            for (int j = 0; j <= NY + 1; ++j) {
                GRAD1[j] = (A(HD, i, j, k) - A(HD, i, j - 1, k)) * (A(HD, i, j + 1, k) - A(HD, i, j, k));
                GRAD[j] = 0.0;
                if (GRAD1[j] > 0.0)
                    GRAD[j] = two * GRAD1[j] / ((A(HD, i, j + 1, k) - A(HD, i, j, k)) * DYB(j) +
                                                (A(HD, i, j, k) - A(HD, i, j - 1, k)) * DYB(j + 1));
            }
            for (int j = 1; j <= NY + 1; ++j)
                A(FLY, i, j, k) = (DDZ1(k) * A(FLMY, i, j, k) + DDZ0(k) * A(FLMY, i, j, k - 1)) *
                                  (A(HD, i, IND[j], k) + (XYA(j) - XYB(IND[j]) - half * S[j]) * GRAD[IND[j]]);
        }
#pragma omp parallel for
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            LINEBUF
            for (int k = 1; k <= NZ; ++k) {
                S[k] = half * (A(O.H, i, j, k) + A(O.H, i, j, k + 1)) * delt;
                IND[k] = k + 1;
                if (S[k] > 0.0) IND[k] = k;
            }
            for (int k = 1; k <= NZ + 1; ++k) {
                GRAD1[k] = (A(HD, i, j, k) - A(HD, i, j, k - 1)) * (A(HD, i, j, k + 1) - A(HD, i, j, k));
                GRAD[k] = 0.0;
                if (GRAD1[k] > 0.0)
                    GRAD[k] = two * GRAD1[k] / ((A(HD, i, j, k + 1) - A(HD, i, j, k)) * DZA(k) +
                                                (A(HD, i, j, k) - A(HD, i, j, k - 1)) * DZA(k + 1));
            }
            for (int k = 1; k <= NZ; ++k)
                A(FLZ, i, j, k) = half * (A(FLMZ, i, j, k) + A(FLMZ, i, j, k + 1)) *
                                  (A(HD, i, j, IND[k]) + (XZB(k) - XZA(IND[k]) - half * S[k]) * GRAD[IND[k]]);
        }
#pragma omp parallel for
    for (int k = 2; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(HD, i, j, k) = A(HD, i, j, k) * A(O.RHQZ, i, j, k) -
                                 (A(FLX, i + 1, j, k) - A(FLX, i, j, k)) / VOLXB(i) -
                                 (A(FLY, i, j + 1, k) - A(FLY, i, j, k)) / VOLYB(j) -
                                 (A(FLZ, i, j, k) - A(FLZ, i, j, k - 1)) / VOLZA(k);
    /* ---- density and new primitive variables ---- */
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.RH, i, j, k) = A(O.RH, i, j, k) - (A(FLMX, i + 1, j, k) - A(FLMX, i, j, k)) / VOLXB(i) -
                                   (A(FLMY, i, j + 1, k) - A(FLMY, i, j, k)) / VOLYB(j) -
                                   (A(FLMZ, i, j, k + 1) - A(FLMZ, i, j, k)) / VOLZB(k);
    rc = tor_bounds();
    if (rc) goto out;
    tor_qdens();
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j) {
            for (int i = 2; i <= NX; ++i) A(O.V, i, j, k) = A(VD, i, j, k) / A(O.RHQX, i, j, k);
            for (int i = 1; i <= NX; ++i)
                A(O.G, i, j, k) = A(O.GD, i, j, k) / (A(O.RHQY, i, j, k) * (GEOXG(i) * GEOXG(i)) *
                                                      (GEOZG(k) * GEOZG(k))) - O.p.omrot;
        }
#pragma omp parallel for
    for (int k = 2; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.H, i, j, k) = A(HD, i, j, k) / (A(O.RHQZ, i, j, k) * (GEOXH(i) * GEOXH(i)));
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i)
                A(O.T, i, j, k) = A(ED, i, j, k) / (A(O.RH, i, j, k) * A(O.CV, i, j, k));
out:
#undef LINEBUF
    free(FLMX); free(FLMY); free(FLMZ); free(FLX); free(FLY); free(FLZ);
    free(ED); free(VD); free(HD);
    return rc;
}

/* --------------------------------------------------------------- ppfor.f */
static void centrifcoriolis(void) /* CENTRIFCORIOLIS, ppfor.f:400-492, NCOSYS=2 */
{
    const double om = O.p.omrot;
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                double G1 = DCY0(j + 1) * A(O.G, i - 1, j, k) + DCY1(j + 1) * A(O.G, i - 1, j + 1, k) + om;
                double G2 = DCY0(j + 1) * A(O.G, i, j, k) + DCY1(j + 1) * A(O.G, i, j + 1, k) + om;
                double GQk = DDX0KEP(i) * G1 + DDX1KEP(i) * G2;
                double H1 = DDZ0(k + 1) * A(O.H, i - 1, j, k) + DDZ1(k + 1) * A(O.H, i - 1, j, k + 1);
                double H2 = DDZ0(k + 1) * A(O.H, i, j, k) + DDZ1(k + 1) * A(O.H, i, j, k + 1);
                double HQ = DDX0(i) * H1 + DDX1(i) * H2;
                double t = GQk / XXA(i) * GEOZG(k);
                A(O.CCX, i, j, k) = t * t + XXA(i) * (HQ * HQ);
                A(O.CCY, i, j, k) = 0.0;
                G1 = DCY0(j + 1) * A(O.G, i, j, k - 1) + DCY1(j + 1) * A(O.G, i, j + 1, k - 1);
                G2 = DCY0(j + 1) * A(O.G, i, j, k) + DCY1(j + 1) * A(O.G, i, j + 1, k);
                double GQ = DDZ1(k) * G2 + DDZ0(k) * G1 + om;
                A(O.CCZ, i, j, k) = -XXB(i) * GQ * GQ * sin(XZA(k)) * cos(XZA(k));
            }
}

/* ETACALC, ppvisc.f:375-484: the dynamic viscosity for IETA 0 and 5
 * (constant kinematic viscosity) and 4 (molecular, capped at XNUE) */
static int etacalc(void)
{
    const double xnue = g_visc.xnue;
    if (g_visc.ieta == 0 || g_visc.ieta == 5) {
        for (size_t n = 0; n < O.n3; ++n) O.ETA[n] = O.RH[n] * xnue;
        return 0;
    }
    if (g_visc.ieta == 2) { /* ramp in time, ppvisc.f:402-420 */
        const double day = (double)8.64e4f, xs = 9.0;
        const double t0 = (double)282.5f * day, t1 = t0 + 10.0 * day, ht = g_lim.hydrotime;
        double xv;
        if (ht < t0) xv = pow(10.0, xs);
        else if (ht >= t0 && ht < t1) {
            const double ms = (log10(xnue) - xs) / (t1 - t0);
            xv = pow(10.0, ms * (ht - t0) + xs);
        } else xv = xnue;
        for (size_t n = 0; n < O.n3; ++n) O.ETA[n] = O.RH[n] * xv;
        return 0;
    }
    if (g_visc.ieta == 3) { /* raised at the bottom boundary, ppvisc.f:427-441 */
        if (NX < 13) return seterr(1, "ETACALC: IETA 3 needs nx >= 13");
        for (size_t n = 0; n < O.n3; ++n) O.ETA[n] = O.RH[n] * xnue;
        const double extra = 4.0, ms = extra / (1.0 - 1.0 * 15);
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j) {
                for (int i = 1; i <= 15; ++i) A(O.ETA, i, j, k) = A(O.RH, i, j, k) * pow(10.0, ms * (double)(i - 15) + log10(xnue));
                for (int i = -1; i <= 0; ++i) A(O.ETA, i, j, k) = A(O.RH, i, j, k) * pow(10.0, extra + log10(xnue));
            }
        return 0;
    }
    if (g_visc.ieta != 4) return seterr(1, "ETACALC: IETA must be 0, 2, 3, 4 or 5");
    tor_sound();
    const double c = (double)1.67f * pow(10.0, -8.0); /* 1.67*10.d0**(-8.d0), 1.67 a REAL */
    for (size_t n = 0; n < O.n3; ++n) {
        double e = O.CS[n] * O.p.mu_gas * c / O.RH[n];
        e = e < xnue ? e : xnue;
        O.ETA[n] = e * O.RH[n];
    }
    return 0;
}

/* VISC, ppvisc.f:2-373, IVISTYP=1 in spherical coordinates (NCOSYS=2, no
 * poles, alprat = 1): the viscous stress tensor from the velocities, its
 * divergence divtx/y/z on the faces and the dissipation function DISFN. */
static int visc(void)
{
    if (g_visc.ivistyp == 0) {
        memset(O.DIVTX, 0, O.n3 * sizeof(double));
        memset(O.DIVTY, 0, O.n3 * sizeof(double));
        memset(O.DIVTZ, 0, O.n3 * sizeof(double));
        return 0;
    }
    const double alprat = 1.0;
    double *DIVV = alloc3(), *TXX = alloc3(), *TYY = alloc3(), *TZZ = alloc3(), *TXY = alloc3(),
           *TXZ = alloc3(), *TYZ = alloc3(), *TXYAV = alloc3(), *TYZAV = alloc3(), *TXZAV = alloc3();
    double *V = O.V, *G = O.G, *H = O.H, *ETA = O.ETA;
    for (int k = 0; k <= NZ + 1; ++k) /* :30-42 */
        for (int j = 0; j <= NY + 1; ++j)
            for (int i = 0; i <= NX + 1; ++i) {
                A(TXX, i, j, k) = 2.0 * A(ETA, i, j, k) * (A(V, i + 1, j, k) - A(V, i, j, k)) / (XXA(i + 1) - XXA(i));
                A(DIVV, i, j, k) = 2.0 / 3.0 * A(ETA, i, j, k) *
                                   ((SURXA(i + 1) * A(V, i + 1, j, k) - SURXA(i) * A(V, i, j, k)) / VOLXB(i) +
                                    (SURYA(j + 1) * A(G, i, j + 1, k) - SURYA(j) * A(G, i, j, k)) / VOLYB(j) +
                                    (SURZA(k + 1) * A(H, i, j, k + 1) - SURZA(k) * A(H, i, j, k)) / VOLZB(k));
            }
    for (int k = 0; k <= NZ + 1; ++k) /* :75-88 */
        for (int j = 0; j <= NY + 1; ++j)
            for (int i = 0; i <= NX + 1; ++i) {
                A(TYY, i, j, k) = 2.0 * A(ETA, i, j, k) *
                                  ((A(G, i, j + 1, k) - A(G, i, j, k)) / (XYA(j + 1) - XYA(j)) +
                                   (A(V, i + 1, j, k) + A(V, i, j, k)) / 2.0 / XXB(i) +
                                   (A(H, i, j, k + 1) + A(H, i, j, k)) / 2.0 * (-tan(XZB(k))));
                A(TZZ, i, j, k) = 2.0 * A(ETA, i, j, k) *
                                  ((A(H, i, j, k + 1) - A(H, i, j, k)) / (XZA(k + 1) - XZA(k)) +
                                   (A(V, i + 1, j, k) + A(V, i, j, k)) / 2.0 / XXB(i));
            }
    for (int k = 0; k <= NZ + 2; ++k) /* :95-118 */
        for (int j = 0; j <= NY + 2; ++j)
            for (int i = 0; i <= NX + 2; ++i) {
                A(TXY, i, j, k) = 0.5 * ((A(V, i, j, k) - A(V, i, j - 1, k)) / (XYB(j) - XYB(j - 1)) / GEOXGA(i) / GEOZGA(k) +
                                         (A(G, i, j, k) - A(G, i - 1, j, k)) / (XXB(i) - XXB(i - 1)) * GEOXGA(i) *
                                             GEOZGA(k) / alprat);
                A(TXZ, i, j, k) = 0.5 * ((A(V, i, j, k) - A(V, i, j, k - 1)) / (XZB(k) - XZB(k - 1)) / GEOXHA(i) +
                                         (A(H, i, j, k) - A(H, i - 1, j, k)) / (XXB(i) - XXB(i - 1)) * GEOXHA(i));
                if (GEOXHA(i) < (double)1.e-10f) A(TXZ, i, j, k) = 0.0;
                A(TYZ, i, j, k) =
                    0.5 * ((A(H, i, j, k) - A(H, i, j - 1, k)) / (XYB(j) - XYB(j - 1)) / GEOXG(i) / GEOZGA(k) * GEOXH(i) +
                           (A(G, i, j, k) - A(G, i, j, k - 1)) / (XZB(k) - XZB(k - 1)) * GEOXG(i) * GEOZGA(k) / GEOXH(i) /
                               alprat);
            }
    for (int k = 0; k <= NZ + 1; ++k) /* :127-139 */
        for (int j = 0; j <= NY + 1; ++j)
            for (int i = 0; i <= NX + 1; ++i) {
                A(TXYAV, i, j, k) = A(ETA, i, j, k) *
                                    (A(TXY, i, j, k) + A(TXY, i + 1, j, k) + A(TXY, i, j + 1, k) + A(TXY, i + 1, j + 1, k)) / 2.0;
                A(TYZAV, i, j, k) = A(ETA, i, j, k) *
                                    (A(TYZ, i, j, k) + A(TYZ, i, j + 1, k) + A(TYZ, i, j, k + 1) + A(TYZ, i, j + 1, k + 1)) / 2.0;
                A(TXZAV, i, j, k) = A(ETA, i, j, k) *
                                    (A(TXZ, i, j, k) + A(TXZ, i + 1, j, k) + A(TXZ, i, j, k + 1) + A(TXZ, i + 1, j, k + 1)) / 2.0;
            }
    for (int k = 0; k <= NZ + 2; ++k) /* :141-152 */
        for (int j = 0; j <= NY + 2; ++j)
            for (int i = 0; i <= NX + 2; ++i) {
                A(TXY, i, j, k) = A(TXY, i, j, k) / 2.0 *
                                  (A(ETA, i, j, k) + A(ETA, i - 1, j, k) + A(ETA, i, j - 1, k) + A(ETA, i - 1, j - 1, k));
                A(TXZ, i, j, k) = A(TXZ, i, j, k) / 2.0 *
                                  (A(ETA, i, j, k) + A(ETA, i - 1, j, k) + A(ETA, i, j, k - 1) + A(ETA, i - 1, j, k - 1));
                A(TYZ, i, j, k) = A(TYZ, i, j, k) / 2.0 *
                                  (A(ETA, i, j, k) + A(ETA, i, j, k - 1) + A(ETA, i, j - 1, k) + A(ETA, i, j - 1, k - 1));
            }
    if (g_visc.idiss != 0) /* :155-167 */
        for (int k = 0; k <= NZ + 1; ++k)
            for (int j = 0; j <= NY + 1; ++j)
                for (int i = 0; i <= NX + 1; ++i)
                    A(O.DISFN, i, j, k) =
                        (3.0 * A(DIVV, i, j, k) * A(DIVV, i, j, k) + A(TXX, i, j, k) * A(TXX, i, j, k) +
                         2.0 * A(TXYAV, i, j, k) * A(TXYAV, i, j, k) * alprat + A(TYY, i, j, k) * A(TYY, i, j, k) +
                         2.0 * A(TYZAV, i, j, k) * A(TYZAV, i, j, k) * alprat + A(TZZ, i, j, k) * A(TZZ, i, j, k) +
                         2.0 * A(TXZAV, i, j, k) * A(TXZAV, i, j, k)) /
                        2.0 / A(ETA, i, j, k);
    for (int k = 0; k <= NZ + 1; ++k) /* :272-286 */
        for (int j = 0; j <= NY + 1; ++j)
            for (int i = 1; i <= NX + 1; ++i)
                A(O.DIVTX, i, j, k) =
                    (SURXB(i - 1) * A(DIVV, i - 1, j, k) - SURXB(i) * A(DIVV, i, j, k)) / VOLXA(i) +
                    (SURXB(i) * A(TXX, i, j, k) - SURXB(i - 1) * A(TXX, i - 1, j, k)) / VOLXA(i) +
                    (SURYA(j + 1) * A(TXY, i, j + 1, k) - SURYA(j) * A(TXY, i, j, k)) / VOLYB(j) / GEOXGA(i) / GEOZG(k) +
                    (SURZA(k + 1) * A(TXZ, i, j, k + 1) - SURZA(k) * A(TXZ, i, j, k)) / VOLZB(k) / GEOXHA(i) -
                    (A(TYY, i, j, k) + A(TYY, i - 1, j, k) + A(TZZ, i, j, k) + A(TZZ, i - 1, j, k) -
                     (A(DIVV, i, j, k) + A(DIVV, i - 1, j, k)) * 2.0) /
                        2.0 / XXA(i);
    for (int k = 0; k <= NZ + 1; ++k) /* :288-303 */
        for (int j = 1; j <= NY + 1; ++j)
            for (int i = 0; i <= NX + 1; ++i)
                A(O.DIVTY, i, j, k) =
                    (SURYB(j - 1) * A(DIVV, i, j - 1, k) - SURYB(j) * A(DIVV, i, j, k)) / VOLYA(j) / GEOXG(i) / GEOZG(k) +
                    (SURXA(i + 1) * A(TXY, i + 1, j, k) - SURXA(i) * A(TXY, i, j, k)) / VOLXB(i) +
                    (SURYB(j) * A(TYY, i, j, k) - SURYB(j - 1) * A(TYY, i, j - 1, k)) / VOLYA(j) / GEOXG(i) / GEOZG(k) +
                    (SURZA(k + 1) * A(TYZ, i, j, k + 1) - SURZA(k) * A(TYZ, i, j, k)) / VOLZB(k) / GEOXH(i) +
                    ((A(TYZAV, i, j, k) + A(TYZAV, i, j - 1, k)) * (-tan(XZB(k))) +
                     (A(TXYAV, i, j, k) + A(TXYAV, i, j - 1, k))) /
                        2.0 / XXB(i);
    for (int k = 1; k <= NZ + 1; ++k) /* :305-321, no poles */
        for (int j = 0; j <= NY + 1; ++j)
            for (int i = 0; i <= NX + 1; ++i)
                A(O.DIVTZ, i, j, k) =
                    (SURZB(k - 1) * A(DIVV, i, j, k - 1) - SURZB(k) * A(DIVV, i, j, k)) / VOLZA(k) / GEOXH(i) +
                    (SURXA(i + 1) * A(TXZ, i + 1, j, k) - SURXA(i) * A(TXZ, i, j, k)) / VOLXB(i) +
                    (SURYA(j + 1) * A(TYZ, i, j + 1, k) - SURYA(j) * A(TYZ, i, j, k)) / VOLYB(j) / GEOXG(i) / GEOZGA(k) +
                    (SURZB(k) * A(TZZ, i, j, k) - SURZB(k - 1) * A(TZZ, i, j, k - 1)) / VOLZA(k) / GEOXH(i) +
                    ((A(TYY, i, j, k) + A(TYY, i, j, k - 1) - (A(DIVV, i, j, k) + A(DIVV, i, j, k - 1))) * tan(XZA(k)) +
                     A(TXZAV, i, j, k) + A(TXZAV, i, j, k - 1)) /
                        2.0 / XXB(i);
    free(DIVV); free(TXX); free(TYY); free(TZZ); free(TXY); free(TXZ); free(TYZ);
    free(TXYAV); free(TYZAV); free(TXZAV);
    return 0;
}

/* NR_ARTIFICIAL_VISC, ppfor.f:1032-1230, one rank without poles: the
 * reference's three sweeps in its loop order, each cell's velocity and T
 * updated in place as the sweep passes it */
int tor_sphericalvellimit(int ivel, double velmax, double hydrotime);

static void nr_artificial_visc(double delt)
{
    const double artvisc_coeff = 3.0, tiny = (double)1.0e-30f; /* ppmod.f:392,1092 */
    if (!g_qvx) {
        g_qvx = (double*)calloc(O.n3, sizeof(double));
        g_qvy = (double*)calloc(O.n3, sizeof(double));
        g_qvz = (double*)calloc(O.n3, sizeof(double));
    }
    double dt = 0.0, d;
#define DMIN(x) do { double x_ = (x); if (x_ < dt) dt = x_; } while (0)
    for (int k = 1; k <= NZ; ++k) /* X, :1066-1099 */
        for (int j = 1; j <= NY; ++j) {
            d = A(O.V, 1, j, k) - A(O.V, 0, j, k);
            if (d > 0.0) d = 0.0;
            A(g_qvx, 0, j, k) = artvisc_coeff * delt * A(O.RH, 0, j, k) * (d * d);
            DMIN(d / DXB(0));
            const int ub = UBND(j, k);
            for (int i = 1; i <= ub - 2; ++i) {
                d = A(O.V, i + 1, j, k) - A(O.V, i, j, k);
                if (d > 0.0) d = 0.0;
                A(g_qvx, i, j, k) = artvisc_coeff * delt * A(O.RH, i, j, k) * (d * d);
                A(O.V, i, j, k) = A(O.V, i, j, k) - (A(g_qvx, i, j, k) - A(g_qvx, i - 1, j, k)) / (DXB(i) * A(O.RHQX, i, j, k));
                A(O.T, i, j, k) = A(O.T, i, j, k) - A(g_qvx, i, j, k) * d / (DXA(i) * A(O.RH, i, j, k) * A(O.CV, i, j, k));
                DMIN(d / DXB(i));
            }
        }
    for (int k = 1; k <= NZ; ++k) /* Y, :1100-1142 */
        for (int j = 1; j <= NY; ++j) {
            const int ub = UBND(j, k);
            for (int i = 1; i <= ub - 2; ++i) {
                if (j == 1) {
                    d = A(O.G, i, 1, k) - A(O.G, i, 0, k);
                    if (d > 0.0) d = 0.0;
                    A(g_qvy, i, 0, k) = artvisc_coeff * delt * A(O.RH, i, 0, k) * (d * d);
                    DMIN(d / DYB(0));
                }
                d = A(O.G, i, j + 1, k) - A(O.G, i, j, k);
                if (d > 0.0) d = 0.0;
                A(g_qvy, i, j, k) = artvisc_coeff * delt * A(O.RH, i, j, k) * (d * d);
                A(O.G, i, j, k) = A(O.G, i, j, k) - (A(g_qvy, i, j, k) - A(g_qvy, i, j - 1, k)) / (DYB(j) * A(O.RHQY, i, j, k));
                A(O.T, i, j, k) = A(O.T, i, j, k) - A(g_qvy, i, j, k) * d / (DYA(j) * A(O.RH, i, j, k) * A(O.CV, i, j, k));
                DMIN(d / DYB(j));
            }
        }
    for (int k = 1; k <= NZ; ++k) /* Z, :1154-1197 */
        for (int j = 1; j <= NY; ++j) {
            const int ub = UBND(j, k);
            for (int i = 1; i <= ub - 2; ++i) {
                if (k == 1) {
                    d = A(O.H, i, j, 1) - A(O.H, i, j, 0);
                    if (d > 0.0) d = 0.0;
                    A(g_qvz, i, j, 0) = artvisc_coeff * delt * A(O.RH, i, j, 0) * (d * d);
                    DMIN(d / DZB(0));
                }
                d = A(O.H, i, j, k + 1) - A(O.H, i, j, k);
                if (d > 0.0) d = 0.0;
                A(g_qvz, i, j, k) = artvisc_coeff * delt * A(O.RH, i, j, k) * (d * d);
                A(O.H, i, j, k) = A(O.H, i, j, k) - (A(g_qvz, i, j, k) - A(g_qvz, i, j, k - 1)) / (DZB(k) * A(O.RHQZ, i, j, k));
                A(O.T, i, j, k) = A(O.T, i, j, k) - A(g_qvz, i, j, k) * d / (DZA(k) * A(O.RH, i, j, k) * A(O.CV, i, j, k));
                DMIN(d / DZB(k));
            }
        }
#undef DMIN
    dt = 1.0 / fabs(8.0 * artvisc_coeff * (dt - tiny)); /* :1211-1212 */
    g_dt_artvis = dt < O.p.dtmax ? dt : O.p.dtmax;
}

int tor_forgas3d(double delt) /* FORGAS3D, ppfor.f:142-397 (IGRAV=1) */
{
    int rc = tor_bounds();
    if (rc) return rc;
    tor_pressg();
    const int vis = g_visc.ivis > 0;
    if (vis) { /* ppfor.f:170-174 */
        rc = etacalc();
        if (rc) return rc;
        rc = visc();
        if (rc) return rc;
    }
    const int diss = vis && g_visc.idiss != 0;
    centrifcoriolis();
    double* TOLD = alloc3();
    double* VH = alloc3(); double* GH = alloc3(); double* HH = alloc3();
    memcpy(TOLD, O.T, O.n3 * sizeof(double));
    if (O.p.isoth == 0) {
#pragma omp parallel for
        for (int k = 0; k <= NZ + 1; ++k)
            for (int j = 0; j <= NY + 1; ++j)
                for (int i = 0; i <= NX + 1; ++i)
                    A(O.T, i, j, k) =
                        A(O.T, i, j, k) - delt / 2.0 * A(O.PG, i, j, k) / A(O.RH, i, j, k) / A(O.CV, i, j, k) *
                                              ((SURXA(i + 1) * A(O.V, i + 1, j, k) - SURXA(i) * A(O.V, i, j, k)) / VOLXB(i) +
                                               (SURYA(j + 1) * A(O.G, i, j + 1, k) - SURYA(j) * A(O.G, i, j, k)) / VOLYB(j) +
                                               (SURZA(k + 1) * A(O.H, i, j, k + 1) - SURZA(k) * A(O.H, i, j, k)) / VOLZB(k));
        if (diss) /* ppfor.f:202-210 */
            for (int k = 1; k <= NZ; ++k)
                for (int j = 1; j <= NY; ++j)
                    for (int i = 1; i <= NX; ++i)
                        A(O.T, i, j, k) = A(O.T, i, j, k) + delt / 2.0 * A(O.DISFN, i, j, k) / A(O.CV, i, j, k) / A(O.RH, i, j, k);
        rc = tor_bounds();
        if (rc) goto out;
        tor_pressg();
    }
    /* divtx/y/z: the viscous stress divergence (zero without viscosity) */
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                A(VH, i, j, k) = A(O.V, i, j, k) +
                                 delt * ((-(A(O.PG, i, j, k) - A(O.PG, i - 1, j, k)) / DXB(i) + (vis ? A(O.DIVTX, i, j, k) : 0.0)) / A(O.RHQX, i, j, k) +
                                         A(O.CCX, i, j, k) + A(O.GRAVX, i, j, k));
                A(GH, i, j, k) = A(O.G, i, j, k) +
                                 delt * (((-(A(O.PG, i, j, k) - A(O.PG, i, j - 1, k)) / DYB(j) / GEOXG(i) / GEOZG(k) + (vis ? A(O.DIVTY, i, j, k) : 0.0)) /
                                              A(O.RHQY, i, j, k) +
                                          A(O.CCY, i, j, k) + A(O.GRAVY, i, j, k)) /
                                         GEOXG(i) / GEOZG(k));
                A(HH, i, j, k) = A(O.H, i, j, k) +
                                 delt * (((-(A(O.PG, i, j, k) - A(O.PG, i, j, k - 1)) / DZB(k) / GEOXH(i) + (vis ? A(O.DIVTZ, i, j, k) : 0.0)) /
                                              A(O.RHQZ, i, j, k) +
                                          A(O.CCZ, i, j, k) + A(O.GRAVZ, i, j, k)) /
                                         GEOXH(i));
            }
    memcpy(O.T, TOLD, O.n3 * sizeof(double));
    if (O.p.isoth == 0) {
#pragma omp parallel for
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i) {
                    A(O.V, i, j, k) = (A(O.V, i, j, k) + A(VH, i, j, k)) / 2.0;
                    A(O.G, i, j, k) = (A(O.G, i, j, k) + A(GH, i, j, k)) / 2.0;
                    A(O.H, i, j, k) = (A(O.H, i, j, k) + A(HH, i, j, k)) / 2.0;
                }
        /* the velocity ceiling on the intermediate velocities, ppfor.f:303-315 */
        if (g_lim.ivel != 0 && (rc = tor_sphericalvellimit(g_lim.ivel, g_lim.velmax, g_lim.hydrotime))) goto out;
        rc = tor_bounds();
        if (rc) goto out;
#pragma omp parallel for
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i)
                    A(O.T, i, j, k) =
                        A(O.T, i, j, k) - delt * A(O.PG, i, j, k) / A(O.RH, i, j, k) / A(O.CV, i, j, k) *
                                              ((SURXA(i + 1) * A(O.V, i + 1, j, k) - SURXA(i) * A(O.V, i, j, k)) / VOLXB(i) +
                                               (SURYA(j + 1) * A(O.G, i, j + 1, k) - SURYA(j) * A(O.G, i, j, k)) / VOLYB(j) +
                                               (SURZA(k + 1) * A(O.H, i, j, k + 1) - SURZA(k) * A(O.H, i, j, k)) / VOLZB(k));
        if (diss) /* ppfor.f:335-343 */
            for (int k = 1; k <= NZ; ++k)
                for (int j = 1; j <= NY; ++j)
                    for (int i = 1; i <= NX; ++i)
                        A(O.T, i, j, k) = A(O.T, i, j, k) + delt * A(O.DISFN, i, j, k) / A(O.CV, i, j, k) / A(O.RH, i, j, k);
    }
#pragma omp parallel for
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                A(O.V, i, j, k) = A(VH, i, j, k);
                A(O.G, i, j, k) = A(GH, i, j, k);
                A(O.H, i, j, k) = A(HH, i, j, k);
            }
    /* the velocity ceiling, ppfor.f:361-371 (SPHERICALVELLIMIT for NCOSYS=2) */
    if (g_lim.ivel != 0 && (rc = tor_sphericalvellimit(g_lim.ivel, g_lim.velmax, g_lim.hydrotime))) goto out;
    if (g_visc.iartvis > 0) { /* ppfor.f:374-395 */
        rc = tor_bounds();
        if (rc) goto out;
        nr_artificial_visc(delt);
    }
    rc = tor_bounds();
out:
    free(TOLD); free(VH); free(GH); free(HH);
    return rc;
}

/* ---------------------------------------------------------------- ppsub.f */
int tor_schritt(double* delt) /* SCHRITT, ppsub.f:1-137 (IGASTYP=1) */
{
    double mrh = O.RH[0], mt = O.T[0];
    for (size_t n = 0; n < O.n3; ++n) {
        if (O.RH[n] < mrh) mrh = O.RH[n];
        if (O.T[n] < mt) mt = O.T[n];
    }
    if (mrh <= 0.0 || mt <= 0.0)
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i)
                    if (A(O.RH, i, j, k) <= 0.0 || A(O.T, i, j, k) <= 0.0 || A(O.PG, i, j, k) <= 0.0)
                        return seterr(2, "non-positive value of rh/pg/t");
    if (O.p.fdelt == 0.0) { *delt = O.p.dtmax; return 0; }
    tor_sound();
    double ldt = O.p.dtmax;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                double v1 = fabs(A(O.V, i + 1, j, k)), v2 = fabs(A(O.G, i, j + 1, k)), v3 = fabs(A(O.H, i, j, k + 1));
                double cs = A(O.CS, i, j, k);
                double z1 = DXA(i + 1) / (v1 + cs);
                double z2 = DYA(j + 1) / (v2 + cs / GEOXG(i) / GEOZG(k));
                double z3 = DZA(k + 1) / (v3 + cs / GEOXH(i));
                double m = z1;
                if (z2 < m) m = z2;
                if (z3 < m) m = z3;
                if (ldt < m) m = ldt;
                ldt = m;
                if (ldt <= 0.0) return seterr(2, "local dt negative");
            }
    if (g_visc.ivis > 0) /* the viscous condition, ppsub.f:97-119 */
        for (int k = 1; k <= NZ; ++k)
            for (int j = 1; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i) {
                    const double xnu = A(O.ETA, i, j, k) / A(O.RH, i, j, k);
                    const double zcf = (1.0 / 8.0) / xnu /
                                       (pow(DXA(i + 1), -2.0) + pow(DYA(j + 1) * GEOXG(i) * GEOZG(k), -2.0) +
                                        pow(DZA(k + 1) * GEOXH(i), -2.0));
                    if (zcf < ldt) ldt = zcf;
                }
    if (g_visc.iartvis > 0 && g_dt_artvis < ldt) ldt = g_dt_artvis; /* ppsub.f:127-128 */
    ldt = (ldt < O.p.dtmax ? ldt : O.p.dtmax) * O.p.fdelt;
    *delt = ldt;
    return 0;
}

/* ---------------------------------------------------------------- ppdif.f */
int tor_diffusion(double delt, double xnue, double pr_num, int* iters) /* DIFFUSION, ppdif.f:1-309, on T */
{
    const double c = xnue / pr_num;
    double* field = O.T;
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) A(O.RHS, i, j, k) = A(field, i, j, k);
    /* DIFFUSION_COEF (ppdif.f:226-275) into the DIFR arrays */
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX + 1; ++i) A(O.DIFRX, i, j, k) = c / DXB(i);
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY + 1; ++j)
            for (int i = 1; i <= NX; ++i) A(O.DIFRY, i, j, k) = c / (DYB(j) * GEOXG(i) * GEOZG(k));
    for (int k = 1; k <= NZ + 1; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) A(O.DIFRZ, i, j, k) = c / (DZB(k) * GEOXH(i));
    /* FILL_DVALUES (ppdif.f:278-308) and DIFFUSION_MAT_ELEM (:169-223) */
    for (int k = 1; k <= NZ; ++k)
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                size_t n = I3(i, j, k);
                double d1 = delt * SURXA(i + 1) * A(O.DIFRX, i + 1, j, k) / (SURXB(i) * DXA(i));
                double d2 = delt * SURXA(i) * A(O.DIFRX, i, j, k) / (SURXB(i) * DXA(i));
                double d3 = delt * A(O.DIFRY, i, j + 1, k) / (GEOXG(i) * SURZB(k) * DYA(j));
                double d4 = delt * A(O.DIFRY, i, j, k) / (GEOXG(i) * SURZB(k) * DYA(j));
                double d5 = delt * A(O.DIFRZ, i, j, k + 1) * SURZA(k + 1) / (GEOXG(i) * SURZB(k) * DZA(k));
                double d6 = delt * A(O.DIFRZ, i, j, k) * SURZA(k) / (GEOXG(i) * SURZB(k) * DZA(k));
                double d16 = d1 + d2 + d3 + d4 + d5 + d6;
                O.SA[n] = 1.0 + d16;
                O.SB[n] = -d2; O.SC[n] = -d1; O.SD[n] = -d4; O.SE[n] = -d3; O.SF[n] = -d6; O.SG[n] = -d5;
                if (i == NX) {
                    O.SA[n] = 1.0 + d2;
                    O.SB[n] = -d2;
                    O.SC[n] = 0.0; O.SD[n] = 0.0; O.SE[n] = 0.0; O.SF[n] = 0.0; O.SG[n] = 0.0;
                }
            }
    double rhsnorm = l1_interior(O.RHS);
    int it = 0, rc = 0;
    for (;;) {
        ++it;
        double rnorm = sor_iteration(field, O.RHS, O.p.sorparam, 10);
        if (it == 500) { rc = seterr(-1, "Max iteration reached: DIF SOR matrix solver"); break; }
        if (!(rnorm / rhsnorm > O.p.epsmat || it < 5)) break;
    }
    if (iters) *iters = it;
    return rc;
}

/* One step of the pp.f:483-821 loop for the configs' switches. */
/* VARLIMIT, ppsub.f:140-154: feld = min(feld, varmax); feld = max(feld,
 * varmin) over every element of the actual array (its declared bounds). */
int tor_varlimit(int f, double vmin, double vmax)
{
    double* a;
    switch (f) {
    case TT_T: a = O.T; break;
    case TT_RH: a = O.RH; break;
    case TT_V: a = O.V; break;
    case TT_G: a = O.G; break;
    case TT_H: a = O.H; break;
    default: return seterr(1, "VARLIMIT: field must be T, RH, V, G or H");
    }
    int32_t lo[3], hi[3], rank, isint;
    tt_field_bounds(f, NX, NY, NZ, lo, hi, &rank, &isint);
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                double x = A(a, i, j, k);
                if (!(x < vmax)) x = vmax;
                if (!(x > vmin)) x = vmin;
                A(a, i, j, k) = x;
            }
    return 0;
}

/* SPHERICALVELLIMIT, ppsub.f:316-391 (NCOSYS=2, no poles). */
int tor_sphericalvellimit(int ivel, double velmax, double hydrotime)
{
    const double day = (double)8.64e4f; /* DAY = 8.64 E 4, a REAL constant (ppmod.f:1083) */
    const double pi = 3.141592653589793;
    double varmin, varmax;
    if (ivel == 1 || ivel == 3) {
        varmin = -velmax;
        varmax = velmax;
    } else if (ivel == 2 || ivel == 4) {
        double time_start = 0.0 * day, damp_time = 25.0 * day, veldamp_start = 0.0;
        if (hydrotime >= time_start && hydrotime <= time_start + damp_time) {
            varmax = (velmax - veldamp_start) *
                         pow(sin((pi / 2.0) * (hydrotime - time_start) / damp_time), (double)4.0f) +
                     veldamp_start;
            varmin = -varmax;
        } else if (hydrotime > time_start + damp_time) {
            varmax = velmax;
            varmin = -velmax;
        } else {
            varmax = veldamp_start;
            varmin = -veldamp_start;
        }
    } else {
        return seterr(1, "SPHERICALVELLIMIT: ivel must be 1, 2, 3 or 4");
    }
    if (ivel == 1 || ivel == 2) {
        for (int k = -1; k <= NZ + 2; ++k)
            for (int j = -1; j <= NY + 2; ++j)
                for (int i = -1; i <= NX + 2; ++i) {
                    if (j != -1) {
                        double g = A(O.G, i, j, k);
                        double c = varmax / (XXB(i) * GEOZG(k));
                        if (!(g < c)) g = c;
                        c = varmin / (XXB(i) * GEOZG(k));
                        if (!(g > c)) g = c;
                        A(O.G, i, j, k) = g;
                    }
                    if (k != -1) {
                        double h = A(O.H, i, j, k);
                        double c = varmax / XXB(i);
                        if (!(h < c)) h = c;
                        c = varmin / XXB(i);
                        if (!(h > c)) h = c;
                        A(O.H, i, j, k) = h;
                    }
                }
    }
    return tor_varlimit(TT_V, varmin, varmax);
}

void tor_set_limits(const tt_limits* l) { g_lim = *l; }
void tor_get_limits(tt_limits* l) { *l = g_lim; }

int tor_step(double delt, int iii, tt_step_stats* st)
{
    int rc;
    if (st) memset(st, 0, sizeof *st);
    g_iii = iii;
    /* pp.f:543-562: velocity and temperature limits, then the time advance
     * (:637-639) */
    if (g_lim.ivel != 0) {
        if ((rc = tor_varlimit(TT_V, -g_lim.velmax, g_lim.velmax))) return rc;
        if (g_lim.ivel < 3 && (rc = tor_sphericalvellimit(g_lim.ivel, g_lim.velmax, g_lim.hydrotime))) return rc;
    }
    if (g_lim.itemp != 0 && (rc = tor_varlimit(TT_T, g_lim.tempmin, g_lim.tempmax))) return rc;
    g_lim.hydrotime = g_lim.hydrotime + delt;
    if ((rc = tor_advect3d(delt))) return rc;
    /* pp.f:673-678 */
    if (g_lim.irho != 0 && (rc = tor_varlimit(TT_RH, g_lim.rhmin, 1.0e10))) return rc;
    if (g_lim.itemp != 0 && (rc = tor_varlimit(TT_T, g_lim.tempmin, g_lim.tempmax))) return rc;
    if ((rc = tor_forgas3d(delt))) return rc;
    /* pp.f:708-710 */
    if (g_lim.itemp != 0 && (rc = tor_varlimit(TT_T, g_lim.tempmin, g_lim.tempmax))) return rc;
    rc = tor_coupled_energy(delt, iii, st ? &st->energy : NULL);
    if (rc > 0) return rc;
    int rc2 = tor_bounds();
    if (rc2) return rc2;
    rc2 = tor_bounds();
    if (rc2) return rc2;
    tor_pressg();
    return rc;
}

void tor_set_iii(int iii) { g_iii = iii; }
int tor_energy_iter(void) { return O.energy_iter; }
