/*
 * ttgpu.h -- C ABI of the MI355X (gfx950) implementation of ttcode's per-step
 * hot path: ADVECT3D, FORGAS3D, COUPLED_ENERGY_EQS (FLD build + red-black SOR),
 * BOUNDS, SCHRITT and DIFFUSION, on one rank's whole subdomain held in HBM.
 *
 * The reference exposes parameterless Fortran subroutines that read and write
 * module globals (ppmod.f).  This library owns a device-resident copy of those
 * globals; the host moves arrays across with tt_upload/tt_download in the
 * reference's own wire format (Fortran column-major, r fastest, with each
 * array's declared lower/upper bounds, ghosts included), and calls the entry
 * points below in place of the reference subroutines.
 *
 * Conventions
 *  - Every entry point returns 0 on success.  A positive return is fatal (the
 *    reference would have called CLEAN_STOP or STOP); a negative return is
 *    informational (an iteration cap was reached, which the reference treats
 *    as non-fatal, ppsor.f:161-174, ppenergy.f:750-775).  tt_last_error()
 *    describes the last non-zero return.
 *  - One host thread; all device work is ordered on one HIP stream owned by
 *    the library.  Host-side module arrays are stale between tt_download
 *    calls.
 *  - Everything here is plain C: pointers, sizes, ints and doubles.
 */
#ifndef TTGPU_H
#define TTGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Run parameters.  Replaces MODULE input_init (ppmod.f:2-33) as read from the
 * positional control file ipp by readinitfile (ppmod.f:38-78), restricted to
 * the switches the hot path reads, plus the power-law opacity that ttcode's
 * KAPPA dispatcher (pprphy.f:1-360) does not have.
 * Only NCOSYS=2 (spherical polar), MODTYP=2 (planet), IGASTYP=1 (ideal gas),
 * ICSTYP=0, IRAD=3 (SOR) and inc_poles=0 are supported; viscosity (IVIS,
 * IARTVIS) is switched on by tt_set_viscosity.
 * ------------------------------------------------------------------------ */
typedef struct tt_params {
    /* integers (keep the order: a Fortran bind(C) type mirrors it) */
    int32_t nx, ny, nz;      /* locNX, locNY, locNZ  (r, phi, theta)          */
    int32_t bcxmin;          /* BOUNDXMIN type at r_min: 12 (fixed dT/dr =
                                D_T) or 9 (the flux FBottom carried by
                                radiative diffusion; ppbou.f:1029-1066,
                                ppenergy.f:150-167)                           */
    int32_t ibdrad;          /* BOUNDXMAX(2) top-T version: 3, 4, 7 or 9      */
    int32_t iflimtyp;        /* flux limiter: 1 = 1/3, 2 = Levermore-Pomraning */
    int32_t f_incident;      /* 0 = no irradiation, 2 = implicit cos heating  */
    int32_t ikaptyp;         /* 1 = constant ZKAPMULT, 4/5 and 9-12/14 tables,
                                6/7 column-depth tables, 8 = H- bound-free,
                                13 = constant with gamma_opc, 20/21/23-25
                                wavelength-binned tables, 26 = power law      */
    int32_t isigtyp;         /* 0 = no scattering, 1 = 0.4 cm^2/g             */
    int32_t isoth;           /* 0 = energy equation on                        */
    int32_t itmax;           /* ITMAX, shared by SOR and Newton (ipp:13)      */
    int32_t igrav;           /* FGRAV as pp.f:461 calls it once at set-up
                                (ppfor.f:646-867): 0 and 1 leave gravx/y/z
                                zero for the host to upload (IGRAV 1 is the
                                planet's constant SETPLANETXGRAV, ppplanets.f:
                                3-33); 4 sets the point mass's radial gravity
                                gravx(i) = -GRAV MSOL MSTAR2 / xxa(i)^2
                                (ppfor.f:753-766)                             */
    /* doubles */
    double xmin, xmax;       /* r range, cm                                   */
    double ymin, ymax;       /* phi range in units of 2 pi (ipp YMIN,YMAX)    */
    double zmin, zmax;       /* latitude range in units of pi/2 (ipp ZMIN..)  */
    double gamma, mu_gas;    /* GAMMA, mu_gas                                 */
    double zkapmult;         /* IKAPTYP=1 opacity                             */
    double kap0, dkap0;      /* IKAPTYP=26: kappa = kap0 (rho/kap_rho0)^a (T/kap_t0)^b,
                                dkappa/dT = dkap0 (rho/kap_rho0)^a (T/kap_t0)^(b-1) */
    double kap_rho0, kap_t0, kap_a, kap_b;
    double tirr;             /* TIRR, K                                       */
    double fbottom;          /* FBottom, bottom radiative flux (FBOT, ipp:12) */
    double d_t;              /* D_T, fixed bottom dT/dr for BCXMIN=12         */
    double omrot;            /* OMROT, frame rotation rate (0 if IROTATE=0)   */
    double sorparam;         /* SORPARAM (omega)                              */
    double epsmat;           /* EPSMAT, L1 during-sweep tolerance             */
    double fdelt, dtmax;     /* FDELT (CFL factor), DTMAX                     */
    double gamma_opc;        /* gamma_OPC = (kappa_A/kappa_P)^(1/4) of IKAPTYP 13
                                (and of the 12/14 tables' rewrite, ipp:11)    */
    double mstar2;           /* MSTAR2, the planet's mass in solar masses
                                (IGRAV 4; ipp:23)                             */
} tt_params;

/* ------------------------------------------------------------------------
 * Field ids for tt_upload / tt_download.  Ids 1..11 follow PLANETSHIFT's
 * numbering (ppshift.f:1013-1022).  Bounds are the reference's declared
 * bounds (ppmod.f allocate*var); n = (nx,ny,nz).
 * ------------------------------------------------------------------------ */
enum {
    TT_T = 1,          /* T   (-1:nx+2,-1:ny+2,-1:nz+2)  ppmod.f:322 */
    TT_RH = 2,         /* RH  (-1:nx+2,-1:ny+2,-1:nz+2)  ppmod.f:325 */
    TT_V = 3,          /* V   ( 0:nx+2,-1:ny+2,-1:nz+2)  ppmod.f:326 */
    TT_G = 4,          /* G   (-1:nx+2, 0:ny+2,-1:nz+2)  ppmod.f:327 */
    TT_H = 5,          /* H   (-1:nx+2,-1:ny+2, 0:nz+2)  ppmod.f:328 */
    TT_ER = 6,         /* ER  ghosted                    ppmod.f:323 */
    TT_DELTA_ER = 8,   /* delta_ER ghosted               ppmod.f:543 */
    TT_GD = 11,        /* GD ghosted                     ppmod.f:329 */
    TT_PG = 20,        /* PG ghosted */
    TT_CS = 21,        /* CS ghosted */
    TT_CV = 22,        /* CV ghosted (constant) */
    TT_RHQX = 23,      /* RHQX ( 0:nx+2,-1:ny+2,-1:nz+2) */
    TT_RHQY = 24,      /* RHQY (-1:nx+2, 0:ny+2,-1:nz+2) */
    TT_RHQZ = 25,      /* RHQZ (-1:nx+2,-1:ny+2, 0:nz+2) */
    TT_XKAPR = 26, TT_XKAPP = 27, TT_XKAPA = 28,   /* ghosted */
    TT_DXKAPP_DT = 29, TT_DXKAPA_DT = 30,          /* ghosted */
    TT_SIG = 31,       /* sig (0:nx+1,0:ny+1,0:nz+1) */
    TT_GRAVX = 32, TT_GRAVY = 33, TT_GRAVZ = 34,   /* ghosted */
    TT_STELLARINPUT = 35, /* (1:nx,-1:ny+2,-1:nz+2) */
    TT_STELLARTAU = 36,   /* (1:nx+1,-1:ny+2,-1:nz+2) */
    TT_DTAUS = 37,        /* (1:nx+1,-1:ny+2,-1:nz+2) */
    TT_ER_OLD = 38,       /* ghosted */
    TT_RHS_ER = 40,       /* (1:nx,1:ny,1:nz) */
    TT_SOR_A = 41, TT_SOR_B, TT_SOR_C, TT_SOR_D, TT_SOR_E, TT_SOR_F, TT_SOR_G, /* 41..47 (1:n) */
    TT_EVAL1 = 50, TT_EVAL2, TT_EVAL3, TT_EVAL4, TT_EVAL5, TT_EVAL6,          /* 50..55 (1:n) */
    TT_EVAL7 = 56, TT_EVAL8 = 57, TT_EVAL_1_6 = 58, TT_JACOB_ETA = 59,
    TT_FQ = 60, TT_FR = 61,                         /* (1:n) */
    TT_FLIMX = 62, TT_FLIMY = 63, TT_FLIMZ = 64,    /* (1:nx+1,1:ny,1:nz) etc. */
    TT_DIFRX = 65, TT_DIFRY = 66, TT_DIFRZ = 67,
    TT_FLUX_X = 68, TT_FLUX_Y = 69, TT_FLUX_Z = 70,
    TT_CENTCORX = 71, TT_CENTCORY = 72, TT_CENTCORZ = 73, /* (1:nx+1,1:ny+1,1:nz+1) */
    TT_UPPERBND = 80,     /* int32 (-1:ny+2,-1:nz+2)  ppmod.f:321 */
    TT_MU_STAR = 81,      /* (-1:ny+2,-1:nz+2)        ppmod.f:549 */
    TT_Q = 82,            /* Q of COUPLED_ENERGY_EQS, ghosted (local there) */
    TT_ETA = 83,          /* eta, dynamic viscosity, ghosted  ppmod.f:443 */
    TT_DIVTX = 84,        /* divtx (1:nx+1,0:ny+1,0:nz+1)     ppmod.f:439 */
    TT_DIVTY = 85,        /* divty (0:nx+1,1:ny+1,0:nz+1)     ppmod.f:440 */
    TT_DIVTZ = 86,        /* divtz (0:nx+1,0:ny+1,1:nz+1)     ppmod.f:441 */
    TT_DISFN = 87,        /* DISFN (0:nx+1,0:ny+1,0:nz+1)     ppmod.f:442 */
    /* conv_flux(c,:,:,:), c = 1..3: the thermal-energy advection fluxes
     * FLX/FLY/FLZ of the last ADVECT3D through each cell's low face
     * (ppadv.f:857-865; declared conv_flux(3,0:nx+1,0:ny+1,0:nz+1),
     * ppmod.f:346).  The library holds the cells ADVECT3D's loops define and
     * PRINTOUT writes (ppprt.f:51,78,96): (1:nx,1:ny,1:nz) per component. */
    TT_CONV_FLUX1 = 88, TT_CONV_FLUX2 = 89, TT_CONV_FLUX3 = 90,
    TT_FIELD_MAX = 92
};

/* Declared bounds of a field: lo[3], hi[3] (hi[2]=lo[2]=0 and rank 2 for
 * upperbnd/mu_star); *is_int = 1 for int32 fields.  Returns 0, or 1 for an
 * unknown id.  Needs tt_init (sizes come from it). */
int tt_field_info(int field, int32_t lo[3], int32_t hi[3], int32_t* rank, int32_t* is_int);

/* Device set-up.  Replaces the device side of allocate*var (ppmod.f:310-911)
 * plus GITTER (ppinit.f:1028-1406), GASINIT (ppinit.f:1408) and CALCMUSTAR
 * (ppsub.f:156): allocates every device array, computes the 1-D metric on the
 * host and uploads it, sets CV and mu_star.  Device 0 of the visible set. */
int tt_init(const tt_params* p);
/* Release all device memory and the stream. */
void tt_finalize(void);
const char* tt_last_error(void);

/* Host <-> device transfer of one module array in Fortran order. */
int tt_upload(int field, const void* host);
int tt_download(int field, void* host);
/* 1-D metric override: id per TT_M_* below, data in the declared bounds. */
int tt_upload_metric(int metric, const double* host);
int tt_download_metric(int metric, double* host);
enum {
    TT_M_XXA = 0, TT_M_XXB, TT_M_DXA, TT_M_DXB, TT_M_XYA, TT_M_XYB, TT_M_DYA, TT_M_DYB,
    TT_M_XZA, TT_M_XZB, TT_M_DZA, TT_M_DZB, TT_M_DDX0, TT_M_DDX1, TT_M_DDY0, TT_M_DDY1,
    TT_M_DDZ0, TT_M_DDZ1, TT_M_DCY0, TT_M_DCY1, TT_M_DDX0KEP, TT_M_DDX1KEP,
    TT_M_SURXA, TT_M_SURXB, TT_M_VOLXA, TT_M_VOLXB, TT_M_CCXA, TT_M_CCXB,
    TT_M_SURYA, TT_M_SURYB, TT_M_VOLYA, TT_M_VOLYB, TT_M_CCYA, TT_M_CCYB,
    TT_M_SURZA, TT_M_SURZB, TT_M_VOLZA, TT_M_VOLZB, TT_M_CCZA, TT_M_CCZB,
    TT_M_GEOXGA, TT_M_GEOXHA, TT_M_GEOXG, TT_M_GEOXH, TT_M_GEOZG, TT_M_GEOZGA,
    TT_M_COUNT
};
/* Declared bounds of a metric array (lo..hi). */
int tt_metric_info(int metric, int32_t* lo, int32_t* hi);

/* ------------------------------------------------------------------------
 * Solver options (new: ttcode has one fixed path).  precision 0 is the
 * reference algorithm: FP64 red-black SOR from delta_ER = 0 (ppsor.f:29) until
 * the L1 during-sweep ratio <= epsmat (ppsor.f:176).  precision 1 runs FP32
 * SOR sweeps inside an FP64 defect-correction loop whose FP64 L1 residual
 * ratio sum|RHS - A d| / sum|RHS| must reach epsmat.
 * guess 0 = cold (delta_ER = 0), 1 = warm (previous solve's delta_ER times
 * warm_scale), 2 = learned (CNN prediction on the first Newton iteration; after
 * 3 rejections in a row the network runs on every 16th first solve only, and
 * the others take the warm or the cold start, whichever has taken fewer SOR
 * iterations on recent first solves -- a moving average with weight 1/4 --
 * trying the other on every 16th); a warm or learned guess is kept only if its
 * FP64 residual does not exceed the cold start's.
 * inner (precision 1 only) picks the FP32 sweeps of the inner solves: 0 =
 * point red-black SOR with omega = SORPARAM, the reference's own sweep in
 * FP32; 1 = red-black radial-line SOR with omega = line_omega (0: chosen by
 * a search on the first line solve, see tt_get_line_omega): each radial
 * column (j,k) is solved exactly as a tridiagonal system in r, the columns
 * coloured by (j+k) parity, so the strong radial couplings of the FLD
 * operator are inside the line solves and the sweep count to the same FP64
 * tolerance falls by more than an order of magnitude.
 * ------------------------------------------------------------------------ */
typedef struct tt_solver_opts {
    int32_t precision;       /* 0 = FP64 SOR, 1 = FP32 SOR + FP64 correction */
    int32_t guess;           /* 0 cold, 1 warm, 2 learned                    */
    int32_t batch;           /* sweeps queued between host convergence checks */
    int32_t inner_max;       /* FP32 sweeps per correction (precision 1)     */
    int32_t adjust_omega;    /* 1: ADJUST_SOR_PARAM at the first Newton
                                iteration of steps III = 1 and III % 10000 = 0
                                (ppenergy.f:372-374); 0 (default): omega fixed */
    int32_t inner;           /* precision 1: 0 point sweeps, 1 radial lines  */
    double inner_tol;        /* FP32 inner L1 ratio target (precision 1)     */
    double warm_scale;       /* scale of the previous delta_ER for guess 1   */
    double line_omega;       /* line sweeps' over-relaxation in (0, 2); 0 =
                                chosen by a search (default)                 */
} tt_solver_opts;

typedef struct tt_sor_stats {
    int32_t sor_iter;        /* SOR_ITER: full red+black iterations          */
    int32_t outer_iter;      /* FP64 corrections (precision 1), else 0       */
    int32_t fp32_sweeps;     /* FP32 red+black iterations (precision 1)      */
    int32_t guess_used;      /* 0 cold, 1 warm, 2 learned actually used      */
    double rhsnorm;          /* global_RHSNORM                               */
    double rnorm;            /* global_RNORM of the last iteration           */
    double r0_guess, r0_cold;/* FP64 L1 residual of the guess and of zero    */
    double guess_scale;      /* residual-minimising factor applied to a guess */
} tt_sor_stats;

int tt_set_solver(const tt_solver_opts* o);
int tt_get_solver(tt_solver_opts* o);

/* SOLVE_WITH_SOR(ppsor.f:2) on the system already on the device:
 * fill_matrix != 0 runs FILL_SOR_MATRIX + B_VECTOR_BOUNDARIES from EVAL1..8,
 * EVAL_1_6, RHS_ER and upperbnd (ppsor.f:18-27) first; fill_matrix == 0 takes
 * SOR_A..G and RHS_ER as uploaded.  Result in delta_ER. */
int tt_solve_sor(int fill_matrix, tt_sor_stats* st);

/* ADJUST_SOR_PARAM (ppsor.f:332-354) on the system already on the device:
 * 50 cold-start solves with SORPARAM = 1.01, 1.03, ..., 1.99 (REAL literals:
 * the start is the float nearest 1.01, the step the float nearest 0.02,
 * added in double), then SORPARAM = the first omega with the fewest
 * iterations.  fill_matrix as for tt_solve_sor (the solves refill from the
 * same EVALs, so once is the same).  Each solve runs in the current solver
 * precision (FP64: SOR_ITER; precision 1: FP32 iterations).  omega and
 * iters[50] (may be null) receive the chosen omega and each solve's count;
 * delta_ER holds the last solve's result, as in the reference. */
int tt_adjust_sor_param(int fill_matrix, double* omega, int32_t* iters);

/* SORPARAM of the running library (set by tt_init from tt_params.sorparam,
 * changed by ADJUST_SOR_PARAM). */
int tt_set_sorparam(double omega);
double tt_get_sorparam(void);

/* The over-relaxation the radial-line sweeps use (tt_solver_opts.inner 1):
 * line_omega where it is set; where it is 0 (the default), the value a search
 * chose on the first line solve after tt_init or tt_set_solver -- cold-start
 * solves of that system to the FP64 target over omega in [1, 1.95], the
 * fewest line iterations winning, as ADJUST_SOR_PARAM (ppsor.f:332-354)
 * chooses SORPARAM -- or 0 before that solve.  *trials: the search's trial
 * solves (0 with line_omega set).  No counterpart in the reference: the line
 * sweeps are this library's own inner solver. */
int tt_get_line_omega(double* omega, int32_t* trials);

/* ------------------------------------------------------------------------
 * Physics entry points, in the step order of pp.f:483-821.
 * ------------------------------------------------------------------------ */
int tt_bounds(void);          /* BOUNDS      ppbou.f:1      */
int tt_pressg(void);          /* PRESSG      ppphy.f:1      */
int tt_qdens(void);           /* QDENS       ppphy.f:47     */
int tt_sound(void);           /* SOUND       ppphy.f:79     */
int tt_advect3d(double delt); /* ADVECT3D    ppadv.f:525    */
int tt_forgas3d(double delt); /* FORGAS3D    ppfor.f:142    */

typedef struct tt_energy_stats {
    int32_t energy_iter;     /* Newton iterations                             */
    int32_t sor_iter_total;  /* sum of SOR_ITER over the Newton iterations    */
    int32_t sor_iter_last;   /* SOR_ITER of the last solve                    */
    int32_t sor_iter_first;  /* SOR_ITER of the first solve                   */
    int32_t fp32_sweeps;     /* FP32 iterations (precision 1)                 */
    int32_t fallbacks;       /* guesses rejected for the cold start           */
    int32_t learned_used;    /* solves that used the learned guess            */
    int32_t learned_skipped; /* first solves that did not run the network:
                                after 3 rejections in a row it runs on every
                                16th first solve only, until it is kept again */
    double max_dq, max_de;   /* global_deltaQ, global_deltaER of last iter    */
    double guess_scale_first;/* least-residual factor of the first solve's guess */
    double guess_r0_first;   /* its L1 residual / the cold start's (0: no guess) */
    int32_t first_start;     /* start the first solve used: 0 cold, 1 warm,
                                2 learned                                     */
    int32_t first_choice;    /* start it was given (a rejected guess falls
                                back to cold): 0 cold, 1 warm, 2 learned      */
    /* device time of the call's phases, summed over the Newton iterations,
     * ms: the FLD build (FLUX_BC, CALCULATEFLUX with DIFF/FLIM, the Newton
     * terms, FILL_EVALUES, FILL_SOR_MATRIX), SOLVE_WITH_SOR (with the
     * initial guess and, when switched on, ADJUST_SOR_PARAM), and the
     * delta_Q / Q, E update with the Q/E boundaries */
    double ms_build, ms_sor, ms_update;
    double ms_sor_first;     /* SOLVE_WITH_SOR phase of the first Newton
                                iteration, ms (guess 2 picks warm or cold by
                                its moving average)                           */
} tt_energy_stats;

/* COUPLED_ENERGY_EQS (ppenergy.f:2-791).  iii is the step counter III (read
 * by IBDRAD=9, and by the ADJUST_SOR_PARAM rule of ppenergy.f:372-374 when
 * tt_solver_opts.adjust_omega is set; off by default). */
int tt_coupled_energy(double delt, int iii, tt_energy_stats* st);

/* SCHRITT (ppsub.f:1-137): CFL time step into *delt (the positivity check
 * returns 2 when rho, T or PG is non-positive). */
int tt_schritt(double* delt);

/* SCHRITT's Courant bound on caller-owned device arrays, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc) and hold their own device state; it needs no tt_init.  rh, t,
 * pg, v, g, h: device arrays over the ghosted box (-1:nx+2, -1:ny+2,
 * -1:nz+2), column-major (r fastest), V, G and H with their unused low ghost
 * layer present; dxa, dya, dza (declared 1:n+2), geoxg, geoxh (-1:nx+2),
 * geozg (-1:nz+2): device arrays of the GITTER metric (tt_download_metric
 * gives them).  Computes SOUND's CS = sqrt(gamma PG / RH) in place and
 * writes to device memory out[0] = FDELT min(DTMAX, min over the interior of
 * the three Courant terms) (ppsub.f:49-78,129-131; IVIS = 0, IARTVIS = 0) and
 * out[1] = 1 where the reference stops on a non-positive RH, T or PG
 * (ppsub.f:26-46), 2 where the bound is <= 0, else 0.  Work is queued on
 * `stream` (a hipStream_t; null: the null stream); the caller synchronises.
 * The reduction's block partials live in a fixed 4 x 1024-double array in
 * device memory (32 KB per GPU, held until the process ends); calls that
 * overlap on one GPU share it and race.  Returns 1 for nx, ny or nz < 1, a
 * box of 2^31 cells or more, or a null pointer; 0 when queued. */
int tt_schritt_noalloc(int nx, int ny, int nz, const double* rh, const double* t, const double* pg,
                       const double* v, const double* g, const double* h, const double* dxa, const double* dya,
                       const double* dza, const double* geoxg, const double* geoxh, const double* geozg,
                       double gamma, double fdelt, double dtmax, double* out, void* stream);

/* DIFFUSION (ppdif.f:1-167) applied to T with constant diffusivity
 * xnue/pr_num (the IDIF=2 call of pp.f:765-770); returns SOR iterations in
 * *iters.  ITMIN=5, SORITMAX=500 as in ppdif.f:14-15. */
int tt_diffusion(double delt, double xnue, double pr_num, int* iters);

/* ------------------------------------------------------------------------
 * Clamps of the step loop (pp.f:543-562, 673-678, 708-710): VARLIMIT
 * (ppsub.f:140-154) and, for NCOSYS=2, SPHERICALVELLIMIT (ppsub.f:316-391),
 * switched by IRHO, ITEMP and IVEL.  All off by default (tt_init).
 * ------------------------------------------------------------------------ */
typedef struct tt_limits {
    int32_t irho;            /* != 0: RH into [rhmin, 1e10] after ADVECT3D     */
    int32_t itemp;           /* != 0: T into [tempmin, tempmax] before
                                ADVECT3D, after ADVECT3D, after FORGAS3D       */
    int32_t ivel;            /* 0 off; 1-4: VARLIMIT of V to +-velmax before
                                ADVECT3D, then for 1 and 2 SPHERICALVELLIMIT
                                (2 ramps the ceiling over 25 days)             */
    int32_t pad0;
    double rhmin, tempmax, tempmin, velmax;
    double hydrotime;        /* model time seen by the next tt_step's clamps
                                (ivel 2); each tt_step then adds delt, as
                                pp.f:637-639 does for IADV > 0                 */
} tt_limits;

int tt_set_limits(const tt_limits* l);
int tt_get_limits(tt_limits* l);

/* VARLIMIT(ilimit, feld, varmin, varmax), ppsub.f:140-154: every element of
 * the field's declared array (TT_T, TT_RH, TT_V, TT_G or TT_H) set to
 * max(min(x, varmax), varmin). */
int tt_varlimit(int field, double varmin, double varmax);

/* SPHERICALVELLIMIT, ppsub.f:316-391 (NCOSYS=2, no poles) for ivel 1-4:
 * for 1 and 2, G into +-ceiling/(r cos(lat)) and H into +-ceiling/r; then V
 * into +-ceiling.  The ceiling is velmax, or for ivel 2 and 4 the ramp
 * velmax sin^4(pi/2 t/(25 days)) while 0 <= t <= 25 days (t = hydrotime). */
int tt_sphericalvellimit(int ivel, double velmax, double hydrotime);

/* ------------------------------------------------------------------------
 * Viscosity in FORGAS3D (ppfor.f:170-174): ETACALC (ppvisc.f:375-484) then
 * VISC (ppvisc.f:2-373, IVISTYP 1 in spherical coordinates), whose stress
 * divergence divtx/y/z enters the momentum updates (ppfor.f:222-262) and
 * whose dissipation DISFN, with IDISS, the two temperature updates
 * (ppfor.f:202-210,335-343).  Off by default (IVIS = 0), as in the configs.
 * With IVIS > 0 SCHRITT also limits the time step by the viscous condition
 * (ppsub.f:97-119).
 * IARTVIS = 1 adds the Neumann-Richtmyer artificial viscosity at the end of
 * FORGAS3D (ppfor.f:374-395, NR_ARTIFICIAL_VISC ppfor.f:1032-1230), whose
 * time-step bound dt_artvis SCHRITT takes (ppsub.f:127-128).  dt_artvis is
 * DTMAX until the first such FORGAS3D (pp.f:436-440).  The compression
 * fluxes QviscY/QviscZ stay on the device between calls, zero at first use:
 * where upperbnd varies between neighbouring columns the reference reads
 * entries the current call did not write, and so does the library.
 * ------------------------------------------------------------------------ */
typedef struct tt_viscosity {
    int32_t ivis;            /* > 0: viscosity on                             */
    int32_t ivistyp;         /* 0: divtx/y/z = 0; 1: the full stress tensor   */
    int32_t ieta;            /* ETACALC: 0 and 5 eta = RH XNUE; 2 XNUE ramped
                                from 1e9 over days 282.5-292.5 of hydrotime
                                (tt_set_limits); 3 raised at the bottom (nx >=
                                13); 4 molecular, eta = RH min(CS mu_gas
                                1.67e-8 / RH, XNUE)                           */
    int32_t idiss;           /* != 0: dissipative heating DISFN               */
    double xnue;             /* XNUE, kinematic viscosity, cm^2/s             */
    int32_t iartvis;         /* 0 off; 1 NR_ARTIFICIAL_VISC                   */
    int32_t pad0;
} tt_viscosity;

int tt_set_viscosity(const tt_viscosity* v);
int tt_get_viscosity(tt_viscosity* v);
/* dt_artvis (artvis_var_init, ppmod.f:388-395): min(1/|8 artvisc_coeff
 * (m - TINY)|, DTMAX) with m the most negative compression rate dv/dx of the
 * last NR_ARTIFICIAL_VISC (ppfor.f:1211-1212); DTMAX before the first. */
int tt_get_dt_artvis(double* dt);

typedef struct tt_step_stats {
    tt_energy_stats energy;
    double ms_advect, ms_force, ms_energy, ms_total;  /* device time, ms */
} tt_step_stats;

/* One step of the pp.f:483-821 loop for IADV=IFOR=1, IRAD=3: ADVECT3D,
 * FORGAS3D, COUPLED_ENERGY_EQS, BOUNDS, BOUNDS, PRESSG (time step held fixed),
 * with the clamps tt_set_limits switched on at their places in that loop. */
int tt_step(double delt, int iii, tt_step_stats* st);

/* III of the host's step loop as BOUNDS reads it (BOUNDXMAX(2) IBDRAD 9,
 * and BOUNDXMIN(9), which at III <= 0 -- the set-up calls of LOCALINIT and
 * pp.f:419 -- evaluates BurrowsOPC at cell (1,1,1) and needs
 * tt_set_burrows_table).  tt_init sets 0; tt_step and tt_coupled_energy set
 * it from their argument.  No counterpart function: the reference's III is a
 * module variable (ppmod.f, fluid_var_init). */
int tt_set_step(int iii);

/* Block until all queued device work is done. */
int tt_sync(void);

/* Device memory (bytes) the library holds for the current tt_init: fields,
 * temporaries and the SOR working set (the learned guess's model and
 * training set are not included). */
int64_t tt_device_bytes(void);

/* Diagnostics: volume integrals over the interior cells (dV = volxb volyb
 * volzb) of the thermal energy rho c_v T, the radiation energy E and the
 * irradiation heating rate STELLARINPUT, into out3[0..2]. */
int tt_energy_totals(double* out3);

/* Per-launch device timing of the SOR half-sweep kernels (HIP events on the
 * library's stream).  tt_sweep_timing(period) resets the counters and, for
 * period >= 1, brackets every period-th half-sweep launch with a pair of
 * events (0 disables; sampling keeps the events' own cost out of the step
 * time).  The getter returns the timed launches and their summed
 * milliseconds since the reset, split by FP64 and FP32 sweeps. */
int tt_sweep_timing(int period);
int tt_sweep_timing_get(int64_t* n64, double* ms64, int64_t* n32, double* ms32);

/* Speculative convergence tests of the FP64 SOR solve.  The test of an
 * iteration normally runs in a kernel of its own, so the solve stops exactly
 * where the reference's does.  Where the L1 ratios of the two latest batches,
 * extrapolated margin times faster than observed, cannot reach EPSMAT
 * before the end of the batch being queued, that batch's tests run inside
 * the next red sweep instead; should one of those tests end the solve after
 * all, the solve is rerun from its start values with tests of their own, so
 * results never depend on the speculation.  margin > 0 sets the factor
 * (default 2.5), 0 speculates on every batch but the first two (testing),
 * < 0 switches speculation off.  The getter returns the speculative
 * batches and the reruns since the last set call. */
int tt_sor_speculation(double margin);
int tt_sor_speculation_get(int64_t* spec_batches, int64_t* reruns);

/* ------------------------------------------------------------------------
 * Table opacities, IKAPTYP 4 and 5 (KAPPA, pprphy.f:84-120; FreedmanOPC,
 * pprphy.f:800-942).  The tables hold log10 of the Planck, Rosseland and
 * absorption mean opacities on the reference's fixed grid of 42 temperatures
 * (75 ... 4000 K) x 18 pressures (0.001 ... 300000 mbar) (ppmod.f:690-699),
 * temperature fastest: kap[it + 42 ip], as the reference's
 * FreedmanXKAP_*(42,18) lie in memory.  Entries its table file leaves out
 * hold -7 (READFREEDMANOPCTABLE, pprphy.f:944-1002, whose text format
 * ttcode_amd/opacity.py reads).  IKAPTYP 4 and 5 differ only in the file.
 * ------------------------------------------------------------------------ */
#define TT_FREEDMAN_NT 42
#define TT_FREEDMAN_NP 18
/* Replaces READFREEDMANOPCTABLE's fill of FreedmanXKAP_plnk/_ross/_abs
 * (pprphy.f:944-1002).  The library keeps a copy, across tt_init and
 * tt_finalize, until the next call; it may be called before tt_init. */
int tt_set_opacity_table(const double* kap_planck, const double* kap_ross, const double* kap_abs);
/* IKAPTYP 6 and 7 (pprphy.f:123-175, columnOPC pprphy.f:1543-1745) take the
 * Planck and Rosseland means from the tables above and the absorption mean
 * from a column-depth table: log10 kappa_abs on the same (T, P) grid x 8
 * column levels log10 kappa_cut = -26 ... -19, kap[it + 42 (ip + 18 ic)], as
 * ColumnXKAP_abs(42,18,8) after READCOLUMNDEPTHOPCTABLE (pprphy.f:1747-1833;
 * ttcode_amd/opacity.py read_column_table reads the text format and caps
 * each level at level 8's value, as the reference does).  Replaces that
 * reader's fill; kept like the other tables. */
int tt_set_column_table(const double* kap_abs);

/* ------------------------------------------------------------------------
 * Table opacities, IKAPTYP 9, 10, 11, 12 and 14 (KAPPA, pprphy.f:199-234,
 * 242-274; BurrowsOPC, pprphy.f:1005-1225): log10 of the Planck, Rosseland
 * and absorption means on the reference's grid of 50 log10 T (1.69896 ...
 * 3.69889) x 50 log10 rho (-11.9996 ... -1.99993) (ppmod.f:724-745, REAL
 * values), temperature fastest: kap[it + 50 ir], as BurrowsXKAP_*(50,50).
 * The tables are the ones READBURROWSOPCTABLE (pprphy.f:1227-1520) leaves:
 * for IKAPTYP 12 and 14 the absorption table is log kP + 4 log10(gamma_OPC),
 * for 14 the Rosseland table is log kP, for 11 the absorption table carries
 * the planet's extra absorber below 0.03 bar (ttcode_amd/opacity.py
 * read_burrows_table applies these).  IKAPTYP 9 reads the table with TiO/VO,
 * the others the one without.  The wavelength-binned families 20-25 take
 * their means from these tables too (below).
 * ------------------------------------------------------------------------ */
#define TT_BURROWS_NT 50
#define TT_BURROWS_NR 50
/* Replaces READBURROWSOPCTABLE's fill of BurrowsXKAP_plnk/_ross/_abs; kept
 * like the IKAPTYP 4/5 tables, and may be called before tt_init. */
int tt_set_burrows_table(const double* kap_planck, const double* kap_ross, const double* kap_abs);

/* ------------------------------------------------------------------------
 * IKAPTYP 20, 21, 23, 24 and 25: the Burrows tables with 30 wavelength bins
 * (KAPPA, pprphy.f:321-349; BurrowsOPC, pprphy.f:1005-1225).  The Planck,
 * Rosseland and absorption means come from tt_set_burrows_table as for 9-14
 * (READBURROWSOPCTABLE's columns 3-5), and the cells from upperbnd up take
 * upperbnd-1's values.  For 20, 21, 23 and 24 STELLARINPUT_HEATING sums the
 * stellar flux over the bins (pprad.f:429-447): per bin the optical depth of
 * xkapW rho DR from upperbnd down (WAVELENGTH_CALCSTELLARTAU, pprad.f:
 * 519-547), its central difference (WAVELENGTH_STELLAR_DTAU, :549-569), and
 * STELLARINPUT = sum of BSTAR_BIN exp(-tau/mu*) dtau, times the dilution,
 * over dxb; 25 uses the grey irradiation.  xkapW, the bins' opacity, is
 * formed where that sum needs it and not kept.  The two-stream bins of
 * 22-25 (xkapLW, local_BB) feed only the wavelength radiative transfer of
 * IRAD 5-7 (ppwavelth.f) and are not formed.  IKAPTYP 22 is refused: the
 * reference's reader opens no file for it and stops (pprphy.f:1265-1275).
 * ------------------------------------------------------------------------ */
#define TT_NWAVE 30
/* Replaces READBURROWSOPCTABLE's fill of BurrowsXKAP_waveopac and BSTAR_BIN
 * (pprphy.f:1297-1360,1438-1469,1509-1538): kap_wave = log10 of the 30 bins'
 * irradiation opacity, kap[it + 50 (ir + 50 l)] as BurrowsXKAP_waveopac(50,
 * 50,30) (with the planet's extra absorber for 20, 24 and 25), bstar[30] =
 * BSTAR_BIN, dilution = (RSTAR RSOL / semi)^2 with semi from pp.f:204-206
 * (ttcode_amd/opacity.py reads both files and computes it).  Kept like the
 * other tables; may be called before tt_init. */
int tt_set_burrows_wave_table(const double* kap_wave, const double* bstar, double dilution);

/* KAPPA + SUBSIG (pprphy.f:1-360,567-583) on the device state, as FLIM calls
 * them (pprphy.f:362-380).  IKAPTYP 4 and 5 call PRESSG first and stop with
 * a positive code on T <= 0 (pprphy.f:825-828); IKAPTYP 9-14 stop likewise
 * on T <= 0 or a density outside the table's log range (pprphy.f:1040-1105). */
int tt_kappa(void);

/* ------------------------------------------------------------------------
 * Learned initial guess (new: no counterpart in the reference).  A 3-D CNN
 * maps five per-cell features -- log E, log T, log rho, log kappa_R,
 * asinh(STELLARINPUT), each normalised by a mean and a scale -- to the first
 * Newton correction of COUPLED_ENERGY_EQS as a rate, delta_ER / (E dt).  Layers
 * 5 -> 16 -> 16 -> 16 -> 16 -> 1, 3x3x3 convolutions (zero padding in r and
 * theta, periodic in phi), GELU (tanh form) after the hidden layers; bf16
 * weights and activations, fp32 accumulation on the matrix cores.
 * Parameters are fp32: weights W[tap][ci][co] per layer, tap = (dk+1)*9 +
 * (dj+1)*3 + (di+1) over (theta, phi, r) offsets, layers concatenated; biases
 * per layer.  n_r must be a multiple of 16.  The model and the training set
 * live across tt_init/tt_finalize until tt_cnn_release.
 * ------------------------------------------------------------------------ */
int64_t tt_cnn_num_weights(void);   /* 23,328 */
int64_t tt_cnn_num_biases(void);    /* 65 */
int tt_cnn_set_weights(const float* w, int64_t nw, const float* b, int64_t nb);
int tt_cnn_get_weights(float* w, int64_t nw, float* b, int64_t nb);
/* norm12 = mean[6], scale[6] of the five features and of the target */
int tt_cnn_set_norm(const float* norm12);
int tt_cnn_get_norm(float* norm12);
/* Forward (and, with target non-null, the mean squared error and its
 * gradients) on host data: x0[cells][5] normalised features, cells ordered
 * r fastest, then phi, then theta; y[cells]. */
int tt_cnn_eval_host(int nx, int ny, int nz, const float* x0, const float* target, float* y, float* loss,
                     float* gw, float* gb);
/* Training set on the device: reserve room for max_samples samples of one
 * grid; with collection on, the first Newton iteration of every
 * tt_coupled_energy / tt_step call appends its features and its
 * delta_ER / (E dt) (from the FP64 SOR solve).  Samples beyond the reserve are dropped. */
int tt_cnn_reserve(int max_samples, int nx, int ny, int nz);
int tt_cnn_collect(int enable);
int tt_cnn_dataset_size(void);
/* Normalisation from the training set's means and standard deviations. */
int tt_cnn_fit_norm(void);
/* Adam over the training set (minibatches of `batch` samples, shuffled with
 * seed); epoch_loss[epochs] receives the mean normalised loss per epoch. */
int tt_cnn_train(int epochs, int batch, float lr, uint64_t seed, float* epoch_loss);
/* Free the model, its workspace and the training set. */
int tt_cnn_release(void);

#ifdef __cplusplus
}
#endif
#endif /* TTGPU_H */
