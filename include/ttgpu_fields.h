/*
 * ttgpu_fields.h -- declared Fortran bounds of every array that crosses the
 * ttgpu C ABI (tt_upload / tt_download), following the ALLOCATE statements of
 * ppmod.f (allocatefluidvar :310-352, allocateforcevar :440-448,
 * allocateradvar :495-552, allocatedeltaERvar :860-879, allocatesorvar
 * :900-911, allocategridvar :605-642) and the automatic arrays of
 * COUPLED_ENERGY_EQS (ppenergy.f:19-20).
 *
 * Header-only so that the device library and any host tool agree on the wire
 * format.  tt_field_bounds() returns 0 and fills lo/hi (inclusive) for a
 * known field, 1 otherwise.  rank is 3, or 2 for the (j,k) surface arrays.
 */
#ifndef TTGPU_FIELDS_H
#define TTGPU_FIELDS_H

#include "ttgpu.h"

static inline int tt_field_bounds(int f, int nx, int ny, int nz,
                                  int32_t lo[3], int32_t hi[3],
                                  int32_t* rank, int32_t* is_int)
{
    int n[3] = {nx, ny, nz};
    int d;
    *rank = 3;
    *is_int = 0;
    /* default: ghosted (-1:n+2) */
    for (d = 0; d < 3; ++d) { lo[d] = -1; hi[d] = n[d] + 2; }
    switch (f) {
    case TT_T: case TT_RH: case TT_ER: case TT_DELTA_ER: case TT_GD:
    case TT_PG: case TT_CS: case TT_CV: case TT_XKAPR: case TT_XKAPP:
    case TT_XKAPA: case TT_DXKAPP_DT: case TT_DXKAPA_DT: case TT_GRAVX:
    case TT_GRAVY: case TT_GRAVZ: case TT_ER_OLD: case TT_Q: case TT_ETA:
        return 0;
    case TT_V: case TT_RHQX: lo[0] = 0; return 0;
    case TT_G: case TT_RHQY: lo[1] = 0; return 0;
    case TT_H: case TT_RHQZ: lo[2] = 0; return 0;
    case TT_SIG: case TT_DISFN:
        for (d = 0; d < 3; ++d) { lo[d] = 0; hi[d] = n[d] + 1; }
        return 0;
    case TT_DIVTX: case TT_DIVTY: case TT_DIVTZ:
        for (d = 0; d < 3; ++d) { lo[d] = 0; hi[d] = n[d] + 1; }
        lo[f - TT_DIVTX] = 1;
        return 0;
    case TT_STELLARINPUT: lo[0] = 1; hi[0] = nx; return 0;
    case TT_STELLARTAU: case TT_DTAUS: lo[0] = 1; hi[0] = nx + 1; return 0;
    case TT_RHS_ER:
    case TT_SOR_A: case TT_SOR_B: case TT_SOR_C: case TT_SOR_D:
    case TT_SOR_E: case TT_SOR_F: case TT_SOR_G:
    case TT_EVAL1: case TT_EVAL2: case TT_EVAL3: case TT_EVAL4: case TT_EVAL5:
    case TT_EVAL6: case TT_EVAL7: case TT_EVAL8: case TT_EVAL_1_6:
    case TT_JACOB_ETA: case TT_FQ: case TT_FR:
    case TT_CONV_FLUX1: case TT_CONV_FLUX2: case TT_CONV_FLUX3:
        for (d = 0; d < 3; ++d) { lo[d] = 1; hi[d] = n[d]; }
        return 0;
    case TT_FLIMX: case TT_DIFRX: case TT_FLUX_X:
        for (d = 0; d < 3; ++d) { lo[d] = 1; hi[d] = n[d]; }
        hi[0] = nx + 1; return 0;
    case TT_FLIMY: case TT_DIFRY: case TT_FLUX_Y:
        for (d = 0; d < 3; ++d) { lo[d] = 1; hi[d] = n[d]; }
        hi[1] = ny + 1; return 0;
    case TT_FLIMZ: case TT_DIFRZ: case TT_FLUX_Z:
        for (d = 0; d < 3; ++d) { lo[d] = 1; hi[d] = n[d]; }
        hi[2] = nz + 1; return 0;
    case TT_CENTCORX: case TT_CENTCORY: case TT_CENTCORZ:
        for (d = 0; d < 3; ++d) { lo[d] = 1; hi[d] = n[d] + 1; }
        return 0;
    case TT_UPPERBND: case TT_MU_STAR:
        *rank = 2;
        *is_int = (f == TT_UPPERBND);
        lo[0] = -1; hi[0] = ny + 2;
        lo[1] = -1; hi[1] = nz + 2;
        lo[2] = 0; hi[2] = 0;
        return 0;
    default:
        return 1;
    }
}

/* Declared bounds of the 1-D metric arrays (ppmod.f:605-642). */
static inline int tt_metric_bounds(int m, int nx, int ny, int nz, int32_t* lo, int32_t* hi)
{
    switch (m) {
    case TT_M_XXA: *lo = 0; *hi = nx + 2; return 0;
    case TT_M_XXB: *lo = -1; *hi = nx + 2; return 0;
    case TT_M_DXA: *lo = 1; *hi = nx + 2; return 0;
    case TT_M_DXB: *lo = 0; *hi = nx + 2; return 0;
    case TT_M_XYA: *lo = 0; *hi = ny + 2; return 0;
    case TT_M_XYB: *lo = -1; *hi = ny + 2; return 0;
    case TT_M_DYA: *lo = 1; *hi = ny + 2; return 0;
    case TT_M_DYB: *lo = 0; *hi = ny + 2; return 0;
    case TT_M_XZA: *lo = 0; *hi = nz + 2; return 0;
    case TT_M_XZB: *lo = -1; *hi = nz + 2; return 0;
    case TT_M_DZA: *lo = 1; *hi = nz + 2; return 0;
    case TT_M_DZB: *lo = 0; *hi = nz + 2; return 0;
    case TT_M_DDX0: case TT_M_DDX1: *lo = 0; *hi = nx + 2; return 0;
    case TT_M_DDY0: case TT_M_DDY1: *lo = 0; *hi = ny + 2; return 0;
    case TT_M_DDZ0: case TT_M_DDZ1: *lo = 0; *hi = nz + 2; return 0;
    case TT_M_DCY0: case TT_M_DCY1: *lo = 1; *hi = ny + 2; return 0;
    case TT_M_DDX0KEP: case TT_M_DDX1KEP: *lo = 1; *hi = nx + 2; return 0;
    case TT_M_SURXA: case TT_M_SURXB: case TT_M_VOLXA: case TT_M_VOLXB:
    case TT_M_CCXA: case TT_M_CCXB: case TT_M_GEOXGA: case TT_M_GEOXHA:
        *lo = 0; *hi = nx + 2; return 0;
    case TT_M_SURYA: case TT_M_SURYB: case TT_M_VOLYA: case TT_M_VOLYB:
    case TT_M_CCYA: case TT_M_CCYB:
        *lo = 0; *hi = ny + 2; return 0;
    case TT_M_SURZA: case TT_M_SURZB: case TT_M_VOLZA: case TT_M_VOLZB:
    case TT_M_CCZA: case TT_M_CCZB: case TT_M_GEOZGA:
        *lo = 0; *hi = nz + 2; return 0;
    case TT_M_GEOXG: case TT_M_GEOXH: *lo = -1; *hi = nx + 2; return 0;
    case TT_M_GEOZG: *lo = -1; *hi = nz + 2; return 0;
    default: return 1;
    }
}

#endif /* TTGPU_FIELDS_H */
