// tt_state.hip -- device state of libttgpu.so: allocation, the grid (GITTER),
// host<->device transfer in the reference's wire format, errors.
//
// Replaces the device side of ppmod.f's allocate*var routines, GITTER
// (ppinit.f:1028-1406), GASINIT (ppinit.f:1408-1417) and CALCMUSTAR
// (ppsub.f:156-179).
#include <cmath>
#include <cstring>
#include <vector>

#include "tt_common.h"

namespace tt {

__global__ __launch_bounds__(256) void k_fold(const double* __restrict__ part, int nb, int nvec, int op, double* o0,
                                              double* o1, double* o2, const int* run) {
    if (skip_run(run)) return;
    __shared__ double sh[4];
    for (int v = 0; v < nvec; ++v) {
        const double id = op == 0 ? 0.0 : (op == 1 ? -1e308 : 1e308);
        double s = id;
        for (int q = threadIdx.x; q < nb; q += 256) {
            double u = part[(long long)v * nb + q];
            s = op == 0 ? s + u : (op == 1 ? fmax(s, u) : fmin(s, u));
        }
        s = block_reduce256(s, sh, op);
        if (threadIdx.x == 0) *(v == 0 ? o0 : (v == 1 ? o1 : o2)) = s;
    }
}

// the control block word by word into the mapped host mirror
__global__ void k_ctrl_snap(const Ctrl* __restrict__ src, Ctrl* dst) {
    constexpr int nw = (int)(sizeof(Ctrl) / sizeof(unsigned int));
    const unsigned int* s = (const unsigned int*)src;
    unsigned int* d = (unsigned int*)dst;
    for (int q = threadIdx.x; q < nw; q += blockDim.x) d[q] = s[q];
}

int ctrl_to_host() {
    static_assert(sizeof(Ctrl) % sizeof(unsigned int) == 0, "control block of whole words");
    k_ctrl_snap<<<1, 64, 0, L.s>>>(L.ctrl, L.hctrl_d);
    TT_LAUNCH();
    TT_HIP(hipStreamSynchronize(L.s));
    return 0;
}

int fold_partials(const double* part, int nb, int nvec, int op, double* out0, double* out1, double* out2) {
    k_fold<<<1, 256, 0, L.s>>>(part, nb, nvec, op, out0, out1, out2, L.run_if);
    TT_LAUNCH();
    return 0;
}


Lib L;

int fail(int code, const std::string& msg) {
    L.err = msg;
    return code;
}

// ---------------------------------------------------------------- GITTER
// Host restatement of GITTER for NCOSYS=2 and uniform grids, in the
// reference's operation order (the metric is an input of every kernel, so it
// is computed once on the host and uploaded).
static void gitter_host(double* mb, int len, const tt_params& p) {
    const int NX = p.nx, NY = p.ny, NZ = p.nz;
    auto M = [&](int m, int i) -> double& { return mb[(size_t)m * len + i + 2]; };
    double YMAX1 = p.ymax * 2.0 * PI, YMIN1 = p.ymin * 2.0 * PI;
    double ZMAX1 = (p.zmax / 2.0) * PI, ZMIN1 = (p.zmin / 2.0) * PI;
    auto axis = [&](int n, double lo, double hi, int A_, int B_, int DA, int DB) {
        double del = (hi - lo) / (double)n;
        for (int i = 1; i <= n + 2; ++i) M(DA, i) = del;
        M(A_, 1) = lo + (0.0 * (hi - lo) / 1.0);
        M(A_, 0) = M(A_, 1) - M(DA, 1);
        for (int i = 2; i <= n + 2; ++i) M(A_, i) = M(A_, i - 1) + M(DA, i);
        for (int i = 0; i <= n + 1; ++i) M(B_, i) = (M(A_, i) + M(A_, i + 1)) / 2.0;
        double d3 = M(DA, n + 2) * 1.0;
        double a3 = M(A_, n + 2) + d3;
        M(B_, n + 2) = (M(A_, n + 2) + a3) / 2.0;
        d3 = M(DA, 1) / 1.0;
        a3 = M(A_, 0) - d3;
        M(B_, -1) = (M(A_, 0) + a3) / 2.0;
        for (int i = 0; i <= n + 2; ++i) M(DB, i) = M(B_, i) - M(B_, i - 1);
    };
    axis(NX, p.xmin, p.xmax, TT_M_XXA, TT_M_XXB, TT_M_DXA, TT_M_DXB);
    axis(NY, YMIN1, YMAX1, TT_M_XYA, TT_M_XYB, TT_M_DYA, TT_M_DYB);
    axis(NZ, ZMIN1, ZMAX1, TT_M_XZA, TT_M_XZB, TT_M_DZA, TT_M_DZB);
    for (int i = 0; i <= NX + 2; ++i) { M(TT_M_DDX0, i) = 0.5; M(TT_M_DDX1, i) = 0.5; }
    for (int i = 0; i <= NY + 2; ++i) { M(TT_M_DDY0, i) = 0.5; M(TT_M_DDY1, i) = 0.5; }
    for (int i = 1; i <= NY + 2; ++i) { M(TT_M_DCY0, i) = 0.5; M(TT_M_DCY1, i) = 0.5; }
    for (int i = 0; i <= NZ + 2; ++i) { M(TT_M_DDZ0, i) = 0.5; M(TT_M_DDZ1, i) = 0.5; }
    for (int i = 1; i <= NX + 2; ++i) {
        M(TT_M_DDX0KEP, i) = M(TT_M_DDX0, i) * std::pow(M(TT_M_XXB, i - 1), 1.5);
        M(TT_M_DDX1KEP, i) = M(TT_M_DDX1, i) * std::pow(M(TT_M_XXB, i), 1.5);
    }
    double xxa1 = M(TT_M_XXA, 1);
    if (xxa1 == 0.0) xxa1 = (double)1.e-20f;
    for (int i = 0; i <= NX + 1; ++i) {
        double a1 = M(TT_M_XXA, i + 1), a0 = M(TT_M_XXA, i);
        M(TT_M_VOLXB, i) = (a1 * a1 * a1 - a0 * a0 * a0) / 3.0;
    }
    for (int i = 0; i <= NX + 2; ++i) {
        double a = M(TT_M_XXA, i), b = M(TT_M_XXB, i), bm = M(TT_M_XXB, i - 1);
        M(TT_M_SURXA, i) = a * a;
        M(TT_M_CCXA, i) = a;
        M(TT_M_VOLXA, i) = (b * b * b - bm * bm * bm) / 3.0;
        M(TT_M_SURXB, i) = b * b;
        M(TT_M_CCXB, i) = b;
        M(TT_M_GEOXGA, i) = a;
        M(TT_M_GEOXHA, i) = a;
    }
    M(TT_M_GEOXGA, 1) = xxa1;
    M(TT_M_GEOXHA, 1) = xxa1;
    for (int i = -1; i <= NX + 2; ++i) { M(TT_M_GEOXG, i) = M(TT_M_XXB, i); M(TT_M_GEOXH, i) = M(TT_M_XXB, i); }
    for (int i = 0; i <= NY + 1; ++i) M(TT_M_VOLYB, i) = M(TT_M_XYA, i + 1) - M(TT_M_XYA, i);
    for (int i = 0; i <= NY + 2; ++i) {
        M(TT_M_SURYA, i) = 1.0;
        M(TT_M_CCYA, i) = 0.0;
        M(TT_M_VOLYA, i) = M(TT_M_XYB, i) - M(TT_M_XYB, i - 1);
        M(TT_M_SURYB, i) = 1.0;
        M(TT_M_CCYB, i) = 0.0;
    }
    for (int i = 0; i <= NZ + 1; ++i) M(TT_M_VOLZB, i) = std::sin(M(TT_M_XZA, i + 1)) - std::sin(M(TT_M_XZA, i));
    for (int i = 0; i <= NZ + 2; ++i) {
        M(TT_M_SURZA, i) = std::cos(M(TT_M_XZA, i));
        M(TT_M_CCZA, i) = std::sin(M(TT_M_XZA, i)) / 2.0;
        M(TT_M_VOLZA, i) = std::fabs(std::sin(M(TT_M_XZB, i)) - std::sin(M(TT_M_XZB, i - 1)));
        M(TT_M_SURZB, i) = std::cos(M(TT_M_XZB, i));
        M(TT_M_CCZB, i) = std::sin(M(TT_M_XZB, i)) / 2.0;
        M(TT_M_GEOZGA, i) = std::cos(M(TT_M_XZA, i));
    }
    for (int i = -1; i <= NZ + 2; ++i) M(TT_M_GEOZG, i) = std::cos(M(TT_M_XZB, i));
    if (M(TT_M_GEOZGA, NZ + 1) == 0.0) M(TT_M_GEOZGA, NZ + 1) = (double)1.e-10f;
    if (M(TT_M_GEOZGA, 1) == 0.0) M(TT_M_GEOZGA, 1) = (double)1.e-10f;
}

// ------------------------------------------------------------ wire format
__global__ void k_scatter(double* box, const double* wire, Geo g, int lo0, int lo1, int lo2, int ni, int nj,
                          long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, ni, nj, lo0, lo1, lo2, i, j, k);
    box[g.idx(i, j, k)] = wire[t];
}

__global__ void k_gather(const double* box, double* wire, Geo g, int lo0, int lo1, int lo2, int ni, int nj,
                         long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k;
    sub_ijk(t, ni, nj, lo0, lo1, lo2, i, j, k);
    wire[t] = box[g.idx(i, j, k)];
}

__global__ void k_fill(double* a, double v, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// Device copies and zero fills inside the step, as plain kernels: between
// two kernels a hipMemcpyAsync/hipMemsetAsync left the queue idle 5-20 us
// (the runtime's own blit kernel and its host path), a launch 1-2 us.
// 16-byte words where both ends are aligned to them, else 4-byte words.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void k_copy16(u32x4* __restrict__ d, const u32x4* __restrict__ s, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        st_nt(&d[t], s[t]);
}
__global__ void k_copy4(uint32_t* __restrict__ d, const uint32_t* __restrict__ s, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        st_nt(&d[t], s[t]);
}
__global__ void k_zero16(u32x4* d, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        st_nt(&d[t], u32x4{0u, 0u, 0u, 0u});
}
__global__ void k_zero4(uint32_t* d, long long n) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        st_nt(&d[t], 0u);
}

// sin and cos of every theta face coordinate XZA(k), by the device's
// functions: CENTRIFCORIOLIS's cell loop evaluated them per cell
__global__ void k_scza(Met m, double* sc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.len) return;
    const double za = m(TT_M_XZA, t - 2);
    sc[2 * t] = sin(za);
    sc[2 * t + 1] = cos(za);
}

int scza_impl() {
    k_scza<<<nblk(L.metlen, 256), 256, 0, L.s>>>(met(), L.scza);
    TT_LAUNCH();
    return 0;
}

static unsigned copy_blocks(long long n) {
    const long long cap = 8192;  // 32 per CU; each thread then takes a few words
    long long b = (n + 255) / 256;
    return (unsigned)(b < cap ? (b > 0 ? b : 1) : cap);
}

int dev_copy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return 0;
    if (bytes % 4) return fail(1, "dev_copy: size not a multiple of 4 bytes");
    if (((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0) {
        const long long n = (long long)(bytes / 16);
        k_copy16<<<copy_blocks(n), 256, 0, L.s>>>((u32x4*)dst, (const u32x4*)src, n);
    } else {
        const long long n = (long long)(bytes / 4);
        k_copy4<<<copy_blocks(n), 256, 0, L.s>>>((uint32_t*)dst, (const uint32_t*)src, n);
    }
    TT_LAUNCH();
    return 0;
}

int dev_zero(void* dst, size_t bytes) {
    if (bytes == 0) return 0;
    if (bytes % 4) return fail(1, "dev_zero: size not a multiple of 4 bytes");
    if (((uintptr_t)dst | bytes) % 16 == 0) {
        const long long n = (long long)(bytes / 16);
        k_zero16<<<copy_blocks(n), 256, 0, L.s>>>((u32x4*)dst, n);
    } else {
        const long long n = (long long)(bytes / 4);
        k_zero4<<<copy_blocks(n), 256, 0, L.s>>>((uint32_t*)dst, n);
    }
    TT_LAUNCH();
    return 0;
}

// a radial profile col[i + 1], i = -1..nx+2, copied to every (j, k) of the box
__global__ void k_fill_radial(Geo g, const double* __restrict__ col, double* a) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.n3) return;
    int i, j, k;
    box_ijk(g, t, i, j, k);
    a[t] = col[i + 1];
}

static double* g_stage = nullptr;  // device staging buffer (one box)

static int xfer(int f, void* host, bool to_host) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int32_t lo[3], hi[3], rank, isint;
    if (tt_field_bounds(f, L.g.nx, L.g.ny, L.g.nz, lo, hi, &rank, &isint)) return fail(1, "unknown field id");
    if (!to_host && (f == TT_RH || isint)) L.ub_repeat = false;  // RH or upperbnd from the host
    if (!to_host && f == TT_CENTCORY) L.ccy_zero = false;
    if (rank == 2) {
        size_t n = (size_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1);
        void* d = isint ? (void*)L.ub : (void*)L.mustar;
        size_t bytes = n * (isint ? sizeof(int32_t) : sizeof(double));
        if (to_host) TT_HIP(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, L.s));
        else TT_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, L.s));
        TT_HIP(hipStreamSynchronize(L.s));
        return 0;
    }
    double* a = L.f[f];
    if (!a) return fail(1, "field not held on the device");
    if (int rc = evals_before_access(f)) return rc;
    if (to_host) {
        int rc = sor_sys_before_download(f);
        if (rc) return rc;
    } else if ((f >= TT_SOR_A && f <= TT_SOR_G) || f == TT_RHS_ER) {
        sor_sys_module_written();
    }
    int ni = hi[0] - lo[0] + 1, nj = hi[1] - lo[1] + 1, nk = hi[2] - lo[2] + 1;
    long long n = (long long)ni * nj * nk;
    if (!to_host && (f == TT_GRAVY || f == TT_GRAVZ)) {
        bool zero = true;  // +0.0 everywhere, bit for bit
        const uint64_t* w = (const uint64_t*)host;
        for (long long t = 0; t < n && zero; ++t) zero = w[t] == 0;
        (f == TT_GRAVY ? L.gy_zero : L.gz_zero) = zero;
    }
    if (to_host) {
        k_gather<<<nblk(n, 256), 256, 0, L.s>>>(a, g_stage, L.g, lo[0], lo[1], lo[2], ni, nj, n);
        TT_LAUNCH();
        TT_HIP(hipMemcpyAsync(host, g_stage, n * sizeof(double), hipMemcpyDeviceToHost, L.s));
    } else {
        TT_HIP(hipMemcpyAsync(g_stage, host, n * sizeof(double), hipMemcpyHostToDevice, L.s));
        k_scatter<<<nblk(n, 256), 256, 0, L.s>>>(a, g_stage, L.g, lo[0], lo[1], lo[2], ni, nj, n);
        TT_LAUNCH();
    }
    TT_HIP(hipStreamSynchronize(L.s));
    return 0;
}

static void free_all() {
    for (int f = 0; f < TT_FIELD_MAX; ++f) {
        if (L.f[f]) hipFree(L.f[f]);
        L.f[f] = nullptr;
    }
    double** tmp[] = {&L.ED,  &L.VD,      &L.GDN,    &L.HDN, &L.FLMX,   &L.FLMY, &L.FLMZ, &L.TOLD, &L.VH,
                      &L.GH,  &L.HH,      &L.DFQ,    &L.DE1, &L.partial, &L.mustar, &L.sincoef, &L.scza, &L.met, &L.ftab, &L.btab, &L.ctab, &L.wtab, &g_stage};
    for (auto p : tmp) {
        if (*p) hipFree(*p);
        *p = nullptr;
    }
    if (L.ub) hipFree(L.ub);
    L.ub = nullptr;
    if (L.ctrl) hipFree(L.ctrl);
    L.ctrl = nullptr;
    if (L.hctrl) hipHostFree(L.hctrl);
    L.hctrl = nullptr;
    L.hctrl_d = nullptr;
    delete[] L.hmet;
    L.hmet = nullptr;
    for (auto& e : L.step_ev) {
        if (e) hipEventDestroy(e);
        e = nullptr;
    }
    for (auto& e : L.en_ev) {
        if (e) hipEventDestroy(e);
        e = nullptr;
    }
    if (L.s) hipStreamDestroy(L.s);
    L.s = nullptr;
    L.ready = false;
}

void sor_free();
int sor_alloc();
void cnn_free();

}  // namespace tt

using namespace tt;

extern "C" {

const char* tt_last_error(void) { return L.err.c_str(); }

int64_t tt_device_bytes(void) { return L.ready ? L.dev_bytes : 0; }

int tt_field_info(int field, int32_t lo[3], int32_t hi[3], int32_t* rank, int32_t* is_int) {
    return tt_field_bounds(field, L.g.nx, L.g.ny, L.g.nz, lo, hi, rank, is_int);
}

int tt_metric_info(int metric, int32_t* lo, int32_t* hi) {
    return tt_metric_bounds(metric, L.g.nx, L.g.ny, L.g.nz, lo, hi);
}

void tt_finalize(void) {
    if (L.s) hipStreamSynchronize(L.s);
    sor_free();
    cnn_free();
    visc_free();
    free_all();
}

int tt_init(const tt_params* p) {
    if (L.ready) tt_finalize();
    L = Lib();
    if (!p) return fail(1, "null params");
    if (p->nx < 4 || p->ny < 4 || p->nz < 4) return fail(1, "nx, ny, nz must be >= 4");
    if (p->nx % 2 || p->ny % 2) return fail(1, "nx and ny must be even (red-black colouring across the phi wrap)");
    // the switch values the kernels implement (everything else is rejected,
    // not silently treated as one of these)
    if (p->bcxmin != 12 && p->bcxmin != 9) return fail(1, "BCXMIN must be 9 or 12");
    if (p->ibdrad != 3 && p->ibdrad != 4 && p->ibdrad != 7 && p->ibdrad != 9)
        return fail(1, "IBDRAD must be 3, 4, 7 or 9");
    if (p->ikaptyp != 1 && !(p->ikaptyp >= 4 && p->ikaptyp <= 14) && p->ikaptyp != 26 &&
        !(p->ikaptyp >= 20 && p->ikaptyp <= 25 && p->ikaptyp != 22))
        return fail(1, "IKAPTYP must be 1, 4-14, 20, 21, 23-25 or 26");
    if (p->iflimtyp != 1 && p->iflimtyp != 2) return fail(1, "IFLIMTYP must be 1 or 2");
    if (p->f_incident != 0 && p->f_incident != 2) return fail(1, "F_INCIDENT must be 0 or 2");
    if (p->isigtyp != 0 && p->isigtyp != 1) return fail(1, "ISIGTYP must be 0 or 1");
    if (p->isoth != 0) return fail(1, "ISOTH must be 0");
    if (p->igrav != 0 && p->igrav != 1 && p->igrav != 4) return fail(1, "IGRAV must be 0, 1 or 4");
    if (p->igrav == 4 && !(p->mstar2 > 0.0)) return fail(1, "IGRAV 4 needs MSTAR2 > 0");
    if (p->nx + 4 > 16384) return fail(1, "nx must be below 16380");
    // The host waits on the device many times per step (SOR batches, the
    // defect correction's residuals, Newton and SCHRITT reads); a spinning
    // wait wakes in microseconds where a sleeping one takes tens.  The flag
    // only takes effect before the process's first HIP context, so its
    // refusal afterwards is ignored.
    (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
    (void)hipGetLastError();
    int ndev = 0;
    TT_HIP(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(10, "no HIP device");
    TT_HIP(hipSetDevice(0));
    L.p = *p;
    Geo& g = L.g;
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
    g.gx = g.nx + 4; g.gy = g.ny + 4; g.gz = g.nz + 4;
    g.n3 = (long long)g.gx * g.gy * g.gz;
    if (g.n3 >= (1LL << 31)) return fail(1, "grid too large: the ghosted box must hold fewer than 2^31 cells");
    TT_HIP(hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking));
    for (auto& e : L.step_ev) TT_HIP(hipEventCreate(&e));
    for (auto& e : L.en_ev) TT_HIP(hipEventCreate(&e));
    L.npartial = 8192;  // per-block partials: 2 x npartial doubles, allocated below
    // the SOR working set first: the hot loop's buffers get the first pages
    int rc = sor_alloc();
    if (rc) return rc;
    const size_t bytes = (size_t)g.n3 * sizeof(double);
    for (int f = 1; f < TT_FIELD_MAX; ++f) {
        int32_t lo[3], hi[3], rank, isint;
        if (tt_field_bounds(f, g.nx, g.ny, g.nz, lo, hi, &rank, &isint) || rank != 3) continue;
        TT_HIP(dev_alloc(&L.f[f], bytes));
        TT_HIP(hipMemsetAsync(L.f[f], 0, bytes, L.s));
    }
    double** tmp[] = {&L.ED, &L.VD, &L.GDN, &L.HDN, &L.FLMX, &L.FLMY, &L.FLMZ, &L.TOLD, &L.VH, &L.GH, &L.HH,
                      &L.DFQ, &L.DE1, &g_stage};
    for (auto q : tmp) {
        TT_HIP(dev_alloc(q, bytes));
        TT_HIP(hipMemsetAsync(*q, 0, bytes, L.s));
    }
    TT_HIP(dev_alloc(&L.partial, (size_t)L.npartial * 2 * sizeof(double)));
    TT_HIP(dev_alloc(&L.ub, (size_t)g.gy * g.gz * sizeof(int32_t)));
    TT_HIP(hipMemsetAsync(L.ub, 0, (size_t)g.gy * g.gz * sizeof(int32_t), L.s));
    TT_HIP(dev_alloc(&L.mustar, (size_t)g.gy * g.gz * sizeof(double)));
    TT_HIP(dev_alloc(&L.sincoef, (size_t)g.gy * g.gz * sizeof(double)));
    TT_HIP(dev_alloc(&L.ftab, (size_t)6 * TT_FREEDMAN_NT * TT_FREEDMAN_NP * sizeof(double)));
    TT_HIP(dev_alloc(&L.ctab, (size_t)16 * TT_FREEDMAN_NT * TT_FREEDMAN_NP * sizeof(double)));
    TT_HIP(dev_alloc(&L.btab, (size_t)2 * (3 * TT_BURROWS_NT * TT_BURROWS_NR + TT_BURROWS_NT) * sizeof(double)));
    TT_HIP(dev_alloc(&L.wtab, (size_t)TT_NWAVE * TT_BURROWS_NT * TT_BURROWS_NR * sizeof(double)));
    TT_HIP(dev_alloc(&L.ctrl, sizeof(Ctrl)));
    TT_HIP(hipMemsetAsync(L.ctrl, 0, sizeof(Ctrl), L.s));
    TT_HIP(hipHostMalloc((void**)&L.hctrl, sizeof(Ctrl), hipHostMallocDefault));
    TT_HIP(hipHostGetDevicePointer((void**)&L.hctrl_d, L.hctrl, 0));
    int nmax = g.nx > g.ny ? g.nx : g.ny;
    if (g.nz > nmax) nmax = g.nz;
    L.metlen = nmax + 6;
    L.hmet = new double[(size_t)TT_M_COUNT * L.metlen]();
    gitter_host(L.hmet, L.metlen, *p);
    TT_HIP(dev_alloc(&L.met, (size_t)TT_M_COUNT * L.metlen * sizeof(double)));
    TT_HIP(hipMemcpyAsync(L.met, L.hmet, (size_t)TT_M_COUNT * L.metlen * sizeof(double), hipMemcpyHostToDevice, L.s));
    TT_HIP(dev_alloc(&L.scza, (size_t)2 * L.metlen * sizeof(double)));
    if ((rc = scza_impl())) return rc;
    // CALCMUSTAR (ppsub.f:156-179)
    std::vector<double> mu((size_t)g.gy * g.gz);
    for (int k = -1; k <= g.nz + 2; ++k)
        for (int j = -1; j <= g.ny + 2; ++j) {
            double v = 1.0;
            if (p->f_incident == 2 || p->f_incident == 3 || p->f_incident == 8) {
                double a = std::cos(L.hmet[(size_t)TT_M_XYB * L.metlen + j + 2]) *
                           std::cos(L.hmet[(size_t)TT_M_XZB * L.metlen + k + 2]);
                double lim = (double)1.e-7f;
                v = a > lim ? a : lim;
            }
            mu[g.idx2(j, k)] = v;
        }
    TT_HIP(hipMemcpyAsync(L.mustar, mu.data(), mu.size() * sizeof(double), hipMemcpyHostToDevice, L.s));
    // GASINIT: cv = Rgas/(mu (gamma-1)) (ppinit.f:1415)
    double cv = RGAS / (p->mu_gas * (p->gamma - 1.0));
    k_fill<<<nblk(g.n3, 256), 256, 0, L.s>>>(L.f[TT_CV], cv, g.n3);
    TT_LAUNCH();
    // FGRAV (pp.f:461).  IGRAV 4, NCOSYS 2 (ppfor.f:753-766): the point mass
    // MPLANET = MSOL MSTAR2; gravx(-1) uses xxa(0) - dxa(1) (xxa(-1) is not
    // defined), then gravx(i) = -GRAV MPLANET / xxa(i)^2, the same in every
    // (j, k); gravy = gravz = 0 (zero since allocation).  GRAV and MSOL are
    // REAL literals in DOUBLE PRECISION parameters (ppmod.f:1075,1079).
    if (p->igrav == 4) {
        const double grav = (double)6.672e-8f, msol = (double)1.989e33f;
        const double mplanet = msol * p->mstar2;
        const double* xxa = L.hmet + (size_t)TT_M_XXA * L.metlen + 2;  // xxa[i], i >= 0
        const double* dxa = L.hmet + (size_t)TT_M_DXA * L.metlen + 2;
        std::vector<double> col((size_t)g.gx);
        const double r0 = xxa[0] - dxa[1];
        col[0] = -(grav * mplanet / (r0 * r0));
        for (int i = 0; i <= g.nx + 2; ++i) col[i + 1] = -(grav * mplanet / (xxa[i] * xxa[i]));
        TT_HIP(hipMemcpyAsync(L.partial, col.data(), col.size() * sizeof(double), hipMemcpyHostToDevice, L.s));
        k_fill_radial<<<nblk(g.n3, 256), 256, 0, L.s>>>(g, L.partial, L.f[TT_GRAVX]);
        TT_LAUNCH();
        TT_HIP(hipStreamSynchronize(L.s));  // col is released after this block
    }
    L.opts.precision = 0;
    L.opts.guess = 0;
    L.opts.batch = 16;
    L.opts.inner_max = 4000;
    L.opts.adjust_omega = 0;
    L.lim = tt_limits{};
    L.learned_rejects = 0;
    L.learned_idle = 0;
    L.start_avg[0] = L.start_avg[1] = 0.0;
    L.start_picks = 0;
    L.opts.inner_tol = 1e-5;
    L.opts.warm_scale = 0.9;
    L.opts.inner = 0;
    L.opts.line_omega = 0.0;  // chosen by a search on the first line solve
    L.line_om_auto = 0.0;
    L.line_om_trials = 0;
    TT_HIP(hipStreamSynchronize(L.s));
    evals_reset();
    L.ub_repeat = false;
    L.ccy_zero = true;  // every field was zeroed above
    L.ready = true;
    return 0;
}

int tt_upload(int field, const void* host) { return xfer(field, const_cast<void*>(host), false); }
int tt_download(int field, void* host) { return xfer(field, host, true); }

int tt_upload_metric(int metric, const double* host) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int32_t lo, hi;
    if (tt_metric_bounds(metric, L.g.nx, L.g.ny, L.g.nz, &lo, &hi)) return fail(1, "unknown metric");
    if (int rc = evals_materialize()) return rc;  // the pending EVALs are those of the old metrics
    for (int i = lo; i <= hi; ++i) L.hmet[(size_t)metric * L.metlen + i + 2] = host[i - lo];
    TT_HIP(hipMemcpyAsync(L.met + (size_t)metric * L.metlen, L.hmet + (size_t)metric * L.metlen,
                          L.metlen * sizeof(double), hipMemcpyHostToDevice, L.s));
    if (metric == TT_M_XZA)
        if (int rc = scza_impl()) return rc;
    TT_HIP(hipStreamSynchronize(L.s));
    return 0;
}

int tt_download_metric(int metric, double* host) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    int32_t lo, hi;
    if (tt_metric_bounds(metric, L.g.nx, L.g.ny, L.g.nz, &lo, &hi)) return fail(1, "unknown metric");
    for (int i = lo; i <= hi; ++i) host[i - lo] = L.hmet[(size_t)metric * L.metlen + i + 2];
    return 0;
}

int tt_sync(void) {
    if (!L.ready) return fail(1, "tt_init has not been called");
    TT_HIP(hipStreamSynchronize(L.s));
    return 0;
}

int tt_set_solver(const tt_solver_opts* o) {
    if (!o) return fail(1, "null options");
    if (o->precision < 0 || o->precision > 1 || o->guess < 0 || o->guess > 2 || o->batch < 1 ||
        o->adjust_omega < 0 || o->adjust_omega > 1 || o->inner < 0 || o->inner > 1)
        return fail(1, "invalid solver options");
    if (o->inner == 1 && !(o->line_omega == 0.0 || (o->line_omega > 0.0 && o->line_omega < 2.0)))
        return fail(1, "invalid solver options: line_omega must be 0 (chosen by a search) or lie in (0, 2)");
    if (o->guess != L.opts.guess || o->precision != L.opts.precision || o->inner != L.opts.inner) {
        // sweep counts and rejections of another mode say nothing about this one
        L.start_avg[0] = L.start_avg[1] = 0.0;
        L.start_picks = 0;
        L.learned_rejects = 0;
        L.learned_idle = 0;
    }
    L.opts = *o;
    L.line_om_auto = 0.0;  // line_omega 0: a new search on the next line solve
    L.line_om_trials = 0;
    return 0;
}

int tt_get_solver(tt_solver_opts* o) {
    if (!o) return fail(1, "null options");
    *o = L.opts;
    return 0;
}

}  // extern "C"
