// heat_lib.hip -- host side of the MI355X heat solver: the C ABI of heat.h.
//
// One handle = one mesh on the current HIP device with its own stream.  The
// hot path (operator apply, PCG vector passes, inner products, ABFT) runs in
// the kernels of heat_kernels.h; alpha, beta and the convergence state stay
// in device memory and the host only polls a flag every few iterations.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "heat.h"
#include "heat_kernels.h"
#include "heat_pcg_fused.h"
#include "heat_mrhs.h"
#include "heat_mrhs_solve.h"
#include "heat_cg1.h"

using namespace heat;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(HEAT_E_HIP, "%s:%d: %s: %s", __FILE__, __LINE__, #call,         \
                        hipGetErrorString(e_));                                         \
    } while (0)

#define CHECK_LAUNCH() HIP_TRY(hipGetLastError())

// HIP_TRY with a clean-up action in place of the plain return
#define HIP_TRY_OR(call, action)                                                        \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fail(HEAT_E_HIP, "%s:%d: %s: %s", __FILE__, __LINE__, #call,                \
                 hipGetErrorString(e_));                                                \
            action;                                                                     \
        }                                                                               \
    } while (0)

inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct heat_op {
    heat_config cfg;
    Grid g;
    hipStream_t stream = nullptr;
    int colored_impl = HEAT_COLORED_TDMA;
    int atomic_impl = HEAT_ATOMIC_TILED;   // --scatter atomic: tiled scatter, or the element kernel
    int march_cx = 8;
    int tdma_res[2][2][3] = {};      // resident blocks of k_apply_tdma [float][mass][tile shape]
    int tile_shape = 0;
    int pcg_chunk = 8;
    int use_graph = 1;
    double m3 = 0, k2 = 0;
    double Mh[64], Kh[64];
    double stats[4];
    double chk_tol = 1e-10;
    double nwM = 0, nwA = 0;   // ||M 1_int||, ||A 1_int||
    unsigned nbn = 0;          // node blocks
    // device memory
    double* kcell = nullptr;
    double* rhoc = nullptr;
    double* cr = nullptr;
    double* ck = nullptr;
    float* cr32 = nullptr;
    float* ck32 = nullptr;
    double* vec[11] = {};
    double* partials = nullptr;
    DevState* st = nullptr;
    DevState* hst = nullptr;    // pinned host mirror
    int* hflag = nullptr;       // pinned done-flag mailboxes
    int* hdone = nullptr;       // pinned word that the fused K2 stores the done flag into
    int* hdone_dev = nullptr;   // its device address
    double* dhist = nullptr;
    int hist_cap = 0;
    double* cur_hist = nullptr;  // history rows written by the running solve, or null
    // FP32 inner-solve vectors
    float *e32 = nullptr, *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *p32b = nullptr, *ap32 = nullptr, *d32 = nullptr;
    double* p2 = nullptr;       // second search-direction buffer of the fused PCG
    // level R arming (host copy, passed by value)
    FiArm fi{0, 0, 0};
    int fi_level = 0;
    // fused PCG (three launches per iteration)
    int pcg_mode = HEAT_PCG_FUSED;
    struct { int armed = 0, vec = 0, bit = 0, iter = 0; long long idx = 0; } pcg_fi;   // one-shot PCG-state SEU
    int pcg_fi_fired = 0;
    int mixed_status = 0;          // how the last refined solve ended (heat_pcg_mixed_status)
    FusedState* fst = nullptr;
    FusedState* hfst = nullptr;   // pinned mirror
    double* fsp = nullptr;        // super-block partials, 5 x nsb
    double* ftp = nullptr;        // fused-apply tile partials, 4 x ntiles
    int nsb = 0, fnby = 0, fnbz = 0, fnch = 0, fntiles = 0;
    // tile geometry of the PCG kernels: target planes per chunk and the
    // whole-wave fill (fused_chunks) of [0] the fused iteration and [1] the
    // one-launch inner iteration; the partial arrays hold every chunk count
    int geo_planes[2] = {FCX, FCX}, geo_wave[2] = {FWAVE, FWAVE};
    int batch_refs = 0;            // batches whose views hold this handle's geometry
    int mrhs_cx = 0;               // node planes per block of the 16-RHS apply (0: not chosen yet)
    // the one-launch inner iteration of the refined solve (heat_cg1.h)
    int inner_mode = HEAT_INNER_ONE_LAUNCH;
    float *c1_r = nullptr, *c1_s[2] = {nullptr, nullptr}, *c1_w[2] = {nullptr, nullptr};
    double* c1_tp = nullptr;               // 2 x C1ROWS x tiles
    float* c1_dinv = nullptr;              // FP32 reciprocal of the diagonal
    int c1_nby = 0, c1_nbz = 0, c1_nch = 0, c1_ntiles = 0;
    // captured PCG chunk
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void* graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace {

// build_reference_matrices (src/heat_solver_het.cu:661-708): 2x2x2 Gauss
// integration of the Q1 mass and stiffness matrices and their closed-form
// checks.  Also sets the closed-form entries used by the fast rows.
int build_reference_matrices(heat_op* op) {
    const double h = op->g.h;
    const double detJ = h * h * h / 8.0;
    const double iJ = 2.0 / h;
    const double gp = 1.0 / 1.7320508075688772;
    const double pts[2] = {-gp, gp};
    static const int sg[8][3] = {{-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},
                                 {-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1}};
    double M[64] = {0}, K[64] = {0};
    for (int a = 0; a < 2; ++a) for (int b = 0; b < 2; ++b) for (int c = 0; c < 2; ++c) {
        double N[8], dN[8][3];
        for (int i = 0; i < 8; ++i) {
            double u = 1 + sg[i][0] * pts[a], v = 1 + sg[i][1] * pts[b], w = 1 + sg[i][2] * pts[c];
            N[i] = 0.125 * u * v * w;
            dN[i][0] = 0.125 * sg[i][0] * v * w;
            dN[i][1] = 0.125 * u * sg[i][1] * w;
            dN[i][2] = 0.125 * u * v * sg[i][2];
        }
        for (int i = 0; i < 8; ++i) for (int j = 0; j < 8; ++j) {
            M[i * 8 + j] += detJ * N[i] * N[j];
            K[i * 8 + j] += detJ * (iJ * dN[i][0] * iJ * dN[j][0] + iJ * dN[i][1] * iJ * dN[j][1] +
                                    iJ * dN[i][2] * iJ * dN[j][2]);
        }
    }
    const double V = h * h * h;
    const double Kd = (h * h / h + h * h / h + h * h / h) / 9.0;
    double errM = 0, sumM = 0, errKrow = 0, errKdiag = 0, errSym = 0;
    for (int i = 0; i < 8; ++i) {
        double rs = 0;
        for (int j = 0; j < 8; ++j) {
            int d = (sg[i][0] != sg[j][0]) + (sg[i][1] != sg[j][1]) + (sg[i][2] != sg[j][2]);
            errM = std::fmax(errM, std::fabs(M[i * 8 + j] - V / 216.0 * (8 >> d)) / (V / 216.0 * 8));
            sumM += M[i * 8 + j];
            rs += K[i * 8 + j];
            errSym = std::fmax(errSym, std::fabs(K[i * 8 + j] - K[j * 8 + i]) + std::fabs(M[i * 8 + j] - M[j * 8 + i]));
        }
        errKrow = std::fmax(errKrow, std::fabs(rs) / Kd);
        errKdiag = std::fmax(errKdiag, std::fabs(K[i * 9] - Kd) / Kd);
    }
    const double errSum = std::fabs(sumM - V) / V;
    memcpy(op->Mh, M, sizeof M);
    memcpy(op->Kh, K, sizeof K);
    op->m3 = (h * h * h) / 216.0;
    op->k2 = -h / 12.0;
    if (!(errM < 1e-12 && errSum < 1e-12 && errKrow < 1e-12 && errKdiag < 1e-12 && errSym < 1e-12 * Kd))
        return fail(HEAT_E_SELFCHECK,
                    "Reference matrices: M pattern %.1e, sum(M)-V %.1e, K rowsum %.1e, K diag %.1e, "
                    "symmetry %.1e -> FAIL", errM, errSum, errKrow, errKdiag, errSym);
    return HEAT_OK;
}

// the projected colour-ordered fast apply runs in one node-centric launch
bool gather_path(const heat_op* op) {
    return op->cfg.kernel == HEAT_KERNEL_FAST && op->cfg.scatter == HEAT_SCATTER_COLORED &&
           op->colored_impl != HEAT_COLORED_SWEEP;
}

// the tiled atomic scatter of the fast rows (k_apply_scatter)
bool tiled_atomic_path(const heat_op* op) {
    return op->cfg.kernel == HEAT_KERNEL_FAST && op->cfg.scatter == HEAT_SCATTER_ATOMIC &&
           op->atomic_impl == HEAT_ATOMIC_TILED;
}

// PCG iterations of these paths run in captured chunks with no host check
bool chunked_path(const heat_op* op) { return gather_path(op) || tiled_atomic_path(op); }

// Resident blocks of one kernel on the whole device (CUs x blocks per CU).
int resident_blocks(const void* kernel, int threads) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess)
        return 0;
    return ncu * per;
}

// Chunks per (y,z) column for k_apply_tdma: about `planes` node planes per
// chunk, then more (shorter) chunks until the last wave of resident blocks
// is as full as the column count allows.  At most one chunk per plane pair.
int tdma_chunks(int nn1, int ncol, int planes, int resident) {
    const int npair = (nn1 + 1) / 2;
    int nch = std::min(npair, std::max(1, (nn1 + planes - 1) / planes));
    if (resident > 0) {
        const long long waves = ((long long)ncol * nch + resident - 1) / resident;
        nch = (int)std::min<long long>(npair, std::max<long long>(nch, waves * resident / ncol));
    }
    return nch;
}

template <typename T, bool MASS>
void tdma_resident(int (&r)[3]) {
    r[0] = resident_blocks((const void*)k_apply_tdma<T, MASS, 8, 32>, 256);
    r[1] = resident_blocks((const void*)k_apply_tdma<T, MASS, 4, 64>, 256);
    r[2] = resident_blocks((const void*)k_apply_tdma<T, MASS, 16, 16>, 256);
}

template <typename T, bool MASS, int TY, int TZ>
void tdma_launch(heat_op* op, const T* x, T* y, const T* cr, const T* ck, FiArm f, int check_done) {
    const int resident = op->tdma_res[sizeof(T) == 4][MASS][op->tile_shape];
    const Grid& g = op->g;
    const int nby = (g.n + TY - 1) / TY, nbz = (g.n + TZ - 1) / TZ;
    const int nch = tdma_chunks(g.nn1, nby * nbz, op->march_cx, resident);
    const int nblocks = nby * nbz * nch;
    const unsigned grid = (unsigned)(((nblocks + NXCD - 1) / NXCD) * NXCD);
    k_apply_tdma<T, MASS, TY, TZ><<<grid, TY * TZ, 0, op->stream>>>(x, y, cr, ck, g, f, nby, nbz, nch, nblocks, op->st,
                                                                   check_done);
}

template <typename T, int TY, int TZ, int CX>
void tile_launch1(heat_op* op, const T* x, T* y, const T* cr, const T* ck, bool mass, FiArm f,
                  int check_done) {
    const Grid& g = op->g;
    const int nby = (g.n + TY - 1) / TY, nbz = (g.n + TZ - 1) / TZ, nch = (g.nn1 + CX - 1) / CX;
    const int nblocks = nby * nbz * nch;
    const unsigned grid = (unsigned)(((nblocks + NXCD - 1) / NXCD) * NXCD);
    hipStream_t s = op->stream;
    if (mass)
        k_apply_tile<T, true, TY, TZ, CX><<<grid, TY * TZ, 0, s>>>(x, y, cr, ck, g, f, nby, nbz, nblocks, op->st, check_done);
    else
        k_apply_tile<T, false, TY, TZ, CX><<<grid, TY * TZ, 0, s>>>(x, y, cr, ck, g, f, nby, nbz, nblocks, op->st, check_done);
}

template <typename T, int TY, int TZ>
void tile_launch_cx(heat_op* op, const T* x, T* y, const T* cr, const T* ck, bool mass, FiArm f,
                    int check_done) {
    switch (op->march_cx) {
        case 4: tile_launch1<T, TY, TZ, 4>(op, x, y, cr, ck, mass, f, check_done); break;
        case 16: tile_launch1<T, TY, TZ, 16>(op, x, y, cr, ck, mass, f, check_done); break;
        default: tile_launch1<T, TY, TZ, 8>(op, x, y, cr, ck, mass, f, check_done); break;
    }
}

template <typename T>
int launch_tile(heat_op* op, const T* x, T* y, const T* cr, const T* ck, bool mass, FiArm f,
                int check_done) {
    if (op->colored_impl == HEAT_COLORED_TDMA) {
#define TDMA(TY, TZ)                                                                  \
    do {                                                                              \
        if (mass) tdma_launch<T, true, TY, TZ>(op, x, y, cr, ck, f, check_done);      \
        else tdma_launch<T, false, TY, TZ>(op, x, y, cr, ck, f, check_done);          \
    } while (0)
        switch (op->tile_shape) {
            case 1: TDMA(4, 64); break;
            case 2: TDMA(16, 16); break;
            default: TDMA(8, 32); break;
        }
#undef TDMA
        CHECK_LAUNCH();
        return HEAT_OK;
    }
    switch (op->tile_shape) {
        case 1: tile_launch_cx<T, 4, 64>(op, x, y, cr, ck, mass, f, check_done); break;
        case 2: tile_launch_cx<T, 16, 16>(op, x, y, cr, ck, mass, f, check_done); break;
        default: tile_launch_cx<T, 8, 32>(op, x, y, cr, ck, mass, f, check_done); break;
    }
    CHECK_LAUNCH();
    return HEAT_OK;
}

// y = P A x, colour order, node-centric (tile, march or one-thread-per-node gather)
template <typename T>
int launch_colored(heat_op* op, const T* x, T* y, const T* cr, const T* ck, bool mass, FiArm f,
                   int check_done) {
    const Grid& g = op->g;
    hipStream_t s = op->stream;
    // an armed level-R fault goes through the gather kernel (same sums, same
    // bits), so the tiled kernels carry no fault code
    if (op->colored_impl == HEAT_COLORED_GATHER || f.armed) {
        if (mass)
            k_apply_gather<T, true, EPI_NONE><<<op->nbn, NT, 0, s>>>(x, y, cr, ck, g, f, nullptr, 0.0, op->partials, op->st, check_done);
        else
            k_apply_gather<T, false, EPI_NONE><<<op->nbn, NT, 0, s>>>(x, y, cr, ck, g, f, nullptr, 0.0, op->partials, op->st, check_done);
        CHECK_LAUNCH();
        return HEAT_OK;
    }
    if (op->colored_impl == HEAT_COLORED_TILE || op->colored_impl == HEAT_COLORED_TDMA)
        return launch_tile<T>(op, x, y, cr, ck, mass, f, check_done);
    const int nqb = (g.sx + NT - 1) / NT;
    const int cx = op->march_cx;
    const int nchunks = (g.nn1 + cx - 1) / cx;
    const int nblocks = nqb * nchunks;
    const unsigned grid = (unsigned)(((nblocks + NXCD - 1) / NXCD) * NXCD);
#define MARCH(CXV)                                                                                   \
    do {                                                                                             \
        if (mass) k_apply_march<T, true, CXV><<<grid, NT, 0, s>>>(x, y, cr, ck, g, f, nqb, nblocks, op->st, check_done);  \
        else k_apply_march<T, false, CXV><<<grid, NT, 0, s>>>(x, y, cr, ck, g, f, nqb, nblocks, op->st, check_done);     \
    } while (0)
    switch (cx) {
        case 2: MARCH(2); break;
        case 4: MARCH(4); break;
        case 16: MARCH(16); break;
        case 32: MARCH(32); break;
        default: MARCH(8); break;
    }
#undef MARCH
    CHECK_LAUNCH();
    return HEAT_OK;
}

// Chunks of element planes per tile column of the tiled atomic scatter:
// about SCX planes each, never more chunks than planes.
inline int scatter_chunks(int n) { return std::max(1, std::min(n, (n + SCX - 1) / SCX)); }

// y = A x through the element-centric kernels (colour sweep or atomic).
// *projected (optional) is set when the kernel has applied P itself (the
// tiled atomic scatter, asked with project); otherwise y is unprojected.
int element_apply(heat_op* op, const double* x, double* y, double dt_eff, FiArm f, bool project = false,
                  bool* projected = nullptr, int check_done = 0) {
    const Grid& g = op->g;
    hipStream_t s = op->stream;
    if (projected) *projected = false;
    if (check_done) k_zero<true><<<zero_blocks(g.nn), NT, 0, s>>>(y, g.nn, op->st);
    else k_zero<false><<<zero_blocks(g.nn), NT, 0, s>>>(y, g.nn, op->st);
    CHECK_LAUNCH();
    const bool gauss = op->cfg.kernel == HEAT_KERNEL_GAUSS;
    if (tiled_atomic_path(op) && !f.armed) {
        // the tiled scatter (an armed level-R fault goes through the
        // element kernel, which carries the flip)
        const int tz = scatter_tile_z(g.n), ty = SNT / tz;
        const int nby = (g.n + ty - 1) / ty, nbz = (g.n + tz - 1) / tz, nch = scatter_chunks(g.n);
        const int nblocks = nby * nbz * nch;
        const unsigned grid = (unsigned)(((nblocks + NXCD - 1) / NXCD) * NXCD);
        const bool mass = dt_eff == 0.0;
        const DevState* st = op->st;
        auto launch = [&](auto tzc) {
            constexpr int TZ = decltype(tzc)::value;
            if (mass && project)
                k_apply_scatter<true, true, TZ><<<grid, SNT, 0, s>>>(x, y, op->cr, op->ck, g, nby, nbz, nch, nblocks, st, check_done);
            else if (mass)
                k_apply_scatter<true, false, TZ><<<grid, SNT, 0, s>>>(x, y, op->cr, op->ck, g, nby, nbz, nch, nblocks, st, check_done);
            else if (project)
                k_apply_scatter<false, true, TZ><<<grid, SNT, 0, s>>>(x, y, op->cr, op->ck, g, nby, nbz, nch, nblocks, st, check_done);
            else
                k_apply_scatter<false, false, TZ><<<grid, SNT, 0, s>>>(x, y, op->cr, op->ck, g, nby, nbz, nch, nblocks, st, check_done);
        };
        if (tz == 32) launch(std::integral_constant<int, 32>{});
        else if (tz == 64) launch(std::integral_constant<int, 64>{});
        else launch(std::integral_constant<int, 128>{});
        CHECK_LAUNCH();
        if (projected) *projected = project;
        return HEAT_OK;
    }
    if (check_done) return fail(HEAT_E_STATE, "internal: element path inside a captured solve");
    if (op->cfg.scatter == HEAT_SCATTER_ATOMIC) {
        if (gauss)
            k_apply_elem<double, 1, true><<<blocks(g.ne, ET), ET, 0, s>>>(x, y, op->kcell, op->rhoc, g, dt_eff, op->m3, op->k2, f, -1);
        else
            k_apply_elem<double, 0, true><<<blocks(g.ne, ET), ET, 0, s>>>(x, y, op->kcell, op->rhoc, g, dt_eff, op->m3, op->k2, f, -1);
        CHECK_LAUNCH();
    } else {
        for (int c = 0; c < 8; ++c) {
            const int cx = (c >> 2) & 1, cy = (c >> 1) & 1, cz = c & 1;
            const long long cnt = (long long)((g.n - cx + 1) / 2) * ((g.n - cy + 1) / 2) * ((g.n - cz + 1) / 2);
            if (cnt == 0) continue;
            if (gauss)
                k_apply_elem<double, 1, false><<<blocks(cnt, ET), ET, 0, s>>>(x, y, op->kcell, op->rhoc, g, dt_eff, op->m3, op->k2, f, c);
            else
                k_apply_elem<double, 0, false><<<blocks(cnt, ET), ET, 0, s>>>(x, y, op->kcell, op->rhoc, g, dt_eff, op->m3, op->k2, f, c);
            CHECK_LAUNCH();
        }
    }
    return HEAT_OK;
}

// alpha and/or the D1 record from the partials of an apply pass
int launch_fin_apply(heat_op* op, int epi, double nw, int check_done) {
    const int nb = (int)op->nbn;
    hipStream_t s = op->stream;
    switch (epi) {
        case EPI_NONE: return HEAT_OK;
        case EPI_D1: k_fin_apply<EPI_D1><<<1, FT, 0, s>>>(op->partials, nb, op->st, nw, check_done); break;
        case EPI_PCG: k_fin_apply<EPI_PCG><<<1, FT, 0, s>>>(op->partials, nb, op->st, nw, check_done); break;
        default: k_fin_apply<EPI_PCG_D1><<<1, FT, 0, s>>>(op->partials, nb, op->st, nw, check_done); break;
    }
    CHECK_LAUNCH();
    return HEAT_OK;
}

// apply_operator with an epilogue: EPI_NONE / EPI_D1 outside PCG, EPI_PCG /
// EPI_PCG_D1 inside (check_done: skip once the solve has finished).
int apply_impl(heat_op* op, const double* x, double* y, double dt_eff, bool allow_fi,
               bool project, int epi, int check_done) {
    const Grid& g = op->g;
    hipStream_t s = op->stream;
    FiArm f = op->fi;
    if (!allow_fi || op->fi_level != 2) f.armed = 0;
    const bool mass = (dt_eff == 0.0);
    if (!mass && dt_eff != op->cfg.dt)
        return fail(HEAT_E_ARG, "dt_eff must be 0 or the handle's dt (%g), got %g", op->cfg.dt, dt_eff);
    const double* w = mass ? op->vec[HEAT_VEC_WM] : op->vec[HEAT_VEC_WA];
    const int use_d1 = (epi == EPI_D1 || epi == EPI_PCG_D1);
    if (use_d1 && !w) return fail(HEAT_E_STATE, "D1 requested without detector vectors");
    // ||w|| was read back to the host at the end of setup
    const double nw = use_d1 ? (mass ? op->nwM : op->nwA) : 0.0;
    int rc0 = HEAT_OK;
    if (gather_path(op) && project) {
        rc0 = launch_colored<double>(op, x, y, op->cr, op->ck, mass, f, check_done);
        if (rc0) return rc0;
        if (epi != EPI_NONE) {
            // inner products of the pass, in the reference's block order
            switch (epi) {
                case EPI_D1: k_apply_post<double, EPI_D1><<<op->nbn, NT, 0, s>>>(x, y, g, 0, w, nw, op->partials, op->st, check_done); break;
                case EPI_PCG: k_apply_post<double, EPI_PCG><<<op->nbn, NT, 0, s>>>(x, y, g, 0, w, nw, op->partials, op->st, check_done); break;
                default: k_apply_post<double, EPI_PCG_D1><<<op->nbn, NT, 0, s>>>(x, y, g, 0, w, nw, op->partials, op->st, check_done); break;
            }
            CHECK_LAUNCH();
            rc0 = launch_fin_apply(op, epi, nw, check_done);
            if (rc0) return rc0;
        }
    } else {
        bool projected = false;
        int rc = element_apply(op, x, y, dt_eff, f, project, &projected, check_done);
        if (rc) return rc;
        const int pj = (project && !projected) ? 1 : 0;
        const int cd = check_done;
        switch (epi) {
            case EPI_NONE:
                if (pj) k_apply_post<double, EPI_NONE><<<op->nbn, NT, 0, s>>>(x, y, g, pj, w, nw, op->partials, op->st, cd);
                break;
            case EPI_D1: k_apply_post<double, EPI_D1><<<op->nbn, NT, 0, s>>>(x, y, g, pj, w, nw, op->partials, op->st, cd); break;
            case EPI_PCG: k_apply_post<double, EPI_PCG><<<op->nbn, NT, 0, s>>>(x, y, g, pj, w, nw, op->partials, op->st, cd); break;
            default: k_apply_post<double, EPI_PCG_D1><<<op->nbn, NT, 0, s>>>(x, y, g, pj, w, nw, op->partials, op->st, cd); break;
        }
        CHECK_LAUNCH();
        rc0 = launch_fin_apply(op, epi, nw, cd);
        if (rc0) return rc0;
    }
    if (allow_fi && op->fi.armed && op->fi_level == 2) op->fi.armed = 0;   // one-shot (:902)
    return HEAT_OK;
}

int read_state(heat_op* op) {
    HIP_TRY(hipMemcpyAsync(op->hst, op->st, sizeof(DevState), hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(hipStreamSynchronize(op->stream));
    return HEAT_OK;
}

int read_fused(heat_op* op) {
    HIP_TRY(hipMemcpyAsync(op->hfst, op->fst, sizeof(FusedState), hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(hipStreamSynchronize(op->stream));
    return HEAT_OK;
}

__global__ void k_pcg_params(DevState* st, double tol, int max_iter) {
    st->tol = tol;
    st->max_iter = max_iter;
}
__global__ void k_mixed_params(DevState* st, double tol, double tol_inner, int max_inner) {
    st->tol_outer = tol;
    st->tol = tol_inner;
    st->max_iter = max_inner;
    st->inner_total = 0;
}
__global__ void k_abft_reset(DevState* st) {
    st->abft_max = 0.0;
    st->abft_calls = 0;
}
__global__ void k_set_norms(DevState* st, int slotM, int slotA) {
    st->nwM = sqrt(st->res[slotM]);
    st->nwA = sqrt(st->res[slotA]);
}

// One PCG iteration: Ap = A p (+ p.Ap -> alpha), x/r/z update (-> beta,
// convergence), p = z + beta p.  Every kernel returns at once when done.
int launch_iteration(heat_op* op, double* x, bool detect) {
    hipStream_t s = op->stream;
    double* p = op->vec[HEAT_VEC_P];
    double* Ap = op->vec[HEAT_VEC_AP];
    double* r = op->vec[HEAT_VEC_R];
    double* z = op->vec[HEAT_VEC_Z];
    const double* d = op->vec[HEAT_VEC_DIAG];
    int rc = apply_impl(op, p, Ap, op->cfg.dt, false, true, detect ? EPI_PCG_D1 : EPI_PCG, 1);
    if (rc) return rc;
    k_pcg_update<<<op->nbn, NT, 0, s>>>(x, p, r, Ap, d, z, op->g.nn, op->partials, op->st);
    CHECK_LAUNCH();
    k_fin_update<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, op->cur_hist);
    CHECK_LAUNCH();
    k_pcg_dir<<<op->nbn, NT, 0, s>>>(p, z, op->g.nn, op->st);
    CHECK_LAUNCH();
    return HEAT_OK;
}

__global__ void k_fused_params(FusedState* fs, double tol, int max_iter, int check_outer,
                               double tol_outer) {
    fs->tol = tol;
    fs->max_iter = max_iter;
    fs->check_outer = check_outer;
    fs->tol_outer = tol_outer;
    fs->inner_total = 0;
    fs->inner_base = 0;
    fs->outer_done = 0;
    fs->done = 0;
}

__global__ void k_fused_fault(FusedState* fs, int armed, int vec, long long idx, int bit, int iter) {
    fs->fi_armed = armed;
    fs->fi_vec = vec;
    fs->fi_idx = idx;
    fs->fi_bit = bit;
    fs->fi_iter = iter;
    fs->fi_fired = 0;
}

// The fused-kernel view of a handle's solve on vectors of type T.
template <typename T>
FusedView<T> make_view(heat_op* op, T* x, T* p0, T* p1, T* r, T* z, T* Ap, T* b, const T* d, const T* cr,
                       const T* ck, double* hist) {
    FusedView<T> v;
    v.x = x; v.P0 = p0; v.P1 = p1; v.r = r; v.z = z; v.Ap = Ap; v.b = b;
    v.d = d; v.cr = cr; v.ck = ck;
    v.w = op->vec[HEAT_VEC_WA];
    v.F = op->vec[HEAT_VEC_F];
    v.g = op->g;
    v.sp = op->fsp; v.sp_stride = op->nsb; v.nsb = op->nsb;
    v.tp = op->ftp; v.tp_stride = op->fntiles; v.ntiles = op->fntiles;
    v.nby = op->fnby; v.nbz = op->fnbz; v.nch = op->fnch;
    v.fs = op->fst; v.st = op->st;
    v.nwA = op->nwA; v.dt = op->cfg.dt;
    v.hist = hist;
    v.add_source = !(op->cfg.verify || op->cfg.no_source);
    v.active = 1;
    v.pre_sp = v.nsb > FUSED_INBLOCK_MAX;
    v.pre_tp = v.ntiles > FUSED_INBLOCK_MAX;
    v.hflag = op->hdone_dev;
    return v;
}

// one fused iteration (K1 + K2) of a view
template <typename T>
int launch_fused_iteration(heat_op* op, const FusedView<T>& v, bool detect) {
    hipStream_t s = op->stream;
    const unsigned grid1 = (unsigned)(((v.ntiles + NXCD - 1) / NXCD) * NXCD);
    if (v.pre_sp) {
        k_fused_fin_sp<T><<<1, NT, 0, s>>>(v);
        CHECK_LAUNCH();
    }
    if (op->pcg_fi.armed) {
        // an upset in the PCG state is injected by the one-launch K1
        if (detect) k_fused_apply<T, true><<<grid1, FTY * FTZ, 0, s>>>(v);
        else k_fused_apply<T, false><<<grid1, FTY * FTZ, 0, s>>>(v);
        CHECK_LAUNCH();
    } else {
        if (detect) k_fused_dir<T, true><<<grid1, FTY * FTZ, 0, s>>>(v);
        else k_fused_dir<T, false><<<grid1, FTY * FTZ, 0, s>>>(v);
        CHECK_LAUNCH();
        if (detect) k_fused_tdma<T, true><<<grid1, FTY * FTZ, 0, s>>>(v);
        else k_fused_tdma<T, false><<<grid1, FTY * FTZ, 0, s>>>(v);
        CHECK_LAUNCH();
    }
    if (v.pre_tp) {
        if (detect) k_fused_fin_tp<T, true><<<1, NT, 0, s>>>(v);
        else k_fused_fin_tp<T, false><<<1, NT, 0, s>>>(v);
        CHECK_LAUNCH();
    }
    if (detect) k_fused_update<T, true><<<v.nsb, NT, 0, s>>>(v);
    else k_fused_update<T, false><<<v.nsb, NT, 0, s>>>(v);
    CHECK_LAUNCH();
    return HEAT_OK;
}

// Launch `pcg_chunk` iterations of `body`, captured once as a hipGraph per key.
template <typename Body>
int launch_chunk_generic(heat_op* op, const void* const key[6], Body body) {
    if (op->use_graph) {
        if (!op->gexec || memcmp(key, op->graph_key, 6 * sizeof(void*)) != 0) {
            if (op->gexec) { (void)hipGraphExecDestroy(op->gexec); op->gexec = nullptr; }
            if (op->graph) { (void)hipGraphDestroy(op->graph); op->graph = nullptr; }
            HIP_TRY(hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
            int rc = HEAT_OK;
            for (int i = 0; i < op->pcg_chunk && rc == HEAT_OK; ++i) rc = body();
            hipGraph_t graph = nullptr;
            hipError_t ec = hipStreamEndCapture(op->stream, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            if (ec != hipSuccess) return fail(HEAT_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
            op->graph = graph;
            HIP_TRY(hipGraphInstantiate(&op->gexec, op->graph, nullptr, nullptr, 0));
            memcpy(op->graph_key, key, 6 * sizeof(void*));
        }
        HIP_TRY(hipGraphLaunch(op->gexec, op->stream));
        return HEAT_OK;
    }
    for (int i = 0; i < op->pcg_chunk; ++i) {
        int rc = body();
        if (rc) return rc;
    }
    return HEAT_OK;
}

int launch_chunk(heat_op* op, double* x, bool detect) {
    const void* key6[6] = {x, (const void*)(intptr_t)((detect ? 1 : 0) + 2 * op->pcg_fi.armed),
                           (const void*)(intptr_t)op->pcg_chunk,
                           op->cur_hist,
                           (const void*)(intptr_t)(op->pcg_mode + 10 * op->colored_impl + 100 * op->atomic_impl),
                           (const void*)(intptr_t)(op->march_cx + 100 * op->tile_shape)};
    if (op->pcg_mode == HEAT_PCG_FUSED && gather_path(op)) {
        return launch_chunk_generic(op, key6, [&]() {
            const FusedView<double> v = make_view<double>(op, x, op->vec[HEAT_VEC_P], op->p2, op->vec[HEAT_VEC_R],
                                                          op->vec[HEAT_VEC_Z], op->vec[HEAT_VEC_AP], nullptr,
                                                          op->vec[HEAT_VEC_DIAG], op->cr, op->ck, op->cur_hist);
            return launch_fused_iteration<double>(op, v, detect);
        });
    }
    return launch_chunk_generic(op, key6, [&]() { return launch_iteration(op, x, detect); });
}

// Launch chunks until the device reports done; the flags of chunk k are read
// while chunk k+1 runs.  With hdone (the fused solves), the update kernels
// store the flag into that pinned word themselves and no copy is queued.
template <typename LaunchFn>
int run_until_done(heat_op* op, int max_iter, LaunchFn launch, const int* dflag = nullptr,
                   const volatile int* hdone = nullptr) {
    if (!dflag) dflag = &op->st->done;
    const int max_chunks = max_iter / (op->pcg_chunk > 0 ? op->pcg_chunk : 1) + 3;
    int k = 0;
    if (hdone) {
        int rc = launch();
        if (rc) return rc;
        HIP_TRY(hipEventRecord(op->ev[0], op->stream));
        for (;;) {
            if (++k > max_chunks) return fail(HEAT_E_STATE, "PCG did not report completion");
            rc = launch();
            if (rc) return rc;
            HIP_TRY(hipEventRecord(op->ev[k & 1], op->stream));
            HIP_TRY(hipEventSynchronize(op->ev[(k - 1) & 1]));
            // every chunk ends with an update launch, which stored the flag
            // as of its iteration; a done solve stays done, so a newer value
            // from the chunk now running is as good
            if (*hdone) break;
        }
        HIP_TRY(hipStreamSynchronize(op->stream));
        return HEAT_OK;
    }
    int rc = launch();
    if (rc) return rc;
    int* slot[2] = {&op->hflag[0], &op->hflag[1]};   // pinned mailboxes, one per chunk in flight
    HIP_TRY(hipMemcpyAsync(slot[0], dflag, sizeof(int), hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(hipEventRecord(op->ev[0], op->stream));
    for (;;) {
        if (++k > max_chunks) return fail(HEAT_E_STATE, "PCG did not report completion");
        rc = launch();
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(slot[k & 1], dflag, sizeof(int), hipMemcpyDeviceToHost, op->stream));
        HIP_TRY(hipEventRecord(op->ev[k & 1], op->stream));
        HIP_TRY(hipEventSynchronize(op->ev[(k - 1) & 1]));
        if (*slot[(k - 1) & 1]) break;
    }
    HIP_TRY(hipStreamSynchronize(op->stream));
    return HEAT_OK;
}

void free_all(heat_op* op) {
    if (op->gexec) (void)hipGraphExecDestroy(op->gexec);
    if (op->graph) (void)hipGraphDestroy(op->graph);
    for (int q = 0; q < 2; ++q) {
        (void)hipFree(op->c1_s[q]);
        (void)hipFree(op->c1_w[q]);
    }
    (void)hipFree(op->c1_r);
    (void)hipFree(op->c1_tp);
    (void)hipFree(op->c1_dinv);
    for (auto& e : op->ev) if (e) (void)hipEventDestroy(e);
    (void)hipFree(op->kcell); (void)hipFree(op->rhoc); (void)hipFree(op->cr); (void)hipFree(op->ck);
    (void)hipFree(op->cr32); (void)hipFree(op->ck32);
    for (auto& v : op->vec) (void)hipFree(v);
    (void)hipFree(op->partials); (void)hipFree(op->st); (void)hipFree(op->dhist);
    (void)hipFree(op->e32); (void)hipFree(op->r32); (void)hipFree(op->z32); (void)hipFree(op->p32); (void)hipFree(op->p32b); (void)hipFree(op->p2); (void)hipFree(op->ap32); (void)hipFree(op->d32);
    if (op->hst) (void)hipHostFree(op->hst);
    if (op->hflag) (void)hipHostFree(op->hflag);
    if (op->hdone) (void)hipHostFree(op->hdone);
    if (op->hfst) (void)hipHostFree(op->hfst);
    (void)hipFree(op->fst); (void)hipFree(op->fsp); (void)hipFree(op->ftp);
    if (op->stream) (void)hipStreamDestroy(op->stream);
}

int setup_impl(const heat_config* cfg, const double* k, const double* rhoc, heat_op* op) {
    op->cfg = *cfg;
    const int n = cfg->n;
    Grid& g = op->g;
    g.n = n; g.nn1 = n + 1; g.sy = n + 1; g.sx = (n + 1) * (n + 1);
    g.nn = (n + 1) * (n + 1) * (n + 1); g.ne = n * n * n; g.h = 1.0 / n;
    op->nbn = blocks(g.nn, NT);
    // load_fields checks (src/heat_solver_het.cu:734-746)
    double st4[4] = {k[0], k[0], rhoc[0], rhoc[0]};
    for (int e = 0; e < g.ne; ++e) {
        if (!(k[e] > 0) || !(rhoc[e] > 0))
            return fail(HEAT_E_FIELD, "rank 0: non-positive or NaN coefficient at element %d", e);
        st4[0] = std::fmin(st4[0], k[e]); st4[1] = std::fmax(st4[1], k[e]);
        st4[2] = std::fmin(st4[2], rhoc[e]); st4[3] = std::fmax(st4[3], rhoc[e]);
    }
    memcpy(op->stats, st4, sizeof st4);
    int rc = build_reference_matrices(op);
    if (rc) return rc;
    HIP_TRY(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
    // occupancy queries stay out of any later stream capture
    tdma_resident<double, false>(op->tdma_res[0][0]);
    tdma_resident<double, true>(op->tdma_res[0][1]);
    tdma_resident<float, false>(op->tdma_res[1][0]);
    tdma_resident<float, true>(op->tdma_res[1][1]);
    HIP_TRY(hipEventCreateWithFlags(&op->ev[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&op->ev[1], hipEventDisableTiming));
    const size_t nb8 = sizeof(double) * (size_t)g.nn, eb8 = sizeof(double) * (size_t)g.ne;
    HIP_TRY(hipMalloc(&op->kcell, eb8));
    HIP_TRY(hipMalloc(&op->rhoc, eb8));
    HIP_TRY(hipMalloc(&op->cr, eb8));
    HIP_TRY(hipMalloc(&op->ck, eb8));
    for (int v = HEAT_VEC_T; v <= HEAT_VEC_F; ++v) HIP_TRY(hipMalloc(&op->vec[v], nb8));
    HIP_TRY(hipMalloc(&op->partials, sizeof(double) * MAXSUM * (size_t)op->nbn));
    op->nsb = (int)((g.nn + SB - 1) / SB);
    op->fnby = (n + FTY - 1) / FTY;
    op->fnbz = (n + FTZ - 1) / FTZ;
    op->fnch = fused_chunks(g.nn1, op->fnby * op->fnbz, op->geo_planes[0], op->geo_wave[0]);
    op->fntiles = op->fnby * op->fnbz * op->fnch;
    HIP_TRY(hipMalloc(&op->fsp, sizeof(double) * 5 * (size_t)op->nsb));
    HIP_TRY(hipMalloc(&op->p2, sizeof(double) * (size_t)g.nn));
    // room for the partials of every geometry (at most one chunk per plane pair)
    HIP_TRY(hipMalloc(&op->ftp, sizeof(double) * 4 * (size_t)op->fnby * op->fnbz * fused_chunks_max(g.nn1)));
    HIP_TRY(hipMalloc(&op->fst, sizeof(FusedState)));
    HIP_TRY(hipHostMalloc(&op->hfst, sizeof(FusedState), hipHostMallocDefault));
    memset(op->hfst, 0, sizeof(FusedState));
    HIP_TRY(hipMalloc(&op->st, sizeof(DevState)));
    HIP_TRY(hipHostMalloc(&op->hst, sizeof(DevState), hipHostMallocDefault));
    memset(op->hst, 0, sizeof(DevState));
    HIP_TRY(hipHostMalloc(&op->hflag, 2 * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&op->hdone, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
    *op->hdone = 0;
    HIP_TRY(hipHostGetDevicePointer((void**)&op->hdone_dev, op->hdone, 0));
    hipStream_t s = op->stream;
    HIP_TRY(hipMemcpyAsync(op->st, op->hst, sizeof(DevState), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(op->fst, op->hfst, sizeof(FusedState), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(op->kcell, k, eb8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(op->rhoc, rhoc, eb8, hipMemcpyHostToDevice, s));
    k_coef64<<<blocks(g.ne, 256), 256, 0, s>>>(op->kcell, op->rhoc, cfg->dt, op->m3, op->k2, op->cr, op->ck, g.ne);
    CHECK_LAUNCH();
    for (int v = HEAT_VEC_T; v <= HEAT_VEC_F; ++v) {
        k_fill<double><<<blocks(g.nn, NT), NT, 0, s>>>(op->vec[v], g.nn, 0.0);
        CHECK_LAUNCH();
    }
    // Jacobi diagonal (src/heat_solver_het.cu:798-804)
    const bool gauss = cfg->kernel == HEAT_KERNEL_GAUSS;
    const double mdiag = gauss ? op->Mh[0] : 8.0 * op->m3;
    const double kdiag = gauss ? op->Kh[0] : -4.0 * op->k2;
    double* d = op->vec[HEAT_VEC_DIAG];
    double* F = op->vec[HEAT_VEC_F];
    const bool atomic = cfg->scatter == HEAT_SCATTER_ATOMIC;
    for (int c = atomic ? -1 : 0; c < (atomic ? 0 : 8); ++c) {
        long long cnt = g.ne;
        if (c >= 0) {
            const int cx = (c >> 2) & 1, cy = (c >> 1) & 1, cz = c & 1;
            cnt = (long long)((n - cx + 1) / 2) * ((n - cy + 1) / 2) * ((n - cz + 1) / 2);
        }
        if (cnt == 0) continue;
        if (atomic) k_diag<true><<<blocks(cnt, ET), ET, 0, s>>>(d, op->kcell, op->rhoc, g, cfg->dt, mdiag, kdiag, c);
        else k_diag<false><<<blocks(cnt, ET), ET, 0, s>>>(d, op->kcell, op->rhoc, g, cfg->dt, mdiag, kdiag, c);
        CHECK_LAUNCH();
        if (!cfg->verify) {
            if (atomic) k_source<true><<<blocks(cnt, ET), ET, 0, s>>>(F, g, cfg->source_mode, cfg->source_value, c);
            else k_source<false><<<blocks(cnt, ET), ET, 0, s>>>(F, g, cfg->source_mode, cfg->source_value, c);
            CHECK_LAUNCH();
        }
    }
    k_dirichlet<double><<<op->nbn, NT, 0, s>>>(d, g, 1.0);
    CHECK_LAUNCH();
    k_dirichlet<double><<<op->nbn, NT, 0, s>>>(F, g, 0.0);
    CHECK_LAUNCH();
    // detectors (src/heat_solver_het.cu:940-953), computed with detect still off
    if (cfg->detect) {
        for (int v = HEAT_VEC_WM; v <= HEAT_VEC_ONES; ++v) HIP_TRY(hipMalloc(&op->vec[v], nb8));
        double* ones = op->vec[HEAT_VEC_ONES];
        k_ones_interior<<<op->nbn, NT, 0, s>>>(ones, g);
        CHECK_LAUNCH();
        FiArm none{0, 0, 0};
        rc = element_apply(op, ones, op->vec[HEAT_VEC_WM], 0.0, none);
        if (rc) return rc;
        rc = element_apply(op, ones, op->vec[HEAT_VEC_WA], cfg->dt, none);
        if (rc) return rc;
        k_dot<<<op->nbn, NT, 0, s>>>(op->vec[HEAT_VEC_WM], op->vec[HEAT_VEC_WM], g.nn, op->partials, op->st, 0);
        CHECK_LAUNCH();
        k_fin_dot<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, 0);
        CHECK_LAUNCH();
        k_dot<<<op->nbn, NT, 0, s>>>(op->vec[HEAT_VEC_WA], op->vec[HEAT_VEC_WA], g.nn, op->partials, op->st, 1);
        CHECK_LAUNCH();
        k_fin_dot<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, 1);
        CHECK_LAUNCH();
        k_set_norms<<<1, 1, 0, s>>>(op->st, 0, 1);
        CHECK_LAUNCH();
    }
    int rc2 = read_state(op);
    if (rc2) return rc2;
    op->nwM = op->hst->nwM;
    op->nwA = op->hst->nwA;
    return HEAT_OK;
}

int ensure_mixed(heat_op* op);

// Node planes per block of the 16-RHS apply of a handle's mesh, chosen once:
// the length in [MCX_MIN, MCX_MAX] that minimises (waves of resident blocks)
// x (planes + 1), the extra plane being the halo plane each chunk loads
// again; ties go to the shorter length.  The tile partition (and with it the
// summation order of the tile sums) is a function of the mesh size and the
// device's resident block count.
int mrhs_chunk(heat_op* op) {
    if (op->mrhs_cx) return op->mrhs_cx;
    const Grid& g = op->g;
    const int slots = resident_blocks((const void*)k_mrhs_apply<false>, MNT);
    const long long cols = (long long)((g.n + MTY) / MTY) * ((g.n + MTZ) / MTZ);
    int best = 16;
    if (slots > 0) {
        long long bcost = -1;
        for (int cx = MCX_MIN; cx <= MCX_MAX; ++cx) {
            const long long nblk = cols * ((g.nn1 + cx - 1) / cx);
            const long long cost = ((nblk + slots - 1) / slots) * (cx + 1);
            if (bcost < 0 || cost < bcost) {
                bcost = cost;
                best = cx;
            }
        }
    }
    op->mrhs_cx = best;
    return best;
}

// tiles of the matrix-core 16-RHS apply of a mesh with cx node planes per block
long long mrhs_tiles(const Grid& g, int cx, int* nby, int* nbz) {
    const int n = g.n;
    *nby = (n + 1 + MTY - 1) / MTY;
    *nbz = (n + 1 + MTZ - 1) / MTZ;
    const int nch = (n + 1 + cx - 1) / cx;
    return (long long)*nby * *nbz * nch;
}

// the matrix-core 16-RHS apply of a handle's mesh on stream s (FP32 coefficients ready);
// SUMS (1, 2): per-tile column sums into part (see k_mrhs_apply)
int launch_mrhs(heat_op* op, const float* X, float* Y, bool mass, hipStream_t s, int sums = 0,
                double* part = nullptr, const double* w = nullptr) {
    int nby, nbz;
    const int cx = mrhs_chunk(op);
    const long long nt = mrhs_tiles(op->g, cx, &nby, &nbz);
    if (nt > (1LL << 30)) return fail(HEAT_E_ARG, "mesh too large for the multi-RHS tiling");
    const int ntiles = (int)nt;
    const unsigned grid = (unsigned)(((ntiles + NXCD - 1) / NXCD) * NXCD);
    const Grid& g = op->g;
    if (mass) k_mrhs_apply<true><<<grid, MNT, 0, s>>>(X, Y, op->cr32, op->ck32, g, nby, nbz, ntiles, cx);
    else if (sums == 1) k_mrhs_apply<false, 1><<<grid, MNT, 0, s>>>(X, Y, op->cr32, op->ck32, g, nby, nbz, ntiles, cx, part, w);
    else if (sums == 2) k_mrhs_apply<false, 2><<<grid, MNT, 0, s>>>(X, Y, op->cr32, op->ck32, g, nby, nbz, ntiles, cx, part, w);
    else k_mrhs_apply<false><<<grid, MNT, 0, s>>>(X, Y, op->cr32, op->ck32, g, nby, nbz, ntiles, cx);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int ensure_mixed(heat_op* op) {
    if (op->e32) return HEAT_OK;
    const Grid& g = op->g;
    hipStream_t s = op->stream;
    const size_t nb4 = sizeof(float) * (size_t)g.nn, eb4 = sizeof(float) * (size_t)g.ne;
    HIP_TRY(hipMalloc(&op->cr32, eb4));
    HIP_TRY(hipMalloc(&op->ck32, eb4));
    HIP_TRY(hipMalloc(&op->e32, nb4));
    HIP_TRY(hipMalloc(&op->r32, nb4));
    HIP_TRY(hipMalloc(&op->z32, nb4));
    HIP_TRY(hipMalloc(&op->p32, nb4));
    HIP_TRY(hipMalloc(&op->p32b, nb4));
    HIP_TRY(hipMalloc(&op->ap32, nb4));
    HIP_TRY(hipMalloc(&op->d32, nb4));
    k_coef32<<<blocks(g.ne, 256), 256, 0, s>>>(op->kcell, op->rhoc, op->cfg.dt, op->m3, op->k2, op->cr32, op->ck32, g.ne);
    CHECK_LAUNCH();
    k_convert<float, double><<<blocks(g.nn, 256), 256, 0, s>>>(op->d32, op->vec[HEAT_VEC_DIAG], g.nn);
    CHECK_LAUNCH();
    return HEAT_OK;
}

// Buffers and tile geometry of the one-launch inner iteration (made on first use).
int ensure_cg1(heat_op* op) {
    int rc = ensure_mixed(op);
    if (rc) return rc;
    if (op->c1_r) return HEAT_OK;
    const Grid& g = op->g;
    hipStream_t s = op->stream;
    const size_t nb4 = sizeof(float) * (size_t)g.nn;
    const int m = g.n - 1 > 0 ? g.n - 1 : 1;
    op->c1_nby = (m + C1TY - 1) / C1TY;
    op->c1_nbz = (m + C1TZ - 1) / C1TZ;
    op->c1_nch = cg1_chunks(g.n, op->c1_nby * op->c1_nbz, op->geo_planes[1], op->geo_wave[1]);
    op->c1_ntiles = op->c1_nby * op->c1_nbz * op->c1_nch;
    HIP_TRY(hipMalloc(&op->c1_r, nb4));
    for (int q = 0; q < 2; ++q) {
        HIP_TRY(hipMalloc(&op->c1_s[q], nb4));
        HIP_TRY(hipMalloc(&op->c1_w[q], nb4));
    }
    // room for the partials of every geometry (at most one chunk per plane pair)
    HIP_TRY(hipMalloc(&op->c1_tp, sizeof(double) * 2 * C1ROWS * (size_t)op->c1_nby * op->c1_nbz * fused_chunks_max(m)));
    HIP_TRY(hipMalloc(&op->c1_dinv, nb4));
    k_recip<<<blocks(g.nn, NT), NT, 0, s>>>(op->c1_dinv, op->d32, g.nn);
    CHECK_LAUNCH();
    // the correction is added to x at every node: zero where the walk never
    // writes it (the boundary)
    k_fill<float><<<blocks(g.nn, NT), NT, 0, s>>>(op->e32, g.nn, 0.0f);
    CHECK_LAUNCH();
    return HEAT_OK;
}

Cg1View make_cg1_view(heat_op* op, const double* b, const double* Ax) {
    Cg1View v;
    v.e = op->e32;
    v.p = op->p32;
    v.r[0] = op->r32;
    v.r[1] = op->c1_r;
    for (int q = 0; q < 2; ++q) {
        v.s[q] = op->c1_s[q];
        v.w[q] = op->c1_w[q];
        v.tp[q] = op->c1_tp + (size_t)q * C1ROWS * op->c1_ntiles;
    }
    v.d = op->c1_dinv;
    v.cr = op->cr32;
    v.ck = op->ck32;
    v.b = b;
    v.Ax = Ax;
    v.wA = op->vec[HEAT_VEC_WA];
    v.g = op->g;
    v.ntiles = op->c1_ntiles;
    v.nby = op->c1_nby;
    v.nbz = op->c1_nbz;
    v.nch = op->c1_nch;
    v.pre = op->c1_ntiles > FUSED_INBLOCK_MAX;
    v.fs = op->fst;
    v.st = op->st;
    v.nwA = op->nwA;
    v.hflag = op->hdone_dev;
    return v;
}

// one step of parity par (with its one-block fold for many tiles)
int cg1_launch_step(heat_op* op, const Cg1View& v, int par, bool first_of_pass) {
    hipStream_t s = op->stream;
    const bool detect = op->cfg.detect != 0;
    const unsigned grid = (unsigned)(((v.ntiles + NXCD - 1) / NXCD) * NXCD);
    if (v.pre) {
        if (detect) k_cg1_fin<true><<<1, NT, 0, s>>>(v, par, first_of_pass ? 1 : 0);
        else k_cg1_fin<false><<<1, NT, 0, s>>>(v, par, first_of_pass ? 1 : 0);
        CHECK_LAUNCH();
    }
    if (first_of_pass) {
        if (detect) k_cg1_step<true, true><<<grid, C1NT, 0, s>>>(v, par);
        else k_cg1_step<false, true><<<grid, C1NT, 0, s>>>(v, par);
    } else {
        if (detect) k_cg1_step<true, false><<<grid, C1NT, 0, s>>>(v, par);
        else k_cg1_step<false, false><<<grid, C1NT, 0, s>>>(v, par);
    }
    CHECK_LAUNCH();
    return HEAT_OK;
}

// Block partials of heat_dot_static: a file-scope device array, so that the
// entry point allocates nothing.  Each GPU holds its own copy until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (static_scratch_ptr).
static __device__ __align__(256) double scratch[HEAT_STATIC_DOT_MAX_BLOCKS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.
int static_scratch_ptr(double** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(HEAT_E_STATE, "device %d outside 0..%d", dev, MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HEAT_E_HIP, "hipGetSymbolAddress returned a null address");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return HEAT_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* heat_last_error(void) { return g_err.c_str(); }

int heat_setup(const heat_config* cfg, const double* k_cell, const double* rhoc_cell, heat_op** out) {
    if (!cfg || !k_cell || !rhoc_cell || !out) return fail(HEAT_E_ARG, "null argument");
    *out = nullptr;
    if (cfg->n < 1 || cfg->n > 1000) return fail(HEAT_E_ARG, "N must be in [1, 1000], got %d", cfg->n);
    if (!(cfg->dt > 0)) return fail(HEAT_E_ARG, "dt must be > 0");
    if (cfg->kernel != HEAT_KERNEL_FAST && cfg->kernel != HEAT_KERNEL_GAUSS) return fail(HEAT_E_ARG, "bad kernel");
    if (cfg->scatter != HEAT_SCATTER_ATOMIC && cfg->scatter != HEAT_SCATTER_COLORED) return fail(HEAT_E_ARG, "bad scatter");
    heat_op* op = new heat_op();
    int rc = setup_impl(cfg, k_cell, rhoc_cell, op);
    if (rc) {
        std::string keep = g_err;
        free_all(op);
        delete op;
        g_err = keep;
        return rc;
    }
    *out = op;
    return HEAT_OK;
}

int heat_cleanup(heat_op* op) {
    if (!op) return HEAT_OK;
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    free_all(op);
    delete op;
    return HEAT_OK;
}

int heat_sizes(const heat_op* op, int64_t* nn, int64_t* ne) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    if (nn) *nn = op->g.nn;
    if (ne) *ne = op->g.ne;
    return HEAT_OK;
}

int heat_field_stats(const heat_op* op, double out4[4]) {
    if (!op || !out4) return fail(HEAT_E_ARG, "null argument");
    memcpy(out4, op->stats, sizeof op->stats);
    return HEAT_OK;
}

int heat_vector(heat_op* op, int which, double** dev_ptr) {
    if (!op || !dev_ptr || which < 0 || which > HEAT_VEC_ONES) return fail(HEAT_E_ARG, "bad vector id");
    if (!op->vec[which]) return fail(HEAT_E_STATE, "vector %d needs detect", which);
    *dev_ptr = op->vec[which];
    return HEAT_OK;
}

int heat_stream(heat_op* op, void** stream) {
    if (!op || !stream) return fail(HEAT_E_ARG, "null argument");
    *stream = (void*)op->stream;
    return HEAT_OK;
}

int heat_synchronize(heat_op* op) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(op->stream));
    return HEAT_OK;
}

int heat_upload(heat_op* op, double* dev_dst, const double* host_src, int64_t count) {
    if (!op || !dev_dst || !host_src || count < 0) return fail(HEAT_E_ARG, "bad upload");
    HIP_TRY(hipMemcpyAsync(dev_dst, host_src, sizeof(double) * count, hipMemcpyHostToDevice, op->stream));
    HIP_TRY(hipStreamSynchronize(op->stream));
    return HEAT_OK;
}

int heat_download(heat_op* op, double* host_dst, const double* dev_src, int64_t count) {
    if (!op || !dev_src || !host_dst || count < 0) return fail(HEAT_E_ARG, "bad download");
    HIP_TRY(hipMemcpyAsync(host_dst, dev_src, sizeof(double) * count, hipMemcpyDeviceToHost, op->stream));
    HIP_TRY(hipStreamSynchronize(op->stream));
    return HEAT_OK;
}

int heat_set_march_chunk(heat_op* op, int planes) {
    if (!op || (planes != 2 && planes != 4 && planes != 8 && planes != 16 && planes != 32))
        return fail(HEAT_E_ARG, "march chunk must be 2, 4, 8, 16 or 32 planes");
    op->march_cx = planes;
    return HEAT_OK;
}

int heat_set_colored_impl(heat_op* op, int impl) {
    if (!op || impl < HEAT_COLORED_GATHER || impl > HEAT_COLORED_TDMA) return fail(HEAT_E_ARG, "bad impl");
    op->colored_impl = impl;
    return HEAT_OK;
}

int heat_set_atomic_impl(heat_op* op, int impl) {
    if (!op || (impl != HEAT_ATOMIC_ELEMENT && impl != HEAT_ATOMIC_TILED)) return fail(HEAT_E_ARG, "bad atomic impl");
    op->atomic_impl = impl;
    return HEAT_OK;
}

int heat_apply_operator(heat_op* op, const double* x, double* y, double dt_eff, int allow_fi, int project) {
    if (!op || !x || !y) return fail(HEAT_E_ARG, "null argument");
    const int epi = (op->cfg.detect && project) ? EPI_D1 : EPI_NONE;
    return apply_impl(op, x, y, dt_eff, allow_fi != 0, project != 0, epi, 0);
}

int heat_apply_operator_f32(heat_op* op, const float* x, float* y, double dt_eff, int project) {
    if (!op || !x || !y) return fail(HEAT_E_ARG, "null argument");
    if (!gather_path(op) || !project)
        return fail(HEAT_E_STATE, "the FP32 operator is the projected colour-ordered fast operator");
    const bool mass = (dt_eff == 0.0);
    if (!mass && dt_eff != op->cfg.dt) return fail(HEAT_E_ARG, "dt_eff must be 0 or the handle's dt");
    int rc = ensure_mixed(op);
    if (rc) return rc;
    FiArm none{0, 0, 0};
    return launch_colored<float>(op, x, y, op->cr32, op->ck32, mass, none, 0);
}

int heat_apply_matvec(heat_op* op, const double* x, double* y) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    return heat_apply_operator(op, x, y, op->cfg.dt, 1, 1);
}

int heat_apply_preconditioner(heat_op* op, const double* r, double* z) {
    if (!op || !r || !z) return fail(HEAT_E_ARG, "null argument");
    k_jacobi<<<op->nbn, NT, 0, op->stream>>>(r, op->vec[HEAT_VEC_DIAG], z, op->g.nn);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_build_rhs(heat_op* op, const double* Tn, double* b) {
    if (!op || !Tn || !b) return fail(HEAT_E_ARG, "null argument");
    int rc = heat_apply_operator(op, Tn, b, 0.0, 0, 1);
    if (rc) return rc;
    if (!op->cfg.verify && !op->cfg.no_source) {
        k_axpy<<<op->nbn, NT, 0, op->stream>>>(op->cfg.dt, op->vec[HEAT_VEC_F], b, op->g.nn);
        CHECK_LAUNCH();
    }
    return HEAT_OK;
}

int heat_dot(heat_op* op, const double* x, const double* y, double* out) {
    if (!op || !x || !y || !out) return fail(HEAT_E_ARG, "null argument");
    k_dot<<<op->nbn, NT, 0, op->stream>>>(x, y, op->g.nn, op->partials, op->st, 2);
    CHECK_LAUNCH();
    k_fin_dot<<<1, FT, 0, op->stream>>>(op->partials, (int)op->nbn, op->st, 2);
    CHECK_LAUNCH();
    int rc = read_state(op);
    if (rc) return rc;
    *out = op->hst->res[2];
    return HEAT_OK;
}

int heat_norm2(heat_op* op, const double* x, double* out) {
    double d2 = 0;
    int rc = heat_dot(op, x, x, &d2);
    if (rc) return rc;
    *out = std::sqrt(d2);
    return HEAT_OK;
}

int heat_dot_static(const double* x, const double* y, int64_t n, double* out, void* stream) {
    if (!x || !y || !out) return fail(HEAT_E_ARG, "null argument");
    if (n < 0 || n > (int64_t)NT * HEAT_STATIC_DOT_MAX_BLOCKS)
        return fail(HEAT_E_ARG, "n = %lld outside 0..%lld (the static partials hold %d blocks of %d)",
                    (long long)n, (long long)NT * HEAT_STATIC_DOT_MAX_BLOCKS, HEAT_STATIC_DOT_MAX_BLOCKS, NT);
    double* part = nullptr;
    int rc = static_scratch_ptr(&part);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = (int)blocks(n, NT);
    // every block writes its partial, so nothing of the last call's contents is read
    if (nb > 0) k_dot<<<nb, NT, 0, s>>>(x, y, (int)n, part, nullptr, 0);
    k_fin_dot_to<<<1, FT, 0, s>>>(part, nb, out);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_minmax(heat_op* op, const double* x, double* vmin, double* vmax) {
    if (!op || !x || !vmin || !vmax) return fail(HEAT_E_ARG, "null argument");
    k_minmax<<<op->nbn, NT, 0, op->stream>>>(x, op->g.nn, op->partials, op->st, 4);
    CHECK_LAUNCH();
    k_fin_minmax<<<1, FT, 0, op->stream>>>(op->partials, (int)op->nbn, op->st, 4);
    CHECK_LAUNCH();
    int rc = read_state(op);
    if (rc) return rc;
    *vmin = op->hst->res[4];
    *vmax = op->hst->res[5];
    return HEAT_OK;
}

int heat_set_dirichlet(heat_op* op, double* v, double value) {
    if (!op || !v) return fail(HEAT_E_ARG, "null argument");
    k_dirichlet<double><<<op->nbn, NT, 0, op->stream>>>(v, op->g, value);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_copy(heat_op* op, const double* x, double* y) {
    if (!op || !x || !y) return fail(HEAT_E_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(y, x, sizeof(double) * (size_t)op->g.nn, hipMemcpyDeviceToDevice, op->stream));
    return HEAT_OK;
}

int heat_axpy(heat_op* op, double a, const double* x, double* y) {
    if (!op || !x || !y) return fail(HEAT_E_ARG, "null argument");
    k_axpy<<<op->nbn, NT, 0, op->stream>>>(a, x, y, op->g.nn);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_initial_condition(heat_op* op, double* v, int mode, double t, double alpha) {
    if (!op || !v || mode < 0 || mode > 2) return fail(HEAT_E_ARG, "bad initial condition");
    k_initial<<<op->nbn, NT, 0, op->stream>>>(v, op->g, mode, t, alpha);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_pcg_solve(heat_op* op, const double* b, double* x, int max_iter, double tol, int* iters,
                   double* hist) {
    if (!op || !b || !x) return fail(HEAT_E_ARG, "null argument");
    if (max_iter < 0) return fail(HEAT_E_ARG, "max_iter must be >= 0");
    hipStream_t s = op->stream;
    if (hist && op->hist_cap < max_iter + 1) {
        if (op->dhist) HIP_TRY(hipFree(op->dhist));
        op->dhist = nullptr;
        op->hist_cap = 0;
        HIP_TRY(hipMalloc(&op->dhist, sizeof(double) * 3 * (size_t)(max_iter + 1)));
        op->hist_cap = max_iter + 1;
    }
    op->cur_hist = hist ? op->dhist : nullptr;
    const bool detect = op->cfg.detect != 0;
    double* Ap = op->vec[HEAT_VEC_AP];
    int rc = HEAT_OK;
    if (op->pcg_fi.armed && op->pcg_fi.vec > HEAT_SEU_AP)
        return fail(HEAT_E_STATE, "upsets of x64 / Ax64 are sites of the refined solve (heat_pcg_solve_mixed)");
    if (gather_path(op) && op->pcg_mode == HEAT_PCG_FUSED) {
        k_fused_params<<<1, 1, 0, s>>>(op->fst, tol, max_iter, 0, 0.0);
        CHECK_LAUNCH();
        k_fused_fault<<<1, 1, 0, s>>>(op->fst, op->pcg_fi.armed, op->pcg_fi.vec, op->pcg_fi.idx, op->pcg_fi.bit,
                                      op->pcg_fi.iter);
        CHECK_LAUNCH();
        // r = b - A x0, level R firing here if armed (src/heat_solver_het.cu:999)
        rc = apply_impl(op, x, Ap, op->cfg.dt, true, true, detect ? EPI_D1 : EPI_NONE, 0);
        if (rc) return rc;
        k_fused_init<double><<<op->nsb, NT, 0, s>>>(make_view<double>(
            op, x, op->vec[HEAT_VEC_P], op->p2, op->vec[HEAT_VEC_R], op->vec[HEAT_VEC_Z], Ap, const_cast<double*>(b),
            op->vec[HEAT_VEC_DIAG], op->cr, op->ck, nullptr));
        CHECK_LAUNCH();
        rc = run_until_done(op, max_iter, [&]() { return launch_chunk(op, x, detect); }, &op->fst->done, op->hdone);
        if (rc) return rc;
        rc = read_fused(op);
        if (rc) return rc;
        op->pcg_fi_fired = op->pcg_fi.armed ? op->hfst->fi_fired : 0;
        op->pcg_fi.armed = 0;
        if (iters) *iters = op->hfst->iter;
        if (hist) {
            const int rows = op->hfst->iter + 1;
            for (int i = 0; i < 3 * (max_iter + 1); ++i) hist[i] = std::nan("");
            HIP_TRY(hipMemcpy(hist, op->dhist, sizeof(double) * 3 * rows, hipMemcpyDeviceToHost));
        }
        return HEAT_OK;
    }
    if (op->pcg_fi.armed) return fail(HEAT_E_STATE, "PCG-state faults need the fused PCG on the colour-ordered fast path");
    k_pcg_params<<<1, 1, 0, s>>>(op->st, tol, max_iter);
    CHECK_LAUNCH();
    // r = b - A x0, with level R firing here if armed (src/heat_solver_het.cu:999)
    rc = apply_impl(op, x, Ap, op->cfg.dt, true, true, detect ? EPI_D1 : EPI_NONE, 0);
    if (rc) return rc;
    k_pcg_init<<<op->nbn, NT, 0, s>>>(b, Ap, op->vec[HEAT_VEC_DIAG], op->vec[HEAT_VEC_R],
                                       op->vec[HEAT_VEC_Z], op->vec[HEAT_VEC_P], op->g.nn,
                                       op->partials);
    CHECK_LAUNCH();
    k_fin_init<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, op->cur_hist);
    CHECK_LAUNCH();
    if (chunked_path(op)) {
        rc = run_until_done(op, max_iter, [&]() { return launch_chunk(op, x, detect); });
    } else {
        // element-centric paths: iterate with a host check per iteration
        rc = HEAT_OK;
        for (int it = 0; it <= max_iter && rc == HEAT_OK; ++it) {
            rc = read_state(op);
            if (rc || op->hst->done) break;
            rc = apply_impl(op, op->vec[HEAT_VEC_P], Ap, op->cfg.dt, false, true, detect ? EPI_PCG_D1 : EPI_PCG, 0);
            if (rc) break;
            k_pcg_update<<<op->nbn, NT, 0, s>>>(x, op->vec[HEAT_VEC_P], op->vec[HEAT_VEC_R], Ap,
                                                 op->vec[HEAT_VEC_DIAG], op->vec[HEAT_VEC_Z], op->g.nn,
                                                 op->partials, op->st);
            CHECK_LAUNCH();
            k_fin_update<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, op->cur_hist);
            CHECK_LAUNCH();
            k_pcg_dir<<<op->nbn, NT, 0, s>>>(op->vec[HEAT_VEC_P], op->vec[HEAT_VEC_Z], op->g.nn, op->st);
            CHECK_LAUNCH();
        }
    }
    if (rc) return rc;
    rc = read_state(op);
    if (rc) return rc;
    if (iters) *iters = op->hst->iter;
    if (hist) {
        const int rows = op->hst->iter + 1;
        for (int i = 0; i < 3 * (max_iter + 1); ++i) hist[i] = std::nan("");
        HIP_TRY(hipMemcpy(hist, op->dhist, sizeof(double) * 3 * rows, hipMemcpyDeviceToHost));
    }
    return HEAT_OK;
}

// iterations per polled chunk of the refined solve's inner passes (at most)
constexpr int MIXED_CHUNK = 4;
// steps of the one-launch inner iteration queued beyond the one the device
// last reported (each returns at once after the pass's end)
constexpr int CG1_AHEAD = 2;

int heat_pcg_solve_mixed(heat_op* op, const double* b, double* x, int max_outer, int max_inner,
                         double tol, double tol_inner, int* outer, int* inner_iters) {
    if (!op || !b || !x) return fail(HEAT_E_ARG, "null argument");
    if (!gather_path(op)) return fail(HEAT_E_STATE, "the FP32-refined solve needs --kernel fast --scatter colored");
    if (max_outer < 0 || max_inner < 0) return fail(HEAT_E_ARG, "iteration limits must be >= 0");
    int rc = ensure_mixed(op);
    if (rc) return rc;
    hipStream_t s = op->stream;
    const Grid& g = op->g;
    if (op->pcg_fi.armed && (op->pcg_mode == HEAT_PCG_REFERENCE_ORDER ||
                             (op->pcg_fi.vec <= HEAT_SEU_AP && op->pcg_fi.bit > 31)))
        return fail(HEAT_E_STATE, "PCG-state faults of the refined solve need the fused order and BIT in 0..31 "
                                  "(0..63 for x64 and Ax64)");
    // an upset of the FP64 outer state is injected by the host loop below;
    // the FP32 inner kernels see none
    const bool fi64 = op->pcg_fi.armed && op->pcg_fi.vec > HEAT_SEU_AP;
    const int fi_armed_inner = (op->pcg_fi.armed && !fi64) ? 1 : 0;
    int fired64 = 0;
    if (op->pcg_mode == HEAT_PCG_REFERENCE_ORDER) {
        k_mixed_params<<<1, 1, 0, s>>>(op->st, tol, tol_inner, max_inner);
        CHECK_LAUNCH();
        FiArm none{0, 0, 0};
        int nout = 0;
        for (;;) {
            k_mixed_residual<<<op->nbn, NT, 0, s>>>(x, b, op->cr, op->ck, g, op->e32, op->r32, op->z32,
                                                     op->p32, op->d32, op->partials, op->st);
            CHECK_LAUNCH();
            k_fin_mixed_resid<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st);
            CHECK_LAUNCH();
            auto launch = [&]() -> int {
                for (int i = 0; i < op->pcg_chunk; ++i) {
                    int rci = launch_colored<float>(op, op->p32, op->ap32, op->cr32, op->ck32, false, none, 1);
                    if (rci) return rci;
                    k_apply_post<float, EPI_PCG><<<op->nbn, NT, 0, s>>>(op->p32, op->ap32, g, 0, nullptr, 0.0,
                                                                       op->partials, op->st, 1);
                    CHECK_LAUNCH();
                    k_fin_apply<EPI_PCG><<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, 0.0, 1, 1);
                    CHECK_LAUNCH();
                    k_mixed_update<<<op->nbn, NT, 0, s>>>(op->e32, op->p32, op->r32, op->ap32, op->d32,
                                                           op->z32, g.nn, op->partials, op->st);
                    CHECK_LAUNCH();
                    k_fin_mixed_update<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st);
                    CHECK_LAUNCH();
                    k_mixed_dir<<<op->nbn, NT, 0, s>>>(op->p32, op->z32, g.nn, op->st);
                    CHECK_LAUNCH();
                }
                return HEAT_OK;
            };
            rc = run_until_done(op, max_inner, launch);
            if (rc) return rc;
            rc = read_state(op);
            if (rc) return rc;
            if (op->hst->outer_done || nout >= max_outer) break;
            k_mixed_correct<<<op->nbn, NT, 0, s>>>(x, op->e32, g.nn);
            CHECK_LAUNCH();
            ++nout;
        }
        op->mixed_status = op->hst->outer_done == 1 || op->hst->outer_done == 2 ? op->hst->outer_done : 0;
        if (outer) *outer = nout;
        if (inner_iters) *inner_iters = op->hst->inner_total;
        return HEAT_OK;
    }
    k_fused_params<<<1, 1, 0, s>>>(op->fst, tol_inner, max_inner, 1, tol);
    CHECK_LAUNCH();
    k_fused_fault<<<1, 1, 0, s>>>(op->fst, fi_armed_inner, op->pcg_fi.vec, op->pcg_fi.idx, op->pcg_fi.bit,
                                  op->pcg_fi.iter);
    CHECK_LAUNCH();
    const bool detect = op->cfg.detect != 0;
    // the one-launch walk addresses every array with 32-bit byte offsets:
    // meshes up to N = 644 (nn * 8 bytes below 2^31), and its status word
    // counts steps below 2^20; larger meshes or limits run the standard
    // iteration
    const bool cg1_fits = (long long)g.nn * 8 < (1LL << 31) && max_inner < CG1_MAX_STEP - 1;
    if (op->inner_mode == HEAT_INNER_ONE_LAUNCH && !op->pcg_fi.armed && cg1_fits) {
        // the inner solve in one launch per iteration (heat_cg1.h); an armed
        // upset of the inner state runs the standard two-kernel iteration below
        rc = ensure_cg1(op);
        if (rc) return rc;
        double* Ax = op->vec[HEAT_VEC_AP];
        const Cg1View v = make_cg1_view(op, b, Ax);
        const unsigned grid = (unsigned)(((v.ntiles + NXCD - 1) / NXCD) * NXCD);
        // The passes follow each other on the device, one plain launch per
        // step (a graph launch would add a gap of its own per chunk): the
        // head of every step stores cg1_flag(pass, step, status) in a pinned
        // word, and the host keeps CG1_AHEAD steps queued beyond the step
        // the device last reported.  When a step reports the pass's end the
        // host queues the correction and the next pass behind the steps
        // still in flight, which return at once; no event, no read-back.
        int nout = 0, pass = 1, launched = 0, status = 0;
        auto start_pass = [&]() -> int {
            int r2 = apply_impl(op, x, Ax, op->cfg.dt, false, true, EPI_NONE, 0);
            if (r2) return r2;
            if (detect) k_cg1_start<true><<<grid, C1NT, 0, s>>>(v, pass);
            else k_cg1_start<false><<<grid, C1NT, 0, s>>>(v, pass);
            CHECK_LAUNCH();
            launched = 0;
            return HEAT_OK;
        };
        volatile int* hflag = op->hdone;
        // pass 0 is no pass of this solve: the fused solves' done flag, the
        // word's other writer, stores pass 0, and every earlier one-launch
        // solve ended with a read-back, so none of its steps is in flight
        *hflag = 0;
        rc = start_pass();
        for (long spins = 0; rc == HEAT_OK; ++spins) {
            const int f = *hflag;
            const bool mine = cg1_flag_pass(f) == cg1_flag_pass(cg1_flag(pass, 0, 0));
            if (mine && cg1_flag_status(f) != 0) {
                status = cg1_flag_status(f);
                if (status != 1 || nout >= max_outer) break;
                // the inner solve of this pass ended and the outer test was not met: correct, next pass
                k_mixed_correct<<<op->nbn, NT, 0, s>>>(x, op->e32, g.nn);
                if (hipGetLastError() != hipSuccess) { rc = fail(HEAT_E_HIP, "kernel launch failed"); break; }
                ++nout;
                ++pass;
                rc = start_pass();
                continue;
            }
            const int running = mine ? cg1_flag_step(f) : -1;
            if (launched - running <= CG1_AHEAD && launched <= max_inner + 1) {
                // the pass's first step folds the start's eight partial rows
                rc = cg1_launch_step(op, v, launched & 1, launched == 0);
                ++launched;
                continue;
            }
            if ((spins & 1023) == 0) {
                // a fault ends the wait; so does a stream that ran dry without the pass's end
                const hipError_t q = hipStreamQuery(s);
                if (q == hipSuccess && *hflag == f) {
                    rc = fail(HEAT_E_STATE, "inner solve did not report completion");
                    break;
                }
                if (q != hipSuccess && q != hipErrorNotReady) {
                    rc = fail(HEAT_E_HIP, "stream query failed: %s", hipGetErrorString(q));
                    break;
                }
            }
            __builtin_ia32_pause();
        }
        if (rc == HEAT_OK) rc = read_fused(op);
        if (rc == HEAT_OK && status == 3) rc = fail(HEAT_E_STATE, "inner iteration launched out of step");
        if (rc) return rc;
        op->mixed_status = op->hfst->outer_done == 1 || op->hfst->outer_done == 2 ? op->hfst->outer_done : 0;
        op->pcg_fi_fired = 0;
        if (outer) *outer = nout;
        if (inner_iters) *inner_iters = op->hfst->inner_total;
        return HEAT_OK;
    }
    int nout = 0;
    // Inner passes poll in chunks of at most MIXED_CHUNK iterations: every
    // pass ends with the rest of its last chunk and one more chunk of
    // iterations that return at once, and a refined solve has several passes.
    const int saved_chunk = op->pcg_chunk;
    op->pcg_chunk = std::min(saved_chunk, MIXED_CHUNK);
    const void* key6[6] = {op->e32, (const void*)(intptr_t)((detect ? 9 : 7) + 16 * op->pcg_fi.armed),
                           (const void*)(intptr_t)op->pcg_chunk,
                           nullptr, (const void*)(intptr_t)(100 + op->colored_impl), nullptr};
    auto inner_iteration = [&]() {
        const FusedView<float> v = make_view<float>(op, op->e32, op->p32, op->p32b, op->r32, op->z32, op->ap32,
                                                    nullptr, op->d32, op->cr32, op->ck32, nullptr);
        return launch_fused_iteration<float>(op, v, detect);
    };
    for (;;) {
        double* Ax = op->vec[HEAT_VEC_AP];
        rc = apply_impl(op, x, Ax, op->cfg.dt, false, true, EPI_NONE, 0);
        if (rc) break;
        if (fi64 && op->pcg_fi.vec == HEAT_SEU_AX64 && op->pcg_fi.iter == nout) {
            k_flip<<<1, 1, 0, s>>>(Ax, op->pcg_fi.idx, op->pcg_fi.bit);   // the pass's stored A x
            fired64 = 1;
        }
        k_fused_mixed_init<<<op->nsb, NT, 0, s>>>(b, Ax, op->d32, op->e32, op->r32, op->z32, g.nn, op->fsp,
                                                   op->nsb, op->fst);
        if (hipGetLastError() != hipSuccess) { rc = fail(HEAT_E_HIP, "kernel launch failed"); break; }
        // After a correction the first inner iteration goes alone and is read
        // back: its first step makes the outer test, so a solve that the
        // correction finished stops here instead of polling chunks.
        bool finished = false;
        if (nout > 0) {
            rc = inner_iteration();
            if (rc == HEAT_OK) rc = read_fused(op);
            if (rc) break;
            finished = op->hfst->done != 0;
        }
        if (!finished) {
            rc = run_until_done(op, max_inner, [&]() { return launch_chunk_generic(op, key6, inner_iteration); },
                                &op->fst->done, op->hdone);
            if (rc == HEAT_OK) rc = read_fused(op);
            if (rc) break;
        }
        if (op->hfst->outer_done || nout >= max_outer) break;
        k_mixed_correct<<<op->nbn, NT, 0, s>>>(x, op->e32, g.nn);
        if (hipGetLastError() != hipSuccess) { rc = fail(HEAT_E_HIP, "kernel launch failed"); break; }
        if (fi64 && op->pcg_fi.vec == HEAT_SEU_X64 && op->pcg_fi.iter == nout) {
            k_flip<<<1, 1, 0, s>>>(x, op->pcg_fi.idx, op->pcg_fi.bit);    // x as this correction stored it
            fired64 = 1;
        }
        ++nout;
    }
    op->pcg_chunk = saved_chunk;
    if (rc) return rc;
    op->mixed_status = op->hfst->outer_done == 1 || op->hfst->outer_done == 2 ? op->hfst->outer_done : 0;
    op->pcg_fi_fired = fi64 ? fired64 : (op->pcg_fi.armed ? op->hfst->fi_fired : 0);
    op->pcg_fi.armed = 0;
    if (outer) *outer = nout;
    if (inner_iters) *inner_iters = op->hfst->inner_total;
    return HEAT_OK;
}

int heat_arm_pcg_fault(heat_op* op, int vec, int64_t node, int bit, int iteration) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    if (vec < HEAT_SEU_U || vec > HEAT_SEU_AX64) return fail(HEAT_E_ARG, "VEC must be one of HEAT_SEU_*");
    if (node < 0 || node >= op->g.nn) return fail(HEAT_E_ARG, "node %lld out of range", (long long)node);
    if (bit < 0 || bit > 63) return fail(HEAT_E_ARG, "BIT must be in 0..63");
    if (iteration < 0) return fail(HEAT_E_ARG, "iteration must be >= 0");
    op->pcg_fi.armed = 1;
    op->pcg_fi.vec = vec;
    op->pcg_fi.idx = node;
    op->pcg_fi.bit = bit;
    op->pcg_fi.iter = iteration;
    return HEAT_OK;
}

int heat_pcg_fault_fired(heat_op* op, int* fired) {
    if (!op || !fired) return fail(HEAT_E_ARG, "null argument");
    *fired = op->pcg_fi_fired;
    return HEAT_OK;
}

int heat_pcg_mixed_status(heat_op* op, int* status) {
    if (!op || !status) return fail(HEAT_E_ARG, "null argument");
    *status = op->mixed_status;
    return HEAT_OK;
}

int heat_apply_mrhs(heat_op* op, const float* X, float* Y, double dt_eff) {
    if (!op || !X || !Y) return fail(HEAT_E_ARG, "null argument");
    if (!gather_path(op)) return fail(HEAT_E_STATE, "the multi-RHS apply needs --kernel fast --scatter colored");
    const bool mass = (dt_eff == 0.0);
    if (!mass && dt_eff != op->cfg.dt)
        return fail(HEAT_E_ARG, "dt_eff must be 0 or the handle's dt (%g), got %g", op->cfg.dt, dt_eff);
    int rc = ensure_mixed(op);
    if (rc) return rc;
    return launch_mrhs(op, X, Y, mass, op->stream);
}

int heat_time_mrhs(heat_op* op, const float* X, float* Y, int reps, double* avg_ms) {
    if (!op || !X || !Y || reps < 1 || !avg_ms) return fail(HEAT_E_ARG, "bad timing request");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = heat_apply_mrhs(op, X, Y, op->cfg.dt);          // warm-up
    if (rc) return rc;
    HIP_TRY(hipEventRecord(e0, op->stream));
    for (int i = 0; i < reps && rc == HEAT_OK; ++i) rc = heat_apply_mrhs(op, X, Y, op->cfg.dt);
    HIP_TRY(hipEventRecord(e1, op->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return rc;
}

int heat_set_inner_solver(heat_op* op, int mode) {
    if (!op || (mode != HEAT_INNER_ONE_LAUNCH && mode != HEAT_INNER_STANDARD))
        return fail(HEAT_E_ARG, "inner solver must be HEAT_INNER_ONE_LAUNCH or HEAT_INNER_STANDARD");
    op->inner_mode = mode;
    return HEAT_OK;
}

int heat_set_pcg_mode(heat_op* op, int mode) {
    if (!op || (mode != HEAT_PCG_FUSED && mode != HEAT_PCG_REFERENCE_ORDER)) return fail(HEAT_E_ARG, "bad PCG mode");
    op->pcg_mode = mode;
    return HEAT_OK;
}

int heat_arm_fault(heat_op* op, int level, int64_t target, int bit) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    if (bit < 0 || bit > 63) return fail(HEAT_E_ARG, "BIT must be in 0..63");
    op->fi_level = level;
    op->fi.armed = 1;
    op->fi.bit = bit;
    op->fi.target = target;
    return HEAT_OK;
}

int heat_disarm_fault(heat_op* op) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    op->fi.armed = 0;
    op->fi_level = 0;
    return HEAT_OK;
}

int heat_flip_bit(heat_op* op, double* v, int64_t idx, int bit) {
    if (!op || !v) return fail(HEAT_E_ARG, "null argument");
    if (bit < 0 || bit > 63) return fail(HEAT_E_ARG, "BIT must be in 0..63");
    if (idx < 0 || idx >= op->g.nn) return fail(HEAT_E_ARG, "index %lld outside [0, %d)", (long long)idx, op->g.nn);
    k_flip<<<1, 1, 0, op->stream>>>(v, idx, bit);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_abft_take(heat_op* op, double* step_max, int64_t* calls) {
    if (!op) return fail(HEAT_E_ARG, "null handle");
    int rc = read_state(op);
    if (rc) return rc;
    if (step_max) *step_max = op->hst->abft_max;
    if (calls) *calls = op->hst->abft_calls;
    k_abft_reset<<<1, 1, 0, op->stream>>>(op->st);
    CHECK_LAUNCH();
    return HEAT_OK;
}

int heat_abft_threshold(heat_op* op, double* chk_tol) {
    if (!op || !chk_tol) return fail(HEAT_E_ARG, "null argument");
    *chk_tol = op->chk_tol;
    return HEAT_OK;
}

// The apply's algorithmic traffic as a plain stream: x, cr, ck read, y
// written, in 16-byte units (pairs of doubles; the odd last node by the first
// thread).  Each thread moves two units of every stream, all loads issued
// before the stores, with non-temporal loads and stores; of the forms timed
// on MI355X (grid-stride loops, one to eight units per thread, with and
// without the non-temporal hint) this one is the fastest once the data no
// longer fits the Infinity Cache and within 1 % of the fastest below that.
typedef double stream_v2 __attribute__((ext_vector_type(2)));
constexpr int STREAM_U = 2;
__global__ __launch_bounds__(256) void k_stream_ceiling(const double* __restrict__ x, const double* __restrict__ cr,
                                                        const double* __restrict__ ck, double* __restrict__ y,
                                                        long long nn, long long ne) {
    const stream_v2* __restrict__ x2 = reinterpret_cast<const stream_v2*>(x);
    const stream_v2* __restrict__ r2 = reinterpret_cast<const stream_v2*>(cr);
    const stream_v2* __restrict__ k2 = reinterpret_cast<const stream_v2*>(ck);
    stream_v2* __restrict__ y2 = reinterpret_cast<stream_v2*>(y);
    const long long nn2 = nn / 2, ne2 = ne / 2;
    const long long base = (long long)blockIdx.x * (256 * STREAM_U) + threadIdx.x;
    stream_v2 a[STREAM_U], r[STREAM_U], k[STREAM_U];
#pragma unroll
    for (int u = 0; u < STREAM_U; ++u) {
        const long long i = base + u * 256;
        a[u] = stream_v2{0.0, 0.0};
        r[u] = a[u];
        k[u] = a[u];
        if (i < nn2) a[u] = __builtin_nontemporal_load(x2 + i);
        if (i < ne2) {
            r[u] = __builtin_nontemporal_load(r2 + i);
            k[u] = __builtin_nontemporal_load(k2 + i);
        }
    }
#pragma unroll
    for (int u = 0; u < STREAM_U; ++u) {
        const long long i = base + u * 256;
        if (i < nn2) __builtin_nontemporal_store(a[u] + 0.5 * r[u] + 0.25 * k[u], y2 + i);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (nn & 1)) y[nn - 1] = x[nn - 1];
}

int heat_time_stream(heat_op* op, const double* x, double* y, int reps, double* avg_ms) {
    if (!op || !x || !y || reps < 1 || !avg_ms) return fail(HEAT_E_ARG, "bad timing request");
    if (((uintptr_t)x | (uintptr_t)y) & 15) return fail(HEAT_E_ARG, "x and y must be 16-byte aligned");
    hipStream_t s = op->stream;
    const long long nn = op->g.nn, ne = op->g.ne;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < reps; ++i) {
        k_stream_ceiling<<<blocks(nn / 2, 256 * STREAM_U), 256, 0, s>>>(x, op->cr, op->ck, y, nn, ne);
    }
    (void)hipEventRecord(e1, s);
    int rc = HEAT_OK;
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(HEAT_E_HIP, "event sync failed");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == HEAT_OK && hipGetLastError() != hipSuccess) rc = fail(HEAT_E_HIP, "stream kernel launch failed");
    return rc;
}

int heat_time_matvec(heat_op* op, const double* x, double* y, int reps, int64_t flush_bytes,
                     double* avg_ms, double* min_ms) {
    if (!op || !x || !y || reps < 1 || flush_bytes < 0) return fail(HEAT_E_ARG, "bad timing request");
    hipStream_t s = op->stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    FiArm none = op->fi;
    const int saved_level = op->fi_level;
    op->fi_level = 0;   // never fire a pending fault from a timing loop
    (void)none;
    double total = 0.0, best = 1e300;
    int rc = HEAT_OK;
    unsigned* flush = nullptr;
    const long long nflush = flush_bytes / 4;
    if (nflush > 0 && hipMalloc(&flush, sizeof(unsigned) * nflush) != hipSuccess) {
        op->fi_level = saved_level;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return fail(HEAT_E_HIP, "cannot allocate %lld flush bytes", (long long)flush_bytes);
    }
    if (nflush > 0) {
        for (int i = 0; i < reps && rc == HEAT_OK; ++i) {
            k_fill<unsigned><<<blocks(nflush, 256), 256, 0, s>>>(flush, (int)std::min<long long>(nflush, INT32_MAX), (unsigned)i);
            (void)hipEventRecord(e0, s);
            rc = apply_impl(op, x, y, op->cfg.dt, false, true, EPI_NONE, 0);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) rc = fail(HEAT_E_HIP, "event sync failed");
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            total += ms;
            best = std::min(best, (double)ms);
        }
    } else {
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < reps && rc == HEAT_OK; ++i) rc = apply_impl(op, x, y, op->cfg.dt, false, true, EPI_NONE, 0);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess && rc == HEAT_OK) rc = fail(HEAT_E_HIP, "event sync failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        total = ms;
        best = ms / reps;
    }
    op->fi_level = saved_level;
    if (flush) (void)hipFree(flush);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (avg_ms) *avg_ms = total / reps;
    if (min_ms) *min_ms = best;
    return HEAT_OK;
}

int heat_autotune(heat_op* op, int reps, int* impl, int* planes, int* shape, double* best_ms) {
    if (!op || reps < 1) return fail(HEAT_E_ARG, "bad autotune request");
    if (!gather_path(op)) {
        if (impl) *impl = op->colored_impl;
        if (planes) *planes = op->march_cx;
        if (shape) *shape = op->tile_shape;
        if (best_ms) *best_ms = 0.0;
        return HEAT_OK;   // only the colour-ordered fast apply has schedules to choose from
    }
    struct Cand { int impl, cx, shape; };
    const Cand cands[] = {{HEAT_COLORED_GATHER, 8, 0}, {HEAT_COLORED_MARCH, 4, 0}, {HEAT_COLORED_MARCH, 8, 0},
                          {HEAT_COLORED_TILE, 4, 0},   {HEAT_COLORED_TILE, 8, 0},  {HEAT_COLORED_TILE, 16, 0},
                          {HEAT_COLORED_TILE, 4, 1},   {HEAT_COLORED_TILE, 8, 1},  {HEAT_COLORED_TILE, 16, 1},
                          {HEAT_COLORED_TILE, 4, 2},   {HEAT_COLORED_TILE, 8, 2},  {HEAT_COLORED_TILE, 16, 2},
                          {HEAT_COLORED_TDMA, 4, 0},   {HEAT_COLORED_TDMA, 8, 0},  {HEAT_COLORED_TDMA, 16, 0},
                          {HEAT_COLORED_TDMA, 4, 1},   {HEAT_COLORED_TDMA, 8, 1},  {HEAT_COLORED_TDMA, 16, 1},
                          {HEAT_COLORED_TDMA, 4, 2},   {HEAT_COLORED_TDMA, 8, 2},  {HEAT_COLORED_TDMA, 16, 2}};
    double* x = op->vec[HEAT_VEC_P];
    double* y = op->vec[HEAT_VEC_AP];
    double best = 1e300;
    Cand pick = {op->colored_impl, op->march_cx, op->tile_shape};
    for (const Cand& c : cands) {
        op->colored_impl = c.impl;
        op->march_cx = c.cx;
        op->tile_shape = c.shape;
        double avg = 0, mn = 0;
        int rc = heat_time_matvec(op, x, y, 3, 0, &avg, &mn);
        if (rc == HEAT_OK) rc = heat_time_matvec(op, x, y, reps, 0, &avg, &mn);
        if (rc) return rc;
        if (avg < best) { best = avg; pick = c; }
    }
    op->colored_impl = pick.impl;
    op->march_cx = pick.cx;
    op->tile_shape = pick.shape;
    if (impl) *impl = pick.impl;
    if (planes) *planes = pick.cx;
    if (shape) *shape = pick.shape;
    if (best_ms) *best_ms = best;
    return HEAT_OK;
}

int heat_set_tile_shape(heat_op* op, int shape) {
    if (!op || shape < 0 || shape > 2) return fail(HEAT_E_ARG, "tile shape must be 0, 1 or 2");
    op->tile_shape = shape;
    return HEAT_OK;
}

}  // extern "C"

namespace {

// Set geometry `which` (0 the fused iteration, 1 the one-launch inner
// iteration) of a handle; a captured chunk holds the old views and is dropped.
int set_geometry(heat_op* op, int which, int planes, int wave) {
    if (planes == 0) { planes = FCX; wave = FWAVE; }
    if (planes < 2 || planes > (1 << 20) || wave < 0 || wave > (1 << 20))
        return fail(HEAT_E_ARG, "geometry: planes must be in 2..2^20 (0: the default) and wave in 0..2^20");
    op->geo_planes[which] = planes;
    op->geo_wave[which] = wave;
    if (which == 0) {
        op->fnch = fused_chunks(op->g.nn1, op->fnby * op->fnbz, planes, wave);
        op->fntiles = op->fnby * op->fnbz * op->fnch;
    } else if (op->c1_r) {
        op->c1_nch = cg1_chunks(op->g.n, op->c1_nby * op->c1_nbz, planes, wave);
        op->c1_ntiles = op->c1_nby * op->c1_nbz * op->c1_nch;
    }
    if (op->gexec) { (void)hipGraphExecDestroy(op->gexec); op->gexec = nullptr; }
    if (op->graph) { (void)hipGraphDestroy(op->graph); op->graph = nullptr; }
    return HEAT_OK;
}

}  // namespace

extern "C" {

int heat_set_pcg_geometry(heat_op* op, int which, int planes, int wave) {
    if (!op || which < 0 || which > 1) return fail(HEAT_E_ARG, "geometry 0 (fused) or 1 (one-launch inner)");
    if (op->batch_refs > 0)
        return fail(HEAT_E_STATE, "the handle's geometry is held by a batch (heat_batch_set_geometry sets it)");
    return set_geometry(op, which, planes, wave);
}

int heat_pcg_geometry(const heat_op* op, int which, int* planes, int* wave, int* chunks, int* tiles) {
    if (!op || which < 0 || which > 1) return fail(HEAT_E_ARG, "geometry 0 (fused) or 1 (one-launch inner)");
    if (planes) *planes = op->geo_planes[which];
    if (wave) *wave = op->geo_wave[which];
    int nch, nt;
    if (which == 0) {
        nch = op->fnch;
        nt = op->fntiles;
    } else {
        const int m = op->g.n - 1 > 0 ? op->g.n - 1 : 1;
        const int ncol = ((m + C1TY - 1) / C1TY) * ((m + C1TZ - 1) / C1TZ);
        nch = cg1_chunks(op->g.n, ncol, op->geo_planes[1], op->geo_wave[1]);
        nt = ncol * nch;
    }
    if (chunks) *chunks = nch;
    if (tiles) *tiles = nt;
    return HEAT_OK;
}

// Time `iters` PCG iterations for every candidate geometry and keep the
// fastest: which = 0 the fused FP64 iteration (heat_pcg_solve), 1 the
// one-launch FP32 inner iteration (one pass of heat_pcg_solve_mixed).  The
// solves run on scratch vectors (right-hand side 1 at interior nodes, x from
// 0, tolerance 0); the handle's work vectors r, z, p, Ap are overwritten, its
// D1 record is restored.  Candidates giving the same chunk count are timed once.
int heat_autotune_pcg(heat_op* op, int which, int iters, int* planes, int* wave, double* best_ms) {
    if (!op || which < 0 || which > 1 || iters < 1) return fail(HEAT_E_ARG, "bad PCG autotune request");
    if (op->batch_refs > 0)
        return fail(HEAT_E_STATE, "the handle's geometry is held by a batch (heat_batch_autotune tunes it)");
    if (!gather_path(op)) return fail(HEAT_E_STATE, "PCG geometry applies to the colour-ordered fast path only");
    if (op->pcg_fi.armed || op->fi.armed) return fail(HEAT_E_STATE, "a fault is armed: tune before arming it");
    if (which == 1) {
        int rc = ensure_cg1(op);
        if (rc) return rc;
    }
    const Grid& g = op->g;
    hipStream_t s = op->stream;
    double *bv = nullptr, *xv = nullptr;
    HIP_TRY(hipMalloc(&bv, sizeof(double) * (size_t)g.nn));
    if (hipMalloc(&xv, sizeof(double) * (size_t)g.nn) != hipSuccess) {
        (void)hipFree(bv);
        return fail(HEAT_E_HIP, "autotune: out of device memory");
    }
    k_ones_interior<<<op->nbn, NT, 0, s>>>(bv, g);
    int rc = read_state(op);                       // the D1 record, restored at the end
    DevState saved = *op->hst;
    const int keep_mode = op->pcg_mode, keep_inner = op->inner_mode;
    const int keep_planes = op->geo_planes[which], keep_wave = op->geo_wave[which];
    op->pcg_mode = HEAT_PCG_FUSED;
    op->inner_mode = HEAT_INNER_ONE_LAUNCH;
    const int cand_planes[] = {4, 8, 12, 16, 24, 32, 64};
    // whole waves of 3 resident blocks per CU (the fused kernels) or, for the
    // one-launch step (its LDS ring and registers allow 4), of 4
    const int cand_wave[] = {0, FWAVE, which == 1 ? C1WAVE4 : FWAVE};
    int pick_p = keep_planes, pick_w = keep_wave;
    double best = 1e300;
    std::vector<int> seen;
    for (int cp : cand_planes) {
        for (int cw : cand_wave) {
            if (rc) break;
            const int nch = which == 0 ? fused_chunks(g.nn1, op->fnby * op->fnbz, cp, cw)
                                       : cg1_chunks(g.n, op->c1_nby * op->c1_nbz, cp, cw);
            if (std::find(seen.begin(), seen.end(), nch) != seen.end()) continue;
            seen.push_back(nch);
            rc = set_geometry(op, which, cp, cw);
            double t_best = 1e300;
            for (int rep = 0; rep < 3 && rc == HEAT_OK; ++rep) {
                k_fill<double><<<op->nbn, NT, 0, s>>>(xv, g.nn, 0.0);
                if (hipStreamSynchronize(s) != hipSuccess) { rc = fail(HEAT_E_HIP, "autotune: stream error"); break; }
                const auto t0 = std::chrono::steady_clock::now();
                int it = 0, outer = 0;
                rc = which == 0 ? heat_pcg_solve(op, bv, xv, iters, 0.0, &it, nullptr)
                                : heat_pcg_solve_mixed(op, bv, xv, 1, iters, 0.0, 0.0, &outer, &it);
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (ms < t_best) t_best = ms;
            }
            if (rc == HEAT_OK && t_best < best) { best = t_best; pick_p = cp; pick_w = cw; }
        }
    }
    op->pcg_mode = keep_mode;
    op->inner_mode = keep_inner;
    const int rc2 = set_geometry(op, which, rc ? keep_planes : pick_p, rc ? keep_wave : pick_w);
    (void)hipStreamSynchronize(s);
    (void)hipFree(bv);
    (void)hipFree(xv);
    *op->hst = saved;
    if (rc == HEAT_OK) rc = rc2;
    if (rc == HEAT_OK && hipMemcpy(op->st, op->hst, sizeof(DevState), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(HEAT_E_HIP, "autotune: could not restore the D1 record");
    if (rc) return rc;
    if (planes) *planes = pick_p;
    if (wave) *wave = pick_w;
    if (best_ms) *best_ms = best / iters;
    return HEAT_OK;
}

int heat_set_tuning(heat_op* op, int pcg_chunk, int use_graph) {
    if (!op || pcg_chunk < 1 || pcg_chunk > 256) return fail(HEAT_E_ARG, "pcg_chunk must be in [1, 256]");
    op->pcg_chunk = pcg_chunk;
    op->use_graph = use_graph ? 1 : 0;
    return HEAT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched ensembles: many independent solves advanced together.
// ---------------------------------------------------------------------------
namespace {

__global__ void k_batch_iters(const FusedView<double>* __restrict__ views, int count, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = views[i].active ? views[i].fs->iter : -1;
}

// inner iterations of each refined solve of the step (-1: masked)
__global__ void k_batch_inner(const FusedView<double>* __restrict__ views, int count, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = views[i].active ? views[i].fs->inner_total : -1;
}

}  // namespace

struct heat_batch {
    std::vector<heat_op*> ops;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int* hflag = nullptr;                    // pinned mailboxes, one per chunk in flight
    int* dflag = nullptr;                    // device: all solves done
    int* hdone = nullptr;                    // pinned host word: all solves done after the last chunk
    int* hdone_dev = nullptr;                // its device address
    int* diters = nullptr;
    int* hiters = nullptr;                   // pinned
    std::vector<FusedView<double>> views;
    FusedView<double>* dviews = nullptr;
    std::vector<FusedView<float>> views32;   // refined steps: the FP32 inner solves (made on first use)
    FusedView<float>* dviews32 = nullptr;
    std::vector<int> tile_off, sb_off, node_off;
    int *d_tile_off = nullptr, *d_sb_off = nullptr, *d_node_off = nullptr;
    double *sp = nullptr, *tp = nullptr;
    int pcg_chunk = 8, use_graph = 1;
    int persistent = 0;                      // one persistent launch per solve (k_batch_persistent)
    int persist_iters = 0;                   // iterations that launch may run
    int* dpflag = nullptr;                   // device: its all-done decisions by iteration parity
    int detect = 0;                          // members built with detectors (all or none)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_chunk = 0;
    hipGraph_t graph32 = nullptr;            // the FP32 inner chunk of refined steps
    hipGraphExec_t gexec32 = nullptr;
    int graph_chunk32 = 0;
    heat_op* mesh = nullptr;                 // shared-mesh batch: the borrowed handle
    // 16-column refined solve of a shared-mesh batch (node-major FP32 columns)
    float *mrE = nullptr, *mrR = nullptr, *mrP = nullptr, *mrAP = nullptr;
    double* mrpart = nullptr;
    double* mrtile = nullptr;                // the apply's tile partials (4 x 16 x tiles)
    MrState* mrst = nullptr;
    MrState* hmrst = nullptr;                // pinned
    std::vector<void*> owned;                // shared-mesh batch: member vectors and states
};

namespace {

int batch_upload_views(heat_batch* b);

void batch_free(heat_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->gexec32) (void)hipGraphExecDestroy(b->gexec32);
    if (b->graph32) (void)hipGraphDestroy(b->graph32);
    (void)hipFree(b->dviews32);
    (void)hipFree(b->mrE); (void)hipFree(b->mrR); (void)hipFree(b->mrP); (void)hipFree(b->mrAP);
    (void)hipFree(b->mrpart); (void)hipFree(b->mrtile); (void)hipFree(b->mrst);
    if (b->hmrst) (void)hipHostFree(b->hmrst);
    for (auto& e : b->ev) if (e) (void)hipEventDestroy(e);
    if (b->hflag) (void)hipHostFree(b->hflag);
    if (b->hiters) (void)hipHostFree(b->hiters);
    (void)hipFree(b->dflag); (void)hipFree(b->diters); (void)hipFree(b->dviews); (void)hipFree(b->dpflag);
    if (b->hdone) (void)hipHostFree(b->hdone);
    (void)hipFree(b->d_tile_off); (void)hipFree(b->d_sb_off); (void)hipFree(b->d_node_off);
    (void)hipFree(b->sp); (void)hipFree(b->tp);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (heat_op* op : b->ops) heat_cleanup(op);
    for (void* p : b->owned) (void)hipFree(p);
    if (b->mesh) --b->mesh->batch_refs;
}

// device tables and partial arrays of a batch whose views are filled in
int batch_finish(heat_batch* b) {
    const int count = (int)b->views.size();
    const int total_tiles = b->tile_off[count], total_sb = b->sb_off[count];
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&b->ev[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&b->ev[1], hipEventDisableTiming));
    HIP_TRY(hipHostMalloc(&b->hflag, 2 * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&b->hiters, count * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipMalloc(&b->dflag, sizeof(int)));
    HIP_TRY(hipHostMalloc(&b->hdone, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
    *b->hdone = 0;
    HIP_TRY(hipHostGetDevicePointer((void**)&b->hdone_dev, b->hdone, 0));
    HIP_TRY(hipMalloc(&b->diters, count * sizeof(int)));
    HIP_TRY(hipMalloc(&b->dviews, count * sizeof(FusedView<double>)));
    HIP_TRY(hipMalloc(&b->d_tile_off, (count + 1) * sizeof(int)));
    HIP_TRY(hipMalloc(&b->d_sb_off, (count + 1) * sizeof(int)));
    HIP_TRY(hipMalloc(&b->d_node_off, (count + 1) * sizeof(int)));
    HIP_TRY(hipMalloc(&b->sp, 5 * sizeof(double) * (size_t)total_sb));   // 5 rows: refined steps
    HIP_TRY(hipMalloc(&b->tp, 4 * sizeof(double) * (size_t)total_tiles));
    for (int i = 0; i < count; ++i) {
        FusedView<double>& v = b->views[i];
        v.sp = b->sp + b->sb_off[i];
        v.sp_stride = total_sb;
        v.tp = b->tp + b->tile_off[i];
        v.tp_stride = total_tiles;
    }
    HIP_TRY(hipMemcpy(b->d_tile_off, b->tile_off.data(), (count + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_sb_off, b->sb_off.data(), (count + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_node_off, b->node_off.data(), (count + 1) * sizeof(int), hipMemcpyHostToDevice));
    return batch_upload_views(b);
}

int batch_upload_views(heat_batch* b) {
    HIP_TRY(hipMemcpyAsync(b->dviews, b->views.data(), sizeof(FusedView<double>) * b->views.size(),
                           hipMemcpyHostToDevice, b->stream));
    if (!b->views32.empty()) {
        for (size_t i = 0; i < b->views.size(); ++i) b->views32[i].active = b->views[i].active;
        HIP_TRY(hipMemcpyAsync(b->dviews32, b->views32.data(), sizeof(FusedView<float>) * b->views32.size(),
                               hipMemcpyHostToDevice, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return HEAT_OK;
}

// The whole batched solve as one cooperative launch of k_batch_persistent:
// as many blocks as stay resident together (a multiple of NXCD), at most
// `iters` iterations; the kernel stores the all-done flag in b->hdone.
template <typename T>
int batch_persistent_t(heat_batch* b, const std::vector<FusedView<T>>& views, const FusedView<T>* dv, int iters) {
    int count = (int)views.size();
    int pre_sp = 0, pre_tp = 0;
    for (const auto& v : views) { pre_sp |= v.pre_sp; pre_tp |= v.pre_tp; }
    const bool detect = b->detect && sizeof(T) == sizeof(double);
    const void* fn = detect ? reinterpret_cast<const void*>(&k_batch_persistent<T, true>)
                            : reinterpret_cast<const void*>(&k_batch_persistent<T, false>);
    int per_cu = 0, dev = 0, ncu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, 0));
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = (per_cu * ncu / NXCD) * NXCD;
    if (grid < NXCD) return fail(HEAT_E_HIP, "the persistent batch kernel does not fit on the device");
    if (!b->dpflag) HIP_TRY(hipMalloc(&b->dpflag, 2 * sizeof(int)));
    const FusedView<T>* a_views = dv;
    const int* a_tile = b->d_tile_off;
    const int* a_sb = b->d_sb_off;
    int* a_flag = b->dpflag;
    int* a_hdone = b->hdone_dev;
    void* args[] = {&a_views, &a_tile, &a_sb, &count, &iters, &pre_sp, &pre_tp, &a_flag, &a_hdone};
    HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(NT), args, 0, b->stream));
    return HEAT_OK;
}

// One chunk of pcg_chunk batched fused iterations over the views `dv` (FP64
// solves, or the FP32 inner solves of refined ones), then the all-done flag;
// captured once per value type as a hipGraph.  A persistent batch runs the
// whole solve in one launch instead.
template <typename T>
int batch_chunk_t(heat_batch* b, const std::vector<FusedView<T>>& views, const FusedView<T>* dv,
                  hipGraph_t& graph_, hipGraphExec_t& gexec_, int& graph_chunk_) {
    if (b->persistent) return batch_persistent_t<T>(b, views, dv, b->persist_iters);
    const int count = (int)views.size();
    hipStream_t s = b->stream;
    const int total_tiles = b->tile_off[count], total_sb = b->sb_off[count];
    const unsigned grid1 = (unsigned)(((total_tiles + NXCD - 1) / NXCD) * NXCD);
    bool pre_sp = false, pre_tp = false;
    for (const auto& v : views) { pre_sp = pre_sp || v.pre_sp; pre_tp = pre_tp || v.pre_tp; }
    const bool detect = b->detect && sizeof(T) == sizeof(double);   // refined batches run without
    auto body = [&]() -> int {
        for (int i = 0; i < b->pcg_chunk; ++i) {
            if (pre_sp) {
                k_batch_fin_sp<T><<<count, NT, 0, s>>>(dv);
                CHECK_LAUNCH();
            }
            if (detect) {
                k_batch_dir<T, true><<<grid1, FTY * FTZ, 0, s>>>(dv, b->d_tile_off, count);
                CHECK_LAUNCH();
                k_batch_tdma<T, true><<<grid1, FTY * FTZ, 0, s>>>(dv, b->d_tile_off, count);
                CHECK_LAUNCH();
                if (pre_tp) {
                    k_batch_fin_tp<T, true><<<count, NT, 0, s>>>(dv);
                    CHECK_LAUNCH();
                }
                k_batch_update<T, true><<<total_sb, NT, 0, s>>>(dv, b->d_sb_off, count);
                CHECK_LAUNCH();
            } else {
                k_batch_dir<T, false><<<grid1, FTY * FTZ, 0, s>>>(dv, b->d_tile_off, count);
                CHECK_LAUNCH();
                k_batch_tdma<T, false><<<grid1, FTY * FTZ, 0, s>>>(dv, b->d_tile_off, count);
                CHECK_LAUNCH();
                if (pre_tp) {
                    k_batch_fin_tp<T, false><<<count, NT, 0, s>>>(dv);
                    CHECK_LAUNCH();
                }
                k_batch_update<T, false><<<total_sb, NT, 0, s>>>(dv, b->d_sb_off, count);
                CHECK_LAUNCH();
            }
        }
        k_batch_all_done<T><<<1, NT, 0, s>>>(dv, count, b->hdone_dev);
        CHECK_LAUNCH();
        return HEAT_OK;
    };
    if (!b->use_graph) return body();
    if (!gexec_ || graph_chunk_ != b->pcg_chunk) {
        if (gexec_) { (void)hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
        if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = body();
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (ec != hipSuccess) return fail(HEAT_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        graph_ = graph;
        HIP_TRY(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
        graph_chunk_ = b->pcg_chunk;
    }
    HIP_TRY(hipGraphLaunch(gexec_, s));
    return HEAT_OK;
}

int batch_chunk(heat_batch* b) {
    return batch_chunk_t<double>(b, b->views, b->dviews, b->graph, b->gexec, b->graph_chunk);
}

// chunks until every solve of the batch reports done; the flag of chunk k is
// read while chunk k+1 runs.  Each chunk ends with k_batch_all_done storing
// the flag into a pinned host word; a done batch stays done, so a newer value
// from the chunk now running is as good.
template <typename Chunk>
int batch_until_done(heat_batch* b, int max_iter, Chunk chunk) {
    hipStream_t s = b->stream;
    const volatile int* hdone = b->hdone;
    if (b->persistent) {
        // one launch: every solve stops by iteration max_iter (+1 for its last test)
        b->persist_iters = max_iter + 2;
        const int rc = chunk();
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(s));
        if (!*hdone) return fail(HEAT_E_STATE, "batch PCG did not report completion");
        return HEAT_OK;
    }
    const int max_chunks = max_iter / b->pcg_chunk + 3;
    int rc = chunk();
    if (rc) return rc;
    HIP_TRY(hipEventRecord(b->ev[0], s));
    for (int kk = 1;; ++kk) {
        if (kk > max_chunks) return fail(HEAT_E_STATE, "batch PCG did not report completion");
        rc = chunk();
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b->ev[kk & 1], s));
        HIP_TRY(hipEventSynchronize(b->ev[(kk - 1) & 1]));
        if (*hdone) break;
    }
    return HEAT_OK;
}

// the FP32 views of a batch of handles (refined steps), made on first use
int batch_views32(heat_batch* b) {
    if (!b->views32.empty()) return HEAT_OK;
    if (b->mesh) return fail(HEAT_E_STATE, "refined batch steps need a batch of handles (heat_batch_setup)");
    const int count = (int)b->views.size();
    for (int i = 0; i < count; ++i) {
        heat_op* op = b->ops[i];
        int rc = ensure_mixed(op);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(op->stream));
        FusedView<float> v = make_view<float>(op, op->e32, op->p32, op->p32b, op->r32, op->z32, op->ap32, nullptr,
                                              op->d32, op->cr32, op->ck32, nullptr);
        v.hflag = nullptr;                       // a batch polls its own all-done flag
        v.sp = b->views[i].sp;
        v.sp_stride = b->views[i].sp_stride;
        v.tp = b->views[i].tp;
        v.tp_stride = b->views[i].tp_stride;
        v.active = b->views[i].active;
        b->views32.push_back(v);
    }
    HIP_TRY(hipMalloc(&b->dviews32, count * sizeof(FusedView<float>)));
    HIP_TRY(hipMemcpy(b->dviews32, b->views32.data(), count * sizeof(FusedView<float>), hipMemcpyHostToDevice));
    return HEAT_OK;
}

}  // namespace

extern "C" {

int heat_batch_setup(int count, const heat_config* cfgs, const double* const* k, const double* const* rhoc,
                     heat_batch** out) {
    if (!out || !cfgs || !k || !rhoc || count < 1) return fail(HEAT_E_ARG, "bad batch arguments");
    *out = nullptr;
    heat_batch* b = new heat_batch();
    auto bail = [&](int rc) { batch_free(b); delete b; return rc; };
    for (int i = 0; i < count; ++i) {
        const heat_config& c = cfgs[i];
        if (c.kernel != HEAT_KERNEL_FAST || c.scatter != HEAT_SCATTER_COLORED)
            return bail(fail(HEAT_E_ARG, "batch solve %d: the batch runs the colour-ordered fast path only", i));
        if ((c.detect != 0) != (cfgs[0].detect != 0))
            return bail(fail(HEAT_E_ARG, "batch solve %d: all solves of a batch have detectors, or none", i));
        heat_op* op = nullptr;
        const int rc = heat_setup(&c, k[i], rhoc[i], &op);
        if (rc) {
            char msg[512];
            snprintf(msg, sizeof msg, "%s", heat_last_error());
            return bail(fail(rc, "batch solve %d: %s", i, msg));
        }
        op->batch_refs = 1;                      // its geometry is the batch's (heat_batch_set_geometry)
        b->ops.push_back(op);
        HIP_TRY_OR(hipStreamSynchronize(op->stream), return bail(HEAT_E_HIP));
    }
    b->detect = cfgs[0].detect != 0;
    b->tile_off.assign(count + 1, 0);
    b->sb_off.assign(count + 1, 0);
    b->node_off.assign(count + 1, 0);
    long long nodes = 0;
    for (int i = 0; i < count; ++i) {
        b->tile_off[i + 1] = b->tile_off[i] + b->ops[i]->fntiles;
        b->sb_off[i + 1] = b->sb_off[i] + b->ops[i]->nsb;
        nodes += b->ops[i]->g.nn;
        if (nodes > (1LL << 30)) return bail(fail(HEAT_E_ARG, "batch too large: %lld nodes", nodes));
        b->node_off[i + 1] = (int)nodes;
    }
    for (int i = 0; i < count; ++i) {
        heat_op* op = b->ops[i];
        b->views.push_back(make_view<double>(op, op->vec[HEAT_VEC_T], op->vec[HEAT_VEC_P], op->p2,
                                             op->vec[HEAT_VEC_R], op->vec[HEAT_VEC_Z], op->vec[HEAT_VEC_AP],
                                             op->vec[HEAT_VEC_B], op->vec[HEAT_VEC_DIAG], op->cr, op->ck, nullptr));
        b->views.back().hflag = nullptr;         // a batch polls its own all-done flag
    }
    const int rc = batch_finish(b);
    if (rc) return bail(rc);
    *out = b;
    return HEAT_OK;
}

int heat_batch_setup_rhs(heat_op* op, int count, const double* F, heat_batch** out) {
    if (!op || !out || count < 1) return fail(HEAT_E_ARG, "bad batch arguments");
    *out = nullptr;
    if (!gather_path(op)) return fail(HEAT_E_STATE, "the batch runs the colour-ordered fast path only");
    // (a handle with detectors is accepted: heat_batch_step runs without the D1
    // check, heat_batch_step_mrhs applies it with the handle's w_A)
    heat_batch* b = new heat_batch();
    b->mesh = op;
    ++op->batch_refs;                            // its fused geometry is held by the batch's views
    auto bail = [&](int rc) { batch_free(b); delete b; return rc; };
    const size_t nn = (size_t)op->g.nn;
    if ((long long)nn * count > (1LL << 30)) return bail(fail(HEAT_E_ARG, "batch too large"));
    b->tile_off.assign(count + 1, 0);
    b->sb_off.assign(count + 1, 0);
    b->node_off.assign(count + 1, 0);
    HIP_TRY_OR(hipStreamSynchronize(op->stream), return bail(HEAT_E_HIP));
    for (int i = 0; i < count; ++i) {
        // T, b, r, z, p, p2, Ap, F of this right-hand side; the mesh arrays are shared
        double* v = nullptr;
        HIP_TRY_OR(hipMalloc(&v, 8 * nn * sizeof(double)), return bail(HEAT_E_HIP));
        b->owned.push_back(v);
        FusedState* fs = nullptr;
        HIP_TRY_OR(hipMalloc(&fs, sizeof(FusedState)), return bail(HEAT_E_HIP));
        b->owned.push_back(fs);
        HIP_TRY_OR(hipMemset(v, 0, 8 * nn * sizeof(double)), return bail(HEAT_E_HIP));
        HIP_TRY_OR(hipMemset(fs, 0, sizeof(FusedState)), return bail(HEAT_E_HIP));
        if (F)
            HIP_TRY_OR(hipMemcpy(v + 7 * nn, F + (size_t)i * nn, nn * sizeof(double), hipMemcpyDeviceToDevice),
                       return bail(HEAT_E_HIP));
        FusedView<double> w = make_view<double>(op, v, v + 4 * nn, v + 5 * nn, v + 2 * nn, v + 3 * nn, v + 6 * nn,
                                                v + nn, op->vec[HEAT_VEC_DIAG], op->cr, op->ck, nullptr);
        w.F = v + 7 * nn;
        w.fs = fs;
        w.hflag = nullptr;                       // a batch polls its own all-done flag
        b->views.push_back(w);
        b->tile_off[i + 1] = b->tile_off[i] + op->fntiles;
        b->sb_off[i + 1] = b->sb_off[i] + op->nsb;
        b->node_off[i + 1] = b->node_off[i] + (int)nn;
    }
    const int rc = batch_finish(b);
    if (rc) return bail(rc);
    *out = b;
    return HEAT_OK;
}

int heat_batch_vector(heat_batch* b, int i, int which, double** ptr) {
    if (!b || !ptr || i < 0 || i >= (int)b->views.size()) return fail(HEAT_E_ARG, "bad batch member");
    const FusedView<double>& v = b->views[i];
    double* p = nullptr;
    switch (which) {
        case HEAT_VEC_T: p = v.x; break;
        case HEAT_VEC_B: p = v.b; break;
        case HEAT_VEC_R: p = v.r; break;
        case HEAT_VEC_Z: p = v.z; break;
        case HEAT_VEC_P: p = v.P0; break;
        case HEAT_VEC_AP: p = v.Ap; break;
        case HEAT_VEC_F: p = const_cast<double*>(v.F); break;
        default: return fail(HEAT_E_ARG, "vector %d is not held per batch member", which);
    }
    *ptr = p;
    return HEAT_OK;
}

int heat_batch_cleanup(heat_batch* b) {
    if (!b) return HEAT_OK;
    batch_free(b);
    delete b;
    return HEAT_OK;
}

int heat_batch_instance(heat_batch* b, int i, heat_op** op) {
    if (!b || !op || i < 0 || i >= (int)b->ops.size())
        return fail(HEAT_E_ARG, "bad batch instance (a shared-mesh batch has no member handles)");
    *op = b->ops[i];
    return HEAT_OK;
}

int heat_batch_set_active(heat_batch* b, const int* active) {
    if (!b || !active) return fail(HEAT_E_ARG, "null argument");
    for (size_t i = 0; i < b->views.size(); ++i) b->views[i].active = active[i] ? 1 : 0;
    return batch_upload_views(b);
}

int heat_batch_set_tuning(heat_batch* b, int pcg_chunk, int use_graph) {
    if (!b || pcg_chunk < 1 || pcg_chunk > 64) return fail(HEAT_E_ARG, "pcg_chunk must be in 1..64");
    b->pcg_chunk = pcg_chunk;
    b->use_graph = use_graph ? 1 : 0;
    return HEAT_OK;
}

int heat_batch_set_persistent(heat_batch* b, int on) {
    if (!b) return fail(HEAT_E_ARG, "null argument");
    b->persistent = on ? 1 : 0;
    return HEAT_OK;
}

}  // extern "C"

namespace {

// Lay the batch's tiles out again after its members' fused geometry changed:
// tile offsets, the shared tile-partial rows, every view (FP64 and FP32) and
// the captured chunks.
int batch_relayout(heat_batch* b) {
    const int count = (int)b->views.size();
    for (int i = 0; i < count; ++i) {
        const heat_op* op = b->mesh ? b->mesh : b->ops[i];
        b->tile_off[i + 1] = b->tile_off[i] + op->fntiles;
    }
    const int total_tiles = b->tile_off[count];
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipFree(b->tp));
    b->tp = nullptr;
    HIP_TRY(hipMalloc(&b->tp, 4 * sizeof(double) * (size_t)total_tiles));
    for (int i = 0; i < count; ++i) {
        const heat_op* op = b->mesh ? b->mesh : b->ops[i];
        auto relay = [&](auto& v) {
            v.tp = b->tp + b->tile_off[i];
            v.tp_stride = total_tiles;
            v.ntiles = op->fntiles;
            v.nch = op->fnch;
            v.pre_tp = v.ntiles > FUSED_INBLOCK_MAX;
        };
        relay(b->views[i]);
        if (!b->views32.empty()) relay(b->views32[i]);
    }
    HIP_TRY(hipMemcpy(b->d_tile_off, b->tile_off.data(), (count + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    if (b->gexec32) { (void)hipGraphExecDestroy(b->gexec32); b->gexec32 = nullptr; }
    if (b->graph32) { (void)hipGraphDestroy(b->graph32); b->graph32 = nullptr; }
    return batch_upload_views(b);
}

int batch_set_geometry(heat_batch* b, int planes, int wave) {
    if (b->mesh) {
        int rc = set_geometry(b->mesh, 0, planes, wave);
        if (rc) return rc;
    } else {
        for (heat_op* op : b->ops) {
            int rc = set_geometry(op, 0, planes, wave);
            if (rc) return rc;
        }
    }
    return batch_relayout(b);
}

}  // namespace

extern "C" {

int heat_batch_set_geometry(heat_batch* b, int planes, int wave) {
    if (!b) return fail(HEAT_E_ARG, "null argument");
    return batch_set_geometry(b, planes, wave);
}

int heat_batch_geometry(heat_batch* b, int i, int* planes, int* wave, int* chunks, int* tiles) {
    if (!b || i < 0 || i >= (int)b->views.size()) return fail(HEAT_E_ARG, "bad batch member");
    const heat_op* op = b->mesh ? b->mesh : b->ops[i];
    return heat_pcg_geometry(op, 0, planes, wave, chunks, tiles);
}

// Time one batched step of `iters` iterations (mode 0: FP64 steps, 1: refined
// steps with one outer pass) for every candidate geometry, every member
// starting from the T it holds, and keep the fastest.  Every member's T is
// restored afterwards.  Batches with detectors are not tuned (their D1
// records would change).
int heat_batch_autotune(heat_batch* b, int mode, int iters, int* planes, int* wave, double* best_ms) {
    if (!b || mode < 0 || mode > 1 || iters < 1) return fail(HEAT_E_ARG, "bad batch autotune request");
    if (b->detect) return fail(HEAT_E_STATE, "a batch with detectors is not tuned (tune one without)");
    if (mode == 1 && b->mesh) return fail(HEAT_E_STATE, "refined batch steps need a batch of handles");
    const int count = (int)b->views.size();
    const size_t total = (size_t)b->node_off[count];
    double* saved = nullptr;
    HIP_TRY(hipMalloc(&saved, sizeof(double) * total));
    hipStream_t s = b->stream;
    auto copy_T = [&](bool save) -> int {
        for (int i = 0; i < count; ++i) {
            double* T = b->views[i].x;
            double* keep = saved + b->node_off[i];
            const size_t nb = sizeof(double) * (size_t)(b->node_off[i + 1] - b->node_off[i]);
            HIP_TRY(hipMemcpyAsync(save ? keep : T, save ? T : keep, nb, hipMemcpyDeviceToDevice, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        return HEAT_OK;
    };
    const heat_op* op0 = b->mesh ? b->mesh : b->ops[0];
    const int keep_p = op0->geo_planes[0], keep_w = op0->geo_wave[0];
    int rc = copy_T(true);
    const int cand_planes[] = {4, 8, 12, 16, 24, 32, 64};
    const int cand_wave[] = {0, FWAVE};
    std::vector<int> its(count);
    std::vector<std::vector<int>> seen;
    int pick_p = keep_p, pick_w = keep_w;
    double best = 1e300;
    for (int cp : cand_planes) {
        for (int cw : cand_wave) {
            if (rc) break;
            std::vector<int> key;
            for (int i = 0; i < count; ++i) {
                const heat_op* op = b->mesh ? b->mesh : b->ops[i];
                key.push_back(fused_chunks(op->g.nn1, op->fnby * op->fnbz, cp, cw));
            }
            if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
            seen.push_back(key);
            rc = batch_set_geometry(b, cp, cw);
            double t_best = 1e300;
            for (int rep = 0; rep < 2 && rc == HEAT_OK; ++rep) {
                rc = copy_T(false);
                if (rc) break;
                const auto t0 = std::chrono::steady_clock::now();
                rc = mode == 0 ? heat_batch_step(b, 1, iters, 0.0, its.data())
                               : heat_batch_step_mixed(b, 1, 1, iters, 0.0, 0.0, its.data());
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (ms < t_best) t_best = ms;
            }
            if (rc == HEAT_OK && t_best < best) { best = t_best; pick_p = cp; pick_w = cw; }
        }
    }
    int rc2 = batch_set_geometry(b, rc ? keep_p : pick_p, rc ? keep_w : pick_w);
    if (rc2 == HEAT_OK) rc2 = copy_T(false);
    (void)hipFree(saved);
    if (rc == HEAT_OK) rc = rc2;
    if (rc) return rc;
    if (planes) *planes = pick_p;
    if (wave) *wave = pick_w;
    if (best_ms) *best_ms = best / iters;
    return HEAT_OK;
}

int heat_batch_step(heat_batch* b, int nsteps, int max_iter, double tol, int* iters) {
    if (!b || nsteps < 0 || max_iter < 0) return fail(HEAT_E_ARG, "bad batch step arguments");
    const int count = (int)b->views.size();
    hipStream_t s = b->stream;
    for (heat_op* op : b->ops) HIP_TRY(hipStreamSynchronize(op->stream));
    if (b->mesh) HIP_TRY(hipStreamSynchronize(b->mesh->stream));
    const int total_sb = b->sb_off[count];
    const int nblk_nodes = (b->node_off[count] + NT - 1) / NT;
    for (int step = 0; step < nsteps; ++step) {
        if (!b->detect) {
            k_batch_rhs<<<nblk_nodes, NT, 0, s>>>(b->dviews, b->d_node_off, count);
            CHECK_LAUNCH();
        }
        k_batch_params<<<(count + NT - 1) / NT, NT, 0, s>>>(b->dviews, count, tol, max_iter);
        CHECK_LAUNCH();
        if (b->detect) {
            // with detectors each active solve forms b and r = b - A x0 as
            // heat_build_rhs and heat_pcg_solve do: the projected applies with
            // their D1 records (and an armed level-R fault firing in the second,
            // :999), then the fused start
            HIP_TRY(hipStreamSynchronize(s));
            for (int i = 0; i < count; ++i) {
                if (!b->views[i].active) continue;
                heat_op* op = b->ops[i];
                double* Ap = op->vec[HEAT_VEC_AP];
                int rc = heat_build_rhs(op, op->vec[HEAT_VEC_T], op->vec[HEAT_VEC_B]);
                if (rc) return rc;
                rc = apply_impl(op, op->vec[HEAT_VEC_T], Ap, op->cfg.dt, true, true, EPI_D1, 0);
                if (rc) return rc;
                // the batch's view: its super-block partials live in the batch's rows
                k_fused_init<double><<<op->nsb, NT, 0, op->stream>>>(b->views[i]);
                CHECK_LAUNCH();
            }
            for (int i = 0; i < count; ++i)
                if (b->views[i].active) HIP_TRY(hipStreamSynchronize(b->ops[i]->stream));
        } else {
            k_batch_residual<<<total_sb, NT, 0, s>>>(b->dviews, b->d_sb_off, count);
            CHECK_LAUNCH();
            // a member with a level-R fault armed forms its first residual
            // alone, as heat_pcg_solve does, so that the fault fires in that
            // apply (:999): the batch's residual is replaced by the member's
            bool armed = false;
            for (size_t i = 0; i < b->ops.size(); ++i)
                armed = armed || (b->views[i].active && b->ops[i]->fi.armed && b->ops[i]->fi_level == 2);
            if (armed) {
                HIP_TRY(hipStreamSynchronize(s));
                for (size_t i = 0; i < b->ops.size(); ++i) {
                    heat_op* op = b->ops[i];
                    if (!b->views[i].active || !op->fi.armed || op->fi_level != 2) continue;
                    int rc = apply_impl(op, op->vec[HEAT_VEC_T], op->vec[HEAT_VEC_AP], op->cfg.dt, true, true,
                                        EPI_NONE, 0);
                    if (rc) return rc;
                    k_fused_init<double><<<op->nsb, NT, 0, op->stream>>>(b->views[i]);
                    CHECK_LAUNCH();
                    HIP_TRY(hipStreamSynchronize(op->stream));
                }
            }
        }
        const int rc = batch_until_done(b, max_iter, [&]() { return batch_chunk(b); });
        if (rc) return rc;
        k_batch_iters<<<(count + NT - 1) / NT, NT, 0, s>>>(b->dviews, count, b->diters);
        CHECK_LAUNCH();
        HIP_TRY(hipMemcpyAsync(b->hiters, b->diters, count * sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (iters) memcpy(iters + (size_t)step * count, b->hiters, count * sizeof(int));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return HEAT_OK;
}

__global__ void k_abft_save(DevState* st, int slot) { st->res[slot] = st->abft_max; }

int heat_batch_detect(heat_batch* b, double* d1, double* d2, double* tmin, double* tmax) {
    if (!b || !d1 || !d2 || !tmin || !tmax) return fail(HEAT_E_ARG, "null argument");
    if (!b->detect) return fail(HEAT_E_STATE, "the batch's solves have no detectors");
    const int count = (int)b->views.size();
    // every member's sequence on its own stream, then one wait for all:
    // the D1 step maximum and its reset (heat_abft_take), the D2 apply and
    // residual (heat_apply_operator, heat_copy, heat_axpy, two heat_norm2),
    // the extremes of T (heat_minmax), the reset of the D1 record the D2 apply
    // made; the results land in res[0], res[1], res[2], res[4], res[5]
    for (int i = 0; i < count; ++i) {
        if (!b->views[i].active) continue;
        heat_op* op = b->ops[i];
        hipStream_t s = op->stream;
        const int nn = op->g.nn;
        double *T = op->vec[HEAT_VEC_T], *B = op->vec[HEAT_VEC_B], *R = op->vec[HEAT_VEC_R];
        double* Ap = op->vec[HEAT_VEC_AP];
        k_abft_save<<<1, 1, 0, s>>>(op->st, 0);
        k_abft_reset<<<1, 1, 0, s>>>(op->st);
        int rc = heat_apply_operator(op, T, Ap, op->cfg.dt, 0, 1);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(R, B, sizeof(double) * (size_t)nn, hipMemcpyDeviceToDevice, s));
        k_axpy<<<op->nbn, NT, 0, s>>>(-1.0, Ap, R, nn);
        k_dot<<<op->nbn, NT, 0, s>>>(B, B, nn, op->partials, op->st, 1);
        k_fin_dot<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, 1);
        k_dot<<<op->nbn, NT, 0, s>>>(R, R, nn, op->partials, op->st, 2);
        k_fin_dot<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, 2);
        k_minmax<<<op->nbn, NT, 0, s>>>(T, nn, op->partials, op->st, 4);
        k_fin_minmax<<<1, FT, 0, s>>>(op->partials, (int)op->nbn, op->st, 4);
        k_abft_reset<<<1, 1, 0, s>>>(op->st);
        CHECK_LAUNCH();
        HIP_TRY(hipMemcpyAsync(op->hst, op->st, sizeof(DevState), hipMemcpyDeviceToHost, s));
    }
    for (int i = 0; i < count; ++i) {
        if (!b->views[i].active) {
            d1[i] = d2[i] = tmin[i] = tmax[i] = std::nan("");
            continue;
        }
        heat_op* op = b->ops[i];
        HIP_TRY(hipStreamSynchronize(op->stream));
        const DevState* h = op->hst;
        d1[i] = h->res[0];
        double bn = std::sqrt(h->res[1]);                    // heat_norm2(b), clamped as the driver does
        if (bn < 1e-300) bn = 1.0;
        d2[i] = std::sqrt(h->res[2]) / bn;
        tmin[i] = h->res[4];
        tmax[i] = h->res[5];
    }
    return HEAT_OK;
}

int heat_batch_step_mrhs(heat_batch* b, int nsteps, int max_outer, int max_inner, double tol, double tol_inner,
                         int* inner_iters, double* d1max) {
    if (!b || nsteps < 0 || max_outer < 0 || max_inner < 0) return fail(HEAT_E_ARG, "bad batch step arguments");
    const int count = (int)b->views.size();
    if (!b->mesh || count != MRC)
        return fail(HEAT_E_STATE, "the 16-column refined solve needs a shared-mesh batch of exactly %d members", MRC);
    heat_op* m = b->mesh;
    for (int i = 0; i < count; ++i)
        if (!b->views[i].active) return fail(HEAT_E_STATE, "the 16-column refined solve has no member masks");
    int rc = ensure_mixed(m);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(m->stream));
    const Grid& g = m->g;
    hipStream_t s = b->stream;
    if (!b->mrE) {
        const size_t bytes = sizeof(float) * (size_t)g.nn * MRC;
        HIP_TRY(hipMalloc(&b->mrE, bytes));
        HIP_TRY(hipMalloc(&b->mrR, bytes));
        HIP_TRY(hipMalloc(&b->mrP, bytes));
        HIP_TRY(hipMalloc(&b->mrAP, bytes));
        HIP_TRY(hipMalloc(&b->mrpart, sizeof(double) * MRK * MRC * MRB));
        int nby_, nbz_;
        HIP_TRY(hipMalloc(&b->mrtile, sizeof(double) * MRK * MRC * (size_t)mrhs_tiles(g, mrhs_chunk(m), &nby_, &nbz_)));
        HIP_TRY(hipMalloc(&b->mrst, sizeof(MrState)));
        HIP_TRY(hipHostMalloc(&b->hmrst, sizeof(MrState), hipHostMallocDefault));
    }
    const bool detect = m->cfg.detect != 0;
    const double* w = m->vec[HEAT_VEC_WA];
    int nby_mr, nbz_mr;
    const int ntiles_mr = (int)mrhs_tiles(g, mrhs_chunk(m), &nby_mr, &nbz_mr);
    const int nblk_nodes = (b->node_off[count] + NT - 1) / NT;
    auto read_state = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(b->hmrst, b->mrst, sizeof(MrState), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return HEAT_OK;
    };
    for (int step = 0; step < nsteps; ++step) {
        k_batch_rhs<<<nblk_nodes, NT, 0, s>>>(b->dviews, b->d_node_off, count);
        CHECK_LAUNCH();
        k_mr_params<<<1, 1, 0, s>>>(b->mrst, tol, tol_inner, max_inner, detect ? m->nwA : 0.0);
        CHECK_LAUNCH();
        for (int nout = 0;; ++nout) {
            k_mr_init<<<MRB, NT, 0, s>>>(b->dviews, m->cr, m->ck, m->d32, g, b->mrE, b->mrR, b->mrP,
                                         b->mrpart, b->mrst);
            CHECK_LAUNCH();
            k_mr_fin_init<<<MRC, NT, 0, s>>>(b->mrpart, b->mrst);
            CHECK_LAUNCH();
            k_mr_flags<<<1, 1, 0, s>>>(b->mrst);
            CHECK_LAUNCH();
            if ((rc = read_state())) return rc;
            // inner iterations in chunks of pcg_chunk, until every column is done
            for (int it = 0; !b->hmrst->all_done; it += b->pcg_chunk) {
                if (it > max_inner + b->pcg_chunk) return fail(HEAT_E_STATE, "16-column PCG did not report completion");
                for (int k = 0; k < b->pcg_chunk; ++k) {
                    // Ap = P A p with the tiles' p.Ap [and D1] sums, then alpha
                    if ((rc = launch_mrhs(m, b->mrP, b->mrAP, false, s, detect ? 2 : 1, b->mrtile, w))) return rc;
                    if (detect) k_mr_fin_pap<true><<<MRC, NT, 0, s>>>(b->mrtile, ntiles_mr, b->mrst);
                    else k_mr_fin_pap<false><<<MRC, NT, 0, s>>>(b->mrtile, ntiles_mr, b->mrst);
                    CHECK_LAUNCH();
                    k_mr_update<<<MRB, NT, 0, s>>>(b->mrR, b->mrAP, m->d32, g, b->mrpart,
                                                   b->mrst);
                    CHECK_LAUNCH();
                    k_mr_fin_update<<<MRC, NT, 0, s>>>(b->mrpart, b->mrst);
                    CHECK_LAUNCH();
                    k_mr_dir<<<MRB, NT, 0, s>>>(b->mrE, b->mrR, m->d32, b->mrP, g, b->mrst);
                    CHECK_LAUNCH();
                }
                k_mr_flags<<<1, 1, 0, s>>>(b->mrst);
                CHECK_LAUNCH();
                if ((rc = read_state())) return rc;
            }
            if (nout >= max_outer) break;
            k_mr_correct<<<MRB, NT, 0, s>>>(b->dviews, b->mrE, g, b->mrst);
            CHECK_LAUNCH();
            k_mr_flags<<<1, 1, 0, s>>>(b->mrst);
            CHECK_LAUNCH();
            if ((rc = read_state())) return rc;
            if (b->hmrst->all_outer_done) break;
        }
        if ((rc = read_state())) return rc;
        for (int c = 0; c < MRC; ++c) {
            if (inner_iters) inner_iters[(size_t)step * MRC + c] = b->hmrst->inner_total[c];
            if (d1max) d1max[(size_t)step * MRC + c] = b->hmrst->d1max[c];
        }
    }
    return HEAT_OK;
}

int heat_batch_step_mixed(heat_batch* b, int nsteps, int max_outer, int max_inner, double tol, double tol_inner,
                          int* inner_iters) {
    if (!b || nsteps < 0 || max_outer < 0 || max_inner < 0) return fail(HEAT_E_ARG, "bad batch step arguments");
    if (b->detect) return fail(HEAT_E_STATE, "refined batch steps run without detectors");
    int rc = batch_views32(b);
    if (rc) return rc;
    const int count = (int)b->views.size();
    hipStream_t s = b->stream;
    for (heat_op* op : b->ops) HIP_TRY(hipStreamSynchronize(op->stream));
    const int total_sb = b->sb_off[count];
    const int nblk_nodes = (b->node_off[count] + NT - 1) / NT;
    for (int step = 0; step < nsteps; ++step) {
        k_batch_rhs<<<nblk_nodes, NT, 0, s>>>(b->dviews, b->d_node_off, count);
        CHECK_LAUNCH();
        k_batch_params_mixed<<<(count + NT - 1) / NT, NT, 0, s>>>(b->dviews, count, tol, tol_inner, max_inner);
        CHECK_LAUNCH();
        // outer passes in step for every open solve, as heat_pcg_solve_mixed's
        // loop: FP64 residual, FP32 inner solve (whose first check may find the
        // outer tolerance met), then x += e unless the pass limit is reached
        for (int nout = 0;; ++nout) {
            k_batch_mixed_init<<<total_sb, NT, 0, s>>>(b->dviews, b->dviews32, b->d_sb_off, count);
            CHECK_LAUNCH();
            rc = batch_until_done(b, max_inner, [&]() {
                return batch_chunk_t<float>(b, b->views32, b->dviews32, b->graph32, b->gexec32, b->graph_chunk32);
            });
            if (rc) return rc;
            if (nout >= max_outer) break;
            k_batch_mixed_correct<<<nblk_nodes, NT, 0, s>>>(b->dviews, b->dviews32, b->d_node_off, count);
            CHECK_LAUNCH();
            k_batch_outer_done<<<1, NT, 0, s>>>(b->dviews, count, b->dflag);
            CHECK_LAUNCH();
            HIP_TRY(hipMemcpyAsync(&b->hflag[0], b->dflag, sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (b->hflag[0]) break;
        }
        k_batch_inner<<<(count + NT - 1) / NT, NT, 0, s>>>(b->dviews, count, b->diters);
        CHECK_LAUNCH();
        HIP_TRY(hipMemcpyAsync(b->hiters, b->diters, count * sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (inner_iters) memcpy(inner_iters + (size_t)step * count, b->hiters, count * sizeof(int));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return HEAT_OK;
}

}  // extern "C"
