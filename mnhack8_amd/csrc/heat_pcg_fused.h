// heat_pcg_fused.h -- Jacobi-PCG in two kernels per iteration (K1 runs as two
// launches, k_fused_dir and k_fused_tdma, unless a PCG-state upset is armed).
//
// The reference's iteration (src/heat_solver_het.cu:1009-1036) runs matvec,
// p.Ap, two axpys, ||r||, z = r/d, r.z and p = z + beta p with a host round
// trip per inner product.  Here one iteration is:
//   K1 k_fused_apply : beta and convergence from K2's partials, p = z + beta p
//                      (fused into the tile fetch), Ap = P A p (the LDS-tiled
//                      colour-ordered apply), partials of p.Ap [and D1]
//   K2 k_fused_update: alpha = rho / p.Ap from K1's partials, x += alpha p,
//                      r -= alpha Ap, z = r/d, partials of r.r and r.z
// Every block reduces the other kernel's few hundred partials itself, in the
// same fixed order, so no separate finalize launch is needed (beyond
// FUSED_INBLOCK_MAX partials a one-block launch reduces them once).
//
// Reduction order (deterministic, independent of the autotuned apply schedule):
//   K1 partials: tile (TY x TZ = 8 x 32 node columns, a chunk of consecutive
//     planes: fused_chunks() chunks per column); each thread sums its column's
//     products over its planes in plane order, then a 256-wide tree gives one
//     partial per tile;
//   K2 / init partials: super-blocks of SB = 2048 consecutive nodes; thread t
//     sums nodes base + 256 s + t for s = 0..7 in order, then a 256-wide tree;
//   final: thread t sums partials t, t+256, ... in order, then a 256-wide tree.
// The CPU oracle restates exactly this order (ho_pcg_solve_fused).
//
// State: values written by K1 are read only by K2 and vice versa (rho is
// double-buffered by iteration parity), so no block reads a value that another
// block of the same launch writes.  The search direction is double-buffered
// the same way: iteration k reads p_{k-1} from P[(k+1)&1] (tile halos included)
// and writes p_k to P[k&1], so no tile reads a direction another tile has
// already overwritten.
//
// D1 of the fused apply: s = 1_int.Ap sums each thread's own interior nodes;
// c = w.p and xx = p.p are summed where the tile writes p (every node of the
// mesh exactly once, boundary included), in plane order, window slot order.
#pragma once
#include <hip/hip_cooperative_groups.h>

#include "heat_kernels.h"

namespace heat {

constexpr int FTY = 8, FTZ = 32;              // fused apply tile (y, z)
constexpr int FCX = 16;                       // target node planes per tile
constexpr int FCX_MIN = 8;                    // shortest chunk the wave fill may make
constexpr int FWAVE = 768;                    // tiles per wave: 256 CUs x 3 resident blocks

// Chunks per (y,z) column of the fused apply: ceil(nn1 / planes), raised
// until the tiles fill whole waves of `wave` tiles (wave = 0: no fill), but
// never to chunks shorter than FCX_MIN planes (small meshes, whose tiles a
// batch launches together, keep long chunks).  The chunks split the
// column's plane pairs as evenly as possible (at most one pair apart; the
// result never exceeds the pair count for planes >= 2).  A function of the
// mesh and of (planes, wave) alone, so the partial sums' order (and every
// bit of the solve) does not depend on the device.  The default geometry is
// (FCX, FWAVE); an autotuner may pick another pair, and the oracle takes
// the same pair.
__host__ __device__ inline int fused_chunks(int nn1, int ncol, int planes = FCX, int wave = FWAVE) {
    const int nch = (nn1 + planes - 1) / planes;
    if (wave <= 0) return nch;
    const long long waves = ((long long)ncol * nch + wave - 1) / wave;
    long long fill = waves * wave / ncol;
    if (fill > nn1 / FCX_MIN) fill = nn1 / FCX_MIN;
    return fill > nch ? (int)fill : nch;
}
// the largest chunk count fused_chunks can return for nn1 planes (planes >= 2)
__host__ __device__ inline int fused_chunks_max(int nn1) { return (nn1 + 1) / 2; }
constexpr int SB = 2048;                      // nodes per update super-block
constexpr int SBS = SB / NT;                  // strided nodes per thread

struct FusedState {
    double rho[2];          // r.z of the last two K1 launches (slot iter & 1)
    double bnorm, alpha, beta, tol;
    int iter;               // completed updates, written by K2 (and init)
    int k1_iter;            // iteration index seen by the current K1, written by K1
    int max_iter, done;
    int outer_done, inner_total, check_outer, pad1;
    double bnorm_outer, rnorm_outer, tol_outer;
    // single-event upset in the PCG state: at iteration fi_iter (counted over
    // inner_base + iter), the producer of vector fi_vec stores node fi_idx with
    // bit fi_bit flipped; sums formed in the same kernel use the clean value
    int fi_armed, fi_vec, fi_bit, fi_iter;
    long long fi_idx;
    int inner_base, fi_fired;
    // sums of the other kernel's partials, when a one-block finalize launch
    // forms them (large meshes) instead of every block of K1 / K2
    double red_sp[5], red_tp[4];
    // the one-launch inner iteration (heat_cg1.h), slots by iteration parity:
    // gamma = r.u and alpha of the step, the index of the next step, and the
    // folded partials when a one-block launch folds them
    double cg_gamma[2], cg_alpha[2];
    double cg_red[2][8];
    int cg_it[2];
    // pass number of the refined solve (set by the pass's start launch) and
    // the status its inner solve ended with: 1 inner done, 2 the solve is
    // finished (outer test met or residual not finite), 3 launched out of step
    int cg_pass, cg_status;
};

enum { SEU_U = 0, SEU_R = 1, SEU_P = 2, SEU_Z = 3, SEU_AP = 4 };

__device__ __forceinline__ double seu_flip(double v, int bit) {
    return __longlong_as_double(__double_as_longlong(v) ^ (long long)(1ull << bit));
}
__device__ __forceinline__ float seu_flip(float v, int bit) {
    return __int_as_float(__float_as_int(v) ^ (int)(1u << bit));
}

// One solve's vectors, mesh and reduction slots.  A single solve launches
// with its view by value; a batch of solves passes a device table of views,
// and each block finds its solve from its tile or super-block index.  The
// partial arrays of a batch are shared: row q of solve i lives at
// sp[q * sp_stride + j] with sp pointing at the solve's first super-block.
template <typename T>
struct FusedView {
    T *x, *P0, *P1, *r, *z, *Ap, *b;
    const T *d, *cr, *ck;
    const double *w, *F;
    Grid g;
    double* sp;                 // super-block partials (rows of sp_stride)
    double* tp;                 // tile partials (rows of tp_stride)
    int sp_stride, nsb, tp_stride, ntiles, nby, nbz, nch;
    FusedState* fs;
    DevState* st;
    double nwA, dt;             // ||w_A||; time step of the right-hand side
    double* hist;
    int add_source, active;     // b = M T (+ dt F); a masked solve does no work
    int pre_sp, pre_tp;         // partial sums come finalized in fs->red_sp / red_tp
    int* hflag;                 // pinned host word: K2 block 0 stores fs->done there (or null)
};

// Partial counts up to this are reduced by every block that needs them;
// beyond it a one-block launch reduces them once (same order, same bits).
constexpr int FUSED_INBLOCK_MAX = 2048;

// 256-wide tree over one value per thread; result valid in thread 0.
// The pairs and their order are those of the plain LDS tree (at stride st,
// element t += element t + st, for st = 128, 64, ..., 1); the strides below
// 64 run inside wave 0 with lane shuffles instead of LDS and barriers, which
// adds the same values in the same order.
__device__ __forceinline__ double wave0_tail(double v) {
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) v += __shfl_down(v, st, 64);
    return v;
}

__device__ __forceinline__ double tree256(double (*sd)[NT], int k, double v) {
    static_assert(NT == 256, "two LDS strides, then one wave");
    const int t = threadIdx.x;
    sd[k][t] = v;
    __syncthreads();
    if (t < 128) sd[k][t] += sd[k][t + 128];
    __syncthreads();
    if (t < 64) {
        const double w = wave0_tail(sd[k][t] + sd[k][t + 64]);
        if (t == 0) sd[k][0] = w;
    }
    __syncthreads();
    return sd[k][0];
}

// K independent 256-wide trees (row k over v[k]) with one set of barriers;
// each row's pairs and order are tree256's.  Results valid in every thread.
template <int K>
__device__ __forceinline__ void tree_rows(double (*sd)[NT], const double (&v)[K], double* out) {
    static_assert(NT == 256, "two LDS strides, then one wave");
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) sd[k][t] = v[k];
    __syncthreads();
    if (t < 128) {
#pragma unroll
        for (int k = 0; k < K; ++k) sd[k][t] += sd[k][t + 128];
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double w = wave0_tail(sd[k][t] + sd[k][t + 64]);
            if (t == 0) sd[k][0] = w;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = sd[k][0];
    __syncthreads();
}

// Final stage over `count` partials of K sums (row k at part + k * stride):
// thread t adds partials t, t + NT, t + 2 NT, ... of each row in that order,
// then tree_rows; every thread gets the results.  The first four partials of
// each row are loaded together (partials_load4) before any is added, so a
// row of up to 4 NT partials costs one memory round trip.
template <int K>
__device__ __forceinline__ void partials_load4(const double* __restrict__ part, int count, int stride,
                                               double (&x)[K][4]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = t + j * NT;
            x[k][j] = i < count ? part[(size_t)k * stride + i] : 0.0;
        }
}
template <int K, int KX>
__device__ __forceinline__ void partials_fold(double (*sd)[NT], const double (&x)[KX][4],
                                              const double* __restrict__ part, int count, int stride,
                                              double* out) {
    static_assert(K <= KX, "rows loaded");
    const int t = threadIdx.x;
    double a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        a[k] = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (t + j * NT < count) a[k] += x[k][j];
        for (int i = t + 4 * NT; i < count; i += NT) a[k] += part[(size_t)k * stride + i];
    }
    tree_rows<K>(sd, a, out);
}
template <int K>
__device__ __forceinline__ void reduce_partials(double (*sd)[NT], const double* __restrict__ part, int count,
                                                int stride, double* out) {
    double x[K][4];
    partials_load4<K>(part, count, stride, x);
    partials_fold<K, K>(sd, x, part, count, stride, out);
}

// r = b - Ap; z = r/d; super-block partials of b.b, r.r, r.z (super-block j).
template <typename T>
__device__ __forceinline__ void fused_init_body(const FusedView<T>& v, int j) {
    __shared__ double sd[3][NT];
    const T* __restrict__ b = v.b;
    const T* __restrict__ Ap = v.Ap;
    const T* __restrict__ d = v.d;
    T* __restrict__ r = v.r;
    T* __restrict__ z = v.z;
    const int nn = v.g.nn;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const size_t base = (size_t)j * SB + threadIdx.x;
    for (int s = 0; s < SBS; ++s) {
        const size_t i = base + (size_t)s * NT;
        if (i < (size_t)nn) {
            const T bi = b[i];
            const T ri = (T)fma(-1.0, (double)Ap[i], (double)bi);
            const T zi = ri / d[i];
            r[i] = ri;
            z[i] = zi;
            a0 += (double)bi * (double)bi;
            a1 += (double)ri * (double)ri;
            a2 += (double)ri * (double)zi;
        }
    }
    const double s0 = tree256(sd, 0, a0), s1 = tree256(sd, 1, a1), s2 = tree256(sd, 2, a2);
    if (threadIdx.x == 0) {
        v.sp[j] = s0;
        v.sp[v.sp_stride + j] = s1;
        v.sp[2 * v.sp_stride + j] = s2;
        if (j == 0) {
            v.fs->iter = 0;
            v.fs->done = v.active ? 0 : 1;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_fused_init(FusedView<T> v) {
    fused_init_body<T>(v, blockIdx.x);
}

// Start of K1 (tile b): beta and the convergence test from K2's partials (or,
// at iteration 0, the init partials), in the reference's order of checks
// (:1020-1035); block 0 records the history and the done flag.  False: the
// solve is done and the block does nothing more (uniform over the grid).
// The done flag as one value for the whole block.  PERSIST (the persistent
// batched kernel, whose blocks go on to other tiles after a body returns):
// thread 0 reads it and hands it on through LDS, since block 0 of a solve
// may set it while another block of the same phase reads it, and a block
// whose waves read different values would leave the body at different
// barriers.  Otherwise each wave reads it (a wave that leaves ends).
template <bool PERSIST>
__device__ __forceinline__ int block_done(const FusedState* fs) {
    if constexpr (!PERSIST) {
        return fs->done;
    } else {
        __shared__ int s_done;
        if (threadIdx.x == 0) s_done = __hip_atomic_load(&fs->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int d = s_done;
        __syncthreads();
        return d;
    }
}

template <typename T, bool PERSIST = false>
__device__ __forceinline__ bool fused_head(const FusedView<T>& v, int b, double (*sd)[NT], double& beta) {
    const double* __restrict__ sp = v.sp;
    const int nsb = v.nsb, sp_stride = v.sp_stride;
    FusedState* fs = v.fs;
    double* __restrict__ hist = v.hist;
    if (block_done<PERSIST>(fs)) return false;
    const int iter = fs->iter;
    // ---- beta, convergence (reference order of checks, :1020-1035)
    beta = 0.0;
    if (iter == 0) {
        double o[5];
        if (fs->check_outer) {
            if (v.pre_sp) { for (int k = 0; k < 5; ++k) o[k] = fs->red_sp[k]; }
            else reduce_partials<5>(sd, sp, nsb, sp_stride, o);
            double bo = sqrt(o[3]);
            if (bo < 1e-30) bo = 1.0;
            // 1: converged; 2: the FP64 residual is not finite (stop: no
            // correction can recover it); 0: run the inner solve
            const double rn = sqrt(o[4]);
            const int outer_done = (rn / bo < fs->tol_outer) ? 1 : (isfinite(rn) ? 0 : 2);
            if (threadIdx.x == 0 && b == 0) {
                fs->bnorm_outer = bo;
                fs->rnorm_outer = rn;
                fs->outer_done = outer_done;
                if (outer_done) fs->done = 1;
            }
            if (outer_done) return false;
        } else {
            if (v.pre_sp) { for (int k = 0; k < 3; ++k) o[k] = fs->red_sp[k]; }
            else reduce_partials<3>(sd, sp, nsb, sp_stride, o);
        }
        double bnorm = sqrt(o[0]);
        if (bnorm < 1e-30) bnorm = 1.0;
        if (threadIdx.x == 0 && b == 0) {
            fs->bnorm = bnorm;
            fs->rho[0] = o[2];
            fs->k1_iter = 0;
            if (hist) { hist[0] = sqrt(o[1]) / bnorm; hist[1] = __builtin_nan(""); hist[2] = __builtin_nan(""); }
            if (fs->max_iter <= 0) fs->done = 1;
        }
        if (fs->max_iter <= 0) return false;
    } else {
        double o[2];
        if (v.pre_sp) { o[0] = fs->red_sp[0]; o[1] = fs->red_sp[1]; }
        else reduce_partials<2>(sd, sp, nsb, sp_stride, o);
        const double bnorm = fs->bnorm;
        const double rnorm = sqrt(o[0]);
        const bool conv = (rnorm / bnorm < fs->tol);
        if (!conv) beta = o[1] / fs->rho[(iter - 1) & 1];
        if (threadIdx.x == 0 && b == 0) {
            if (hist) {
                double* h = hist + 3 * iter;
                h[0] = rnorm / bnorm; h[1] = fs->alpha; h[2] = conv ? 0.0 : beta;
            }
            if (conv || iter >= fs->max_iter) fs->done = 1;
            else { fs->rho[iter & 1] = o[1]; fs->beta = beta; fs->k1_iter = iter; }
        }
        if (conv || iter >= fs->max_iter) return false;      // uniform over the grid
    }
    return true;
}

// K1.  sp: K2 partials (r.r, r.z) or, at iteration 0, init partials (b.b, r.r, r.z).
// tp: output partials of this launch [p.Ap | 1.Ap | w.p | p.p] x ntiles.
// (tile b of the solve's ntiles)
template <typename T, bool DETECT>
__device__ __forceinline__ void fused_apply_body(const FusedView<T>& v, int b) {
    T* __restrict__ P0 = v.P0;
    T* __restrict__ P1 = v.P1;
    const T* __restrict__ z = v.z;
    T* __restrict__ Ap = v.Ap;
    const T* __restrict__ cr = v.cr;
    const T* __restrict__ ck = v.ck;
    const Grid g = v.g;
    const double* __restrict__ w = v.w;
    double* __restrict__ tp = v.tp;
    const int nby = v.nby, nbz = v.nbz, tp_stride = v.tp_stride;
    FusedState* fs = v.fs;
    constexpr int TY = FTY, TZ = FTZ, NTH = TY * TZ;
    constexpr int XW = TZ + 2, XH = TY + 2, XN = XW * XH;
    constexpr int EW = TZ + 1, EH = TY + 1, EN = EW * EH;
    constexpr int XR = (XN + NTH - 1) / NTH, ER = (EN + NTH - 1) / NTH;
    __shared__ T xs[2][XN];
    __shared__ T es_r[2][EN];
    __shared__ T es_k[2][EN];
    __shared__ double sd[5][NT];         // reductions (partials in, tile sums out)
    if (fs->done) return;
    const int iter = fs->iter;
    double beta;
    if (!fused_head<T>(v, b, sd, beta)) return;
    // ---- tile apply on p = z + beta p (iteration 0: p = z)
    const int bz = b % nbz;
    const int by = (b / nbz) % nby;
    const int chunk = b / (nbz * nby);
    const int n = g.n, S = g.sx, nn1 = g.nn1, nsq = n * n;
    const int tid = threadIdx.x;
    const int ty = tid / TZ, tz = tid - (tid / TZ) * TZ;
    const int y0 = 1 + by * TY, z0 = 1 + bz * TZ;
    const int iy = y0 + ty, iz = z0 + tz;
    const bool inmesh = (iy <= n) && (iz <= n);
    const bool interior = (iy < n) && (iz < n);
    // nch chunks of whole plane pairs per column, lengths differing by at
    // most one pair
    const int npair = (nn1 + 1) >> 1;
    const int ix0 = 2 * (int)(((long long)chunk * npair) / v.nch);
    const int ix1 = min(2 * (int)(((long long)(chunk + 1) * npair) / v.nch), nn1);
    const int py = iy & 1, pz = iz & 1;
    const int sy = py ? -1 : 1, sz = pz ? -1 : 1;
    const bool first = (iter == 0);
    const bool fi_now = fs->fi_armed && (fs->inner_base + iter == fs->fi_iter);
    const int fi_vec = fi_now ? fs->fi_vec : -1;
    const long long fi_idx = fs->fi_idx;
    const int fi_bit = fs->fi_bit;
    const T* __restrict__ pold = (iter & 1) ? P0 : P1;
    T* __restrict__ p = (iter & 1) ? P1 : P0;
    auto pnew = [&](size_t i) -> T {
        return first ? z[i] : (T)fma(beta, (double)pold[i], (double)z[i]);
    };
    double a_pap = 0.0, a_s = 0.0, a_c = 0.0, a_xx = 0.0;
    int xoff[XR], eoff_t[ER];
    bool xok[XR], xown[XR], eok[ER];
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int i = tid + k * NTH;
        const int r = i / XW, c = i - (i / XW) * XW;
        const int yy = y0 - 1 + r, zz = z0 - 1 + c;
        xok[k] = (i < XN) && yy <= n && zz <= n;
        // the tile writes the new p of its own nodes, and of row 0 / column 0 on the first tiles
        xown[k] = xok[k] && r <= TY && c <= TZ && (r >= 1 || by == 0) && (c >= 1 || bz == 0);
        xoff[k] = yy * nn1 + zz;
    }
#pragma unroll
    for (int k = 0; k < ER; ++k) {
        const int i = tid + k * NTH;
        const int r = i / EW, c = i - (i / EW) * EW;
        const int ey = y0 - 1 + r, ez = z0 - 1 + c;
        eok[k] = (i < EN) && ey < n && ez < n;
        eoff_t[k] = ey * n + ez;
    }
    // fetch issues the loads of a plane into registers; park (after the
    // plane in flight has been used) forms p = z + beta p_old, writes the
    // owned part of it to memory and the window to LDS, so that the load
    // latency is covered by the compute in between.
    T zf[XR], pf[XR], ef_r[ER], ef_k[ER];
    double wf[XR];
    int fplane = 0;
    auto fetch = [&](int xplane, int eplane) {
        fplane = xplane;
#pragma unroll
        for (int k = 0; k < XR; ++k) {
            T zv = T(0), pv = T(0);
            double wv = 0.0;
            if (xplane <= n && xok[k]) {
                const size_t i = (size_t)xplane * S + xoff[k];
                zv = z[i];
                if (!first) pv = pold[i];
                if (DETECT && xown[k]) wv = w[i];
            }
            zf[k] = zv;
            pf[k] = pv;
            wf[k] = wv;
        }
#pragma unroll
        for (int k = 0; k < ER; ++k) {
            const bool ok = eplane < n && eok[k];
            const size_t e = (size_t)eplane * nsq + eoff_t[k];
            ef_r[k] = ok ? cr[e] : T(0);
            ef_k[k] = ok ? ck[e] : T(0);
        }
    };
    auto park = [&](int xslot, int eslot) {
        const bool own_plane = fplane >= ix0 && fplane < ix1 && fplane <= n;
#pragma unroll
        for (int k = 0; k < XR; ++k) {
            const T v = first ? zf[k] : (T)fma(beta, (double)pf[k], (double)zf[k]);
            if (own_plane && xown[k]) {
                const size_t i = (size_t)fplane * S + xoff[k];
                if (fi_vec == SEU_P && (long long)i == fi_idx) {
                    p[i] = seu_flip(v, fi_bit);
                    fs->fi_fired = 1;
                } else {
                    p[i] = v;
                }
                if (DETECT) {
                    a_c += wf[k] * (double)v;
                    a_xx += (double)v * (double)v;
                }
            }
            if (tid + k * NTH < XN) xs[xslot][tid + k * NTH] = v;
        }
#pragma unroll
        for (int k = 0; k < ER; ++k)
            if (tid + k * NTH < EN) {
                es_r[eslot][tid + k * NTH] = ef_r[k];
                es_k[eslot][tid + k * NTH] = ef_k[k];
            }
    };
    int eidx[2][2], eoff[2][2], xidx[3][3];
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
            eidx[cy][cz] = (ty + 1 - (cy ^ py)) * EW + (tz + 1 - (cz ^ pz));
            eoff[cy][cz] = (iy - (cy ^ py)) * n + (iz - (cz ^ pz));
        }
#pragma unroll
    for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) xidx[bb][c] = (ty + 1 + (bb - 1) * sy) * XW + (tz + 1 + (c - 1) * sz);
    PlaneT<T> PT[3];                            // terms of window planes ix-1, ix, ix+1
    T Xc[3];                                    // the node's own value of each window plane
    Coef<T> Eprev[2][2], Enew[2][2];
    // prologue: plane ix0-1 of the window and element plane ix0-1 straight from memory
    if (interior && ix0 >= 1) {
        const size_t q = (size_t)iy * nn1 + iz;
        T Wn[3][3];
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                Wn[bb][c] = pnew((size_t)(ix0 - 1) * S + q + (bb - 1) * sy * nn1 + (c - 1) * sz);
        plane_terms(Wn, PT[0]);
        Xc[0] = Wn[1][1];
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) {
                const size_t e = (size_t)(ix0 - 1) * nsq + eoff[cy][cz];
                Eprev[cy][cz].cr = cr[e];
                Eprev[cy][cz].ck = ck[e];
            }
    }
    // plane ix0 through the tile (it writes the new p of plane ix0), then ix0+1
    fetch(ix0, ix0);
    park(0, 0);
    __syncthreads();
    if (interior) {
        T Wn[3][3];
#pragma unroll
        for (int bb = 0; bb < 3; ++bb)
#pragma unroll
            for (int c = 0; c < 3; ++c) Wn[bb][c] = xs[0][xidx[bb][c]];
        plane_terms(Wn, PT[1]);
        Xc[1] = Wn[1][1];
    }
    __syncthreads();
    fetch(ix0 + 1, ix0);
    park((ix0 + 1) & 1, ix0 & 1);
    __syncthreads();
    for (int ix = ix0; ix < ix1; ++ix) {
        fetch(ix + 2, ix + 1);
        if (interior && ix <= n - 1) {
            const T* xp = xs[(ix + 1) & 1];
            T Wn[3][3];
#pragma unroll
            for (int bb = 0; bb < 3; ++bb)
#pragma unroll
                for (int c = 0; c < 3; ++c) Wn[bb][c] = xp[xidx[bb][c]];
            plane_terms(Wn, PT[2]);
            Xc[2] = Wn[1][1];
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cz = 0; cz < 2; ++cz) {
                    Enew[cy][cz].cr = es_r[ix & 1][eidx[cy][cz]];
                    Enew[cy][cz].ck = es_k[ix & 1][eidx[cy][cz]];
                }
        }
        if (inmesh) {
            T acc = T(0);
            if (interior && ix >= 1 && ix <= n - 1) {
                if (ix & 1) acc = node_rows<T, false, 1>(PT[0], PT[1], PT[2], Eprev, Enew);
                else acc = node_rows<T, false, 0>(PT[0], PT[1], PT[2], Eprev, Enew);
            }
            const size_t base = (size_t)ix * S;
            const size_t me = base + (size_t)iy * nn1 + iz;
            auto store_ap = [&](size_t i, T v) {
                if (fi_vec == SEU_AP && (long long)i == fi_idx) {
                    Ap[i] = seu_flip(v, fi_bit);
                    fs->fi_fired = 1;
                } else {
                    Ap[i] = v;
                }
            };
            store_ap(me, acc);
            if (by == 0 && ty == 0) store_ap(base + iz, T(0));
            if (bz == 0 && tz == 0) store_ap(base + (size_t)iy * nn1, T(0));
            if (by == 0 && bz == 0 && tid == 0) store_ap(base, T(0));
            // p.Ap over interior nodes (Ap is 0 elsewhere, and adding a zero
            // product leaves a sum that starts at +0 unchanged)
            const bool int_node = interior && ix >= 1 && ix <= n - 1;
            if (int_node) {
                a_pap += (double)Xc[1] * (double)acc;
                if (DETECT) a_s += (double)acc;
            }
        }
        park(ix & 1, (ix + 1) & 1);
        __syncthreads();
        PT[0] = PT[1];
        PT[1] = PT[2];
        Xc[0] = Xc[1];
        Xc[1] = Xc[2];
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz];
    }
    const double t0 = tree256(sd, 0, a_pap);
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (DETECT) { t1 = tree256(sd, 1, a_s); t2 = tree256(sd, 2, a_c); t3 = tree256(sd, 3, a_xx); }
    if (tid == 0) {
        tp[b] = t0;
        if (DETECT) { tp[tp_stride + b] = t1; tp[2 * tp_stride + b] = t2; tp[3 * tp_stride + b] = t3; }
    }
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(FTY* FTZ, 3) void k_fused_apply(FusedView<T> v) {
    const int b = xcd_slab_block(v.ntiles);
    if (b >= v.ntiles) return;
    fused_apply_body<T, DETECT>(v, b);
}

// K1 in two launches (no upset armed): K1a forms the new direction, K1b
// applies the operator to it.  Same tiles, same partials, same bits as K1.
//
// K1a: beta and convergence as K1, then p = z + beta p_old.  With the
// detectors, over the tile's own nodes (each node of the mesh exactly once),
// in plane order, slot order, with the D1 sums w.p and p.p of K1 (partial
// rows 2 and 3); without them, as a plain stream.
template <typename T, bool DETECT, bool PERSIST = false>
__device__ __forceinline__ void fused_dir_body(const FusedView<T>& v, int b, double (*sd)[NT]) {
    double beta;
    if (!fused_head<T, PERSIST>(v, b, sd, beta)) return;
    constexpr int NTH = FTY * FTZ, XW = FTZ + 2, XL = XW * (FTY + 2);
    constexpr int XR = (XL + NTH - 1) / NTH;
    const FusedState* fs = v.fs;
    const int iter = fs->iter;
    const bool first = (iter == 0);
    const T* __restrict__ z = v.z;
    const T* __restrict__ pold = (iter & 1) ? v.P0 : v.P1;
    T* __restrict__ p = (iter & 1) ? v.P1 : v.P0;
    const double* __restrict__ w = v.w;
    if (!DETECT) {
        // no sums: a contiguous stream of 16-byte units (grid stride), the
        // last nn mod VPU nodes by the first block
        constexpr int VPU = (int)(16 / sizeof(T));
        struct alignas(16) U { T e[VPU]; };
        const long long nn = v.g.nn, nu = nn / VPU;
        const long long stride = (long long)v.ntiles * NTH;
        for (long long u = (long long)b * NTH + threadIdx.x; u < nu; u += stride) {
            const U zu = reinterpret_cast<const U*>(z)[u];
            U pu;
            if (first) {
                pu = zu;
            } else {
                const U ou = reinterpret_cast<const U*>(pold)[u];
#pragma unroll
                for (int e = 0; e < VPU; ++e) pu.e[e] = (T)fma(beta, (double)ou.e[e], (double)zu.e[e]);
            }
            reinterpret_cast<U*>(p)[u] = pu;
        }
        if (b == 0 && threadIdx.x < nn - nu * VPU) {
            const long long i = nu * VPU + threadIdx.x;
            p[i] = first ? z[i] : (T)fma(beta, (double)pold[i], (double)z[i]);
        }
        return;
    }
    const int nby = v.nby, nbz = v.nbz;
    const int bz = b % nbz, by = (b / nbz) % nby, chunk = b / (nbz * nby);
    const int n = v.g.n, S = v.g.sx, nn1 = v.g.nn1;
    const int tid = threadIdx.x;
    const int y0 = 1 + by * FTY, z0 = 1 + bz * FTZ;
    const int npair = (nn1 + 1) >> 1;
    const int ix0 = 2 * (int)(((long long)chunk * npair) / v.nch);
    const int ix1 = min(2 * (int)(((long long)(chunk + 1) * npair) / v.nch), nn1);
    int xoff[XR];
    bool xown[XR];
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int i = tid + k * NTH;
        const int r = i / XW, c = i - (i / XW) * XW;
        const int yy = y0 - 1 + r, zz = z0 - 1 + c;
        xown[k] = (i < XL) && yy <= n && zz <= n && r <= FTY && c <= FTZ && (r >= 1 || by == 0) &&
                  (c >= 1 || bz == 0);
        xoff[k] = yy * nn1 + zz;
    }
    double a_c = 0.0, a_xx = 0.0;
    // UP planes of loads in flight per thread, then the stores and the sums
    // in plane order, slot order
    constexpr int UP = 4;
    for (int xb = ix0; xb < ix1; xb += UP) {
        T zv[UP][XR], ov[UP][XR];
        double wv[UP][XR];
#pragma unroll
        for (int u = 0; u < UP; ++u)
#pragma unroll
            for (int k = 0; k < XR; ++k) {
                const bool on = xb + u < ix1 && xown[k];
                const size_t i = (size_t)(xb + u) * S + xoff[k];
                zv[u][k] = on ? z[i] : T(0);
                ov[u][k] = (on && !first) ? pold[i] : T(0);
                wv[u][k] = (DETECT && on) ? w[i] : 0.0;
            }
#pragma unroll
        for (int u = 0; u < UP; ++u)
#pragma unroll
            for (int k = 0; k < XR; ++k)
                if (xb + u < ix1 && xown[k]) {
                    const T val = first ? zv[u][k] : (T)fma(beta, (double)ov[u][k], (double)zv[u][k]);
                    p[(size_t)(xb + u) * S + xoff[k]] = val;
                    if (DETECT) {
                        a_c += wv[u][k] * (double)val;
                        a_xx += (double)val * (double)val;
                    }
                }
    }
    if (DETECT) {
        const double t2 = tree256(sd, 2, a_c), t3 = tree256(sd, 3, a_xx);
        if (tid == 0) { v.tp[2 * v.tp_stride + b] = t2; v.tp[3 * v.tp_stride + b] = t3; }
    }
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(FTY* FTZ) void k_fused_dir(FusedView<T> v) {
    __shared__ double sd[5][NT];
    const int b = xcd_slab_block(v.ntiles);
    if (b >= v.ntiles) return;
    fused_dir_body<T, DETECT>(v, b, sd);
}

// K1b: Ap = P A p through the LDS-tiled walk with asynchronous copies, and
// the partials p.Ap (and, with the detectors, 1.Ap) of each tile (rows 0, 1).
template <typename T, bool DETECT>
__device__ __forceinline__ void fused_tdma_body(const FusedView<T>& v, int b, TdmaHalf<T, false, FTY, FTZ>& H0,
                                                TdmaHalf<T, false, FTY, FTZ>& H1, double (*sd)[NT]) {
    const FusedState* fs = v.fs;
    if (fs->done) return;
    const T* p = (fs->iter & 1) ? v.P1 : v.P0;
    double s_pap = 0.0, s_ap = 0.0;
    tdma_body<T, false, FTY, FTZ, true>(p, v.Ap, v.cr, v.ck, v.g, v.nby, v.nbz, v.nch, b, H0, H1, s_pap, s_ap);
    const double t0 = tree256(sd, 0, s_pap);
    double t1 = 0.0;
    if (DETECT) t1 = tree256(sd, 1, s_ap);
    if (threadIdx.x == 0) {
        v.tp[b] = t0;
        if (DETECT) v.tp[v.tp_stride + b] = t1;
    }
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(FTY* FTZ, 3) void k_fused_tdma(FusedView<T> v) {
    __shared__ TdmaHalf<T, false, FTY, FTZ> H0, H1;
    __shared__ double sd[2][NT];
    const int b = xcd_slab_block(v.ntiles);
    if (b >= v.ntiles) return;
    fused_tdma_body<T, DETECT>(v, b, H0, H1, sd);
}

// K2.  tp: K1 partials.  sp: output super-block partials (r.r, r.z).
// (super-block j of the solve's nsb)
template <typename T, bool DETECT, bool PERSIST = false>
__device__ __forceinline__ void fused_update_body(const FusedView<T>& v, int j) {
    T* __restrict__ x = v.x;
    const T* __restrict__ P0 = v.P0;
    const T* __restrict__ P1 = v.P1;
    T* __restrict__ r = v.r;
    const T* __restrict__ Ap = v.Ap;
    const T* __restrict__ d = v.d;
    T* __restrict__ z = v.z;
    const int nn = v.g.nn;
    const double* __restrict__ tp = v.tp;
    const int ntiles = v.ntiles, tp_stride = v.tp_stride;
    double* __restrict__ sp = v.sp;
    const int sp_stride = v.sp_stride;
    FusedState* fs = v.fs;
    DevState* st = v.st;
    const double nwA = v.nwA;
    __shared__ double sd[4][NT];
    // the host polls this word after a chunk of iterations instead of copying
    // the flag back (a system-scope store to pinned host memory)
    if (v.hflag && j == 0 && threadIdx.x == 0)
        __hip_atomic_store(v.hflag, fs->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (block_done<PERSIST>(fs)) return;
    const int k1 = fs->k1_iter;
    double o[4];
    if (v.pre_tp) { for (int k = 0; k < 4; ++k) o[k] = fs->red_tp[k]; }
    else if (DETECT) reduce_partials<4>(sd, tp, ntiles, tp_stride, o);
    else reduce_partials<1>(sd, tp, ntiles, tp_stride, o);
    const double pAp = o[0];
    // D1 of this iteration's apply is recorded before the breakdown check, as
    // the reference's apply_operator does (:908-919 run before :1013)
    if (DETECT && threadIdx.x == 0 && j == 0) abft_record(st, o[1], o[2], o[3], nwA);
    // breakdown (:1013): the reference's |p.Ap| < 1e-30 for the FP64 solve; an
    // inner FP32 solve of a refined one solves A e = r for a residual r that
    // shrinks with every correction, so its test is scaled by |r|^2
    const double brk = (sizeof(T) == sizeof(float)) ? 1e-30 * fs->bnorm * fs->bnorm : 1e-30;
    if (fabs(pAp) < brk) {
        if (threadIdx.x == 0 && j == 0) fs->done = 1;
        return;
    }
    const double alpha = fs->rho[k1 & 1] / pAp;
    const T* __restrict__ p = (k1 & 1) ? P1 : P0;
    const bool fi_now = fs->fi_armed && (fs->inner_base + k1 == fs->fi_iter);
    const int fi_vec = fi_now ? fs->fi_vec : -1;
    const long long fi_idx = fs->fi_idx;
    const int fi_bit = fs->fi_bit;
    double a_rr = 0.0, a_rz = 0.0;
    const size_t base = (size_t)j * SB + threadIdx.x;
    for (int s = 0; s < SBS; ++s) {
        const size_t i = base + (size_t)s * NT;
        if (i < (size_t)nn) {
            const T xi = (T)fma(alpha, (double)p[i], (double)x[i]);
            const T ri = (T)fma(-alpha, (double)Ap[i], (double)r[i]);
            const T zi = ri / d[i];
            if (fi_vec >= 0 && (long long)i == fi_idx) {
                x[i] = fi_vec == SEU_U ? seu_flip(xi, fi_bit) : xi;
                r[i] = fi_vec == SEU_R ? seu_flip(ri, fi_bit) : ri;
                z[i] = fi_vec == SEU_Z ? seu_flip(zi, fi_bit) : zi;
                if (fi_vec == SEU_U || fi_vec == SEU_R || fi_vec == SEU_Z) fs->fi_fired = 1;
            } else {
                x[i] = xi;
                r[i] = ri;
                z[i] = zi;
            }
            a_rr += (double)ri * (double)ri;
            a_rz += (double)ri * (double)zi;
        }
    }
    const double s0 = tree256(sd, 0, a_rr), s1 = tree256(sd, 1, a_rz);
    if (threadIdx.x == 0) {
        sp[j] = s0;
        sp[sp_stride + j] = s1;
        if (j == 0) {
            fs->alpha = alpha;
            fs->iter = k1 + 1;
            fs->inner_total += 1;
        }
    }
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(NT) void k_fused_update(FusedView<T> v) {
    fused_update_body<T, DETECT>(v, blockIdx.x);
}

// One-block finalize of the super-block partials K1 will read (the rows its
// iteration needs) and of the tile partials K2 will read.
template <typename T>
__device__ __forceinline__ void fused_fin_sp_body(const FusedView<T>& v, double (*sd)[NT]) {
    FusedState* fs = v.fs;
    if (fs->done) return;
    double o[5];
    if (fs->iter == 0) {
        if (fs->check_outer) reduce_partials<5>(sd, v.sp, v.nsb, v.sp_stride, o);
        else reduce_partials<3>(sd, v.sp, v.nsb, v.sp_stride, o);
    } else {
        reduce_partials<2>(sd, v.sp, v.nsb, v.sp_stride, o);
    }
    if (threadIdx.x == 0) {
        const int K = fs->iter == 0 ? (fs->check_outer ? 5 : 3) : 2;
        for (int k = 0; k < K; ++k) fs->red_sp[k] = o[k];
    }
}

template <typename T, bool DETECT>
__device__ __forceinline__ void fused_fin_tp_body(const FusedView<T>& v, double (*sd)[NT]) {
    FusedState* fs = v.fs;
    if (fs->done) return;
    double o[4];
    if (DETECT) reduce_partials<4>(sd, v.tp, v.ntiles, v.tp_stride, o);
    else reduce_partials<1>(sd, v.tp, v.ntiles, v.tp_stride, o);
    if (threadIdx.x == 0)
        for (int k = 0; k < (DETECT ? 4 : 1); ++k) fs->red_tp[k] = o[k];
}

template <typename T>
__global__ __launch_bounds__(NT) void k_fused_fin_sp(FusedView<T> v) {
    __shared__ double sd[5][NT];
    fused_fin_sp_body<T>(v, sd);
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(NT) void k_fused_fin_tp(FusedView<T> v) {
    __shared__ double sd[4][NT];
    fused_fin_tp_body<T, DETECT>(v, sd);
}

// FP64 outer residual of the refined solve from Ax (the projected FP64 apply):
// r64 = b - Ax, inner start e = 0, r32 = (float) r64, z32 = r32 / d32;
// super-block partials [r32.r32, r32.r32, r32.z32, b.b, r64.r64].
__global__ __launch_bounds__(NT) void k_fused_mixed_init(
    const double* __restrict__ b, const double* __restrict__ Ax, const float* __restrict__ d32,
    float* __restrict__ e32, float* __restrict__ r32, float* __restrict__ z32, int nn,
    double* __restrict__ sp, int nsb, FusedState* fs) {
    __shared__ double sd[5][NT];
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const size_t base = (size_t)blockIdx.x * SB + threadIdx.x;
    for (int s = 0; s < SBS; ++s) {
        const size_t i = base + (size_t)s * NT;
        if (i < (size_t)nn) {
            const double bi = b[i];
            const double ri = fma(-1.0, Ax[i], bi);
            const float rf = (float)ri;
            const float zf = rf / d32[i];
            e32[i] = 0.0f;
            r32[i] = rf;
            z32[i] = zf;
            a[0] += (double)rf * (double)rf;
            a[1] += (double)rf * (double)rf;
            a[2] += (double)rf * (double)zf;
            a[3] += bi * bi;
            a[4] += ri * ri;
        }
    }
    double t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = tree256(sd, k, a[k]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) sp[(size_t)k * nsb + blockIdx.x] = t[k];
        if (blockIdx.x == 0) {
            fs->iter = 0;
            fs->done = 0;
            fs->inner_base = fs->inner_total;
        }
    }
}

// ---------------------------------------------------------------------------
// Batches of independent solves (an ensemble of meshes and fields).  Each
// solve keeps the tile and super-block decomposition it has on its own, so
// its sums, and therefore its iterates, are bit-identical to a single solve;
// the batch only packs all solves' tiles (super-blocks) into one grid.  A
// solve that has converged, or is masked, sets its done flag and its blocks
// return at once.
// ---------------------------------------------------------------------------

// i with off[i] <= gidx < off[i + 1] (off ascending, off[0] = 0)
__device__ __forceinline__ int find_solve(const int* __restrict__ off, int count, int gidx) {
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= gidx) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Colour-ordered row sum at any node, boundary included: the contributions of
// the elements that exist around node (ix, iy, iz), in colour order from 0
// (the colour sweep's sum, src/heat_solver_het.cu:896-899).  MASS: dt_eff = 0.
template <typename T, bool MASS>
__device__ __forceinline__ T node_apply_any(const T* __restrict__ x, const T* __restrict__ cr,
                                           const T* __restrict__ ck, const Grid& g, int ix, int iy, int iz) {
    const int n = g.n, nsq = n * n;
    const int px = ix & 1, py = iy & 1, pz = iz & 1;
    const int sx = px ? -1 : 1, sy = py ? -1 : 1, sz = pz ? -1 : 1;
    T m[3][3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int X = ix + (a - 1) * sx, Y = iy + (b - 1) * sy, Z = iz + (c - 1) * sz;
                const bool ok = X >= 0 && X <= n && Y >= 0 && Y <= n && Z >= 0 && Z <= n;
                m[a][b][c] = ok ? x[(size_t)X * g.sx + (size_t)Y * g.sy + Z] : T(0);
            }
    T crc[8], ckc[8];
    int emask = 0;
#pragma unroll
    for (int col = 0; col < 8; ++col) {
        const int cx = (col >> 2) & 1, cy = (col >> 1) & 1, cz = col & 1;
        const int ex = ix - (cx ^ px), ey = iy - (cy ^ py), ez = iz - (cz ^ pz);
        crc[col] = ckc[col] = T(0);
        if (ex < 0 || ex >= n || ey < 0 || ey >= n || ez < 0 || ez >= n) continue;
        const size_t e = (size_t)ex * nsq + (size_t)ey * n + ez;
        crc[col] = cr[e];
        ckc[col] = MASS ? T(0) : ck[e];
        emask |= 1 << col;
    }
    return gather_rows<T, MASS>(m, crc, ckc, emask);
}

// b = P(M T) [+ dt F] for every node of every active solve: build_rhs's
// apply_operator(T, b, 0, false) projects by default
// (src/heat_solver_het.cu:885, 963-971); one thread per node of the batch.
__global__ __launch_bounds__(NT) void k_batch_rhs(const FusedView<double>* __restrict__ views,
                                                  const int* __restrict__ node_off, int count) {
    const int gid = blockIdx.x * NT + threadIdx.x;
    if (gid >= node_off[count]) return;
    const int i = find_solve(node_off, count, gid);
    const FusedView<double>& v = views[i];
    if (!v.active) return;
    const int nid = gid - node_off[i];
    int ix, iy, iz;
    node_coords(v.g, nid, ix, iy, iz);
    const double acc = node_is_boundary(v.g, ix, iy, iz) ? 0.0
                                                         : node_apply_any<double, true>(v.x, v.cr, v.ck, v.g, ix, iy, iz);
    v.b[nid] = v.add_source ? fma(v.dt, v.F[nid], acc) : acc;
}

// PCG start of every solve: r = b - P A x, z = r/d, super-block partials of
// b.b, r.r, r.z (the order of k_fused_init); grid = all super-blocks.
__global__ __launch_bounds__(NT) void k_batch_residual(const FusedView<double>* __restrict__ views,
                                                       const int* __restrict__ sb_off, int count) {
    __shared__ double sd[3][NT];
    if ((int)blockIdx.x >= sb_off[count]) return;
    const int i = find_solve(sb_off, count, blockIdx.x);
    const FusedView<double>& v = views[i];
    const int j = blockIdx.x - sb_off[i];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const size_t base = (size_t)j * SB + threadIdx.x;
    if (v.active) {
        for (int s = 0; s < SBS; ++s) {
            const size_t n = base + (size_t)s * NT;
            if (n < (size_t)v.g.nn) {
                int ix, iy, iz;
                node_coords(v.g, (int)n, ix, iy, iz);
                const double ax = node_is_boundary(v.g, ix, iy, iz)
                                      ? 0.0 : node_apply_any<double, false>(v.x, v.cr, v.ck, v.g, ix, iy, iz);
                const double bi = v.b[n];
                const double ri = fma(-1.0, ax, bi);
                const double zi = ri / v.d[n];
                v.r[n] = ri;
                v.z[n] = zi;
                a0 += bi * bi;
                a1 += ri * ri;
                a2 += ri * zi;
            }
        }
    }
    const double s0 = tree256(sd, 0, a0), s1 = tree256(sd, 1, a1), s2 = tree256(sd, 2, a2);
    if (threadIdx.x == 0) {
        v.sp[j] = s0;
        v.sp[v.sp_stride + j] = s1;
        v.sp[2 * v.sp_stride + j] = s2;
        if (j == 0) {
            v.fs->iter = 0;
            v.fs->done = v.active ? 0 : 1;
        }
    }
}

// the two launches of K1 over a batch's tiles (no upset is armed in a batch);
// T = double: the FP64 solves, T = float: the inner solves of refined ones
template <typename T, bool DETECT>
__global__ __launch_bounds__(FTY* FTZ) void k_batch_dir(const FusedView<T>* __restrict__ views,
                                                        const int* __restrict__ tile_off, int count) {
    __shared__ double sd[5][NT];
    const int total = tile_off[count];
    const int gb = xcd_slab_block(total);
    if (gb >= total) return;
    const int i = find_solve(tile_off, count, gb);
    fused_dir_body<T, DETECT>(views[i], gb - tile_off[i], sd);
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(FTY* FTZ, 3) void k_batch_tdma(const FusedView<T>* __restrict__ views,
                                                            const int* __restrict__ tile_off, int count) {
    __shared__ TdmaHalf<T, false, FTY, FTZ> H0, H1;
    __shared__ double sd[2][NT];
    const int total = tile_off[count];
    const int gb = xcd_slab_block(total);
    if (gb >= total) return;
    const int i = find_solve(tile_off, count, gb);
    fused_tdma_body<T, DETECT>(views[i], gb - tile_off[i], H0, H1, sd);
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(NT) void k_batch_update(const FusedView<T>* __restrict__ views,
                                                     const int* __restrict__ sb_off, int count) {
    if ((int)blockIdx.x >= sb_off[count]) return;
    const int i = find_solve(sb_off, count, blockIdx.x);
    fused_update_body<T, DETECT>(views[i], blockIdx.x - sb_off[i]);
}

// one block per solve (solves with few partials skip: their blocks reduce in place)
template <typename T>
__global__ __launch_bounds__(NT) void k_batch_fin_sp(const FusedView<T>* __restrict__ views) {
    __shared__ double sd[5][NT];
    const FusedView<T>& v = views[blockIdx.x];
    if (v.pre_sp) fused_fin_sp_body<T>(v, sd);
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(NT) void k_batch_fin_tp(const FusedView<T>* __restrict__ views) {
    __shared__ double sd[4][NT];
    const FusedView<T>& v = views[blockIdx.x];
    if (v.pre_tp) fused_fin_tp_body<T, DETECT>(v, sd);
}

// The batched iteration as ONE persistent (cooperative) launch: every block
// stays resident and walks the phases of up to `iters` iterations -- [one-
// block folds of the super-block partials], the direction pass, the operator
// walk, [one-block folds of the tile partials], the update -- over all the
// batch's tiles and super-blocks, a grid-wide barrier between phases where
// the launches end in the graph-replayed form; after each iteration block 0
// decides whether every solve is done (flag[it & 1]) and all blocks read that
// decision after one more barrier, so all leave together.  Same bodies, same
// order of operations, same bits as the launches of batch_chunk_t.
template <typename F>
__device__ __forceinline__ void persist_blocks(int total, F&& f) {
    // tile t of `total` for this block: each XCD (blocks are dealt round-robin
    // to the NXCD XCDs; the grid is a multiple of NXCD) works through one
    // contiguous slab, as xcd_slab_block arranges for a launch of `total` blocks
    const int per = (total + NXCD - 1) / NXCD;
    const int xcd = (int)(blockIdx.x % NXCD), lane = (int)(blockIdx.x / NXCD), lanes = (int)(gridDim.x / NXCD);
    for (int t = lane; t < per; t += lanes) {
        const int vb = xcd * per + t;
        if (vb < total) f(vb);
    }
}

template <typename T, bool DETECT>
__global__ __launch_bounds__(NT) void k_batch_persistent(const FusedView<T>* __restrict__ views,
                                                         const int* __restrict__ tile_off,
                                                         const int* __restrict__ sb_off, int count, int iters,
                                                         int pre_sp, int pre_tp, int* __restrict__ flag,
                                                         int* hflag) {
    static_assert(FTY * FTZ == NT, "one block shape for every phase");
    __shared__ TdmaHalf<T, false, FTY, FTZ> H0, H1;
    __shared__ double sd[5][NT];
    __shared__ int open;
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int total_t = tile_off[count], total_sb = sb_off[count];
    int all = 0;
    for (int it = 0; it < iters; ++it) {
        if (pre_sp) {
            for (int i = blockIdx.x; i < count; i += gridDim.x) {
                if (views[i].pre_sp) fused_fin_sp_body<T>(views[i], sd);
                __syncthreads();
            }
            grid.sync();
        }
        persist_blocks(total_t, [&](int gb) {
            const int i = find_solve(tile_off, count, gb);
            fused_dir_body<T, DETECT, true>(views[i], gb - tile_off[i], sd);
            __syncthreads();
        });
        grid.sync();
        persist_blocks(total_t, [&](int gb) {
            const int i = find_solve(tile_off, count, gb);
            fused_tdma_body<T, DETECT>(views[i], gb - tile_off[i], H0, H1, sd);
            __builtin_amdgcn_s_waitcnt(0);              // no copy into H0/H1 still in flight
            __syncthreads();
        });
        grid.sync();
        if (pre_tp) {
            for (int i = blockIdx.x; i < count; i += gridDim.x) {
                if (views[i].pre_tp) fused_fin_tp_body<T, DETECT>(views[i], sd);
                __syncthreads();
            }
            grid.sync();
        }
        persist_blocks(total_sb, [&](int j) {
            const int i = find_solve(sb_off, count, j);
            fused_update_body<T, DETECT, true>(views[i], j - sb_off[i]);
            __syncthreads();
        });
        grid.sync();
        if (blockIdx.x == 0) {
            if (threadIdx.x == 0) open = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < count; i += NT)
                if (!__hip_atomic_load(&views[i].fs->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    atomicOr(&open, 1);
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(&flag[it & 1], open ? 0 : 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        grid.sync();
        all = __hip_atomic_load(&flag[it & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (all) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && hflag) __hip_atomic_store(hflag, all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- FP32-refined solves of a batch (each solve as heat_pcg_solve_mixed) ----
// v64: the solve's FP64 vectors (x = T, b, cr, ck); v32: its FP32 inner
// solve (x = e, r, z, d, the partial arrays) -- both views share the state.

// per-solve parameters of a refined batch step (one thread per solve)
__global__ void k_batch_params_mixed(const FusedView<double>* __restrict__ v64, int count, double tol,
                                     double tol_inner, int max_inner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    FusedState* fs = v64[i].fs;
    fs->tol = tol_inner;
    fs->max_iter = max_inner;
    fs->check_outer = 1;
    fs->tol_outer = tol;
    fs->inner_total = 0;
    fs->inner_base = 0;
    fs->fi_armed = 0;
    fs->fi_fired = 0;
    fs->outer_done = v64[i].active ? 0 : 1;
    fs->done = v64[i].active ? 0 : 1;
}

// Outer pass start of every refined solve that is still open: the FP64
// residual r = b - P A x at each node (the colour-ordered sum, as the single
// solve's apply), then e = 0, r32, z32 and the five super-block partials of
// k_fused_mixed_init; grid = all super-blocks.
__global__ __launch_bounds__(NT) void k_batch_mixed_init(const FusedView<double>* __restrict__ v64,
                                                         const FusedView<float>* __restrict__ v32,
                                                         const int* __restrict__ sb_off, int count) {
    __shared__ double sd[5][NT];
    if ((int)blockIdx.x >= sb_off[count]) return;
    const int i = find_solve(sb_off, count, blockIdx.x);
    const FusedView<double>& V = v64[i];
    const FusedView<float>& F = v32[i];
    FusedState* fs = V.fs;
    if (fs->outer_done) return;                      // finished (or masked): keep its state
    const int j = blockIdx.x - sb_off[i];
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const size_t base = (size_t)j * SB + threadIdx.x;
    for (int s = 0; s < SBS; ++s) {
        const size_t n = base + (size_t)s * NT;
        if (n < (size_t)V.g.nn) {
            int ix, iy, iz;
            node_coords(V.g, (int)n, ix, iy, iz);
            const double ax = node_is_boundary(V.g, ix, iy, iz)
                                  ? 0.0 : node_apply_any<double, false>(V.x, V.cr, V.ck, V.g, ix, iy, iz);
            const double bi = V.b[n];
            const double ri = fma(-1.0, ax, bi);
            const float rf = (float)ri;
            const float zf = rf / F.d[n];
            F.x[n] = 0.0f;
            F.r[n] = rf;
            F.z[n] = zf;
            a[0] += (double)rf * (double)rf;
            a[1] += (double)rf * (double)rf;
            a[2] += (double)rf * (double)zf;
            a[3] += bi * bi;
            a[4] += ri * ri;
        }
    }
    double t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = tree256(sd, k, a[k]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) F.sp[(size_t)k * F.sp_stride + j] = t[k];
        if (j == 0) {
            fs->iter = 0;
            fs->done = 0;
            fs->inner_base = fs->inner_total;
        }
    }
}

// x += e for every node of every solve still open (k_mixed_correct)
__global__ __launch_bounds__(NT) void k_batch_mixed_correct(const FusedView<double>* __restrict__ v64,
                                                            const FusedView<float>* __restrict__ v32,
                                                            const int* __restrict__ node_off, int count) {
    const int gid = blockIdx.x * NT + threadIdx.x;
    if (gid >= node_off[count]) return;
    const int i = find_solve(node_off, count, gid);
    if (v64[i].fs->outer_done) return;
    const int n = gid - node_off[i];
    v64[i].x[n] += (double)v32[i].x[n];
}

// *flag = 1 when every refined solve of the batch has met the outer tolerance
__global__ void k_batch_outer_done(const FusedView<double>* __restrict__ views, int count, int* flag) {
    __shared__ int open;
    if (threadIdx.x == 0) open = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += blockDim.x)
        if (!views[i].fs->outer_done) atomicOr(&open, 1);
    __syncthreads();
    if (threadIdx.x == 0) *flag = open ? 0 : 1;
}

// per-solve PCG parameters of a batch step (one thread per solve)
__global__ void k_batch_params(const FusedView<double>* __restrict__ views, int count, double tol,
                               int max_iter) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    FusedState* fs = views[i].fs;
    fs->tol = tol;
    fs->max_iter = max_iter;
    fs->check_outer = 0;
    fs->inner_total = 0;
    fs->inner_base = 0;
    fs->outer_done = 0;
    fs->fi_armed = 0;
    fs->fi_fired = 0;
    fs->done = views[i].active ? 0 : 1;
}

// *flag = 1 when every solve of the batch is done (flag: a pinned host word,
// stored at system scope, which the host polls)
template <typename T>
__global__ void k_batch_all_done(const FusedView<T>* __restrict__ views, int count, int* flag) {
    __shared__ int open;
    if (threadIdx.x == 0) open = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += blockDim.x)
        if (!views[i].fs->done) atomicOr(&open, 1);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, open ? 0 : 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace heat
