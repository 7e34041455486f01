// heat_cg1.h -- the inner FP32 solve of the refined solve in ONE launch per
// iteration: preconditioned CG in the Chronopoulos-Gear arrangement, whose two
// inner products of an iteration are taken together, so one global reduction
// (a kernel boundary) per iteration suffices.  With u = D^-1 r and w = A u:
//
//   start:   r0 = (float)(b - A x) [FP64 residual], u0 = r0 / d, w0 = A u0
//   step i:  beta_i  = gamma_i / gamma_{i-1}              (beta_0 = 0)
//            alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1})
//                                                         (alpha_0 = gamma_0 / delta_0)
//            p_i = u_i + beta_i p_{i-1};   s_i = w_i + beta_i s_{i-1}
//            e  += alpha_i p_i;            r_{i+1} = r_i - alpha_i s_i
//            u_{i+1} = r_{i+1} / d;        w_{i+1} = A u_{i+1}
//            gamma_{i+1} = r_{i+1}.u_{i+1}, delta_{i+1} = w_{i+1}.u_{i+1}, |r_{i+1}|^2
//
// In exact arithmetic these are the iterates of the standard PCG
// (src/heat_solver_het.cu:994-1039): s_i = A p_i and gamma_i / (p_i . A p_i)
// is alpha_i.  One launch does step i for a tile of nodes: the stencil
// w_{i+1} = A u_{i+1} needs u_{i+1} on the tile's one-node halo, which the
// tile forms itself from r_i, w_i, s_{i-1} and d there (the neighbours form
// the same values for their own nodes), so r, s and w are double-buffered
// by iteration parity and no tile reads a value another tile of the same
// launch writes; e and p are read and written only at a tile's own nodes.
//
// Tiles: 8 x 32 interior nodes in (y, z) by a chunk of interior x planes
// (cg1_chunks); a block of 256 threads walks its chunk plane by plane,
// forming u of the incoming plane (tile and halo) in registers, parking it
// in LDS, and applying the colour-ordered stencil (the same sums as the
// FP32 apply) to the plane behind it.  The steps take their inputs from an
// LDS ring filled by asynchronous copies; the start launch loads its FP64
// inputs per thread.  Thread t owns tile node (t / 32,
// t % 32): it updates that node's vectors, forms its stencil and sums its
// terms; threads 0..83 also form u at one node of the tile's halo ring.
// Sums are FP64 sums of FP32 products, per thread in plane order, then a
// 256-wide tree per tile; every block of the next launch folds the tiles'
// partials in the same fixed order (or, for many tiles, one block does it
// once).  The CPU oracle restates all of this (ho_pcg_solve_mixed_cg1).
#pragma once
#include "heat_pcg_fused.h"

namespace heat {

constexpr int C1TY = 8, C1TZ = 32, C1NT = C1TY * C1TZ;
constexpr int C1XW = C1TZ + 2, C1XN = C1XW * (C1TY + 2);     // u tile with its halo: 34 x 10
constexpr int C1HN = C1XN - C1NT;                            // halo ring: 84 nodes
static_assert(C1HN <= C1NT, "one halo node per thread at most");
constexpr int C1EW = C1TZ + 1, C1EN = C1EW * (C1TY + 1);     // element tile: 33 x 9
constexpr int C1ER = (C1EN + C1NT - 1) / C1NT;
constexpr int C1ROWS = 8;                                    // partial rows (see cg1_head)
constexpr int C1WAVE4 = 1024;                                // tiles per wave at 4 resident step blocks per CU

// chunks of interior planes per (y, z) column: the fused solve's rule on the
// n - 1 interior planes (a function of the mesh and the geometry alone)
__host__ __device__ inline int cg1_chunks(int n, int ncol, int planes = FCX, int wave = FWAVE) {
    return fused_chunks(n - 1 > 0 ? n - 1 : 1, ncol, planes, wave);
}

// The pinned status word every step's head stores: the pass (mod 1024), the
// step of the pass and its status (0 running, 1 the inner solve ended, 2 the
// outer test ended the solve, 3 a launch out of step).
constexpr int CG1_MAX_STEP = (1 << 20) - 1;
__host__ __device__ constexpr int cg1_flag(int pass, int step, int status) {
    return ((pass & 1023) << 22) | ((step & CG1_MAX_STEP) << 2) | status;
}
__host__ __device__ constexpr int cg1_flag_pass(int f) { return (f >> 22) & 1023; }
__host__ __device__ constexpr int cg1_flag_step(int f) { return (f >> 2) & CG1_MAX_STEP; }
__host__ __device__ constexpr int cg1_flag_status(int f) { return f & 3; }

struct Cg1View {
    float *e, *p;                 // iterate (the correction) and direction
    float *r[2], *s[2], *w[2];    // by iteration parity
    const float *d, *cr, *ck;     // d: the reciprocal diagonal 1/d in FP32
    const double *b, *Ax;         // start: FP64 right-hand side and A x
    const double* wA;             // D1: (M + dt K) 1_int
    Grid g;
    double* tp[2];                // tile partials by parity: C1ROWS rows of ntiles
    int ntiles, nby, nbz, nch;
    int pre;                      // 1: the partials come folded in fs->cg_red
    FusedState* fs;
    DevState* st;
    double nwA;
    int* hflag;                   // pinned host word: cg1_flag at each step's head (or null)
};

// Partial rows: 0 r.u, 1 w.u, 2 r.r of the launch's outputs; with the
// detectors 3 1_int.w, 4 wA.u, 5 u.u (the D1 sums of its apply); the start
// launch adds 6 b.b and 7 r64.r64 (the FP64 residual of the outer pass).
template <bool DETECT, bool START>
constexpr int cg1_rows() { return START ? 8 : (DETECT ? 6 : 3); }

// One-block fold of the partials of parity PAR into fs->cg_red[par] (many tiles).
template <bool DETECT>
__global__ __launch_bounds__(NT) void k_cg1_fin(Cg1View v, int par, int start) {
    __shared__ double sd[C1ROWS][NT];
    if (v.fs->done) return;
    double o[C1ROWS];
    if (start) reduce_partials<8>(sd, v.tp[par], v.ntiles, v.ntiles, o);
    else if (DETECT) reduce_partials<6>(sd, v.tp[par], v.ntiles, v.ntiles, o);
    else reduce_partials<3>(sd, v.tp[par], v.ntiles, v.ntiles, o);
    if (threadIdx.x == 0) {
        const int K = start ? 8 : (DETECT ? 6 : 3);
        for (int k = 0; k < K; ++k) v.fs->cg_red[par][k] = o[k];
    }
}

// Head of step i (parity par): the tests in the order of the standard solve
// (outer residual at i = 0, then convergence and the iteration limit), D1
// of the apply that formed w_i, then alpha_i and beta_i.  x: this thread's
// first four tile partials of the first KX rows (partials_load4), loaded
// before the head so that their round trip overlaps the walk's first loads;
// for many tiles (pre) k_cg1_fin has folded them into fs->cg_red.  False:
// the solve is done (the same decision in every block).
template <bool DETECT, int KX>
__device__ __forceinline__ bool cg1_head(const Cg1View& v, int par, double (*sd)[NT], const double (&x)[KX][4],
                                         double& alpha, double& beta, bool& first) {
    FusedState* fs = v.fs;
    const bool b0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (fs->done) {
        if (b0 && v.hflag)
            __hip_atomic_store(v.hflag, cg1_flag(fs->cg_pass, fs->iter, fs->cg_status), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        return false;
    }
    const int it = fs->cg_it[par];
    double o[8];
    first = (it == 0);
    constexpr int KS = DETECT ? 6 : 3;
    const bool need8 = first && fs->check_outer;
    const double* part = v.tp[par];
    const int nt = v.ntiles;
    if (v.pre) {
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = fs->cg_red[par][k];
    } else if (need8) {
        if constexpr (KX >= 8) partials_fold<8, KX>(sd, x, part, nt, nt, o);
        else reduce_partials<8>(sd, part, nt, nt, o);     // a first step launched without the 8-row load
    } else {
        if constexpr (KX >= KS) partials_fold<KS, KX>(sd, x, part, nt, nt, o);
        else reduce_partials<KS>(sd, part, nt, nt, o);
    }
    if (DETECT && b0) abft_record(v.st, o[3], o[4], o[5], v.nwA);
    bool stop = false;
    double den = 0.0;
    int outer = 0;
    if ((it & 1) != par) {
        stop = true;                       // a launch out of step with the state: stop, reported by the host
        outer = 3;
    } else if (first) {
        if (fs->check_outer) {
            double bo = sqrt(o[6]);
            if (bo < 1e-30) bo = 1.0;
            const double rn = sqrt(o[7]);
            outer = (rn / bo < fs->tol_outer) ? 1 : (isfinite(rn) ? 0 : 2);
            if (b0) { fs->bnorm_outer = bo; fs->rnorm_outer = rn; }
        }
        double bn = sqrt(o[2]);
        if (bn < 1e-30) bn = 1.0;
        if (b0) fs->bnorm = bn;
        beta = 0.0;
        den = o[1];
        stop = outer != 0 || fs->max_iter <= 0 || fabs(den) < 1e-30 * bn * bn;
    } else {
        const double bn = fs->bnorm;
        const bool conv = sqrt(o[2]) / bn < fs->tol;
        beta = o[0] / fs->cg_gamma[par ^ 1];
        den = fma(-beta, o[0] / fs->cg_alpha[par ^ 1], o[1]);
        stop = conv || it >= fs->max_iter || fabs(den) < 1e-30 * bn * bn;
    }
    alpha = stop ? 0.0 : o[0] / den;
    const int status = !stop ? 0 : (outer == 3 ? 3 : (outer != 0 ? 2 : 1));
    if (b0) {
        if (first) fs->outer_done = outer;
        if (outer == 3) fs->outer_done = 3;
        if (stop) {
            fs->done = 1;
            fs->cg_status = status;
        } else {
            fs->cg_gamma[par] = o[0];
            fs->cg_alpha[par] = alpha;
            fs->cg_it[par ^ 1] = it + 1;
            fs->iter = it + 1;
            fs->inner_total += 1;
        }
        if (v.hflag)
            __hip_atomic_store(v.hflag, cg1_flag(fs->cg_pass, it, status), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return !stop;
}

// Raw buffer access: an offset past the resource's size reads 0 and drops a
// store, so the walk needs no branch around a load or store (its wait counts
// stay exact).  C1_OOB is such an offset for every resource here.
constexpr unsigned C1_OOB = 0x7ffffff0u;
constexpr int C1_RSRC3 = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t c1_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, C1_RSRC3);
}
__device__ __forceinline__ float c1_ldf(__amdgpu_buffer_rsrc_t r, unsigned off) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
#else
    return 0.0f;
#endif
}
__device__ __forceinline__ double c1_ldd(__amdgpu_buffer_rsrc_t r, unsigned off) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
#else
    return 0.0;
#endif
}
__device__ __forceinline__ void c1_stf(float v, __amdgpu_buffer_rsrc_t r, unsigned off) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
#endif
}

// The start launch's walk of tile b: the pass's first launch forms r0, u0
// and w0 and their sums from b and A x (par = 1, outputs in slot 0); u =
// r * (1/d) with the FP32 reciprocal diagonal dinv.  (Step i runs the same
// walk with its inputs copied to LDS, cg1_step_ring below.)
//
// Every load and store of the walk is issued on every path: one that falls
// off the tile's planes or the mesh gets an out-of-range offset instead of a
// branch (a load reads 0, a store is dropped).  So the memory-counter waits
// are the same on every path, and each plane's loads, issued one plane
// ahead, are waited for only when that plane is formed.
template <bool DETECT>
__device__ __forceinline__ void cg1_start_body(const Cg1View& v, int b, float (*us)[C1XN],
                                               float (*er)[C1EN], float (*ek)[C1EN], double (*sd)[NT]) {
    constexpr int par = 1, out = 0;
    const Grid g = v.g;
    const int n = g.n, S = g.sx, nn1 = g.nn1, nsq = n * n, m = n - 1;
    const int nby = v.nby, nbz = v.nbz;
    const int bz = b % nbz, by = (b / nbz) % nby, chunk = b / (nbz * nby);
    const int ix0 = 1 + (int)(((long long)chunk * m) / v.nch);
    const int ix1 = 1 + (int)(((long long)(chunk + 1) * m) / v.nch);
    const int tid = threadIdx.x;
    const int ty = tid / C1TZ, tz = tid - (tid / C1TZ) * C1TZ;
    const int y0 = 1 + by * C1TY, z0 = 1 + bz * C1TZ;
    const int iy = y0 + ty, iz = z0 + tz;
    const bool node_in = iy <= m && iz <= m;
    const int py = iy & 1, pz = iz & 1;
    const int sy = py ? -1 : 1, sz = pz ? -1 : 1;
    // this thread's two slots of the u tile: slot 0 its own node (ty, tz),
    // slot 1 one node of the halo ring (threads below C1HN: rows 0 and 9,
    // then columns 0 and 33 of rows 1..8); per slot the in-plane node offset
    // (elements), the LDS index and whether the node is interior.  Every
    // array is one buffer resource, and a slot's byte offset is formed once
    // per plane (C1_OOB off the interior)
    int xq[2], xi[2];
    bool xin[2];
    xq[0] = iy * nn1 + iz;
    xi[0] = (ty + 1) * C1XW + (tz + 1);
    xin[0] = node_in;
    {
        int r, c;
        if (tid < 2 * C1XW) {
            r = tid < C1XW ? 0 : C1TY + 1;
            c = tid < C1XW ? tid : tid - C1XW;
        } else {
            const int k = tid - 2 * C1XW;
            r = 1 + (k >> 1);
            c = (k & 1) ? C1XW - 1 : 0;
        }
        const int yy = y0 - 1 + r, zz = z0 - 1 + c;
        xq[1] = yy * nn1 + zz;
        xi[1] = r * C1XW + c;
        xin[1] = tid < C1HN && yy >= 1 && yy <= m && zz >= 1 && zz <= m;
    }
    int eq[C1ER];
    bool eok[C1ER];
#pragma unroll
    for (int k = 0; k < C1ER; ++k) {
        const int i = tid + k * C1NT;
        const int r = i / C1EW, c = i - (i / C1EW) * C1EW;
        const int ey = y0 - 1 + r, ez = z0 - 1 + c;
        eok[k] = i < C1EN && ey < n && ez < n;
        eq[k] = ey * n + ez;
    }
    const int wq = xq[0];
    int eidx[2][2];
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) eidx[cy][cz] = (ty + 1 - (cy ^ py)) * C1EW + (tz + 1 - (cz ^ pz));
    int widx[3][3];
#pragma unroll
    for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
        for (int c = 0; c < 3; ++c) widx[b2][c] = (ty + 1 + (b2 - 1) * sy) * C1XW + (tz + 1 + (c - 1) * sz);
    struct Stage {
        float d[2];                                     // by slot
        double bb[2], ax[2];
        float cr[C1ER], ck[C1ER];
        double wa;                                      // D1 weight at the own node of the plane behind
    };
    // whole-array resources (the host keeps every byte offset below 2^31)
    const int NB = g.nn * 4, EB = g.ne * 4;
    const __amdgpu_buffer_rsrc_t R_d = c1_rsrc(v.d, NB), R_e = c1_rsrc(v.e, NB);
    const __amdgpu_buffer_rsrc_t R_rout = c1_rsrc(v.r[out], NB), R_wout = c1_rsrc(v.w[out], NB);
    const __amdgpu_buffer_rsrc_t R_cr = c1_rsrc(v.cr, EB), R_ck = c1_rsrc(v.ck, EB);
    const __amdgpu_buffer_rsrc_t R_b = c1_rsrc(v.b, 2 * NB), R_ax = c1_rsrc(v.Ax, 2 * NB);
    const __amdgpu_buffer_rsrc_t R_wA = c1_rsrc(v.wA, 2 * NB);
    // plane P's loads: u-tile slots (planes 0 and n, and planes past the
    // walk, read as zeros), element plane P - 1 and, with the detectors, the
    // D1 weight of plane P - 1's own node
    auto fetch = [&](int P, Stage& St) __attribute__((always_inline)) {
        const bool pin = P >= 1 && P <= m && P <= ix1;
        const int pq = P * S;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const unsigned o4 = (pin && xin[k]) ? (unsigned)(pq + xq[k]) * 4u : C1_OOB;
            const unsigned o8 = (pin && xin[k]) ? (unsigned)(pq + xq[k]) * 8u : C1_OOB;
            St.d[k] = c1_ldf(R_d, o4);
            St.bb[k] = c1_ldd(R_b, o8);
            St.ax[k] = c1_ldd(R_ax, o8);
        }
        const int ep = P - 1;
        const bool epin = ep >= 0 && ep < n && P <= ix1;
        const int eqp = (epin ? ep : 0) * nsq;
#pragma unroll
        for (int k = 0; k < C1ER; ++k) {
            const unsigned o = (epin && eok[k]) ? (unsigned)(eqp + eq[k]) * 4u : C1_OOB;
            St.cr[k] = c1_ldf(R_cr, o);
            St.ck[k] = c1_ldf(R_ck, o);
        }
        if (DETECT) {
            const bool xact = ep >= ix0 && ep < ix1;
            St.wa = c1_ldd(R_wA, (xact && node_in) ? (unsigned)(ep * S + wq) * 8u : C1_OOB);
        }
    };
    // one stage one plane ahead (it holds the FP64 right-hand side and A x)
    Stage SA;
    fetch(ix0 - 1, SA);
    double a_g = 0.0, a_rr = 0.0, a_bb = 0.0, a_r64 = 0.0;
    double a_dl = 0.0, a_1w = 0.0, a_c = 0.0, a_uu = 0.0;
    // u of plane P (zero off the interior) at both slots; the own node
    // updated, stored and summed
    auto form = [&](int P, int slot, const Stage& St) __attribute__((always_inline)) {
        const bool own = P >= ix0 && P < ix1;           // uniform over the block
        const bool mine = own && node_in;
        const unsigned ow = mine ? (unsigned)(P * S + xq[0]) * 4u : C1_OOB;
        const double r64 = fma(-1.0, St.ax[0], St.bb[0]);
        const float rf = (float)r64;
        const float u0 = rf * St.d[0];
        c1_stf(rf, R_rout, ow);
        c1_stf(0.0f, R_e, ow);
        a_g += mine ? (double)rf * (double)u0 : 0.0;
        a_rr += mine ? (double)rf * (double)rf : 0.0;
        a_bb += mine ? St.bb[0] * St.bb[0] : 0.0;
        a_r64 += mine ? r64 * r64 : 0.0;
        // the halo node: the same u its own tile forms
        const float u1 = (float)fma(-1.0, St.ax[1], St.bb[1]) * St.d[1];
        us[slot][xi[0]] = u0;
        if (tid < C1HN) us[slot][xi[1]] = u1;
#pragma unroll
        for (int k = 0; k < C1ER; ++k) {
            const int i = tid + k * C1NT;
            if (i < C1EN) {
                er[slot][i] = St.cr[k];
                ek[slot][i] = St.ck[k];
            }
        }
    };
    PlaneT<float> PT[3];                            // terms of the u window planes X-1, X, X+1
    float Xc[3];                                    // u at the own node of each window plane
    Coef<float> Eprev[2][2], Enew[2][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        Xc[a] = 0.0f;
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) PT[a].F[cy][cz] = PT[a].DC[cy][cz] = PT[a].C4[cy][cz] = 0.0f;
    }
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz] = Coef<float>{0.0f, 0.0f};
    // one incoming plane P: form it, load plane P + 1 into the same stage,
    // then the stencil of plane P - 1 (stored and summed if it is one of the
    // chunk's planes)
    constexpr int AHEAD = 1;
    auto step = [&](int P, Stage& St) __attribute__((always_inline)) {
        const int slot = P & 1;
        form(P, slot, St);
        const double wa = DETECT ? St.wa : 0.0;
        fetch(P + AHEAD, St);
        __syncthreads();
        {
            float Wn[3][3];
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
                for (int c = 0; c < 3; ++c) Wn[b2][c] = us[slot][widx[b2][c]];
            plane_terms(Wn, PT[2]);
            Xc[2] = Wn[1][1];
        }
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) {
                Enew[cy][cz].cr = er[slot][eidx[cy][cz]];
                Enew[cy][cz].ck = ek[slot][eidx[cy][cz]];
            }
        const int X = P - 1;                            // the plane whose stencil is now complete
        const bool xact = X >= ix0 && X < ix1;          // uniform over the block
        const float wn = (X & 1) ? node_rows<float, false, 1>(PT[0], PT[1], PT[2], Eprev, Enew)
                                 : node_rows<float, false, 0>(PT[0], PT[1], PT[2], Eprev, Enew);
        // u is 0 at a position outside the mesh, so such a lane adds exact zeros
        const float uc = Xc[1];
        c1_stf(wn, R_wout, (xact && node_in) ? (unsigned)(X * S + wq) * 4u : C1_OOB);
        const double wd = (double)wn, ud = (double)uc;
        a_dl += xact ? wd * ud : 0.0;
        if (DETECT) {
            a_1w += (xact && node_in) ? wd : 0.0;
            a_c += xact ? wa * ud : 0.0;
            a_uu += xact ? ud * ud : 0.0;
        }
        PT[0] = PT[1];
        PT[1] = PT[2];
        Xc[0] = Xc[1];
        Xc[1] = Xc[2];
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz];
    };
    // planes ix0 - 1 .. ix1
    for (int P = ix0 - 1; P <= ix1; ++P) step(P, SA);
    double t[8];
    const double a8[8] = {a_g, a_dl, a_rr, a_1w, a_c, a_uu, a_bb, a_r64};
    tree_rows<8>(sd, a8, t);
    if (tid == 0) {
        double* tpo = v.tp[out];
        const int nt = v.ntiles;
#pragma unroll
        for (int k = 0; k < 8; ++k) tpo[k * nt + b] = t[k];
    }
}

// ---------------------------------------------------------------------------
// The step walk fed by asynchronous copies.  Step i of the walk above (the
// same u, stencil, vector updates and sums, in the same order per thread),
// with the tile's inputs brought to LDS by 16-byte buffer-to-LDS copies
// instead of per-thread 4-byte loads into registers: per plane slot, r, w,
// s and 1/d over the u tile with its halo (10 rows of 34 nodes, rows padded
// to 36 = nine copies) and cr, ck of the element plane behind it (9 rows of
// 33, padded to 36).  Two planes form a batch; while a batch is formed, the
// next batch's copies are in flight into the other half of a two-half ring
// (one wait and one barrier per batch, as in tdma_body).  Wave k copies
// array k of the four node arrays, waves 0 and 1 also cr and ck; a copy
// past the plane reads zeros, and positions outside the mesh or off the
// walk's planes are zeroed by selects where they are read, so every value
// the arithmetic sees is the one the register walk loads.  e and p (own
// nodes) and the D1 weight stay per-thread loads two planes ahead.
// ---------------------------------------------------------------------------
constexpr int C1RW = 36;                                     // padded row of a ring tile (floats)
constexpr int C1RN = (C1TY + 2) * C1RW;                      // node tile: 10 rows
constexpr int C1RE = (C1TY + 1) * C1RW;                      // element tile: 9 rows
constexpr int C1UN = C1RN / 4, C1UE = C1RE / 4;              // 16-byte copies per tile: 90, 81
static_assert(C1UN <= 128 && C1UE <= 128, "two copy instructions per wave and array");

struct Cg1Half {
    float nd[2][4][C1RN];     // [plane of the batch][r, w, s, 1/d]
    float ce[2][2][C1RE];     // [plane of the batch][cr, ck]
};

template <bool DETECT, bool FOLD8>
__device__ __forceinline__ void cg1_step_ring(const Cg1View& v, int b, int par, Cg1Half (&H)[2],
                                              float (*us)[C1XN], double (*sd)[NT]) {
    const int out = par ^ 1;
    const Grid g = v.g;
    const int n = g.n, S = g.sx, nn1 = g.nn1, nsq = n * n, m = n - 1;
    const int nby = v.nby, nbz = v.nbz;
    const int bz = b % nbz, by = (b / nbz) % nby, chunk = b / (nbz * nby);
    const int ix0 = 1 + (int)(((long long)chunk * m) / v.nch);
    const int ix1 = 1 + (int)(((long long)(chunk + 1) * m) / v.nch);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);      // uniform: copy targets are per wave
    const int ty = tid / C1TZ, tz = tid - (tid / C1TZ) * C1TZ;
    const int y0 = 1 + by * C1TY, z0 = 1 + bz * C1TZ;
    const int iy = y0 + ty, iz = z0 + tz;
    const bool node_in = iy <= m && iz <= m;
    const int py = iy & 1, pz = iz & 1;
    const int sy = py ? -1 : 1, sz = pz ? -1 : 1;
    constexpr int KX = FOLD8 ? 8 : (DETECT ? 6 : 3);
    double xpart[KX][4];
    partials_load4<KX>(v.tp[par], v.pre ? 0 : v.ntiles, v.ntiles, xpart);
    // the two u-tile slots of this thread (own node; one halo-ring node for
    // tid < C1HN): in-plane node offset, ring index, LDS u index, interior
    int xq0 = iy * nn1 + iz;
    const int ri0 = (ty + 1) * C1RW + (tz + 1);
    const int xi0 = (ty + 1) * C1XW + (tz + 1);
    int ri1, xi1;
    bool xin1;
    {
        int r, c;
        if (tid < 2 * C1XW) {
            r = tid < C1XW ? 0 : C1TY + 1;
            c = tid < C1XW ? tid : tid - C1XW;
        } else {
            const int k = tid - 2 * C1XW;
            r = 1 + (k >> 1);
            c = (k & 1) ? C1XW - 1 : 0;
        }
        const int yy = y0 - 1 + r, zz = z0 - 1 + c;
        ri1 = r * C1RW + c;
        xi1 = r * C1XW + c;
        xin1 = tid < C1HN && yy >= 1 && yy <= m && zz >= 1 && zz <= m;
    }
    int eidx[2][2];
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) eidx[cy][cz] = (ty + 1 - (cy ^ py)) * C1RW + (tz + 1 - (cz ^ pz));
    int widx[3][3];
#pragma unroll
    for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
        for (int c = 0; c < 3; ++c) widx[b2][c] = (ty + 1 + (b2 - 1) * sy) * C1XW + (tz + 1 + (c - 1) * sz);
    // this wave's copy units: node array `wave` (two instructions), and for
    // waves 0 and 1 element array `wave`; byte offsets in a plane, -1 = none
    int no[2], eo[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int u = 64 * j + lane;
        const int r = u / 9, c4 = u - (u / 9) * 9;
        no[j] = u < C1UN ? ((y0 - 1 + r) * nn1 + (z0 - 1) + 4 * c4) * 4 : -1;
        eo[j] = (u < C1UE && wave < 2) ? ((y0 - 1 + r) * n + (z0 - 1) + 4 * c4) * 4 : -1;
    }
    const float* narr = wave == 0 ? v.r[par] : (wave == 1 ? v.w[par] : (wave == 2 ? v.s[out] : v.d));
    const float* earr = wave == 0 ? v.cr : v.ck;
    // copies of the batch starting at plane P0 into half h: node planes P0,
    // P0 + 1 (those in 1..m and at most ix1) and element planes P0 - 1, P0
    // (those in 0..n-1, for planes at most ix1)
    auto issue = [&](Cg1Half& Hh, int P0) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int P = P0 + q;
            if (P >= 1 && P <= m && P <= ix1) {
                const __amdgpu_buffer_rsrc_t R = c1_rsrc(narr + (size_t)P * S, S * 4);
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (no[j] >= 0)
                        lds_copy16(R, (lds_void*)(reinterpret_cast<char*>(&Hh.nd[q][wave][0]) + 16 * 64 * j), no[j]);
            }
            const int ep = P - 1;
            if (wave < 2 && ep >= 0 && ep < n && P <= ix1) {
                const __amdgpu_buffer_rsrc_t R = c1_rsrc(earr + (size_t)ep * nsq, nsq * 4);
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (eo[j] >= 0)
                        lds_copy16(R, (lds_void*)(reinterpret_cast<char*>(&Hh.ce[q][wave][0]) + 16 * 64 * j), eo[j]);
            }
        }
    };
    // per-thread loads two planes ahead: e and p of the own node of plane
    // P and, with the detectors, the D1 weight of plane P - 1's own node
    const int NB = g.nn * 4;
    const __amdgpu_buffer_rsrc_t R_e = c1_rsrc(v.e, NB), R_p = c1_rsrc(v.p, NB);
    const __amdgpu_buffer_rsrc_t R_rout = c1_rsrc(v.r[out], NB), R_sout = c1_rsrc(v.s[par], NB);
    const __amdgpu_buffer_rsrc_t R_wout = c1_rsrc(v.w[out], NB);
    const __amdgpu_buffer_rsrc_t R_wA = c1_rsrc(v.wA, 2 * NB);
    struct Own {
        float e, p;
        double wa;
    };
    auto fetch_own = [&](int P, Own& St) __attribute__((always_inline)) {
        const bool pown = P >= ix0 && P < ix1;
        const unsigned ow = (pown && node_in) ? (unsigned)(P * S + xq0) * 4u : C1_OOB;
        St.e = c1_ldf(R_e, ow);
        St.p = c1_ldf(R_p, ow);
        if (DETECT) {
            const int ep = P - 1;
            const bool xact = ep >= ix0 && ep < ix1;
            St.wa = c1_ldd(R_wA, (xact && node_in) ? (unsigned)(ep * S + xq0) * 8u : C1_OOB);
        }
    };
    issue(H[0], ix0 - 1);
    Own SA, SB;
    fetch_own(ix0 - 1, SA);
    fetch_own(ix0, SB);
    double alpha = 0.0, beta = 0.0;
    bool first = true;
    if (!cg1_head<DETECT, KX>(v, par, sd, xpart, alpha, beta, first)) {
        __builtin_amdgcn_s_waitcnt(0);     // no copy outlives the block
        return;
    }
    double a_g = 0.0, a_rr = 0.0, a_dl = 0.0, a_1w = 0.0, a_c = 0.0, a_uu = 0.0;
    const float af = (float)alpha, bf = (float)beta;
    PlaneT<float> PT[3];
    float Xc[3];
    Coef<float> Eprev[2][2], Enew[2][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        Xc[a] = 0.0f;
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) PT[a].F[cy][cz] = PT[a].DC[cy][cz] = PT[a].C4[cy][cz] = 0.0f;
    }
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz] = Coef<float>{0.0f, 0.0f};
    // plane P from slot q of half Hh: u formed (tile and halo), the own
    // node's vectors updated, stored and summed; then the stencil of plane
    // P - 1, stored and summed if it is one of the chunk's planes
    auto step = [&](const Cg1Half& Hh, int q, int P, Own& St) __attribute__((always_inline)) {
        const bool pin = P >= 1 && P <= m && P <= ix1;             // uniform over the block
        const bool own = P >= ix0 && P < ix1;
        const bool mine = own && node_in;
        const unsigned ow = mine ? (unsigned)(P * S + xq0) * 4u : C1_OOB;
        const bool in0 = pin && node_in, in1 = pin && xin1;
        const float r0 = in0 ? Hh.nd[q][0][ri0] : 0.0f, w0 = in0 ? Hh.nd[q][1][ri0] : 0.0f;
        const float s0 = in0 ? Hh.nd[q][2][ri0] : 0.0f, d0 = in0 ? Hh.nd[q][3][ri0] : 0.0f;
        const float sn = first ? w0 : fmaf(bf, s0, w0);
        const float rn = fmaf(-af, sn, r0);
        const float u0 = rn * d0;
        const float uo = r0 * d0;
        const float pn = first ? uo : fmaf(bf, St.p, uo);
        c1_stf(fmaf(af, pn, St.e), R_e, ow);
        c1_stf(pn, R_p, ow);
        c1_stf(sn, R_sout, ow);
        c1_stf(rn, R_rout, ow);
        a_g += mine ? (double)rn * (double)u0 : 0.0;
        a_rr += mine ? (double)rn * (double)rn : 0.0;
        const int slot = P & 1;
        us[slot][xi0] = u0;
        if (tid < C1HN) {
            const float r1 = in1 ? Hh.nd[q][0][ri1] : 0.0f, w1 = in1 ? Hh.nd[q][1][ri1] : 0.0f;
            const float s1 = in1 ? Hh.nd[q][2][ri1] : 0.0f, d1 = in1 ? Hh.nd[q][3][ri1] : 0.0f;
            const float sh = first ? w1 : fmaf(bf, s1, w1);
            us[slot][xi1] = fmaf(-af, sh, r1) * d1;
        }
        const double wa = DETECT ? St.wa : 0.0;
        fetch_own(P + 2, St);
        __syncthreads();
        {
            float Wn[3][3];
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
                for (int c = 0; c < 3; ++c) Wn[b2][c] = us[slot][widx[b2][c]];
            plane_terms(Wn, PT[2]);
            Xc[2] = Wn[1][1];
        }
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) {
                Enew[cy][cz].cr = Hh.ce[q][0][eidx[cy][cz]];
                Enew[cy][cz].ck = Hh.ce[q][1][eidx[cy][cz]];
            }
        const int X = P - 1;
        const bool xact = X >= ix0 && X < ix1;
        const float wn = (X & 1) ? node_rows<float, false, 1>(PT[0], PT[1], PT[2], Eprev, Enew)
                                 : node_rows<float, false, 0>(PT[0], PT[1], PT[2], Eprev, Enew);
        const float uc = Xc[1];
        const bool xmine = xact && node_in;
        c1_stf(wn, R_wout, xmine ? (unsigned)(X * S + xq0) * 4u : C1_OOB);
        // positions outside the mesh hold u = 0; their stencil values (which
        // may read coefficients beyond the mesh) are never summed
        const double wd = (double)wn, ud = (double)uc;
        a_dl += xmine ? wd * ud : 0.0;
        if (DETECT) {
            a_1w += xmine ? wd : 0.0;
            a_c += xmine ? wa * ud : 0.0;
            a_uu += xmine ? ud * ud : 0.0;
        }
        PT[0] = PT[1];
        PT[1] = PT[2];
        Xc[0] = Xc[1];
        Xc[1] = Xc[2];
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz];
    };
    // planes ix0 - 1 .. ix1 (and one more when their count is odd), two per batch
    int h = 0;
    for (int P = ix0 - 1; P <= ix1; P += 2) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (P + 2 <= ix1) issue(H[h ^ 1], P + 2);
        step(H[h], 0, P, SA);
        step(H[h], 1, P + 1, SB);
        h ^= 1;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();                       // every thread is done with the ring (sd shares half 1)
    double t[6];
    if (DETECT) {
        const double a6[6] = {a_g, a_dl, a_rr, a_1w, a_c, a_uu};
        tree_rows<6>(sd, a6, t);
    } else {
        const double a3[3] = {a_g, a_dl, a_rr};
        tree_rows<3>(sd, a3, t);
    }
    if (tid == 0) {
        double* tpo = v.tp[out];
        const int nt = v.ntiles;
        tpo[b] = t[0];
        tpo[nt + b] = t[1];
        tpo[2 * nt + b] = t[2];
        if (DETECT) {
            tpo[3 * nt + b] = t[3];
            tpo[4 * nt + b] = t[4];
            tpo[5 * nt + b] = t[5];
        }
    }
}

// dinv = 1 / d in FP32 (the preconditioner of the one-launch iteration)
__global__ void k_recip(float* __restrict__ dinv, const float* __restrict__ d, int nn) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i < nn) dinv[i] = 1.0f / d[i];
}

// The pass's first launch: r0, u0, w0 and their sums (outputs of parity 0).
template <bool DETECT>
__global__ __launch_bounds__(C1NT) void k_cg1_start(Cg1View v, int pass) {
    __shared__ float us[2][C1XN];
    __shared__ float er[2][C1EN], ek[2][C1EN];
    __shared__ double sd[C1ROWS][NT];
    const int b = xcd_slab_block(v.ntiles);
    if (b == 0 && threadIdx.x == 0) {
        FusedState* fs = v.fs;
        fs->iter = 0;
        fs->done = 0;
        fs->cg_it[0] = 0;
        fs->cg_pass = pass;
        fs->cg_status = 0;
        fs->inner_base = fs->inner_total;
    }
    if (b >= v.ntiles) return;
    cg1_start_body<DETECT>(v, b, us, er, ek, sd);
}

// Step i of the inner solve (par = i & 1); FOLD8: the pass's first step.
template <bool DETECT, bool FOLD8>
__global__ __launch_bounds__(C1NT) void k_cg1_step(Cg1View v, int par) {
    __shared__ Cg1Half H[2];
    __shared__ float us[2][C1XN];
    // the reduction rows (the head may fold all eight partial rows) share
    // ring half 1: the head runs while only half 0 has copies in flight,
    // and the final tree after the last wait and a barrier
    static_assert(sizeof(Cg1Half) >= sizeof(double) * C1ROWS * NT, "reduction rows fit in a ring half");
    double (*sd)[NT] = reinterpret_cast<double (*)[NT]>(&H[1]);
    const int b = xcd_slab_block(v.ntiles);
    if (b >= v.ntiles) return;
    cg1_step_ring<DETECT, FOLD8>(v, b, par, H, us, sd);
}

}  // namespace heat
