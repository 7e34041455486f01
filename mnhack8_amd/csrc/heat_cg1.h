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
// FP32 apply) to the plane behind it.  Sums are FP64 sums of FP32 products,
// per thread in plane order (for the vector terms, the thread's two halo-tile
// slots in slot order), then a 256-wide tree per tile; every block of the
// next launch folds the tiles' partials in the same fixed order (or, for
// many tiles, one block does it once).  The CPU oracle restates all of this
// (ho_pcg_solve_mixed_cg1).
#pragma once
#include "heat_pcg_fused.h"

namespace heat {

constexpr int C1TY = 8, C1TZ = 32, C1NT = C1TY * C1TZ;
constexpr int C1XW = C1TZ + 2, C1XN = C1XW * (C1TY + 2);     // u tile with its halo: 34 x 10
constexpr int C1XR = (C1XN + C1NT - 1) / C1NT;               // slots per thread
constexpr int C1EW = C1TZ + 1, C1EN = C1EW * (C1TY + 1);     // element tile: 33 x 9
constexpr int C1ER = (C1EN + C1NT - 1) / C1NT;
constexpr int C1ROWS = 8;                                    // partial rows (see cg1_head)

// chunks of interior planes per (y, z) column: the fused solve's rule on the
// n - 1 interior planes (a function of the mesh alone)
__host__ __device__ inline int cg1_chunks(int n, int ncol) {
    return fused_chunks(n - 1 > 0 ? n - 1 : 1, ncol);
}

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
    int* hflag;                   // pinned host word: done after each launch's decision (or null)
};

// Partial rows: 0 r.u, 1 w.u, 2 r.r of the launch's outputs; with the
// detectors 3 1_int.w, 4 wA.u, 5 u.u (the D1 sums of its apply); the start
// launch adds 6 b.b and 7 r64.r64 (the FP64 residual of the outer pass).
template <bool DETECT, bool START>
constexpr int cg1_rows() { return START ? 8 : (DETECT ? 6 : 3); }

template <int K>
__device__ __forceinline__ void cg1_fold(const Cg1View& v, int par, double (*sd)[NT], double o[K]) {
    if (v.pre) {
#pragma unroll
        for (int k = 0; k < K; ++k) o[k] = v.fs->cg_red[par][k];
    } else {
        reduce_partials<K>(sd, v.tp[par], v.ntiles, v.ntiles, o);
    }
}

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
// of the apply that formed w_i, then alpha_i and beta_i.  False: the solve is
// done (the same decision in every block).
template <bool DETECT>
__device__ __forceinline__ bool cg1_head(const Cg1View& v, int par, double (*sd)[NT], double& alpha, double& beta,
                                         bool& first) {
    FusedState* fs = v.fs;
    const bool b0 = blockIdx.x == 0 && threadIdx.x == 0;
    if (fs->done) {
        if (b0 && v.hflag) __hip_atomic_store(v.hflag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return false;
    }
    const int it = fs->cg_it[par];
    double o[8];
    first = (it == 0);
    if (first && fs->check_outer) cg1_fold<8>(v, par, sd, o);
    else if (DETECT) cg1_fold<6>(v, par, sd, o);
    else cg1_fold<3>(v, par, sd, o);
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
    if (b0) {
        if (first) fs->outer_done = outer;
        if (outer == 3) fs->outer_done = 3;
        if (stop) {
            fs->done = 1;
        } else {
            fs->cg_gamma[par] = o[0];
            fs->cg_alpha[par] = alpha;
            fs->cg_it[par ^ 1] = it + 1;
            fs->iter = it + 1;
            fs->inner_total += 1;
        }
        if (v.hflag) __hip_atomic_store(v.hflag, stop ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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

// The walk of tile b.  START: the pass's first launch (u0, w0 and their
// sums from b and A x; par = 1, outputs in slot 0).  Otherwise step i with
// par = i & 1: inputs r[par], w[par], s[par ^ 1], outputs r[par ^ 1],
// s[par], w[par ^ 1], partials tp[par ^ 1].  Vector updates are FP32 with
// alpha and beta rounded to FP32 once; u = r * (1/d) with the FP32
// reciprocal diagonal dinv.  Each plane's values are loaded two planes ahead.
template <bool DETECT, bool START>
__device__ __forceinline__ void cg1_body(const Cg1View& v, int b, int par, double alpha, double beta, bool first,
                                         float (*us)[C1XN], float (*er)[C1EN], float (*ek)[C1EN],
                                         double (*sd)[NT]) {
    const int out = par ^ 1;
    const Grid g = v.g;
    const int n = g.n, S = g.sx, nn1 = g.nn1, nsq = n * n, m = n - 1;
    const int PB = S * (int)sizeof(float);                 // bytes of an FP32 node plane
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
    // byte offsets of this thread's slots: the u tile (interior nodes, else
    // past the end), its own nodes, the element tile, and its stencil node
    unsigned xo[C1XR], xw[C1XR], xd[C1XR];
    bool xown[C1XR];
#pragma unroll
    for (int k = 0; k < C1XR; ++k) {
        const int i = tid + k * C1NT;
        const int r = i / C1XW, c = i - (i / C1XW) * C1XW;
        const int yy = y0 - 1 + r, zz = z0 - 1 + c;
        const bool in = i < C1XN && yy >= 1 && yy <= m && zz >= 1 && zz <= m;
        xown[k] = in && r >= 1 && r <= C1TY && c >= 1 && c <= C1TZ;
        xo[k] = in ? (unsigned)((yy * nn1 + zz) * 4) : C1_OOB;
        xw[k] = xown[k] ? xo[k] : C1_OOB;
        xd[k] = in ? (unsigned)((yy * nn1 + zz) * 8) : C1_OOB;
    }
    unsigned eo[C1ER];
#pragma unroll
    for (int k = 0; k < C1ER; ++k) {
        const int i = tid + k * C1NT;
        const int r = i / C1EW, c = i - (i / C1EW) * C1EW;
        const int ey = y0 - 1 + r, ez = z0 - 1 + c;
        eo[k] = (i < C1EN && ey < n && ez < n) ? (unsigned)((ey * n + ez) * 4) : C1_OOB;
    }
    const unsigned wo = node_in ? (unsigned)((iy * nn1 + iz) * 4) : C1_OOB;
    const unsigned wod = node_in ? (unsigned)((iy * nn1 + iz) * 8) : C1_OOB;
    int eidx[2][2], eoff[2][2];
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
            eidx[cy][cz] = (ty + 1 - (cy ^ py)) * C1EW + (tz + 1 - (cz ^ pz));
            eoff[cy][cz] = (iy - (cy ^ py)) * n + (iz - (cz ^ pz));
        }
    int widx[3][3];
#pragma unroll
    for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
        for (int c = 0; c < 3; ++c) widx[b2][c] = (ty + 1 + (b2 - 1) * sy) * C1XW + (tz + 1 + (c - 1) * sz);
    struct Stage {
        float r[C1XR], w[C1XR], s[C1XR], d[C1XR], e[C1XR], p[C1XR];
        double bb[C1XR], ax[C1XR];
        float cr[C1ER], ck[C1ER];
    };
    // plane P (0..n) of an FP32 node array; planes off the interior (0 and n)
    // read as zeros: a zero-size resource (the walk never stores there)
    auto plane = [&](const float* a, int P) __attribute__((always_inline)) {
        return c1_rsrc(a + (size_t)P * S, (P >= 1 && P <= m) ? PB : 0);
    };
    auto fetch = [&](int P, Stage& St) __attribute__((always_inline)) {
        const __amdgpu_buffer_rsrc_t rd = plane(v.d, P);
#pragma unroll
        for (int k = 0; k < C1XR; ++k) St.d[k] = c1_ldf(rd, xo[k]);
        if (START) {
            const int dbytes = (P >= 1 && P <= m) ? 2 * PB : 0;
            const __amdgpu_buffer_rsrc_t rb = c1_rsrc(v.b + (size_t)P * S, dbytes);
            const __amdgpu_buffer_rsrc_t ra = c1_rsrc(v.Ax + (size_t)P * S, dbytes);
#pragma unroll
            for (int k = 0; k < C1XR; ++k) {
                St.bb[k] = c1_ldd(rb, xd[k]);
                St.ax[k] = c1_ldd(ra, xd[k]);
            }
        } else {
            const __amdgpu_buffer_rsrc_t rr = plane(v.r[par], P), rw = plane(v.w[par], P), rs = plane(v.s[out], P);
            const __amdgpu_buffer_rsrc_t re = plane(v.e, P), rp = plane(v.p, P);
#pragma unroll
            for (int k = 0; k < C1XR; ++k) {
                St.r[k] = c1_ldf(rr, xo[k]);
                St.w[k] = c1_ldf(rw, xo[k]);
                St.s[k] = c1_ldf(rs, xo[k]);
                St.e[k] = c1_ldf(re, xw[k]);
                St.p[k] = c1_ldf(rp, xw[k]);
            }
        }
        // element plane P - 1 (none below plane 1: a zero-size resource)
        const int ep = P - 1;
        const bool epin = ep >= 0 && ep < n;
        const __amdgpu_buffer_rsrc_t rcr = c1_rsrc(v.cr + (size_t)(epin ? ep : 0) * nsq, epin ? nsq * 4 : 0);
        const __amdgpu_buffer_rsrc_t rck = c1_rsrc(v.ck + (size_t)(epin ? ep : 0) * nsq, epin ? nsq * 4 : 0);
#pragma unroll
        for (int k = 0; k < C1ER; ++k) {
            St.cr[k] = c1_ldf(rcr, eo[k]);
            St.ck[k] = c1_ldf(rck, eo[k]);
        }
    };
    double a_g = 0.0, a_rr = 0.0, a_bb = 0.0, a_r64 = 0.0;
    double a_dl = 0.0, a_1w = 0.0, a_c = 0.0, a_uu = 0.0;
    const float af = (float)alpha, bf = (float)beta;
    // u of plane P (zero off the interior), own nodes updated and stored
    auto form = [&](int P, int slot, const Stage& St) __attribute__((always_inline)) {
        const bool own = P >= ix0 && P < ix1;           // uniform over the block
        const __amdgpu_buffer_rsrc_t ro = plane(v.r[out], P);
        const __amdgpu_buffer_rsrc_t eo2 = plane(v.e, P);
#pragma unroll
        for (int k = 0; k < C1XR; ++k) {
            const int i = tid + k * C1NT;
            float un;
            if (START) {
                const double r64 = fma(-1.0, St.ax[k], St.bb[k]);
                const float rf = (float)r64;
                un = rf * St.d[k];
                if (own) {
                    c1_stf(rf, ro, xw[k]);
                    c1_stf(0.0f, eo2, xw[k]);
                    if (xown[k]) {
                        a_g += (double)rf * (double)un;
                        a_rr += (double)rf * (double)rf;
                        a_bb += St.bb[k] * St.bb[k];
                        a_r64 += r64 * r64;
                    }
                }
            } else {
                const float sn = first ? St.w[k] : fmaf(bf, St.s[k], St.w[k]);
                const float rn = fmaf(-af, sn, St.r[k]);
                un = rn * St.d[k];
                if (own) {
                    const float uo = St.r[k] * St.d[k];
                    const float pn = first ? uo : fmaf(bf, St.p[k], uo);
                    c1_stf(fmaf(af, pn, St.e[k]), eo2, xw[k]);
                    c1_stf(pn, plane(v.p, P), xw[k]);
                    c1_stf(sn, plane(v.s[par], P), xw[k]);
                    c1_stf(rn, ro, xw[k]);
                    if (xown[k]) {
                        a_g += (double)rn * (double)un;
                        a_rr += (double)rn * (double)rn;
                    }
                }
            }
            if (i < C1XN) us[slot][i] = un;
        }
#pragma unroll
        for (int k = 0; k < C1ER; ++k) {
            const int i = tid + k * C1NT;
            if (i < C1EN) {
                er[slot][i] = St.cr[k];
                ek[slot][i] = St.ck[k];
            }
        }
    };
    float W[3][3][3];
    Coef<float> Eprev[2][2], Enew[2][2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
            for (int c = 0; c < 3; ++c) W[a][b2][c] = 0.0f;
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz] = Coef<float>{0.0f, 0.0f};
    const int par_yz = (py << 1) | pz;
    const FiArm nofi{0, 0, 0};
    // one incoming plane P: form it, load plane P + 2 into the same stage,
    // then the stencil of plane P - 1
    auto step = [&](int P, Stage& St) __attribute__((always_inline)) {
        const int slot = P & 1;
        form(P, slot, St);
        if (P + 2 <= ix1) fetch(P + 2, St);
        __syncthreads();
#pragma unroll
        for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
            for (int c = 0; c < 3; ++c) W[2][b2][c] = us[slot][widx[b2][c]];
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) {
                Enew[cy][cz].cr = er[slot][eidx[cy][cz]];
                Enew[cy][cz].ck = ek[slot][eidx[cy][cz]];
            }
        const int X = P - 1;                            // the plane whose stencil is now complete
        if (X >= ix0) {                                 // uniform over the block
            const float wn = (X & 1) ? march_node<float, false, 1>(W, Eprev, Enew, par_yz, X, eoff, nsq, nofi)
                                     : march_node<float, false, 0>(W, Eprev, Enew, par_yz, X, eoff, nsq, nofi);
            // u is 0 at a position outside the mesh, so such a lane adds exact zeros
            const float uc = W[1][1][1];
            c1_stf(wn, plane(v.w[out], X), wo);
            const double wd = (double)wn, ud = (double)uc;
            a_dl += wd * ud;
            if (DETECT) {
                const double wa = c1_ldd(c1_rsrc(v.wA + (size_t)X * S, 2 * PB), wod);
                a_1w += node_in ? wd : 0.0;
                a_c += wa * ud;
                a_uu += ud * ud;
            }
        }
#pragma unroll
        for (int b2 = 0; b2 < 3; ++b2)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                W[0][b2][c] = W[1][b2][c];
                W[1][b2][c] = W[2][b2][c];
            }
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz) Eprev[cy][cz] = Enew[cy][cz];
    };
    Stage SA, SB;
    fetch(ix0 - 1, SA);
    fetch(ix0, SB);
    for (int P = ix0 - 1; P <= ix1; P += 2) {
        step(P, SA);
        if (P + 1 <= ix1) step(P + 1, SB);
    }
    double* tpo = v.tp[out];
    const int nt = v.ntiles;
    const double t0 = tree256(sd, 0, a_g), t1 = tree256(sd, 1, a_dl), t2 = tree256(sd, 2, a_rr);
    double t3 = 0.0, t4 = 0.0, t5 = 0.0, t6 = 0.0, t7 = 0.0;
    if (DETECT || START) {
        t3 = tree256(sd, 3, a_1w);
        t4 = tree256(sd, 4, a_c);
        t5 = tree256(sd, 5, a_uu);
    }
    if (START) {
        t6 = tree256(sd, 6, a_bb);
        t7 = tree256(sd, 7, a_r64);
    }
    if (tid == 0) {
        tpo[b] = t0;
        tpo[nt + b] = t1;
        tpo[2 * nt + b] = t2;
        if (DETECT || START) {
            tpo[3 * nt + b] = t3;
            tpo[4 * nt + b] = t4;
            tpo[5 * nt + b] = t5;
        }
        if (START) {
            tpo[6 * nt + b] = t6;
            tpo[7 * nt + b] = t7;
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
__global__ __launch_bounds__(C1NT) void k_cg1_start(Cg1View v) {
    __shared__ float us[2][C1XN];
    __shared__ float er[2][C1EN], ek[2][C1EN];
    __shared__ double sd[C1ROWS][NT];
    const int b = xcd_slab_block(v.ntiles);
    if (b == 0 && threadIdx.x == 0) {
        FusedState* fs = v.fs;
        fs->iter = 0;
        fs->done = 0;
        fs->cg_it[0] = 0;
        fs->inner_base = fs->inner_total;
    }
    if (b >= v.ntiles) return;
    cg1_body<DETECT, true>(v, b, 1, 0.0, 0.0, true, us, er, ek, sd);
}

// Step i of the inner solve (par = i & 1).
template <bool DETECT>
__global__ __launch_bounds__(C1NT) void k_cg1_step(Cg1View v, int par) {
    __shared__ float us[2][C1XN];
    __shared__ float er[2][C1EN], ek[2][C1EN];
    __shared__ double sd[C1ROWS][NT];
    const int b = xcd_slab_block(v.ntiles);
    if (b >= v.ntiles) return;
    double alpha, beta;
    bool first;
    if (!cg1_head<DETECT>(v, par, sd, alpha, beta, first)) return;
    cg1_body<DETECT, false>(v, b, par, alpha, beta, first, us, er, ek, sd);
}

}  // namespace heat
