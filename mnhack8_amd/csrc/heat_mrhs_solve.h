// heat_mrhs_solve.h -- sixteen right-hand sides on one mesh solved together:
// FP32 Jacobi-PCG on all sixteen columns at once, with the matrix-core
// multi-RHS apply (heat_mrhs.h) as its operator, inside FP64 defect
// correction per column.
//
// It is the refined solve of heat_pcg_solve_mixed (FP32 operator and inner
// PCG to the inner tolerance; FP64 outer residual b - A x, FP64 alpha / beta
// inner products of FP32 values, FP64 D1 comparison) for sixteen columns,
// each with its own alpha, beta, convergence and outer state:
//   outer pass: r = b - P A x in FP64 (the colour-ordered sum), e = 0,
//     r32 = (float) r, z32 = r32 / d, p = z; stop a column whose FP64
//     residual meets the outer tolerance (or is not finite);
//   inner iteration: Ap = P A p (matrix cores; p.Ap [and the D1 sums] per
//     tile in the same kernel); alpha;
//     r -= alpha Ap; r.r, r.z with z = r / d; convergence, beta;
//     e += alpha p and p = r / d + beta p in one pass (z is formed again there
//     from r and d, not stored, and p is read once for both: two node vectors
//     of 16 columns less to move per iteration);
//   then x += e for every column still open.
// Vectors are node-major, right-hand-side-minor (V[node * 16 + c]), the
// layout of the multi-RHS apply; x and b stay the batch members' FP64 vectors.
// Inner products: each thread of a vector pass owns one column and a stride
// of nodes, a block sums its 16 node lanes per column, and one block per
// column folds the blocks' partials in a fixed order (deterministic).
#pragma once
#include "heat_pcg_fused.h"

namespace heat {

constexpr int MRC = 16;              // columns (= MR)
constexpr int MRB = 2048;            // blocks of a vector pass
constexpr int MRK = 4;               // partial rows per column

struct MrState {
    double rz[MRC], alpha[MRC], beta[MRC], bnorm[MRC];
    double bnorm_outer[MRC], rnorm_outer[MRC], d1max[MRC];
    int done[MRC], outer_done[MRC], iter[MRC], inner_total[MRC];
    int stepped[MRC];                // 1: the column took this iteration's step (e += alpha p still due)
    double tol, tol_outer, nwA;
    int max_inner, all_done, all_outer_done, pad;
};

// per-column sums of one block: thread t holds column t & 15; after the
// tree, thread c < 16 writes partial row k of column c (layout [k][c][MRB])
template <int K>
__device__ __forceinline__ void mr_block_partials(double (*sd)[NT], const double (&a)[K], double* __restrict__ part) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) sd[k][t] = a[k];
    __syncthreads();
#pragma unroll
    for (int st = NT / 2; st >= MRC; st >>= 1) {
        if (t < st) {
#pragma unroll
            for (int k = 0; k < K; ++k) sd[k][t] += sd[k][t + st];
        }
        __syncthreads();
    }
    if (t < MRC) {
#pragma unroll
        for (int k = 0; k < K; ++k) part[((size_t)k * MRC + t) * MRB + blockIdx.x] = sd[k][t];
    }
}

// one block per column: the COUNT partials of rows 0..K-1 (layout
// [k][c][count]) folded (thread t sums t, t + 256, ... in order, then a
// 256-wide tree); all threads get the sums
template <int K>
__device__ __forceinline__ void mr_column_sums(double (*sd)[NT], const double* __restrict__ part, int c,
                                               double out[K], int count = MRB) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double a = 0.0;
        for (int i = t; i < count; i += NT) a += part[((size_t)k * MRC + c) * count + i];
        sd[k][t] = a;
    }
    __syncthreads();
#pragma unroll
    for (int st = NT / 2; st > 0; st >>= 1) {
        if (t < st) {
#pragma unroll
            for (int k = 0; k < K; ++k) sd[k][t] += sd[k][t + st];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = sd[k][0];
}

// Outer pass start: FP64 residual of every open column at every node, the
// FP32 inner start vectors (e = 0, r32, p = z32 = r32 / d) and the sums
// [r32.r32, r32.z32, b.b, r.r].
__global__ __launch_bounds__(NT) void k_mr_init(const FusedView<double>* __restrict__ views, const double* __restrict__ cr,
                                                const double* __restrict__ ck, const float* __restrict__ d32, Grid g,
                                                float* __restrict__ E, float* __restrict__ R,
                                                float* __restrict__ P, double* __restrict__ part,
                                                const MrState* __restrict__ st) {
    __shared__ double sd[MRK][NT];
    const int c = threadIdx.x & (MRC - 1);
    double a[MRK] = {0.0, 0.0, 0.0, 0.0};
    if (!st->outer_done[c]) {
        const double* __restrict__ x = views[c].x;
        const double* __restrict__ b = views[c].b;
        for (int n = blockIdx.x * (NT / MRC) + (threadIdx.x >> 4); n < g.nn; n += MRB * (NT / MRC)) {
            int ix, iy, iz;
            node_coords(g, n, ix, iy, iz);
            const double ax = node_is_boundary(g, ix, iy, iz) ? 0.0
                                                              : node_apply_any<double, false>(x, cr, ck, g, ix, iy, iz);
            const double bi = b[n];
            const double ri = fma(-1.0, ax, bi);
            const float rf = (float)ri;
            const float zf = rf / d32[n];
            const size_t i = (size_t)n * MRC + c;
            E[i] = 0.0f;
            R[i] = rf;
            P[i] = zf;
            a[0] += (double)rf * (double)rf;
            a[1] += (double)rf * (double)zf;
            a[2] += bi * bi;
            a[3] += ri * ri;
        }
    }
    mr_block_partials<MRK>(sd, a, part);
}

__global__ __launch_bounds__(NT) void k_mr_fin_init(const double* __restrict__ part, MrState* st) {
    __shared__ double sd[MRK][NT];
    const int c = blockIdx.x;
    double o[MRK];
    mr_column_sums<MRK>(sd, part, c, o);
    if (threadIdx.x == 0) {
        st->iter[c] = 0;
        if (st->outer_done[c]) {
            st->done[c] = 1;
            return;
        }
        double bn = sqrt(o[0]);
        if (bn < 1e-30) bn = 1.0;
        double bo = sqrt(o[2]);
        if (bo < 1e-30) bo = 1.0;
        const double rn = sqrt(o[3]);
        st->bnorm[c] = bn;
        st->rz[c] = o[1];
        st->bnorm_outer[c] = bo;
        st->rnorm_outer[c] = rn;
        // 1: the FP64 residual meets the outer tolerance; 2: it is not finite
        const int od = (rn / bo < st->tol_outer) ? 1 : (isfinite(rn) ? 0 : 2);
        st->outer_done[c] = od;
        st->done[c] = (od || st->max_inner <= 0) ? 1 : 0;
    }
}

// part: the matrix-core apply's tile partials [p.Ap | 1.Ap | w.p | p.p] x ntiles
template <bool DETECT>
__global__ __launch_bounds__(NT) void k_mr_fin_pap(const double* __restrict__ part, int ntiles, MrState* st) {
    __shared__ double sd[MRK][NT];
    const int c = blockIdx.x;
    double o[MRK] = {0.0, 0.0, 0.0, 0.0};
    if (DETECT) mr_column_sums<MRK>(sd, part, c, o, ntiles);
    else mr_column_sums<1>(sd, part, c, o, ntiles);
    if (threadIdx.x == 0) st->stepped[c] = 0;
    if (threadIdx.x == 0 && !st->done[c]) {
        if (DETECT) {
            // D1 of this apply, recorded before the breakdown check (as abft_record)
            const double rel = fabs(o[1] - o[2]) / (st->nwA * sqrt(o[3]) + 1e-300);
            const double cur = st->d1max[c];
            if (!isnan(cur) && (isnan(rel) || rel > cur)) st->d1max[c] = rel;   // a NaN stays
        }
        // breakdown, scaled by the squared norm of the column's inner right-hand
        // side (the FP64 residual, which shrinks with every correction)
        if (fabs(o[0]) < 1e-30 * st->bnorm[c] * st->bnorm[c]) {
            st->done[c] = 1;
            st->alpha[c] = 0.0;
        } else {
            st->alpha[c] = st->rz[c] / o[0];
            st->stepped[c] = 1;
        }
    }
}

// r -= alpha Ap of every open column; sums r.r, r.z (z = r / d)
__global__ __launch_bounds__(NT) void k_mr_update(float* __restrict__ R, const float* __restrict__ AP,
                                                  const float* __restrict__ d32, Grid g,
                                                  double* __restrict__ part, const MrState* __restrict__ st) {
    __shared__ double sd[2][NT];
    const int c = threadIdx.x & (MRC - 1);
    double a[2] = {0.0, 0.0};
    if (!st->done[c]) {
        const double alpha = st->alpha[c];
        for (int n = blockIdx.x * (NT / MRC) + (threadIdx.x >> 4); n < g.nn; n += MRB * (NT / MRC)) {
            const size_t i = (size_t)n * MRC + c;
            const float ri = (float)fma(-alpha, (double)AP[i], (double)R[i]);
            const float zi = ri / d32[n];
            R[i] = ri;
            a[0] += (double)ri * (double)ri;
            a[1] += (double)ri * (double)zi;
        }
    }
    mr_block_partials<2>(sd, a, part);
}

__global__ __launch_bounds__(NT) void k_mr_fin_update(const double* __restrict__ part, MrState* st) {
    __shared__ double sd[2][NT];
    const int c = blockIdx.x;
    double o[2];
    mr_column_sums<2>(sd, part, c, o);
    if (threadIdx.x == 0 && !st->done[c]) {
        const int it = st->iter[c] + 1;
        st->iter[c] = it;
        st->inner_total[c] += 1;
        const double rn = sqrt(o[0]);
        if (rn / st->bnorm[c] < st->tol || it >= st->max_inner) {
            st->done[c] = 1;
        } else {
            st->beta[c] = o[1] / st->rz[c];
            st->rz[c] = o[1];
        }
    }
}

// e += alpha p of every column that took this iteration's step, then
// p = z + beta p of every column still open (z = r / d formed as k_mr_update
// forms it: the same float division of the same values)
__global__ __launch_bounds__(NT) void k_mr_dir(float* __restrict__ E, const float* __restrict__ R,
                                               const float* __restrict__ d32, float* __restrict__ P, Grid g,
                                               const MrState* __restrict__ st) {
    const int c = threadIdx.x & (MRC - 1);
    if (!st->stepped[c]) return;
    const bool open = !st->done[c];
    const double alpha = st->alpha[c], beta = st->beta[c];
    for (int n = blockIdx.x * (NT / MRC) + (threadIdx.x >> 4); n < g.nn; n += MRB * (NT / MRC)) {
        const size_t i = (size_t)n * MRC + c;
        const float pi = P[i];
        E[i] = (float)fma(alpha, (double)pi, (double)E[i]);
        if (open) {
            const float zi = R[i] / d32[n];
            P[i] = (float)fma(beta, (double)pi, (double)zi);
        }
    }
}

__global__ void k_mr_flags(MrState* st) {
    if (threadIdx.x != 0) return;
    int all = 1, all_outer = 1;
    for (int c = 0; c < MRC; ++c) {
        if (!st->done[c]) all = 0;
        if (!st->outer_done[c]) all_outer = 0;
    }
    st->all_done = all;
    st->all_outer_done = all_outer;
}

// x += e of every column still open
__global__ __launch_bounds__(NT) void k_mr_correct(const FusedView<double>* __restrict__ views,
                                                   const float* __restrict__ E, Grid g,
                                                   const MrState* __restrict__ st) {
    const int c = threadIdx.x & (MRC - 1);
    if (st->outer_done[c]) return;
    double* __restrict__ x = views[c].x;
    for (int n = blockIdx.x * (NT / MRC) + (threadIdx.x >> 4); n < g.nn; n += MRB * (NT / MRC))
        x[n] += (double)E[(size_t)n * MRC + c];
}

// per-step parameters (one thread)
__global__ void k_mr_params(MrState* st, double tol, double tol_inner, int max_inner, double nwA) {
    st->tol = tol_inner;
    st->tol_outer = tol;
    st->max_inner = max_inner;
    st->nwA = nwA;
    for (int c = 0; c < MRC; ++c) {
        st->outer_done[c] = 0;
        st->done[c] = 0;
        st->stepped[c] = 0;
        st->inner_total[c] = 0;
        st->iter[c] = 0;
        st->d1max[c] = 0.0;
        st->alpha[c] = 0.0;
        st->beta[c] = 0.0;
    }
    st->all_done = 0;
    st->all_outer_done = 0;
}

}  // namespace heat
