/*
 * heat_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the matrix-free
 * Q1 hexahedral backward-Euler heat solve with Jacobi-PCG, ABFT detectors and
 * SEU injection.  A plain-C restatement of src/heat_solver_het.cu for one rank
 * (nprocs = 1: every node is owned, the halo exchange is a no-op,
 * src/heat_solver_het.cu:225, 828).  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py use it, as the checker.  Built with
 * -ffp-contract=off so every rounding is the one written here.
 *
 * Parity pins (checked by tests/test_golden_pins.py and tests/test_oracle_math.py
 * against the reference's committed outputs under tests/baseline_validation/v10
 * and against closed forms): V0 eigenmode errors
 * (11 digits), per-step CG counts and D1/D2/D3 traces of the N=32 lognormal
 * run, fault classes/detected_by/first-firing steps of the G and R campaign,
 * the V1 quadrants norm, and the closed forms of the reference matrices.
 * ho_proto_run restates the MNHack7 prototype run (src/heat_solver.cu: the
 * quadrant material at each quadrature point) for V1; every printed step of
 * its recorded N=15 run (CG count and ||T|| to 7 digits) is pinned.
 */
#include "heat_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* grid, src/heat_solver_het.cu:155-167, 202-227                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n;            /* cells per axis */
    double h;
    int nn1;          /* nodes per axis */
    long nn, ne;
    long sx, sy;      /* node strides (x slowest, z fastest) */
    /* tile geometry of the GPU library's fused order (fused_chunks): target
     * planes per chunk and whole-wave fill, [0] the fused iteration, [1] the
     * one-launch inner iteration */
    int geo_planes[2], geo_wave[2];
} ho_grid;

typedef struct { int level, step; long target; int bit, rank, armed; } ho_fi;

struct ho_op {
    ho_grid g;
    double dt;
    int kernel, colored, verify, no_source, source_mode;
    int proto;                  /* MNHack7 prototype: material of each quadrature point (ho_proto_run) */
    double source_value;
    double* k;
    double* rhoc;
    double Mh[64], Kh[64];
    double m3, k2;              /* closed-form M(d=3) = V/216 and K(d=2) = -h/12 */
    double* diag;
    double* F;
    int detect;
    double chk_tol;
    double *ones, *wM, *wA;
    double norm_wM, norm_wA;
    double abft_step_max;
    long abft_step_calls;
    ho_fi fi;
    double kmin, kmax, rhocmin, rhocmax;
    double* partial;
    long nblk;
    double *r, *z, *p, *Ap;     /* PCG work vectors */
    int nthreads;
    int pcg_order;              /* 0: the reference's sums, 1: the GPU library's fused order */
    /* one-shot PCG-state upset of the fused solves: vector, node, bit, iteration */
    int sfi_armed, sfi_vec, sfi_bit, sfi_iter, sfi_fired;
    long sfi_idx;
    int mixed_status;           /* how the last refined solve ended (ho_mixed_status) */
};

static char g_err[512];
const char* ho_last_error(void) { return g_err; }

static const int OFF[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},
                              {0,0,1},{1,0,1},{1,1,1},{0,1,1}};   /* :261-263 */
static int LOC[2][2][2];   /* inverse of OFF */

static void init_loc(void) {
    for (int i = 0; i < 8; ++i) LOC[OFF[i][0]][OFF[i][1]][OFF[i][2]] = i;
}

static inline long lnid(const ho_grid* g, int ix, int iy, int iz) {      /* :203-205 */
    return (long)ix * g->sx + (long)iy * g->sy + iz;
}
static inline int is_boundary(const ho_grid* g, int ix, int iy, int iz) { /* :216-220 */
    return ix == 0 || ix == g->n || iy == 0 || iy == g->n || iz == 0 || iz == g->n;
}
static inline int node_is_boundary(const ho_grid* g, long nid) {
    int ix = (int)(nid / g->sx), iy = (int)((nid / g->sy) % g->nn1), iz = (int)(nid % g->nn1);
    return is_boundary(g, ix, iy, iz);
}

/* color_count / resolve_element, src/heat_solver_het.cu:274-304 */
static long color_count(const ho_grid* g, int color) {
    int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    long a = (g->n - cx + 1) / 2, b = (g->n - cy + 1) / 2, c = (g->n - cz + 1) / 2;
    return a * b * c;
}
static void resolve_element(const ho_grid* g, long tid, int color,
                            long* e, int* ix, int* iy, int* iz) {
    if (color < 0) {
        *e = tid;
        *ix = (int)(tid / ((long)g->n * g->n));
        *iy = (int)((tid / g->n) % g->n);
        *iz = (int)(tid % g->n);
        return;
    }
    int cx = (color >> 2) & 1, cy = (color >> 1) & 1, cz = color & 1;
    long nyb = (g->n - cy + 1) / 2, nzc = (g->n - cz + 1) / 2;
    int a = (int)(tid / (nyb * nzc)), b = (int)((tid / nzc) % nyb), c = (int)(tid % nzc);
    *ix = 2 * a + cx; *iy = 2 * b + cy; *iz = 2 * c + cz;
    *e = (long)(*ix) * g->n * g->n + (long)(*iy) * g->n + *iz;
}
static void element_nodes(const ho_grid* g, int ix, int iy, int iz, long gid[8]) {
    for (int i = 0; i < 8; ++i)
        gid[i] = lnid(g, ix + OFF[i][0], iy + OFF[i][1], iz + OFF[i][2]);
}

/* ------------------------------------------------------------------------- */
/* shape functions and reference matrices, src/heat_solver_het.cu:242-257,   */
/* 661-708                                                                   */
/* ------------------------------------------------------------------------- */
static void shape_func_grad(double xi, double eta, double zeta, double N[8], double dN[8][3]) {
    static const int sgn[8][3] = {{-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},
                                  {-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1}};
    for (int i = 0; i < 8; ++i) {
        double sx = sgn[i][0], sy = sgn[i][1], sz = sgn[i][2];
        double a = 1 + sx * xi, b = 1 + sy * eta, c = 1 + sz * zeta;
        N[i] = 0.125 * a * b * c;
        dN[i][0] = 0.125 * sx * b * c;
        dN[i][1] = 0.125 * a * sy * c;
        dN[i][2] = 0.125 * a * b * sz;
    }
}

static int build_reference_matrices(ho_op* op, int verbose) {
    const double hx = op->g.h, hy = op->g.h, hz = op->g.h;
    double Mh[64] = {0}, Kh[64] = {0};
    const double detJ = hx * hy * hz / 8.0;
    const double iJ[3] = {2.0 / hx, 2.0 / hy, 2.0 / hz};
    const double gp = 1.0 / 1.7320508075688772;
    const double pts[2] = {-gp, gp};
    for (int a = 0; a < 2; ++a) for (int b = 0; b < 2; ++b) for (int c = 0; c < 2; ++c) {
        double N[8], dN[8][3];
        shape_func_grad(pts[a], pts[b], pts[c], N, dN);
        for (int i = 0; i < 8; ++i) for (int j = 0; j < 8; ++j) {
            Mh[i * 8 + j] += detJ * N[i] * N[j];
            Kh[i * 8 + j] += detJ * (iJ[0] * dN[i][0] * iJ[0] * dN[j][0]
                                   + iJ[1] * dN[i][1] * iJ[1] * dN[j][1]
                                   + iJ[2] * dN[i][2] * iJ[2] * dN[j][2]);
        }
    }
    static const int sgn[8][3] = {{-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},
                                  {-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1}};
    const double V = hx * hy * hz;
    double errM = 0, sumM = 0, errKrow = 0, errKdiag = 0, errSym = 0;
    const double Kdiag = (hy * hz / hx + hx * hz / hy + hx * hy / hz) / 9.0;
    for (int i = 0; i < 8; ++i) {
        double rs = 0;
        for (int j = 0; j < 8; ++j) {
            int d = (sgn[i][0] != sgn[j][0]) + (sgn[i][1] != sgn[j][1]) + (sgn[i][2] != sgn[j][2]);
            double Mref = V / 216.0 * (double)(8 >> d);
            errM = fmax(errM, fabs(Mh[i * 8 + j] - Mref) / (V / 216.0 * 8));
            sumM += Mh[i * 8 + j];
            rs += Kh[i * 8 + j];
            errSym = fmax(errSym, fabs(Kh[i * 8 + j] - Kh[j * 8 + i]) + fabs(Mh[i * 8 + j] - Mh[j * 8 + i]));
        }
        errKrow = fmax(errKrow, fabs(rs) / Kdiag);
        errKdiag = fmax(errKdiag, fabs(Kh[i * 9] - Kdiag) / Kdiag);
    }
    double errSum = fabs(sumM - V) / V;
    int ok = (errM < 1e-12 && errSum < 1e-12 && errKrow < 1e-12 &&
              errKdiag < 1e-12 && errSym < 1e-12 * Kdiag);
    if (verbose)
        printf("Reference matrices: M pattern %.1e, sum(M)-V %.1e, K rowsum %.1e, "
               "K diag %.1e, symmetry %.1e  -> %s\n",
               errM, errSum, errKrow, errKdiag, errSym, ok ? "OK" : "FAIL");
    memcpy(op->Mh, Mh, sizeof(Mh));
    memcpy(op->Kh, Kh, sizeof(Kh));
    /* closed forms of the cube's distinct entries (used by the STRUCT rows) */
    op->m3 = (op->g.h * op->g.h * op->g.h) / 216.0;
    op->k2 = -op->g.h / 12.0;
    return ok;
}

/* ------------------------------------------------------------------------- */
/* element kernels                                                           */
/* ------------------------------------------------------------------------- */
/* src/heat_solver_het.cu:340-349 */
static void elem_apply_fast(const ho_op* op, const double xe[8], double ye[8],
                            double k_e, double rhoc_e, double dt) {
    for (int i = 0; i < 8; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 8; ++j)
            acc += (rhoc_e * op->Mh[i * 8 + j] + dt * k_e * op->Kh[i * 8 + j]) * xe[j];
        ye[i] = acc;
    }
}

/* The MNHack7 prototype's material (src/heat_solver.cu:101-111): four
 * regions split at 0.5, kappa 100..400 and rho*c 1, 1.2, 1.2, 1.25. */
static void proto_material(double x, double y, double z, double* k, double* rc) {
    double kk, rho, c;
    if (x < 0.5 && y < 0.5 && z < 0.5)       { kk = 100.0; rho = 1.0; c = 1.0; }
    else if (x >= 0.5 && y < 0.5 && z < 0.5) { kk = 200.0; rho = 0.8; c = 1.5; }
    else if (x < 0.5 && y >= 0.5 && z < 0.5) { kk = 300.0; rho = 0.6; c = 2.0; }
    else                                     { kk = 400.0; rho = 0.5; c = 2.5; }
    *k = kk;
    *rc = rho * c;
}

/* src/heat_solver_het.cu:310-338; with op->proto the material of each
 * quadrature point instead of the element's (src/heat_solver.cu:193-232) */
static void elem_apply_gauss(const ho_op* op, const double xe[8], double ye[8],
                             double k_e, double rhoc_e, double dt, int ex, int ey, int ez) {
    const double h = op->g.h;
    const double det_J = h * h * h / 8.0;
    const double inv_Jx = 2.0 / h, inv_Jy = 2.0 / h, inv_Jz = 2.0 / h;
    const double gp = 1.0 / 1.7320508075688772;
    const double pts[2] = {-gp, gp};
    for (int i = 0; i < 8; ++i) ye[i] = 0.0;
    for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj)
    for (int gk = 0; gk < 2; ++gk) {
        double N[8], dNref[8][3];
        shape_func_grad(pts[gi], pts[gj], pts[gk], N, dNref);
        double dNx[8], dNy[8], dNz[8];
        for (int i = 0; i < 8; ++i) {
            dNx[i] = inv_Jx * dNref[i][0];
            dNy[i] = inv_Jy * dNref[i][1];
            dNz[i] = inv_Jz * dNref[i][2];
        }
        double kq = k_e, rq = rhoc_e;
        if (op->proto)
            proto_material(ex * h + (pts[gi] + 1.0) * 0.5 * h, ey * h + (pts[gj] + 1.0) * 0.5 * h,
                           ez * h + (pts[gk] + 1.0) * 0.5 * h, &kq, &rq);
        for (int i = 0; i < 8; ++i) {
            double Mx = 0.0, Kx = 0.0;
            for (int j = 0; j < 8; ++j) {
                Mx += rq * N[i] * N[j] * xe[j];
                Kx += kq * (dNx[i] * dNx[j] + dNy[i] * dNy[j] + dNz[i] * dNz[j]) * xe[j];
            }
            ye[i] += det_J * (Mx + dt * Kx);
        }
    }
}

/* Closed-form rows of the cube element matrix, plane-factored.  With d(i,j)
 * the Hamming distance of the local nodes, M_ij = m3 * {8,4,2,1}[d] and
 * K_ij = {-4 k2, 0, k2, k2}[d] (the values checked at
 * src/heat_solver_het.cu:677-705), so row i of (rhoc M + dt k K) x is
 * cr*Mr + ck*Kr with cr = rhoc*m3, ck = (dt*k)*k2 and, x_p(b,c) the element's
 * value in x-plane p at y-row b and z-column c (own = node i's, o = the other):
 *   Mr = 2 F(own plane) + F(other plane),  F(p) = 2 G(p, own row) + G(p, o row),
 *        G(p, b) = 2 x_p(b, own) + x_p(b, o)          (= 8 S0 + 4 S1 + 2 S2 + S3)
 *   Kr = DC(other plane) + C4(own plane),  DC(p) = (x_p(o, own) + x_p(own, o)) + x_p(o, o),
 *        C4(p) = x_p(o, o) - 4 x_p(own, own)           (= S2 + S3 - 4 S0)
 * with S_d the sum of x over the nodes at distance d from i.  In colour order
 * a row is added as y = fma(ck, Kr, fma(cr, Mr, y)) (mass-only, dt_eff = 0:
 * y = fma(cr, Mr, y)); a row taken on its own (the atomic order, a flipped row)
 * is fma(ck, Kr, cr*Mr) (mass-only: cr*Mr).  The operation order is the one
 * the HIP kernels use, so results agree bit for bit. */
static void row_terms(const double xe[8], int i, double* Mr, double* Kr) {
    const int ox = OFF[i][0], oy = OFF[i][1], oz = OFF[i][2];
    double F[2], DC[2], C4[2];
    for (int s = 0; s < 2; ++s) {               /* s = 0: the own plane, 1: the other */
        const int p = s ? 1 - ox : ox;
        const double aa = xe[LOC[p][oy][oz]], ab = xe[LOC[p][oy][1 - oz]];
        const double ba = xe[LOC[p][1 - oy][oz]], bb = xe[LOC[p][1 - oy][1 - oz]];
        F[s] = fma(2.0, fma(2.0, aa, ab), fma(2.0, ba, bb));
        DC[s] = (ba + ab) + bb;
        C4[s] = fma(-4.0, aa, bb);
    }
    *Mr = fma(2.0, F[0], F[1]);
    *Kr = DC[1] + C4[0];
}
static void row_terms_f(const float xe[8], int i, float* Mr, float* Kr) {
    const int ox = OFF[i][0], oy = OFF[i][1], oz = OFF[i][2];
    float F[2], DC[2], C4[2];
    for (int s = 0; s < 2; ++s) {
        const int p = s ? 1 - ox : ox;
        const float aa = xe[LOC[p][oy][oz]], ab = xe[LOC[p][oy][1 - oz]];
        const float ba = xe[LOC[p][1 - oy][oz]], bb = xe[LOC[p][1 - oy][1 - oz]];
        F[s] = fmaf(2.0f, fmaf(2.0f, aa, ab), fmaf(2.0f, ba, bb));
        DC[s] = (ba + ab) + bb;
        C4[s] = fmaf(-4.0f, aa, bb);
    }
    *Mr = fmaf(2.0f, F[0], F[1]);
    *Kr = DC[1] + C4[0];
}

static inline uint64_t d2u(double v) { uint64_t u; memcpy(&u, &v, 8); return u; }
static inline double u2d(uint64_t u) { double v; memcpy(&v, &u, 8); return v; }

/* operator_kernel, src/heat_solver_het.cu:351-378.  One element. */
static void operator_element(ho_op* op, const double* x, double* y, double dt_eff,
                             const ho_fi* fi, long e, int ix, int iy, int iz) {
    long gid[8];
    element_nodes(&op->g, ix, iy, iz, gid);
    double xe[8], ye[8];
    for (int i = 0; i < 8; ++i) xe[i] = x[gid[i]];
    const double k_e = op->k[e], rhoc_e = op->rhoc[e];
    const int flip = fi->armed && fi->level == 2 && fi->rank == 0 && e == fi->target;
    if (op->kernel == HO_KERNEL_STRUCT) {
        /* the closed-form rows: in colour order each row goes into y as
         * fma(ck, Kr, fma(cr, Mr, y)); the atomic order and a flipped row
         * (local node 0 of the target) add the row as a value */
        const double cr = rhoc_e * op->m3;
        const double ck = (dt_eff * k_e) * op->k2;
        const int mass = dt_eff == 0.0;
        for (int i = 0; i < 8; ++i) {
            double Mr, Kr;
            row_terms(xe, i, &Mr, &Kr);
            double* yi = &y[gid[i]];
            if (op->colored && !(i == 0 && flip)) {
                *yi = mass ? fma(cr, Mr, *yi) : fma(ck, Kr, fma(cr, Mr, *yi));
            } else {
                double v = mass ? cr * Mr : fma(ck, Kr, cr * Mr);
                if (i == 0 && flip) v = u2d(d2u(v) ^ (1ULL << fi->bit));
                *yi += v;
            }
        }
        return;
    }
    if (op->kernel == HO_KERNEL_FAST) elem_apply_fast(op, xe, ye, k_e, rhoc_e, dt_eff);
    else elem_apply_gauss(op, xe, ye, k_e, rhoc_e, dt_eff, ix, iy, iz);
    if (flip) ye[0] = u2d(d2u(ye[0]) ^ (1ULL << fi->bit));
    for (int i = 0; i < 8; ++i) y[gid[i]] += ye[i];
}

/* y += A x over all elements, colour order 0..7 (colored) or element order
 * (the atomic scatter's sums, taken in one fixed order). */
static void operator_sweep(ho_op* op, const double* x, double* y, double dt_eff, const ho_fi* fi) {
    const ho_grid* g = &op->g;
    if (!op->colored) {
        for (long e = 0; e < g->ne; ++e) {
            long ee; int ix, iy, iz;
            resolve_element(g, e, -1, &ee, &ix, &iy, &iz);
            operator_element(op, x, y, dt_eff, fi, ee, ix, iy, iz);
        }
        return;
    }
    for (int c = 0; c < 8; ++c) {
        long cnt = color_count(g, c);
#pragma omp parallel for schedule(static) num_threads(op->nthreads) if (op->nthreads > 1)
        for (long t = 0; t < cnt; ++t) {
            long e; int ix, iy, iz;
            resolve_element(g, t, c, &e, &ix, &iy, &iz);
            operator_element(op, x, y, dt_eff, fi, e, ix, iy, iz);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* reductions, src/heat_solver_het.cu:458-492, 506-534                       */
/* ------------------------------------------------------------------------- */
/* 256-wide block tree -> partial[block]; 1024-wide strided fixed-order sum and
 * tree.  `vals(i)` is x[i]*y[i]. */
static double dot_raw(const ho_op* op, const double* x, const double* y, double* partial) {
    const long n = op->g.nn, nb = op->nblk;
#pragma omp parallel for schedule(static) num_threads(op->nthreads) if (op->nthreads > 1)
    for (long b = 0; b < nb; ++b) {
        double s[256];
        for (int t = 0; t < 256; ++t) {
            long i = b * 256 + t;
            s[t] = (i < n) ? x[i] * y[i] : 0.0;
        }
        for (int st = 128; st > 0; st >>= 1)
            for (int t = 0; t < st; ++t) s[t] += s[t + st];
        partial[b] = s[0];
    }
    double f[1024];
    for (int t = 0; t < 1024; ++t) {
        double a = 0.0;
        for (long i = t; i < nb; i += 1024) a += partial[i];
        f[t] = a;
    }
    for (int st = 512; st > 0; st >>= 1)
        for (int t = 0; t < st; ++t) f[t] += f[t + st];
    return f[0];
}

double ho_dot(ho_op* op, const double* x, const double* y) { return dot_raw(op, x, y, op->partial); }
static double norm2(ho_op* op, const double* x) { return sqrt(ho_dot(op, x, x)); }

void ho_minmax(ho_op* op, const double* x, double* vmin, double* vmax) {
    double a = 1e300, b = -1e300;
    for (long i = 0; i < op->g.nn; ++i) { a = fmin(a, x[i]); b = fmax(b, x[i]); }
    *vmin = a; *vmax = b;
}

/* ------------------------------------------------------------------------- */
/* vector kernels, src/heat_solver_het.cu:427-456.  axpy/aypx are written as */
/* the fused multiply-add the CUDA compiler forms from y += a*x.             */
/* ------------------------------------------------------------------------- */
static void dirichlet_value(const ho_op* op, double* v, double val) {
    const ho_grid* g = &op->g;
    for (int ix = 0; ix <= g->n; ++ix) for (int iy = 0; iy <= g->n; ++iy) for (int iz = 0; iz <= g->n; ++iz)
        if (is_boundary(g, ix, iy, iz)) v[lnid(g, ix, iy, iz)] = val;
}
static void axpy(const ho_op* op, double a, const double* x, double* y) {
    const long n = op->g.nn;
#pragma omp parallel for schedule(static) num_threads(op->nthreads) if (op->nthreads > 1)
    for (long i = 0; i < n; ++i) y[i] = fma(a, x[i], y[i]);
}
static void aypx(const ho_op* op, double a, const double* x, double* y) {
    const long n = op->g.nn;
#pragma omp parallel for schedule(static) num_threads(op->nthreads) if (op->nthreads > 1)
    for (long i = 0; i < n; ++i) y[i] = fma(a, y[i], x[i]);
}
void ho_apply_preconditioner(ho_op* op, const double* r, double* z) {
    const long n = op->g.nn;
#pragma omp parallel for schedule(static) num_threads(op->nthreads) if (op->nthreads > 1)
    for (long i = 0; i < n; ++i) z[i] = r[i] / op->diag[i];
}

/* ------------------------------------------------------------------------- */
/* apply_operator, src/heat_solver_het.cu:885-921                            */
/* ------------------------------------------------------------------------- */
void ho_apply_operator(ho_op* op, const double* x, double* y, double dt_eff,
                       int allow_fi, int project) {
    memset(y, 0, sizeof(double) * op->g.nn);
    ho_fi f = op->fi;
    if (!allow_fi) f.armed = 0;
    operator_sweep(op, x, y, dt_eff, &f);
    if (allow_fi && op->fi.armed && op->fi.level == 2) op->fi.armed = 0;
    if (project) dirichlet_value(op, y, 0.0);
    if (op->detect && project) {
        const double* w = (dt_eff == 0.0) ? op->wM : op->wA;
        const double nw = (dt_eff == 0.0) ? op->norm_wM : op->norm_wA;
        double s = ho_dot(op, y, op->ones);
        double c = ho_dot(op, w, x);
        double nx = norm2(op, x);
        double rel = fabs(s - c) / (nw * nx + 1e-300);
        if (!isfinite(rel)) rel = 1e300;
        op->abft_step_max = fmax(op->abft_step_max, rel);
        op->abft_step_calls++;
    }
}

/* FP32 structured operator: FP32 loads, FP32 arithmetic, colour order. */
void ho_apply_operator_f32(ho_op* op, const float* x, float* y, double dt_eff, int project) {
    const ho_grid* g = &op->g;
    const float m3 = (float)op->m3, k2 = (float)op->k2;
    const int mass = dt_eff == 0.0;
    memset(y, 0, sizeof(float) * g->nn);
    for (int c = 0; c < 8; ++c) {
        long cnt = color_count(g, c);
        for (long t = 0; t < cnt; ++t) {
            long e; int ix, iy, iz;
            resolve_element(g, t, c, &e, &ix, &iy, &iz);
            long gid[8];
            element_nodes(g, ix, iy, iz, gid);
            float xe[8];
            for (int i = 0; i < 8; ++i) xe[i] = x[gid[i]];
            const float cr = (float)op->rhoc[e] * m3;
            const float ck = ((float)dt_eff * (float)op->k[e]) * k2;
            for (int i = 0; i < 8; ++i) {
                float Mr, Kr;
                row_terms_f(xe, i, &Mr, &Kr);
                float* yi = &y[gid[i]];
                *yi = mass ? fmaf(cr, Mr, *yi) : fmaf(ck, Kr, fmaf(cr, Mr, *yi));
            }
        }
    }
    if (project)
        for (long i = 0; i < g->nn; ++i) if (node_is_boundary(g, i)) y[i] = 0.0f;
}

/* b = M T^n + dt F, src/heat_solver_het.cu:963-971 */
void ho_build_rhs(ho_op* op, const double* Tn, double* b) {
    ho_apply_operator(op, Tn, b, 0.0, 0, 1);
    if (!op->verify && !op->no_source) axpy(op, op->dt, op->F, b);
}

/* ------------------------------------------------------------------------- */
/* setup, src/heat_solver_het.cu:756-816, 940-953, 380-422                   */
/* ------------------------------------------------------------------------- */
static double source_at(const ho_op* op, double x, double y, double z) {
    if (op->verify) return 0.0;
    if (op->source_mode == HO_SOURCE_CONST) return op->source_value;
    const double cx = 0.5, cy = 0.5, cz = 0.5, sigma = 0.1, coeff = 1000.0;
    double r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
    return coeff * exp(-r2 / (2.0 * sigma * sigma));
}

static void source_element(ho_op* op, double* F, int ix, int iy, int iz) {
    const ho_grid* g = &op->g;
    long gid[8];
    element_nodes(g, ix, iy, iz, gid);
    const double x0 = ix * g->h, y0 = iy * g->h, z0 = iz * g->h;
    const double det_J = g->h * g->h * g->h / 8.0;
    const double gp = 1.0 / 1.7320508075688772;
    const double pts[2] = {-gp, gp};
    double Fe[8] = {0};
    for (int gi = 0; gi < 2; ++gi)
    for (int gj = 0; gj < 2; ++gj)
    for (int gk = 0; gk < 2; ++gk) {
        double xg = x0 + (pts[gi] + 1.0) * 0.5 * g->h;
        double yg = y0 + (pts[gj] + 1.0) * 0.5 * g->h;
        double zg = z0 + (pts[gk] + 1.0) * 0.5 * g->h;
        double Q = source_at(op, xg, yg, zg);
        double N[8], dN[8][3];
        shape_func_grad(pts[gi], pts[gj], pts[gk], N, dN);
        for (int i = 0; i < 8; ++i) Fe[i] += det_J * Q * N[i];
    }
    for (int i = 0; i < 8; ++i) F[gid[i]] += Fe[i];
}

/* The HIP library's separable form of the same Gauss sum (k_source): both
 * source models factor over the axes, so node (a,b,c) of an element gets
 * ((pref * X_a) * Y_b) * Z_c with X_a = sum_p phi_a(t_p) G(x_p) over the two
 * Gauss points t_p of the element's x interval, pref = 1000 h^3/8 (Gaussian)
 * or Q h^3/8 with X = Y = Z = 1 (constant).  Used with the closed-form
 * ("struct") element rows, whose arithmetic it shares with the GPU. */
static const double GT0 = 0.21132486540518713, GT1 = 0.7886751345948129;

static void source_axis_sep(double lo, double h, double out[2]) {
    const double s0 = fma(h, GT0, lo) - 0.5, s1 = fma(h, GT1, lo) - 0.5;
    const double g0 = exp(-(s0 * s0) / (2.0 * 0.1 * 0.1)), g1 = exp(-(s1 * s1) / (2.0 * 0.1 * 0.1));
    out[0] = fma(GT1, g0, GT0 * g1);
    out[1] = fma(GT0, g0, GT1 * g1);
}

static void source_element_sep(ho_op* op, double* F, int ix, int iy, int iz) {
    const ho_grid* g = &op->g;
    if (op->verify) return;
    long gid[8];
    element_nodes(g, ix, iy, iz, gid);
    const double vol8 = g->h * g->h * g->h / 8.0;
    double X[2] = {1.0, 1.0}, Y[2] = {1.0, 1.0}, Z[2] = {1.0, 1.0};
    double pref = op->source_value * vol8;
    if (op->source_mode != HO_SOURCE_CONST) {
        source_axis_sep(ix * g->h, g->h, X);
        source_axis_sep(iy * g->h, g->h, Y);
        source_axis_sep(iz * g->h, g->h, Z);
        pref = 1000.0 * vol8;
    }
    for (int i = 0; i < 8; ++i)
        F[gid[i]] += ((pref * X[OFF[i][0]]) * Y[OFF[i][1]]) * Z[OFF[i][2]];
}

static void diag_element(ho_op* op, double* d, long e, int ix, int iy, int iz) {
    long gid[8];
    element_nodes(&op->g, ix, iy, iz, gid);
    const double k_e = op->k[e], rhoc_e = op->rhoc[e];
    if (op->proto) {
        /* the prototype's Gauss sum with the material of each point (src/heat_solver.cu:297-343) */
        const double h = op->g.h, det_J = h * h * h / 8.0, ij = 2.0 / h;
        const double gp = 1.0 / 1.7320508075688772;
        const double pts[2] = {-gp, gp};
        double de[8] = {0};
        for (int gi = 0; gi < 2; ++gi)
        for (int gj = 0; gj < 2; ++gj)
        for (int gk = 0; gk < 2; ++gk) {
            double kq, rq, N[8], dN[8][3];
            proto_material(ix * h + (pts[gi] + 1.0) * 0.5 * h, iy * h + (pts[gj] + 1.0) * 0.5 * h,
                           iz * h + (pts[gk] + 1.0) * 0.5 * h, &kq, &rq);
            shape_func_grad(pts[gi], pts[gj], pts[gk], N, dN);
            for (int i = 0; i < 8; ++i) {
                const double gx = ij * dN[i][0], gy = ij * dN[i][1], gz = ij * dN[i][2];
                de[i] += det_J * (rq * N[i] * N[i] + op->dt * kq * (gx * gx + gy * gy + gz * gz));
            }
        }
        for (int i = 0; i < 8; ++i) d[gid[i]] += de[i];
        return;
    }
    for (int i = 0; i < 8; ++i) {
        double de;
        if (op->kernel == HO_KERNEL_STRUCT)
            de = rhoc_e * (8.0 * op->m3) + (op->dt * k_e) * (-4.0 * op->k2);
        else
            de = rhoc_e * op->Mh[i * 9] + op->dt * k_e * op->Kh[i * 9];
        d[gid[i]] += de;
    }
}

static void element_loop(ho_op* op, double* v, int what) {
    const ho_grid* g = &op->g;
    for (int c = (op->colored ? 0 : -1); c < (op->colored ? 8 : 0); ++c) {
        long cnt = (c < 0) ? g->ne : color_count(g, c);
        for (long t = 0; t < cnt; ++t) {
            long e; int ix, iy, iz;
            resolve_element(g, t, c, &e, &ix, &iy, &iz);
            if (what == 0) diag_element(op, v, e, ix, iy, iz);
            else if (op->kernel == HO_KERNEL_STRUCT) source_element_sep(op, v, ix, iy, iz);
            else source_element(op, v, ix, iy, iz);
        }
    }
}

static void setup_detectors(ho_op* op) {
    const long n = op->g.nn;
    op->ones = (double*)malloc(sizeof(double) * n);
    op->wM = (double*)malloc(sizeof(double) * n);
    op->wA = (double*)malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) op->ones[i] = node_is_boundary(&op->g, i) ? 0.0 : 1.0;
    ho_apply_operator(op, op->ones, op->wM, 0.0, 0, 0);
    ho_apply_operator(op, op->ones, op->wA, op->dt, 0, 0);
    op->norm_wM = norm2(op, op->wM);
    op->norm_wA = norm2(op, op->wA);
    op->detect = 1;
}

static ho_op* create_impl(int n, double dt, const double* k, const double* rhoc, int kernel,
                          int colored, int verify, int source_mode, double source_value,
                          int detect, int verbose) {
    init_loc();
    if (n < 1) { snprintf(g_err, sizeof g_err, "N must be >= 1"); return NULL; }
    ho_op* op = (ho_op*)calloc(1, sizeof(ho_op));
    op->g.n = n; op->g.h = 1.0 / n; op->g.nn1 = n + 1;
    op->g.nn = (long)(n + 1) * (n + 1) * (n + 1);
    op->g.ne = (long)n * n * n;
    op->g.sx = (long)(n + 1) * (n + 1); op->g.sy = n + 1;
    op->g.geo_planes[0] = op->g.geo_planes[1] = 16;
    op->g.geo_wave[0] = op->g.geo_wave[1] = 768;
    op->dt = dt; op->kernel = kernel; op->colored = colored; op->verify = verify;
    op->source_mode = source_mode; op->source_value = source_value;
    op->chk_tol = 1e-10; op->nthreads = 1;
    const long ne = op->g.ne, nn = op->g.nn;
    op->k = (double*)malloc(sizeof(double) * ne);
    op->rhoc = (double*)malloc(sizeof(double) * ne);
    memcpy(op->k, k, sizeof(double) * ne);
    memcpy(op->rhoc, rhoc, sizeof(double) * ne);
    op->kmin = op->kmax = k[0]; op->rhocmin = op->rhocmax = rhoc[0];
    for (long i = 0; i < ne; ++i) {                               /* :734-742 */
        if (!(k[i] > 0) || !(rhoc[i] > 0)) {
            snprintf(g_err, sizeof g_err, "rank 0: non-positive or NaN coefficient at element %ld", i);
            ho_destroy(op); return NULL;
        }
        op->kmin = fmin(op->kmin, k[i]); op->kmax = fmax(op->kmax, k[i]);
        op->rhocmin = fmin(op->rhocmin, rhoc[i]); op->rhocmax = fmax(op->rhocmax, rhoc[i]);
    }
    op->nblk = (nn + 255) / 256;
    op->partial = (double*)malloc(sizeof(double) * op->nblk);
    op->diag = (double*)calloc(nn, sizeof(double));
    op->F = (double*)calloc(nn, sizeof(double));
    op->r = (double*)calloc(nn, sizeof(double));
    op->z = (double*)calloc(nn, sizeof(double));
    op->p = (double*)calloc(nn, sizeof(double));
    op->Ap = (double*)calloc(nn, sizeof(double));
    if (!build_reference_matrices(op, verbose)) {
        snprintf(g_err, sizeof g_err, "reference matrix self-check failed");
        ho_destroy(op); return NULL;
    }
    element_loop(op, op->diag, 0);
    dirichlet_value(op, op->diag, 1.0);
    if (!op->verify) {
        element_loop(op, op->F, 1);
        dirichlet_value(op, op->F, 0.0);
    }
    if (detect) setup_detectors(op);
    return op;
}

ho_op* ho_create(int n, double dt, const double* k, const double* rhoc, int kernel,
                 int colored, int verify, int source_mode, double source_value, int detect) {
    return create_impl(n, dt, k, rhoc, kernel, colored, verify, source_mode, source_value, detect, 0);
}

void ho_destroy(ho_op* op) {
    if (!op) return;
    free(op->k); free(op->rhoc); free(op->partial); free(op->diag); free(op->F);
    free(op->ones); free(op->wM); free(op->wA);
    free(op->r); free(op->z); free(op->p); free(op->Ap);
    free(op);
}

static int pcg_impl(ho_op* op, const double* b, double* x, int max_iter, double tol,
                    double* hist, FILE* diag_log);

/* The MNHack7 prototype run (src/heat_solver.cu:679-758): the unit cube at
 * N, T = 0, n_steps = (int)(t_final / dt + 0.5) backward-Euler steps of the
 * Gauss-quadrature operator with the quadrant material at each quadrature
 * point, atomic scatter, the Gaussian source, each step PCG to 1e-8 in at
 * most 500 iterations from the previous T.  iters[s] and norms[s] receive
 * step s's iteration count and ||T||_2 (the prototype prints them at
 * s % 10 == 0 and the last step).  Returns n_steps, or -1. */
int ho_proto_run(int n, double t_final, double dt, int nthreads, int max_steps, int* iters, double* norms) {
    const int n_steps = (int)(t_final / dt + 0.5);
    if (n < 1 || n_steps > max_steps) return -1;
    const long ne = (long)n * n * n;
    double* k = (double*)malloc(sizeof(double) * ne);
    double* rc = (double*)malloc(sizeof(double) * ne);
    const double h = 1.0 / n;
    for (int ix = 0; ix < n; ++ix)                      /* element-centre values: kmin/kmax only */
        for (int iy = 0; iy < n; ++iy)
            for (int iz = 0; iz < n; ++iz) {
                const long e = ((long)ix * n + iy) * n + iz;
                proto_material((ix + 0.5) * h, (iy + 0.5) * h, (iz + 0.5) * h, &k[e], &rc[e]);
            }
    ho_op* op = create_impl(n, dt, k, rc, HO_KERNEL_GAUSS, 0, 0, HO_SOURCE_GAUSSIAN, 0.0, 0, 0);
    free(k);
    free(rc);
    if (!op) return -1;
    op->nthreads = nthreads < 1 ? 1 : nthreads;
    op->proto = 1;
    const long nn = op->g.nn;
    memset(op->diag, 0, sizeof(double) * nn);           /* the diagonal again, point by point */
    element_loop(op, op->diag, 0);
    dirichlet_value(op, op->diag, 1.0);
    double* T = (double*)calloc(nn, sizeof(double));
    double* b = (double*)malloc(sizeof(double) * nn);
    for (int s = 0; s < n_steps; ++s) {
        ho_build_rhs(op, T, b);
        iters[s] = pcg_impl(op, b, T, 500, 1e-8, NULL, NULL);
        norms[s] = sqrt(ho_dot(op, T, T));
    }
    free(T);
    free(b);
    ho_destroy(op);
    return n_steps;
}

void ho_set_threads(ho_op* op, int nthreads) { op->nthreads = nthreads < 1 ? 1 : nthreads; }
void ho_set_no_source(ho_op* op, int no_source) { op->no_source = no_source; }
long ho_nn(const ho_op* op) { return op->g.nn; }
long ho_ne(const ho_op* op) { return op->g.ne; }
int ho_get_vector(const ho_op* op, int which, double* out) {
    const double* src = which == 0 ? op->diag : which == 1 ? op->F : which == 2 ? op->wM
                      : which == 3 ? op->wA : which == 4 ? op->ones : NULL;
    if (!src) return 1;
    memcpy(out, src, sizeof(double) * op->g.nn);
    return 0;
}
/* Replace the source vector F (which = 1) or the diagonal (which = 0): lets a
 * check feed the oracle the GPU's own setup vectors. */
int ho_set_vector(ho_op* op, int which, const double* in) {
    double* dst = which == 0 ? op->diag : which == 1 ? op->F : NULL;
    if (!dst) return 1;
    memcpy(dst, in, sizeof(double) * op->g.nn);
    return 0;
}
void ho_get_ref_matrices(const ho_op* op, double* Mh, double* Kh) {
    memcpy(Mh, op->Mh, sizeof(op->Mh)); memcpy(Kh, op->Kh, sizeof(op->Kh));
}
void ho_get_stats(const ho_op* op, double* o) {
    o[0] = op->kmin; o[1] = op->kmax; o[2] = op->rhocmin; o[3] = op->rhocmax;
}
void ho_arm_fault(ho_op* op, int level, long target, int bit) {
    op->fi.level = level; op->fi.target = target; op->fi.bit = bit; op->fi.rank = 0; op->fi.armed = 1;
}
void ho_flip_bit(double* v, long idx, int bit) { v[idx] = u2d(d2u(v[idx]) ^ (1ULL << bit)); }
void ho_abft_take(ho_op* op, double* step_max, long* calls) {
    if (step_max) *step_max = op->abft_step_max;
    if (calls) *calls = op->abft_step_calls;
    op->abft_step_max = 0; op->abft_step_calls = 0;
}

/* ------------------------------------------------------------------------- */
/* pcg_solve, src/heat_solver_het.cu:994-1039                                */
/* ------------------------------------------------------------------------- */
static int pcg_impl(ho_op* op, const double* b, double* x, int max_iter, double tol,
                    double* hist, FILE* diag_log) {
    const long n = op->g.nn;
    double *r = op->r, *z = op->z, *p = op->p, *Ap = op->Ap;
    ho_apply_operator(op, x, Ap, op->dt, 1, 1);          /* level R fires here */
    memcpy(r, b, sizeof(double) * n);
    axpy(op, -1.0, Ap, r);
    double bnorm = norm2(op, b);
    if (bnorm < 1e-30) bnorm = 1.0;
    if (hist || diag_log) {
        double r0 = norm2(op, r) / bnorm;
        if (hist) { hist[0] = r0; hist[1] = NAN; hist[2] = NAN; }
        if (diag_log) fprintf(diag_log, "0,%.10e,,\n", r0);
    }
    ho_apply_preconditioner(op, r, z);
    memcpy(p, z, sizeof(double) * n);
    double rho = ho_dot(op, r, z);
    int iter;
    for (iter = 0; iter < max_iter; ++iter) {
        ho_apply_operator(op, p, Ap, op->dt, 1, 1);
        double pAp = ho_dot(op, p, Ap);
        if (fabs(pAp) < 1e-30) break;
        double alpha = rho / pAp;
        axpy(op, alpha, p, x);
        axpy(op, -alpha, Ap, r);
        double rnorm = norm2(op, r);
        int conv = (rnorm / bnorm < tol);
        double beta = 0.0;
        if (!conv) {
            ho_apply_preconditioner(op, r, z);
            double rho_new = ho_dot(op, r, z);
            beta = rho_new / rho;
            aypx(op, beta, z, p);
            rho = rho_new;
        }
        if (hist) { double* hrow = hist + 3 * (iter + 1); hrow[0] = rnorm / bnorm; hrow[1] = alpha; hrow[2] = beta; }
        if (diag_log) fprintf(diag_log, "%d,%.10e,%.10e,%.10e\n", iter + 1, rnorm / bnorm, alpha, beta);
        if (conv) { iter++; break; }
    }
    return iter;
}

int ho_pcg_solve(ho_op* op, const double* b, double* x, int max_iter, double tol, double* hist) {
    return pcg_impl(op, b, x, max_iter, tol, hist, NULL);
}

/* ------------------------------------------------------------------------- */
/* pcg_solve in the GPU library's two-launch order.  The iteration is the    */
/* reference's (src/heat_solver_het.cu:994-1039); only the fixed order of    */
/* the inner-product sums differs:                                            */
/*  - p.Ap and the D1 sums of the apply: tiles of 8 (y) x 32 (z) node columns */
/*    by a chunk of consecutive planes (x; fused_chunks() chunks per column  */
/*    of whole plane pairs, lengths differing by at most one pair); each of   */
/*    the tile's 256 threads sums its column, then a 256-wide tree;           */
/*  - r.r, r.z (and b.b at the start): super-blocks of 2048 nodes; thread t   */
/*    sums nodes base + 256 s + t, s = 0..7, then a 256-wide tree;            */
/*  - final: thread t sums partials t, t+256, ... then a 256-wide tree.       */
/* ------------------------------------------------------------------------- */
enum { F_TY = 8, F_TZ = 32, F_CX = 16, F_CX_MIN = 8, F_WAVE = 768, F_NT = 256, F_SB = 2048 };
enum { F_XW = F_TZ + 2, F_XN = (F_TZ + 2) * (F_TY + 2), F_XR = (F_XN + F_NT - 1) / F_NT };

/* chunks per (y,z) column: ceil(nn1 / planes), raised until the tiles fill  */
/* whole waves of `wave` tiles (wave 0: no fill) but not to chunks shorter    */
/* than 8 planes; the default geometry is (16, 768)                           */
/* (mnhack8_amd/csrc/heat_pcg_fused.h fused_chunks)                           */
static int fused_chunks(int nn1, int ncol, int planes, int wave) {
    const int nch = (nn1 + planes - 1) / planes;
    if (wave <= 0) return nch;
    const long long waves = ((long long)ncol * nch + wave - 1) / wave;
    long long fill = waves * wave / ncol;
    if (fill > nn1 / F_CX_MIN) fill = nn1 / F_CX_MIN;
    return fill > nch ? (int)fill : nch;
}

static int fused_nch(const ho_grid* g, int ncol) {
    return fused_chunks(g->nn1, ncol, g->geo_planes[0], g->geo_wave[0]);
}

static long fused_ntiles(const ho_grid* g) {
    const int n = g->n;
    const int ncol = ((n + F_TY - 1) / F_TY) * ((n + F_TZ - 1) / F_TZ);
    return (long)ncol * fused_nch(g, ncol);
}

int ho_set_pcg_geometry(ho_op* op, int which, int planes, int wave) {
    if (which < 0 || which > 1) return -1;
    if (planes == 0) { planes = F_CX; wave = F_WAVE; }
    if (planes < 2 || wave < 0) return -1;
    op->g.geo_planes[which] = planes;
    op->g.geo_wave[which] = wave;
    return 0;
}

static double tree_256(double* s) {
    for (int st = F_NT / 2; st > 0; st >>= 1)
        for (int t = 0; t < st; ++t) s[t] += s[t + st];
    return s[0];
}

static double fold_partials(const double* part, long count) {
    double s[F_NT];
    for (int t = 0; t < F_NT; ++t) {
        double a = 0.0;
        for (long i = t; i < count; i += F_NT) a += part[i];
        s[t] = a;
    }
    return tree_256(s);
}

/* super-block order of sum_i x_i y_i */
static int g_fused_threads = 1;

static double sb_dot(long nn, const double* x, const double* y, double* part) {
    const long nsb = (nn + F_SB - 1) / F_SB;
#pragma omp parallel for schedule(static) num_threads(g_fused_threads) if (g_fused_threads > 1)
    for (long sb = 0; sb < nsb; ++sb) {
        double s[F_NT];
        for (int t = 0; t < F_NT; ++t) {
            double a = 0.0;
            for (int k = 0; k < F_SB / F_NT; ++k) {
                const long i = sb * F_SB + (long)k * F_NT + t;
                if (i < nn) a += x[i] * y[i];
            }
            s[t] = a;
        }
        part[sb] = tree_256(s);
    }
    return fold_partials(part, nsb);
}

/* tile order of p.Ap, 1_int.Ap, w.p and p.p (out[0..3]) */
static void tile_sums(const ho_grid* g, const double* p, const double* Ap, const double* w,
                      double* part, double out[4]) {
    const int n = g->n;
    const int nby = (n + F_TY - 1) / F_TY, nbz = (n + F_TZ - 1) / F_TZ;
    const int nch = fused_nch(g, nby * nbz);
    const long ntiles = (long)nby * nbz * nch;
#pragma omp parallel for schedule(static) num_threads(g_fused_threads) if (g_fused_threads > 1)
    for (long b = 0; b < ntiles; ++b) {
        const int bz = (int)(b % nbz), by = (int)((b / nbz) % nby), chunk = (int)(b / ((long)nbz * nby));
        /* chunks of whole plane pairs, lengths differing by at most one pair */
        const int npair = (g->nn1 + 1) / 2;
        const int ix0 = 2 * (int)(((long long)chunk * npair) / nch);
        int ix1 = 2 * (int)(((long long)(chunk + 1) * npair) / nch);
        if (ix1 > g->nn1) ix1 = g->nn1;
        double s[4][F_NT];
        for (int tid = 0; tid < F_NT; ++tid) {
            const int ty = tid / F_TZ, tz = tid % F_TZ;
            const int iy = 1 + by * F_TY + ty, iz = 1 + bz * F_TZ + tz;
            double a_pap = 0.0, a_s = 0.0, a_c = 0.0, a_xx = 0.0;
            /* where the tile writes p: window rows 1..8 and columns 1..32 (row and
             * column 0 too on the first tiles), in plane order, slot order */
            for (int xp = ix0; xp < ix1; ++xp)
                for (int k = 0; k < F_XR; ++k) {
                    const int i = tid + k * F_NT;
                    if (i >= F_XN) continue;
                    const int r = i / F_XW, c = i % F_XW;
                    const int yy = by * F_TY + r, zz = bz * F_TZ + c;
                    if (yy > n || zz > n || r > F_TY || c > F_TZ) continue;
                    if ((r < 1 && by != 0) || (c < 1 && bz != 0)) continue;
                    const double v = p[lnid(g, xp, yy, zz)];
                    if (w) a_c += w[lnid(g, xp, yy, zz)] * v;
                    a_xx += v * v;
                }
            for (int ix = ix0; ix < ix1; ++ix)
                if (ix >= 1 && ix <= n - 1 && iy <= n - 1 && iz <= n - 1) {
                    const long i = lnid(g, ix, iy, iz);
                    a_pap += p[i] * Ap[i];
                    a_s += Ap[i];
                }
            s[0][tid] = a_pap; s[1][tid] = a_s; s[2][tid] = a_c; s[3][tid] = a_xx;
        }
        for (int q = 0; q < 4; ++q) part[q * ntiles + b] = tree_256(s[q]);
    }
    for (int q = 0; q < 4; ++q) out[q] = fold_partials(part + q * ntiles, ntiles);
}

static long fused_part_len(const ho_grid* g) {
    const long ntiles = fused_ntiles(g);
    const long nsb = (g->nn + F_SB - 1) / F_SB;
    return 4 * (ntiles > nsb ? ntiles : nsb);
}

/* the two orders on their own, for the checks of the restatement */
double ho_fused_sb_dot(ho_op* op, const double* x, const double* y) {
    double* part = (double*)malloc(sizeof(double) * fused_part_len(&op->g));
    const double v = sb_dot(op->g.nn, x, y, part);
    free(part);
    return v;
}
void ho_fused_tile_sums(ho_op* op, const double* p, const double* Ap, const double* w, double out[4]) {
    double* part = (double*)malloc(sizeof(double) * fused_part_len(&op->g));
    tile_sums(&op->g, p, Ap, w, part, out);
    free(part);
}

static void apply_projected(ho_op* op, const double* x, double* y) {
    ho_fi none = op->fi;
    none.armed = 0;
    memset(y, 0, sizeof(double) * op->g.nn);
    operator_sweep(op, x, y, op->dt, &none);
    dirichlet_value(op, y, 0.0);
}

static void flip_f32(float* v, long idx, int bit) {
    uint32_t u;
    memcpy(&u, &v[idx], 4);
    u ^= 1u << bit;
    memcpy(&v[idx], &u, 4);
}

void ho_arm_pcg_fault(ho_op* op, int vec, long idx, int bit, int iter) {
    op->sfi_armed = 1; op->sfi_vec = vec; op->sfi_idx = idx; op->sfi_bit = bit; op->sfi_iter = iter;
}
int ho_pcg_fault_fired(const ho_op* op) { return op->sfi_fired; }
int ho_mixed_status(const ho_op* op) { return op->mixed_status; }

int ho_pcg_solve_fused(ho_op* op, const double* b, double* x, int max_iter, double tol, double* hist) {
    const ho_grid* g = &op->g;
    const long n = g->nn;
    g_fused_threads = op->nthreads;
    double *r = op->r, *z = op->z, *p = op->p, *Ap = op->Ap;
    double* part = (double*)malloc(sizeof(double) * fused_part_len(g));
    ho_apply_operator(op, x, Ap, op->dt, 1, 1);          /* level R fires here; D1 in the reference order */
    for (long i = 0; i < n; ++i) { r[i] = fma(-1.0, Ap[i], b[i]); z[i] = r[i] / op->diag[i]; }
    double bnorm = sqrt(sb_dot(n, b, b, part));
    if (bnorm < 1e-30) bnorm = 1.0;
    const double rr0 = sb_dot(n, r, r, part);
    double rho = sb_dot(n, r, z, part);
    if (hist) { hist[0] = sqrt(rr0) / bnorm; hist[1] = NAN; hist[2] = NAN; }
    int iter = 0;
    const int fi = op->sfi_armed;
    op->sfi_armed = 0;
    op->sfi_fired = 0;
    if (max_iter > 0) {
        memcpy(p, z, sizeof(double) * n);
        for (;;) {
            apply_projected(op, p, Ap);
            double t[4];
            tile_sums(g, p, Ap, op->detect ? op->wA : NULL, part, t);
            /* the apply stores p and Ap; an armed upset hits the stored copy */
            const int hit = fi && iter == op->sfi_iter;
            if (hit && op->sfi_vec == HO_SEU_P) { ho_flip_bit(p, op->sfi_idx, op->sfi_bit); op->sfi_fired = 1; }
            if (hit && op->sfi_vec == HO_SEU_AP) { ho_flip_bit(Ap, op->sfi_idx, op->sfi_bit); op->sfi_fired = 1; }
            if (op->detect) {
                double rel = fabs(t[1] - t[2]) / (op->norm_wA * sqrt(t[3]) + 1e-300);
                if (!isfinite(rel)) rel = 1e300;
                op->abft_step_max = fmax(op->abft_step_max, rel);
                op->abft_step_calls++;
            }
            if (fabs(t[0]) < 1e-30) break;
            const double alpha = rho / t[0];
            for (long i = 0; i < n; ++i) {
                x[i] = fma(alpha, p[i], x[i]);
                r[i] = fma(-alpha, Ap[i], r[i]);
                z[i] = r[i] / op->diag[i];
            }
            const double rr = sb_dot(n, r, r, part), rz = sb_dot(n, r, z, part);
            if (hit && (op->sfi_vec == HO_SEU_U || op->sfi_vec == HO_SEU_R || op->sfi_vec == HO_SEU_Z)) {
                double* v = op->sfi_vec == HO_SEU_U ? x : op->sfi_vec == HO_SEU_R ? r : z;
                ho_flip_bit(v, op->sfi_idx, op->sfi_bit);
                op->sfi_fired = 1;
            }
            ++iter;
            const double rnorm = sqrt(rr);
            const int conv = (rnorm / bnorm < tol);
            const double beta = conv ? 0.0 : rz / rho;
            if (hist) { double* h = hist + 3 * iter; h[0] = rnorm / bnorm; h[1] = alpha; h[2] = beta; }
            if (conv || iter >= max_iter) break;
            rho = rz;
            for (long i = 0; i < n; ++i) p[i] = fma(beta, p[i], z[i]);
        }
    }
    free(part);
    return iter;
}

/* FP32-refined solve in the same order: FP64 outer residual r = b - A x, an
 * FP32 Jacobi-PCG on A e = r (FP32 vectors and operator; alpha, beta and the
 * inner products in FP64) to a relative tolerance tol_inner, x += e, until
 * ||b - A x|| / ||b|| < tol, the residual is not finite, or max_outer
 * corrections. */
int ho_pcg_solve_mixed_fused(ho_op* op, const double* b, double* x, int max_outer, int max_inner,
                             double tol, double tol_inner, int* outer, int* inner_total) {
    const ho_grid* g = &op->g;
    const long n = g->nn;
    float* f = (float*)malloc(sizeof(float) * 6 * n);
    float *e = f, *r = f + n, *z = f + 2 * n, *p = f + 3 * n, *ap = f + 4 * n, *d = f + 5 * n;
    double* dv = (double*)malloc(sizeof(double) * 5 * n);
    double *r64 = dv, *dr = dv + n, *dz = dv + 2 * n, *dp = dv + 3 * n, *dap = dv + 4 * n;
    double* part = (double*)malloc(sizeof(double) * fused_part_len(g));
    for (long i = 0; i < n; ++i) d[i] = (float)op->diag[i];
    int nout = 0, total = 0;
    const int fi = op->sfi_armed;
    op->sfi_armed = 0;
    op->sfi_fired = 0;
    for (;;) {
        apply_projected(op, x, r64);
        if (fi && op->sfi_vec == HO_SEU_AX64 && nout == op->sfi_iter) {   /* the pass's stored A x */
            ho_flip_bit(r64, op->sfi_idx, op->sfi_bit);
            op->sfi_fired = 1;
        }
        for (long i = 0; i < n; ++i) {
            r64[i] = fma(-1.0, r64[i], b[i]);
            r[i] = (float)r64[i];
            z[i] = r[i] / d[i];
            e[i] = 0.0f;
            dr[i] = r[i]; dz[i] = z[i];
        }
        const double rr = sb_dot(n, dr, dr, part), rz = sb_dot(n, dr, dz, part);
        double bo = sqrt(sb_dot(n, b, b, part));
        if (bo < 1e-30) bo = 1.0;
        const double rn = sqrt(sb_dot(n, r64, r64, part));
        const int outer_done = (rn / bo < tol) ? 1 : (isfinite(rn) ? 0 : 2);   /* 2: not finite, stop */
        if (!outer_done && max_inner > 0) {
            double bnorm = sqrt(rr);
            if (bnorm < 1e-30) bnorm = 1.0;
            double rho = rz;
            int it = 0;
            memcpy(p, z, sizeof(float) * n);
            for (;;) {
                ho_apply_operator_f32(op, p, ap, op->dt, 1);
                for (long i = 0; i < n; ++i) { dp[i] = p[i]; dap[i] = ap[i]; }
                double t[4];
                tile_sums(g, dp, dap, op->detect ? op->wA : NULL, part, t);
                if (op->detect) {
                    double rel = fabs(t[1] - t[2]) / (op->norm_wA * sqrt(t[3]) + 1e-300);
                    if (!isfinite(rel)) rel = 1e300;
                    op->abft_step_max = fmax(op->abft_step_max, rel);
                    op->abft_step_calls++;
                }
                const int hit = fi && total == op->sfi_iter;
                if (hit && op->sfi_vec == HO_SEU_P) { flip_f32(p, op->sfi_idx, op->sfi_bit); op->sfi_fired = 1; }
                if (hit && op->sfi_vec == HO_SEU_AP) { flip_f32(ap, op->sfi_idx, op->sfi_bit); op->sfi_fired = 1; }
                if (fabs(t[0]) < 1e-30 * bnorm * bnorm) break;   /* scaled to the inner right-hand side */
                const double alpha = rho / t[0];
                for (long i = 0; i < n; ++i) {
                    e[i] = (float)fma(alpha, (double)p[i], (double)e[i]);
                    r[i] = (float)fma(-alpha, (double)ap[i], (double)r[i]);
                    z[i] = r[i] / d[i];
                    dr[i] = r[i]; dz[i] = z[i];
                }
                const double rr2 = sb_dot(n, dr, dr, part), rz2 = sb_dot(n, dr, dz, part);
                if (hit && (op->sfi_vec == HO_SEU_U || op->sfi_vec == HO_SEU_R || op->sfi_vec == HO_SEU_Z)) {
                    float* v = op->sfi_vec == HO_SEU_U ? e : op->sfi_vec == HO_SEU_R ? r : z;
                    flip_f32(v, op->sfi_idx, op->sfi_bit);
                    op->sfi_fired = 1;
                }
                ++it;
                ++total;
                const int conv = sqrt(rr2) / bnorm < tol_inner;
                const double beta = conv ? 0.0 : rz2 / rho;
                if (conv || it >= max_inner) break;
                rho = rz2;
                for (long i = 0; i < n; ++i) p[i] = (float)fma(beta, (double)p[i], (double)z[i]);
            }
        }
        op->mixed_status = outer_done;
        if (outer_done || nout >= max_outer) break;
        for (long i = 0; i < n; ++i) x[i] += (double)e[i];
        if (fi && op->sfi_vec == HO_SEU_X64 && nout == op->sfi_iter) {    /* x as this correction stores it */
            ho_flip_bit(x, op->sfi_idx, op->sfi_bit);
            op->sfi_fired = 1;
        }
        ++nout;
    }
    if (outer) *outer = nout;
    if (inner_total) *inner_total = total;
    free(f);
    free(dv);
    free(part);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* The one-launch inner iteration of the refined solve (mnhack8_amd/csrc/     */
/* heat_cg1.h): preconditioned CG in the Chronopoulos-Gear arrangement, its   */
/* tiles and its summation order.                                            */
/* ------------------------------------------------------------------------- */
enum { C1_TY = 8, C1_TZ = 32 };

typedef struct { int m, nby, nbz, nch; long ntiles; } c1_geom;

/* tiles of 8 x 32 interior nodes by chunks of the n - 1 interior planes */
static c1_geom cg1_geom(const ho_grid* g) {
    c1_geom q;
    q.m = g->n - 1 > 0 ? g->n - 1 : 1;
    q.nby = (q.m + C1_TY - 1) / C1_TY;
    q.nbz = (q.m + C1_TZ - 1) / C1_TZ;
    q.nch = fused_chunks(q.m, q.nby * q.nbz, g->geo_planes[1], g->geo_wave[1]);
    q.ntiles = (long)q.nby * q.nbz * q.nch;
    return q;
}

/* A row of tile sums: the term A[i] * B[i] (or A[i] when B is NULL) over the
 * tile's own interior nodes, per thread over its node (y0 + tid / 32,
 * z0 + tid % 32) plane by plane, then the 256-wide tree, then the fold of
 * the tiles' partials. */
typedef struct { const double* a; const double* b; } c1_row;

static void cg1_sums(const ho_grid* g, const c1_geom* q, const c1_row* rows, int nrows, double* part,
                     double* out) {
#pragma omp parallel for schedule(static) num_threads(g_fused_threads) if (g_fused_threads > 1)
    for (long t = 0; t < q->ntiles; ++t) {
        const int bz = (int)(t % q->nbz), by = (int)((t / q->nbz) % q->nby), chunk = (int)(t / ((long)q->nbz * q->nby));
        const int ix0 = 1 + (int)(((long long)chunk * q->m) / q->nch);
        const int ix1 = 1 + (int)(((long long)(chunk + 1) * q->m) / q->nch);
        const int y0 = 1 + by * C1_TY, z0 = 1 + bz * C1_TZ;
        for (int k = 0; k < nrows; ++k) {
            double s[F_NT];
            for (int tid = 0; tid < F_NT; ++tid) {
                double acc = 0.0;
                const int yy = y0 + tid / C1_TZ, zz = z0 + tid % C1_TZ;
                if (yy <= q->m && zz <= q->m) {
                    for (int xp = ix0; xp < ix1; ++xp) {
                        const long id = lnid(g, xp, yy, zz);
                        acc += rows[k].b ? rows[k].a[id] * rows[k].b[id] : rows[k].a[id];
                    }
                }
                s[tid] = acc;
            }
            part[(long)k * q->ntiles + t] = tree_256(s);
        }
    }
    for (int k = 0; k < nrows; ++k) out[k] = fold_partials(part + (long)k * q->ntiles, q->ntiles);
}

/* The refined solve with the one-launch inner iteration: as
 * ho_pcg_solve_mixed_fused (FP64 residual, correction, the same stopping
 * rules), with the inner solve
 *   start   r0 = (float)(b - A x), u0 = r0 * dinv (dinv = 1/d in FP32), w0 = A u0 (FP32 apply)
 *   step i  beta = gamma_i / gamma_{i-1}, alpha = gamma_i / fma(-beta, gamma_i / alpha_{i-1}, delta_i)
 *           (step 0: beta = 0, alpha = gamma_0 / delta_0), then in FP32 with alpha, beta rounded:
 *           s = w + beta s, r -= alpha s, p = r_old dinv + beta p, e += alpha p, u = r dinv, w = A u,
 *           gamma = r.u, delta = w.u, r.r in the tile order of cg1_sums;
 * the D1 check of each inner apply is taken at the next step's head. */
int ho_pcg_solve_mixed_cg1(ho_op* op, const double* b, double* x, int max_outer, int max_inner,
                           double tol, double tol_inner, int* outer, int* inner_total) {
    const ho_grid* g = &op->g;
    const long n = g->nn;
    const c1_geom q = cg1_geom(g);
    float* f = (float*)malloc(sizeof(float) * 7 * n);
    float *e = f, *r = f + n, *s = f + 2 * n, *p = f + 3 * n, *u = f + 4 * n, *w = f + 5 * n, *dinv = f + 6 * n;
    double* dv = (double*)malloc(sizeof(double) * 6 * n);
    double *ax = dv, *r64 = dv + n, *dr = dv + 2 * n, *du = dv + 3 * n, *dw = dv + 4 * n, *bcp = dv + 5 * n;
    double* part = (double*)malloc(sizeof(double) * 8 * (size_t)q.ntiles);
    unsigned char* in = (unsigned char*)malloc((size_t)n);
    for (long i = 0; i < n; ++i) {
        dinv[i] = 1.0f / (float)op->diag[i];
        in[i] = !node_is_boundary(g, i);
        e[i] = 0.0f;
    }
    int nout = 0, total = 0;
    op->sfi_armed = 0;
    op->sfi_fired = 0;
    for (;;) {
        apply_projected(op, x, ax);
        for (long i = 0; i < n; ++i) {
            if (in[i]) {
                r64[i] = fma(-1.0, ax[i], b[i]);
                r[i] = (float)r64[i];
                u[i] = r[i] * dinv[i];
                e[i] = 0.0f;
                bcp[i] = b[i];
            } else {
                r64[i] = 0.0; r[i] = 0.0f; u[i] = 0.0f; bcp[i] = 0.0;
            }
            dr[i] = r[i]; du[i] = u[i];
        }
        ho_apply_operator_f32(op, u, w, op->dt, 1);
        for (long i = 0; i < n; ++i) dw[i] = w[i];
        double t[8];
        const c1_row start_rows[8] = {{dr, du}, {dw, du}, {dr, dr}, {dw, NULL},
                                      {op->detect ? op->wA : du, du}, {du, du}, {bcp, bcp}, {r64, r64}};
        cg1_sums(g, &q, start_rows, 8, part, t);
        double bn = 1.0, gamma_prev = 0.0, alpha_prev = 0.0;
        int outer_done = 0;
        for (int it = 0;; ++it) {
            if (op->detect) {
                double rel = fabs(t[3] - t[4]) / (op->norm_wA * sqrt(t[5]) + 1e-300);
                if (!isfinite(rel)) rel = 1e300;
                op->abft_step_max = fmax(op->abft_step_max, rel);
                op->abft_step_calls++;
            }
            double beta, den;
            int stop;
            if (it == 0) {
                double bo = sqrt(t[6]);
                if (bo < 1e-30) bo = 1.0;
                const double rn = sqrt(t[7]);
                outer_done = (rn / bo < tol) ? 1 : (isfinite(rn) ? 0 : 2);
                bn = sqrt(t[2]);
                if (bn < 1e-30) bn = 1.0;
                beta = 0.0;
                den = t[1];
                stop = outer_done != 0 || max_inner <= 0 || fabs(den) < 1e-30 * bn * bn;
            } else {
                const int conv = sqrt(t[2]) / bn < tol_inner;
                beta = t[0] / gamma_prev;
                den = fma(-beta, t[0] / alpha_prev, t[1]);
                stop = conv || it >= max_inner || fabs(den) < 1e-30 * bn * bn;
            }
            if (stop) break;
            const double alpha = t[0] / den;
            gamma_prev = t[0];
            alpha_prev = alpha;
            const int first = it == 0;
            const float af = (float)alpha, bf = (float)beta;
            for (long i = 0; i < n; ++i) {
                if (!in[i]) continue;
                const float sn = first ? w[i] : fmaf(bf, s[i], w[i]);
                const float rn = fmaf(-af, sn, r[i]);
                const float uo = r[i] * dinv[i];
                const float pn = first ? uo : fmaf(bf, p[i], uo);
                e[i] = fmaf(af, pn, e[i]);
                p[i] = pn;
                s[i] = sn;
                r[i] = rn;
                u[i] = rn * dinv[i];
                dr[i] = r[i]; du[i] = u[i];
            }
            ho_apply_operator_f32(op, u, w, op->dt, 1);
            for (long i = 0; i < n; ++i) dw[i] = w[i];
            const c1_row step_rows[6] = {{dr, du}, {dw, du}, {dr, dr}, {dw, NULL},
                                         {op->detect ? op->wA : du, du}, {du, du}};
            cg1_sums(g, &q, step_rows, op->detect ? 6 : 3, part, t);
            ++total;
        }
        op->mixed_status = outer_done;
        if (outer_done || nout >= max_outer) break;
        for (long i = 0; i < n; ++i) x[i] += (double)e[i];
        ++nout;
    }
    if (outer) *outer = nout;
    if (inner_total) *inner_total = total;
    free(f);
    free(dv);
    free(part);
    free(in);
    return 0;
}

/* the solve the command line runs, in the order chosen by --pcg-order */
static int solve_cli(ho_op* op, const double* b, double* x, int max_iter, double tol, FILE* diag_log) {
    /* the GPU library runs the fused iteration on the colour-ordered fast path only */
    if (op->pcg_order == 0 || !op->colored || op->kernel == HO_KERNEL_GAUSS)
        return pcg_impl(op, b, x, max_iter, tol, NULL, diag_log);
    double* hist = diag_log ? (double*)malloc(sizeof(double) * 3 * (size_t)(max_iter + 1)) : NULL;
    const int it = ho_pcg_solve_fused(op, b, x, max_iter, tol, hist);
    if (diag_log) {
        fprintf(diag_log, "0,%.10e,,\n", hist[0]);
        for (int i = 1; i <= it; ++i)
            fprintf(diag_log, "%d,%.10e,%.10e,%.10e\n", i, hist[3 * i], hist[3 * i + 1], hist[3 * i + 2]);
        free(hist);
    }
    return it;
}

/* src/heat_solver_het.cu:556-584 */
void ho_initial_condition(ho_op* op, double* v, int mode, double t, double alpha) {
    const ho_grid* g = &op->g;
    const double pi = 3.14159265358979323846;
    for (int ix = 0; ix <= g->n; ++ix) for (int iy = 0; iy <= g->n; ++iy) for (int iz = 0; iz <= g->n; ++iz) {
        long nid = lnid(g, ix, iy, iz);
        double x = ix * g->h, y = iy * g->h, z = iz * g->h;
        double val = 0.0;
        if (mode == 1) {
            double r2 = (x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) + (z - 0.5) * (z - 0.5);
            val = exp(-r2 / (2.0 * 0.1 * 0.1));
        } else if (mode == 2) {
            double amp = exp(-3.0 * pi * pi * alpha * t);
            val = amp * sin(pi * x) * sin(pi * y) * sin(pi * z);
        }
        if (is_boundary(g, ix, iy, iz)) val = 0.0;
        v[nid] = val;
    }
}

/* ------------------------------------------------------------------------- */
/* driver: run_case, twin run, verify, CSV outputs,                          */
/* src/heat_solver_het.cu:1054-1629                                          */
/* ------------------------------------------------------------------------- */
typedef struct { double* v; long n, cap; } dvec;
typedef struct { int* v; long n, cap; } ivec;
static void dpush(dvec* a, double x) {
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 64; a->v = (double*)realloc(a->v, sizeof(double) * a->cap); }
    a->v[a->n++] = x;
}
static void ipush(ivec* a, int x) {
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 64; a->v = (int*)realloc(a->v, sizeof(int) * a->cap); }
    a->v[a->n++] = x;
}

typedef struct {
    ivec snap_steps; dvec E, spread; ivec cg_iters;
    double solve_ms; int nonfinite;
    dvec d1_abft, d2_resid, d3_tmin, d3_tmax;
    double d1_max, d2_max;
    int d1_first, d2_first, d3_first;
} run_result;

static void rr_init(run_result* r) { memset(r, 0, sizeof *r); r->d1_first = r->d2_first = r->d3_first = -1; }
static void rr_free(run_result* r) {
    free(r->snap_steps.v); free(r->E.v); free(r->spread.v); free(r->cg_iters.v);
    free(r->d1_abft.v); free(r->d2_resid.v); free(r->d3_tmin.v); free(r->d3_tmax.v);
}

static double now_ms(void) {
    struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* global_rel_err, src/heat_solver_het.cu:1054-1080 */
static double global_rel_err(const ho_op* op, const double* a, const double* b, double* spread) {
    double num = 0.0, den = 0.0, dmax = 0.0; long nown = 0;
    for (long n = 0; n < op->g.nn; ++n) {
        double d = a[n] - b[n];
        num += d * d; den += b[n] * b[n];
        if (fabs(d) > dmax) dmax = fabs(d);
        ++nown;
    }
    if (spread) {
        long cnt = 0;
        for (long n = 0; n < op->g.nn; ++n) if (fabs(a[n] - b[n]) > 1e-3 * dmax) ++cnt;
        *spread = (dmax > 0 && nown > 0) ? (double)cnt / (double)nown : 0.0;
    }
    if (den < 1e-300) den = 1.0;
    return sqrt(num / den);
}

typedef struct {
    int max_iter; double tol;
    int ic_mode;
    const char* dump_dir; const char* dump_tag;
} run_opts;

/* run_case, src/heat_solver_het.cu:1107-1259 */
static void run_case(ho_op* op, int n_steps, int snap_every, ho_fi fi_cfg, int reproject,
                     double* T, double* b, double** keep, long* nkeep, double* const* ref,
                     run_result* res, const run_opts* ro, const char* case_tag,
                     const run_result* d3_ref) {
    const long n = op->g.nn;
    double* tmp = (double*)malloc(sizeof(double) * n);
    memset(T, 0, sizeof(double) * n);
    if (ro->ic_mode == 1) ho_initial_condition(op, T, 1, 0.0, 0.0);
    if (ro->ic_mode == 2) ho_initial_condition(op, T, 2, 0.0, op->kmin / op->rhocmin);
    memset(&op->fi, 0, sizeof op->fi);
    op->fi.step = -1;
    rr_init(res);
    if (nkeep) *nkeep = 0;
    long si = 0;
    for (int step = 0; step < n_steps; ++step) {
        if (fi_cfg.level > 0 && step == fi_cfg.step) {
            if (fi_cfg.level == 1) { if (fi_cfg.rank == 0) ho_flip_bit(T, fi_cfg.target % n, fi_cfg.bit); }
            else { op->fi = fi_cfg; op->fi.armed = 1; }
        }
        double t0 = now_ms();
        ho_build_rhs(op, T, b);
        int it = solve_cli(op, b, T, ro->max_iter, ro->tol, NULL);
        if (reproject) dirichlet_value(op, T, 0.0);
        res->solve_ms += now_ms() - t0;
        op->fi.armed = 0;
        ipush(&res->cg_iters, it);
        if (op->detect) {
            double d1 = op->abft_step_max;
            ho_apply_operator(op, T, op->Ap, op->dt, 0, 1);
            memcpy(op->r, b, sizeof(double) * n);
            axpy(op, -1.0, op->Ap, op->r);
            double bn = norm2(op, b); if (bn < 1e-300) bn = 1.0;
            double d2 = norm2(op, op->r) / bn;
            double tmin, tmax; ho_minmax(op, T, &tmin, &tmax);
            int d3_fire = !isfinite(tmin) || !isfinite(tmax);
            if (d3_ref && step < d3_ref->d3_tmax.n) {
                double ref_max = d3_ref->d3_tmax.v[step], ref_min = d3_ref->d3_tmin.v[step];
                double margin = 1e-6 * fmax(fabs(ref_max), 1e-300) + 1e-12;
                if (tmax > ref_max + margin || tmin < ref_min - margin) d3_fire = 1;
            }
            dpush(&res->d1_abft, d1); dpush(&res->d2_resid, d2);
            dpush(&res->d3_tmin, tmin); dpush(&res->d3_tmax, tmax);
            if (!(d1 <= op->chk_tol) && res->d1_first < 0) res->d1_first = step;
            if (!(d2 <= 1e-6) && res->d2_first < 0) res->d2_first = step;
            if (d3_fire && res->d3_first < 0) res->d3_first = step;
            if (isfinite(d1)) res->d1_max = fmax(res->d1_max, d1); else res->d1_max = 1e300;
            if (isfinite(d2)) res->d2_max = fmax(res->d2_max, d2); else res->d2_max = 1e300;
            op->abft_step_max = 0; op->abft_step_calls = 0;
        }
        int snap = (snap_every > 0 && step % snap_every == 0) || step == n_steps - 1;
        if (snap) {
            if (keep) { keep[*nkeep] = (double*)malloc(sizeof(double) * n); memcpy(keep[*nkeep], T, sizeof(double) * n); }
            if (ro->dump_dir) {
                char pth[1024];
                snprintf(pth, sizeof pth, "%s/snap_%s_%s_%05d_rank%d.bin", ro->dump_dir, ro->dump_tag, case_tag, step, 0);
                FILE* f = fopen(pth, "wb"); if (f) { fwrite(T, sizeof(double), n, f); fclose(f); }
            }
            if (keep) (*nkeep)++;
            if (ref) {
                double sp = 0;
                double E = global_rel_err(op, T, ref[si], &sp);
                if (!isfinite(E)) res->nonfinite = 1;
                dpush(&res->E, E); dpush(&res->spread, sp);
            }
            ipush(&res->snap_steps, step);
            ++si;
        }
    }
    free(tmp);
}

static int write_field(const double* T, long n, const char* outdir, const char* tag) {
    char path[1024];
    snprintf(path, sizeof path, "%s/T_final_%s_rank0.bin", outdir, tag);
    FILE* f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path); return 1; }
    fwrite(T, sizeof(double), n, f);
    fclose(f);
    return 0;
}

/* target_on_boundary, src/heat_solver_het.cu:1272-1290 */
static int target_on_boundary(const ho_op* op, const ho_fi* fi) {
    const ho_grid* g = &op->g;
    if (fi->level == 1) return node_is_boundary(g, fi->target % g->nn);
    if (fi->level == 2) {
        long e = fi->target % g->ne;
        int ix = (int)(e / ((long)g->n * g->n)), iy = (int)((e / g->n) % g->n), iz = (int)(e % g->n);
        return is_boundary(g, ix, iy, iz);
    }
    return 0;
}

static void usage(const char* prog) {
    fprintf(stderr,
      "usage: %s N t_final dt field_prefix [outdir]\n"
      "   [--fi LEVEL STEP TARGET BIT] [--fi-rank R] [--snap S] [--tag NAME]\n"
      "   [--kernel gauss|fast|struct] [--scatter atomic|colored] [--thr X]\n"
      "   [--reproject-bc] [--detect] [--verify] [--no-source] [--ic zero|gaussian|sine] [--dump-snaps]\n"
      "   [--diag-cg FILE]   (one CG solve, logs iteration,rel_residual to FILE, then exits)\n"
      "   [--tol X] [--max-iter N] [--source-const Q] [--threads T]\n"
      "  LEVEL: 1=G global memory, 2=R accumulator, 3=H halo staging buffer\n"
      "  BIT:   0-51 mantissa, 52-62 exponent, 63 sign\n", prog);
}

static int load_fields(const char* prefix, long ne, double** k, double** rhoc) {
    char fk[1024], fr[1024];
    snprintf(fk, sizeof fk, "%s_k.bin", prefix);
    snprintf(fr, sizeof fr, "%s_rhoc.bin", prefix);
    FILE* f1 = fopen(fk, "rb");
    FILE* f2 = fopen(fr, "rb");
    if (!f1 || !f2) {
        fprintf(stderr, "Cannot open %s / %s (run gen_field.py)\n", fk, fr);
        if (f1) fclose(f1);
        if (f2) fclose(f2);
        return 1;
    }
    *k = (double*)malloc(sizeof(double) * ne);
    *rhoc = (double*)malloc(sizeof(double) * ne);
    int bad = fread(*k, sizeof(double), ne, f1) != (size_t)ne || fread(*rhoc, sizeof(double), ne, f2) != (size_t)ne;
    fclose(f1); fclose(f2);
    if (bad) {
        fprintf(stderr, "Field file too small: expected %ld doubles for this rank "
                "(grid mismatch? regenerate with gen_field.py --n %d)\n", ne, (int)0);
        return 1;
    }
    return 0;
}

int ho_main(int argc, char** argv) {
    if (argc < 5) { usage(argv[0]); return 1; }
    int N = atoi(argv[1]);
    double t_final = atof(argv[2]);
    double dt = atof(argv[3]);
    const char* field_prefix = argv[4];
    const char* outdir = "out";
    const char* tag = "run";
    ho_fi fi_cfg; memset(&fi_cfg, 0, sizeof fi_cfg); fi_cfg.step = -1;
    int snap_every = 5, verify_mode = 0, kernel = HO_KERNEL_FAST, colored = 0, reproject = 0, detect = 0;
    int no_source = 0, ic_mode = 0, dump_snaps = 0, nthreads = 1, max_iter = 500, pcg_order = 0;
    int source_mode = HO_SOURCE_GAUSSIAN; double source_value = 0.0;
    double tol = 1e-8;
    const char* diag_cg_path = NULL;
    double thr_user = 1e-13;
    for (int i = 5; i < argc; ++i) {
        if (!strcmp(argv[i], "--fi") && i + 4 < argc) {
            fi_cfg.level = atoi(argv[++i]); fi_cfg.step = atoi(argv[++i]);
            fi_cfg.target = atol(argv[++i]); fi_cfg.bit = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--fi-rank") && i + 1 < argc) { fi_cfg.rank = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--snap") && i + 1 < argc) { snap_every = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--tag") && i + 1 < argc) { tag = argv[++i];
        } else if (!strcmp(argv[i], "--thr") && i + 1 < argc) { thr_user = atof(argv[++i]);
        } else if (!strcmp(argv[i], "--kernel") && i + 1 < argc) {
            ++i; kernel = !strcmp(argv[i], "gauss") ? HO_KERNEL_GAUSS : !strcmp(argv[i], "struct") ? HO_KERNEL_STRUCT : HO_KERNEL_FAST;
        } else if (!strcmp(argv[i], "--scatter") && i + 1 < argc) { colored = !strcmp(argv[++i], "colored");
        } else if (!strcmp(argv[i], "--reproject-bc")) { reproject = 1;
        } else if (!strcmp(argv[i], "--detect")) { detect = 1;
        } else if (!strcmp(argv[i], "--diag-cg") && i + 1 < argc) { diag_cg_path = argv[++i];
        } else if (!strcmp(argv[i], "--no-source")) { no_source = 1;
        } else if (!strcmp(argv[i], "--dump-snaps")) { dump_snaps = 1;
        } else if (!strcmp(argv[i], "--ic") && i + 1 < argc) {
            ++i; ic_mode = !strcmp(argv[i], "gaussian") ? 1 : !strcmp(argv[i], "sine") ? 2 : 0;
        } else if (!strcmp(argv[i], "--verify")) { verify_mode = 1;
        } else if (!strcmp(argv[i], "--tol") && i + 1 < argc) { tol = atof(argv[++i]);
        } else if (!strcmp(argv[i], "--max-iter") && i + 1 < argc) { max_iter = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--source-const") && i + 1 < argc) {
            source_mode = HO_SOURCE_CONST; source_value = atof(argv[++i]);
        } else if (!strcmp(argv[i], "--threads") && i + 1 < argc) { nthreads = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--pcg-order") && i + 1 < argc) { pcg_order = !strcmp(argv[++i], "fused");
        } else if (argv[i][0] != '-') { outdir = argv[i];
        } else { fprintf(stderr, "unknown option %s\n", argv[i]); usage(argv[0]); return 1; }
    }
    const int nprocs = 1;
    if (fi_cfg.level > 0 && (fi_cfg.rank < 0 || fi_cfg.rank >= nprocs)) {
        fprintf(stderr, "--fi-rank must be in [0,%d)\n", nprocs); return 1;
    }
    if (fi_cfg.level == 3) { fprintf(stderr, "level H needs >= 2 ranks and --fi-rank < nprocs-1\n"); return 1; }
    if (fi_cfg.level == 4) { fprintf(stderr, "level 4 (handoff transit) needs --qrank --qblock\n"); return 1; }
    if (fi_cfg.level > 0 && (fi_cfg.bit < 0 || fi_cfg.bit > 63)) { fprintf(stderr, "BIT must be in 0..63\n"); return 1; }

    int n_steps = (int)(t_final / dt + 0.5);
    const char* kname = kernel == HO_KERNEL_GAUSS ? "gauss" : kernel == HO_KERNEL_STRUCT ? "struct" : "fast";
    printf("=== Assembly-Free FEM Heat Solver v2 (Backward Euler + Jacobi-PCG) ===\n");
    printf("Grid %d^3 (DOFs %ld)  ranks %d (1D x-decomp, %d cells/rank)\n",
           N, (long)(N + 1) * (N + 1) * (N + 1), nprocs, N / nprocs);
    printf("dt %.4e  t_final %.4e  steps %d  kernel %s  scatter %s  mode %s\n",
           dt, t_final, n_steps, kname, colored ? "colored" : "atomic", verify_mode ? "VERIFY" : "production");

    long ne = (long)N * N * N;
    double *hk = NULL, *hr = NULL;
    if (load_fields(field_prefix, ne, &hk, &hr)) { free(hk); free(hr); return 1; }
    ho_op* op = create_impl(N, dt, hk, hr, kernel, colored, verify_mode, source_mode, source_value, detect, 1);
    free(hk); free(hr);
    if (!op) { fprintf(stderr, "%s\n", g_err); return g_err[0] == 'r' && strstr(g_err, "self-check") ? 2 : 1; }
    op->no_source = no_source;
    ho_set_threads(op, nthreads);
    op->pcg_order = pcg_order;
    printf("Material: k in [%g, %g] (contrast %.3g), rho*c in [%g, %g]\n",
           op->kmin, op->kmax, op->kmax / op->kmin, op->rhocmin, op->rhocmax);

    const long n = op->g.nn;
    double* T = (double*)calloc(n, sizeof(double));
    double* b = (double*)calloc(n, sizeof(double));
    int rc = 0;
    run_opts ro = {max_iter, tol, ic_mode, dump_snaps ? outdir : NULL, tag};

    if (verify_mode) {
        /* run_verify, src/heat_solver_het.cu:1293-1338 */
        if (op->kmin != op->kmax || op->rhocmin != op->rhocmax) {
            fprintf(stderr, "--verify needs a homogeneous field (k %g..%g, rhoc %g..%g)\n",
                    op->kmin, op->kmax, op->rhocmin, op->rhocmax);
            rc = 1; goto done;
        }
        const double alpha = op->kmin / op->rhocmin;
        ho_initial_condition(op, T, 2, 0.0, alpha);
        double t0 = now_ms();
        long total_it = 0;
        for (int step = 0; step < n_steps; ++step) {
            ho_build_rhs(op, T, b);
            total_it += solve_cli(op, b, T, max_iter, tol, NULL);
        }
        double ms = now_ms() - t0;
        double t_end = n_steps * dt;
        double* ex = (double*)malloc(sizeof(double) * n);
        ho_initial_condition(op, ex, 2, t_end, alpha);
        double exact_norm = norm2(op, ex);
        axpy(op, -1.0, T, ex);
        double rel_err = norm2(op, ex) / exact_norm;
        free(ex);
        printf("VERIFY: N=%d dt=%.4e t_end=%.6e alpha=%g rel_L2_err=%.6e  (%s/%s, %.1f ms, %.1f CG it/step)\n",
               N, dt, t_end, alpha, rel_err, kname, colored ? "colored" : "atomic", ms, (double)total_it / n_steps);
        char p[1024]; snprintf(p, sizeof p, "%s/verify.csv", outdir);
        FILE* vf = fopen(p, "r"); int fresh = (vf == NULL); if (vf) fclose(vf);
        vf = fopen(p, "a");
        if (!vf) { fprintf(stderr, "cannot write %s\n", p); rc = 1; goto done; }
        if (fresh) fprintf(vf, "N,dt,t_end,nprocs,kernel,scatter,rel_L2_err,solve_ms,avg_cg_iters\n");
        fprintf(vf, "%d,%.6e,%.6e,%d,%s,%s,%.10e,%.3f,%.3f\n", N, dt, t_end, nprocs, kname,
                colored ? "colored" : "atomic", rel_err, ms, (double)total_it / n_steps);
        fclose(vf);
    } else if (diag_cg_path) {
        FILE* lf = fopen(diag_cg_path, "w");
        if (!lf) { fprintf(stderr, "cannot write %s\n", diag_cg_path); rc = 1; goto done; }
        fprintf(lf, "iter,rel_residual,alpha,beta\n");
        memset(T, 0, sizeof(double) * n);
        if (ic_mode == 1) ho_initial_condition(op, T, 1, 0.0, 0.0);
        if (ic_mode == 2) ho_initial_condition(op, T, 2, 0.0, op->kmin / op->rhocmin);
        ho_build_rhs(op, T, b);
        int it = solve_cli(op, b, T, max_iter, tol, lf);
        fclose(lf);
        printf("diag-cg: %d iterations to %g, contrast k_max/k_min=%.3g -> %s\n", it, tol, op->kmax / op->kmin, diag_cg_path);
    } else {
        /* twin-run protocol, src/heat_solver_het.cu:1505-1621 */
        long max_snaps = n_steps + 1;
        double** snapA = (double**)calloc(max_snaps, sizeof(double*));
        long nA = 0;
        run_result rA, rB, rF;
        rr_init(&rF);
        ho_fi none; memset(&none, 0, sizeof none); none.step = -1;
        run_case(op, n_steps, snap_every, none, reproject, T, b, snapA, &nA, NULL, &rA, &ro, "cleanA", NULL);
        char tg[512];
        snprintf(tg, sizeof tg, "%s_clean", tag);
        if (write_field(snapA[nA - 1], n, outdir, tg)) { rc = 1; goto twin_done; }
        run_case(op, n_steps, snap_every, none, reproject, T, b, NULL, NULL, snapA, &rB, &ro, "cleanB", &rA);
        int on_bnd = -1;
        if (fi_cfg.level > 0) {
            on_bnd = target_on_boundary(op, &fi_cfg);
            run_case(op, n_steps, snap_every, fi_cfg, reproject, T, b, NULL, NULL, snapA, &rF, &ro, "fault", &rA);
            snprintf(tg, sizeof tg, "%s_fault", tag);
            if (write_field(T, n, outdir, tg)) { rc = 1; goto twin_done; }
        }
        {
            double Efl_max = 0, Efa_max = 0, Efa_int = 0, Efa_final = 0, Efl_final = 0, spread_at_max = 0, spread_final = 0;
            int step_max = -1, iters_delta_max = 0;
            for (long s = 0; s < rB.E.n; ++s) {
                Efl_max = fmax(Efl_max, rB.E.v[s]); Efl_final = rB.E.v[s];
                if (fi_cfg.level > 0) {
                    double E = rF.E.v[s];
                    if (isfinite(E) && E > Efa_max) { Efa_max = E; step_max = rB.snap_steps.v[s]; spread_at_max = rF.spread.v[s]; }
                    Efa_final = E; spread_final = rF.spread.v[s];
                    double w = (s + 1 < rB.E.n) ? (rB.snap_steps.v[s + 1] - rB.snap_steps.v[s]) * dt : 0.0;
                    if (isfinite(E)) Efa_int += E * w;
                }
            }
            for (long s = 0; fi_cfg.level > 0 && s < rF.cg_iters.n; ++s) {
                int d = rF.cg_iters.v[s] - rA.cg_iters.v[s];
                if (d > iters_delta_max) iters_delta_max = d;
            }
            double thr = fmax(thr_user, 10.0 * Efl_max);
            const char* cls = "clean";
            if (fi_cfg.level > 0) {
                if (rF.nonfinite) cls = "detected";
                else if (Efa_max <= thr) cls = "benign";
                else if (Efa_final <= thr) cls = "transient";
                else cls = "persistent";
            }
            char tp[1024]; snprintf(tp, sizeof tp, "%s/fi_trace_%s.csv", outdir, tag);
            FILE* tf = fopen(tp, "w");
            if (!tf) { fprintf(stderr, "cannot write %s\n", tp); rc = 1; goto twin_done; }
            fprintf(tf, "step,t,E_noise_floor,E_fault,spread_F,cg_A,cg_B,cg_F,d1_abft_A,d1_abft_F,d2_resid_A,d2_resid_F,d3_tmin_F,d3_tmax_F,hyb_block_err_A,hyb_block_err_F,hyb_d4_F\n");
            const run_result* rX = fi_cfg.level > 0 ? &rF : &rA;
            for (long s = 0; s < rB.E.n; ++s) {
                int st = rB.snap_steps.v[s];
                fprintf(tf, "%d,%.6e,%.10e,%.10e,%.4f,%d,%d,%d", st, st * dt, rB.E.v[s],
                        fi_cfg.level > 0 ? rF.E.v[s] : 0.0, fi_cfg.level > 0 ? rF.spread.v[s] : 0.0,
                        rA.cg_iters.v[st], rB.cg_iters.v[st], fi_cfg.level > 0 ? rF.cg_iters.v[st] : -1);
                if (detect)
                    fprintf(tf, ",%.3e,%.3e,%.3e,%.3e,%.6e,%.6e", rA.d1_abft.v[st], rX->d1_abft.v[st],
                            rA.d2_resid.v[st], rX->d2_resid.v[st], rX->d3_tmin.v[st], rX->d3_tmax.v[st]);
                else fprintf(tf, ",,,,,,");
                fprintf(tf, ",,,");
                fprintf(tf, "\n");
            }
            fclose(tf);
            double meanA = 0; for (long s = 0; s < rA.cg_iters.n; ++s) meanA += rA.cg_iters.v[s];
            meanA /= (rA.cg_iters.n > 0 ? rA.cg_iters.n : 1);
            char sp[1024]; snprintf(sp, sizeof sp, "%s/fi_summary_%s.csv", outdir, tag);
            FILE* sf = fopen(sp, "w");
            if (!sf) { fprintf(stderr, "cannot write %s\n", sp); rc = 1; goto twin_done; }
            fprintf(sf, "tag,N,dt,n_steps,nprocs,kernel,scatter,reproject_bc,kmin,kmax,rhocmin,rhocmax,"
                        "fi_level,fi_rank,fi_step,fi_target,fi_bit,target_on_boundary,"
                        "E_floor_max,E_floor_final,E_fault_max,step_at_max,E_fault_int,E_fault_final,"
                        "thr,iters_delta_max,mean_cg_iters_A,solve_ms_A,solve_ms_B,solve_ms_F,class,spread_at_max,spread_final,"
                        "detect,d1_first,d2_first,d3_first,d1_max_F,d2_max_F,d1_max_A,d2_max_A,detected_by,"
                        "hybrid,hyb_block_err_max_A,hyb_block_err_max_F,d4_first_F\n");
            fprintf(sf, "%s,%d,%.3e,%d,%d,%s,%s,%d,%g,%g,%g,%g,%d,%d,%d,%ld,%d,%d,"
                        "%.6e,%.6e,%.6e,%d,%.6e,%.6e,%.3e,%d,%.2f,%.3f,%.3f,%.3f,%s",
                    tag, N, dt, n_steps, nprocs, kname, colored ? "colored" : "atomic", reproject,
                    op->kmin, op->kmax, op->rhocmin, op->rhocmax,
                    fi_cfg.level, fi_cfg.rank, fi_cfg.step, fi_cfg.target, fi_cfg.bit, on_bnd,
                    Efl_max, Efl_final, Efa_max, step_max, Efa_int, Efa_final, thr, iters_delta_max, meanA,
                    rA.solve_ms, rB.solve_ms, rF.solve_ms, cls);
            fprintf(sf, ",%.4f,%.4f", spread_at_max, spread_final);
            char by[32] = "";
            if (detect) {
                if (rX->d1_first >= 0) strcat(by, "D1");
                if (rX->d2_first >= 0) { if (by[0]) strcat(by, "+"); strcat(by, "D2"); }
                if (rX->d3_first >= 0) { if (by[0]) strcat(by, "+"); strcat(by, "D3"); }
                if (!by[0]) strcpy(by, "none");
            } else strcpy(by, "off");
            fprintf(sf, ",%d,%d,%d,%d,%.3e,%.3e,%.3e,%.3e,%s", detect, rX->d1_first, rX->d2_first, rX->d3_first,
                    rX->d1_max, rX->d2_max, rA.d1_max, rA.d2_max, by);
            fprintf(sf, ",%d,%.3e,%.3e,%d\n", 0, 0.0, 0.0, -1);
            if (detect)
                printf("detectors    D1 first %d (max %.2e)  D2 first %d (max %.2e)  D3 first %d  -> %s   [clean A: D1 %.2e D2 %.2e]\n",
                       rX->d1_first, rX->d1_max, rX->d2_first, rX->d2_max, rX->d3_first, by, rA.d1_max, rA.d2_max);
            fclose(sf);
            printf("\n=== Done (%s) ===\n", tag);
            printf("solve wall A/B/F (ms): %.1f / %.1f / %.1f   mean CG iters A: %.2f\n", rA.solve_ms, rB.solve_ms, rF.solve_ms, meanA);
            printf("noise floor  max %.3e  final %.3e\n", Efl_max, Efl_final);
            if (fi_cfg.level > 0)
                printf("fault        max %.3e @step %d  int %.3e  final %.3e  iters+%d  bnd %d  -> %s\n",
                       Efa_max, step_max, Efa_int, Efa_final, iters_delta_max, on_bnd, cls);
            char mp[1024]; snprintf(mp, sizeof mp, "%s/meta_%s.txt", outdir, tag);
            FILE* mf = fopen(mp, "w");
            if (mf) {
                fprintf(mf, "N %d\nnprocs %d\ndt %g\nn_steps %d\nfield %s\nkernel %s\nscatter %s\n",
                        N, nprocs, dt, n_steps, field_prefix, kname, colored ? "colored" : "atomic");
                fclose(mf);
            }
        }
    twin_done:
        for (long s = 0; s < nA; ++s) free(snapA[s]);
        free(snapA);
        rr_free(&rA); rr_free(&rB); rr_free(&rF);
    }
done:
    free(T); free(b);
    ho_destroy(op);
    return rc;
}
