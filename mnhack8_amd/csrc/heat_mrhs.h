// heat_mrhs.h -- sixteen right-hand sides at once: Y_r = P (rho*c M + dt kappa K) X_r
// in FP32, with the element contraction on the matrix cores.
//
// Per element e the sixteen columns of X_e (8 nodes x 16 right-hand sides)
// are contracted with one constant 16 x 8 integer matrix whose rows are, for
// each local node i, the Hamming-distance weights of the closed-form element
// rows (src/heat_solver_het.cu:677-705):
//   P_i = sum_j {8,4,2,1}[d(i,j)] x_j      (rho*c M row = m3 P_i)
//   Q_i = sum_j {-4,0,1,1}[d(i,j)] x_j     (kappa K row = k2 Q_i)
// so y_i = cr_e P_i + ck_e Q_i with cr = rho*c m3 and ck = dt kappa k2, the
// same coefficients and final fused multiply-add as the single-vector
// kernels.  The 16 x 8 by 8 x 16 product is one v_mfma_f32_16x16x32_bf16:
// each FP32 value x is split exactly into three bfloat16 terms by truncation
// (hi = the top 16 bits of x, mid = the top 16 bits of x - hi, lo = x - hi -
// mid, which has at most 8 significant bits), the contraction matrix's small
// integers are exact in bfloat16, so every product is exact and the matrix
// core sums the 24 of them in FP32: K = 32 holds, per node, (hi, mid, lo, 0).
// Rows are ordered (P_a, P_b, Q_a, Q_b) per pair of nodes a = (0, y, z),
// b = (1, y, z), so each lane holds both terms of its two outputs and no
// lane exchange is needed.
//
// Blocks own a TY x TZ tile of node columns and walk cx node planes in x
// (a launch argument: the host picks it per mesh to fill the last wave).
// Element plane ex is contracted in two phases of element-row parity: rows
// of one parity share no node row, and one wave takes a row, adds each pair
// of z-neighbouring elements' contributions in registers (a cross-lane swap)
// and updates each LDS accumulator word once, so every node receives its
// eight contributions in a fixed order.  Node plane ex is complete after
// element plane ex and is written out (boundary rows 0: the Dirichlet
// projection).
//
// Layout: node-major, right-hand-side-minor: X[node * 16 + r].
#pragma once
#include "heat_kernels.h"

namespace heat {

constexpr int MR = 16;                                   // right-hand sides
constexpr int MTY = 15, MTZ = 16;                        // node tile per block
constexpr int MCX_MIN = 8, MCX_MAX = 64;                 // node planes per block (host-chosen)
constexpr int MNT = 512;                                 // threads: 8 waves, one element row per wave and phase

typedef float mr_v4f __attribute__((ext_vector_type(4)));
typedef unsigned mr_v4u __attribute__((ext_vector_type(4)));
typedef __bf16 mr_v8bf __attribute__((ext_vector_type(8)));

// the exact three-term bfloat16 split of x as two words of the B operand:
// (hi | mid << 16, lo)
__device__ __forceinline__ void mrhs_split(float x, unsigned& w0, unsigned& w1) {
    const unsigned xb = __float_as_uint(x);
    const float r1 = x - __uint_as_float(xb & 0xffff0000u);          // exact: same sign and exponent
    const unsigned r1b = __float_as_uint(r1);
    const float r2 = r1 - __uint_as_float(r1b & 0xffff0000u);        // exact, at most 8 significant bits
    w0 = __builtin_amdgcn_perm(r1b, xb, 0x07060302u);                // upper halves of x and r1
    w1 = __float_as_uint(r2) >> 16;
}

// entry (row, j) of the contraction matrix: row = 4 g + t, g = (y | z << 1) of
// the node pair, t = 0: P_a, 1: P_b, 2: Q_a, 3: Q_b; column j = local node
// (j & 1, (j >> 1) & 1, j >> 2) in (x, y, z)
__device__ __forceinline__ float mrhs_coef(int row, int j) {
    const int g = row >> 2, t = row & 3;
    const int ix = t & 1, iy = g & 1, iz = g >> 1;
    const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
    const int d = (ix != jx) + (iy != jy) + (iz != jz);
    if (t < 2) return (float)(8 >> d);
    return d == 0 ? -4.0f : (d == 1 ? 0.0f : 1.0f);
}

// SUMS = 0: Y = P A X only.  SUMS = 1: also each tile's per-column sums of
// X.Y over its own nodes; SUMS = 2: also 1.Y, w.X and X.X (the D1 terms).
// Partial row k of column c of tile b lands in part[(k * 16 + c) * ntiles + b]:
// each thread sums its nodes in plane order, a wave combines its lanes of one
// column, and the block adds its waves in order (a fixed order).
template <bool MASS, int SUMS = 0>
__global__ __launch_bounds__(MNT) void k_mrhs_apply(const float* __restrict__ X, float* __restrict__ Y,
                                                    const float* __restrict__ cr, const float* __restrict__ ck,
                                                    Grid g, int nby, int nbz, int ntiles, int cx,
                                                    double* __restrict__ part = nullptr,
                                                    const double* __restrict__ w = nullptr) {
    constexpr int TY = MTY, TZ = MTZ;
    constexpr int WY = TY + 2, WZ = TZ + 2, WN = WY * WZ;          // node window with halo
    constexpr int EY = TY + 1, EZ = TZ + 1, EN = EY * EZ;          // elements per plane
    constexpr int XSTRIDE = WN * MR + 16;                          // slot stride: odd multiple of 16 floats
    // accumulator rows of TZ nodes plus a 16-float pad, with a pad row (and 16
    // floats) before and a pad row after the tile: the halo outputs of the
    // tile's edge elements (nodes another tile owns) land in pads, so no
    // output needs a mask.  Plane slot s starts at s YSTRIDE + YBASE.
    constexpr int YROW = TZ * MR + 16;                              // accumulator row stride
    constexpr int YSTRIDE = (TY + 2) * YROW;
    constexpr int YBASE = YROW + 16;
    constexpr int YALL = 2 * YSTRIDE + 32;
    constexpr int V4 = WN * (MR / 4);                               // float4 per node window
    constexpr int XR = (V4 + MNT - 1) / MNT;
    constexpr int ER = (EN + MNT - 1) / MNT;
    __shared__ float xs[2 * XSTRIDE];
    __shared__ __align__(16) float ys[YALL];
    __shared__ float ec[2][2][EN];
    const int b = xcd_slab_block(ntiles);
    if (b >= ntiles) return;
    const int bz = b % nbz, by = (b / nbz) % nby, chunk = b / (nbz * nby);
    const int n = g.n, nn1 = g.nn1, nsq = n * n;
    const size_t S = (size_t)g.sx;
    const int y0 = by * TY, z0 = bz * TZ;
    const int ix0 = chunk * cx, ix1 = min(ix0 + cx, n + 1);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, q = lane >> 4;                        // RHS column; k / row group
    const int qx = q & 1, qy = q >> 1;                               // B operand node (x, y) of group q
    // A: row = lane % 16; lane group q holds in one half of its K slots node
    // q's three terms and a zero, in the other node 4 + q's (the same slots
    // as the B operand's).  Element kz of a row keeps its lower-z nodes in
    // the half its upper-z nodes had in element kz - 1, so the B operand is
    // updated in place: even kz use amat[0] (lower z first), odd kz amat[1].
    mr_v8bf amat[2];
    {
        const unsigned c0 = __float_as_uint(mrhs_coef(col, q)) >> 16, c1 = __float_as_uint(mrhs_coef(col, 4 + q)) >> 16;
        const mr_v4u ae = {c0 | (c0 << 16), c0, c1 | (c1 << 16), c1};
        const mr_v4u ao = {c1 | (c1 << 16), c1, c0 | (c0 << 16), c0};
        amat[0] = __builtin_bit_cast(mr_v8bf, ae);
        amat[1] = __builtin_bit_cast(mr_v8bf, ao);
    }
    constexpr int NK = SUMS == 2 ? 4 : 1;                            // sums per column
    double acc_s[NK][4];                                             // this thread's 4 columns: part = tid & 3
#pragma unroll
    for (int k = 0; k < NK; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc_s[k][j] = 0.0;

    // ---- staging: node window of plane p and element coefficients of plane e
    // This thread's window units and elements are the same on every plane:
    // their in-plane offsets (floats of X, elements of cr / ck) and whether
    // they lie on the mesh are formed once; a plane adds its base.
    int xoff[XR], eoff[ER];
    bool xok[XR], eok[ER];
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int v = tid + k * MNT;
        const int w = v >> 2, part = v & 3;
        const int wy = w / WZ, wz = w - (w / WZ) * WZ;
        const int yy = y0 - 1 + wy, zz = z0 - 1 + wz;
        xok[k] = v < V4 && yy >= 0 && yy <= n && zz >= 0 && zz <= n;
        xoff[k] = xok[k] ? (yy * nn1 + zz) * MR + part * 4 : 0;
    }
#pragma unroll
    for (int k = 0; k < ER; ++k) {
        const int v = tid + k * MNT;
        const int ky = v / EZ, kz = v - (v / EZ) * EZ;
        const int ey = y0 - 1 + ky, ez = z0 - 1 + kz;
        eok[k] = v < EN && ey >= 0 && ey < n && ez >= 0 && ez < n;
        eoff[k] = eok[k] ? ey * n + ez : 0;
    }
    float4 xf[XR];
    float ef[2][ER];
    auto fetch = [&](int p, int e) {
        const bool pin = p >= 0 && p <= n, ein = e >= 0 && e < n;          // uniform over the block
        const float* Xp = X + (size_t)(pin ? p : 0) * S * MR;
        const float* crp = cr + (size_t)(ein ? e : 0) * nsq;
        const float* ckp = ck + (size_t)(ein ? e : 0) * nsq;
#pragma unroll
        for (int k = 0; k < XR; ++k) {
            float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
            if (pin && xok[k]) val = *reinterpret_cast<const float4*>(Xp + xoff[k]);
            xf[k] = val;
        }
#pragma unroll
        for (int k = 0; k < ER; ++k) {
            float c0 = 0.f, c1 = 0.f;
            if (ein && eok[k]) {
                c0 = crp[eoff[k]];
                c1 = MASS ? 0.f : ckp[eoff[k]];
            }
            ef[0][k] = c0;
            ef[1][k] = c1;
        }
    };
    auto park_x = [&](int slot) {
#pragma unroll
        for (int k = 0; k < XR; ++k) {
            const int v = tid + k * MNT;
            if (v < V4) *reinterpret_cast<float4*>(&xs[slot * XSTRIDE + v * 4]) = xf[k];
        }
    };
    auto park_e = [&](int slot) {
#pragma unroll
        for (int k = 0; k < ER; ++k) {
            const int v = tid + k * MNT;
            if (v < EN) { ec[slot][0][v] = ef[0][k]; ec[slot][1][v] = ef[1][k]; }
        }
    };
    // accumulators of both node planes start at 0
    for (int i = tid; i < YALL; i += MNT) ys[i] = 0.f;
    // prologue: node planes ix0-1 and ix0, element plane ix0-1
    fetch(ix0 - 1, ix0 - 1);
    park_x((ix0 - 1) & 1);
    park_e((ix0 - 1) & 1);
    fetch(ix0, ix0);
    park_x(ix0 & 1);
    __syncthreads();

    for (int ex = ix0 - 1; ex < ix1; ++ex) {
        fetch(ex + 2, ex + 1);                                   // in flight during the contraction
        if (ex >= 0 && ex < n) {
            const float* xlo = xs + (ex & 1) * XSTRIDE;          // node plane ex
            const float* xhi = xs + ((ex + 1) & 1) * XSTRIDE;    // node plane ex + 1
            const float* xlane = (qx ? xhi : xlo) + qy * WZ * MR + col;      // B operand base of this lane
            float* ylo = ys + (ex & 1) * YSTRIDE + YBASE;
            float* yhi = ys + ((ex + 1) & 1) * YSTRIDE + YBASE;
            // Element rows of one parity share no node row: per phase each
            // wave takes one row and walks its EZ elements in z order.  The
            // MFMA outputs of element kz hold, in lanes 0-31 (gz = 0), its
            // contributions to node column kz - 1 and, in lanes 32-63 (gz = 1),
            // to node column kz, each for the lower (ya) and the upper (yb)
            // node plane.  One v_permlane32_swap of (ya, yb) gives column
            // kz - 1's lower- and upper-plane parts side by side (lanes 0-31
            // and 32-63) and, in a second register, column kz's, which waits
            // for the next element.  Each accumulator word then takes one
            // read-modify-write per phase: lane group q = 0, 1 the lower node
            // plane and rows gy = q, q = 2, 3 the upper plane and rows q - 2.
            // Node (qx, qy, 1) of element kz is node (qx, qy, 0) of element
            // kz + 1, so each B operand is read and split once.  Elements outside the
            // mesh have zero coefficients (staging) and add zeros; columns -1
            // and TZ are the tile's pads and are not written.  Both planes'
            // sums are always formed: a plane this block does not own is never
            // written out and its accumulator is cleared after the plane.
            float* const yq = ((q >> 1) ? yhi : ylo) + (q & 1) * YROW + col;
            auto run_row = [&](int ky) __attribute__((always_inline)) {
                const float* __restrict__ er_row = &ec[ex & 1][0][ky * EZ];
                const float* __restrict__ ek_row = &ec[ex & 1][1][ky * EZ];
                const float* __restrict__ x_row = xlane + ky * WZ * MR;
                float* __restrict__ y_row = yq + (ky - 1) * YROW;          // column 0 of its rows
                mr_v4u bu;                              // the B operand: two halves of split words
                {
                    unsigned w0, w1;
                    mrhs_split(x_row[0], w0, w1);
                    bu[0] = w0;
                    bu[1] = w1;
                    bu[2] = 0u;
                    bu[3] = 0u;
                }
                float carry = 0.f;
                auto batch = [&](auto k0c, auto nbc) __attribute__((always_inline)) {
                    constexpr int K0 = decltype(k0c)::value, NB = decltype(nbc)::value;
                    mr_v4f c[NB];
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        const int kz = K0 + i;             // a constant once the loop is unrolled
                        unsigned w0, w1;
                        mrhs_split(x_row[(kz + 1) * MR], w0, w1);
                        if (kz & 1) {
                            bu[0] = w0;
                            bu[1] = w1;
                        } else {
                            bu[2] = w0;
                            bu[3] = w1;
                        }
                        const mr_v4f zero = {0.f, 0.f, 0.f, 0.f};
                        c[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(amat[kz & 1], __builtin_bit_cast(mr_v8bf, bu),
                                                                       zero, 0, 0, 0);
                    }
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        const int kz = K0 + i;
                        const float ce_r = er_row[kz];
                        const float ce_k = ek_row[kz];
                        const float ya = MASS ? ce_r * c[i][0] : fmaf(ce_r, c[i][0], ce_k * c[i][2]);
                        const float yb = MASS ? ce_r * c[i][1] : fmaf(ce_r, c[i][1], ce_k * c[i][3]);
                        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(ya), __float_as_uint(yb),
                                                                         false, false);
                        const float cur = __uint_as_float(sw[0]) + carry;   // column kz - 1, complete
                        carry = __uint_as_float(sw[1]);                     // column kz, this element's part
                        if (kz >= 1) y_row[(kz - 1) * MR] += cur;
                    }
                };
                static_assert(EZ == 17, "two batches of the row's elements");
                batch(std::integral_constant<int, 0>{}, std::integral_constant<int, 9>{});
                batch(std::integral_constant<int, 9>{}, std::integral_constant<int, 8>{});
            };
#pragma unroll 1
            for (int cy = 0; cy < 2; ++cy) {
                const int ky0 = (cy - (y0 - 1)) & 1;
                const int nyc = (EY - ky0 + 1) >> 1;
#pragma unroll 1
                for (int kyi = wave; kyi < nyc; kyi += MNT / 64) run_row(ky0 + 2 * kyi);
                __syncthreads();
            }
        } else {
            __syncthreads();
        }
        {                                                        // node plane ex is complete
            const bool own = ex >= ix0;                          // (plane ix0 - 1 belongs to the previous chunk)
            float* yp = ys + (ex & 1) * YSTRIDE + YBASE;
            for (int i = tid; i < TY * TZ * (MR / 4); i += MNT) {
                const int node = i >> 2, part = i & 3;
                const int ty = node / TZ, tz = node - (node / TZ) * TZ;
                const int iy = y0 + ty, iz = z0 + tz;
                float4* src = reinterpret_cast<float4*>(yp + ty * YROW + tz * MR + part * 4);
                if (own && iy <= n && iz <= n) {
                    const bool bnd = ex == 0 || ex == n || iy == 0 || iy == n || iz == 0 || iz == n;
                    const float4 v = bnd ? make_float4(0.f, 0.f, 0.f, 0.f) : *src;
                    *reinterpret_cast<float4*>(Y + ((size_t)ex * S + (size_t)iy * nn1 + iz) * MR + part * 4) = v;
                    if (SUMS) {
                        // X of this node: window node (ty + 1, tz + 1) of plane ex (not yet replaced)
                        const float4 xv = *reinterpret_cast<const float4*>(
                            xs + (ex & 1) * XSTRIDE + ((ty + 1) * WZ + (tz + 1)) * MR + part * 4);
                        const float yv[4] = {v.x, v.y, v.z, v.w}, xw[4] = {xv.x, xv.y, xv.z, xv.w};
                        const double wn = SUMS == 2 ? w[(size_t)ex * S + (size_t)iy * nn1 + iz] : 0.0;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            acc_s[0][j] += (double)xw[j] * (double)yv[j];
                            if (SUMS == 2) {
                                acc_s[NK > 1 ? 1 : 0][j] += (double)yv[j];
                                acc_s[NK > 2 ? 2 : 0][j] += wn * (double)xw[j];
                                acc_s[NK > 3 ? 3 : 0][j] += (double)xw[j] * (double)xw[j];
                            }
                        }
                    }
                }
                *src = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        // with the sums, every thread must have read plane ex's window before
        // plane ex + 2 replaces it
        if (SUMS) __syncthreads();
        park_x(ex & 1);                                          // plane ex + 2 replaces plane ex
        park_e((ex + 1) & 1);                                    // element plane ex + 1
        __syncthreads();
    }
    if (SUMS) {
        // lanes of one wave with the same column group (lane & 3) combine, then
        // the block's waves are added in order; ys is free after the walk
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int off = 4; off < 64; off <<= 1) acc_s[k][j] += __shfl_xor(acc_s[k][j], off);
        double* red = reinterpret_cast<double*>(ys);            // [wave][k][column]
        constexpr int NW = MNT / 64;
        if (lane < 4) {
#pragma unroll
            for (int k = 0; k < NK; ++k)
#pragma unroll
                for (int j = 0; j < 4; ++j) red[(wave * NK + k) * MR + lane * 4 + j] = acc_s[k][j];
        }
        __syncthreads();
        if (tid < NK * MR) {
            const int k = tid / MR, c = tid - k * MR;
            double t = 0.0;
            for (int wv = 0; wv < NW; ++wv) t += red[(wv * NK + k) * MR + c];
            part[((size_t)k * MR + c) * ntiles + b] = t;
        }
    }
}

}  // namespace heat
