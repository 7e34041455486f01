/*
 * heat.h -- C ABI of the MI355X matrix-free Q1 hexahedral heat solver.
 *
 * One handle owns one structured mesh ([0,1]^3, N cells per axis), its
 * per-element coefficients, the Jacobi diagonal, the source vector, the PCG
 * work vectors and the ABFT checksum vectors, all in device memory of the
 * current HIP device, plus one HIP stream on which every call is ordered.
 * A handle is not thread-safe.  Every entry point returns 0 on success and a
 * non-zero HEAT_E* code on failure; heat_last_error() then describes it.  No
 * entry point aborts the host process.
 *
 * Vector arguments are DEVICE pointers of length heat_nn() (nodes, x slowest:
 * id = ix*(N+1)^2 + iy*(N+1) + iz) holding FP64 values, unless noted.  They
 * may be handle-owned (heat_vector) or caller-owned (hipMalloc, a torch
 * tensor's data_ptr()).  Calls are asynchronous on the handle's stream except
 * those that return a scalar to the host, which synchronise it.
 *
 * Each declaration cites the reference interface it replaces
 * (src/heat_solver_het.cu of mariagarciapenalva/mnhack8).
 */
#ifndef MNHACK8_AMD_HEAT_H
#define MNHACK8_AMD_HEAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEAT_OK          0
#define HEAT_E_ARG       1   /* invalid argument */
#define HEAT_E_HIP       2   /* HIP runtime error */
#define HEAT_E_FIELD     3   /* non-positive or NaN coefficient */
#define HEAT_E_SELFCHECK 4   /* reference matrices failed their closed-form check */
#define HEAT_E_STATE     5   /* call not valid in the handle's state */

/* --kernel */
#define HEAT_KERNEL_GAUSS 0  /* 2x2x2 Gauss quadrature on the fly (elem_apply_gauss) */
#define HEAT_KERNEL_FAST  1  /* element-congruent closed-form rows (elem_apply_fast) */
/* --scatter */
#define HEAT_SCATTER_ATOMIC  0
#define HEAT_SCATTER_COLORED 1
/* source models */
#define HEAT_SOURCE_GAUSSIAN 0  /* 1000 exp(-|x-c|^2 / (2*0.1^2)) (source_at) */
#define HEAT_SOURCE_CONST    1  /* Q = source_value everywhere */
/* colored-scatter implementations (all give the same bits) */
#define HEAT_COLORED_GATHER 0   /* one thread per node, contributions summed in colour order */
#define HEAT_COLORED_SWEEP  1   /* eight element launches, one per colour */
#define HEAT_COLORED_MARCH  2   /* threads walk x-columns of nodes with a register window */
#define HEAT_COLORED_TILE   3   /* blocks walk x with LDS-staged plane tiles */
#define HEAT_COLORED_TDMA   4   /* as TILE, tiles copied to LDS asynchronously two planes ahead (default) */
#define HEAT_ATOMIC_ELEMENT 0   /* --scatter atomic: one thread per element, 8 global atomic adds each */
#define HEAT_ATOMIC_TILED   1   /* element tiles summed in LDS, global atomics only where tiles meet (default) */

/* PCG execution (both run the reference's algorithm; they differ in the fixed
 * order in which inner products are summed) */
#define HEAT_PCG_FUSED           0  /* three launches per iteration, tile/super-block order (default) */
#define HEAT_PCG_REFERENCE_ORDER 1  /* six launches per iteration, the reference's 256/1024 order */

/* handle-owned vectors for heat_vector() */
#define HEAT_VEC_T    0
#define HEAT_VEC_B    1
#define HEAT_VEC_R    2
#define HEAT_VEC_Z    3
#define HEAT_VEC_P    4
#define HEAT_VEC_AP   5
#define HEAT_VEC_DIAG 6
#define HEAT_VEC_F    7
#define HEAT_VEC_WM   8   /* M 1_int, unprojected (needs detect) */
#define HEAT_VEC_WA   9   /* (M + dt K) 1_int, unprojected (needs detect) */
#define HEAT_VEC_ONES 10  /* 1 on interior nodes, 0 on the boundary (needs detect) */

typedef struct heat_op heat_op;

typedef struct heat_config {
    int    n;              /* cells per axis (cube [0,1]^3, h = 1/n) */
    double dt;
    int    kernel;         /* HEAT_KERNEL_* */
    int    scatter;        /* HEAT_SCATTER_* */
    int    verify;         /* 1: source off (Grid::verify) */
    int    no_source;      /* 1: F is built but not added by build_rhs (Grid::no_source) */
    int    source_mode;    /* HEAT_SOURCE_* */
    double source_value;   /* Q for HEAT_SOURCE_CONST */
    int    detect;         /* 1: build ABFT vectors and check every projected apply */
} heat_config;

/* Last error message of the calling thread. */
const char* heat_last_error(void);

/* HeatOperator::setup + load_fields + setup_detectors
 * (src/heat_solver_het.cu:756-816, 711-753, 940-953).  k_cell and rhoc_cell
 * are HOST arrays of N^3 doubles, element e = ix*N*N + iy*N + iz. */
int heat_setup(const heat_config* cfg, const double* k_cell, const double* rhoc_cell,
               heat_op** out);
/* HeatOperator::cleanup (src/heat_solver_het.cu:818-824). */
int heat_cleanup(heat_op* op);

int heat_sizes(const heat_op* op, int64_t* nn, int64_t* ne);
/* kmin, kmax, rhocmin, rhocmax (HeatOperator::kmin... src/heat_solver_het.cu:658) */
int heat_field_stats(const heat_op* op, double out4[4]);
/* Device pointer of a handle-owned vector (HEAT_VEC_*). */
int heat_vector(heat_op* op, int which, double** dev_ptr);
/* hipStream_t of the handle, as void*. */
int heat_stream(heat_op* op, void** stream);
int heat_synchronize(heat_op* op);
int heat_upload(heat_op* op, double* dev_dst, const double* host_src, int64_t count);
int heat_download(heat_op* op, double* host_dst, const double* dev_src, int64_t count);
/* Select how --scatter colored is executed (HEAT_COLORED_*). */
int heat_set_colored_impl(heat_op* op, int impl);
/* Select how --scatter atomic is executed (new): HEAT_ATOMIC_ELEMENT is the
 * reference's operator_kernel with colour -1 (src/heat_solver_het.cu:351-378,
 * 892-893), HEAT_ATOMIC_TILED (default) the same element rows with the
 * tile-interior sums in LDS.  Both add in an unfixed order, so they agree
 * with each other and with the colour order to rounding. */
int heat_set_atomic_impl(heat_op* op, int impl);

/* HeatOperator::apply_operator (src/heat_solver_het.cu:885-921):
 * y = P (rhoc M + dt_eff k K) x, P zeroing Dirichlet rows when project != 0,
 * a pending level-R fault fires when allow_fi != 0, D1 accumulates when the
 * handle detects and project != 0.  dt_eff must be 0 or the handle's dt. */
int heat_apply_operator(heat_op* op, const double* x, double* y, double dt_eff,
                        int allow_fi, int project);
/* FP32 operator of the refined solve (new): x and y are FP32 device vectors,
 * FP32 coefficients and FP32 accumulation, summed in colour order. */
int heat_apply_operator_f32(heat_op* op, const float* x, float* y, double dt_eff, int project);
/* HeatOperator::apply_matvec (src/heat_solver_het.cu:955). */
int heat_apply_matvec(heat_op* op, const double* x, double* y);
/* HeatOperator::apply_preconditioner (src/heat_solver_het.cu:957-960): z = r / diag. */
int heat_apply_preconditioner(heat_op* op, const double* r, double* z);
/* HeatOperator::build_rhs (src/heat_solver_het.cu:963-971): b = P M Tn + dt F. */
int heat_build_rhs(heat_op* op, const double* Tn, double* b);
/* HeatOperator::dot / norm2 (src/heat_solver_het.cu:973-988): fixed-order
 * two-stage sum (256-node blocks, then a 1024-way strided sum and tree). */
int heat_dot(heat_op* op, const double* x, const double* y, double* out);
int heat_norm2(heat_op* op, const double* x, double* out);
/* HeatOperator::minmax (src/heat_solver_het.cu:924-936). */
int heat_minmax(heat_op* op, const double* x, double* vmin, double* vmax);
/* apply_dirichlet_value_kernel (src/heat_solver_het.cu:427-434). */
int heat_set_dirichlet(heat_op* op, double* v, double value);
/* y = x  (copy_kernel, src/heat_solver_het.cu:444-447). */
int heat_copy(heat_op* op, const double* x, double* y);
/* y += a x  (axpy_kernel, src/heat_solver_het.cu:436-439). */
int heat_axpy(heat_op* op, double a, const double* x, double* y);
/* initial conditions (src/heat_solver_het.cu:554-584, 1115-1116):
 * mode 0 zero, 1 Gaussian bump, 2 sine eigenmode exp(-3 pi^2 alpha t) sin sin sin. */
int heat_initial_condition(heat_op* op, double* v, int mode, double t, double alpha);

/* pcg_solve (src/heat_solver_het.cu:994-1039): Jacobi PCG on the handle's
 * operator, x0 = x, stop when ||r||/||b|| < tol or after max_iter iterations
 * or when |p.Ap| < 1e-30.  alpha/beta and every inner product stay on the
 * device; the host polls a convergence flag every few iterations.  *iters
 * receives the reference's return value.  hist (HOST, optional) receives
 * 3*(max_iter+1) doubles: row 0 {||r0||/||b||, NaN, NaN}, row i
 * {||r_i||/||b||, alpha_i, beta_i} for i = 1..*iters (the --diag-cg log). */
int heat_pcg_solve(heat_op* op, const double* b, double* x, int max_iter, double tol,
                   int* iters, double* hist);

/* FP32-refined solve (new; no reference counterpart): FP64 defect correction
 * x += e, r = b - A x (FP64 operator) around an FP32 Jacobi-PCG on A e = r
 * (FP32 operator and vectors, FP64 alpha/beta inner products) to the relative
 * inner tolerance tol_inner, until ||b - A x|| / ||b|| < tol, max_outer
 * corrections, or (fused order) an FP64 residual that is not finite.
 * *outer and *inner_iters receive the outer and the total inner iteration
 * counts. */
int heat_pcg_solve_mixed(heat_op* op, const double* b, double* x, int max_outer,
                         int max_inner, double tol, double tol_inner,
                         int* outer, int* inner_iters);

/* Inner solver of the fused-order refined solve (new; no reference
 * counterpart).  HEAT_INNER_ONE_LAUNCH (default): preconditioned CG in the
 * Chronopoulos-Gear arrangement, one launch per iteration (both inner
 * products of an iteration reduced together; the same iterates as the
 * standard PCG in exact arithmetic, other roundings).  HEAT_INNER_STANDARD:
 * the standard PCG in two kernels per iteration, the order of the batched
 * refined solves (heat_batch_step_mixed).  A solve with a PCG-state upset
 * armed, on a mesh above N = 644 or with max_inner above 2^20 - 3 runs the
 * standard iteration. */
#define HEAT_INNER_ONE_LAUNCH 0
#define HEAT_INNER_STANDARD   1
int heat_set_inner_solver(heat_op* op, int mode);

/* Level R: the next apply with allow_fi != 0 flips bit `bit` of element
 * `target`'s accumulator ye[0] before it is scattered, once
 * (src/heat_solver_het.cu:369-374, 902, 1137). */
int heat_arm_fault(heat_op* op, int level, int64_t target, int bit);
int heat_disarm_fault(heat_op* op);
/* Level G: XOR bit `bit` of v[idx] (flip_bit_global_kernel, src/heat_solver_het.cu:181-186). */
int heat_flip_bit(heat_op* op, double* v, int64_t idx, int bit);
/* D1 state accumulated since the last call (abft_step_max, abft_step_calls,
 * src/heat_solver_het.cu:642-643, 1204, 1233); reading resets it. */
int heat_abft_take(heat_op* op, double* step_max, int64_t* calls);
/* PCG execution mode (HEAT_PCG_*), new. */
int heat_set_pcg_mode(heat_op* op, int mode);

/* Single-event upset in the PCG state (new; the reference injects only into
 * T before build_rhs, level G, and into one element contribution of the
 * first matvec, level R: src/heat_solver_het.cu:181-186, 369-374).  Arms a
 * one-shot fault for the next heat_pcg_solve or heat_pcg_solve_mixed (fused
 * order only): at PCG iteration ITERATION (0-based; the refined solve counts
 * its inner iterations over all corrections), the kernel that produces
 * vector VEC stores node NODE with bit BIT flipped (FP64 words: 0..63; the
 * refined solve's FP32 inner state: 0..31).  Sums taken in that same kernel
 * see the clean value.  VEC: the iterate (x, or the inner correction e),
 * r, p, z or Ap. */
#define HEAT_SEU_U  0
#define HEAT_SEU_R  1
#define HEAT_SEU_P  2
#define HEAT_SEU_Z  3
#define HEAT_SEU_AP 4
/* The refined solve's FP64 state (heat_pcg_solve_mixed only; bits 0..63;
 * ITERATION counts its outer passes from 0): the iterate x as the
 * correction of that pass stores it, and A x as that pass's FP64 apply
 * stores it for the outer residual b - A x. */
#define HEAT_SEU_X64  5
#define HEAT_SEU_AX64 6
int heat_arm_pcg_fault(heat_op* op, int vec, int64_t node, int bit, int iteration);
/* 1 if the fault armed for the last solve fired (its iteration was reached). */
int heat_pcg_fault_fired(heat_op* op, int* fired);
/* How the last heat_pcg_solve_mixed ended (new): 1 the outer tolerance was
 * met, 2 the outer residual was not finite (the solve stops: no correction
 * can recover it), 0 the pass limit was reached. */
int heat_pcg_mixed_status(heat_op* op, int* status);

/* Sixteen right-hand sides at once (new; the reference applies its operator
 * to one vector, src/heat_solver_het.cu:885-921).  X and Y are FP32 device
 * arrays of nn x 16, node-major (X[node * 16 + r]).  Y_r = P (rho*c M +
 * dt_eff kappa K) X_r with FP32 coefficients (dt_eff = 0 or the handle's
 * dt); the element contraction runs on the matrix cores.  Each node's eight
 * element contributions are summed in a fixed order (element plane below,
 * then above, each in (y, z) parity order), not the colour order, so results
 * agree with the single-vector FP32 apply to FP32 rounding, not bitwise. */
int heat_apply_mrhs(heat_op* op, const float* X, float* Y, double dt_eff);
/* Mean time of REPS back-to-back multi-RHS applies (A, not M) on the handle's stream. */
int heat_time_mrhs(heat_op* op, const float* X, float* Y, int reps, double* avg_ms);

/* Batched ensembles (new; replaces the reference's one process per ensemble
 * member, scripts/ensemble/run_ensemble.py:53-104, each running
 * src/heat_solver_het.cu's main).  COUNT independent solves, each with its own
 * configuration and fields (fast kernel, colour-ordered scatter; detectors
 * on all solves or none), advanced together: every PCG iteration of all
 * solves is one set of launches over all their tiles, and a converged or
 * masked solve's blocks return at once.  Each solve's iterates are
 * bit-identical to the same solve run alone in the fused order.  With
 * detectors, each solve forms b and its first residual alone, as
 * heat_build_rhs and heat_pcg_solve do (D1 records, an armed level-R fault),
 * and its D1 record after a step equals the single solve's. */
typedef struct heat_batch heat_batch;
int heat_batch_setup(int count, const heat_config* cfgs, const double* const* k,
                     const double* const* rhoc, heat_batch** out);
int heat_batch_cleanup(heat_batch* b);
/* The handle of solve I (owned by the batch): upload T^0, download T, etc. */
int heat_batch_instance(heat_batch* b, int i, heat_op** op);
/* Per-solve mask (nonzero = active); masked solves keep their T. */
int heat_batch_set_active(heat_batch* b, const int* active);
int heat_batch_set_tuning(heat_batch* b, int pcg_chunk, int use_graph);
/* Nonzero ON (new; no reference counterpart): each batched PCG solve of
 * heat_batch_step and heat_batch_step_mixed runs as ONE persistent
 * (cooperative) launch whose resident blocks walk every iteration's phases
 * with grid-wide barriers between them, instead of graph-replayed chunks of
 * launches.  Same iterates, same bits. */
int heat_batch_set_persistent(heat_batch* b, int on);
/* A batch of COUNT right-hand sides on one mesh (new): the members share
 * OP's coefficients and diagonal (OP stays owned by the caller and must
 * outlive the batch) and own their T, b, PCG vectors and source vector F,
 * copied from the device array F (COUNT x nn doubles, member-major; NULL:
 * zero sources).  Steps and masks as for heat_batch_setup. */
int heat_batch_setup_rhs(heat_op* op, int count, const double* F, heat_batch** out);
/* Device pointer of vector WHICH (HEAT_VEC_T, _B, _R, _Z, _P, _AP or _F) of member I. */
int heat_batch_vector(heat_batch* b, int i, int which, double** ptr);
/* NSTEPS backward-Euler steps of every active solve: b = M T + dt F, then
 * Jacobi-PCG from x0 = T (src/heat_solver_het.cu:1308-1311).  ITERS
 * (optional) receives NSTEPS x COUNT iteration counts (-1: masked). */
int heat_batch_step(heat_batch* b, int nsteps, int max_iter, double tol, int* iters);
/* The detectors of every active member of a batch with detectors after a
 * step (new, no counterpart in the reference, which reads them for its one
 * solve, src/heat_solver_het.cu:1202-1234): D1[i] = the member's largest D1
 * ratio since the last reset (then reset), D2[i] = ||b - P A T|| / ||b|| from a
 * fresh projected apply (||b|| below 1e-300 taken as 1), TMIN[i] and TMAX[i]
 * = the extremes of T; the D1 record that the D2 apply makes is reset.  The
 * same values as heat_abft_take, heat_apply_operator, heat_copy, heat_axpy,
 * heat_norm2 and heat_minmax called in that order on the member, with all
 * members' work queued before one wait.  Masked members get NaN. */
int heat_batch_detect(heat_batch* b, double* d1, double* d2, double* tmin, double* tmax);
/* NSTEPS backward-Euler steps of every active solve of a batch of handles
 * (heat_batch_setup), each solve as heat_pcg_solve_mixed (FP32 inner
 * Jacobi-PCG in fused order inside FP64 defect correction to TOL, inner
 * tolerance TOL_INNER; new, no counterpart in the reference, whose solves are
 * FP64 only): every member bit-identical to that single solve.  INNER_ITERS
 * (optional) receives NSTEPS x COUNT inner-iteration totals (-1: masked). */
int heat_batch_step_mixed(heat_batch* b, int nsteps, int max_outer, int max_inner, double tol, double tol_inner,
                          int* inner_iters);
/* NSTEPS backward-Euler steps of a shared-mesh batch of exactly 16 right-hand
 * sides (heat_batch_setup_rhs) solved together (new, no counterpart in the
 * reference): FP32 Jacobi-PCG on all sixteen columns with the matrix-core
 * operator of heat_apply_mrhs, per-column alpha, beta and convergence, inside
 * per-column FP64 defect correction to TOL (inner tolerance TOL_INNER).  With
 * a handle built with detectors, every inner apply is checked by D1 per
 * column.  INNER_ITERS and D1MAX (optional) receive NSTEPS x 16 inner-iteration
 * totals and largest D1 ratios. */
int heat_batch_step_mrhs(heat_batch* b, int nsteps, int max_outer, int max_inner, double tol, double tol_inner,
                         int* inner_iters, double* d1max);
/* Launch tuning (new): PCG iterations per convergence poll (1..256) and
 * whether a chunk is replayed as a captured hipGraph. */
int heat_set_tuning(heat_op* op, int pcg_chunk, int use_graph);
/* x-planes walked per thread block by HEAT_COLORED_MARCH / _TILE (2, 4, 8,
 * 16 or 32; the tiled kernel takes 4, 8 or 16). */
int heat_set_march_chunk(heat_op* op, int planes);
/* Node tile of HEAT_COLORED_TILE: 0 = 8 x 32 (y x z), 1 = 4 x 64, 2 = 16 x 16. */
int heat_set_tile_shape(heat_op* op, int shape);
/* Benchmark helper (new): time `reps` back-to-back apply_matvec launches
 * (y = P A x, dt_eff = dt, no fault, no D1) with HIP events on the handle's
 * stream.  With flush_bytes > 0 a buffer of that size is overwritten before
 * every rep, outside the timed interval, to evict the caches; each rep is then
 * timed alone.  *avg_ms and *min_ms receive the mean and minimum per apply. */
int heat_time_matvec(heat_op* op, const double* x, double* y, int reps, int64_t flush_bytes,
                     double* avg_ms, double* min_ms);
/* Benchmark helper (new, no reference counterpart): the bandwidth ceiling of
 * the apply's memory traffic.  Times `reps` back-to-back launches of a plain
 * streaming kernel that moves exactly the FP64 apply's algorithmic bytes --
 * reads x (nn values) and the two element coefficient arrays (ne values
 * each), writes y (nn values) -- with 16-byte loads and stores and no
 * stencil, halo or reduction.  y receives x plus a mix of the coefficients
 * (a value of no meaning).  *avg_ms receives the mean per launch. */
int heat_time_stream(heat_op* op, const double* x, double* y, int reps, double* avg_ms);
/* Autotuner (new): time every colour-ordered apply schedule (one thread per
 * node, and plane marching with 2..32 planes per block) on this handle's
 * mesh with HIP events, keep the fastest for all later applies, and report it
 * (*impl = HEAT_COLORED_*, *planes, *shape, *best_ms per apply). */
int heat_autotune(heat_op* op, int reps, int* impl, int* planes, int* shape, double* best_ms);
/* Tile geometry of the PCG kernels (new; the reference's solve has no tiles):
 * WHICH = 0 the fused iteration (FP64 solves, the standard FP32 inner solve
 * and batches), 1 the one-launch FP32 inner iteration.  Each (y, z) column of
 * tiles is split into ceil(planes_in_column / PLANES) chunks of whole plane
 * pairs, raised until the tiles fill whole waves of WAVE tiles (WAVE = 0: no
 * fill) but not to chunks shorter than 8 planes.  PLANES = 0 restores the
 * default (16 planes, waves of 768 tiles).  The chunking fixes the order of
 * the per-tile inner-product partials, so a solve's bits depend on it (the
 * oracle takes the same two numbers).  HEAT_E_STATE for a handle whose
 * geometry a batch holds (heat_batch_set_geometry). */
int heat_set_pcg_geometry(heat_op* op, int which, int planes, int wave);
/* The geometry in force: PLANES, WAVE and the resulting chunks per column and
 * tile count (any pointer may be NULL). */
int heat_pcg_geometry(const heat_op* op, int which, int* planes, int* wave, int* chunks, int* tiles);
/* Autotuner of the PCG tile geometry (new): time ITERS iterations of the
 * solve of geometry WHICH (0: heat_pcg_solve in fused order; 1: one outer
 * pass of heat_pcg_solve_mixed with the one-launch inner iteration) for
 * PLANES in {4, 8, 12, 16, 24, 32, 64} and WAVE in {0, 768} on scratch
 * vectors, keep the fastest and report it (*best_ms: per iteration).  The
 * work vectors r, z, p and Ap are overwritten; the D1 record is kept. */
int heat_autotune_pcg(heat_op* op, int which, int iters, int* planes, int* wave, double* best_ms);
/* The fused geometry of every solve of a batch (new): as
 * heat_set_pcg_geometry(member, 0, PLANES, WAVE) for each member (for a
 * shared-mesh batch, the mesh handle), with the batch's tile layout remade. */
int heat_batch_set_geometry(heat_batch* b, int planes, int wave);
/* The fused geometry of member I (see heat_pcg_geometry). */
int heat_batch_geometry(heat_batch* b, int i, int* planes, int* wave, int* chunks, int* tiles);
/* Autotuner of a batch's geometry (new): time one step of ITERS iterations
 * (MODE 0: heat_batch_step, 1: heat_batch_step_mixed with one outer pass),
 * every member starting from its T, for the candidates of
 * heat_autotune_pcg; keep the fastest (*best_ms: per iteration).  Every T is
 * restored.  HEAT_E_STATE for a batch with detectors. */
int heat_batch_autotune(heat_batch* b, int mode, int iters, int* planes, int* wave, double* best_ms);
/* The D1 threshold (chk_tol, src/heat_solver_het.cu:637) used by the driver. */
int heat_abft_threshold(heat_op* op, double* chk_tol);

/* heat_dot without a handle and without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, such as an external CG loop built on heat_apply_operator,
 * tight serving loops, harnesses that forbid hipMalloc).  *out (a DEVICE
 * address) receives sum_i x[i] y[i] over the n device values of x and y, in
 * heat_dot's fixed order (256-value block partials, then a 1024-way strided
 * sum and tree), so the bits equal heat_dot's for the same vectors.  The block
 * partials live in a file-scope device array of HEAT_STATIC_DOT_MAX_BLOCKS
 * doubles (256 KiB, held by each GPU the library runs on until the process
 * ends): n must be in 0..256 * HEAT_STATIC_DOT_MAX_BLOCKS (8,388,608 values;
 * N <= 202), else HEAT_E_ARG.  Queued on STREAM (a hipStream_t; NULL: the
 * null stream) with no host wait; no hipMalloc, hipFree or memset.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race. */
#define HEAT_STATIC_DOT_MAX_BLOCKS 32768
int heat_dot_static(const double* x, const double* y, int64_t n, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
