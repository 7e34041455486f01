/*
 * heat_oracle.h -- CPU oracle for the matrix-free Q1 hexahedral heat solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver src/heat_solver_het.cu (single rank: the MPI slab decomposition
 * collapses to one rank that owns every node).  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it, and only as the checker.
 * The product (mnhack8_amd/) never links, imports or runs anything here.
 *
 * Every function cites the reference lines it restates.
 */
#ifndef HEAT_ORACLE_H
#define HEAT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* element kernel variants */
#define HO_KERNEL_GAUSS  0   /* 2x2x2 Gauss on the fly, src/heat_solver_het.cu:310-338 */
#define HO_KERNEL_FAST   1   /* 8x8 reference matrices, src/heat_solver_het.cu:340-349 */
#define HO_KERNEL_STRUCT 2   /* closed-form Hamming-distance rows (the GPU's fast form) */

/* source models */
#define HO_SOURCE_GAUSSIAN 0 /* 1000 exp(-r^2/(2*0.1^2)), src/heat_solver_het.cu:229-237 */
#define HO_SOURCE_CONST    1 /* Q = const (new: f = 1 workloads) */

typedef struct ho_op ho_op;

/* setup: src/heat_solver_het.cu:756-816 (load_fields replaced by caller arrays) */
ho_op* ho_create(int n, double dt, const double* k, const double* rhoc,
                 int kernel, int colored, int verify, int source_mode,
                 double source_value, int detect);
void   ho_destroy(ho_op* op);
const char* ho_last_error(void);
void   ho_set_threads(ho_op* op, int nthreads);
void   ho_set_no_source(ho_op* op, int no_source);

long   ho_nn(const ho_op* op);
long   ho_ne(const ho_op* op);
/* copies of internal vectors: 0 diag, 1 F, 2 wM, 3 wA, 4 ones_int */
int    ho_get_vector(const ho_op* op, int which, double* out);
int    ho_set_vector(ho_op* op, int which, const double* in);   /* 0 diag, 1 F */
void   ho_get_ref_matrices(const ho_op* op, double* Mh, double* Kh);
void   ho_get_stats(const ho_op* op, double* out4); /* kmin kmax rhocmin rhocmax */

/* HeatOperator methods, src/heat_solver_het.cu:885-988 */
void   ho_apply_operator(ho_op* op, const double* x, double* y, double dt_eff,
                         int allow_fi, int project);
void   ho_build_rhs(ho_op* op, const double* Tn, double* b);
double ho_dot(ho_op* op, const double* x, const double* y);
void   ho_minmax(ho_op* op, const double* x, double* vmin, double* vmax);
void   ho_apply_preconditioner(ho_op* op, const double* r, double* z);

/* FP32 restatement of the structured operator (FP32 inputs, FP32 accumulation) */
void   ho_apply_operator_f32(ho_op* op, const float* x, float* y, double dt_eff,
                             int project);

/* fault injection: src/heat_solver_het.cu:181-186, 369-374 */
void   ho_arm_fault(ho_op* op, int level, long target, int bit);
void   ho_flip_bit(double* v, long idx, int bit);
void   ho_abft_take(ho_op* op, double* step_max, long* calls);

/* pcg_solve, src/heat_solver_het.cu:994-1039.  hist (optional) receives
 * (max_iter+1) rows of {rel_residual, alpha, beta}. */
int    ho_pcg_solve(ho_op* op, const double* b, double* x, int max_iter, double tol,
                    double* hist);
/* The same solve with the inner products summed in the GPU library's fused
 * two-launch order (tile / super-block partials, 256-wide trees). */
int    ho_pcg_solve_fused(ho_op* op, const double* b, double* x, int max_iter, double tol,
                          double* hist);
/* The GPU library's tile geometry (heat_set_pcg_geometry): WHICH 0 the fused
 * iteration, 1 the one-launch inner iteration; PLANES 0 restores the default
 * (16, 768).  Returns 0, or -1 for a bad argument. */
int    ho_set_pcg_geometry(ho_op* op, int which, int planes, int wave);
/* One-shot upset of the PCG state for the next fused solve: at iteration
 * ITER the stored copy of VEC (HO_SEU_*) at IDX gets bit BIT flipped after
 * the sums of the kernel that produced it (the refined solve: its FP32 inner
 * state, iterations counted over all corrections). */
#define HO_SEU_U  0
#define HO_SEU_R  1
#define HO_SEU_P  2
#define HO_SEU_Z  3
#define HO_SEU_AP 4
/* the refined solve's FP64 state, ITER = outer pass (0-based): x as the
 * pass's correction stores it, A x as the pass's apply stores it */
#define HO_SEU_X64  5
#define HO_SEU_AX64 6
void   ho_arm_pcg_fault(ho_op* op, int vec, long idx, int bit, int iter);
int    ho_pcg_fault_fired(const ho_op* op);
/* how the last refined solve ended: 1 tolerance met, 2 residual not finite, 0 pass limit */
int    ho_mixed_status(const ho_op* op);
/* The fused order's sums alone: super-block x.y, and the tile sums
 * [p.Ap (interior), 1_int.Ap, w.p, p.p]. */
double ho_fused_sb_dot(ho_op* op, const double* x, const double* y);
void   ho_fused_tile_sums(ho_op* op, const double* p, const double* Ap, const double* w, double out[4]);
/* FP32-refined solve (FP64 outer residual, FP32 inner Jacobi-PCG) in that order. */
int    ho_pcg_solve_mixed_fused(ho_op* op, const double* b, double* x, int max_outer, int max_inner,
                                double tol, double tol_inner, int* outer, int* inner_total);
int    ho_pcg_solve_mixed_cg1  (ho_op* op, const double* b, double* x, int max_outer, int max_inner,
                                double tol, double tol_inner, int* outer, int* inner_total);

/* initial conditions, src/heat_solver_het.cu:556-584.  mode 1 gaussian, 2 sine mode at t */
void   ho_initial_condition(ho_op* op, double* v, int mode, double t, double alpha);

/* the whole command-line program (src/heat_solver_het.cu:1355-1629) */
/* the MNHack7 prototype run: per-step CG iterations and ||T||_2; returns the step count or -1 */
int    ho_proto_run(int n, double t_final, double dt, int nthreads, int max_steps, int* iters, double* norms);
int    ho_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif
#endif
