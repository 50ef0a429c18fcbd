/*
 * libdfmi — MI355X-native SIMP inner loop (hex8 K.u, GMG-PCG, compliance,
 * sensitivity, density filter, OC/MMA updates) behind a plain C ABI.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero status on failure;
 *     dfmi_last_error() then holds a message (thread-local).
 *   - Pointers named *_dev are device pointers (e.g. torch tensor data_ptr());
 *     the caller owns them.  The library owns its hierarchy and workspace.
 *   - Work is queued on the stream set by dfmi_set_stream (default: the null
 *     stream).  Functions that return host values synchronise that stream.
 *   - Flat DOF vectors use the reference numbering 3*node + component with nodes
 *     in C order over (nx+1, ny+1, nz+1); a multi-vector with k columns is
 *     stored column after column: v[r * ndof + dof].
 *   - Element fields are (nx, ny, nz) in C order.
 *   - One handle per thread; calls on one handle must not overlap.
 */
#ifndef DFMI_H
#define DFMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dfmi_mg dfmi_mg;
typedef struct dfmi_filter dfmi_filter;
typedef struct dfmi_net dfmi_net;

/* Settings of deep_frame/topology_multigrid.py:8-10 (MULTIGRID) that the solver uses. */
typedef struct dfmi_settings {
  int32_t coarsest_dofs;     /* 80000: add levels while the next-coarser level has more DOFs */
  int32_t max_levels;        /* 8 */
  int32_t levels;            /* 0 = choose from coarsest_dofs/max_levels, else force this count */
  int32_t chebyshev_degree;  /* 3 */
  double chebyshev_ratio;    /* 20.0 */
  double growth;             /* 2.0: degree at level l = round(degree * growth^l) */
  int32_t power_iterations;  /* 20 */
  int32_t precision;         /* V-cycle arithmetic: 0 = float64, 1 = float32 */
  double tolerance;          /* 1e-8, relative residual per column */
  int32_t max_iterations;    /* 2000 */
  double dead_ratio;         /* 1e-12 */
  int32_t seed;              /* power-iteration start vector seed */
  int32_t refine;            /* 0 = flexible PCG with FP64 outer operator (reference scheme);
                                1 = FP32 PCG with reliable FP64 updates of solution and residual */
  double inner_tolerance;    /* refine = 1: reliable-update factor (FP32 residual drop between FP64 updates) */
  int32_t fine_smoother;     /* precision = 1: fine-level smoother and residual of the V-cycle on the FP32 vector
                                units (0, default) or on the matrix cores (1, split-bf16 MFMA); the Krylov
                                operator is always the FP32 / FP64 vector kernel */
  int32_t retries;           /* 2: rounds of FP64 re-solves of columns that did not converge (warm start when
                                their relative residual is below 1), topology_multigrid.py:376-406 */
  int32_t coarse_krylov;     /* 0 (default, the reference's V-cycle) or 1 or 2: the fine level's coarse-grid
                                correction solves the level-1 system by that many flexible-CG steps, each
                                preconditioned by the level-1 V-cycle, instead of one level-1 V-cycle (a
                                Krylov "K-cycle"; needs three or more levels, and a flexible outer PCG, which
                                both solve schemes are; hierarchies with a free rigid kernel keep the V-cycle) */
  int32_t block_krylov;      /* refine = 1: 0 (default) solves the columns as separate PCG recurrences, 1 as one
                                breakdown-free block PCG over up to 48 columns (more columns, or a free rigid
                                kernel, keep the separate recurrences) */
  int32_t smoother;          /* MULTIGRID["smoother"] (topology_multigrid.py:284-305): 0 = "chebyshev" (default),
                                1 = "jacobi" (damped, step damping / lambda_max, round(sweeps * growth^l) sweeps) */
  int32_t sweeps;            /* 2: Jacobi sweeps on the fine level */
  double damping;            /* 1.0: Jacobi damping */
  int32_t cycle;             /* MULTIGRID["cycle"] (topology_multigrid.py:312): 0 = "V" (default), 1 = "W" (two
                                coarse-grid corrections on every level whose coarser level is not the coarsest) */
  int32_t coarse;            /* MULTIGRID["coarse"] (topology_multigrid.py:160-166): 0 = "galerkin" (default),
                                1 = "rediscretize" (coarse moduli by 2x2x2 averaging, Ke scaled by the spacing
                                ratio, mask injected from the finer level and cut by the dead-DOF ratio) */
  int32_t krylov_levels;     /* coarse_krylov > 0: the coarse-grid corrections of levels 0 .. krylov_levels - 1
                                (1, default: the fine level only) are Krylov solves on the next coarser level, so
                                the V-cycles inside a Krylov solve use one again (a recursive K-cycle); levels
                                whose coarser level is the coarsest keep the direct solve */
  int32_t chebyshev_kind;    /* smoother = chebyshev: 1 (default) the reference's first-kind polynomial on
                                [1.1 lambda_max / chebyshev_ratio, 1.1 lambda_max] (topology_multigrid.py:291-304);
                                4 the fourth-kind polynomial (Lottes 2023), which needs only the upper bound
                                1.1 lambda_max: error factor sin((2k+1) asin(sqrt t)) / ((2k+1) sqrt t) at
                                t = eigenvalue of Dinv A / (1.1 lambda_max) after k steps; no counterpart in the
                                reference */
} dfmi_settings;

/* Per-column outcome of a solve. */
typedef struct dfmi_report {
  int32_t iterations;        /* PCG iterations (all outer steps for refine = 1) */
  double relative_residual;  /* true FP64 |b - K x| / |b| at exit */
  int32_t converged;         /* relative_residual < tolerance */
  int32_t retried;           /* FP64 retry rounds this column took (0 if the first solve converged) */
} dfmi_report;

const char* dfmi_last_error(void);
void dfmi_default_settings(dfmi_settings* out);

/* GeometricMultigrid(shape, spacing, ke, settings) — deep_frame/topology_multigrid.py:47-79.
 * ke: 24x24 row-major FP64 element stiffness at unit modulus, reference DOF order
 * (3*corner + component, corners as deep_frame/topology_optimization.py:246-249). */
int dfmi_create(const int32_t shape[3], const double spacing[3], const double* ke, const dfmi_settings* settings,
                dfmi_mg** out);
int dfmi_destroy(dfmi_mg* mg);
int dfmi_set_stream(dfmi_mg* mg, void* hip_stream);
/* Level count and padded element shape (topology_multigrid.py:58-63). */
int dfmi_levels(dfmi_mg* mg, int32_t* levels, int32_t padded[3]);

/* GeometricMultigrid.update(moduli) — topology_multigrid.py:97-100.  moduli_dev: FP64 element
 * field (E_e, 0 = inactive).  Drops every cached hierarchy. */
int dfmi_update(dfmi_mg* mg, const double* moduli_dev);

/* GeometricMultigrid.hierarchy(key, fixed) — topology_multigrid.py:145-208: masks, Galerkin
 * levels, lambda_max, coarsest factor.  fixed_dev: uint8 per flat DOF (1 = fixed).  slot < 8
 * plays the role of the reference's cache key.  setup_seconds may be NULL. */
int dfmi_build_hierarchy(dfmi_mg* mg, int32_t slot, const uint8_t* fixed_dev, double* setup_seconds);

/* Hierarchy of a fixture whose Dirichlet DOFs leave rigid modes free (topology_multigrid.py:151-195,
 * 233-246): coefficients (6 x nk row-major, 1 <= nk <= 6; nk = 0 is dfmi_build_hierarchy) combine the
 * rigid modes about the centre shape * spacing / 2 (translations x, y, z; rotations (0, -z, y), (z, 0, -x),
 * (-y, x, 0)) into the free kernel.  The solves then project it out of right-hand sides, true residuals and
 * V-cycle outputs, and the coarsest level pins nk DOFs (pivoted QR) and applies P A^-1 P; such slots solve
 * with the flexible FP64 PCG (refine is ignored). */
int dfmi_build_hierarchy_kernel(dfmi_mg* mg, int32_t slot, const uint8_t* fixed_dev, const double* coefficients,
                                int32_t nk, double* setup_seconds);

/* rhs_kernel_component of solve (topology_multigrid.py:355-357): nk of the slot, and for k flat FP64 columns
 * (rhs_dev may be NULL) component[r] = |B^T M b_r| / |M b_r| with B the orthonormal fine kernel. */
int dfmi_kernel_component(dfmi_mg* mg, int32_t slot, const double* rhs_dev, int32_t k, int32_t* nk,
                          double* component);

/* GeometricMultigrid.solve(key, fixed, forces) for a built slot — topology_multigrid.py:336-375
 * (batched flexible PCG :407-454, FP64 retries of failed columns :376-406).  rhs_dev, x_dev: k flat FP64
 * columns.  report: host array of k. */
int dfmi_solve(dfmi_mg* mg, int32_t slot, const double* rhs_dev, double* x_dev, int32_t k, dfmi_report* report);

/* GeometricMultigrid.operator / product — topology_multigrid.py:105-109, :455-457.
 * slot < 0: unmasked K(E) x; slot >= 0: M K M x with that hierarchy's fine mask.
 * precision: 0 = FP64, 1 = FP32, 2 = FP32 vectors with split-bf16 matrix-core products (the V-cycle's
 * fine-level kernel); inputs and outputs stay FP64. */
int dfmi_apply(dfmi_mg* mg, int32_t slot, int32_t precision, const double* x_dev, double* y_dev, int32_t k);

/* Unit-density element mass me (24x24 row-major FP64, reference DOF order), the `me` argument of
 * GeometricMultigrid(shape, spacing, ke, settings, me) (topology_multigrid.py:47-79; hexahedron_matrices,
 * topology_optimization.py:269-299).  Must be the same scalar matrix for each component. */
int dfmi_set_mass(dfmi_mg* mg, const double* me);

/* Mass operator of GeometricMultigrid.eigenpairs (topology_multigrid.py:491): y = M (sum_e m_e Me +
 * diag(lumped)) M x with a built slot's fine mask M; mass_dev: element field (Ne), lumped_dev: flat DOF
 * vector or NULL; x_dev, y_dev: k flat FP64 columns. */
int dfmi_mass_apply(dfmi_mg* mg, int32_t slot, const double* mass_dev, const double* lumped_dev, const double* x_dev,
                    double* y_dev, int32_t k);

/* x *= M in place for k flat FP64 columns, M the built slot's fine mask (free DOFs of active nodes): the
 * `grid * mask` of GeometricMultigrid.eigenpairs (topology_multigrid.py:501-503, :518). */
int dfmi_mask(dfmi_mg* mg, int32_t slot, double* x_dev, int32_t k);

/* Sensitivity of ModalConstraint.aggregate (deep_frame/topology_optimization.py:956-964) for k (1..16)
 * flat FP64 mode shapes u_dev (M-normalised): out_dev[e] = sum_r weight[r] (stiffness_slope_dev[e] u_re^T Ke
 * u_re - value[r] mass_slope_dev[e] u_re^T Me u_re) / target; weight and value are host arrays of k, the
 * slopes element fields (zero on inactive elements).  Needs dfmi_set_mass. */
int dfmi_modal_sensitivity(dfmi_mg* mg, const double* u_dev, int32_t k, const double* weight, const double* value,
                           const double* stiffness_slope_dev, const double* mass_slope_dev, double target,
                           double* out_dev);

/* GeometricMultigrid.element_energy — topology_multigrid.py:458-459: energy_dev[r*Ne + e] = u_e^T Ke u_e. */
int dfmi_element_energy(dfmi_mg* mg, const double* u_dev, int32_t k, double* energy_dev);

/* solve_elasticity / HexElasticity._collect — topology_multigrid.py:568-579,
 * deep_frame/topology_optimization.py:721-734, :707: compliance_dev[r] = factor * f_r . u_r and
 * sensitivity_dev[e] = -dmoduli[e] * factor * sum_r u_re^T Ke u_re. */
int dfmi_compliance_sensitivity(dfmi_mg* mg, const double* u_dev, const double* f_dev, int32_t k,
                                const double* dmoduli_dev, double factor, double* compliance_dev,
                                double* sensitivity_dev);

/* Operator of level `level` of a built slot on its padded node grid (x, y: k columns of
 * 3 * nodes_l FP64, node-major); level 0 is M K M, coarser levels the Galerkin P^T A P
 * (deep_frame/topology_multigrid.py:110-117).  For tests and diagnostics. */
int dfmi_level_apply(dfmi_mg* mg, int32_t slot, int32_t level, const double* x_dev, double* y_dev, int32_t k);

/* One application of the V-cycle preconditioner (topology_multigrid.py:318-324) to k flat FP64
 * columns, in the handle's V-cycle precision.  For tests and diagnostics. */
int dfmi_precondition(dfmi_mg* mg, int32_t slot, const double* r_dev, double* z_dev, int32_t k);

/* Transfer between level `level` and level + 1 of a built slot on padded node grids:
 * direction 0 = full-weighting restriction masked by the coarse mask, 1 = trilinear prolongation
 * masked by the fine mask (topology_multigrid.py:314-316).  For tests and diagnostics. */
int dfmi_transfer(dfmi_mg* mg, int32_t slot, int32_t level, int32_t direction, const double* in_dev,
                  double* out_dev, int32_t k);

/* Average duration (ms, HIP events on the handle's stream) of `reps` launches of the operator kernel of
 * `level` (0 = matrix-free fine level, >= 1 = Galerkin stencil level, not the coarsest) on k columns of a
 * built slot, precision 0 = FP64, 1 = FP32, 2 = FP32 on the matrix cores (fine level).  For benchmarks. */
int dfmi_time_apply(dfmi_mg* mg, int32_t slot, int32_t level, int32_t precision, int32_t k, int32_t reps,
                    double* ms_per_launch);

/* Autotuner of the matrix-core fine-level kernel (no counterpart in the reference, whose fine operator is
 * a cuDNN convolution, deep_frame/topology_multigrid.py:101-109): times each candidate workgroup tile on k
 * columns of a built slot with HIP events and keeps the fastest for k.  tile[0] = element rows per tile,
 * tile[1] = node planes per workgroup along x; ms_per_launch = the winner's time.  Either may be NULL.
 * With tile[0] > 0 on entry the given tile is set for k without timing (tile[0] in {8, 16}). */
int dfmi_tune_fine(dfmi_mg* mg, int32_t slot, int32_t k, int32_t tile[2], double* ms_per_launch);

/* Diagnostics of a built slot: lambda_max per smoothed level and the coarsest live DOF count. */
int dfmi_diagnostics(dfmi_mg* mg, int32_t slot, double* lambda_max, int32_t* coarsest_dofs);

/* report["history"] of GeometricMultigrid.solve (topology_multigrid.py:421-426, :400, :367): the largest
 * relative recursive residual over the active columns after every PCG iteration of the handle's last
 * dfmi_solve, FP64 retries appended.  Copies min(count, capacity) values into out (host, may be NULL). */
int dfmi_history(dfmi_mg* mg, double* out, int32_t capacity, int32_t* count);

/* hierarchy["kernel_check"] (topology_multigrid.py:325-335, :206): for a slot with a free rigid kernel, per
 * level |A_l B_l| / |A_l R_l| with B_l the orthonormal masked kernel modes of the level and R_l uniform [0, 1)
 * values on its live DOFs, in the V-cycle precision; count = levels (0 for a clamped slot).  out: host array
 * of at least the level count, may be NULL. */
int dfmi_kernel_check(dfmi_mg* mg, int32_t slot, double* out, int32_t* count);

/* InterfaceCovariance.matrix (deep_frame/topology_problem.py:639-646) and the Gram matrices of
 * GeometricMultigrid.eigenpairs (topology_multigrid.py:462-500): out_dev[i * kb + j] =
 * sum_n a_dev[i * n + n'] b_dev[j * n + n'] for ka load columns and kb solution columns (1..64 each), FP64,
 * deterministic.  work_dev holds work_blocks * ka * kb doubles (per-block partial sums). */
int dfmi_gram(const double* a_dev, const double* b_dev, int64_t n, int32_t ka, int32_t kb, double* out_dev,
              double* work_dev, int32_t work_blocks, void* hip_stream);

/* dfmi_gram without the work buffer, for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference.  The
 * partial sums live in a file-scope 8 MiB device array of the library (256 workgroups x 64 x 64 doubles, one
 * per GPU, held until the process ends): ka, kb <= 64.  Calls that overlap on one GPU share that array and
 * race; calls in one stream are ordered. */
int dfmi_gram_static(const double* a_dev, const double* b_dev, int64_t n, int32_t ka, int32_t kb, double* out_dev,
                     void* hip_stream);

/* Combined displacement fields of InterfaceCovariance.measure / energy (topology_problem.py:647-652, 675-684)
 * and the block combinations of GeometricMultigrid.eigenpairs (topology_multigrid.py:476-480):
 * y_dev[m * n + i] (+ y_dev[m * n + i] if accumulate) = sum_r u_dev[r * n + i] m_dev[r * mcols + m] for k
 * (1..64) columns, FP64; y must not overlap u. */
int dfmi_combine(const double* u_dev, int64_t n, int32_t k, const double* m_dev, int32_t mcols, double* y_dev,
                 int32_t accumulate, void* hip_stream);

/* Passes over the DOFs of the block PCG recurrence that dfmi_solve runs with settings.block_krylov (no
 * counterpart in the reference, whose _pcg, topology_multigrid.py:407-454, treats the columns separately;
 * exported for tests and for callers that run their own block recurrence).  Operands are FP32 column-major
 * multi-vectors (column j at dev + j * n), at most 48 columns; coefficient matrices are FP32 with row stride 48
 * (entries past the used columns zero).  Sums over DOFs are FP32 within a wave's share and FP64 across waves
 * and workgroups, deterministic.
 *   dfmi_block_work_doubles: size of the work buffer of a Gram with ka x kb outputs and of an update;
 *   dfmi_block_gram: out_dev (ka x (kb1 + kb2), FP64 row-major) = A^T [B1 | B2]; b2_dev may be NULL with kb2 = 0,
 *     otherwise ceil(ka / 16), ceil(kb1 / 16) and ceil(kb2 / 16) must be equal (and so must the first two);
 *   dfmi_block_combine: Y = X + U M for kin input and kout output columns (y may alias x, not u);
 *   dfmi_block_update: D += V M, R -= W M (k columns) and rr_dev[j] = |R_j|^2 (FP64) after the update;
 *   dfmi_block_combine2: Y0 = X0 + U M0 and Y1 = X1 + U M1 (k columns each) in one pass over U (each y may alias
 *     its x, neither may alias u);
 *   dfmi_block_combine_flush: D += U M (k columns), then for every column j with flag_dev[j] != 0 also
 *     x_j += scale_dev[j] D_j into the FP64 multi-vector x (column j at x_dev + j * n) and D_j = 0, in the same
 *     pass (the reliable update of the block recurrence); d must not alias u. */
int dfmi_block_work_doubles(int64_t n, int32_t ka, int32_t kb, int64_t* count);
int dfmi_block_gram(const float* a_dev, int32_t ka, const float* b1_dev, int32_t kb1, const float* b2_dev,
                    int32_t kb2, int64_t n, double* out_dev, double* work_dev, void* hip_stream);
int dfmi_block_combine(const float* u_dev, const float* x_dev, float* y_dev, const float* m_dev, int32_t kin,
                       int32_t kout, int64_t n, void* hip_stream);
int dfmi_block_update(const float* v_dev, float* d_dev, const float* w_dev, float* r_dev, const float* m_dev,
                      int32_t k, int64_t n, double* rr_dev, double* work_dev, void* hip_stream);
int dfmi_block_combine2(const float* u_dev, const float* x0_dev, float* y0_dev, const float* m0_dev,
                        const float* x1_dev, float* y1_dev, const float* m1_dev, int32_t k, int64_t n,
                        void* hip_stream);
int dfmi_block_combine_flush(const float* u_dev, float* d_dev, const float* m_dev, int32_t k, int64_t n,
                             double* x_dev, const double* scale_dev, const int32_t* flag_dev, void* hip_stream);

/* SIMP interpolation E = E0 (eps + (1 - eps) rho^p), dE — topology_optimization.py:657-661;
 * inactive cells (active_dev[e] == 0) get 0. */
int dfmi_simp(const double* rho_dev, const uint8_t* active_dev, int64_t n, double young, double penal, double eps,
              double* moduli_dev, double* dmoduli_dev, void* hip_stream);

/* DensityMap.project / fields (deep_frame/topology_optimization.py:1145-1171): Heaviside projection of the
 * filtered field at count (1..3) thresholds eta (host array; robust: eroded eta + delta, intermediate eta,
 * dilated eta - delta) -> physical_dev[q * n + i] (clipped to [0, 1]) and derivative_dev[q * n + i];
 * beta = 0 is the identity; flags_dev bit 0 preserve (physical 1), bit 1 forbidden (physical 0), bit 2
 * free (derivative kept, else 0). */
int dfmi_project(const double* filtered_dev, const uint8_t* flags_dev, int64_t n, double beta, const double* eta,
                 int32_t count, int32_t robust, double* physical_dev, double* derivative_dev, void* hip_stream);

/* ConeFilter / DensityMap.filtered / pullback — topology_optimization.py:1082-1092, 1135-1142,
 * 1171-1178.  allowed/preserve: uint8 per element (host). */
int dfmi_filter_create(const int32_t shape[3], const double spacing[3], double radius, const uint8_t* allowed,
                       const uint8_t* preserve, dfmi_filter** out);
int dfmi_filter_destroy(dfmi_filter* f);
int dfmi_filter_set_stream(dfmi_filter* f, void* hip_stream);
int dfmi_filter_forward(dfmi_filter* f, const double* x_dev, double* out_dev);
int dfmi_filter_pullback(dfmi_filter* f, const double* sens_dev, const double* dproj_dev, double* out_dev);
/* Corridor blend of DensityMap (topology_optimization.py:1125-1129, :1142, :1176): a second cone of radius
 * `radius` whose normalised field is blended in with per-cell weights w = clip(weights, 0, 1) on allowed cells
 * (host array of Ne): forward (1 - w) x~ + w x~_wide, pullback through both cones.  Without a positive weight
 * on an allowed cell the filter stays plain. */
int dfmi_filter_set_corridor(dfmi_filter* f, double radius, const double* weights);
/* Sum of the filtered field over elements with weights (NULL = 1): the volume of DensityMap.volume
 * with the identity projection (topology_optimization.py:1148-1151). */
int dfmi_filter_volume(dfmi_filter* f, const double* x_dev, double* volume);

/* _oc_update — topology_optimization.py:1197-1233 (volume through the filter). */
int dfmi_oc_update(dfmi_filter* f, const double* x_dev, const double* dc_dev, const double* dv_dev, double target,
                   double xmin, double move, double* out_dev);

/* One MMA step (Svanberg's method as MMAOptimizer.step calls mmapy.mmasub, deep_frame/topology_optimizers.py:55-67,
 * with a0 = 1, a = 0, c = 1e4, d = 1, bounds [0, 1]) for m (1..16) constraints g_j <= 0 (m = len(g), :58):
 * g host array of m, dg_dev the m x n constraint gradients row-major (device), f0/df0 and g/dg multiplied by
 * `scale` inside; state arrays low/upp/xold1/xold2 (n, device) are updated in place; x_dev receives the new
 * design. */
int dfmi_mma_step(int64_t n, int32_t m, int32_t iteration, double* x_dev, double* xold1_dev, double* xold2_dev,
                  double* low_dev, double* upp_dev, double f0, const double* df0_dev, const double* g,
                  const double* dg_dev, double move, double scale, void* hip_stream);

/* ---- Learned coarse correction (no counterpart in the reference; the nearest pattern is the
 * hand-written NumPy MLP with backward and Adam of deep_frame/topology_neural.py:59-113).
 * A 4-level 3D U-Net (channels 16-32-64-64, BF16 weights, FP32 activations and accumulation) on the
 * level-1 node grid maps the level-1 residual g = R (b - A x) and a stiffness feature to the correction
 * delta = Dinv g + Dinv^(1/2) s N(Dinv^(1/2) g / s, f) that the fp32 V-cycle adds, with an energy-optimal step per column,
 * between the fine coarse-grid correction and the post-smoother.  Parameters: one flat FP32 array
 * (per layer: weights [c_out][c_in][27], then biases; up layers [c_coarse][c_fine][27]). */
int dfmi_net_create(dfmi_net** out);
int dfmi_net_destroy(dfmi_net* net);
int dfmi_net_set_stream(dfmi_net* net, void* hip_stream);
int dfmi_net_param_count(dfmi_net* net, int64_t* count);
/* Load FP32 master weights (device) and reset the Adam state. */
int dfmi_net_set_params(dfmi_net* net, const float* params_dev);
/* which = 0: master weights, 1: gradients of the last backward / train step (device out). */
int dfmi_net_get(dfmi_net* net, int32_t which, float* out_dev);
/* u_dev (n, 3, X, Y, Z) = N(x_dev (n, 6, X, Y, Z)) on an X x Y x Z node grid. */
int dfmi_net_forward(dfmi_net* net, const float* x_dev, int32_t n, const int32_t dims[3], float* u_dev);
/* Gradients of a loss with dL/du = du_dev, for the last forward. */
int dfmi_net_backward(dfmi_net* net, const float* du_dev);
int dfmi_net_adam(dfmi_net* net, double lr, double beta1, double beta2, double eps);
/* One step on the normalised energy loss sum_n [0.5 delta_n.A_1 delta_n - delta_n.g_n] / (g_n.Dinv g_n)
 * with the level-1 Galerkin operator of a built slot; g_dev: n level-1 residuals [n][nodes][3] (as
 * harvested).  lr = 0 computes loss and gradients only.  loss may be NULL. */
int dfmi_net_train_step(dfmi_net* net, dfmi_mg* mg, int32_t slot, const float* g_dev, int32_t n, double lr,
                        double* loss);
/* Use the network inside the handle's fp32 V-cycle (NULL detaches). */
int dfmi_attach_net(dfmi_mg* mg, dfmi_net* net);
/* Record the level-1 residual g of every fine V-cycle into buffer_dev ([capacity][nodes_1][3], FP32),
 * zeros for inactive columns, until it is full; capacity 0 stops.  The count restarts at 0. */
int dfmi_harvest(dfmi_mg* mg, float* buffer_dev, int64_t capacity);
int dfmi_harvest_count(dfmi_mg* mg, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* DFMI_H */
