/*
 * mf3d.h -- C ABI of the MI355X-native Darcy head solver (libmf3d_hip.so).
 *
 * Plain C: POD descriptors, caller-owned device pointers with explicit element
 * counts, an explicit HIP stream (hipStream_t passed as void*, NULL = null
 * stream), and an int status (0 = MF3D_OK). No exception crosses this ABI;
 * mf3d_last_error() gives the detail text of the calling thread's last
 * failure. Every entry point enqueues its work on the given stream and
 * returns without synchronising unless its comment says otherwise.
 *
 * Each entry point names the MacroFlow3D function it replaces (paths relative
 * to the MacroFlow3D source root, file:line of the declaration/definition).
 */
#ifndef MF3D_H
#define MF3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF3D_ABI_VERSION 1

typedef void* mf3d_stream; /* hipStream_t */

/* Status codes. */
enum {
    MF3D_OK = 0,
    MF3D_ERR_INVALID_ARGUMENT = 1, /* bad pointer, BC pair, non-positive size ... */
    MF3D_ERR_SIZE_MISMATCH = 2,    /* buffer shorter than the grid, wrong 2:1 sizes */
    MF3D_ERR_BREAKDOWN = 3,        /* (p, Ap) zero or not finite */
    MF3D_ERR_HIP = 4,              /* a HIP runtime call failed */
    MF3D_ERR_ALLOC = 5,            /* device allocation failed */
    MF3D_ERR_SINGULAR = 6,         /* coarsest operator not definite (direct solve); a
                                      single solve reports it at its first coarsest solve */
    MF3D_ERR_UNSUPPORTED = 7       /* combination not implemented */
};

/* Grid3D (src/core/Grid3D.hpp:32-55). Cell (i,j,k) is at i + nx*(j + ny*k). */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
} mf3d_grid;

/* BCType (src/core/BCSpec.hpp:13) */
enum { MF3D_BC_DIRICHLET = 0, MF3D_BC_NEUMANN = 1, MF3D_BC_PERIODIC = 2 };

/* BCSpecDevice (src/core/BCSpecDevice.cuh:12-20): faces xmin, xmax, ymin, ymax, zmin, zmax. */
typedef struct {
    uint8_t type[6];
    double value[6];
} mf3d_bc;

/* HeadSolverType (src/physics/flow/solve_head.cuh:52-56) */
enum { MF3D_SOLVER_CG = 0, MF3D_SOLVER_MG = 1, MF3D_SOLVER_PCG_MG = 2 };
/* PinMode (src/numerics/pin_spec.hpp:56-60) */
enum { MF3D_PIN_AUTO = 0, MF3D_PIN_ON = 1, MF3D_PIN_OFF = 2 };

/* Extension enums (value 0 = reference behaviour). */
enum { MF3D_GUESS_REFERENCE = 0, MF3D_GUESS_ZERO = 1, MF3D_GUESS_PROVIDED = 2 };
enum { MF3D_COARSE_GSRB = 0, MF3D_COARSE_DIRECT = 1 };
enum { MF3D_STOP_INITIAL_RESIDUAL = 0, MF3D_STOP_RHS_NORM = 1 };
enum { MF3D_PRECOND_FP64 = 0, MF3D_PRECOND_FP32 = 1 };
enum { MF3D_COEF_FP32 = 0, MF3D_COEF_BF16 = 1, MF3D_COEF_FP16 = 2 };

/*
 * HeadSolveConfig (src/physics/flow/solve_head.cuh:64-131), field for field,
 * followed by extensions whose zero value keeps the reference behaviour.
 */
typedef struct {
    int solver_type;     /* MF3D_SOLVER_* (default PCG_MG) */
    int mg_levels;       /* 4 */
    int mg_pre_smooth;   /* 2 */
    int mg_post_smooth;  /* 2 */
    int mg_max_cycles;   /* 20 */
    int mg_coarse_iters; /* 50 */
    int cg_max_iter;     /* 1000 */
    int cg_check_every;  /* 10 */
    double cg_rtol;      /* 1e-8 */
    double rtol;         /* 1e-6 */
    int pin_mode;        /* MF3D_PIN_* */
    /* extensions */
    int init_guess;        /* MF3D_GUESS_* */
    int coarse_solver;     /* MF3D_COARSE_* */
    int stop_norm;         /* MF3D_STOP_* */
    int precond_precision; /* MF3D_PRECOND_* */
    int coef_format;       /* MF3D_COEF_* (FP32 preconditioner only) */
    int true_residual;     /* 1: report ||b - A h|| of the returned h in FP64 (the sums of a
                              final true-residual check when the solve ended on one) */
    int flexible;          /* 1: flexible-CG beta (r_new, z_new - z_old) / (r_old, z_old) */
    int reserved[4];
} mf3d_head_solve_config;

/* HeadSolveResult (src/physics/flow/solve_head.cuh:136-141) + extensions. */
typedef struct {
    int num_iterations;
    int converged;
    double initial_residual;
    double final_residual;
    double bnorm;         /* ||b|| of the head system */
    double true_residual; /* ||b - A h|| when requested, else -1 */
    double setup_ms;      /* device time of hierarchy/coefficient setup */
    double solve_ms;      /* device time of the iteration */
    int restarts;         /* true-residual restarts of the PCG (extension) */
    int reserved[3];
} mf3d_head_solve_result;

/* Opaque FlowWorkspace (src/physics/common/workspaces.cuh:111-167): MG
 * hierarchy, PCG vectors, reduction scratch. Device memory is allocated
 * lazily on first use for a grid and reused while the grid is unchanged
 * (solve_head.cu:89-94). Not re-entrant: one stream at a time per workspace. */
typedef struct mf3d_workspace mf3d_workspace;

/* ---- library --------------------------------------------------------- */
int mf3d_abi_version(void);
const char* mf3d_status_string(int status);
const char* mf3d_last_error(void);
/* Fills cfg with the reference defaults of HeadSolveConfig (solve_head.cuh:64-96). */
void mf3d_default_config(mf3d_head_solve_config* cfg);

int mf3d_workspace_create(mf3d_workspace** ws);
int mf3d_workspace_destroy(mf3d_workspace* ws);
/* Device bytes currently held by the workspace. */
size_t mf3d_workspace_bytes(const mf3d_workspace* ws);
/* FlowWorkspace::allocate(grid, mg_levels) (src/physics/common/workspaces.cuh:146-151)
 * with MGHierarchy(grid, num_levels) (src/multigrid/mg_types.hpp:75-98): the
 * hierarchy (halving rule) and every solve vector for this grid, allocated now
 * rather than by the first solve; a no-op when the workspace already fits. */
int mf3d_workspace_reserve(mf3d_workspace* ws, const mf3d_grid* grid, int mg_levels);

/* ---- physics --------------------------------------------------------- */
/* physics::solve_head (src/physics/flow/solve_head.cuh:158-160,
 * solve_head.cu:75-243). h: in/out (n >= grid cells); K: cell conductivity.
 * Synchronises the stream before returning (the result is host data). */
int mf3d_solve_head(mf3d_workspace* ws, mf3d_stream stream, const mf3d_grid* grid,
                    const mf3d_bc* bc, const mf3d_head_solve_config* cfg, const double* K,
                    double* h, size_t n, mf3d_head_solve_result* result);

/* Extension (no counterpart in MacroFlow3D; replaces the sequential solve
 * loop of EnsembleRunner.cu:259-311): `count` independent head solves, one
 * workspace per realisation (distinct, reused across calls), grids of any
 * sizes, shared BCs and configuration. Every kernel launch of the iteration
 * covers all realisations; one PCG iteration is captured once into a hipGraph
 * and replayed between host checks (flags & MF3D_BATCH_NO_GRAPH: plain
 * launches). Requires PCG_MG with the FP32 V-cycle, the direct coarsest solve,
 * pre/post smoothing >= 1, flexible = 0, nx % 8 == 0 on every
 * colour-swept level and 16-byte aligned heads (other cases: status
 * MF3D_ERR_UNSUPPORTED). Each realisation stops on its own tolerance; results[r]
 * holds its iterations and residuals (setup_ms/solve_ms: the whole batch).
 * Every coarsest operator is factored on the device; one that is not
 * definite makes the call return MF3D_ERR_SINGULAR after the solve, with the
 * realisation and pivot in mf3d_last_error. Synchronises. */
enum { MF3D_BATCH_NO_GRAPH = 1 };
int mf3d_solve_head_batched(int count, mf3d_workspace* const* ws, mf3d_stream stream,
                            const mf3d_grid* grids, const mf3d_bc* bc,
                            const mf3d_head_solve_config* cfg, const double* const* K,
                            double* const* h, const size_t* n, int flags,
                            mf3d_head_solve_result* results);

/* physics::init_head_guess (solve_head.cuh:173-174, solve_head.cu:56-73). */
int mf3d_init_head_guess(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc, double* h,
                         size_t n);

/* physics::build_rhs_head (src/physics/flow/rhs_head.cuh:50-51, rhs_head.cu:121-194). */
int mf3d_build_rhs_head(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc,
                        const double* K, double* rhs, size_t n);

/* physics::coarsen_K (src/physics/flow/coarsen_K.cuh:46-47, coarsen_K.cu:55-78). */
int mf3d_coarsen_K(mf3d_stream stream, const mf3d_grid* coarse, const mf3d_grid* fine,
                   const double* K_fine, size_t n_fine, double* K_coarse, size_t n_coarse);

/* physics::compute_velocity_from_head, compact MAC layout
 * (src/physics/flow/velocity_from_head.cuh:54-55, velocity_from_head.cu:601-691).
 * U: (nx+1)*ny*nz, V: nx*(ny+1)*nz, W: nx*ny*(nz+1). */
int mf3d_velocity_compact(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc,
                          const double* h, const double* K, size_t n, double* U, size_t nU,
                          double* V, size_t nV, double* W, size_t nW);

/* physics::compute_velocity_from_head, padded layout
 * (velocity_from_head.cuh:72-73, velocity_from_head.cu:697-777). Each of U, V,
 * W holds (nx+1)*(ny+1)*(nz+1) values indexed k*(ny+1)*(nx+1) + j*(nx+1) + i
 * (padded_idx, src/physics/flow/padded_layout.cuh:30-33). Only physical faces
 * are written; the caller zero-fills the padding once (fields.cuh:268-279). */
int mf3d_velocity_padded(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc,
                         const double* h, const double* K, size_t n, double* U, double* V,
                         double* W, size_t n_padded);

/* physics::compute_velocity_diagnostics (src/physics/flow/velocity_diagnostics.cuh:61-67,
 * velocity_diagnostics.cu:24-177) on a padded field (layout of mf3d_velocity_padded):
 * div, vort_mag and helicity (n >= nx*ny*nz cell values each). The divergence
 * covers every cell; |curl| and helicity are computed at interior cells
 * (1 <= i <= nx-2, same in j and k) and are 0 elsewhere, and everywhere when a
 * dimension is below 3 (the reference also prints a warning then; callers of
 * this entry point do that themselves). d_stats (nullable, device, 9 values):
 * min, max, mean of div, vort_mag and helicity, the numbers
 * print_velocity_diagnostics prints (velocity_diagnostics.cu:179-199); it
 * needs ws for the reduction scratch (ws may be null when d_stats is null). */
int mf3d_velocity_diagnostics(mf3d_workspace* ws, mf3d_stream stream, const mf3d_grid* grid,
                              const double* U, const double* V, const double* W,
                              size_t n_padded, double* div, double* vort_mag, double* helicity,
                              size_t n, double* d_stats);

/* compute_field_stats (src/physics/stochastic/stochastic.cu:369-414): d_out
 * (device, 3 values) = min, max, mean of x[0..n); zeros when n == 0. */
int mf3d_field_stats(mf3d_workspace* ws, mf3d_stream stream, const double* x, size_t n,
                     double* d_out);

/* ---- multigrid ------------------------------------------------------- */
/* multigrid::gsrb_smooth_3d (src/multigrid/smoothers/gsrb_3d.cuh:29-31, gsrb_3d.cu:368-576). */
int mf3d_gsrb_smooth_3d(mf3d_stream stream, const mf3d_grid* grid, double* x, const double* b,
                        const double* K, size_t n, int num_iters, const mf3d_bc* bc, int pin);

/* multigrid::compute_residual_3d (residual_3d.cuh:30-32, residual_3d.cu:378-503). */
int mf3d_compute_residual_3d(mf3d_stream stream, const mf3d_grid* grid, const double* x,
                             const double* b, const double* K, double* r, size_t n,
                             const mf3d_bc* bc, int pin);

/* multigrid::restrict_3d (src/multigrid/transfer/restrict_3d.cuh:14-15, restrict_3d.cu:55-81). */
int mf3d_restrict_3d(mf3d_stream stream, const mf3d_grid* fine, const mf3d_grid* coarse,
                     const double* r_fine, size_t n_fine, double* b_coarse, size_t n_coarse);

/* multigrid::prolong_3d_add (src/multigrid/transfer/prolong_3d.cuh:14-15, prolong_3d.cu:520-621). */
int mf3d_prolong_3d_add(mf3d_stream stream, const mf3d_grid* coarse, const mf3d_grid* fine,
                        const double* x_coarse, size_t n_coarse, double* x_fine, size_t n_fine);

/* MGHierarchy set-up (src/multigrid/mg_types.hpp:75-98) inside the workspace,
 * with K copied to level 0 and coarsened (solve_head.cu:89-114); builds the
 * dense coarsest inverse when cfg->coarse_solver is DIRECT. */
int mf3d_mg_setup(mf3d_workspace* ws, mf3d_stream stream, const mf3d_grid* grid,
                  const mf3d_bc* bc, const mf3d_head_solve_config* cfg, const double* K,
                  size_t n);
/* Tuning knobs of a workspace (extension; the autotuner sets them). Keys
 * 1-3 and 6-9 of earlier versions (one-pass march smoother, overlapped x
 * step) no longer exist and are rejected as unknown. */
enum {
    MF3D_TUNE_SPLIT = 4,                /* 1 (default): colour-split FP32 V-cycle vectors */
    MF3D_TUNE_SWEEP_BLOCK = 5,          /* threads per workgroup of the colour-split sweeps
                                           (64..512, multiple of 64; default 256) */
    MF3D_TUNE_FOLD_FIRST = 10,          /* 1 (default): a colour-split coarse level's first red
                                           half-sweep is computed by the restriction kernel */
    MF3D_TUNE_FOLD_R = 11,              /* 1 (default): the FP32-preconditioned PCG's residual
                                           update runs in the first finest-level half-sweep */
    MF3D_TUNE_FOLD_RZ = 12,             /* 1 (default): ... and its (r, z) in the last one */
    MF3D_TUNE_OP_TILE = 13,             /* 1 (default): the FP64 operator stages plane tiles in
                                           LDS when nx <= 512 (even, 16-byte aligned vectors) */
    MF3D_TUNE_GRAPH = 14,               /* 1: a mixed-precision single solve captures one PCG
                                           iteration in a hipGraph and replays it; 0: it
                                           launches the kernels; -1 (default): replay when the
                                           grid has at most 2^21 cells */
    MF3D_TUNE_OP_CHUNK = 15,            /* 16 (default): the FP64 operator's z-planes per
                                           workgroup, at most (1..64; fewer where the launch
                                           would have too few workgroups) */
    MF3D_TUNE_MARCH_BLOCKS = 16         /* 2048 (default): workgroups an FP64 colour-split
                                           sweep aims for (256..8192) by shortening its chunks
                                           of planes */
};
int mf3d_workspace_set_tuning(mf3d_workspace* ws, int key, long value);

/* Extension, no counterpart in MacroFlow3D: the finest-level FP64 operator
 * applications (A p and residuals) and the preconditioner applications
 * enqueued on ws since the last mf3d_solve_head began (it resets both; later
 * mf3d_mg_precondition calls on ws add to them), replays of a captured PCG
 * iteration included, iterations enqueued past convergence too. A check on
 * the kernel sequence for tests and tuning: a PCG solve applies the operator
 * once per preconditioner application, plus the initial and true residuals. */
int mf3d_workspace_solve_counts(const mf3d_workspace* ws, long* op_applications,
                                long* precond_applications);

/* Extension: num_iters GSRB iterations on level `level` of the current
 * hierarchy, in place on the workspace's own correction and right-hand side
 * of that level (level 0: the preconditioner's input/output buffers), in the
 * set-up's precision. For timing and testing the smoother in situ. */
int mf3d_mg_smooth(mf3d_workspace* ws, mf3d_stream stream, int level, int num_iters);

/* Number of levels of the current hierarchy (0 before set-up). */
int mf3d_mg_num_levels(const mf3d_workspace* ws);

/* solvers::MultigridPreconditioner::apply (src/numerics/solvers/mg_preconditioner.cuh:44-71,
 * mg_preconditioner.cu:16-45): z = one V-cycle (v_cycle.cu:36-74) from x = 0 with b = r.
 * r and z must be 16-byte aligned (MF3D_ERR_INVALID_ARGUMENT otherwise). */
int mf3d_mg_precondition(mf3d_workspace* ws, mf3d_stream stream, const double* r, double* z,
                         size_t n);

/* ---- numerics -------------------------------------------------------- */
/* operators::VarCoeffLaplacian::apply (src/numerics/operators/varcoeff_laplacian.cuh:47-55,
 * varcoeff_laplacian.cu:241-277). */
int mf3d_varcoeff_apply(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc, int pin,
                        const double* K, const double* x, double* y, size_t n);

/* blas::dot_device (src/numerics/blas/dot.cuh, dot.cu:9-36): *d_result = x.y
 * (device pointer). Deterministic two-stage reduction; scratch from ws. */
int mf3d_dot(mf3d_workspace* ws, mf3d_stream stream, const double* x, const double* y, size_t n,
             double* d_result);
/* No counterpart in the reference: mf3d_dot without a workspace, for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). Same kernels and summation order as
 * mf3d_dot, so the result is bitwise the same. Its partial sums live in a
 * fixed 512 KiB device array inside the library, one per GPU, allocated with
 * the library's code and held until the process ends: n is limited to
 * 2^27 (= 512^3) elements, larger n returns MF3D_ERR_INVALID_ARGUMENT. Calls
 * that overlap on one GPU (two streams, two threads) share that array and
 * race; serialise them. Allocates nothing, ever. */
int mf3d_dot_noalloc(mf3d_stream stream, const double* x, const double* y, size_t n,
                     double* d_result);
/* blas::nrm2_device (nrm2.cu:9-36) */
int mf3d_nrm2(mf3d_workspace* ws, mf3d_stream stream, const double* x, size_t n,
              double* d_result);
/* blas::axpy (axpy.cu:7-29): y = a x + y */
int mf3d_axpy(mf3d_stream stream, double a, const double* x, double* y, size_t n);
/* blas::axpby (axpby.cu:7-29): y = a x + b y */
int mf3d_axpby(mf3d_stream stream, double a, const double* x, double b, double* y, size_t n);
/* blas::copy (copy.cu:7-29) */
int mf3d_copy(mf3d_stream stream, const double* x, double* y, size_t n);
/* blas::fill (fill.cu:7-27) */
int mf3d_fill(mf3d_stream stream, double* x, size_t n, double value);
/* blas::scal (scal.cu:7-27): x = a x */
int mf3d_scal(mf3d_stream stream, double* x, size_t n, double a);

/* ---- learned warm start (extension, no counterpart in MacroFlow3D) ----
 * A 3-D convolutional encoder-decoder (3 levels, 16/32/64 channels, 3^3
 * kernels, BF16 weights/activations, FP32 accumulation on MFMA) mapping ln K
 * to an initial head h0 = h_lin + J trilinear(y) for the Darcy set-up with
 * Dirichlet x faces (h_lin linear between them, J = (h_W - h_E) / nx). It
 * runs on the grid coarsened 4x per axis (inputs: block means of ln K, K and
 * 1/K); its levels are 4, 8 and 16 times coarser than the grid, so sizes must
 * be divisible by 16. Parameters are FP32 masters: per layer (conv_a..conv_g)
 * the weights [cout][27][cin] (tap = dx + 3 dy + 9 dz over offsets -1..1,
 * +1) followed by the bias [cout]. */
typedef struct mf3d_wsnet mf3d_wsnet;
long mf3d_wsnet_param_count(void);
int mf3d_wsnet_create(mf3d_wsnet** net);  /* parameters zero: set them before use */
int mf3d_wsnet_destroy(mf3d_wsnet* net);
int mf3d_wsnet_set_params(mf3d_wsnet* net, const float* host_params, long n); /* resets Adam */
int mf3d_wsnet_get_params(const mf3d_wsnet* net, float* host_params, long n);
int mf3d_wsnet_get_grads(const mf3d_wsnet* net, float* host_grads, long n);
/* h0 (n >= cells) from K (k_is_log = 0) or ln K (k_is_log = 1). */
int mf3d_wsnet_predict(mf3d_wsnet* net, mf3d_stream stream, const mf3d_grid* grid,
                       const mf3d_bc* bc, const double* K, int k_is_log, double* h0, size_t n);
/* Loss mean((y - (h - h_lin)/J)^2) for the solved head h and the gradient of
 * every parameter (read with mf3d_wsnet_get_grads). Synchronises if loss != NULL. */
int mf3d_wsnet_gradients(mf3d_wsnet* net, mf3d_stream stream, const mf3d_grid* grid,
                         const mf3d_bc* bc, const double* K, int k_is_log, const double* h,
                         size_t n, double* loss);
/* mf3d_wsnet_gradients followed by one Adam step (beta 0.9/0.999, eps 1e-8). */
int mf3d_wsnet_train_step(mf3d_wsnet* net, mf3d_stream stream, const mf3d_grid* grid,
                          const mf3d_bc* bc, const double* K, int k_is_log, const double* h,
                          size_t n, double lr, double* loss);
/* Single-layer kernels (BF16 channels-last tensors [cells][channels]; zero
 * padding; synchronous). Forward: Y = relu?(conv(X, w) + bias + R), cin a
 * multiple of 8, cout 16/32/64. Data gradient: dX = conv^T(dY, w) * (M > 0).
 * Weight gradient: gw [cout][27][cin], gb [cout] (FP32). */
int mf3d_conv3d_forward(mf3d_stream stream, const mf3d_grid* grid, int cin, int cout,
                        const void* X, const float* w, const float* bias, const void* R,
                        int relu, void* Y);
int mf3d_conv3d_backward_data(mf3d_stream stream, const mf3d_grid* grid, int cin, int cout,
                              const void* dY, const float* w, const void* M, void* dX);
int mf3d_conv3d_backward_weights(mf3d_stream stream, const mf3d_grid* grid, int cin, int cout,
                                 const void* X, const void* dY, float* gw, float* gb);

#ifdef __cplusplus
}
#endif

#endif /* MF3D_H */
