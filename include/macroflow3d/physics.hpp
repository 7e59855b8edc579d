// C++ host surface of libmf3d_hip: the flow physics.
//
//   ScalarField / KField / HeadField        src/physics/common/fields.cuh:37-92
//   VelocityField                            src/physics/common/fields.cuh:108-212
//   PaddedVelocityField                      src/physics/common/fields.cuh:248-319
//   FlowWorkspace                            src/physics/common/workspaces.cuh:111-167
//   HeadSolverType/HeadSolveConfig/Result    src/physics/flow/solve_head.cuh:52-141
//   solve_head, init_head_guess              src/physics/flow/solve_head.cuh:158-174
//   build_rhs_head                           src/physics/flow/rhs_head.cuh:50-51
//   coarsen_K                                src/physics/flow/coarsen_K.cuh:46-47
//   compute_velocity_from_head (x2), compute_norm2/sum/mean, check_no_nans
//                                            src/physics/flow/velocity_from_head.cuh:54-107
//   VelocityDiagnostics, compute_velocity_diagnostics, print_velocity_diagnostics
//                                            src/physics/flow/velocity_diagnostics.cuh:31-92
//   compute_field_stats                      src/physics/stochastic/stochastic.cu:369-414
//   solve_head_batched, WarmStartNet         extensions without a counterpart
//
// HeadSolveConfig keeps the reference fields and defaults; the extension
// fields after them default to the reference behaviour.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "numerics.hpp"

namespace macroflow3d {

// ---- fields -------------------------------------------------------------------
struct ScalarField {
    int nx = 0, ny = 0, nz = 0;
    real dx = 1.0, dy = 1.0, dz = 1.0;
    DeviceBuffer<real> data;

    ScalarField() = default;
    explicit ScalarField(const Grid3D& g)
        : nx(g.nx), ny(g.ny), nz(g.nz), dx(g.dx), dy(g.dy), dz(g.dz), data(g.num_cells()) {}
    ScalarField(int nx_, int ny_, int nz_, real dx_ = 1.0, real dy_ = 1.0, real dz_ = 1.0)
        : nx(nx_), ny(ny_), nz(nz_), dx(dx_), dy(dy_), dz(dz_), data((size_t)nx_ * ny_ * nz_) {}

    void resize(const Grid3D& g) {
        nx = g.nx, ny = g.ny, nz = g.nz, dx = g.dx, dy = g.dy, dz = g.dz;
        data.resize(g.num_cells());
    }
    void resize(int nx_, int ny_, int nz_) {
        nx = nx_, ny = ny_, nz = nz_;
        data.resize((size_t)nx_ * ny_ * nz_);
    }
    size_t size() const { return (size_t)nx * ny * nz; }
    bool empty() const { return data.size() == 0; }
    DeviceSpan<real> span() { return data.span(); }
    DeviceSpan<const real> span() const { return data.span(); }
    real* device_ptr() { return data.data(); }
    const real* device_ptr() const { return data.data(); }
    Grid3D grid() const { return Grid3D(nx, ny, nz, dx, dy, dz); }
    size_t idx(int i, int j, int k) const { return (size_t)i + (size_t)nx * (j + (size_t)ny * k); }
};
using KField = ScalarField;
using HeadField = ScalarField;

// Compact MAC faces: U (nx+1)*ny*nz, V nx*(ny+1)*nz, W nx*ny*(nz+1).
struct VelocityField {
    int nx = 0, ny = 0, nz = 0;
    real dx = 1.0, dy = 1.0, dz = 1.0;
    DeviceBuffer<real> U, V, W;

    VelocityField() = default;
    explicit VelocityField(const Grid3D& g) { resize(g); }
    VelocityField(int nx_, int ny_, int nz_, real dx_ = 1.0, real dy_ = 1.0, real dz_ = 1.0)
        : nx(nx_), ny(ny_), nz(nz_), dx(dx_), dy(dy_), dz(dz_) {
        allocate();
    }
    void allocate() {
        if (nx > 0 && ny > 0 && nz > 0) {
            U.resize(size_U());
            V.resize(size_V());
            W.resize(size_W());
        }
    }
    void resize(const Grid3D& g) {
        nx = g.nx, ny = g.ny, nz = g.nz, dx = g.dx, dy = g.dy, dz = g.dz;
        allocate();
    }
    void resize(int nx_, int ny_, int nz_) {
        nx = nx_, ny = ny_, nz = nz_;
        allocate();
    }
    size_t size_U() const { return (size_t)(nx + 1) * ny * nz; }
    size_t size_V() const { return (size_t)nx * (ny + 1) * nz; }
    size_t size_W() const { return (size_t)nx * ny * (nz + 1); }
    size_t total_size() const { return size_U() + size_V() + size_W(); }
    bool empty() const { return U.size() == 0; }
    DeviceSpan<real> U_span() { return U.span(); }
    DeviceSpan<real> V_span() { return V.span(); }
    DeviceSpan<real> W_span() { return W.span(); }
    DeviceSpan<const real> U_span() const { return U.span(); }
    DeviceSpan<const real> V_span() const { return V.span(); }
    DeviceSpan<const real> W_span() const { return W.span(); }
    int stride_U_j() const { return nx + 1; }
    int stride_U_k() const { return (nx + 1) * ny; }
    int stride_V_j() const { return nx; }
    int stride_V_k() const { return nx * (ny + 1); }
    int stride_W_j() const { return nx; }
    int stride_W_k() const { return nx * ny; }
    size_t idx_U(int i, int j, int k) const { return i + (size_t)j * stride_U_j() + (size_t)k * stride_U_k(); }
    size_t idx_V(int i, int j, int k) const { return i + (size_t)j * stride_V_j() + (size_t)k * stride_V_k(); }
    size_t idx_W(int i, int j, int k) const { return i + (size_t)j * stride_W_j() + (size_t)k * stride_W_k(); }
};

// Padded faces: each component (nx+1)(ny+1)(nz+1), index i + (nx+1)(j + (ny+1) k),
// zero-filled once at allocation (padding entries stay zero).
struct PaddedVelocityField {
    int nx = 0, ny = 0, nz = 0;
    real dx = 1.0, dy = 1.0, dz = 1.0;
    DeviceBuffer<real> U, V, W;

    PaddedVelocityField() = default;
    explicit PaddedVelocityField(const Grid3D& g) { resize(g); }
    void allocate() {
        if (nx > 0 && ny > 0 && nz > 0) {
            const size_t s = field_size();
            for (DeviceBuffer<real>* b : {&U, &V, &W}) {
                b->resize(s);
                hip_check(hipMemset(b->data(), 0, s * sizeof(real)), "PaddedVelocityField");
            }
            // hipMemset runs on the null stream, which does not order against
            // the contexts' non-blocking streams: finish it before any kernel
            // writes the faces
            hip_check(hipDeviceSynchronize(), "PaddedVelocityField");
        }
    }
    void resize(const Grid3D& g) {
        nx = g.nx, ny = g.ny, nz = g.nz, dx = g.dx, dy = g.dy, dz = g.dz;
        allocate();
    }
    size_t field_size() const { return (size_t)(nx + 1) * (ny + 1) * (nz + 1); }
    size_t total_size() const { return 3 * field_size(); }
    bool empty() const { return U.size() == 0; }
    size_t idx(int i, int j, int k) const {
        return (size_t)i + (size_t)(nx + 1) * ((size_t)j + (size_t)(ny + 1) * k);
    }
    DeviceSpan<real> U_span() { return U.span(); }
    DeviceSpan<real> V_span() { return V.span(); }
    DeviceSpan<real> W_span() { return W.span(); }
    DeviceSpan<const real> U_span() const { return U.span(); }
    DeviceSpan<const real> V_span() const { return V.span(); }
    DeviceSpan<const real> W_span() const { return W.span(); }
};

// ---- workspace ------------------------------------------------------------------
// FlowWorkspace (src/physics/common/workspaces.cuh:111-167). The MG hierarchy,
// the Krylov vectors and the right-hand side of solve_head all live in
// mg_hierarchy's device workspace, sized for the grid it last saw: solve_head
// sizes it on first use (solve_head.cu:89-94), allocate() does so up front.
struct FlowWorkspace {
    multigrid::MGHierarchy mg_hierarchy;
    size_t n_cells = 0;

    FlowWorkspace() = default;
    // records the grid; the device memory is taken by the first solve
    void allocate(const Grid3D& grid) { n_cells = grid.num_cells(); }
    // the hierarchy of mg_levels levels and every solve vector, now
    void allocate(const Grid3D& grid, int mg_levels) {
        allocate(grid);
        if (mg_levels > 0) mg_hierarchy.reserve(grid, mg_levels);
    }
    bool is_allocated() const { return n_cells > 0; }
    void clear() {
        mg_hierarchy = multigrid::MGHierarchy();
        n_cells = 0;
    }
    size_t device_bytes() const { return mg_hierarchy.device_bytes(); }
    void set_tuning(int key, long value) { mg_hierarchy.set_tuning(key, value); }
};

namespace physics {

enum class HeadSolverType { CG, MG, PCG_MG };

// extensions (value 0 = reference behaviour)
enum class InitGuess { Reference, Zero, Provided };
enum class CoarseSolver { GSRB, Direct };
enum class StopNorm { InitialResidual, RhsNorm };
enum class PrecondPrecision { FP64, FP32 };
enum class CoefFormat { FP32, BF16, FP16 };

struct HeadSolveConfig {
    HeadSolverType solver_type = HeadSolverType::PCG_MG;
    int mg_levels = 4;
    int mg_pre_smooth = 2;
    int mg_post_smooth = 2;
    int mg_max_cycles = 20;
    int mg_coarse_iters = 50;
    int cg_max_iter = 1000;
    int cg_check_every = 10;
    real cg_rtol = 1e-8;
    real rtol = 1e-6;
    io::PinConfig pin;
    // extensions
    InitGuess init_guess = InitGuess::Reference;
    CoarseSolver coarse_solver = CoarseSolver::GSRB;
    StopNorm stop_norm = StopNorm::InitialResidual;
    PrecondPrecision precond_precision = PrecondPrecision::FP64;
    CoefFormat coef_format = CoefFormat::BF16;
    bool true_residual = true;
    bool flexible = false;

    HeadSolveConfig() = default;

    // HeadSolveConfig::from_yaml (solve_head.cuh:98-130): "cg" -> CG, "mg" ->
    // MG, "pcg_mg" / "mg_cg" and any other string -> PCG_MG; the numbers and
    // the pin mode are copied; the extensions keep their defaults.
    static HeadSolveConfig from_yaml(const io::FlowYamlConfig& flow_cfg) {
        HeadSolveConfig cfg;
        if (flow_cfg.solver == "cg") cfg.solver_type = HeadSolverType::CG;
        else if (flow_cfg.solver == "mg") cfg.solver_type = HeadSolverType::MG;
        else cfg.solver_type = HeadSolverType::PCG_MG;
        cfg.mg_levels = flow_cfg.mg_levels;
        cfg.mg_pre_smooth = flow_cfg.mg_pre_smooth;
        cfg.mg_post_smooth = flow_cfg.mg_post_smooth;
        cfg.mg_max_cycles = flow_cfg.mg_max_cycles;
        cfg.mg_coarse_iters = flow_cfg.mg_coarse_iters;
        cfg.cg_max_iter = flow_cfg.cg_max_iter;
        cfg.cg_rtol = flow_cfg.cg_rtol;
        cfg.cg_check_every = flow_cfg.cg_check_every;
        cfg.rtol = flow_cfg.rtol;
        cfg.pin = flow_cfg.pin;
        return cfg;
    }

    // direct coarsest solve, ||b - A h|| <= rtol ||b|| tested every iteration, zero guess
    static HeadSolveConfig fp64_direct(real rtol = 1e-10, int levels = 6) {
        HeadSolveConfig c;
        c.mg_levels = levels;
        c.cg_rtol = rtol;
        c.cg_check_every = 1;
        c.init_guess = InitGuess::Zero;
        c.coarse_solver = CoarseSolver::Direct;
        c.stop_norm = StopNorm::RhsNorm;
        return c;
    }
    // fp64_direct with the FP32 V-cycle and 16-bit face coefficients
    static HeadSolveConfig mixed(real rtol = 1e-10, int levels = 6,
                                 CoefFormat coef = CoefFormat::BF16) {
        HeadSolveConfig c = fp64_direct(rtol, levels);
        c.precond_precision = PrecondPrecision::FP32;
        c.coef_format = coef;
        return c;
    }

    mf3d_head_solve_config c() const {
        mf3d_head_solve_config o;
        mf3d_default_config(&o);
        o.solver_type = (int)solver_type;
        o.mg_levels = mg_levels;
        o.mg_pre_smooth = mg_pre_smooth;
        o.mg_post_smooth = mg_post_smooth;
        o.mg_max_cycles = mg_max_cycles;
        o.mg_coarse_iters = mg_coarse_iters;
        o.cg_max_iter = cg_max_iter;
        o.cg_check_every = cg_check_every;
        o.cg_rtol = cg_rtol;
        o.rtol = rtol;
        o.pin_mode = (int)pin.mode;
        o.init_guess = (int)init_guess;
        o.coarse_solver = (int)coarse_solver;
        o.stop_norm = (int)stop_norm;
        o.precond_precision = (int)precond_precision;
        o.coef_format = (int)coef_format;
        o.true_residual = true_residual ? 1 : 0;
        o.flexible = flexible ? 1 : 0;
        return o;
    }
};

struct HeadSolveResult {
    int num_iterations = 0;
    real initial_residual = 0.0;
    real final_residual = 0.0;
    bool converged = false;
    // extensions
    real bnorm = 0.0;
    real true_residual = -1.0;
    real setup_ms = 0.0;
    real solve_ms = 0.0;
    int restarts = 0;
};

inline HeadSolveResult from_c(const mf3d_head_solve_result& r) {
    HeadSolveResult o;
    o.num_iterations = r.num_iterations;
    o.initial_residual = r.initial_residual;
    o.final_residual = r.final_residual;
    o.converged = r.converged != 0;
    o.bnorm = r.bnorm;
    o.true_residual = r.true_residual;
    o.setup_ms = r.setup_ms;
    o.solve_ms = r.solve_ms;
    o.restarts = r.restarts;
    return o;
}

// Solves -div(K grad h) = 0 with the boundary conditions; h is overwritten
// by the initial guess unless cfg.init_guess is Provided. Non-convergence is
// reported in the result, not thrown.
inline HeadSolveResult solve_head(DeviceSpan<real> h, DeviceSpan<const real> K,
                                  const Grid3D& grid, const BCSpec& bc,
                                  const HeadSolveConfig& cfg, CudaContext& ctx,
                                  FlowWorkspace& workspace) {
    bc.validate();
    if (K.size() < h.size()) throw std::runtime_error("solve_head: K shorter than h");
    const mf3d_grid g = to_c(grid);
    const mf3d_bc b = to_c(bc);
    const mf3d_head_solve_config c = cfg.c();
    mf3d_head_solve_result r{};
    mf3d_check(mf3d_solve_head(workspace.mg_hierarchy.handle(), ctx.c_stream(), &g, &b, &c,
                               K.data(), h.data(), h.size(), &r),
               "solve_head");
    return from_c(r);
}

inline void init_head_guess(DeviceSpan<real> h, const Grid3D& grid, const BCSpec& bc,
                            const CudaContext& ctx) {
    const mf3d_grid g = to_c(grid);
    const mf3d_bc b = to_c(bc);
    mf3d_check(mf3d_init_head_guess(ctx.c_stream(), &g, &b, h.data(), h.size()),
               "init_head_guess");
}

inline void build_rhs_head(DeviceSpan<real> rhs, DeviceSpan<const real> K, const Grid3D& grid,
                           const BCSpec& bc, const CudaContext& ctx) {
    const mf3d_grid g = to_c(grid);
    const mf3d_bc b = to_c(bc);
    if (K.size() < rhs.size()) throw std::runtime_error("build_rhs_head: K shorter than rhs");
    mf3d_check(mf3d_build_rhs_head(ctx.c_stream(), &g, &b, K.data(), rhs.data(), rhs.size()),
               "build_rhs_head");
}

inline void coarsen_K(CudaContext& ctx, const Grid3D& coarse_grid, const Grid3D& fine_grid,
                      DeviceSpan<const real> K_fine, DeviceSpan<real> K_coarse) {
    const mf3d_grid c = to_c(coarse_grid), f = to_c(fine_grid);
    mf3d_check(mf3d_coarsen_K(ctx.c_stream(), &c, &f, K_fine.data(), K_fine.size(),
                              K_coarse.data(), K_coarse.size()),
               "coarsen_K");
}

inline void compute_velocity_from_head(VelocityField& vel, const HeadField& head, const KField& K,
                                       const Grid3D& grid, const BCSpec& bc, CudaContext& ctx) {
    const mf3d_grid g = to_c(grid);
    const mf3d_bc b = to_c(bc);
    if (vel.nx != grid.nx || vel.ny != grid.ny || vel.nz != grid.nz) vel.resize(grid);
    mf3d_check(mf3d_velocity_compact(ctx.c_stream(), &g, &b, head.device_ptr(), K.device_ptr(),
                                     head.data.size(), vel.U.data(), vel.U.size(), vel.V.data(),
                                     vel.V.size(), vel.W.data(), vel.W.size()),
               "compute_velocity_from_head");
}

inline void compute_velocity_from_head(PaddedVelocityField& vel, const HeadField& head,
                                       const KField& K, const Grid3D& grid, const BCSpec& bc,
                                       CudaContext& ctx) {
    const mf3d_grid g = to_c(grid);
    const mf3d_bc b = to_c(bc);
    if (vel.nx != grid.nx || vel.ny != grid.ny || vel.nz != grid.nz) vel.resize(grid);
    mf3d_check(mf3d_velocity_padded(ctx.c_stream(), &g, &b, head.device_ptr(), K.device_ptr(),
                                    head.data.size(), vel.U.data(), vel.V.data(), vel.W.data(),
                                    vel.field_size()),
               "compute_velocity_from_head");
}

// ---- checksum utilities (synchronous) ------------------------------------------
inline real compute_norm2(DeviceSpan<const real> data, CudaContext& ctx) {
    if (data.empty()) return 0.0;
    blas::ReductionWorkspace red;
    return blas::nrm2_host(ctx, data, red);
}

inline real compute_sum(DeviceSpan<const real> data, CudaContext& ctx) {
    if (data.empty()) return 0.0;
    DeviceBuffer<real> ones(data.size());
    blas::fill(ctx, ones.span(), 1.0);
    blas::ReductionWorkspace red;
    return blas::dot_host(ctx, data, ones.span(), red);
}

inline real compute_mean(DeviceSpan<const real> data, CudaContext& ctx) {
    return data.empty() ? 0.0 : compute_sum(data, ctx) / (real)data.size();
}

inline bool check_no_nans(DeviceSpan<const real> data, CudaContext& ctx) {
    return !std::isnan(compute_norm2(data, ctx));
}

inline void print_velocity_checksums(const VelocityField& vel, CudaContext& ctx) {
    const char* names[3] = {"U", "V", "W"};
    DeviceSpan<const real> spans[3] = {vel.U_span(), vel.V_span(), vel.W_span()};
    for (int c = 0; c < 3; ++c) {
        const real nrm = compute_norm2(spans[c], ctx);
        std::printf("[velocity] %s: sum=%.17g norm2=%.17g nan=%s\n", names[c],
                    compute_sum(spans[c], ctx), nrm, std::isnan(nrm) ? "yes" : "no");
    }
}

// ---- velocity diagnostics ------------------------------------------------------
// min, max and mean of a device field (synchronous); zeros when empty
inline void compute_field_stats(DeviceSpan<const real> data, real& min_val, real& max_val,
                                real& mean_val, const CudaContext& ctx) {
    if (data.empty()) {
        min_val = max_val = mean_val = 0.0;
        return;
    }
    blas::ReductionWorkspace red;
    DeviceBuffer<real> out(3);
    mf3d_check(mf3d_field_stats(red.handle(), ctx.c_stream(), data.data(), data.size(),
                                out.data()),
               "compute_field_stats");
    real h[3];
    hip_check(hipMemcpyAsync(h, out.data(), sizeof(h), hipMemcpyDeviceToHost, ctx.stream()),
              "compute_field_stats");
    ctx.synchronize();
    min_val = h[0], max_val = h[1], mean_val = h[2];
}

// Cell-centred divergence, |curl| and helicity (nx*ny*nz values each).
struct VelocityDiagnostics {
    ScalarField divergence;
    ScalarField vorticity_mag;
    ScalarField helicity;

    VelocityDiagnostics() = default;
    explicit VelocityDiagnostics(const Grid3D& grid)
        : divergence(grid), vorticity_mag(grid), helicity(grid) {}
    void resize(const Grid3D& grid) {
        divergence.resize(grid);
        vorticity_mag.resize(grid);
        helicity.resize(grid);
    }
    bool empty() const { return divergence.empty(); }
};

// One launch over all cells: divergence everywhere, |curl| and helicity at
// interior cells (0 at the border, and everywhere when a dimension is below
// 3, with a warning). diag must be allocated to the grid size.
inline void compute_velocity_diagnostics(const PaddedVelocityField& vel, VelocityDiagnostics& diag,
                                         const Grid3D& grid, const CudaContext& ctx) {
    const size_t total = grid.num_cells();
    if (total == 0) return;
    if (grid.nx < 3 || grid.ny < 3 || grid.nz < 3)
        std::printf("       [diag] WARNING: grid too small for curl/helicity "
                    "(%dx%dx%d, need >=3 in each dim). Only divergence computed.\n",
                    grid.nx, grid.ny, grid.nz);
    const mf3d_grid g = to_c(grid);
    mf3d_check(mf3d_velocity_diagnostics(nullptr, ctx.c_stream(), &g, vel.U.data(), vel.V.data(),
                                         vel.W.data(), vel.U.size(),
                                         diag.divergence.device_ptr(),
                                         diag.vorticity_mag.device_ptr(),
                                         diag.helicity.device_ptr(), diag.divergence.data.size(),
                                         nullptr),
               "compute_velocity_diagnostics");
}

inline void print_velocity_diagnostics(const VelocityDiagnostics& diag, int realization_id,
                                       const CudaContext& ctx) {
    real lo, hi, mean;
    compute_field_stats(diag.divergence.span(), lo, hi, mean, ctx);
    std::printf("       [diag] r=%d  div:  min=%.4e  max=%.4e  mean=%.4e\n", realization_id, lo,
                hi, mean);
    compute_field_stats(diag.vorticity_mag.span(), lo, hi, mean, ctx);
    std::printf("       [diag] r=%d  |w|:  min=%.4e  max=%.4e  mean=%.4e  (incl. border=0)\n",
                realization_id, lo, hi, mean);
    compute_field_stats(diag.helicity.span(), lo, hi, mean, ctx);
    std::printf("       [diag] r=%d  h:    min=%.4e  max=%.4e  mean=%.4e  (incl. border=0)\n",
                realization_id, lo, hi, mean);
}

// ---- extensions without a counterpart in MacroFlow3D ---------------------------

// Independent head solves of many realisations at once (the realisation loop
// of EnsembleRunner.cu:259-311 solved together): one workspace per
// realisation, grids of any sizes, shared BCs and configuration. Each
// realisation gets the head and iteration count of its own solve_head call.
// Needs PCG_MG with the FP32 V-cycle (HeadSolveConfig::mixed) and nx % 8 == 0
// on every colour-swept level; other configurations throw (status
// MF3D_ERR_UNSUPPORTED) and the caller loops over solve_head instead.
inline std::vector<HeadSolveResult> solve_head_batched(
    const std::vector<DeviceSpan<real>>& h, const std::vector<DeviceSpan<const real>>& K,
    const std::vector<Grid3D>& grids, const BCSpec& bc, const HeadSolveConfig& cfg,
    CudaContext& ctx, const std::vector<FlowWorkspace*>& workspaces, bool use_graph = true) {
    bc.validate();
    const size_t R = h.size();
    if (K.size() != R || grids.size() != R || workspaces.size() != R)
        throw std::runtime_error("solve_head_batched: h, K, grids and workspaces differ in count");
    std::vector<mf3d_workspace*> ws(R);
    std::vector<mf3d_grid> g(R);
    std::vector<const double*> kp(R);
    std::vector<double*> hp(R);
    std::vector<size_t> n(R);
    for (size_t r = 0; r < R; ++r) {
        if (!workspaces[r]) throw std::runtime_error("solve_head_batched: null workspace");
        if (K[r].size() < h[r].size())
            throw std::runtime_error("solve_head_batched: K shorter than h");
        ws[r] = workspaces[r]->mg_hierarchy.handle();
        g[r] = to_c(grids[r]);
        kp[r] = K[r].data();
        hp[r] = h[r].data();
        n[r] = h[r].size();
    }
    const mf3d_bc b = to_c(bc);
    const mf3d_head_solve_config c = cfg.c();
    std::vector<mf3d_head_solve_result> res(R);
    mf3d_check(mf3d_solve_head_batched((int)R, ws.data(), ctx.c_stream(), g.data(), &b, &c,
                                       kp.data(), hp.data(), n.data(),
                                       use_graph ? 0 : MF3D_BATCH_NO_GRAPH, res.data()),
               "solve_head_batched");
    std::vector<HeadSolveResult> out(R);
    for (size_t r = 0; r < R; ++r) out[r] = from_c(res[r]);
    return out;
}

// Learned initial head for the Darcy set-up (Dirichlet x faces): a 3-D
// convolutional encoder-decoder on the grid coarsened 4x per axis predicts
// h0 = h_lin + J trilinear(y); solve from it with InitGuess::Provided. Grid
// sizes must be divisible by 16. Parameters: the FP32 vector of
// macroflow3d_amd/data/warmstart_128.npy (load_npy) or any vector of
// param_count() values.
class WarmStartNet {
public:
    WarmStartNet() { mf3d_check(mf3d_wsnet_create(&net_), "mf3d_wsnet_create"); }
    explicit WarmStartNet(const std::string& npy_path) : WarmStartNet() { load_npy(npy_path); }
    ~WarmStartNet() {
        if (net_) mf3d_wsnet_destroy(net_);
    }
    WarmStartNet(const WarmStartNet&) = delete;
    WarmStartNet& operator=(const WarmStartNet&) = delete;

    static long param_count() { return mf3d_wsnet_param_count(); }

    void set_params(const std::vector<float>& p) {
        mf3d_check(mf3d_wsnet_set_params(net_, p.data(), (long)p.size()), "wsnet_set_params");
    }
    std::vector<float> params() const {
        std::vector<float> p((size_t)param_count());
        mf3d_check(mf3d_wsnet_get_params(net_, p.data(), (long)p.size()), "wsnet_get_params");
        return p;
    }

    // Reads a 1-D little-endian float32 .npy file (format versions 1-3).
    void load_npy(const std::string& path) {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("WarmStartNet: cannot open " + path);
        std::vector<float> p;
        std::string err;
        unsigned char pre[10];
        if (std::fread(pre, 1, 10, f) != 10 || std::string((const char*)pre + 1, 5) != "NUMPY" ||
            pre[0] != 0x93) {
            err = "not a .npy file";
        } else {
            size_t hlen = (size_t)pre[8] | ((size_t)pre[9] << 8);
            if (pre[6] >= 2) {  // 4-byte header length
                unsigned char ext[2];
                if (std::fread(ext, 1, 2, f) != 2) err = "truncated header";
                hlen |= ((size_t)ext[0] << 16) | ((size_t)ext[1] << 24);
            }
            std::string hdr(hlen, '\0');
            if (err.empty() && std::fread(&hdr[0], 1, hlen, f) != hlen) err = "truncated header";
            if (err.empty() && (hdr.find("'<f4'") == std::string::npos ||
                                hdr.find("'fortran_order': False") == std::string::npos))
                err = "expected a C-order float32 array";
            const size_t sp = hdr.find("'shape': (");
            if (err.empty() && sp == std::string::npos) err = "no shape";
            if (err.empty()) {
                const long count = std::strtol(hdr.c_str() + sp + 10, nullptr, 10);
                if (count != param_count()) {
                    err = "holds " + std::to_string(count) + " values, the network has " +
                          std::to_string(param_count());
                } else {
                    p.resize((size_t)count);
                    if (std::fread(p.data(), sizeof(float), p.size(), f) != p.size())
                        err = "truncated data";
                }
            }
        }
        std::fclose(f);
        if (!err.empty()) throw std::runtime_error("WarmStartNet: " + path + ": " + err);
        set_params(p);
    }

    // h0 <- prediction from K (or ln K with k_is_log); enqueued on ctx's stream.
    void predict(CudaContext& ctx, const Grid3D& grid, const BCSpec& bc, DeviceSpan<const real> K,
                 DeviceSpan<real> h0, bool k_is_log = false) {
        if (K.size() < grid.num_cells() || h0.size() < grid.num_cells())
            throw std::runtime_error("WarmStartNet::predict: buffers shorter than the grid");
        const mf3d_grid g = to_c(grid);
        const mf3d_bc b = to_c(bc);
        mf3d_check(mf3d_wsnet_predict(net_, ctx.c_stream(), &g, &b, K.data(), k_is_log ? 1 : 0,
                                      h0.data(), h0.size()),
                   "wsnet_predict");
    }

private:
    mf3d_wsnet* net_ = nullptr;
};

}  // namespace physics
}  // namespace macroflow3d
