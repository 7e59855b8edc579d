// C++ host surface of libmf3d_hip: numerics.
//
//   blas::ReductionWorkspace       src/numerics/blas/reduction_workspace.cuh:9-31
//   blas::dot_device / dot_host    src/numerics/blas/dot.cuh
//   blas::nrm2_device / nrm2_host  src/numerics/blas/nrm2.cuh
//   blas::axpy/axpby/copy/fill/scal src/numerics/blas/{axpy,axpby,copy,fill,scal}.cuh
//   operators::VarCoeffLaplacian   src/numerics/operators/varcoeff_laplacian.cuh:43-67
//   solvers::MultigridPreconditioner src/numerics/solvers/mg_preconditioner.cuh:34-81
//   solvers::PCGConfig/Result/Workspace, pcg_solve  src/numerics/solvers/pcg.cuh:29-185
//
// pcg_solve here is the generic host-driven loop for any operator and
// preconditioner with apply(ctx, in, out); physics::solve_head runs its own
// device-resident PCG and is the fast path for the head system.
#pragma once

#include <cmath>

#include "multigrid.hpp"

namespace macroflow3d {
namespace blas {

// Reduction scratch (a workspace of its own) and one device scalar.
class ReductionWorkspace {
  public:
    ReductionWorkspace() : d_scalar(1) {
        mf3d_check(mf3d_workspace_create(&ws_), "mf3d_workspace_create");
    }
    ~ReductionWorkspace() {
        if (ws_) mf3d_workspace_destroy(ws_);
    }
    ReductionWorkspace(const ReductionWorkspace&) = delete;
    ReductionWorkspace& operator=(const ReductionWorkspace&) = delete;
    mf3d_workspace* handle() const { return ws_; }
    void ensure_scalar() {
        if (d_scalar.size() < 1) d_scalar.resize(1);
    }

    DeviceBuffer<real> d_scalar;

  private:
    mf3d_workspace* ws_ = nullptr;
};

inline void dot_device(CudaContext& ctx, DeviceSpan<const real> x, DeviceSpan<const real> y,
                       DeviceSpan<real> d_result, ReductionWorkspace& ws) {
    if (x.size() != y.size()) throw std::runtime_error("dot: size mismatch");
    mf3d_check(mf3d_dot(ws.handle(), ctx.c_stream(), x.data(), y.data(), x.size(),
                        d_result.data()),
               "dot");
}

inline real dot_host(CudaContext& ctx, DeviceSpan<const real> x, DeviceSpan<const real> y,
                     ReductionWorkspace& ws) {
    ws.ensure_scalar();
    dot_device(ctx, x, y, ws.d_scalar.span(), ws);
    real v = 0.0;
    hip_check(hipMemcpyAsync(&v, ws.d_scalar.data(), sizeof(real), hipMemcpyDeviceToHost,
                             ctx.stream()),
              "dot_host");
    ctx.synchronize();
    return v;
}

inline void nrm2_device(CudaContext& ctx, DeviceSpan<const real> x, DeviceSpan<real> d_result,
                        ReductionWorkspace& ws) {
    mf3d_check(mf3d_nrm2(ws.handle(), ctx.c_stream(), x.data(), x.size(), d_result.data()),
               "nrm2");
}

inline real nrm2_host(CudaContext& ctx, DeviceSpan<const real> x, ReductionWorkspace& ws) {
    ws.ensure_scalar();
    nrm2_device(ctx, x, ws.d_scalar.span(), ws);
    real v = 0.0;
    hip_check(hipMemcpyAsync(&v, ws.d_scalar.data(), sizeof(real), hipMemcpyDeviceToHost,
                             ctx.stream()),
              "nrm2_host");
    ctx.synchronize();
    return v;
}

inline void axpy(CudaContext& ctx, real a, DeviceSpan<const real> x, DeviceSpan<real> y) {
    mf3d_check(mf3d_axpy(ctx.c_stream(), a, x.data(), y.data(), y.size()), "axpy");
}
inline void axpby(CudaContext& ctx, real a, DeviceSpan<const real> x, real b, DeviceSpan<real> y) {
    mf3d_check(mf3d_axpby(ctx.c_stream(), a, x.data(), b, y.data(), y.size()), "axpby");
}
inline void copy(CudaContext& ctx, DeviceSpan<const real> x, DeviceSpan<real> y) {
    mf3d_check(mf3d_copy(ctx.c_stream(), x.data(), y.data(), y.size()), "copy");
}
inline void fill(CudaContext& ctx, DeviceSpan<real> y, real value) {
    mf3d_check(mf3d_fill(ctx.c_stream(), y.data(), y.size(), value), "fill");
}
inline void scal(CudaContext& ctx, DeviceSpan<real> x, real a) {
    mf3d_check(mf3d_scal(ctx.c_stream(), x.data(), x.size(), a), "scal");
}

}  // namespace blas

namespace operators {

// y = A x = -sum_f K_f (x_C - x_nb) / dx^2 (negative definite; pin as in the reference).
class VarCoeffLaplacian {
  public:
    VarCoeffLaplacian() = default;
    VarCoeffLaplacian(const Grid3D& grid, DeviceSpan<const real> K, const BCSpec& bc,
                      PinSpec pin = {})
        : grid_(grid), K_(K), bc_(bc), pin_(pin) {}

    void apply(CudaContext& ctx, DeviceSpan<const real> x, DeviceSpan<real> y) const {
        const mf3d_grid g = to_c(grid_);
        const mf3d_bc c = to_c(bc_);
        if (x.size() < y.size()) throw std::runtime_error("VarCoeffLaplacian: x shorter than y");
        mf3d_check(mf3d_varcoeff_apply(ctx.c_stream(), &g, &c, pin_.enabled ? 1 : 0, K_.data(),
                                       x.data(), y.data(), y.size()),
                   "VarCoeffLaplacian::apply");
    }
    const Grid3D& grid() const { return grid_; }
    size_t size() const { return grid_.num_cells(); }
    const PinSpec& pin() const { return pin_; }

  private:
    Grid3D grid_;
    DeviceSpan<const real> K_;
    BCSpec bc_;
    PinSpec pin_;
};

// y = -A x, for plain CG on the positive definite -A (cg.cuh).
template <typename Op>
class NegatedOperator {
  public:
    explicit NegatedOperator(const Op& A) : A_(A) {}
    void apply(CudaContext& ctx, DeviceSpan<const real> x, DeviceSpan<real> y) const {
        A_.apply(ctx, x, y);
        blas::scal(ctx, y, -1.0);
    }

  private:
    const Op& A_;
};

}  // namespace operators

namespace solvers {

struct IdentityPreconditioner {
    void apply(CudaContext& ctx, DeviceSpan<const real> r, DeviceSpan<real> z) const {
        blas::copy(ctx, r, z);
    }
};

// z = one V-cycle from x = 0 with b = r on a set-up hierarchy (non-owning).
class MultigridPreconditioner {
  public:
    MultigridPreconditioner() = default;
    MultigridPreconditioner(multigrid::MGHierarchy& hierarchy, const BCSpec& bc,
                            const multigrid::MGConfig& config, PinSpec pin = {})
        : hierarchy_(&hierarchy), bc_(bc), config_(config), pin_(pin) {}
    MultigridPreconditioner(const MultigridPreconditioner&) = delete;
    MultigridPreconditioner& operator=(const MultigridPreconditioner&) = delete;
    MultigridPreconditioner(MultigridPreconditioner&&) = default;
    MultigridPreconditioner& operator=(MultigridPreconditioner&&) = default;

    void apply(CudaContext& ctx, DeviceSpan<const real> r, DeviceSpan<real> z) const {
        if (!hierarchy_) throw std::runtime_error("MultigridPreconditioner: no hierarchy");
        mf3d_check(mf3d_mg_precondition(hierarchy_->handle(), ctx.c_stream(), r.data(), z.data(),
                                        z.size()),
                   "MultigridPreconditioner::apply");
    }
    bool is_valid() const { return hierarchy_ != nullptr; }

  private:
    multigrid::MGHierarchy* hierarchy_ = nullptr;
    BCSpec bc_;
    multigrid::MGConfig config_;
    PinSpec pin_;
};

struct PCGConfig {
    int max_iter = 1000;
    int check_every = 10;
    real rtol = 1e-6;
    bool verbose = false;
};

struct PCGResult {
    int iterations = 0;
    real initial_residual = 0.0;
    real final_residual = 0.0;
    bool converged = false;
};

struct PCGWorkspace {
    DeviceBuffer<real> r, z, p, Ap;
    blas::ReductionWorkspace red;
    void ensure(size_t n) {
        if (r.size() != n) {
            r.resize(n);
            z.resize(n);
            p.resize(n);
            Ap.resize(n);
        }
    }
};

// Preconditioned CG (pcg.cuh:95-185): r = b - A x, z = M r, p = z;
// alpha = (r,z)/(p,Ap); x += alpha p; r -= alpha Ap; every check_every
// iterations stop on ||r|| / ||r0|| < rtol; z = M r; beta = (r,z)_new/(r,z);
// p = z + beta p. Early exit when ||r0|| < 1e-15. A and M may both be negative
// definite (the head operator and its V-cycle), which keeps alpha positive.
template <typename Operator, typename Preconditioner>
PCGResult pcg_solve(CudaContext& ctx, const Operator& A, const Preconditioner& M,
                    DeviceSpan<const real> b, DeviceSpan<real> x, const PCGConfig& cfg,
                    PCGWorkspace& ws) {
    PCGResult res;
    const size_t n = x.size();
    ws.ensure(n);
    auto r = ws.r.span(), z = ws.z.span(), p = ws.p.span(), Ap = ws.Ap.span();
    A.apply(ctx, x, Ap);
    blas::copy(ctx, b, r);
    blas::axpy(ctx, -1.0, Ap, r);
    const real r0 = blas::nrm2_host(ctx, r, ws.red);
    res.initial_residual = r0;
    res.final_residual = r0;
    if (r0 < 1e-15) {
        res.converged = true;
        return res;
    }
    M.apply(ctx, r, z);
    blas::copy(ctx, z, p);
    real rz = blas::dot_host(ctx, r, z, ws.red);
    for (int it = 0; it < cfg.max_iter; ++it) {
        A.apply(ctx, p, Ap);
        const real pAp = blas::dot_host(ctx, p, Ap, ws.red);
        if (pAp == 0.0 || !std::isfinite(pAp)) break;
        const real alpha = rz / pAp;
        blas::axpy(ctx, alpha, p, x);
        blas::axpy(ctx, -alpha, Ap, r);
        res.iterations = it + 1;
        if ((it + 1) % cfg.check_every == 0 || it + 1 == cfg.max_iter) {
            const real rn = blas::nrm2_host(ctx, r, ws.red);
            res.final_residual = rn;
            if (rn / r0 < cfg.rtol) {
                res.converged = true;
                break;
            }
        }
        M.apply(ctx, r, z);
        const real rz_new = blas::dot_host(ctx, r, z, ws.red);
        const real beta = rz_new / rz;
        rz = rz_new;
        blas::axpby(ctx, 1.0, z, beta, p);
    }
    return res;
}

}  // namespace solvers
}  // namespace macroflow3d
