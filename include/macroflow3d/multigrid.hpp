// C++ host surface of libmf3d_hip: multigrid.
//
//   MGConfig / MGHierarchy  src/multigrid/mg_types.hpp:16-104
//   gsrb_smooth_3d          src/multigrid/smoothers/gsrb_3d.cuh:29-31
//   compute_residual_3d     src/multigrid/residual/residual_3d.cuh:30-32
//   restrict_3d             src/multigrid/transfer/restrict_3d.cuh:14-15
//   prolong_3d_add          src/multigrid/transfer/prolong_3d.cuh:14-15
//
// The hierarchy's level vectors, coarsened K and face coefficients live in a
// device workspace owned by MGHierarchy; setup() fills it for one K field
// (what solve_head.cu:89-114 does before the solve).
#pragma once

#include "core.hpp"

namespace macroflow3d {
namespace multigrid {

struct MGConfig {
    int num_levels = 4;
    int pre_smooth = 2;
    int post_smooth = 2;
    int coarse_solve_iters = 50;
};

// Owns the device hierarchy (an mf3d_workspace). Move-only.
class MGHierarchy {
  public:
    MGHierarchy() { mf3d_check(mf3d_workspace_create(&ws_), "mf3d_workspace_create"); }
    ~MGHierarchy() {
        if (ws_) mf3d_workspace_destroy(ws_);
    }
    MGHierarchy(const MGHierarchy&) = delete;
    MGHierarchy& operator=(const MGHierarchy&) = delete;
    MGHierarchy(MGHierarchy&& o) noexcept : ws_(o.ws_) { o.ws_ = nullptr; }
    // MGHierarchy(grid, num_levels) (mg_types.hpp:75-98): levels and vectors
    // allocated now (halving rule, stopping before a dimension below 2)
    MGHierarchy(const Grid3D& grid, int num_levels) : MGHierarchy() { reserve(grid, num_levels); }
    MGHierarchy& operator=(MGHierarchy&& o) noexcept {
        std::swap(ws_, o.ws_);
        return *this;
    }

    // Levels, coarsened K and coefficients for K on `grid` (c: full solver
    // configuration, only its multigrid, pin and precision fields matter).
    void setup(CudaContext& ctx, const Grid3D& grid, const BCSpec& bc,
               DeviceSpan<const real> K, const mf3d_head_solve_config& c) {
        bc.validate();
        const mf3d_grid g = to_c(grid);
        const mf3d_bc b = to_c(bc);
        mf3d_check(mf3d_mg_setup(ws_, ctx.c_stream(), &g, &b, &c, K.data(), K.size()),
                   "mf3d_mg_setup");
    }
    void setup(CudaContext& ctx, const Grid3D& grid, const BCSpec& bc,
               DeviceSpan<const real> K, const MGConfig& cfg, PinMode pin = PinMode::Auto) {
        mf3d_head_solve_config c;
        mf3d_default_config(&c);
        c.mg_levels = cfg.num_levels;
        c.mg_pre_smooth = cfg.pre_smooth;
        c.mg_post_smooth = cfg.post_smooth;
        c.mg_coarse_iters = cfg.coarse_solve_iters;
        c.pin_mode = pin == PinMode::On ? MF3D_PIN_ON : pin == PinMode::Off ? MF3D_PIN_OFF
                                                                            : MF3D_PIN_AUTO;
        setup(ctx, grid, bc, K, c);
    }

    // sizes the hierarchy and the solve vectors for grid (a no-op when they fit)
    void reserve(const Grid3D& grid, int num_levels) {
        const mf3d_grid g = to_c(grid);
        mf3d_check(mf3d_workspace_reserve(ws_, &g, num_levels), "mf3d_workspace_reserve");
    }

    int num_levels() const { return mf3d_mg_num_levels(ws_); }
    size_t device_bytes() const { return mf3d_workspace_bytes(ws_); }
    void set_tuning(int key, long value) {
        mf3d_check(mf3d_workspace_set_tuning(ws_, key, value), "mf3d_workspace_set_tuning");
    }
    mf3d_workspace* handle() const { return ws_; }

  private:
    mf3d_workspace* ws_ = nullptr;
};

// num_iters red-then-black Gauss-Seidel iterations of -div(K grad x) = b.
inline void gsrb_smooth_3d(CudaContext& ctx, const Grid3D& grid, DeviceSpan<real> x,
                           DeviceSpan<const real> b, DeviceSpan<const real> K, int num_iters,
                           const BCSpec& bc, PinSpec pin = {}) {
    const mf3d_grid g = to_c(grid);
    const mf3d_bc c = to_c(bc);
    if (b.size() < x.size() || K.size() < x.size())
        throw std::runtime_error("gsrb_smooth_3d: b/K shorter than x");
    mf3d_check(mf3d_gsrb_smooth_3d(ctx.c_stream(), &g, x.data(), b.data(), K.data(), x.size(),
                                   num_iters, &c, pin.enabled ? 1 : 0),
               "gsrb_smooth_3d");
}

// r = b - A x in the smoother's convention.
inline void compute_residual_3d(CudaContext& ctx, const Grid3D& grid, DeviceSpan<const real> x,
                                DeviceSpan<const real> b, DeviceSpan<const real> K,
                                DeviceSpan<real> r, const BCSpec& bc, PinSpec pin = {}) {
    const mf3d_grid g = to_c(grid);
    const mf3d_bc c = to_c(bc);
    if (x.size() < r.size() || b.size() < r.size() || K.size() < r.size())
        throw std::runtime_error("compute_residual_3d: input shorter than r");
    mf3d_check(mf3d_compute_residual_3d(ctx.c_stream(), &g, x.data(), b.data(), K.data(),
                                        r.data(), r.size(), &c, pin.enabled ? 1 : 0),
               "compute_residual_3d");
}

// b_coarse = mean of the 8 children of each coarse cell.
inline void restrict_3d(CudaContext& ctx, const Grid3D& fine, const Grid3D& coarse,
                        DeviceSpan<const real> r_fine, DeviceSpan<real> b_coarse) {
    const mf3d_grid f = to_c(fine), c = to_c(coarse);
    mf3d_check(mf3d_restrict_3d(ctx.c_stream(), &f, &c, r_fine.data(), r_fine.size(),
                                b_coarse.data(), b_coarse.size()),
               "restrict_3d");
}

// x_fine += x_coarse of the parent cell (piecewise constant).
inline void prolong_3d_add(CudaContext& ctx, const Grid3D& coarse, const Grid3D& fine,
                           DeviceSpan<const real> x_coarse, DeviceSpan<real> x_fine) {
    const mf3d_grid c = to_c(coarse), f = to_c(fine);
    mf3d_check(mf3d_prolong_3d_add(ctx.c_stream(), &c, &f, x_coarse.data(), x_coarse.size(),
                                   x_fine.data(), x_fine.size()),
               "prolong_3d_add");
}

}  // namespace multigrid
}  // namespace macroflow3d
