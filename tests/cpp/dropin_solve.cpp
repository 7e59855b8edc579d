// Drop-in check of the C++ host surface (include/macroflow3d/*.hpp): the call
// sequence of the pipeline (EnsembleRunner.cu:286-309: solve_head, then
// compute_velocity_from_head, then the checksums) and of a caller that builds
// its own PCG from VarCoeffLaplacian + MultigridPreconditioner (pcg.cuh).
//
// Usage: dropin_solve <n> <out_prefix> [warm_start.npy]
// K = exp(0.8 sin(0.31 i + 0.17 j) cos(0.23 k - 0.11 i)) on an n^3 grid, h = 1
// at x = 0, h = 0 at x = L, no-flow elsewhere. Writes <out_prefix>.K.bin,
// .h_ref.bin (reference semantics), .h_ns.bin (direct coarsest, ||r||/||b||),
// .h_pcg.bin (generic pcg_solve), .U.bin/.V.bin/.W.bin (velocity of h_ns),
// .Up.bin (padded U of h_ns) as raw float64, and prints one JSON line.
// Extensions: a mixed-precision solve, the same solve batched with a second
// realisation (K^2) and compared bit for bit with its single solve, and, given
// the network file, the learned initial head (.h0.bin) and a solve from it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "macroflow3d/physics.hpp"

using namespace macroflow3d;

static void write_bin(const std::string& path, const std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static std::vector<double> to_host(DeviceSpan<const real> d, CudaContext& ctx) {
    std::vector<double> h(d.size());
    ctx.synchronize();
    hip_check(hipMemcpy(h.data(), d.data(), d.size() * sizeof(double), hipMemcpyDeviceToHost),
              "to_host");
    return h;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <n> <out_prefix>\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const std::string out = argv[2];
    try {
        CudaContext ctx(0);
        const Grid3D grid(n, n, n, 1.0, 1.0, 1.0);
        const BCSpec bc = darcy_x(1.0, 0.0);
        const size_t N = grid.num_cells();

        std::vector<double> Kh(N);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    Kh[grid.idx(i, j, k)] =
                        std::exp(0.8 * std::sin(0.31 * i + 0.17 * j) * std::cos(0.23 * k - 0.11 * i));
        KField K(grid);
        hip_check(hipMemcpy(K.device_ptr(), Kh.data(), N * sizeof(double), hipMemcpyHostToDevice),
                  "upload K");

        // pipeline sequence, reference semantics (defaults + 1e-10, 3 levels)
        FlowWorkspace ws;
        HeadField h(grid);
        physics::HeadSolveConfig cfg;
        cfg.mg_levels = 3;
        cfg.cg_rtol = 1e-10;
        const physics::HeadSolveResult r_ref =
            physics::solve_head(h.span(), K.span(), grid, bc, cfg, ctx, ws);
        write_bin(out + ".h_ref.bin", to_host(h.span(), ctx));

        // extension semantics in the same workspace
        const physics::HeadSolveConfig ns = physics::HeadSolveConfig::fp64_direct(1e-10, 3);
        const physics::HeadSolveResult r_ns =
            physics::solve_head(h.span(), K.span(), grid, bc, ns, ctx, ws);
        write_bin(out + ".h_ns.bin", to_host(h.span(), ctx));

        VelocityField vel(grid);
        physics::compute_velocity_from_head(vel, h, K, grid, bc, ctx);
        PaddedVelocityField pv(grid);
        physics::compute_velocity_from_head(pv, h, K, grid, bc, ctx);
        write_bin(out + ".U.bin", to_host(vel.U_span(), ctx));
        write_bin(out + ".V.bin", to_host(vel.V_span(), ctx));
        write_bin(out + ".W.bin", to_host(vel.W_span(), ctx));
        write_bin(out + ".Up.bin", to_host(pv.U_span(), ctx));
        const double usum = physics::compute_sum(vel.U_span(), ctx);
        const bool finite = physics::check_no_nans(vel.U_span(), ctx);
        // mass-conservation check of the padded hand-off (velocity_diagnostics.cu:147-199)
        physics::VelocityDiagnostics diag(grid);
        physics::compute_velocity_diagnostics(pv, diag, grid, ctx);
        physics::print_velocity_diagnostics(diag, 0, ctx);
        write_bin(out + ".div.bin", to_host(diag.divergence.span(), ctx));
        write_bin(out + ".omag.bin", to_host(diag.vorticity_mag.span(), ctx));
        write_bin(out + ".hel.bin", to_host(diag.helicity.span(), ctx));
        real dmin = 0, dmax = 0, dmean = 0;
        physics::compute_field_stats(diag.divergence.span(), dmin, dmax, dmean, ctx);

        // generic PCG: operator + MG preconditioner of the reference interfaces
        multigrid::MGHierarchy hier;
        multigrid::MGConfig mgc;
        mgc.num_levels = 3;
        hier.setup(ctx, grid, bc, K.span(), mgc);
        const PinSpec pin(pin_enabled(PinMode::Auto, bc));
        operators::VarCoeffLaplacian A(grid, K.span(), bc, pin);
        solvers::MultigridPreconditioner M(hier, bc, mgc, pin);
        DeviceBuffer<real> rhs(N), x(N);
        physics::build_rhs_head(rhs.span(), K.span(), grid, bc, ctx);
        physics::init_head_guess(x.span(), grid, bc, ctx);
        solvers::PCGConfig pc;
        pc.rtol = 1e-10;
        pc.check_every = 10;
        solvers::PCGWorkspace pws;
        const solvers::PCGResult r_pcg = solvers::pcg_solve(ctx, A, M, rhs.span(), x.span(), pc, pws);
        write_bin(out + ".h_pcg.bin", to_host(x.span(), ctx));
        write_bin(out + ".K.bin", Kh);

        // extensions: mixed precision, batched, learned warm start
        const physics::HeadSolveConfig mx = physics::HeadSolveConfig::mixed(1e-10, 3);
        HeadField hm(grid);
        const physics::HeadSolveResult r_mx =
            physics::solve_head(hm.span(), K.span(), grid, bc, mx, ctx, ws);
        std::vector<double> K2h(N);
        for (size_t c = 0; c < N; ++c) K2h[c] = Kh[c] * Kh[c];
        KField K2(grid);
        hip_check(hipMemcpy(K2.device_ptr(), K2h.data(), N * sizeof(double), hipMemcpyHostToDevice),
                  "upload K2");
        HeadField h2(grid);
        const physics::HeadSolveResult r_2 =
            physics::solve_head(h2.span(), K2.span(), grid, bc, mx, ctx, ws);
        FlowWorkspace wa, wb;
        HeadField ha(grid), hb(grid);
        const std::vector<physics::HeadSolveResult> rb = physics::solve_head_batched(
            {ha.span(), hb.span()}, {K.span(), K2.span()}, {grid, grid}, bc, mx, ctx, {&wa, &wb});
        const bool batched_equal = to_host(ha.span(), ctx) == to_host(hm.span(), ctx) &&
                                   to_host(hb.span(), ctx) == to_host(h2.span(), ctx) &&
                                   rb[0].num_iterations == r_mx.num_iterations &&
                                   rb[1].num_iterations == r_2.num_iterations;
        int ws_iters = -1, ws_converged = 0;
        if (argc > 3) {
            physics::WarmStartNet net(argv[3]);
            HeadField hw(grid);
            net.predict(ctx, grid, bc, K.span(), hw.span());
            write_bin(out + ".h0.bin", to_host(hw.span(), ctx));
            physics::HeadSolveConfig wc = mx;
            wc.init_guess = physics::InitGuess::Provided;
            const physics::HeadSolveResult r_w =
                physics::solve_head(hw.span(), K.span(), grid, bc, wc, ctx, ws);
            ws_iters = r_w.num_iterations;
            ws_converged = r_w.converged && r_w.true_residual <= 1e-10 * r_w.bnorm;
        }

        std::printf("{\"mixed_iters\": %d, \"mixed_converged\": %d, \"batched_equal\": %d, "
                    "\"warm_iters\": %d, \"warm_converged\": %d, ",
                    r_mx.num_iterations, (int)(r_mx.converged && r_mx.true_residual <= 1e-10 * r_mx.bnorm),
                    (int)batched_equal, ws_iters, ws_converged);
        std::printf("\"n\": %d, \"ref_iters\": %d, \"ref_converged\": %d, \"ref_r0\": %.17g, "
                    "\"ref_rfinal\": %.17g, \"ns_iters\": %d, \"ns_converged\": %d, "
                    "\"ns_true_residual\": %.17g, \"ns_bnorm\": %.17g, \"pcg_iters\": %d, "
                    "\"pcg_converged\": %d, \"pcg_r0\": %.17g, \"U_sum\": %.17g, "
                    "\"U_finite\": %d, \"levels\": %d, \"div_min\": %.17g, "
                    "\"div_max\": %.17g, \"div_mean\": %.17g}\n",
                    n, r_ref.num_iterations, (int)r_ref.converged, r_ref.initial_residual,
                    r_ref.final_residual, r_ns.num_iterations, (int)r_ns.converged,
                    r_ns.true_residual, r_ns.bnorm, r_pcg.iterations, (int)r_pcg.converged,
                    r_pcg.initial_residual, usum, (int)finite, hier.num_levels(), dmin, dmax,
                    dmean);

        // errors surface as exceptions (BCSpec validation, cuda_check.cuh:12-18)
        BCSpec bad = bc;
        bad.ymin = BCFace(BCType::Periodic, 0.0);
        bool threw = false;
        try {
            physics::solve_head(h.span(), K.span(), grid, bad, cfg, ctx, ws);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!threw) {
            std::fprintf(stderr, "unpaired periodic BC was accepted\n");
            return 1;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
