// The pipeline's realisation loop written against include/macroflow3d/: the
// head-solve calls of EnsembleRunner.cu:146-309 (configuration from the YAML
// flow block, one-shot allocation of fields and workspaces, then per
// realisation: K, its statistics, init_head_guess + solve_head, compact and
// padded velocities, velocity diagnostics), with the same call expressions, so
// that a maintainer swaps the headers and this sequence compiles unchanged.
// Transport, I/O and the reference's curand K generator are not on this path:
// K comes from a seeded host field here (see seeded_lnK).
//
// Usage: ensemble_dropin <n> <realisations> <base_seed> <out_prefix>
// Writes <out_prefix>.r<r>.K.bin and .h.bin (float64) and prints one JSON line
// per realisation.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "macroflow3d/physics.hpp"

using namespace macroflow3d;
using namespace macroflow3d::physics;

namespace {

// The parts of the pipeline's configuration tree this path reads.
struct PipelineConfig {
    struct {
        int nx = 32, ny = 32, nz = 32;
        real dx = 1.0;
    } grid;
    io::FlowYamlConfig flow;
    struct {
        std::string method = "par2";
    } transport;
    struct {
        bool velocity_field = true;
    } diagnostics;
    std::uint64_t seed = 0;
};

// Seeded smooth log-conductivity: a sum of 8 cosine modes whose wave vectors
// and phases come from a 64-bit LCG seeded with `seed`.
std::vector<double> seeded_lnK(const Grid3D& g, std::uint64_t seed) {
    std::uint64_t s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto next = [&s]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    double kx[8], ky[8], kz[8], ph[8];
    for (int m = 0; m < 8; ++m) {
        kx[m] = 0.05 + 0.5 * next();
        ky[m] = 0.05 + 0.5 * next();
        kz[m] = 0.05 + 0.5 * next();
        ph[m] = 6.283185307179586 * next();
    }
    std::vector<double> y(g.num_cells());
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                double v = 0.0;
                for (int m = 0; m < 8; ++m) v += std::cos(kx[m] * i + ky[m] * j + kz[m] * k + ph[m]);
                y[g.idx(i, j, k)] = 0.5 * v;
            }
    return y;
}

void write_bin(const std::string& path, const std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

std::vector<double> to_host(DeviceSpan<const real> d, CudaContext& ctx) {
    std::vector<double> h(d.size());
    ctx.synchronize();
    hip_check(hipMemcpy(h.data(), d.data(), d.size() * sizeof(double), hipMemcpyDeviceToHost),
              "to_host");
    return h;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <n> <realisations> <base_seed> <out_prefix>\n", argv[0]);
        return 2;
    }
    PipelineConfig cfg;
    cfg.grid.nx = cfg.grid.ny = cfg.grid.nz = std::atoi(argv[1]);
    const int NR = std::atoi(argv[2]);
    cfg.seed = std::strtoull(argv[3], nullptr, 10);
    const std::string out = argv[4];
    // flow block of a pipeline YAML: MG-preconditioned CG, reference semantics
    cfg.flow.solver = "mg_cg";
    cfg.flow.mg_levels = 3;
    cfg.flow.cg_rtol = 1e-10;
    cfg.flow.cg_check_every = 10;
    cfg.flow.bc = darcy_x(1.0, 0.0);
    try {
        CudaContext ctx(0);

        // ---- one-shot allocation (EnsembleRunner.cu:137-162) ----
        const int nx = cfg.grid.nx, ny = cfg.grid.ny, nz = cfg.grid.nz;
        const real dx = cfg.grid.dx;
        const Grid3D grid(nx, ny, nz, dx, dx, dx);
        const size_t N = grid.num_cells();

        HeadSolveConfig head_cfg = HeadSolveConfig::from_yaml(cfg.flow);

        ScalarField K_field(grid);
        ScalarField head_field(grid);

        FlowWorkspace flow_ws;
        flow_ws.allocate(grid, head_cfg.mg_levels);

        PaddedVelocityField vel(grid);
        VelocityField vel_compact;
        if (cfg.transport.method == "pspta") {
            vel_compact = VelocityField(grid);
        }
        VelocityDiagnostics vel_diag;
        if (cfg.diagnostics.velocity_field) {
            vel_diag.resize(grid);
        }
        std::fprintf(stderr, "fields: K=%zu head=%zu vel=3x%zu, workspace %zu bytes\n", N, N,
                     vel.field_size(), flow_ws.device_bytes());

        // ---- realisation loop (EnsembleRunner.cu:259-309) ----
        for (int r = 0; r < NR; ++r) {
            const std::uint64_t r_seed = cfg.seed + static_cast<std::uint64_t>(r);
            std::vector<double> Kh = seeded_lnK(grid, r_seed);
            for (double& v : Kh) v = std::exp(v);
            hip_check(hipMemcpy(K_field.device_ptr(), Kh.data(), N * sizeof(double),
                                hipMemcpyHostToDevice),
                      "upload K");

            real kmin, kmax, kmean;
            compute_field_stats(K_field.span(), kmin, kmax, kmean, ctx);

            init_head_guess(head_field.span(), grid, cfg.flow.bc, ctx);
            const ScalarField& K_const = K_field;
            HeadSolveResult result = solve_head(head_field.span(), K_const.span(), grid,
                                                cfg.flow.bc, head_cfg, ctx, flow_ws);

            if (cfg.transport.method == "pspta") {
                compute_velocity_from_head(vel_compact, head_field, K_field, grid, cfg.flow.bc, ctx);
                if (cfg.diagnostics.velocity_field) {
                    compute_velocity_from_head(vel, head_field, K_field, grid, cfg.flow.bc, ctx);
                }
            } else {
                compute_velocity_from_head(vel, head_field, K_field, grid, cfg.flow.bc, ctx);
            }
            if (cfg.diagnostics.velocity_field) {
                compute_velocity_diagnostics(vel, vel_diag, grid, ctx);
                print_velocity_diagnostics(vel_diag, r, ctx);
            }
            real div_lo, div_hi, div_mean;
            compute_field_stats(vel_diag.divergence.span(), div_lo, div_hi, div_mean, ctx);

            const std::string pre = out + ".r" + std::to_string(r);
            write_bin(pre + ".K.bin", Kh);
            write_bin(pre + ".h.bin", to_host(head_field.span(), ctx));
            std::printf("{\"r\": %d, \"seed\": %llu, \"iters\": %d, \"converged\": %s, \"r0\": %.17g, "
                        "\"r_final\": %.17g, \"kmin\": %.17g, \"kmax\": %.17g, \"kmean\": %.17g, "
                        "\"div_min\": %.17g, \"div_max\": %.17g, \"solver_type\": %d, "
                        "\"mg_levels\": %d, \"cg_check_every\": %d}\n",
                        r, (unsigned long long)r_seed, result.num_iterations,
                        result.converged ? "true" : "false", result.initial_residual,
                        result.final_residual, kmin, kmax, kmean, div_lo, div_hi,
                        (int)head_cfg.solver_type, head_cfg.mg_levels, head_cfg.cg_check_every);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
