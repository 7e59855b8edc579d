// Address/undefined-behaviour sanitizer run of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: built with -fsanitize=address,undefined together
// with mf3d_oracle.cpp by `make -C oracle sanitize-check`, it calls every
// exported oracle routine on small grids (even, odd and one-cell-thick sizes,
// every boundary mix, pin on and off, both solver semantics) so that an
// out-of-bounds index or an undefined operation in a boundary branch aborts
// the run. It checks a few results for sanity only; the numerical pins of the
// oracle are the Python tests.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
typedef struct { int nx, ny, nz; double dx, dy, dz; } orc_grid;
typedef struct { int type[6]; double value[6]; } orc_bc;
typedef struct {
    int solver_type, mg_levels, mg_pre_smooth, mg_post_smooth, mg_max_cycles, mg_coarse_iters,
        cg_max_iter, cg_check_every;
    double cg_rtol, rtol;
    int pin_mode, init_guess, coarse_solver, stop_norm;
} orc_config;
typedef struct {
    int num_iterations, converged;
    double initial_residual, final_residual, bnorm, true_residual;
} orc_result;
int orc_num_threads(void);
void orc_set_num_threads(int);
void orc_init_head_guess(const orc_grid*, const orc_bc*, double*);
void orc_build_rhs_head(const orc_grid*, const orc_bc*, const double*, double*);
void orc_coarsen_K(const orc_grid*, const double*, double*);
void orc_gsrb_smooth_3d(const orc_grid*, const orc_bc*, int, double*, const double*,
                        const double*, int);
void orc_compute_residual_3d(const orc_grid*, const orc_bc*, int, const double*, const double*,
                             const double*, double*);
void orc_restrict_3d(const orc_grid*, const double*, double*);
void orc_prolong_3d_add(const orc_grid*, const double*, double*);
void orc_varcoeff_apply(const orc_grid*, const orc_bc*, int, const double*, const double*,
                        double*);
double orc_dot(size_t, const double*, const double*);
double orc_nrm2(size_t, const double*);
int orc_dense_inverse(const orc_grid*, const orc_bc*, int, const double*, double*);
int orc_v_cycle_precond(const orc_grid*, const orc_bc*, const orc_config*, const double*,
                        const double*, double*);
int orc_solve_head(const orc_grid*, const orc_bc*, const orc_config*, const double*, double*,
                   orc_result*);
void orc_velocity_compact(const orc_grid*, const orc_bc*, const double*, const double*, double*,
                          double*, double*);
void orc_velocity_padded(const orc_grid*, const orc_bc*, const double*, const double*, double*,
                         double*, double*);
void orc_velocity_diagnostics(const orc_grid*, const double*, const double*, const double*,
                              double*, double*, double*);
void orc_field_stats(size_t, const double*, double*);
int orc_lognormal_lnK(int, int, int, double, double, uint64_t, double*, double*);
void orc_gauss_pair(uint64_t, uint64_t, double*);
}

static int failures = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

static orc_bc make_bc(const int t[6]) {
    orc_bc b;
    for (int f = 0; f < 6; ++f) {
        b.type[f] = t[f];
        b.value[f] = t[f] == 0 ? 1.0 - 0.25 * f : 0.0;
    }
    return b;
}

static bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

static void run_case(const orc_grid& g, const orc_bc& bc, int pin, std::mt19937_64& rng) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    std::normal_distribution<double> nd;
    std::vector<double> K(n), x(n), b(n), r(n), y(n), h(n);
    for (size_t i = 0; i < n; ++i) {
        K[i] = std::exp(0.8 * nd(rng));
        x[i] = nd(rng);
        b[i] = nd(rng);
    }
    orc_init_head_guess(&g, &bc, h.data());
    orc_build_rhs_head(&g, &bc, K.data(), r.data());
    orc_varcoeff_apply(&g, &bc, pin, K.data(), x.data(), y.data());
    CHECK(finite(y));
    orc_compute_residual_3d(&g, &bc, pin, x.data(), b.data(), K.data(), r.data());
    CHECK(finite(r));
    orc_gsrb_smooth_3d(&g, &bc, pin, x.data(), b.data(), K.data(), 2);
    CHECK(finite(x));
    CHECK(orc_dot(n, x.data(), x.data()) >= 0.0 && orc_nrm2(n, x.data()) >= 0.0);
    // transfers and coarsening where every dimension halves
    if (g.nx % 2 == 0 && g.ny % 2 == 0 && g.nz % 2 == 0) {
        orc_grid c{g.nx / 2, g.ny / 2, g.nz / 2, 2 * g.dx, 2 * g.dy, 2 * g.dz};
        std::vector<double> xc((size_t)c.nx * c.ny * c.nz), Kc(xc.size());
        orc_restrict_3d(&c, x.data(), xc.data());
        orc_prolong_3d_add(&c, xc.data(), y.data());
        orc_coarsen_K(&c, K.data(), Kc.data());
        CHECK(finite(Kc));
    }
    if (n <= 512) {
        // nonsingular with a Dirichlet face or the pin; without either the
        // matrix is singular and only the memory accesses are checked
        bool dirichlet = false;
        for (int f = 0; f < 6; ++f) dirichlet = dirichlet || bc.type[f] == 0;
        std::vector<double> inv(n * n);
        const int rc = orc_dense_inverse(&g, &bc, pin, K.data(), inv.data());
        if (dirichlet || pin) CHECK(rc == 0 && finite(inv));
    }
    // velocities and diagnostics, compact and padded
    const size_t nu = (size_t)(g.nx + 1) * g.ny * g.nz, nv = (size_t)g.nx * (g.ny + 1) * g.nz,
                 nw = (size_t)g.nx * g.ny * (g.nz + 1),
                 np = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    std::vector<double> U(nu), V(nv), W(nw);
    orc_velocity_compact(&g, &bc, x.data(), K.data(), U.data(), V.data(), W.data());
    CHECK(finite(U) && finite(V) && finite(W));
    std::vector<double> Up(np, 0.0), Vp(np, 0.0), Wp(np, 0.0), dv(n), om(n), he(n);
    orc_velocity_padded(&g, &bc, x.data(), K.data(), Up.data(), Vp.data(), Wp.data());
    orc_velocity_diagnostics(&g, Up.data(), Vp.data(), Wp.data(), dv.data(), om.data(), he.data());
    CHECK(finite(dv) && finite(om) && finite(he));
    double st[8] = {0};
    orc_field_stats(n, dv.data(), st);
}

static void run_solves(const orc_grid& g, const orc_bc& bc, std::mt19937_64& rng) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    std::normal_distribution<double> nd;
    std::vector<double> K(n), h(n), r(n), z(n);
    for (size_t i = 0; i < n; ++i) K[i] = std::exp(nd(rng));
    for (int solver = 0; solver < 3; ++solver)
        for (int mode = 0; mode < 2; ++mode) {
            orc_config c{solver, 3, 2, 2, 20, 50, 200, mode ? 1 : 10, 1e-8, 1e-6, 0, mode ? 1 : 0,
                         mode, mode};
            orc_result res;
            CHECK(orc_solve_head(&g, &bc, &c, K.data(), h.data(), &res) == 0);
            CHECK(finite(h));
            if (solver == 2) {
                for (size_t i = 0; i < n; ++i) r[i] = nd(rng);
                CHECK(orc_v_cycle_precond(&g, &bc, &c, K.data(), r.data(), z.data()) == 0);
                CHECK(finite(z));
            }
        }
}

int main() {
    orc_set_num_threads(2);
    CHECK(orc_num_threads() >= 1);
    std::mt19937_64 rng(12345);
    const int mixes[6][6] = {{0, 0, 1, 1, 1, 1}, {0, 0, 0, 0, 0, 0}, {0, 0, 2, 2, 2, 2},
                             {2, 2, 1, 1, 0, 0}, {1, 1, 1, 1, 1, 1}, {2, 2, 2, 2, 2, 2}};
    const int shapes[6][3] = {{8, 6, 4}, {5, 7, 3}, {1, 1, 9}, {9, 1, 1}, {2, 2, 2}, {4, 3, 1}};
    int cases = 0;
    for (const auto& m : mixes) {
        const orc_bc bc = make_bc(m);
        for (const auto& s : shapes) {
            orc_grid g{s[0], s[1], s[2], 0.5, 0.5, 0.5};
            for (int pin = 0; pin < 2; ++pin) {
                run_case(g, bc, pin, rng);
                ++cases;
            }
        }
        orc_grid gs{16, 8, 8, 1.0, 1.0, 1.0};
        run_solves(gs, bc, rng);
    }
    std::vector<double> lnK(6 * 5 * 4);
    double stats[5];
    CHECK(orc_lognormal_lnK(6, 5, 4, 1.0, 2.0, 7, lnK.data(), stats) == 0);
    CHECK(finite(lnK));
    double g2[2];
    orc_gauss_pair(7, 3, g2);
    CHECK(std::isfinite(g2[0]) && std::isfinite(g2[1]));
    std::printf("sanitize check: %d field cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
