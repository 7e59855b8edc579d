// CPU FP64 oracle for the Darcy head-solve path of MacroFlow3D.
//
// TEST INFRASTRUCTURE ONLY. This file is a host restatement of the reference
// algorithm, used by the tests, by __graft_entry__.smoke() and by bench.py's
// CPU-baseline leg, to check the GPU results against. The product
// (macroflow3d_amd) never links, loads or calls it.
//
// Every function names the reference file:line it restates (paths relative to
// the MacroFlow3D source root). The reference path is CUDA C++ and cannot be
// compiled here, so this restatement is pinned instead by closed forms,
// discrete invariants and dense brute-force solves (tests/test_oracle_*.py).
//
// Arithmetic: plain IEEE FP64, built with -ffp-contract=off so that every
// multiply and add rounds separately; summation orders follow the reference
// kernels term by term (interior kernels and boundary kernels use different
// orders in the reference, and so do they here).
//
// Two solver semantics are provided:
//   * reference semantics ("mode R"): linear initial guess, GSRB coarsest solve,
//     ||r||/||r0|| checked every cg_check_every iterations;
//   * extension semantics ("mode N"): dense direct coarsest solve, ||r||/||b||
//     stopping rule, choice of initial guess, true residual recomputed.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Mirrors Grid3D (src/core/Grid3D.hpp:32-55).
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz;
} orc_grid;

// Mirrors BCSpecDevice (src/core/BCSpecDevice.cuh:12-20). Face order
// xmin, xmax, ymin, ymax, zmin, zmax; type follows BCType
// (src/core/BCSpec.hpp:13): 0 Dirichlet, 1 Neumann, 2 Periodic.
typedef struct {
    int type[6];
    double value[6];
} orc_bc;

// HeadSolveConfig (src/physics/flow/solve_head.cuh:64-131) plus extensions.
typedef struct {
    int solver_type; // HeadSolverType: 0 CG, 1 MG, 2 PCG_MG
    int mg_levels;
    int mg_pre_smooth;
    int mg_post_smooth;
    int mg_max_cycles;
    int mg_coarse_iters;
    int cg_max_iter;
    int cg_check_every;
    double cg_rtol;
    double rtol;
    int pin_mode;      // PinMode: 0 Auto, 1 On, 2 Off
    int init_guess;    // 0 reference guess (linear in x / zero), 1 zero, 2 caller's h
    int coarse_solver; // 0 GSRB iterations (reference), 1 dense direct
    int stop_norm;     // 0 ||r||/||r0|| (reference), 1 ||r||/||b||
} orc_config;

typedef struct {
    int num_iterations;
    int converged;
    double initial_residual;
    double final_residual;
    double bnorm;         // ||b||
    double true_residual; // ||b - A h|| after the solve (operator convention)
} orc_result;

} // extern "C"

namespace {

enum { BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_PERIODIC = 2 };

inline size_t cell(const orc_grid& g, int i, int j, int k) {
    return (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * (size_t)k);
}
inline size_t ncells(const orc_grid& g) { return (size_t)g.nx * g.ny * g.nz; }
inline double hmean_op(double a, double b) { return 2.0 / (1.0 / a + 1.0 / b); }

// ---------------------------------------------------------------------------
// Boundary-aware neighbour helper.
// Follows bc_helpers::neighbor_{x,y,z}{minus,plus}
// (src/multigrid/smoothers/bc_stencil_helpers.cuh:33-267):
//   interior or periodic neighbour -> coef = harmonic mean, value = x[nb]
//   Neumann                        -> coef = 0, value = x[C]
//   Dirichlet                      -> coef = 2*K_C, value = 0
// dir: 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+.
// ---------------------------------------------------------------------------
inline void neighbour(const orc_grid& g, const orc_bc& bc, const double* x, const double* K,
                      int i, int j, int k, int dir, double& val, double& coef) {
    const size_t c = cell(g, i, j, k);
    const double Kc = K[c];
    int ni = i, nj = j, nk = k;
    bool at_bnd = false;
    switch (dir) {
    case 0: at_bnd = (i == 0); ni = at_bnd ? g.nx - 1 : i - 1; break;
    case 1: at_bnd = (i == g.nx - 1); ni = at_bnd ? 0 : i + 1; break;
    case 2: at_bnd = (j == 0); nj = at_bnd ? g.ny - 1 : j - 1; break;
    case 3: at_bnd = (j == g.ny - 1); nj = at_bnd ? 0 : j + 1; break;
    case 4: at_bnd = (k == 0); nk = at_bnd ? g.nz - 1 : k - 1; break;
    default: at_bnd = (k == g.nz - 1); nk = at_bnd ? 0 : k + 1; break;
    }
    if (at_bnd) {
        const int t = bc.type[dir];
        if (t == BC_NEUMANN) {
            val = x[c];
            coef = 0.0;
            return;
        }
        if (t == BC_DIRICHLET) {
            val = 0.0;
            coef = 2.0 * Kc;
            return;
        }
    }
    const size_t n = cell(g, ni, nj, nk);
    coef = hmean_op(Kc, K[n]);
    val = x[n];
}

inline bool is_interior(const orc_grid& g, int i, int j, int k) {
    return i > 0 && i < g.nx - 1 && j > 0 && j < g.ny - 1 && k > 0 && k < g.nz - 1;
}

// Boundary cells in the reference's launch order: 6 faces (without their
// edges), 12 edges, 8 vertices (gsrb_3d.cu:392-575, residual_3d.cu:392-500).
// Vertices appear once; the reference also touches them from the edge
// kernels first, which is idempotent for a colour pass.
void boundary_cells(const orc_grid& g, std::vector<int>& out) {
    out.clear();
    const int X = g.nx - 1, Y = g.ny - 1, Z = g.nz - 1;
    auto push = [&](int i, int j, int k) {
        out.push_back(i);
        out.push_back(j);
        out.push_back(k);
    };
    // faces: xmin, xmax (j,k interior), ymin, ymax (i,k interior), zmin, zmax (i,j interior)
    for (int f = 0; f < 2; ++f)
        for (int k = 1; k < Z; ++k)
            for (int j = 1; j < Y; ++j) push(f ? X : 0, j, k);
    for (int f = 0; f < 2; ++f)
        for (int k = 1; k < Z; ++k)
            for (int i = 1; i < X; ++i) push(i, f ? Y : 0, k);
    for (int f = 0; f < 2; ++f)
        for (int j = 1; j < Y; ++j)
            for (int i = 1; i < X; ++i) push(i, j, f ? Z : 0);
    // edges along z: (x,y) in (min,min),(min,max),(max,min),(max,max)
    const int ex[4] = {0, 0, X, X}, ey[4] = {0, Y, 0, Y};
    for (int e = 0; e < 4; ++e)
        for (int k = 1; k < Z; ++k) push(ex[e], ey[e], k);
    for (int e = 0; e < 4; ++e) // edges along y: (x,z)
        for (int j = 1; j < Y; ++j) push(ex[e], j, ey[e] == 0 ? 0 : Z);
    for (int e = 0; e < 4; ++e) // edges along x: (y,z)
        for (int i = 1; i < X; ++i) push(i, ex[e] == 0 ? 0 : Y, ey[e] == 0 ? 0 : Z);
    // vertices: XMIN_YMIN_ZMIN first (the pinned one)
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) push(a ? X : 0, b ? Y : 0, c ? Z : 0);
    // de-duplicate (degenerate grids with a dimension < 3)
    std::vector<int> uniq;
    std::vector<char> seen(ncells(g), 0);
    for (size_t p = 0; p < out.size(); p += 3) {
        const size_t c = cell(g, out[p], out[p + 1], out[p + 2]);
        if (!seen[c]) {
            seen[c] = 1;
            uniq.push_back(out[p]);
            uniq.push_back(out[p + 1]);
            uniq.push_back(out[p + 2]);
        }
    }
    out.swap(uniq);
}

// ---------------------------------------------------------------------------
// GSRB smoother: gsrb_smooth_3d (src/multigrid/smoothers/gsrb_3d.cu:368-576)
// interior update: gsrb_interior_kernel (gsrb_3d.cu:29-73)
// boundary update: gsrb_face/edge/vertex_kernel (gsrb_3d.cu:76-366)
// ---------------------------------------------------------------------------
inline void gsrb_interior_update(const orc_grid& g, double* x, const double* b, const double* K,
                                 size_t idx, double dx2) {
    const size_t Nx = g.nx, st = (size_t)g.nx * g.ny;
    const double KC = K[idx];
    const double K_xp = 2.0 / (1.0 / KC + 1.0 / K[idx + 1]);
    const double K_xm = 2.0 / (1.0 / KC + 1.0 / K[idx - 1]);
    const double K_yp = 2.0 / (1.0 / KC + 1.0 / K[idx + Nx]);
    const double K_ym = 2.0 / (1.0 / KC + 1.0 / K[idx - Nx]);
    const double K_zp = 2.0 / (1.0 / KC + 1.0 / K[idx + st]);
    const double K_zm = 2.0 / (1.0 / KC + 1.0 / K[idx - st]);
    double result = x[idx + 1] * K_xp;
    result += x[idx - 1] * K_xm;
    result += x[idx + Nx] * K_yp;
    result += x[idx - Nx] * K_ym;
    result += x[idx + st] * K_zp;
    result += x[idx - st] * K_zm;
    const double aC = K_xp + K_xm + K_yp + K_ym + K_zp + K_zm;
    x[idx] = (result - b[idx] * dx2) / aC;
}

inline void gsrb_boundary_update(const orc_grid& g, const orc_bc& bc, bool pin, double* x,
                                 const double* b, const double* K, int i, int j, int k,
                                 double dx2) {
    double v[6], c[6];
    for (int d = 0; d < 6; ++d) neighbour(g, bc, x, K, i, j, k, d, v[d], c[d]);
    const double result =
        c[0] * v[0] + c[1] * v[1] + c[2] * v[2] + c[3] * v[3] + c[4] * v[4] + c[5] * v[5];
    double aC = c[0] + c[1] + c[2] + c[3] + c[4] + c[5];
    const size_t idx = cell(g, i, j, k);
    const double b_adj = b[idx] + (0.0 + 0.0 + 0.0 + 0.0 + 0.0 + 0.0);
    if (pin && idx == 0) aC *= 2.0;
    x[idx] = (result - b_adj * dx2) / aC;
}

void gsrb_smooth(const orc_grid& g, const orc_bc& bc, bool pin, double* x, const double* b,
                 const double* K, int iters) {
    const double dx2 = g.dx * g.dx;
    std::vector<int> bcells;
    boundary_cells(g, bcells);
    for (int it = 0; it < iters; ++it) {
        for (int colour = 0; colour < 2; ++colour) { // red: (i+j+k) even
#pragma omp parallel for schedule(static)
            for (int k = 1; k < g.nz - 1; ++k)
                for (int j = 1; j < g.ny - 1; ++j) {
                    int i0 = 1 + ((1 + j + k + colour) & 1);
                    for (int i = i0; i < g.nx - 1; i += 2)
                        gsrb_interior_update(g, x, b, K, cell(g, i, j, k), dx2);
                }
            for (size_t p = 0; p < bcells.size(); p += 3) {
                const int i = bcells[p], j = bcells[p + 1], k = bcells[p + 2];
                if (((i + j + k) & 1) != colour) continue;
                gsrb_boundary_update(g, bc, pin, x, b, K, i, j, k, dx2);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Residual r = b - A x: compute_residual_3d (src/multigrid/smoothers/residual_3d.cu:378-503)
// interior: residual_interior_kernel (residual_3d.cu:30-80)
// boundary: residual_{face,edge,vertex}_kernel (residual_3d.cu:83-376)
// ---------------------------------------------------------------------------
void residual(const orc_grid& g, const orc_bc& bc, bool pin, const double* x, const double* b,
              const double* K, double* r) {
    const double inv_dx2 = 1.0 / (g.dx * g.dx);
    const double dx2 = g.dx * g.dx;
    const size_t Nx = g.nx, st = (size_t)g.nx * g.ny;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g.nz - 1; ++k)
        for (int j = 1; j < g.ny - 1; ++j)
            for (int i = 1; i < g.nx - 1; ++i) {
                const size_t o = cell(g, i, j, k);
                const double xc = x[o], Kc = K[o];
                double Ax = 0.0;
                Ax -= 2.0 * (xc - x[o + 1]) / (1.0 / Kc + 1.0 / K[o + 1]);
                Ax -= 2.0 * (xc - x[o + Nx]) / (1.0 / Kc + 1.0 / K[o + Nx]);
                Ax -= 2.0 * (xc - x[o - 1]) / (1.0 / Kc + 1.0 / K[o - 1]);
                Ax -= 2.0 * (xc - x[o - Nx]) / (1.0 / Kc + 1.0 / K[o - Nx]);
                Ax -= 2.0 * (xc - x[o + st]) / (1.0 / Kc + 1.0 / K[o + st]);
                Ax -= 2.0 * (xc - x[o - st]) / (1.0 / Kc + 1.0 / K[o - st]);
                r[o] = b[o] - Ax * inv_dx2;
            }
    std::vector<int> bcells;
    boundary_cells(g, bcells);
    for (size_t p = 0; p < bcells.size(); p += 3) {
        const int i = bcells[p], j = bcells[p + 1], k = bcells[p + 2];
        const size_t idx = cell(g, i, j, k);
        double v[6], c[6];
        for (int d = 0; d < 6; ++d) neighbour(g, bc, x, K, i, j, k, d, v[d], c[d]);
        const double xc = x[idx];
        double result = c[0] * (xc - v[0]) + c[1] * (xc - v[1]) + c[2] * (xc - v[2]) +
                        c[3] * (xc - v[3]) + c[4] * (xc - v[4]) + c[5] * (xc - v[5]);
        if (pin && idx == 0) {
            const double aC = c[0] + c[1] + c[2] + c[3] + c[4] + c[5];
            result += aC * xc;
        }
        const double b_adj = b[idx] + (0.0 + 0.0 + 0.0 + 0.0 + 0.0 + 0.0);
        r[idx] = b_adj - (-result / dx2);
    }
}

// ---------------------------------------------------------------------------
// Operator y = A x: VarCoeffLaplacian::apply
// (src/numerics/operators/varcoeff_laplacian.cu:38-277). The pin's diagonal
// sum here counts only in-domain neighbours (varcoeff_laplacian.cu:110-234).
// ---------------------------------------------------------------------------
void varcoeff_apply(const orc_grid& g, const orc_bc& bc, bool pin, const double* K,
                    const double* x, double* y) {
    const int Nx = g.nx, Ny = g.ny, Nz = g.nz;
    const double inv_dx2 = 1.0 / (g.dx * g.dx);
    const size_t st = (size_t)Nx * Ny;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                const size_t idx = cell(g, i, j, k);
                const double xC = x[idx], KC = K[idx];
                if (is_interior(g, i, j, k)) {
                    const double Kxm = 2.0 / (1.0 / KC + 1.0 / K[idx - 1]);
                    const double Kxp = 2.0 / (1.0 / KC + 1.0 / K[idx + 1]);
                    const double Kym = 2.0 / (1.0 / KC + 1.0 / K[idx - Nx]);
                    const double Kyp = 2.0 / (1.0 / KC + 1.0 / K[idx + Nx]);
                    const double Kzm = 2.0 / (1.0 / KC + 1.0 / K[idx - st]);
                    const double Kzp = 2.0 / (1.0 / KC + 1.0 / K[idx + st]);
                    const double Ax = Kxm * (xC - x[idx - 1]) + Kxp * (xC - x[idx + 1]) +
                                      Kym * (xC - x[idx - Nx]) + Kyp * (xC - x[idx + Nx]) +
                                      Kzm * (xC - x[idx - st]) + Kzp * (xC - x[idx + st]);
                    y[idx] = -Ax * inv_dx2;
                    continue;
                }
                double Ax = 0.0, aC = 0.0;
                const int pos[3] = {i, j, k};
                const int lim[3] = {Nx, Ny, Nz};
                const size_t stride[3] = {1, (size_t)Nx, st};
                for (int d = 0; d < 6; ++d) {
                    const int ax = d >> 1;
                    const bool plus = d & 1;
                    const bool inside = plus ? (pos[ax] < lim[ax] - 1) : (pos[ax] > 0);
                    if (inside) {
                        const size_t n = plus ? idx + stride[ax] : idx - stride[ax];
                        const double Kh = 2.0 / (1.0 / KC + 1.0 / K[n]);
                        Ax += Kh * (xC - x[n]);
                        aC += Kh;
                    } else if (bc.type[d] == BC_PERIODIC) {
                        const size_t n = plus ? idx - (size_t)(lim[ax] - 1) * stride[ax]
                                              : idx + (size_t)(lim[ax] - 1) * stride[ax];
                        const double Kh = 2.0 / (1.0 / KC + 1.0 / K[n]);
                        Ax += Kh * (xC - x[n]);
                    } else if (bc.type[d] == BC_DIRICHLET) {
                        Ax += 2.0 * KC * xC;
                    }
                }
                if (pin && idx == 0) Ax += aC * xC;
                y[idx] = -Ax * inv_dx2;
            }
}

// restrict_kernel (src/multigrid/transfer/restrict_3d.cu:12-53)
void restrict_avg(const orc_grid& cg, const double* f, double* c) {
    const int NX = cg.nx, NY = cg.ny, NZ = cg.nz;
    const size_t Nx = 2 * (size_t)NX, Ny = 2 * (size_t)NY, st = Nx * Ny;
#pragma omp parallel for schedule(static)
    for (int K_ = 0; K_ < NZ; ++K_)
        for (int J = 0; J < NY; ++J)
            for (int I = 0; I < NX; ++I) {
                const size_t in = 2 * (size_t)I + 2 * (size_t)J * Nx + 2 * (size_t)K_ * st;
                double result = 0.0;
                result += f[in];
                result += f[in + 1];
                result += f[in + Nx];
                result += f[in + 1 + Nx];
                result += f[in + st];
                result += f[in + 1 + st];
                result += f[in + Nx + st];
                result += f[in + 1 + Nx + st];
                c[I + (size_t)NX * (J + (size_t)NY * K_)] = result / 8.0;
            }
}

// prolong_3d_add (src/multigrid/transfer/prolong_3d.cu:30-621): every fine
// cell (i,j,k) adds coarse cell (i/2, j/2, k/2).
void prolong_add(const orc_grid& cg, const double* c, double* f) {
    const int Nx = 2 * cg.nx, Ny = 2 * cg.ny, Nz = 2 * cg.nz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i)
                f[(size_t)i + (size_t)Nx * (j + (size_t)Ny * k)] +=
                    c[(size_t)(i / 2) + (size_t)cg.nx * (j / 2 + (size_t)cg.ny * (k / 2))];
}

// coarsen_K_kernel (src/physics/flow/coarsen_K.cu:15-53): geometric mean of
// the 2x2x2 children.
void coarsen_K(const orc_grid& cg, const double* Kf, double* Kc) {
    const int NX = cg.nx, NY = cg.ny, NZ = cg.nz;
    const size_t Nx = 2 * (size_t)NX, Ny = 2 * (size_t)NY, st = Nx * Ny;
#pragma omp parallel for schedule(static)
    for (int IZ = 0; IZ < NZ; ++IZ)
        for (int IY = 0; IY < NY; ++IY)
            for (int IX = 0; IX < NX; ++IX) {
                const size_t i000 = 2 * (size_t)IX + 2 * (size_t)IY * Nx + 2 * (size_t)IZ * st;
                const double s = std::log(Kf[i000]) + std::log(Kf[i000 + 1]) +
                                 std::log(Kf[i000 + Nx]) + std::log(Kf[i000 + 1 + Nx]) +
                                 std::log(Kf[i000 + st]) + std::log(Kf[i000 + st + 1]) +
                                 std::log(Kf[i000 + st + Nx]) + std::log(Kf[i000 + st + 1 + Nx]);
                Kc[IX + (size_t)NX * (IY + (size_t)NY * IZ)] = std::exp(s / 8.0);
            }
}

// build_rhs_head (src/physics/flow/rhs_head.cu:121-194)
void build_rhs(const orc_grid& g, const orc_bc& bc, const double* K, double* rhs) {
    const size_t n = ncells(g);
    const double inv_dx2 = 1.0 / (g.dx * g.dx);
    for (size_t i = 0; i < n; ++i) rhs[i] = 0.0;
    for (int f = 0; f < 6; ++f) {
        if (bc.type[f] != BC_DIRICHLET) continue;
        const double h_bc = bc.value[f];
        const int ax = f >> 1;
        const bool plus = f & 1;
        for (int k = 0; k < g.nz; ++k)
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const int p = ax == 0 ? i : (ax == 1 ? j : k);
                    const int lim = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
                    if (p != (plus ? lim - 1 : 0)) continue;
                    const size_t idx = cell(g, i, j, k);
                    rhs[idx] -= 2.0 * h_bc * K[idx] * inv_dx2;
                }
    }
}

// init_head_guess (src/physics/flow/solve_head.cu:29-73)
void init_guess(const orc_grid& g, const orc_bc& bc, double* h) {
    if (bc.type[0] == BC_DIRICHLET && bc.type[1] == BC_DIRICHLET) {
        const double hw = bc.value[0], he = bc.value[1];
        for (int k = 0; k < g.nz; ++k)
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const double t = ((double)i + 0.5) / (double)g.nx;
                    h[cell(g, i, j, k)] = hw + (he - hw) * t;
                }
    } else {
        std::fill(h, h + ncells(g), 0.0);
    }
}

// Reductions: extended-precision accumulation, deterministic order.
double dot(size_t n, const double* a, const double* b) {
    long double s = 0.0L;
    for (size_t i = 0; i < n; ++i) s += (long double)a[i] * (long double)b[i];
    return (double)s;
}
double nrm2(size_t n, const double* a) { return std::sqrt(dot(n, a, a)); }

// pin_enabled (src/numerics/pin_spec.hpp:93-120)
bool pin_enabled(int mode, const orc_bc& bc) {
    if (mode == 1) return true;
    if (mode == 2) return false;
    for (int f = 0; f < 6; ++f)
        if (bc.type[f] == BC_DIRICHLET) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Dense coarse operator: column j of A = -(residual of e_j with b = 0), so the
// matrix is exactly the smoother/residual operator of that level, pin included.
// Inverse through Cholesky of the SPD matrix -A.
// ---------------------------------------------------------------------------
bool dense_inverse(const orc_grid& g, const orc_bc& bc, bool pin, const double* K,
                   std::vector<double>& Ainv) {
    const size_t n = ncells(g);
    std::vector<double> M(n * n), e(n, 0.0), zero(n, 0.0), r(n);
    for (size_t j = 0; j < n; ++j) {
        e[j] = 1.0;
        residual(g, bc, pin, e.data(), zero.data(), K, r.data());
        for (size_t i = 0; i < n; ++i) M[i * n + j] = r[i]; // -A = residual(e, 0)
        e[j] = 0.0;
    }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            const double s = 0.5 * (M[i * n + j] + M[j * n + i]);
            M[i * n + j] = M[j * n + i] = s;
        }
    // Cholesky M = L L^T (lower, in place)
    for (size_t j = 0; j < n; ++j) {
        long double d = M[j * n + j];
        for (size_t p = 0; p < j; ++p) d -= (long double)M[j * n + p] * M[j * n + p];
        if (!(d > 0)) return false;
        const double Ljj = std::sqrt((double)d);
        M[j * n + j] = Ljj;
        for (size_t i = j + 1; i < n; ++i) {
            long double s = M[i * n + j];
            for (size_t p = 0; p < j; ++p) s -= (long double)M[i * n + p] * M[j * n + p];
            M[i * n + j] = (double)(s / Ljj);
        }
    }
    // inverse of M via L^{-1}: M^{-1} = L^{-T} L^{-1}; A^{-1} = -M^{-1}
    std::vector<double> Linv(n * n, 0.0);
    for (size_t c = 0; c < n; ++c) {
        for (size_t i = c; i < n; ++i) {
            long double s = (i == c) ? 1.0L : 0.0L;
            for (size_t p = c; p < i; ++p) s -= (long double)M[i * n + p] * Linv[p * n + c];
            Linv[i * n + c] = (double)(s / M[i * n + i]);
        }
    }
    Ainv.assign(n * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j <= i; ++j) {
            long double s = 0.0L;
            for (size_t p = i; p < n; ++p) s += (long double)Linv[p * n + i] * Linv[p * n + j];
            Ainv[i * n + j] = Ainv[j * n + i] = -(double)s;
        }
    return true;
}

// ---------------------------------------------------------------------------
// Multigrid hierarchy: MGHierarchy (src/multigrid/mg_types.hpp:75-98)
// ---------------------------------------------------------------------------
struct Level {
    orc_grid g;
    std::vector<double> x, b, r, K;
};

struct Hierarchy {
    std::vector<Level> lv;
    std::vector<double> coarse_inv; // dense inverse of the coarsest operator (mode N)
};

void build_hierarchy(Hierarchy& H, const orc_grid& fine, int num_levels) {
    H.lv.clear();
    orc_grid cur = fine;
    for (int l = 0; l < num_levels; ++l) {
        Level L;
        L.g = cur;
        const size_t n = ncells(cur);
        L.x.assign(n, 0.0);
        L.b.assign(n, 0.0);
        L.r.assign(n, 0.0);
        L.K.assign(n, 0.0);
        H.lv.push_back(std::move(L));
        if (l < num_levels - 1) {
            cur.nx /= 2;
            cur.ny /= 2;
            cur.nz /= 2;
            cur.dx *= 2.0;
            cur.dy *= 2.0;
            cur.dz *= 2.0;
            if (cur.nx < 2 || cur.ny < 2 || cur.nz < 2) break;
        }
    }
}

struct MGParams {
    int pre, post, coarse_iters;
    int coarse_solver;
    bool pin;
    orc_bc bc;
};

void coarse_solve(Hierarchy& H, const MGParams& P) {
    Level& L = H.lv.back();
    if (P.coarse_solver == 1) {
        const size_t n = ncells(L.g);
        for (size_t i = 0; i < n; ++i) {
            long double s = 0.0L;
            for (size_t j = 0; j < n; ++j) s += (long double)H.coarse_inv[i * n + j] * L.b[j];
            L.x[i] = (double)s;
        }
    } else {
        gsrb_smooth(L.g, P.bc, P.pin, L.x.data(), L.b.data(), L.K.data(), P.coarse_iters);
    }
}

// v_cycle_recursive (src/multigrid/cycle/v_cycle.cu:36-74)
void v_cycle(Hierarchy& H, int level, const MGParams& P) {
    const int nl = (int)H.lv.size();
    if (level == nl - 1) {
        coarse_solve(H, P);
        return;
    }
    Level& F = H.lv[level];
    Level& C = H.lv[level + 1];
    gsrb_smooth(F.g, P.bc, P.pin, F.x.data(), F.b.data(), F.K.data(), P.pre);
    residual(F.g, P.bc, P.pin, F.x.data(), F.b.data(), F.K.data(), F.r.data());
    restrict_avg(C.g, F.r.data(), C.b.data());
    std::fill(C.x.begin(), C.x.end(), 0.0);
    v_cycle(H, level + 1, P);
    prolong_add(C.g, C.x.data(), F.x.data());
    gsrb_smooth(F.g, P.bc, P.pin, F.x.data(), F.b.data(), F.K.data(), P.post);
}

} // namespace

// ===========================================================================
// C entry points (ctypes)
// ===========================================================================
extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void orc_init_head_guess(const orc_grid* g, const orc_bc* bc, double* h) { init_guess(*g, *bc, h); }

void orc_build_rhs_head(const orc_grid* g, const orc_bc* bc, const double* K, double* rhs) {
    build_rhs(*g, *bc, K, rhs);
}

void orc_coarsen_K(const orc_grid* coarse, const double* Kf, double* Kc) {
    coarsen_K(*coarse, Kf, Kc);
}

void orc_gsrb_smooth_3d(const orc_grid* g, const orc_bc* bc, int pin, double* x, const double* b,
                        const double* K, int num_iters) {
    gsrb_smooth(*g, *bc, pin != 0, x, b, K, num_iters);
}

void orc_compute_residual_3d(const orc_grid* g, const orc_bc* bc, int pin, const double* x,
                             const double* b, const double* K, double* r) {
    residual(*g, *bc, pin != 0, x, b, K, r);
}

void orc_restrict_3d(const orc_grid* coarse, const double* f, double* c) {
    restrict_avg(*coarse, f, c);
}

void orc_prolong_3d_add(const orc_grid* coarse, const double* c, double* f) {
    prolong_add(*coarse, c, f);
}

void orc_varcoeff_apply(const orc_grid* g, const orc_bc* bc, int pin, const double* K,
                        const double* x, double* y) {
    varcoeff_apply(*g, *bc, pin != 0, K, x, y);
}

double orc_dot(size_t n, const double* a, const double* b) { return dot(n, a, b); }
double orc_nrm2(size_t n, const double* a) { return nrm2(n, a); }

int orc_dense_inverse(const orc_grid* g, const orc_bc* bc, int pin, const double* K,
                      double* Ainv) {
    std::vector<double> inv;
    if (!dense_inverse(*g, *bc, pin != 0, K, inv)) return 1;
    std::memcpy(Ainv, inv.data(), inv.size() * sizeof(double));
    return 0;
}

// One V-cycle from x = 0 with b given on the finest level
// (MultigridPreconditioner::apply, src/numerics/solvers/mg_preconditioner.cu:16-45).
int orc_v_cycle_precond(const orc_grid* g, const orc_bc* bc, const orc_config* cfg,
                        const double* K, const double* r_in, double* z_out) {
    Hierarchy H;
    build_hierarchy(H, *g, cfg->mg_levels);
    MGParams P{cfg->mg_pre_smooth, cfg->mg_post_smooth, cfg->mg_coarse_iters, cfg->coarse_solver,
               pin_enabled(cfg->pin_mode, *bc), *bc};
    const size_t n = ncells(*g);
    std::memcpy(H.lv[0].K.data(), K, n * sizeof(double));
    for (size_t l = 0; l + 1 < H.lv.size(); ++l)
        coarsen_K(H.lv[l + 1].g, H.lv[l].K.data(), H.lv[l + 1].K.data());
    if (P.coarse_solver == 1 &&
        !dense_inverse(H.lv.back().g, *bc, P.pin, H.lv.back().K.data(), H.coarse_inv))
        return 2;
    std::memcpy(H.lv[0].b.data(), r_in, n * sizeof(double));
    std::fill(H.lv[0].x.begin(), H.lv[0].x.end(), 0.0);
    v_cycle(H, 0, P);
    std::memcpy(z_out, H.lv[0].x.data(), n * sizeof(double));
    return 0;
}

// solve_head (src/physics/flow/solve_head.cu:75-243) with pcg_solve
// (src/numerics/solvers/pcg.cuh:95-185), mg_solve (v_cycle.cu:76-116) and
// cg_solve (src/numerics/solvers/cg.cuh:15-172) on the negated operator.
int orc_solve_head(const orc_grid* gp, const orc_bc* bcp, const orc_config* cfg,
                   const double* K, double* h, orc_result* res) {
    const orc_grid& g = *gp;
    const orc_bc& bc = *bcp;
    const size_t n = ncells(g);
    std::memset(res, 0, sizeof(orc_result));

    Hierarchy H;
    build_hierarchy(H, g, cfg->mg_levels);
    const bool pin = pin_enabled(cfg->pin_mode, bc);
    MGParams P{cfg->mg_pre_smooth, cfg->mg_post_smooth, cfg->mg_coarse_iters, cfg->coarse_solver,
               pin, bc};

    std::memcpy(H.lv[0].K.data(), K, n * sizeof(double));
    for (size_t l = 0; l + 1 < H.lv.size(); ++l)
        coarsen_K(H.lv[l + 1].g, H.lv[l].K.data(), H.lv[l + 1].K.data());
    if (cfg->coarse_solver == 1 &&
        !dense_inverse(H.lv.back().g, bc, pin, H.lv.back().K.data(), H.coarse_inv))
        return 2;

    std::vector<double> rhs(n);
    build_rhs(g, bc, H.lv[0].K.data(), rhs.data());
    if (cfg->init_guess == 0)
        init_guess(g, bc, h);
    else if (cfg->init_guess == 1)
        std::fill(h, h + n, 0.0);
    // init_guess == 2: caller's h is the initial guess
    res->bnorm = nrm2(n, rhs.data());

    std::vector<double> r(n), z(n), p(n), Ap(n);
    auto stop_den = [&](double r0) { return cfg->stop_norm == 1 ? res->bnorm : r0; };

    if (cfg->solver_type == 1) { // MG: standalone V-cycles on level-0 x
        Level& F = H.lv[0];
        std::memcpy(F.x.data(), h, n * sizeof(double));
        std::memcpy(F.b.data(), rhs.data(), n * sizeof(double));
        residual(g, bc, pin, F.x.data(), F.b.data(), F.K.data(), F.r.data());
        res->initial_residual = nrm2(n, F.r.data());
        const double den = stop_den(res->initial_residual);
        for (int c = 0; c < cfg->mg_max_cycles; ++c) {
            res->num_iterations = c + 1;
            v_cycle(H, 0, P);
            residual(g, bc, pin, F.x.data(), F.b.data(), F.K.data(), F.r.data());
            res->final_residual = nrm2(n, F.r.data());
            if (res->final_residual / den < cfg->rtol) {
                res->converged = 1;
                break;
            }
        }
        std::memcpy(h, F.x.data(), n * sizeof(double));
    } else if (cfg->solver_type == 0) { // CG on -A with -b
        double* x = h;
        std::vector<double> bneg(n);
        for (size_t i = 0; i < n; ++i) bneg[i] = -1.0 * rhs[i];
        varcoeff_apply(g, bc, pin, K, x, r.data());
        for (size_t i = 0; i < n; ++i) r[i] = -1.0 * r[i];                  // A_pos x
        for (size_t i = 0; i < n; ++i) r[i] = 1.0 * bneg[i] + -1.0 * r[i]; // axpby
        p = r;
        double rr = dot(n, r.data(), r.data());
        const double r0 = std::sqrt(rr);
        res->initial_residual = r0;
        const double tol = (cfg->stop_norm == 1 ? res->bnorm : r0) * cfg->cg_rtol;
        if (r0 <= tol) {
            res->converged = 1;
            res->final_residual = r0;
        } else {
            int it = 0;
            bool done = false;
            for (; it < cfg->cg_max_iter; ++it) {
                varcoeff_apply(g, bc, pin, K, p.data(), Ap.data());
                for (size_t i = 0; i < n; ++i) Ap[i] = -1.0 * Ap[i];
                const double pAp = dot(n, p.data(), Ap.data());
                if ((it + 1) % cfg->cg_check_every == 0 &&
                    (!std::isfinite(pAp) || std::fabs(pAp) < 1e-30)) {
                    res->num_iterations = it + 1;
                    res->final_residual = -1.0;
                    done = true;
                    break;
                }
                const double alpha = rr / pAp;
                for (size_t i = 0; i < n; ++i) {
                    x[i] = x[i] + alpha * p[i];
                    r[i] = r[i] - alpha * Ap[i];
                }
                const double rr_new = dot(n, r.data(), r.data());
                if ((it + 1) % cfg->cg_check_every == 0 && std::sqrt(rr_new) <= tol) {
                    res->converged = 1;
                    res->num_iterations = it + 1;
                    res->final_residual = std::sqrt(rr_new);
                    done = true;
                    break;
                }
                const double beta = rr_new / rr;
                for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
                rr = rr_new;
            }
            if (!done) {
                res->num_iterations = cfg->cg_max_iter;
                res->final_residual = std::sqrt(rr);
            }
        }
    } else { // PCG_MG
        double* x = h;
        Level& F = H.lv[0];
        auto precond = [&](const double* rin, double* zout) {
            std::memcpy(F.b.data(), rin, n * sizeof(double));
            std::fill(F.x.begin(), F.x.end(), 0.0);
            v_cycle(H, 0, P);
            std::memcpy(zout, F.x.data(), n * sizeof(double));
        };
        varcoeff_apply(g, bc, pin, K, x, Ap.data());
        for (size_t i = 0; i < n; ++i) r[i] = rhs[i];
        for (size_t i = 0; i < n; ++i) r[i] = -1.0 * Ap[i] + r[i];
        res->initial_residual = nrm2(n, r.data());
        const double den = stop_den(res->initial_residual);
        if (res->initial_residual < 1e-15) {
            res->converged = 1;
            res->final_residual = res->initial_residual;
        } else {
            precond(r.data(), z.data());
            p = z;
            double rz_old = dot(n, r.data(), z.data());
            bool done = false;
            for (int it = 0; it < cfg->cg_max_iter; ++it) {
                res->num_iterations = it + 1;
                varcoeff_apply(g, bc, pin, K, p.data(), Ap.data());
                const double pAp = dot(n, p.data(), Ap.data());
                const double alpha = rz_old / pAp;
                for (size_t i = 0; i < n; ++i) x[i] = alpha * p[i] + x[i];
                for (size_t i = 0; i < n; ++i) r[i] = -alpha * Ap[i] + r[i];
                if ((it + 1) % cfg->cg_check_every == 0 || it == cfg->cg_max_iter - 1) {
                    res->final_residual = nrm2(n, r.data());
                    if (res->final_residual / den < cfg->cg_rtol) {
                        res->converged = 1;
                        done = true;
                        break;
                    }
                }
                precond(r.data(), z.data());
                const double rz_new = dot(n, r.data(), z.data());
                const double beta = rz_new / rz_old;
                for (size_t i = 0; i < n; ++i) p[i] = 1.0 * z[i] + beta * p[i];
                rz_old = rz_new;
            }
            if (!done) res->final_residual = nrm2(n, r.data());
        }
    }
    // true residual with the operator convention: ||b - A h||
    varcoeff_apply(g, bc, pin, K, h, Ap.data());
    for (size_t i = 0; i < n; ++i) r[i] = rhs[i] - Ap[i];
    res->true_residual = nrm2(n, r.data());
    return 0;
}

// ---------------------------------------------------------------------------
// Face velocities: compute_velocity_from_head, compact MAC layout
// (src/physics/flow/velocity_from_head.cu:102-381, 601-691) and padded layout
// (velocity_from_head.cu:389-551, 697-777; padded_idx, padded_layout.cuh:30-33).
// Harmonic mean here is 2*Ka*Kb/(Ka+Kb) (velocity_from_head.cu:36-38).
// ---------------------------------------------------------------------------
static inline double hmean_vel(double a, double b) { return 2.0 * a * b / (a + b); }

static void velocity_impl(const orc_grid& g, const orc_bc& bc, const double* H, const double* K,
                          double* U, double* V, double* W, bool padded) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    auto cidx = [&](int i, int j, int k) { return (size_t)i + (size_t)nx * (j + (size_t)ny * k); };
    auto pidx = [&](int i, int j, int k) {
        return (size_t)k * (size_t)(ny + 1) * (size_t)(nx + 1) + (size_t)j * (size_t)(nx + 1) +
               (size_t)i;
    };
    auto Uidx = [&](int i, int j, int k) {
        return padded ? pidx(i, j, k) : (size_t)i + (size_t)(nx + 1) * (j + (size_t)ny * k);
    };
    auto Vidx = [&](int i, int j, int k) {
        return padded ? pidx(i, j, k) : (size_t)i + (size_t)nx * (j + (size_t)(ny + 1) * k);
    };
    auto Widx = [&](int i, int j, int k) {
        return padded ? pidx(i, j, k) : (size_t)i + (size_t)nx * (j + (size_t)ny * k);
    };
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    // interior faces
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (i >= 1) {
                    const size_t l = cidx(i - 1, j, k), r = cidx(i, j, k);
                    U[Uidx(i, j, k)] = -hmean_vel(K[l], K[r]) * (H[r] - H[l]) / dx;
                }
                if (j >= 1) {
                    const size_t s = cidx(i, j - 1, k), n_ = cidx(i, j, k);
                    V[Vidx(i, j, k)] = -hmean_vel(K[s], K[n_]) * (H[n_] - H[s]) / dy;
                }
                if (k >= 1) {
                    const size_t b = cidx(i, j, k - 1), t = cidx(i, j, k);
                    W[Widx(i, j, k)] = -hmean_vel(K[b], K[t]) * (H[t] - H[b]) / dz;
                }
            }
    // boundary faces
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            { // west
                const size_t c = cidx(0, j, k);
                double u = 0.0;
                if (bc.type[0] == BC_DIRICHLET)
                    u = -K[c] * (H[c] - bc.value[0]) / (dx * 0.5);
                else if (bc.type[0] == BC_PERIODIC) {
                    const size_t w = cidx(nx - 1, j, k);
                    u = -hmean_vel(K[w], K[c]) * (H[c] - H[w]) / dx;
                }
                U[Uidx(0, j, k)] = u;
            }
            { // east
                const size_t c = cidx(nx - 1, j, k);
                double u = 0.0;
                if (bc.type[1] == BC_DIRICHLET)
                    u = -K[c] * (bc.value[1] - H[c]) / (dx * 0.5);
                else if (bc.type[1] == BC_PERIODIC) {
                    const size_t w = cidx(0, j, k);
                    u = -hmean_vel(K[c], K[w]) * (H[w] - H[c]) / dx;
                }
                U[Uidx(nx, j, k)] = u;
            }
        }
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i) {
            {
                const size_t c = cidx(i, 0, k);
                double v = 0.0;
                if (bc.type[2] == BC_DIRICHLET)
                    v = -K[c] * (H[c] - bc.value[2]) / (dy * 0.5);
                else if (bc.type[2] == BC_PERIODIC) {
                    const size_t w = cidx(i, ny - 1, k);
                    v = -hmean_vel(K[w], K[c]) * (H[c] - H[w]) / dy;
                }
                V[Vidx(i, 0, k)] = v;
            }
            {
                const size_t c = cidx(i, ny - 1, k);
                double v = 0.0;
                if (bc.type[3] == BC_DIRICHLET)
                    v = -K[c] * (bc.value[3] - H[c]) / (dy * 0.5);
                else if (bc.type[3] == BC_PERIODIC) {
                    const size_t w = cidx(i, 0, k);
                    v = -hmean_vel(K[c], K[w]) * (H[w] - H[c]) / dy;
                }
                V[Vidx(i, ny, k)] = v;
            }
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            {
                const size_t c = cidx(i, j, 0);
                double w = 0.0;
                if (bc.type[4] == BC_DIRICHLET)
                    w = -K[c] * (H[c] - bc.value[4]) / (dz * 0.5);
                else if (bc.type[4] == BC_PERIODIC) {
                    const size_t o = cidx(i, j, nz - 1);
                    w = -hmean_vel(K[o], K[c]) * (H[c] - H[o]) / dz;
                }
                W[Widx(i, j, 0)] = w;
            }
            {
                const size_t c = cidx(i, j, nz - 1);
                double w = 0.0;
                if (bc.type[5] == BC_DIRICHLET)
                    w = -K[c] * (bc.value[5] - H[c]) / (dz * 0.5);
                else if (bc.type[5] == BC_PERIODIC) {
                    const size_t o = cidx(i, j, 0);
                    w = -hmean_vel(K[c], K[o]) * (H[o] - H[c]) / dz;
                }
                W[Widx(i, j, nz)] = w;
            }
        }
}

void orc_velocity_compact(const orc_grid* g, const orc_bc* bc, const double* H, const double* K,
                          double* U, double* V, double* W) {
    velocity_impl(*g, *bc, H, K, U, V, W, false);
}

// U, V, W must be zero-initialised by the caller (PaddedVelocityField::allocate,
// src/physics/common/fields.cuh:268-279).
void orc_velocity_padded(const orc_grid* g, const orc_bc* bc, const double* H, const double* K,
                         double* U, double* V, double* W) {
    velocity_impl(*g, *bc, H, K, U, V, W, true);
}

// Cell-centred diagnostics of a padded velocity field, compute_velocity_diagnostics
// (src/physics/flow/velocity_diagnostics.cu:46-141 kernel, 147-177 host):
//   div = (U(i+1)-U(i))/dx + (V(j+1)-V(j))/dy + (W(k+1)-W(k))/dz at every cell;
//   at interior cells (1 <= i <= nx-2, same in j and k) the face averages
//   uc = (U(i)+U(i+1))/2 (and vc, wc) are differenced centrally,
//   w = (dwc/dy - dvc/dz, duc/dz - dwc/dx, dvc/dx - duc/dy), and
//   |w| and the helicity uc*wx + vc*wy + wc*wz are stored; elsewhere, and
//   everywhere when a dimension is below 3 (:157), both are 0.
void orc_velocity_diagnostics(const orc_grid* gp, const double* U, const double* V,
                              const double* W, double* div, double* omag, double* hel) {
    const orc_grid& g = *gp;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool curl = nx >= 3 && ny >= 3 && nz >= 3;
    const double ix = 1.0 / g.dx, iy = 1.0 / g.dy, iz = 1.0 / g.dz;
    const double hx = 0.5 * ix, hy = 0.5 * iy, hz = 0.5 * iz;
    auto P = [&](int i, int j, int k) {
        return ((size_t)k * (size_t)(ny + 1) + (size_t)j) * (size_t)(nx + 1) + (size_t)i;
    };
    // face averages at the centre of cell (i, j, k)
    auto uc = [&](int i, int j, int k) { return 0.5 * (U[P(i, j, k)] + U[P(i + 1, j, k)]); };
    auto vc = [&](int i, int j, int k) { return 0.5 * (V[P(i, j, k)] + V[P(i, j + 1, k)]); };
    auto wc = [&](int i, int j, int k) { return 0.5 * (W[P(i, j, k)] + W[P(i, j, k + 1)]); };
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                div[c] = (U[P(i + 1, j, k)] - U[P(i, j, k)]) * ix +
                         (V[P(i, j + 1, k)] - V[P(i, j, k)]) * iy +
                         (W[P(i, j, k + 1)] - W[P(i, j, k)]) * iz;
                const bool inner = curl && i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 &&
                                   k >= 1 && k <= nz - 2;
                if (!inner) {
                    omag[c] = 0.0;
                    hel[c] = 0.0;
                    continue;
                }
                const double wx = (wc(i, j + 1, k) - wc(i, j - 1, k)) * hy -
                                  (vc(i, j, k + 1) - vc(i, j, k - 1)) * hz;
                const double wy = (uc(i, j, k + 1) - uc(i, j, k - 1)) * hz -
                                  (wc(i + 1, j, k) - wc(i - 1, j, k)) * hx;
                const double wz = (vc(i + 1, j, k) - vc(i - 1, j, k)) * hx -
                                  (uc(i, j + 1, k) - uc(i, j - 1, k)) * hy;
                omag[c] = std::sqrt(wx * wx + wy * wy + wz * wz);
                hel[c] = uc(i, j, k) * wx + vc(i, j, k) * wy + wc(i, j, k) * wz;
            }
}

// compute_field_stats (src/physics/stochastic/stochastic.cu:369-414):
// out = {min, max, mean}; all zero for n == 0. The sum runs in index order.
void orc_field_stats(size_t n, const double* x, double* out) {
    if (n == 0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    double lo = x[0], hi = x[0], sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        lo = std::min(lo, x[i]);
        hi = std::max(hi, x[i]);
        sum += x[i];
    }
    out[0] = lo;
    out[1] = hi;
    out[2] = sum / (double)n;
}

// ---------------------------------------------------------------------------
// Log-normal conductivity by FFT circulant embedding (Dietrich & Newsam 1997).
// Not a reference function: the reference draws K with a randomized spectral
// method (src/physics/stochastic/stochastic.cu:56-197). The workload asks for
// an exponential covariance C(r) = sigma2 * exp(-r / corr_len), r in cell units.
//
// Embedding extent per axis m = next power of two >= 2 n (doubled until the
// circulant's eigenvalues are non-negative, at most three times; remaining
// negative eigenvalues are clipped to zero and their mass reported).
// Draws: counter-based splitmix64, Box-Muller; xi_k = (n1, n2) for embedding
// index k (x fastest). Z = FFT(sqrt(lambda/M) * xi); lnK = Re Z on the grid.
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static const uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

static void gauss_pair(uint64_t key, uint64_t k, double& n1, double& n2) {
    const uint64_t a = mix64(key + (2 * k + 1) * kGolden);
    const uint64_t b = mix64(key + (2 * k + 2) * kGolden);
    const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(b >> 11) * 0x1.0p-53;
    const double rad = std::sqrt(-2.0 * std::log(u1));
    const double th = 6.283185307179586 * u2;
    n1 = rad * std::cos(th);
    n2 = rad * std::sin(th);
}

typedef std::complex<double> cplx;

static void fft1d(cplx* a, int n) { // in-place radix-2, forward (e^{-i})
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < len / 2; ++j) {
                const cplx w(std::cos(ang * j), std::sin(ang * j));
                const cplx u = a[i + j], v = a[i + j + len / 2] * w;
                a[i + j] = u + v;
                a[i + j + len / 2] = u - v;
            }
    }
}

static void fft3d(std::vector<cplx>& a, int mx, int my, int mz) {
    const size_t sx = 1, sy = mx, sz = (size_t)mx * my;
    const int dims[3] = {mx, my, mz};
    const size_t strides[3] = {sx, sy, sz};
    for (int ax = 0; ax < 3; ++ax) {
        const int n = dims[ax];
        const size_t s = strides[ax];
        const int o1 = dims[(ax + 1) % 3], o2 = dims[(ax + 2) % 3];
        const size_t s1 = strides[(ax + 1) % 3], s2 = strides[(ax + 2) % 3];
#pragma omp parallel
        {
            std::vector<cplx> line(n);
#pragma omp for schedule(static)
            for (long q = 0; q < (long)o1 * o2; ++q) {
                const size_t base = (size_t)(q % o1) * s1 + (size_t)(q / o1) * s2;
                for (int t = 0; t < n; ++t) line[t] = a[base + t * s];
                fft1d(line.data(), n);
                for (int t = 0; t < n; ++t) a[base + t * s] = line[t];
            }
        }
    }
}

static int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// stats[0..2] = embedding extents, stats[3] = min eigenvalue / max eigenvalue,
// stats[4] = clipped negative mass / total mass.
int orc_lognormal_lnK(int nx, int ny, int nz, double sigma2, double corr_len, uint64_t seed,
                      double* lnK, double* stats) {
    int mx = next_pow2(2 * nx), my = next_pow2(2 * ny), mz = next_pow2(2 * nz);
    std::vector<cplx> lam;
    double min_ratio = 0.0, clipped = 0.0;
    for (int attempt = 0; attempt < 4; ++attempt) {
        lam.assign((size_t)mx * my * mz, cplx(0, 0));
        for (int c = 0; c < mz; ++c)
            for (int b = 0; b < my; ++b)
                for (int a = 0; a < mx; ++a) {
                    const double da = std::min(a, mx - a), db = std::min(b, my - b),
                                 dc = std::min(c, mz - c);
                    const double r = std::sqrt(da * da + db * db + dc * dc);
                    lam[(size_t)a + (size_t)mx * (b + (size_t)my * c)] =
                        cplx(sigma2 * std::exp(-r / corr_len), 0.0);
                }
        fft3d(lam, mx, my, mz);
        double mn = 1e300, mxv = -1e300, neg = 0.0, tot = 0.0;
        for (auto& v : lam) {
            mn = std::min(mn, v.real());
            mxv = std::max(mxv, v.real());
            if (v.real() < 0) neg -= v.real();
            tot += std::fabs(v.real());
        }
        min_ratio = mn / mxv;
        clipped = neg / tot;
        if (mn >= -1e-12 * mxv || attempt == 3) break;
        mx *= 2;
        my *= 2;
        mz *= 2;
    }
    const size_t M = (size_t)mx * my * mz;
    const uint64_t key = mix64(seed + kGolden);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)M; ++q) {
        double n1, n2;
        gauss_pair(key, (uint64_t)q, n1, n2);
        const double s = std::sqrt(std::max(lam[q].real(), 0.0) / (double)M);
        lam[q] = cplx(s * n1, s * n2);
    }
    fft3d(lam, mx, my, mz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                lnK[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] =
                    lam[(size_t)i + (size_t)mx * (j + (size_t)my * k)].real();
    if (stats) {
        stats[0] = mx;
        stats[1] = my;
        stats[2] = mz;
        stats[3] = min_ratio;
        stats[4] = clipped;
    }
    return 0;
}

// Standard normal pair for counter k (exposed so the product generator's draws
// can be compared one by one).
void orc_gauss_pair(uint64_t seed, uint64_t k, double* out2) {
    gauss_pair(mix64(seed + kGolden), k, out2[0], out2[1]);
}

} // extern "C"
