// Host-side launchers of the gfx950 kernels (all enqueue on the given stream).
#pragma once

#include "common.hpp"

namespace mf3d {
namespace kern {

// ---- stencil (stencil.hip) -------------------------------------------------
int gsrb(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double* x,
         const double* b, const double* K, int iters, bool zero_init);
int residual(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, const double* x,
             const double* b, const double* K, double* r);
int residual_restrict(hipStream_t s, const mf3d_grid& fine, const BcTypes& bc, bool pin,
                      const double* x, const double* b, const double* K, double* b_coarse);
// first-stage block count of the fused reductions: one partial per kReduceCells
// consecutive cells (8 per thread of a 256-thread block)
unsigned reduce_blocks(size_t n);
int apply(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, const double* K,
          const double* x, double* y);
int restrict_avg(hipStream_t s, const mf3d_grid& coarse, const double* f, double* c);
int prolong_add(hipStream_t s, const mf3d_grid& fine, const double* c, double* f);
int coarsen(hipStream_t s, const mf3d_grid& coarse, const double* Kf, double* Kc);
int rhs(hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc, const double* K, double* out);
int guess(hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc, int mode, double* h);
int dense_gemv(hipStream_t s, int n, const double* Ainv, const double* b, double* x);
// banded Cholesky factors of m coarsest operators, assembled from K: one
// workgroup per factor (wmax = max band width)
int band_factor_batch(hipStream_t s, const FactorDesc* D, int m, int wmax);
// A^{-1} = -L^{-T} L^{-1} from the band factor (Y: n * n scratch for L^{-1})
int band_inverse(hipStream_t s, int n, int w, const double* Lb, double* Y, double* Ainv);
// band_inverse for m factors in one launch (nblocks = sum of n, nmax = max n,
// wmax = max bandwidth)
int band_inverse_batch(hipStream_t s, const BandDesc* D, const int* start, int m, int nblocks,
                       int nmax, int wmax);

// ---- BLAS-1 and reductions (blas.hip) ---------------------------------------
int fill(hipStream_t s, double* x, size_t n, double v);
int copy(hipStream_t s, const double* x, double* y, size_t n);
int axpy(hipStream_t s, double a, const double* x, double* y, size_t n);
int axpby(hipStream_t s, double a, const double* x, double b, double* y, size_t n);
int scal(hipStream_t s, double* x, size_t n, double a);
// first stage of (x, y) into partials[reduce_blocks(n)]
int dot_partials(hipStream_t s, const double* x, const double* y, size_t n, double* partials);
// second stage: sum partials[0..np) in a fixed order and finalise into st / out
int finalize(hipStream_t s, const double* partials, unsigned np, Finalize op, IterState* st,
             double* out);
// finalize(p1, n1, op1) then finalize(p2, n2, op2) in one launch (state ops)
int finalize2(hipStream_t s, const double* p1, unsigned n1, Finalize op1, const double* p2,
              unsigned n2, Finalize op2, IterState* st);
// PCG: if !st->stop { x += alpha p; r -= alpha Ap } with partials of (r, r)
int pcg_update_xr(hipStream_t s, const IterState* st, const double* p, const double* Ap,
                  double* x, double* r, size_t n, double* partials);
// PCG: p = z + beta p
// r -= alpha Ap with (r, r) partials (x is updated later, with p); the x +=
// alpha p still owed after the last such update (IterState::xpend)
int pcg_update_r(hipStream_t s, const IterState* st, const double* Ap, double* r, size_t n,
                 double* partials);
int pcg_x_pending(hipStream_t s, const IterState* st, const double* p, double* x, size_t n);
int pcg_update_p(hipStream_t s, const IterState* st, const double* z, double* p, size_t n);
int pcg_update_xp(hipStream_t s, const IterState* st, const double* z, double* x, double* p,
                  size_t n);
// CG on -A: if !st->stop { x += alpha p; r -= alpha (-Ap) } with partials of (r, r)
int cg_update_xr(hipStream_t s, const IterState* st, const double* p, const double* Ap,
                 double* x, double* r, size_t n, double* partials);
// CG: p = r + beta p (beta from st)
int cg_update_p(hipStream_t s, const IterState* st, const double* r, double* p, size_t n);
// y = a * x (used to negate vectors for the CG path)
int scale_copy(hipStream_t s, double a, const double* x, double* y, size_t n);

// ---- batched Krylov steps (blas.hip; BatchVec descriptors, nblocks = sum of
// reduce_blocks(n) over the realisations, one partial per block) -------------
// second stage per realisation (one block each); guard 0: all, 1: only
// running realisations (stop == 0), 2: only converged ones (stop == 1)
int batch_finalize(hipStream_t s, const BatchVec* V, int m, Finalize op, int guard,
                   bool op_count = false);
// mode 0: (r, z) partials (guarded), mode 1: (rhs, rhs) partials
int batch_dot(hipStream_t s, int mode, const BatchVec* V, const int* start, int m, long nblocks);
// x += alpha p, r -= alpha Ap with (r, r) partials (guarded)
// r -= alpha Ap with (r, r) partials; the owed x += alpha p of stopped
// realisations (IterState::xpend, cleared afterwards)
int batch_update_r(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks);
int batch_x_pending(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks);
// mode 0: p = z + beta p, mode 1: p = z (guarded)

// ---- velocity (velocity.hip) -----------------------------------------------
int velocity(hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc, const double* h,
             const double* K, double* U, double* V, double* W, bool padded);
// divergence, |curl| and helicity at cell centres from padded U, V, W; with
// stats (device, 9 values: min, max, mean of each field) the partials buffer
// holds diag_partials(g) values
size_t diag_partials(const mf3d_grid& g);
int velocity_diagnostics(hipStream_t s, const mf3d_grid& g, const double* U, const double* V,
                         const double* W, double* div, double* omag, double* hel,
                         double* partials, double* stats);
// out (device, 3 values) = min, max, mean of x[0..n), n > 0; partials holds
// stats_partials(n) values
size_t stats_partials(size_t n);
int field_stats(hipStream_t s, const double* x, size_t n, double* partials, double* out);

}  // namespace kern
}  // namespace mf3d
