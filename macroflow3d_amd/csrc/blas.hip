// BLAS-1 kernels, deterministic two-stage reductions and the fused Krylov
// vector updates (FP64), gfx950.
//
// Reductions: every first stage writes one partial per 256-thread block (wave64
// shuffle tree, then LDS across the 4 waves); the second stage is a single
// 1024-thread block summing the partials in a fixed order, so a result depends
// only on the data and n. The second stage also finalises the Krylov scalars
// (alpha, beta, convergence flag) on the device, so an iteration needs no host
// round trip (the reference syncs the stream 2-3 times per PCG iteration,
// src/numerics/blas/dot.cu:38-50, nrm2.cu:38-49).

#include "kernels.hpp"
#include "reduce.hpp"

namespace mf3d {
namespace {

constexpr int kBlock = kReduceBlock;

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }


// blas::fill (src/numerics/blas/fill.cu:7-14)
__global__ void __launch_bounds__(kBlock) fill_kernel(double* __restrict__ x, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
// blas::copy (copy.cu:7-14)
__global__ void __launch_bounds__(kBlock)
    copy_kernel(const double* __restrict__ x, double* __restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}
// blas::axpy (axpy.cu:7-14): y = a x + y
__global__ void __launch_bounds__(kBlock)
    axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + y[i];
}
// blas::axpby (axpby.cu:7-14): y = a x + b y
__global__ void __launch_bounds__(kBlock)
    axpby_kernel(double a, const double* __restrict__ x, double b, double* __restrict__ y,
                 size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
// blas::scal (scal.cu:7-14): x = a x
__global__ void __launch_bounds__(kBlock) scal_kernel(double* __restrict__ x, size_t n, double a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = a * x[i];
}
__global__ void __launch_bounds__(kBlock)
    scale_copy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

// First-stage reductions: block blk owns cells [blk * kReduceCells, + kReduceCells),
// thread t its cells blk * kReduceCells + u * kBlock + t (u = 0..7, summed in
// that order), then the block tree; partials[reduce_blocks(n)].
// All 8 slots are loaded before any is used (a slot past n reads the last
// cell and is masked), so a thread keeps 8 loads per array in flight.
constexpr int kPer = kReduceCells / kBlock;
__device__ __forceinline__ long red_cell(long blk, int u) {
    return blk * kReduceCells + (long)u * kBlock + threadIdx.x;
}
__device__ __forceinline__ long red_load(long blk, int u, long n) {
    return min(red_cell(blk, u), n - 1);
}

__global__ void __launch_bounds__(kBlock)
    dot_partials_kernel(const double* __restrict__ x, const double* __restrict__ y, size_t n,
                        double* __restrict__ partials) {
    double xv[kPer], yv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_load(blockIdx.x, u, (long)n);
        xv[u] = x[i];
        yv[u] = y[i];
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u)
        if (red_cell(blockIdx.x, u) < (long)n) v += xv[u] * yv[u];
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kFinalBlock)
    finalize_kernel(const double* __restrict__ partials, unsigned np, int op,
                    IterState* __restrict__ st, double* __restrict__ out) {
    finalize_body(partials, np, op, st, out);
}

// two second stages in one launch: op1 on the sum of p1, then op2 on that of p2
__global__ void __launch_bounds__(kFinalBlock)
    finalize2_kernel(const double* __restrict__ p1, unsigned n1, int op1,
                     const double* __restrict__ p2, unsigned n2, int op2,
                     IterState* __restrict__ st) {
    const double S1 = finalize_sum(p1, n1);
    __syncthreads();  // the block tree's LDS is reused
    const double S2 = finalize_sum(p2, n2);
    if (threadIdx.x == 0) {
        finalize_apply(op1, S1, st, nullptr);
        finalize_apply(op2, S2, st, nullptr);
    }
}

// ---- batched (one launch for many realisations, descriptors BatchVec) --------
// second stage: one block per realisation; guard 0: always, 1: only running
// realisations (stop == 0), 2: only converged ones (stop == 1); op_count:
// the partials come from the operator (np_op of them)
__global__ void __launch_bounds__(kFinalBlock)
    finalize_batch_kernel(const BatchVec* __restrict__ Vs, int op, int guard, int op_count) {
    const BatchVec& V = Vs[blockIdx.x];
    const int stop = V.st->stop;
    if ((guard == 1 && stop != 0) || (guard == 2 && stop != 1)) return;
    finalize_body(V.partials, op_count ? V.np_op : V.np, op, V.st, V.out);
}

// MODE 0: (r, z), skipped once stopped; MODE 1: (rhs, rhs)
template <int MODE>
__global__ void __launch_bounds__(kBlock)
    dot_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m) {
    const int e = batch_entry(start, m, blockIdx.x);
    const BatchVec& V = Vs[e];
    if (MODE == 0 && V.st->stop) return;
    const long blk = (long)blockIdx.x - start[e];
    const long n = (long)V.nx * V.ny * V.nz;
    const double* a = MODE == 0 ? V.r : V.rhs;
    const double* b = MODE == 0 ? V.z : V.rhs;
    double av[kPer], bv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_load(blk, u, n);
        av[u] = a[i];
        bv[u] = b[i];
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u)
        if (red_cell(blk, u) < n) v += av[u] * bv[u];
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) V.partials[blk] = s;
}

// r -= alpha Ap with (r, r) partials for the running realisations (x is
// updated with p, xp_split8_batch_kernel)
__global__ void __launch_bounds__(kBlock)
    r_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m) {
    const int e = batch_entry(start, m, blockIdx.x);
    const BatchVec& V = Vs[e];
    if (V.st->stop) return;
    const double alpha = V.st->alpha;
    const long blk = (long)blockIdx.x - start[e];
    const long n = (long)V.nx * V.ny * V.nz;
    double av[kPer], rv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_load(blk, u, n);
        av[u] = V.Ap[i];
        rv[u] = V.r[i];
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_cell(blk, u);
        if (i < n) {
            const double ri = -alpha * av[u] + rv[u];
            V.r[i] = ri;
            v += ri * ri;
        }
    }
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) V.partials[blk] = s;
}

// the x += alpha p still owed by realisations that stopped after their r update
__global__ void __launch_bounds__(kBlock)
    x_pending_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m) {
    const int e = batch_entry(start, m, blockIdx.x);
    const BatchVec& V = Vs[e];
    if (!V.st->xpend) return;
    const double alpha = V.st->alpha;
    const long blk = (long)blockIdx.x - start[e];
    const long n = (long)V.nx * V.ny * V.nz;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_cell(blk, u);
        if (i < n) V.x[i] = alpha * V.p[i] + V.x[i];
    }
}
__global__ void xpend_clear_kernel(const BatchVec* __restrict__ Vs, int m) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) Vs[e].st->xpend = 0;
}

// PCG x/r update (pcg.cuh:149-153 as two axpy) fused with (r, r).
__global__ void __launch_bounds__(kBlock)
    pcg_update_xr_kernel(const IterState* __restrict__ st, const double* __restrict__ p,
                         const double* __restrict__ Ap, double* __restrict__ x,
                         double* __restrict__ r, size_t n, double* __restrict__ partials) {
    if (st->stop) return;  // uniform across the grid
    const double alpha = st->alpha;
    double pv[kPer], av[kPer], xv[kPer], rv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_load(blockIdx.x, u, (long)n);
        pv[u] = p[i];
        av[u] = Ap[i];
        xv[u] = x[i];
        rv[u] = r[i];
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_cell(blockIdx.x, u);
        if (i < (long)n) {
            x[i] = alpha * pv[u] + xv[u];
            const double ri = -alpha * av[u] + rv[u];
            r[i] = ri;
            v += ri * ri;
        }
    }
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// r -= alpha Ap with (r, r) partials: the r half of pcg_update_xr_kernel, for
// the iteration that updates x together with p (x += alpha p; p = z + beta p).
__global__ void __launch_bounds__(kBlock)
    pcg_update_r_kernel(const IterState* __restrict__ st, const double* __restrict__ Ap,
                        double* __restrict__ r, size_t n, double* __restrict__ partials) {
    if (st->stop) return;  // uniform across the grid
    const double alpha = st->alpha;
    double av[kPer], rv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_load(blockIdx.x, u, (long)n);
        av[u] = Ap[i];
        rv[u] = r[i];
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_cell(blockIdx.x, u);
        if (i < (long)n) {
            const double ri = -alpha * av[u] + rv[u];
            r[i] = ri;
            v += ri * ri;
        }
    }
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// the x += alpha p still owed when the iteration stopped after its r update
__global__ void __launch_bounds__(kBlock)
    pcg_x_pending_kernel(const IterState* __restrict__ st, const double* __restrict__ p,
                         double* __restrict__ x, size_t n) {
    if (!st->xpend) return;
    const double alpha = st->alpha;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = alpha * p[i] + x[i];
}

// PCG p = z + beta p (pcg.cuh:179, axpby(1, z, beta, p)).
__global__ void __launch_bounds__(kBlock)
    pcg_update_p_kernel(const IterState* __restrict__ st, const double* __restrict__ z,
                        double* __restrict__ p, size_t n) {
    if (st->stop) return;
    const double beta = st->beta;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1.0 * z[i] + beta * p[i];
}

// x += alpha p and p = z + beta p in one pass (the x half of
// pcg_update_xr_kernel and pcg_update_p_kernel), for the iteration whose r
// update rode on the preconditioner's first half-sweep; skipped once stopped
// (pcg_x_pending_kernel then applies the last x step).
__global__ void __launch_bounds__(kBlock)
    pcg_update_xp_kernel(const IterState* __restrict__ st, const double* __restrict__ z,
                         double* __restrict__ x, double* __restrict__ p, size_t n) {
    if (st->stop) return;  // uniform across the grid
    const double alpha = st->alpha, beta = st->beta;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        // x and p are touched once here: non-temporal, to keep the
        // V-cycle's arrays in the Infinity Cache
        const double pv = __builtin_nontemporal_load(p + i);
        __builtin_nontemporal_store(alpha * pv + __builtin_nontemporal_load(x + i), x + i);
        __builtin_nontemporal_store(1.0 * z[i] + beta * pv, p + i);
    }
}

// CG on the negated operator (update_x_and_r_kernel, src/numerics/blas/cg_kernels.cu:19-31):
// Ap holds A p, the CG operator is -A, so r - alpha (-Ap) = r + alpha Ap.
__global__ void __launch_bounds__(kBlock)
    cg_update_xr_kernel(const IterState* __restrict__ st, const double* __restrict__ p,
                        const double* __restrict__ Ap, double* __restrict__ x,
                        double* __restrict__ r, size_t n, double* __restrict__ partials) {
    if (st->stop) return;
    const double alpha = st->alpha;
    double pv[kPer], av[kPer], xv[kPer], rv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_load(blockIdx.x, u, (long)n);
        pv[u] = p[i];
        av[u] = Ap[i];
        xv[u] = x[i];
        rv[u] = r[i];
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const long i = red_cell(blockIdx.x, u);
        if (i < (long)n) {
            x[i] = xv[u] + alpha * pv[u];
            const double ri = rv[u] - alpha * (-av[u]);
            r[i] = ri;
            v += ri * ri;
        }
    }
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// update_p_kernel (cg_kernels.cu:33-44): p = r + beta p
__global__ void __launch_bounds__(kBlock)
    cg_update_p_kernel(const IterState* __restrict__ st, const double* __restrict__ r,
                       double* __restrict__ p, size_t n) {
    if (st->stop) return;
    const double beta = st->beta;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = r[i] + beta * p[i];
}

}  // namespace

namespace kern {

#define MF3D_LAUNCH_CHECK() MF3D_HIP_TRY(hipGetLastError())

int fill(hipStream_t s, double* x, size_t n, double v) {
    if (n == 0) return MF3D_OK;
    fill_kernel<<<grid_for(n), kBlock, 0, s>>>(x, n, v);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int copy(hipStream_t s, const double* x, double* y, size_t n) {
    if (n == 0) return MF3D_OK;
    copy_kernel<<<grid_for(n), kBlock, 0, s>>>(x, y, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int axpy(hipStream_t s, double a, const double* x, double* y, size_t n) {
    if (n == 0) return MF3D_OK;
    axpy_kernel<<<grid_for(n), kBlock, 0, s>>>(a, x, y, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int axpby(hipStream_t s, double a, const double* x, double b, double* y, size_t n) {
    if (n == 0) return MF3D_OK;
    axpby_kernel<<<grid_for(n), kBlock, 0, s>>>(a, x, b, y, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int scal(hipStream_t s, double* x, size_t n, double a) {
    if (n == 0) return MF3D_OK;
    scal_kernel<<<grid_for(n), kBlock, 0, s>>>(x, n, a);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int scale_copy(hipStream_t s, double a, const double* x, double* y, size_t n) {
    if (n == 0) return MF3D_OK;
    scale_copy_kernel<<<grid_for(n), kBlock, 0, s>>>(a, x, y, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int dot_partials(hipStream_t s, const double* x, const double* y, size_t n, double* partials) {
    dot_partials_kernel<<<reduce_blocks(n), kBlock, 0, s>>>(x, y, n, partials);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int finalize(hipStream_t s, const double* partials, unsigned np, Finalize op, IterState* st,
             double* out) {
    finalize_kernel<<<1, kFinalBlock, 0, s>>>(partials, np, (int)op, st, out);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int finalize2(hipStream_t s, const double* p1, unsigned n1, Finalize op1, const double* p2,
              unsigned n2, Finalize op2, IterState* st) {
    finalize2_kernel<<<1, kFinalBlock, 0, s>>>(p1, n1, (int)op1, p2, n2, (int)op2, st);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int pcg_update_xr(hipStream_t s, const IterState* st, const double* p, const double* Ap,
                  double* x, double* r, size_t n, double* partials) {
    pcg_update_xr_kernel<<<reduce_blocks(n), kBlock, 0, s>>>(st, p, Ap, x, r, n, partials);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int pcg_update_r(hipStream_t s, const IterState* st, const double* Ap, double* r, size_t n,
                 double* partials) {
    pcg_update_r_kernel<<<reduce_blocks(n), kBlock, 0, s>>>(st, Ap, r, n, partials);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int pcg_x_pending(hipStream_t s, const IterState* st, const double* p, double* x, size_t n) {
    pcg_x_pending_kernel<<<grid_for(n), kBlock, 0, s>>>(st, p, x, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int pcg_update_xp(hipStream_t s, const IterState* st, const double* z, double* x, double* p,
                  size_t n) {
    pcg_update_xp_kernel<<<grid_for(n), kBlock, 0, s>>>(st, z, x, p, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int pcg_update_p(hipStream_t s, const IterState* st, const double* z, double* p, size_t n) {
    pcg_update_p_kernel<<<grid_for(n), kBlock, 0, s>>>(st, z, p, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int cg_update_xr(hipStream_t s, const IterState* st, const double* p, const double* Ap,
                 double* x, double* r, size_t n, double* partials) {
    cg_update_xr_kernel<<<reduce_blocks(n), kBlock, 0, s>>>(st, p, Ap, x, r, n, partials);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int cg_update_p(hipStream_t s, const IterState* st, const double* r, double* p, size_t n) {
    cg_update_p_kernel<<<grid_for(n), kBlock, 0, s>>>(st, r, p, n);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}

int batch_finalize(hipStream_t s, const BatchVec* V, int m, Finalize op, int guard,
                   bool op_count) {
    finalize_batch_kernel<<<m, kFinalBlock, 0, s>>>(V, (int)op, guard, op_count ? 1 : 0);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int batch_dot(hipStream_t s, int mode, const BatchVec* V, const int* start, int m, long nblocks) {
    if (mode == 0) dot_batch_kernel<0><<<(unsigned)nblocks, kBlock, 0, s>>>(V, start, m);
    else dot_batch_kernel<1><<<(unsigned)nblocks, kBlock, 0, s>>>(V, start, m);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int batch_update_r(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks) {
    r_batch_kernel<<<(unsigned)nblocks, kBlock, 0, s>>>(V, start, m);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
int batch_x_pending(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks) {
    x_pending_batch_kernel<<<(unsigned)nblocks, kBlock, 0, s>>>(V, start, m);
    MF3D_LAUNCH_CHECK();
    xpend_clear_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(V, m);
    MF3D_LAUNCH_CHECK();
    return MF3D_OK;
}
}  // namespace kern
}  // namespace mf3d
