// Cell-centred 7-point variable-K stencil kernels (FP64), gfx950.
//
// One thread per cell (or per own-colour cell for the red-black smoother),
// all boundary types handled in the same launch: the reference spends 27
// launches per colour on interior/face/edge/vertex kernels
// (src/multigrid/smoothers/gsrb_3d.cu:392-575); here a colour is one launch.
// These kernels serve every level and every boundary mix of the
// reference-semantics entry points (mf3d_gsrb_smooth_3d, ...).

#include "kernels.hpp"

namespace mf3d {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block = kBlock) {
    return (unsigned)((n + block - 1) / block);
}

// GSRB value of one cell: (sum K_f x_nb - b dx^2) / sum K_f.
// Interior: gsrb_interior_kernel (src/multigrid/smoothers/gsrb_3d.cu:29-73).
// Boundary: gsrb_{face,edge,vertex}_kernel (gsrb_3d.cu:76-366), pin doubles
// the diagonal of cell 0 (gsrb_3d.cu:360-362).
template <bool ZERO>
__device__ inline double gsrb_value(int nx, int ny, int nz, const BcTypes& bc, bool pin,
                                    const double* __restrict__ x, const double* __restrict__ b,
                                    const double* __restrict__ K, int i, int j, int k, long idx,
                                    double dx2) {
    const double KC = K[idx];
    if (interior_cell(nx, ny, nz, i, j, k)) {
        const long Nx = nx, st = (long)nx * ny;
        const double iKC = 1.0 / KC;
        const double K_xp = 2.0 / (iKC + 1.0 / K[idx + 1]);
        const double K_xm = 2.0 / (iKC + 1.0 / K[idx - 1]);
        const double K_yp = 2.0 / (iKC + 1.0 / K[idx + Nx]);
        const double K_ym = 2.0 / (iKC + 1.0 / K[idx - Nx]);
        const double K_zp = 2.0 / (iKC + 1.0 / K[idx + st]);
        const double K_zm = 2.0 / (iKC + 1.0 / K[idx - st]);
        double result = 0.0;
        if (!ZERO) {
            result = x[idx + 1] * K_xp;
            result += x[idx - 1] * K_xm;
            result += x[idx + Nx] * K_yp;
            result += x[idx - Nx] * K_ym;
            result += x[idx + st] * K_zp;
            result += x[idx - st] * K_zm;
        }
        const double aC = K_xp + K_xm + K_yp + K_ym + K_zp + K_zm;
        return (result - b[idx] * dx2) / aC;
    }
    double c[6], v[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const long n = neighbour_coef(nx, ny, nz, bc, K, i, j, k, d, KC, c[d]);
        v[d] = (ZERO || n < 0) ? 0.0 : x[n];
    }
    const double result =
        c[0] * v[0] + c[1] * v[1] + c[2] * v[2] + c[3] * v[3] + c[4] * v[4] + c[5] * v[5];
    double aC = c[0] + c[1] + c[2] + c[3] + c[4] + c[5];
    if (pin && idx == 0) aC *= 2.0;
    return (result - b[idx] * dx2) / aC;
}

// One colour of one GSRB iteration. Thread t covers the pair of cells
// (2*ii, 2*ii+1) of row (j,k) and updates the one of colour `colour`
// ((i+j+k) even = red = 0). ZERO: x is taken as zero on entry, the partner
// (other colour) cell is written 0 -- a fill(0) fused into the first sweep.
template <bool ZERO>
__global__ void __launch_bounds__(kBlock)
    gsrb_colour_kernel(int nx, int ny, int nz, double dx2, BcTypes bc, int pin,
                       double* __restrict__ x, const double* __restrict__ b,
                       const double* __restrict__ K, int colour) {
    const int nxh = (nx + 1) >> 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)nxh * ny * nz;
    if (t >= total) return;
    const int ii = (int)(t % nxh);
    const long row = t / nxh;
    const int j = (int)(row % ny);
    const int k = (int)(row / ny);
    const int par = (j + k + colour) & 1;
    const int i = 2 * ii + par;
    const long rowbase = (long)nx * row;
    if (ZERO) {
        const int ip = 2 * ii + (par ^ 1);
        if (ip < nx) x[rowbase + ip] = 0.0;
    }
    if (i >= nx) return;
    const long idx = rowbase + i;
    x[idx] = gsrb_value<ZERO>(nx, ny, nz, bc, pin != 0, x, b, K, i, j, k, idx, dx2);
}

// Residual value r = b - A x of one cell.
// Interior: residual_interior_kernel (src/multigrid/smoothers/residual_3d.cu:30-80).
// Boundary: residual_{face,edge,vertex}_kernel (residual_3d.cu:83-376), pin at
// cell 0 adds aC*x_C (residual_3d.cu:365-368).
__device__ inline double residual_value(int nx, int ny, int nz, const BcTypes& bc, bool pin,
                                        const double* __restrict__ x,
                                        const double* __restrict__ b,
                                        const double* __restrict__ K, int i, int j, int k,
                                        long o, double dx2, double inv_dx2) {
    const double xc = x[o], Kc = K[o];
    if (interior_cell(nx, ny, nz, i, j, k)) {
        const long Nx = nx, st = (long)nx * ny;
        const double iKc = 1.0 / Kc;
        double Ax = 0.0;
        Ax -= 2.0 * (xc - x[o + 1]) / (iKc + 1.0 / K[o + 1]);
        Ax -= 2.0 * (xc - x[o + Nx]) / (iKc + 1.0 / K[o + Nx]);
        Ax -= 2.0 * (xc - x[o - 1]) / (iKc + 1.0 / K[o - 1]);
        Ax -= 2.0 * (xc - x[o - Nx]) / (iKc + 1.0 / K[o - Nx]);
        Ax -= 2.0 * (xc - x[o + st]) / (iKc + 1.0 / K[o + st]);
        Ax -= 2.0 * (xc - x[o - st]) / (iKc + 1.0 / K[o - st]);
        return b[o] - Ax * inv_dx2;
    }
    double c[6], v[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const long n = neighbour_coef(nx, ny, nz, bc, K, i, j, k, d, Kc, c[d]);
        v[d] = n < 0 ? (bc.t[d] == MF3D_BC_NEUMANN ? xc : 0.0) : x[n];
    }
    double result = c[0] * (xc - v[0]) + c[1] * (xc - v[1]) + c[2] * (xc - v[2]) +
                    c[3] * (xc - v[3]) + c[4] * (xc - v[4]) + c[5] * (xc - v[5]);
    if (pin && o == 0) result += (c[0] + c[1] + c[2] + c[3] + c[4] + c[5]) * xc;
    return b[o] - (-result / dx2);
}

__global__ void __launch_bounds__(kBlock)
    residual_kernel(int nx, int ny, int nz, double dx2, double inv_dx2, BcTypes bc, int pin,
                    const double* __restrict__ x, const double* __restrict__ b,
                    const double* __restrict__ K, double* __restrict__ r) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)nx * ny * nz;
    if (o >= n) return;
    const int i = (int)(o % nx);
    const long q = o / nx;
    const int j = (int)(q % ny);
    const int k = (int)(q / ny);
    r[o] = residual_value(nx, ny, nz, bc, pin != 0, x, b, K, i, j, k, o, dx2, inv_dx2);
}

// Residual fused with restriction: one thread per coarse cell evaluates the
// residual of its 8 children and averages them in the order of
// restrict_kernel (src/multigrid/transfer/restrict_3d.cu:36-49); the fine
// residual is never stored.
__global__ void __launch_bounds__(kBlock)
    residual_restrict_kernel(int nx, int ny, int nz, double dx2, double inv_dx2, BcTypes bc,
                             int pin, const double* __restrict__ x,
                             const double* __restrict__ b, const double* __restrict__ K,
                             double* __restrict__ bc_out) {
    const int NX = nx >> 1, NY = ny >> 1, NZ = nz >> 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)NX * NY * NZ) return;
    const int I = (int)(t % NX);
    const long q = t / NX;
    const int J = (int)(q % NY);
    const int Kc = (int)(q / NY);
    const int i = 2 * I, j = 2 * J, k = 2 * Kc;
    const long Nx = nx, st = (long)nx * ny;
    const long in = i + Nx * j + st * k;
    const bool p = pin != 0;
    double result = 0.0;
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i, j, k, in, dx2, inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i + 1, j, k, in + 1, dx2, inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i, j + 1, k, in + Nx, dx2, inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i + 1, j + 1, k, in + 1 + Nx, dx2,
                             inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i, j, k + 1, in + st, dx2, inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i + 1, j, k + 1, in + 1 + st, dx2,
                             inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i, j + 1, k + 1, in + Nx + st, dx2,
                             inv_dx2);
    result += residual_value(nx, ny, nz, bc, p, x, b, K, i + 1, j + 1, k + 1,
                             in + 1 + Nx + st, dx2, inv_dx2);
    bc_out[t] = result / 8.0;
}

// Operator value y = A x of one cell, VarCoeffLaplacian::apply
// (src/numerics/operators/varcoeff_laplacian.cu:38-239). The pin's diagonal
// counts only in-domain neighbours (varcoeff_laplacian.cu:110-119, 232-234).
__device__ inline double apply_value(int nx, int ny, int nz, const BcTypes& bc, bool pin,
                                     const double* __restrict__ K,
                                     const double* __restrict__ x, int i, int j, int k, long idx,
                                     double inv_dx2) {
    const double xC = x[idx], KC = K[idx];
    const long Nx = nx, st = (long)nx * ny;
    if (interior_cell(nx, ny, nz, i, j, k)) {
        const double iKC = 1.0 / KC;
        const double Kxm = 2.0 / (iKC + 1.0 / K[idx - 1]);
        const double Kxp = 2.0 / (iKC + 1.0 / K[idx + 1]);
        const double Kym = 2.0 / (iKC + 1.0 / K[idx - Nx]);
        const double Kyp = 2.0 / (iKC + 1.0 / K[idx + Nx]);
        const double Kzm = 2.0 / (iKC + 1.0 / K[idx - st]);
        const double Kzp = 2.0 / (iKC + 1.0 / K[idx + st]);
        const double Ax = Kxm * (xC - x[idx - 1]) + Kxp * (xC - x[idx + 1]) +
                          Kym * (xC - x[idx - Nx]) + Kyp * (xC - x[idx + Nx]) +
                          Kzm * (xC - x[idx - st]) + Kzp * (xC - x[idx + st]);
        return -Ax * inv_dx2;
    }
    double Ax = 0.0, aC = 0.0;
    const int pos[3] = {i, j, k};
    const int lim[3] = {nx, ny, nz};
    const long stride[3] = {1, Nx, st};
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int ax = d >> 1;
        const bool plus = d & 1;
        const bool inside = plus ? (pos[ax] < lim[ax] - 1) : (pos[ax] > 0);
        if (inside) {
            const long n = plus ? idx + stride[ax] : idx - stride[ax];
            const double Kh = 2.0 / (1.0 / KC + 1.0 / K[n]);
            Ax += Kh * (xC - x[n]);
            aC += Kh;
        } else if (bc.t[d] == MF3D_BC_PERIODIC) {
            const long n = plus ? idx - (long)(lim[ax] - 1) * stride[ax]
                                : idx + (long)(lim[ax] - 1) * stride[ax];
            const double Kh = 2.0 / (1.0 / KC + 1.0 / K[n]);
            Ax += Kh * (xC - x[n]);
        } else if (bc.t[d] == MF3D_BC_DIRICHLET) {
            Ax += 2.0 * KC * xC;
        }
    }
    if (pin && idx == 0) Ax += aC * xC;
    return -Ax * inv_dx2;
}

__device__ inline double block_sum(double v) {
    __shared__ double part[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (w == 0) {
        v = lane < kBlock / 64 ? part[lane] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    }
    return v;
}

// y = A x
__global__ void __launch_bounds__(kBlock)
    apply_kernel(int nx, int ny, int nz, double inv_dx2, BcTypes bc, int pin,
                 const double* __restrict__ K, const double* __restrict__ x,
                 double* __restrict__ y) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)nx * ny * nz;
    if (o < n) {
        const int i = (int)(o % nx);
        const long q = o / nx;
        const int j = (int)(q % ny);
        const int k = (int)(q / ny);
        y[o] = apply_value(nx, ny, nz, bc, pin != 0, K, x, i, j, k, o, inv_dx2);
    }
}

// restrict_kernel (src/multigrid/transfer/restrict_3d.cu:12-53)
__global__ void __launch_bounds__(kBlock)
    restrict_kernel(int NX, int NY, int NZ, const double* __restrict__ f,
                    double* __restrict__ c) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)NX * NY * NZ) return;
    const int I = (int)(t % NX);
    const long q = t / NX;
    const int J = (int)(q % NY);
    const int Kk = (int)(q / NY);
    const long Nx = 2L * NX, st = Nx * 2L * NY;
    const long in = 2L * I + Nx * 2L * J + st * 2L * Kk;
    double result = 0.0;
    result += f[in];
    result += f[in + 1];
    result += f[in + Nx];
    result += f[in + 1 + Nx];
    result += f[in + st];
    result += f[in + 1 + st];
    result += f[in + Nx + st];
    result += f[in + 1 + Nx + st];
    c[t] = result / 8.0;
}

// prolong_3d_add (src/multigrid/transfer/prolong_3d.cu:30-621): x_f += x_c[i/2, j/2, k/2].
__global__ void __launch_bounds__(kBlock)
    prolong_add_kernel(int nx, int ny, int nz, const double* __restrict__ c,
                       double* __restrict__ f) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)nx * ny * nz) return;
    const int i = (int)(o % nx);
    const long q = o / nx;
    const int j = (int)(q % ny);
    const int k = (int)(q / ny);
    const int NX = nx >> 1, NY = ny >> 1;
    f[o] += c[(long)(i >> 1) + (long)NX * ((long)(j >> 1) + (long)NY * (long)(k >> 1))];
}

// coarsen_K_kernel (src/physics/flow/coarsen_K.cu:15-53): geometric mean.
__global__ void __launch_bounds__(kBlock)
    coarsen_kernel(int NX, int NY, int NZ, const double* __restrict__ Kf,
                   double* __restrict__ Kc) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)NX * NY * NZ) return;
    const int I = (int)(t % NX);
    const long q = t / NX;
    const int J = (int)(q % NY);
    const int Kk = (int)(q / NY);
    const long Nx = 2L * NX, st = Nx * 2L * NY;
    const long i0 = 2L * I + Nx * 2L * J + st * 2L * Kk;
    const double s = log(Kf[i0]) + log(Kf[i0 + 1]) + log(Kf[i0 + Nx]) + log(Kf[i0 + 1 + Nx]) +
                     log(Kf[i0 + st]) + log(Kf[i0 + 1 + st]) + log(Kf[i0 + Nx + st]) +
                     log(Kf[i0 + 1 + Nx + st]);
    Kc[t] = exp(s / 8.0);
}

// build_rhs_head (src/physics/flow/rhs_head.cu:21-194): b = 0, then
// b -= 2 h_bc K / dx^2 on each Dirichlet face, faces in the order
// xmin, xmax, ymin, ymax, zmin, zmax.
__global__ void __launch_bounds__(kBlock)
    rhs_kernel(int nx, int ny, int nz, double inv_dx2, mf3d_bc bc,
               const double* __restrict__ K, double* __restrict__ rhs) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)nx * ny * nz) return;
    const int i = (int)(o % nx);
    const long q = o / nx;
    const int j = (int)(q % ny);
    const int k = (int)(q / ny);
    const int pos[3] = {i, j, k};
    const int lim[3] = {nx, ny, nz};
    double v = 0.0;
    for (int f = 0; f < 6; ++f) {
        if (bc.type[f] != MF3D_BC_DIRICHLET) continue;
        const int ax = f >> 1;
        if (pos[ax] != ((f & 1) ? lim[ax] - 1 : 0)) continue;
        v -= 2.0 * bc.value[f] * K[o] * inv_dx2;
    }
    rhs[o] = v;
}

// kernel_init_head_linear_x / kernel_fill_value (src/physics/flow/solve_head.cu:29-50)
__global__ void __launch_bounds__(kBlock)
    guess_kernel(int nx, int ny, int nz, int linear, double hw, double he,
                 double* __restrict__ h) {
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)nx * ny * nz) return;
    if (linear) {
        const int i = (int)(o % nx);
        const double t = ((double)i + 0.5) / (double)nx;
        h[o] = hw + (he - hw) * t;
    } else {
        h[o] = 0.0;
    }
}

// Whole GSRB smoothing of a small level by one workgroup: x and 1/K in LDS,
// each thread's own K and b in registers. Same per-cell values as
// gsrb_colour_kernel up to the summation order of boundary cells; replaces
// 2*iters launches (the coarsest level runs 50 iterations, v_cycle.cu:41-45).
constexpr int kSmallBlock = 1024;
constexpr int kSmallPer = 4;
constexpr int kSmallMaxCells = kSmallBlock * kSmallPer;

__global__ void __launch_bounds__(kSmallBlock)
    gsrb_small_kernel(int nx, int ny, int nz, double dx2, BcTypes bc, int pin,
                      double* __restrict__ x, const double* __restrict__ b,
                      const double* __restrict__ K, int iters, int zero_init) {
    __shared__ double xs[kSmallMaxCells];
    __shared__ double iks[kSmallMaxCells];
    const int n = nx * ny * nz;
    double bl[kSmallPer], Kl[kSmallPer];
#pragma unroll
    for (int m = 0; m < kSmallPer; ++m) {
        const int o = threadIdx.x + m * kSmallBlock;
        bl[m] = 0.0;
        Kl[m] = 1.0;
        if (o < n) {
            xs[o] = zero_init ? 0.0 : x[o];
            Kl[m] = K[o];
            iks[o] = 1.0 / Kl[m];
            bl[m] = b[o];
        }
    }
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        for (int colour = 0; colour < 2; ++colour) {
#pragma unroll
            for (int m = 0; m < kSmallPer; ++m) {
                const int o = threadIdx.x + m * kSmallBlock;
                if (o >= n) continue;
                const int i = o % nx, q = o / nx, j = q % ny, k = q / ny;
                if (((i + j + k) & 1) != colour) continue;
                const int pos[3] = {i, j, k};
                const int lim[3] = {nx, ny, nz};
                const int str[3] = {1, nx, nx * ny};
                double result = 0.0, aC = 0.0;
                for (int d = 0; d < 6; ++d) {
                    const int ax = d >> 1;
                    const bool plus = d & 1;
                    const bool edge = plus ? pos[ax] == lim[ax] - 1 : pos[ax] == 0;
                    double c;
                    if (edge && bc.t[d] != MF3D_BC_PERIODIC) {
                        c = bc.t[d] == MF3D_BC_DIRICHLET ? 2.0 * Kl[m] : 0.0;
                    } else {
                        const int nb = edge ? (plus ? o - (lim[ax] - 1) * str[ax]
                                                    : o + (lim[ax] - 1) * str[ax])
                                            : (plus ? o + str[ax] : o - str[ax]);
                        c = 2.0 / (iks[o] + iks[nb]);
                        result += c * xs[nb];
                    }
                    aC += c;
                }
                if (pin && o == 0) aC *= 2.0;
                xs[o] = (result - bl[m] * dx2) / aC;
            }
            __syncthreads();
        }
    }
    for (int o = threadIdx.x; o < n; o += kSmallBlock) x[o] = xs[o];
}

// Coarsest direct solve x = A^{-1} b, one wave per row (dense FP64 GEMV).
__global__ void __launch_bounds__(256)
    dense_gemv_kernel(int n, const double* __restrict__ Ainv, const double* __restrict__ b,
                      double* __restrict__ x) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* a = Ainv + (long)row * n;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += a[c] * b[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) x[row] = s;
}

// Inverse of the coarsest operator from its banded Cholesky factor M = -A =
// L L^T (row-major band storage, Lb[i*(w+1) + (i-j)] = L[i][j] for
// i-w <= j <= i), as LAPACK's potri does it: the columns of L^{-1}, then
// A^{-1} = -M^{-1} = -L^{-T} L^{-1} as one symmetric product.
//
// band_lower_inverse: one wave per column c solves L y = e_c and writes
// Y[c][i] = y[i] = L^{-1}[i][c] (zero for i < c); y lives in LDS.
// band_lower_slab (w <= 64): the substitution goes in blocks of 32 rows, one
// row per lane. The band rows of a block are first copied to LDS with
// coalesced loads (transposed: S[d * 32 + r] = L[R0 + r][R0 + r - d]); a lane
// then gathers its row's terms from the blocks already solved, and the block
// is solved row by row, each new value broadcast from its lane with
// v_readlane and folded into the later rows' sums, so a sequential step is a
// multiply, a broadcast and a fused multiply-add. LDS: n + kSlab doubles.
// band_lower_row (any w): one row per step, the band dot product spread over
// the lanes and reduced across the wave. LDS: n doubles.
// syrk_neg_kernel: A^{-1}[i][j] = -sum_{k >= max(i, j)} Y[i][k] Y[j][k].
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

constexpr int kSlabRows = 32;
constexpr int kSlab = kSlabRows * 65;  // doubles per staged block of band rows (w <= 64)

// S[d * 32 + r] = Lb[(R0 + r) * W1 + d] (0 outside the matrix); all loads of a
// block are in flight together
__device__ __forceinline__ void stage_band(const double* __restrict__ Lb, int W1, int n, int R0,
                                           double* __restrict__ S) {
    const int lane = threadIdx.x;
    double v[kSlabRows];
#pragma unroll
    for (int r = 0; r < kSlabRows; ++r) {
        const int row = R0 + r;
        const bool ok = row >= 0 && row < n;
        v[r] = ok && lane < W1 ? Lb[(long)row * W1 + lane] : 0.0;
    }
    double v64 = 0.0;  // d = 64 (w = 64): lane r < 32 takes row R0 + r
    if (W1 > 64 && lane < kSlabRows) {
        const int row = R0 + lane;
        v64 = row >= 0 && row < n ? Lb[(long)row * W1 + 64] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kSlabRows; ++r)
        if (lane < W1) S[lane * kSlabRows + r] = v[r];
    if (W1 > 64 && lane < kSlabRows) S[64 * kSlabRows + lane] = v64;
}

__device__ void band_lower_slab(int n, int w, const double* __restrict__ Lb,
                                double* __restrict__ Y, int c, double* __restrict__ y,
                                double* __restrict__ S0) {
    constexpr int B = kSlabRows;
    const int lane = threadIdx.x;
    const bool part = lane < B;  // lanes B..63 only help with the copies
    const int W1 = w + 1;
    // forward: L y = e_c; rows before c are zero
    for (int i0 = c; i0 < n; i0 += B) {
        const int i = i0 + lane;
        const bool act = part && i < n;
        __syncthreads();  // S0 of the previous block has been read
        stage_band(Lb, W1, n, i0, S0);
        __syncthreads();
        // terms j = i - d (d = lane + 1 .. min(w, i - c)) from the blocks already solved
        const int dmax = act ? min(w, i - c) : 0;
        double p[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int d0 = 1; d0 <= w; d0 += 8) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int d = d0 + k;
                const bool ok = d > lane && d <= dmax;
                const double L = S0[min(d, w) * B + (lane & (B - 1))];
                const double yv = y[ok ? i - d : 0];
                p[k] += ok ? L * yv : 0.0;
            }
        }
        double acc = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
        const double rd = act ? 1.0 / S0[lane] : 0.0;
        const double e = i == c ? 1.0 : 0.0;
        double yi = 0.0;
        const int m = min(B, n - i0);
        for (int s = 0; s < m; ++s) {
            const int d = lane - s;
            const double L = S0[(d > 0 && d <= w ? d : 0) * B + (lane & (B - 1))];
            const double ys = readlane_f64((e - acc) * rd, s);
            if (lane == s) yi = ys;
            if (act && d > 0 && d <= w) acc += L * ys;
        }
        if (act) y[i] = yi;
    }
    __syncthreads();
    double* out = Y + (long)c * n;
    for (int i = lane; i < n; i += 64) out[i] = i < c ? 0.0 : y[i];
}

__device__ void band_lower_row(int n, int w, const double* __restrict__ Lb,
                               double* __restrict__ Y, int c, double* __restrict__ y) {
    const int lane = threadIdx.x;
    const int W1 = w + 1;
    for (int i = lane; i < n; i += 64) y[i] = 0.0;
    __syncthreads();
    for (int i = c; i < n; ++i) {
        double s = 0.0;
        for (int t = lane; t < w; t += 64) {
            const int j = i - 1 - t;
            if (j >= c) s += Lb[(long)i * W1 + (i - j)] * y[j];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        const double yi = ((i == c ? 1.0 : 0.0) - s) / Lb[(long)i * W1];
        if (lane == 0) y[i] = yi;
        __syncthreads();
    }
    double* out = Y + (long)c * n;
    for (int i = lane; i < n; i += 64) out[i] = y[i];
}

// the slab variant when its LDS fits in 64 KB
__host__ __device__ inline bool band_slab_ok(int n, int w) {
    return w <= 64 && (long)(n + kSlab) * 8 <= 65536;
}

__device__ __forceinline__ void band_lower_body(int n, int w, const double* __restrict__ Lb,
                                                double* __restrict__ Y, int c, int nmax,
                                                bool slab) {
    extern __shared__ double y[];
    if (slab) band_lower_slab(n, w, Lb, Y, c, y, y + nmax);
    else band_lower_row(n, w, Lb, Y, c, y);
}

__global__ void __launch_bounds__(64)
    band_lower_kernel(int n, int w, const double* __restrict__ Lb, double* __restrict__ Y,
                      int slab) {
    band_lower_body(n, w, Lb, Y, blockIdx.x, n, slab != 0);
}

// many factors at once: entry e owns the blocks [start[e], start[e+1]), one per column
__global__ void __launch_bounds__(64)
    band_lower_batch_kernel(const BandDesc* __restrict__ D, const int* __restrict__ start, int m,
                            int nmax, int slab) {
    const int e = batch_entry(start, m, blockIdx.x);
    const BandDesc& B = D[e];
    band_lower_body(B.n, B.w, B.Lb, B.Y, (int)blockIdx.x - start[e], nmax, slab != 0);
}

// Ainv = -Y Y^T in 32 x 32 tiles, 2 x 2 entries per thread: only tiles with
// I >= J are computed, each entry written to (i, j) and (j, i), so the result
// is exactly symmetric. k runs from the tile's first nonzero column
// (Y[i][k] = 0 for k < i) in steps of 32 through LDS.
constexpr int kSyrkTile = 32, kSyrkK = 32;

__device__ __forceinline__ void syrk_neg_tile(int n, const double* __restrict__ Y,
                                              double* __restrict__ Ainv, int I0, int J0) {
    __shared__ double As[kSyrkK][kSyrkTile + 1], Bs[kSyrkK][kSyrkTile + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int kb = I0 & ~(kSyrkK - 1); kb < n; kb += kSyrkK) {  // I0 >= J0
        for (int q = threadIdx.x; q < kSyrkTile * kSyrkK; q += 256) {
            const int r = q / kSyrkK, kk = q % kSyrkK;
            const int gk = kb + kk;
            As[kk][r] = I0 + r < n && gk < n ? Y[(long)(I0 + r) * n + gk] : 0.0;
            Bs[kk][r] = J0 + r < n && gk < n ? Y[(long)(J0 + r) * n + gk] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kSyrkK; ++kk) {
            const double a0 = As[kk][ty], a1 = As[kk][ty + 16];
            const double b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int i = I0 + ty + 16 * u, j = J0 + tx + 16 * v;
            if (i < n && j < n && j <= i) {
                Ainv[(long)i * n + j] = -acc[u][v];
                Ainv[(long)j * n + i] = -acc[u][v];
            }
        }
}

__global__ void __launch_bounds__(256)
    syrk_neg_kernel(int n, const double* __restrict__ Y, double* __restrict__ Ainv) {
    if (blockIdx.x > blockIdx.y) return;
    syrk_neg_tile(n, Y, Ainv, blockIdx.y * kSyrkTile, blockIdx.x * kSyrkTile);
}

// one realisation per blockIdx.z; tiles beyond its n return
__global__ void __launch_bounds__(256) syrk_neg_batch_kernel(const BandDesc* __restrict__ D) {
    const BandDesc& B = D[blockIdx.z];
    const int I0 = blockIdx.y * kSyrkTile, J0 = blockIdx.x * kSyrkTile;
    if (blockIdx.x > blockIdx.y || I0 >= B.n) return;
    syrk_neg_tile(B.n, B.Y, B.Ainv, I0, J0);
}

// ---------------------------------------------------------------------------
// Banded Cholesky factors M = L L^T of the coarsest operators M = -A of a
// batched solve, assembled from K on the device. One workgroup per factor:
//  1. the thread of each cell writes that cell's column of M into column
//     storage (C, global scratch); the factor's band entries above the matrix
//     are zeroed;
//  2. right-looking factorization, one column j per step: column j is scaled
//     by 1 / sqrt(M[j][j]) and written to the row-major factor, and its outer
//     product is subtracted from the next w columns, each entry with one fused
//     multiply-add (entry (j + dk + t, j + dk) -= L[j + dk + t][j] L[j + dk][j]),
//     steps in ascending j.
// Bands of w + 1 <= kFactorWin entries per column factor in an LDS window of
// w + 2 columns (column k in slot k mod (w + 2)) with one barrier per step:
// during step j, wave 0 updates column j + 1 and scales it at once (its pivot
// broadcast within the wave), waves 1..15 update columns j + 2 .. j + w, and
// the last two waves store the column entering the window (loaded a step
// earlier, into the slot column j - 1 left) and load the next one; barriers
// wait for LDS only, so that load stays in flight, and no wave waits for the
// factor's global stores. Wider bands factor in place in C, two barriers per
// step. Both variants round identically (the same operations in the same
// order).
// ---------------------------------------------------------------------------
constexpr int kFactorThreads = 1024;
constexpr int kFactorWin = 90;    // (w + 2)(w + 1) doubles of LDS: 65,520 bytes at w + 1 = 90
constexpr int kFactorPairs = 5;   // update pairs per thread of waves 1..15: w (w - 1) / 2 <= 5 * 960

// barrier for LDS data only: outstanding global loads and stores stay in flight
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// M into C (and columns 0..w into window slots 0..w when win is given). C and
// the window are cleared first; then the thread of cell c writes column c:
// the diagonal sum_f K_f / dx^2 (doubled at a pinned cell 0) plus -K_f / dx^2
// for each face that wraps onto the cell itself, and -K_f / dx^2 at offset
// q - c for each face joining c to a cell q > c (faces to the same q
// accumulate in face order). The face coefficients are symmetric, so column
// q's entry at c is the same sum seen from q.
__device__ void band_assemble(const FactorDesc& D, int n, int W1, double* __restrict__ win) {
    constexpr int nt = kFactorThreads;
    const int tid = threadIdx.x;
    const long nW = (long)n * W1;
    for (long e = tid; e < nW; e += nt) D.C[e] = 0.0;
    if (win)
        for (int e = tid; e < W1 * W1; e += nt) win[e] = 0.0;
    for (int e = tid; e < W1 * W1; e += nt) {  // L[i][i - d] with i - d < 0
        const int i = e / W1, d = e - i * W1;
        if (d > i && i < n) D.Lb[e] = 0.0;
    }
    __syncthreads();
    const int nx = D.nx, ny = D.ny, nz = D.nz;
    for (int c = tid; c < n; c += nt) {
        const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        const double Kc = D.K[c];
        double coef[6], asum = 0.0;
        long nb[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            nb[d] = neighbour_coef(nx, ny, nz, D.bc, D.K, i, j, k, d, Kc, coef[d]);
            asum += coef[d];
        }
        if (D.pin && c == 0) asum *= 2.0;
        double e[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) e[d] = -coef[d] / D.dx2;
        double diag = asum / D.dx2;
#pragma unroll
        for (int d = 0; d < 6; ++d)
            if (nb[d] == c) diag += e[d];
        double* col = D.C + (long)c * W1;
        double* wcol = win && c < W1 ? win + c * W1 : nullptr;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            bool first = nb[d] > c;  // the first face to this neighbour writes its entry
            for (int d2 = 0; d2 < d; ++d2) first = first && nb[d2] != nb[d];
            if (!first) continue;
            double acc = e[d];
            for (int d2 = d + 1; d2 < 6; ++d2)
                if (nb[d2] == nb[d]) acc += e[d2];
            col[nb[d] - c] = acc;
            if (wcol) wcol[nb[d] - c] = acc;
        }
        col[0] = diag;
        if (wcol) wcol[0] = diag;
    }
}

__device__ void band_factor_window(const FactorDesc& D, double* __restrict__ win) {
    constexpr int nt = kFactorThreads;
    const int n = D.nx * D.ny * D.nz;
    const int w = D.w, W1 = w + 1, WS = w + 2;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    band_assemble(D, n, W1, win);
    __syncthreads();
    // waves 1..15: this thread's update pairs (dk >= 2, t), dk = 0 for none
    int pdk[kFactorPairs], pt[kFactorPairs];
#pragma unroll
    for (int r = 0; r < kFactorPairs; ++r) pdk[r] = pt[r] = 0;
    if (wave > 0) {
        int e0 = 0;
        for (int dk = 2; dk <= w; ++dk) {
            const int len = W1 - dk;
#pragma unroll
            for (int r = 0; r < kFactorPairs; ++r) {
                const int idx = tid - 64 + r * (nt - 64) - e0;
                if (idx >= 0 && idx < len) {
                    pdk[r] = dk;
                    pt[r] = idx;
                }
            }
            e0 += len;
        }
    }
    // the last two waves: entry pf of the column entering the window next
    const int pf = tid - (nt - 128);
    const bool pfl = pf >= 0 && pf < W1;
    double pre = 0.0;
    if (pfl && w + 1 < n) pre = D.C[(long)(w + 1) * W1 + pf];
    int bad = 0;
    if (wave == 0) {  // column 0
        const int m0 = min(w, n - 1);
        const double d0 = win[0];
        const double L0 = sqrt(d0);
        if (!(d0 > 0.0)) bad = 1;
        for (int t = lane; t <= m0; t += 64) {
            const double v = t == 0 ? L0 : win[t] / L0;
            if (t > 0) win[t] = v;
            D.Lb[(long)t * W1 + t] = v;
        }
    }
    lds_barrier();
    int sj = 0;  // slot of column j
    for (int j = 0; j < n; ++j) {
        const int m = min(w, n - 1 - j);
        const double* cj = win + sj * W1;
        const int s1 = sj + 1 == WS ? 0 : sj + 1;
        if (wave == 0) {
            if (m >= 1) {  // column j + 1: entries t = lane and lane + 64
                const int m1 = min(w, n - 2 - j);
                double* c1 = win + s1 * W1;
                const double l1 = cj[1];
                const int t0 = lane, t1 = lane + 64;
                double v0 = 0.0, v1 = 0.0;
                if (t0 <= m1) {
                    v0 = c1[t0];
                    if (t0 <= m - 1) v0 = fma(-cj[1 + t0], l1, v0);
                }
                if (t1 <= m1) {
                    v1 = c1[t1];
                    if (t1 <= m - 1) v1 = fma(-cj[1 + t1], l1, v1);
                }
                const double piv = readlane_f64(v0, 0);
                const double L = sqrt(piv);
                if (!(piv > 0.0) && bad == 0) bad = j + 2;
                if (t0 <= m1) {
                    const double v = t0 == 0 ? L : v0 / L;
                    if (t0 > 0) c1[t0] = v;
                    D.Lb[(long)(j + 1 + t0) * W1 + t0] = v;
                }
                if (t1 <= m1) {
                    const double v = v1 / L;
                    c1[t1] = v;
                    D.Lb[(long)(j + 1 + t1) * W1 + t1] = v;
                }
            }
        } else {
            if (pfl) {  // column j + 1 + w into the slot of column j - 1
                if (j + 1 + w < n) win[(sj == 0 ? WS - 1 : sj - 1) * W1 + pf] = pre;
                if (j + 2 + w < n) pre = D.C[(long)(j + 2 + w) * W1 + pf];
            }
            double a[kFactorPairs], l[kFactorPairs], c[kFactorPairs];
            int at[kFactorPairs];
            bool ok[kFactorPairs];
#pragma unroll
            for (int r = 0; r < kFactorPairs; ++r) {
                const int dk = pdk[r], t = pt[r];
                ok[r] = dk > 0 && dk + t <= m;
                const int s = sj + dk < WS ? sj + dk : sj + dk - WS;
                at[r] = s * W1 + t;
                a[r] = ok[r] ? cj[dk + t] : 0.0;
                l[r] = ok[r] ? cj[dk] : 0.0;
                c[r] = ok[r] ? win[at[r]] : 0.0;
            }
#pragma unroll
            for (int r = 0; r < kFactorPairs; ++r)
                if (ok[r]) win[at[r]] = fma(-a[r], l[r], c[r]);
        }
        lds_barrier();
        sj = s1;
    }
    if (tid == 0) *D.flag = bad;
}

__device__ void band_factor_global(const FactorDesc& D) {
    constexpr int nt = kFactorThreads;
    const int n = D.nx * D.ny * D.nz;
    const int w = D.w, W1 = w + 1;
    const int tid = threadIdx.x;
    band_assemble(D, n, W1, nullptr);
    __syncthreads();
    int bad = 0;
    for (int j = 0; j < n; ++j) {
        const int m = min(w, n - 1 - j);
        double* cj = D.C + (long)j * W1;
        const double djj = cj[0];
        const double Ljj = sqrt(djj);
        if (!(djj > 0.0) && bad == 0) bad = j + 1;
        for (int t = tid; t <= m; t += nt) {
            const double v = t == 0 ? Ljj : cj[t] / Ljj;
            if (t > 0) cj[t] = v;  // the pivot entry is read by every thread and stays
            D.Lb[(long)(j + t) * W1 + t] = v;
        }
        __syncthreads();
        for (int dk = 1 + tid / 64; dk <= m; dk += nt / 64) {
            double* ck = D.C + (long)(j + dk) * W1;
            const double l = cj[dk];
            for (int t = tid & 63; t <= m - dk; t += 64) ck[t] = fma(-cj[dk + t], l, ck[t]);
        }
        __syncthreads();
    }
    if (tid == 0) *D.flag = bad;
}

template <bool WIN>
__device__ __forceinline__ void band_factor_body(const FactorDesc& D, double* win) {
    if (WIN) band_factor_window(D, win);
    else band_factor_global(D);
}

template <bool WIN>
__global__ void __launch_bounds__(kFactorThreads)
    band_factor_kernel(const FactorDesc* __restrict__ D) {
    extern __shared__ double win[];
    band_factor_body<WIN>(D[blockIdx.x], win);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace kern {

int gsrb(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double* x,
         const double* b, const double* K, int iters, bool zero_init) {
    if (iters <= 0) return zero_init ? fill(s, x, cells(g), 0.0) : MF3D_OK;
    const double dx2 = g.dx * g.dx;
    const size_t n = cells(g);
    if (n <= (size_t)kSmallMaxCells && iters > 1) {
        gsrb_small_kernel<<<1, kSmallBlock, 0, s>>>(g.nx, g.ny, g.nz, dx2, bc, pin, x, b, K,
                                                    iters, zero_init ? 1 : 0);
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    }
    const size_t half = (size_t)((g.nx + 1) / 2) * g.ny * g.nz;
    const unsigned grid = grid_for(half);
    for (int it = 0; it < iters; ++it) {
        for (int colour = 0; colour < 2; ++colour) {
            if (zero_init && it == 0 && colour == 0)
                gsrb_colour_kernel<true><<<grid, kBlock, 0, s>>>(g.nx, g.ny, g.nz, dx2, bc, pin,
                                                                 x, b, K, colour);
            else
                gsrb_colour_kernel<false><<<grid, kBlock, 0, s>>>(g.nx, g.ny, g.nz, dx2, bc, pin,
                                                                  x, b, K, colour);
        }
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int residual(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, const double* x,
             const double* b, const double* K, double* r) {
    const size_t n = cells(g);
    residual_kernel<<<grid_for(n), kBlock, 0, s>>>(g.nx, g.ny, g.nz, g.dx * g.dx,
                                                   1.0 / (g.dx * g.dx), bc, pin, x, b, K, r);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int residual_restrict(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                      const double* x, const double* b, const double* K, double* b_coarse) {
    const size_t nc = (size_t)(g.nx / 2) * (g.ny / 2) * (g.nz / 2);
    residual_restrict_kernel<<<grid_for(nc), kBlock, 0, s>>>(
        g.nx, g.ny, g.nz, g.dx * g.dx, 1.0 / (g.dx * g.dx), bc, pin, x, b, K, b_coarse);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

unsigned reduce_blocks(size_t n) { return (unsigned)((n + kReduceCells - 1) / kReduceCells); }

int apply(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, const double* K,
          const double* x, double* y) {
    const size_t n = cells(g);
    apply_kernel<<<grid_for(n), kBlock, 0, s>>>(g.nx, g.ny, g.nz, 1.0 / (g.dx * g.dx), bc, pin, K,
                                                x, y);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int restrict_avg(hipStream_t s, const mf3d_grid& coarse, const double* f, double* c) {
    const size_t nc = cells(coarse);
    restrict_kernel<<<grid_for(nc), kBlock, 0, s>>>(coarse.nx, coarse.ny, coarse.nz, f, c);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int prolong_add(hipStream_t s, const mf3d_grid& fine, const double* c, double* f) {
    prolong_add_kernel<<<grid_for(cells(fine)), kBlock, 0, s>>>(fine.nx, fine.ny, fine.nz, c, f);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int coarsen(hipStream_t s, const mf3d_grid& coarse, const double* Kf, double* Kc) {
    coarsen_kernel<<<grid_for(cells(coarse)), kBlock, 0, s>>>(coarse.nx, coarse.ny, coarse.nz,
                                                              Kf, Kc);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int rhs(hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc, const double* K, double* out) {
    rhs_kernel<<<grid_for(cells(g)), kBlock, 0, s>>>(g.nx, g.ny, g.nz, 1.0 / (g.dx * g.dx), bc, K,
                                                     out);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int guess(hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc, int mode, double* h) {
    // mode: MF3D_GUESS_REFERENCE (linear in x when both x faces are Dirichlet,
    // else zero) or MF3D_GUESS_ZERO.
    const int linear = (mode == MF3D_GUESS_REFERENCE && bc.type[0] == MF3D_BC_DIRICHLET &&
                        bc.type[1] == MF3D_BC_DIRICHLET)
                           ? 1
                           : 0;
    guess_kernel<<<grid_for(cells(g)), kBlock, 0, s>>>(g.nx, g.ny, g.nz, linear, bc.value[0],
                                                       bc.value[1], h);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int dense_gemv(hipStream_t s, int n, const double* Ainv, const double* b, double* x) {
    dense_gemv_kernel<<<(n + 3) / 4, 256, 0, s>>>(n, Ainv, b, x);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int band_factor_batch(hipStream_t s, const FactorDesc* D, int m, int wmax) {
    const int W1 = wmax + 1;
    if (W1 <= kFactorWin)
        band_factor_kernel<true>
            <<<m, kFactorThreads, (size_t)(W1 + 1) * W1 * sizeof(double), s>>>(D);
    else
        band_factor_kernel<false><<<m, kFactorThreads, 0, s>>>(D);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int band_inverse(hipStream_t s, int n, int w, const double* Lb, double* Y, double* Ainv) {
    const bool slab = band_slab_ok(n, w);
    const size_t lds = ((size_t)n + (slab ? kSlab : 0)) * sizeof(double);
    band_lower_kernel<<<n, 64, lds, s>>>(n, w, Lb, Y, slab ? 1 : 0);
    MF3D_HIP_TRY(hipGetLastError());
    const unsigned t = (unsigned)((n + kSyrkTile - 1) / kSyrkTile);
    syrk_neg_kernel<<<dim3(t, t), 256, 0, s>>>(n, Y, Ainv);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

// wmax: the largest bandwidth of the batch (the slab variant needs every w <= 64)
int band_inverse_batch(hipStream_t s, const BandDesc* D, const int* start, int m, int nblocks,
                       int nmax, int wmax) {
    const bool slab = band_slab_ok(nmax, wmax);
    const size_t lds = ((size_t)nmax + (slab ? kSlab : 0)) * sizeof(double);
    band_lower_batch_kernel<<<nblocks, 64, lds, s>>>(D, start, m, nmax, slab ? 1 : 0);
    MF3D_HIP_TRY(hipGetLastError());
    const unsigned t = (unsigned)((nmax + kSyrkTile - 1) / kSyrkTile);
    syrk_neg_batch_kernel<<<dim3(t, t, (unsigned)m), 256, 0, s>>>(D);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace kern
}  // namespace mf3d
