// Darcy face velocities from head (FP64), gfx950.
//
// One launch covers every face, interior and boundary, in either layout: each
// thread owns cell (i,j,k) and writes its lower x-, y- and z-faces, plus the
// upper boundary face when the cell touches xmax/ymax/zmax. The reference
// launches 9 kernels on the legacy default stream and then syncs a different
// stream (src/physics/flow/velocity_from_head.cu:637-690); here everything is
// on the caller's stream.
//
// Formulas (velocity_from_head.cu:102-381):
//   interior face: q = -K_h (h_R - h_L) / dx, K_h = 2 Ka Kb / (Ka + Kb)
//   Dirichlet face: q = -K_C (h_in - h_bc) / (dx/2) (lower face),
//                   q = -K_C (h_bc - h_in) / (dx/2) (upper face)
//   Neumann face: 0; periodic face: harmonic mean across the wrap.

#include <algorithm>

#include "kernels.hpp"

namespace mf3d {
namespace {

constexpr int kBlock = 256;

__device__ inline double hmean_v(double a, double b) { return 2.0 * a * b / (a + b); }

struct FaceIdx {
    int nx, ny, nz;
    bool padded;
    __device__ long pidx(int i, int j, int k) const {
        return (long)k * (long)(ny + 1) * (long)(nx + 1) + (long)j * (long)(nx + 1) + (long)i;
    }
    __device__ long U(int i, int j, int k) const {
        return padded ? pidx(i, j, k) : (long)i + (long)(nx + 1) * ((long)j + (long)ny * k);
    }
    __device__ long V(int i, int j, int k) const {
        return padded ? pidx(i, j, k) : (long)i + (long)nx * ((long)j + (long)(ny + 1) * k);
    }
    __device__ long W(int i, int j, int k) const {
        return padded ? pidx(i, j, k) : (long)i + (long)nx * ((long)j + (long)ny * k);
    }
};

__global__ void __launch_bounds__(kBlock)
    velocity_kernel(FaceIdx fi, double dx, double dy, double dz, mf3d_bc bc,
                    const double* __restrict__ H, const double* __restrict__ K,
                    double* __restrict__ U, double* __restrict__ V, double* __restrict__ W) {
    const int nx = fi.nx, ny = fi.ny, nz = fi.nz;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long)nx * ny * nz) return;
    const int i = (int)(o % nx);
    const long q = o / nx;
    const int j = (int)(q % ny);
    const int k = (int)(q / ny);
    const long sy = nx, sz = (long)nx * ny;
    const double Hc = H[o], Kc = K[o];

    // ---- x faces ----
    if (i > 0) {
        const long l = o - 1;
        U[fi.U(i, j, k)] = -hmean_v(K[l], Kc) * (Hc - H[l]) / dx;
    } else {  // kernel_U_west (velocity_from_head.cu:196-227)
        double u = 0.0;
        if (bc.type[0] == MF3D_BC_DIRICHLET) {
            u = -Kc * (Hc - bc.value[0]) / (dx * 0.5);
        } else if (bc.type[0] == MF3D_BC_PERIODIC) {
            const long w = o + (nx - 1);
            u = -hmean_v(K[w], Kc) * (Hc - H[w]) / dx;
        }
        U[fi.U(0, j, k)] = u;
    }
    if (i == nx - 1) {  // kernel_U_east (velocity_from_head.cu:232-258)
        double u = 0.0;
        if (bc.type[1] == MF3D_BC_DIRICHLET) {
            u = -Kc * (bc.value[1] - Hc) / (dx * 0.5);
        } else if (bc.type[1] == MF3D_BC_PERIODIC) {
            const long w = o - (nx - 1);
            u = -hmean_v(Kc, K[w]) * (H[w] - Hc) / dx;
        }
        U[fi.U(nx, j, k)] = u;
    }
    // ---- y faces ----
    if (j > 0) {
        const long s = o - sy;
        V[fi.V(i, j, k)] = -hmean_v(K[s], Kc) * (Hc - H[s]) / dy;
    } else {
        double v = 0.0;
        if (bc.type[2] == MF3D_BC_DIRICHLET) {
            v = -Kc * (Hc - bc.value[2]) / (dy * 0.5);
        } else if (bc.type[2] == MF3D_BC_PERIODIC) {
            const long w = o + (long)(ny - 1) * sy;
            v = -hmean_v(K[w], Kc) * (Hc - H[w]) / dy;
        }
        V[fi.V(i, 0, k)] = v;
    }
    if (j == ny - 1) {
        double v = 0.0;
        if (bc.type[3] == MF3D_BC_DIRICHLET) {
            v = -Kc * (bc.value[3] - Hc) / (dy * 0.5);
        } else if (bc.type[3] == MF3D_BC_PERIODIC) {
            const long w = o - (long)(ny - 1) * sy;
            v = -hmean_v(Kc, K[w]) * (H[w] - Hc) / dy;
        }
        V[fi.V(i, ny, k)] = v;
    }
    // ---- z faces ----
    if (k > 0) {
        const long b = o - sz;
        W[fi.W(i, j, k)] = -hmean_v(K[b], Kc) * (Hc - H[b]) / dz;
    } else {
        double w = 0.0;
        if (bc.type[4] == MF3D_BC_DIRICHLET) {
            w = -Kc * (Hc - bc.value[4]) / (dz * 0.5);
        } else if (bc.type[4] == MF3D_BC_PERIODIC) {
            const long t = o + (long)(nz - 1) * sz;
            w = -hmean_v(K[t], Kc) * (Hc - H[t]) / dz;
        }
        W[fi.W(i, j, 0)] = w;
    }
    if (k == nz - 1) {
        double w = 0.0;
        if (bc.type[5] == MF3D_BC_DIRICHLET) {
            w = -Kc * (bc.value[5] - Hc) / (dz * 0.5);
        } else if (bc.type[5] == MF3D_BC_PERIODIC) {
            const long t = o - (long)(nz - 1) * sz;
            w = -hmean_v(Kc, K[t]) * (H[t] - Hc) / dz;
        }
        W[fi.W(i, j, nz)] = w;
    }
}

// ---- velocity diagnostics (velocity_diagnostics.cu:46-177) -------------------
// One workgroup owns a tile of 64 x 4 cell columns (one wave per x row) and
// marches kz planes up in z. Along z each thread carries the face values it
// needs again one plane later (U and V of its column at k-1, k, k+1; W at k
// and k+1 for itself and its four x/y neighbours), so a cell costs 9 marching
// loads plus 8 loads of the x/y neighbours' U and V at plane k, which the
// neighbouring lanes and rows bring into L1/L2. Algorithmic traffic: the three
// padded components read once and the three cell fields written once.
constexpr int kDiagX = 64, kDiagY = 4;
constexpr int kDiagQ = 9;  // per-workgroup partials: (min, max, sum) x 3 fields

struct DiagDims {
    int nx, ny, nz;
    int tx, kz;    // tiles along x; planes per workgroup
    int tiles;     // tiles per plane chunk
    long nblocks;  // tiles x chunks
    double ix, iy, iz;
    int curl;      // all dimensions >= 3
};

// padded_idx (src/physics/flow/padded_layout.cuh:30-33)
__device__ __forceinline__ long pad_idx(const DiagDims& d, int i, int j, int k) {
    return ((long)k * (d.ny + 1) + j) * (long)(d.nx + 1) + i;
}

struct MinMaxSum {
    double lo, hi, sum;
    __device__ void init() {
        lo = __builtin_huge_val();
        hi = -__builtin_huge_val();
        sum = 0.0;
    }
    __device__ void add(double v) {
        lo = fmin(lo, v);
        hi = fmax(hi, v);
        sum += v;
    }
};

// min, max and sum over the workgroup (kBlock threads), valid in thread 0;
// wv: LDS of 3 * kBlock / 64 doubles
__device__ inline void block_mms(MinMaxSum& a, double* wv) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a.lo = fmin(a.lo, __shfl_xor(a.lo, off, 64));
        a.hi = fmax(a.hi, __shfl_xor(a.hi, off, 64));
        a.sum += __shfl_xor(a.sum, off, 64);
    }
    constexpr int W = kBlock / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();  // wv may still be read by a previous call
    if (lane == 0) {
        wv[w] = a.lo;
        wv[W + w] = a.hi;
        wv[2 * W + w] = a.sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < W; ++q) {
            a.lo = fmin(a.lo, wv[q]);
            a.hi = fmax(a.hi, wv[W + q]);
        }
        double s = wv[2 * W];
        for (int q = 1; q < W; ++q) s += wv[2 * W + q];
        a.sum = s;
    }
}

__global__ void __launch_bounds__(kBlock)
    diag_kernel(DiagDims d, const double* __restrict__ U, const double* __restrict__ V,
                const double* __restrict__ W, double* __restrict__ dv, double* __restrict__ om,
                double* __restrict__ he, double* __restrict__ partials) {
    const long blk = xcd_block(d.nblocks);
    if (blk < 0) return;  // padding block: uniform exit before the reductions
    const int tile = (int)(blk % d.tiles), chunk = (int)(blk / d.tiles);
    const int i = (tile % d.tx) * kDiagX + (int)(threadIdx.x & 63);
    const int j = (tile / d.tx) * kDiagY + (int)(threadIdx.x >> 6);
    const int k0 = chunk * d.kz, k1 = min(k0 + d.kz, d.nz);
    const bool active = i < d.nx && j < d.ny;
    // every load stays inside the padded arrays: columns clamped to the grid,
    // neighbours to [0, n] (the padded extent), planes to [0, nz]
    const int ic = min(i, d.nx - 1), jc = min(j, d.ny - 1);
    const int im = max(ic - 1, 0), ip = ic + 1, jm = max(jc - 1, 0), jp = jc + 1;
    const bool colin = d.curl && ic >= 1 && ic <= d.nx - 2 && jc >= 1 && jc <= d.ny - 2;
    const double hx = 0.5 * d.ix, hy = 0.5 * d.iy, hz = 0.5 * d.iz;

    // U(i), U(i+1) and V(j), V(j+1) of the column at planes k-1, k, k+1
    const int kb = max(k0 - 1, 0);
    double ua0 = U[pad_idx(d, ic, jc, kb)], ub0 = U[pad_idx(d, ip, jc, kb)];
    double va0 = V[pad_idx(d, ic, jc, kb)], vb0 = V[pad_idx(d, ic, jp, kb)];
    double ua1 = U[pad_idx(d, ic, jc, k0)], ub1 = U[pad_idx(d, ip, jc, k0)];
    double va1 = V[pad_idx(d, ic, jc, k0)], vb1 = V[pad_idx(d, ic, jp, k0)];
    // W at plane k of the column and of its x/y neighbours
    double w0 = W[pad_idx(d, ic, jc, k0)];
    double wxm0 = W[pad_idx(d, im, jc, k0)], wxp0 = W[pad_idx(d, ip, jc, k0)];
    double wym0 = W[pad_idx(d, ic, jm, k0)], wyp0 = W[pad_idx(d, ic, jp, k0)];

    MinMaxSum sd, so, sh;
    sd.init();
    so.init();
    sh.init();
    for (int k = k0; k < k1; ++k) {
        const int kn = k + 1;  // <= nz
        const double ua2 = U[pad_idx(d, ic, jc, kn)], ub2 = U[pad_idx(d, ip, jc, kn)];
        const double va2 = V[pad_idx(d, ic, jc, kn)], vb2 = V[pad_idx(d, ic, jp, kn)];
        const double w1 = W[pad_idx(d, ic, jc, kn)];
        const double wxm1 = W[pad_idx(d, im, jc, kn)], wxp1 = W[pad_idx(d, ip, jc, kn)];
        const double wym1 = W[pad_idx(d, ic, jm, kn)], wyp1 = W[pad_idx(d, ic, jp, kn)];
        const double uym_a = U[pad_idx(d, ic, jm, k)], uym_b = U[pad_idx(d, ip, jm, k)];
        const double uyp_a = U[pad_idx(d, ic, jp, k)], uyp_b = U[pad_idx(d, ip, jp, k)];
        const double vxm_a = V[pad_idx(d, im, jc, k)], vxm_b = V[pad_idx(d, im, jp, k)];
        const double vxp_a = V[pad_idx(d, ip, jc, k)], vxp_b = V[pad_idx(d, ip, jp, k)];

        const double divv = (ub1 - ua1) * d.ix + (vb1 - va1) * d.iy + (w1 - w0) * d.iz;
        double omv = 0.0, hev = 0.0;
        if (colin && k >= 1 && k <= d.nz - 2) {
            const double uc = 0.5 * (ua1 + ub1), vc = 0.5 * (va1 + vb1), wc = 0.5 * (w0 + w1);
            const double wx = (0.5 * (wyp0 + wyp1) - 0.5 * (wym0 + wym1)) * hy -
                              (0.5 * (va2 + vb2) - 0.5 * (va0 + vb0)) * hz;
            const double wy = (0.5 * (ua2 + ub2) - 0.5 * (ua0 + ub0)) * hz -
                              (0.5 * (wxp0 + wxp1) - 0.5 * (wxm0 + wxm1)) * hx;
            const double wz = (0.5 * (vxp_a + vxp_b) - 0.5 * (vxm_a + vxm_b)) * hx -
                              (0.5 * (uyp_a + uyp_b) - 0.5 * (uym_a + uym_b)) * hy;
            omv = sqrt(wx * wx + wy * wy + wz * wz);
            hev = uc * wx + vc * wy + wc * wz;
        }
        if (active) {
            const long c = (long)i + (long)d.nx * ((long)j + (long)d.ny * k);
            dv[c] = divv;
            om[c] = omv;
            he[c] = hev;
            if (partials) {
                sd.add(divv);
                so.add(omv);
                sh.add(hev);
            }
        }
        ua0 = ua1, ub0 = ub1, va0 = va1, vb0 = vb1;
        ua1 = ua2, ub1 = ub2, va1 = va2, vb1 = vb2;
        w0 = w1, wxm0 = wxm1, wxp0 = wxp1, wym0 = wym1, wyp0 = wyp1;
    }
    if (!partials) return;  // uniform across the grid
    __shared__ double wv[3 * kBlock / 64];
    block_mms(sd, wv);
    block_mms(so, wv);
    block_mms(sh, wv);
    if (threadIdx.x == 0) {
        const MinMaxSum* f[3] = {&sd, &so, &sh};
        for (int q = 0; q < 3; ++q) {
            partials[(3 * q + 0) * d.nblocks + blk] = f[q]->lo;
            partials[(3 * q + 1) * d.nblocks + blk] = f[q]->hi;
            partials[(3 * q + 2) * d.nblocks + blk] = f[q]->sum;
        }
    }
}

// compute_field_stats first stage (stochastic.cu:369-414): min, max and sum of
// 2048 consecutive values per workgroup (8 per thread, strided by kBlock);
// partials[q * nblocks + block] for q = min, max, sum
__global__ void __launch_bounds__(kBlock)
    stats_kernel(const double* __restrict__ x, long n, double* __restrict__ partials) {
    constexpr int kPer = 8;
    const long base = (long)blockIdx.x * kPer * kBlock + threadIdx.x;
    double v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) v[u] = x[min(base + (long)u * kBlock, n - 1)];
    MinMaxSum a;
    a.init();
#pragma unroll
    for (int u = 0; u < kPer; ++u)
        if (base + (long)u * kBlock < n) a.add(v[u]);
    __shared__ double wv[3 * kBlock / 64];
    block_mms(a, wv);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = a.lo;
        partials[gridDim.x + blockIdx.x] = a.hi;
        partials[2 * (long)gridDim.x + blockIdx.x] = a.sum;
    }
}

// second stage: nq fields, each with (min, max, sum) partials of np blocks;
// out[3q..3q+2] = min, max, sum / n (fixed summation order)
__global__ void __launch_bounds__(kBlock)
    stats_final_kernel(const double* __restrict__ partials, long np, int nq, double n,
                       double* __restrict__ out) {
    __shared__ double wv[3 * kBlock / 64];
    for (int q = 0; q < nq; ++q) {
        const double* lo = partials + (3 * q + 0) * np;
        const double* hi = partials + (3 * q + 1) * np;
        const double* sm = partials + (3 * q + 2) * np;
        MinMaxSum a;
        a.init();
        for (long b = threadIdx.x; b < np; b += kBlock) {
            a.lo = fmin(a.lo, lo[b]);
            a.hi = fmax(a.hi, hi[b]);
            a.sum += sm[b];
        }
        block_mms(a, wv);
        if (threadIdx.x == 0) {
            out[3 * q + 0] = a.lo;
            out[3 * q + 1] = a.hi;
            out[3 * q + 2] = a.sum / n;
        }
    }
}

DiagDims diag_dims(const mf3d_grid& g) {
    DiagDims d{};
    d.nx = g.nx;
    d.ny = g.ny;
    d.nz = g.nz;
    d.tx = (g.nx + kDiagX - 1) / kDiagX;
    d.tiles = d.tx * ((g.ny + kDiagY - 1) / kDiagY);
    // enough plane chunks for ~2048 workgroups, at most 16 planes each
    const long want = 2048;
    d.kz = (int)std::max(1L, std::min(16L, (long)g.nz * d.tiles / want));
    d.nblocks = (long)d.tiles * ((g.nz + d.kz - 1) / d.kz);
    d.ix = 1.0 / g.dx;
    d.iy = 1.0 / g.dy;
    d.iz = 1.0 / g.dz;
    d.curl = g.nx >= 3 && g.ny >= 3 && g.nz >= 3;
    return d;
}

}  // namespace

namespace kern {

size_t diag_partials(const mf3d_grid& g) { return (size_t)kDiagQ * diag_dims(g).nblocks; }

int velocity_diagnostics(hipStream_t s, const mf3d_grid& g, const double* U, const double* V,
                         const double* W, double* div, double* omag, double* hel,
                         double* partials, double* stats) {
    const DiagDims d = diag_dims(g);
    diag_kernel<<<xcd_round((size_t)d.nblocks), kBlock, 0, s>>>(d, U, V, W, div, omag, hel,
                                                               stats ? partials : nullptr);
    MF3D_HIP_TRY(hipGetLastError());
    if (stats) {
        stats_final_kernel<<<1, kBlock, 0, s>>>(partials, d.nblocks, 3, (double)cells(g), stats);
        MF3D_HIP_TRY(hipGetLastError());
    }
    return MF3D_OK;
}

size_t stats_partials(size_t n) { return 3 * ((n + 2047) / 2048); }

int field_stats(hipStream_t s, const double* x, size_t n, double* partials, double* out) {
    const long nb = (long)((n + 2047) / 2048);
    stats_kernel<<<(unsigned)nb, kBlock, 0, s>>>(x, (long)n, partials);
    MF3D_HIP_TRY(hipGetLastError());
    stats_final_kernel<<<1, kBlock, 0, s>>>(partials, nb, 1, (double)n, out);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int velocity(hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc, const double* h,
             const double* K, double* U, double* V, double* W, bool padded) {
    const size_t n = cells(g);
    FaceIdx fi{g.nx, g.ny, g.nz, padded};
    velocity_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        fi, g.dx, g.dy, g.dz, bc, h, K, U, V, W);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace kern
}  // namespace mf3d
