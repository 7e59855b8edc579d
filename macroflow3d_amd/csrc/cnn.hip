// Learned warm start for the head solve: a 3-D convolutional encoder-decoder
// that maps ln K to an initial head, trained and run with BF16 MFMA kernels
// (FP32 accumulation) on gfx950. No counterpart in MacroFlow3D.
//
// The network runs on the grid coarsened 4x per axis (block statistics of K
// in, a smooth correction out): MG-PCG removes high-frequency error in one
// V-cycle, so only the smooth part of a guess shortens the solve, and the
// coarse grid makes inference 64x cheaper and the receptive field 4x wider.
// Network (3 levels at 4x, 8x and 16x coarsening, 16/32/64 channels, 3^3
// kernels, zero padding), on the coarse grid G (cell V = a 4x4x4 block):
//   x0  = [<lnK>/2, log<K>/2, -log<1/K>/2, 2 (I + 1/2)/NX - 1]  (4 channels, stored as 16)
//   e0  = relu(conv_a(x0))          4 -> 16  G
//   e1  = relu(conv_b(pool(e0)))   16 -> 32  G/2
//   e2  = relu(conv_c(pool(e1)))   32 -> 64  G/4
//   e2b = relu(conv_d(e2))         64 -> 64  G/4
//   d1  = relu(conv_e(up(e2b)) + e1)   64 -> 32  G/2
//   d0  = relu(conv_f(up(d1)) + e0)    32 -> 16  G
//   y   = conv_g(d0)               16 -> 1   G  (FP32)
//   h0  = h_lin + J trilinear(y),  h_lin linear between the x faces, J = (h_W - h_E) / nx
// (<.> = mean over the block, pool = mean of 2x2x2 children, up = copy to the
// 8 children). Training minimises mean over G of (y - <(h - h_lin) / J>)^2
// with Adam on FP32 master weights.
//
// Activations are BF16, channels-last: a[v * C + c], v = i + nx (j + ny k),
// C a multiple of 16. Convolution = implicit GEMM: for 16 output channels x
// 16 voxels, K runs over (tap, input channel) pairs, tap = dx + 3 dy + 9 dz
// (d = -1, 0, 1 offsets + 1), k = tap * Cin + c, 32 per MFMA.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace mf3d {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kVT = 4;  // 16-voxel tiles per wave in the convolution (large grids)
constexpr long kConvWaves = 2048;  // fewer tiles per wave below this many waves

__device__ __forceinline__ float bf2f(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ __hip_bfloat16 f2bf(float v) { return __float2bfloat16(v); }

struct Geo {
    int nx, ny, nz;
    __device__ long n() const { return (long)nx * ny * nz; }
};

__device__ __forceinline__ bf16x8 load8(const __hip_bfloat16* p) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, u);
}

// ---- convolution (forward, and backward-data with transposed weights) --------
// W: [Cout_pad][Kpad] BF16, Kpad = 32 * ceil(27 Cin / 32), zero beyond 27 Cin
// and for output channels >= Cout. FLIP reads the taps mirrored (backward
// data with the forward tap order in W). Epilogue, per output (v, c):
//   s = acc + bias[c] (if bias) + R[v*Cs + c] (if R); s = relu(s) (if relu);
//   s *= (M[v*Cs + c] > 0) (if M); stored to out (BF16 or FP32) for c < Cs.
template <int CT, int VT, bool F32OUT>
__global__ void __launch_bounds__(256)
    conv_kernel(Geo g, int cin, int kpad, int cs, const __hip_bfloat16* __restrict__ X,
                const __hip_bfloat16* __restrict__ W, const float* __restrict__ bias,
                const __hip_bfloat16* __restrict__ R, const __hip_bfloat16* __restrict__ M,
                int relu, int flip, void* __restrict__ out) {
    const long nvox = g.n();
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long vbase = wave * (16L * VT);
    if (vbase >= nvox) return;  // whole wave
    const int col = lane & 15;   // voxel within a tile / output channel row of A
    const int grp = lane >> 4;   // k group: k = 8 grp .. 8 grp + 7 of each 32-chunk

    int vi[VT], vj[VT], vk[VT];
    bool live[VT];
#pragma unroll
    for (int t = 0; t < VT; ++t) {
        const long v = vbase + 16 * t + col;
        live[t] = v < nvox;
        const long vv = live[t] ? v : 0;
        vi[t] = (int)(vv % g.nx);
        const long q = vv / g.nx;
        vj[t] = (int)(q % g.ny);
        vk[t] = (int)(q / g.ny);
    }
    f32x4 acc[VT][CT];
#pragma unroll
    for (int t = 0; t < VT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int K = 27 * cin;
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
    // the operands of one 32-wide k step: weight rows (A) and the tap's
    // input channels at each voxel (B, zero outside the grid or past K)
    auto load_step = [&](int k0, bf16x8 (&a)[CT], bf16x8 (&b)[VT]) {
        const int k = k0 + 8 * grp;
#pragma unroll
        for (int c = 0; c < CT; ++c) a[c] = load8(W + (long)(c * 16 + col) * kpad + k);
        if (k < K) {
            int tap = k / cin;
            const int c0 = k - tap * cin;
            if (flip) tap = 26 - tap;
            const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
#pragma unroll
            for (int t = 0; t < VT; ++t) {
                const int ii = vi[t] + dx, jj = vj[t] + dy, kk = vk[t] + dz;
                const bool in = live[t] && ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny &&
                                kk >= 0 && kk < g.nz;
                b[t] = in ? load8(X + ((long)ii + (long)g.nx * (jj + (long)g.ny * kk)) * cin + c0)
                          : zero8;
            }
        } else {
#pragma unroll
            for (int t = 0; t < VT; ++t) b[t] = zero8;
        }
    };
    // the next step's operands are loaded before this step's products are
    // issued, so the load latency overlaps the MFMAs (same accumulation order)
    bf16x8 a[CT], b[VT];
    load_step(0, a, b);
    for (int k0 = 0; k0 < kpad; k0 += 32) {
        bf16x8 an[CT], bn[VT];
        if (k0 + 32 < kpad) load_step(k0 + 32, an, bn);
#pragma unroll
        for (int t = 0; t < VT; ++t)
#pragma unroll
            for (int c = 0; c < CT; ++c)
                acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b[t], acc[t][c], 0, 0, 0);
        if (k0 + 32 < kpad) {
#pragma unroll
            for (int c = 0; c < CT; ++c) a[c] = an[c];
#pragma unroll
            for (int t = 0; t < VT; ++t) b[t] = bn[t];
        }
    }
    // acc[t][c][r] = out(channel c*16 + 4 grp + r, voxel vbase + 16 t + col)
#pragma unroll
    for (int t = 0; t < VT; ++t) {
        if (!live[t]) continue;
        const long v = vbase + 16 * t + col;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int ch0 = c * 16 + 4 * grp;
            if (ch0 >= cs) continue;
            float s[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float x = acc[t][c][r];
                if (bias) x += bias[ch0 + r];
                if (R) x += bf2f(R[v * cs + ch0 + r]);
                if (relu) x = fmaxf(x, 0.0f);
                if (M) x = bf2f(M[v * cs + ch0 + r]) > 0.0f ? x : 0.0f;
                s[r] = x;
            }
            if (F32OUT) {
                float* o = static_cast<float*>(out) + v * cs + ch0;
                *reinterpret_cast<float4*>(o) = make_float4(s[0], s[1], s[2], s[3]);
            } else {
                __hip_bfloat16* o = static_cast<__hip_bfloat16*>(out) + v * cs + ch0;
                o[0] = f2bf(s[0]);
                o[1] = f2bf(s[1]);
                o[2] = f2bf(s[2]);
                o[3] = f2bf(s[3]);
            }
        }
    }
}

// ---- convolution, weight gradient ---------------------------------------------
// dW[co][tap*Cin + ci] = sum_v dY[v][co] X[v + d(tap)][ci]; one wave per
// (tap, voxel slice) accumulates all Cout x Cin of that tap (CT x IT tiles)
// over its slice with the voxels as the MFMA K dimension; partial sums go to
// P[slice][co][tap*Cin + ci] (FP32), reduced by wgrad_reduce_kernel.
template <int CT, int IT>
__global__ void __launch_bounds__(64)
    wgrad_kernel(Geo g, int cin, int cout_ch, const __hip_bfloat16* __restrict__ dY,
                 const __hip_bfloat16* __restrict__ X, int slices, float* __restrict__ P) {
    const int tap = blockIdx.x % 27;
    const int slice = blockIdx.x / 27;
    const long nvox = g.n();
    const long per = ((nvox + slices - 1) / slices + 31) / 32 * 32;
    const long v0 = (long)slice * per;
    const long v1 = v0 + per < nvox ? v0 + per : nvox;
    const int lane = threadIdx.x;
    const int col = lane & 15, grp = lane >> 4;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    f32x4 acc[CT][IT];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int i = 0; i < IT; ++i) acc[c][i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long vb = v0; vb < v1; vb += 32) {
        // this lane's 8 voxels: vb + 8 grp + e
        long nb[8];
        bool okv[8], okn[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long v = vb + 8 * grp + e;
            okv[e] = v < v1;
            const long vv = okv[e] ? v : v0;
            const int i = (int)(vv % g.nx);
            const long q = vv / g.nx;
            const int j = (int)(q % g.ny), k = (int)(q / g.ny);
            const int ii = i + dx, jj = j + dy, kk = k + dz;
            okn[e] = okv[e] && ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz;
            nb[e] = okn[e] ? (long)ii + (long)g.nx * (jj + (long)g.ny * kk) : 0;
        }
        bf16x8 a[CT], b[IT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            alignas(16) __hip_bfloat16 tmp[8];
#pragma unroll
            for (int e = 0; e < 8; ++e)
                tmp[e] = okv[e] ? dY[(vb + 8 * grp + e) * cout_ch + c * 16 + col] : f2bf(0.0f);
            a[c] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(tmp));
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            alignas(16) __hip_bfloat16 tmp[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) tmp[e] = okn[e] ? X[nb[e] * cin + i * 16 + col] : f2bf(0.0f);
            b[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(tmp));
        }
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int i = 0; i < IT; ++i)
                acc[c][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c], b[i], acc[c][i], 0, 0, 0);
    }
    // acc[c][i][r] = dW(co = c*16 + 4 grp + r, ci = i*16 + col)
    const long K = 27L * cin;
    float* p = P + (long)slice * (CT * 16) * K;
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int i = 0; i < IT; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = c * 16 + 4 * grp + r;
                p[(long)co * K + (long)tap * cin + i * 16 + col] = acc[c][i][r];
            }
}

// G[co][k] (FP32, [Cout][27 Cin] of the layer) = sum over slices of P
__global__ void __launch_bounds__(256)
    wgrad_reduce_kernel(const float* __restrict__ P, int slices, int cout, int cout_pad, long K,
                        float* __restrict__ G) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)cout * K) return;
    const long co = e / K, k = e - co * K;
    float s = 0.0f;
    for (int sl = 0; sl < slices; ++sl) s += P[((long)sl * cout_pad + co) * K + k];
    G[e] = s;
}

// bias gradient db[c] = sum_v dY[v][c], one block per channel (deterministic)
__global__ void __launch_bounds__(256)
    bgrad_kernel(Geo g, int ch, int cout, const __hip_bfloat16* __restrict__ dY,
                 float* __restrict__ db) {
    const int c = blockIdx.x;
    if (c >= cout) return;
    const long n = g.n();
    float s = 0.0f;
    for (long v = threadIdx.x; v < n; v += blockDim.x) s += bf2f(dY[v * ch + c]);
    __shared__ float part[256];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = part[0];
}

// ---- pooling / upsampling (channels-last), factor kF per axis ------------------
constexpr int kS = 4;  // fine cells per network cell, per axis
constexpr int kF = 2;
constexpr float kInvF3 = 1.0f / (kF * kF * kF);

__device__ __forceinline__ long parent_of(const Geo& g, long v) {
    const int i = (int)(v % g.nx);
    const long q = v / g.nx;
    const int j = (int)(q % g.ny), k = (int)(q / g.ny);
    return (long)(i / kF) + (long)(g.nx / kF) * ((j / kF) + (long)(g.ny / kF) * (k / kF));
}

// sum over the kF^3 children of coarse cell V (fine g), channel c
__device__ __forceinline__ float child_sum(const Geo& g, int ch, const __hip_bfloat16* X, long V,
                                           int c) {
    const int NX = g.nx / kF, NY = g.ny / kF;
    const int I = (int)(V % NX);
    const long q = V / NX;
    const int J = (int)(q % NY), K = (int)(q / NY);
    float s = 0.0f;
    for (int dz = 0; dz < kF; ++dz)
        for (int dy = 0; dy < kF; ++dy)
            for (int dx = 0; dx < kF; ++dx)
                s += bf2f(X[((long)(kF * I + dx) +
                             (long)g.nx * ((kF * J + dy) + (long)g.ny * (kF * K + dz))) * ch + c]);
    return s;
}

// pool: Y[V][c] = mean of the children of coarse cell V (fine g)
__global__ void __launch_bounds__(256)
    pool_kernel(Geo g, int ch, const __hip_bfloat16* __restrict__ X, __hip_bfloat16* __restrict__ Y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.n() / (kF * kF * kF) * ch) return;
    Y[e] = f2bf(child_sum(g, ch, X, e / ch, (int)(e % ch)) * kInvF3);
}

// pool backward with the skip gradient and the ReLU mask of the pooled input:
// dX[v][c] = (dY[parent][c] / kF^3 + (S ? S[v][c] : 0)) * (M[v][c] > 0)
__global__ void __launch_bounds__(256)
    pool_bwd_kernel(Geo g, int ch, const __hip_bfloat16* __restrict__ dY,
                    const __hip_bfloat16* __restrict__ S, const __hip_bfloat16* __restrict__ M,
                    __hip_bfloat16* __restrict__ dX) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.n() * ch) return;
    const int c = (int)(e % ch);
    float s = bf2f(dY[parent_of(g, e / ch) * ch + c]) * kInvF3;
    if (S) s += bf2f(S[e]);
    if (M) s = bf2f(M[e]) > 0.0f ? s : 0.0f;
    dX[e] = f2bf(s);
}

// up: Y[v][c] = X[parent(v)][c] (fine g)
__global__ void __launch_bounds__(256)
    up_kernel(Geo g, int ch, const __hip_bfloat16* __restrict__ X, __hip_bfloat16* __restrict__ Y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.n() * ch) return;
    Y[e] = X[parent_of(g, e / ch) * ch + (e % ch)];
}

// up backward with the ReLU mask of the coarse input:
// dX[V][c] = sum of dY over the children, times (M[V][c] > 0)
__global__ void __launch_bounds__(256)
    up_bwd_kernel(Geo g, int ch, const __hip_bfloat16* __restrict__ dY,
                  const __hip_bfloat16* __restrict__ M, __hip_bfloat16* __restrict__ dX) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.n() / (kF * kF * kF) * ch) return;
    float s = child_sum(g, ch, dY, e / ch, (int)(e % ch));
    if (M) s = bf2f(M[e]) > 0.0f ? s : 0.0f;
    dX[e] = f2bf(s);
}

// ---- input, output, loss -------------------------------------------------------
// x0 on the coarse grid from the fine K (or ln K) of the kS^3 block of each cell:
// 0.5 mean ln K, 0.5 ln mean K, -0.5 ln mean 1/K, the x coordinate in [-1, 1].
// The features end in BF16, so they are accumulated in FP32. A workgroup holds
// 64 coarse cells x kS z-slices (one thread each); the slices' sums are added
// in LDS in a fixed order.
__global__ void __launch_bounds__(256)
    input_kernel(Geo g, const double* __restrict__ K, int is_log, __hip_bfloat16* __restrict__ x0) {
    static_assert(kS == 4, "64 cells x 4 slices per workgroup");
    const int NX = g.nx / kS, NY = g.ny / kS, NZ = g.nz / kS;
    const long nv = (long)NX * NY * NZ;
    const int lane = (int)(threadIdx.x & 63), dz = (int)(threadIdx.x >> 6);
    const long V = (long)blockIdx.x * 64 + lane;
    float sl = 0.0f, sk = 0.0f, si = 0.0f;
    int I = 0, J = 0, Kc = 0;
    if (V < nv) {
        I = (int)(V % NX);
        const long q = V / NX;
        J = (int)(q % NY);
        Kc = (int)(q / NY);
#pragma unroll
        for (int dy = 0; dy < kS; ++dy)
#pragma unroll
            for (int dx = 0; dx < kS; ++dx) {
                const long o = (long)(kS * I + dx) +
                               (long)g.nx * ((kS * J + dy) + (long)g.ny * (kS * Kc + dz));
                const float v = (float)K[o];
                const float l = is_log ? v : __logf(v);
                const float kv = is_log ? __expf(v) : v;
                sl += l;
                sk += kv;
                si += 1.0f / kv;
            }
    }
    __shared__ float part[3][kS][64];
    part[0][dz][lane] = sl;
    part[1][dz][lane] = sk;
    part[2][dz][lane] = si;
    __syncthreads();
    if (dz != 0 || V >= nv) return;
    sl = ((part[0][0][lane] + part[0][1][lane]) + part[0][2][lane]) + part[0][3][lane];
    sk = ((part[1][0][lane] + part[1][1][lane]) + part[1][2][lane]) + part[1][3][lane];
    si = ((part[2][0][lane] + part[2][1][lane]) + part[2][2][lane]) + part[2][3][lane];
    const float w = 1.0f / (kS * kS * kS);
    __hip_bfloat16* o = x0 + V * 16;
    o[0] = f2bf(0.5f * sl * w);
    o[1] = f2bf(0.5f * __logf(sk * w));
    o[2] = f2bf(-0.5f * __logf(si * w));
    o[3] = f2bf((float)(2.0 * (I + 0.5) / NX - 1.0));
#pragma unroll
    for (int c = 4; c < 16; ++c) o[c] = f2bf(0.0f);
}

// linear interpolation weights of fine index i on a coarse axis of n cells
// (cell centres at (I + 1/2) kS; constant beyond the outer centres)
__device__ __forceinline__ void lin_axis(int i, int n, int& i0, int& i1, float& w1) {
    const float u = (i + 0.5f) / kS - 0.5f;
    if (u <= 0.0f) {
        i0 = i1 = 0;
        w1 = 0.0f;
    } else if (u >= n - 1) {
        i0 = i1 = n - 1;
        w1 = 0.0f;
    } else {
        i0 = (int)u;
        i1 = i0 + 1;
        w1 = u - i0;
    }
}

__global__ void __launch_bounds__(256)
    channel0_kernel(const float* __restrict__ y, int ych, long n, float* __restrict__ yd) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) yd[v] = y[v * ych];
}

// h0 = h_lin + J trilinear(yd) on the fine grid g (yd: the coarse field's
// channel 0, stored densely). A thread writes 4 consecutive cells of a row
// (two 16-byte stores; the network's sizes are multiples of 16); the coarse
// values it interpolates, at most 3 along x for each of the 4 y/z corner rows,
// are read directly (the coarse field is 1/64 of the fine one and stays in L2).
__global__ void __launch_bounds__(256)
    output_kernel(Geo g, double hw, double he, double inv_nx, const float* __restrict__ yd,
                  double* __restrict__ h) {
    const int q4 = g.nx / 4;
    const long total = (long)q4 * g.ny * g.nz;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int q = (int)(t % q4);
    const long row = t / q4;
    const int j = (int)(row % g.ny), k = (int)(row / g.ny);
    const int NX = g.nx / kS, NY = g.ny / kS, NZ = g.nz / kS;
    int y0, y1, z0, z1;
    float wy, wz;
    lin_axis(j, NY, y0, y1, wy);
    lin_axis(k, NZ, z0, z1, wz);
    const float* r00 = yd + (long)NX * (y0 + (long)NY * z0);
    const float* r10 = yd + (long)NX * (y1 + (long)NY * z0);
    const float* r01 = yd + (long)NX * (y0 + (long)NY * z1);
    const float* r11 = yd + (long)NX * (y1 + (long)NY * z1);
    const double J = (hw - he) * inv_nx;
    double out[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = 4 * q + e;
        int x0, x1;
        float wx;
        lin_axis(i, NX, x0, x1, wx);
        const float c00 = r00[x0] * (1.f - wx) + r00[x1] * wx;
        const float c10 = r10[x0] * (1.f - wx) + r10[x1] * wx;
        const float c01 = r01[x0] * (1.f - wx) + r01[x1] * wx;
        const float c11 = r11[x0] * (1.f - wx) + r11[x1] * wx;
        const float c0 = c00 * (1.f - wy) + c10 * wy, c1 = c01 * (1.f - wy) + c11 * wy;
        const float yi = c0 * (1.f - wz) + c1 * wz;
        out[e] = hw + (he - hw) * ((i + 0.5) * inv_nx) + J * (double)yi;
    }
    double* o = h + row * g.nx + 4L * q;
    *reinterpret_cast<double2*>(o) = make_double2(out[0], out[1]);
    *reinterpret_cast<double2*>(o + 2) = make_double2(out[2], out[3]);
}

// per coarse cell V: target t = mean over its block of (h - h_lin) / J;
// dOut[V][0] = 2 (y - t) / N_coarse (channels 1..dch-1 zero); loss partials
__global__ void __launch_bounds__(256)
    loss_kernel(Geo g, double hw, double he, const float* __restrict__ y, int ych,
                const double* __restrict__ h, __hip_bfloat16* __restrict__ dOut, int dch,
                float* __restrict__ partials) {
    const int NX = g.nx / kS, NY = g.ny / kS, NZ = g.nz / kS;
    const long nc = (long)NX * NY * NZ;
    const long V = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.0f;
    if (V < nc) {
        const int I = (int)(V % NX);
        const long q = V / NX;
        const int J = (int)(q % NY), Kc = (int)(q / NY);
        const double Jg = (hw - he) / g.nx;
        double st = 0.0;
        for (int dz = 0; dz < kS; ++dz)
            for (int dy = 0; dy < kS; ++dy)
                for (int dx = 0; dx < kS; ++dx) {
                    const int i = kS * I + dx;
                    const long o = (long)i + (long)g.nx * ((kS * J + dy) + (long)g.ny * (kS * Kc + dz));
                    st += (h[o] - (hw + (he - hw) * (i + 0.5) / g.nx)) / Jg;
                }
        const float t = (float)(st / (kS * kS * kS));
        const float d = y[V * ych] - t;
        e2 = d * d;
        dOut[V * dch] = f2bf(2.0f * d / (float)nc);
        for (int c = 1; c < dch; ++c) dOut[V * dch + c] = f2bf(0.0f);
    }
    __shared__ float part[256];
    part[threadIdx.x] = e2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = part[0];
}

// ---- parameters ------------------------------------------------------------------
// master weights of a layer: w[co][tap][ci] (FP32, tap as above), bias[co]
// W  [Cout_pad][Kpad]: W[co][tap*Cin_pad + ci] = w[co][tap][ci] (zero padded)
// Wt [Cin_pad][Kpad_t]: Wt[ci][tap*Cout_pad + co] = w[co][tap][ci], used with
//   FLIP (tap mirrored) for the data gradient
__global__ void __launch_bounds__(256)
    pack_kernel(const float* __restrict__ w, int cin, int cout, int cin_pad, int cout_pad,
                int kpad, int kpad_t, __hip_bfloat16* __restrict__ W,
                __hip_bfloat16* __restrict__ Wt) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nW = (long)cout_pad * kpad, nWt = (long)cin_pad * kpad_t;
    if (e < nW) {
        const int co = (int)(e / kpad), k = (int)(e % kpad);
        const int tap = k / cin_pad, ci = k % cin_pad;
        float v = 0.0f;
        if (co < cout && tap < 27 && ci < cin) v = w[((long)co * 27 + tap) * cin + ci];
        W[e] = f2bf(v);
    }
    if (e < nWt) {
        const int ci = (int)(e / kpad_t), k = (int)(e % kpad_t);
        const int tap = k / cout_pad, co = k % cout_pad;
        float v = 0.0f;
        if (ci < cin && tap < 27 && co < cout) v = w[((long)co * 27 + tap) * cin + ci];
        Wt[e] = f2bf(v);
    }
}

// layer weight gradient in the master layout from G[co][tap*Cin_pad + ci]
__global__ void __launch_bounds__(256)
    unpack_grad_kernel(const float* __restrict__ G, int cin, int cout, int cin_pad,
                       float* __restrict__ gw) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)cout * 27 * cin) return;
    const int ci = (int)(e % cin);
    const long q = e / cin;
    const int tap = (int)(q % 27), co = (int)(q / 27);
    gw[e] = G[(long)co * 27 * cin_pad + (long)tap * cin_pad + ci];
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size
__global__ void __launch_bounds__(256)
    adam_kernel(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                float* __restrict__ v, long n, float lr_t, float b1, float b2, float eps) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float g = gr[e];
    const float mm = b1 * m[e] + (1.0f - b1) * g;
    const float vv = b2 * v[e] + (1.0f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr_t * mm / (sqrtf(vv) + eps);
}

inline unsigned blocks(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
namespace cnn {

struct LayerDef {
    int cin, cout, level;
};
// conv_a .. conv_g
constexpr LayerDef kLayers[7] = {{4, 16, 0},  {16, 32, 1}, {32, 64, 2}, {64, 64, 2},
                                 {64, 32, 1}, {32, 16, 0}, {16, 1, 0}};
constexpr int kNumLayers = 7;

inline int pad16(int c) { return (c + 15) & ~15; }
inline int kpad_of(int cin_pad) { return (27 * cin_pad + 31) / 32 * 32; }

long param_count() {
    long n = 0;
    for (const LayerDef& L : kLayers) n += 27L * L.cin * L.cout + L.cout;
    return n;
}

}  // namespace cnn
}  // namespace mf3d

using namespace mf3d;
using namespace mf3d::cnn;

struct mf3d_wsnet {
    float* params = nullptr;  // FP32 master: per layer w then bias
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    long nparam = 0;
    long step = 0;
    long woff[kNumLayers], boff[kNumLayers];
    int cin_pad[kNumLayers], cout_pad[kNumLayers], kpad[kNumLayers], kpad_t[kNumLayers];
    __hip_bfloat16* W[kNumLayers] = {};
    __hip_bfloat16* Wt[kNumLayers] = {};
    bool packed = false;
    // activations for the current grid
    int gx = 0, gy = 0, gz = 0;
    std::vector<void*> act;
    __hip_bfloat16 *x0 = nullptr, *e0 = nullptr, *p1 = nullptr, *e1 = nullptr, *p2 = nullptr,
                   *e2 = nullptr, *e2b = nullptr, *u1 = nullptr, *d1 = nullptr, *u0 = nullptr,
                   *d0 = nullptr;
    float* y = nullptr;  // [N0][16] FP32 (channel 0 used)
    float* yd = nullptr; // [N0] channel 0 of y, dense (read by the output interpolation)
    // backward scratch
    __hip_bfloat16 *g0a = nullptr, *g0b = nullptr, *g0c = nullptr, *g1a = nullptr,
                   *g1b = nullptr, *g2a = nullptr, *g2b = nullptr;
    float* wpart = nullptr;
    long wpart_cap = 0;
    float* G = nullptr;
    long G_cap = 0;
    float* lpart = nullptr;
    float* host_lpart = nullptr;
    long lpart_cap = 0;
};

namespace {

int net_alloc(void** p, size_t bytes) {
    MF3D_HIP_TRY(hipMalloc(p, bytes ? bytes : 8));
    return MF3D_OK;
}

void free_acts(mf3d_wsnet* net) {
    for (void* p : net->act) (void)hipFree(p);
    net->act.clear();
    net->gx = net->gy = net->gz = 0;
    if (net->wpart) (void)hipFree(net->wpart);
    if (net->G) (void)hipFree(net->G);
    if (net->lpart) (void)hipFree(net->lpart);
    if (net->host_lpart) (void)hipHostFree(net->host_lpart);
    net->wpart = net->G = net->lpart = net->host_lpart = nullptr;
    net->wpart_cap = net->G_cap = net->lpart_cap = 0;
}

int check_grid(const mf3d_grid* g) {
    MF3D_TRY(validate_grid(g));
    constexpr int q = kS * kF * kF;
    if ((g->nx % q) || (g->ny % q) || (g->nz % q))
        return fail(MF3D_ERR_UNSUPPORTED, "warm-start network needs sizes divisible by 16");
    return MF3D_OK;
}

constexpr int kSlices = 64;

int ensure_acts(mf3d_wsnet* net, const mf3d_grid& g) {
    if (net->gx == g.nx && net->gy == g.ny && net->gz == g.nz) return MF3D_OK;
    free_acts(net);
    const size_t n0 = cells(g), n1 = n0 / (kF * kF * kF), n2 = n1 / (kF * kF * kF);
    auto bf = [&](__hip_bfloat16** p, size_t elems) -> int {
        void* q = nullptr;
        MF3D_TRY(net_alloc(&q, elems * sizeof(__hip_bfloat16)));
        net->act.push_back(q);
        *p = static_cast<__hip_bfloat16*>(q);
        return MF3D_OK;
    };
    MF3D_TRY(bf(&net->x0, n0 * 16));
    MF3D_TRY(bf(&net->e0, n0 * 16));
    MF3D_TRY(bf(&net->p1, n1 * 16));
    MF3D_TRY(bf(&net->e1, n1 * 32));
    MF3D_TRY(bf(&net->p2, n2 * 32));
    MF3D_TRY(bf(&net->e2, n2 * 64));
    MF3D_TRY(bf(&net->e2b, n2 * 64));
    MF3D_TRY(bf(&net->u1, n1 * 64));
    MF3D_TRY(bf(&net->d1, n1 * 32));
    MF3D_TRY(bf(&net->u0, n0 * 32));
    MF3D_TRY(bf(&net->d0, n0 * 16));
    MF3D_TRY(bf(&net->g0a, n0 * 32));
    MF3D_TRY(bf(&net->g0b, n0 * 32));
    MF3D_TRY(bf(&net->g0c, n0 * 16));
    MF3D_TRY(bf(&net->g1a, n1 * 64));
    MF3D_TRY(bf(&net->g1b, n1 * 64));
    MF3D_TRY(bf(&net->g2a, n2 * 64));
    MF3D_TRY(bf(&net->g2b, n2 * 64));
    void* q = nullptr;
    MF3D_TRY(net_alloc(&q, n0 * 16 * sizeof(float)));
    net->act.push_back(q);
    net->y = static_cast<float*>(q);
    MF3D_TRY(net_alloc(&q, n0 * sizeof(float)));
    net->act.push_back(q);
    net->yd = static_cast<float*>(q);
    long wp = 0, gmax = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        wp = std::max(wp, (long)kSlices * net->cout_pad[l] * 27L * net->cin_pad[l]);
        gmax = std::max(gmax, (long)net->cout_pad[l] * 27L * net->cin_pad[l]);
    }
    MF3D_TRY(net_alloc((void**)&net->wpart, wp * sizeof(float)));
    MF3D_TRY(net_alloc((void**)&net->G, gmax * sizeof(float)));
    net->wpart_cap = wp;
    net->G_cap = gmax;
    const long nb = (long)blocks((long)n0);
    MF3D_TRY(net_alloc((void**)&net->lpart, nb * sizeof(float)));
    MF3D_HIP_TRY(hipHostMalloc((void**)&net->host_lpart, nb * sizeof(float), hipHostMallocDefault));
    net->lpart_cap = nb;
    net->gx = g.nx;
    net->gy = g.ny;
    net->gz = g.nz;
    return MF3D_OK;
}

int pack(mf3d_wsnet* net, hipStream_t s) {
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& L = kLayers[l];
        const long n = std::max((long)net->cout_pad[l] * net->kpad[l],
                                (long)net->cin_pad[l] * net->kpad_t[l]);
        pack_kernel<<<blocks(n), 256, 0, s>>>(net->params + net->woff[l], L.cin, L.cout,
                                              net->cin_pad[l], net->cout_pad[l], net->kpad[l],
                                              net->kpad_t[l], net->W[l], net->Wt[l]);
    }
    MF3D_HIP_TRY(hipGetLastError());
    net->packed = true;
    return MF3D_OK;
}

// one convolution: out[v][0..cs) from X[v][0..cin_ch) with W (cout_pad rows)
int conv(hipStream_t s, const mf3d_grid& g, int cin_ch, int cout_pad, int kp, int cs,
         const __hip_bfloat16* X, const __hip_bfloat16* W, const float* bias,
         const __hip_bfloat16* R, const __hip_bfloat16* M, bool relu, bool flip, void* out,
         bool f32out) {
    const Geo G{g.nx, g.ny, g.nz};
    // 16-voxel tiles per wave: kVT on large grids, fewer where that would
    // leave the launch with fewer than kConvWaves waves (the small grids of
    // the deeper layers; same products and order per output either way)
    const long nv = (long)cells(g);
    const int vt = (nv + 16 * kVT - 1) / (16 * kVT) >= kConvWaves ? kVT
                   : (nv + 31) / 32 >= kConvWaves                   ? 2
                                                                    : 1;
    const long waves = (nv + 16L * vt - 1) / (16L * vt);
    const unsigned nb = (unsigned)((waves + 3) / 4);
    const int ct = cout_pad / 16;
#define MF3D_CONV(CTV, VTV, F32)                                                               \
    conv_kernel<CTV, VTV, F32><<<nb, 256, 0, s>>>(G, cin_ch, kp, cs, X, W, bias, R, M,        \
                                                  relu ? 1 : 0, flip ? 1 : 0, out)
#define MF3D_CONV_VT(CTV, F32)                                                                 \
    do {                                                                                       \
        if (vt == kVT) MF3D_CONV(CTV, kVT, F32);                                               \
        else if (vt == 2) MF3D_CONV(CTV, 2, F32);                                              \
        else MF3D_CONV(CTV, 1, F32);                                                           \
    } while (0)
    if (f32out) {
        if (ct != 1) return fail(MF3D_ERR_UNSUPPORTED, "FP32 output only for 16 channels");
        MF3D_CONV_VT(1, true);
    } else if (ct == 1) MF3D_CONV_VT(1, false);
    else if (ct == 2) MF3D_CONV_VT(2, false);
    else if (ct == 4) MF3D_CONV_VT(4, false);
    else return fail(MF3D_ERR_UNSUPPORTED, "convolution: unsupported channel count");
#undef MF3D_CONV_VT
#undef MF3D_CONV
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

// weight gradient (G[co][tap*cin_ch + ci], FP32) of a layer from dY (cout_ch
// channels) and its input X (cin_ch channels)
int wgrad(mf3d_wsnet* net, hipStream_t s, const mf3d_grid& g, int cin_ch, int cout_ch,
          int cout, const __hip_bfloat16* dY, const __hip_bfloat16* X, float* G) {
    const Geo geo{g.nx, g.ny, g.nz};
    const int ct = cout_ch / 16, it = cin_ch / 16;
    const long K = 27L * cin_ch;
    if ((long)kSlices * cout_ch * K > net->wpart_cap)
        return fail(MF3D_ERR_UNSUPPORTED, "weight-gradient scratch too small");
    const unsigned nb = 27 * kSlices;
#define MF3D_WG(C, I) wgrad_kernel<C, I><<<nb, 64, 0, s>>>(geo, cin_ch, cout_ch, dY, X, kSlices, net->wpart)
    if (ct == 1 && it == 1) MF3D_WG(1, 1);
    else if (ct == 1 && it == 2) MF3D_WG(1, 2);
    else if (ct == 2 && it == 1) MF3D_WG(2, 1);
    else if (ct == 2 && it == 4) MF3D_WG(2, 4);
    else if (ct == 4 && it == 2) MF3D_WG(4, 2);
    else if (ct == 4 && it == 4) MF3D_WG(4, 4);
    else return fail(MF3D_ERR_UNSUPPORTED, "weight gradient: unsupported channel counts");
#undef MF3D_WG
    wgrad_reduce_kernel<<<blocks((long)cout * K), 256, 0, s>>>(net->wpart, kSlices, cout, cout_ch,
                                                               K, G);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

mf3d_grid coarse(const mf3d_grid& g, int f) {
    mf3d_grid c = g;
    c.nx /= f;
    c.ny /= f;
    c.nz /= f;
    return c;
}

// forward pass for the fine grid gf (the network runs on gf / kS)
int forward(mf3d_wsnet* net, hipStream_t s, const mf3d_grid& gf, const double* K, int is_log) {
    const mf3d_grid g = coarse(gf, kS);
    MF3D_TRY(ensure_acts(net, g));
    if (!net->packed) MF3D_TRY(pack(net, s));
    const mf3d_grid g1 = coarse(g, kF), g2 = coarse(g, kF * kF);
    const Geo G0{g.nx, g.ny, g.nz}, G1{g1.nx, g1.ny, g1.nz};
    const long n0 = (long)cells(g), n1 = (long)cells(g1), n2 = (long)cells(g2);
    float* P = net->params;
    input_kernel<<<(unsigned)((n0 + 63) / 64), 256, 0, s>>>(Geo{gf.nx, gf.ny, gf.nz}, K, is_log,
                                                          net->x0);
    MF3D_TRY(conv(s, g, 16, 16, net->kpad[0], 16, net->x0, net->W[0], P + net->boff[0], nullptr,
                  nullptr, true, false, net->e0, false));
    pool_kernel<<<blocks(n1 * 16), 256, 0, s>>>(G0, 16, net->e0, net->p1);
    MF3D_TRY(conv(s, g1, 16, 32, net->kpad[1], 32, net->p1, net->W[1], P + net->boff[1], nullptr,
                  nullptr, true, false, net->e1, false));
    pool_kernel<<<blocks(n2 * 32), 256, 0, s>>>(G1, 32, net->e1, net->p2);
    MF3D_TRY(conv(s, g2, 32, 64, net->kpad[2], 64, net->p2, net->W[2], P + net->boff[2], nullptr,
                  nullptr, true, false, net->e2, false));
    MF3D_TRY(conv(s, g2, 64, 64, net->kpad[3], 64, net->e2, net->W[3], P + net->boff[3], nullptr,
                  nullptr, true, false, net->e2b, false));
    up_kernel<<<blocks(n1 * 64), 256, 0, s>>>(Geo{g1.nx, g1.ny, g1.nz}, 64, net->e2b, net->u1);
    MF3D_TRY(conv(s, g1, 64, 32, net->kpad[4], 32, net->u1, net->W[4], P + net->boff[4], net->e1,
                  nullptr, true, false, net->d1, false));
    up_kernel<<<blocks(n0 * 32), 256, 0, s>>>(G0, 32, net->d1, net->u0);
    MF3D_TRY(conv(s, g, 32, 16, net->kpad[5], 16, net->u0, net->W[5], P + net->boff[5], net->e0,
                  nullptr, true, false, net->d0, false));
    MF3D_TRY(conv(s, g, 16, 16, net->kpad[6], 16, net->d0, net->W[6], P + net->boff[6], nullptr,
                  nullptr, false, false, net->y, true));
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

// loss and all parameter gradients (net->grads) for target head h
int backward(mf3d_wsnet* net, hipStream_t s, const mf3d_grid& gf, double hw, double he,
             const double* h, double* loss) {
    const mf3d_grid g = coarse(gf, kS);
    const mf3d_grid g1 = coarse(g, kF), g2 = coarse(g, kF * kF);
    const Geo G0{g.nx, g.ny, g.nz}, G1{g1.nx, g1.ny, g1.nz};
    const long n0 = (long)cells(g), n1 = (long)cells(g1), n2 = (long)cells(g2);
    const unsigned lb = blocks(n0);
    float* GR = net->grads;
    // dY of conv_g (16 channels, channel 0 live) -> g0c
    loss_kernel<<<lb, 256, 0, s>>>(Geo{gf.nx, gf.ny, gf.nz}, hw, he, net->y, 16, h, net->g0c, 16,
                                   net->lpart);
    auto grads_of = [&](int l, const mf3d_grid& gg, int cin_ch, int cout_ch,
                        const __hip_bfloat16* dY, const __hip_bfloat16* X) -> int {
        const LayerDef& L = kLayers[l];
        MF3D_TRY(wgrad(net, s, gg, cin_ch, cout_ch, L.cout, dY, X, net->G));
        unpack_grad_kernel<<<blocks(27L * L.cin * L.cout), 256, 0, s>>>(net->G, L.cin, L.cout,
                                                                       cin_ch, GR + net->woff[l]);
        bgrad_kernel<<<L.cout, 256, 0, s>>>(Geo{gg.nx, gg.ny, gg.nz}, cout_ch, L.cout, dY,
                                            GR + net->boff[l]);
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    };
    // conv_g: input d0 (16); data gradient masked by d0 > 0 -> g0b (16 ch)
    MF3D_TRY(grads_of(6, g, 16, 16, net->g0c, net->d0));
    MF3D_TRY(conv(s, g, 16, 16, net->kpad_t[6], 16, net->g0c, net->Wt[6], nullptr, nullptr,
                  net->d0, false, true, net->g0b, false));
    // conv_f (dPre_f = g0b): input u0 (32); skip gradient to e0 = g0b
    MF3D_TRY(grads_of(5, g, 32, 16, net->g0b, net->u0));
    MF3D_TRY(conv(s, g, 16, 32, net->kpad_t[5], 32, net->g0b, net->Wt[5], nullptr, nullptr,
                  nullptr, false, true, net->g0a, false));  // du0 (32 ch)
    // up bwd into d1, masked by d1 > 0 -> dPre_e (g1a, 32 ch)
    up_bwd_kernel<<<blocks(n1 * 32), 256, 0, s>>>(G0, 32, net->g0a, net->d1, net->g1a);
    MF3D_TRY(grads_of(4, g1, 64, 32, net->g1a, net->u1));
    MF3D_TRY(conv(s, g1, 32, 64, net->kpad_t[4], 64, net->g1a, net->Wt[4], nullptr, nullptr,
                  nullptr, false, true, net->g1b, false));  // du1 (64 ch)
    // up bwd into e2b, masked -> dPre_d (g2a)
    up_bwd_kernel<<<blocks(n2 * 64), 256, 0, s>>>(G1, 64, net->g1b, net->e2b, net->g2a);
    MF3D_TRY(grads_of(3, g2, 64, 64, net->g2a, net->e2));
    MF3D_TRY(conv(s, g2, 64, 64, net->kpad_t[3], 64, net->g2a, net->Wt[3], nullptr, nullptr,
                  net->e2, false, true, net->g2b, false));  // dPre_c (64 ch)
    MF3D_TRY(grads_of(2, g2, 32, 64, net->g2b, net->p2));
    MF3D_TRY(conv(s, g2, 64, 32, net->kpad_t[2], 32, net->g2b, net->Wt[2], nullptr, nullptr,
                  nullptr, false, true, net->g2a, false));  // dp2 (32 ch)
    // pool bwd into e1 + skip gradient (g1a), masked by e1 -> dPre_b (g1b, 32 ch)
    pool_bwd_kernel<<<blocks(n1 * 32), 256, 0, s>>>(G1, 32, net->g2a, net->g1a, net->e1, net->g1b);
    MF3D_TRY(grads_of(1, g1, 16, 32, net->g1b, net->p1));
    MF3D_TRY(conv(s, g1, 32, 16, net->kpad_t[1], 16, net->g1b, net->Wt[1], nullptr, nullptr,
                  nullptr, false, true, net->g1a, false));  // dp1 (16 ch)
    // pool bwd into e0 + skip gradient (g0b), masked by e0 -> dPre_a (g0c, 16 ch)
    pool_bwd_kernel<<<blocks(n0 * 16), 256, 0, s>>>(G0, 16, net->g1a, net->g0b, net->e0, net->g0c);
    MF3D_TRY(grads_of(0, g, 16, 16, net->g0c, net->x0));
    if (loss) {
        MF3D_HIP_TRY(hipMemcpyAsync(net->host_lpart, net->lpart, lb * sizeof(float),
                                    hipMemcpyDeviceToHost, s));
        MF3D_HIP_TRY(hipStreamSynchronize(s));
        double t = 0.0;
        for (unsigned b = 0; b < lb; ++b) t += net->host_lpart[b];
        *loss = t / (double)n0;
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int x_dirichlet(const mf3d_bc* bc, double* hw, double* he) {
    if (bc->type[0] != MF3D_BC_DIRICHLET || bc->type[1] != MF3D_BC_DIRICHLET)
        return fail(MF3D_ERR_UNSUPPORTED, "warm start needs Dirichlet x faces");
    if (bc->value[0] == bc->value[1])
        return fail(MF3D_ERR_UNSUPPORTED, "warm start needs h_west != h_east");
    *hw = bc->value[0];
    *he = bc->value[1];
    return MF3D_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int mf3d_wsnet_destroy(mf3d_wsnet* net);

long mf3d_wsnet_param_count(void) { return param_count(); }

int mf3d_wsnet_create(mf3d_wsnet** out) {
    if (!out) return fail(MF3D_ERR_INVALID_ARGUMENT, "null output pointer");
    mf3d_wsnet* net = new mf3d_wsnet();
    long off = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        const LayerDef& L = kLayers[l];
        net->woff[l] = off;
        off += 27L * L.cin * L.cout;
        net->boff[l] = off;
        off += L.cout;
        net->cin_pad[l] = pad16(L.cin);
        net->cout_pad[l] = pad16(L.cout);
        net->kpad[l] = kpad_of(net->cin_pad[l]);
        net->kpad_t[l] = kpad_of(net->cout_pad[l]);
    }
    net->nparam = off;
    int st = MF3D_OK;
    auto al = [&](float** p) {
        if (st == MF3D_OK) st = net_alloc((void**)p, off * sizeof(float));
    };
    al(&net->params);
    al(&net->grads);
    al(&net->adam_m);
    al(&net->adam_v);
    for (int l = 0; l < kNumLayers && st == MF3D_OK; ++l) {
        st = net_alloc((void**)&net->W[l], (size_t)net->cout_pad[l] * net->kpad[l] * 2);
        if (st == MF3D_OK)
            st = net_alloc((void**)&net->Wt[l], (size_t)pad16(net->cin_pad[l]) * net->kpad_t[l] * 2);
    }
    if (st == MF3D_OK) {
        hipError_t e = hipMemset(net->params, 0, off * sizeof(float));
        if (e == hipSuccess) e = hipMemset(net->adam_m, 0, off * sizeof(float));
        if (e == hipSuccess) e = hipMemset(net->adam_v, 0, off * sizeof(float));
        if (e != hipSuccess) st = fail(MF3D_ERR_HIP, hipGetErrorString(e));
    }
    if (st != MF3D_OK) {
        mf3d_wsnet_destroy(net);
        return st;
    }
    *out = net;
    return MF3D_OK;
}

int mf3d_wsnet_destroy(mf3d_wsnet* net) {
    if (!net) return MF3D_OK;
    free_acts(net);
    for (float* p : {net->params, net->grads, net->adam_m, net->adam_v})
        if (p) (void)hipFree(p);
    for (int l = 0; l < kNumLayers; ++l) {
        if (net->W[l]) (void)hipFree(net->W[l]);
        if (net->Wt[l]) (void)hipFree(net->Wt[l]);
    }
    delete net;
    return MF3D_OK;
}

int mf3d_wsnet_set_params(mf3d_wsnet* net, const float* host, long n) {
    if (!net || !host) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    if (n != net->nparam) return fail(MF3D_ERR_SIZE_MISMATCH, "parameter count");
    MF3D_HIP_TRY(hipMemcpy(net->params, host, n * sizeof(float), hipMemcpyHostToDevice));
    MF3D_HIP_TRY(hipMemset(net->adam_m, 0, n * sizeof(float)));
    MF3D_HIP_TRY(hipMemset(net->adam_v, 0, n * sizeof(float)));
    net->step = 0;
    net->packed = false;
    return MF3D_OK;
}

int mf3d_wsnet_get_params(const mf3d_wsnet* net, float* host, long n) {
    if (!net || !host) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    if (n != net->nparam) return fail(MF3D_ERR_SIZE_MISMATCH, "parameter count");
    MF3D_HIP_TRY(hipMemcpy(host, net->params, n * sizeof(float), hipMemcpyDeviceToHost));
    return MF3D_OK;
}

int mf3d_wsnet_get_grads(const mf3d_wsnet* net, float* host, long n) {
    if (!net || !host) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    if (n != net->nparam) return fail(MF3D_ERR_SIZE_MISMATCH, "parameter count");
    MF3D_HIP_TRY(hipMemcpy(host, net->grads, n * sizeof(float), hipMemcpyDeviceToHost));
    return MF3D_OK;
}

int mf3d_wsnet_predict(mf3d_wsnet* net, mf3d_stream stream, const mf3d_grid* grid,
                       const mf3d_bc* bc, const double* K, int k_is_log, double* h0, size_t n) {
    if (!net || !grid || !bc || !K || !h0) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    MF3D_TRY(check_grid(grid));
    if (n < cells(*grid)) return fail(MF3D_ERR_SIZE_MISMATCH, "h0/K shorter than the grid");
    double hw, he;
    MF3D_TRY(x_dirichlet(bc, &hw, &he));
    (void)hipGetLastError();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    MF3D_TRY(forward(net, s, *grid, K, k_is_log));
    // channel 0 densely first: the interpolation's 8 gathers per cell then
    // share cache lines between neighbouring cells
    const long n0 = cells(*grid) / ((long)kS * kS * kS);
    channel0_kernel<<<blocks(n0), 256, 0, s>>>(net->y, 16, n0, net->yd);
    // h0 must be 16-byte aligned for the paired stores (the network's sizes
    // make every row start so)
    if (reinterpret_cast<uintptr_t>(h0) & 15)
        return fail(MF3D_ERR_UNSUPPORTED, "warm-start output needs a 16-byte aligned head");
    const long quads = cells(*grid) / 4;
    output_kernel<<<(unsigned)((quads + 255) / 256), 256, 0, s>>>(
        Geo{grid->nx, grid->ny, grid->nz}, hw, he, 1.0 / grid->nx, net->yd, h0);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int mf3d_wsnet_gradients(mf3d_wsnet* net, mf3d_stream stream, const mf3d_grid* grid,
                         const mf3d_bc* bc, const double* K, int k_is_log, const double* h,
                         size_t n, double* loss) {
    if (!net || !grid || !bc || !K || !h) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    MF3D_TRY(check_grid(grid));
    if (n < cells(*grid)) return fail(MF3D_ERR_SIZE_MISMATCH, "h/K shorter than the grid");
    double hw, he;
    MF3D_TRY(x_dirichlet(bc, &hw, &he));
    (void)hipGetLastError();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    MF3D_TRY(forward(net, s, *grid, K, k_is_log));
    return backward(net, s, *grid, hw, he, h, loss);
}

int mf3d_wsnet_train_step(mf3d_wsnet* net, mf3d_stream stream, const mf3d_grid* grid,
                          const mf3d_bc* bc, const double* K, int k_is_log, const double* h,
                          size_t n, double lr, double* loss) {
    MF3D_TRY(mf3d_wsnet_gradients(net, stream, grid, bc, K, k_is_log, h, n, loss));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float b1 = 0.9f, b2 = 0.999f;
    net->step += 1;
    const double c1 = 1.0 - std::pow((double)b1, (double)net->step);
    const double c2 = 1.0 - std::pow((double)b2, (double)net->step);
    const float lr_t = (float)(lr * std::sqrt(c2) / c1);
    adam_kernel<<<blocks(net->nparam), 256, 0, s>>>(net->params, net->grads, net->adam_m,
                                                    net->adam_v, net->nparam, lr_t, b1, b2, 1e-8f);
    MF3D_HIP_TRY(hipGetLastError());
    net->packed = false;
    return MF3D_OK;
}

// single convolution layer for tests: X [N][cin] BF16 (cin a multiple of 8),
// w [cout][27][cin] FP32, bias [cout] FP32 (nullable), R [N][cout] (nullable,
// added before the ReLU), relu; Y [N][cout] BF16 (cout a multiple of 16, <= 64)
int mf3d_conv3d_forward(mf3d_stream stream, const mf3d_grid* grid, int cin, int cout,
                        const void* X, const float* w, const float* bias, const void* R,
                        int relu, void* Y) {
    if (!grid || !X || !w || !Y) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    MF3D_TRY(validate_grid(grid));
    if (cin % 8 || cin > 128 || cout % 16 || cout > 64)
        return fail(MF3D_ERR_UNSUPPORTED, "cin must be a multiple of 8, cout 16/32/64");
    (void)hipGetLastError();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int kp = kpad_of(cin), kpt = kpad_of(cout);
    __hip_bfloat16 *W = nullptr, *Wt = nullptr;
    MF3D_TRY(net_alloc((void**)&W, (size_t)cout * kp * 2));
    int st = net_alloc((void**)&Wt, (size_t)pad16(cin) * kpt * 2);
    if (st == MF3D_OK) {
        pack_kernel<<<blocks(std::max((long)cout * kp, (long)pad16(cin) * kpt)), 256, 0, s>>>(
            w, cin, cout, cin, cout, kp, kpt, W, Wt);
        st = conv(s, *grid, cin, cout, kp, cout, (const __hip_bfloat16*)X, W, bias,
                  (const __hip_bfloat16*)R, nullptr, relu != 0, false, Y, false);
    }
    if (st == MF3D_OK) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(MF3D_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(W);
    if (Wt) (void)hipFree(Wt);
    return st;
}

// data gradient of a convolution: dX [N][cin] = sum over taps/cout of w and
// dY [N][cout], times (M > 0) if M [N][cin] is given (the ReLU of X)
int mf3d_conv3d_backward_data(mf3d_stream stream, const mf3d_grid* grid, int cin, int cout,
                              const void* dY, const float* w, const void* M, void* dX) {
    if (!grid || !dY || !w || !dX) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    MF3D_TRY(validate_grid(grid));
    if (cin % 16 || cin > 64 || cout % 16 || cout > 64)
        return fail(MF3D_ERR_UNSUPPORTED, "cin and cout must be 16/32/64");
    (void)hipGetLastError();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int kp = kpad_of(cin), kpt = kpad_of(cout);
    __hip_bfloat16 *W = nullptr, *Wt = nullptr;
    MF3D_TRY(net_alloc((void**)&W, (size_t)cout * kp * 2));
    int st = net_alloc((void**)&Wt, (size_t)cin * kpt * 2);
    if (st == MF3D_OK) {
        pack_kernel<<<blocks(std::max((long)cout * kp, (long)cin * kpt)), 256, 0, s>>>(
            w, cin, cout, cin, cout, kp, kpt, W, Wt);
        st = conv(s, *grid, cout, cin, kpt, cin, (const __hip_bfloat16*)dY, Wt, nullptr, nullptr,
                  (const __hip_bfloat16*)M, false, true, dX, false);
    }
    if (st == MF3D_OK) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(MF3D_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(W);
    if (Wt) (void)hipFree(Wt);
    return st;
}

// weight and bias gradients: gw [cout][27][cin] = sum_v dY[v][co] X[v + d][ci],
// gb [cout] = sum_v dY[v][co] (FP32)
int mf3d_conv3d_backward_weights(mf3d_stream stream, const mf3d_grid* grid, int cin, int cout,
                                 const void* X, const void* dY, float* gw, float* gb) {
    if (!grid || !X || !dY || !gw || !gb) return fail(MF3D_ERR_INVALID_ARGUMENT, "null pointer");
    MF3D_TRY(validate_grid(grid));
    if (cin % 16 || cin > 64 || cout % 16 || cout > 64)
        return fail(MF3D_ERR_UNSUPPORTED, "cin and cout must be 16/32/64");
    (void)hipGetLastError();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    mf3d_wsnet tmp;
    const long K = 27L * cin;
    tmp.wpart_cap = (long)kSlices * cout * K;
    MF3D_TRY(net_alloc((void**)&tmp.wpart, tmp.wpart_cap * sizeof(float)));
    float* G = nullptr;
    int st = net_alloc((void**)&G, (size_t)cout * K * sizeof(float));
    if (st == MF3D_OK)
        st = wgrad(&tmp, s, *grid, cin, cout, cout, (const __hip_bfloat16*)dY,
                   (const __hip_bfloat16*)X, G);
    if (st == MF3D_OK) {
        unpack_grad_kernel<<<blocks((long)cout * K), 256, 0, s>>>(G, cin, cout, cin, gw);
        bgrad_kernel<<<cout, 256, 0, s>>>(Geo{grid->nx, grid->ny, grid->nz}, cout, cout,
                                          (const __hip_bfloat16*)dY, gb);
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = fail(MF3D_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(tmp.wpart);
    tmp.wpart = nullptr;
    if (G) (void)hipFree(G);
    return st;
}

}  // extern "C"
