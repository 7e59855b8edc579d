// Face-coefficient kernels of the solve path (FP64 and mixed precision), gfx950.
//
// Each multigrid level stores its harmonic face conductances once per solve,
// on the MAC face grid of the compact velocity layout:
//   FX[i + (nx+1)(j + ny k)]  face between cells i-1 and i   (i = 0..nx)
//   FY[i + nx(j + (ny+1) k)]  face between cells j-1 and j   (j = 0..ny)
//   FZ[i + nx(j + ny k)]      face between cells k-1 and k   (k = 0..nz)
// Interior faces hold 2/(1/K_a + 1/K_b) (src/multigrid/smoothers/gsrb_3d.cu:53-58);
// boundary faces hold the bc_helpers coefficient (bc_stencil_helpers.cuh:33-267):
// Dirichlet 2 K_cell, Neumann 0, periodic the harmonic mean across the wrap.
// A boundary face's neighbour value is 0 unless the axis is periodic.
//
// With the coefficients precomputed, the smoother needs one divide per cell
// (the reference recomputes six harmonic means, ~13 FP64 divides, per update)
// and the operator and residual need none.
//
// T is the arithmetic type of x and b (double, or float for the mixed-precision
// V-cycle); TC is the coefficient storage type (double, float, bf16, fp16).
// Stored coefficients are multiplied by a per-level power of two `cs` (chosen
// so that fp16 keeps its range); the kernels take bs = cs * dx^2 and
// rs = 1 / (cs * dx^2) so every result is unchanged by the scaling.

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

#include "faces.hpp"
#include "hface.hpp"
#include "reduce.hpp"

namespace mf3d {
namespace {

constexpr int kBlock = 256;
inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// grid of a 1-D launch remapped by xcd_block: padded to a multiple of 8 blocks
inline unsigned xcd_grid(size_t n) { return xcd_round(grid_for(n)); }

// ---- coefficient storage conversions ---------------------------------------
__device__ __forceinline__ double cload(const double* p, long i) { return p[i]; }
__device__ __forceinline__ float cload(const float* p, long i) { return p[i]; }
__device__ __forceinline__ float cload(const __hip_bfloat16* p, long i) {
    return __bfloat162float(p[i]);
}
__device__ __forceinline__ float cload(const __half* p, long i) { return __half2float(p[i]); }
__device__ __forceinline__ void cstore(double* p, long i, double v) { p[i] = v; }
__device__ __forceinline__ void cstore(float* p, long i, double v) { p[i] = (float)v; }
__device__ __forceinline__ void cstore(__hip_bfloat16* p, long i, double v) {
    p[i] = __float2bfloat16((float)v);
}
__device__ __forceinline__ void cstore(__half* p, long i, double v) { p[i] = __float2half((float)v); }

struct Dims {
    int nx, ny, nz;
    __device__ long fx(int i, int j, int k) const {
        return (long)i + (long)(nx + 1) * ((long)j + (long)ny * k);
    }
    __device__ long fy(int i, int j, int k) const {
        return (long)i + (long)nx * ((long)j + (long)(ny + 1) * k);
    }
    __device__ long fz(int i, int j, int k) const {
        return (long)i + (long)nx * ((long)j + (long)ny * k);
    }
};

__device__ __forceinline__ double bface(int t, double Kc) {
    return t == MF3D_BC_DIRICHLET ? 2.0 * Kc : 0.0;
}

// ---- face coefficients from K ------------------------------------------------
template <typename TC>
__global__ void __launch_bounds__(kBlock)
    faces_build_kernel(Dims d, BcTypes bc, double cs, const double* __restrict__ K,
                       TC* __restrict__ FX, TC* __restrict__ FY, TC* __restrict__ FZ) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long n = (long)nx * ny * nz;
    const long blk = xcd_block((n + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long o = blk * kBlock + threadIdx.x;
    if (o >= n) return;
    const int i = (int)(o % nx);
    const long q = o / nx;
    const int j = (int)(q % ny);
    const int k = (int)(q / ny);
    const long sy = nx, sz = (long)nx * ny;
    const double Kc = K[o];
    const double iKc = 1.0 / Kc;
    // lower faces
    double f;
    if (i > 0) f = 2.0 / (iKc + 1.0 / K[o - 1]);
    else f = bc.t[0] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / K[o + nx - 1]) : bface(bc.t[0], Kc);
    cstore(FX, d.fx(i, j, k), f * cs);
    if (j > 0) f = 2.0 / (iKc + 1.0 / K[o - sy]);
    else f = bc.t[2] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / K[o + (long)(ny - 1) * sy])
                                         : bface(bc.t[2], Kc);
    cstore(FY, d.fy(i, j, k), f * cs);
    if (k > 0) f = 2.0 / (iKc + 1.0 / K[o - sz]);
    else f = bc.t[4] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / K[o + (long)(nz - 1) * sz])
                                         : bface(bc.t[4], Kc);
    cstore(FZ, d.fz(i, j, k), f * cs);
    // upper boundary faces
    if (i == nx - 1) {
        f = bc.t[1] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / K[o - (nx - 1)]) : bface(bc.t[1], Kc);
        cstore(FX, d.fx(nx, j, k), f * cs);
    }
    if (j == ny - 1) {
        f = bc.t[3] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / K[o - (long)(ny - 1) * sy])
                                        : bface(bc.t[3], Kc);
        cstore(FY, d.fy(i, ny, k), f * cs);
    }
    if (k == nz - 1) {
        f = bc.t[5] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / K[o - (long)(nz - 1) * sz])
                                        : bface(bc.t[5], Kc);
        cstore(FZ, d.fz(i, j, nz), f * cs);
    }
}

// Neighbour values of cell (i,j,k) in the order x+, x-, y+, y-, z+, z-; a
// non-periodic boundary neighbour is 0.
template <typename T>
__device__ __forceinline__ void neighbours(const Dims& d, const BcTypes& bc, const T* x, long o,
                                           int i, int j, int k, T (&v)[6]) {
    const long sy = d.nx, sz = (long)d.nx * d.ny;
    v[0] = i < d.nx - 1 ? x[o + 1] : (bc.t[1] == MF3D_BC_PERIODIC ? x[o - (d.nx - 1)] : T(0));
    v[1] = i > 0 ? x[o - 1] : (bc.t[0] == MF3D_BC_PERIODIC ? x[o + (d.nx - 1)] : T(0));
    v[2] = j < d.ny - 1 ? x[o + sy] : (bc.t[3] == MF3D_BC_PERIODIC ? x[o - (long)(d.ny - 1) * sy] : T(0));
    v[3] = j > 0 ? x[o - sy] : (bc.t[2] == MF3D_BC_PERIODIC ? x[o + (long)(d.ny - 1) * sy] : T(0));
    v[4] = k < d.nz - 1 ? x[o + sz] : (bc.t[5] == MF3D_BC_PERIODIC ? x[o - (long)(d.nz - 1) * sz] : T(0));
    v[5] = k > 0 ? x[o - sz] : (bc.t[4] == MF3D_BC_PERIODIC ? x[o + (long)(d.nz - 1) * sz] : T(0));
}

template <typename T, typename TC>
__device__ __forceinline__ void cell_faces(const Dims& d, const TC* FX, const TC* FY,
                                           const TC* FZ, int i, int j, int k, T (&f)[6]) {
    f[0] = (T)cload(FX, d.fx(i + 1, j, k));
    f[1] = (T)cload(FX, d.fx(i, j, k));
    f[2] = (T)cload(FY, d.fy(i, j + 1, k));
    f[3] = (T)cload(FY, d.fy(i, j, k));
    f[4] = (T)cload(FZ, d.fz(i, j, k + 1));
    f[5] = (T)cload(FZ, d.fz(i, j, k));
}

// GSRB cell value (sum f x_nb - b cs dx^2) / sum f, summed x+, x-, y+, y-,
// z+, z- as in gsrb_interior_kernel (gsrb_3d.cu:53-69); pin doubles the
// diagonal of cell 0 (gsrb_3d.cu:360-362). The multiply-adds are written as
// fma: the first product is rounded, each later one is fused into the running
// sum, and b bs into the last step. Left to contraction, the compiler may
// fuse either product of a pair depending on the code around a call, and two
// kernels calling this would then differ in the last bit.
template <typename T>
__device__ __forceinline__ T gs_value(const T (&f)[6], const T (&v)[6], T b, T bs, bool pin0) {
    T result = v[0] * f[0];
    result = fma(v[1], f[1], result);
    result = fma(v[2], f[2], result);
    result = fma(v[3], f[3], result);
    result = fma(v[4], f[4], result);
    result = fma(v[5], f[5], result);
    T aC = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
    if (pin0) aC *= T(2);
    return fma(-b, bs, result) / aC;
}

// ---- one colour of GSRB on any level -----------------------------------------
template <typename T, typename TC, bool ZERO>
__global__ void __launch_bounds__(kBlock)
    colour_kernel(Dims d, BcTypes bc, int pin, T bs, T* __restrict__ x,
                  const T* __restrict__ b, const TC* __restrict__ FX,
                  const TC* __restrict__ FY, const TC* __restrict__ FZ, int colour) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int nxh = (nx + 1) >> 1;
    const long total = (long)nxh * ny * nz;
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long t = blk * kBlock + threadIdx.x;
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
        if (ip < nx) x[rowbase + ip] = T(0);
    }
    if (i >= nx) return;
    const long o = rowbase + i;
    T f[6], v[6];
    cell_faces(d, FX, FY, FZ, i, j, k, f);
    if (ZERO) {
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] = T(0);
    } else {
        neighbours(d, bc, x, o, i, j, k, v);
    }
    x[o] = gs_value(f, v, b[o], bs, pin && o == 0);
}

// ---- colour-split storage of the FP32 V-cycle levels -----------------------------
// With even nx, ny, nz a level's x and b may be stored colour by colour: the
// cells of colour c = (i+j+k) & 1 at c*N/2 + (i >> 1) + (nx/2)(j + ny k). A
// half-sweep then reads only the opposite colour's x and its own colour's b
// and writes only its own colour, all with unit stride.
// Every neighbour of a cell has the other colour: the y and z neighbours sit at
// the same (i >> 1) one row or plane away, the x neighbours at ((i +- 1) >> 1)
// of the same row.
// The face coefficients of such a level are stored the same way, each face
// with the cell on its upper side: LX/LY/LZ[split(i,j,k)] hold the faces
// between (i-1,j,k)|(i,j,k), (i,j-1,k)|(i,j,k), (i,j,k-1)|(i,j,k) (the MAC
// entries FX(i,j,k), FY(i,j,k), FZ(i,j,k)); the faces on the upper domain
// boundary, owned by no cell, go to UX[j + ny k], UY[i + nx k], UZ[i + nx j]
// (FX(nx,j,k), FY(i,ny,k), FZ(i,j,nz)). A cell's lower faces are then at its
// own index, its upper faces at its +x/+y/+z neighbours' indices.
struct Split {
    int h;       // nx / 2
    long half;   // N / 2
    __device__ long at(const Dims& d, int i, int j, int k) const {
        return (long)((i + j + k) & 1) * half + (long)(i >> 1) + (long)h * ((long)j + (long)d.ny * k);
    }
};
__device__ __forceinline__ Split split_of(const Dims& d) {
    return Split{d.nx >> 1, (long)(d.nx >> 1) * d.ny * d.nz};
}

// neighbours (x+, x-, y+, y-, z+, z-) of cell (i,j,k) from the opposite
// colour's array O; r0 = (j + ny k) * nx/2 is the cell's row offset
template <typename T>
__device__ __forceinline__ void neighbours_split(const Dims& d, const BcTypes& bc, const T* O,
                                                 long r0, int h, int i, int j, int k, T (&v)[6]) {
    const int m = i >> 1;
    const long sy = h, sz = (long)h * d.ny;
    v[0] = i < d.nx - 1 ? O[r0 + ((i + 1) >> 1)] : (bc.t[1] == MF3D_BC_PERIODIC ? O[r0] : T(0));
    v[1] = i > 0 ? O[r0 + ((i - 1) >> 1)] : (bc.t[0] == MF3D_BC_PERIODIC ? O[r0 + h - 1] : T(0));
    v[2] = j < d.ny - 1 ? O[r0 + sy + m]
                        : (bc.t[3] == MF3D_BC_PERIODIC ? O[r0 - (long)(d.ny - 1) * sy + m] : T(0));
    v[3] = j > 0 ? O[r0 - sy + m]
                 : (bc.t[2] == MF3D_BC_PERIODIC ? O[r0 + (long)(d.ny - 1) * sy + m] : T(0));
    v[4] = k < d.nz - 1 ? O[r0 + sz + m]
                        : (bc.t[5] == MF3D_BC_PERIODIC ? O[r0 - (long)(d.nz - 1) * sz + m] : T(0));
    v[5] = k > 0 ? O[r0 - sz + m]
                 : (bc.t[4] == MF3D_BC_PERIODIC ? O[r0 + (long)(d.nz - 1) * sz + m] : T(0));
}

// index of the coarse parent of fine cell (i,j,k) in a natural or split coarse array
template <bool CS>
__device__ __forceinline__ long parent_index(const Dims& dc, int i, int j, int k) {
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    if (CS) return split_of(dc).at(dc, I, J, K);
    return (long)I + (long)dc.nx * ((long)J + (long)dc.ny * K);
}

// wrapped neighbour coordinates (x+, x-, y+, y-, z+, z-); false for a
// non-periodic boundary neighbour
__device__ __forceinline__ bool nb_coord(const Dims& d, const BcTypes& bc, int q, int& i, int& j,
                                         int& k) {
    switch (q) {
    case 0: if (++i == d.nx) { if (bc.t[1] != MF3D_BC_PERIODIC) return false; i = 0; } break;
    case 1: if (--i < 0) { if (bc.t[0] != MF3D_BC_PERIODIC) return false; i = d.nx - 1; } break;
    case 2: if (++j == d.ny) { if (bc.t[3] != MF3D_BC_PERIODIC) return false; j = 0; } break;
    case 3: if (--j < 0) { if (bc.t[2] != MF3D_BC_PERIODIC) return false; j = d.ny - 1; } break;
    case 4: if (++k == d.nz) { if (bc.t[5] != MF3D_BC_PERIODIC) return false; k = 0; } break;
    default: if (--k < 0) { if (bc.t[4] != MF3D_BC_PERIODIC) return false; k = d.nz - 1; } break;
    }
    return true;
}

// The FP64 Krylov vectors that the mixed-precision iteration's r-update sweep
// and x/p update touch (r, A p, x, p: 128 MiB each at 256^3) are each read or
// written once per kernel, so they move with non-temporal loads and stores:
// that keeps them from displacing the FP32 V-cycle's x, b and faces (224 MiB
// at 256^3) from the 256 MB Infinity Cache between that V-cycle's
// half-sweeps. (The (r, z) sweep's read of r measured faster without it.)
typedef double __attribute__((ext_vector_type(2))) ntd2;
__device__ __forceinline__ double2 ldnt(const double* p) {
    const ntd2 v = __builtin_nontemporal_load(reinterpret_cast<const ntd2*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void stnt(double* p, double a, double b) {
    ntd2 v;
    v.x = a;
    v.y = b;
    __builtin_nontemporal_store(v, reinterpret_cast<ntd2*>(p));
}

template <typename TC>
struct SFaces {
    const TC* lx;
    const TC* ly;
    const TC* lz;
    const TC* ux;
    const TC* uy;
    const TC* uz;
};

__device__ __forceinline__ float cvt(float v) { return v; }
__device__ __forceinline__ float cvt(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ float cvt(__half v) { return __half2float(v); }

// faces (x+, x-, y+, y-, z+, z-) of cell (i,j,k) of a colour-split level
template <typename TC>
__device__ __forceinline__ void cell_faces_split(const Dims& d, const Split& sp,
                                                 const SFaces<TC>& F, int i, int j, int k,
                                                 float (&f)[6]) {
    const long own = sp.at(d, i, j, k);
    f[1] = cvt(F.lx[own]);
    f[3] = cvt(F.ly[own]);
    f[5] = cvt(F.lz[own]);
    f[0] = i < d.nx - 1 ? cvt(F.lx[sp.at(d, i + 1, j, k)]) : cvt(F.ux[j + (long)d.ny * k]);
    f[2] = j < d.ny - 1 ? cvt(F.ly[sp.at(d, i, j + 1, k)]) : cvt(F.uy[i + (long)d.nx * k]);
    f[4] = k < d.nz - 1 ? cvt(F.lz[sp.at(d, i, j, k + 1)]) : cvt(F.uz[i + (long)d.nx * j]);
}

// adds the parent coarse value to each in-domain neighbour value (MODE 2)
template <bool CS>
__device__ __forceinline__ void add_parents(const Dims& d, const BcTypes& bc, const float* xc,
                                            int i, int j, int k, float (&v)[6]) {
    const Dims dc{d.nx >> 1, d.ny >> 1, d.nz >> 1};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        int a = i, bb = j, c = k;
        if (nb_coord(d, bc, q, a, bb, c)) v[q] += xc[parent_index<CS>(dc, a, bb, c)];
    }
}

// One colour of GSRB on a colour-split level, one cell per thread (any even
// sizes). MODE 0: plain. MODE 1: x is zero on entry (no neighbour reads).
// MODE 2: the coarse correction xc is prolonged (piecewise constant) into the
// neighbours as they are read; used for the first colour of post-smoothing,
// whose own-colour values GS never reads (the other colour is rewritten by
// the next half-sweep).
template <typename TC, int MODE, bool CS>
__global__ void __launch_bounds__(kBlock)
    colour_split_kernel(Dims d, BcTypes bc, int pin, float bs, float* __restrict__ x,
                        const float* __restrict__ b, SFaces<TC> F, int colour,
                        const float* __restrict__ xc) {
    const Split sp = split_of(d);
    const long blk = xcd_block((sp.half + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long t = blk * kBlock + threadIdx.x;
    if (t >= sp.half) return;
    const int m = (int)(t % sp.h);
    const long row = t / sp.h;
    const int j = (int)(row % d.ny);
    const int k = (int)(row / d.ny);
    const int i = 2 * m + ((j + k + colour) & 1);
    float f[6], v[6];
    cell_faces_split(d, sp, F, i, j, k, f);
    if (MODE == 1) {
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] = 0.0f;
    } else {
        const float* O = x + (long)(colour ^ 1) * sp.half;
        neighbours_split(d, bc, O, row * sp.h, sp.h, i, j, k, v);
        if (MODE == 2) add_parents<CS>(d, bc, xc, i, j, k, v);
    }
    const long o = (long)colour * sp.half + t;
    x[o] = gs_value(f, v, b[o], bs, pin && colour == 0 && t == 0);
}

// 4 consecutive same-colour cells of a row per thread (nx a multiple of 8):
// x, b and the faces move as 16-byte / 8-byte vectors; the x-neighbour one
// place outside the thread's 4 comes from the adjacent lane.
template <typename TC>
struct alignas(4 * sizeof(TC)) Pack4 {
    TC v[4];
};

template <typename TC>
__device__ __forceinline__ void load4(const TC* p, float (&o)[4]) {
    const Pack4<TC> q = *reinterpret_cast<const Pack4<TC>*>(p);
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = cvt(q.v[e]);
}

// MODE 3: as MODE 1, with b taken from the FP64 natural-order r64 (both
// colours of b written as FP32 on the way: the conversion at the entry of the
// FP32 V-cycle folded into its first half-sweep).
// sweep workgroups: any block size up to kMaxSweepBlock (a tuning choice:
// results do not depend on it)
constexpr int kMaxSweepBlock = 512;

// (i, j, k) of a natural cell index, advanced along the natural order without
// divisions: the unit-stride passes below visit e, e + 128, ... and divide once
struct CellWalk {
    int i, j, k;
    __device__ CellWalk(const Dims& d, long e) {
        const long row = e / d.nx;
        i = (int)(e - row * d.nx);
        j = (int)(row % d.ny);
        k = (int)(row / d.ny);
    }
    __device__ void advance(const Dims& d, int by) {
        i += by;
        while (i >= d.nx) {
            i -= d.nx;
            if (++j == d.ny) j = 0, ++k;
        }
    }
};

// RZ (MODE 0, the last black half-sweep of a V-cycle): with the red values it
// loads as neighbours and the black values it writes, the thread holds z of
// its 8 natural cells and adds (r, z) to the workgroup's partial.
template <typename TC, int MODE, bool CS, bool RU = false, bool RZ = false>
__device__ __forceinline__ void split4_body(const Dims& d, const BcTypes& bc, int pin, float bs,
                                            float* __restrict__ x, const float* __restrict__ b,
                                            const SFaces<TC>& F, int colour,
                                            const float* __restrict__ xc,
                                            const double* r64, long blk,
                                            const RUpd* ru = nullptr,
                                            const RzDot* rz = nullptr) {
    const Split sp = split_of(d);
    const int h4 = sp.h >> 2;
    const long total = sp.half >> 2;
    const long t = blk * (long)blockDim.x + threadIdx.x;
    const bool valid = t < total;
    const long tt = valid ? t : total - 1;  // lanes past the end mirror the last one
    const int lane = threadIdx.x & 63;
    const int q = (int)(tt % h4);
    const long row = tt / h4;
    const int j = (int)(row % d.ny);
    const int k = (int)(row / d.ny);
    const int p = (j + k + colour) & 1;
    const long r0 = row * sp.h;
    const long base = r0 + 4L * q;
    const long own = (long)colour * sp.half + base;
    const long oth = (long)(colour ^ 1) * sp.half + base;
    const long sz = (long)sp.h * d.ny;
    const bool xlo = q == 0, xhi = q == h4 - 1;  // thread holds i = 0 / i = nx - 1

    // MODE 3 with the r update: the r and Ap loads of its unit-stride pass
    // (below) are issued before the face loads, so that their latency
    // overlaps; gfx9's in-order vmcnt would otherwise serialize each pair's
    // loads behind the previous pair's stores. Indices past the end read the
    // last pair (in bounds) and are not used.
    double2 rpre[4], apre[4];
    long epre[4];
    int cpre[4];
    if constexpr (MODE == 3 && RU) {
        const long n = (long)d.nx * d.ny * d.nz;
        const long w0 = 8L * (blk * (long)blockDim.x + (long)(threadIdx.x & ~63));
        CellWalk cw(d, min(w0 + 2L * lane, n - 2));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long e = w0 + 128L * u + 2L * lane;
            if (u > 0) cw.advance(d, 128);
            const long ec = e < n ? e : n - 2;
            epre[u] = e;
            cpre[u] = (cw.j + cw.k) & 1;  // colour of cell 2m
            rpre[u] = ldnt(r64 + ec);
            apre[u] = ldnt(ru->Ap + ec);
        }
    }

    // Every global load of the thread is issued before the first value is
    // used: addresses are chosen without branches (a lane at a boundary reads
    // an in-bounds stand-in, replaced or zeroed afterwards), so that a wave
    // waits for its memory once instead of once per branch (gfx9's vmcnt
    // counts in order).
    constexpr bool XR = !(MODE == 1 || MODE == 3);  // reads the other colour
    const bool ytop = j == d.ny - 1, ztop = k == d.nz - 1;
    // faces: lower ones at the own index, upper ones at the +x/+y/+z neighbours
    // (the last row / plane reads its own other-colour row and takes the
    // boundary faces below)
    const Pack4<TC> pxm = *reinterpret_cast<const Pack4<TC>*>(F.lx + own);
    const Pack4<TC> pym = *reinterpret_cast<const Pack4<TC>*>(F.ly + own);
    const Pack4<TC> pzm = *reinterpret_cast<const Pack4<TC>*>(F.lz + own);
    const Pack4<TC> pyp = *reinterpret_cast<const Pack4<TC>*>(F.ly + oth + (ytop ? 0 : sp.h));
    const Pack4<TC> pzp = *reinterpret_cast<const Pack4<TC>*>(F.lz + oth + (ztop ? 0 : sz));
    const Pack4<TC> plo = *reinterpret_cast<const Pack4<TC>*>(F.lx + oth);  // other colour
    const TC fnext = F.lx[oth + (xhi ? 0 : 4)];  // the next thread's first place (lane 63)
    const TC fux = F.ux[row];                    // upper x boundary face of the row
    // the other colour's values: own places, the places either side of the
    // thread's 4 (x edges and periodic wraps) and the y/z neighbour rows (the
    // wrapped row at a boundary: periodic, or zeroed below)
    const float* O = x + (long)(colour ^ 1) * sp.half;
    const long sy = sp.h;
    float a[4], oyp[4], oym[4], ozp[4], ozm[4], aprv = 0.0f, anxt = 0.0f;
    if (XR) {
        load4(O + base, a);
        aprv = O[xlo ? r0 + sp.h - 1 : base - 1];
        anxt = O[xhi ? r0 : base + 4];
        load4(O + base + (j < d.ny - 1 ? sy : -(long)(d.ny - 1) * sy), oyp);
        load4(O + base + (j > 0 ? -sy : (long)(d.ny - 1) * sy), oym);
        load4(O + base + (k < d.nz - 1 ? sz : -(long)(d.nz - 1) * sz), ozp);
        load4(O + base + (k > 0 ? -sz : (long)(d.nz - 1) * sz), ozm);
    }
    float bpre[4];  // own b (MODE 3 takes it from r below)
    if constexpr (MODE != 3) load4(b + own, bpre);
    // RZ: the r of the thread's 8 natural cells for the (r, z) partial
    double2 rzr[4];
    if constexpr (RZ) {
        const double2* rp = reinterpret_cast<const double2*>(rz->r + row * d.nx + 8L * q);
#pragma unroll
        for (int e = 0; e < 4; ++e) rzr[e] = rp[e];
    }

    float fxm[4], fym[4], fzm[4], fxp[4], fyp[4], fzp[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        fxm[e] = cvt(pxm.v[e]);
        fym[e] = cvt(pym.v[e]);
        fzm[e] = cvt(pzm.v[e]);
        fyp[e] = cvt(pyp.v[e]);
        fzp[e] = cvt(pzp.v[e]);
    }
    {
        float lo[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) lo[e] = cvt(plo.v[e]);
        float nxt = __shfl_down(lo[0], 1, 64);
        if (lane == 63 && !xhi) nxt = cvt(fnext);
        if (p == 0) {  // i = 2m: the +x neighbour is other-colour place m
#pragma unroll
            for (int e = 0; e < 4; ++e) fxp[e] = lo[e];
        } else {       // i = 2m + 1: place m + 1
            fxp[0] = lo[1];
            fxp[1] = lo[2];
            fxp[2] = lo[3];
            fxp[3] = xhi ? cvt(fux) : nxt;
        }
    }
    if (ytop) {
#pragma unroll
        for (int e = 0; e < 4; ++e) fyp[e] = cvt(F.uy[2 * (4 * q + e) + p + (long)d.nx * k]);
    }
    if (ztop) {
#pragma unroll
        for (int e = 0; e < 4; ++e) fzp[e] = cvt(F.uz[2 * (4 * q + e) + p + (long)d.nx * j]);
    }

    float vxp[4], vxm[4], vyp[4], vym[4], vzp[4], vzm[4];
    float oc[4];  // the other colour at the thread's 4 places (MODE 0 and 2)
    if (!XR) {
#pragma unroll
        for (int e = 0; e < 4; ++e) vxp[e] = vxm[e] = vyp[e] = vym[e] = vzp[e] = vzm[e] = 0.0f;
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) oc[e] = a[e];
        float prv = __shfl_up(a[3], 1, 64);
        float nxt = __shfl_down(a[0], 1, 64);
        if (lane == 0 && !xlo) prv = aprv;   // O[base - 1]
        if (lane == 63 && !xhi) nxt = anxt;  // O[base + 4]
        if (p == 0) {  // x+ at place m, x- at m - 1
            vxp[0] = a[0], vxp[1] = a[1], vxp[2] = a[2], vxp[3] = a[3];
            vxm[1] = a[0], vxm[2] = a[1], vxm[3] = a[2];
            // (at xlo, aprv is O[r0 + h - 1])
            vxm[0] = xlo ? (bc.t[0] == MF3D_BC_PERIODIC ? aprv : 0.0f) : prv;
        } else {       // x- at place m, x+ at m + 1
            vxm[0] = a[0], vxm[1] = a[1], vxm[2] = a[2], vxm[3] = a[3];
            vxp[0] = a[1], vxp[1] = a[2], vxp[2] = a[3];
            // (at xhi, anxt is O[r0])
            vxp[3] = xhi ? (bc.t[1] == MF3D_BC_PERIODIC ? anxt : 0.0f) : nxt;
        }
        const bool yp_in = j < d.ny - 1 || bc.t[3] == MF3D_BC_PERIODIC;
        const bool ym_in = j > 0 || bc.t[2] == MF3D_BC_PERIODIC;
        const bool zp_in = k < d.nz - 1 || bc.t[5] == MF3D_BC_PERIODIC;
        const bool zm_in = k > 0 || bc.t[4] == MF3D_BC_PERIODIC;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            vyp[e] = yp_in ? oyp[e] : 0.0f;
            vym[e] = ym_in ? oym[e] : 0.0f;
            vzp[e] = zp_in ? ozp[e] : 0.0f;
            vzm[e] = zm_in ? ozm[e] : 0.0f;
        }
    }
    float bb[4];
    if (MODE == 3 && RU) {
        // The wave's threads own the 512 natural cells [w0, w0 + 512) (8 per
        // thread, rows are contiguous). First pass over them in unit stride:
        // the cell pairs (2m, 2m + 1) of a row, one of each colour at place m,
        // pair P = 64 u + lane: r -= alpha Ap as pcg_update_r_kernel, (r, r),
        // both colours of b (FP32) out, and the own colour's b into LDS at the
        // pair's index, where the GS pass's thread finds its 4 places
        // (pairs 4 lane .. 4 lane + 3) as one float4.
        __shared__ float bown[kMaxSweepBlock * 4];
        const long n = (long)d.nx * d.ny * d.nz;
        const bool go = !ru->st->stop;  // uniform across the grid
        const double alpha = ru->st->alpha;
        float* bw = const_cast<float*>(b);
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long e = epre[u];
            if (e < n) {
                const int c0 = cpre[u];
                const long pl = e >> 1;
                double2 rv = rpre[u];
                const double2 av = apre[u];
                if (go) {
                    rv.x = -alpha * av.x + rv.x;
                    rv.y = -alpha * av.y + rv.y;
                    stnt(ru->r + e, rv.x, rv.y);
                }
                acc += rv.x * rv.x;
                acc += rv.y * rv.y;
                const float b0 = (float)rv.x, b1 = (float)rv.y;
                bw[(long)c0 * sp.half + pl] = b0;
                bw[(long)(c0 ^ 1) * sp.half + pl] = b1;
                bown[(threadIdx.x & ~63) * 4 + 64 * u + lane] = c0 == colour ? b0 : b1;
            }
        }
        const double sb = block_sum_dyn(acc);  // (synchronizes the workgroup)
        if (threadIdx.x == 0) ru->partials[blk] = sb;
        __syncthreads();
        const float4 o4 = *reinterpret_cast<const float4*>(bown + 4 * threadIdx.x);
        bb[0] = o4.x, bb[1] = o4.y, bb[2] = o4.z, bb[3] = o4.w;
    } else if (MODE == 3) {
        // the thread's 8 natural cells of the row: own colour at 2e + p, the other at 2e + 1 - p
        const double2* rp = reinterpret_cast<const double2*>(r64 + row * d.nx + 8L * q);
        double rv[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double2 v2 = rp[e];
            rv[2 * e] = v2.x;
            rv[2 * e + 1] = v2.y;
        }
        float bo[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            bb[e] = (float)rv[2 * e + p];
            bo[e] = (float)rv[2 * e + 1 - p];
        }
        if (valid) {
            float* bw = const_cast<float*>(b);
            *reinterpret_cast<float4*>(bw + own) = make_float4(bb[0], bb[1], bb[2], bb[3]);
            *reinterpret_cast<float4*>(bw + oth) = make_float4(bo[0], bo[1], bo[2], bo[3]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) bb[e] = bpre[e];
    }
    if (MODE == 2) {
        // prolongation: each in-domain neighbour gains its coarse parent's value.
        // Own cells i = 8q + 2e + p have parents I = 4q + e; their y/z neighbours
        // too (in the parent rows of j +- 1, k +- 1); x neighbours (i +- 1) >> 1,
        // i.e. I - 1 .. I + 1 of the own parent row. A parent row's 4 values
        // I = 4q .. 4q + 3 are one float4 (natural coarse) or two float2 (one
        // per colour: places 2q, 2q + 1) of a colour-split coarse level.
        const int NXc = d.nx >> 1, NYc = d.ny >> 1, NZc = d.nz >> 1;
        const int J0 = j >> 1, K0 = k >> 1;
        const long hc = NXc >> 1, HC = hc * NYc * NZc;
        auto cv = [&](int I, int J, int K) -> float {
            if (CS) return xc[(long)((I + J + K) & 1) * HC + (I >> 1) + hc * ((long)J + (long)NYc * K)];
            return xc[(long)I + (long)NXc * ((long)J + (long)NYc * K)];
        };
        auto row4 = [&](int J, int K, float (&o)[4]) {
            if (CS) {
                const long rb = 2L * q + hc * ((long)J + (long)NYc * K);
                const long c0 = (long)((J + K) & 1) * HC, c1 = (long)((J + K + 1) & 1) * HC;
                const float2 A = *reinterpret_cast<const float2*>(xc + c0 + rb);
                const float2 B = *reinterpret_cast<const float2*>(xc + c1 + rb);
                o[0] = A.x, o[1] = B.x, o[2] = A.y, o[3] = B.y;
            } else {
                const float4 A = *reinterpret_cast<const float4*>(
                    xc + 4L * q + (long)NXc * ((long)J + (long)NYc * K));
                o[0] = A.x, o[1] = A.y, o[2] = A.z, o[3] = A.w;
            }
        };
        const bool yp = j + 1 < d.ny || bc.t[3] == MF3D_BC_PERIODIC;
        const bool ym = j > 0 || bc.t[2] == MF3D_BC_PERIODIC;
        const bool zp = k + 1 < d.nz || bc.t[5] == MF3D_BC_PERIODIC;
        const bool zm = k > 0 || bc.t[4] == MF3D_BC_PERIODIC;
        const int Jp = j + 1 < d.ny ? (j + 1) >> 1 : 0, Jm = j > 0 ? (j - 1) >> 1 : NYc - 1;
        const int Kp = k + 1 < d.nz ? (k + 1) >> 1 : 0, Km = k > 0 ? (k - 1) >> 1 : NZc - 1;
        // own parent row, with I = 4q - 1 and I = 4q + 4 from the adjacent lanes
        float R[4];
        row4(J0, K0, R);
        // sizes are even: an even j shares its parent row with j + 1 (always in
        // the domain), an odd j with j - 1, so one of the two y neighbours'
        // rows is R and only the other is loaded; the same in z
        const bool jodd = j & 1, kodd = k & 1;
        float cy[4], cz[4];
        row4(jodd ? Jp : Jm, K0, cy);
        row4(J0, kodd ? Kp : Km, cz);
        const bool ylo = jodd || ym, yhi = !jodd || yp;  // in-domain or periodic
        const bool zlo = kodd || zm, zhi = !kodd || zp;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (yhi) vyp[e] += jodd ? cy[e] : R[e];
            if (ylo) vym[e] += jodd ? R[e] : cy[e];
            if (zhi) vzp[e] += kodd ? cz[e] : R[e];
            if (zlo) vzm[e] += kodd ? R[e] : cz[e];
        }
        float rm = __shfl_up(R[3], 1, 64), rp = __shfl_down(R[0], 1, 64);
        if (lane == 0 && !xlo) rm = cv(4 * q - 1, J0, K0);
        if (lane == 63 && !xhi) rp = cv(4 * q + 4, J0, K0);
        if (p == 0) {  // x+ neighbour 2I + 1: parent I; x- neighbour 2I - 1: parent I - 1
#pragma unroll
            for (int e = 0; e < 4; ++e) vxp[e] += R[e];
            vxm[1] += R[0], vxm[2] += R[1], vxm[3] += R[2];
            if (!xlo) vxm[0] += rm;
            else if (bc.t[0] == MF3D_BC_PERIODIC) vxm[0] += cv(NXc - 1, J0, K0);
        } else {       // x- neighbour 2I: parent I; x+ neighbour 2I + 2: parent I + 1
#pragma unroll
            for (int e = 0; e < 4; ++e) vxm[e] += R[e];
            vxp[0] += R[1], vxp[1] += R[2], vxp[2] += R[3];
            if (!xhi) vxp[3] += rp;
            else if (bc.t[1] == MF3D_BC_PERIODIC) vxp[3] += cv(0, J0, K0);
        }
    }
    float res[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        float f[6] = {fxp[e], fxm[e], fyp[e], fym[e], fzp[e], fzm[e]};
        float v[6] = {vxp[e], vxm[e], vyp[e], vym[e], vzp[e], vzm[e]};
        res[e] = gs_value(f, v, bb[e], bs, pin && colour == 0 && row == 0 && q == 0 && e == 0);
    }
    if (valid) *reinterpret_cast<float4*>(x + own) = make_float4(res[0], res[1], res[2], res[3]);
    if (RZ) {  // own cells at natural 2e + p, the other colour at 2e + 1 - p
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double2 v2 = rzr[e];
            const double ro = p ? v2.y : v2.x, rt = p ? v2.x : v2.y;
            acc += ro * (double)res[e];
            acc += rt * (double)oc[e];
        }
        const double sb = block_sum_dyn(valid ? acc : 0.0);
        if (threadIdx.x == 0) rz->partials[blk] = sb;
    }
}

template <typename TC, int MODE, bool CS>
__global__ void __launch_bounds__(kMaxSweepBlock)
    colour_split4_kernel(Dims d, BcTypes bc, int pin, float bs, float* __restrict__ x,
                         const float* __restrict__ b, SFaces<TC> F, int colour,
                         const float* __restrict__ xc, const double* __restrict__ r64) {
    const long total = ((long)(d.nx >> 1) * d.ny * d.nz) >> 2;
    const long blk = xcd_block((total + blockDim.x - 1) / blockDim.x);
    if (blk < 0) return;  // whole workgroup
    split4_body<TC, MODE, CS>(d, bc, pin, bs, x, b, F, colour, xc, r64, blk);
}

// the first half-sweep from x = 0 with b from r, after r -= alpha Ap (RUpd)
template <typename TC>
__global__ void __launch_bounds__(kMaxSweepBlock)
    colour_split4_r_kernel(Dims d, BcTypes bc, int pin, float bs, float* __restrict__ x,
                           const float* __restrict__ b, SFaces<TC> F, int colour,
                           const double* r64, RUpd ru) {
    const long total = ((long)(d.nx >> 1) * d.ny * d.nz) >> 2;
    const long blk = xcd_block((total + blockDim.x - 1) / blockDim.x);
    if (blk < 0) return;  // whole workgroup
    split4_body<TC, 3, false, true>(d, bc, pin, bs, x, b, F, colour, nullptr, r64, blk, &ru);
}

// the last (black) half-sweep of the V-cycle with the (r, z) partials (RzDot)
template <typename TC>
__global__ void __launch_bounds__(kMaxSweepBlock)
    colour_split4_rz_kernel(Dims d, BcTypes bc, int pin, float bs, float* __restrict__ x,
                            const float* __restrict__ b, SFaces<TC> F, int colour, RzDot rz) {
    const long total = ((long)(d.nx >> 1) * d.ny * d.nz) >> 2;
    const long blk = xcd_block((total + blockDim.x - 1) / blockDim.x);
    if (blk < 0) return;  // whole workgroup
    split4_body<TC, 0, false, false, true>(d, bc, pin, bs, x, b, F, colour, nullptr, nullptr,
                                           blk, nullptr, &rz);
}

// prolong-add between any two orders (used when a level has no post-smoothing)
template <bool FS, bool CS>
__global__ void __launch_bounds__(kBlock)
    prolong_order_kernel(Dims d, const float* __restrict__ c, float* __restrict__ f) {
    const long n = (long)d.nx * d.ny * d.nz;
    const long blk = xcd_block((n + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long o = blk * kBlock + threadIdx.x;
    if (o >= n) return;
    const int i = (int)(o % d.nx);
    const long q = o / d.nx;
    const int j = (int)(q % d.ny);
    const int k = (int)(q / d.ny);
    const Dims dc{d.nx >> 1, d.ny >> 1, d.nz >> 1};
    const long fo = FS ? split_of(d).at(d, i, j, k) : o;
    f[fo] += c[parent_index<CS>(dc, i, j, k)];
}

// ---- whole smoothing of a small level in one workgroup -----------------------
constexpr int kSmallBlock = 1024;
constexpr int kSmallPer = 4;

// Thread t owns cells o = 2t + (m & 1) + 2048 (m >> 1), m = 0..3: two
// x-adjacent pairs, one cell of each colour per pair when nx is even, so every
// lane has work in every half-sweep. The neighbours' LDS slots (-1 for a
// non-periodic boundary, whose ghost value is 0), the faces and b are
// gathered once; the iterations then touch LDS only. Per cell the same sums
// as colour_kernel (neighbours x+, x-, y+, y-, z+, z-; gs_value).
template <typename T, typename TC>
__device__ __forceinline__ void small_body(T* xs, const Dims& d, const BcTypes& bc, int pin, T bs,
                                           T* __restrict__ x, const T* __restrict__ b,
                                           const TC* __restrict__ FX, const TC* __restrict__ FY,
                                           const TC* __restrict__ FZ, int iters, int zero_init,
                                           const T* __restrict__ xc) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int n = nx * ny * nz;
    const long sy = nx, sz = (long)nx * ny;
    T f[kSmallPer][6], bl[kSmallPer];
    int nb[kSmallPer][6];
    int colour[kSmallPer];  // -1: no cell
    const int NXc = nx >> 1, NYc = ny >> 1;
#pragma unroll
    for (int m = 0; m < kSmallPer; ++m) {
        const int o = 2 * (int)threadIdx.x + (m & 1) + 2 * kSmallBlock * (m >> 1);
        bl[m] = T(0);
        colour[m] = -1;
        if (o < n) {
            const int i = o % nx, q = o / nx, j = q % ny, k = q / ny;
            colour[m] = (i + j + k) & 1;
            T v = zero_init ? T(0) : x[o];
            if (xc) v += xc[(i >> 1) + NXc * ((j >> 1) + NYc * (k >> 1))];
            xs[o] = v;
            cell_faces(d, FX, FY, FZ, i, j, k, f[m]);
            bl[m] = b[o];
            nb[m][0] = i < nx - 1 ? o + 1 : (bc.t[1] == MF3D_BC_PERIODIC ? o - (nx - 1) : -1);
            nb[m][1] = i > 0 ? o - 1 : (bc.t[0] == MF3D_BC_PERIODIC ? o + (nx - 1) : -1);
            nb[m][2] = j < ny - 1 ? (int)(o + sy)
                                  : (bc.t[3] == MF3D_BC_PERIODIC ? (int)(o - (ny - 1) * sy) : -1);
            nb[m][3] = j > 0 ? (int)(o - sy)
                             : (bc.t[2] == MF3D_BC_PERIODIC ? (int)(o + (ny - 1) * sy) : -1);
            nb[m][4] = k < nz - 1 ? (int)(o + sz)
                                  : (bc.t[5] == MF3D_BC_PERIODIC ? (int)(o - (nz - 1) * sz) : -1);
            nb[m][5] = k > 0 ? (int)(o - sz)
                             : (bc.t[4] == MF3D_BC_PERIODIC ? (int)(o + (nz - 1) * sz) : -1);
        }
    }
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int m = 0; m < kSmallPer; ++m) {
                if (colour[m] != c) continue;
                const int o = 2 * (int)threadIdx.x + (m & 1) + 2 * kSmallBlock * (m >> 1);
                T v[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) v[q] = nb[m][q] >= 0 ? xs[nb[m][q]] : T(0);
                xs[o] = gs_value(f[m], v, bl[m], bs, pin && o == 0);
            }
            __syncthreads();
        }
    }
    for (int o = threadIdx.x; o < n; o += kSmallBlock) x[o] = xs[o];
}

template <typename T, typename TC>
__global__ void __launch_bounds__(kSmallBlock)
    small_kernel(Dims d, BcTypes bc, int pin, T bs, T* __restrict__ x, const T* __restrict__ b,
                 const TC* __restrict__ FX, const TC* __restrict__ FY,
                 const TC* __restrict__ FZ, int iters, int zero_init, const T* __restrict__ xc) {
    __shared__ T xs[kSmallBlock * kSmallPer];
    small_body<T, TC>(xs, d, bc, pin, bs, x, b, FX, FY, FZ, iters, zero_init, xc);
}

// ---- residual fused with restriction -----------------------------------------
// r = b - A x (smoother convention, pin adds aC x_C at cell 0,
// residual_3d.cu:365-368) of the 8 children of one coarse cell, averaged in
// the order of restrict_kernel (restrict_3d.cu:36-49).
template <typename T, typename TC>
__device__ __forceinline__ T residual_cell(const Dims& d, const BcTypes& bc, bool pin,
                                           const T* x, const T* b, const TC* FX, const TC* FY,
                                           const TC* FZ, int i, int j, int k, T rs) {
    const long o = (long)i + (long)d.nx * ((long)j + (long)d.ny * k);
    T f[6], v[6];
    cell_faces(d, FX, FY, FZ, i, j, k, f);
    neighbours(d, bc, x, o, i, j, k, v);
    const T xc = x[o];
    T result = f[0] * (xc - v[0]);  // fma as in gs_value
    result = fma(f[1], xc - v[1], result);
    result = fma(f[2], xc - v[2], result);
    result = fma(f[3], xc - v[3], result);
    result = fma(f[4], xc - v[4], result);
    result = fma(f[5], xc - v[5], result);
    if (pin && o == 0) result = fma(f[0] + f[1] + f[2] + f[3] + f[4] + f[5], xc, result);
    return fma(result, rs, b[o]);
}

template <typename T, typename TC>
__device__ __forceinline__ void rr_body(const Dims& d, const BcTypes& bc, int pin, T rs,
                                        const T* __restrict__ x, const T* __restrict__ b,
                                        const TC* __restrict__ FX, const TC* __restrict__ FY,
                                        const TC* __restrict__ FZ, T* __restrict__ bcoarse,
                                        long blk) {
    const int NX = d.nx >> 1, NY = d.ny >> 1, NZ = d.nz >> 1;
    const long total = (long)NX * NY * NZ;
    const long t = blk * kBlock + threadIdx.x;
    if (t >= total) return;
    const int I = (int)(t % NX);
    const long q = t / NX;
    const int J = (int)(q % NY);
    const int Kc = (int)(q / NY);
    const int i = 2 * I, j = 2 * J, k = 2 * Kc;
    const bool p = pin != 0;
    T s = T(0);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i, j, k, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i + 1, j, k, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i, j + 1, k, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i + 1, j + 1, k, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i, j, k + 1, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i + 1, j, k + 1, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i, j + 1, k + 1, rs);
    s += residual_cell(d, bc, p, x, b, FX, FY, FZ, i + 1, j + 1, k + 1, rs);
    bcoarse[t] = s / T(8);
}

template <typename T, typename TC>
__global__ void __launch_bounds__(kBlock)
    residual_restrict_kernel(Dims d, BcTypes bc, int pin, T rs, const T* __restrict__ x,
                             const T* __restrict__ b, const TC* __restrict__ FX,
                             const TC* __restrict__ FY, const TC* __restrict__ FZ,
                             T* __restrict__ bcoarse) {
    const long total = (long)(d.nx >> 1) * (d.ny >> 1) * (d.nz >> 1);
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;
    rr_body<T, TC>(d, bc, pin, rs, x, b, FX, FY, FZ, bcoarse, blk);
}

// residual of one cell of a colour-split level (same sums as residual_cell)
template <typename TC>
__device__ __forceinline__ float residual_cell_split(const Dims& d, const BcTypes& bc, bool pin,
                                                     const Split& sp, const float* x,
                                                     const float* b, const SFaces<TC>& F, int i,
                                                     int j, int k, float rs) {
    const int c = (i + j + k) & 1;
    const long r0 = (long)sp.h * ((long)j + (long)d.ny * k);
    const long o = (long)c * sp.half + r0 + (i >> 1);
    float f[6], v[6];
    cell_faces_split(d, sp, F, i, j, k, f);
    neighbours_split(d, bc, x + (long)(c ^ 1) * sp.half, r0, sp.h, i, j, k, v);
    const float xc = x[o];
    float result = f[0] * (xc - v[0]);  // fma as in gs_value
    result = fma(f[1], xc - v[1], result);
    result = fma(f[2], xc - v[2], result);
    result = fma(f[3], xc - v[3], result);
    result = fma(f[4], xc - v[4], result);
    result = fma(f[5], xc - v[5], result);
    if (pin && (i | j | k) == 0) result = fma(f[0] + f[1] + f[2] + f[3] + f[4] + f[5], xc, result);
    return fma(result, rs, b[o]);
}

// residual + restriction of a colour-split fine level into a coarse level of
// either order. RED: after a red-black GS sweep that ended with the black
// half-sweep, each black cell satisfies its own equation given its (final)
// red neighbours, so its residual is zero up to FP32 rounding; only the four
// red children of a coarse cell are evaluated.
// Coarse first half-sweep folded in (FUSE, colour-split coarse level whose x
// starts at zero): a red coarse cell's first GS value depends only on its own
// b and faces, x = (0 - b cs dx^2) / sum f, so it is written here (values as
// the zero-start half-sweep) and the coarse smoothing starts with black.
struct CoarseFirst {
    float* x;     // coarse x (split order); null: not folded
    float bs;     // coarse cs dx^2
};

template <typename TC, bool CS, bool RED, bool FUSE = false>
__device__ __forceinline__ void rr_split_body(const Dims& d, const BcTypes& bc, int pin, float rs,
                                              const float* __restrict__ x,
                                              const float* __restrict__ b, const SFaces<TC>& F,
                                              float* __restrict__ bcoarse, long blk,
                                              const SFaces<TC>* Fc = nullptr,
                                              CoarseFirst cf = {}) {
    const Dims dc{d.nx >> 1, d.ny >> 1, d.nz >> 1};
    const long total = (long)dc.nx * dc.ny * dc.nz;
    const long t = blk * kBlock + threadIdx.x;
    if (t >= total) return;
    const int I = (int)(t % dc.nx);
    const long q = t / dc.nx;
    const int J = (int)(q % dc.ny);
    const int Kc = (int)(q / dc.ny);
    const int i = 2 * I, j = 2 * J, k = 2 * Kc;
    const bool p = pin != 0;
    const Split sp = split_of(d);
    float s = 0.0f;
    if (RED) {  // children of colour 0: (0,0,0), (1,1,0), (1,0,1), (0,1,1)
        s += residual_cell_split(d, bc, p, sp, x, b, F, i, j, k, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i + 1, j + 1, k, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i + 1, j, k + 1, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i, j + 1, k + 1, rs);
    } else {
        s += residual_cell_split(d, bc, p, sp, x, b, F, i, j, k, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i + 1, j, k, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i, j + 1, k, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i + 1, j + 1, k, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i, j, k + 1, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i + 1, j, k + 1, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i, j + 1, k + 1, rs);
        s += residual_cell_split(d, bc, p, sp, x, b, F, i + 1, j + 1, k + 1, rs);
    }
    const long oc = CS ? split_of(dc).at(dc, I, J, Kc) : t;
    const float bcv = s / 8.0f;
    bcoarse[oc] = bcv;
    if (FUSE && CS && ((I + J + Kc) & 1) == 0) {
        float f[6], v[6];
        cell_faces_split(dc, split_of(dc), *Fc, I, J, Kc, f);
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] = 0.0f;
        cf.x[oc] = gs_value(f, v, bcv, cf.bs, pin && (I | J | Kc) == 0);
    }
}

template <typename TC, bool CS, bool RED, bool FUSE>
__global__ void __launch_bounds__(kBlock)
    residual_restrict_split_kernel(Dims d, BcTypes bc, int pin, float rs,
                                   const float* __restrict__ x, const float* __restrict__ b,
                                   SFaces<TC> F, float* __restrict__ bcoarse, SFaces<TC> Fc,
                                   CoarseFirst cf) {
    const long total = (long)(d.nx >> 1) * (d.ny >> 1) * (d.nz >> 1);
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;
    rr_split_body<TC, CS, RED, FUSE>(d, bc, pin, rs, x, b, F, bcoarse, blk, &Fc, cf);
}

// face coefficients of a colour-split level from K (values as faces_build_kernel);
// with HK, also H = 1 / (2 K) of the FP64 operator (the finest level of the
// mixed-precision solve)
template <typename TC, bool HK>
__global__ void __launch_bounds__(kBlock)
    faces_build_split_kernel(Dims d, BcTypes bc, double cs, const double* __restrict__ K,
                             TC* __restrict__ LX, TC* __restrict__ LY, TC* __restrict__ LZ,
                             TC* __restrict__ UX, TC* __restrict__ UY, TC* __restrict__ UZ,
                             double* __restrict__ H) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long n = (long)nx * ny * nz;
    const long blk = xcd_block((n + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long o = blk * kBlock + threadIdx.x;
    if (o >= n) return;
    const int i = (int)(o % nx);
    const long q = o / nx;
    const int j = (int)(q % ny);
    const int k = (int)(q / ny);
    const long sy = nx, sz = (long)nx * ny;
    const long so = split_of(d).at(d, i, j, k);
    // every load first, from indices that are in range whether or not the
    // face uses them (the lower neighbour or the periodic wrap; the upper wrap
    // at the last cell of a line, the cell itself elsewhere), so that they
    // are in flight together
    const double Kc = K[o];
    const double Kw = K[i > 0 ? o - 1 : o + nx - 1];
    const double Ks = K[j > 0 ? o - sy : o + (long)(ny - 1) * sy];
    const double Kd = K[k > 0 ? o - sz : o + (long)(nz - 1) * sz];
    const double Ke = K[i == nx - 1 ? o - (nx - 1) : o];
    const double Kn = K[j == ny - 1 ? o - (long)(ny - 1) * sy : o];
    const double Ku = K[k == nz - 1 ? o - (long)(nz - 1) * sz : o];
    const double iKc = 1.0 / Kc;
    if (HK) H[o] = 0.5 * iKc;
    double f;
    if (i > 0) f = 2.0 / (iKc + 1.0 / Kw);
    else f = bc.t[0] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / Kw) : bface(bc.t[0], Kc);
    cstore(LX, so, f * cs);
    if (j > 0) f = 2.0 / (iKc + 1.0 / Ks);
    else f = bc.t[2] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / Ks) : bface(bc.t[2], Kc);
    cstore(LY, so, f * cs);
    if (k > 0) f = 2.0 / (iKc + 1.0 / Kd);
    else f = bc.t[4] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / Kd) : bface(bc.t[4], Kc);
    cstore(LZ, so, f * cs);
    if (i == nx - 1) {
        f = bc.t[1] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / Ke) : bface(bc.t[1], Kc);
        cstore(UX, j + (long)ny * k, f * cs);
    }
    if (j == ny - 1) {
        f = bc.t[3] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / Kn) : bface(bc.t[3], Kc);
        cstore(UY, i + (long)nx * k, f * cs);
    }
    if (k == nz - 1) {
        f = bc.t[5] == MF3D_BC_PERIODIC ? 2.0 / (iKc + 1.0 / Ku) : bface(bc.t[5], Kc);
        cstore(UZ, i + (long)nx * j, f * cs);
    }
}

// H = 1 / (2 K) (0.5 times the rounded 1 / K, exact scaling)
__global__ void __launch_bounds__(kBlock)
    half_recip_kernel(const double* __restrict__ K, double* __restrict__ H, long n) {
    const long o = (long)blockIdx.x * kBlock + threadIdx.x;
    if (o < n) H[o] = 0.5 * (1.0 / K[o]);
}

// ---- prolongation for natural-order levels ------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock)
    prolong_kernel(Dims d, const T* __restrict__ c, T* __restrict__ f) {
    const long n = (long)d.nx * d.ny * d.nz;
    const long blk = xcd_block((n + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long o = blk * kBlock + threadIdx.x;
    if (o >= n) return;
    const int i = (int)(o % d.nx);
    const long q = o / d.nx;
    const int j = (int)(q % d.ny);
    const int k = (int)(q / d.ny);
    f[o] += c[(long)(i >> 1) + (long)(d.nx >> 1) * ((long)(j >> 1) + (long)(d.ny >> 1) * (k >> 1))];
}

// ---- FP64 operator on the finest level ----------------------------------------

// A x = -sum f (x_C - x_nb) / dx^2 (varcoeff_laplacian.cu:38-239), summed
// x-, x+, y-, y+, z-, z+; the pin's diagonal counts in-domain neighbours only
// (varcoeff_laplacian.cu:110-234).
//
// The face coefficients are not stored: with h = 1 / (2 K) per cell (8 bytes
// per cell instead of 24 for three FP64 face arrays) the harmonic face is
// f = 2 Ka Kb / (Ka + Kb) = 1 / (ha + hb), a Dirichlet face (2 K_C, the
// ghost value lives in b) is 1 / h_C, i.e. the same form with h_nb = 0, and
// a Neumann face is 0. Periodic faces pair with the wrapped cell.
//
// MODE 0: y = A x, partials of (x, y).  MODE 1: r = b - A x (y may be null),
// partials of (r, r). One partial per workgroup: the column sums in z order,
// then the block tree.
//
// Single-cell variant (odd nx): a workgroup owns kBlock consecutive cells of
// an xy plane and marches kz planes up in z, each thread keeping its column's
// x and h at k-1, k, k+1 and the z face below k in registers; the next plane's
// loads are issued before the current plane is computed.
template <int MODE>
__device__ __forceinline__ void opz_body(const Dims& d, const BcTypes& bc, int pin, double inv_dx2,
                                         int tiles, int kz, const double* __restrict__ x,
                                         const double* __restrict__ H,
                                         const double* __restrict__ b, double* __restrict__ y,
                                         double* __restrict__ partials, long blk) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int P = nx * ny;
    const int chunk = (int)(blk / tiles);
    const int p = (int)(blk - (long)chunk * tiles) * kBlock + (int)threadIdx.x;
    double acc = 0.0;
    if (p < P) {
        const int j = p / nx;
        const int i = p - j * nx;
        const long sz = P;
        const int k0 = chunk * kz;
        const int k1 = min(nz, k0 + kz);
        const Side E = side(i < nx - 1, 1, bc.t[1], -(long)(nx - 1));
        const Side Wt = side(i > 0, -1, bc.t[0], (long)(nx - 1));
        const Side N = side(j < ny - 1, nx, bc.t[3], -(long)(ny - 1) * nx);
        const Side S = side(j > 0, -(long)nx, bc.t[2], (long)(ny - 1) * nx);
        long o = (long)k0 * sz + p;
        const Side D0 = side(k0 > 0, -sz, bc.t[4], (long)(nz - 1) * sz);
        double xc = x[o], hc = H[o];
        const double xd = x[o + D0.off], hd = H[o + D0.off];
        double xm = D0.nb ? xd : 0.0;
        double fzl = hface(D0, hc, hd);
        struct Ld {
            double xp, hp, v0, v1, v2, v3, h0, h1, h2, h3, bo;
        };
        auto load = [&](int k, long o) {
            const Side U = side(k < nz - 1, sz, bc.t[5], -(long)(nz - 1) * sz);
            Ld l;
            l.xp = x[o + U.off];
            l.hp = H[o + U.off];
            l.v0 = x[o + E.off];
            l.v1 = x[o + Wt.off];
            l.v2 = x[o + N.off];
            l.v3 = x[o + S.off];
            l.h0 = H[o + E.off];
            l.h1 = H[o + Wt.off];
            l.h2 = H[o + N.off];
            l.h3 = H[o + S.off];
            l.bo = MODE == 1 ? b[o] : 0.0;
            return l;
        };
        Ld cur = load(k0, o);
        for (int k = k0; k < k1; ++k) {
            Ld nxt = cur;
            if (k + 1 < k1) nxt = load(k + 1, o + sz);
            const Side U = side(k < nz - 1, sz, bc.t[5], -(long)(nz - 1) * sz);
            const double f0 = hface(E, hc, cur.h0), f1 = hface(Wt, hc, cur.h1);
            const double f2 = hface(N, hc, cur.h2), f3 = hface(S, hc, cur.h3);
            const double fzu = hface(U, hc, cur.hp);
            const double v0 = E.nb ? cur.v0 : 0.0, v1 = Wt.nb ? cur.v1 : 0.0;
            const double v2 = N.nb ? cur.v2 : 0.0, v3 = S.nb ? cur.v3 : 0.0;
            const double xp = U.nb ? cur.xp : 0.0;
            double Ax = f1 * (xc - v1) + f0 * (xc - v0) + f3 * (xc - v3) + f2 * (xc - v2) +
                        fzl * (xc - xm) + fzu * (xc - xp);
            if (pin && o == 0) {  // in-domain faces of cell 0: x+, y+, z+
                double aC = 0.0;
                if (nx > 1) aC += f0;
                if (ny > 1) aC += f2;
                if (nz > 1) aC += fzu;
                Ax += aC * xc;
            }
            Ax = -Ax * inv_dx2;
            if (MODE == 0) {
                y[o] = Ax;
                acc += xc * Ax;
            } else {
                const double r = -1.0 * Ax + cur.bo;
                if (y) y[o] = r;
                acc += r * r;
            }
            xm = xc;
            xc = xp;
            hc = U.nb ? cur.hp : 0.0;
            fzl = fzu;
            o += sz;
            cur = nxt;
        }
    }
    const double s = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) partials[blk] = s;
}

// Paired variant (even nx): each thread owns two x-adjacent cells and moves
// them with 16-byte loads and stores, which halves the vector-memory
// instructions per cell (the single-cell kernel is bound by address
// processing, not by HBM). A workgroup owns 2 kBlock consecutive cells of an
// xy plane and marches kz planes up in z. The x neighbours across a pair
// boundary come from the adjacent lanes; the first and last lane of a wave,
// and periodic wraps, load them. Per cell: the same faces and the same sum
// order as opz_body.
typedef double __attribute__((ext_vector_type(2))) dbl2;
__device__ __forceinline__ dbl2 ld2(const double* __restrict__ q, long off) {
    return *reinterpret_cast<const dbl2*>(q + off);
}

template <int MODE>
__device__ __forceinline__ void opz2_body(const Dims& d, const BcTypes& bc, int pin,
                                          double inv_dx2, int tiles, int kz,
                                          const double* __restrict__ x,
                                          const double* __restrict__ H,
                                          const double* __restrict__ b, double* __restrict__ y,
                                          double* __restrict__ partials, long blk) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int P = nx * ny;
    const int chunk = (int)(blk / tiles);
    const int p = (int)(blk - (long)chunk * tiles) * (2 * kBlock) + 2 * (int)threadIdx.x;
    const int lane = (int)(threadIdx.x & 63);
    const bool act = p < P;
    const int pc = act ? p : 0;  // idle lanes read a valid pair and store nothing
    const int j = pc / nx;
    const int i = pc - j * nx;  // even
    const long sz = P;
    const int k0 = chunk * kz;
    const int k1 = min(nz, k0 + kz);
    // x sides of the pair: west of cell i, east of cell i + 1
    const Side Wt = side(i > 0, -1, bc.t[0], (long)(nx - 1));
    const Side E = side(i + 2 < nx, 2, bc.t[1], 2 - (long)nx);
    // lanes that load the x neighbour instead of taking it from the next lane
    const bool wload = Wt.nb && (lane == 0 || i == 0);
    const bool eload = E.nb && (lane == 63 || i + 2 == nx);
    const Side N = side(j < ny - 1, nx, bc.t[3], -(long)(ny - 1) * nx);
    const Side S = side(j > 0, -(long)nx, bc.t[2], (long)(ny - 1) * nx);
    long o = (long)k0 * sz + pc;
    const Side D0 = side(k0 > 0, -sz, bc.t[4], (long)(nz - 1) * sz);
    dbl2 xc = ld2(x, o), hc = ld2(H, o);
    dbl2 xm = ld2(x, o + D0.off);
    const dbl2 hd = ld2(H, o + D0.off);
    if (!D0.nb) xm = dbl2{0.0, 0.0};
    dbl2 fzl = dbl2{hface(D0, hc.x, hd.x), hface(D0, hc.y, hd.y)};
    struct Ld {
        dbl2 xp, hp, xn, xs, hn, hs, bo;
        double xw, hw, xe, he;
    };
    auto load = [&](int k, long o) {
        const Side U = side(k < nz - 1, sz, bc.t[5], -(long)(nz - 1) * sz);
        Ld l;
        l.xp = ld2(x, o + U.off);
        l.hp = ld2(H, o + U.off);
        l.xn = ld2(x, o + N.off);
        l.hn = ld2(H, o + N.off);
        l.xs = ld2(x, o + S.off);
        l.hs = ld2(H, o + S.off);
        l.bo = MODE == 1 ? ld2(b, o) : dbl2{0.0, 0.0};
        l.xw = wload ? x[o + Wt.off] : 0.0;
        l.hw = wload ? H[o + Wt.off] : 0.0;
        l.xe = eload ? x[o + E.off] : 0.0;
        l.he = eload ? H[o + E.off] : 0.0;
        return l;
    };
    double acc = 0.0;
    Ld cur = load(k0, o);
    for (int k = k0; k < k1; ++k) {
        Ld nxt = cur;
        if (k + 1 < k1) nxt = load(k + 1, o + sz);
        const Side U = side(k < nz - 1, sz, bc.t[5], -(long)(nz - 1) * sz);
        double xw = __shfl_up(xc.y, 1, 64), hw = __shfl_up(hc.y, 1, 64);
        double xe = __shfl_down(xc.x, 1, 64), he = __shfl_down(hc.x, 1, 64);
        if (wload) xw = cur.xw, hw = cur.hw;
        if (eload) xe = cur.xe, he = cur.he;
        if (!Wt.nb) xw = 0.0;
        if (!E.nb) xe = 0.0;
        const double fw = hface(Wt, hc.x, hw);
        const double fm = rcp_nr(hc.x + hc.y);  // between the two cells
        const double fe = hface(E, hc.y, he);
        const double fn0 = hface(N, hc.x, cur.hn.x), fn1 = hface(N, hc.y, cur.hn.y);
        const double fs0 = hface(S, hc.x, cur.hs.x), fs1 = hface(S, hc.y, cur.hs.y);
        const double fu0 = hface(U, hc.x, cur.hp.x), fu1 = hface(U, hc.y, cur.hp.y);
        const double n0 = N.nb ? cur.xn.x : 0.0, n1 = N.nb ? cur.xn.y : 0.0;
        const double s0 = S.nb ? cur.xs.x : 0.0, s1 = S.nb ? cur.xs.y : 0.0;
        const double u0 = U.nb ? cur.xp.x : 0.0, u1 = U.nb ? cur.xp.y : 0.0;
        double A0 = fw * (xc.x - xw) + fm * (xc.x - xc.y) + fs0 * (xc.x - s0) + fn0 * (xc.x - n0) +
                    fzl.x * (xc.x - xm.x) + fu0 * (xc.x - u0);
        double A1 = fm * (xc.y - xc.x) + fe * (xc.y - xe) + fs1 * (xc.y - s1) + fn1 * (xc.y - n1) +
                    fzl.y * (xc.y - xm.y) + fu1 * (xc.y - u1);
        if (pin && o == 0) {  // in-domain faces of cell 0: x+ (fm), y+, z+
            double aC = fm;
            if (ny > 1) aC += fn0;
            if (nz > 1) aC += fu0;
            A0 += aC * xc.x;
        }
        A0 = -A0 * inv_dx2;
        A1 = -A1 * inv_dx2;
        if (MODE == 0) {
            if (act) {
                *reinterpret_cast<dbl2*>(y + o) = dbl2{A0, A1};
                acc += xc.x * A0;
                acc += xc.y * A1;
            }
        } else {
            const double r0 = -1.0 * A0 + cur.bo.x, r1 = -1.0 * A1 + cur.bo.y;
            if (act) {
                if (y) *reinterpret_cast<dbl2*>(y + o) = dbl2{r0, r1};
                acc += r0 * r0;
                acc += r1 * r1;
            }
        }
        xm = xc;
        xc = dbl2{u0, u1};
        hc = U.nb ? cur.hp : dbl2{0.0, 0.0};
        fzl = dbl2{fu0, fu1};
        o += sz;
        cur = nxt;
    }
    const double s = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) partials[blk] = s;
}

// Plane-tile variant (even nx <= kOpLdsMaxNx): a workgroup owns the kOpTile
// consecutive cells [t0, t0 + kOpTile) of an xy plane and marches kz planes up
// in z. Each plane's tile and the nx cells on either side of it (the y
// neighbours) are staged in LDS, slot L holding plane cell (t0 - nx + L) mod
// nx*ny, so that every x and y neighbour of a tile cell, periodic wraps
// included, is the slot at a fixed distance: one unit-stride pass per array
// and plane replaces the three row loads of the paired kernel. Thread tid
// owns the tile's cell pairs tid + m kBlock (m < kOpPairs) and loads them
// itself (their plane-k+1 values are the up neighbours); the halo pairs are
// dealt to the threads in order. Per cell: the faces and the sum order of
// opz2_body.
constexpr int kOpPairs = 2;                     // cell pairs per thread
constexpr int kOpTile = 2 * kOpPairs * kBlock;  // cells
constexpr int kOpLdsMaxNx = 512;
constexpr int kOpHalo = kOpLdsMaxNx / kBlock;  // halo pairs per thread, at most

template <int MODE>
__device__ __forceinline__ void opz_lds_body(const Dims& d, const BcTypes& bc, int pin,
                                             double inv_dx2, int tiles, int kz,
                                             const double* __restrict__ x,
                                             const double* __restrict__ H,
                                             const double* __restrict__ b,
                                             double* __restrict__ y,
                                             double* __restrict__ partials, long blk) {
    extern __shared__ __align__(16) double op_lds[];  // x then h, kOpTile + 2 nx slots each
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long P = (long)nx * ny;
    const int slots = kOpTile + 2 * nx;
    double* lx = op_lds;
    double* lh = op_lds + slots;
    const int chunk = (int)(blk / tiles);
    const long t0 = (blk - (long)chunk * tiles) * kOpTile;
    const int tid = (int)threadIdx.x;
    const int k0 = chunk * kz;
    const int k1 = min(nz, k0 + kz);
    const int nh = nx;  // halo pairs: nx / 2 below the tile, nx / 2 above
    auto wrapP = [&](long c) { c %= P; return c < 0 ? c + P : c; };
    // plane cell of slot L (even L: the pair's left cell, same row)
    auto cell_of = [&](int L) { return wrapP(t0 - nx + L); };
    auto plane_off = [&](int k) { return (long)((k % nz + nz) % nz) * P; };
    // own pairs: tile pairs tid + m kBlock (slots nx + 2 tid + 2 m kBlock)
    int oi[kOpPairs], oj[kOpPairs];
    bool act[kOpPairs];
    long oc[kOpPairs];
#pragma unroll
    for (int m = 0; m < kOpPairs; ++m) {
        const long c = t0 + 2L * (tid + m * kBlock);
        act[m] = c < P;
        // pairs past the plane's end stage their slot's wrapped cells (the
        // periodic y neighbours of the last row) and store nothing
        oc[m] = wrapP(c);
        oj[m] = (int)(oc[m] / nx);
        oi[m] = (int)(oc[m] - (long)oj[m] * nx);
    }
    long hc_[kOpHalo];  // halo pairs of this thread: plane cells and slots
    int hs_[kOpHalo];
#pragma unroll
    for (int m = 0; m < kOpHalo; ++m) {
        const int q = tid + m * kBlock;  // halo pair index, < nh when used
        const int L = q < nh / 2 ? 2 * q : kOpTile + nx + 2 * (q - nh / 2);
        hs_[m] = q < nh ? L : -1;
        hc_[m] = q < nh ? cell_of(L) : 0;
    }
    struct Plane {
        dbl2 x[kOpPairs], h[kOpPairs], hx[kOpHalo], hh[kOpHalo], bo[kOpPairs];
    };
    auto load = [&](int k, bool halo) {
        const long po = plane_off(k);
        Plane l;
#pragma unroll
        for (int m = 0; m < kOpPairs; ++m) {
            l.x[m] = ld2(x, po + oc[m]);
            l.h[m] = ld2(H, po + oc[m]);
            l.bo[m] = MODE == 1 ? ld2(b, po + oc[m]) : dbl2{0.0, 0.0};
        }
#pragma unroll
        for (int m = 0; m < kOpHalo; ++m) {
            if (halo && hs_[m] >= 0) {
                l.hx[m] = ld2(x, po + hc_[m]);
                l.hh[m] = ld2(H, po + hc_[m]);
            }
        }
        return l;
    };
    auto stage = [&](const Plane& l) {
#pragma unroll
        for (int m = 0; m < kOpPairs; ++m) {
            const int L = nx + 2 * (tid + m * kBlock);
            *reinterpret_cast<dbl2*>(lx + L) = l.x[m];
            *reinterpret_cast<dbl2*>(lh + L) = l.h[m];
        }
#pragma unroll
        for (int m = 0; m < kOpHalo; ++m) {
            if (hs_[m] >= 0) {
                *reinterpret_cast<dbl2*>(lx + hs_[m]) = l.hx[m];
                *reinterpret_cast<dbl2*>(lh + hs_[m]) = l.hh[m];
            }
        }
    };
    // the plane below k0 (own pairs) and the first plane
    const Side D0 = side(k0 > 0, 0, bc.t[4], 0);
    dbl2 xm[kOpPairs], fzl[kOpPairs];
    Plane cur = load(k0, true);
#pragma unroll
    for (int m = 0; m < kOpPairs; ++m) {
        const long po = plane_off(k0 - 1);
        const dbl2 xd = ld2(x, po + oc[m]), hd = ld2(H, po + oc[m]);
        xm[m] = D0.nb ? xd : dbl2{0.0, 0.0};
        fzl[m] = dbl2{hface(D0, cur.h[m].x, hd.x), hface(D0, cur.h[m].y, hd.y)};
    }
    stage(cur);
    double acc = 0.0;
    for (int k = k0; k < k1; ++k) {
        // the next plane (wraps to plane 0 past the top, used if periodic); its
        // halo only if it is staged
        Plane nxt = load(k + 1, k + 1 < k1);
        __syncthreads();          // plane k staged
        const Side U = side(k < nz - 1, 0, bc.t[5], 0);
#pragma unroll
        for (int m = 0; m < kOpPairs; ++m) {
            const int i = oi[m], j = oj[m];
            const int L = nx + 2 * (tid + m * kBlock);
            const Side Wt = side(i > 0, 0, bc.t[0], 0);
            const Side E = side(i + 2 < nx, 0, bc.t[1], 0);
            const Side N = side(j < ny - 1, 0, bc.t[3], 0);
            const Side S = side(j > 0, 0, bc.t[2], 0);
            const dbl2 xc = cur.x[m], hc = cur.h[m];
            // x neighbours: the slot before the pair (west; at i = 0 the row's
            // last cell, nx - 1 slots on) and after it (east; wrap: nx - 2 back)
            const int Lw = i > 0 ? L - 1 : L + nx - 1;
            const int Le = i + 2 < nx ? L + 2 : L + 2 - nx;
            const double xw = Wt.nb ? lx[Lw] : 0.0, hw = lh[Lw];
            const double xe = E.nb ? lx[Le] : 0.0, he = lh[Le];
            const dbl2 xn = *reinterpret_cast<const dbl2*>(lx + L + nx);
            const dbl2 hn = *reinterpret_cast<const dbl2*>(lh + L + nx);
            const dbl2 xs = *reinterpret_cast<const dbl2*>(lx + L - nx);
            const dbl2 hs = *reinterpret_cast<const dbl2*>(lh + L - nx);
            const dbl2 xp = nxt.x[m], hp = nxt.h[m];
            const double fw = hface(Wt, hc.x, hw);
            const double fm = rcp_nr(hc.x + hc.y);  // between the two cells
            const double fe = hface(E, hc.y, he);
            const double fn0 = hface(N, hc.x, hn.x), fn1 = hface(N, hc.y, hn.y);
            const double fs0 = hface(S, hc.x, hs.x), fs1 = hface(S, hc.y, hs.y);
            const double fu0 = hface(U, hc.x, hp.x), fu1 = hface(U, hc.y, hp.y);
            const double n0 = N.nb ? xn.x : 0.0, n1 = N.nb ? xn.y : 0.0;
            const double s0 = S.nb ? xs.x : 0.0, s1 = S.nb ? xs.y : 0.0;
            const double u0 = U.nb ? xp.x : 0.0, u1 = U.nb ? xp.y : 0.0;
            double A0 = fw * (xc.x - xw) + fm * (xc.x - xc.y) + fs0 * (xc.x - s0) +
                        fn0 * (xc.x - n0) + fzl[m].x * (xc.x - xm[m].x) + fu0 * (xc.x - u0);
            double A1 = fm * (xc.y - xc.x) + fe * (xc.y - xe) + fs1 * (xc.y - s1) +
                        fn1 * (xc.y - n1) + fzl[m].y * (xc.y - xm[m].y) + fu1 * (xc.y - u1);
            if (pin && k == 0 && oc[m] == 0 && act[m]) {  // cell 0: x+ (fm), y+, z+
                double aC = fm;
                if (ny > 1) aC += fn0;
                if (nz > 1) aC += fu0;
                A0 += aC * xc.x;
            }
            A0 = -A0 * inv_dx2;
            A1 = -A1 * inv_dx2;
            const long o = (long)k * P + oc[m];
            if (MODE == 0) {
                if (act[m]) {
                    *reinterpret_cast<dbl2*>(y + o) = dbl2{A0, A1};
                    acc += xc.x * A0;
                    acc += xc.y * A1;
                }
            } else {
                const double r0 = -1.0 * A0 + cur.bo[m].x, r1 = -1.0 * A1 + cur.bo[m].y;
                if (act[m]) {
                    if (y) *reinterpret_cast<dbl2*>(y + o) = dbl2{r0, r1};
                    acc += r0 * r0;
                    acc += r1 * r1;
                }
            }
            xm[m] = xc;
            fzl[m] = dbl2{fu0, fu1};
        }
        __syncthreads();  // every read of plane k done
        if (k + 1 < k1) stage(nxt);
        // the next plane's own values: a non-periodic top has no neighbour there
        // (u = 0, h unused: the march ends at nz - 1)
        cur = nxt;
    }
    const double s = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) partials[blk] = s;
}

template <int MODE>
__global__ void __launch_bounds__(kBlock)
    op_lds_kernel(Dims d, BcTypes bc, int pin, double inv_dx2, int tiles, int kz, long nblocks,
                  const double* __restrict__ x, const double* __restrict__ H,
                  const double* __restrict__ b, double* __restrict__ y,
                  double* __restrict__ partials) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;  // padding block: uniform exit before the block reduction
    opz_lds_body<MODE>(d, bc, pin, inv_dx2, tiles, kz, x, H, b, y, partials, blk);
}

template <int MODE, bool PAIRS>
__global__ void __launch_bounds__(kBlock)
    op_kernel(Dims d, BcTypes bc, int pin, double inv_dx2, int tiles, int kz, long nblocks,
              const double* __restrict__ x, const double* __restrict__ H,
              const double* __restrict__ b, double* __restrict__ y,
              double* __restrict__ partials) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;  // padding block: uniform exit before the block reduction
    if (PAIRS) opz2_body<MODE>(d, bc, pin, inv_dx2, tiles, kz, x, H, b, y, partials, blk);
    else opz_body<MODE>(d, bc, pin, inv_dx2, tiles, kz, x, H, b, y, partials, blk);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) fill_t_kernel(T* __restrict__ x, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = T(0);
}

// partials[reduce_blocks(n)] of ln K: 8 cells per thread, strided by kBlock
__global__ void __launch_bounds__(kBlock)
    sum_log_kernel(const double* __restrict__ K, long n, double* __restrict__ partials) {
    constexpr int P8 = kReduceCells / kBlock;
    double kv[P8];
#pragma unroll
    for (int u = 0; u < P8; ++u)
        kv[u] = K[min((long)blockIdx.x * kReduceCells + (long)u * kBlock + threadIdx.x, n - 1)];
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < P8; ++u)
        if ((long)blockIdx.x * kReduceCells + (long)u * kBlock + threadIdx.x < n) v += log(kv[u]);
    const double s = block_sum<kBlock>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ---- precision conversions at the V-cycle boundary -----------------------------
__global__ void __launch_bounds__(kBlock)
    to_f32_kernel(const double* __restrict__ a, float* __restrict__ o, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)a[i];
}
__global__ void __launch_bounds__(kBlock)
    to_f64_kernel(const float* __restrict__ a, double* __restrict__ o, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (double)a[i];
}

// FP64 natural -> FP32 colour-split, and back
__device__ __forceinline__ void to_f32_split_body(const Dims& d, const double* __restrict__ a,
                                                  float* __restrict__ o, long blk) {
    const long n = (long)d.nx * d.ny * d.nz;
    const long e = blk * kBlock + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e % d.nx);
    const long q = e / d.nx;
    const int j = (int)(q % d.ny);
    const int k = (int)(q / d.ny);
    o[split_of(d).at(d, i, j, k)] = (float)a[e];
}
// z = x32 (natural order); with r, also the (r, z) partial of the block (the
// same blocks and in-block order as dot_partials_kernel, so the same sums)
template <bool DOT>
__device__ __forceinline__ void to_f64_split_body(const Dims& d, const float* __restrict__ a,
                                                  double* __restrict__ o,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ partials, long blk) {
    const long n = (long)d.nx * d.ny * d.nz;
    double prod = 0.0;
#pragma unroll
    for (int u = 0; u < kReduceCells / kBlock; ++u) {
        const long e = blk * kReduceCells + (long)u * kBlock + threadIdx.x;
        if (e < n) {
            const int i = (int)(e % d.nx);
            const long q = e / d.nx;
            const int j = (int)(q % d.ny);
            const int k = (int)(q / d.ny);
            const double z = (double)a[split_of(d).at(d, i, j, k)];
            o[e] = z;
            if (DOT) prod += r[e] * z;
        }
    }
    if (DOT) {
        const double s = block_sum<kBlock>(prod);
        if (threadIdx.x == 0) partials[blk] = s;
    }
}
// nx % 8 == 0: the FP32 colour-split x of the V-cycle against FP64 natural
// vectors, 8 consecutive cells of a row per thread (one float4 of each colour,
// places 4q .. 4q + 3). KIND:
//   S8_Z     o = z                      (z = (double) x32)
//   S8_ZDOT  o = z, (r, z) partials
//   S8_DOT   (r, z) partials only
//   S8_P     o = 1 z + beta o           (PCG p update, pcg.cuh:179; skipped once stopped)
//   S8_P0    o = z                      (first p), skipped once stopped
// The (r, z) partial of each block (2048 consecutive cells): its threads'
// 8-cell sums, then the block tree; partials[reduce_blocks(n)].
enum { S8_Z = 0, S8_ZDOT = 1, S8_DOT = 2, S8_P = 3, S8_P0 = 4 };
template <int KIND>
__device__ __forceinline__ void split8_body(const Dims& d, const float* __restrict__ a,
                                            double* __restrict__ o, const double* __restrict__ r,
                                            double* __restrict__ partials, double beta, long blk) {
    constexpr bool DOT = KIND == S8_ZDOT || KIND == S8_DOT;
    const long total = ((long)d.nx * d.ny * d.nz) >> 3;
    const long t = blk * kBlock + threadIdx.x;
    double prod = 0.0;
    if (t < total) {
        const int h8 = d.nx >> 3;
        const long row = t / h8;
        const int q = (int)(t - row * h8);
        const int j = (int)(row % d.ny);
        const int k = (int)(row / d.ny);
        const int c0 = (j + k) & 1;  // colour of the even cells i = 8q + 2e
        const long half = ((long)(d.nx >> 1) * d.ny) * d.nz;
        const long pl = (long)(d.nx >> 1) * row + 4L * q;
        const float4 A = *reinterpret_cast<const float4*>(a + (long)c0 * half + pl);
        const float4 B = *reinterpret_cast<const float4*>(a + (long)(c0 ^ 1) * half + pl);
        const double z[8] = {A.x, B.x, A.y, B.y, A.z, B.z, A.w, B.w};
        const long e0 = row * d.nx + 8L * q;
        double2* op = reinterpret_cast<double2*>(o + e0);
        if (KIND == S8_Z || KIND == S8_ZDOT || KIND == S8_P0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) op[e] = make_double2(z[2 * e], z[2 * e + 1]);
        } else if (KIND == S8_P) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double2 pv = op[e];
                op[e] = make_double2(1.0 * z[2 * e] + beta * pv.x, 1.0 * z[2 * e + 1] + beta * pv.y);
            }
        }
        if (DOT) {
            const double2* rp = reinterpret_cast<const double2*>(r + e0);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double2 rv = rp[e];
                prod += rv.x * z[2 * e];
                prod += rv.y * z[2 * e + 1];
            }
        }
    }
    if (DOT) {  // the block's 2048 consecutive cells: partials[reduce_blocks(n)]
        const double sb = block_sum<kBlock>(prod);
        if (threadIdx.x == 0) partials[blk] = sb;
    }
}
// x += alpha p; p = z + beta p (z = x32): the x half of pcg_update_xr_kernel
// moved to the p update, so the iteration reads and writes p once for both.
// Skipped once stopped. A workgroup covers the 2048 natural cells
// [2048 blk, 2048 blk + 2048) (nx % 8 == 0) in pairs (2m, 2m + 1) of a row,
// one cell of each colour at colour-split place m; a wave takes 4 runs of 64
// consecutive pairs, so every load and store of a wave is unit-stride.
__device__ __forceinline__ void xp_split8_body(const Dims& d, const float* __restrict__ a,
                                               double* __restrict__ x, double* __restrict__ p,
                                               double alpha, double beta, long blk) {
    const long n = (long)d.nx * d.ny * d.nz;
    const long half = n >> 1;
    const long e0 = blk * (8L * kBlock) + (long)(threadIdx.x >> 6) * 512 +
                    2L * (threadIdx.x & 63);
    if (e0 >= n) return;
    CellWalk cw(d, e0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long e = e0 + 128L * u;  // even: i = 2m
        if (e >= n) break;
        if (u > 0) cw.advance(d, 128);
        const long pl = e >> 1;  // (nx / 2) row + m
        const int c0 = (cw.j + cw.k) & 1;
        const double z0 = a[(long)c0 * half + pl], z1 = a[(long)(c0 ^ 1) * half + pl];
        const double2 pv = ldnt(p + e), xv = ldnt(x + e);
        stnt(x + e, alpha * pv.x + xv.x, alpha * pv.y + xv.y);
        stnt(p + e, 1.0 * z0 + beta * pv.x, 1.0 * z1 + beta * pv.y);
    }
}
__global__ void __launch_bounds__(kBlock)
    xp_split8_kernel(Dims d, const float* __restrict__ a, double* __restrict__ x,
                     double* __restrict__ p, const IterState* __restrict__ st) {
    if (st->stop) return;  // uniform across the grid
    const long total = ((long)d.nx * d.ny * d.nz) >> 3;
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;
    xp_split8_body(d, a, x, p, st->alpha, st->beta, blk);
}

template <int KIND>
__global__ void __launch_bounds__(kBlock)
    split8_kernel(Dims d, const float* __restrict__ a, double* __restrict__ o,
                  const double* __restrict__ r, double* __restrict__ partials,
                  const IterState* __restrict__ st) {
    if ((KIND == S8_P || KIND == S8_P0) && st->stop) return;  // uniform across the grid
    const long total = ((long)d.nx * d.ny * d.nz) >> 3;
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;  // whole padding block
    split8_body<KIND>(d, a, o, r, partials, KIND == S8_P ? st->beta : 0.0, blk);
}

__global__ void __launch_bounds__(kBlock)
    to_f32_split_kernel(Dims d, const double* __restrict__ a, float* __restrict__ o) {
    const long blk = xcd_block(((long)d.nx * d.ny * d.nz + kBlock - 1) / kBlock);
    if (blk < 0) return;
    to_f32_split_body(d, a, o, blk);
}
template <bool DOT>
__global__ void __launch_bounds__(kBlock)
    to_f64_split_kernel(Dims d, const float* __restrict__ a, double* __restrict__ o,
                        const double* __restrict__ r, double* __restrict__ partials) {
    const long blk = xcd_block(((long)d.nx * d.ny * d.nz + kReduceCells - 1) / kReduceCells);
    if (blk < 0) return;  // whole padding block, before the block reduction
    to_f64_split_body<DOT>(d, a, o, r, partials, blk);
}

// coarsest direct solve in FP64 for any x/b type: x = A^{-1} b, one wave per row
template <typename T>
__device__ __forceinline__ void gemv_body(int n, const double* __restrict__ Ainv,
                                          const T* __restrict__ b, T* __restrict__ x, long blk) {
    const int row = (int)blk * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* a = Ainv + (long)row * n;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += a[c] * (double)b[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) x[row] = (T)s;
}
template <typename T>
__global__ void __launch_bounds__(256)
    gemv_kernel(int n, const double* __restrict__ Ainv, const T* __restrict__ b,
                T* __restrict__ x) {
    gemv_body<T>(n, Ainv, b, x, blockIdx.x);
}

// ---- batched V-cycle launches -------------------------------------------------
template <typename TC>
__device__ __forceinline__ SFaces<TC> batch_sfaces(const BatchLevel& L) {
    return SFaces<TC>{(const TC*)L.fx, (const TC*)L.fy, (const TC*)L.fz,
                      (const TC*)L.ux, (const TC*)L.uy, (const TC*)L.uz};
}

template <typename TC, int MODE, bool CS>
__global__ void __launch_bounds__(kMaxSweepBlock)
    split4_batch_kernel(const BatchLevel* __restrict__ Ls, const int* __restrict__ start, int m,
                        long nblocks, int colour) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchLevel& L = Ls[e];
    if (L.st->stop) return;
    split4_body<TC, MODE, CS>(Dims{L.nx, L.ny, L.nz}, L.bc, L.pin, L.bs, L.x, L.b,
                              batch_sfaces<TC>(L), colour, L.xc, L.r64, blk - start[e]);
}

template <typename TC, bool CS>
__global__ void __launch_bounds__(kBlock)
    rr_split_batch_kernel(const BatchLevel* __restrict__ Ls, const int* __restrict__ start, int m,
                          long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchLevel& L = Ls[e];
    if (L.st->stop) return;
    // batched solves smooth at least once before each restriction
    rr_split_body<TC, CS, true>(Dims{L.nx, L.ny, L.nz}, L.bc, L.pin, L.rs, L.x, L.b,
                                batch_sfaces<TC>(L), L.bcoarse, blk - start[e]);
}

// small (single-workgroup) levels: one workgroup per entry
template <typename TC>
__global__ void __launch_bounds__(kSmallBlock)
    small_batch_kernel(const BatchLevel* __restrict__ Ls, int iters, int zero_init, int use_xc) {
    __shared__ float xs[kSmallBlock * kSmallPer];
    const BatchLevel& L = Ls[blockIdx.x];
    if (L.st->stop) return;
    small_body<float, TC>(xs, Dims{L.nx, L.ny, L.nz}, L.bc, L.pin, L.bs, L.x, L.b,
                          (const TC*)L.fx, (const TC*)L.fy, (const TC*)L.fz, iters, zero_init,
                          use_xc ? L.xc : nullptr);
}

template <typename TC>
__global__ void __launch_bounds__(kBlock)
    rr_batch_kernel(const BatchLevel* __restrict__ Ls, const int* __restrict__ start, int m,
                    long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchLevel& L = Ls[e];
    if (L.st->stop) return;
    rr_body<float, TC>(Dims{L.nx, L.ny, L.nz}, L.bc, L.pin, L.rs, L.x, L.b, (const TC*)L.fx,
                       (const TC*)L.fy, (const TC*)L.fz, L.bcoarse, blk - start[e]);
}

__global__ void __launch_bounds__(256)
    gemv_batch_kernel(const BatchLevel* __restrict__ Ls, const int* __restrict__ start, int m) {
    const int e = batch_entry(start, m, blockIdx.x);
    const BatchLevel& L = Ls[e];
    if (L.st->stop) return;
    gemv_body<float>(L.nc, L.Ainv, L.b, L.x, (long)blockIdx.x - start[e]);
}

// finest level: FP64 <-> FP32 colour-split conversions, FP64 operator
__global__ void __launch_bounds__(kBlock)
    tof32_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m,
                       long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchVec& V = Vs[e];
    if (V.st->stop) return;
    to_f32_split_body(Dims{V.nx, V.ny, V.nz}, V.r, V.b32, blk - start[e]);
}

// the (r, z) partials of each 256 cells, z = x32 not stored (p is updated
// from x32 directly); start[] counts ceil(n / 2048) workgroups per
// realisation (8 cells per thread)
__global__ void __launch_bounds__(kBlock)
    tof64_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m,
                       long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchVec& V = Vs[e];
    if (V.st->stop) return;
    split8_body<S8_DOT>(Dims{V.nx, V.ny, V.nz}, V.x32, nullptr, V.r, V.partials, 0.0,
                        blk - start[e]);
}

// PCG p update from x32 (MODE 0: p = x32 + beta p; MODE 1: p = x32) for the
// running realisations; start[] as tof64_batch_kernel's
template <int MODE>
__global__ void __launch_bounds__(kBlock)
    p_split8_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m,
                          long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchVec& V = Vs[e];
    if (V.st->stop) return;
    split8_body<MODE == 0 ? S8_P : S8_P0>(Dims{V.nx, V.ny, V.nz}, V.x32, V.p, nullptr, nullptr,
                                          MODE == 0 ? V.st->beta : 0.0, blk - start[e]);
}

// x += alpha p; p = x32 + beta p for the running realisations; start[] as
// tof64_batch_kernel's
__global__ void __launch_bounds__(kBlock)
    xp_split8_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m,
                           long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchVec& V = Vs[e];
    if (V.st->stop) return;
    xp_split8_body(Dims{V.nx, V.ny, V.nz}, V.x32, V.x, V.p, V.st->alpha, V.st->beta,
                   blk - start[e]);
}

// MODE 0: Ap = A p with (p, Ap) partials, skipped once stopped. MODE 1:
// r = rhs - A x with (r, r) partials, always (the true-residual check).
// start[] counts each realisation's operator workgroups (its op_plan geometry).
template <int MODE>
__global__ void __launch_bounds__(kBlock)
    op_batch_kernel(const BatchVec* __restrict__ Vs, const int* __restrict__ start, int m,
                    long nblocks) {
    const long blk = xcd_block(nblocks);
    if (blk < 0) return;
    const int e = batch_entry(start, m, blk);
    const BatchVec& V = Vs[e];
    if (MODE == 0 && V.st->stop) return;
    const Dims d{V.nx, V.ny, V.nz};
    const long b = blk - start[e];
    if (MODE == 0) {
        if (V.opairs == 2)
            opz_lds_body<0>(d, V.bc, V.pin, V.inv_dx2, V.otiles, V.okz, V.p, V.hk, nullptr, V.Ap,
                            V.partials, b);
        else if (V.opairs)
            opz2_body<0>(d, V.bc, V.pin, V.inv_dx2, V.otiles, V.okz, V.p, V.hk, nullptr, V.Ap,
                         V.partials, b);
        else
            opz_body<0>(d, V.bc, V.pin, V.inv_dx2, V.otiles, V.okz, V.p, V.hk, nullptr, V.Ap,
                        V.partials, b);
    } else {
        if (V.opairs == 2)
            opz_lds_body<1>(d, V.bc, V.pin, V.inv_dx2, V.otiles, V.okz, V.x, V.hk, V.rhs, V.r,
                            V.partials, b);
        else if (V.opairs)
            opz2_body<1>(d, V.bc, V.pin, V.inv_dx2, V.otiles, V.okz, V.x, V.hk, V.rhs, V.r,
                         V.partials, b);
        else
            opz_body<1>(d, V.bc, V.pin, V.inv_dx2, V.otiles, V.okz, V.x, V.hk, V.rhs, V.r,
                        V.partials, b);
    }
}


inline Dims dims(const mf3d_grid& g) { return Dims{g.nx, g.ny, g.nz}; }

template <typename TC>
const TC* cptr(const FaceSet& f, int a) {
    return static_cast<const TC*>(a == 0 ? f.fx : a == 1 ? f.fy : f.fz);
}

}  // namespace

// ===========================================================================
// host launchers
// ===========================================================================
namespace kern {

size_t face_count(const mf3d_grid& g) {
    return (size_t)(g.nx + 1) * g.ny * g.nz + (size_t)g.nx * (g.ny + 1) * g.nz +
           (size_t)g.nx * g.ny * (g.nz + 1);
}

size_t coef_bytes(CoefType t) {
    return t == COEF_F64 ? 8 : t == COEF_F32 ? 4 : 2;
}

namespace {
template <typename TC>
void launch_faces_split(hipStream_t s, const Dims& d, const BcTypes& bc, double cs,
                        const double* K, FaceSet& f, unsigned gr, double* H = nullptr) {
    if (H)
        faces_build_split_kernel<TC, true><<<gr, kBlock, 0, s>>>(
            d, bc, cs, K, (TC*)f.fx, (TC*)f.fy, (TC*)f.fz, (TC*)f.ux, (TC*)f.uy, (TC*)f.uz, H);
    else
        faces_build_split_kernel<TC, false><<<gr, kBlock, 0, s>>>(
            d, bc, cs, K, (TC*)f.fx, (TC*)f.fy, (TC*)f.fz, (TC*)f.ux, (TC*)f.uy, (TC*)f.uz,
            nullptr);
}
}  // namespace

int faces_build(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, double cs,
                const double* K, FaceSet& f) {
    const Dims d = dims(g);
    const unsigned gr = xcd_grid(cells(g));
    if (f.split) {
        if (!split_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "split faces need even sizes");
        switch (f.type) {
        case COEF_F32: launch_faces_split<float>(s, d, bc, cs, K, f, gr); break;
        case COEF_BF16: launch_faces_split<__hip_bfloat16>(s, d, bc, cs, K, f, gr); break;
        case COEF_F16: launch_faces_split<__half>(s, d, bc, cs, K, f, gr); break;
        default: return fail(MF3D_ERR_UNSUPPORTED, "split faces are 32- or 16-bit");
        }
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    }
    switch (f.type) {
    case COEF_F64:
        faces_build_kernel<double><<<gr, kBlock, 0, s>>>(d, bc, cs, K, (double*)f.fx,
                                                         (double*)f.fy, (double*)f.fz);
        break;
    case COEF_F32:
        faces_build_kernel<float><<<gr, kBlock, 0, s>>>(d, bc, cs, K, (float*)f.fx, (float*)f.fy,
                                                        (float*)f.fz);
        break;
    case COEF_BF16:
        faces_build_kernel<__hip_bfloat16><<<gr, kBlock, 0, s>>>(
            d, bc, cs, K, (__hip_bfloat16*)f.fx, (__hip_bfloat16*)f.fy, (__hip_bfloat16*)f.fz);
        break;
    case COEF_F16:
        faces_build_kernel<__half><<<gr, kBlock, 0, s>>>(d, bc, cs, K, (__half*)f.fx,
                                                         (__half*)f.fy, (__half*)f.fz);
        break;
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int faces_build_split_h(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, double cs,
                        const double* K, FaceSet& flo, double* H) {
    if (!flo.split || !split_ok(g) || !H)
        return fail(MF3D_ERR_UNSUPPORTED, "faces with H: colour-split low-precision faces");
    const Dims d = dims(g);
    const unsigned gr = xcd_grid(cells(g));
    switch (flo.type) {
    case COEF_F32: launch_faces_split<float>(s, d, bc, cs, K, flo, gr, H); break;
    case COEF_BF16: launch_faces_split<__hip_bfloat16>(s, d, bc, cs, K, flo, gr, H); break;
    case COEF_F16: launch_faces_split<__half>(s, d, bc, cs, K, flo, gr, H); break;
    default: return fail(MF3D_ERR_UNSUPPORTED, "split faces are 32- or 16-bit");
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int half_recip(hipStream_t s, const double* K, double* H, size_t n) {
    half_recip_kernel<<<grid_for(n), kBlock, 0, s>>>(K, H, (long)n);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}


namespace {

template <typename T, typename TC>
int smooth_t(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
             T* x, const T* b, const FaceSet& F, int iters, bool zero_init,
             const T* xc, const SmoothPlan& plan) {
    const Dims d = dims(g);
    const TC* FX = cptr<TC>(F, 0);
    const TC* FY = cptr<TC>(F, 1);
    const TC* FZ = cptr<TC>(F, 2);
    const size_t n = cells(g);
    const T tbs = (T)bs;
    if (iters <= 0) {
        if (zero_init) {
            fill_t_kernel<T><<<grid_for(n), kBlock, 0, s>>>(x, (long)n);
        } else if (xc) {
            prolong_kernel<T><<<xcd_grid(n), kBlock, 0, s>>>(d, xc, x);
        }
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    }
    if (plan.kind == SMOOTH_SMALL && n <= (size_t)kSmallBlock * kSmallPer) {
        small_kernel<T, TC><<<1, kSmallBlock, 0, s>>>(d, bc, pin ? 1 : 0, tbs, x, b, FX, FY, FZ,
                                                      iters, zero_init ? 1 : 0, xc);
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    }
    if (xc) prolong_kernel<T><<<xcd_grid(n), kBlock, 0, s>>>(d, xc, x);
    const unsigned gh = xcd_grid((size_t)((g.nx + 1) / 2) * g.ny * g.nz);
    for (int it = 0; it < iters; ++it)
        for (int colour = 0; colour < 2; ++colour) {
            if (it == 0 && colour == 0 && zero_init)
                colour_kernel<T, TC, true><<<gh, kBlock, 0, s>>>(d, bc, pin ? 1 : 0, tbs, x, b, FX,
                                                                 FY, FZ, colour);
            else
                colour_kernel<T, TC, false><<<gh, kBlock, 0, s>>>(d, bc, pin ? 1 : 0, tbs, x, b,
                                                                  FX, FY, FZ, colour);
        }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

template <typename T, typename TC>
int rr_t(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double rs, const T* x,
         const T* b, const FaceSet& F, T* bcoarse) {
    const size_t nc = (size_t)(g.nx / 2) * (g.ny / 2) * (g.nz / 2);
    residual_restrict_kernel<T, TC><<<xcd_grid(nc), kBlock, 0, s>>>(
        dims(g), bc, pin ? 1 : 0, (T)rs, x, b, cptr<TC>(F, 0), cptr<TC>(F, 1), cptr<TC>(F, 2),
        bcoarse);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace

int smooth_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                 double* x, const double* b, const FaceSet& F, int iters,
                 bool zero_init, const double* xc, const SmoothPlan& plan, int order,
                 int corder, const double* r64, bool red_done, const RUpd* ru,
                 const RzDot* rz) {
    if (red_done || ru || rz)
        return fail(MF3D_ERR_UNSUPPORTED, "FP64 levels do not fold the first sweep");
    if (F.type != COEF_F64) return fail(MF3D_ERR_UNSUPPORTED, "FP64 smoothing needs FP64 faces");
    if (r64) return fail(MF3D_ERR_UNSUPPORTED, "FP64 levels read b directly");
    if (order != ORDER_NATURAL || corder != ORDER_NATURAL)
        return fail(MF3D_ERR_UNSUPPORTED, "FP64 levels are stored in natural order");
    return smooth_t<double, double>(s, g, bc, pin, bs, x, b, F, iters, zero_init, xc, plan);
}

namespace {

template <typename TC>
SFaces<TC> sfaces(const FaceSet& F) {
    return SFaces<TC>{(const TC*)F.fx, (const TC*)F.fy, (const TC*)F.fz,
                      (const TC*)F.ux, (const TC*)F.uy, (const TC*)F.uz};
}

template <typename TC, int MODE, bool CS>
void launch_split(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, int pin, float bs,
                  float* x, const float* b, const SFaces<TC>& F, int colour, const float* xc,
                  int tb, const double* r64 = nullptr, const RUpd* ru = nullptr,
                  const RzDot* rz = nullptr) {
    const Dims d = dims(g);
    const size_t n = cells(g);
    if (MODE == 3 && ru) {  // nx % 8 == 0 (MODE 3 needs it)
        const size_t nb = (n / 8 + tb - 1) / tb;
        colour_split4_r_kernel<TC><<<(unsigned)((nb + 7) / 8 * 8), tb, 0, s>>>(
            d, bc, pin, bs, x, b, F, colour, r64, *ru);
        return;
    }
    if (MODE == 0 && rz) {  // nx % 8 == 0 (checked by smooth_faces)
        const size_t nb = (n / 8 + tb - 1) / tb;
        colour_split4_rz_kernel<TC><<<(unsigned)((nb + 7) / 8 * 8), tb, 0, s>>>(
            d, bc, pin, bs, x, b, F, colour, *rz);
        return;
    }
    if ((g.nx & 7) == 0) {
        const size_t nb = (n / 8 + tb - 1) / tb;
        colour_split4_kernel<TC, MODE, CS><<<(unsigned)((nb + 7) / 8 * 8), tb, 0, s>>>(
            d, bc, pin, bs, x, b, F, colour, xc, r64);
    } else {
        colour_split_kernel<TC, MODE, CS><<<xcd_grid(n / 2), kBlock, 0, s>>>(d, bc, pin, bs, x, b,
                                                                           F, colour, xc);
    }
}

template <typename TC>
int smooth_split(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                 float* x, const float* b, const FaceSet& F, int iters, bool zero_init,
                 const float* xc, bool csplit, int tb, const double* r64, bool red_done,
                 const RUpd* ru, const RzDot* rz) {
    const Dims d = dims(g);
    const SFaces<TC> SF = sfaces<TC>(F);
    const size_t n = cells(g);
    const float fbs = (float)bs;
    const int p = pin ? 1 : 0;
    if (zero_init && xc) {  // not used by the V-cycle; keep the meaning x = 0 + P xc
        fill_t_kernel<float><<<grid_for(n), kBlock, 0, s>>>(x, (long)n);
        zero_init = false;
    }
    if (iters <= 0) {
        if (zero_init) {
            fill_t_kernel<float><<<grid_for(n), kBlock, 0, s>>>(x, (long)n);
        } else if (xc) {
            if (csplit) prolong_order_kernel<true, true><<<xcd_grid(n), kBlock, 0, s>>>(d, xc, x);
            else prolong_order_kernel<true, false><<<xcd_grid(n), kBlock, 0, s>>>(d, xc, x);
        }
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    }
    for (int it = 0; it < iters; ++it)
        for (int colour = 0; colour < 2; ++colour) {
            if (it == 0 && colour == 0 && zero_init && red_done) continue;  // folded upstream
            if (it == 0 && colour == 0 && zero_init && r64)
                launch_split<TC, 3, false>(s, g, bc, p, fbs, x, b, SF, colour, nullptr, tb, r64,
                                           ru);
            else if (it == 0 && colour == 0 && zero_init)
                launch_split<TC, 1, false>(s, g, bc, p, fbs, x, b, SF, colour, nullptr, tb);
            else if (it == 0 && colour == 0 && xc && csplit)
                launch_split<TC, 2, true>(s, g, bc, p, fbs, x, b, SF, colour, xc, tb);
            else if (it == 0 && colour == 0 && xc)
                launch_split<TC, 2, false>(s, g, bc, p, fbs, x, b, SF, colour, xc, tb);
            else
                launch_split<TC, 0, false>(s, g, bc, p, fbs, x, b, SF, colour, nullptr, tb,
                                           nullptr, nullptr,
                                           it == iters - 1 && colour == 1 ? rz : nullptr);
        }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

template <typename TC, bool RED>
void launch_rr_split(hipStream_t s, const Dims& d, const BcTypes& bc, bool pin, double rs,
                     const float* x, const float* b, const FaceSet& F, float* bcoarse,
                     bool csplit, size_t nc, const FaceSet* Fc, CoarseFirst cf) {
    const SFaces<TC> sf = sfaces<TC>(F);
    const SFaces<TC> sfc = Fc ? sfaces<TC>(*Fc) : sf;
    if (csplit && cf.x)
        residual_restrict_split_kernel<TC, true, RED, true><<<xcd_grid(nc), kBlock, 0, s>>>(
            d, bc, pin ? 1 : 0, (float)rs, x, b, sf, bcoarse, sfc, cf);
    else if (csplit)
        residual_restrict_split_kernel<TC, true, RED, false><<<xcd_grid(nc), kBlock, 0, s>>>(
            d, bc, pin ? 1 : 0, (float)rs, x, b, sf, bcoarse, sfc, cf);
    else
        residual_restrict_split_kernel<TC, false, RED, false><<<xcd_grid(nc), kBlock, 0, s>>>(
            d, bc, pin ? 1 : 0, (float)rs, x, b, sf, bcoarse, sfc, cf);
}

template <typename TC>
int rr_split(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double rs,
             const float* x, const float* b, const FaceSet& F, float* bcoarse, bool csplit,
             bool red, const FaceSet* Fc, CoarseFirst cf) {
    const size_t nc = (size_t)(g.nx / 2) * (g.ny / 2) * (g.nz / 2);
    const Dims d = dims(g);
    if (red) launch_rr_split<TC, true>(s, d, bc, pin, rs, x, b, F, bcoarse, csplit, nc, Fc, cf);
    else launch_rr_split<TC, false>(s, d, bc, pin, rs, x, b, F, bcoarse, csplit, nc, Fc, cf);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace

bool split_ok(const mf3d_grid& g) {
    return g.nx >= 2 && g.ny >= 2 && g.nz >= 2 && !(g.nx & 1) && !(g.ny & 1) && !(g.nz & 1);
}

int smooth_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                 float* x, const float* b, const FaceSet& F, int iters,
                 bool zero_init, const float* xc, const SmoothPlan& plan, int order, int corder,
                 const double* r64, bool red_done, const RUpd* ru, const RzDot* rz) {
    if (ru && !r64) return fail(MF3D_ERR_UNSUPPORTED, "the r update rides on the b-from-r sweep");
    if (rz && !(order == ORDER_SPLIT && (g.nx & 7) == 0 && iters > 0))
        return fail(MF3D_ERR_UNSUPPORTED, "(r, z) partials need a split4 sweep to end on");
    if (r64 && !(order == ORDER_SPLIT && (g.nx & 7) == 0 && zero_init && iters > 0 && !xc))
        return fail(MF3D_ERR_UNSUPPORTED, "b from FP64 needs a zero-start split4 sweep");
    if (red_done && !(order == ORDER_SPLIT && zero_init && iters > 0 && !xc && !r64))
        return fail(MF3D_ERR_UNSUPPORTED, "a folded first half-sweep needs a zero-start split sweep");
    if ((order == ORDER_SPLIT) != F.split)
        return fail(MF3D_ERR_UNSUPPORTED, "vector order and face layout differ");
    if (order == ORDER_SPLIT) {
        if (!split_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "colour-split order needs even sizes");
        const bool cs = corder == ORDER_SPLIT;
        const int tb = plan.block;
        if (tb < 64 || tb > kMaxSweepBlock || (tb & 63))
            return fail(MF3D_ERR_INVALID_ARGUMENT, "sweep block must be 64..512, a multiple of 64");
        switch (F.type) {
        case COEF_F32:
            return smooth_split<float>(s, g, bc, pin, bs, x, b, F, iters, zero_init, xc, cs, tb,
                                      r64, red_done, ru, rz);
        case COEF_BF16:
            return smooth_split<__hip_bfloat16>(s, g, bc, pin, bs, x, b, F, iters, zero_init, xc,
                                                cs, tb, r64, red_done, ru, rz);
        case COEF_F16:
            return smooth_split<__half>(s, g, bc, pin, bs, x, b, F, iters, zero_init, xc, cs, tb,
                                       r64, red_done, ru, rz);
        default:
            return fail(MF3D_ERR_UNSUPPORTED, "FP32 smoothing needs 32- or 16-bit faces");
        }
    }
    if (corder == ORDER_SPLIT)
        return fail(MF3D_ERR_UNSUPPORTED, "a natural-order level cannot take a split coarse level");
    switch (F.type) {
    case COEF_F32:
        return smooth_t<float, float>(s, g, bc, pin, bs, x, b, F, iters, zero_init, xc, plan);
    case COEF_BF16:
        return smooth_t<float, __hip_bfloat16>(s, g, bc, pin, bs, x, b, F, iters, zero_init,
                                               xc, plan);
    case COEF_F16:
        return smooth_t<float, __half>(s, g, bc, pin, bs, x, b, F, iters, zero_init, xc,
                                       plan);
    default:
        return fail(MF3D_ERR_UNSUPPORTED, "FP32 smoothing needs 32- or 16-bit faces");
    }
}

int residual_restrict_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                            double rs, const double* x, const double* b, const FaceSet& F,
                            double* bcoarse, int order, int corder, bool red_only,
                            double* xcoarse, const FaceSet* Fc, double bsc) {
    (void)red_only;  // natural-order FP64 levels evaluate every child
    (void)Fc;
    (void)bsc;
    if (xcoarse) return fail(MF3D_ERR_UNSUPPORTED, "FP64 levels do not fold the first sweep");
    if (order != ORDER_NATURAL || corder != ORDER_NATURAL)
        return fail(MF3D_ERR_UNSUPPORTED, "FP64 levels are stored in natural order");
    return rr_t<double, double>(s, g, bc, pin, rs, x, b, F, bcoarse);
}

int residual_restrict_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                            double rs, const float* x, const float* b, const FaceSet& F,
                            float* bcoarse, int order, int corder, bool red_only,
                            float* xcoarse, const FaceSet* Fc, double bsc) {
    if ((order == ORDER_SPLIT) != F.split)
        return fail(MF3D_ERR_UNSUPPORTED, "vector order and face layout differ");
    if (order == ORDER_SPLIT) {
        if (!split_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "colour-split order needs even sizes");
        const bool cs = corder == ORDER_SPLIT;
        if (xcoarse && (!cs || !Fc || !Fc->split || Fc->type != F.type))
            return fail(MF3D_ERR_UNSUPPORTED, "folded coarse sweep needs a split coarse level");
        const CoarseFirst cf{xcoarse, (float)bsc};
        switch (F.type) {
        case COEF_F32:
            return rr_split<float>(s, g, bc, pin, rs, x, b, F, bcoarse, cs, red_only, Fc, cf);
        case COEF_BF16:
            return rr_split<__hip_bfloat16>(s, g, bc, pin, rs, x, b, F, bcoarse, cs, red_only, Fc,
                                            cf);
        case COEF_F16:
            return rr_split<__half>(s, g, bc, pin, rs, x, b, F, bcoarse, cs, red_only, Fc, cf);
        default: return fail(MF3D_ERR_UNSUPPORTED, "FP32 residual needs 32- or 16-bit faces");
        }
    }
    if (corder == ORDER_SPLIT)
        return fail(MF3D_ERR_UNSUPPORTED, "a natural-order level cannot take a split coarse level");
    switch (F.type) {
    case COEF_F32: return rr_t<float, float>(s, g, bc, pin, rs, x, b, F, bcoarse);
    case COEF_BF16: return rr_t<float, __hip_bfloat16>(s, g, bc, pin, rs, x, b, F, bcoarse);
    case COEF_F16: return rr_t<float, __half>(s, g, bc, pin, rs, x, b, F, bcoarse);
    default: return fail(MF3D_ERR_UNSUPPORTED, "FP32 residual needs 32- or 16-bit faces");
    }
}

OpGeom op_geom(const mf3d_grid& g, bool pairs, int kz_max) {
    OpGeom o;
    o.pairs = pairs && !(g.nx & 1);
    const long P = (long)g.nx * g.ny;
    const long per = o.pairs ? 2 * kBlock : kBlock;
    o.tiles = (int)((P + per - 1) / per);
    // z-chunks of up to kz_max planes, shortened while the launch has fewer
    // than 2048 workgroups (8 per CU)
    o.kz = (int)std::max(1L, std::min((long)kz_max, (long)o.tiles * g.nz / 2048));
    o.chunks = (g.nz + o.kz - 1) / o.kz;
    return o;
}

// the plane-tile variant: kOpTile cells per workgroup
static OpGeom op_tile_geom(const mf3d_grid& g, int kz_max) {
    OpGeom o;
    o.pairs = true;
    const long P = (long)g.nx * g.ny;
    o.tiles = (int)((P + kOpTile - 1) / kOpTile);
    o.kz = (int)std::max(1L, std::min((long)kz_max, (long)o.tiles * g.nz / 1024));
    o.chunks = (g.nz + o.kz - 1) / o.kz;
    return o;
}

OpGeom op_plan(const mf3d_grid& g, bool pairs, bool allow_tile, int kz_max) {
    if (pairs && allow_tile && !(g.nx & 1) && g.nx <= kOpLdsMaxNx) {
        OpGeom o = op_tile_geom(g, kz_max);
        o.tile = true;
        return o;
    }
    return op_geom(g, pairs, kz_max);
}

size_t op_tile_lds(const mf3d_grid& g) {
    return 2 * (size_t)(kOpTile + 2 * g.nx) * sizeof(double);
}

unsigned op_partials(const mf3d_grid& g) {
    const OpGeom a = op_geom(g, false, 1), b = op_geom(g, true, 1), c = op_tile_geom(g, 1);
    return (unsigned)std::max({(long)a.tiles * a.chunks, (long)b.tiles * b.chunks,
                               (long)c.tiles * c.chunks});
}

bool op_pairs_ok(const mf3d_grid& g, const void* x, const void* b, const void* y,
                 const void* H) {
    auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    return !(g.nx & 1) && al(x) && al(b) && al(y) && al(H);
}

int op_apply(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, const double* H,
             int mode, const double* x, const double* b, double* y, double* partials,
             unsigned* np, bool allow_pairs, bool allow_tile, int kz_max) {
    if (kz_max < 1 || kz_max > kOpChunkMax)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "operator chunk depth out of range");
    const double inv_dx2 = 1.0 / (g.dx * g.dx);
    const bool pairs = allow_pairs && op_pairs_ok(g, x, b, y, H);
    const Dims d = dims(g);
    const int pn = pin ? 1 : 0;
    const OpGeom o = op_plan(g, pairs, allow_tile, kz_max);
    const long nb = (long)o.tiles * o.chunks;
    const unsigned grid = xcd_round((size_t)nb);
    if (o.tile) {
        const size_t lds = op_tile_lds(g);
        if (mode == 0)
            op_lds_kernel<0><<<grid, kBlock, lds, s>>>(d, bc, pn, inv_dx2, o.tiles, o.kz, nb, x,
                                                       H, b, y, partials);
        else
            op_lds_kernel<1><<<grid, kBlock, lds, s>>>(d, bc, pn, inv_dx2, o.tiles, o.kz, nb, x,
                                                       H, b, y, partials);
        MF3D_HIP_TRY(hipGetLastError());
        if (np) *np = (unsigned)nb;
        return MF3D_OK;
    }
    if (mode == 0) {
        if (o.pairs)
            op_kernel<0, true><<<grid, kBlock, 0, s>>>(d, bc, pn, inv_dx2, o.tiles, o.kz, nb, x, H,
                                                       b, y, partials);
        else
            op_kernel<0, false><<<grid, kBlock, 0, s>>>(d, bc, pn, inv_dx2, o.tiles, o.kz, nb, x,
                                                        H, b, y, partials);
    } else {
        if (o.pairs)
            op_kernel<1, true><<<grid, kBlock, 0, s>>>(d, bc, pn, inv_dx2, o.tiles, o.kz, nb, x, H,
                                                       b, y, partials);
        else
            op_kernel<1, false><<<grid, kBlock, 0, s>>>(d, bc, pn, inv_dx2, o.tiles, o.kz, nb, x,
                                                        H, b, y, partials);
    }
    MF3D_HIP_TRY(hipGetLastError());
    if (np) *np = (unsigned)nb;
    return MF3D_OK;
}

int fill_f32(hipStream_t s, float* x, size_t n, float v) {
    (void)v;
    fill_t_kernel<float><<<grid_for(n), kBlock, 0, s>>>(x, (long)n);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int sum_log_partials(hipStream_t s, const double* K, size_t n, double* partials) {
    sum_log_kernel<<<reduce_blocks(n), kBlock, 0, s>>>(K, (long)n, partials);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int to_f32(hipStream_t s, const double* a, float* o, size_t n) {
    to_f32_kernel<<<grid_for(n), kBlock, 0, s>>>(a, o, (long)n);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int to_f64(hipStream_t s, const float* a, double* o, size_t n) {
    to_f64_kernel<<<grid_for(n), kBlock, 0, s>>>(a, o, (long)n);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int to_f32(hipStream_t s, const double* a, float* o, const mf3d_grid& g, int order) {
    if (order != ORDER_SPLIT) return to_f32(s, a, o, cells(g));
    if (!split_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "colour-split order needs even sizes");
    to_f32_split_kernel<<<xcd_grid(cells(g)), kBlock, 0, s>>>(dims(g), a, o);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

// grid of split8_kernel (8 cells per thread, padded to a multiple of 8 blocks)
unsigned split8_grid(const mf3d_grid& g) {
    return (unsigned)((((cells(g) / 8 + kBlock - 1) / kBlock + 7) / 8) * 8);
}

int to_f64(hipStream_t s, const float* a, double* o, const mf3d_grid& g, int order,
           const double* r, double* partials) {
    if (order != ORDER_SPLIT) {
        MF3D_TRY(to_f64(s, a, o, cells(g)));
        return r ? dot_partials(s, r, o, cells(g), partials) : MF3D_OK;
    }
    if (!split_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "colour-split order needs even sizes");
    if (!(g.nx & 7)) {
        const unsigned gr = split8_grid(g);
        if (r)
            split8_kernel<S8_ZDOT><<<gr, kBlock, 0, s>>>(dims(g), a, o, r, partials, nullptr);
        else
            split8_kernel<S8_Z><<<gr, kBlock, 0, s>>>(dims(g), a, o, nullptr, nullptr, nullptr);
        MF3D_HIP_TRY(hipGetLastError());
        return MF3D_OK;
    }
    const unsigned gr = ((reduce_blocks(cells(g)) + 7u) / 8u) * 8u;
    if (r)
        to_f64_split_kernel<true><<<gr, kBlock, 0, s>>>(dims(g), a, o, r, partials);
    else
        to_f64_split_kernel<false><<<gr, kBlock, 0, s>>>(dims(g), a, o, nullptr, nullptr);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int rz_split8(hipStream_t s, const float* x32, const mf3d_grid& g, const double* r,
              double* partials) {
    if (g.nx & 7) return fail(MF3D_ERR_UNSUPPORTED, "8-cell conversion needs nx % 8 == 0");
    split8_kernel<S8_DOT><<<split8_grid(g), kBlock, 0, s>>>(dims(g), x32, nullptr, r, partials,
                                                            nullptr);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int update_xp_split8(hipStream_t s, const IterState* st, const float* x32, const mf3d_grid& g,
                     double* x, double* p) {
    if (g.nx & 7) return fail(MF3D_ERR_UNSUPPORTED, "8-cell conversion needs nx % 8 == 0");
    xp_split8_kernel<<<split8_grid(g), kBlock, 0, s>>>(dims(g), x32, x, p, st);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int update_p_split8(hipStream_t s, const IterState* st, const float* x32, const mf3d_grid& g,
                    double* p, bool first) {
    if (g.nx & 7) return fail(MF3D_ERR_UNSUPPORTED, "8-cell conversion needs nx % 8 == 0");
    if (first)
        split8_kernel<S8_P0><<<split8_grid(g), kBlock, 0, s>>>(dims(g), x32, p, nullptr, nullptr,
                                                               st);
    else
        split8_kernel<S8_P><<<split8_grid(g), kBlock, 0, s>>>(dims(g), x32, p, nullptr, nullptr,
                                                              st);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int gemv(hipStream_t s, int n, const double* Ainv, const double* b, double* x) {
    gemv_kernel<double><<<(n + 3) / 4, 256, 0, s>>>(n, Ainv, b, x);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int gemv(hipStream_t s, int n, const double* Ainv, const float* b, float* x) {
    gemv_kernel<float><<<(n + 3) / 4, 256, 0, s>>>(n, Ainv, b, x);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

// ---- batched launchers ----------------------------------------------------------
namespace {
template <typename TC>
void launch_split4_batch(hipStream_t s, int mode, bool cs, const BatchLevel* L, const int* start,
                         int m, long nblocks, int colour, int tb) {
    const unsigned g = (unsigned)(((nblocks + 7) / 8) * 8);
    if (mode == 3)
        split4_batch_kernel<TC, 3, false><<<g, tb, 0, s>>>(L, start, m, nblocks, colour);
    else if (mode == 1)
        split4_batch_kernel<TC, 1, false><<<g, tb, 0, s>>>(L, start, m, nblocks, colour);
    else if (mode == 2 && cs)
        split4_batch_kernel<TC, 2, true><<<g, tb, 0, s>>>(L, start, m, nblocks, colour);
    else if (mode == 2)
        split4_batch_kernel<TC, 2, false><<<g, tb, 0, s>>>(L, start, m, nblocks, colour);
    else
        split4_batch_kernel<TC, 0, false><<<g, tb, 0, s>>>(L, start, m, nblocks, colour);
}
template <typename TC>
void launch_rr_split_batch(hipStream_t s, bool cs, const BatchLevel* L, const int* start, int m,
                           long nblocks) {
    const unsigned g = (unsigned)(((nblocks + 7) / 8) * 8);
    if (cs) rr_split_batch_kernel<TC, true><<<g, kBlock, 0, s>>>(L, start, m, nblocks);
    else rr_split_batch_kernel<TC, false><<<g, kBlock, 0, s>>>(L, start, m, nblocks);
}
}  // namespace

long batch_blocks(long threads) { return (threads + kBlock - 1) / kBlock; }

int batch_sweep(hipStream_t s, CoefType t, int mode, bool cs, const BatchLevel* L,
                const int* start, int m, long nblocks, int colour, int tb) {
    if (tb < 64 || tb > kMaxSweepBlock || (tb & 63))
        return fail(MF3D_ERR_INVALID_ARGUMENT, "sweep block must be 64..512, a multiple of 64");
    switch (t) {
    case COEF_F32: launch_split4_batch<float>(s, mode, cs, L, start, m, nblocks, colour, tb); break;
    case COEF_BF16:
        launch_split4_batch<__hip_bfloat16>(s, mode, cs, L, start, m, nblocks, colour, tb);
        break;
    case COEF_F16: launch_split4_batch<__half>(s, mode, cs, L, start, m, nblocks, colour, tb); break;
    default: return fail(MF3D_ERR_UNSUPPORTED, "batched sweep: coefficient type");
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_residual_restrict_split(hipStream_t s, CoefType t, bool cs, const BatchLevel* L,
                                  const int* start, int m, long nblocks) {
    switch (t) {
    case COEF_F32: launch_rr_split_batch<float>(s, cs, L, start, m, nblocks); break;
    case COEF_BF16: launch_rr_split_batch<__hip_bfloat16>(s, cs, L, start, m, nblocks); break;
    case COEF_F16: launch_rr_split_batch<__half>(s, cs, L, start, m, nblocks); break;
    default: return fail(MF3D_ERR_UNSUPPORTED, "batched residual: coefficient type");
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_small(hipStream_t s, CoefType t, const BatchLevel* L, int m, int iters, bool zero_init,
                bool use_xc) {
    switch (t) {
    case COEF_F32:
        small_batch_kernel<float><<<m, kSmallBlock, 0, s>>>(L, iters, zero_init, use_xc);
        break;
    case COEF_BF16:
        small_batch_kernel<__hip_bfloat16><<<m, kSmallBlock, 0, s>>>(L, iters, zero_init, use_xc);
        break;
    case COEF_F16:
        small_batch_kernel<__half><<<m, kSmallBlock, 0, s>>>(L, iters, zero_init, use_xc);
        break;
    default: return fail(MF3D_ERR_UNSUPPORTED, "batched small level: coefficient type");
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_residual_restrict(hipStream_t s, CoefType t, const BatchLevel* L, const int* start,
                            int m, long nblocks) {
    const unsigned g = (unsigned)(((nblocks + 7) / 8) * 8);
    switch (t) {
    case COEF_F32: rr_batch_kernel<float><<<g, kBlock, 0, s>>>(L, start, m, nblocks); break;
    case COEF_BF16:
        rr_batch_kernel<__hip_bfloat16><<<g, kBlock, 0, s>>>(L, start, m, nblocks);
        break;
    case COEF_F16: rr_batch_kernel<__half><<<g, kBlock, 0, s>>>(L, start, m, nblocks); break;
    default: return fail(MF3D_ERR_UNSUPPORTED, "batched residual: coefficient type");
    }
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_gemv(hipStream_t s, const BatchLevel* L, const int* start, int m, long nblocks) {
    gemv_batch_kernel<<<(unsigned)nblocks, 256, 0, s>>>(L, start, m);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_to_f32(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks) {
    tof32_batch_kernel<<<(unsigned)(((nblocks + 7) / 8) * 8), kBlock, 0, s>>>(V, start, m, nblocks);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_to_f64(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks) {
    tof64_batch_kernel<<<(unsigned)(((nblocks + 7) / 8) * 8), kBlock, 0, s>>>(V, start, m, nblocks);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_update_p_split8(hipStream_t s, int mode, const BatchVec* V, const int* start, int m,
                          long nblocks) {
    const unsigned g = (unsigned)(((nblocks + 7) / 8) * 8);
    if (mode == 0) p_split8_batch_kernel<0><<<g, kBlock, 0, s>>>(V, start, m, nblocks);
    else p_split8_batch_kernel<1><<<g, kBlock, 0, s>>>(V, start, m, nblocks);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_update_xp_split8(hipStream_t s, const BatchVec* V, const int* start, int m,
                           long nblocks) {
    xp_split8_batch_kernel<<<(unsigned)(((nblocks + 7) / 8) * 8), kBlock, 0, s>>>(V, start, m,
                                                                                 nblocks);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int batch_op(hipStream_t s, int mode, const BatchVec* V, const int* start, int m, long nblocks,
             size_t lds) {
    const unsigned g = (unsigned)(((nblocks + 7) / 8) * 8);
    if (mode == 0) op_batch_kernel<0><<<g, kBlock, lds, s>>>(V, start, m, nblocks);
    else op_batch_kernel<1><<<g, kBlock, lds, s>>>(V, start, m, nblocks);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace kern
}  // namespace mf3d
