// Fine-level red-black Gauss-Seidel: one pass per iteration (FP64), gfx950.
//
// A workgroup owns a 64 x 16 column tile (one thread per (x, y) column) and
// marches a z-chunk. At plane k it updates the red cells of plane k on the
// tile minus a 1-column rim, then the black cells of plane k-1 on the tile
// minus a 2-column rim, and writes plane k-1. Red needs only old black values
// and black only new red values, so each iteration reads x, K and b once and
// writes x once (32 B/cell instead of 2 x 32 B for two colour passes). The
// red rim is recomputed by neighbouring tiles; reads come from x_in and writes
// go to x_out, so overlapping tiles never see each other's updates (the
// result equals the reference's red pass followed by its black pass,
// gsrb_3d.cu:392-575, because cells of one colour do not depend on each
// other).
//
// Planes of x and 1/K live in an LDS ring of two; the column's own z-window
// lives in registers. Workgroups are ordered so that each XCD (blockIdx % 8)
// gets a contiguous run of neighbouring tiles, keeping halo re-reads in that
// XCD's L2.
//
// Variants: ZERO (x_in is zero: first sweep of a V-cycle level) and PROLONG
// (x_in + P x_coarse, prolong_3d.cu:30-621 folded into the first post-sweep).

#include <type_traits>

#include "kernels.hpp"

namespace mf3d {
namespace {

constexpr int TX = 64, TY = 16, HALO = 2;
constexpr int OX = TX - 2 * HALO, OY = TY - 2 * HALO;

struct MarchGeom {
    int tiles_x, tiles_y, chunks, zchunk, nblocks, per_xcd;
};

// Workgroup barrier that orders LDS only: global loads stay in flight across
// it (__syncthreads also waits for every outstanding global load, which would
// expose the next-plane prefetch at each barrier).
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ double bc_coef(const BcTypes& bc, int face, double Kc) {
    return bc.t[face] == MF3D_BC_DIRICHLET ? 2.0 * Kc : 0.0;
}

// One GSRB cell update from its six face coefficients (x+, x-, y+, y-, z+,
// z-; a boundary face has value 0 and the bc_helpers coefficient,
// bc_stencil_helpers.cuh:33-267): (sum f x_nb - b dx^2) / sum f, summed in
// the order of gsrb_interior_kernel (gsrb_3d.cu:53-69).
__device__ __forceinline__ double rb_cell(const double (&f)[6], const double (&xv)[6], double b,
                                          double dx2, bool pin0) {
    double result = xv[0] * f[0];
    result += xv[1] * f[1];
    result += xv[2] * f[2];
    result += xv[3] * f[3];
    result += xv[4] * f[4];
    result += xv[5] * f[5];
    double aC = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
    if (pin0) aC *= 2.0;
    return (result - b * dx2) / aC;
}

template <bool ZERO, bool PROLONG>
__global__ void __launch_bounds__(TX * TY)
    gsrb_march_kernel(int nx, int ny, int nz, double dx2, BcTypes bc, int pin, MarchGeom geo,
                      const double* __restrict__ xin, double* __restrict__ xout,
                      const double* __restrict__ b, const double* __restrict__ K,
                      const double* __restrict__ xc) {
    __shared__ double sik[TY][TX];      // 1/K of the plane being faced
    __shared__ double sx[2][TY][TX];    // x of planes k, k-1 (red values new)
    __shared__ double sfx[2][TY][TX];   // x-face coefficient to the i+1 neighbour
    __shared__ double sfy[2][TY][TX];   // y-face coefficient to the j+1 neighbour

    // XCD-aware tile order: blocks on one XCD take a contiguous run of tiles
    const int bid = blockIdx.x;
    const int t = (bid & 7) * geo.per_xcd + (bid >> 3);
    if (t >= geo.nblocks) return;  // padding blocks (uniform exit, no barrier reached)
    const int tile_x = t % geo.tiles_x;
    const int rem = t / geo.tiles_x;
    const int tile_y = rem % geo.tiles_y;
    const int chunk = rem / geo.tiles_y;
    const int z0 = chunk * geo.zchunk;
    const int z1 = min(nz, z0 + geo.zchunk);

    const int tx = threadIdx.x, ty = threadIdx.y;
    int i = tile_x * OX - HALO + tx;
    int j = tile_y * OY - HALO + ty;
    bool active = true;
    if (i < 0 || i >= nx) {
        if (bc.t[0] == MF3D_BC_PERIODIC) i = (i + nx) % nx; else active = false;
    }
    if (j < 0 || j >= ny) {
        if (bc.t[2] == MF3D_BC_PERIODIC) j = (j + ny) % ny; else active = false;
    }
    const bool in_red = tx >= 1 && tx < TX - 1 && ty >= 1 && ty < TY - 1;
    const bool in_out = tx >= HALO && tx < TX - HALO && ty >= HALO && ty < TY - HALO &&
                        (tile_x * OX - HALO + tx) < nx && (tile_y * OY - HALO + ty) < ny;
    const bool face_x = tx < TX - 1;  // the i+1 neighbour is in the tile
    const bool face_y = ty < TY - 1;
    // non-periodic domain edges of this column
    const bool bxp = (i == nx - 1) && bc.t[1] != MF3D_BC_PERIODIC;
    const bool bxm = (i == 0) && bc.t[0] != MF3D_BC_PERIODIC;
    const bool byp = (j == ny - 1) && bc.t[3] != MF3D_BC_PERIODIC;
    const bool bym = (j == 0) && bc.t[2] != MF3D_BC_PERIODIC;
    const long plane = (long)nx * ny;
    const long col = (long)i + (long)nx * j;
    const int NXc = nx >> 1, NYc = ny >> 1;
    const long ccol = (long)(i >> 1) + (long)NXc * (j >> 1);
    const long cplane = (long)NXc * NYc;

    auto load = [&](int p, double& xv, double& Kv, double& bv) {
        if (active && p >= 0 && p < nz) {
            const long o = col + plane * p;
            Kv = K[o];
            bv = b[o];
            double v = 0.0;
            if (!ZERO) v = xin[o];
            if (PROLONG) v += xc[ccol + cplane * (p >> 1)];
            xv = v;
        } else {
            Kv = 1.0;
            bv = 0.0;
            xv = 0.0;
        }
    };

    // Register windows over planes, slot of plane p = (p - kstart) mod 5. The
    // march is unrolled by 5 so every slot index is a compile-time constant:
    // loads land in their final registers and nothing is shifted, so a
    // prefetch (issued two planes ahead) is only waited for where it is used.
    // FZ[slot p] = z-face (p, p+1); FX/FY[slot p] = own in-plane faces of p.
    const int kstart = z0 - 1;
    double X[5], KK[5], B[5], IK[5], FZ[5], FX[5], FY[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        X[c] = 0.0; KK[c] = 1.0; B[c] = 0.0; IK[c] = 1.0; FZ[c] = 0.0; FX[c] = 0.0; FY[c] = 0.0;
    }
    load(kstart - 1, X[4], KK[4], B[4]);
    load(kstart, X[0], KK[0], B[0]);
    load(kstart + 1, X[1], KK[1], B[1]);
    IK[4] = 1.0 / KK[4];
    IK[0] = 1.0 / KK[0];
    FZ[4] = 2.0 / (IK[4] + IK[0]);

    auto step = [&](auto cc, int k) {
        constexpr int c = decltype(cc)::value;
        constexpr int m2 = (c + 3) % 5, m1 = (c + 4) % 5, p1 = (c + 1) % 5, p2 = (c + 2) % 5;
        load(k + 2, X[p2], KK[p2], B[p2]);  // consumed two steps later
        // z-face (k, k+1) and 1/K of plane k+1: every lane, no masking
        IK[p1] = 1.0 / KK[p1];
        FZ[c] = 2.0 / (IK[c] + IK[p1]);
        sik[ty][tx] = IK[c];
        lds_sync();
        // in-plane faces of plane k, published with the plane's x values
        const int s = k & 1;
        FX[c] = face_x ? 2.0 / (IK[c] + sik[ty][tx + 1]) : 0.0;
        FY[c] = face_y ? 2.0 / (IK[c] + sik[ty + 1][tx]) : 0.0;
        sfx[s][ty][tx] = FX[c];
        sfy[s][ty][tx] = FY[c];
        sx[s][ty][tx] = X[c];
        lds_sync();
        // red cells of plane k on the tile minus its outer ring
        if (k >= 0 && k < nz && in_red && active && ((i + j + k) & 1) == 0) {
            const bool bzp = k == nz - 1, bzm = k == 0;
            const double Kc = KK[c];
            const double f[6] = {bxp ? bc_coef(bc, 1, Kc) : FX[c],
                                 bxm ? bc_coef(bc, 0, Kc) : sfx[s][ty][tx - 1],
                                 byp ? bc_coef(bc, 3, Kc) : FY[c],
                                 bym ? bc_coef(bc, 2, Kc) : sfy[s][ty - 1][tx],
                                 bzp ? bc_coef(bc, 5, Kc) : FZ[c],
                                 bzm ? bc_coef(bc, 4, Kc) : FZ[m1]};
            const double xv[6] = {bxp ? 0.0 : sx[s][ty][tx + 1], bxm ? 0.0 : sx[s][ty][tx - 1],
                                  byp ? 0.0 : sx[s][ty + 1][tx], bym ? 0.0 : sx[s][ty - 1][tx],
                                  bzp ? 0.0 : X[p1], bzm ? 0.0 : X[m1]};
            X[c] = rb_cell(f, xv, B[c], dx2, pin && i == 0 && j == 0 && k == 0);
            sx[s][ty][tx] = X[c];
        }
        lds_sync();
        // black cells of plane p = k-1 on the output tile, then write plane p
        const int p = k - 1;
        if (p >= z0 && p < z1 && active && in_out) {
            if (((i + j + p) & 1) == 1) {
                const int q = p & 1;
                const bool bzp = p == nz - 1, bzm = p == 0;
                const double Kc = KK[m1];
                const double f[6] = {bxp ? bc_coef(bc, 1, Kc) : FX[m1],
                                     bxm ? bc_coef(bc, 0, Kc) : sfx[q][ty][tx - 1],
                                     byp ? bc_coef(bc, 3, Kc) : FY[m1],
                                     bym ? bc_coef(bc, 2, Kc) : sfy[q][ty - 1][tx],
                                     bzp ? bc_coef(bc, 5, Kc) : FZ[m1],
                                     bzm ? bc_coef(bc, 4, Kc) : FZ[m2]};
                const double xv[6] = {bxp ? 0.0 : sx[q][ty][tx + 1],
                                      bxm ? 0.0 : sx[q][ty][tx - 1],
                                      byp ? 0.0 : sx[q][ty + 1][tx],
                                      bym ? 0.0 : sx[q][ty - 1][tx], bzp ? 0.0 : X[c],
                                      bzm ? 0.0 : X[m2]};
                X[m1] = rb_cell(f, xv, B[m1], dx2, pin && i == 0 && j == 0 && p == 0);
            }
            xout[col + plane * p] = X[m1];
        }
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    for (int k = kstart; k <= z1; k += 5) {
        step(I0{}, k);
        if (k + 1 > z1) break;
        step(I1{}, k + 1);
        if (k + 2 > z1) break;
        step(I2{}, k + 2);
        if (k + 3 > z1) break;
        step(I3{}, k + 3);
        if (k + 4 > z1) break;
        step(I4{}, k + 4);
    }
}

MarchGeom geometry(const mf3d_grid& g, int zchunk) {
    MarchGeom m;
    m.tiles_x = (g.nx + OX - 1) / OX;
    m.tiles_y = (g.ny + OY - 1) / OY;
    m.zchunk = zchunk;
    m.chunks = (g.nz + zchunk - 1) / zchunk;
    m.nblocks = m.tiles_x * m.tiles_y * m.chunks;
    m.per_xcd = (m.nblocks + 7) / 8;
    return m;
}

int default_zchunk(const mf3d_grid& g) {
    // enough workgroups to cover 256 CUs several times, chunks of >= 8 planes
    const long tiles = (long)((g.nx + OX - 1) / OX) * ((g.ny + OY - 1) / OY);
    int z = 32;
    while (z > 8 && tiles * ((g.nz + z - 1) / z) < 1024) z >>= 1;
    return z;
}

template <bool ZERO, bool PROLONG>
int launch_march(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                 const double* xin, double* xout, const double* b, const double* K,
                 const double* xc, int zchunk) {
    const MarchGeom geo = geometry(g, zchunk);
    gsrb_march_kernel<ZERO, PROLONG><<<geo.per_xcd * 8, dim3(TX, TY), 0, s>>>(
        g.nx, g.ny, g.nz, g.dx * g.dx, bc, pin ? 1 : 0, geo, xin, xout, b, K, xc);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace

namespace kern {

bool fast_ok(const mf3d_grid& g, const BcTypes& bc) {
    if (g.nx < 32 || g.ny < 16 || g.nz < 8) return false;
    if (bc.t[4] == MF3D_BC_PERIODIC) return false;  // the z march does not wrap
    if ((bc.t[0] == MF3D_BC_PERIODIC && (g.nx & 1)) || (bc.t[2] == MF3D_BC_PERIODIC && (g.ny & 1)))
        return false;  // odd periodic sizes break the two-colouring across the wrap
    return true;
}

// iters GSRB iterations with the march kernel. Result lands in x; tmp is an
// n-element scratch. zero_init: x is zero on entry. xc (nullable): coarse
// correction to prolong-add into x first (PROLONG variant, x, xc 2:1).
int gsrb_march(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double* x,
               double* tmp, const double* b, const double* K, int iters, bool zero_init,
               const double* xc, int zchunk) {
    if (zchunk <= 0) zchunk = default_zchunk(g);
    if (iters <= 0) {
        if (zero_init) return fill(s, x, cells(g), 0.0);
        if (xc) return prolong_add(s, g, xc, x);
        return MF3D_OK;
    }
    // outputs alternate so that the last sweep writes x; an odd count whose
    // first input is x itself ends in tmp and is copied back
    // with a virtual first input (zero) an odd count can start writing x;
    // otherwise the first sweep writes tmp
    double* out = ((iters & 1) && zero_init) ? x : tmp;
    const double* in = x;
    for (int it = 0; it < iters; ++it) {
        if (it == 0 && zero_init)
            MF3D_TRY((launch_march<true, false>(s, g, bc, pin, nullptr, out, b, K, nullptr, zchunk)));
        else if (it == 0 && xc)
            MF3D_TRY((launch_march<false, true>(s, g, bc, pin, in, out, b, K, xc, zchunk)));
        else
            MF3D_TRY((launch_march<false, false>(s, g, bc, pin, in, out, b, K, nullptr, zchunk)));
        in = out;
        out = (out == x) ? tmp : x;
    }
    if (in != x) MF3D_TRY(copy(s, in, x, cells(g)));
    return MF3D_OK;
}

}  // namespace kern
}  // namespace mf3d
