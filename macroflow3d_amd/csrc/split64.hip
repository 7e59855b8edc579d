// FP64 V-cycle levels in colour-split storage, faces from h = 1 / (2 K), gfx950.
//
// The FP64 V-cycle of the head solve (v_cycle.cu:36-74 with gsrb_smooth_3d,
// gsrb_3d.cu:29-366; compute_residual_3d + restrict_3d, residual_3d.cu:30-376
// and restrict_3d.cu:12-53; prolong_3d_add, prolong_3d.cu:30-621) on levels
// whose sizes allow it (nx a multiple of 4, ny and nz even).
//
// Storage: x, b and h of such a level are kept colour by colour, cell (i,j,k)
// of colour c = (i + j + k) & 1 at  c N/2 + (i >> 1) + (nx/2)(j + ny k)  (the
// "place" i >> 1 of its row). A half-sweep then streams the other colour's x
// and h, its own colour's h and b in and its own colour's x out, all with
// unit stride: 20 bytes per cell. No face is stored: the face between cells a
// and b is 1 / (h_a + h_b) (= 2 K_a K_b / (K_a + K_b), the harmonic mean of
// gsrb_3d.cu:53-58), a Dirichlet face 1 / h_C (= 2 K_C, ghost value 0, the
// boundary value living in b), a Neumann face 0, and a periodic face pairs
// with the wrapped cell (bc_stencil_helpers.cuh:33-267).
//
// A thread updates two consecutive places of one row (cells i = 4q + 2e + p,
// e = 0, 1, p = (j + k + colour) & 1) with 16-byte loads and stores. Their x
// neighbours sit at the other colour's places 2q - 1 .. 2q + 2 of the same
// row; the two outside the thread's pair come from the adjacent lanes
// (__shfl), the first and last lane of a wave and periodic wraps load them.
// Each cell's sums follow the reference's interior update term by term
// (x+, x-, y+, y-, z+, z-; (sum f x_nb - b dx^2) / sum f), the pin doubling
// the diagonal of cell 0 (gsrb_3d.cu:360-362).
//
// Fusions: the first half-sweep of a zero-start smoothing reads no x (MODE 1),
// and on the finest level also takes b from the natural-order FP64 PCG
// residual r and writes both colours of the split b (MODE 3); the first
// post-smoothing half-sweep adds the coarse correction to the neighbour values
// it reads (MODE 2: prolongation folded in; the own colour is overwritten and
// the other colour rewritten by the next half-sweep); the last half-sweep of
// the finest level writes z in natural order from the values it computes and
// the other colour it reads, with the (r, z) partials (RZ).

#include <hip/hip_runtime.h>

#include <algorithm>

#include "faces.hpp"
#include "hface.hpp"
#include "reduce.hpp"

namespace mf3d {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlock = 512;

typedef double __attribute__((ext_vector_type(2))) d2;
__device__ __forceinline__ d2 ld2(const double* __restrict__ p) {
    return *reinterpret_cast<const d2*>(p);
}
__device__ __forceinline__ void st2(double* __restrict__ p, double a, double b) {
    *reinterpret_cast<d2*>(p) = d2{a, b};
}
// r and A p of the PCG update folded into the first half-sweep are each
// touched once per launch: non-temporal loads and stores keep them from
// displacing the V-cycle's own arrays from the Infinity Cache
__device__ __forceinline__ d2 ldnt2(const double* __restrict__ p) {
    return __builtin_nontemporal_load(reinterpret_cast<const d2*>(p));
}
__device__ __forceinline__ void stnt2(double* __restrict__ p, double a, double b) {
    __builtin_nontemporal_store(d2{a, b}, reinterpret_cast<d2*>(p));
}

// one colour-split level
struct Geo {
    int nx, ny, nz;
    int h;      // places per row and colour: nx / 2
    int h2;     // threads per row of a sweep: nx / 4
    long half;  // N / 2
};
Geo geo_of(const mf3d_grid& g) {
    Geo G;
    G.nx = g.nx, G.ny = g.ny, G.nz = g.nz;
    G.h = g.nx / 2;
    G.h2 = g.nx / 4;
    G.half = (long)G.h * g.ny * g.nz;
    return G;
}

// (sum f x_nb - b bs) / sum f, summed x+, x-, y+, y-, z+, z- (gsrb_3d.cu:53-69)
__device__ __forceinline__ double gs64(const double (&f)[6], const double (&v)[6], double b,
                                       double bs, bool pin0) {
    double result = v[0] * f[0];
    result += v[1] * f[1];
    result += v[2] * f[2];
    result += v[3] * f[3];
    result += v[4] * f[4];
    result += v[5] * f[5];
    double aC = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
    if (pin0) aC *= 2.0;
    return (result - b * bs) / aC;
}

// Other-colour values at the x neighbours of the thread's two cells, from the
// other colour's pair a = (places base, base + 1) of row r0 (array O):
// xp[e] / xm[e] at x+ / x- of cell e; the places outside the pair come from
// the adjacent lanes, the first and last lane of a wave load them; a missing
// neighbour (non-periodic boundary) reads 0.
__device__ __forceinline__ void x_nb(const double* __restrict__ O, const d2 a, long base, long r0,
                                     int h, int p, bool xlo, bool xhi, int lane, bool per_lo,
                                     bool per_hi, double (&xp)[2], double (&xm)[2]) {
    double prv = __shfl_up(a.y, 1, 64);
    double nxt = __shfl_down(a.x, 1, 64);
    if (lane == 0 && !xlo) prv = O[base - 1];
    if (lane == 63 && !xhi) nxt = O[base + 2];
    if (p == 0) {  // i = 4q + 2e: x+ at place 2q + e, x- at 2q + e - 1
        xp[0] = a.x, xp[1] = a.y;
        xm[1] = a.x;
        xm[0] = xlo ? (per_lo ? O[r0 + h - 1] : 0.0) : prv;
    } else {       // i = 4q + 2e + 1: x- at place 2q + e, x+ at 2q + e + 1
        xm[0] = a.x, xm[1] = a.y;
        xp[0] = a.y;
        xp[1] = xhi ? (per_hi ? O[r0] : 0.0) : nxt;
    }
}

// Chunks of planes a sweep marches through: a workgroup owns `tiles`-th of a
// plane's thread rows and kz consecutive planes.
struct March {
    int tiles;   // workgroups per plane
    int kz;      // planes per chunk
    int chunks;  // ceil(nz / kz)
};
// at least `blocks` workgroups where the level allows, else one plane per chunk
March march_of(const Geo& G, int tb, long blocks) {
    March m;
    const long P = (long)G.h2 * G.ny;
    m.tiles = (int)((P + tb - 1) / tb);
    const long want = (blocks + m.tiles - 1) / m.tiles;
    const int chunks = (int)std::max(1L, std::min((long)G.nz, want));
    m.kz = (G.nz + chunks - 1) / chunks;
    m.chunks = (G.nz + m.kz - 1) / m.kz;
    return m;
}

// What one plane of the march loads ahead (issued before the previous plane
// is computed): the streams that come from HBM, i.e. the other colour's pair
// one plane up (z+; it becomes the x neighbours of the next plane and the z-
// of the one after) and the own colour's h and b. The y-neighbour rows (the
// other colour's rows j +- 1, loaded one plane earlier as other threads' z+
// pairs, so L2 hits) are loaded when the plane is computed.
struct PlaneLd {
    d2 xzp, hzp, hown, bown;
};

// One colour of GSRB on a colour-split level, marching in z. A thread owns
// places (2q, 2q + 1) of row j in every plane of its chunk; the cells it
// updates in plane k are i = 4q + 2e + p (p = (j + k + colour) & 1). The
// other colour's values at those places in plane k +- 1 are exactly its z
// neighbours, and in plane k its x neighbours (with the adjacent places from
// the adjacent lanes), so each other-colour pair is loaded once (as z+) and
// reused from registers twice.
// MODE 0: plain; 1: x = 0 on entry; 2: coarse correction xc (split coarse
// order if CS) prolonged into the neighbour values; 3: as 1 with b taken from
// the natural-order r64 and both colours of the split b written to bw; with
// ru.st set, the PCG residual update r -= alpha Ap (pcg_update_xr_kernel's
// form, skipped once stopped) comes first, in place in ru.r (then r64), b
// is taken from the new r and the workgroup's (r, r) partial goes to
// ru.partials.
// RZ (MODE 0, last half-sweep): z of the thread's 4 natural cells written
// instead of x, and the workgroup's (r64, z) partial when partials is given.
template <int MODE, bool CS, bool RZ>
__device__ __forceinline__ void march_body(const Geo& G, const BcTypes& bc, int pin, double bs,
                                           double* __restrict__ x, const double* __restrict__ b,
                                           const double* __restrict__ hs, int colour,
                                           const double* __restrict__ xc, const double* r64,
                                           double* __restrict__ bw, double* __restrict__ z,
                                           double* __restrict__ partials, long blk,
                                           const March& M, const RUpd& ru) {
    constexpr bool XR = MODE == 0 || MODE == 2;  // reads the other colour's x
    const long P = (long)G.h2 * G.ny;
    const int chunk = (int)(blk / M.tiles);
    const long tp = (blk - (long)chunk * M.tiles) * (long)blockDim.x + threadIdx.x;
    const bool valid = tp < P;
    const long tq = valid ? tp : P - 1;  // lanes past the plane mirror its last thread
    const int lane = threadIdx.x & 63;
    const int q = (int)(tq % G.h2);
    const int j = (int)(tq / G.h2);
    const int k0 = chunk * M.kz, k1 = min(G.nz, k0 + M.kz);
    const long sy = G.h, sz = (long)G.h * G.ny;
    const long ownc = (long)colour * G.half, othc = (long)(colour ^ 1) * G.half;
    const bool xlo = q == 0, xhi = q == G.h2 - 1;
    const bool per_lo = bc.t[0] == MF3D_BC_PERIODIC, per_hi = bc.t[1] == MF3D_BC_PERIODIC;
    const Side Yp = side(j < G.ny - 1, sy, bc.t[3], -(long)(G.ny - 1) * sy);
    const Side Ym = side(j > 0, -sy, bc.t[2], (long)(G.ny - 1) * sy);
    const Side Zb = side(k0 > 0, -sz, bc.t[4], (long)(G.nz - 1) * sz);  // z- of plane k0

    auto load = [&](int k, long pb) {
        const Side Zp = side(k < G.nz - 1, sz, bc.t[5], -(long)(G.nz - 1) * sz);
        PlaneLd l;
        l.hzp = ld2(hs + othc + pb + Zp.off);
        l.hown = ld2(hs + ownc + pb);
        l.xzp = XR ? ld2(x + othc + pb + Zp.off) : d2{0.0, 0.0};
        l.bown = MODE == 3 ? d2{0.0, 0.0} : ld2(b + ownc + pb);
        return l;
    };

    long pb = ((long)j + (long)G.ny * k0) * G.h + 2L * q;  // the pair's place in plane k0
    // the other colour in plane k0 (x neighbours) and below it (z-)
    d2 xo = XR ? ld2(x + othc + pb) : d2{0.0, 0.0};
    d2 ho = ld2(hs + othc + pb);
    d2 xzm = XR ? ld2(x + othc + pb + Zb.off) : d2{0.0, 0.0};
    d2 hzm = ld2(hs + othc + pb + Zb.off);
    PlaneLd cur = load(k0, pb);
    double acc = 0.0;
    const bool upd = MODE == 3 && ru.st != nullptr;  // uniform
    const bool go = upd && !ru.st->stop;
    const double alpha = upd ? ru.st->alpha : 0.0;
    for (int k = k0; k < k1; ++k) {
        PlaneLd nxt = cur;
        if (k + 1 < k1) nxt = load(k + 1, pb + sz);
        const Side Zp = side(k < G.nz - 1, sz, bc.t[5], -(long)(G.nz - 1) * sz);
        const Side Zm = k == k0 ? Zb : side(true, -sz, bc.t[4], 0);
        const int p = (j + k + colour) & 1;
        const long r0 = ((long)j + (long)G.ny * k) * G.h;
        const long row = (long)j + (long)G.ny * k;
        Side Ex[2], Wx[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            Ex[e] = side(!(p == 1 && e == 1 && xhi), 0, bc.t[1], 0);
            Wx[e] = side(!(p == 0 && e == 0 && xlo), 0, bc.t[0], 0);
        }
        const d2 hyp = ld2(hs + othc + pb + Yp.off), hym = ld2(hs + othc + pb + Ym.off);
        double hxp[2], hxm[2];
        x_nb(hs + othc, ho, pb, r0, G.h, p, xlo, xhi, lane, per_lo, per_hi, hxp, hxm);
        double vxp[2] = {0.0, 0.0}, vxm[2] = {0.0, 0.0}, vyp[2] = {0.0, 0.0}, vym[2] = {0.0, 0.0};
        double vzp[2] = {0.0, 0.0}, vzm[2] = {0.0, 0.0};
        if (XR) {
            const d2 xyp = ld2(x + othc + pb + Yp.off), xym = ld2(x + othc + pb + Ym.off);
            x_nb(x + othc, xo, pb, r0, G.h, p, xlo, xhi, lane, per_lo, per_hi, vxp, vxm);
            vyp[0] = Yp.nb ? xyp.x : 0.0, vyp[1] = Yp.nb ? xyp.y : 0.0;
            vym[0] = Ym.nb ? xym.x : 0.0, vym[1] = Ym.nb ? xym.y : 0.0;
            vzp[0] = Zp.nb ? cur.xzp.x : 0.0, vzp[1] = Zp.nb ? cur.xzp.y : 0.0;
            vzm[0] = Zm.nb ? xzm.x : 0.0, vzm[1] = Zm.nb ? xzm.y : 0.0;
        }
        double bb[2];
        if (MODE == 3) {
            // natural cells 4q .. 4q + 3 of the row: own colour at 2e + p
            const long rn = row * G.nx + 4L * q;
            d2 n01, n23;
            if (upd) {
                n01 = ldnt2(ru.r + rn), n23 = ldnt2(ru.r + rn + 2);
                const d2 a01 = ldnt2(ru.Ap + rn), a23 = ldnt2(ru.Ap + rn + 2);
                if (go) {
                    n01.x = -alpha * a01.x + n01.x;
                    n01.y = -alpha * a01.y + n01.y;
                    n23.x = -alpha * a23.x + n23.x;
                    n23.y = -alpha * a23.y + n23.y;
                    if (valid) {
                        stnt2(ru.r + rn, n01.x, n01.y);
                        stnt2(ru.r + rn + 2, n23.x, n23.y);
                    }
                }
                if (valid) {
                    acc += n01.x * n01.x;
                    acc += n01.y * n01.y;
                    acc += n23.x * n23.x;
                    acc += n23.y * n23.y;
                }
            } else {
                n01 = ld2(r64 + rn), n23 = ld2(r64 + rn + 2);
            }
            const double nv[4] = {n01.x, n01.y, n23.x, n23.y};
            bb[0] = nv[p], bb[1] = nv[2 + p];
            if (valid) {
                st2(bw + ownc + pb, bb[0], bb[1]);
                st2(bw + othc + pb, nv[1 - p], nv[3 - p]);
            }
        } else {
            bb[0] = cur.bown.x, bb[1] = cur.bown.y;
        }
        if (MODE == 2) {
            // each in-domain (or periodic) neighbour gains its coarse parent's
            // value: own cells' parents I = 2q + e of parent row (J0, K0); the
            // y/z neighbours' parent rows; the x neighbours' parents I - 1 .. I + 1
            const int NXc = G.nx >> 1, NYc = G.ny >> 1, NZc = G.nz >> 1;
            const int J0 = j >> 1, K0 = k >> 1;
            const long hcw = NXc >> 1, HC = hcw * NYc * NZc;
            auto cval = [&](int I, int J, int K) -> double {
                if (CS) return xc[(long)((I + J + K) & 1) * HC + (I >> 1) + hcw * ((long)J + (long)NYc * K)];
                return xc[(long)I + (long)NXc * ((long)J + (long)NYc * K)];
            };
            auto rowpair = [&](int J, int K, double (&o)[2]) {
                if (CS) {
                    const long rb = (long)q + hcw * ((long)J + (long)NYc * K);
                    o[0] = xc[(long)((J + K) & 1) * HC + rb];
                    o[1] = xc[(long)((J + K + 1) & 1) * HC + rb];
                } else {
                    const d2 A = ld2(xc + 2L * q + (long)NXc * ((long)J + (long)NYc * K));
                    o[0] = A.x, o[1] = A.y;
                }
            };
            double R[2], cy[2], cz[2];
            rowpair(J0, K0, R);
            // even sizes: an even j shares its parent row with j + 1, an odd j
            // with j - 1; only the other y neighbour's parent row is loaded (z alike)
            const bool jodd = j & 1, kodd = k & 1;
            const int Jp = j + 1 < G.ny ? (j + 1) >> 1 : 0, Jm = j > 0 ? (j - 1) >> 1 : NYc - 1;
            const int Kp = k + 1 < G.nz ? (k + 1) >> 1 : 0, Km = k > 0 ? (k - 1) >> 1 : NZc - 1;
            rowpair(jodd ? Jp : Jm, K0, cy);
            rowpair(J0, kodd ? Kp : Km, cz);
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if (Yp.nb) vyp[e] += jodd ? cy[e] : R[e];
                if (Ym.nb) vym[e] += jodd ? R[e] : cy[e];
                if (Zp.nb) vzp[e] += kodd ? cz[e] : R[e];
                if (Zm.nb) vzm[e] += kodd ? R[e] : cz[e];
            }
            double rm = __shfl_up(R[1], 1, 64), rpn = __shfl_down(R[0], 1, 64);
            if (lane == 0 && !xlo) rm = cval(2 * q - 1, J0, K0);
            if (lane == 63 && !xhi) rpn = cval(2 * q + 2, J0, K0);
            if (p == 0) {  // x+ of cell e: parent 2q + e; x-: 2q + e - 1
                vxp[0] += R[0], vxp[1] += R[1];
                vxm[1] += R[0];
                if (!xlo) vxm[0] += rm;
                else if (per_lo) vxm[0] += cval(NXc - 1, J0, K0);
            } else {       // x- of cell e: parent 2q + e; x+: 2q + e + 1
                vxm[0] += R[0], vxm[1] += R[1];
                vxp[0] += R[1];
                if (!xhi) vxp[1] += rpn;
                else if (per_hi) vxp[1] += cval(0, J0, K0);
            }
        }
        const double hc[2] = {cur.hown.x, cur.hown.y};
        const double hy[2][2] = {{hyp.x, hyp.y}, {hym.x, hym.y}};
        const double hz[2][2] = {{cur.hzp.x, cur.hzp.y}, {hzm.x, hzm.y}};
        double res[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const double f[6] = {hface(Ex[e], hc[e], hxp[e]), hface(Wx[e], hc[e], hxm[e]),
                                 hface(Yp, hc[e], hy[0][e]),  hface(Ym, hc[e], hy[1][e]),
                                 hface(Zp, hc[e], hz[0][e]),  hface(Zm, hc[e], hz[1][e])};
            const double v[6] = {vxp[e], vxm[e], vyp[e], vym[e], vzp[e], vzm[e]};
            res[e] = gs64(f, v, bb[e], bs, pin && colour == 0 && row == 0 && q == 0 && e == 0);
        }
        if (!RZ) {
            if (valid) st2(x + ownc + pb, res[0], res[1]);
        } else {
            // z of natural cells 4q .. 4q + 3: own at 2e + p, the other colour
            // (final after the previous half-sweep) at 2e + 1 - p
            const long zn = row * G.nx + 4L * q;
            const double z0 = p ? xo.x : res[0], z1 = p ? res[0] : xo.x;
            const double z2 = p ? xo.y : res[1], z3 = p ? res[1] : xo.y;
            if (valid) {
                st2(z + zn, z0, z1);
                st2(z + zn + 2, z2, z3);
                if (partials) {  // r64 is the r of the (r, z) partials here
                    const d2 ra = ld2(r64 + zn), rb = ld2(r64 + zn + 2);
                    acc += ra.x * z0;
                    acc += ra.y * z1;
                    acc += rb.x * z2;
                    acc += rb.y * z3;
                }
            }
        }
        // march: this plane's other colour becomes the next plane's z-, the
        // z+ pair the next plane's x neighbours
        xzm = xo, hzm = ho;
        xo = cur.xzp, ho = cur.hzp;
        pb += sz;
        cur = nxt;
    }
    if (RZ && partials) {  // uniform
        const double sb = block_sum_dyn(acc);
        if (threadIdx.x == 0) partials[blk] = sb;
    }
    if (upd) {  // uniform
        const double sb = block_sum_dyn(acc);
        if (threadIdx.x == 0) ru.partials[blk] = sb;
    }
}

template <int MODE, bool CS, bool RZ>
__global__ void __launch_bounds__(kMaxBlock)
    march64_kernel(Geo G, BcTypes bc, int pin, double bs, double* __restrict__ x,
                   const double* __restrict__ b, const double* __restrict__ hs, int colour,
                   const double* __restrict__ xc, const double* r64,
                   double* __restrict__ bw, double* __restrict__ z,
                   double* __restrict__ partials, March M, RUpd ru) {
    const long blk = xcd_block((long)M.tiles * M.chunks);
    if (blk < 0) return;  // whole workgroup
    march_body<MODE, CS, RZ>(G, bc, pin, bs, x, b, hs, colour, xc, r64, bw, z, partials, blk, M,
                             ru);
}

// residual r = b - A x of one cell of a split level (compute_residual_3d's
// sums: x+, x-, y+, y-, z+, z- of f (x_C - x_nb); the pin adds the diagonal
// times x_C at cell 0, residual_3d.cu:365-368), times rs = 1 / dx^2
__device__ __forceinline__ double residual64(const Geo& G, const BcTypes& bc, bool pin,
                                             const double* __restrict__ x,
                                             const double* __restrict__ b,
                                             const double* __restrict__ hs, int i, int j, int k,
                                             double rs) {
    const int c = (i + j + k) & 1;
    const long r0 = (long)G.h * ((long)j + (long)G.ny * k);
    const long o = (long)c * G.half + r0 + (i >> 1);
    const long O = (long)(c ^ 1) * G.half;
    const long sy = G.h, sz = (long)G.h * G.ny;
    const int m = i >> 1;
    const Side E = side(i < G.nx - 1, r0 + ((i + 1) >> 1), bc.t[1], r0);
    const Side W = side(i > 0, r0 + ((i - 1) >> 1), bc.t[0], r0 + G.h - 1);
    const Side N = side(j < G.ny - 1, r0 + sy + m, bc.t[3], r0 - (long)(G.ny - 1) * sy + m);
    const Side S = side(j > 0, r0 - sy + m, bc.t[2], r0 + (long)(G.ny - 1) * sy + m);
    const Side U = side(k < G.nz - 1, r0 + sz + m, bc.t[5], r0 - (long)(G.nz - 1) * sz + m);
    const Side D = side(k > 0, r0 - sz + m, bc.t[4], r0 + (long)(G.nz - 1) * sz + m);
    const Side sd[6] = {E, W, N, S, U, D};
    const double xc = x[o], hc = hs[o];
    double f[6], v[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const long a = O + (sd[q].nb ? sd[q].off : r0);
        f[q] = hface(sd[q], hc, hs[a]);
        v[q] = sd[q].nb ? x[a] : 0.0;
    }
    double result = f[0] * (xc - v[0]);
    result += f[1] * (xc - v[1]);
    result += f[2] * (xc - v[2]);
    result += f[3] * (xc - v[3]);
    result += f[4] * (xc - v[4]);
    result += f[5] * (xc - v[5]);
    if (pin && (i | j | k) == 0) result += (f[0] + f[1] + f[2] + f[3] + f[4] + f[5]) * xc;
    return b[o] + result * rs;
}

// residual + restriction (1/8 of the children's sum, restrict_3d.cu:12-53)
// into a coarse level of either order. RED: after a red-black sweep ending
// with the black half-sweep each black cell satisfies its own equation given
// its final red neighbours, so its residual vanishes up to rounding; only the
// four red children of a coarse cell are evaluated.
template <bool CS, bool RED>
__global__ void __launch_bounds__(kBlock)
    rr64_kernel(Geo G, BcTypes bc, int pin, double rs, const double* __restrict__ x,
                const double* __restrict__ b, const double* __restrict__ hs,
                double* __restrict__ bcoarse) {
    const int NXc = G.nx >> 1, NYc = G.ny >> 1, NZc = G.nz >> 1;
    const long total = (long)NXc * NYc * NZc;
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long t = blk * kBlock + threadIdx.x;
    if (t >= total) return;
    const int I = (int)(t % NXc);
    const long qq = t / NXc;
    const int J = (int)(qq % NYc);
    const int Kc = (int)(qq / NYc);
    const int i = 2 * I, j = 2 * J, k = 2 * Kc;
    const bool pn = pin != 0;
    double s = 0.0;
    if (RED) {  // children of colour 0: (0,0,0), (1,1,0), (1,0,1), (0,1,1)
        s += residual64(G, bc, pn, x, b, hs, i, j, k, rs);
        s += residual64(G, bc, pn, x, b, hs, i + 1, j + 1, k, rs);
        s += residual64(G, bc, pn, x, b, hs, i + 1, j, k + 1, rs);
        s += residual64(G, bc, pn, x, b, hs, i, j + 1, k + 1, rs);
    } else {
        s += residual64(G, bc, pn, x, b, hs, i, j, k, rs);
        s += residual64(G, bc, pn, x, b, hs, i + 1, j, k, rs);
        s += residual64(G, bc, pn, x, b, hs, i, j + 1, k, rs);
        s += residual64(G, bc, pn, x, b, hs, i + 1, j + 1, k, rs);
        s += residual64(G, bc, pn, x, b, hs, i, j, k + 1, rs);
        s += residual64(G, bc, pn, x, b, hs, i + 1, j, k + 1, rs);
        s += residual64(G, bc, pn, x, b, hs, i, j + 1, k + 1, rs);
        s += residual64(G, bc, pn, x, b, hs, i + 1, j + 1, k + 1, rs);
    }
    long oc = t;
    if (CS) {
        const long hcw = NXc >> 1;
        oc = (long)((I + J + Kc) & 1) * (hcw * NYc * NZc) + (I >> 1) + hcw * ((long)J + (long)NYc * Kc);
    }
    bcoarse[oc] = s / 8.0;
}

// h = 1 / (2 K) in split order (hs) and, if hn is given, natural order too;
// one thread per 4 natural cells of a row (cells 4q, 4q + 2 of colour c0 =
// (j + k) & 1 at places 2q, 2q + 1, cells 4q + 1, 4q + 3 of the other colour)
__global__ void __launch_bounds__(kBlock)
    split_h_kernel(Geo G, const double* __restrict__ K, double* __restrict__ hs,
                   double* __restrict__ hn) {
    const long total = G.half >> 1;
    const long blk = xcd_block((total + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long t = blk * kBlock + threadIdx.x;
    if (t >= total) return;
    const int q = (int)(t % G.h2);
    const long row = t / G.h2;
    const int j = (int)(row % G.ny), k = (int)(row / G.ny);
    const long nat = row * G.nx + 4L * q;
    const d2 a = ld2(K + nat), c = ld2(K + nat + 2);
    const double h0 = 0.5 * (1.0 / a.x), h1 = 0.5 * (1.0 / a.y);
    const double h2 = 0.5 * (1.0 / c.x), h3 = 0.5 * (1.0 / c.y);
    const int c0 = (j + k) & 1;
    const long pl = row * G.h + 2L * q;
    st2(hs + (long)c0 * G.half + pl, h0, h2);
    st2(hs + (long)(c0 ^ 1) * G.half + pl, h1, h3);
    if (hn) {
        st2(hn + nat, h0, h1);
        st2(hn + nat + 2, h2, h3);
    }
}

// natural -> split (TO) or split -> natural with optional (r, nat) partials
template <bool TO>
__global__ void __launch_bounds__(kBlock)
    reorder64_kernel(Geo G, const double* __restrict__ a, double* __restrict__ o,
                     const double* __restrict__ r, double* __restrict__ partials) {
    const long total = G.half >> 1;
    const long nb = (total + kBlock - 1) / kBlock;
    const long blk = xcd_block(nb);
    if (blk < 0) return;
    const long t = blk * kBlock + threadIdx.x;
    double acc = 0.0;
    if (t < total) {
        const int q = (int)(t % G.h2);
        const long row = t / G.h2;
        const int j = (int)(row % G.ny), k = (int)(row / G.ny);
        const long nat = row * G.nx + 4L * q;
        const int c0 = (j + k) & 1;
        const long pl = row * G.h + 2L * q;
        const long s0 = (long)c0 * G.half + pl, s1 = (long)(c0 ^ 1) * G.half + pl;
        if (TO) {
            const d2 u = ld2(a + nat), w = ld2(a + nat + 2);
            st2(o + s0, u.x, w.x);
            st2(o + s1, u.y, w.y);
        } else {
            const d2 u = ld2(a + s0), w = ld2(a + s1);
            st2(o + nat, u.x, w.x);
            st2(o + nat + 2, u.y, w.y);
            if (partials) {
                const d2 ra = ld2(r + nat), rb = ld2(r + nat + 2);
                acc = ra.x * u.x;
                acc += ra.y * w.x;
                acc += rb.x * u.y;
                acc += rb.y * w.y;
            }
        }
    }
    if (!TO && partials) {
        const double sb = block_sum<kBlock>(acc);
        if (threadIdx.x == 0) partials[blk] = sb;
    }
}

// x += P xc on a split level (every cell; used when there is no
// post-smoothing to fold the prolongation into)
template <bool CS>
__global__ void __launch_bounds__(kBlock)
    prolong64_kernel(Geo G, const double* __restrict__ xc, double* __restrict__ x) {
    const long n = 2 * G.half;
    const long blk = xcd_block((n + kBlock - 1) / kBlock);
    if (blk < 0) return;
    const long o = blk * kBlock + threadIdx.x;
    if (o >= n) return;
    const int c = o >= G.half;
    const long pl = o - (long)c * G.half;
    const int m = (int)(pl % G.h);
    const long row = pl / G.h;
    const int j = (int)(row % G.ny), k = (int)(row / G.ny);
    const int i = 2 * m + ((j + k + c) & 1);
    const int NXc = G.nx >> 1, NYc = G.ny >> 1, NZc = G.nz >> 1;
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    long po = (long)I + (long)NXc * ((long)J + (long)NYc * K);
    if (CS) {
        const long hcw = NXc >> 1;
        po = (long)((I + J + K) & 1) * (hcw * NYc * NZc) + (I >> 1) + hcw * ((long)J + (long)NYc * K);
    }
    x[o] += xc[po];
}

__global__ void __launch_bounds__(kBlock) fill64_kernel(double* __restrict__ x, long n) {
    const long o = (long)blockIdx.x * kBlock + threadIdx.x;
    if (o < n) x[o] = 0.0;
}

inline unsigned blocks_for(long threads, int tb) {
    return xcd_round((size_t)((threads + tb - 1) / tb));
}

template <int MODE, bool CS, bool RZ>
int launch_sweep(hipStream_t s, const Geo& G, const BcTypes& bc, bool pin, double bs, double* x,
                 const double* b, const double* hs, int colour, const double* xc,
                 const double* r64, double* bw, double* z, double* partials, int tb,
                 long blocks, const RUpd* ru = nullptr) {
    const March M = march_of(G, tb, blocks);
    march64_kernel<MODE, CS, RZ><<<xcd_round((size_t)M.tiles * M.chunks), tb, 0, s>>>(
        G, bc, pin ? 1 : 0, bs, x, b, hs, colour, xc, r64, bw, z, partials, M,
        ru ? *ru : RUpd{nullptr, nullptr, nullptr, nullptr});
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int sweep(hipStream_t s, const Geo& G, const BcTypes& bc, bool pin, double bs, double* x,
          const double* b, const double* hs, int colour, int mode, bool cs, const double* xc,
          const double* r64, double* bw, int tb, long blocks, const RUpd* ru) {
    switch (mode) {
    case 0: return launch_sweep<0, false, false>(s, G, bc, pin, bs, x, b, hs, colour, nullptr,
                                                 nullptr, nullptr, nullptr, nullptr, tb, blocks);
    case 1: return launch_sweep<1, false, false>(s, G, bc, pin, bs, x, b, hs, colour, nullptr,
                                                 nullptr, nullptr, nullptr, nullptr, tb, blocks);
    case 3: return launch_sweep<3, false, false>(s, G, bc, pin, bs, x, bw, hs, colour, nullptr,
                                                 r64, bw, nullptr, nullptr, tb, blocks, ru);
    default:
        return cs ? launch_sweep<2, true, false>(s, G, bc, pin, bs, x, b, hs, colour, xc, nullptr,
                                                 nullptr, nullptr, nullptr, tb, blocks)
                  : launch_sweep<2, false, false>(s, G, bc, pin, bs, x, b, hs, colour, xc,
                                                  nullptr, nullptr, nullptr, nullptr, tb, blocks);
    }
}

}  // namespace

namespace kern {

bool split64_ok(const mf3d_grid& g) {
    return g.nx % 4 == 0 && g.ny % 2 == 0 && g.nz % 2 == 0 && g.nx >= 4 && g.ny >= 2 &&
           g.nz >= 2;
}

unsigned split64_partials(const mf3d_grid& g, int tb, long blocks) {
    const March M = march_of(geo_of(g), tb, blocks);
    return (unsigned)M.tiles * (unsigned)M.chunks;
}

int split_h(hipStream_t s, const mf3d_grid& g, const double* K, double* hs, double* hn) {
    if (!split64_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "split_h: sizes not split-capable");
    const Geo G = geo_of(g);
    split_h_kernel<<<blocks_for(G.half >> 1, kBlock), kBlock, 0, s>>>(G, K, hs, hn);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int to_split64(hipStream_t s, const mf3d_grid& g, const double* nat, double* sp) {
    const Geo G = geo_of(g);
    reorder64_kernel<true><<<blocks_for(G.half >> 1, kBlock), kBlock, 0, s>>>(G, nat, sp, nullptr,
                                                                              nullptr);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

int from_split64(hipStream_t s, const mf3d_grid& g, const double* sp, double* nat,
                 const double* r, double* partials, unsigned* np) {
    const Geo G = geo_of(g);
    reorder64_kernel<false><<<blocks_for(G.half >> 1, kBlock), kBlock, 0, s>>>(G, sp, nat, r,
                                                                               partials);
    MF3D_HIP_TRY(hipGetLastError());
    if (np) *np = (unsigned)(((G.half >> 1) + kBlock - 1) / kBlock);
    return MF3D_OK;
}

int smooth_split64(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                   double* x, double* b, const double* hs, int iters, bool zero_init,
                   const double* xc, int corder, int tb, const double* r64, double* z,
                   const double* rdot, double* partials, unsigned* np, long blocks,
                   const RUpd* ru) {
    if (!split64_ok(g)) return fail(MF3D_ERR_UNSUPPORTED, "smooth_split64: sizes not split-capable");
    if (tb < 64 || tb > kMaxBlock || tb % 64)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "sweep workgroup size must be 64..512, a multiple of 64");
    if (blocks < kMarchMinBlocks || blocks > kMarchMaxBlocks)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "march workgroup target out of range");
    if (zero_init && xc) return fail(MF3D_ERR_INVALID_ARGUMENT, "zero start with a coarse correction");
    if (r64 && !zero_init) return fail(MF3D_ERR_INVALID_ARGUMENT, "b from r64 needs a zero start");
    if (ru && !(r64 == ru->r && iters > 0))
        return fail(MF3D_ERR_INVALID_ARGUMENT, "the r update rides on the b-from-r half-sweep");
    const Geo G = geo_of(g);
    const bool cs = corder == ORDER_SPLIT;
    const size_t n = cells(g);
    if (iters == 0) {
        if (r64) MF3D_TRY(to_split64(s, g, r64, b));
        if (zero_init) {
            fill64_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(x, (long)n);
            MF3D_HIP_TRY(hipGetLastError());
        }
        if (xc) {
            if (cs) prolong64_kernel<true><<<xcd_round((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(G, xc, x);
            else prolong64_kernel<false><<<xcd_round((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(G, xc, x);
            MF3D_HIP_TRY(hipGetLastError());
        }
        if (z) MF3D_TRY(from_split64(s, g, x, z, rdot, rdot ? partials : nullptr, np));
        return MF3D_OK;
    }
    for (int it = 0; it < iters; ++it) {
        for (int colour = 0; colour < 2; ++colour) {
            const bool first = it == 0 && colour == 0;
            const bool last = it == iters - 1 && colour == 1;
            int mode = 0;
            if (first && zero_init) mode = r64 ? 3 : 1;
            else if (first && xc) mode = 2;
            if (last && z) {
                // (the red half-sweep before it is never the MODE 2 one unless
                // iters == 1, where it ran just above)
                MF3D_TRY((launch_sweep<0, false, true>(s, G, bc, pin, bs, x, b, hs, 1, nullptr,
                                                       rdot, nullptr, z,
                                                       rdot ? partials : nullptr, tb,
                                                       blocks)));
                if (np) *np = split64_partials(g, tb, blocks);
                continue;
            }
            MF3D_TRY(sweep(s, G, bc, pin, bs, x, b, hs, colour, mode, cs, xc, r64, b, tb, blocks,
                           mode == 3 ? ru : nullptr));
        }
    }
    return MF3D_OK;
}

int residual_restrict_split64(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                              double rs, const double* x, const double* b, const double* hs,
                              double* bcoarse, int corder, bool red_only) {
    const Geo G = geo_of(g);
    const long total = (long)(g.nx / 2) * (g.ny / 2) * (g.nz / 2);
    const unsigned nb = xcd_round((size_t)((total + kBlock - 1) / kBlock));
    const int p = pin ? 1 : 0;
    const bool cs = corder == ORDER_SPLIT;
    if (cs && red_only) rr64_kernel<true, true><<<nb, kBlock, 0, s>>>(G, bc, p, rs, x, b, hs, bcoarse);
    else if (cs) rr64_kernel<true, false><<<nb, kBlock, 0, s>>>(G, bc, p, rs, x, b, hs, bcoarse);
    else if (red_only) rr64_kernel<false, true><<<nb, kBlock, 0, s>>>(G, bc, p, rs, x, b, hs, bcoarse);
    else rr64_kernel<false, false><<<nb, kBlock, 0, s>>>(G, bc, p, rs, x, b, hs, bcoarse);
    MF3D_HIP_TRY(hipGetLastError());
    return MF3D_OK;
}

}  // namespace kern
}  // namespace mf3d
