// Face-coefficient fields of a multigrid level and their kernels (faces.hip).
#pragma once

#include "kernels.hpp"

namespace mf3d {

enum CoefType : int { COEF_F64 = 0, COEF_F32 = 1, COEF_BF16 = 2, COEF_F16 = 3 };

// Face conductances of one level, values of `type` stored times a
// power-of-two scale. Natural (MAC) layout: fx (nx+1)*ny*nz, fy nx*(ny+1)*nz,
// fz nx*ny*(nz+1). Split layout (colour-split levels): fx/fy/fz hold N values,
// the lower faces of each cell at the cell's colour-split index, and ux
// (ny*nz), uy (nx*nz), uz (nx*ny) the faces on the upper domain boundary.
struct FaceSet {
    CoefType type = COEF_F64;
    bool split = false;
    void* fx = nullptr;
    void* fy = nullptr;
    void* fz = nullptr;
    void* ux = nullptr;
    void* uy = nullptr;
    void* uz = nullptr;
};

// Storage order of a level's FP32 V-cycle vectors x and b. NATURAL:
// i + nx (j + ny k). SPLIT (even nx, ny, nz): the cells of colour
// c = (i+j+k) & 1 at c*N/2 + (i >> 1) + (nx/2)(j + ny k).
enum Order : int { ORDER_NATURAL = 0, ORDER_SPLIT = 1 };

enum SmoothKind : int { SMOOTH_COLOUR = 0, SMOOTH_SMALL = 1 };

// How one level is smoothed (chosen by defaults or the autotuner).
struct SmoothPlan {
    int kind = SMOOTH_COLOUR;
    int block = 256;   // colour-split sweeps: threads per workgroup (64..512)
};

namespace kern {

size_t face_count(const mf3d_grid& g);
size_t coef_bytes(CoefType t);
// faces from K (cs: storage scale)
int faces_build(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, double cs, const double* K,
                FaceSet& f);
// one pass for the finest mixed-precision level: flo colour-split (scaled by
// cs) and H = 1 / (2 K) of the FP64 operator
int faces_build_split_h(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, double cs,
                        const double* K, FaceSet& flo, double* H);
// H = 1 / (2 K)
int half_recip(hipStream_t s, const double* K, double* H, size_t n);
// iters GSRB iterations (red then black) with face coefficients; result in x.
// bs = cs * dx^2. zero_init: x is zero on entry. xc (nullable): coarse
// correction prolonged and added first.
int smooth_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                 double* x, const double* b, const FaceSet& F, int iters,
                 bool zero_init, const double* xc, const SmoothPlan& plan,
                 int order = ORDER_NATURAL, int corder = ORDER_NATURAL,
                 const double* r64 = nullptr, bool red_done = false,
                 const RUpd* ru = nullptr, const RzDot* rz = nullptr);
// order: storage order of x/b, corder: of xc (FP64 levels are natural only);
// a split level's coarse level may be natural, not the reverse.
// r64 (nullable): b is the FP32 copy of this FP64 natural-order vector, made
// by the first half-sweep (colour-split level, nx % 8 == 0, zero_init, no xc)
// red_done: the first (red) half-sweep from x = 0 was folded into the
// restriction that produced b (residual_restrict_faces with xcoarse);
// ru (with r64): the PCG r update rides on the first half-sweep
// rz (colour-split level, nx % 8 == 0): the last half-sweep also writes the
// (r, x) partials of its natural-order FP64 r, one per sweep workgroup
int smooth_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                 float* x, const float* b, const FaceSet& F, int iters,
                 bool zero_init, const float* xc, const SmoothPlan& plan,
                 int order = ORDER_NATURAL, int corder = ORDER_NATURAL,
                 const double* r64 = nullptr, bool red_done = false,
                 const RUpd* ru = nullptr, const RzDot* rz = nullptr);
bool split_ok(const mf3d_grid& g);
// coarse b = restriction of (b - A x), rs = 1 / (cs * dx^2)
int residual_restrict_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                            double rs, const double* x, const double* b, const FaceSet& F,
                            double* bcoarse, int order = ORDER_NATURAL,
                            int corder = ORDER_NATURAL, bool red_only = false,
                            double* xcoarse = nullptr, const FaceSet* Fc = nullptr,
                            double bsc = 0.0);
// red_only (colour-split levels after at least one red-black sweep): the
// black children's residuals are taken as zero (see rr_split_body)
// xcoarse (split coarse level only): also the coarse level's first red
// half-sweep from x = 0 (faces Fc, bsc = its cs dx^2); the coarse smoothing
// then runs with red_done
int residual_restrict_faces(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                            double rs, const float* x, const float* b, const FaceSet& F,
                            float* bcoarse, int order = ORDER_NATURAL,
                            int corder = ORDER_NATURAL, bool red_only = false,
                            float* xcoarse = nullptr, const FaceSet* Fc = nullptr,
                            double bsc = 0.0);
// FP64 operator with FP64 faces (cs = 1). mode 0: y = A x, partials (x, y);
// mode 1: y = b - A x (y nullable), partials (y, y). It writes op_partials(g)
// partials: one per (plane tile of 256 cells, z-chunk of kz planes).
// FP64 operator of the finest level from h = 1 / (2 K): workgroups of
// tiles x chunks (one partial each); pairs: two x-adjacent cells per thread
// with 16-byte accesses (nx even and every vector 16-byte aligned)
struct OpGeom {
    int tiles, kz, chunks;
    bool pairs;
    bool tile = false;  // plane tiles staged in LDS (pairs too)
};
// z-chunks of at most kz_max planes per workgroup (kOpChunkDefault unless tuned)
constexpr int kOpChunkDefault = 16, kOpChunkMax = 64;
OpGeom op_geom(const mf3d_grid& g, bool pairs, int kz_max = kOpChunkDefault);
// the variant op_apply runs: plane tiles with pairs when allowed and nx <= 512
OpGeom op_plan(const mf3d_grid& g, bool pairs, bool allow_tile, int kz_max = kOpChunkDefault);
// dynamic LDS bytes of the plane-tile variant
size_t op_tile_lds(const mf3d_grid& g);
// the most partials any variant writes, at any chunk depth (one plane per chunk)
unsigned op_partials(const mf3d_grid& g);
bool op_pairs_ok(const mf3d_grid& g, const void* x, const void* b, const void* y,
                 const void* H);
// mode 0: y = A x with (x, y) partials; mode 1: y = b - A x (y nullable) with
// (y, y) partials; np (nullable): the number of partials written; pairs only
// when allowed and op_pairs_ok; with pairs and nx <= 512 (and allow_tile) the
// plane-tile variant with the y neighbours staged in LDS
int op_apply(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, const double* H,
             int mode, const double* x, const double* b, double* y, double* partials,
             unsigned* np, bool allow_pairs = true, bool allow_tile = true,
             int kz_max = kOpChunkDefault);
int to_f32(hipStream_t s, const double* a, float* o, size_t n);
int to_f64(hipStream_t s, const float* a, double* o, size_t n);
// conversions between a natural FP64 vector and an FP32 vector of `order`
int to_f32(hipStream_t s, const double* a, float* o, const mf3d_grid& g, int order);
// with r (nullable): also the (r, o) first-stage partials, as dot_partials
// colour-split x32 with nx % 8 == 0: (r, x32) partials[reduce_blocks(n)] only,
// and the PCG p update from x32 (first: p = x32; else p = x32 + beta p;
// skipped once st->stop is set)
int rz_split8(hipStream_t s, const float* x32, const mf3d_grid& g, const double* r,
              double* partials);
int update_p_split8(hipStream_t s, const IterState* st, const float* x32, const mf3d_grid& g,
                    double* p, bool first);
// x += alpha p; p = x32 + beta p (skipped once st->stop is set)
int update_xp_split8(hipStream_t s, const IterState* st, const float* x32, const mf3d_grid& g,
                     double* x, double* p);
int to_f64(hipStream_t s, const float* a, double* o, const mf3d_grid& g, int order,
           const double* r = nullptr, double* partials = nullptr);
int gemv(hipStream_t s, int n, const double* Ainv, const double* b, double* x);
int gemv(hipStream_t s, int n, const double* Ainv, const float* b, float* x);
int fill_f32(hipStream_t s, float* x, size_t n, float v);
// partials of sum(log K) for the coefficient scale
int sum_log_partials(hipStream_t s, const double* K, size_t n, double* partials);

// ---- FP64 V-cycle levels in colour-split order, faces from h = 1 / (2 K) -------
// (split64.hip) x, b and h in the split order of ORDER_SPLIT; sizes: nx % 4 == 0,
// ny and nz even
bool split64_ok(const mf3d_grid& g);
// (r, z) partials written by the finest level's last half-sweep (tb threads
// per workgroup)
// blocks: the workgroup count a sweep aims for (planes per chunk shortened
// until the launch reaches it, where the level allows)
constexpr long kMarchDefaultBlocks = 2048, kMarchMinBlocks = 256, kMarchMaxBlocks = 8192;
unsigned split64_partials(const mf3d_grid& g, int tb, long blocks = kMarchDefaultBlocks);
// hs = 1 / (2 K) in split order, and in natural order into hn when not null
int split_h(hipStream_t s, const mf3d_grid& g, const double* K, double* hs, double* hn);
int to_split64(hipStream_t s, const mf3d_grid& g, const double* nat, double* sp);
// split -> natural; with r and partials also the (r, nat) partials (np of them)
int from_split64(hipStream_t s, const mf3d_grid& g, const double* sp, double* nat,
                 const double* r, double* partials, unsigned* np);
// iters GSRB iterations (red then black), bs = dx^2, tb threads per workgroup.
// zero_init: x is zero on entry; xc: coarse correction (order corder) prolonged
// and added (folded into the first half-sweep); r64 (zero_init): b is taken
// from this natural-order vector and written to b in split order; z: the
// last half-sweep writes x to z in natural order instead, with the (rdot, z)
// partials when rdot is given (np: their count); ru (r64 == ru->r, iters > 0):
// the first half-sweep first updates r -= alpha Ap in place and writes the
// (r, r) partials, split64_partials(g, tb, blocks) of them
int smooth_split64(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double bs,
                   double* x, double* b, const double* hs, int iters, bool zero_init,
                   const double* xc, int corder, int tb, const double* r64, double* z,
                   const double* rdot, double* partials, unsigned* np,
                   long blocks = kMarchDefaultBlocks, const RUpd* ru = nullptr);
// coarse b = restriction of (b - A x) (rs = 1 / dx^2) into a coarse level of
// order corder; red_only after a sweep that ended with the black half-sweep
int residual_restrict_split64(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin,
                              double rs, const double* x, const double* b, const double* hs,
                              double* bcoarse, int corder, bool red_only);

// ---- batched launches (descriptors and block starts in device memory) --------
// kBlock-thread blocks needed for `threads` threads
long batch_blocks(long threads);
// one colour of GSRB on colour-split levels (nx % 8 == 0), mode as
// colour_split4_kernel: 0 plain, 1 zero start, 3 zero start with b from the
// descriptor's r64, 2 prolongation folded in (cs:
// coarse level colour-split); tb threads per workgroup, nblocks = sum of
// ceil(n / 8 / tb)
int batch_sweep(hipStream_t s, CoefType t, int mode, bool cs, const BatchLevel* L,
                const int* start, int m, long nblocks, int colour, int tb);
// residual + restriction of colour-split levels (nblocks: sum of batch_blocks(n / 8))
int batch_residual_restrict_split(hipStream_t s, CoefType t, bool cs, const BatchLevel* L,
                                  const int* start, int m, long nblocks);
// whole smoothing of small natural-order levels, one workgroup per entry
int batch_small(hipStream_t s, CoefType t, const BatchLevel* L, int m, int iters, bool zero_init,
                bool use_xc);
// residual + restriction of natural-order levels (nblocks: sum of batch_blocks(n / 8))
int batch_residual_restrict(hipStream_t s, CoefType t, const BatchLevel* L, const int* start,
                            int m, long nblocks);
// coarsest direct solve (nblocks: sum of ceil(nc / 4))
int batch_gemv(hipStream_t s, const BatchLevel* L, const int* start, int m, long nblocks);
// finest level (nblocks: sum of batch_blocks(n)): r -> b32 split, x32 split -> z,
// mode 0: Ap = A p (+ (p, Ap) partials), mode 1: r = rhs - A x (+ (r, r) partials)
int batch_to_f32(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks);
// x32 split -> z with the (r, z) partials (replaces a batch_dot mode 0)
int batch_to_f64(hipStream_t s, const BatchVec* V, const int* start, int m, long nblocks);
int batch_update_p_split8(hipStream_t s, int mode, const BatchVec* V, const int* start, int m,
                          long nblocks);
int batch_update_xp_split8(hipStream_t s, const BatchVec* V, const int* start, int m,
                           long nblocks);
// lds: dynamic LDS bytes, the largest op_tile_lds of the entries that use
// plane tiles (BatchVec::opairs == 2), else 0
int batch_op(hipStream_t s, int mode, const BatchVec* V, const int* start, int m, long nblocks,
             size_t lds);

}  // namespace kern
}  // namespace mf3d
