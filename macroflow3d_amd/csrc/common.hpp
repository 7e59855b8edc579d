// Shared host/device definitions of libmf3d_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/mf3d.h"

namespace mf3d {

// ---------------------------------------------------------------------------
// Error plumbing: internal functions return an mf3d status; the detail text of
// the last failure is kept per thread (mf3d_last_error).
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);

#define MF3D_HIP_TRY(expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::mf3d::fail(MF3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MF3D_TRY(expr)                                                                       \
    do {                                                                                     \
        int s_ = (expr);                                                                     \
        if (s_ != MF3D_OK) return s_;                                                        \
    } while (0)

inline size_t cells(const mf3d_grid& g) { return (size_t)g.nx * (size_t)g.ny * (size_t)g.nz; }

// Workgroups are dealt round-robin to the 8 XCDs (blockIdx % 8), each with its
// own L2. xcd_block maps the launch so that the blocks of one XCD cover a
// contiguous range of the logical grid (a slab of rows and planes), keeping
// the y- and z-neighbour rows of a stencil in that XCD's L2. The grid is
// rounded up to a multiple of 8 blocks (xcd_round); the padding blocks get -1.
inline unsigned xcd_round(size_t blocks) { return (unsigned)((blocks + 7u) / 8u * 8u); }
__device__ __forceinline__ long xcd_block(long nblocks) {
    const long per = (nblocks + 7) >> 3;
    const long b = blockIdx.x;
    const long t = (b & 7) * per + (b >> 3);
    return t < nblocks ? t : -1;
}

// Boundary types only (values enter through the RHS and velocity kernels).
struct BcTypes {
    int t[6];
};
inline BcTypes bc_types(const mf3d_bc& bc) {
    BcTypes o;
    for (int i = 0; i < 6; ++i) o.t[i] = bc.type[i];
    return o;
}

// Validation used by every entry point (BCSpec::validate, src/core/BCSpec.hpp:63-87).
int validate_grid(const mf3d_grid* g);
int validate_bc(const mf3d_bc* bc);

// pin_enabled (src/numerics/pin_spec.hpp:110-120)
inline bool pin_enabled(int mode, const mf3d_bc& bc) {
    if (mode == MF3D_PIN_ON) return true;
    if (mode == MF3D_PIN_OFF) return false;
    for (int f = 0; f < 6; ++f)
        if (bc.type[f] == MF3D_BC_DIRICHLET) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Per-cell stencil coefficients with boundary conditions, shared by host and
// device. Semantics of bc_helpers::neighbor_* (src/multigrid/smoothers/
// bc_stencil_helpers.cuh:33-267): periodic/interior -> harmonic mean of K,
// Neumann -> 0, Dirichlet -> 2*K_C with ghost value 0.
// dir: 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+. Returns the neighbour's linear
// index (or -1 for a Neumann/Dirichlet face) and its coefficient.
// ---------------------------------------------------------------------------
// neighbour of cell (i, j, k) across face dir, or -1 where the face is a
// non-periodic boundary
__host__ __device__ inline long neighbour_index(int nx, int ny, int nz, const BcTypes& bc, int i,
                                                int j, int k, int dir) {
    int ni = i, nj = j, nk = k;
    bool at_bnd;
    switch (dir) {
    case 0: at_bnd = (i == 0); ni = at_bnd ? nx - 1 : i - 1; break;
    case 1: at_bnd = (i == nx - 1); ni = at_bnd ? 0 : i + 1; break;
    case 2: at_bnd = (j == 0); nj = at_bnd ? ny - 1 : j - 1; break;
    case 3: at_bnd = (j == ny - 1); nj = at_bnd ? 0 : j + 1; break;
    case 4: at_bnd = (k == 0); nk = at_bnd ? nz - 1 : k - 1; break;
    default: at_bnd = (k == nz - 1); nk = at_bnd ? 0 : k + 1; break;
    }
    if (at_bnd && bc.t[dir] != MF3D_BC_PERIODIC) return -1;
    return (long)ni + (long)nx * ((long)nj + (long)ny * (long)nk);
}

__host__ __device__ inline long neighbour_coef(int nx, int ny, int nz, const BcTypes& bc,
                                               const double* K, int i, int j, int k, int dir,
                                               double Kc, double& coef) {
    const long n = neighbour_index(nx, ny, nz, bc, i, j, k, dir);
    if (n < 0) {
        coef = (bc.t[dir] == MF3D_BC_DIRICHLET) ? 2.0 * Kc : 0.0;
        return -1;
    }
    coef = 2.0 / (1.0 / Kc + 1.0 / K[n]);
    return n;
}

__host__ __device__ inline bool interior_cell(int nx, int ny, int nz, int i, int j, int k) {
    return i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1;
}

// ---------------------------------------------------------------------------
// Device-resident iteration state of the Krylov solvers. Scalars never leave
// the device inside the iteration; the host reads this struct only at its
// convergence checks.
// ---------------------------------------------------------------------------
struct IterState {
    double rz;     // (r, z) of the current iterate
    double pAp;    // (p, A p)
    double alpha;  // rz / pAp
    double beta;   // rz_new / rz
    double rr;     // ||r||^2 of the recursive residual
    double den;    // stopping denominator: ||r0|| or ||b||
    double r0;     // ||r0||
    double bnorm;  // ||b||
    double rtol;   // relative tolerance
    double rzo;    // (r_new, z_old) of the flexible variant
    int iter;      // completed iterations
    int stop;      // 0 running, 1 converged, 2 breakdown, 3 iteration cap
    int check_every;
    int max_iter;
    int stop_norm; // MF3D_STOP_*
    int flexible;  // beta = (r_new, z_new - z_old) / (r_old, z_old)
    int xpend;     // x += alpha p still owed (x updated with the next p, see FIN_RR)
    double xa;     // this iteration's x step: alpha once (r, r) was formed, else 0 (FIN_RR)
};

// The PCG residual update folded into the first half-sweep of the FP32
// V-cycle on the finest level: r -= alpha Ap (skipped once stopped) before r
// becomes b, with (r, r) partials, one per sweep workgroup.
struct RUpd {
    const IterState* st;
    const double* Ap;
    double* r;
    double* partials;
};

// The (r, z) partials of PCG folded into the last half-sweep of the FP32
// V-cycle on the finest level (z = x32 of both colours once it ends): one
// partial per sweep workgroup.
struct RzDot {
    const double* r;
    double* partials;
};

// Reduction finalisers (what thread 0 does with the reduced sum S).
enum Finalize : int {
    FIN_STORE = 0,   // *out = S
    FIN_SQRT = 1,    // *out = sqrt(S)
    FIN_R0 = 2,      // r0 = sqrt(S); den from stop_norm; early exit if r0 < 1e-15
    FIN_RZ0 = 3,     // rz = S
    FIN_PAP = 4,     // pAp = S, alpha = rz / pAp
    FIN_RR = 5,      // rr = S; ++iter; convergence / cap test
    FIN_RZ = 6,      // beta = S / rz; rz = S
    FIN_BNORM = 7,   // bnorm = sqrt(S)
    FIN_CG_RR0 = 8,  // plain CG: rz = S (= r.r), r0 = sqrt(S), den
    FIN_CG_PAP = 9,  // plain CG: pAp = S, breakdown test every check_every, alpha
    FIN_CG_RR = 10,  // plain CG: rr_new = S, ++iter, test, beta = S / rz, rz = S
    FIN_RZO = 11,    // rzo = S
    FIN_RESUME = 12  // S = ||b - A x||^2: rr = S; keep "converged" only if it meets rtol
};

constexpr int kReduceBlock = 256;   // threads per first-stage block
constexpr int kFinalBlock = 256;    // threads of the second stage (= first-stage workgroups)
constexpr int kReduceCells = 2048;  // cells per first-stage partial (8 per thread, 256 threads)

// ---------------------------------------------------------------------------
// Batched solves: one launch covers many realisations. Entry e of a launch
// owns the blocks [start[e], start[e+1]) (start in device memory, m + 1
// values); each block finds its entry by binary search and works on that
// entry's descriptor. A realisation whose state says stop skips guarded work.
// ---------------------------------------------------------------------------
struct BatchLevel {  // one FP32 V-cycle level of one realisation
    int nx, ny, nz;
    BcTypes bc;
    int pin;
    float bs, rs;          // cs dx^2 and 1 / (cs dx^2)
    float* x;
    float* b;
    float* bcoarse;        // right-hand side of the next level (restriction target)
    const float* xc;       // correction of the next level (prolongation source)
    const double* r64;     // finest level: FP64 right-hand side converted into b
    const void* fx;        // faces (natural MAC or colour-split, type of the solve)
    const void* fy;
    const void* fz;
    const void* ux;
    const void* uy;
    const void* uz;
    const double* Ainv;    // coarsest level: dense inverse
    int nc;                // coarsest level: cells
    const IterState* st;
};

struct BatchVec {  // FP64 finest level and Krylov vectors of one realisation
    int nx, ny, nz;
    BcTypes bc;
    int pin;
    double inv_dx2;
    const double* hk;  // 1 / (2 K) of the finest level (FP64 operator)
    double* x;
    double* r;
    double* z;
    double* p;
    double* Ap;
    double* rhs;
    float* b32;
    float* x32;
    double* partials;
    unsigned np;      // partials of the 256-cell reductions (reduce_blocks)
    unsigned np_op;   // partials of the operator (op_partials)
    int otiles, okz;  // operator geometry (op_geom)
    int opairs;       // operator variant: 0 single cells, 1 pairs, 2 plane tiles (op_plan)
    IterState* st;
    double* out;  // scalar result of a FIN_STORE / FIN_SQRT second stage
};

struct BandDesc {  // banded Cholesky factor of one coarsest operator and its inverse
    int n, w;
    const double* Lb;
    double* Y;     // n * n: row c = column c of L^{-1}
    double* Ainv;  // n * n
};

// One coarsest operator M = -A to assemble from K and factor on the device
// (natural order, band width w)
struct FactorDesc {
    int nx, ny, nz, w;
    double dx2;
    BcTypes bc;
    int pin;
    const double* K;
    double* C;   // n * (w + 1) scratch: M in column storage, C[j*(w+1) + (i-j)] = M[i][j]
    double* Lb;  // n * (w + 1) factor, row-major band: Lb[i*(w+1) + (i-j)] = L[i][j]
    int* flag;   // written last: 0, or j + 1 for the first pivot j that is not positive
};

__device__ __forceinline__ int batch_entry(const int* __restrict__ start, int m, long blk) {
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace mf3d
