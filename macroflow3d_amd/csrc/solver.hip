// Host orchestration of the head solve and the C ABI (include/mf3d.h).
//
// The iteration keeps every Krylov scalar on the device (IterState): the host
// enqueues whole chunks of iterations and reads the state only at its
// convergence checks. Device kernels stop updating x and r as soon as the
// device-side test fires, so the reported iteration count is exact whatever
// the chunk length.

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "faces.hpp"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace mf3d {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }
int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int validate_grid(const mf3d_grid* g) {
    if (!g) return fail(MF3D_ERR_INVALID_ARGUMENT, "grid is null");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "grid dimensions must be positive");
    if (!(g->dx > 0.0) || !std::isfinite(g->dx))
        return fail(MF3D_ERR_INVALID_ARGUMENT, "grid spacing must be positive and finite");
    if ((double)g->nx * g->ny * g->nz > 4.0e9)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "grid larger than 2^32 cells");
    return MF3D_OK;
}

// BCSpec::validate (src/core/BCSpec.hpp:63-87)
int validate_bc(const mf3d_bc* bc) {
    if (!bc) return fail(MF3D_ERR_INVALID_ARGUMENT, "bc is null");
    for (int f = 0; f < 6; ++f) {
        if (bc->type[f] > MF3D_BC_PERIODIC)
            return fail(MF3D_ERR_INVALID_ARGUMENT, "unknown boundary type");
        if (!std::isfinite(bc->value[f]))
            return fail(MF3D_ERR_INVALID_ARGUMENT, "BC values must be finite");
    }
    static const char* axis[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; ++a) {
        const bool p0 = bc->type[2 * a] == MF3D_BC_PERIODIC;
        const bool p1 = bc->type[2 * a + 1] == MF3D_BC_PERIODIC;
        if (p0 != p1)
            return fail(MF3D_ERR_INVALID_ARGUMENT,
                        std::string("Periodic BC must be specified on both ") + axis[a] +
                            "min and " + axis[a] + "max");
    }
    return MF3D_OK;
}
}  // namespace mf3d

using namespace mf3d;

// ---------------------------------------------------------------------------
// workspace
// ---------------------------------------------------------------------------
namespace {

struct Level {
    mf3d_grid g{};
    size_t n = 0;
    double* x = nullptr;    // FP64 V-cycle correction (levels > 0)
    double* b = nullptr;    // FP64 V-cycle right-hand side (levels > 0)
    double* K = nullptr;    // coarsened conductivity (levels > 0)
    float* x32 = nullptr;   // FP32 V-cycle buffers
    float* b32 = nullptr;
    FaceSet f64;            // FP64 faces (operator on level 0, FP64 V-cycle)
    FaceSet flo;            // low-precision faces of the FP32 V-cycle
    bool has_f64 = false;
    bool has_flo = false;
    double cs = 1.0;        // storage scale of flo
    int order = ORDER_NATURAL;  // storage order of x32/b32
    // FP64 V-cycle: colour-split x, b and h = 1 / (2 K) (no stored faces), or
    // natural order with the FP64 faces f64
    int order64 = ORDER_NATURAL;
    double* hs = nullptr;
    SmoothPlan plan;
};

inline hipStream_t as_stream(mf3d_stream s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

struct mf3d_workspace {
    // hierarchy (MGHierarchy, src/multigrid/mg_types.hpp:75-98)
    mf3d_grid fine{};
    int req_levels = 0;
    std::vector<Level> lv;
    // dense coarsest inverse
    double* coarse_inv = nullptr;  // n*n inverse, the n*(w+1) band factor, n*n for L^{-1}
    int coarse_n = 0;
    size_t coarse_cap = 0;         // elements allocated at coarse_inv
    // single solves: the coarsest K travels to the host when the set-up is
    // enqueued, a host thread factors it, and the inverse is built on
    // coarse_stream; the solve's stream waits for it (inv_ev) only before the
    // first coarsest solve
    bool coarse_pending = false;
    double* host_kc = nullptr;     // pinned coarsest K
    double* host_band = nullptr;   // pinned band factor (source of the upload)
    size_t host_kc_cap = 0, host_band_cap = 0;
    hipEvent_t kc_ev = nullptr;
    hipEvent_t inv_ev = nullptr;   // the inverse is complete (recorded on coarse_stream)
    hipStream_t coarse_stream = nullptr;
    std::thread coarse_worker;     // factors as soon as the coarsest K has arrived
    std::vector<double> coarse_band_h;
    int coarse_status = 0;
    std::string coarse_msg;
    // batched solves factor on the device: pinned host int the factor kernel
    // writes (0: definite, j + 1: first pivot j that is not positive)
    int* coarse_flag = nullptr;
    // batched solves led by this workspace: their stream and phase events
    hipStream_t batch_stream = nullptr;
    hipEvent_t batch_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    void join_coarse() {
        if (coarse_worker.joinable()) coarse_worker.join();
    }
    // the coarse stream's work reads and writes workspace memory
    void drain_coarse() {
        join_coarse();
        if (coarse_stream) (void)hipStreamSynchronize(coarse_stream);
    }
    // Krylov vectors (PCGWorkspace, src/numerics/solvers/pcg.cuh:45-62) + rhs
    size_t nvec = 0;
    double *rhs = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr;
    // reductions
    double* partials = nullptr;
    size_t partials_cap = 0;
    IterState* st = nullptr;
    double* dscalar = nullptr;
    IterState* host_st = nullptr;  // pinned
    double* host_scalar = nullptr; // pinned
    // set-up snapshot
    mf3d_bc bc{};
    BcTypes bct{};
    bool pin = false;
    int pre = 2, post = 2, coarse_iters = 50, coarse_solver = 0;
    int precision = MF3D_PRECOND_FP64;
    CoefType lo_type = COEF_BF16;
    const double* K0 = nullptr;
    double* h0 = nullptr;  // 1 / (2 K) of the finest level: the FP64 operator's faces
    // the operator may run on cell pairs (decided per solve from the vectors'
    // alignment, the same way for single and batched solves)
    bool op_pairs = true;
    bool op_tile = true;  // plane-tile operator where it applies (MF3D_TUNE_OP_TILE)
    int op_chunk = kern::kOpChunkDefault;        // operator planes per workgroup, at most
    long march_blocks = kern::kMarchDefaultBlocks;  // FP64 split sweeps' workgroup target
    // single solves replay one captured PCG iteration (MF3D_TUNE_GRAPH: 1
    // always, 0 never, -1 when the finest level has at most kGraphMaxCells
    // cells); the capture runs on this private stream (created on first use)
    int graph = -1;
    hipStream_t capture_stream = nullptr;
    // finest-level operator and preconditioner applications enqueued since
    // the last single solve began (replays of a captured iteration included)
    long n_op = 0, n_pc = 0;
    // partials of ||b - A x||^2 that a passing true-residual check left in
    // `partials` for the solve's final x (0: none), so that the reported true
    // residual needs no second operator pass
    unsigned true_rr_np = 0;
    // aligned stand-in for a head vector that is not 16-byte aligned
    double* hcopy = nullptr;
    size_t hcopy_n = 0;
    bool ready = false;
    // tuning (mf3d_workspace_set_tuning / autotuner)
    bool split = true;  // colour-split FP32 levels where the sizes allow
    int sweep_block = 256;  // threads per workgroup of the colour-split sweeps
    bool fold_first = true;  // coarse first half-sweeps folded into the restrictions
    // PCG with z kept in x32: the r update (and its (r, r) partials, in
    // rr_partials) done by the V-cycle's first finest-level half-sweep
    bool fold_r = true;
    double* rr_partials = nullptr;
    // ... and the (r, z) partials by its last finest-level half-sweep
    bool fold_rz = true;
    size_t bytes = 0;
    std::vector<void*> owned;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

    int alloc_bytes(void** p, size_t nbytes) {
        void* q = nullptr;
        if (nbytes == 0) nbytes = 8;
        hipError_t e = hipMalloc(&q, nbytes);
        if (e != hipSuccess)
            return fail(MF3D_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
        owned.push_back(q);
        bytes += nbytes;
        *p = q;
        return MF3D_OK;
    }
    int alloc(double** p, size_t n) { return alloc_bytes((void**)p, n * sizeof(double)); }
    int alloc(float** p, size_t n) { return alloc_bytes((void**)p, n * sizeof(float)); }
    int alloc_faces(FaceSet& f, const mf3d_grid& g, CoefType t, bool split = false) {
        f.type = t;
        f.split = split;
        const size_t b = kern::coef_bytes(t);
        if (split) {
            const size_t n = cells(g);
            MF3D_TRY(alloc_bytes(&f.fx, n * b));
            MF3D_TRY(alloc_bytes(&f.fy, n * b));
            MF3D_TRY(alloc_bytes(&f.fz, n * b));
            MF3D_TRY(alloc_bytes(&f.ux, (size_t)g.ny * g.nz * b));
            MF3D_TRY(alloc_bytes(&f.uy, (size_t)g.nx * g.nz * b));
            MF3D_TRY(alloc_bytes(&f.uz, (size_t)g.nx * g.ny * b));
            return MF3D_OK;
        }
        f.ux = f.uy = f.uz = nullptr;
        MF3D_TRY(alloc_bytes(&f.fx, (size_t)(g.nx + 1) * g.ny * g.nz * b));
        MF3D_TRY(alloc_bytes(&f.fy, (size_t)g.nx * (g.ny + 1) * g.nz * b));
        MF3D_TRY(alloc_bytes(&f.fz, (size_t)g.nx * g.ny * (g.nz + 1) * b));
        return MF3D_OK;
    }
    void release() {
        drain_coarse();
        for (void* q : owned) (void)hipFree(q);
        owned.clear();
        bytes = 0;
        lv.clear();
        coarse_inv = nullptr;
        coarse_pending = false;
        coarse_n = 0;
        coarse_cap = 0;
        nvec = 0;
        rhs = r = z = p = Ap = nullptr;
        partials = nullptr;
        partials_cap = 0;
        rr_partials = nullptr;
        st = nullptr;
        dscalar = nullptr;
        hcopy = nullptr;
        hcopy_n = 0;
        h0 = nullptr;
        ready = false;
        req_levels = 0;
        fine = mf3d_grid{};
    }
    ~mf3d_workspace() {
        release();
        if (host_st) (void)hipHostFree(host_st);
        if (host_scalar) (void)hipHostFree(host_scalar);
        if (host_kc) (void)hipHostFree(host_kc);
        if (host_band) (void)hipHostFree(host_band);
        if (coarse_flag) (void)hipHostFree(coarse_flag);
        if (kc_ev) (void)hipEventDestroy(kc_ev);
        if (inv_ev) (void)hipEventDestroy(inv_ev);
        if (coarse_stream) (void)hipStreamDestroy(coarse_stream);
        if (capture_stream) (void)hipStreamDestroy(capture_stream);
        if (batch_stream) (void)hipStreamDestroy(batch_stream);
        for (auto& e : batch_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace {

bool same_grid(const mf3d_grid& a, const mf3d_grid& b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx && a.dy == b.dy &&
           a.dz == b.dz;
}

// Builds (or reuses) the hierarchy and vectors for this grid and level count.
int ensure_workspace(mf3d_workspace* ws, const mf3d_grid& g, int levels) {
    if (levels < 1) return fail(MF3D_ERR_INVALID_ARGUMENT, "mg_levels must be >= 1");
    if (ws->ready && same_grid(ws->fine, g) && ws->req_levels == levels) return MF3D_OK;
    ws->release();
    // halving rule of MGHierarchy (mg_types.hpp:75-98): stop before a dim < 2
    std::vector<mf3d_grid> grids;
    mf3d_grid cur = g;
    for (int l = 0; l < levels; ++l) {
        grids.push_back(cur);
        if (l < levels - 1) {
            cur.nx /= 2;
            cur.ny /= 2;
            cur.nz /= 2;
            cur.dx *= 2.0;
            cur.dy *= 2.0;
            cur.dz *= 2.0;
            if (cur.nx < 2 || cur.ny < 2 || cur.nz < 2) break;
        }
    }
    for (size_t l = 0; l + 1 < grids.size(); ++l) {
        const mf3d_grid& f = grids[l];
        if ((f.nx & 1) || (f.ny & 1) || (f.nz & 1))
            return fail(MF3D_ERR_SIZE_MISMATCH,
                        "grid level " + std::to_string(l) + " (" + std::to_string(f.nx) + "x" +
                            std::to_string(f.ny) + "x" + std::to_string(f.nz) +
                            ") is odd but not the coarsest: 2:1 restriction needs even sizes");
    }
    ws->lv.resize(grids.size());
    for (size_t l = 0; l < grids.size(); ++l) {
        Level& L = ws->lv[l];
        L.g = grids[l];
        L.n = cells(grids[l]);
        if (l > 0) {
            MF3D_TRY(ws->alloc(&L.x, L.n));
            MF3D_TRY(ws->alloc(&L.b, L.n));
            MF3D_TRY(ws->alloc(&L.K, L.n));
        }
    }
    const size_t n = cells(g);
    ws->nvec = n;
    MF3D_TRY(ws->alloc(&ws->rhs, n));
    MF3D_TRY(ws->alloc(&ws->r, n));
    MF3D_TRY(ws->alloc(&ws->z, n));
    MF3D_TRY(ws->alloc(&ws->p, n));
    MF3D_TRY(ws->alloc(&ws->Ap, n));
    // (the finest split4 sweeps write one per workgroup of >= 64 threads, the
    // FP64 split sweeps too, with 4 cells per thread, in at most
    // kMarchMaxBlocks + one plane's workgroups where chunks are shortened)
    const size_t march_cap =
        (size_t)kern::kMarchMaxBlocks + ((size_t)g.nx * g.ny / 4 + 63) / 64;
    ws->partials_cap = std::max<size_t>({kern::reduce_blocks(n), kern::op_partials(g), n / 256,
                                         march_cap}) + 64;
    MF3D_TRY(ws->alloc(&ws->partials, ws->partials_cap));
    // one per workgroup of the finest sweep (8 cells per thread, >= 64 threads;
    // FP64: one per march workgroup)
    MF3D_TRY(ws->alloc(&ws->rr_partials, std::max<size_t>(n / 512, march_cap) + 64));
    double* stp = nullptr;
    MF3D_TRY(ws->alloc(&stp, (sizeof(IterState) + 7) / 8));
    ws->st = reinterpret_cast<IterState*>(stp);
    MF3D_TRY(ws->alloc(&ws->dscalar, 8));
    if (!ws->host_st) MF3D_HIP_TRY(hipHostMalloc((void**)&ws->host_st, sizeof(IterState)));
    if (!ws->host_scalar) MF3D_HIP_TRY(hipHostMalloc((void**)&ws->host_scalar, sizeof(double)));
    for (auto& e : ws->ev)
        if (!e) MF3D_HIP_TRY(hipEventCreate(&e));
    ws->fine = g;
    ws->req_levels = levels;
    ws->ready = true;
    return MF3D_OK;
}

// reduction scratch of at least need doubles (grown, never shrunk)
int scratch_count(mf3d_workspace* ws, size_t need) {
    if (ws->partials && ws->partials_cap >= need) return MF3D_OK;
    double* p = nullptr;
    MF3D_TRY(ws->alloc(&p, need));
    ws->partials = p;
    ws->partials_cap = need;
    if (!ws->dscalar) MF3D_TRY(ws->alloc(&ws->dscalar, 8));
    return MF3D_OK;
}
int scratch_partials(mf3d_workspace* ws, size_t n) {
    return scratch_count(ws, kern::reduce_blocks(n) + 64);
}

int host_scalar_sync(mf3d_workspace* ws, hipStream_t s, double* out) {
    MF3D_HIP_TRY(hipMemcpyAsync(ws->host_scalar, ws->dscalar, sizeof(double),
                                hipMemcpyDeviceToHost, s));
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    *out = *ws->host_scalar;
    return MF3D_OK;
}

// ---------------------------------------------------------------------------
// Dense coarsest solve: M = -A of the coarsest level (smoother / residual
// operator, pin included), its banded Cholesky factor, and the explicit
// inverse -L^{-T} L^{-1} on the device (band_lower_kernel, syrk_neg_kernel);
// the V-cycle's coarsest solve is then one GEMV. A single solve factors on a
// host thread while its set-up runs on the GPU; a batched solve factors every
// realisation on the device (band_factor_kernel, one workgroup each), with no
// host round trip.
// ---------------------------------------------------------------------------
const double* coarsest_K(const mf3d_workspace* ws) {
    return ws->lv.size() == 1 ? ws->K0 : ws->lv.back().K;
}

// band width of the coarsest operator in natural order: the largest index
// distance between neighbouring cells (periodic faces included)
int coarse_bandwidth(const mf3d_workspace* ws) {
    const mf3d_grid& g = ws->lv.back().g;
    int w = 0;
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                const int c = i + g.nx * (j + g.ny * k);
                for (int d = 0; d < 6; ++d) {
                    const long nb = neighbour_index(g.nx, g.ny, g.nz, ws->bct, i, j, k, d);
                    if (nb >= 0) w = std::max(w, std::abs((int)nb - c));
                }
            }
    return w;
}

// Banded Cholesky factor of M = -A (coarsest level) from its host K: returns
// the band width w and fills band (row-major, band[i*(w+1) + (i-j)] = L[i][j]).
// The factorization runs right-looking on column storage (C[j*(w+1) + (i-j)] =
// M[i][j], i >= j): column j is scaled, then its outer product is subtracted
// from the next w columns with unit-stride updates. Every entry receives its
// terms in the same order as the row-wise formula L[i][j] = (M[i][j] -
// sum_p L[i][p] L[j][p]) / L[j][j], p ascending.
int coarse_band(const mf3d_workspace* ws, const double* Kc, std::vector<double>& band, int* wout) {
    const Level& L = ws->lv.back();
    const int n = (int)L.n;
    const mf3d_grid& g = L.g;
    const double dx2 = g.dx * g.dx;
    const int w = coarse_bandwidth(ws);
    const int W1 = w + 1;
    std::vector<double> C((size_t)n * W1, 0.0);
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                const int c = i + g.nx * (j + g.ny * k);
                double coef[6], asum = 0.0;
                long nb[6];
                for (int d = 0; d < 6; ++d) {
                    nb[d] = neighbour_coef(g.nx, g.ny, g.nz, ws->bct, Kc, i, j, k, d, Kc[c],
                                           coef[d]);
                    asum += coef[d];
                }
                if (ws->pin && c == 0) asum *= 2.0;
                C[(size_t)c * W1] += asum / dx2;
                for (int d = 0; d < 6; ++d) {  // symmetric part: half of each (c, nb) entry
                    if (nb[d] < 0) continue;
                    const int q = (int)nb[d];
                    const double e = -coef[d] / dx2;
                    if (q < c) C[(size_t)q * W1 + (c - q)] += 0.5 * e;
                    if (q > c) C[(size_t)c * W1 + (q - c)] += 0.5 * e;
                    if (q == c) C[(size_t)c * W1] += e;  // self-wrap (dimension 1)
                }
            }
    double* cm = C.data();
    for (int j = 0; j < n; ++j) {
        double* cj = cm + (size_t)j * W1;
        if (!(cj[0] > 0.0))
            return fail(MF3D_ERR_SINGULAR, "coarsest operator is not definite at pivot " +
                                               std::to_string(j) +
                                               " (pure Neumann/periodic problem without pin?)");
        const double Ljj = std::sqrt(cj[0]);
        cj[0] = Ljj;
        const int m = std::min(w, n - 1 - j);
        for (int d = 1; d <= m; ++d) cj[d] /= Ljj;
        for (int dk = 1; dk <= m; ++dk) {  // column j + dk -= L[.][j] L[j+dk][j]
            double* ck = cm + (size_t)(j + dk) * W1;
            const double lkj = cj[dk];
            const double* src = cj + dk;
            const int len = m - dk + 1;
            for (int t = 0; t < len; ++t) ck[t] -= src[t] * lkj;
        }
    }
    band.assign((size_t)n * W1, 0.0);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d <= std::min(w, i); ++d) band[(size_t)i * W1 + d] = C[(size_t)(i - d) * W1 + d];
    *wout = w;
    return MF3D_OK;
}

// the n*n inverse, the n*(w+1) band factor and n*n for L^{-1}, whose room
// holds the assembled operator while a device factor is computed (w < n)
int reserve_coarse(mf3d_workspace* ws, int n, int w) {
    const size_t need = 2 * (size_t)n * n + (size_t)n * (w + 1);
    if (!ws->coarse_inv || ws->coarse_cap < need) {
        MF3D_TRY(ws->alloc(&ws->coarse_inv, need));
        ws->coarse_cap = need;
    }
    ws->coarse_n = n;
    return MF3D_OK;
}

double* coarse_band_dev(mf3d_workspace* ws, int n) { return ws->coarse_inv + (size_t)n * n; }
double* coarse_lower(mf3d_workspace* ws, int n, int w) {
    return ws->coarse_inv + (size_t)n * n + (size_t)n * (w + 1);
}

int grow_pinned(double** p, size_t* cap, size_t need) {
    if (*p && *cap >= need) return MF3D_OK;
    if (*p) MF3D_HIP_TRY(hipHostFree(*p));
    *p = nullptr;
    *cap = 0;
    MF3D_HIP_TRY(hipHostMalloc((void**)p, need * sizeof(double)));
    *cap = need;
    return MF3D_OK;
}

// a pending inverse that no solve will use: s waits for the coarse stream's
// writes to the workspace before it reuses the memory
int discard_coarse_inverse(mf3d_workspace* ws, hipStream_t s) {
    if (!ws->coarse_pending) return MF3D_OK;
    ws->coarse_pending = false;
    ws->join_coarse();
    if (ws->inv_ev) MF3D_HIP_TRY(hipStreamWaitEvent(s, ws->inv_ev, 0));
    return MF3D_OK;
}

// Single solve: start_coarse_inverse enqueues the download of the coarsest K
// (after the coarsening kernels) and starts the worker thread, which waits for
// the download, factors, uploads the factor and launches the inverse on the
// workspace's second stream; the solve's stream waits for that
// (finish_coarse_inverse) only before its first coarsest solve, so the
// factorization and the inverse overlap the set-up and the first descent of
// the V-cycle. Allocation stays on this thread.
int start_coarse_inverse(mf3d_workspace* ws, hipStream_t s) {
    const int n = (int)ws->lv.back().n;
    if (n > 4096) return fail(MF3D_ERR_UNSUPPORTED, "dense coarsest solve limited to 4096 cells");
    // a set-up not followed by a solve: s waits for its inverse, so that the
    // download below (kc_ev) completes after that inverse's upload from
    // host_band, which the new worker overwrites once kc_ev has completed
    MF3D_TRY(discard_coarse_inverse(ws, s));
    const int w = coarse_bandwidth(ws);
    MF3D_TRY(reserve_coarse(ws, n, w));
    // an earlier consumed inverse's upload from host_band was waited for by the
    // stream work that consumed it, which precedes the download below; growing
    // host_band frees it, so then no upload from it may be in flight at all
    if (ws->coarse_stream && ws->host_band_cap < (size_t)n * (w + 1))
        MF3D_HIP_TRY(hipStreamSynchronize(ws->coarse_stream));
    MF3D_TRY(grow_pinned(&ws->host_kc, &ws->host_kc_cap, (size_t)n));
    MF3D_TRY(grow_pinned(&ws->host_band, &ws->host_band_cap, (size_t)n * (w + 1)));
    if (!ws->kc_ev) MF3D_HIP_TRY(hipEventCreateWithFlags(&ws->kc_ev, hipEventDisableTiming));
    if (!ws->inv_ev) MF3D_HIP_TRY(hipEventCreateWithFlags(&ws->inv_ev, hipEventDisableTiming));
    if (!ws->coarse_stream)
        MF3D_HIP_TRY(hipStreamCreateWithFlags(&ws->coarse_stream, hipStreamNonBlocking));
    MF3D_HIP_TRY(hipMemcpyAsync(ws->host_kc, coarsest_K(ws), n * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    MF3D_HIP_TRY(hipEventRecord(ws->kc_ev, s));
    // the inverse overwrites coarse_inv: the previous solve's stream work is done
    ws->coarse_pending = true;
    ws->coarse_worker = std::thread([ws, n, w]() {
        ws->coarse_status = MF3D_OK;
        auto failed = [ws](int st, const char* msg) {
            ws->coarse_status = st;
            ws->coarse_msg = msg;
        };
        if (hipEventSynchronize(ws->kc_ev) != hipSuccess)
            return failed(MF3D_ERR_HIP, "coarsest K download failed");
        int wb = 0;
        const int st = coarse_band(ws, ws->host_kc, ws->coarse_band_h, &wb);
        if (st != MF3D_OK) return failed(st, last_error_cstr());
        if (wb != w) return failed(MF3D_ERR_HIP, "coarsest band width changed");
        std::memcpy(ws->host_band, ws->coarse_band_h.data(),
                    ws->coarse_band_h.size() * sizeof(double));
        double* dband = coarse_band_dev(ws, n);
        if (hipMemcpyAsync(dband, ws->host_band, ws->coarse_band_h.size() * sizeof(double),
                           hipMemcpyHostToDevice, ws->coarse_stream) != hipSuccess)
            return failed(MF3D_ERR_HIP, "band factor upload failed");
        const int bs = kern::band_inverse(ws->coarse_stream, n, w, dband, coarse_lower(ws, n, w),
                                          ws->coarse_inv);
        if (bs != MF3D_OK) return failed(bs, last_error_cstr());
        if (hipEventRecord(ws->inv_ev, ws->coarse_stream) != hipSuccess)
            return failed(MF3D_ERR_HIP, "event record failed");
    });
    return MF3D_OK;
}

int finish_coarse_inverse(mf3d_workspace* ws, hipStream_t s) {
    if (!ws->coarse_pending) return MF3D_OK;
    ws->coarse_pending = false;
    ws->join_coarse();
    if (ws->coarse_status != MF3D_OK) return fail(ws->coarse_status, ws->coarse_msg);
    MF3D_HIP_TRY(hipStreamWaitEvent(s, ws->inv_ev, 0));
    return MF3D_OK;
}

// Batched solves: the factor descriptor of ws's coarsest operator (after
// reserve_coarse)
int coarse_desc(mf3d_workspace* ws, int w, FactorDesc* D) {
    if (!ws->coarse_flag)
        MF3D_HIP_TRY(hipHostMalloc((void**)&ws->coarse_flag, sizeof(int), hipHostMallocCoherent));
    const Level& L = ws->lv.back();
    const int n = (int)L.n;
    *D = FactorDesc{L.g.nx, L.g.ny, L.g.nz, w, L.g.dx * L.g.dx, ws->bct, ws->pin ? 1 : 0,
                    coarsest_K(ws), coarse_lower(ws, n, w), coarse_band_dev(ws, n),
                    ws->coarse_flag};
    return MF3D_OK;
}

CoefType coef_type_of(int fmt) {
    return fmt == MF3D_COEF_FP32 ? COEF_F32 : fmt == MF3D_COEF_FP16 ? COEF_F16 : COEF_BF16;
}

inline const double* level_K(const mf3d_workspace* ws, size_t l) {
    return l == 0 ? ws->K0 : ws->lv[l].K;
}

// Hierarchy coefficients for a K field (solve_head.cu:106-114): coarsened K
// (coarsen_K.cu:15-53), face coefficients per level, smoothing plans and, with
// a DIRECT coarse solver, the dense coarsest inverse.
int mg_setup(mf3d_workspace* ws, hipStream_t s, const mf3d_grid& g, const mf3d_bc& bc,
             const mf3d_head_solve_config& cfg, const double* K, bool defer_coarse = false) {
    MF3D_TRY(ensure_workspace(ws, g, cfg.mg_levels));
    ws->bc = bc;
    ws->bct = bc_types(bc);
    ws->pin = pin_enabled(cfg.pin_mode, bc);
    ws->pre = cfg.mg_pre_smooth;
    ws->post = cfg.mg_post_smooth;
    ws->coarse_iters = cfg.mg_coarse_iters;
    ws->coarse_solver = cfg.coarse_solver;
    ws->precision = cfg.precond_precision;
    ws->lo_type = coef_type_of(cfg.coef_format);
    ws->K0 = K;
    const bool mixed = ws->precision == MF3D_PRECOND_FP32;
    // fp16 coefficients are stored divided by the geometric mean of K (a power
    // of two), so both tails of a log-normal field stay inside fp16's range
    double cs = 1.0;
    if (mixed && ws->lo_type == COEF_F16) {
        const size_t n0 = cells(g);
        MF3D_TRY(kern::sum_log_partials(s, K, n0, ws->partials));
        MF3D_TRY(kern::finalize(s, ws->partials, kern::reduce_blocks(n0), FIN_STORE, nullptr,
                                ws->dscalar));
        double slog = 0.0;
        MF3D_TRY(host_scalar_sync(ws, s, &slog));
        const double mean_log2 = slog / (double)n0 / std::log(2.0);
        cs = std::ldexp(1.0, -(int)std::lround(mean_log2));
    }
    for (size_t l = 1; l < ws->lv.size(); ++l)
        MF3D_TRY(kern::coarsen(s, ws->lv[l].g, level_K(ws, l - 1), ws->lv[l].K));
    if (cfg.coarse_solver == MF3D_COARSE_DIRECT && !defer_coarse) {
        MF3D_TRY(start_coarse_inverse(ws, s));
    } else {
        MF3D_TRY(discard_coarse_inverse(ws, s));
    }
    if (!ws->h0) MF3D_TRY(ws->alloc(&ws->h0, cells(g)));
    // FP64 V-cycle: colour-split levels (x, b, h = 1 / (2 K) in split order,
    // no stored faces) on a prefix of colour-swept levels, the rest natural
    // with stored FP64 faces; the finest split level also writes the
    // operator's natural-order h in the same pass
    bool prefix64 = !mixed && ws->split;
    for (size_t l = 0; l < ws->lv.size(); ++l) {
        Level& L = ws->lv[l];
        L.plan.kind = L.n <= 4096 ? SMOOTH_SMALL : SMOOTH_COLOUR;
        L.plan.block = ws->sweep_block;
        prefix64 = prefix64 && l + 1 < ws->lv.size() && L.plan.kind == SMOOTH_COLOUR &&
                   kern::split64_ok(L.g);
        L.order64 = prefix64 ? ORDER_SPLIT : ORDER_NATURAL;
        if (mixed) continue;
        if (L.order64 == ORDER_SPLIT) {
            if (!L.hs) MF3D_TRY(ws->alloc(&L.hs, L.n));
            if (l == 0) {  // level 0's x and b (the others have theirs)
                if (!L.x) MF3D_TRY(ws->alloc(&L.x, L.n));
                if (!L.b) MF3D_TRY(ws->alloc(&L.b, L.n));
            }
            MF3D_TRY(kern::split_h(s, L.g, level_K(ws, l), L.hs, l == 0 ? ws->h0 : nullptr));
            continue;
        }
        if (!L.has_f64) {  // FP64 faces of a natural-order level
            MF3D_TRY(ws->alloc_faces(L.f64, L.g, COEF_F64));
            L.has_f64 = true;
        }
        MF3D_TRY(kern::faces_build(s, L.g, ws->bct, 1.0, level_K(ws, l), L.f64));
    }
    // colour-split FP32 vectors and faces on a prefix of colour-swept levels
    // (never the coarsest, whose direct or small-level solve is natural order)
    bool prefix = mixed && ws->split;
    for (size_t l = 0; l < ws->lv.size(); ++l) {
        Level& L = ws->lv[l];
        prefix = prefix && l + 1 < ws->lv.size() && L.plan.kind == SMOOTH_COLOUR &&
                 kern::split_ok(L.g);
        L.order = prefix ? ORDER_SPLIT : ORDER_NATURAL;
        if (!mixed) continue;
        const bool split = L.order == ORDER_SPLIT;
        if (!L.has_flo || L.flo.type != ws->lo_type || L.flo.split != split) {
            MF3D_TRY(ws->alloc_faces(L.flo, L.g, ws->lo_type, split));
            L.has_flo = true;
        }
        L.cs = cs;
        if (l == 0 && split) {
            if (!(ws->split && kern::split_ok(g)))
                return fail(MF3D_ERR_UNSUPPORTED, "finest level split without split_ok");
            MF3D_TRY(kern::faces_build_split_h(s, L.g, ws->bct, cs, level_K(ws, l), L.flo,
                                               ws->h0));
        } else {
            MF3D_TRY(kern::faces_build(s, L.g, ws->bct, cs, level_K(ws, l), L.flo));
        }
        if (!L.x32) MF3D_TRY(ws->alloc(&L.x32, L.n));
        if (!L.b32) MF3D_TRY(ws->alloc(&L.b32, L.n));
    }
    // the FP64 operator's h = 1 / (2 K): built together with the 16-bit faces
    // when level 0 is the colour-split level of the FP32 V-cycle (decided
    // above), on its own otherwise (FP64 V-cycle, a finest level too small
    // for colour-split storage, a single level, odd sizes)
    const bool h0_built = mixed ? ws->lv[0].order == ORDER_SPLIT
                                : ws->lv[0].order64 == ORDER_SPLIT;
    if (!h0_built) MF3D_TRY(kern::half_recip(s, K, ws->h0, cells(g)));
    return MF3D_OK;
}

// The coarsest factors and inverses of many workspaces, two launches on s
// with no host round trip; the pinned host descriptors (keep_host) and their
// device copy (keep_dev) must live until s completes
int build_coarse_inverses(mf3d_workspace* const* ws, int count, hipStream_t s,
                          std::vector<void*>& keep_host, std::vector<void*>& keep_dev) {
    const size_t bytes = sizeof(FactorDesc) * count + sizeof(BandDesc) * count +
                         sizeof(int) * (count + 1);
    void* hbuf = nullptr;
    MF3D_HIP_TRY(hipHostMalloc(&hbuf, bytes, hipHostMallocDefault));
    keep_host.push_back(hbuf);
    FactorDesc* fh = static_cast<FactorDesc*>(hbuf);
    BandDesc* dh = reinterpret_cast<BandDesc*>(fh + count);
    int* sh = reinterpret_cast<int*>(dh + count);
    int nmax = 0, wmax = 0;
    sh[0] = 0;
    for (int r = 0; r < count; ++r) {
        const int n = (int)ws[r]->lv.back().n;
        if (n > 4096) return fail(MF3D_ERR_UNSUPPORTED, "dense coarsest solve limited to 4096 cells");
        const int w = coarse_bandwidth(ws[r]);
        MF3D_TRY(reserve_coarse(ws[r], n, w));
        MF3D_TRY(coarse_desc(ws[r], w, &fh[r]));
        dh[r] = BandDesc{n, w, fh[r].Lb, fh[r].C, ws[r]->coarse_inv};
        sh[r + 1] = sh[r] + n;
        nmax = std::max(nmax, n);
        wmax = std::max(wmax, w);
    }
    void* dbuf = nullptr;
    MF3D_HIP_TRY(hipMalloc(&dbuf, bytes));
    keep_dev.push_back(dbuf);
    MF3D_HIP_TRY(hipMemcpyAsync(dbuf, hbuf, bytes, hipMemcpyHostToDevice, s));
    const FactorDesc* F = static_cast<const FactorDesc*>(dbuf);
    const BandDesc* D = reinterpret_cast<const BandDesc*>(F + count);
    const int* st = reinterpret_cast<const int*>(D + count);
    MF3D_TRY(kern::band_factor_batch(s, F, count, wmax));
    return kern::band_inverse_batch(s, D, st, count, sh[count], nmax, wmax);
}

// per-type views of a level's V-cycle buffers
template <typename T> struct LevelView;
template <> struct LevelView<double> {
    static double* x(Level& L) { return L.x; }
    static double* b(Level& L) { return L.b; }
    static const FaceSet& F(const Level& L) { return L.f64; }
    static double cs(const Level&) { return 1.0; }
    static int order(const Level&) { return ORDER_NATURAL; }
};
template <> struct LevelView<float> {
    static float* x(Level& L) { return L.x32; }
    static float* b(Level& L) { return L.b32; }
    static const FaceSet& F(const Level& L) { return L.flo; }
    static double cs(const Level& L) { return L.cs; }
    static int order(const Level& L) { return L.order; }
};

// v_cycle_recursive (src/multigrid/cycle/v_cycle.cu:36-74) in precision T.
// zero_init: x is zero on entry (the fills of v_cycle.cu:62 and
// mg_preconditioner.cu:37 are folded into the first smoothing sweep); the
// prolongation (v_cycle.cu:69) is folded into the first post-smoothing sweep.
// red_done: level l's first red half-sweep from x = 0 was folded into the
// restriction from level l - 1
template <typename T>
int vcycle(mf3d_workspace* ws, hipStream_t s, size_t l, T* x, const T* b, bool zero_init,
           const double* r64 = nullptr, bool red_done = false, const RUpd* ru = nullptr,
           const RzDot* rz = nullptr) {
    using V = LevelView<T>;
    const size_t nl = ws->lv.size();
    Level& L = ws->lv[l];
    const double dx2 = L.g.dx * L.g.dx;
    const double cs = V::cs(L);
    const FaceSet& F = V::F(L);
    if (l == nl - 1) {
        if (ws->coarse_solver == MF3D_COARSE_DIRECT) {
            MF3D_TRY(finish_coarse_inverse(ws, s));
            return kern::gemv(s, (int)L.n, ws->coarse_inv, b, x);
        }
        return kern::smooth_faces(s, L.g, ws->bct, ws->pin, cs * dx2, x, b, F,
                                  ws->coarse_iters, zero_init, (const T*)nullptr, L.plan);
    }
    Level& C = ws->lv[l + 1];
    const int fo = V::order(L), co = V::order(C);
    MF3D_TRY(kern::smooth_faces(s, L.g, ws->bct, ws->pin, cs * dx2, x, b, F, ws->pre,
                                zero_init, (const T*)nullptr, L.plan, fo, co, r64, red_done,
                                ru));
    // after pre-smoothing that ends with a black half-sweep, the split
    // levels evaluate only the red children's residuals; a colour-split
    // coarse level (not the coarsest) gets its first red half-sweep here too
    const bool fold = std::is_same<T, float>::value && ws->fold_first && ws->pre > 0 &&
                      fo == ORDER_SPLIT && co == ORDER_SPLIT && l + 2 < nl &&
                      C.plan.kind == SMOOTH_COLOUR;
    const double bsc = V::cs(C) * C.g.dx * C.g.dx;
    MF3D_TRY(kern::residual_restrict_faces(s, L.g, ws->bct, ws->pin, 1.0 / (cs * dx2), x, b, F,
                                           V::b(C), fo, co, ws->pre > 0,
                                           fold ? V::x(C) : (T*)nullptr, fold ? &V::F(C) : nullptr,
                                           bsc));
    MF3D_TRY(vcycle<T>(ws, s, l + 1, V::x(C), V::b(C), true, nullptr, fold));
    return kern::smooth_faces(s, L.g, ws->bct, ws->pin, cs * dx2, x, b, F, ws->post,
                              false, (const T*)V::x(C), L.plan, fo, co, nullptr, false, nullptr,
                              rz);
}

// The FP64 V-cycle from colour-split level l (v_cycle.cu:36-74): pre-smoothing
// (zero start; on the finest level b taken from the natural-order r64 by the
// first half-sweep), residual of the red children + restriction, the coarser
// levels (split while the sizes allow, then natural with stored faces),
// post-smoothing with the prolongation folded into its first half-sweep. z:
// the last half-sweep writes the result there in natural order (with the
// (rdot, z) partials when rdot is given) instead of into the split x.
int vcycle64(mf3d_workspace* ws, hipStream_t s, size_t l, bool zero_init, const double* r64,
             double* z, const double* rdot, double* partials, unsigned* np,
             const RUpd* ru = nullptr) {
    Level& L = ws->lv[l];
    Level& C = ws->lv[l + 1];  // a split level is never the coarsest
    const double dx2 = L.g.dx * L.g.dx;
    const int tb = L.plan.block;
    MF3D_TRY(kern::smooth_split64(s, L.g, ws->bct, ws->pin, dx2, L.x, L.b, L.hs, ws->pre,
                                  zero_init, nullptr, ORDER_NATURAL, tb, r64, nullptr, nullptr,
                                  nullptr, nullptr, ws->march_blocks, ru));
    MF3D_TRY(kern::residual_restrict_split64(s, L.g, ws->bct, ws->pin, 1.0 / dx2, L.x, L.b, L.hs,
                                             C.b, C.order64, ws->pre > 0));
    if (C.order64 == ORDER_SPLIT)
        MF3D_TRY(vcycle64(ws, s, l + 1, true, nullptr, nullptr, nullptr, nullptr, nullptr));
    else
        MF3D_TRY(vcycle<double>(ws, s, l + 1, C.x, C.b, true));
    return kern::smooth_split64(s, L.g, ws->bct, ws->pin, dx2, L.x, L.b, L.hs, ws->post, false,
                                C.x, C.order64, tb, nullptr, z, rdot, partials, np,
                                ws->march_blocks);
}

// PCG may leave z = x32 unstored (the (r, z) partials and the p update read the
// FP32 V-cycle result directly): FP32 V-cycle, colour-split finest level with
// nx % 8 == 0, and not flexible CG (which keeps z_old).
bool z_in_x32(const mf3d_workspace* ws, bool flexible) {
    const Level& L0 = ws->lv[0];
    return !flexible && ws->precision == MF3D_PRECOND_FP32 && ws->lv.size() > 1 &&
           L0.order == ORDER_SPLIT && !(L0.g.nx & 7);
}

// partials of a finest-level split4 sweep: one per workgroup
unsigned sweep_partials(const mf3d_workspace* ws) {
    const Level& L0 = ws->lv[0];
    return (unsigned)((L0.n / 8 + L0.plan.block - 1) / L0.plan.block);
}

// the FP64 V-cycle's colour-split finest level takes r directly in its first
// half-sweep (b from r), which can update r first
bool r_in_first_sweep64(const mf3d_workspace* ws) {
    return ws->precision == MF3D_PRECOND_FP64 && ws->lv.size() > 1 &&
           ws->lv[0].order64 == ORDER_SPLIT && ws->pre > 0;
}

// the FP32 V-cycle's colour-split finest level takes r directly in its first
// half-sweep
bool r_in_first_sweep(const mf3d_workspace* ws) {
    const Level& L0 = ws->lv[0];
    return ws->precision == MF3D_PRECOND_FP32 && ws->lv.size() > 1 &&
           L0.order == ORDER_SPLIT && !(L0.g.nx & 7) && ws->pre > 0 &&
           L0.plan.kind == SMOOTH_COLOUR;
}

// MultigridPreconditioner::apply (mg_preconditioner.cu:16-45): z = V-cycle(b = r, x = 0),
// in FP64 or with the FP32 V-cycle between two conversions.
// rz_partials (nullable): also leave the first-stage partials of (r, z) there.
// z == nullptr (z_in_x32 only): z stays in the finest level's x32.
// ru: r -= alpha Ap is done by that first half-sweep (r_in_first_sweep only)
// np_rz: the number of (r, z) partials written (z == nullptr: the last
// finest-level half-sweep writes one per workgroup when it can, else
// reduce_blocks(n))
int precondition(mf3d_workspace* ws, hipStream_t s, const double* r, double* z,
                 double* rz_partials = nullptr, const RUpd* ru = nullptr,
                 unsigned* np_rz = nullptr) {
    Level& L0 = ws->lv[0];
    ++ws->n_pc;
    if (!z && !z_in_x32(ws, false))
        return fail(MF3D_ERR_UNSUPPORTED, "z can stay in x32 only on a colour-split FP32 level");
    if (np_rz) *np_rz = kern::reduce_blocks(L0.n);
    if (ws->precision == MF3D_PRECOND_FP64) {
        if (L0.order64 == ORDER_SPLIT) {  // z and the (r, z) partials from the last sweep
            if (ru && !r_in_first_sweep64(ws))
                return fail(MF3D_ERR_UNSUPPORTED, "the r update needs a pre-smoothing half-sweep");
            unsigned np = 0;
            MF3D_TRY(vcycle64(ws, s, 0, true, r, z, rz_partials ? r : nullptr, rz_partials, &np,
                              ru));
            if (np_rz) *np_rz = np;
            return MF3D_OK;
        }
        if (ru) return fail(MF3D_ERR_UNSUPPORTED, "the r update rides on a colour-split sweep");
        MF3D_TRY(vcycle<double>(ws, s, 0, z, r, true));
        return rz_partials ? kern::dot_partials(s, r, z, L0.n, rz_partials) : MF3D_OK;
    }
    const bool fold = r_in_first_sweep(ws);
    if (ru && !fold) return fail(MF3D_ERR_UNSUPPORTED, "the r update needs the folded r sweep");
    if (!fold) MF3D_TRY(kern::to_f32(s, r, L0.b32, L0.g, L0.order));
    // the last finest half-sweep forms (r, z): z in x32, post-smoothing on level 0
    const bool rz_fold = !z && rz_partials && ws->fold_rz && ws->post > 0 && fold;
    const RzDot rzd{r, rz_partials};
    MF3D_TRY(vcycle<float>(ws, s, 0, L0.x32, L0.b32, true, fold ? r : nullptr, false, ru,
                           rz_fold ? &rzd : nullptr));
    if (rz_fold) {
        if (np_rz) *np_rz = sweep_partials(ws);
        return MF3D_OK;
    }
    if (!z) return rz_partials ? kern::rz_split8(s, L0.x32, L0.g, r, rz_partials) : MF3D_OK;
    return kern::to_f64(s, L0.x32, z, L0.g, L0.order, rz_partials ? r : nullptr, rz_partials);
}

// whether the finest-level operator runs on cell pairs for a solve into x
bool op_pairs_for(const mf3d_workspace* ws, const double* x) {
    const mf3d_grid& g = ws->lv[0].g;
    return kern::op_pairs_ok(g, x, ws->rhs, ws->r, ws->h0) &&
           kern::op_pairs_ok(g, ws->p, nullptr, ws->Ap, ws->h0);
}

// FP64 operator on the finest level (faces from h0); np: partials written
int apply_op(mf3d_workspace* ws, hipStream_t s, const double* x, double* y, double* partials,
             unsigned* np = nullptr) {
    Level& L0 = ws->lv[0];
    ++ws->n_op;
    return kern::op_apply(s, L0.g, ws->bct, ws->pin, ws->h0, 0, x, nullptr, y, partials, np,
                          ws->op_pairs, ws->op_tile, ws->op_chunk);
}
int residual_op(mf3d_workspace* ws, hipStream_t s, const double* x, const double* b, double* r,
                double* partials, unsigned* np = nullptr) {
    Level& L0 = ws->lv[0];
    ++ws->n_op;
    return kern::op_apply(s, L0.g, ws->bct, ws->pin, ws->h0, 1, x, b, r, partials, np,
                          ws->op_pairs, ws->op_tile, ws->op_chunk);
}

IterState make_state(const mf3d_head_solve_config& cfg, double rtol) {
    IterState st;
    std::memset(&st, 0, sizeof(st));
    st.rtol = rtol;
    st.check_every = std::max(1, cfg.cg_check_every);
    st.max_iter = std::max(0, cfg.cg_max_iter);
    st.stop_norm = cfg.stop_norm;
    st.flexible = cfg.flexible ? 1 : 0;
    st.den = 1.0;
    if (st.max_iter == 0) st.stop = 3;
    return st;
}

int read_state(mf3d_workspace* ws, hipStream_t s) {
    MF3D_HIP_TRY(hipMemcpyAsync(ws->host_st, ws->st, sizeof(IterState), hipMemcpyDeviceToHost, s));
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    return MF3D_OK;
}

// Chunk length before the next host check: with the reference stopping rule
// the device tests only every check_every iterations, so the host follows that
// cadence; otherwise the length follows the observed contraction rate.
struct Chunker {
    int fixed = 0;
    double prev_log = 0.0;
    int prev_iter = 0;
    bool have_prev = false;
    int next(const IterState& st, double target) {
        if (fixed > 0) return fixed - (st.iter % fixed);
        int c = 4;
        const double cur = 0.5 * std::log(std::max(st.rr, 1e-300));
        if (have_prev && st.iter > prev_iter) {
            const double rate = (cur - prev_log) / (st.iter - prev_iter);
            if (rate < -1e-3) {
                // iterations enqueued past convergence still run their
                // V-cycle (a whole iteration), while stopping short costs one
                // more host check; PCG tends to speed up as it goes, so the
                // extrapolated count is scaled down
                const double left = (std::log(target) - cur) / rate;
                c = (int)std::ceil(0.8 * left);
            }
        }
        prev_log = cur;
        prev_iter = st.iter;
        have_prev = true;
        return std::max(1, std::min(c, 32));
    }
};

constexpr size_t kGraphMaxCells = size_t(1) << 21;  // graph replay by default up to 128^3

struct GraphHolder {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphHolder() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

int state_upload(mf3d_workspace* ws, hipStream_t s, const IterState& st) {
    *ws->host_st = st;
    MF3D_HIP_TRY(hipMemcpyAsync(ws->st, ws->host_st, sizeof(IterState), hipMemcpyHostToDevice, s));
    // host_st is reused as the download buffer: finish the upload first
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    return MF3D_OK;
}

// pcg_solve (src/numerics/solvers/pcg.cuh:95-185) with the MG preconditioner.
// Extensions: flexible beta (r_new, z_new - z_old) / (r_old, z_old) for an
// inexact (FP32) preconditioner, and, once the recursive residual meets the
// tolerance, a true-residual check: if ||b - A x|| misses it, the iteration
// restarts from the true residual (at most 3 times).
int run_pcg(mf3d_workspace* ws, hipStream_t s, const mf3d_grid& g,
            const mf3d_head_solve_config& cfg, const double* K, double* x,
            mf3d_head_solve_result* res) {
    (void)K;
    const size_t n = cells(g);
    const unsigned np = kern::reduce_blocks(n);
    unsigned npo = 0;  // partials of the last operator launch
    const bool flexible = cfg.flexible != 0;
    const bool zx = z_in_x32(ws, flexible);  // z kept in the finest level's x32
    double* z = zx ? nullptr : ws->z;
    Level& L0 = ws->lv[0];
    // r update in the V-cycle's first sweep (its FIN_RR then follows the V-cycle):
    // the FP32 V-cycle with z in x32, or the FP64 one on a colour-split finest
    // level (z natural; flexible CG needs (r_new, z_old) before the V-cycle)
    const bool fuse64 = !flexible && ws->fold_r && r_in_first_sweep64(ws);
    const bool fuse_r = (zx && ws->fold_r && r_in_first_sweep(ws)) || fuse64;
    const unsigned np_sweep =
        fuse64 ? kern::split64_partials(L0.g, L0.plan.block, ws->march_blocks)
               : sweep_partials(ws);
    unsigned nprz = 0;  // (r, z) partials of the last preconditioner application
    // the reference's own rule (||r||/||r0||, FP64) keeps its exact semantics
    const bool true_check =
        cfg.stop_norm == MF3D_STOP_RHS_NORM || ws->precision == MF3D_PRECOND_FP32;
    IterState st0 = make_state(cfg, cfg.cg_rtol);
    MF3D_TRY(state_upload(ws, s, st0));
    MF3D_TRY(kern::dot_partials(s, ws->rhs, ws->rhs, n, ws->partials));
    MF3D_TRY(kern::finalize(s, ws->partials, np, FIN_BNORM, ws->st, nullptr));
    MF3D_TRY(residual_op(ws, s, x, ws->rhs, ws->r, ws->partials, &npo));
    MF3D_TRY(kern::finalize(s, ws->partials, npo, FIN_R0, ws->st, nullptr));
    // one iteration of the folded form: A p with (p, Ap), the V-cycle whose
    // first half-sweep updates r (with (r, r)) and whose last forms (r, z),
    // both finalisations, then x and p. Every scalar it needs lives in the
    // device state, so the sequence can be captured once and replayed.
    auto iteration = [&](hipStream_t q) -> int {
        unsigned npa = 0, nrz = 0;
        MF3D_TRY(apply_op(ws, q, ws->p, ws->Ap, ws->partials, &npa));
        MF3D_TRY(kern::finalize(q, ws->partials, npa, FIN_PAP, ws->st, nullptr));
        const RUpd ru{ws->st, ws->Ap, ws->r, ws->rr_partials};
        MF3D_TRY(precondition(ws, q, ws->r, z, ws->partials, &ru, &nrz));
        MF3D_TRY(kern::finalize2(q, ws->rr_partials, np_sweep, FIN_RR, ws->partials, nrz, FIN_RZ,
                                 ws->st));
        if (fuse64) return kern::pcg_update_xp(q, ws->st, ws->z, x, ws->p, n);
        return kern::update_xp_split8(q, ws->st, L0.x32, L0.g, x, ws->p);
    };
    // replaying the captured iteration saves host launch time, which bounds
    // small grids; from 128^3 up the device is the bound and plain launches
    // are as fast or faster (64^3: 164 vs 172 us per iteration, 256^3: 844 vs
    // 831 us on MI355X)
    const bool use_graph = ws->graph == 1 || (ws->graph == -1 && n <= kGraphMaxCells);
    GraphHolder G;
    long g_op = 0, g_pc = 0;  // applications in one replay of G
    int restarts = 0;
    for (;;) {
        MF3D_TRY(precondition(ws, s, ws->r, z, ws->partials, nullptr, &nprz));
        MF3D_TRY(kern::finalize(s, ws->partials, nprz, FIN_RZ0, ws->st, nullptr));
        if (fuse_r && use_graph && !G.x) {
            // captured on the workspace's private stream (the caller's may be
            // the null stream, which cannot capture), replayed on the caller's;
            // the coarsest inverse has joined in the V-cycle above
            if (!ws->capture_stream)
                MF3D_HIP_TRY(hipStreamCreateWithFlags(&ws->capture_stream, hipStreamNonBlocking));
            MF3D_HIP_TRY(hipStreamBeginCapture(ws->capture_stream, hipStreamCaptureModeThreadLocal));
            const long op0 = ws->n_op, pc0 = ws->n_pc;
            const int cst = iteration(ws->capture_stream);
            g_op = ws->n_op - op0;  // captured, not run
            g_pc = ws->n_pc - pc0;
            ws->n_op = op0;
            ws->n_pc = pc0;
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(ws->capture_stream, &gr);
            if (cst != MF3D_OK) {
                if (gr) (void)hipGraphDestroy(gr);
                return cst;
            }
            if (e != hipSuccess)
                return fail(MF3D_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            G.g = gr;
            MF3D_HIP_TRY(hipGraphInstantiate(&G.x, gr, nullptr, nullptr, 0));
        }
        if (zx) MF3D_TRY(kern::update_p_split8(s, ws->st, L0.x32, L0.g, ws->p, true));
        else MF3D_TRY(kern::copy(s, ws->z, ws->p, n));
        MF3D_TRY(read_state(ws, s));
        Chunker ch;
        if (cfg.stop_norm == MF3D_STOP_INITIAL_RESIDUAL && cfg.cg_check_every > 1)
            ch.fixed = cfg.cg_check_every;
        while (!ws->host_st->stop) {
            const double target = cfg.cg_rtol * ws->host_st->den;
            const int chunk = ch.next(*ws->host_st, target);
            for (int c = 0; c < chunk; ++c) {
                if (fuse_r) {  // the whole iteration, A p included
                    if (G.x) {
                        MF3D_HIP_TRY(hipGraphLaunch(G.x, s));
                        ws->n_op += g_op;
                        ws->n_pc += g_pc;
                    } else {
                        MF3D_TRY(iteration(s));
                    }
                    continue;
                }
                MF3D_TRY(apply_op(ws, s, ws->p, ws->Ap, ws->partials, &npo));
                MF3D_TRY(kern::finalize(s, ws->partials, npo, FIN_PAP, ws->st, nullptr));
                if (zx)  // x is updated together with p below
                    MF3D_TRY(kern::pcg_update_r(s, ws->st, ws->Ap, ws->r, n, ws->partials));
                else
                    MF3D_TRY(kern::pcg_update_xr(s, ws->st, ws->p, ws->Ap, x, ws->r, n,
                                                 ws->partials));
                MF3D_TRY(kern::finalize(s, ws->partials, np, FIN_RR, ws->st, nullptr));
                if (flexible) {  // (r_new, z_old) before z is overwritten
                    MF3D_TRY(kern::dot_partials(s, ws->r, ws->z, n, ws->partials));
                    MF3D_TRY(kern::finalize(s, ws->partials, np, FIN_RZO, ws->st, nullptr));
                }
                MF3D_TRY(precondition(ws, s, ws->r, z, ws->partials, nullptr, &nprz));
                MF3D_TRY(kern::finalize(s, ws->partials, nprz, FIN_RZ, ws->st, nullptr));
                if (zx) {
                    MF3D_TRY(kern::update_xp_split8(s, ws->st, L0.x32, L0.g, x, ws->p));
                } else {
                    MF3D_TRY(kern::pcg_update_p(s, ws->st, ws->z, ws->p, n));
                }
            }
            MF3D_TRY(read_state(ws, s));
        }
        // the x += alpha p of the iteration that stopped after its r update
        if (zx || fuse64) MF3D_TRY(kern::pcg_x_pending(s, ws->st, ws->p, x, n));
        if (ws->host_st->stop != 1 || restarts >= 3 || !true_check) break;
        // converged on the recursive residual: decide by the true one
        MF3D_TRY(residual_op(ws, s, x, ws->rhs, ws->r, ws->partials, &npo));
        MF3D_TRY(kern::finalize(s, ws->partials, npo, FIN_RESUME, ws->st, nullptr));
        MF3D_TRY(read_state(ws, s));
        if (ws->host_st->stop == 1) {  // the true residual meets the tolerance
            ws->true_rr_np = npo;
            break;
        }
        ++restarts;
    }
    const IterState& st = *ws->host_st;
    res->num_iterations = st.iter;
    res->converged = st.stop == 1;
    res->initial_residual = st.r0;
    res->final_residual = std::sqrt(st.rr);
    res->bnorm = st.bnorm;
    res->restarts = restarts;
    return st.stop == 2 ? fail(MF3D_ERR_BREAKDOWN, "PCG breakdown: (p, Ap) not finite") : MF3D_OK;
}

// cg_solve on NegatedOperator (src/numerics/solvers/cg.cuh:15-172,
// solve_head.cu:137-180).
int run_cg(mf3d_workspace* ws, hipStream_t s, const mf3d_grid& g,
           const mf3d_head_solve_config& cfg, const double* K, double* x,
           mf3d_head_solve_result* res) {
    const size_t n = cells(g);
    const unsigned np = kern::reduce_blocks(n);
    IterState st0 = make_state(cfg, cfg.cg_rtol);
    MF3D_TRY(state_upload(ws, s, st0));
    MF3D_TRY(kern::dot_partials(s, ws->rhs, ws->rhs, n, ws->partials));
    MF3D_TRY(kern::finalize(s, ws->partials, np, FIN_BNORM, ws->st, nullptr));
    double* bneg = ws->z;
    MF3D_TRY(kern::scale_copy(s, -1.0, ws->rhs, bneg, n));
    MF3D_TRY(apply_op(ws, s, x, ws->r, ws->partials));
    MF3D_TRY(kern::scal(s, ws->r, n, -1.0));
    MF3D_TRY(kern::axpby(s, 1.0, bneg, -1.0, ws->r, n));
    MF3D_TRY(kern::copy(s, ws->r, ws->p, n));
    MF3D_TRY(kern::dot_partials(s, ws->r, ws->r, n, ws->partials));
    MF3D_TRY(kern::finalize(s, ws->partials, np, FIN_CG_RR0, ws->st, nullptr));
    MF3D_TRY(read_state(ws, s));
    Chunker ch;
    if (cfg.stop_norm == MF3D_STOP_INITIAL_RESIDUAL && cfg.cg_check_every > 1)
        ch.fixed = cfg.cg_check_every;
    while (!ws->host_st->stop) {
        const int chunk = ch.next(*ws->host_st, cfg.cg_rtol * ws->host_st->den);
        for (int c = 0; c < chunk; ++c) {
            unsigned npa = 0;
            MF3D_TRY(apply_op(ws, s, ws->p, ws->Ap, ws->partials, &npa));
            MF3D_TRY(kern::finalize(s, ws->partials, npa, FIN_CG_PAP, ws->st, nullptr));
            MF3D_TRY(kern::cg_update_xr(s, ws->st, ws->p, ws->Ap, x, ws->r, n, ws->partials));
            MF3D_TRY(kern::finalize(s, ws->partials, np, FIN_CG_RR, ws->st, nullptr));
            MF3D_TRY(kern::cg_update_p(s, ws->st, ws->r, ws->p, n));
        }
        MF3D_TRY(read_state(ws, s));
    }
    const IterState& st = *ws->host_st;
    res->num_iterations = st.iter;
    res->converged = st.stop == 1;
    res->initial_residual = st.r0;
    res->final_residual = st.stop == 2 ? -1.0 : std::sqrt(st.rr);
    res->bnorm = st.bnorm;
    return MF3D_OK;
}

// host-synchronous ||v|| with the workspace reduction scratch
int host_nrm2(mf3d_workspace* ws, hipStream_t s, const double* v, size_t n, double* out) {
    MF3D_TRY(kern::dot_partials(s, v, v, n, ws->partials));
    MF3D_TRY(kern::finalize(s, ws->partials, kern::reduce_blocks(n), FIN_SQRT, nullptr,
                            ws->dscalar));
    MF3D_HIP_TRY(hipMemcpyAsync(out, ws->dscalar, sizeof(double), hipMemcpyDeviceToHost, s));
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    return MF3D_OK;
}

// mg_solve (src/multigrid/cycle/v_cycle.cu:76-116)
int run_mg(mf3d_workspace* ws, hipStream_t s, const mf3d_grid& g,
           const mf3d_head_solve_config& cfg, const double* K, double* x,
           mf3d_head_solve_result* res) {
    const size_t n = cells(g);
    double bnorm = 0.0, r0 = 0.0, rn = 0.0;
    MF3D_TRY(host_nrm2(ws, s, ws->rhs, n, &bnorm));
    MF3D_TRY(kern::residual(s, g, ws->bct, ws->pin, x, ws->rhs, K, ws->r));
    MF3D_TRY(host_nrm2(ws, s, ws->r, n, &r0));
    const double den = cfg.stop_norm == MF3D_STOP_RHS_NORM ? bnorm : r0;
    res->initial_residual = r0;
    res->bnorm = bnorm;
    Level& L0 = ws->lv[0];
    const bool split0 = L0.order64 == ORDER_SPLIT;
    if (split0) {  // the iterate lives in split order between the cycles
        MF3D_TRY(kern::to_split64(s, g, x, L0.x));
        MF3D_TRY(kern::to_split64(s, g, ws->rhs, L0.b));
    }
    for (int c = 0; c < cfg.mg_max_cycles; ++c) {
        res->num_iterations = c + 1;
        if (split0) {
            MF3D_TRY(vcycle64(ws, s, 0, false, nullptr, nullptr, nullptr, nullptr, nullptr));
            MF3D_TRY(kern::from_split64(s, g, L0.x, x, nullptr, nullptr, nullptr));
        } else {
            MF3D_TRY(vcycle<double>(ws, s, 0, x, ws->rhs, false));
        }
        MF3D_TRY(kern::residual(s, g, ws->bct, ws->pin, x, ws->rhs, K, ws->r));
        MF3D_TRY(host_nrm2(ws, s, ws->r, n, &rn));
        res->final_residual = rn;
        if (rn / den < cfg.rtol) {
            res->converged = 1;
            break;
        }
    }
    return MF3D_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.f;
    return ms;
}

// ---------------------------------------------------------------------------
// batched solves: many realisations (own workspaces, any grid sizes) advanced
// together, every kernel launch covering all of them; one PCG iteration is
// captured once into a hipGraph and replayed between host checks.
// ---------------------------------------------------------------------------
struct BLaunch {
    int m = 0;
    long nblocks = 0;
    void* desc = nullptr;  // device BatchLevel[m] or BatchVec[m]
    int* start = nullptr;  // device, m + 1 block offsets
    long snblocks = 0;     // colour-split sweeps: blocks of sweep_block threads
    int* sstart = nullptr;
    long onblocks = 0;  // FP64 operator: op_partials workgroups
    int* ostart = nullptr;
    long cnblocks = 0;  // FP32 -> FP64 conversion: 2048 cells per workgroup
    int* cstart = nullptr;
};

struct BDepth {
    BLaunch s4_cs, s4_cn;  // colour-split levels whose next level is split / natural
    BLaunch small, small_rr;  // small natural-order levels (non-coarsest)
    BLaunch direct;         // coarsest direct solves
};

struct BPlan {
    std::vector<void*> mem;
    BLaunch vec;  // finest level, blocks of reduce_blocks(n)
    std::vector<BDepth> depth;
    CoefType type = COEF_BF16;
    int pre = 2, post = 2;
    int tb = 256;  // sweep threads per workgroup
    size_t op_lds = 0;  // dynamic LDS of the batched operator (plane-tile entries)
    ~BPlan() {
        for (void* q : mem) (void)hipFree(q);
    }
};

int upload_starts(BPlan& P, const std::vector<long>& blocks, int** dst, long* total) {
    std::vector<int> start(blocks.size() + 1, 0);
    for (size_t e = 0; e < blocks.size(); ++e) {
        if (start[e] + blocks[e] > (long)INT32_MAX)
            return fail(MF3D_ERR_UNSUPPORTED, "batched launch larger than 2^31 blocks");
        start[e + 1] = start[e] + (int)blocks[e];
    }
    *total = start.back();
    void* q = nullptr;
    MF3D_HIP_TRY(hipMalloc(&q, start.size() * sizeof(int)));
    P.mem.push_back(q);
    MF3D_HIP_TRY(hipMemcpy(q, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
    *dst = static_cast<int*>(q);
    return MF3D_OK;
}

template <typename D>
int upload_launch(BPlan& P, BLaunch& L, const std::vector<D>& d, const std::vector<long>& blocks,
                  const std::vector<long>* sweep_blocks = nullptr) {
    L.m = (int)d.size();
    if (L.m == 0) return MF3D_OK;
    void* q = nullptr;
    MF3D_HIP_TRY(hipMalloc(&q, d.size() * sizeof(D)));
    P.mem.push_back(q);
    MF3D_HIP_TRY(hipMemcpy(q, d.data(), d.size() * sizeof(D), hipMemcpyHostToDevice));
    L.desc = q;
    MF3D_TRY(upload_starts(P, blocks, &L.start, &L.nblocks));
    if (sweep_blocks) MF3D_TRY(upload_starts(P, *sweep_blocks, &L.sstart, &L.snblocks));
    return MF3D_OK;
}

// Launch plan of a batch whose workspaces are set up (mixed precision, direct
// coarsest solve). Levels must be colour-split with nx % 8 == 0, small, or the
// coarsest.
int build_plan(BPlan& P, mf3d_workspace* const* ws, int count, double* const* h) {
    P.type = ws[0]->lo_type;
    P.pre = ws[0]->pre;
    P.post = ws[0]->post;
    P.tb = ws[0]->sweep_block;
    size_t maxl = 0;
    for (int r = 0; r < count; ++r) maxl = std::max(maxl, ws[r]->lv.size());
    P.depth.resize(maxl);
    std::vector<BatchVec> vd;
    std::vector<long> vb, ob, cb;
    struct G {
        std::vector<BatchLevel> d;
        std::vector<long> b, sb;  // blocks of 256 threads, sweep blocks
    };
    std::vector<G> s4c(maxl), s4n(maxl), sm(maxl), smr(maxl), dir(maxl);
    for (int r = 0; r < count; ++r) {
        mf3d_workspace* w = ws[r];
        Level& L0 = w->lv[0];
        if (L0.order != ORDER_SPLIT || (L0.g.nx & 7))
            return fail(MF3D_ERR_UNSUPPORTED,
                        "batched solve: the finest level must be colour-split with nx % 8 == 0");
        if (reinterpret_cast<uintptr_t>(h[r]) & 15)
            return fail(MF3D_ERR_UNSUPPORTED, "batched solve: heads must be 16-byte aligned");
        BatchVec v{};
        v.nx = L0.g.nx, v.ny = L0.g.ny, v.nz = L0.g.nz;
        v.bc = w->bct;
        v.pin = w->pin ? 1 : 0;
        v.inv_dx2 = 1.0 / (L0.g.dx * L0.g.dx);
        v.hk = w->h0;
        v.x = h[r];
        v.r = w->r, v.z = w->z, v.p = w->p, v.Ap = w->Ap, v.rhs = w->rhs;
        v.b32 = L0.b32, v.x32 = L0.x32;
        v.partials = w->partials;
        v.np = kern::reduce_blocks(L0.n);
        // the variant and geometry single solves use for the same vectors (op_apply)
        const kern::OpGeom og =
            kern::op_plan(L0.g, op_pairs_for(w, v.x), w->op_tile, w->op_chunk);
        v.np_op = (unsigned)og.tiles * (unsigned)og.chunks;
        v.otiles = og.tiles;
        v.okz = og.kz;
        v.opairs = og.tile ? 2 : og.pairs ? 1 : 0;
        if (og.tile) P.op_lds = std::max(P.op_lds, kern::op_tile_lds(L0.g));
        v.st = w->st;
        v.out = w->dscalar;
        vd.push_back(v);
        vb.push_back((long)kern::reduce_blocks(L0.n));
        ob.push_back((long)v.np_op);
        cb.push_back((long)((L0.n / 8 + 255) / 256));
        const size_t nl = w->lv.size();
        for (size_t l = 0; l < nl; ++l) {
            Level& L = w->lv[l];
            const double dx2 = L.g.dx * L.g.dx;
            BatchLevel b{};
            b.nx = L.g.nx, b.ny = L.g.ny, b.nz = L.g.nz;
            b.bc = w->bct;
            b.pin = w->pin ? 1 : 0;
            b.bs = (float)(L.cs * dx2);
            b.rs = (float)(1.0 / (L.cs * dx2));
            b.x = L.x32, b.b = L.b32;
            b.r64 = l == 0 ? w->r : nullptr;
            b.fx = L.flo.fx, b.fy = L.flo.fy, b.fz = L.flo.fz;
            b.ux = L.flo.ux, b.uy = L.flo.uy, b.uz = L.flo.uz;
            b.st = w->st;
            if (l + 1 == nl) {
                b.Ainv = w->coarse_inv;
                b.nc = (int)L.n;
                dir[l].d.push_back(b);
                dir[l].b.push_back((long)((L.n + 3) / 4));
                continue;
            }
            Level& C = w->lv[l + 1];
            b.bcoarse = C.b32;
            b.xc = C.x32;
            const long nb8 = kern::batch_blocks((long)(L.n / 8));
            if (L.order == ORDER_SPLIT) {
                if (L.g.nx & 7)
                    return fail(MF3D_ERR_UNSUPPORTED,
                                "batched solve: colour-split levels need nx % 8 == 0");
                G& grp = C.order == ORDER_SPLIT ? s4c[l] : s4n[l];
                grp.d.push_back(b);
                grp.b.push_back(nb8);
                grp.sb.push_back((long)((L.n / 8 + P.tb - 1) / P.tb));
            } else if (L.plan.kind == SMOOTH_SMALL) {
                sm[l].d.push_back(b);
                sm[l].b.push_back(1);
                smr[l].d.push_back(b);
                smr[l].b.push_back(nb8);
            } else {
                return fail(MF3D_ERR_UNSUPPORTED, "batched solve: level kind not batched");
            }
        }
    }
    MF3D_TRY(upload_launch(P, P.vec, vd, vb));
    MF3D_TRY(upload_starts(P, ob, &P.vec.ostart, &P.vec.onblocks));
    MF3D_TRY(upload_starts(P, cb, &P.vec.cstart, &P.vec.cnblocks));
    for (size_t l = 0; l < maxl; ++l) {
        MF3D_TRY(upload_launch(P, P.depth[l].s4_cs, s4c[l].d, s4c[l].b, &s4c[l].sb));
        MF3D_TRY(upload_launch(P, P.depth[l].s4_cn, s4n[l].d, s4n[l].b, &s4n[l].sb));
        MF3D_TRY(upload_launch(P, P.depth[l].small, sm[l].d, sm[l].b));
        MF3D_TRY(upload_launch(P, P.depth[l].small_rr, smr[l].d, smr[l].b));
        MF3D_TRY(upload_launch(P, P.depth[l].direct, dir[l].d, dir[l].b));
    }
    return MF3D_OK;
}

#define BL(L) static_cast<const BatchLevel*>((L).desc), (L).start, (L).m, (L).nblocks
#define BS(L) static_cast<const BatchLevel*>((L).desc), (L).sstart, (L).m, (L).snblocks
#define BV(L) static_cast<const BatchVec*>((L).desc), (L).start, (L).m, (L).nblocks
#define BO(L) static_cast<const BatchVec*>((L).desc), (L).ostart, (L).m, (L).onblocks
#define BC8(L) static_cast<const BatchVec*>((L).desc), (L).cstart, (L).m, (L).cnblocks

// V-cycle from x = 0 at depth d for every realisation that has that level
int vcycle_batch(const BPlan& P, hipStream_t s, size_t d) {
    const BDepth& D = P.depth[d];
    for (int q = 0; q < 2; ++q) {
        const BLaunch& L = q == 0 ? D.s4_cs : D.s4_cn;
        if (!L.m) continue;
        for (int it = 0; it < P.pre; ++it)
            for (int c = 0; c < 2; ++c)
                MF3D_TRY(kern::batch_sweep(s, P.type, (it == 0 && c == 0) ? (d == 0 ? 3 : 1) : 0,
                                           q == 0, BS(L), c, P.tb));
        MF3D_TRY(kern::batch_residual_restrict_split(s, P.type, q == 0, BL(L)));
    }
    if (D.small.m) {
        MF3D_TRY(kern::batch_small(s, P.type, static_cast<const BatchLevel*>(D.small.desc),
                                   D.small.m, P.pre, true, false));
        MF3D_TRY(kern::batch_residual_restrict(s, P.type, BL(D.small_rr)));
    }
    if (D.direct.m) {
        const BLaunch& L = D.direct;
        MF3D_TRY(kern::batch_gemv(s, static_cast<const BatchLevel*>(L.desc), L.start, L.m,
                                  L.nblocks));
    }
    if (d + 1 < P.depth.size()) MF3D_TRY(vcycle_batch(P, s, d + 1));
    for (int q = 0; q < 2; ++q) {
        const BLaunch& L = q == 0 ? D.s4_cs : D.s4_cn;
        if (!L.m) continue;
        for (int it = 0; it < P.post; ++it)
            for (int c = 0; c < 2; ++c)
                MF3D_TRY(kern::batch_sweep(s, P.type, (it == 0 && c == 0) ? 2 : 0, q == 0, BS(L), c,
                                           P.tb));
    }
    if (D.small.m)
        MF3D_TRY(kern::batch_small(s, P.type, static_cast<const BatchLevel*>(D.small.desc),
                                   D.small.m, P.post, false, true));
    return MF3D_OK;
}

// the finest levels' first half-sweep reads r (FP64) and fills b32
int precondition_batch(const BPlan& P, hipStream_t s) {
    MF3D_TRY(vcycle_batch(P, s, 0));
    return kern::batch_to_f64(s, BC8(P.vec));
}

// one PCG iteration for every running realisation (pcg.cuh:139-180)
int pcg_iteration_batch(const BPlan& P, hipStream_t s) {
    const BatchVec* V = static_cast<const BatchVec*>(P.vec.desc);
    const int m = P.vec.m;
    MF3D_TRY(kern::batch_op(s, 0, BO(P.vec), P.op_lds));
    MF3D_TRY(kern::batch_finalize(s, V, m, FIN_PAP, 1, true));
    MF3D_TRY(kern::batch_update_r(s, BV(P.vec)));  // x follows with p
    MF3D_TRY(kern::batch_finalize(s, V, m, FIN_RR, 1));
    MF3D_TRY(precondition_batch(P, s));  // also the (r, z) partials
    MF3D_TRY(kern::batch_finalize(s, V, m, FIN_RZ, 1));
    return kern::batch_update_xp_split8(s, BC8(P.vec));
}

int read_states(mf3d_workspace* const* ws, int count, hipStream_t s) {
    for (int r = 0; r < count; ++r)
        MF3D_HIP_TRY(hipMemcpyAsync(ws[r]->host_st, ws[r]->st, sizeof(IterState),
                                    hipMemcpyDeviceToHost, s));
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    return MF3D_OK;
}


int capture_iteration(const BPlan& P, hipStream_t s, GraphHolder& G) {
    MF3D_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = pcg_iteration_batch(P, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st != MF3D_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(MF3D_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                       hipGetErrorString(e));
    G.g = g;
    MF3D_HIP_TRY(hipGraphInstantiate(&G.x, g, nullptr, nullptr, 0));
    return MF3D_OK;
}

// PCG of run_pcg for all realisations (FP32 preconditioner: convergence is
// decided on the true residual, with up to 3 restarts). It returns after a
// true-residual pass over every realisation's final x, which leaves that
// residual's partial sums in each workspace.
int run_pcg_batch(mf3d_workspace* const* ws, int count, hipStream_t s,
                  const mf3d_head_solve_config& cfg, const BPlan& P, bool use_graph,
                  std::vector<int>& restarts) {
    const BatchVec* V = static_cast<const BatchVec*>(P.vec.desc);
    const int m = P.vec.m;
    const IterState st0 = make_state(cfg, cfg.cg_rtol);
    for (int r = 0; r < count; ++r) {
        *ws[r]->host_st = st0;
        MF3D_HIP_TRY(hipMemcpyAsync(ws[r]->st, ws[r]->host_st, sizeof(IterState),
                                    hipMemcpyHostToDevice, s));
    }
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    MF3D_TRY(kern::batch_dot(s, 1, BV(P.vec)));
    MF3D_TRY(kern::batch_finalize(s, V, m, FIN_BNORM, 0));
    MF3D_TRY(kern::batch_op(s, 1, BO(P.vec), P.op_lds));
    MF3D_TRY(kern::batch_finalize(s, V, m, FIN_R0, 0, true));
    GraphHolder G;
    if (use_graph) MF3D_TRY(capture_iteration(P, s, G));
    restarts.assign(count, 0);
    int cycles = 0;
    for (;;) {
        // start of a cycle for the running realisations: z = M r, rz, p = z
        MF3D_TRY(precondition_batch(P, s));  // also the (r, z) partials
        MF3D_TRY(kern::batch_finalize(s, V, m, FIN_RZ0, 1));
        MF3D_TRY(kern::batch_update_p_split8(s, 1, BC8(P.vec)));
        MF3D_TRY(read_states(ws, count, s));
        for (;;) {
            bool running = false;
            for (int r = 0; r < count; ++r) running = running || !ws[r]->host_st->stop;
            if (!running) break;
            for (int c = 0; c < 4; ++c) {
                if (use_graph) MF3D_HIP_TRY(hipGraphLaunch(G.x, s));
                else MF3D_TRY(pcg_iteration_batch(P, s));
            }
            MF3D_TRY(read_states(ws, count, s));
        }
        // the x += alpha p owed by the last iteration, then decide by the true residual
        MF3D_TRY(kern::batch_x_pending(s, BV(P.vec)));
        MF3D_TRY(kern::batch_op(s, 1, BO(P.vec), P.op_lds));
        MF3D_TRY(kern::batch_finalize(s, V, m, FIN_RESUME, 2, true));
        MF3D_TRY(read_states(ws, count, s));
        // a realisation whose true residual misses the tolerance resumes from
        // it (FIN_RESUME cleared its stop flag): that is its own restart
        bool again = false;
        for (int r = 0; r < count; ++r) again = again || !ws[r]->host_st->stop;
        if (!again || cycles >= 3) break;
        ++cycles;
        for (int r = 0; r < count; ++r)
            if (!ws[r]->host_st->stop) ++restarts[r];
    }
    return MF3D_OK;
}

#undef BL
#undef BS
#undef BV

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// ---- dot product without allocation -----------------------------------
// The first-stage partials of mf3d_dot_noalloc live in this file-scope device
// array (one copy per GPU, held until the process ends). Host code never
// passes the array's name to a kernel: its device address is looked up once
// per device with hipGetSymbolAddress and kept in noalloc_ptr.
namespace {
constexpr size_t kNoAllocPartials = size_t(1) << 16;  // 512 KiB of doubles
constexpr size_t kNoAllocMaxCells = kNoAllocPartials * kReduceCells;  // 2^27 = 512^3
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kNoAllocMaxDevices = 64;
std::mutex noalloc_mutex;
double* noalloc_ptr[kNoAllocMaxDevices] = {};

// device address of this GPU's scratch; the first call on a device resolves it
int noalloc_scratch(double** out) {
    int dev = 0;
    MF3D_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(MF3D_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MF3D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                          hipGetErrorString(e));
        noalloc_ptr[dev] = static_cast<double*>(p);
    }
    *out = noalloc_ptr[dev];
    return MF3D_OK;
}
}  // namespace

extern "C" {

int mf3d_abi_version(void) { return MF3D_ABI_VERSION; }

const char* mf3d_status_string(int status) {
    switch (status) {
    case MF3D_OK: return "ok";
    case MF3D_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MF3D_ERR_SIZE_MISMATCH: return "size mismatch";
    case MF3D_ERR_BREAKDOWN: return "breakdown";
    case MF3D_ERR_HIP: return "HIP runtime error";
    case MF3D_ERR_ALLOC: return "device allocation failed";
    case MF3D_ERR_SINGULAR: return "singular coarsest operator";
    case MF3D_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

const char* mf3d_last_error(void) { return mf3d::last_error_cstr(); }

void mf3d_default_config(mf3d_head_solve_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->solver_type = MF3D_SOLVER_PCG_MG;
    cfg->mg_levels = 4;
    cfg->mg_pre_smooth = 2;
    cfg->mg_post_smooth = 2;
    cfg->mg_max_cycles = 20;
    cfg->mg_coarse_iters = 50;
    cfg->cg_max_iter = 1000;
    cfg->cg_check_every = 10;
    cfg->cg_rtol = 1e-8;
    cfg->rtol = 1e-6;
    cfg->pin_mode = MF3D_PIN_AUTO;
}

int mf3d_workspace_create(mf3d_workspace** ws) {
    if (!ws) return fail(MF3D_ERR_INVALID_ARGUMENT, "ws is null");
    *ws = new (std::nothrow) mf3d_workspace();
    return *ws ? MF3D_OK : fail(MF3D_ERR_ALLOC, "host allocation failed");
}

int mf3d_workspace_destroy(mf3d_workspace* ws) {
    delete ws;
    return MF3D_OK;
}

size_t mf3d_workspace_bytes(const mf3d_workspace* ws) { return ws ? ws->bytes : 0; }

// An error left by an unrelated earlier HIP call (this thread's last-error
// slot) must not be reported as a failure of the launches below.
static void clear_stale_error() { (void)hipGetLastError(); }

#define MF3D_CHECK_PTR(p)                                                       \
    do {                                                                        \
        if (!(p)) return fail(MF3D_ERR_INVALID_ARGUMENT, #p " is null");        \
    } while (0)
#define MF3D_CHECK_LEN(len, need, what)                                         \
    do {                                                                        \
        if ((len) < (need))                                                     \
            return fail(MF3D_ERR_SIZE_MISMATCH, std::string(what) + " buffer too small"); \
    } while (0)

int mf3d_workspace_reserve(mf3d_workspace* ws, const mf3d_grid* grid, int mg_levels) {
    MF3D_CHECK_PTR(ws);
    MF3D_TRY(validate_grid(grid));
    return ensure_workspace(ws, *grid, mg_levels);
}

static int check_pair(const mf3d_grid* coarse, const mf3d_grid* fine) {
    MF3D_TRY(validate_grid(coarse));
    MF3D_TRY(validate_grid(fine));
    if (fine->nx != 2 * coarse->nx || fine->ny != 2 * coarse->ny || fine->nz != 2 * coarse->nz)
        return fail(MF3D_ERR_SIZE_MISMATCH, "fine grid must be exactly 2x the coarse grid");
    return MF3D_OK;
}

static int check_config(const mf3d_head_solve_config* cfg) {
    MF3D_CHECK_PTR(cfg);
    if (cfg->solver_type < 0 || cfg->solver_type > 2)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "unknown solver_type");
    if (cfg->mg_levels < 1 || cfg->mg_pre_smooth < 0 || cfg->mg_post_smooth < 0 ||
        cfg->mg_coarse_iters < 0 || cfg->cg_max_iter < 0 || cfg->mg_max_cycles < 0)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "negative iteration count or mg_levels < 1");
    if (cfg->precond_precision != MF3D_PRECOND_FP64 && cfg->precond_precision != MF3D_PRECOND_FP32)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "unknown precond_precision");
    if (cfg->precond_precision == MF3D_PRECOND_FP32 && cfg->solver_type != MF3D_SOLVER_PCG_MG)
        return fail(MF3D_ERR_UNSUPPORTED, "the FP32 V-cycle serves the PCG_MG solver only");
    if (cfg->coef_format < MF3D_COEF_FP32 || cfg->coef_format > MF3D_COEF_FP16)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "unknown coef_format");
    return MF3D_OK;
}

int mf3d_solve_head(mf3d_workspace* ws, mf3d_stream stream, const mf3d_grid* grid,
                    const mf3d_bc* bc, const mf3d_head_solve_config* cfg, const double* K,
                    double* h, size_t n, mf3d_head_solve_result* result) {
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(result);
    std::memset(result, 0, sizeof(*result));
    result->true_residual = -1.0;
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_TRY(check_config(cfg));
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(h);
    const size_t nc = cells(*grid);
    // solve_head.cu:85-87
    if (n < nc) return fail(MF3D_ERR_SIZE_MISMATCH, "Head or K buffer too small");
    clear_stale_error();
    hipStream_t s = as_stream(stream);
    MF3D_TRY(ensure_workspace(ws, *grid, cfg->mg_levels));
    ws->n_op = ws->n_pc = 0;
    MF3D_HIP_TRY(hipEventRecord(ws->ev[0], s));
    MF3D_TRY(mg_setup(ws, s, *grid, *bc, *cfg, K));
    MF3D_TRY(kern::rhs(s, *grid, *bc, K, ws->rhs));
    // the kernels move FP64 vectors with 16-byte accesses: a head that is not
    // 16-byte aligned (a view at an odd offset) is solved in a workspace copy
    double* const hu = h;
    if (reinterpret_cast<uintptr_t>(h) & 15) {
        if (!ws->hcopy || ws->hcopy_n < nc) {
            MF3D_TRY(ws->alloc(&ws->hcopy, nc));
            ws->hcopy_n = nc;
        }
        if (cfg->init_guess == MF3D_GUESS_PROVIDED) MF3D_TRY(kern::copy(s, hu, ws->hcopy, nc));
        h = ws->hcopy;
    }
    // solve_head.cu:123 overwrites h with the reference guess; the extension
    // keeps a caller-provided guess.
    if (cfg->init_guess != MF3D_GUESS_PROVIDED)
        MF3D_TRY(kern::guess(s, *grid, *bc, cfg->init_guess, h));
    MF3D_HIP_TRY(hipEventRecord(ws->ev[1], s));
    ws->op_pairs = op_pairs_for(ws, h);
    int st = MF3D_OK;
    ws->true_rr_np = 0;
    if (cfg->solver_type == MF3D_SOLVER_PCG_MG)
        st = run_pcg(ws, s, *grid, *cfg, K, h, result);
    else if (cfg->solver_type == MF3D_SOLVER_CG)
        st = run_cg(ws, s, *grid, *cfg, K, h, result);
    else
        st = run_mg(ws, s, *grid, *cfg, K, h, result);
    if (st != MF3D_OK) return st;
    MF3D_HIP_TRY(hipEventRecord(ws->ev[2], s));
    if (cfg->true_residual) {
        // the same partials, from the same operator pass over this h, when the
        // solve ended on a passing true-residual check
        unsigned npt = ws->true_rr_np;
        if (!npt) MF3D_TRY(residual_op(ws, s, h, ws->rhs, nullptr, ws->partials, &npt));
        MF3D_TRY(kern::finalize(s, ws->partials, npt, FIN_SQRT, nullptr, ws->dscalar));
        MF3D_HIP_TRY(hipMemcpyAsync(&result->true_residual, ws->dscalar, sizeof(double),
                                    hipMemcpyDeviceToHost, s));
    }
    if (h != hu) MF3D_TRY(kern::copy(s, h, hu, nc));
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    result->setup_ms = elapsed(ws->ev[0], ws->ev[1]);
    result->solve_ms = elapsed(ws->ev[1], ws->ev[2]);
    return MF3D_OK;
}

int mf3d_solve_head_batched(int count, mf3d_workspace* const* ws, mf3d_stream stream,
                            const mf3d_grid* grids, const mf3d_bc* bc,
                            const mf3d_head_solve_config* cfg, const double* const* K,
                            double* const* h, const size_t* n, int flags,
                            mf3d_head_solve_result* results) {
    if (count <= 0) return fail(MF3D_ERR_INVALID_ARGUMENT, "count must be positive");
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(grids);
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(h);
    MF3D_CHECK_PTR(n);
    MF3D_CHECK_PTR(results);
    MF3D_TRY(validate_bc(bc));
    MF3D_TRY(check_config(cfg));
    if (cfg->solver_type != MF3D_SOLVER_PCG_MG || cfg->precond_precision != MF3D_PRECOND_FP32 ||
        cfg->coarse_solver != MF3D_COARSE_DIRECT || cfg->flexible || cfg->mg_pre_smooth < 1 ||
        cfg->mg_post_smooth < 1)
        return fail(MF3D_ERR_UNSUPPORTED,
                    "batched solve: PCG_MG with the FP32 V-cycle, direct coarsest solve, "
                    "pre/post smoothing >= 1 and the standard beta");
    for (int r = 0; r < count; ++r) {
        MF3D_CHECK_PTR(ws[r]);
        MF3D_CHECK_PTR(K[r]);
        MF3D_CHECK_PTR(h[r]);
        MF3D_TRY(validate_grid(&grids[r]));
        if (n[r] < cells(grids[r])) return fail(MF3D_ERR_SIZE_MISMATCH, "Head or K buffer too small");
        for (int q = 0; q < r; ++q)
            if (ws[q] == ws[r]) return fail(MF3D_ERR_INVALID_ARGUMENT, "workspaces must differ");
        std::memset(&results[r], 0, sizeof(results[r]));
        results[r].true_residual = -1.0;
    }
    clear_stale_error();
    hipStream_t caller = as_stream(stream);
    // the iteration is captured into a graph, which needs a stream of its own
    // (not the null stream a caller may pass): kept by the first workspace
    // and reused by its later batches, together with the phase events
    mf3d_workspace* w0 = ws[0];
    if (!w0->batch_stream)
        MF3D_HIP_TRY(hipStreamCreateWithFlags(&w0->batch_stream, hipStreamNonBlocking));
    for (auto& e : w0->batch_ev)
        if (!e) MF3D_HIP_TRY(hipEventCreate(&e));
    hipStream_t s = w0->batch_stream;
    hipEvent_t* ev = w0->batch_ev;
    MF3D_HIP_TRY(hipEventRecord(ev[3], caller));
    MF3D_HIP_TRY(hipStreamWaitEvent(s, ev[3], 0));
    MF3D_HIP_TRY(hipEventRecord(ev[0], s));
    // host-side phase times on stderr when MF3D_BATCH_TIMING is set
    const bool timing = std::getenv("MF3D_BATCH_TIMING") != nullptr;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const auto t0 = now();
    for (int r = 0; r < count; ++r) {
        mf3d_workspace* w = ws[r];
        MF3D_TRY(ensure_workspace(w, grids[r], cfg->mg_levels));
        MF3D_TRY(mg_setup(w, s, grids[r], *bc, *cfg, K[r], true));
        MF3D_TRY(kern::rhs(s, grids[r], *bc, K[r], w->rhs));
        if (cfg->init_guess != MF3D_GUESS_PROVIDED)
            MF3D_TRY(kern::guess(s, grids[r], *bc, cfg->init_guess, h[r]));
    }
    struct Keep {
        std::vector<void*> host, dev;
        ~Keep() {
            for (void* q : host) (void)hipHostFree(q);
            for (void* q : dev) (void)hipFree(q);
        }
    } keep;
    const auto t1 = now();
    MF3D_TRY(build_coarse_inverses(ws, count, s, keep.host, keep.dev));
    const auto t2 = now();
    BPlan P;
    MF3D_TRY(build_plan(P, ws, count, h));
    const auto t3 = now();
    MF3D_HIP_TRY(hipEventRecord(ev[1], s));
    std::vector<int> restarts;
    MF3D_TRY(run_pcg_batch(ws, count, s, *cfg, P, !(flags & MF3D_BATCH_NO_GRAPH), restarts));
    if (timing) {
        MF3D_HIP_TRY(hipStreamSynchronize(s));
        std::fprintf(stderr,
                     "[batch] enqueue setup %.2f ms, coarse inverses %.2f ms, plan %.2f ms, "
                     "iterations %.2f ms (host clock)\n",
                     ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, now()));
    }
    MF3D_HIP_TRY(hipEventRecord(ev[2], s));
    if (cfg->true_residual) {
        // run_pcg_batch ends on a true-residual pass over every realisation's
        // final x, whose partial sums are still in place: only the second
        // stage runs again
        const BatchVec* V = static_cast<const BatchVec*>(P.vec.desc);
        MF3D_TRY(kern::batch_finalize(s, V, P.vec.m, FIN_SQRT, 0, true));
        for (int r = 0; r < count; ++r)
            MF3D_HIP_TRY(hipMemcpyAsync(&results[r].true_residual, ws[r]->dscalar, sizeof(double),
                                        hipMemcpyDeviceToHost, s));
    }
    MF3D_HIP_TRY(hipStreamSynchronize(s));
    for (int r = 0; r < count; ++r)
        if (const int bad = *ws[r]->coarse_flag)
            return fail(MF3D_ERR_SINGULAR, "realisation " + std::to_string(r) +
                                               ": coarsest operator is not definite at pivot " +
                                               std::to_string(bad - 1));
    const float setup = elapsed(ev[0], ev[1]), solve = elapsed(ev[1], ev[2]);
    int status = MF3D_OK;
    for (int r = 0; r < count; ++r) {
        const IterState& st = *ws[r]->host_st;
        results[r].num_iterations = st.iter;
        results[r].converged = st.stop == 1;
        results[r].initial_residual = st.r0;
        results[r].final_residual = std::sqrt(st.rr);
        results[r].bnorm = st.bnorm;
        results[r].restarts = restarts[r];
        results[r].setup_ms = setup;
        results[r].solve_ms = solve;
        if (st.stop == 2) status = fail(MF3D_ERR_BREAKDOWN, "PCG breakdown: (p, Ap) not finite");
    }
    MF3D_HIP_TRY(hipEventRecord(ev[3], s));
    MF3D_HIP_TRY(hipStreamWaitEvent(caller, ev[3], 0));
    return status;
}

int mf3d_init_head_guess(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc, double* h,
                         size_t n) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(h);
    MF3D_CHECK_LEN(n, cells(*grid), "head");
    clear_stale_error();
    return kern::guess(as_stream(stream), *grid, *bc, MF3D_GUESS_REFERENCE, h);
}

int mf3d_build_rhs_head(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc,
                        const double* K, double* rhs, size_t n) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(rhs);
    // rhs_head.cu:124-126
    if (n < cells(*grid)) return fail(MF3D_ERR_SIZE_MISMATCH, "RHS or K buffer too small");
    clear_stale_error();
    return kern::rhs(as_stream(stream), *grid, *bc, K, rhs);
}

int mf3d_coarsen_K(mf3d_stream stream, const mf3d_grid* coarse, const mf3d_grid* fine,
                   const double* K_fine, size_t n_fine, double* K_coarse, size_t n_coarse) {
    MF3D_TRY(check_pair(coarse, fine));
    MF3D_CHECK_PTR(K_fine);
    MF3D_CHECK_PTR(K_coarse);
    MF3D_CHECK_LEN(n_fine, cells(*fine), "fine K");
    MF3D_CHECK_LEN(n_coarse, cells(*coarse), "coarse K");
    clear_stale_error();
    return kern::coarsen(as_stream(stream), *coarse, K_fine, K_coarse);
}

int mf3d_velocity_compact(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc,
                          const double* h, const double* K, size_t n, double* U, size_t nU,
                          double* V, size_t nV, double* W, size_t nW) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(h);
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(U);
    MF3D_CHECK_PTR(V);
    MF3D_CHECK_PTR(W);
    const mf3d_grid& g = *grid;
    MF3D_CHECK_LEN(n, cells(g), "head/K");
    MF3D_CHECK_LEN(nU, (size_t)(g.nx + 1) * g.ny * g.nz, "U");
    MF3D_CHECK_LEN(nV, (size_t)g.nx * (g.ny + 1) * g.nz, "V");
    MF3D_CHECK_LEN(nW, (size_t)g.nx * g.ny * (g.nz + 1), "W");
    clear_stale_error();
    return kern::velocity(as_stream(stream), g, *bc, h, K, U, V, W, false);
}

int mf3d_velocity_padded(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc,
                         const double* h, const double* K, size_t n, double* U, double* V,
                         double* W, size_t n_padded) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(h);
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(U);
    MF3D_CHECK_PTR(V);
    MF3D_CHECK_PTR(W);
    const mf3d_grid& g = *grid;
    MF3D_CHECK_LEN(n, cells(g), "head/K");
    MF3D_CHECK_LEN(n_padded, (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1), "padded velocity");
    clear_stale_error();
    return kern::velocity(as_stream(stream), g, *bc, h, K, U, V, W, true);
}

int mf3d_velocity_diagnostics(mf3d_workspace* ws, mf3d_stream stream, const mf3d_grid* grid,
                              const double* U, const double* V, const double* W,
                              size_t n_padded, double* div, double* vort_mag, double* helicity,
                              size_t n, double* d_stats) {
    MF3D_TRY(validate_grid(grid));
    MF3D_CHECK_PTR(U);
    MF3D_CHECK_PTR(V);
    MF3D_CHECK_PTR(W);
    MF3D_CHECK_PTR(div);
    MF3D_CHECK_PTR(vort_mag);
    MF3D_CHECK_PTR(helicity);
    const mf3d_grid& g = *grid;
    MF3D_CHECK_LEN(n_padded, (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1), "padded velocity");
    MF3D_CHECK_LEN(n, cells(g), "diagnostic field");
    if (d_stats) {
        MF3D_CHECK_PTR(ws);
        MF3D_TRY(scratch_count(ws, kern::diag_partials(g)));
    }
    clear_stale_error();
    return kern::velocity_diagnostics(as_stream(stream), g, U, V, W, div, vort_mag, helicity,
                                      d_stats ? ws->partials : nullptr, d_stats);
}

int mf3d_field_stats(mf3d_workspace* ws, mf3d_stream stream, const double* x, size_t n,
                     double* d_out) {
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(d_out);
    clear_stale_error();
    hipStream_t s = as_stream(stream);
    if (n == 0) return kern::fill(s, d_out, 3, 0.0);  // stochastic.cu:372-375
    MF3D_CHECK_PTR(x);
    MF3D_TRY(scratch_count(ws, kern::stats_partials(n)));
    return kern::field_stats(s, x, n, ws->partials, d_out);
}

int mf3d_gsrb_smooth_3d(mf3d_stream stream, const mf3d_grid* grid, double* x, const double* b,
                        const double* K, size_t n, int num_iters, const mf3d_bc* bc, int pin) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(x);
    MF3D_CHECK_PTR(b);
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_LEN(n, cells(*grid), "x/b/K");
    if (num_iters < 0) return fail(MF3D_ERR_INVALID_ARGUMENT, "num_iters < 0");
    clear_stale_error();
    return kern::gsrb(as_stream(stream), *grid, bc_types(*bc), pin != 0, x, b, K, num_iters,
                      false);
}

int mf3d_compute_residual_3d(mf3d_stream stream, const mf3d_grid* grid, const double* x,
                             const double* b, const double* K, double* r, size_t n,
                             const mf3d_bc* bc, int pin) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(x);
    MF3D_CHECK_PTR(b);
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(r);
    MF3D_CHECK_LEN(n, cells(*grid), "x/b/K/r");
    clear_stale_error();
    return kern::residual(as_stream(stream), *grid, bc_types(*bc), pin != 0, x, b, K, r);
}

int mf3d_restrict_3d(mf3d_stream stream, const mf3d_grid* fine, const mf3d_grid* coarse,
                     const double* r_fine, size_t n_fine, double* b_coarse, size_t n_coarse) {
    MF3D_TRY(check_pair(coarse, fine));
    MF3D_CHECK_PTR(r_fine);
    MF3D_CHECK_PTR(b_coarse);
    MF3D_CHECK_LEN(n_fine, cells(*fine), "fine");
    MF3D_CHECK_LEN(n_coarse, cells(*coarse), "coarse");
    clear_stale_error();
    return kern::restrict_avg(as_stream(stream), *coarse, r_fine, b_coarse);
}

int mf3d_prolong_3d_add(mf3d_stream stream, const mf3d_grid* coarse, const mf3d_grid* fine,
                        const double* x_coarse, size_t n_coarse, double* x_fine, size_t n_fine) {
    MF3D_TRY(check_pair(coarse, fine));
    MF3D_CHECK_PTR(x_coarse);
    MF3D_CHECK_PTR(x_fine);
    MF3D_CHECK_LEN(n_fine, cells(*fine), "fine");
    MF3D_CHECK_LEN(n_coarse, cells(*coarse), "coarse");
    clear_stale_error();
    return kern::prolong_add(as_stream(stream), *fine, x_coarse, x_fine);
}

int mf3d_mg_setup(mf3d_workspace* ws, mf3d_stream stream, const mf3d_grid* grid,
                  const mf3d_bc* bc, const mf3d_head_solve_config* cfg, const double* K,
                  size_t n) {
    MF3D_CHECK_PTR(ws);
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_TRY(check_config(cfg));
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_LEN(n, cells(*grid), "K");
    clear_stale_error();
    return mg_setup(ws, as_stream(stream), *grid, *bc, *cfg, K);
}

int mf3d_workspace_set_tuning(mf3d_workspace* ws, int key, long value) {
    MF3D_CHECK_PTR(ws);
    switch (key) {
    case MF3D_TUNE_SPLIT: ws->split = value != 0; break;
    case MF3D_TUNE_FOLD_FIRST: ws->fold_first = value != 0; break;
    case MF3D_TUNE_FOLD_R: ws->fold_r = value != 0; break;
    case MF3D_TUNE_FOLD_RZ: ws->fold_rz = value != 0; break;
    case MF3D_TUNE_OP_TILE: ws->op_tile = value != 0; break;
    case MF3D_TUNE_OP_CHUNK:
        if (value < 1 || value > kern::kOpChunkMax)
            return fail(MF3D_ERR_INVALID_ARGUMENT, "operator chunk depth must be 1..64 planes");
        ws->op_chunk = (int)value;
        break;
    case MF3D_TUNE_MARCH_BLOCKS:
        if (value < kern::kMarchMinBlocks || value > kern::kMarchMaxBlocks)
            return fail(MF3D_ERR_INVALID_ARGUMENT, "march workgroup target must be 256..8192");
        ws->march_blocks = value;
        break;
    case MF3D_TUNE_GRAPH:
        if (value < -1 || value > 1) return fail(MF3D_ERR_INVALID_ARGUMENT, "graph must be -1, 0 or 1");
        ws->graph = (int)value;
        break;
    case MF3D_TUNE_SWEEP_BLOCK:
        if (value < 64 || value > 512 || (value & 63))
            return fail(MF3D_ERR_INVALID_ARGUMENT, "sweep block must be 64..512, a multiple of 64");
        ws->sweep_block = (int)value;
        break;
    default: return fail(MF3D_ERR_INVALID_ARGUMENT, "unknown tuning key");
    }
    return MF3D_OK;
}

int mf3d_workspace_solve_counts(const mf3d_workspace* ws, long* op_applications,
                                long* precond_applications) {
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(op_applications);
    MF3D_CHECK_PTR(precond_applications);
    *op_applications = ws->n_op;
    *precond_applications = ws->n_pc;
    return MF3D_OK;
}

int mf3d_mg_smooth(mf3d_workspace* ws, mf3d_stream stream, int level, int num_iters) {
    MF3D_CHECK_PTR(ws);
    if (!ws->ready || !ws->K0)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "mf3d_mg_setup has not been called");
    if (level < 0 || level >= (int)ws->lv.size())
        return fail(MF3D_ERR_INVALID_ARGUMENT, "level out of range");
    if (num_iters < 0) return fail(MF3D_ERR_INVALID_ARGUMENT, "num_iters < 0");
    clear_stale_error();
    hipStream_t s = as_stream(stream);
    Level& L = ws->lv[level];
    const double dx2 = L.g.dx * L.g.dx;
    if (ws->precision == MF3D_PRECOND_FP32)
        return kern::smooth_faces(s, L.g, ws->bct, ws->pin, L.cs * dx2, L.x32, L.b32, L.flo,
                                  num_iters, false, (const float*)nullptr, L.plan, L.order);
    if (L.order64 == ORDER_SPLIT)  // in place on the level's split x and b
        return kern::smooth_split64(s, L.g, ws->bct, ws->pin, dx2, L.x, L.b, L.hs, num_iters,
                                    false, nullptr, ORDER_NATURAL, L.plan.block, nullptr,
                                    nullptr, nullptr, nullptr, nullptr, ws->march_blocks);
    double* x = level == 0 ? ws->z : L.x;
    const double* b = level == 0 ? ws->r : L.b;
    return kern::smooth_faces(s, L.g, ws->bct, ws->pin, dx2, x, b, L.f64, num_iters, false,
                              (const double*)nullptr, L.plan);
}

int mf3d_mg_num_levels(const mf3d_workspace* ws) {
    return ws && ws->ready ? (int)ws->lv.size() : 0;
}

int mf3d_mg_precondition(mf3d_workspace* ws, mf3d_stream stream, const double* r, double* z,
                         size_t n) {
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(r);
    MF3D_CHECK_PTR(z);
    if (!ws->ready || !ws->K0)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "mf3d_mg_setup has not been called");
    MF3D_CHECK_LEN(n, cells(ws->fine), "r/z");
    // the FP32 V-cycle reads r and writes z with 16-byte accesses
    if ((reinterpret_cast<uintptr_t>(r) | reinterpret_cast<uintptr_t>(z)) & 15)
        return fail(MF3D_ERR_INVALID_ARGUMENT, "r and z must be 16-byte aligned");
    clear_stale_error();
    return precondition(ws, as_stream(stream), r, z);
}

int mf3d_varcoeff_apply(mf3d_stream stream, const mf3d_grid* grid, const mf3d_bc* bc, int pin,
                        const double* K, const double* x, double* y, size_t n) {
    MF3D_TRY(validate_grid(grid));
    MF3D_TRY(validate_bc(bc));
    MF3D_CHECK_PTR(K);
    MF3D_CHECK_PTR(x);
    MF3D_CHECK_PTR(y);
    MF3D_CHECK_LEN(n, cells(*grid), "x/y/K");
    clear_stale_error();
    return kern::apply(as_stream(stream), *grid, bc_types(*bc), pin != 0, K, x, y);
}

int mf3d_dot(mf3d_workspace* ws, mf3d_stream stream, const double* x, const double* y, size_t n,
             double* d_result) {
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(d_result);
    clear_stale_error();
    hipStream_t s = as_stream(stream);
    if (n == 0) return kern::fill(s, d_result, 1, 0.0);  // dot.cu:13-18
    MF3D_CHECK_PTR(x);
    MF3D_CHECK_PTR(y);
    MF3D_TRY(scratch_partials(ws, n));
    MF3D_TRY(kern::dot_partials(s, x, y, n, ws->partials));
    return kern::finalize(s, ws->partials, kern::reduce_blocks(n), FIN_STORE, nullptr, d_result);
}

int mf3d_dot_noalloc(mf3d_stream stream, const double* x, const double* y, size_t n,
                     double* d_result) {
    MF3D_CHECK_PTR(d_result);
    if (n > kNoAllocMaxCells)
        return fail(MF3D_ERR_INVALID_ARGUMENT,
                    "mf3d_dot_noalloc: n = " + std::to_string(n) + " exceeds the maximum " +
                        std::to_string(kNoAllocMaxCells) + " (use mf3d_dot)");
    clear_stale_error();
    hipStream_t s = as_stream(stream);
    if (n == 0) return kern::fill(s, d_result, 1, 0.0);  // dot.cu:13-18
    MF3D_CHECK_PTR(x);
    MF3D_CHECK_PTR(y);
    double* part = nullptr;
    MF3D_TRY(noalloc_scratch(&part));
    // every partial the finalisation reads is written by this call's first stage
    MF3D_TRY(kern::dot_partials(s, x, y, n, part));
    return kern::finalize(s, part, kern::reduce_blocks(n), FIN_STORE, nullptr, d_result);
}

int mf3d_nrm2(mf3d_workspace* ws, mf3d_stream stream, const double* x, size_t n,
              double* d_result) {
    MF3D_CHECK_PTR(ws);
    MF3D_CHECK_PTR(d_result);
    clear_stale_error();
    hipStream_t s = as_stream(stream);
    if (n == 0) return kern::fill(s, d_result, 1, 0.0);
    MF3D_CHECK_PTR(x);
    MF3D_TRY(scratch_partials(ws, n));
    MF3D_TRY(kern::dot_partials(s, x, x, n, ws->partials));
    return kern::finalize(s, ws->partials, kern::reduce_blocks(n), FIN_SQRT, nullptr, d_result);
}

int mf3d_axpy(mf3d_stream stream, double a, const double* x, double* y, size_t n) {
    if (n && (!x || !y)) return fail(MF3D_ERR_INVALID_ARGUMENT, "null vector");
    clear_stale_error();
    return kern::axpy(as_stream(stream), a, x, y, n);
}
int mf3d_axpby(mf3d_stream stream, double a, const double* x, double b, double* y, size_t n) {
    if (n && (!x || !y)) return fail(MF3D_ERR_INVALID_ARGUMENT, "null vector");
    clear_stale_error();
    return kern::axpby(as_stream(stream), a, x, b, y, n);
}
int mf3d_copy(mf3d_stream stream, const double* x, double* y, size_t n) {
    if (n && (!x || !y)) return fail(MF3D_ERR_INVALID_ARGUMENT, "null vector");
    clear_stale_error();
    return kern::copy(as_stream(stream), x, y, n);
}
int mf3d_fill(mf3d_stream stream, double* x, size_t n, double value) {
    if (n && !x) return fail(MF3D_ERR_INVALID_ARGUMENT, "null vector");
    clear_stale_error();
    return kern::fill(as_stream(stream), x, n, value);
}
int mf3d_scal(mf3d_stream stream, double* x, size_t n, double a) {
    if (n && !x) return fail(MF3D_ERR_INVALID_ARGUMENT, "null vector");
    clear_stale_error();
    return kern::scal(as_stream(stream), x, n, a);
}

}  // extern "C"
