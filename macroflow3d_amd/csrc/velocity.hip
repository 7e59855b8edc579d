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

}  // namespace

namespace kern {

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
