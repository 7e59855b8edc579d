// Fine-level z-marching smoother (placeholder: generic kernels serve all levels).
#include "kernels.hpp"

namespace mf3d {
namespace kern {

bool fast_ok(const mf3d_grid&, const BcTypes&) { return false; }

int gsrb_fast(hipStream_t s, const mf3d_grid& g, const BcTypes& bc, bool pin, double* x,
              const double* b, const double* K, int iters, bool zero_init) {
    return gsrb(s, g, bc, pin, x, b, K, iters, zero_init);
}

}  // namespace kern
}  // namespace mf3d
