// Face conductances from h = 1 / (2 K) (device helpers shared by the FP64
// operator and the FP64 colour-split V-cycle).
//
// With h per cell the harmonic face is f = 2 Ka Kb / (Ka + Kb) = 1 / (ha + hb),
// a Dirichlet face (2 K_C; the ghost value lives in b) is 1 / h_C, the same
// form with h_nb = 0, and a Neumann face is 0. Periodic faces pair with the
// wrapped cell.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace mf3d {

// 1/s by the hardware reciprocal estimate refined with two Newton-Raphson
// steps r <- r + r (1 - s r) (each step squares the relative error); for the
// positive, finite sums here the result is within a few ulp of 1/s.
__device__ __forceinline__ double rcp_nr(double s) {
    double r = __builtin_amdgcn_rcp(s);
    double e = fma(-s, r, 1.0);
    r = fma(r, e, r);
    e = fma(-s, r, 1.0);
    return fma(r, e, r);
}

// Boundary of one side of a cell: the neighbour's offset (0 without one), and
// whether the neighbour exists (in the domain or periodic) and the face does
// (neighbour or Dirichlet).
struct Side {
    long off;
    bool nb, face;
};
__device__ __forceinline__ Side side(bool inside, long in_off, int t, long wrap_off) {
    Side s;
    s.nb = inside || t == MF3D_BC_PERIODIC;
    s.face = s.nb || t == MF3D_BC_DIRICHLET;
    s.off = inside ? in_off : (t == MF3D_BC_PERIODIC ? wrap_off : 0);
    return s;
}
// face coefficient between a cell (hc) and a side (hn, loaded at the side's
// offset; ignored without a neighbour)
__device__ __forceinline__ double hface(const Side& s, double hc, double hn) {
    return s.face ? rcp_nr(hc + (s.nb ? hn : 0.0)) : 0.0;
}

}  // namespace mf3d
