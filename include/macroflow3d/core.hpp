// C++ host surface of libmf3d_hip: core types.
//
// Same namespace, names and member meanings as MacroFlow3D's src/core and
// src/runtime types, so a caller written against them compiles unchanged:
//   Grid3D        src/core/Grid3D.hpp:32-55
//   BCType/BCFace/BCSpec   src/core/BCSpec.hpp:13-87
//   DeviceSpan    src/core/DeviceSpan.cuh:12-36   (non-owning pointer + size)
//   DeviceBuffer  src/core/DeviceBuffer.cuh:16-108 (owning, move-only)
//   CudaContext   src/runtime/CudaContext.cuh:8-33 (here a HIP stream on one
//                 device; the name is an alias of HipContext)
// Errors follow src/runtime/cuda_check.cuh:12-18: a failed call throws
// std::runtime_error. All device work goes to the context's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>

#include "../mf3d.h"

namespace macroflow3d {

using real = double;

// ---- error plumbing ---------------------------------------------------------
inline void mf3d_check(int status, const char* what) {
    if (status != MF3D_OK)
        throw std::runtime_error(std::string(what) + ": " + mf3d_status_string(status) + " (" +
                                 mf3d_last_error() + ")");
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ---- grid -------------------------------------------------------------------
struct Grid3D {
    int nx = 0, ny = 0, nz = 0;
    real dx = 0.0, dy = 0.0, dz = 0.0;

    Grid3D() = default;
    Grid3D(int nx_, int ny_, int nz_, real dx_, real dy_, real dz_)
        : nx(nx_), ny(ny_), nz(nz_), dx(dx_), dy(dy_), dz(dz_) {}

    size_t num_cells() const { return (size_t)nx * (size_t)ny * (size_t)nz; }
    real Lx() const { return nx * dx; }
    real Ly() const { return ny * dy; }
    real Lz() const { return nz * dz; }
    // cell (i,j,k) at i + nx (j + ny k)
    size_t idx(int i, int j, int k) const {
        return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
    }
};

inline mf3d_grid to_c(const Grid3D& g) { return mf3d_grid{g.nx, g.ny, g.nz, g.dx, g.dy, g.dz}; }

// ---- boundary conditions ------------------------------------------------------
enum class BCType { Dirichlet, Neumann, Periodic };

struct BCFace {
    BCType type = BCType::Dirichlet;
    real value = 0.0;
    BCFace() = default;
    BCFace(BCType t, real v) : type(t), value(v) {}
};

struct BCSpec {
    BCFace xmin, xmax, ymin, ymax, zmin, zmax;  // default: Dirichlet 0 everywhere

    bool is_periodic_x() const { return both(xmin, xmax, BCType::Periodic); }
    bool is_periodic_y() const { return both(ymin, ymax, BCType::Periodic); }
    bool is_periodic_z() const { return both(zmin, zmax, BCType::Periodic); }
    bool is_all_homog_neumann() const {
        for (const BCFace* f : faces())
            if (f->type != BCType::Neumann || f->value != 0.0) return false;
        return true;
    }
    // periodic faces come in pairs; values are finite
    void validate() const {
        const char* axis[3] = {"x", "y", "z"};
        const BCFace* const* fs = faces().data();
        for (int a = 0; a < 3; ++a) {
            const bool lo = fs[2 * a]->type == BCType::Periodic;
            const bool hi = fs[2 * a + 1]->type == BCType::Periodic;
            if (lo != hi)
                throw std::runtime_error(std::string("Periodic BC must be specified on both ") +
                                         axis[a] + "min and " + axis[a] + "max");
        }
        for (const BCFace* f : faces())
            if (!std::isfinite(f->value)) throw std::runtime_error("BC values must be finite");
    }

    struct FaceList {
        const BCFace* f[6];
        const BCFace* const* begin() const { return f; }
        const BCFace* const* end() const { return f + 6; }
        const BCFace* const* data() const { return f; }
    };
    FaceList faces() const { return FaceList{{&xmin, &xmax, &ymin, &ymax, &zmin, &zmax}}; }

  private:
    static bool both(const BCFace& a, const BCFace& b, BCType t) {
        return a.type == t && b.type == t;
    }
};

inline mf3d_bc to_c(const BCSpec& bc) {
    mf3d_bc o{};
    int q = 0;
    for (const BCFace* f : bc.faces()) {
        o.type[q] = f->type == BCType::Dirichlet ? MF3D_BC_DIRICHLET
                    : f->type == BCType::Neumann ? MF3D_BC_NEUMANN
                                                 : MF3D_BC_PERIODIC;
        o.value[q] = f->value;
        ++q;
    }
    return o;
}

// Darcy set-up of the pipeline: h = h_west at x = 0, h = h_east at x = L,
// no-flow on the y and z faces.
inline BCSpec darcy_x(real h_west = 1.0, real h_east = 0.0) {
    BCSpec bc;
    bc.xmin = BCFace(BCType::Dirichlet, h_west);
    bc.xmax = BCFace(BCType::Dirichlet, h_east);
    bc.ymin = bc.ymax = bc.zmin = bc.zmax = BCFace(BCType::Neumann, 0.0);
    return bc;
}

// ---- device memory ----------------------------------------------------------
template <typename T>
class DeviceSpan {
  public:
    DeviceSpan() = default;
    DeviceSpan(T* p, size_t n) : ptr_(p), n_(n) {}
    template <typename U, typename = decltype(static_cast<T*>(std::declval<U*>()))>
    DeviceSpan(const DeviceSpan<U>& o) : ptr_(o.data()), n_(o.size()) {}

    T* data() const { return ptr_; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    DeviceSpan subspan(size_t off, size_t n) const { return DeviceSpan(ptr_ + off, n); }

  private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

template <typename T>
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    ~DeviceBuffer() { reset(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            reset();
            swap(o);
        }
        return *this;
    }

    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    size_t size() const { return n_; }
    DeviceSpan<T> span() { return DeviceSpan<T>(ptr_, n_); }
    DeviceSpan<const T> span() const { return DeviceSpan<const T>(ptr_, n_); }

    void reset() noexcept {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        n_ = cap_ = 0;
    }
    // keeps the allocation when it is large enough; contents are not preserved
    void resize(size_t n) {
        if (n > cap_) {
            reset();
            void* p = nullptr;
            hip_check(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "DeviceBuffer::resize");
            ptr_ = static_cast<T*>(p);
            cap_ = n;
        }
        n_ = n;
    }
    void ensure_capacity(size_t n) { resize(n); }
    void swap(DeviceBuffer& o) noexcept {
        std::swap(ptr_, o.ptr_);
        std::swap(n_, o.n_);
        std::swap(cap_, o.cap_);
    }

  private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
    size_t cap_ = 0;
};

// ---- execution context ------------------------------------------------------
// One device and one stream (default priority). Move-only.
class HipContext {
  public:
    explicit HipContext(int device_id = 0) : device_(device_id) {
        hip_check(hipSetDevice(device_id), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
    ~HipContext() {
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;
    HipContext(HipContext&& o) noexcept : device_(o.device_), stream_(o.stream_) {
        o.stream_ = nullptr;
    }
    HipContext& operator=(HipContext&& o) noexcept {
        std::swap(device_, o.device_);
        std::swap(stream_, o.stream_);
        return *this;
    }

    int device() const noexcept { return device_; }
    hipStream_t stream() const noexcept { return stream_; }
    mf3d_stream c_stream() const noexcept { return reinterpret_cast<mf3d_stream>(stream_); }
    void synchronize() const { hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

  private:
    int device_ = 0;
    hipStream_t stream_ = nullptr;
};

using CudaContext = HipContext;

// ---- pin (src/numerics/pin_spec.hpp:56-120) ----------------------------------
enum class PinMode { Auto, On, Off };

struct PinSpec {
    bool enabled = false;
    size_t index = 0;  // always cell (0,0,0)
    PinSpec() = default;
    explicit PinSpec(bool en) : enabled(en) {}
};

inline bool needs_pin(const BCSpec& bc) {
    for (const BCFace* f : bc.faces())
        if (f->type == BCType::Dirichlet) return false;
    return true;
}

inline bool pin_enabled(PinMode mode, const BCSpec& bc) {
    return mode == PinMode::On ? true : mode == PinMode::Off ? false : needs_pin(bc);
}

namespace io {
// PinConfig (src/io/config/Config.hpp:52-54)
struct PinConfig {
    PinMode mode = PinMode::Auto;
};

// FlowYamlConfig (src/io/config/Config.hpp:86-114): the `flow:` block of the
// pipeline's YAML as the parser leaves it; HeadSolveConfig::from_yaml turns it
// into solver settings. Same members and defaults (no YAML parser here: the
// caller fills the struct).
struct FlowYamlConfig {
    std::string solver = "mg";  // "mg", "cg", "mg_cg" / "pcg_mg"
    int mg_levels = 4;
    int mg_pre_smooth = 2;
    int mg_post_smooth = 2;
    int mg_coarse_iters = 50;
    int mg_max_cycles = 20;
    int cg_max_iter = 1000;
    real cg_rtol = 1e-8;
    int cg_check_every = 10;
    real rtol = 1e-6;
    BCSpec bc;
    PinConfig pin;
    bool verify_velocity = false;
};
}  // namespace io

}  // namespace macroflow3d
