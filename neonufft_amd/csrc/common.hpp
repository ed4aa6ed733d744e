// Internal types shared by the host orchestration and the HIP kernels of the
// MI355X NeoNUFFT library.
#pragma once

#include <hip/hip_runtime.h>

#include <neonufft/allocator.hpp>
#include <neonufft/exceptions.hpp>
#include <neonufft/gpu/types.hpp>
#include <neonufft/types.hpp>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>

namespace nnf {

using neonufft::IntType;

// ---------------------------------------------------------------------------
// Error handling. The public exception types only hold string literals (as in
// the reference); a per-thread detail string keeps the runtime's message for
// the C ABI's nnf_last_error().
// ---------------------------------------------------------------------------
void set_error_detail(const std::string& msg);
const std::string& error_detail();

void throw_hip(hipError_t err, const char* what);

inline void check_hip(hipError_t err, const char* what) {
  if (err != hipSuccess) throw_hip(err, what);
}

// Kernel launches are checked through hipGetLastError so that an invalid
// configuration surfaces as GPUError at the call site.
inline void check_launch(const char* what) { check_hip(hipGetLastError(), what); }

// ---------------------------------------------------------------------------
// Device memory owned through the plan's injected Allocator
// (include/neonufft/allocator.hpp).
// ---------------------------------------------------------------------------
template <typename E>
class DeviceBuffer {
public:
  DeviceBuffer() = default;
  DeviceBuffer(std::size_t count, std::shared_ptr<neonufft::Allocator> alloc) { reset(count, alloc); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    release();
    ptr_ = o.ptr_;
    count_ = o.count_;
    alloc_ = std::move(o.alloc_);
    o.ptr_ = nullptr;
    o.count_ = 0;
    return *this;
  }
  ~DeviceBuffer() { release(); }

  void reset(std::size_t count, std::shared_ptr<neonufft::Allocator> alloc) {
    release();
    alloc_ = std::move(alloc);
    count_ = count;
    if (count) {
      ptr_ = static_cast<E*>(alloc_->allocate(count * sizeof(E)));
      if (!ptr_) throw neonufft::MemoryAllocError("Neonufft:: device allocation returned null");
    }
  }
  // Grow-only: keeps the allocation when it is already large enough.
  void ensure(std::size_t count, std::shared_ptr<neonufft::Allocator> alloc) {
    if (count > count_ || !alloc_) reset(count, std::move(alloc));
  }
  void release() noexcept {
    if (ptr_ && alloc_) alloc_->deallocate(ptr_);
    ptr_ = nullptr;
    count_ = 0;
  }

  E* data() const { return ptr_; }
  std::size_t size() const { return count_; }
  std::size_t bytes() const { return count_ * sizeof(E); }

private:
  E* ptr_ = nullptr;
  std::size_t count_ = 0;
  std::shared_ptr<neonufft::Allocator> alloc_;
};

// Temporary device memory carved out of a caller-provided region (no
// allocation): take() hands out 256-byte aligned pieces in order. With a null
// base it only counts the bytes (a dry run that sizes the region).
struct Workspace {
  char* base = nullptr;
  std::size_t cap = 0;
  std::size_t used = 0;

  template <typename E>
  E* take(std::size_t count) {
    const std::size_t off = (used + 255) & ~std::size_t(255);
    const std::size_t bytes = count * sizeof(E);
    if (base && off + bytes > cap) {
      throw neonufft::InputError("Neonufft:: workspace too small");
    }
    used = off + bytes;
    return base ? reinterpret_cast<E*>(base + off) : nullptr;
  }
};

// ---------------------------------------------------------------------------
// Device-side complex pair and per-precision helpers.
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) Cplx {
  T x, y;
};

// ---------------------------------------------------------------------------
// Point records after fold + binning (one record per point, stored in bin
// order). x[d] is the Horner argument in [0,1): x = ceil(xi - ns/2) - (xi - ns/2)
// for the fine-grid coordinate xi = u * n, computed in FP64 and rounded once.
// `cell` packs, per dimension, the first kernel cell relative to the bin's
// tile origin (bin origin - floor(ns/2)); the range is [0, bin_size].
// ---------------------------------------------------------------------------
template <typename T, int DIM>
struct PointRec;

template <>
struct alignas(16) PointRec<float, 3> {
  float x[3];
  uint32_t cell;
};
// 2D records also keep the point's linear bin index (`bin`), from which the
// row-sorted view of the row-band spreading is derived.
template <>
struct alignas(16) PointRec<float, 2> {
  float x[2];
  uint32_t cell;
  uint32_t bin;
};
template <>
struct alignas(8) PointRec<float, 1> {
  float x[1];
  uint32_t cell;
};
template <>
struct alignas(16) PointRec<double, 3> {
  double x[3];
  uint32_t cell;
  uint32_t pad;
};
template <>
struct alignas(16) PointRec<double, 2> {
  double x[2];
  uint32_t cell;
  uint32_t bin;
  uint32_t pad[2];
};
template <>
struct alignas(16) PointRec<double, 1> {
  double x[1];
  uint32_t cell;
  uint32_t pad;
};

// Bit packing of the relative cell per dimension.
template <int DIM>
struct CellPack;
template <>
struct CellPack<3> {
  static constexpr int bits = 10;
  static constexpr uint32_t mask = (1u << bits) - 1u;
  static constexpr int max_bin = 1023;
};
template <>
struct CellPack<2> {
  static constexpr int bits = 16;
  static constexpr uint32_t mask = (1u << bits) - 1u;
  static constexpr int max_bin = 65535;
};
template <>
struct CellPack<1> {
  static constexpr int bits = 32;
  static constexpr uint32_t mask = 0xffffffffu;
  static constexpr int max_bin = 1 << 30;
};

// One unit of spreading/interpolation work: up to `max_points` points of one
// bin; `job` selects the transform in batched launches (0 otherwise).
struct Subprob {
  uint32_t bin;
  uint32_t begin;
  uint32_t count;
  uint32_t job;
};

// Geometry of a grid and its bin (subgrid tile) decomposition.
struct BinGeom {
  int n[3];      // grid size per dim (1 for unused dims)
  int bin[3];    // bin edge per dim
  int nbins[3];  // number of bins per dim
  int tile[3];   // LDS tile pitch per dim = bin + ns (1 for unused dims)
};

// How raw coordinates map to the grid: v = (x - offset) * scale, then the
// fold u = frac((v + pi) / 2 pi), xi = u * n.
struct FoldParams {
  double offset[3];
  double scale[3];
};

// ES kernel evaluation parameters. For float plans the kernel is a per-piece
// polynomial (Horner) in y = 2x - 1, fitted on the host in FP64 for the plan's
// (ns, beta); coef[j * (deg + 1) + k] is the coefficient of y^k of piece j.
// Double plans evaluate exp(beta (sqrt(1 - (2t/ns)^2) - 1)) directly.
constexpr int kMaxHornerCoef = 16 * 18;
struct EsParams {
  float coef[kMaxHornerCoef];
  double beta;
};

// fp32 Horner degree for kernel width NS (chosen so that the fit error sits at
// the sqrt-branch floor ~exp(-beta) for NS <= 7 and at fp32 rounding beyond).
__host__ __device__ constexpr int horner_degree(int ns) { return ns <= 7 ? ns + 1 : ns; }

inline constexpr double kPi = 3.14159265358979323846264338327950288;

}  // namespace nnf
