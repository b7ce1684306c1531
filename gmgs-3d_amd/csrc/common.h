// Shared definitions for the MI355X geometric-multigrid PCG library.
//
// Grid layout (one problem, one level):
//   element grid  ey x ex x ez  (resY, resX, resZ of the reference level),
//   node grid     ny x nx x nz  = element grid + 1,
//   padded grid   py x px x pz  with a one-node void halo on every side and
//                 py rounded up to a multiple of 16 (64-byte FP32 rows).
//   node (y,x,z)  -> pidx = (y+1) + py*((x+1) + px*(z+1))
//   cell (y,x,z)  -> stored at the pidx of its minimum-corner node.
// Row index y follows the reference's voxel rows (row 0 = max-y side,
// common/DiscretizeDesignDomain.m:96); x and z ascend.
// Vectors are structure-of-arrays: component c of node i at v[c*N + i] with
// N = py*px*pz.  Every void and halo entry of every vector is exactly zero.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace gmgs {

constexpr int TILE_Y = 16;   // tile = 16 (y) x 4 (x) x 1 (z) nodes = one wave
constexpr int TILE_X = 4;
constexpr int WAVE = 64;

// node flag bits
constexpr uint32_t F_ELEM_MASK = 0xFFu;      // bit p: incident element p present
constexpr uint32_t F_FIX_SHIFT = 8;          // bits 8..10: fixed component c
constexpr uint32_t F_FIX_MASK = 0x7u << 8;
constexpr uint32_t F_FAST = 1u << 11;        // all 8 incident elements present and universal
constexpr uint32_t F_STORED_SHIFT = 12;      // bits 12..19: incident element p present and stored

// Element position p around a node (bits: py | px<<1 | pz<<2): the element
// whose minimum corner is (y-1+py, x-1+px, z-1+pz).  The node is that
// element's corner with offsets (1-py, 1-px, 1-pz).
// Hex8 local corner m -> (row, x, z) offsets, common/DiscretizeDesignDomain.m:80.
__host__ __device__ constexpr int hex_ro(int m) { return (m == 0 || m == 1 || m == 4 || m == 5) ? 1 : 0; }
__host__ __device__ constexpr int hex_xo(int m) { return (m == 1 || m == 2 || m == 5 || m == 6) ? 1 : 0; }
__host__ __device__ constexpr int hex_zo(int m) { return m >= 4 ? 1 : 0; }
// (ro + 2*xo + 4*zo) -> local corner index
__host__ __device__ constexpr int hex_local(int code) {
  return code == 0 ? 3 : code == 1 ? 0 : code == 2 ? 2 : code == 3 ? 1
       : code == 4 ? 7 : code == 5 ? 4 : code == 6 ? 6 : 5;
}
// local index of the node inside element at position p
__host__ __device__ constexpr int node_local_in(int p) {
  return hex_local((1 - (p & 1)) + 2 * (1 - ((p >> 1) & 1)) + 4 * (1 - ((p >> 2) & 1)));
}

struct Grid {
  int ey, ex, ez;       // element counts
  int ny, nx, nz;       // node counts
  int py, px, pz;       // padded node dims
  long long N;          // py*px*pz
  __host__ __device__ long long pidx(int y, int x, int z) const {
    return (long long)(y + 1) + (long long)py * ((x + 1) + (long long)px * (z + 1));
  }
};

inline Grid make_grid(int ey, int ex, int ez) {
  Grid g;
  g.ey = ey; g.ex = ex; g.ez = ez;
  g.ny = ey + 1; g.nx = ex + 1; g.nz = ez + 1;
  g.py = ((g.ny + 2 + TILE_Y - 1) / TILE_Y) * TILE_Y;
  g.px = g.nx + 2;
  g.pz = g.nz + 2;
  g.N = (long long)g.py * g.px * g.pz;
  return g;
}

// A tile: origin padded index; lanes cover y in [0,16), x in [0,4).
struct Tile {
  long long base;   // pidx of (ty*16, tx*4, z) in padded coordinates (no +1 shift)
};

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define GMGS_HIP(expr)                                                          \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      throw ::gmgs::HipError(std::string("HIP error '") + hipGetErrorString(_e) \
                             + "' at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define GMGS_LAUNCH_CHECK() GMGS_HIP(hipGetLastError())

inline unsigned grid1d(long long n, int block) {
  return (unsigned)((n + block - 1) / block);
}

}  // namespace gmgs

namespace gmgs {

// Owning device allocation (hipMalloc / hipFree).
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; bytes_ = o.bytes_; o.p_ = nullptr; o.bytes_ = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 256;
    GMGS_HIP(hipMalloc(&p_, bytes));
    bytes_ = bytes;
  }
  void alloc_zero(size_t bytes, hipStream_t s) {
    alloc(bytes);
    GMGS_HIP(hipMemsetAsync(p_, 0, bytes_, s));
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p_); }
  size_t bytes() const { return bytes_; }
  bool empty() const { return p_ == nullptr; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// Owning HIP event, stream and pinned host buffer: released on every exit
// path, including a HipError thrown between creation and use.
class HipEvent {
 public:
  explicit HipEvent(unsigned flags = hipEventDefault) { GMGS_HIP(hipEventCreateWithFlags(&e_, flags)); }
  HipEvent(const HipEvent&) = delete;
  HipEvent& operator=(const HipEvent&) = delete;
  ~HipEvent() { if (e_) (void)hipEventDestroy(e_); }
  operator hipEvent_t() const { return e_; }

 private:
  hipEvent_t e_ = nullptr;
};

class HipStream {
 public:
  HipStream() { GMGS_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  HipStream(const HipStream&) = delete;
  HipStream& operator=(const HipStream&) = delete;
  ~HipStream() {
    if (s_) {
      (void)hipStreamSynchronize(s_);
      (void)hipStreamDestroy(s_);
    }
  }
  operator hipStream_t() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

template <typename T>
class PinnedBuf {
 public:
  explicit PinnedBuf(size_t n) { GMGS_HIP(hipHostMalloc((void**)&p_, sizeof(T) * (n ? n : 1), hipHostMallocDefault)); }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { if (p_) (void)hipHostFree(p_); }
  T* get() const { return p_; }
  T& operator[](size_t i) const { return p_[i]; }

 private:
  T* p_ = nullptr;
};

}  // namespace gmgs
