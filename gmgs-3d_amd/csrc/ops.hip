// Level operators on the dense padded grids.
//
// One wavefront owns one 16x4x1 tile of nodes; a lane owns one node and
// pulls everything it needs from its 3x3x3 neighbourhood, so no output is
// ever written by two lanes (no atomics, bitwise-reproducible sums).
//
//   apply  : y = A_l u  (kernel/AtX.m:1-58), optionally fused with the damped
//            Jacobi step u = omega*r./diagK, r' = r - A u (JacobiSmoother.m:5-9)
//   restrict, prolong: span-2 trilinear transfers (RestrictResidual.m,
//            InterpolationDeviation.m) written on the structured grid.
#include "ops.h"
#include "launch.h"

namespace gmgs {

namespace {

__device__ inline double block_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

struct LaneNode {
  long long n;
  uint32_t fl;
  bool active;
  int yp, xp, zp;
};

// node of this lane in tile t (one 64-node tile per wave)
__device__ inline LaneNode lane_node_at(const Grid& g, const uint32_t* __restrict__ flags,
                                        const TileRec* __restrict__ tiles, long long ntiles,
                                        long long t) {
  LaneNode r;
  r.active = false;
  r.n = 0;
  r.fl = 0;
  const int lane = threadIdx.x & 63;
  r.yp = r.xp = r.zp = 0;
  if (t < ntiles) {
    const TileRec tr = tiles[t];
    r.yp = tr.y0 + (lane & 15);
    r.xp = tr.x0 + (lane >> 4);
    r.zp = tr.z;
    if (r.xp < g.px) {
      r.n = r.yp + (long long)g.py * (r.xp + (long long)g.px * r.zp);
      r.fl = flags[r.n];
      r.active = (r.fl & F_ELEM_MASK) != 0;
    }
  }
  return r;
}

// one tile per wave, four per block, tile = 4 block + wave
__device__ inline LaneNode lane_node(const Grid& g, const uint32_t* __restrict__ flags,
                                     const TileRec* __restrict__ tiles, long long ntiles, int block) {
  return lane_node_at(g, flags, tiles, ntiles, (long long)block * 4 + (threadIdx.x >> 6));
}

// tile loop of the pointwise kernels: a fixed grid walks the tile list, so a
// kernel that reduces leaves one partial per block, nblocks_ in all (block_,
// nblocks_: the kernel body's block and block count)
#define GMGS_TILE_LOOP(g, flags, tiles, nt, ln_)                                              \
  for (long long tb_ = (long long)block_ * 4; tb_ < (nt); tb_ += (long long)nblocks_ * 4) \
    if (const LaneNode ln_ = lane_node_at(g, flags, tiles, nt, tb_ + (threadIdx.x >> 6)); true)

template <typename T, typename TR, int MODE>
struct InputU {
  const TR* __restrict__ in;
  const T* __restrict__ invd;
  T omega;
  long long N;
  __device__ inline T operator()(long long m, int d) const {
    if (MODE == AP_AX) return (T)in[d * N + m];
    return omega * invd[d * N + m] * (T)in[d * N + m];
  }
};

template <typename T, typename TR, int MODE, bool HAS_EOP, bool HAS_EMOD>
struct KApply {
  static constexpr int kThreads = 256;
  LevelOp<T> op;
  const TileRec* tiles;
  long long ntiles;
  const TR* in;
  T omega;
  T* out;
  TR* xacc;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    LevelOp<T> op = this->op;
    const TileRec* __restrict__ tiles = this->tiles;
    long long ntiles = this->ntiles;
    const TR* __restrict__ in = this->in;
    T omega = this->omega;
    T* __restrict__ out = this->out;
    TR* __restrict__ xacc = this->xacc;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  const LaneNode ln_ = lane_node(op.g, op.flags, tiles, ntiles, block_);
  double dsum = 0.0;
  if (ln_.active) {
    const long long N = op.g.N;
    const long long n = ln_.n;
    const uint32_t fl = ln_.fl;
    const long long sx = op.g.py, sz = (long long)op.g.py * op.g.px;
    InputU<T, TR, MODE> U{in, op.invd, omega, N};
    T a0 = 0, a1 = 0, a2 = 0;
    if (!HAS_EMOD && (fl & F_FAST)) {
      const T* __restrict__ S = op.stencil;
#pragma unroll
      for (int o = 0; o < 27; ++o) {
        const int dy = o % 3 - 1, dx = (o / 3) % 3 - 1, dz = o / 9 - 1;
        const long long m = n + dy + dx * sx + dz * sz;
        const T u0 = U(m, 0), u1 = U(m, 1), u2 = U(m, 2);
        a0 += S[o * 9 + 0] * u0 + S[o * 9 + 1] * u1 + S[o * 9 + 2] * u2;
        a1 += S[o * 9 + 3] * u0 + S[o * 9 + 4] * u1 + S[o * 9 + 5] * u2;
        a2 += S[o * 9 + 6] * u0 + S[o * 9 + 7] * u1 + S[o * 9 + 8] * u2;
      }
    } else {
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        if (!((fl >> p) & 1u)) continue;
        const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = p >> 2;
        const long long cell = n - (1 - pyb) - (1 - pxb) * sx - (1 - pzb) * sz;
        const T* __restrict__ K = op.Ke;
        if (HAS_EOP) {
          const int e = op.eop[cell];
          if (e > 0) K = op.Ks + (long long)(e - 1) * 576;
        }
        const int ln = node_local_in(p);
        T e0 = 0, e1 = 0, e2 = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const long long mm = n + (pyb - 1 + hex_ro(m)) + (pxb - 1 + hex_xo(m)) * sx
                               + (pzb - 1 + hex_zo(m)) * sz;
          const T u0 = U(mm, 0), u1 = U(mm, 1), u2 = U(mm, 2);
          const T* k0 = K + (3 * ln + 0) * 24 + 3 * m;
          const T* k1 = K + (3 * ln + 1) * 24 + 3 * m;
          const T* k2 = K + (3 * ln + 2) * 24 + 3 * m;
          e0 += k0[0] * u0 + k0[1] * u1 + k0[2] * u2;
          e1 += k1[0] * u0 + k1[1] * u1 + k1[2] * u2;
          e2 += k2[0] * u0 + k2[1] * u1 + k2[2] * u2;
        }
        if (HAS_EMOD) {
          const T s = op.emod[cell];
          e0 *= s; e1 *= s; e2 *= s;
        }
        a0 += e0; a1 += e1; a2 += e2;
      }
    }
    // fixed DOFs: AtX output zeroed (kernel/AtX.m:57)
    if (fl & (1u << (F_FIX_SHIFT + 0))) a0 = 0;
    if (fl & (1u << (F_FIX_SHIFT + 1))) a1 = 0;
    if (fl & (1u << (F_FIX_SHIFT + 2))) a2 = 0;
    if (MODE == AP_AX) {
      out[n] = a0;
      out[N + n] = a1;
      out[2 * N + n] = a2;
      if (partial) {
        dsum = (double)in[n] * (double)a0 + (double)in[N + n] * (double)a1
               + (double)in[2 * N + n] * (double)a2;
      }
    } else {
      const bool z = op.zero_fixed != 0;
      out[n] = (z && (fl & (1u << (F_FIX_SHIFT + 0)))) ? (T)0 : (T)in[n] - a0;
      out[N + n] = (z && (fl & (1u << (F_FIX_SHIFT + 1)))) ? (T)0 : (T)in[N + n] - a1;
      out[2 * N + n] = (z && (fl & (1u << (F_FIX_SHIFT + 2)))) ? (T)0 : (T)in[2 * N + n] - a2;
      if (xacc) {
        xacc[n] += (TR)U(n, 0);
        xacc[N + n] += (TR)U(n, 1);
        xacc[2 * N + n] += (TR)U(n, 2);
      }
    }
  }
  if (partial) {
    const double b = block_sum(dsum);
    if (threadIdx.x == 0) partial[block_] = b;
  }
  }
};

// Level operator from assembled stencil rows (coarse levels).  One wave per
// 16 x 4 x 1 node tile, four tiles per block.  The tile's 18 x 6 x 3
// neighbourhood of the input (u = omega r / d for the smoother) is staged in
// LDS by the wave; each lane then multiplies its 27 neighbours by its row of
// the operator.  Rows are stored one after another, each z-plane's 81
// coefficients padded to whole 16-byte groups (q' = (q / 81) QP + q % 81), so
// a lane reads its row with consecutive 16-byte loads, one z-plane at a time.
// Equal rows are stored once (nodes with the same neighbourhood of elements
// share them), which keeps the rows of a level in the last-level cache.
// F_FAST lanes share the universal row.
// Coefficients read through a constant-address-space pointer kept opaque to
// loop-invariant code motion: scalar loads into SGPR operands where they are
// used, instead of values hoisted into vector registers.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

template <typename T>
__device__ inline cptr<T> opaque(const T* p) {
  // (the pointer is uniform: read from the first lane so that it sits in
  // scalar registers however it was loaded)
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u),
                 hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  cptr<T> q = (cptr<T>)(((unsigned long long)hi << 32) | lo);
  asm volatile("" : "+s"(q));
  return q;
}

template <typename T>
struct alignas(16) CoefGroup {
  T v[16 / sizeof(T)];
};

template <typename T, typename TR, int MODE>
struct KApplyAs {
  static constexpr int kThreads = 256;
  LevelOp<T> op;
  const TileRec* tiles;
  long long ntiles;
  const TR* in;
  T omega;
  const T* in2;
  T* out;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    LevelOp<T> op = this->op;
    const TileRec* __restrict__ tiles = this->tiles;
    long long ntiles = this->ntiles;
    const TR* __restrict__ in = this->in;
    T omega = this->omega;
    const T* __restrict__ in2 = this->in2;
    T* __restrict__ out = this->out;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  constexpr int SY = 18, SX = 6, NS = SY * SX * 3;   // staged nodes per tile
  constexpr int G = 16 / (int)sizeof(T), NGP = stencil_plane_groups((int)sizeof(T));
  __shared__ T sU[4][3][NS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long t = (long long)block_ * 4 + wave;
  const Grid g = op.g;
  const long long N = g.N, sx = g.py, sz = (long long)g.py * g.px;
  double dsum = 0.0;
  TileRec tr{0, 0, 0, 0};
  int arow = 0;                                     // this lane's row (F_FAST lanes: the universal row)
  uint32_t fl = 0;
  TR cen[3] = {0, 0, 0};                            // the lane's own in (SMOOTH, AX dot) or in2 (RESID)
  const int ly = lane & 15, lx = lane >> 4;
  long long n = 0;
  if (t < ntiles) {
    tr = tiles[t];
    arow = op.arow[t * 64 + lane];                  // with the tile record: it needs only t
    // Staging, all loads issued before the first is used.  Index 0 of every
    // padded array is a corner halo entry holding 0: a position outside the
    // grid loads from there (32-bit indices, 3 N < 2^31 checked at set-up).
    // Tiles start at y, x >= 0 and hold solid nodes (z >= 1).
    // Rows (c, zz, xx) of 18 y-values: lane (yy = lane % 16, w = lane / 16)
    // loads y-values 0..15 of the rows with xx = w and xx = 4 + w (w < 2) of
    // every (c, zz), and lane l < 54 the y-values 16 and 17 of row l.
    const int gy0 = tr.y0 - 1, gx0 = tr.x0 - 1, gz0 = tr.z - 1;
    const int yy = lane & 15, w = lane >> 4;
    const int isx = (int)sx, isz = (int)sz, iN = (int)N;
    const bool y_ok = gy0 + yy >= 0 && gy0 + yy < g.py;
    const bool xa_ok = gx0 + w >= 0 && gx0 + w < g.px, xb_ok = w < 2 && gx0 + 4 + w < g.px;
    const int ba = gy0 + yy + isx * (gx0 + w), bb = ba + 4 * isx;
    const int le = lane < 54 ? lane : 0;                   // edge row of this lane
    const int ec = le / 18, ezz = (le % 18) / 6, exx = le % 6;
    const bool e_ok = lane < 54 && gx0 + exx >= 0 && gx0 + exx < g.px && gz0 + ezz < g.pz;
    const int eb = gy0 + 16 + isx * (gx0 + exx) + isz * (gz0 + ezz) + iN * ec;
    const bool e16 = e_ok && gy0 + 16 < g.py, e17 = e_ok && gy0 + 17 < g.py;
    TR va[9], vb[9], ve[2];
    T da[MODE == AP_SMOOTH ? 9 : 1], db[MODE == AP_SMOOTH ? 9 : 1], de[MODE == AP_SMOOTH ? 2 : 1];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int c = k / 3, zz = k % 3;
      const bool z_ok = gz0 + zz < g.pz;                   // gz0 >= 0: solid nodes have z >= 1
      const int u = iN * c + isz * (gz0 + zz);
      const unsigned ma = (y_ok && xa_ok && z_ok) ? (unsigned)(ba + u) : 0u;
      const unsigned mb = (y_ok && xb_ok && z_ok) ? (unsigned)(bb + u) : 0u;
      va[k] = in[ma];
      vb[k] = in[mb];
      if (MODE == AP_SMOOTH) {
        da[k] = op.invd[ma];
        db[k] = op.invd[mb];
      }
    }
    {
      const unsigned m16 = e16 ? (unsigned)eb : 0u, m17 = e17 ? (unsigned)(eb + 1) : 0u;
      ve[0] = in[m16];
      ve[1] = in[m17];
      if (MODE == AP_SMOOTH) {
        de[0] = op.invd[m16];
        de[1] = op.invd[m17];
      }
    }
    // the lane's own flags, loaded with the staging batch
    const int xp = tr.x0 + lx;
    if (xp < g.px) {
      n = (tr.y0 + ly) + sx * xp + sz * tr.z;
      fl = op.flags[n];
      if (MODE == AP_RESID) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cen[c] = (TR)in2[c * N + n];
      } else if (MODE == AP_SMOOTH || partial) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cen[c] = in[c * N + n];
      }
    }
    // every loaded value is taken here, after the last load is issued: the
    // compiler may not sink a load into the branch that stores it (which
    // would put a wait between the loads)
    asm volatile("" : "+v"(fl), "+v"(cen[0]), "+v"(cen[1]), "+v"(cen[2]));
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      asm volatile("" : "+v"(va[k]), "+v"(vb[k]));
      if (MODE == AP_SMOOTH) asm volatile("" : "+v"(da[k]), "+v"(db[k]));
    }
    asm volatile("" : "+v"(ve[0]), "+v"(ve[1]));
    if (MODE == AP_SMOOTH) asm volatile("" : "+v"(de[0]), "+v"(de[1]));
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int c = k / 3, zz = k % 3;
      T a = (T)va[k], b = (T)vb[k];
      if (MODE == AP_SMOOTH) {
        a = omega * da[k] * a;
        b = omega * db[k] * b;
      }
      sU[wave][c][SY * SX * zz + SY * w + yy] = a;
      if (w < 2) sU[wave][c][SY * SX * zz + SY * (4 + w) + yy] = b;
    }
    if (lane < 54) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        T v = (T)ve[k];
        if (MODE == AP_SMOOTH) v = omega * de[k] * v;
        sU[wave][ec][SY * SX * ezz + SY * exx + 16 + k] = v;
      }
    }
  }
  // each wave stages and reads only its own tile: a wave-level barrier (LDS
  // operations of one wave complete in order)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (fl & F_ELEM_MASK) {
    T a[3] = {0, 0, 0};
    // F_FAST nodes (all eight elements present and universal): the universal
    // stencil from its 30 reflection-class coefficients, read as scalar
    // operands, over the three staged planes (input plane z - 1 + zz, so
    // oz = zz - 1 is the input's z offset from the output)
    const bool ucls = op.cls != nullptr;
    if (ucls && (fl & F_FAST)) {
      const cptr<T> C = opaque(op.cls);
      T cf[30];
#pragma unroll
      for (int i = 0; i < 30; ++i) cf[i] = C[i];
#pragma unroll
      for (int zz = 0; zz < 3; ++zz) {
        const int oz = zz - 1, q = oz != 0 ? 1 : 0;
        const T* __restrict__ pl = &sU[wave][0][SY * SX * zz];
        // neighbour (dy, dx) = (a - 1, b - 1), component c
#define GMGS_AU(a_, b_, c_) pl[(c_) * NS + SY * (lx + (b_)) + (ly + (a_))]
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T C0 = GMGS_AU(1, 1, c);
          const T Cy = GMGS_AU(0, 1, c) + GMGS_AU(2, 1, c);
          const T Cx = GMGS_AU(1, 0, c) + GMGS_AU(1, 2, c);
          const T Cxy = (GMGS_AU(0, 0, c) + GMGS_AU(2, 2, c)) + (GMGS_AU(0, 2, c) + GMGS_AU(2, 0, c));
          a[c] += cf[8 * c + 4 * q] * C0 + cf[8 * c + 1 + 4 * q] * Cy + cf[8 * c + 2 + 4 * q] * Cx
                  + cf[8 * c + 3 + 4 * q] * Cxy;
        }
        const T Dxy_x = (GMGS_AU(2, 2, 0) + GMGS_AU(0, 0, 0)) - (GMGS_AU(2, 0, 0) + GMGS_AU(0, 2, 0));
        const T Dxy_y = (GMGS_AU(2, 2, 1) + GMGS_AU(0, 0, 1)) - (GMGS_AU(2, 0, 1) + GMGS_AU(0, 2, 1));
        a[0] += cf[24 + q] * Dxy_y;
        a[1] += cf[24 + q] * Dxy_x;
        if (oz != 0) {
          const T Dx0_x = GMGS_AU(1, 2, 0) - GMGS_AU(1, 0, 0);
          const T Dx1_x = (GMGS_AU(0, 2, 0) - GMGS_AU(0, 0, 0)) + (GMGS_AU(2, 2, 0) - GMGS_AU(2, 0, 0));
          const T Dx0_z = GMGS_AU(1, 2, 2) - GMGS_AU(1, 0, 2);
          const T Dx1_z = (GMGS_AU(0, 2, 2) - GMGS_AU(0, 0, 2)) + (GMGS_AU(2, 2, 2) - GMGS_AU(2, 0, 2));
          const T Dy0_y = GMGS_AU(2, 1, 1) - GMGS_AU(0, 1, 1);
          const T Dy1_y = (GMGS_AU(2, 0, 1) - GMGS_AU(0, 0, 1)) + (GMGS_AU(2, 2, 1) - GMGS_AU(0, 2, 1));
          const T Dy0_z = GMGS_AU(2, 1, 2) - GMGS_AU(0, 1, 2);
          const T Dy1_z = (GMGS_AU(2, 0, 2) - GMGS_AU(0, 0, 2)) + (GMGS_AU(2, 2, 2) - GMGS_AU(0, 2, 2));
          const T sgn = (T)oz;
          a[0] += sgn * (cf[26] * Dx0_z + cf[27] * Dx1_z);
          a[2] += sgn * (cf[26] * Dx0_x + cf[27] * Dx1_x);
          a[1] += sgn * (cf[28] * Dy0_z + cf[29] * Dy1_z);
          a[2] += sgn * (cf[28] * Dy0_y + cf[29] * Dy1_y);
        }
#undef GMGS_AU
      }
    } else {
      // the node's assembled row (F_FAST nodes: the universal row), read
      // group after group; a wave whose nodes are all F_FAST skips these loads
      const char* __restrict__ base =
          reinterpret_cast<const char*>(op.asc) + (long long)arow * (3 * NGP * 16);
#pragma unroll 1
      for (int kz = 0; kz < 3; ++kz) {
        const T* __restrict__ sp = &sU[wave][0][SY * SX * kz + (ly + 1) + SY * (lx + 1)];
#pragma unroll
        for (int g0 = 0; g0 < NGP; ++g0) {
          const CoefGroup<T> cg = *reinterpret_cast<const CoefGroup<T>*>(base + (kz * NGP + g0) * 16);
#pragma unroll
          for (int k = 0; k < G; ++k) {
            const int r = g0 * G + k;                   // coefficient within the z-plane
            if (r < 81) {
              const int j = r / 9, cc = (r / 3) % 3, d = r % 3;
              const int dy = j % 3 - 1, dx = j / 3 - 1;
              a[cc] += cg.v[k] * sp[d * NS + dy + SY * dx];
            }
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool fx = (fl & (1u << (F_FIX_SHIFT + c))) != 0;
      const T ac = fx ? (T)0 : a[c];          // AtX output zeroed on fixed DOFs (kernel/AtX.m:57)
      if (MODE == AP_AX) {
        out[c * N + n] = ac;
        if (partial) dsum += (double)cen[c] * (double)ac;
      } else if (MODE == AP_RESID) {
        out[c * N + n] = (T)cen[c] - ac;
      } else {
        out[c * N + n] = (fx && op.zero_fixed) ? (T)0 : (T)cen[c] - ac;
      }
    }
  }
  if (partial) {
    const double b = block_sum(dsum);
    if (threadIdx.x == 0) partial[block_] = b;
  }
  }
};

// The split operator's second half: one wave per row box.  The wave stages
// the box's neighbourhood (18 x (kRowBoxX + 2) x (kRowBoxZ + 2) nodes, u = omega r / d for
// the smoother) in LDS, 16 loads per lane at a time with out-of-grid positions
// reading index 0 (a zero halo entry), and each lane with an entry multiplies
// its node's 27 neighbours by the node's assembled row, in the row's
// coefficient order (the sums of k_apply_as).
template <typename T, typename TR, int MODE>
struct KApplyRows {
  static constexpr int kThreads = 256;
  LevelOp<T> op;
  const RowTile* boxes;
  const RowEntry* entries;
  long long nbox;
  const TR* in;
  T omega;
  const T* in2;
  T* out;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    LevelOp<T> op = this->op;
    const RowTile* __restrict__ boxes = this->boxes;
    const RowEntry* __restrict__ entries = this->entries;
    long long nbox = this->nbox;
    const TR* __restrict__ in = this->in;
    T omega = this->omega;
    const T* __restrict__ in2 = this->in2;
    T* __restrict__ out = this->out;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  constexpr int SY = 18, SX = kRowBoxX + 2, SZ = kRowBoxZ + 2, NP = SY * SX * SZ, NS = 3 * NP;
  constexpr int G = 16 / (int)sizeof(T), NGP = stencil_plane_groups((int)sizeof(T));
  constexpr int BATCH = 16;
  __shared__ T sU[4][NS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long r = (long long)block_ * 4 + wave;
  const Grid g = op.g;
  const long long N = g.N;
  double dsum = 0.0;
  RowTile bx{0, 0, 0, 0};
  RowEntry en{-1, 0};
  long long n = 0;
  uint32_t fl = 0;
  TR cen[3] = {0, 0, 0};                 // the node's in (SMOOTH, AX dot) or in2 (RESID)
  if (r < nbox) {
    bx = boxes[r];
    en = entries[r * 64 + lane];
    if (en.pos >= 0) {
      n = (bx.y0 + (en.pos & 15)) + g.py * ((bx.x0 + ((en.pos >> 4) & 7)) + g.px * (long long)(bx.z0 + (en.pos >> 7)));
      fl = op.flags[n];
      if (MODE == AP_RESID) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cen[c] = (TR)in2[c * N + n];
      } else if (MODE == AP_SMOOTH || partial) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cen[c] = in[c * N + n];
      }
    }
    // 32-bit indices: 3 N < 2^31 (checked at set-up)
    const int gy0 = bx.y0 - 1, gx0 = bx.x0 - 1, gz0 = bx.z0 - 1;
    const int isx = (int)g.py, isz = (int)(g.py * g.px), iN = (int)N;
    for (int b0 = 0; b0 < NS; b0 += 64 * BATCH) {
      TR v[BATCH];
      T d[MODE == AP_SMOOTH ? BATCH : 1];
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int e = b0 + k * 64 + lane;
        const int c = e / NP, rem = e % NP, zz = rem / (SY * SX), xx = (rem / SY) % SX, yy = rem % SY;
        const int y = gy0 + yy, x = gx0 + xx, z = gz0 + zz;
        const bool ok = e < NS && y >= 0 && y < g.py && x >= 0 && x < g.px && z >= 0 && z < g.pz;
        const unsigned m = ok ? (unsigned)(y + isx * x + isz * z + iN * c) : 0u;
        v[k] = in[m];
        if (MODE == AP_SMOOTH) d[k] = op.invd[m];
      }
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int e = b0 + k * 64 + lane;
        if (e < NS) {
          T a = (T)v[k];
          if (MODE == AP_SMOOTH) a = omega * d[k] * a;
          sU[wave][e] = a;
        }
      }
    }
  }
  // each wave stages and reads only its own box
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (en.pos >= 0) {
    const int ly = en.pos & 15, lx = (en.pos >> 4) & 7, lz = en.pos >> 7;
    T a[3] = {0, 0, 0};
    const char* __restrict__ base = reinterpret_cast<const char*>(op.asc) + (long long)en.row * (3 * NGP * 16);
#pragma unroll 1
    for (int kz = 0; kz < 3; ++kz) {
      const T* __restrict__ sp = &sU[wave][SY * SX * (lz + kz) + SY * (lx + 1) + (ly + 1)];
#pragma unroll
      for (int g0 = 0; g0 < NGP; ++g0) {
        const CoefGroup<T> cg = *reinterpret_cast<const CoefGroup<T>*>(base + (kz * NGP + g0) * 16);
#pragma unroll
        for (int k = 0; k < G; ++k) {
          const int q = g0 * G + k;                     // coefficient within the z-plane
          if (q < 81) {
            const int j = q / 9, cc = (q / 3) % 3, dd = q % 3;
            const int dy = j % 3 - 1, dx = j / 3 - 1;
            a[cc] += cg.v[k] * sp[dd * NP + dy + SY * dx];
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool fx = (fl & (1u << (F_FIX_SHIFT + c))) != 0;
      const T ac = fx ? (T)0 : a[c];          // AtX output zeroed on fixed DOFs (kernel/AtX.m:57)
      if (MODE == AP_AX) {
        out[c * N + n] = ac;
        if (partial) dsum += (double)cen[c] * (double)ac;
      } else if (MODE == AP_RESID) {
        const T v = (T)cen[c] - ac;
        out[c * N + n] = v;
        if (partial) dsum += (double)v * (double)v;
      } else {
        out[c * N + n] = (fx && op.zero_fixed) ? (T)0 : (T)cen[c] - ac;
      }
    }
  }
  if (partial) {
    const double b = block_sum(dsum);
    if (threadIdx.x == 0) partial[block_] = b;
  }
  }
};

// canonical FP64 rows [q][row] (stride 'rows') -> the row-major layout
// [u][q'] of k_apply_as in T for the rows rep[0..nu): q' = (q / 81) QP +
// q % 81, QP = G x (16-byte groups per z-plane), 3 QP elements per row (the
// padding of each z-plane zero)
template <typename T>
__global__ void k_relayout_rows(const double* __restrict__ src, long long rows,
                                const long long* __restrict__ rep, long long nu, T* __restrict__ dst) {
  constexpr int G = 16 / (int)sizeof(T), QP = G * stencil_plane_groups((int)sizeof(T));
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)kStencilCoefs * nu) return;
  const int q = (int)(i % kStencilCoefs);
  const long long u = i / kStencilCoefs;
  const int qp = (q / 81) * QP + q % 81;
  dst[u * (3 * QP) + qp] = (T)src[(long long)q * rows + rep[u]];
}

// two 64-bit fingerprints of every FP64 row [q][row]: the bit patterns of its
// 243 coefficients through two multiply-xorshift chains with different
// constants
__global__ void k_row_hash(const double* __restrict__ src, long long rows,
                           unsigned long long* __restrict__ h) {
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  unsigned long long a = 0x243F6A8885A308D3ull, b = 0x13198A2E03707344ull;
  for (int q = 0; q < kStencilCoefs; ++q) {
    const unsigned long long v = (unsigned long long)__double_as_longlong(src[(long long)q * rows + row]);
    a = (a ^ v) * 0x9E3779B97F4A7C15ull;
    a ^= a >> 29;
    b = (b + v) * 0xC2B2AE3D27D4EB4Full;
    b ^= b >> 31;
  }
  h[2 * row] = a;
  h[2 * row + 1] = b;
}

// *bad = 1 if some row's coefficients differ bitwise from its representative's
__global__ void k_row_verify(const double* __restrict__ src, long long rows, const int* __restrict__ uid,
                             const long long* __restrict__ rep, int* __restrict__ bad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)kStencilCoefs * rows) return;
  const long long row = i % rows;
  const long long q = i / rows;
  const long long r = rep[uid[row]];
  if (__double_as_longlong(src[q * rows + row]) != __double_as_longlong(src[q * rows + r])) atomicOr(bad, 1);
}

// the row of a solid node in the level operator as a 27-point stencil: the
// same element sum as k_apply (universal or stored element matrices, SIMP
// scale), accumulated per neighbour instead of multiplied by the input
template <bool HAS_EOP, bool HAS_EMOD>
struct KAssembleStencil {
  static constexpr int kThreads = 256;
  LevelOp<double> op;
  const TileRec* tiles;
  long long ntiles;
  const int* arow;
  long long stride;
  double* asc;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    LevelOp<double> op = this->op;
    const TileRec* __restrict__ tiles = this->tiles;
    long long ntiles = this->ntiles;
    const int* __restrict__ arow = this->arow;
    long long stride = this->stride;
    double* __restrict__ asc = this->asc;
    (void)block_;
    (void)nblocks_;

  if (block_ == 0 && threadIdx.x < kStencilCoefs)
    asc[(long long)threadIdx.x * stride + stride - 1] = op.stencil[threadIdx.x];
  const LaneNode ln_ = lane_node(op.g, op.flags, tiles, ntiles, block_);
  if (!ln_.active || (ln_.fl & F_FAST)) return;
  const long long slot = arow[((long long)block_ * 4 + (threadIdx.x >> 6)) * 64 + (threadIdx.x & 63)];
  const long long n = ln_.n;
  const long long sx = op.g.py, sz = (long long)op.g.py * op.g.px;
  for (int p = 0; p < 8; ++p) {
    if (!((ln_.fl >> p) & 1u)) continue;
    const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = p >> 2;
    const long long cell = n - (1 - pyb) - (1 - pxb) * sx - (1 - pzb) * sz;
    const double* __restrict__ K = op.Ke;
    if (HAS_EOP) {
      const int e = op.eop[cell];
      if (e > 0) K = op.Ks + (long long)(e - 1) * 576;
    }
    const double sc = HAS_EMOD ? op.emod[cell] : 1.0;
    const int ln = node_local_in(p);
    for (int m = 0; m < 8; ++m) {
      const int dy = pyb - 1 + hex_ro(m), dx = pxb - 1 + hex_xo(m), dz = pzb - 1 + hex_zo(m);
      const int o = (dy + 1) + 3 * (dx + 1) + 9 * (dz + 1);
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d)
          asc[(long long)(9 * o + 3 * c + d) * stride + slot] += sc * K[(3 * ln + c) * 24 + 3 * m + d];
    }
  }
  }
};

template <typename T>
struct KRestrict {
  static constexpr int kThreads = 256;
  Grid gc;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  Grid gf;
  const T* rf;
  T* rc;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid gc = this->gc;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    Grid gf = this->gf;
    const T* __restrict__ rf = this->rf;
    T* __restrict__ rc = this->rc;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  const LaneNode ln_ = lane_node(gc, flags, tiles, nt, block_);
  if (!ln_.active) return;
  const long long Nf = gf.N;
  const long long sx = gf.py, sz = (long long)gf.py * gf.px;
  const long long nf = gf.pidx(2 * (ln_.yp - 1), 2 * (ln_.xp - 1), 2 * (ln_.zp - 1));
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    const int dy = o % 3 - 1, dx = (o / 3) % 3 - 1, dz = o / 9 - 1;
    const T w = (dy ? (T)0.5 : (T)1) * (dx ? (T)0.5 : (T)1) * (dz ? (T)0.5 : (T)1);
    const long long m = nf + dy + dx * sx + dz * sz;
    a0 += w * rf[m];
    a1 += w * rf[Nf + m];
    a2 += w * rf[2 * Nf + m];
  }
  const long long N = gc.N;
  rc[ln_.n] = a0;
  rc[N + ln_.n] = a1;
  rc[2 * N + ln_.n] = a2;
  }
};

template <typename T, typename TR, typename TZ>
struct KProlongSmooth {
  static constexpr int kThreads = 256;
  Grid gf;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  Grid gc;
  const T* xc;
  const T* invd;
  const TR* r;
  const T* rp;
  T omega;
  TZ* x;
  int fix;
  const TZ* zprev;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid gf = this->gf;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    Grid gc = this->gc;
    const T* __restrict__ xc = this->xc;
    const T* __restrict__ invd = this->invd;
    const TR* __restrict__ r = this->r;
    const T* __restrict__ rp = this->rp;
    T omega = this->omega;
    TZ* __restrict__ x = this->x;
    int fix = this->fix;
    const TZ* __restrict__ zprev = this->zprev;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  double d0 = 0.0, d1 = 0.0;
  GMGS_TILE_LOOP(gf, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const int yf = ln_.yp - 1, xf = ln_.xp - 1, zf = ln_.zp - 1;
    const int ya = yf >> 1, yb = (yf + 1) >> 1, xa = xf >> 1, xb = (xf + 1) >> 1, za = zf >> 1,
              zb = (zf + 1) >> 1;
    const T wy = (yf & 1) ? (T)0.5 : (T)1, wx = (xf & 1) ? (T)0.5 : (T)1,
            wz = (zf & 1) ? (T)0.5 : (T)1;
    const long long Nc = gc.N, N = gf.N, n = ln_.n;
    T p[3] = {0, 0, 0};
    const int ys[2] = {ya, yb}, xs[2] = {xa, xb}, zs[2] = {za, zb};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int cy = ys[k & 1], cx = xs[(k >> 1) & 1], cz = zs[k >> 2];
      // odd coordinates average two coarse nodes; even ones take one (weight on 'a')
      const T w = (((k & 1) && !(yf & 1)) ? (T)0 : wy) * ((((k >> 1) & 1) && !(xf & 1)) ? (T)0 : wx)
                  * (((k >> 2) && !(zf & 1)) ? (T)0 : wz);
      const long long m = gc.pidx(cy, cx, cz);
      p[0] += w * xc[m];
      p[1] += w * xc[Nc + m];
      p[2] += w * xc[2 * Nc + m];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T v = omega * invd[c * N + n] * ((T)r[c * N + n] + rp[c * N + n]) + p[c];
      if (fix && (ln_.fl & (1u << (F_FIX_SHIFT + c)))) v = 0;
      x[c * N + n] = (TZ)v;
      if (partial) {
        d0 += (double)r[c * N + n] * (double)(TZ)v;
        if (zprev) d1 += (double)r[c * N + n] * (double)zprev[c * N + n];
      }
    }
  }
  if (partial) {
    const double b0 = block_sum(d0);
    const double b1 = block_sum(d1);
    if (threadIdx.x == 0) {
      partial[block_] = b0;
      partial[nblocks_ + block_] = b1;
    }
  }
  }
};

// e = P xc on the fine level's solid nodes (the prolongation of
// kernel/InterpolationDeviation.m:3-17 on the padded grids), in the order of
// prolong_zplanes / prolong_xy
template <typename T>
struct KProlong {
  static constexpr int kThreads = 256;
  Grid gf;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  Grid gc;
  const T* xc;
  T* e;
  int mask;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid gf = this->gf;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    Grid gc = this->gc;
    const T* __restrict__ xc = this->xc;
    T* __restrict__ e = this->e;
    int mask = this->mask;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  GMGS_TILE_LOOP(gf, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const int yf = ln_.yp - 1, xf = ln_.xp - 1, zf = ln_.zp - 1;
    const int ya = yf >> 1, yb = (yf + 1) >> 1, xa = xf >> 1, xb = (xf + 1) >> 1, za = zf >> 1,
              zb = (zf + 1) >> 1;
    const T wy0 = (yf & 1) ? (T)0.5 : (T)1, wy1 = (yf & 1) ? (T)0.5 : (T)0;
    const T wx0 = (xf & 1) ? (T)0.5 : (T)1, wx1 = (xf & 1) ? (T)0.5 : (T)0;
    const T wz0 = (zf & 1) ? (T)0.5 : (T)1, wz1 = (zf & 1) ? (T)0.5 : (T)0;
    const long long Nc = gc.N, N = gf.N, n = ln_.n;
    const long long m00 = gc.pidx(ya, xa, za), m01 = gc.pidx(ya, xb, za), m10 = gc.pidx(yb, xa, za),
                    m11 = gc.pidx(yb, xb, za), dz = gc.pidx(ya, xa, zb) - m00;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T* __restrict__ v = xc + c * Nc;
      const T z00 = prolong_zplanes(wz0, wz1, v[m00], v[m00 + dz]);
      const T z01 = prolong_zplanes(wz0, wz1, v[m01], v[m01 + dz]);
      const T z10 = prolong_zplanes(wz0, wz1, v[m10], v[m10 + dz]);
      const T z11 = prolong_zplanes(wz0, wz1, v[m11], v[m11 + dz]);
      const T p = prolong_xy(wy0, wy1, wx0, wx1, z00, z01, z10, z11);
      e[c * N + n] = (mask && (ln_.fl & (1u << (F_FIX_SHIFT + c)))) ? (T)0 : p;
    }
  }
  }
};

template <typename T>
struct KMaskFixed {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const T* src;
  T* dst;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const T* __restrict__ src = this->src;
    T* __restrict__ dst = this->dst;
    (void)block_;
    (void)nblocks_;

  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      dst[c * N + ln_.n] = (ln_.fl & (1u << (F_FIX_SHIFT + c))) ? (T)0 : src[c * N + ln_.n];
  }
  }
};

// x = (sum_s w_s (R_s + Q_s)) / d + e on the solid nodes, R_0 = r0, s < NU;
// zero fixed components if 'fix'; partials of r0.x and r0.zprev.  The terms
// are separate restrict arguments so the loads of all three components can
// be issued ahead of the stores.
template <typename T, typename TR, typename TZ, int NU>
struct KCombine {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const T* invd;
  T w0;
  T w1;
  T w2;
  T w3;
  const TR* r0;
  const T* q0;
  const T* r1;
  const T* q1;
  const T* r2;
  const T* q2;
  const T* r3;
  const T* q3;
  const T* e;
  TZ* x;
  int fix;
  const TZ* zprev;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const T* __restrict__ invd = this->invd;
    T w0 = this->w0;
    T w1 = this->w1;
    T w2 = this->w2;
    T w3 = this->w3;
    const TR* __restrict__ r0 = this->r0;
    const T* __restrict__ q0 = this->q0;
    const T* __restrict__ r1 = this->r1;
    const T* __restrict__ q1 = this->q1;
    const T* __restrict__ r2 = this->r2;
    const T* __restrict__ q2 = this->q2;
    const T* __restrict__ r3 = this->r3;
    const T* __restrict__ q3 = this->q3;
    const T* __restrict__ e = this->e;
    TZ* __restrict__ x = this->x;
    int fix = this->fix;
    const TZ* __restrict__ zprev = this->zprev;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  double d0 = 0.0, d1 = 0.0;
  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long n = ln_.n;
    // the sums of all three components before the first store
    T v[3];
    TR r0v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long m = c * N + n;
      r0v[c] = r0[m];
      T acc = w0 * ((T)r0v[c] + q0[m]);
      if (NU > 1) acc += w1 * (r1[m] + q1[m]);
      if (NU > 2) acc += w2 * (r2[m] + q2[m]);
      if (NU > 3) acc += w3 * (r3[m] + q3[m]);
      v[c] = invd[m] * acc + e[m];
      if (fix && (ln_.fl & (1u << (F_FIX_SHIFT + c)))) v[c] = 0;
      if (partial && zprev) d1 += (double)r0v[c] * (double)zprev[m];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c * N + n] = (TZ)v[c];
      if (partial) d0 += (double)r0v[c] * (double)(TZ)v[c];
    }
  }
  if (partial) {
    const double b0 = block_sum(d0);
    const double b1 = block_sum(d1);
    if (threadIdx.x == 0) {
      partial[block_] = b0;
      partial[nblocks_ + block_] = b1;
    }
  }
  }
};

// out = a - out on the solid nodes
template <typename T>
struct KRsub {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const T* a;
  T* out;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const T* __restrict__ a = this->a;
    T* __restrict__ out = this->out;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    T av[3], ov[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      av[c] = a[c * N + ln_.n];
      ov[c] = out[c * N + ln_.n];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c * N + ln_.n] = av[c] - ov[c];
  }
  }
};

template <typename TZ>
struct KUpdateP {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const TZ* z;
  double* p;
  double* x;
  const PcgScalars* sc;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const TZ* __restrict__ z = this->z;
    double* __restrict__ p = this->p;
    double* __restrict__ x = this->x;
    const PcgScalars* __restrict__ sc = this->sc;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  const double beta = sc->beta, alpha = sc->alpha;
  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long n = ln_.n;
    // every load before the first store (the components share the arrays, so
    // the compiler could not move a load above an earlier component's store)
    double po[3], xo[3] = {0, 0, 0};
    TZ zv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      po[c] = p[c * N + n];
      zv[c] = z[c * N + n];
      if (x) xo[c] = x[c * N + n];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (x) x[c * N + n] = xo[c] + alpha * po[c];   // the previous iteration's x update
      p[c * N + n] = (double)zv[c] + beta * po[c];
    }
  }
  }
};

// x += alpha p (the last iteration's deferred x update; runs after the loop)
struct KAddAlphaP {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  double* x;
  const double* p;
  const double* p_odd;
  const PcgScalars* sc;
  int two;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    double* __restrict__ x = this->x;
    const double* __restrict__ p = this->p;
    const double* __restrict__ p_odd = this->p_odd;
    const PcgScalars* __restrict__ sc = this->sc;
    (void)block_;
    (void)nblocks_;

  const double alpha = sc->alpha, alpha2 = sc->alpha_prev;
  // directions alternating between two buffers: iteration k (its = k + 1
  // after it) wrote p[k & 1]; with 'two', an odd last iteration leaves the
  // even one before it pending as well
  const bool odd = ((sc->its - 1) & 1) != 0;
  const double* __restrict__ p2 = (two && odd) ? p : nullptr;
  // with 'two', x is first written by the third iteration's update: after a
  // one- or two-iteration solve it still counts as 0 (the start leaves it)
  const bool x0 = two && sc->its <= 2;
  if (p_odd && odd) p = p_odd;
  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long n = ln_.n;
    double xo[3], po[3], qo[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xo[c] = x0 ? 0.0 : x[c * N + n];
      po[c] = p[c * N + n];
      qo[c] = p2 ? p2[c * N + n] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x1 = p2 ? fma_t(alpha2, qo[c], xo[c]) : xo[c];
      x[c * N + n] = fma_t(alpha, po[c], x1);
    }
  }
  }
};

// Padded index of this lane's node in tile tr (0, a halo entry, for a lane
// beyond the grid's x extent)
__device__ inline long long lane_index(const Grid& g, const TileRec& tr, bool& ok) {
  const int lane = threadIdx.x & 63;
  const int yp = tr.y0 + (lane & 15), xp = tr.x0 + (lane >> 4);
  ok = xp < g.px;
  return ok ? yp + (long long)g.py * (xp + (long long)g.px * tr.z) : 0;
}

// x = (x + omega r / d) + alpha p with the iteration's Jacobi increment
// taken from r before it is replaced (no increment when invd is null; no x
// update at all when x is null); r = r_a - alpha q (r_a may be r itself);
// partials of r.r.  Two tiles per pass of the fixed grid, the loads of both
// issued together and independently of the flags (void and halo entries hold
// zero and read as zero): only the stores and the norm look at the flags.
// Each lane visits its tiles in the same order as with one tile per pass, so
// the partial sums are unchanged.
struct KUpdateXr {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  double* x;
  const double* p;
  const double* ra;
  const double* q;
  double* r;
  const double* invd;
  double omega;
  const PcgScalars* sc;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    double* __restrict__ x = this->x;
    const double* __restrict__ p = this->p;
    const double* ra = this->ra;
    const double* __restrict__ q = this->q;
    double* r = this->r;
    const double* __restrict__ invd = this->invd;
    double omega = this->omega;
    const PcgScalars* __restrict__ sc = this->sc;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  double d = 0.0;
  const double alpha = sc->alpha;
  const long long N = g.N;
  const long long stride = (long long)nblocks_ * 4;
  for (long long t0 = (long long)block_ * 4 + (threadIdx.x >> 6); t0 < nt; t0 += 2 * stride) {
    const long long t1 = t0 + stride;
    const bool has1 = t1 < nt;
    const TileRec tr0 = tiles[t0];
    const TileRec tr1 = tiles[has1 ? t1 : t0];
    bool ok0, ok1;
    const long long n0 = lane_index(g, tr0, ok0);
    long long n1 = lane_index(g, tr1, ok1);
    ok1 = ok1 && has1;
    if (!ok1) n1 = 0;
    const long long nn[2] = {n0, n1};
    const uint32_t fl[2] = {flags[n0], flags[n1]};
    // every load of both tiles before the first store (r_a may be r itself)
    double rav[2][3], qv[2][3], xv[2][3], pv[2][3], uv[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const long long m = c * N + nn[k];
        rav[k][c] = ra[m];
        qv[k][c] = q[m];
        xv[k][c] = pv[k][c] = uv[k][c] = 0.0;
        if (x) {
          xv[k][c] = x[m];
          pv[k][c] = p[m];
          if (invd) uv[k][c] = omega * invd[m] * r[m];
        }
      }
    const bool act[2] = {ok0 && (fl[0] & F_ELEM_MASK) != 0, ok1 && (fl[1] & F_ELEM_MASK) != 0};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!act[k]) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const long long m = c * N + nn[k];
        if (x && invd) x[m] = (xv[k][c] + uv[k][c]) + alpha * pv[k][c];
        else if (x) x[m] = xv[k][c] + alpha * pv[k][c];
        const double rv = rav[k][c] - alpha * qv[k][c];
        r[m] = rv;
        d += rv * rv;
      }
    }
  }
  const double b = block_sum(d);
  if (threadIdx.x == 0) partial[block_] = b;
  }
};

// x += omega r / d on the solid nodes
template <typename T, typename TR>
struct KJacobiAcc {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const T* invd;
  const TR* r;
  T omega;
  TR* x;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const T* __restrict__ invd = this->invd;
    const TR* __restrict__ r = this->r;
    T omega = this->omega;
    TR* __restrict__ x = this->x;
    (void)block_;
    (void)nblocks_;

  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long n = ln_.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c * N + n] += (TR)(omega * invd[c * N + n] * (T)r[c * N + n]);
  }
  }
};

// v = values in [-1, 1) from a fixed integer hash of (padded index, component)
// on the free components of the solid nodes (the lowbias32 mixer of
// C. Wellons' hash-prospector: xorshift-multiply rounds)
template <typename T>
struct KSeed {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  T* v;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    T* __restrict__ v = this->v;
    (void)block_;
    (void)nblocks_;

  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint32_t h = (uint32_t)(3 * ln_.n + c);
      h ^= h >> 16;
      h *= 0x7feb352du;
      h ^= h >> 15;
      h *= 0x846ca68bu;
      h ^= h >> 16;
      const T val = (T)((double)(h >> 8) * (1.0 / 8388608.0) - 1.0);
      v[c * N + ln_.n] = (ln_.fl & (1u << (F_FIX_SHIFT + c))) ? (T)0 : val;
    }
  }
  }
};

// r -= alpha q (q null: r unchanged), z = invd r; partials of r.z
template <typename T>
struct KJcgStep {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const T* invd;
  const T* q;
  double alpha;
  T* r;
  T* z;
  double* partial;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const T* __restrict__ invd = this->invd;
    const T* __restrict__ q = this->q;
    double alpha = this->alpha;
    T* __restrict__ r = this->r;
    T* __restrict__ z = this->z;
    double* __restrict__ partial = this->partial;
    (void)block_;
    (void)nblocks_;

  double d = 0.0;
  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long m = c * N + ln_.n;
      T rv = r[m];
      if (q) {
        rv = (T)((double)rv - alpha * (double)q[m]);
        r[m] = rv;
      }
      const T zv = invd[m] * rv;
      z[m] = zv;
      d += (double)rv * (double)zv;
    }
  }
  const double b = block_sum(d);
  if (threadIdx.x == 0) partial[block_] = b;
  }
};

// p = z + beta p (beta = 0: p = z)
template <typename T>
struct KXpby {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const T* z;
  double beta;
  T* p;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const T* __restrict__ z = this->z;
    double beta = this->beta;
    T* __restrict__ p = this->p;
    (void)block_;
    (void)nblocks_;

  const long long N = g.N;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long m = c * N + ln_.n;
      p[m] = beta == 0.0 ? z[m] : (T)((double)z[m] + beta * (double)p[m]);
    }
  }
  }
};

struct KDot {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const double* a;
  const double* b;
  double* partial;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const double* __restrict__ a = this->a;
    const double* __restrict__ b = this->b;
    double* __restrict__ partial = this->partial;
    (void)block_;
    (void)nblocks_;

  double d = 0.0;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long N = g.N, n = ln_.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) d += a[c * N + n] * b[c * N + n];
  }
  const double s = block_sum(d);
  if (threadIdx.x == 0) partial[block_] = s;
  }
};

// start of a solve (CG_solver.m:11-13 with x0 = 0): x = p = 0, z0 = z1 = 0,
// r = b, partials of b.b; solid nodes only (void entries are never written
// and stay zero)
template <typename TZ>
struct KPcgStart {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const double* b;
  double* x;
  double* p;
  double* p2;
  double* r;
  TZ* z0;
  TZ* z1;
  double* partial;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const double* __restrict__ b = this->b;
    double* __restrict__ x = this->x;
    double* __restrict__ p = this->p;
    double* __restrict__ p2 = this->p2;
    double* __restrict__ r = this->r;
    TZ* __restrict__ z0 = this->z0;
    TZ* __restrict__ z1 = this->z1;
    double* __restrict__ partial = this->partial;
    (void)block_;
    (void)nblocks_;

  double d = 0.0;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long N = g.N, n = ln_.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long m = c * N + n;
      const double v = b[m];
      if (x) x[m] = 0.0;
      if (p) p[m] = 0.0;
      if (p2) p2[m] = 0.0;
      r[m] = v;
      if (z0) z0[m] = (TZ)0;
      if (z1) z1[m] = (TZ)0;
      d += v * v;
    }
  }
  const double s = block_sum(d);
  if (threadIdx.x == 0) partial[block_] = s;
  }
};

// partials of |b - q|^2 (the true residual of x when q = A x)
struct KResidSq {
  static constexpr int kThreads = 256;
  Grid g;
  const uint32_t* flags;
  const TileRec* tiles;
  long long nt;
  const double* b;
  const double* q;
  double* partial;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    Grid g = this->g;
    const uint32_t* __restrict__ flags = this->flags;
    const TileRec* __restrict__ tiles = this->tiles;
    long long nt = this->nt;
    const double* __restrict__ b = this->b;
    const double* __restrict__ q = this->q;
    double* __restrict__ partial = this->partial;
    (void)block_;
    (void)nblocks_;

  double d = 0.0;
  GMGS_TILE_LOOP(g, flags, tiles, nt, ln_) {
    if (!ln_.active) continue;
    const long long N = g.N, n = ln_.n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double e = b[c * N + n] - q[c * N + n];
      d += e * e;
    }
  }
  const double s = block_sum(d);
  if (threadIdx.x == 0) partial[block_] = s;
  }
};

// one block of kFinalizeThreads threads; fixed-order sums (thread t adds
// part[t], part[t + T], ... in turn; the loads are issued eight at a time so
// the chain of adds does not wait for each load in turn)
constexpr int kFinalizeThreads = 1024;
__device__ inline double reduce_parts(const double* __restrict__ part, int n) {
  constexpr int T = kFinalizeThreads;
  double v = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * T < n; i += 8 * T) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = part[i + u * T];
#pragma unroll
    for (int u = 0; u < 8; ++u) v += a[u];
  }
  for (; i < n; i += T) v += part[i];
  // the wave's sum by shuffles, then the block's 16 wave sums in order
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __shared__ double sh[T / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < T / 64; ++w) r += sh[w];
  __syncthreads();
  return r;
}

struct KFinalize {
  static constexpr int kThreads = kFinalizeThreads;
  int which;
  const double* part;
  int nparts;
  PcgScalars* sc;
  double* hist_rel;
  unsigned long long* hist_t;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    int which = this->which;
    const double* __restrict__ part = this->part;
    int nparts = this->nparts;
    PcgScalars* __restrict__ sc = this->sc;
    double* __restrict__ hist_rel = this->hist_rel;
    unsigned long long* __restrict__ hist_t = this->hist_t;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  if (which == 0) {
    const double s = reduce_parts(part, nparts);
    if (threadIdx.x == 0) {
      sc->normB = sqrt(s);
      sc->its = 0;
      sc->done = 0;
      sc->converged = 0;
      sc->rz_prev = 0.0;
      sc->t0 = wall_clock64();
    }
  } else if (which == 1) {
    const double rz = reduce_parts(part, nparts);
    const double rzo = reduce_parts(part + nparts, nparts);
    if (threadIdx.x == 0) {
      sc->rz = rz;
      sc->rz_old = rzo;
      sc->its += 1;
      if (sc->its == 1) sc->beta = 0.0;
      else if (sc->flexible) sc->beta = (rz - rzo) / sc->rz_prev;   // Polak-Ribiere
      else sc->beta = rz / sc->rz_prev;                             // CG_solver.m:26
    }
  } else if (which == 2) {
    const double pq = reduce_parts(part, nparts);
    if (threadIdx.x == 0) {
      sc->pq = pq;
      sc->alpha_prev = sc->alpha;
      sc->alpha = sc->rz / pq;                                      // CG_solver.m:31
    }
  } else {
    const double rr = reduce_parts(part, nparts);
    // (its = 0: the residual update fused into the next iteration's first
    // sweep has nothing to test in the first iteration, r = b)
    if (threadIdx.x == 0 && sc->its > 0) {
      sc->rr = rr;
      const double rel = sqrt(rr) / sc->normB;                      // CG_solver.m:35
      sc->relres = rel;
      const int k = sc->its - 1;
      hist_rel[k] = rel;
      hist_t[k] = wall_clock64();
      if (rel < sc->tol) {
        sc->converged = 1;
        sc->done = 1;
      } else if (!(rel == rel) || sc->its >= sc->max_it + 1) {
        sc->done = 1;
      }
      sc->rz_prev = sc->rz;
    }
  }
  }
};

template <typename TS, typename TD>
__global__ void k_convert(const TS* __restrict__ a, TD* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (TD)a[i];
}

template <typename T>
__global__ void k_fill(T* __restrict__ a, T v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ c, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    c[i] = a[i] - b[i];
}

__global__ void k_scatter_compact(const long long* __restrict__ pidx, long long nn,
                                  const double* __restrict__ comp, double* __restrict__ dense,
                                  long long N) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const long long p = pidx[k];
  dense[p] = comp[3 * k];
  dense[N + p] = comp[3 * k + 1];
  dense[2 * N + p] = comp[3 * k + 2];
}

template <typename T>
__global__ void k_gather_compact(const long long* __restrict__ pidx, long long nn,
                                 const T* __restrict__ dense, long long N, double* __restrict__ comp) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const long long p = pidx[k];
  comp[3 * k] = (double)dense[p];
  comp[3 * k + 1] = (double)dense[N + p];
  comp[3 * k + 2] = (double)dense[2 * N + p];
}

__global__ __launch_bounds__(256) void k_stress(Grid g, const long long* __restrict__ npidx,
                                              long long nn, const uint32_t* __restrict__ flags,
                                              const double* __restrict__ M,
                                              const double* __restrict__ emod,
                                              const double* __restrict__ u,
                                              double* __restrict__ S) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const long long n = npidx[k], N = g.N;
  const uint32_t fl = flags[n];
  const long long sx = g.py, sz = (long long)g.py * g.px;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int p = 0; p < 8; ++p) {
    if (!((fl >> p) & 1u)) continue;
    const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = p >> 2;
    const long long cell = n - (1 - pyb) - (1 - pxb) * sx - (1 - pzb) * sz;
    const int ln = node_local_in(p);
    double e[6] = {0, 0, 0, 0, 0, 0};
    for (int m = 0; m < 8; ++m) {
      const long long mm = cell + hex_ro(m) + hex_xo(m) * sx + hex_zo(m) * sz;
      const double u0 = u[mm], u1 = u[N + mm], u2 = u[2 * N + mm];
      for (int a = 0; a < 6; ++a) {
        const double* r = M + (6 * ln + a) * 24 + 3 * m;
        e[a] += r[0] * u0 + r[1] * u1 + r[2] * u2;
      }
    }
    const double s = emod ? emod[cell] : 1.0;
    for (int a = 0; a < 6; ++a) acc[a] += s * e[a];
  }
  const double cnt = (double)__popc(fl & F_ELEM_MASK);
  for (int a = 0; a < 6; ++a) S[k + nn * a] = acc[a] / cnt;
}

__global__ __launch_bounds__(256) void k_element_energy(Grid g, const long long* __restrict__ epidx,
                                                        long long ne, const double* __restrict__ Ke,
                                                        const double* __restrict__ lam,
                                                        const double* __restrict__ u,
                                                        double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  const long long cell = epidx[k], N = g.N;
  const long long sx = g.py, sz = (long long)g.py * g.px;
  double le[24], ue[24];
  for (int m = 0; m < 8; ++m) {
    const long long mm = cell + hex_ro(m) + hex_xo(m) * sx + hex_zo(m) * sz;
    for (int d = 0; d < 3; ++d) {
      le[3 * m + d] = lam[d * N + mm];
      ue[3 * m + d] = u[d * N + mm];
    }
  }
  double acc = 0.0;
  for (int i = 0; i < 24; ++i) {
    double t = 0.0;
    for (int j = 0; j < 24; ++j) t += Ke[i * 24 + j] * ue[j];
    acc += le[i] * t;
  }
  out[k] = -acc;
}

__global__ __launch_bounds__(256) void k_copy_probe(const double* __restrict__ a,
                                                    double* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ void k_gather_flags(const int* const* __restrict__ flags, int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = *flags[i];
}

}  // namespace

void launch_copy_probe(const double* a, double* b, long long n_doubles, hipStream_t s) {
  // one double per lane, the access width of the solver's kernels
  k_copy_probe<<<(unsigned)std::min<long long>((n_doubles + 255) / 256, 65536), 256, 0, s>>>(
      a, b, n_doubles);
  GMGS_LAUNCH_CHECK();
}

void launch_gather_flags(const int* const* flags, int n, int* out, hipStream_t s) {
  k_gather_flags<<<(n + 255) / 256, 256, 0, s>>>(flags, n, out);
  GMGS_LAUNCH_CHECK();
}

void launch_stress(const Grid& g, const long long* node_pidx, long long nn, const uint32_t* flags,
                   const double* M, const double* emod, const double* u, double* S, hipStream_t s) {
  if (nn <= 0) return;
  k_stress<<<grid1d(nn, 256), 256, 0, s>>>(g, node_pidx, nn, flags, M, emod, u, S);
  GMGS_LAUNCH_CHECK();
}

void launch_element_energy(const Grid& g, const long long* elem_pidx, long long ne, const double* Ke,
                           const double* lam, const double* u, double* out, hipStream_t s) {
  if (ne <= 0) return;
  k_element_energy<<<grid1d(ne, 256), 256, 0, s>>>(g, elem_pidx, ne, Ke, lam, u, out);
  GMGS_LAUNCH_CHECK();
}

template <typename T, typename TR>
void launch_apply(const LevelOp<T>& op, const TileRec* tiles, long long ntiles, int mode,
                  const TR* in, T omega, T* out, TR* xacc, double* partial, const int* done,
                  hipStream_t s) {
  if (ntiles <= 0) return;
  const unsigned nb = apply_blocks(ntiles);
  const bool eop = op.eop != nullptr, emod = op.emod != nullptr;
  if (op.asc && !xacc && mode != AP_RESID) {
    if (mode == AP_AX)
      launch(KApplyAs<T, TR, AP_AX>{op, tiles, ntiles, in, omega, nullptr, out, partial,
                                                          done}, nb, s);
    else
      launch(KApplyAs<T, TR, AP_SMOOTH>{op, tiles, ntiles, in, omega, nullptr, out,
                                                              partial, done}, nb, s);
    GMGS_LAUNCH_CHECK();
    return;
  }
#define GMGS_APPLY(MODE, E, M) \
  launch(KApply<T, TR, MODE, E, M>{op, tiles, ntiles, in, omega, out, xacc, partial, done}, nb, s)
  if (mode == AP_AX) {
    if (eop) GMGS_APPLY(AP_AX, true, false);
    else if (emod) GMGS_APPLY(AP_AX, false, true);
    else GMGS_APPLY(AP_AX, false, false);
  } else {
    if (eop) GMGS_APPLY(AP_SMOOTH, true, false);
    else if (emod) GMGS_APPLY(AP_SMOOTH, false, true);
    else GMGS_APPLY(AP_SMOOTH, false, false);
  }
#undef GMGS_APPLY
  GMGS_LAUNCH_CHECK();
}

template <typename T, typename TR>
int launch_apply_rows(const LevelOp<T>& op, const RowTile* boxes, const RowEntry* entries, long long nbox,
                      int mode, const TR* in, T omega, const T* in2, T* out, double* partial, const int* done,
                      hipStream_t s) {
  if (nbox <= 0) return 0;
  if (!op.asc) throw std::invalid_argument("apply_rows: no assembled rows");
  if (3 * op.g.N >= (1LL << 31)) throw std::invalid_argument("apply_rows: grid too large");
  const unsigned nb = apply_blocks(nbox);
  if (mode == AP_AX)
    launch(KApplyRows<T, TR, AP_AX>{op, boxes, entries, nbox, in, omega, in2, out, partial, done}, nb, s);
  else if (mode == AP_RESID)
    launch(KApplyRows<T, TR, AP_RESID>{op, boxes, entries, nbox, in, omega, in2, out, partial,
                                                     done}, nb, s);
  else if (mode == AP_SMOOTH)
    launch(KApplyRows<T, TR, AP_SMOOTH>{op, boxes, entries, nbox, in, omega, in2, out, partial,
                                                      done}, nb, s);
  else
    throw std::invalid_argument("apply_rows: mode");
  GMGS_LAUNCH_CHECK();
  return (int)nb;
}
template int launch_apply_rows<double, double>(const LevelOp<double>&, const RowTile*, const RowEntry*,
                                               long long, int, const double*, double, const double*, double*,
                                               double*, const int*, hipStream_t);
template int launch_apply_rows<float, float>(const LevelOp<float>&, const RowTile*, const RowEntry*, long long,
                                             int, const float*, float, const float*, float*, double*,
                                             const int*, hipStream_t);
template int launch_apply_rows<float, double>(const LevelOp<float>&, const RowTile*, const RowEntry*,
                                              long long, int, const double*, float, const float*, float*,
                                              double*, const int*, hipStream_t);

template <typename T>
void launch_relayout_rows(const double* src, long long rows, const long long* rep, long long nu, T* dst,
                          hipStream_t s) {
  const long long n = (long long)kStencilCoefs * nu;
  if (n <= 0) return;
  k_relayout_rows<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, rows, rep, nu, dst);
  GMGS_LAUNCH_CHECK();
}
template void launch_relayout_rows<double>(const double*, long long, const long long*, long long, double*,
                                           hipStream_t);
template void launch_relayout_rows<float>(const double*, long long, const long long*, long long, float*,
                                          hipStream_t);

void launch_row_hash(const double* src, long long rows, unsigned long long* h, hipStream_t s) {
  if (rows <= 0) return;
  k_row_hash<<<(unsigned)((rows + 255) / 256), 256, 0, s>>>(src, rows, h);
  GMGS_LAUNCH_CHECK();
}

void launch_row_verify(const double* src, long long rows, const int* uid, const long long* rep, int* bad,
                       hipStream_t s) {
  const long long n = (long long)kStencilCoefs * rows;
  if (n <= 0) return;
  k_row_verify<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, rows, uid, rep, bad);
  GMGS_LAUNCH_CHECK();
}

void launch_assemble_stencil(const LevelOp<double>& op, const TileRec* tiles, long long ntiles,
                             const int* arow, long long stride, double* asc, hipStream_t s) {
  if (ntiles <= 0) return;
  const unsigned nb = apply_blocks(ntiles);
  if (op.eop && op.emod) throw std::invalid_argument("assemble_stencil: eop and emod together");
  if (op.eop) launch(KAssembleStencil<true, false>{op, tiles, ntiles, arow, stride, asc}, nb, s);
  else if (op.emod) launch(KAssembleStencil<false, true>{op, tiles, ntiles, arow, stride, asc}, nb, s);
  else launch(KAssembleStencil<false, false>{op, tiles, ntiles, arow, stride, asc}, nb, s);
  GMGS_LAUNCH_CHECK();
}

template <typename T>
void launch_restrict(const Grid& gc, const uint32_t* flags_c, const TileRec* tiles_c, long long nt,
                     const Grid& gf, const T* rf, T* rc, const int* done, hipStream_t s) {
  if (nt <= 0) return;
  launch(KRestrict<T>{gc, flags_c, tiles_c, nt, gf, rf, rc, done}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}

template <typename T, typename TR, typename TZ>
void launch_prolong_smooth(const Grid& gf, const uint32_t* flags_f, const TileRec* tiles, long long nt,
                           const Grid& gc, const T* xc, const T* invd, const TR* r, const T* rp,
                           T omega, TZ* x, bool fix, const TZ* zprev, double* partial,
                           const int* done, hipStream_t s) {
  if (nt <= 0) return;
  launch(KProlongSmooth<T, TR, TZ>{gf, flags_f, tiles, nt, gc, xc, invd, r,
                                                               rp, omega, x, fix ? 1 : 0, zprev,
                                                               partial, done}, grid_for(nt, partial), s);
  GMGS_LAUNCH_CHECK();
}

template <typename T>
void launch_rsub(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, const T* a,
                 T* out, const int* done, hipStream_t s) {
  if (nt <= 0) return;
  launch(KRsub<T>{g, flags, tiles, nt, a, out, done}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template void launch_rsub<double>(const Grid&, const uint32_t*, const TileRec*, long long, const double*,
                                  double*, const int*, hipStream_t);
template void launch_rsub<float>(const Grid&, const uint32_t*, const TileRec*, long long, const float*,
                                 float*, const int*, hipStream_t);

template <typename T>
void launch_prolong(const Grid& gf, const uint32_t* flags_f, const TileRec* tiles, long long nt,
                    const Grid& gc, const T* xc, T* e, bool mask, const int* done, hipStream_t s) {
  if (nt <= 0) return;
  launch(KProlong<T>{gf, flags_f, tiles, nt, gc, xc, e, mask ? 1 : 0, done}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template void launch_prolong<double>(const Grid&, const uint32_t*, const TileRec*, long long, const Grid&,
                                     const double*, double*, bool, const int*, hipStream_t);
template void launch_prolong<float>(const Grid&, const uint32_t*, const TileRec*, long long, const Grid&,
                                    const float*, float*, bool, const int*, hipStream_t);

template <typename T>
void launch_mask_fixed(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                       const T* src, T* dst, hipStream_t s) {
  if (nt <= 0) return;
  launch(KMaskFixed<T>{g, flags, tiles, nt, src, dst}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template void launch_mask_fixed<double>(const Grid&, const uint32_t*, const TileRec*, long long,
                                        const double*, double*, hipStream_t);
template void launch_mask_fixed<float>(const Grid&, const uint32_t*, const TileRec*, long long,
                                       const float*, float*, hipStream_t);

template <typename T, typename TR, typename TZ>
void launch_combine(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                    const T* invd, const CycleSum<T>& cs, const TR* r0, const T* e, TZ* x, bool fix,
                    const TZ* zprev, double* partial, const int* done, hipStream_t s) {
  if (nt <= 0) return;
  const unsigned nb = grid_for(nt, partial);
  const int f = fix ? 1 : 0;
#define GMGS_COMBINE(NU)                                                                            \
  launch(KCombine<T, TR, TZ, NU>{g, flags, tiles, nt, invd, cs.w[0], cs.w[1], cs.w[2],         \
                                 cs.w[3], r0, cs.Q[0], cs.R[1], cs.Q[1], cs.R[2], cs.Q[2], cs.R[3], \
                                 cs.Q[3], e, x, f, zprev, partial, done},                           \
         nb, s)
  switch (cs.nu) {
    case 1: GMGS_COMBINE(1); break;
    case 2: GMGS_COMBINE(2); break;
    case 3: GMGS_COMBINE(3); break;
    case 4: GMGS_COMBINE(4); break;
    default: throw std::invalid_argument("combine: sweeps out of range");
  }
#undef GMGS_COMBINE
  GMGS_LAUNCH_CHECK();
}
#define GMGS_INST_COMBINE(T, TR, TZ)                                                                \
  template void launch_combine<T, TR, TZ>(const Grid&, const uint32_t*, const TileRec*, long long,  \
                                          const T*, const CycleSum<T>&, const TR*, const T*, TZ*,   \
                                          bool, const TZ*, double*, const int*, hipStream_t);
GMGS_INST_COMBINE(double, double, double)
GMGS_INST_COMBINE(float, float, float)
GMGS_INST_COMBINE(float, double, float)
#undef GMGS_INST_COMBINE

template <typename TZ>
void launch_update_p(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                     const TZ* z, double* p, const PcgScalars* sc, const int* done, hipStream_t s,
                     double* x) {
  launch(KUpdateP<TZ>{g, flags, tiles, nt, z, p, x, sc, done}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}

void launch_add_alpha_p(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, double* x,
                        const double* p, const double* p_odd, const PcgScalars* sc, hipStream_t s, bool two) {
  if (nt <= 0) return;
  if (two && !p_odd) throw std::invalid_argument("add_alpha_p: the two-step update needs both direction buffers");
  launch(KAddAlphaP{g, flags, tiles, nt, x, p, p_odd, sc, two ? 1 : 0}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}

void launch_update_xr(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                      double* x, const double* p, const double* ra, const double* q, double* r,
                      const double* invd, double omega, const PcgScalars* sc, double* partial,
                      const int* done, hipStream_t s) {
  launch(KUpdateXr{g, flags, tiles, nt, x, p, ra, q, r, invd, omega, sc,
                                                   partial, done}, pointwise_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}

template <typename T, typename TR>
void launch_jacobi_acc(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                       const T* invd, const TR* r, T omega, TR* x, hipStream_t s) {
  if (nt <= 0) return;
  launch(KJacobiAcc<T, TR>{g, flags, tiles, nt, invd, r, omega, x}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template void launch_jacobi_acc<double, double>(const Grid&, const uint32_t*, const TileRec*, long long,
                                                const double*, const double*, double, double*,
                                                hipStream_t);
template void launch_jacobi_acc<float, float>(const Grid&, const uint32_t*, const TileRec*, long long,
                                              const float*, const float*, float, float*, hipStream_t);
template void launch_jacobi_acc<float, double>(const Grid&, const uint32_t*, const TileRec*, long long,
                                               const float*, const double*, float, double*, hipStream_t);

template <typename T>
void launch_seed(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, T* v,
                 hipStream_t s) {
  if (nt <= 0) return;
  launch(KSeed<T>{g, flags, tiles, nt, v}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template <typename T>
void launch_jcg_step(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                     const T* invd, const T* q, double alpha, T* r, T* z, double* partial,
                     hipStream_t s) {
  if (nt <= 0) return;
  launch(KJcgStep<T>{g, flags, tiles, nt, invd, q, alpha, r, z, partial}, pointwise_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template <typename T>
void launch_xpby(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, const T* z,
                 double beta, T* p, hipStream_t s) {
  if (nt <= 0) return;
  launch(KXpby<T>{g, flags, tiles, nt, z, beta, p}, apply_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
#define GMGS_INST_LANCZOS(T)                                                                          \
  template void launch_seed<T>(const Grid&, const uint32_t*, const TileRec*, long long, T*, hipStream_t); \
  template void launch_jcg_step<T>(const Grid&, const uint32_t*, const TileRec*, long long, const T*,   \
                                   const T*, double, T*, T*, double*, hipStream_t);                     \
  template void launch_xpby<T>(const Grid&, const uint32_t*, const TileRec*, long long, const T*, double, \
                               T*, hipStream_t);
GMGS_INST_LANCZOS(double)
GMGS_INST_LANCZOS(float)
#undef GMGS_INST_LANCZOS

void launch_dot(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                const double* a, const double* b, double* partial, hipStream_t s) {
  launch(KDot{g, flags, tiles, nt, a, b, partial}, pointwise_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}

template <typename TZ>
void launch_pcg_start(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                      const double* b, double* x, double* p, double* p2, double* r, TZ* z0, TZ* z1,
                      double* partial, hipStream_t s) {
  launch(KPcgStart<TZ>{g, flags, tiles, nt, b, x, p, p2, r, z0, z1, partial}, pointwise_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}
template void launch_pcg_start<double>(const Grid&, const uint32_t*, const TileRec*, long long,
                                       const double*, double*, double*, double*, double*, double*,
                                       double*, double*, hipStream_t);
template void launch_pcg_start<float>(const Grid&, const uint32_t*, const TileRec*, long long,
                                      const double*, double*, double*, double*, double*, float*, float*,
                                      double*, hipStream_t);

void launch_resid_sq(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                     const double* b, const double* q, double* partial, hipStream_t s) {
  launch(KResidSq{g, flags, tiles, nt, b, q, partial}, pointwise_blocks(nt), s);
  GMGS_LAUNCH_CHECK();
}

void launch_finalize(int which, const double* partial, int nparts, PcgScalars* sc,
                     double* hist_relres, unsigned long long* hist_t, const int* done,
                     hipStream_t s) {
  launch(KFinalize{which, partial, nparts, sc, hist_relres, hist_t, done}, 1, s);
  GMGS_LAUNCH_CHECK();
}

template <typename TS, typename TD>
void launch_convert(const TS* a, TD* b, long long n, hipStream_t s) {
  if (n <= 0) return;
  unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  k_convert<TS, TD><<<nb, 256, 0, s>>>(a, b, n);
  GMGS_LAUNCH_CHECK();
}

template <typename T>
void launch_fill(T* a, T v, long long n, hipStream_t s) {
  if (n <= 0) return;
  unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  k_fill<T><<<nb, 256, 0, s>>>(a, v, n);
  GMGS_LAUNCH_CHECK();
}

void launch_sub(const double* a, const double* b, double* c, long long n, hipStream_t s) {
  if (n <= 0) return;
  unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  k_sub<<<nb, 256, 0, s>>>(a, b, c, n);
  GMGS_LAUNCH_CHECK();
}

void launch_scatter_compact(const long long* node_pidx, long long nn, const double* comp,
                            double* dense, long long N, hipStream_t s) {
  if (nn <= 0) return;
  k_scatter_compact<<<grid1d(nn, 256), 256, 0, s>>>(node_pidx, nn, comp, dense, N);
  GMGS_LAUNCH_CHECK();
}

template <typename T>
void launch_gather_compact(const long long* node_pidx, long long nn, const T* dense, long long N,
                           double* comp, hipStream_t s) {
  if (nn <= 0) return;
  k_gather_compact<T><<<grid1d(nn, 256), 256, 0, s>>>(node_pidx, nn, dense, N, comp);
  GMGS_LAUNCH_CHECK();
}

// explicit instantiations
template void launch_apply<double, double>(const LevelOp<double>&, const TileRec*, long long, int,
                                           const double*, double, double*, double*, double*,
                                           const int*, hipStream_t);
template void launch_apply<float, float>(const LevelOp<float>&, const TileRec*, long long, int,
                                         const float*, float, float*, float*, double*, const int*,
                                         hipStream_t);
template void launch_apply<float, double>(const LevelOp<float>&, const TileRec*, long long, int,
                                          const double*, float, float*, double*, double*,
                                          const int*, hipStream_t);
template void launch_restrict<double>(const Grid&, const uint32_t*, const TileRec*, long long,
                                      const Grid&, const double*, double*, const int*, hipStream_t);
template void launch_restrict<float>(const Grid&, const uint32_t*, const TileRec*, long long,
                                     const Grid&, const float*, float*, const int*, hipStream_t);
template void launch_prolong_smooth<double, double, double>(
    const Grid&, const uint32_t*, const TileRec*, long long, const Grid&, const double*,
    const double*, const double*, const double*, double, double*, bool, const double*, double*,
    const int*, hipStream_t);
template void launch_prolong_smooth<float, float, float>(
    const Grid&, const uint32_t*, const TileRec*, long long, const Grid&, const float*, const float*,
    const float*, const float*, float, float*, bool, const float*, double*, const int*, hipStream_t);
template void launch_prolong_smooth<float, double, float>(
    const Grid&, const uint32_t*, const TileRec*, long long, const Grid&, const float*, const float*,
    const double*, const float*, float, float*, bool, const float*, double*, const int*, hipStream_t);
template void launch_update_p<double>(const Grid&, const uint32_t*, const TileRec*, long long,
                                      const double*, double*, const PcgScalars*, const int*,
                                      hipStream_t, double*);
template void launch_update_p<float>(const Grid&, const uint32_t*, const TileRec*, long long,
                                     const float*, double*, const PcgScalars*, const int*,
                                     hipStream_t, double*);
template void launch_convert<double, float>(const double*, float*, long long, hipStream_t);
template void launch_convert<float, double>(const float*, double*, long long, hipStream_t);
template void launch_fill<double>(double*, double, long long, hipStream_t);
template void launch_fill<float>(float*, float, long long, hipStream_t);
template void launch_fill<int>(int*, int, long long, hipStream_t);
template void launch_gather_compact<double>(const long long*, long long, const double*, long long,
                                            double*, hipStream_t);
template void launch_gather_compact<float>(const long long*, long long, const float*, long long,
                                           double*, hipStream_t);

}  // namespace gmgs
