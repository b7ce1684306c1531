// z-marching level operator (levels with a universal element matrix).
//
// A workgroup owns a 16 (y) x TX (x) column of nodes and walks a chunk of
// z-planes.  Each input plane (with a one-node halo) is staged once in LDS;
// every lane reads its 3x3 in-plane neighbourhood from LDS and adds that
// plane's contribution to the three outputs it touches (z-1, z, z+1), so a
// value is fetched from memory about once per plane instead of 27 times.
//
// Void nodes hold zero in every vector, so a staged value outside its
// column's solid z range is not loaded.  Tiles are numbered y-fastest and the
// hardware deals consecutive workgroups to the 8 XCDs in turn; the block ->
// tile map below gives each XCD a contiguous run of tiles instead, so the
// halo rows a tile shares with its neighbours are read while they are still
// in that XCD's L2.
//
// Every solid node gets the stencil of the universal element matrix (its 30
// class coefficients).  A node that misses an incident element, or touches a
// stored (non-universal) element, is then corrected element by element; the
// wave gathers those (node, element) pairs and spreads them over its lanes,
// so the few boundary nodes of a wave do not serialise it.  Together: the
// per-element sum of kernel/AtX.m:1-58.
#include <type_traits>

#include "ops.h"
#include "ops_zm.h"
#include "launch.h"
#include "mesh.h"


namespace gmgs {

namespace {


// Coefficients are read through a constant-address-space pointer kept opaque
// to loop-invariant code motion: every plane re-reads them with scalar loads
// (scalar cache) into SGPR operands instead of hoisting hundreds of values
// into vector registers.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

// LDS written by some lanes of a wave and read by others of the same wave:
// LDS operations of one wave complete in order
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

// E: 0 = universal elements only (corrections for missing elements), 1 =
// also stored elements, 2 = F_FAST nodes only (no corrections; the other
// solid nodes are left to the assembled-row kernel)
template <typename T, typename TR, int MODE, int E, int TX>
struct KApplyZm {
  static constexpr int kThreads = 16 * TX;
  LevelOp<T> op;
  const ColTile* tiles;
  int zc;
  const TR* in;
  T omega;
  const T* in2;
  T* out;
  double* partial;
  const int* done;
  CycleOut<T> co;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    LevelOp<T> op = this->op;
    const ColTile* __restrict__ tiles = this->tiles;
    int zc = this->zc;
    const TR* __restrict__ in = this->in;
    T omega = this->omega;
    const T* __restrict__ in2 = this->in2;
    T* __restrict__ out = this->out;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    CycleOut<T> co = this->co;
    (void)block_;
    (void)nblocks_;

  constexpr int TY = 16, NT = TY * TX, SY = TY + 2, SX = TX + 2;
  constexpr int NSTAGE = 3 * SX * SY;
  constexpr int PER = (NSTAGE + NT - 1) / NT;      // staged values per thread
  constexpr bool PRO = MODE == AP_RESID_SUM_P;       // e = P x_c formed while staged
  constexpr bool SUM = MODE == AP_SMOOTH_SUM, RSUM = MODE == AP_RESID_SUM || PRO;
  constexpr bool RUP = MODE == AP_SMOOTH_R;         // r = in - alpha q staged, written and summed
  constexpr bool SM = MODE == AP_SMOOTH || SUM || RUP;
  constexpr bool AXP = MODE == AP_AX_P || MODE == AP_AX_P_NOX;   // p = z + beta p_old staged
  constexpr bool XUP = MODE == AP_AX_P;              // AX_P: x takes the two deferred updates
  constexpr bool TWO = SUM || RSUM;                 // two dot partials (r0.z, r0.zprev)
  constexpr int ND = (SM || AXP) ? PER : 1;          // second staged operand: 1/d or z
  using TD = typename std::conditional<AXP, float, T>::type;
  // PRO: the coarse patch a staged plane draws on (CY x CX coarse nodes,
  // three components each) in planes (z - 1) >> 1 and z >> 1 of x_c; a thread
  // loads both planes of its share of the patch, PQ positions; the staged
  // positions' flags replace the input's loads
  constexpr int CY = TY / 2 + 2, CX = TX / 2 + 2, NQ = 3 * CY * CX;
  constexpr int PQ = PRO ? (NQ + NT - 1) / NT : 1, PC = 2 * PQ, PF = PRO ? PER : 1;
  constexpr bool HAS_EOP = E == 1, FAST_ONLY = E == 2;
  // output nodes of this kernel
  constexpr uint32_t OUT_MASK = FAST_ONLY ? F_FAST : F_ELEM_MASK;
  if (done && *done) return;
  // four plane buffers: plane k (counted from the tile's first) is staged in
  // buffer k % 4 and read while computing planes k, k + 1 and k + 2 (stencil,
  // output completion, element corrections)
  __shared__ T sU[4][3][SX][SY];
  // SMOOTH: centre values of r (SMOOTH_R: the FP64 r); AX_P: of p_old
  using TC = typename std::conditional<RUP, TR, T>::type;
  __shared__ TC sR[(SM || XUP) ? 2 : 1][3][TX][TY];
  __shared__ T sC[PRO ? NQ : 1];                     // PRO: the patch, its two planes combined
  // per wave: the element corrections' task list and results
  __shared__ int sTask[FAST_ONLY ? 1 : NT / 64][64];
  __shared__ T sRes[FAST_ONLY ? 1 : NT / 64][3][64];
  const int wv = (int)(threadIdx.x >> 6);
  // XCD-contiguous tile order: blocks b = 8 j + k (k = XCD) of the first
  // 8 floor(n / 8) blocks take tile k floor(n / 8) + j; the rest keep their own
  const int nb = (int)nblocks_, q8 = nb >> 3, bid = (int)block_;
  const int ti = bid < (q8 << 3) ? (bid & 7) * q8 + (bid >> 3) : bid;
  const ColTile t = tiles[ti];
  const int tid = threadIdx.x, ly = tid % TY, lx = tid / TY, lane = tid & 63;
  const int lx_w = (tid >> 6) * (64 / TY);           // lx of the wave's first lane
  const Grid g = op.g;
  const long long N = g.N, sx = g.py, sz = (long long)g.py * g.px;
  const int yp = t.y0 + ly, xp = t.x0 + lx;
  const bool lane_ok = xp < g.px;
  const long long colidx = yp + sx * xp;
  const int z0 = t.z0, z1 = min(t.z0 + zc, g.pz - 1);   // outputs [z0, z1)
  // This thread's share of a staged plane: global offset (without z) and
  // that column's solid z range (lo in the low 16 bits, hi - lo above); a
  // position outside the grid or on a column without solid nodes gets the
  // range lo = 0xffff, hi - lo = 0, which no plane index matches.  The
  // column ranges (one per staged value and this lane's own) are all loaded
  // before the first is used: one round trip at the start of the tile, not
  // one per value.  Column 0 is a halo column without solid nodes (lo =
  // 0xffff > hi = 0), so a position outside the grid loads from there and
  // comes out empty, branch-free.
  int scol[PER];
  int zr[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    const int rem = e % (SX * SY), xx = rem / SY, yy = rem % SY;
    const int gy = t.y0 - 1 + yy, gx = t.x0 - 1 + xx;
    const bool in = e < NSTAGE && gy >= 0 && gy < g.py && gx >= 0 && gx < g.px;
    scol[i] = in ? gy + (int)sx * gx : 0;
    zr[i] = op.colz[scol[i]];
  }
  const int own_r = op.colz[lane_ok ? (int)colidx : 0];
  const int own_lo = own_r & 0xffff, own_hi = own_r >> 16;
  int soff[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    const int c = e / (SX * SY);
    const int lo = zr[i] & 0xffff, hi = zr[i] >> 16;
    const bool solid = hi >= lo;
    soff[i] = solid ? scol[i] + (PRO ? 0 : c * (int)N) : 0;
    zr[i] = solid ? lo | ((hi - lo) << 16) : 0xffff;
  }

  // PRO: per staged position its coarse-patch offset at (ya, xa), its
  // component and the parities of its y and x (bits 0-11, 14-15, 12, 13);
  // the patch origin is (cy0, cx0) in coarse node coordinates (y0, x0 even)
  const int cy0 = (t.y0 >> 1) - 1, cx0 = (t.x0 >> 1) - 1;
  int ppos[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) {
    const int e = tid + i * NT;
    const int c = e / (SX * SY), rem = e % (SX * SY), xx = rem / SY, yy = rem % SY;
    const int yf = t.y0 - 2 + yy, xf = t.x0 - 2 + xx;   // fine node coordinates
    const int base = ((yf >> 1) - cy0) + CY * (((xf >> 1) - cx0) + CX * c);
    ppos[i] = e < NSTAGE ? base | ((yf & 1) << 12) | ((xf & 1) << 13) | (c << 14) : 0;
  }
  static_assert(NQ <= 4096, "patch offset in 12 bits");

  // raw loads only (no arithmetic on them until they are staged, two planes
  // later): the input and, for the smoother, the inverse diagonal (AX_P: z).  Index 0 of
  // every padded array is a corner halo entry and holds 0, so a skipped
  // position loads from there: the loads need no branches.
  // The index is formed in 32 bits (3 N < 2^31) and selected, not branched on.
  // AX_P: the first iteration of a solve (its = 1 after its beta step)
  const bool pfirst = AXP && co.sc->its == 1;
  // SMOOTH_R: no update in the first iteration of a solve (its = 0): q is
  // not read (index 0, a halo entry, holds 0), so r = in exactly
  const bool rfirst = RUP && co.sc->its == 0;
  const TR ralpha = RUP ? (TR)co.sc->alpha : (TR)0;
  auto fetch = [&](int zz, TR* vr, TD* vd, uint32_t* vf, TR* vc, TR* vq) {
    const int szz = (int)sz * zz;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const bool ok = (unsigned)(zz - (zr[i] & 0xffff)) <= (unsigned)(zr[i] >> 16);
      const unsigned m = ok ? (unsigned)(soff[i] + szz) : 0u;
      if constexpr (PRO) {
        vf[i] = op.flags[m];
      } else {
        // AX_P, first iteration: beta = 0 and p_old is not read (index 0, a
        // halo entry, holds 0)
        vr[i] = in[pfirst ? 0u : m];
        if (RUP) vq[i] = co.qv[rfirst ? 0u : m];
        if (SM) vd[i] = (TD)op.invd[m];
        if (AXP) vd[i] = (TD)co.zf[m];
      }
    }
    if constexpr (PRO) {
      // coarse planes (zz - 2) >> 1 and (zz - 1) >> 1 (node coordinates) of
      // the fine node plane zz - 1; outside the padded coarse grid: index 0
      const int zf = zz - 1;
#pragma unroll
      for (int j = 0; j < PQ; ++j) {
        const int q = tid + j * NT;
        const int yi = q % CY, xi = (q / CY) % CX, c = q / (CY * CX);
        const int gy = cy0 + yi + 1, gx = cx0 + xi + 1, ga = (zf >> 1) + 1, gb = ((zf + 1) >> 1) + 1;
        const bool okc = q < NQ && gy >= 0 && gy < co.gc.py && gx >= 0 && gx < co.gc.px;
        const int col = gy + co.gc.py * gx + c * (int)co.gc.N, sp = co.gc.py * co.gc.px;
        const bool oka = okc && ga >= 0 && ga < co.gc.pz, okb = okc && gb >= 0 && gb < co.gc.pz;
        vc[2 * j] = in[oka ? (unsigned)(col + sp * ga) : 0u];
        vc[2 * j + 1] = in[okb ? (unsigned)(col + sp * gb) : 0u];
      }
    }
  };

  T aM[3] = {0, 0, 0}, a0[3] = {0, 0, 0}, aP[3] = {0, 0, 0};   // outputs zz-1, zz, zz+1
  uint32_t fM = 0, f0 = 0, fP = 0;
  double dsum = 0.0, dsum2 = 0.0;
  int kb = 0;                                       // buffer of the current plane
  // two planes of loads in flight: plane zz is staged from one register set
  // while the other holds plane zz + 1; z0 < z1 for every tile
  TR rA[PRO ? 1 : PER], rB[PRO ? 1 : PER];
  TD dA[ND], dB[ND];
  uint32_t gA[PF], gB[PF];                          // PRO: flags of the staged positions
  TR cA[PC], cB[PC];                                // PRO: the coarse patch
  TR qA[RUP ? PER : 1], qB[RUP ? PER : 1];          // SMOOTH_R: q
  T beta = 0, alpha = 0, alpha2 = 0;
  bool xup = false, xzero = false;
  if (AXP) {
    beta = (T)co.sc->beta;
    alpha = (T)co.sc->alpha;
    alpha2 = (T)co.sc->alpha_prev;
    // nothing is pending in the first iteration (its = 1 after its beta
    // step), and x is still 0 in the third (its = 3), the first to update it:
    // the start of the solve writes neither
    xup = XUP && co.sc->its != 1;
    xzero = co.sc->its == 3;
  }
  fetch(z0 - 1, rA, dA, gA, cA, qA);
  fetch(z0, rB, dB, gB, cB, qB);
  // flags of the outputs this lane owns, loaded one plane before they are used
  auto flags_at = [&](const int z) -> uint32_t {
    const bool ok = lane_ok && z >= z0 && z < z1 && z >= own_lo && z <= own_hi;
    return op.flags[ok ? (unsigned)((int)colidx + (int)sz * z) : 0u];    // flags[0]: a halo node, 0
  };
  uint32_t fN = flags_at(z0);

  auto plane = [&](const int zz, TR (&sr)[PRO ? 1 : PER], TD (&sd)[ND], uint32_t (&sf)[PF], TR (&sc)[PC],
                   TR (&sq)[RUP ? PER : 1]) {
    T (*const sB)[SX][SY] = sU[kb];
    if constexpr (PRO) {
      // e = P x_c at the staged positions (kernel/InterpolationDeviation.m:3-17)
      // in the order of prolong_zplanes / prolong_xy, as the stand-alone
      // prolongation: the patch's two planes are combined into sC, then every
      // staged value reads its four corners there; 0 at void nodes and fixed
      // components.  sC was last read before the previous plane's barrier.
      const int zo = (zz - 1) & 1;
      const T wz0 = zo ? (T)0.5 : (T)1, wz1 = zo ? (T)0.5 : (T)0;
#pragma unroll
      for (int j = 0; j < PQ; ++j) {
        const int q = tid + j * NT;
        if (q < NQ) sC[q] = prolong_zplanes(wz0, wz1, (T)sc[2 * j], (T)sc[2 * j + 1]);
      }
      __syncthreads();
      // (the position words pass through an empty asm each plane, so that
      // what is derived from them is not kept in registers across planes)
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = tid + i * NT;
        if (e < NSTAGE) {
          int w = ppos[i];
          asm volatile("" : "+v"(w));
          const int c = (w >> 14) & 3, yo = (w >> 12) & 1, xo = (w >> 13) & 1, o = w & 0xfff;
          const uint32_t f = sf[i];
          T v = 0;
          if ((f & F_ELEM_MASK) && !(f & (1u << (F_FIX_SHIFT + c)))) {
            const T wy0 = yo ? (T)0.5 : (T)1, wy1 = yo ? (T)0.5 : (T)0;
            const T wx0 = xo ? (T)0.5 : (T)1, wx1 = xo ? (T)0.5 : (T)0;
            const int ox = CY * xo;
            v = prolong_xy(wy0, wy1, wx0, wx1, sC[o], sC[o + ox], sC[o + yo], sC[o + yo + ox]);
          }
          (&sB[0][0][0])[e] = v;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = tid + i * NT;
        if (e < NSTAGE) {
          // (SMOOTH_R: r = in - alpha q, kept in sr for the centre values)
          if (RUP) sr[i] = fma_t(-ralpha, sq[i], sr[i]);
          T v = (T)sr[i];
          if (SM) v = omega * (T)sd[i] * v;
          if (AXP) v = (T)sd[i] + beta * v;              // p = z + beta p_old (CG_solver.m:28)
          (&sB[0][0][0])[e] = v;
        }
      }
    }
    // one barrier per plane: buffer kb was last read while computing planes
    // zz - 4 .. zz - 2, which every thread finished before the barrier of
    // plane zz - 1
    __syncthreads();
    if (SM || XUP) {
      // centre values of r (p_old) for the output completed at plane zz + 1; buffer
      // zz % 2 was last read at plane zz - 1, before this plane's barrier
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = tid + i * NT;
        if (e < NSTAGE) {
          const int c = e / (SX * SY), rem = e % (SX * SY), xx = rem / SY, yy = rem % SY;
          if (xx >= 1 && xx <= TX && yy >= 1 && yy <= TY) sR[zz & 1][c][xx - 1][yy - 1] = (TC)sr[i];
        }
      }
    }
    // planes zz-2, zz-1, zz are staged in buffers b2, b1, kb
    const int b1 = (kb + 3) & 3, b2 = (kb + 2) & 3;
    // RESID: centre values of in2 at the output completed in this plane
    // (AX_P: of x), loaded now and used after the stencil
    T cen[3] = {0, 0, 0};
    if ((MODE == AP_RESID || RSUM || (AXP && xup && !xzero)) && lane_ok && (fM & OUT_MASK)) {
      const long long n = colidx + sz * (zz - 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) cen[c] = AXP ? (T)co.x[c * N + n] : in2[c * N + n];
    }
    // Element corrections of output zz-1.  An owned node lacking incident
    // element p subtracts that element's universal row block, one touching a
    // stored element p swaps the universal block for the stored one.  The
    // wave lists its (node, element) pairs, spreads them over its 64 lanes,
    // and every owner sums its own pairs in ascending p.  A universal block
    // is a reflection of the canonical corner block (uniform coefficients):
    // reflect the element's displacements, multiply, reflect back.  Done
    // before the next loads are issued and before the class sums exist, where
    // the fewest registers are live; aM is additive.
    uint32_t cm = 0;
    if (!FAST_ONLY && (fM & F_ELEM_MASK)) cm = ((~fM) & F_ELEM_MASK) | ((fM >> F_STORED_SHIFT) & 0xFFu);
    if (!FAST_ONLY && __ballot(cm != 0u)) {
      const int cnt = __popc(cm);
      // exclusive prefix of the pair counts over the wave and their total,
      // from one ballot per element bit (no chain of cross-lane moves)
      int off = 0, total = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const unsigned long long mb = __ballot((cm >> b) & 1u);
        off += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mb >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u));
        total += __popcll(mb);
      }
      T D0 = 0, D1 = 0, D2 = 0;
      for (int rb = 0; rb < total; rb += 64) {
        // each lane lists its pairs of this round in the wave's task slots:
        // owner lane, element present (stored) and element index
        {
          uint32_t mm = cm;
          int idx = off;
          while (mm) {
            const int p = __ffs(mm) - 1;
            mm &= mm - 1;
            if (idx >= rb && idx < rb + 64)
              sTask[wv][idx - rb] = (lane << 4) | ((int)((fM >> p) & 1u) << 3) | p;
            ++idx;
          }
        }
        wave_lds_sync();
        const int j = rb + lane;
        T e0 = 0, e1 = 0, e2 = 0;
        if (j < total) {
          const int task = sTask[wv][lane];
          const int own = task >> 4, p = task & 7;
          const int cn = p ^ 7;                        // corner code of the node in element p
          const T sgx = (cn & 2) ? (T)-1 : (T)1, sgy = (cn & 1) ? (T)-1 : (T)1,
                  sgz = (cn & 4) ? (T)-1 : (T)1;
          const int oly = own % TY + 1, olx = lx_w + own / TY + 1;   // staged coords of the node
          const T* Kst = nullptr;
          if (HAS_EOP && ((task >> 3) & 1)) {
            const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = p >> 2;
            const long long no = (t.y0 + oly - 1) + sx * (t.x0 + olx - 1) + sz * (zz - 1);
            const long long cell = no - (1 - pyb) - (1 - pxb) * sx - (1 - pzb) * sz;
            Kst = op.Ks + (long long)(op.eop[cell] - 1) * 576 + 3 * hex_local(cn) * 24;
          }
          const cptr<T> Cc = opaque(op.corner);
          T h0 = 0, h1 = 0, h2 = 0;                    // stored block times u
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            // canonical corner k is corner k ^ cn of the element
            const int dy = (cn & 1) ? -(k & 1) : (k & 1);
            const int dx = (cn & 2) ? -((k >> 1) & 1) : ((k >> 1) & 1);
            const int dz = (cn & 4) ? -(k >> 2) : (k >> 2);
            const int bz = dz < 0 ? b2 : (dz == 0 ? b1 : kb);
            const T u0 = sU[bz][0][olx + dx][oly + dy], u1 = sU[bz][1][olx + dx][oly + dy],
                    u2 = sU[bz][2][olx + dx][oly + dy];
            const T v0 = sgx * u0, v1 = sgy * u1, v2 = sgz * u2;
            e0 += Cc[9 * k + 0] * v0 + Cc[9 * k + 1] * v1 + Cc[9 * k + 2] * v2;
            e1 += Cc[9 * k + 3] * v0 + Cc[9 * k + 4] * v1 + Cc[9 * k + 5] * v2;
            e2 += Cc[9 * k + 6] * v0 + Cc[9 * k + 7] * v1 + Cc[9 * k + 8] * v2;
            if (HAS_EOP && Kst) {
              const T* ks = Kst + 3 * hex_local(k ^ cn);
              h0 += ks[0] * u0 + ks[1] * u1 + ks[2] * u2;
              h1 += ks[24] * u0 + ks[25] * u1 + ks[26] * u2;
              h2 += ks[48] * u0 + ks[49] * u1 + ks[50] * u2;
            }
          }
          e0 = sgx * e0 - h0;
          e1 = sgy * e1 - h1;
          e2 = sgz * e2 - h2;
        }
        sRes[wv][0][lane] = e0;
        sRes[wv][1][lane] = e1;
        sRes[wv][2][lane] = e2;
        wave_lds_sync();
        // owners collect this round's pairs, in order
        for (int q = 0, idx = off; q < cnt; ++q, ++idx) {
          if (idx >= rb && idx < rb + 64) {
            D0 += sRes[wv][0][idx - rb];
            D1 += sRes[wv][1][idx - rb];
            D2 += sRes[wv][2][idx - rb];
          }
        }
        wave_lds_sync();   // the slots are written again by the next round
      }
      aM[0] -= D0;
      aM[1] -= D1;
      aM[2] -= D2;
    }
    if (zz + 2 <= z1) fetch(zz + 2, sr, sd, sf, sc, sq);
    // flags of the output at zz+1 (only outputs in [z0, z1) are owned)
    fP = fN;
    fN = flags_at(zz + 2);
    // neighbour (dy, dx) = (a-1, b-1), component c, read from the staged plane
#define GMGS_U(a, b, c) sB[c][lx + (b)][ly + (a)]
    // One displacement component at a time: its nine in-plane values, their
    // class sums and differences, and their contributions to the three
    // outputs (z-1, z, z+1) are formed before the next component's values are
    // read, so only nine staged values are live at once (the scheduling
    // barrier keeps the next component's LDS reads behind this one's
    // arithmetic).  The class coefficients: diagonal blocks by
    // |dy| + 2|dx| (cf[8 c + 4 q + 0..3], q = 1 off the output's plane),
    // the xy coupling of the two in-plane diagonals (cf[24 + q]), the xz and
    // yz couplings of the off-plane outputs (cf[26..27], cf[28..29], odd in
    // dz).
    {
      const cptr<T> C = opaque(op.cls);
      T cf[30];
#pragma unroll
      for (int i = 0; i < 30; ++i) cf[i] = C[i];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T u00 = GMGS_U(0, 0, c), u01 = GMGS_U(0, 1, c), u02 = GMGS_U(0, 2, c);
        const T u10 = GMGS_U(1, 0, c), u11 = GMGS_U(1, 1, c), u12 = GMGS_U(1, 2, c);
        const T u20 = GMGS_U(2, 0, c), u21 = GMGS_U(2, 1, c), u22 = GMGS_U(2, 2, c);
        const T cy = u01 + u21, cx = u10 + u12, cxy = (u00 + u22) + (u02 + u20);
        aM[c] += cf[8 * c + 4] * u11 + cf[8 * c + 5] * cy + cf[8 * c + 6] * cx + cf[8 * c + 7] * cxy;
        a0[c] += cf[8 * c + 0] * u11 + cf[8 * c + 1] * cy + cf[8 * c + 2] * cx + cf[8 * c + 3] * cxy;
        aP[c] += cf[8 * c + 4] * u11 + cf[8 * c + 5] * cy + cf[8 * c + 6] * cx + cf[8 * c + 7] * cxy;
        if (c < 2) {
          // in-plane diagonal difference: x feeds the y output, y the x output
          const T dxy = (u22 + u00) - (u20 + u02);
          const int o = 1 - c;
          aM[o] += cf[25] * dxy;
          a0[o] += cf[24] * dxy;
          aP[o] += cf[25] * dxy;
        }
        if (c != 1) {
          // x differences: of x into the z outputs, of z into the x outputs
          const T dx0 = u12 - u10, dx1 = (u02 - u00) + (u22 - u20);
          const T t = cf[26] * dx0 + cf[27] * dx1;
          const int o = c == 0 ? 2 : 0;
          aM[o] += t;
          aP[o] -= t;
        }
        if (c != 0) {
          // y differences: of y into the z outputs, of z into the y outputs
          const T dy0 = u21 - u01, dy1 = (u20 - u00) + (u22 - u02);
          const T t = cf[28] * dy0 + cf[29] * dy1;
          const int o = c == 1 ? 2 : 1;
          aM[o] += t;
          aP[o] -= t;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
#undef GMGS_U
    // write output zz-1
    const int zo_ = zz - 1;
    if (lane_ok && zo_ >= z0 && zo_ < z1 && (fM & OUT_MASK)) {
      const long long n = colidx + sz * zo_;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T a = (fM & (1u << (F_FIX_SHIFT + c))) ? (T)0 : aM[c];
        if (MODE == AP_AX) {
          out[c * N + n] = a;
          if (partial) dsum += (double)sU[b1][c][lx + 1][ly + 1] * (double)a;
        } else if (AXP) {
          // q = A p, the new p, and (AX_P) the two deferred x updates:
          // x = (x + alpha_prev p_2) + alpha p_old, p_2 read from the new p's
          // buffer before it is overwritten
          const long long m = c * N + n;
          const T pn = sU[b1][c][lx + 1][ly + 1];
          out[m] = a;
          if (XUP && xup) {
            const T p2 = co.pn[m];
            co.x[m] = fma_t(alpha, (T)sR[zo_ & 1][c][lx][ly], fma_t(alpha2, p2, cen[c]));
          }
          co.pn[m] = pn;
          if (partial) dsum += (double)pn * (double)a;
        } else if (MODE == AP_RESID) {
          const T v = cen[c] - a;
          out[c * N + n] = v;
          if (partial) dsum += (double)v * (double)v;   // |in2 - A in|^2
        } else if (RSUM) {
          // z = omega (r0 + Q_0) / d + e, Q_0 = in2 - A e; e at the node is
          // the staged input's centre
          const long long m = c * N + n;
          const double r0 = co.r0[m];
          const bool fx = (fM & (1u << (F_FIX_SHIFT + c))) != 0;
          // (fixed components are set to 0 below, whatever their diagonal)
          const T di = (co.ufast && (fM & F_FAST)) ? co.ucinvd[c] : op.invd[m];
          T v = di * (omega * ((T)r0 + (cen[c] - a))) + sU[b1][c][lx + 1][ly + 1];
          if (fx) v = 0;
          co.z[m] = v;
          if (partial) {
            dsum += r0 * (double)v;
            if (co.zprev) dsum2 += r0 * (double)co.zprev[m];
          }
        } else if (SUM) {
          // the other terms of the cycle's sum, loaded here to keep them out
          // of the stencil's registers
          const long long m = c * N + n;
          const double r0 = co.r0[m];
          const bool fx = (fM & (1u << (F_FIX_SHIFT + c))) != 0;
          const T q0 = (T)sR[zo_ & 1][c][lx][ly];
          const T q1 = q0 - a;
          T v = op.invd[m] * (omega * ((T)r0 + q0) + co.w1 * (co.R1[m] + q1)) + co.e[m];
          if (fx) v = 0;
          co.z[m] = v;
          if (partial) {
            dsum += r0 * (double)v;
            if (co.zprev) dsum2 += r0 * (double)co.zprev[m];
          }
        } else {
          const bool fx = (fM & (1u << (F_FIX_SHIFT + c))) != 0;
          out[c * N + n] = (fx && op.zero_fixed) ? (T)0 : (T)sR[zo_ & 1][c][lx][ly] - a;
          if (RUP) {
            // the new residual of the PCG and its square (CG_solver.m:33-35)
            const TR rv = sR[zo_ & 1][c][lx][ly];
            co.rn[c * N + n] = rv;
            dsum += (double)rv * (double)rv;
          }
        }
      }
    }
    // rotate
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      aM[c] = a0[c];
      a0[c] = aP[c];
      aP[c] = 0;
    }
    fM = f0;
    f0 = fP;
    kb = (kb + 1) & 3;
  };
  for (int zz = z0 - 1; zz <= z1; zz += 2) {
    plane(zz, rA, dA, gA, cA, qA);
    if (zz + 1 <= z1) plane(zz + 1, rB, dB, gB, cB, qB);
  }
  if (partial) {
    // block reduction (fixed order)
    double v = dsum, v2 = dsum2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v += __shfl_xor(v, off);
      if (TWO) v2 += __shfl_xor(v2, off);
    }
    __shared__ double sh[TWO ? 2 : 1][NT / 64];
    if ((tid & 63) == 0) {
      sh[0][tid >> 6] = v;
      if (TWO) sh[TWO ? 1 : 0][tid >> 6] = v2;
    }
    __syncthreads();
    if (tid == 0) {
      double r = 0.0, r2 = 0.0;
      for (int w = 0; w < NT / 64; ++w) {
        r += sh[0][w];
        if (TWO) r2 += sh[TWO ? 1 : 0][w];
      }
      partial[block_] = r;
      if (TWO) partial[nblocks_ + block_] = r2;
    }
  }
  }
};

__global__ void k_coltile_mark(Grid g, const uint32_t* __restrict__ flags, int nty, int ntx, int ntz,
                               int tx, int zc, int* __restrict__ mark) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nt = (long long)nty * ntx * ntz;
  if (t >= nt) return;
  const int ty = (int)(t % nty), txi = (int)((t / nty) % ntx), tz = (int)(t / ((long long)nty * ntx));
  int any = 0;
  for (int z = 1 + tz * zc; z < min(1 + (tz + 1) * zc, g.pz - 1) && !any; ++z)
    for (int x = txi * tx; x < min((txi + 1) * tx, g.px) && !any; ++x)
      for (int y = ty * 16; y < ty * 16 + 16; ++y)
        if (flags[y + (long long)g.py * (x + (long long)g.px * z)] & F_ELEM_MASK) { any = 1; break; }
  mark[t] = any;
}

__global__ void k_coltile_emit(const int* __restrict__ mark, const int* __restrict__ scan, int nty,
                               int ntx, int ntz, int tx, int zc, ColTile* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nt = (long long)nty * ntx * ntz;
  if (t >= nt || !mark[t]) return;
  ColTile r;
  r.y0 = (int)(t % nty) * 16;
  r.x0 = (int)((t / nty) % ntx) * tx;
  r.z0 = 1 + (int)(t / ((long long)nty * ntx)) * zc;
  r.pad = 0;
  out[scan[t]] = r;
}

__global__ void k_colrange(Grid g, const uint32_t* __restrict__ flags, int* __restrict__ colz) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncol = (long long)g.py * g.px, sz = ncol;
  if (col >= ncol) return;
  int lo = 0xffff, hi = 0;
  for (int z = 0; z < g.pz; ++z)
    if (flags[col + sz * z] & F_ELEM_MASK) {
      lo = min(lo, z);
      hi = z;
    }
  colz[col] = lo | (hi << 16);
}

}  // namespace

void build_colrange(const Grid& g, const uint32_t* flags, DevBuf& colz, hipStream_t s) {
  if (g.pz > 32767) throw std::invalid_argument("z extent too large for the column range table");
  const long long ncol = (long long)g.py * g.px;
  colz.alloc(sizeof(int) * ncol);
  k_colrange<<<grid1d(ncol, 256), 256, 0, s>>>(g, flags, colz.as<int>());
  GMGS_LAUNCH_CHECK();
}

long long build_coltiles(const Grid& g, const uint32_t* flags, int tx, int zc, DevBuf& tiles,
                         hipStream_t s) {
  const int nty = g.py / 16, ntx = (g.px + tx - 1) / tx, ntz = (g.pz - 2 + zc - 1) / zc;
  const long long nt = (long long)nty * ntx * ntz;
  DevBuf mark, scan;
  mark.alloc(sizeof(int) * nt);
  scan.alloc(sizeof(int) * nt);
  k_coltile_mark<<<grid1d(nt, 256), 256, 0, s>>>(g, flags, nty, ntx, ntz, tx, zc, mark.as<int>());
  GMGS_LAUNCH_CHECK();
  const long long total = exclusive_scan(mark.as<int>(), scan.as<int>(), nt, s);
  tiles.alloc(sizeof(ColTile) * (total > 0 ? total : 1));
  k_coltile_emit<<<grid1d(nt, 256), 256, 0, s>>>(mark.as<int>(), scan.as<int>(), nty, ntx, ntz, tx, zc,
                                                 tiles.as<ColTile>());
  GMGS_LAUNCH_CHECK();
  GMGS_HIP(hipStreamSynchronize(s));
  return total;
}

template <typename T, typename TR>
int launch_apply_zm(const LevelOp<T>& op, const ColTile* tiles, long long ntiles, int tx, int zc,
                    int mode, const TR* in, T omega, const T* in2, T* out, double* partial,
                    const int* done, hipStream_t s, const CycleOut<T>* co, bool fast_only) {
  if (ntiles <= 0) return 0;
  if (3 * op.g.N >= (1LL << 31)) throw std::invalid_argument("grid too large for the z-marching operator");
  const int e = fast_only ? 2 : (op.eop != nullptr ? 1 : 0);
  const CycleOut<T> cov = co ? *co : CycleOut<T>{};
  if ((mode == AP_SMOOTH_SUM || mode == AP_RESID_SUM || mode == AP_RESID_SUM_P) && (!co || e != 0 || !op.invd))
    throw std::invalid_argument("SMOOTH_SUM / RESID_SUM: level 1 only, with its outputs");
  if (mode == AP_RESID_SUM_P &&
      (co->gc.ey * 2 != op.g.ey || co->gc.ex * 2 != op.g.ex || co->gc.ez * 2 != op.g.ez || 3 * co->gc.N >= (1LL << 31)))
    throw std::invalid_argument("RESID_SUM_P: the coarse grid is not this level's next coarser grid");
  if ((mode == AP_AX_P || mode == AP_AX_P_NOX) &&
      (!co || e != 0 || !co->zf || !co->pn || (mode == AP_AX_P && !co->x) || !co->sc || co->pn == (const void*)in))
    throw std::invalid_argument("AX_P: level 1 only, with z, x, the scalars and a separate new p");
  if (mode == AP_SMOOTH_R &&
      (!co || e != 0 || !co->qv || !co->rn || !co->sc || !partial || co->rn == (const void*)in ||
       co->rn == co->qv))
    throw std::invalid_argument("SMOOTH_R: level 1 only, with q, the scalars, the partials and a separate new r");
#define GMGS_ZM(MODE, E, TXV) \
  launch(KApplyZm<T, TR, MODE, E, TXV>{op, tiles, zc, in, omega, in2, out, partial, done, cov}, ntiles, s)
#define GMGS_ZM_TX(MODE, E)               \
  do {                                    \
    if (tx == 4) GMGS_ZM(MODE, E, 4);     \
    else if (tx == 16) GMGS_ZM(MODE, E, 16); \
    else GMGS_ZM(MODE, E, 8);             \
  } while (0)
#define GMGS_ZM_E(MODE)                   \
  do {                                    \
    if (e == 1) GMGS_ZM_TX(MODE, 1);      \
    else if (e == 2) GMGS_ZM_TX(MODE, 2); \
    else GMGS_ZM_TX(MODE, 0);             \
  } while (0)
  if (mode == AP_AX) {
    GMGS_ZM_E(AP_AX);
  } else if (mode == AP_RESID) {
    GMGS_ZM_E(AP_RESID);
  } else if (mode == AP_SMOOTH_SUM) {
    if constexpr (std::is_same<T, TR>::value) GMGS_ZM_TX(AP_SMOOTH_SUM, 0);
    else throw std::invalid_argument("SMOOTH_SUM: input in the cycle precision");
  } else if (mode == AP_RESID_SUM) {
    if constexpr (std::is_same<T, TR>::value) GMGS_ZM_TX(AP_RESID_SUM, 0);
    else throw std::invalid_argument("RESID_SUM: input in the cycle precision");
  } else if (mode == AP_RESID_SUM_P) {
    if constexpr (std::is_same<T, TR>::value) GMGS_ZM_TX(AP_RESID_SUM_P, 0);
    else throw std::invalid_argument("RESID_SUM_P: input in the cycle precision");
  } else if (mode == AP_AX_P) {
    if constexpr (std::is_same<T, double>::value && std::is_same<TR, double>::value) GMGS_ZM_TX(AP_AX_P, 0);
    else throw std::invalid_argument("AX_P: FP64 operator and input");
  } else if (mode == AP_AX_P_NOX) {
    if constexpr (std::is_same<T, double>::value && std::is_same<TR, double>::value) GMGS_ZM_TX(AP_AX_P_NOX, 0);
    else throw std::invalid_argument("AX_P: FP64 operator and input");
  } else if (mode == AP_SMOOTH_R) {
    if constexpr (std::is_same<T, float>::value && std::is_same<TR, double>::value) GMGS_ZM_TX(AP_SMOOTH_R, 0);
    else throw std::invalid_argument("SMOOTH_R: FP32 operator, FP64 residual");
  } else {
    GMGS_ZM_E(AP_SMOOTH);
  }
#undef GMGS_ZM_E
#undef GMGS_ZM_TX
#undef GMGS_ZM
  GMGS_LAUNCH_CHECK();
  return (int)ntiles;
}

template int launch_apply_zm<double, double>(const LevelOp<double>&, const ColTile*, long long, int, int,
                                             int, const double*, double, const double*, double*, double*,
                                             const int*, hipStream_t, const CycleOut<double>*, bool);
template int launch_apply_zm<float, float>(const LevelOp<float>&, const ColTile*, long long, int, int,
                                           int, const float*, float, const float*, float*, double*,
                                           const int*, hipStream_t, const CycleOut<float>*, bool);
template int launch_apply_zm<float, double>(const LevelOp<float>&, const ColTile*, long long, int, int,
                                            int, const double*, float, const float*, float*, double*,
                                            const int*, hipStream_t, const CycleOut<float>*, bool);

}  // namespace gmgs
