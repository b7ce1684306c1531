// z-marching level operator (levels with a universal element matrix).
//
// A workgroup owns a 16 (y) x TX (x) column of nodes and walks a chunk of
// z-planes.  Each input plane (with a one-node halo) is staged once in LDS;
// every lane reads its 3x3 in-plane neighbourhood from LDS and adds that
// plane's contribution to the three outputs it touches (z-1, z, z+1), so a
// value is fetched from memory about once per plane instead of 27 times.
//
// Interior nodes use the assembled 27-point stencil of the universal element
// matrix with its structural zeros skipped at compile time; a node missing an
// element, or touching a stored (non-universal) element, subtracts that
// element's universal contribution plane by plane and adds the stored one.
// Same result as the per-element sum of kernel/AtX.m:1-58.
#include "ops.h"
#include "ops_zm.h"
#include "mesh.h"

namespace gmgs {

namespace {


// Coefficients are read through a constant-address-space pointer kept opaque
// to loop-invariant code motion: every plane re-reads them with scalar loads
// (scalar cache) into SGPR operands instead of hoisting hundreds of values
// into vector registers.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

template <typename T>
__device__ inline cptr<T> opaque(const T* p) {
  cptr<T> q = (cptr<T>)p;
  asm volatile("" : "+s"(q));
  return q;
}

template <typename T, typename TR, int MODE, bool HAS_EOP, int TX>
__global__ __launch_bounds__(16 * TX) void k_apply_zm(LevelOp<T> op, const ColTile* __restrict__ tiles,
                                                      int zc, const TR* __restrict__ in, T omega,
                                                      T* __restrict__ out, TR* __restrict__ xacc,
                                                      double* __restrict__ partial,
                                                      const int* __restrict__ done) {
  constexpr int TY = 16, NT = TY * TX, SY = TY + 2, SX = TX + 2;
  constexpr int NSTAGE = 3 * SX * SY;
  constexpr int PER = (NSTAGE + NT - 1) / NT;      // staged values per thread
  if (done && *done) return;
  __shared__ T sU[3][SX][SY];
  const ColTile t = tiles[blockIdx.x];
  const int tid = threadIdx.x, ly = tid % TY, lx = tid / TY;
  const Grid g = op.g;
  const long long N = g.N, sx = g.py, sz = (long long)g.py * g.px;
  const int yp = t.y0 + ly, xp = t.x0 + lx;
  const bool lane_ok = xp < g.px;
  const long long colidx = yp + sx * xp;
  const int z0 = t.z0, z1 = min(t.z0 + zc, g.pz - 1);   // outputs [z0, z1)

  // this thread's share of a staged plane: global offset (without z) or -1
  long long soff[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    soff[i] = -1;
    if (e < NSTAGE) {
      const int c = e / (SX * SY), rem = e % (SX * SY), xx = rem / SY, yy = rem % SY;
      const int gy = t.y0 - 1 + yy, gx = t.x0 - 1 + xx;
      if (gy >= 0 && gy < g.py && gx >= 0 && gx < g.px) soff[i] = gy + sx * gx + c * N;
    }
  }
  auto fetch = [&](int zz, T* v) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      v[i] = 0;
      if (soff[i] >= 0) {
        const long long m = soff[i] + sz * zz;
        if (MODE == AP_AX) v[i] = (T)in[m];
        else v[i] = omega * op.invd[m] * (T)in[m];
      }
    }
  };

  T aM[3] = {0, 0, 0}, a0[3] = {0, 0, 0}, aP[3] = {0, 0, 0};   // outputs zz-1, zz, zz+1
  uint32_t fM = 0, f0 = 0, fP = 0;
  T uc0[3] = {0, 0, 0};   // centre u of plane zz-1 (the output being written)
  double dsum = 0.0;
  T stage[PER];
  fetch(z0 - 1, stage);

  for (int zz = z0 - 1; zz <= z1; ++zz) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = tid + i * NT;
      if (e < NSTAGE) (&sU[0][0][0])[e] = stage[i];
    }
    __syncthreads();
    if (zz < z1) fetch(zz + 1, stage);               // next plane in flight during compute
    // flags of the output at zz+1 (only outputs in [z0, z1) are owned)
    fP = (lane_ok && zz + 1 >= z0 && zz + 1 < z1) ? op.flags[colidx + sz * (zz + 1)] : 0u;
    // neighbour (dy, dx) = (a-1, b-1), component c, read from the staged plane
#define GMGS_U(a, b, c) sU[c][lx + (b)][ly + (a)]
    // class sums of the in-plane neighbourhood (shared by the three outputs)
    T C0[3], Cy[3], Cx[3], Cxy[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      C0[c] = GMGS_U(1, 1, c);
      Cy[c] = GMGS_U(0, 1, c) + GMGS_U(2, 1, c);
      Cx[c] = GMGS_U(1, 0, c) + GMGS_U(1, 2, c);
      Cxy[c] = (GMGS_U(0, 0, c) + GMGS_U(2, 2, c)) + (GMGS_U(0, 2, c) + GMGS_U(2, 0, c));
    }
    const T Dxy_x = (GMGS_U(2, 2, 0) + GMGS_U(0, 0, 0)) - (GMGS_U(2, 0, 0) + GMGS_U(0, 2, 0));
    const T Dxy_y = (GMGS_U(2, 2, 1) + GMGS_U(0, 0, 1)) - (GMGS_U(2, 0, 1) + GMGS_U(0, 2, 1));
    const T Dx0_x = GMGS_U(1, 2, 0) - GMGS_U(1, 0, 0);
    const T Dx1_x = (GMGS_U(0, 2, 0) - GMGS_U(0, 0, 0)) + (GMGS_U(2, 2, 0) - GMGS_U(2, 0, 0));
    const T Dx0_z = GMGS_U(1, 2, 2) - GMGS_U(1, 0, 2);
    const T Dx1_z = (GMGS_U(0, 2, 2) - GMGS_U(0, 0, 2)) + (GMGS_U(2, 2, 2) - GMGS_U(2, 0, 2));
    const T Dy0_y = GMGS_U(2, 1, 1) - GMGS_U(0, 1, 1);
    const T Dy1_y = (GMGS_U(2, 0, 1) - GMGS_U(0, 0, 1)) + (GMGS_U(2, 2, 1) - GMGS_U(0, 2, 1));
    const T Dy0_z = GMGS_U(2, 1, 2) - GMGS_U(0, 1, 2);
    const T Dy1_z = (GMGS_U(2, 0, 2) - GMGS_U(0, 0, 2)) + (GMGS_U(2, 2, 2) - GMGS_U(0, 2, 2));
    {
      const cptr<T> C = opaque(op.cls);
      T cf[30];
#pragma unroll
      for (int i = 0; i < 30; ++i) cf[i] = C[i];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int oz = 1 - k, q = oz != 0 ? 1 : 0;
        T* acc = k == 0 ? aM : (k == 1 ? a0 : aP);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          acc[c] += cf[8 * c + 4 * q] * C0[c] + cf[8 * c + 1 + 4 * q] * Cy[c]
                    + cf[8 * c + 2 + 4 * q] * Cx[c] + cf[8 * c + 3 + 4 * q] * Cxy[c];
        acc[0] += cf[24 + q] * Dxy_y;
        acc[1] += cf[24 + q] * Dxy_x;
        if (oz != 0) {
          const T sgn = (T)oz;
          acc[0] += sgn * (cf[26] * Dx0_z + cf[27] * Dx1_z);
          acc[2] += sgn * (cf[26] * Dx0_x + cf[27] * Dx1_x);
          acc[1] += sgn * (cf[28] * Dy0_z + cf[29] * Dy1_z);
          acc[2] += sgn * (cf[28] * Dy0_y + cf[29] * Dy1_y);
        }
      }
    }
    // corrections for nodes that miss an element or touch a stored one
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int oz = 1 - k;
      const uint32_t f = k == 0 ? fM : (k == 1 ? f0 : fP);
      if (!(f & F_ELEM_MASK) || (f & F_FAST)) continue;
      const cptr<T> Ke = opaque(op.Ke);
      const int zout = zz - oz;
      const long long nout = colidx + sz * zout;
      T d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = p >> 2;
        if (pzb != oz && pzb != oz + 1) continue;          // element does not reach plane zz
        const bool present = (f >> p) & 1u;
        const T* Kst = nullptr;
        if (present) {
          if (!HAS_EOP) continue;
          const long long cell = nout - (1 - pyb) - (1 - pxb) * sx - (1 - pzb) * sz;
          const int e = op.eop[cell];
          if (e <= 0) continue;
          Kst = op.Ks + (long long)(e - 1) * 576;
        }
        const int ln = node_local_in(p);
        const int zo = oz + 1 - pzb;                       // local z of the plane in the element
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          if (hex_zo(m) != zo) continue;
          const int a = pyb + hex_ro(m), b = pxb + hex_xo(m);   // = dy+1, dx+1
          const T u0 = GMGS_U(a, b, 0), u1 = GMGS_U(a, b, 1), u2 = GMGS_U(a, b, 2);
          const cptr<T> k0 = Ke + (3 * ln + 0) * 24 + 3 * m;
          const cptr<T> k1 = Ke + (3 * ln + 1) * 24 + 3 * m;
          const cptr<T> k2 = Ke + (3 * ln + 2) * 24 + 3 * m;
          T e0 = k0[0] * u0 + k0[1] * u1 + k0[2] * u2;
          T e1 = k1[0] * u0 + k1[1] * u1 + k1[2] * u2;
          T e2 = k2[0] * u0 + k2[1] * u1 + k2[2] * u2;
          if (HAS_EOP && Kst) {
            const T* s0 = Kst + (3 * ln + 0) * 24 + 3 * m;
            const T* s1 = Kst + (3 * ln + 1) * 24 + 3 * m;
            const T* s2 = Kst + (3 * ln + 2) * 24 + 3 * m;
            e0 -= s0[0] * u0 + s0[1] * u1 + s0[2] * u2;
            e1 -= s1[0] * u0 + s1[1] * u1 + s1[2] * u2;
            e2 -= s2[0] * u0 + s2[1] * u1 + s2[2] * u2;
          }
          d0 += e0; d1 += e1; d2 += e2;
        }
      }
      if (k == 0) { aM[0] -= d0; aM[1] -= d1; aM[2] -= d2; }
      else if (k == 1) { a0[0] -= d0; a0[1] -= d1; a0[2] -= d2; }
      else { aP[0] -= d0; aP[1] -= d1; aP[2] -= d2; }
    }
    // output zz-1 is complete
    const int zo_ = zz - 1;
    if (lane_ok && zo_ >= z0 && zo_ < z1 && (fM & F_ELEM_MASK)) {
      const long long n = colidx + sz * zo_;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T a = (fM & (1u << (F_FIX_SHIFT + c))) ? (T)0 : aM[c];
        if (MODE == AP_AX) {
          out[c * N + n] = a;
          if (partial) dsum += (double)uc0[c] * (double)a;
        } else {
          out[c * N + n] = (T)in[c * N + n] - a;
          if (xacc) xacc[c * N + n] += (TR)uc0[c];
        }
      }
    }
    // rotate
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      aM[c] = a0[c];
      a0[c] = aP[c];
      aP[c] = 0;
      uc0[c] = C0[c];
    }
#undef GMGS_U
    fM = f0;
    f0 = fP;
  }
  if (partial) {
    // block reduction (fixed order)
    double v = dsum;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __shared__ double sh[NT / 64];
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double r = 0.0;
      for (int w = 0; w < NT / 64; ++w) r += sh[w];
      partial[blockIdx.x] = r;
    }
  }
}

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

}  // namespace

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
void launch_apply_zm(const LevelOp<T>& op, const ColTile* tiles, long long ntiles, int tx, int zc,
                     int mode, const TR* in, T omega, T* out, TR* xacc, double* partial,
                     const int* done, hipStream_t s) {
  if (ntiles <= 0) return;
  const bool eop = op.eop != nullptr;
#define GMGS_ZM(MODE, E, TXV)                                                                    \
  k_apply_zm<T, TR, MODE, E, TXV><<<(unsigned)ntiles, 16 * TXV, 0, s>>>(op, tiles, zc, in, omega, out, \
                                                                        xacc, partial, done)
#define GMGS_ZM_TX(MODE, E)            \
  do {                                 \
    if (tx == 4) GMGS_ZM(MODE, E, 4);  \
    else if (tx == 16) GMGS_ZM(MODE, E, 16); \
    else GMGS_ZM(MODE, E, 8);          \
  } while (0)
  if (mode == AP_AX) {
    if (eop) GMGS_ZM_TX(AP_AX, true);
    else GMGS_ZM_TX(AP_AX, false);
  } else {
    if (eop) GMGS_ZM_TX(AP_SMOOTH, true);
    else GMGS_ZM_TX(AP_SMOOTH, false);
  }
#undef GMGS_ZM_TX
#undef GMGS_ZM
  GMGS_LAUNCH_CHECK();
}

template void launch_apply_zm<double, double>(const LevelOp<double>&, const ColTile*, long long, int, int,
                                              int, const double*, double, double*, double*, double*,
                                              const int*, hipStream_t);
template void launch_apply_zm<float, float>(const LevelOp<float>&, const ColTile*, long long, int, int,
                                            int, const float*, float, float*, float*, double*,
                                            const int*, hipStream_t);
template void launch_apply_zm<float, double>(const LevelOp<float>&, const ColTile*, long long, int, int,
                                             int, const double*, float, float*, double*, double*,
                                             const int*, hipStream_t);

}  // namespace gmgs
