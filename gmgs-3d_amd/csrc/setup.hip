// Solver set-up kernels (kernel/SetupSolver.m).
#include "setup.h"

#include <cstring>
#include "mesh.h"

namespace gmgs {

namespace {

__constant__ double c_W[8][8][8];     // child-slot weights, see child_weights()
// elementUpwardMap column j -> geometric SubEleNodMat slot (exact placement)
__constant__ int c_colmajor_to_slot[8] = {0, 1, 3, 2, 4, 5, 7, 6};

__global__ void k_tile_mark(Grid g, const uint32_t* __restrict__ flags, int nty, int ntx,
                            long long ntiles, int* __restrict__ mark) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  int ty = (int)(t % nty), tx = (int)((t / nty) % ntx), z = (int)(t / ((long long)nty * ntx));
  int any = 0;
  for (int lx = 0; lx < TILE_X; ++lx) {
    int x = tx * TILE_X + lx;
    if (x >= g.px) break;
    for (int ly = 0; ly < TILE_Y; ++ly) {
      long long n = (ty * TILE_Y + ly) + (long long)g.py * (x + (long long)g.px * z);
      any |= (flags[n] & F_ELEM_MASK) ? 1 : 0;
    }
  }
  mark[t] = any;
}

__global__ void k_tile_emit(const int* __restrict__ mark, const int* __restrict__ scan, int nty,
                            int ntx, long long ntiles, TileRec* __restrict__ out) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles || !mark[t]) return;
  TileRec r;
  r.y0 = (int)(t % nty) * TILE_Y;
  r.x0 = (int)((t / nty) % ntx) * TILE_X;
  r.z = (int)(t / ((long long)nty * ntx));
  r.pad = 0;
  out[scan[t]] = r;
}

__global__ void k_clear_fixed(uint32_t* __restrict__ flags, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] &= ~F_FIX_MASK;
}

__global__ void k_set_fixed(const long long* __restrict__ pidx, const int* __restrict__ ids,
                            long long k, uint32_t* __restrict__ flags) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  atomicOr(&flags[pidx[ids[i] - 1]], F_FIX_MASK);
}

__device__ inline bool lane_pos(const Grid& g, const TileRec* __restrict__ tiles, long long nt,
                                long long& n, int& yp, int& xp, int& zp) {
  const long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= nt) return false;
  const TileRec tr = tiles[t];
  const int lane = threadIdx.x & 63;
  yp = tr.y0 + (lane & 15);
  xp = tr.x0 + (lane >> 4);
  zp = tr.z;
  if (xp >= g.px) return false;
  n = yp + (long long)g.py * (xp + (long long)g.px * zp);
  return true;
}

// per component the smallest and largest bit pattern of a positive float
// inverse diagonal over the F_FAST nodes (mm: min[3], max[3])
__global__ void k_fast_invd_range(Grid g, const TileRec* __restrict__ tiles, long long nt,
                                  const uint32_t* __restrict__ flags, const float* __restrict__ invd,
                                  unsigned* __restrict__ mm) {
  long long n;
  int yp, xp, zp;
  if (!lane_pos(g, tiles, nt, n, yp, xp, zp)) return;
  if (!(flags[n] & F_FAST)) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const unsigned b = __float_as_uint(invd[c * g.N + n]);
    atomicMin(&mm[c], b);
    atomicMax(&mm[3 + c], b);
  }
}

__global__ void k_restrict_fixed(Grid gc, const TileRec* __restrict__ tiles, long long nt,
                                 uint32_t* __restrict__ fc, Grid gf,
                                 const uint32_t* __restrict__ ff) {
  long long n;
  int yp, xp, zp;
  if (!lane_pos(gc, tiles, nt, n, yp, xp, zp)) return;
  uint32_t f = fc[n];
  if (!(f & F_ELEM_MASK)) return;
  const long long nf = gf.pidx(2 * (yp - 1), 2 * (xp - 1), 2 * (zp - 1));
  const long long sx = gf.py, sz = (long long)gf.py * gf.px;
  uint32_t acc = 0;
  for (int o = 0; o < 27; ++o) {
    const int dy = o % 3 - 1, dx = (o / 3) % 3 - 1, dz = o / 9 - 1;
    acc |= ff[nf + dy + dx * sx + dz * sz] & F_FIX_MASK;
  }
  fc[n] = (f & ~F_FIX_MASK) | acc;
}

__global__ void k_mark_fast(Grid g, const TileRec* __restrict__ tiles, long long nt,
                            uint32_t* __restrict__ flags, const int* __restrict__ eop, int allow) {
  long long n;
  int yp, xp, zp;
  if (!lane_pos(g, tiles, nt, n, yp, xp, zp)) return;
  uint32_t f = flags[n] & ~(F_FAST | (0xFFu << F_STORED_SHIFT));
  bool fast = allow && ((f & F_ELEM_MASK) == 0xFFu);
  uint32_t stored = 0;
  if (eop) {
    const long long sx = g.py, sz = (long long)g.py * g.px;
    for (int p = 0; p < 8; ++p) {
      if (!((f >> p) & 1u)) continue;
      const long long cell = n - (1 - (p & 1)) - (1 - ((p >> 1) & 1)) * sx - (1 - (p >> 2)) * sz;
      if (eop[cell] > 0) stored |= 1u << p;
    }
    fast = fast && stored == 0;
  }
  flags[n] = f | (fast ? F_FAST : 0u) | (stored << F_STORED_SHIFT);
}

__global__ void k_child_masks(const long long* __restrict__ epc, long long ne, Grid gc, Grid gf,
                              const uint8_t* __restrict__ esf, int* __restrict__ mask) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  const long long cell = epc[k];
  // cell pidx -> unpadded coords
  const long long sxy = (long long)gc.py * gc.px;
  const int z = (int)(cell / sxy) - 1, x = (int)((cell % sxy) / gc.py) - 1, y = (int)(cell % gc.py) - 1;
  int m = 0;
  for (int j = 0; j < 8; ++j)
    if (esf[gf.pidx(2 * y + (j & 1), 2 * x + ((j >> 1) & 1), 2 * z + (j >> 2))]) m |= 1 << j;
  mask[k] = m;
}

__global__ void k_unique_l3(const long long* __restrict__ epc, long long ne, Grid gc, Grid gf,
                            const int* __restrict__ eopf, int exact, int* __restrict__ flag) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  const long long cell = epc[k];
  const long long sxy = (long long)gc.py * gc.px;
  const int z = (int)(cell / sxy) - 1, x = (int)((cell % sxy) / gc.py) - 1, y = (int)(cell % gc.py) - 1;
  int u = 0;
  for (int j = 0; j < 8; ++j) {
    const int e = eopf[gf.pidx(2 * y + (j & 1), 2 * x + ((j >> 1) & 1), 2 * z + (j >> 2))];
    if (e > 0) u = 1;                 // present child with a void grandchild (B3)
    if (exact && e < 0) u = 1;        // absent child: not counted by the reference
  }
  flag[k] = u;
}

__global__ void k_scatter_eop(const long long* __restrict__ ep, long long ne,
                              const int* __restrict__ v, int* __restrict__ eop) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ne) eop[ep[k]] = v[k];
}

// Galerkin product for one coarse element per block
__device__ void galerkin_block(const double* const kids[8], const double kscale[8], int exact,
                               double* __restrict__ out) {
  __shared__ double Kc[576];
  __shared__ double T[576];
  double acc[3] = {0.0, 0.0, 0.0};
  const int tid = threadIdx.x;
  for (int j = 0; j < 8; ++j) {
    const double* K = kids[j];
    if (!K) continue;                          // uniform across the block
    const int s = exact ? c_colmajor_to_slot[j] : j;   // B12: column j at slot j
    const double sc = kscale[j];
    for (int e = tid; e < 576; e += 256) Kc[e] = K[e] * sc;
    __syncthreads();
    // T = Kc * kron(W_s, I3): T[(3i+c)][3b+d] = sum_jj Kc[(3i+c)][3jj+d] W[s][jj][b]
    for (int e = tid; e < 576; e += 256) {
      const int row = e / 24, col = e % 24, b = col / 3, d = col % 3;
      double v = 0.0;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) v += Kc[row * 24 + 3 * jj + d] * c_W[s][jj][b];
      T[e] = v;
    }
    __syncthreads();
    // out += kron(W_s, I3)' * T: out[(3a+c)][col] = sum_i W[s][i][a] T[(3i+c)][col]
    int q = 0;
    for (int e = tid; e < 576; e += 256, ++q) {
      const int row = e / 24, col = e % 24, a = row / 3, c = row % 3;
      double v = 0.0;
#pragma unroll
      for (int i = 0; i < 8; ++i) v += c_W[s][i][a] * T[(3 * i + c) * 24 + col];
      acc[q] += v;
    }
    __syncthreads();
  }
  int q = 0;
  for (int e = tid; e < 576; e += 256, ++q) out[e] = acc[q];
}

__global__ __launch_bounds__(256) void k_galerkin_cells(
    const long long* __restrict__ cells, long long njobs, Grid gc, Grid gf,
    const uint8_t* __restrict__ esf, const int* __restrict__ eopf, const double* __restrict__ Kef,
    const double* __restrict__ Ksf, const double* __restrict__ emodf, int exact,
    double* __restrict__ out) {
  const long long job = blockIdx.x;
  if (job >= njobs) return;
  const long long cell = cells[job];
  const long long sxy = (long long)gc.py * gc.px;
  const int z = (int)(cell / sxy) - 1, x = (int)((cell % sxy) / gc.py) - 1, y = (int)(cell % gc.py) - 1;
  const double* kids[8];
  double ks[8];
  for (int j = 0; j < 8; ++j) {
    const long long fc = gf.pidx(2 * y + (j & 1), 2 * x + ((j >> 1) & 1), 2 * z + (j >> 2));
    ks[j] = 1.0;
    if (!eopf) {
      kids[j] = esf[fc] ? Kef : nullptr;
      if (emodf) ks[j] = emodf[fc];
    } else {
      const int e = eopf[fc];
      kids[j] = e < 0 ? nullptr : (e == 0 ? Kef : Ksf + (long long)(e - 1) * 576);
    }
  }
  galerkin_block(kids, ks, exact, out + job * 576);
}

__global__ __launch_bounds__(256) void k_galerkin_masks(const double* __restrict__ Kef, int exact,
                                                        double* __restrict__ table) {
  const int m = blockIdx.x + 1;                 // masks 1..255
  const double* kids[8];
  double ks[8];
  for (int j = 0; j < 8; ++j) {
    kids[j] = ((m >> j) & 1) ? Kef : nullptr;
    ks[j] = 1.0;
  }
  galerkin_block(kids, ks, exact, table + (long long)m * 576);
}

__global__ __launch_bounds__(256) void k_inverse_diag(Grid g, const TileRec* __restrict__ tiles,
                                                      long long nt, const uint32_t* __restrict__ flags,
                                                      const int* __restrict__ eop,
                                                      const double* __restrict__ Ke,
                                                      const double* __restrict__ Ks,
                                                      const double* __restrict__ emod,
                                                      double* __restrict__ invd) {
  long long n;
  int yp, xp, zp;
  if (!lane_pos(g, tiles, nt, n, yp, xp, zp)) return;
  const uint32_t f = flags[n];
  if (!(f & F_ELEM_MASK)) return;
  const long long sx = g.py, sz = (long long)g.py * g.px;
  double d[3] = {0.0, 0.0, 0.0};
  for (int p = 0; p < 8; ++p) {
    if (!((f >> p) & 1u)) continue;
    const long long cell = n - (1 - (p & 1)) - (1 - ((p >> 1) & 1)) * sx - (1 - (p >> 2)) * sz;
    const double* K = Ke;
    double s = 1.0;
    if (eop) {
      const int e = eop[cell];
      if (e > 0) K = Ks + (long long)(e - 1) * 576;
    }
    if (emod) s = emod[cell];
    const int ln = node_local_in(p);
    for (int c = 0; c < 3; ++c) d[c] += s * K[(3 * ln + c) * 24 + 3 * ln + c];
  }
  for (int c = 0; c < 3; ++c) invd[c * g.N + n] = 1.0 / d[c];
}

constexpr int BS = 256;

}  // namespace

void set_galerkin_weights(const double W[8][8][8], hipStream_t s) {
  GMGS_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_W), W, sizeof(double) * 512, 0,
                                  hipMemcpyHostToDevice, s));
}

long long build_tiles(const Grid& g, const uint32_t* flags, DevBuf& tiles, hipStream_t s) {
  const int nty = g.py / TILE_Y, ntx = (g.px + TILE_X - 1) / TILE_X;
  const long long nt = (long long)nty * ntx * g.pz;
  DevBuf mark, scan;
  mark.alloc(sizeof(int) * nt);
  scan.alloc(sizeof(int) * nt);
  k_tile_mark<<<grid1d(nt, BS), BS, 0, s>>>(g, flags, nty, ntx, nt, mark.as<int>());
  GMGS_LAUNCH_CHECK();
  long long total = exclusive_scan(mark.as<int>(), scan.as<int>(), nt, s);
  tiles.alloc(sizeof(TileRec) * (total > 0 ? total : 1));
  k_tile_emit<<<grid1d(nt, BS), BS, 0, s>>>(mark.as<int>(), scan.as<int>(), nty, ntx, nt,
                                            tiles.as<TileRec>());
  GMGS_LAUNCH_CHECK();
  GMGS_HIP(hipStreamSynchronize(s));
  return total;
}

void clear_fixed(const Grid& g, uint32_t* flags, hipStream_t s) {
  k_clear_fixed<<<grid1d(g.N, BS), BS, 0, s>>>(flags, g.N);
  GMGS_LAUNCH_CHECK();
}

void set_fixed_nodes(const long long* node_pidx, const int* ids, long long k, uint32_t* flags,
                     hipStream_t s) {
  if (k <= 0) return;
  k_set_fixed<<<grid1d(k, BS), BS, 0, s>>>(node_pidx, ids, k, flags);
  GMGS_LAUNCH_CHECK();
}

void restrict_fixed(const Grid& gc, const TileRec* tiles_c, long long nt, uint32_t* flags_c,
                    const Grid& gf, const uint32_t* flags_f, hipStream_t s) {
  if (nt <= 0) return;
  k_restrict_fixed<<<apply_blocks(nt), 256, 0, s>>>(gc, tiles_c, nt, flags_c, gf, flags_f);
  GMGS_LAUNCH_CHECK();
}

void mark_fast(const Grid& g, const TileRec* tiles, long long nt, uint32_t* flags, const int* eop,
               bool allow, hipStream_t s) {
  if (nt <= 0) return;
  k_mark_fast<<<apply_blocks(nt), 256, 0, s>>>(g, tiles, nt, flags, eop, allow ? 1 : 0);
  GMGS_LAUNCH_CHECK();
}

void child_masks(const long long* elem_pidx_c, long long ne, const Grid& gc, const Grid& gf,
                 const uint8_t* esolid_f, int* mask, hipStream_t s) {
  if (ne <= 0) return;
  k_child_masks<<<grid1d(ne, BS), BS, 0, s>>>(elem_pidx_c, ne, gc, gf, esolid_f, mask);
  GMGS_LAUNCH_CHECK();
}

void galerkin_cells(const long long* cells, long long njobs, const Grid& gc, const Grid& gf,
                    const uint8_t* esolid_f, const int* eop_f, const double* Ke_f,
                    const double* Ks_f, const double* emod_f, int exact, double* Ks_out,
                    hipStream_t s) {
  if (njobs <= 0) return;
  const long long chunk = 1 << 20;
  for (long long a = 0; a < njobs; a += chunk) {
    const long long n = std::min(chunk, njobs - a);
    k_galerkin_cells<<<(unsigned)n, 256, 0, s>>>(cells + a, n, gc, gf, esolid_f, eop_f, Ke_f, Ks_f,
                                                 emod_f, exact, Ks_out + a * 576);
    GMGS_LAUNCH_CHECK();
  }
}

void galerkin_mask_table(const double* Ke_f, int exact, double* table_out, hipStream_t s) {
  k_galerkin_masks<<<255, 256, 0, s>>>(Ke_f, exact, table_out);
  GMGS_LAUNCH_CHECK();
}

void unique_flags_l3(const long long* elem_pidx_c, long long ne, const Grid& gc, const Grid& gf,
                     const int* eop_f, int exact, int* flag, hipStream_t s) {
  if (ne <= 0) return;
  k_unique_l3<<<grid1d(ne, BS), BS, 0, s>>>(elem_pidx_c, ne, gc, gf, eop_f, exact, flag);
  GMGS_LAUNCH_CHECK();
}

void scatter_eop(const long long* elem_pidx, long long ne, const int* value, int* eop,
                 hipStream_t s) {
  if (ne <= 0) return;
  k_scatter_eop<<<grid1d(ne, BS), BS, 0, s>>>(elem_pidx, ne, value, eop);
  GMGS_LAUNCH_CHECK();
}

bool fast_inverse_diagonal(const Grid& g, const TileRec* tiles, long long nt, const uint32_t* flags,
                           const float* invd, float out[3], hipStream_t s) {
  if (nt <= 0) return false;
  const unsigned init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  unsigned h[6];
  DevBuf mm;
  mm.alloc(sizeof(init));
  GMGS_HIP(hipMemcpyAsync(mm.as<void>(), init, sizeof(init), hipMemcpyHostToDevice, s));
  k_fast_invd_range<<<apply_blocks(nt), 256, 0, s>>>(g, tiles, nt, flags, invd, mm.as<unsigned>());
  GMGS_LAUNCH_CHECK();
  GMGS_HIP(hipMemcpyAsync(h, mm.as<void>(), sizeof(h), hipMemcpyDeviceToHost, s));
  GMGS_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < 3; ++c) {
    if (h[c] != h[3 + c] || h[c] == 0u || h[c] == 0xffffffffu) return false;
    unsigned b = h[c];
    std::memcpy(&out[c], &b, sizeof(float));
    if (!(out[c] > 0.0f)) return false;
  }
  return true;
}

void inverse_diagonal(const Grid& g, const TileRec* tiles, long long nt, const uint32_t* flags,
                      const int* eop, const double* Ke, const double* Ks, const double* emod,
                      double* invd, hipStream_t s) {
  if (nt <= 0) return;
  k_inverse_diag<<<apply_blocks(nt), 256, 0, s>>>(g, tiles, nt, flags, eop, Ke, Ks, emod, invd);
  GMGS_LAUNCH_CHECK();
}

}  // namespace gmgs
