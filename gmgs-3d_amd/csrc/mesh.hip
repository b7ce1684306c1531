// Device construction of the mesh hierarchy.
//
// Finest level: common/DiscretizeDesignDomain.m:68-116.
// Coarser levels: kernel/BuildingMeshHierarchy.m:26-140.
// Every compaction the reference performs with find()/unique()/intersect()
// over sorted full-grid ids is an order-preserving stream compaction over
// MATLAB's column-major linear order, done here with a device prefix sum.
#include "mesh.h"

namespace gmgs {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;                 // per thread
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__global__ void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                             int* __restrict__ tile_sums, long long n) {
  __shared__ int sh[SCAN_BLOCK];
  const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int sum = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long long k = base + i;
    v[i] = k < n ? in[k] : 0;
    sum += v[i];
  }
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - sum;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    long long k = base + i;
    if (k < n) out[k] = run;
    run += v[i];
  }
  if (threadIdx.x == SCAN_BLOCK - 1) tile_sums[blockIdx.x] = sh[SCAN_BLOCK - 1];
}

__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ offs, long long n) {
  const long long base = (long long)blockIdx.x * SCAN_TILE;
  const int o = offs[blockIdx.x];
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_BLOCK) {
    long long k = base + i;
    if (k < n) out[k] += o;
  }
}

// ---------------------------------------------------------------- level 1
__global__ void k_level1_esolid(const int* __restrict__ vox, int nely, int nelx, int nelz,
                                Grid g, uint8_t* __restrict__ esolid) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)g.ey * g.ex * g.ez;
  if (i >= tot) return;
  int y = (int)(i % g.ey), x = (int)((i / g.ey) % g.ex), z = (int)(i / ((long long)g.ey * g.ex));
  uint8_t s = 0;
  if (y < nely && x < nelx && z < nelz)
    s = vox[y + (long long)nely * (x + (long long)nelx * z)] == 1 ? 1 : 0;
  esolid[g.pidx(y, x, z)] = s;
}

// coarse cell is non-empty iff any of its 8 children is solid
// (BuildingMeshHierarchy.m:70)
__global__ void k_coarse_esolid(const uint8_t* __restrict__ ef, Grid gf, Grid g,
                                uint8_t* __restrict__ ec) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)g.ey * g.ex * g.ez;
  if (i >= tot) return;
  int y = (int)(i % g.ey), x = (int)((i / g.ey) % g.ex), z = (int)(i / ((long long)g.ey * g.ex));
  uint8_t s = 0;
  for (int c = 0; c < 8; ++c)
    s |= ef[gf.pidx(2 * y + (c & 1), 2 * x + ((c >> 1) & 1), 2 * z + (c >> 2))];
  ec[g.pidx(y, x, z)] = s ? 1 : 0;
}

// node flag: bit p set iff the element at position p around the node is solid
__global__ void k_node_flags(const uint8_t* __restrict__ es, Grid g, uint32_t* __restrict__ flags) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)g.ny * g.nx * g.nz;
  if (i >= tot) return;
  int y = (int)(i % g.ny), x = (int)((i / g.ny) % g.nx), z = (int)(i / ((long long)g.ny * g.nx));
  long long n = g.pidx(y, x, z);
  uint32_t m = 0;
  for (int p = 0; p < 8; ++p) {
    long long c = n - (1 - (p & 1)) - (long long)g.py * (1 - ((p >> 1) & 1))
                  - (long long)g.py * g.px * (1 - (p >> 2));
    if (es[c]) m |= 1u << p;
  }
  flags[n] = m;
}

__global__ void k_mark_cells(const uint8_t* __restrict__ es, Grid g, int* __restrict__ mark) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)g.ey * g.ex * g.ez;
  if (i >= tot) return;
  int y = (int)(i % g.ey), x = (int)((i / g.ey) % g.ex), z = (int)(i / ((long long)g.ey * g.ex));
  mark[i] = es[g.pidx(y, x, z)] ? 1 : 0;
}

__global__ void k_mark_nodes(const uint32_t* __restrict__ flags, Grid g, int* __restrict__ mark) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)g.ny * g.nx * g.nz;
  if (i >= tot) return;
  int y = (int)(i % g.ny), x = (int)((i / g.ny) % g.nx), z = (int)(i / ((long long)g.ny * g.nx));
  mark[i] = (flags[g.pidx(y, x, z)] & F_ELEM_MASK) ? 1 : 0;
}

// back[scan[i]] = i+1 where marked; forward[i] = scan[i]+1 or 0
__global__ void k_emit_maps(const int* __restrict__ mark, const int* __restrict__ scan, long long n,
                            int* __restrict__ back, int* __restrict__ forward) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int m = mark[i];
  if (m) back[scan[i]] = (int)(i + 1);
  if (forward) forward[i] = m ? scan[i] + 1 : 0;
}

// compact element -> padded cell index and eNodMat (DiscretizeDesignDomain.m:76-93)
__global__ void k_elements(const int* __restrict__ back, long long ne, Grid g,
                           const int* __restrict__ nodFwd, long long* __restrict__ epidx,
                           int* __restrict__ enod) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  long long f = back[k] - 1;
  int y = (int)(f % g.ey), x = (int)((f / g.ey) % g.ex), z = (int)(f / ((long long)g.ey * g.ex));
  epidx[k] = g.pidx(y, x, z);
  if (enod) {
    for (int m = 0; m < 8; ++m) {
      long long nid = (long long)(y + hex_ro(m)) + (long long)g.ny * ((x + hex_xo(m))
                      + (long long)g.nx * (z + hex_zo(m)));
      enod[k + ne * m] = nodFwd[nid];
    }
  }
}

// per compact node: padded index, incident-element count, boundary mark and,
// on coarse levels, the finest-grid id (BuildingMeshHierarchy.m:119-122)
__global__ void k_nodes(const int* __restrict__ full, long long nn, Grid g,
                        const uint32_t* __restrict__ flags, long long* __restrict__ npidx,
                        int* __restrict__ count, int* __restrict__ bmark, int* __restrict__ finest,
                        int stride, int fny, int fnx) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  long long f = full[k] - 1;
  int y = (int)(f % g.ny), x = (int)((f / g.ny) % g.nx), z = (int)(f / ((long long)g.ny * g.nx));
  long long p = g.pidx(y, x, z);
  npidx[k] = p;
  int c = __popc(flags[p] & F_ELEM_MASK);
  if (count) count[k] = c;
  if (bmark) bmark[k] = c < 8 ? 1 : 0;
  if (finest)
    finest[k] = (int)((long long)y * stride + (long long)fny * ((long long)x * stride
                      + (long long)fnx * ((long long)z * stride)) + 1);
}

__global__ void k_elem_boundary(const int* __restrict__ enod, long long ne,
                                const int* __restrict__ count, int* __restrict__ mark) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  int b = 0;
  for (int m = 0; m < 8; ++m) b |= count[enod[k + ne * m] - 1] < 8;
  mark[k] = b;
}

// elementUpwardMap: children in column-major (y,x,z) order, BuildingMeshHierarchy.m:45-52
__global__ void k_upward(const int* __restrict__ back, long long ne, Grid g, Grid gf,
                         const int* __restrict__ fineFwd, int* __restrict__ up) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  long long f = back[k] - 1;
  int y = (int)(f % g.ey), x = (int)((f / g.ey) % g.ex), z = (int)(f / ((long long)g.ey * g.ex));
  for (int c = 0; c < 8; ++c) {
    int cy = 2 * y + (c & 1), cx = 2 * x + ((c >> 1) & 1), cz = 2 * z + (c >> 2);
    up[k + ne * c] = fineFwd[cy + (long long)gf.ey * (cx + (long long)gf.ex * cz)];
  }
}

// intermediate nodes: fine nodes in the closure of a non-empty coarse cell,
// with the number of such cells (BuildingMeshHierarchy.m:70-84)
__global__ void k_mark_intermediate(const uint8_t* __restrict__ ec, Grid g, Grid gf,
                                    int* __restrict__ mark, double* __restrict__ coef_full) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)gf.ny * gf.nx * gf.nz;
  if (i >= tot) return;
  int y = (int)(i % gf.ny), x = (int)((i / gf.ny) % gf.nx), z = (int)(i / ((long long)gf.ny * gf.nx));
  int cnt = 0;
  // a fine node at even coordinate 2k lies on coarse cells k-1 and k; at odd 2k+1 only on k
  int ylo = (y % 2 == 0) ? y / 2 - 1 : y / 2, yhi = y / 2;
  int xlo = (x % 2 == 0) ? x / 2 - 1 : x / 2, xhi = x / 2;
  int zlo = (z % 2 == 0) ? z / 2 - 1 : z / 2, zhi = z / 2;
  for (int cz = zlo; cz <= zhi; ++cz) {
    if (cz < 0 || cz >= g.ez) continue;
    for (int cx = xlo; cx <= xhi; ++cx) {
      if (cx < 0 || cx >= g.ex) continue;
      for (int cy = ylo; cy <= yhi; ++cy) {
        if (cy < 0 || cy >= g.ey) continue;
        cnt += ec[g.pidx(cy, cx, cz)] ? 1 : 0;
      }
    }
  }
  mark[i] = cnt > 0 ? 1 : 0;
  coef_full[i] = (double)cnt;
}

__global__ void k_gather_coef(const int* __restrict__ mark, const int* __restrict__ scan,
                              const double* __restrict__ coef_full, long long n,
                              double* __restrict__ coef) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mark[i]) coef[scan[i]] = coef_full[i];
}

// transferMat (27 x ne): intermediate id of the 27 fine nodes of each coarse
// element, 3x3x3 column-major (BuildingMeshHierarchy.m:49-54,75-78)
__global__ void k_transfer(const int* __restrict__ back, long long ne, Grid g, Grid gf,
                           const int* __restrict__ interFwd, int* __restrict__ tm) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ne) return;
  long long f = back[k] - 1;
  int y = (int)(f % g.ey), x = (int)((f / g.ey) % g.ex), z = (int)(f / ((long long)g.ey * g.ex));
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a) {
        long long nid = (long long)(2 * y + a) + (long long)gf.ny * ((2 * x + b)
                        + (long long)gf.nx * (2 * z + c));
        tm[(a + 3 * b + 9 * c) + 27 * k] = interFwd[nid];
      }
}

// solidNodeMapCoarser2Finer: intermediate id of each fine solid node in fine
// compact order (the ia output of intersect(), BuildingMeshHierarchy.m:85-88)
__global__ void k_solid_map(const int* __restrict__ fineFull, long long nnf,
                            const int* __restrict__ interFwd, int* __restrict__ out) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnf) return;
  out[k] = interFwd[fineFull[k] - 1];
}

constexpr int BS = 256;

struct Compaction {
  DevBuf mark, scan;
  long long total = 0;
};

void compact(Compaction& c, long long n, hipStream_t s) {
  c.total = exclusive_scan(c.mark.as<int>(), c.scan.as<int>(), n, s);
}

void finish_level(LevelMesh& L, const Grid& finest, bool tables, hipStream_t s) {
  const Grid& g = L.g;
  // node flags
  L.flags.alloc_zero(sizeof(uint32_t) * g.N, s);
  long long nodes_full = (long long)g.ny * g.nx * g.nz;
  long long cells_full = (long long)g.ey * g.ex * g.ez;
  k_node_flags<<<grid1d(nodes_full, BS), BS, 0, s>>>(L.esolid.as<uint8_t>(), g, L.flags.as<uint32_t>());
  GMGS_LAUNCH_CHECK();
  // elements
  Compaction ce;
  ce.mark.alloc(sizeof(int) * cells_full);
  ce.scan.alloc(sizeof(int) * cells_full);
  k_mark_cells<<<grid1d(cells_full, BS), BS, 0, s>>>(L.esolid.as<uint8_t>(), g, ce.mark.as<int>());
  GMGS_LAUNCH_CHECK();
  compact(ce, cells_full, s);
  L.ne = ce.total;
  L.eleMapBack.alloc(sizeof(int) * L.ne);
  L.eleMapForward.alloc(sizeof(int) * cells_full);
  k_emit_maps<<<grid1d(cells_full, BS), BS, 0, s>>>(ce.mark.as<int>(), ce.scan.as<int>(), cells_full,
                                                    L.eleMapBack.as<int>(), L.eleMapForward.as<int>());
  GMGS_LAUNCH_CHECK();
  // nodes
  Compaction cn;
  cn.mark.alloc(sizeof(int) * nodes_full);
  cn.scan.alloc(sizeof(int) * nodes_full);
  k_mark_nodes<<<grid1d(nodes_full, BS), BS, 0, s>>>(L.flags.as<uint32_t>(), g, cn.mark.as<int>());
  GMGS_LAUNCH_CHECK();
  compact(cn, nodes_full, s);
  L.nn = cn.total;
  L.nodFull.alloc(sizeof(int) * L.nn);
  L.nodMapForward.alloc(sizeof(int) * nodes_full);
  k_emit_maps<<<grid1d(nodes_full, BS), BS, 0, s>>>(cn.mark.as<int>(), cn.scan.as<int>(), nodes_full,
                                                    L.nodFull.as<int>(), L.nodMapForward.as<int>());
  GMGS_LAUNCH_CHECK();
  L.elem_pidx.alloc(sizeof(long long) * L.ne);
  if (tables) L.eNodMat.alloc(sizeof(int) * L.ne * 8);
  k_elements<<<grid1d(L.ne, BS), BS, 0, s>>>(L.eleMapBack.as<int>(), L.ne, g,
                                             L.nodMapForward.as<int>(), L.elem_pidx.as<long long>(),
                                             tables ? L.eNodMat.as<int>() : nullptr);
  GMGS_LAUNCH_CHECK();
  L.node_pidx.alloc(sizeof(long long) * L.nn);
  L.numNod2Eles.alloc(sizeof(int) * L.nn);
  Compaction cb;
  cb.mark.alloc(sizeof(int) * L.nn);
  cb.scan.alloc(sizeof(int) * L.nn);
  int stride = 1 << (L.level - 1);
  if (L.level > 1) L.nodMapBack.alloc(sizeof(int) * L.nn);
  k_nodes<<<grid1d(L.nn, BS), BS, 0, s>>>(L.nodFull.as<int>(), L.nn, g, L.flags.as<uint32_t>(),
                                          L.node_pidx.as<long long>(), L.numNod2Eles.as<int>(),
                                          cb.mark.as<int>(),
                                          L.level > 1 ? L.nodMapBack.as<int>() : nullptr, stride,
                                          finest.ny, finest.nx);
  GMGS_LAUNCH_CHECK();
  if (tables) {
    compact(cb, L.nn, s);
    L.n_nob = cb.total;
    L.nodesOnBoundary.alloc(sizeof(int) * L.n_nob);
    k_emit_maps<<<grid1d(L.nn, BS), BS, 0, s>>>(cb.mark.as<int>(), cb.scan.as<int>(), L.nn,
                                                L.nodesOnBoundary.as<int>(), nullptr);
    GMGS_LAUNCH_CHECK();
    Compaction cq;
    cq.mark.alloc(sizeof(int) * L.ne);
    cq.scan.alloc(sizeof(int) * L.ne);
    k_elem_boundary<<<grid1d(L.ne, BS), BS, 0, s>>>(L.eNodMat.as<int>(), L.ne, L.numNod2Eles.as<int>(),
                                                    cq.mark.as<int>());
    GMGS_LAUNCH_CHECK();
    compact(cq, L.ne, s);
    L.n_eob = cq.total;
    L.elementsOnBoundary.alloc(sizeof(int) * L.n_eob);
    k_emit_maps<<<grid1d(L.ne, BS), BS, 0, s>>>(cq.mark.as<int>(), cq.scan.as<int>(), L.ne,
                                                L.elementsOnBoundary.as<int>(), nullptr);
    GMGS_LAUNCH_CHECK();
  }
  GMGS_HIP(hipStreamSynchronize(s));   // keep temporaries alive until done
}

}  // namespace

long long exclusive_scan(const int* in, int* out, long long n, hipStream_t s) {
  if (n <= 0) return 0;
  long long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  DevBuf sums, sums_scan;
  sums.alloc(sizeof(int) * tiles);
  sums_scan.alloc(sizeof(int) * tiles);
  k_scan_tiles<<<(unsigned)tiles, SCAN_BLOCK, 0, s>>>(in, out, sums.as<int>(), n);
  GMGS_LAUNCH_CHECK();
  long long total;
  if (tiles > 1) {
    total = exclusive_scan(sums.as<int>(), sums_scan.as<int>(), tiles, s);
    k_scan_add<<<(unsigned)tiles, SCAN_BLOCK, 0, s>>>(out, sums_scan.as<int>(), n);
    GMGS_LAUNCH_CHECK();
  } else {
    int t = 0;
    GMGS_HIP(hipMemcpyAsync(&t, sums.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
    GMGS_HIP(hipStreamSynchronize(s));
    total = t;
  }
  GMGS_HIP(hipStreamSynchronize(s));
  return total;
}

void build_level1(LevelMesh& L, const int* d_vox, int nely, int nelx, int nelz, bool tables,
                  hipStream_t s) {
  const Grid& g = L.g;
  L.level = 1;
  L.esolid.alloc_zero(g.N, s);
  long long cells = (long long)g.ey * g.ex * g.ez;
  k_level1_esolid<<<grid1d(cells, BS), BS, 0, s>>>(d_vox, nely, nelx, nelz, g, L.esolid.as<uint8_t>());
  GMGS_LAUNCH_CHECK();
  finish_level(L, g, tables, s);
  // level-1 nodMapBack holds level-1 full ids (DiscretizeDesignDomain.m:85)
  L.nodMapBack.alloc(sizeof(int) * L.nn);
  GMGS_HIP(hipMemcpyAsync(L.nodMapBack.as<int>(), L.nodFull.as<int>(), sizeof(int) * L.nn,
                          hipMemcpyDeviceToDevice, s));
  GMGS_HIP(hipStreamSynchronize(s));
}

void build_coarse(LevelMesh& C, const LevelMesh& F, const Grid& finest, bool tables, hipStream_t s) {
  const Grid& g = C.g;
  const Grid& gf = F.g;
  C.level = F.level + 1;
  C.esolid.alloc_zero(g.N, s);
  long long cells = (long long)g.ey * g.ex * g.ez;
  k_coarse_esolid<<<grid1d(cells, BS), BS, 0, s>>>(F.esolid.as<uint8_t>(), gf, g, C.esolid.as<uint8_t>());
  GMGS_LAUNCH_CHECK();
  finish_level(C, finest, tables, s);
  if (!tables) return;
  C.elementUpwardMap.alloc(sizeof(int) * C.ne * 8);
  k_upward<<<grid1d(C.ne, BS), BS, 0, s>>>(C.eleMapBack.as<int>(), C.ne, g, gf,
                                           F.eleMapForward.as<int>(), C.elementUpwardMap.as<int>());
  GMGS_LAUNCH_CHECK();
  long long fnodes = (long long)gf.ny * gf.nx * gf.nz;
  Compaction ci;
  ci.mark.alloc(sizeof(int) * fnodes);
  ci.scan.alloc(sizeof(int) * fnodes);
  DevBuf coef_full, interFwd;
  coef_full.alloc(sizeof(double) * fnodes);
  k_mark_intermediate<<<grid1d(fnodes, BS), BS, 0, s>>>(C.esolid.as<uint8_t>(), g, gf, ci.mark.as<int>(),
                                                        coef_full.as<double>());
  GMGS_LAUNCH_CHECK();
  compact(ci, fnodes, s);
  C.intermediateNumNodes = ci.total;
  C.transferMatCoeffi.alloc(sizeof(double) * ci.total);
  k_gather_coef<<<grid1d(fnodes, BS), BS, 0, s>>>(ci.mark.as<int>(), ci.scan.as<int>(),
                                                  coef_full.as<double>(), fnodes,
                                                  C.transferMatCoeffi.as<double>());
  GMGS_LAUNCH_CHECK();
  interFwd.alloc(sizeof(int) * fnodes);
  DevBuf dummy_back;
  dummy_back.alloc(sizeof(int) * ci.total);
  k_emit_maps<<<grid1d(fnodes, BS), BS, 0, s>>>(ci.mark.as<int>(), ci.scan.as<int>(), fnodes,
                                                dummy_back.as<int>(), interFwd.as<int>());
  GMGS_LAUNCH_CHECK();
  C.transferMat.alloc(sizeof(int) * 27 * C.ne);
  k_transfer<<<grid1d(C.ne, BS), BS, 0, s>>>(C.eleMapBack.as<int>(), C.ne, g, gf, interFwd.as<int>(),
                                             C.transferMat.as<int>());
  GMGS_LAUNCH_CHECK();
  C.solidNodeMapCoarser2Finer.alloc(sizeof(int) * F.nn);
  k_solid_map<<<grid1d(F.nn, BS), BS, 0, s>>>(F.nodFull.as<int>(), F.nn, interFwd.as<int>(),
                                              C.solidNodeMapCoarser2Finer.as<int>());
  GMGS_LAUNCH_CHECK();
  GMGS_HIP(hipStreamSynchronize(s));
}

}  // namespace gmgs
