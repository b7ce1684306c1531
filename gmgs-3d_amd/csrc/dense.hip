// Coarsest-level direct solve on the device.
#include "dense.h"
#include "launch.h"
#include "mesh.h"

namespace gmgs {

namespace {

constexpr int BS = 256;
constexpr int LD = DB + 1;      // LDS row pitch (bank-conflict padding)

__global__ void k_free_mark(const long long* __restrict__ npidx, long long nn, long long N,
                            const uint32_t* __restrict__ flags, int* __restrict__ mark) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nn) return;
  const long long k = i / 3;
  const int c = (int)(i % 3);
  mark[i] = (flags[npidx[k]] >> (F_FIX_SHIFT + c)) & 1u ? 0 : 1;
}

__global__ void k_free_emit(const long long* __restrict__ npidx, long long nn, long long N,
                            const int* __restrict__ mark, const int* __restrict__ scan,
                            long long* __restrict__ code, int* __restrict__ dofmap) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nn) return;
  if (!mark[i]) return;
  const long long k = i / 3;
  const int c = (int)(i % 3);
  const long long cd = c * N + npidx[k];
  code[scan[i]] = cd;
  dofmap[cd] = scan[i];
}

__global__ void k_assemble(Grid g, const uint32_t* __restrict__ flags, const int* __restrict__ eop,
                           const double* __restrict__ Ke, const double* __restrict__ Ks,
                           const long long* __restrict__ code, const int* __restrict__ dofmap,
                           long long nfree, long long nfp, double* __restrict__ A) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfp) return;
  double* row = A + i * nfp;
  if (i >= nfree) {
    row[i] = 1.0;
    return;
  }
  const long long cd = code[i];
  const int c = (int)(cd / g.N);
  const long long n = cd % g.N;
  const uint32_t f = flags[n];
  const long long sx = g.py, sz = (long long)g.py * g.px;
  for (int p = 0; p < 8; ++p) {
    if (!((f >> p) & 1u)) continue;
    const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = p >> 2;
    const long long cell = n - (1 - pyb) - (1 - pxb) * sx - (1 - pzb) * sz;
    const double* K = Ke;
    if (eop) {
      const int e = eop[cell];
      if (e > 0) K = Ks + (long long)(e - 1) * 576;
    }
    const int ln = node_local_in(p);
    for (int m = 0; m < 8; ++m) {
      const long long mm = n + (pyb - 1 + hex_ro(m)) + (pxb - 1 + hex_xo(m)) * sx
                           + (pzb - 1 + hex_zo(m)) * sz;
      for (int d = 0; d < 3; ++d) {
        const int col = dofmap[d * g.N + mm];
        if (col >= 0) row[col] += K[(3 * ln + c) * 24 + 3 * m + d];
      }
    }
  }
}

// A pivot at or below this fraction of its row's diagonal before
// factoring marks the matrix singular (a rigid-body mode left free, e.g. a
// solid component without a fixed node): exactly singular modes leave
// pivots at rounding level (~1e-16), while a constrained body's scaled
// stiffness keeps its pivots above 1/cond(D^-1/2 K D^-1/2).
constexpr double kPivotTol = 1e-12;

__global__ void k_diag_copy(const double* __restrict__ A, long long n, double* __restrict__ d) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = A[i * n + i];
}

__global__ __launch_bounds__(256) void k_potrf_diag(double* __restrict__ A, long long n, int k,
                                                    const double* __restrict__ diag0,
                                                    int* __restrict__ info) {
  __shared__ double L[DB][LD];
  const int tid = threadIdx.x;
  const long long o = (long long)k * DB;
  for (int e = tid; e < DB * DB; e += BS) {
    const int i = e / DB, j = e % DB;
    L[i][j] = A[(o + i) * n + o + j];
  }
  __syncthreads();
  for (int j = 0; j < DB; ++j) {
    if (tid == 0) {
      double d = L[j][j];
      if (!(d > kPivotTol * diag0[o + j])) {
        if (*info == 0) *info = (int)(o + j + 1);
        d = 1.0;
      }
      L[j][j] = sqrt(d);
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < DB; i += BS) L[i][j] /= L[j][j];
    __syncthreads();
    for (int e = tid; e < DB * DB; e += BS) {
      const int i = e / DB, c = e % DB;
      if (c > j && i >= c) L[i][c] -= L[i][j] * L[c][j];
    }
    __syncthreads();
  }
  for (int e = tid; e < DB * DB; e += BS) {
    const int i = e / DB, j = e % DB;
    A[(o + i) * n + o + j] = j <= i ? L[i][j] : 0.0;
  }
}

// A_ik = A_ik * inv(L_kk)' for block rows i > k
__global__ __launch_bounds__(256) void k_trsm(double* __restrict__ A, long long n, int k) {
  __shared__ double L[DB][LD];
  __shared__ double B[DB][LD];
  const int tid = threadIdx.x;
  const long long ok = (long long)k * DB, oi = (long long)(k + 1 + blockIdx.x) * DB;
  for (int e = tid; e < DB * DB; e += BS) {
    const int i = e / DB, j = e % DB;
    L[i][j] = A[(ok + i) * n + ok + j];
    B[i][j] = A[(oi + i) * n + ok + j];
  }
  __syncthreads();
  if (tid < DB) {
    const int r = tid;
    for (int j = 0; j < DB; ++j) {
      double v = B[r][j];
      for (int t = 0; t < j; ++t) v -= B[r][t] * L[j][t];
      B[r][j] = v / L[j][j];
    }
  }
  __syncthreads();
  for (int e = tid; e < DB * DB; e += BS) {
    const int i = e / DB, j = e % DB;
    A[(oi + i) * n + ok + j] = B[i][j];
  }
}

__device__ inline void tri_index(long long t, int& a, int& b) {
  // t -> (a, b) with 0 <= b <= a
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((long long)(i + 1) * (i + 2) / 2 <= t) ++i;
  while ((long long)i * (i + 1) / 2 > t) --i;
  a = i;
  b = (int)(t - (long long)i * (i + 1) / 2);
}

// trailing update A_ij -= A_ik A_jk' for k < j <= i
__global__ __launch_bounds__(256) void k_syrk(double* __restrict__ A, long long n, int k) {
  __shared__ double Ai[DB][LD];
  __shared__ double Aj[DB][LD];
  int a, b;
  tri_index(blockIdx.x, a, b);
  const int i = k + 1 + a, j = k + 1 + b;
  const int tid = threadIdx.x;
  const long long ok = (long long)k * DB, oi = (long long)i * DB, oj = (long long)j * DB;
  for (int e = tid; e < DB * DB; e += BS) {
    const int r = e / DB, c = e % DB;
    Ai[r][c] = A[(oi + r) * n + ok + c];
    Aj[r][c] = A[(oj + r) * n + ok + c];
  }
  __syncthreads();
  const int r = tid / 4, c0 = (tid % 4) * 16;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (int t = 0; t < DB; ++t) {
    const double av = Ai[r][t];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] += av * Aj[c0 + q][t];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) A[(oi + r) * n + oj + c0 + q] -= acc[q];
}

// W(i, j) for j <= i:  L_ii W_ij = delta_ij I - sum_{k=j}^{i-1} L_ik W_kj
__global__ __launch_bounds__(256) void k_trtri_row(const double* __restrict__ Lm,
                                                   double* __restrict__ W, long long n, int i) {
  __shared__ double P[DB][LD];
  __shared__ double Q[DB][LD];
  const int j = blockIdx.x;
  const int tid = threadIdx.x;
  const int r = tid / 4, c0 = (tid % 4) * 16;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  const long long oi = (long long)i * DB, oj = (long long)j * DB;
  for (int k = j; k < i; ++k) {
    const long long ok = (long long)k * DB;
    __syncthreads();
    for (int e = tid; e < DB * DB; e += BS) {
      const int a = e / DB, b = e % DB;
      P[a][b] = Lm[(oi + a) * n + ok + b];
      Q[a][b] = W[(ok + a) * n + oj + b];
    }
    __syncthreads();
    for (int t = 0; t < DB; ++t) {
      const double pv = P[r][t];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] += pv * Q[t][c0 + q];
    }
  }
  __syncthreads();
  // P <- L_ii, Q <- T
  for (int e = tid; e < DB * DB; e += BS) {
    const int a = e / DB, b = e % DB;
    P[a][b] = Lm[(oi + a) * n + oi + b];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) Q[r][c0 + q] = ((i == j && r == c0 + q) ? 1.0 : 0.0) - acc[q];
  __syncthreads();
  if (tid < DB) {
    const int c = tid;
    for (int a = 0; a < DB; ++a) {
      double v = Q[a][c];
      for (int t = 0; t < a; ++t) v -= P[a][t] * Q[t][c];
      Q[a][c] = v / P[a][a];
    }
  }
  __syncthreads();
  for (int e = tid; e < DB * DB; e += BS) {
    const int a = e / DB, b = e % DB;
    W[(oi + a) * n + oj + b] = (i == j && b > a) ? 0.0 : Q[a][b];
  }
}

// C(i, j) = sum_{k >= i} W_ki' W_kj for j <= i; mirrored into C(j, i)
__global__ __launch_bounds__(256) void k_wtw(const double* __restrict__ W, double* __restrict__ C,
                                             long long n, int nb) {
  __shared__ double P[DB][LD];
  __shared__ double Q[DB][LD];
  int i, j;
  tri_index(blockIdx.x, i, j);
  const int tid = threadIdx.x;
  const int r = tid / 4, c0 = (tid % 4) * 16;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  const long long oi = (long long)i * DB, oj = (long long)j * DB;
  for (int k = i; k < nb; ++k) {
    const long long ok = (long long)k * DB;
    __syncthreads();
    for (int e = tid; e < DB * DB; e += BS) {
      const int a = e / DB, b = e % DB;
      P[a][b] = W[(ok + a) * n + oi + b];
      Q[a][b] = W[(ok + a) * n + oj + b];
    }
    __syncthreads();
    for (int t = 0; t < DB; ++t) {
      const double pv = P[t][r];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] += pv * Q[t][c0 + q];
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    C[(oi + r) * n + oj + c0 + q] = acc[q];
    C[(oj + c0 + q) * n + oi + r] = acc[q];
  }
}

template <typename T>
struct KGatherFree {
  static constexpr int kThreads = BS;
  const long long* code;
  long long nfree;
  long long nfp;
  const T* r;
  double* b;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    const long long* __restrict__ code = this->code;
    long long nfree = this->nfree;
    long long nfp = this->nfp;
    const T* __restrict__ r = this->r;
    double* __restrict__ b = this->b;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  long long i = (long long)block_ * blockDim.x + threadIdx.x;
  if (i >= nfp) return;
  b[i] = i < nfree ? (double)r[code[i]] : 0.0;
  }
};

constexpr int GV_ROWS = 256;
// rows of K_ff^-1 per partial: short enough that the product's serial sums
// (one per block and row) leave hundreds of blocks in flight
constexpr int GV_CHUNK = 64;

// partial[chunk][i] = sum_{k in chunk} Kinv[k][i] b[k]   (Kinv symmetric)
struct KGemv {
  static constexpr int kThreads = 256;
  const double* Kinv;
  long long nfp;
  const double* b;
  double* partial;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    const double* __restrict__ Kinv = this->Kinv;
    long long nfp = this->nfp;
    const double* __restrict__ b = this->b;
    double* __restrict__ partial = this->partial;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  __shared__ double bs[GV_CHUNK];
  const long long nrb = (nfp + GV_ROWS - 1) / GV_ROWS;   // row blocks; block = row block + nrb chunk
  const long long i = (long long)(block_ % nrb) * GV_ROWS + threadIdx.x;
  const long long chunk = (long long)block_ / nrb;
  const long long k0 = chunk * GV_CHUNK;
  if (threadIdx.x < GV_CHUNK) bs[threadIdx.x] = (k0 + threadIdx.x < nfp) ? b[k0 + threadIdx.x] : 0.0;
  __syncthreads();
  if (i >= nfp) return;
  double acc = 0.0;
  const long long kend = (k0 + GV_CHUNK < nfp) ? GV_CHUNK : nfp - k0;
  for (long long k = 0; k < kend; ++k) acc += Kinv[(k0 + k) * nfp + i] * bs[k];
  partial[chunk * nfp + i] = acc;
  }
};

template <typename T>
struct KScatterFree {
  static constexpr int kThreads = BS;
  const long long* code;
  long long nfree;
  long long nfp;
  int nchunks;
  const double* partial;
  T* x;
  const int* done;
  __device__ __forceinline__ void run(int block_, int nblocks_) const {
    const long long* __restrict__ code = this->code;
    long long nfree = this->nfree;
    long long nfp = this->nfp;
    int nchunks = this->nchunks;
    const double* __restrict__ partial = this->partial;
    T* __restrict__ x = this->x;
    const int* __restrict__ done = this->done;
    (void)block_;
    (void)nblocks_;

  if (done && *done) return;
  long long i = (long long)block_ * blockDim.x + threadIdx.x;
  if (i >= nfree) return;
  double v = 0.0;
  for (int c = 0; c < nchunks; ++c) v += partial[(long long)c * nfp + i];
  x[code[i]] = (T)v;
  }
};

}  // namespace

long long coarsest_partial_elems(long long nfp) { return (nfp + GV_CHUNK - 1) / GV_CHUNK * nfp; }

long long coarsest_free_dofs(const Grid& g, const long long* node_pidx, long long nn,
                             const uint32_t* flags, DevBuf& code, DevBuf& dofmap, hipStream_t s) {
  DevBuf mark, scan;
  mark.alloc(sizeof(int) * 3 * nn);
  scan.alloc(sizeof(int) * 3 * nn);
  k_free_mark<<<grid1d(3 * nn, BS), BS, 0, s>>>(node_pidx, nn, g.N, flags, mark.as<int>());
  GMGS_LAUNCH_CHECK();
  const long long nfree = exclusive_scan(mark.as<int>(), scan.as<int>(), 3 * nn, s);
  code.alloc(sizeof(long long) * (nfree > 0 ? nfree : 1));
  dofmap.alloc(sizeof(int) * 3 * g.N);
  GMGS_HIP(hipMemsetAsync(dofmap.as<int>(), 0xFF, sizeof(int) * 3 * g.N, s));
  k_free_emit<<<grid1d(3 * nn, BS), BS, 0, s>>>(node_pidx, nn, g.N, mark.as<int>(), scan.as<int>(),
                                                code.as<long long>(), dofmap.as<int>());
  GMGS_LAUNCH_CHECK();
  GMGS_HIP(hipStreamSynchronize(s));
  return nfree;
}

void assemble_coarsest(const Grid& g, const uint32_t* flags, const int* eop, const double* Ke,
                       const double* Ks, const long long* code, const int* dofmap, long long nfree,
                       long long nfp, double* A, hipStream_t s) {
  GMGS_HIP(hipMemsetAsync(A, 0, sizeof(double) * nfp * nfp, s));
  k_assemble<<<grid1d(nfp, 64), 64, 0, s>>>(g, flags, eop, Ke, Ks, code, dofmap, nfree, nfp, A);
  GMGS_LAUNCH_CHECK();
}

int cholesky(double* A, long long nfp, hipStream_t s) {
  const int nb = (int)(nfp / DB);
  DevBuf info, diag0;
  info.alloc_zero(sizeof(int), s);
  diag0.alloc(sizeof(double) * nfp);
  k_diag_copy<<<grid1d(nfp, BS), BS, 0, s>>>(A, nfp, diag0.as<double>());
  GMGS_LAUNCH_CHECK();
  for (int k = 0; k < nb; ++k) {
    k_potrf_diag<<<1, BS, 0, s>>>(A, nfp, k, diag0.as<double>(), info.as<int>());
    GMGS_LAUNCH_CHECK();
    const int m = nb - k - 1;
    if (m > 0) {
      k_trsm<<<m, BS, 0, s>>>(A, nfp, k);
      GMGS_LAUNCH_CHECK();
      k_syrk<<<(unsigned)((long long)m * (m + 1) / 2), BS, 0, s>>>(A, nfp, k);
      GMGS_LAUNCH_CHECK();
    }
  }
  int h = 0;
  GMGS_HIP(hipMemcpyAsync(&h, info.as<int>(), sizeof(int), hipMemcpyDeviceToHost, s));
  GMGS_HIP(hipStreamSynchronize(s));
  return h;
}

void inverse_from_cholesky(const double* L, double* W, double* Kinv, long long nfp, hipStream_t s) {
  const int nb = (int)(nfp / DB);
  GMGS_HIP(hipMemsetAsync(W, 0, sizeof(double) * nfp * nfp, s));
  for (int i = 0; i < nb; ++i) {
    k_trtri_row<<<i + 1, BS, 0, s>>>(L, W, nfp, i);
    GMGS_LAUNCH_CHECK();
  }
  k_wtw<<<(unsigned)((long long)nb * (nb + 1) / 2), BS, 0, s>>>(W, Kinv, nfp, nb);
  GMGS_LAUNCH_CHECK();
  GMGS_HIP(hipStreamSynchronize(s));
}

template <typename T>
void coarsest_apply(const double* Kinv, long long nfp, long long nfree, const long long* code,
                    const T* r, T* x, long long N, double* bbuf, double* partial, const int* done,
                    hipStream_t s) {
  // x is all-zero except at free DOFs: the caller keeps fixed/void entries zero
  launch(KGatherFree<T>{code, nfree, nfp, r, bbuf, done}, grid1d(nfp, BS), s);
  const long long nrb = (nfp + GV_ROWS - 1) / GV_ROWS, nch = (nfp + GV_CHUNK - 1) / GV_CHUNK;
  launch(KGemv{Kinv, nfp, bbuf, partial, done}, nrb * nch, s);
  launch(KScatterFree<T>{code, nfree, nfp, (int)nch, partial, x, done}, grid1d(nfree, BS), s);
  (void)N;
}

template void coarsest_apply<double>(const double*, long long, long long, const long long*,
                                     const double*, double*, long long, double*, double*,
                                     const int*, hipStream_t);
template void coarsest_apply<float>(const double*, long long, long long, const long long*,
                                    const float*, float*, long long, double*, double*, const int*,
                                    hipStream_t);

}  // namespace gmgs
