// Level operators: matrix-free apply, Jacobi smoothing, transfers and the
// fused PCG vector updates.
#pragma once
#include "common.h"

namespace gmgs {

struct TileRec {
  int y0, x0, z, pad;   // padded coordinates of the tile origin
};

// Operator of one level in precision T.
template <typename T>
struct LevelOp {
  Grid g;
  const uint32_t* flags;     // per padded node
  const int* eop;            // per padded cell: -1 absent, 0 universal, k>0 stored Ks[k-1]; null on level 1
  const T* Ks;               // stored element matrices, 576 each, row-major
  const T* Ke;               // universal element matrix of this level (576, row-major)
  const T* stencil;          // interior 27x3x3 stencil of Ke (for F_FAST nodes)
  const T* cls;              // its 30 class coefficients (z-marching kernel)
  const T* emod;             // per padded cell modulus factor (level-1 SIMP) or null
  const T* invd;             // inverse diagonal, SoA 3N (zero off the solid nodes)
  const int* colz;           // per padded column (y + py x): solid-node z range lo | hi << 16, or null
  const T* asc;              // assembled stencil rows, one after another (stencil_row_elems each), or null
  const int* arow;           // per tile slot (64 tile + lane): the node's row of asc
  long long as_stride;       // rows of asc
  const T* corner;           // canonical row block of Ke (72): every block is a reflection of one
  int zero_fixed;            // SMOOTH: fixed components of the output are 0 instead of the input's
};

struct PcgScalars {
  double normB;
  double rz;          // r1'z2 of this iteration
  double rz_prev;     // r0'z1
  double rz_old;      // r1'z1 (flexible beta)
  double pq;
  double rr;
  double alpha, beta;
  double alpha_prev;  // alpha of the iteration before (the two-step x update of AX_P)
  double relres;
  double tol;
  int its;
  int done;
  int converged;
  int max_it;
  int flexible;
  int pad;
  unsigned long long t0;
};

// AX: out = A in; SMOOTH: out = in - A (omega in / d); RESID: out = in2 - A in;
// SMOOTH_SUM: the smoothing sweep Q_1 = Q_0 - A (omega Q_0 / d) from in = Q_0
// that ends the symmetric V(2, 2) cycle on level 1, writing its sum instead
// of Q_1 (CycleOut).  RESID_SUM: the residual Q_0 = in2 - A e from in = e
// that ends the symmetric V(1, 1) cycle on level 1, writing its sum
// z = omega (r0 + Q_0) / d + e instead of Q_0 (CycleOut without R1, e, w1).
// AX_P: the PCG's direction update fused into its matvec (level 1, FP64, FP32
// z): p = z + beta p_old from in = p_old, out = A p, p written to CycleOut::pn
// (a buffer other than in), and the p.q partials; x takes the two deferred
// updates x = (x + alpha_prev p_2) + alpha p_old, p_2 being the direction
// before p_old, read from CycleOut::pn before it is overwritten (none in the
// first iteration).  AX_P_NOX: the same without the x update (the iterations
// alternate between the two, so x is read and written every other one).
// RESID_SUM_P: RESID_SUM with in = x_c, the coarse level's correction on the
// grid CycleOut::gc: every staged value of e = P x_c (fixed components 0) is
// formed from the coarse values while it is staged, so e is never stored.
// SMOOTH_R: SMOOTH from the PCG residual of the previous iteration: every
// staged value is r = in - alpha q (FP64; q = CycleOut::qv, alpha from the
// scalars; in as it is in the first iteration, its = 0), the sweep runs on
// that r, and the owned nodes write r to CycleOut::rn (a buffer other than in)
// with the partials of r.r
enum ApplyMode {
  AP_AX = 0, AP_SMOOTH = 1, AP_RESID = 2, AP_SMOOTH_SUM = 3, AP_RESID_SUM = 4, AP_AX_P = 5, AP_RESID_SUM_P = 6,
  AP_AX_P_NOX = 7, AP_SMOOTH_R = 8
};

// outputs of AP_SMOOTH_SUM: z = (omega (r0 + Q_0) + w1 (R1 + Q_1)) / d + e,
// fixed components 0, and partials of r0.z and r0.zprev (zprev may be null)
template <typename T>
struct CycleOut {
  const double* r0;
  const T* R1;
  const T* e;
  const T* zprev;
  T* z;
  T w1;
  // AX_P: z of this iteration, the new direction, x, and beta / alpha
  const float* zf;
  double* pn;
  double* x;
  const PcgScalars* sc;
  // RESID_SUM_P: the coarse grid of in = x_c
  Grid gc;
  // RESID_SUM / RESID_SUM_P: with ufast, an F_FAST output node takes its
  // inverse diagonal from ucinvd (the value every F_FAST node holds) instead
  // of loading it
  int ufast;
  T ucinvd[3];
  // SMOOTH_R: q of the previous iteration and the new residual
  const double* qv;
  double* rn;
};

// The prolongation's arithmetic at one fine node, shared by the stand-alone
// prolongation and the residual kernel that forms it while staging, so both
// give the same bits: the two coarse planes are combined first (weights wz0,
// wz1), then the two coarse columns in x, then the two rows in y.  A weight
// is 1 or 0.5 on an odd fine coordinate's two neighbours (0.5 each) and 1, 0
// on an even one's single neighbour, so the unused corner adds an exact 0.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename T>
__device__ __forceinline__ T prolong_zplanes(T wz0, T wz1, T v0, T v1) {
  return fma_t(wz1, v1, wz0 * v0);
}
// z-combined corner values zab (a: y offset, b: x offset)
template <typename T>
__device__ __forceinline__ T prolong_xy(T wy0, T wy1, T wx0, T wx1, T z00, T z01, T z10, T z11) {
  const T r0 = fma_t(wx1, z01, wx0 * z00), r1 = fma_t(wx1, z11, wx0 * z10);
  return fma_t(wy1, r1, wy0 * r0);
}

// ---- launches (all on stream s; 'done' may be null) ----
// assembled stencil rows of a coarse level (FP64): coefficient (o, c, d) of
// row k at asc[(9 o + 3 c + d) * stride + k], o = neighbour (dy + 1) +
// 3 (dx + 1) + 9 (dz + 1).  Row arow[64 t + lane] of every solid node that is
// not F_FAST gets that node's element sum; row stride - 1 is the universal
// stencil (op.stencil), shared by the F_FAST nodes.  asc must be zeroed.
void launch_assemble_stencil(const LevelOp<double>& op, const TileRec* tiles, long long ntiles,
                             const int* arow, long long stride, double* asc, hipStream_t s);
constexpr int kStencilCoefs = 243;
// 16-byte groups of one z-plane's 81 coefficients of an assembled row
__host__ __device__ constexpr int stencil_plane_groups(int elem_bytes) {
  return (81 + 16 / elem_bytes - 1) / (16 / elem_bytes);
}
// elements of one assembled row in the layout of k_apply_as: three z-planes
// of 81 coefficients, each padded to whole 16-byte groups
inline long long stencil_row_elems(int elem_bytes) {
  return 3LL * stencil_plane_groups(elem_bytes) * (16 / elem_bytes);
}
// rows rep[0..nu) of the FP64 rows [q][row] (stride 'rows') in that layout,
// one after another; dst zeroed, nu * stencil_row_elems(sizeof(T)) elements
template <typename T>
void launch_relayout_rows(const double* src, long long rows, const long long* rep, long long nu, T* dst,
                          hipStream_t s);
// two 64-bit fingerprints per FP64 row [q][row] (h[2 row], h[2 row + 1])
void launch_row_hash(const double* src, long long rows, unsigned long long* h, hipStream_t s);
// *bad = 1 (*bad zeroed by the caller) if a row's coefficients differ
// bitwise from those of row rep[uid[row]]
void launch_row_verify(const double* src, long long rows, const int* uid, const long long* rep, int* bad,
                       hipStream_t s);

template <typename T, typename TR>
void launch_apply(const LevelOp<T>& op, const TileRec* tiles, long long ntiles, int mode,
                  const TR* in, T omega, T* out, TR* xacc, double* partial, const int* done,
                  hipStream_t s);

template <typename T>
void launch_restrict(const Grid& gc, const uint32_t* flags_c, const TileRec* tiles_c, long long nt,
                     const Grid& gf, const T* rf, T* rc, const int* done, hipStream_t s);

// x = omega*invd*(r + rp) + P xc on the fine level; if 'fix' zero fixed
// components; optional dot partials vs rdot (FP64): sum rdot.x and rdot.zprev
template <typename T, typename TR, typename TZ>
void launch_prolong_smooth(const Grid& gf, const uint32_t* flags_f, const TileRec* tiles, long long nt,
                           const Grid& gc, const T* xc, const T* invd, const TR* r, const T* rp,
                           T omega, TZ* x, bool fix, const TZ* zprev, double* partial,
                           const int* done, hipStream_t s);

// out = a - out on the solid nodes
template <typename T>
void launch_rsub(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, const T* a,
                 T* out, const int* done, hipStream_t s);

// e = P xc on the fine level (prolongation only); fixed components 0 if 'mask'
template <typename T>
void launch_prolong(const Grid& gf, const uint32_t* flags_f, const TileRec* tiles, long long nt,
                    const Grid& gc, const T* xc, T* e, bool mask, const int* done, hipStream_t s);

// dst = src with the fixed components zeroed, on the solid nodes
template <typename T>
void launch_mask_fixed(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                       const T* src, T* dst, hipStream_t s);

// terms of the symmetric cycle's sum on one level: R_s (1 <= s < nu) and Q_s
// (0 <= s < nu) the smoothed residuals, w_s the sweep weights (R[0] unused:
// R_0 is the level's input, passed on its own)
constexpr int kMaxSweeps = 4;
template <typename T>
struct CycleSum {
  const T* R[kMaxSweeps];
  const T* Q[kMaxSweeps];
  T w[kMaxSweeps];
  int nu;
};

// x = (sum_s w_s (R_s + Q_s)) / d + e, R_0 = r0; fixed components zeroed if
// 'fix'; optional partials of r0.x and r0.zprev (FP64)
template <typename T, typename TR, typename TZ>
void launch_combine(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                    const T* invd, const CycleSum<T>& cs, const TR* r0, const T* e, TZ* x, bool fix,
                    const TZ* zprev, double* partial, const int* done, hipStream_t s);

// p = z + beta p; with x, first x += alpha p (the previous iteration's x
// update, deferred so that p is read once per iteration)
template <typename TZ>
void launch_update_p(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                     const TZ* z, double* p, const PcgScalars* sc, const int* done, hipStream_t s,
                     double* x = nullptr);

// x += alpha p, unconditionally (the last deferred x update, after the PCG
// loop); with p_odd, the directions alternate between p (even iterations)
// and p_odd, and the last iteration's one is taken; with 'two' as well (the
// AX_P / AX_P_NOX alternation), after an odd last iteration the one before it
// is still pending too: x = (x + alpha_prev p) + alpha p_odd
void launch_add_alpha_p(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, double* x,
                        const double* p, const double* p_odd, const PcgScalars* sc, hipStream_t s,
                        bool two = false);

// x = (x + omega r / d) + alpha p (the Jacobi increment of the iteration, from
// r before it is replaced; none when invd is null; no x update when x is
// null), r = r_a - alpha q (r_a may be r), partials of r.r
void launch_update_xr(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                      double* x, const double* p, const double* ra, const double* q, double* r,
                      const double* invd, double omega, const PcgScalars* sc, double* partial,
                      const int* done, hipStream_t s);

// x += omega r / d on the solid nodes
template <typename T, typename TR>
void launch_jacobi_acc(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                       const T* invd, const TR* r, T omega, TR* x, hipStream_t s);

// Jacobi-preconditioned CG steps of the smoother-bound estimate:
// v = fixed pseudo-random values in [-1, 1) on the free components of the solid nodes
template <typename T>
void launch_seed(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, T* v,
                 hipStream_t s);
// r -= alpha q (q null: r kept), z = invd r; pointwise_blocks(nt) partials of r.z
template <typename T>
void launch_jcg_step(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                     const T* invd, const T* q, double alpha, T* r, T* z, double* partial,
                     hipStream_t s);
// p = z + beta p
template <typename T>
void launch_xpby(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt, const T* z,
                 double beta, T* p, hipStream_t s);

// dot / norm partials over tiles (FP64)
void launch_dot(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                const double* a, const double* b, double* partial, hipStream_t s);

// Row boxes of the split operator's second half: the solid nodes that are
// not F_FAST, grouped by box of 16 (y) x kRowBoxX (x) x kRowBoxZ (z) nodes,
// up to 64 per record.  Record r's box starts at padded (y0, x0, z0); entry
// 64 r + lane gives that lane's node in the box (pos = ly | lx << 4 | lz << 7,
// -1 for none) and its assembled row.
constexpr int kRowBoxX = 4, kRowBoxZ = 4;
static_assert(kRowBoxX == 4 || kRowBoxX == 8, "row boxes hold whole 16 x 4 tiles in x, lx in 3 bits");
struct RowTile {
  int y0, x0, z0, n;
};
struct RowEntry {
  int pos, row;
};
// the assembled rows of the nodes the row boxes list: AX, SMOOTH or RESID
// (out = in2 - A in) as launch_apply; returns the number of dot partials
// written (one per block of four records: AX p.Ap, RESID |in2 - A in|^2)
template <typename T, typename TR>
int launch_apply_rows(const LevelOp<T>& op, const RowTile* boxes, const RowEntry* entries, long long nbox,
                      int mode, const TR* in, T omega, const T* in2, T* out, double* partial, const int* done,
                      hipStream_t s);

// start of a solve from x0 = 0: x = 0, r = b, partials of b.b, and p, p2,
// z0, z1 = 0 where given (null: not written)
template <typename TZ>
void launch_pcg_start(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                      const double* b, double* x, double* p, double* p2, double* r, TZ* z0, TZ* z1,
                      double* partial, hipStream_t s);

// partials of |b - q|^2 over tiles
void launch_resid_sq(const Grid& g, const uint32_t* flags, const TileRec* tiles, long long nt,
                     const double* b, const double* q, double* partial, hipStream_t s);

// scalar steps of the PCG iteration; 'which': 0 = init normB, 1 = beta, 2 = alpha, 3 = residual
void launch_finalize(int which, const double* partial, int nparts, PcgScalars* sc,
                     double* hist_relres, unsigned long long* hist_t, const int* done, hipStream_t s);

template <typename TS, typename TD>
void launch_convert(const TS* a, TD* b, long long n, hipStream_t s);

template <typename T>
void launch_fill(T* a, T v, long long n, hipStream_t s);

// c = a - b (dense arrays)
void launch_sub(const double* a, const double* b, double* c, long long n, hipStream_t s);

// compact interleaved (3*nn) <-> dense SoA
void launch_scatter_compact(const long long* node_pidx, long long nn, const double* comp,
                            double* dense, long long N, hipStream_t s);
template <typename T>
void launch_gather_compact(const long long* node_pidx, long long nn, const T* dense, long long N,
                           double* comp, hipStream_t s);

// nodal stress (GetCartesianStress.m:32-41): S (nn x 6, column-major) from the
// dense level-1 displacement; M = OTP*D*B (48x24)
void launch_stress(const Grid& g, const long long* node_pidx, long long nn, const uint32_t* flags,
                   const double* M, const double* emod, const double* u, double* S, hipStream_t s);

// per compact element: out[k] = -lam_e' Ke u_e
void launch_element_energy(const Grid& g, const long long* elem_pidx, long long ne, const double* Ke,
                           const double* lam, const double* u, double* out, hipStream_t s);

// plain device copy of n doubles (HBM counter calibration)
void launch_copy_probe(const double* a, double* b, long long n_doubles, hipStream_t s);

// out[i] = *flags[i] (per-problem 'done' flags of a batch)
void launch_gather_flags(const int* const* flags, int n, int* out, hipStream_t s);

inline unsigned apply_blocks(long long ntiles) { return (unsigned)((ntiles + 3) / 4); }
// grid of the pointwise kernels that reduce (update_xr, dot, the start and
// exit norms, prolong_smooth and combine with dot partials): also their
// number of dot partials.  The others take one wave per tile (apply_blocks),
// no slower and faster for combine (level 2: 61 vs 89 us).
inline unsigned pointwise_blocks(long long ntiles) {
  const unsigned b = apply_blocks(ntiles);
  return b < 2048u ? b : 2048u;
}
inline unsigned grid_for(long long ntiles, const double* partial) {
  return partial ? pointwise_blocks(ntiles) : apply_blocks(ntiles);
}

}  // namespace gmgs
