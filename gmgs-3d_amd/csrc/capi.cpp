// Context, set-up orchestration, PCG driver and the C ABI (include/gmgs.h).
#include "../../include/gmgs.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "dense.h"
#include "fe_host.h"
#include "launch.h"
#include "mesh.h"
#include "ops.h"
#include "ops_zm.h"
#include "setup.h"

using namespace gmgs;

namespace {

struct Status : std::runtime_error {
  int code;
  Status(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Level {
  LevelMesh mesh;
  DevBuf tiles;
  long long ntiles = 0;
  int state = 0;                 // storingState (SetupSolver.m:10,41,70,...)
  DevBuf eop;                    // int32 per padded cell (levels >= 2)
  DevBuf Ks64, Ks32;             // stored element matrices
  long long nKs = 0;
  DevBuf Ke64, Ke32, S64, S32;   // universal matrix and interior stencil of this level
  DevBuf C64, C32;               // class coefficients of the stencil
  DevBuf Kc64, Kc32;             // canonical corner row block of Ke (72)
  bool classes_ok = false;
  DevBuf invd64, invd32;
  DevBuf invdm;                  // symmetric cycle: inverse diagonal on the free DOFs (cycle precision)
  DevBuf uniq;                   // int32 compact ids of unique elements (state 2)
  long long nuniq = 0;
  DevBuf r, rp, x;               // V-cycle work vectors (precision of the cycle)
  DevBuf wk[2 * kMaxSweeps - 2]; // symmetric cycle: smoothed residuals (see Rbuf, Qbuf)
  DevBuf e;                      // symmetric cycle: P x_c
  double scale = 1.0;            // Ke_l = Ke0 * scale
  // z-marching operator (levels with a universal element matrix)
  bool zm = false;
  DevBuf coltiles;
  DevBuf colz;                   // per padded column: solid-node z range
  long long ncol = 0;
  // assembled 27-point stencil rows (coarse levels): one per solid node that
  // is not F_FAST plus the shared universal row; built, and in use
  DevBuf asc64, asc32, arow;
  long long as_stride = 0;
  bool as_built = false, as = false;
  // split operator: the z-marching kernel on the F_FAST nodes and the
  // assembled rows on the others, grouped in row boxes (rboxes, with their
  // nodes and rows in rentries)
  DevBuf rboxes, rentries;
  long long nrbox = 0;
  bool split_ok = false, split = false;
  int col_tx = 8, col_zc = 16;
  // level 1 of the symmetric V(1, 1) cycle: the residual kernel forms
  // e = P x_c while staging it (else a prolongation kernel stores e first)
  bool fuse_prolong = true;
  // the FP32 inverse diagonal every F_FAST node of the level holds, if one
  bool ufast_invd = false;
  float ucinvd[3] = {0.f, 0.f, 0.f};
  // level 1 of a mixed context: the FP64 matvec's own column tiles (the
  // cycle's FP32 sweeps are tuned on coltiles)
  DevBuf coltiles64;
  long long ncol64 = 0;
  int col_tx64 = 0, col_zc64 = 0;
  // level 1 of a mixed solver: column tiles of the first sweep when it carries
  // the PCG's residual update (AP_SMOOTH_R), when tuned apart (else coltiles)
  DevBuf coltilesR;
  long long ncolR = 0;
  int col_txR = 0, col_zcR = 0;
  // symmetric cycle: nu sweeps with weights om, and the estimate of
  // lambda_max(D^-1 A) on the free DOFs they came from (0 = plain Jacobi)
  int nu = 2;
  double om[kMaxSweeps] = {0.0, 0.0, 0.0, 0.0};
  double lmax = 0.0;
};

}  // namespace

namespace gmgs {
Recorder*& current_recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}
}  // namespace gmgs

struct gmgs_ctx {
  gmgs_params prm{};
  hipStream_t stream = nullptr;
  int device = 0;
  int L = 0;
  int nely = 0, nelx = 0, nelz = 0;       // input voxel dims
  double bbox[6] = {0, 0, 0, 0, 0, 0};     // adjusted [min xyz, max xyz]
  double eleSize[3] = {1, 1, 1};
  double Ke0[576];
  double M[48 * 24];
  std::vector<Level> lev;
  bool mixed = false;
  bool flexible = false;
  bool exact = false;
  bool tables = true;
  bool simple_kernels = false;
  bool unfused_prolong = false;   // V(1, 1) level 1: prolongation as its own kernel
  bool unfused_resid = false;     // PCG residual update as its own kernel (see fused_residual)
  bool autotune = false;
  bool sym = false;              // symmetric V(nu, nu) cycle
  int sweeps = 2, sweeps_coarse = 2;   // nu of level 1, of the coarser smoothed levels
  int sweeps_l2 = 2;                   // nu of level 2
  bool cheb = false;             // Chebyshev sweep weights in the symmetric cycle
  bool prefer_as = false;        // assembled stencil on every coarse level where it fits
  bool prefer_split = false;     // split operator wherever it exists, the assembled stencil elsewhere
  bool no_jacobi = false;        // no Jacobi step opening each PCG iteration
  // SIMP
  bool simp = false;
  DevBuf emod64, emod32;
  DevBuf stressM;
  // boundary conditions
  bool bc = false;
  DevBuf b;                 // dense FP64 load vector (level 1)
  std::vector<double> F;    // compact F_
  std::vector<int32_t> bc_fixed;
  std::vector<double> bc_loads;
  int64_t bc_m = 0;
  DevBuf free_code, dofmap, A, W, Kinv, gv_b, gv_part;
  long long nfree = 0, nfp = 0;
  // PCG state
  DevBuf x, r, ra, z[2], p, p2, q, tmp, tmp2, part, sc, hist_rel, hist_t;   // p2: see fused_direction
  long long nparts = 0;
  bool have_solution = false;
  std::string err;
  double clock_khz = 100000.0;
  // kernel timing (HIP events around level-1 launches on this stream)
  bool ktime = false;
  struct KPair { hipEvent_t a, b; int kernel, iter; };
  std::vector<KPair> kpairs;
  std::vector<hipEvent_t> kpool;
  size_t kpool_used = 0;
  double ktot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // captured PCG iteration pair (even, odd) and pinned polling buffers
  bool use_graphs = true;
  hipGraphExec_t gexec = nullptr;
  PcgScalars* pinned = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  long long kcnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int cur_iter = 0;
};

namespace {

Level& L1(gmgs_ctx* c) { return c->lev[0]; }

// the mixed solver without the Jacobi step folds p = z + beta p and the
// deferred x update into its level-1 FP64 matvec when level 1 runs the
// z-marching operator (p then alternates between p and p2)
// the symmetric V(1, 1) cycle's level-1 residual kernel forms e = P x_c
// while it stages it (vcycle_sym)
bool v11_fused(const gmgs_ctx* c) {
  const Level& l = c->lev[0];
  return c->sym && c->L > 1 && l.nu == 1 && l.zm && l.fuse_prolong;
}

bool fused_direction(gmgs_ctx* c) { return c->mixed && c->no_jacobi && !c->p2.empty() && L1(c).zm; }

// With the direction update fused, the residual update r = r - alpha q of
// iteration k (and the test on |r|) moves into the first level-1 sweep of
// iteration k + 1, which stages r - alpha q instead of r (AP_SMOOTH_R); r
// alternates between two buffers, since tiles read their neighbours' old r
bool fused_residual(gmgs_ctx* c) { return fused_direction(c) && c->sym && !c->unfused_resid; }

// iterations a solve may launch: the test on the last one's residual runs in
// the iteration after it when the residual update is fused
int iteration_cap(gmgs_ctx* c) { return c->prm.max_it + 1 + (fused_residual(c) ? 1 : 0); }

hipEvent_t kt_event(gmgs_ctx* c) {
  if (c->kpool_used == c->kpool.size()) {
    hipEvent_t e;
    GMGS_HIP(hipEventCreate(&e));
    c->kpool.push_back(e);
  }
  return c->kpool[c->kpool_used++];
}

// RAII bracket: records an event pair around one kernel when timing is on
struct KTime {
  gmgs_ctx* c;
  int kernel;
  hipEvent_t a = nullptr;
  KTime(gmgs_ctx* cc, int k) : c(cc), kernel(k) {
    if (c->ktime && k >= 0) {   // a negative id times nothing
      a = kt_event(c);
      GMGS_HIP(hipEventRecord(a, c->stream));
    }
  }
  ~KTime() {
    if (a) {
      hipEvent_t b = kt_event(c);
      (void)hipEventRecord(b, c->stream);
      c->kpairs.push_back({a, b, kernel, c->cur_iter});
    }
  }
};

// fold the recorded pairs of one solve into the totals (iterations < its)
void kt_collect(gmgs_ctx* c, int its) {
  for (const auto& p : c->kpairs) {
    if (p.iter >= its) continue;
    float ms = 0.f;
    GMGS_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    c->ktot[p.kernel] += ms;
    c->kcnt[p.kernel] += 1;
  }
  c->kpairs.clear();
  c->kpool_used = 0;
}

template <typename T> const T* Kptr(const Level& l);
template <> const double* Kptr<double>(const Level& l) { return l.Ke64.as<double>(); }
template <> const float* Kptr<float>(const Level& l) { return l.Ke32.as<float>(); }

// the operator of level li; 'cyc' selects the symmetric cycle's free-DOF
// smoother (masked inverse diagonal, fixed outputs zero)
template <typename T>
LevelOp<T> make_op(gmgs_ctx* c, int li, bool cyc = false) {
  Level& l = c->lev[li];
  LevelOp<T> op;
  op.zero_fixed = 0;
  op.asc = nullptr;
  op.arow = l.arow.empty() ? nullptr : l.arow.as<int>();
  op.as_stride = l.as_stride;
  op.g = l.mesh.g;
  op.flags = l.mesh.flags.as<uint32_t>();
  op.colz = l.colz.empty() ? nullptr : l.colz.as<int>();
  op.eop = li == 0 ? nullptr : l.eop.as<int>();
  if constexpr (std::is_same<T, double>::value) {
    op.Ks = l.Ks64.empty() ? nullptr : l.Ks64.as<double>();
    op.Ke = l.Ke64.as<double>();
    op.stencil = l.S64.as<double>();
    op.cls = (l.C64.empty() || !l.classes_ok) ? nullptr : l.C64.as<double>();
    op.corner = l.Kc64.empty() ? nullptr : l.Kc64.as<double>();
    op.invd = l.invd64.empty() ? nullptr : l.invd64.as<double>();
    op.emod = (li == 0 && c->simp) ? c->emod64.as<double>() : nullptr;
  } else {
    op.Ks = l.Ks32.empty() ? nullptr : l.Ks32.as<float>();
    op.Ke = l.Ke32.as<float>();
    op.stencil = l.S32.as<float>();
    op.cls = (l.C32.empty() || !l.classes_ok) ? nullptr : l.C32.as<float>();
    op.corner = l.Kc32.empty() ? nullptr : l.Kc32.as<float>();
    op.invd = l.invd32.empty() ? nullptr : l.invd32.as<float>();
    op.emod = (li == 0 && c->simp) ? c->emod32.as<float>() : nullptr;
  }
  if (l.as || l.split) {
    if constexpr (std::is_same<T, double>::value) op.asc = l.asc64.as<double>();
    else op.asc = l.asc32.as<float>();
  }
  if (cyc && c->sym && !l.invdm.empty()) {
    op.invd = l.invdm.as<T>();
    op.zero_fixed = 1;
  }
  return op;
}

// y = A_l u (or the fused smoother) on level li; returns the number of dot partials
// AX / SMOOTH / RESID (out = in2 - A in) on level li; returns the number of
// dot partials written
template <typename T, typename TR>
int apply_level(gmgs_ctx* c, int li, int mode, const TR* in, T omega, T* out, TR* xacc,
                double* partial, const int* done, const T* in2 = nullptr, bool cyc = false) {
  Level& l = c->lev[li];
  LevelOp<T> op = make_op<T>(c, li, cyc);
  if (l.split) {
    int np = launch_apply_zm<T, TR>(op, l.coltiles.as<ColTile>(), l.ncol, l.col_tx, l.col_zc, mode, in, omega,
                                    in2, out, partial, done, c->stream, nullptr, true);
    np += launch_apply_rows<T, TR>(op, l.rboxes.as<RowTile>(), l.rentries.as<RowEntry>(), l.nrbox, mode, in,
                                   omega, in2, out, partial ? partial + np : nullptr, done, c->stream);
    if (xacc)
      launch_jacobi_acc<T, TR>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles,
                               op.invd, in, omega, xacc, c->stream);
    return np;
  }
  if (mode == AP_RESID && !l.zm) {
    launch_apply<T, TR>(op, l.tiles.as<TileRec>(), l.ntiles, AP_AX, in, omega, out, nullptr, nullptr,
                        done, c->stream);
    launch_rsub<T>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles, in2, out,
                   done, c->stream);
    return 0;
  }
  if (l.zm) {
    const bool own64 = std::is_same<T, double>::value && mode == AP_AX && !l.coltiles64.empty();
    const int np = own64 ? launch_apply_zm<T, TR>(op, l.coltiles64.as<ColTile>(), l.ncol64, l.col_tx64, l.col_zc64,
                                                  mode, in, omega, in2, out, partial, done, c->stream)
                         : launch_apply_zm<T, TR>(op, l.coltiles.as<ColTile>(), l.ncol, l.col_tx, l.col_zc,
                                                  mode, in, omega, in2, out, partial, done, c->stream);
    if (xacc)
      launch_jacobi_acc<T, TR>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles,
                               op.invd, in, omega, xacc, c->stream);
    return np;
  }
  launch_apply<T, TR>(op, l.tiles.as<TileRec>(), l.ntiles, mode, in, omega, out, xacc, partial, done,
                      c->stream);
  return (int)apply_blocks(l.ntiles);
}

template <typename T>
void upload(DevBuf& d, const T* h, size_t n, hipStream_t s) {
  d.alloc(sizeof(T) * n);
  GMGS_HIP(hipMemcpyAsync(d.as<void>(), h, sizeof(T) * n, hipMemcpyHostToDevice, s));
}

template <typename T>
std::vector<T> download(const DevBuf& d, size_t n, hipStream_t s) {
  std::vector<T> h(n);
  if (n) GMGS_HIP(hipMemcpyAsync(h.data(), d.as<void>(), sizeof(T) * n, hipMemcpyDeviceToHost, s));
  GMGS_HIP(hipStreamSynchronize(s));
  return h;
}

// Equal rows of the FP64 assembled rows [q][row] (stride 'rows'): uid[row]
// numbers the distinct rows in order of first appearance, rep[u] is the first
// row of distinct row u.  Rows are grouped by two 64-bit fingerprints and the
// grouping is then checked bitwise on the device; should two different rows
// ever share both fingerprints, every row stays its own.
void distinct_rows(const double* canon, long long rows, std::vector<int>& uid, std::vector<long long>& rep,
                   hipStream_t s) {
  using Key = std::pair<unsigned long long, unsigned long long>;
  struct KeyHash {
    size_t operator()(const Key& k) const { return (size_t)(k.first ^ (k.second * 0x9E3779B97F4A7C15ull)); }
  };
  DevBuf dh;
  dh.alloc(sizeof(unsigned long long) * 2 * (size_t)rows);
  launch_row_hash(canon, rows, dh.as<unsigned long long>(), s);
  const std::vector<unsigned long long> h = download<unsigned long long>(dh, 2 * (size_t)rows, s);
  std::unordered_map<Key, int, KeyHash> first;
  first.reserve((size_t)rows);
  uid.assign((size_t)rows, 0);
  rep.clear();
  for (long long r = 0; r < rows; ++r) {
    const auto it = first.emplace(Key(h[2 * r], h[2 * r + 1]), (int)rep.size());
    if (it.second) rep.push_back(r);
    uid[(size_t)r] = it.first->second;
  }
  DevBuf du, drep, dbad;
  upload(du, uid.data(), uid.size(), s);
  upload(drep, rep.data(), rep.size(), s);
  dbad.alloc_zero(sizeof(int), s);
  launch_row_verify(canon, rows, du.as<int>(), drep.as<long long>(), dbad.as<int>(), s);
  if (download<int>(dbad, 1, s)[0] != 0) {
    rep.resize((size_t)rows);
    for (long long r = 0; r < rows; ++r) {
      uid[(size_t)r] = (int)r;
      rep[(size_t)r] = r;
    }
  }
}

void to_f32(const DevBuf& a, DevBuf& b, long long n, hipStream_t s) {
  b.alloc(sizeof(float) * (n > 0 ? n : 1));
  launch_convert<double, float>(a.as<double>(), b.as<float>(), n, s);
}

// MATLAB a:d:b evaluated from both ends
double colon_value(double a, double d, double b, long long n, long long k) {
  return (2 * k >= n) ? b - (double)(n - k) * d : a + (double)k * d;
}

// ------------------------------------------------------------------ set-up
void build_mesh(gmgs_ctx* c, const int32_t* vox) {
  hipStream_t s = c->stream;
  const gmgs_params& p = c->prm;
  const long long nelx = c->nelx, nely = c->nely, nelz = c->nelz;
  int Lm1;   // levels - 1
  if (p.num_levels > 0) {
    Lm1 = p.num_levels - 1;
  } else {
    double num = (double)nelx * nely * nelz;
    if (num < p.coarsest_resolution_control)
      throw Status(GMGS_ERR_RESOLUTION, "There is no Sufficient Resolution to Build Mesh Hierarchy!");
    Lm1 = 0;
    while (num >= p.coarsest_resolution_control) {
      ++Lm1;
      const double f = std::pow(2.0, Lm1);
      num = nelx / f * nely / f * nelz / f;
    }
  }
  c->L = Lm1 + 1;
  if (c->L < 2) throw Status(GMGS_ERR_LEVELS, "There is no sufficient resolution to invoke GMGS!");
  const long long sp = 1LL << Lm1;
  const long long ax = (nelx + sp - 1) / sp * sp, ay = (nely + sp - 1) / sp * sp,
                  az = (nelz + sp - 1) / sp * sp;
  // bounding box stretch (DiscretizeDesignDomain.m:57-58) and element size (:65)
  const long long res[3] = {ax, ay, az};
  const long long nel[3] = {nelx, nely, nelz};
  for (int d = 0; d < 3; ++d) {
    const double lo = c->bbox[d], hi = c->bbox[3 + d];
    c->bbox[3 + d] = lo + (hi - lo) * (double)res[d] / (double)nel[d];
    c->eleSize[d] = (c->bbox[3 + d] - lo) / (double)res[d];
  }
  {
    bool any = false;
    for (long long i = 0, n = nelx * nely * nelz; i < n && !any; ++i) any = vox[i] == 1;
    if (!any) throw Status(GMGS_ERR_RESOLUTION, "the voxel model has no solid voxel");
  }
  DevBuf dvox;
  upload(dvox, vox, (size_t)(nelx * nely * nelz), s);
  c->lev.resize(c->L);
  c->lev[0].mesh.g = make_grid((int)ay, (int)ax, (int)az);
  build_level1(c->lev[0].mesh, dvox.as<int>(), (int)nely, (int)nelx, (int)nelz, c->tables, s);
  if (c->lev[0].mesh.ne == 0) throw Status(GMGS_ERR_RESOLUTION, "the voxel model has no solid voxel");
  for (int l = 1; l < c->L; ++l) {
    const Grid& gf = c->lev[l - 1].mesh.g;
    c->lev[l].mesh.g = make_grid(gf.ey / 2, gf.ex / 2, gf.ez / 2);
    build_coarse(c->lev[l].mesh, c->lev[l - 1].mesh, c->lev[0].mesh.g, c->tables, s);
  }
  for (int l = 0; l < c->L; ++l) {
    Level& lv = c->lev[l];
    lv.ntiles = build_tiles(lv.mesh.g, lv.mesh.flags.as<uint32_t>(), lv.tiles, s);
  }
}

void probe_coords(gmgs_ctx* c, double probe[8][3]) {
  // nodeCoords_ of the first solid element's corners (SetupStiffnessMatrix.m:11-13)
  const Grid& g = L1(c).mesh.g;
  int e1 = 0;
  GMGS_HIP(hipMemcpyAsync(&e1, L1(c).mesh.eleMapBack.as<int>(), sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
  GMGS_HIP(hipStreamSynchronize(c->stream));
  const long long f = e1 - 1;
  const int y = (int)(f % g.ey), x = (int)((f / g.ey) % g.ex), z = (int)(f / ((long long)g.ey * g.ex));
  const double* bb = c->bbox;
  const double dx = (bb[3] - bb[0]) / g.ex, dy = (bb[1] - bb[4]) / g.ey, dz = (bb[5] - bb[2]) / g.ez;
  for (int m = 0; m < 8; ++m) {
    const int ny_ = y + hex_ro(m), nx_ = x + hex_xo(m), nz_ = z + hex_zo(m);
    probe[m][0] = colon_value(bb[0], dx, bb[3], g.ex, nx_);
    probe[m][1] = colon_value(bb[4], dy, bb[1], g.ey, ny_);   // ySeed descends (DiscretizeDesignDomain.m:96)
    probe[m][2] = colon_value(bb[2], dz, bb[5], g.ez, nz_);
  }
}

void upload_level_constants(gmgs_ctx* c, int li) {
  Level& l = c->lev[li];
  hipStream_t s = c->stream;
  // Ke_l = Ke0 * eleSize_l(1)/eleSize_1(1) (SetupSolver.m:37)
  l.scale = std::ldexp(1.0, li);
  std::vector<double> K(576), S(243);
  for (int i = 0; i < 576; ++i) K[i] = c->Ke0[i] * l.scale;
  double St[27][3][3];
  interior_stencil(K.data(), St);
  std::memcpy(S.data(), St, sizeof(double) * 243);
  double cls[30], Kc[72];
  // the z-marching operator needs the stencil's reflection classes and the
  // element matrix's corner reflection symmetry
  l.classes_ok = stencil_classes(St, cls, 1e-12) && corner_blocks(K.data(), Kc, 1e-12);
  upload(l.Ke64, K.data(), 576, s);
  upload(l.S64, S.data(), 243, s);
  upload(l.C64, cls, 30, s);
  upload(l.Kc64, Kc, 72, s);
  if (c->mixed) {
    std::vector<float> Kf(576), Sf(243), Cf(30), Kcf(72);
    for (int i = 0; i < 576; ++i) Kf[i] = (float)K[i];
    for (int i = 0; i < 243; ++i) Sf[i] = (float)S[i];
    for (int i = 0; i < 30; ++i) Cf[i] = (float)cls[i];
    for (int i = 0; i < 72; ++i) Kcf[i] = (float)Kc[i];
    upload(l.Ke32, Kf.data(), 576, s);
    upload(l.S32, Sf.data(), 243, s);
    upload(l.C32, Cf.data(), 30, s);
    upload(l.Kc32, Kcf.data(), 72, s);
  }
}

// largest FP64 assembled stencil of one level (bytes)
constexpr double kAscBudgetBytes = 4.0e9;

// coarse element matrices and storing states (SetupSolver.m:34-196)
void build_operators(gmgs_ctx* c) {
  hipStream_t s = c->stream;
  const int L = c->L;
  const int exact = c->exact ? 1 : 0;
  for (int li = 1; li < L; ++li) {
    Level& l = c->lev[li];
    Level& f = c->lev[li - 1];
    const Grid& g = l.mesh.g;
    const long long ne = l.mesh.ne;
    const int lvl = li + 1;
    l.eop.alloc(sizeof(int) * g.N);
    launch_fill<int>(l.eop.as<int>(), -1, g.N, s);
    std::vector<int> val;
    DevBuf dval;
    dval.alloc(sizeof(int) * (ne > 0 ? ne : 1));
    const double* emodf = (li == 1 && c->simp) ? c->emod64.as<double>() : nullptr;
    if (c->simp) {
      // per-element Galerkin of scaled children on every coarse level
      l.state = 1;
    } else if ((lvl == 2 && L > 2) || (lvl == 3 && L > 3)) {
      l.state = 2;
    } else {
      l.state = 1;
    }
    if (lvl == 2 && !c->simp) {
      // children are level-1 universal elements: Ks depends on the child mask only
      DevBuf mask;
      mask.alloc(sizeof(int) * (ne > 0 ? ne : 1));
      child_masks(l.mesh.elem_pidx.as<long long>(), ne, g, f.mesh.g, f.mesh.esolid.as<uint8_t>(),
                  mask.as<int>(), s);
      l.Ks64.alloc(sizeof(double) * 256 * 576);
      GMGS_HIP(hipMemsetAsync(l.Ks64.as<double>(), 0, sizeof(double) * 256 * 576, s));
      galerkin_mask_table(f.Ke64.as<double>(), exact, l.Ks64.as<double>(), s);
      l.nKs = 256;
      std::vector<int> m = download<int>(mask, ne, s);
      val.resize(ne);
      std::vector<int> uq;
      for (long long k = 0; k < ne; ++k) {
        if (l.state == 2 && m[k] == 0xFF) val[k] = 0;
        else val[k] = 1 + m[k];
        if (l.state == 2 && m[k] != 0xFF) uq.push_back((int)(k + 1));
      }
      l.nuniq = (long long)uq.size();
      if (!uq.empty()) upload(l.uniq, uq.data(), uq.size(), s);
    } else if (l.state == 2) {
      // level 3 with deeper levels: B3 test, unique ones stored
      DevBuf flag;
      flag.alloc(sizeof(int) * (ne > 0 ? ne : 1));
      unique_flags_l3(l.mesh.elem_pidx.as<long long>(), ne, g, f.mesh.g, f.eop.as<int>(), exact,
                      flag.as<int>(), s);
      std::vector<int> fl = download<int>(flag, ne, s);
      std::vector<long long> epidx = download<long long>(l.mesh.elem_pidx, ne, s);
      std::vector<long long> cells;
      std::vector<int> uq;
      val.resize(ne);
      for (long long k = 0; k < ne; ++k) {
        if (fl[k]) {
          val[k] = 1 + (int)cells.size();
          cells.push_back(epidx[k]);
          uq.push_back((int)(k + 1));
        } else {
          val[k] = 0;
        }
      }
      l.nKs = (long long)cells.size();
      l.nuniq = l.nKs;
      l.Ks64.alloc(sizeof(double) * 576 * (l.nKs > 0 ? l.nKs : 1));
      if (l.nKs) {
        DevBuf dcells;
        upload(dcells, cells.data(), cells.size(), s);
        galerkin_cells(dcells.as<long long>(), l.nKs, g, f.mesh.g, f.mesh.esolid.as<uint8_t>(),
                       f.eop.as<int>(), f.Ke64.as<double>(),
                       f.Ks64.empty() ? nullptr : f.Ks64.as<double>(), nullptr, exact,
                       l.Ks64.as<double>(), s);
        upload(l.uniq, uq.data(), uq.size(), s);
        GMGS_HIP(hipStreamSynchronize(s));
      }
    } else {
      // storing state 1: every element stored
      l.nKs = ne;
      l.Ks64.alloc(sizeof(double) * 576 * (ne > 0 ? ne : 1));
      galerkin_cells(l.mesh.elem_pidx.as<long long>(), ne, g, f.mesh.g, f.mesh.esolid.as<uint8_t>(),
                     li == 1 ? nullptr : f.eop.as<int>(), f.Ke64.as<double>(),
                     f.Ks64.empty() ? nullptr : f.Ks64.as<double>(), emodf, exact,
                     l.Ks64.as<double>(), s);
      val.resize(ne);
      for (long long k = 0; k < ne; ++k) val[k] = (int)(k + 1);
    }
    if (ne) {
      upload(dval, val.data(), (size_t)ne, s);
      scatter_eop(l.mesh.elem_pidx.as<long long>(), ne, dval.as<int>(), l.eop.as<int>(), s);
    }
    GMGS_HIP(hipStreamSynchronize(s));
  }
  // fast-path marks, inverse diagonals (levels 1..L-1), FP32 copies
  for (int li = 0; li < L; ++li) {
    Level& l = c->lev[li];
    mark_fast(l.mesh.g, l.tiles.as<TileRec>(), l.ntiles, l.mesh.flags.as<uint32_t>(),
              li == 0 ? nullptr : l.eop.as<int>(), !(li == 0 && c->simp), c->stream);
    if (li < L - 1) {
      l.invd64.alloc_zero(sizeof(double) * 3 * l.mesh.g.N, s);
      inverse_diagonal(l.mesh.g, l.tiles.as<TileRec>(), l.ntiles, l.mesh.flags.as<uint32_t>(),
                       li == 0 ? nullptr : l.eop.as<int>(), l.Ke64.as<double>(),
                       l.Ks64.empty() ? nullptr : l.Ks64.as<double>(),
                       (li == 0 && c->simp) ? c->emod64.as<double>() : nullptr, l.invd64.as<double>(), s);
    }
    if (c->mixed) {
      if (!l.Ks64.empty()) to_f32(l.Ks64, l.Ks32, (long long)(l.Ks64.bytes() / sizeof(double)), s);
      if (!l.invd64.empty()) to_f32(l.invd64, l.invd32, 3 * l.mesh.g.N, s);
      l.ufast_invd = li == 0 && !l.invd32.empty() &&
                     fast_inverse_diagonal(l.mesh.g, l.tiles.as<TileRec>(), l.ntiles, l.mesh.flags.as<uint32_t>(),
                                           l.invd32.as<float>(), l.ucinvd, s);
    }
    // z-marching operator where a universal element matrix exists
    l.zm = !c->simple_kernels && l.classes_ok && (li == 0 ? !c->simp : l.state == 2);
    l.fuse_prolong = !c->unfused_prolong;
    if (l.zm && l.coltiles.empty()) {
      l.ncol = build_coltiles(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.col_tx, l.col_zc, l.coltiles, s);
      build_colrange(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.colz, s);
    }
    // assembled stencil rows on the smoothed coarse levels (assembled in
    // FP64): one per solid node that is not F_FAST, in tile-slot order, and
    // the universal row last; used where the z-marching kernel is not
    l.as = l.as_built = false;
    l.split = l.split_ok = false;
    l.rboxes.release();
    l.rentries.release();
    l.nrbox = 0;
    l.asc64.release();
    l.asc32.release();
    l.arow.release();
    l.as_stride = 0;
    long long rows = 0;
    std::vector<int> ar;
    std::vector<TileRec> th;
    if (li >= 1 && li < L - 1 && !c->simple_kernels && l.ntiles > 0) {
      const Grid& g = l.mesh.g;
      th = download<TileRec>(l.tiles, (size_t)l.ntiles, s);
      const std::vector<uint32_t> fh = download<uint32_t>(l.mesh.flags, (size_t)g.N, s);
      ar.assign((size_t)l.ntiles * 64, -1);
      for (long long t = 0; t < l.ntiles; ++t)
        for (int lane = 0; lane < 64; ++lane) {
          const int yp = th[t].y0 + (lane & 15), xp = th[t].x0 + (lane >> 4);
          if (xp >= g.px) continue;
          const uint32_t f = fh[yp + (long long)g.py * (xp + (long long)g.px * th[t].z)];
          if ((f & F_ELEM_MASK) && !(f & F_FAST)) ar[t * 64 + lane] = (int)rows++;
        }
    }
    const long long stride = rows + 1;
    const long long asc_n = (long long)kStencilCoefs * stride;
    if (!ar.empty() && (double)asc_n * sizeof(double) <= kAscBudgetBytes && 3 * l.mesh.g.N < (1LL << 31)) {
      for (int& v : ar)
        if (v < 0) v = (int)(stride - 1);
      DevBuf slots;
      upload(slots, ar.data(), ar.size(), s);
      DevBuf canon;
      canon.alloc_zero(sizeof(double) * asc_n, s);
      launch_assemble_stencil(make_op<double>(c, li), l.tiles.as<TileRec>(), l.ntiles, slots.as<int>(), stride,
                              canon.as<double>(), s);
      // equal rows stored once: the nodes' rows become indices of distinct rows
      std::vector<int> uid;
      std::vector<long long> rep;
      distinct_rows(canon.as<double>(), stride, uid, rep, s);
      const long long nu = (long long)rep.size();
      DevBuf drep;
      upload(drep, rep.data(), rep.size(), s);
      if (c->mixed) {
        l.asc32.alloc_zero(sizeof(float) * (size_t)(nu * stencil_row_elems(sizeof(float))), s);
        launch_relayout_rows<float>(canon.as<double>(), stride, drep.as<long long>(), nu, l.asc32.as<float>(), s);
      } else {
        l.asc64.alloc_zero(sizeof(double) * (size_t)(nu * stencil_row_elems(sizeof(double))), s);
        launch_relayout_rows<double>(canon.as<double>(), stride, drep.as<long long>(), nu, l.asc64.as<double>(),
                                     s);
      }
      l.as_stride = nu;
      std::vector<int> au(ar.size());
      for (size_t i = 0; i < ar.size(); ++i) au[i] = uid[(size_t)ar[i]];
      upload(l.arow, au.data(), au.size(), s);
      // row boxes of the split operator: the solid nodes that are not
      // F_FAST, by box of 16 x kRowBoxX x kRowBoxZ nodes (boxes in order of their
      // first node in the tile list, nodes in tile and lane order), up to 64
      // per record
      std::map<std::tuple<int, int, int>, int> box_of;
      std::vector<RowTile> bhead;
      std::vector<std::vector<RowEntry>> bents;
      for (long long t = 0; t < l.ntiles; ++t)
        for (int lane = 0; lane < 64; ++lane) {
          if (ar[t * 64 + lane] == (int)(stride - 1)) continue;
          const int zb = (th[t].z - 1) / kRowBoxZ, z0 = 1 + zb * kRowBoxZ;
          const int x0 = th[t].x0 - th[t].x0 % kRowBoxX;
          const auto key = std::make_tuple(zb, x0, th[t].y0);
          auto it = box_of.find(key);
          if (it == box_of.end()) {
            it = box_of.emplace(key, (int)bhead.size()).first;
            bhead.push_back(RowTile{th[t].y0, x0, z0, 0});
            bents.emplace_back();
          }
          const int pos = (lane & 15) | ((th[t].x0 - x0 + (lane >> 4)) << 4) | ((th[t].z - z0) << 7);
          bents[(size_t)it->second].push_back(RowEntry{pos, au[t * 64 + lane]});
        }
      std::vector<RowTile> rb;
      std::vector<RowEntry> re;
      for (size_t b = 0; b < bhead.size(); ++b)
        for (size_t k0 = 0; k0 < bents[b].size(); k0 += 64) {
          const size_t k1 = std::min(bents[b].size(), k0 + 64);
          RowTile r = bhead[b];
          r.n = (int)(k1 - k0);
          rb.push_back(r);
          for (size_t k = k0; k < k0 + 64; ++k) re.push_back(k < k1 ? bents[b][k] : RowEntry{-1, 0});
        }
      l.nrbox = (long long)rb.size();
      if (!rb.empty()) {
        upload(l.rboxes, rb.data(), rb.size(), s);
        upload(l.rentries, re.data(), re.size(), s);
      }
      GMGS_HIP(hipStreamSynchronize(s));
      l.as_built = true;
      // the split operator needs the z-marching kernel's tiles
      l.split_ok = l.zm && l.nrbox > 0;
      l.split = l.split_ok && c->prefer_split;
      l.as = !l.split && (!l.zm || c->prefer_as || c->prefer_split);
      if (c->prefer_as || c->prefer_split) l.zm = false;
    }
  }
  GMGS_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ autotuner
// per-level choice of the z-marching column tile (TX x ZC), or of the tile
// kernel on coarse levels, timed on the level's own grid with the V-cycle
// smoother; cached per (grid, precision)
struct TuneKey {
  int ey, ex, ez, prec;
  bool operator<(const TuneKey& o) const {
    return std::tie(ey, ex, ez, prec) < std::tie(o.ey, o.ex, o.ez, o.prec);
  }
};
std::map<TuneKey, std::pair<int, int>>& tune_cache() {
  static std::map<TuneKey, std::pair<int, int>> m;
  return m;
}
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}

// the tile kernel on level li, element sums (as = false) or the assembled stencil
template <typename T, typename TR>
float time_v1(gmgs_ctx* c, int li, const TR* in, T* out, bool as = false) {
  Level& l = c->lev[li];
  hipStream_t s = c->stream;
  const bool keep = l.as;
  l.as = as;
  LevelOp<T> op = make_op<T>(c, li);
  l.as = keep;
  HipEvent a, b;
  for (int w = 0; w < 2; ++w)
    launch_apply<T, TR>(op, l.tiles.as<TileRec>(), l.ntiles, AP_SMOOTH, in, (T)0.5, out, nullptr, nullptr,
                        nullptr, s);
  GMGS_HIP(hipEventRecord(a, s));
  const int reps = 5;
  for (int w = 0; w < reps; ++w)
    launch_apply<T, TR>(op, l.tiles.as<TileRec>(), l.ntiles, AP_SMOOTH, in, (T)0.5, out, nullptr, nullptr,
                        nullptr, s);
  GMGS_HIP(hipEventRecord(b, s));
  GMGS_HIP(hipEventSynchronize(b));
  float ms = 0;
  GMGS_HIP(hipEventElapsedTime(&ms, a, b));
  return ms / reps;
}

template <typename T, typename TR>
float time_zm(gmgs_ctx* c, int li, const ColTile* tiles, long long nt, int tx, int zc, const TR* in,
              T* out, int mode = AP_SMOOTH, const CycleOut<T>* co = nullptr, const T* in2 = nullptr,
              double* partial = nullptr) {
  Level& l = c->lev[li];
  hipStream_t s = c->stream;
  LevelOp<T> op = make_op<T>(c, li);
  HipEvent a, b;
  for (int w = 0; w < 2; ++w)
    launch_apply_zm<T, TR>(op, tiles, nt, tx, zc, mode, in, (T)0.5, in2, out, partial, nullptr, s, co);
  // the better of two timed batches: one disturbed batch does not decide
  const int reps = 5;
  float best = 1e30f;
  for (int batch = 0; batch < 2; ++batch) {
    GMGS_HIP(hipEventRecord(a, s));
    for (int w = 0; w < reps; ++w)
      launch_apply_zm<T, TR>(op, tiles, nt, tx, zc, mode, in, (T)0.5, in2, out, partial, nullptr, s, co);
    GMGS_HIP(hipEventRecord(b, s));
    GMGS_HIP(hipEventSynchronize(b));
    float ms = 0;
    GMGS_HIP(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms / reps);
  }
  (void)l;
  return best;
}

// the split operator on (tx, zc) column tiles: z-marching kernel on the
// F_FAST nodes, then the assembled rows
template <typename T, typename TR>
float time_split(gmgs_ctx* c, int li, const ColTile* tiles, long long nt, int tx, int zc, const TR* in,
                 T* out) {
  Level& l = c->lev[li];
  hipStream_t s = c->stream;
  const bool keep = l.split;
  l.split = true;
  LevelOp<T> op = make_op<T>(c, li);
  l.split = keep;
  auto run = [&]() {
    launch_apply_zm<T, TR>(op, tiles, nt, tx, zc, AP_SMOOTH, in, (T)0.5, nullptr, out, nullptr, nullptr, s,
                           nullptr, true);
    launch_apply_rows<T, TR>(op, l.rboxes.as<RowTile>(), l.rentries.as<RowEntry>(), l.nrbox, AP_SMOOTH, in,
                             (T)0.5, nullptr, out, nullptr, nullptr, s);
  };
  HipEvent a, b;
  for (int w = 0; w < 2; ++w) run();
  GMGS_HIP(hipEventRecord(a, s));
  const int reps = 5;
  for (int w = 0; w < reps; ++w) run();
  GMGS_HIP(hipEventRecord(b, s));
  GMGS_HIP(hipEventSynchronize(b));
  float ms = 0;
  GMGS_HIP(hipEventElapsedTime(&ms, a, b));
  return ms / reps;
}

// dot partials any level-1 launch may write: the tile kernels (one per 4
// tiles), the pointwise kernels, or the z-marching operator (one per tile)
// partials of any reduction: the PCG kernels and an AX on any level
long long partial_capacity(gmgs_ctx* c) {
  long long n = c->nparts;
  for (const Level& l : c->lev) {
    n = std::max<long long>(n, std::max<long long>(apply_blocks(l.ntiles), std::max({l.ncol, l.ncol64, l.ncolR})));
    if (l.split_ok) n = std::max<long long>(n, l.ncol + (long long)apply_blocks(l.nrbox));
  }
  return n;
}

// a candidate replaces the best so far only when it is 2% faster: shapes
// within timing noise of each other resolve to the earlier one in the list,
// so repeated set-ups of one problem pick the same shapes
inline bool clearly_faster(float ms, float best_ms) { return ms < 0.98f * best_ms; }

// the level-1 prolongation e = P x_c of the mixed cycle, ms per launch (the
// better of two batches of 5 after 2 untimed launches)
float time_prolong(gmgs_ctx* c) {
  Level& l = c->lev[0];
  Level& cr = c->lev[1];
  hipStream_t s = c->stream;
  auto go = [&]() {
    launch_prolong<float>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles, cr.mesh.g,
                          cr.x.as<float>(), l.e.as<float>(), true, nullptr, s);
  };
  HipEvent a, b;
  for (int w = 0; w < 2; ++w) go();
  const int reps = 5;
  float best = 1e30f;
  for (int batch = 0; batch < 2; ++batch) {
    GMGS_HIP(hipEventRecord(a, s));
    for (int w = 0; w < reps; ++w) go();
    GMGS_HIP(hipEventRecord(b, s));
    GMGS_HIP(hipEventSynchronize(b));
    float ms = 0;
    GMGS_HIP(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms / reps);
  }
  return best;
}

void autotune_levels(gmgs_ctx* c) {
  static const int cand[][2] = {{8, 16}, {4, 16}, {16, 16}, {8, 8}, {4, 8}, {16, 8}, {8, 4},
                                {4, 4}, {16, 4}, {8, 32}, {4, 32}};
  hipStream_t s = c->stream;
  for (int li = 0; li < c->L - 1; ++li) {
    Level& l = c->lev[li];
    if (!l.zm && li == 0) continue;
    // the assembled stencil or the split operator was asked for
    if (li > 0 && (c->prefer_as || c->prefer_split) && l.as_built) continue;
    // (level 1 of a mixed solver whose first sweep carries the PCG's residual
    // update is timed on that sweep and cached under its own key)
    const bool rup = li == 0 && c->mixed && c->no_jacobi && c->sym && !c->unfused_resid;
    const TuneKey key{l.mesh.g.ey, l.mesh.g.ex, l.mesh.g.ez, c->mixed ? (rup ? 6 : 1) : 0};
    struct SweepTime { int tx, zc; float ms; };
    std::vector<SweepTime> sweep_times;   // rup, V(1, 1): the first sweep per shape
    // (tx, zc) of the z-marching kernel; (-tx, zc) the split operator on those
    // tiles; (0, 0) the tile kernel, (0, 1) the assembled stencil
    std::pair<int, int> best{l.zm ? l.col_tx : 0, l.zm ? l.col_zc : (l.as ? 1 : 0)};
    bool cached = false;
    {
      std::lock_guard<std::mutex> lk(tune_mutex());
      auto it = tune_cache().find(key);
      if (it != tune_cache().end()) { best = it->second; cached = true; }
    }
    // level 1 of the mixed V(1, 1) cycle: whether its residual kernel forms
    // e = P x_c itself, cached under its own key
    const bool v11 = li == 0 && c->mixed && c->sym && l.nu == 1 && !l.e.empty() && c->L > 1 &&
                     !c->lev[1].x.empty();
    const TuneKey fkey{l.mesh.g.ey, l.mesh.g.ex, l.mesh.g.ez, c->unfused_prolong ? 5 : 4};
    bool best_fuse = l.fuse_prolong;
    if (cached && v11) {
      std::lock_guard<std::mutex> lk(tune_mutex());
      auto it = tune_cache().find(fkey);
      if (it != tune_cache().end()) best_fuse = it->second.first != 0;
      else cached = false;
    }
    if (!cached) {
      float best_ms = 1e30f;
      const float prolong_ms = v11 ? time_prolong(c) : 0.f;
      // rup: the sweep is timed with the scalars of a later iteration and
      // alpha = 0, so the update is loaded but changes nothing (r_old = tmp,
      // the new r into r_a, unused until a solve starts)
      DevBuf tsc;
      CycleOut<float> rco{};
      if (rup) {
        PcgScalars h{};
        h.its = 1;
        tsc.alloc(sizeof(PcgScalars));
        GMGS_HIP(hipMemcpyAsync(tsc.as<void>(), &h, sizeof(h), hipMemcpyHostToDevice, s));
        GMGS_HIP(hipStreamSynchronize(s));
        rco.sc = tsc.as<PcgScalars>();
        rco.qv = c->q.as<double>();
        rco.rn = c->ra.as<double>();
      }
      if (l.zm) {
        for (const auto& cd : cand) {
          bool fuse = false;
          DevBuf tl;
          const long long nt = build_coltiles(l.mesh.g, l.mesh.flags.as<uint32_t>(), cd[0], cd[1], tl, s);
          float ms;
          if (li == 0) {
            if (c->mixed) {
              // (the sweep's |r| partials, one per tile of this shape)
              DevBuf tpart;
              if (rup) tpart.alloc(sizeof(double) * (size_t)std::max(nt, 1LL));
              const float sweep_ms =
                  rup ? time_zm<float, double>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1], c->tmp.as<double>(),
                                               l.rp.as<float>(), AP_SMOOTH_R, &rco, nullptr, tpart.as<double>())
                      : time_zm<float, double>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1], c->tmp.as<double>(),
                                               l.rp.as<float>());
              // V(1, 1) with the residual update in the sweep: the sweep gets
              // tiles of its own (below), these serve the residual kernel
              if (rup && v11) sweep_times.push_back({cd[0], cd[1], sweep_ms});
              ms = rup && v11 ? 0.f : sweep_ms;
              // the symmetric V(1, 1) cycle runs its residual-and-sum kernel on
              // the same tiles: time it too (scratch: e or level 2's x, R_1 =
              // rp, r0 = tmp, z = z[1]), stand-alone after the prolongation
              // and forming e itself, and keep the faster of the two
              if (v11) {
                CycleOut<float> co{};
                co.r0 = c->tmp.as<double>();
                co.z = c->z[1].as<float>();
                co.gc = c->lev[1].mesh.g;
                const float unf = time_zm<float, float>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1],
                                                        l.e.as<float>(), nullptr, AP_RESID_SUM, &co,
                                                        l.rp.as<float>()) + prolong_ms;
                const float fus = c->unfused_prolong
                                      ? 1e30f
                                      : time_zm<float, float>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1],
                                                              c->lev[1].x.as<float>(), nullptr, AP_RESID_SUM_P,
                                                              &co, l.rp.as<float>());
                fuse = fus <= unf;
                ms += fuse ? fus : unf;
              }
            }
            else ms = time_zm<double, double>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1], c->tmp.as<double>(), l.rp.as<double>());
          } else {
            if (c->mixed) ms = time_zm<float, float>(c, li, tl.as<ColTile>(), nt, cd[0], cd[1], l.r.as<float>(), l.rp.as<float>());
            else ms = time_zm<double, double>(c, li, tl.as<ColTile>(), nt, cd[0], cd[1], l.r.as<double>(), l.rp.as<double>());
          }
          if (clearly_faster(ms, best_ms)) {
            best_ms = ms;
            best = {cd[0], cd[1]};
            best_fuse = fuse;
          }
          if (li > 0 && l.split_ok) {
            ms = c->mixed ? time_split<float, float>(c, li, tl.as<ColTile>(), nt, cd[0], cd[1], l.r.as<float>(),
                                                     l.rp.as<float>())
                          : time_split<double, double>(c, li, tl.as<ColTile>(), nt, cd[0], cd[1],
                                                       l.r.as<double>(), l.rp.as<double>());
            if (clearly_faster(ms, best_ms)) { best_ms = ms; best = {-cd[0], cd[1]}; }
          }
        }
      }
      // coarse levels (small, many stored elements) may run faster on the tile
      // kernel, with element sums or the assembled stencil
      if (li > 0) {
        for (int as = 0; as <= (l.as_built ? 1 : 0); ++as) {
          const float ms = c->mixed ? time_v1<float, float>(c, li, l.r.as<float>(), l.rp.as<float>(), as)
                                    : time_v1<double, double>(c, li, l.r.as<double>(), l.rp.as<double>(), as);
          if (clearly_faster(ms, best_ms)) { best_ms = ms; best = {0, as}; }
        }
      }
      std::lock_guard<std::mutex> lk(tune_mutex());
      tune_cache()[key] = best;
      if (v11) tune_cache()[fkey] = {best_fuse ? 1 : 0, 0};
    }
    if (v11) l.fuse_prolong = best_fuse && !c->unfused_prolong;
    if (best.first == 0) {
      l.zm = l.split = false;
      l.as = best.second == 1 && l.as_built;
      continue;
    }
    if (!l.zm) continue;   // a cached z-marching shape this level cannot use
    if (best.first < 0 && !l.split_ok) continue;
    l.as = false;
    l.split = best.first < 0;
    if (l.split) l.zm = false;
    const int tx = std::abs(best.first);
    if (tx != l.col_tx || best.second != l.col_zc) {
      l.col_tx = tx;
      l.col_zc = best.second;
      l.ncol = build_coltiles(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.col_tx, l.col_zc, l.coltiles, s);
    }
    if (li == 0 && rup && v11 && l.zm && l.mesh.g.N >= (1LL << 22)) {
      // the first sweep, which carries the PCG's residual update, on tiles of
      // its own when another shape suits it better (cached under its own key).
      // Only on large grids: a batch of small problems merges the launches of
      // problems whose kernels have equal shapes, and one more tuned shape
      // per problem splits it into more groups for little gain.
      const TuneKey rkey{l.mesh.g.ey, l.mesh.g.ex, l.mesh.g.ez, 7};
      std::pair<int, int> rbest{l.col_tx, l.col_zc};
      bool have = false;
      {
        std::lock_guard<std::mutex> lk(tune_mutex());
        if (!sweep_times.empty()) {
          float bms = 1e30f;
          for (const auto& st : sweep_times)
            if (clearly_faster(st.ms, bms)) { bms = st.ms; rbest = {st.tx, st.zc}; }
          tune_cache()[rkey] = rbest;
          have = true;
        } else {
          auto it = tune_cache().find(rkey);
          if (it != tune_cache().end()) { rbest = it->second; have = true; }
        }
      }
      l.coltilesR = DevBuf();
      l.ncolR = 0;
      if (have && (rbest.first != l.col_tx || rbest.second != l.col_zc)) {
        l.col_txR = rbest.first;
        l.col_zcR = rbest.second;
        l.ncolR = build_coltiles(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.col_txR, l.col_zcR, l.coltilesR, s);
      }
    }
  }
  // level 1 of a mixed context: the FP64 matvec of the PCG on tiles of its own
  Level& l1 = L1(c);
  if (c->mixed && l1.zm) {
    // the kernel the PCG runs: the matvec, or the matvec with the direction
    // update fused (timed on scratch vectors: p_old = tmp, p = tmp2, x = ra)
    const bool fused = fused_direction(c);
    CycleOut<double> co{};
    co.zf = c->z[0].as<float>();
    co.pn = c->tmp2.as<double>();
    co.x = c->ra.as<double>();
    co.sc = c->sc.as<PcgScalars>();
    const TuneKey key{l1.mesh.g.ey, l1.mesh.g.ex, l1.mesh.g.ez, fused ? 3 : 2};
    std::pair<int, int> best{l1.col_tx, l1.col_zc};
    bool cached = false;
    {
      std::lock_guard<std::mutex> lk(tune_mutex());
      auto it = tune_cache().find(key);
      if (it != tune_cache().end()) { best = it->second; cached = true; }
    }
    if (!cached) {
      float best_ms = 1e30f;
      for (const auto& cd : cand) {
        DevBuf tl;
        const long long nt = build_coltiles(l1.mesh.g, l1.mesh.flags.as<uint32_t>(), cd[0], cd[1], tl, s);
        // (fused: the iterations alternate between the kernel with the x
        // updates and the one without; both are timed)
        float ms = time_zm<double, double>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1], c->tmp.as<double>(),
                                           c->q.as<double>(), fused ? AP_AX_P : AP_AX, fused ? &co : nullptr);
        if (fused)
          ms += time_zm<double, double>(c, 0, tl.as<ColTile>(), nt, cd[0], cd[1], c->tmp.as<double>(),
                                        c->q.as<double>(), AP_AX_P_NOX, &co);
        if (clearly_faster(ms, best_ms)) { best_ms = ms; best = {cd[0], cd[1]}; }
      }
      std::lock_guard<std::mutex> lk(tune_mutex());
      tune_cache()[key] = best;
    }
    if (best.first != l1.col_tx || best.second != l1.col_zc) {
      l1.col_tx64 = best.first;
      l1.col_zc64 = best.second;
      l1.ncol64 = build_coltiles(l1.mesh.g, l1.mesh.flags.as<uint32_t>(), best.first, best.second,
                                 l1.coltiles64, s);
    }
  }
  const long long np_max = partial_capacity(c);
  if ((size_t)(sizeof(double) * 2 * (np_max + 1)) > c->part.bytes())
    c->part.alloc_zero(sizeof(double) * 2 * (np_max + 1), s);
  GMGS_HIP(hipStreamSynchronize(s));
}

void alloc_vectors(gmgs_ctx* c) {
  hipStream_t s = c->stream;
  const size_t es = c->mixed ? sizeof(float) : sizeof(double);
  for (int li = 0; li < c->L; ++li) {
    Level& l = c->lev[li];
    const long long n3 = 3 * l.mesh.g.N;
    if (li == 0) {
      l.rp.alloc_zero(es * n3, s);
    } else {
      l.r.alloc_zero(es * n3, s);
      l.x.alloc_zero(es * n3, s);
      if (li < c->L - 1) l.rp.alloc_zero(es * n3, s);
    }
    if (c->sym && li < c->L - 1) {
      const int nb = std::max(2 * l.nu - 2, 2);
      for (int k = 0; k < nb; ++k) l.wk[k].alloc_zero(es * n3, s);
      l.e.alloc_zero(es * n3, s);
    }
  }
  const long long n3 = 3 * L1(c).mesh.g.N;
  c->x.alloc_zero(sizeof(double) * n3, s);
  c->r.alloc_zero(sizeof(double) * n3, s);
  c->ra.alloc_zero(sizeof(double) * n3, s);
  c->p.alloc_zero(sizeof(double) * n3, s);
  if (c->mixed && c->no_jacobi) c->p2.alloc_zero(sizeof(double) * n3, s);
  c->q.alloc_zero(sizeof(double) * n3, s);
  c->tmp.alloc_zero(sizeof(double) * n3, s);
  c->tmp2.alloc_zero(sizeof(double) * n3, s);
  c->z[0].alloc_zero(es * n3, s);
  c->z[1].alloc_zero(es * n3, s);
  c->b.alloc_zero(sizeof(double) * n3, s);
  c->nparts = pointwise_blocks(L1(c).ntiles);
  const long long np_max = partial_capacity(c);
  c->part.alloc_zero(sizeof(double) * 2 * (np_max + 1), s);
  c->sc.alloc_zero(sizeof(PcgScalars), s);
  const long long hcap = (long long)c->prm.max_it + 2;
  c->hist_rel.alloc_zero(sizeof(double) * hcap, s);
  c->hist_t.alloc_zero(sizeof(unsigned long long) * hcap, s);
  GMGS_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ V-cycle
template <typename T>
T* vec(DevBuf& b) { return b.as<T>(); }

// z = V(rin) (kernel/Vcycle.m:1-36); rin is the level-1 FP64 residual;
// returns the number n of r'z partials (r'zprev at [n, 2 n))
template <typename T>
int vcycle(gmgs_ctx* c, const double* rin, T* z, const T* zprev, double* partial, const int* done) {
  hipStream_t s = c->stream;
  const int L = c->L;
  const T omega = (T)c->prm.jacobi_weight;
  // down: level 1
  {
    Level& l = c->lev[0];
    {
      KTime t(c, 2);
      apply_level<T, double>(c, 0, AP_SMOOTH, rin, omega, vec<T>(l.rp), nullptr, nullptr, done);
    }
    Level& n2 = c->lev[1];
    KTime t(c, 3);
    launch_restrict<T>(n2.mesh.g, n2.mesh.flags.as<uint32_t>(), n2.tiles.as<TileRec>(), n2.ntiles,
                       l.mesh.g, vec<T>(l.rp), vec<T>(n2.r), done, s);
  }
  for (int li = 1; li < L - 1; ++li) {
    Level& l = c->lev[li];
    apply_level<T, T>(c, li, AP_SMOOTH, vec<T>(l.r), omega, vec<T>(l.rp), nullptr, nullptr, done);
    Level& nx = c->lev[li + 1];
    launch_restrict<T>(nx.mesh.g, nx.mesh.flags.as<uint32_t>(), nx.tiles.as<TileRec>(), nx.ntiles,
                       l.mesh.g, vec<T>(l.rp), vec<T>(nx.r), done, s);
  }
  // coarsest
  {
    Level& lc = c->lev[L - 1];
    coarsest_apply<T>(c->Kinv.as<double>(), c->nfp, c->nfree, c->free_code.as<long long>(),
                      vec<T>(lc.r), vec<T>(lc.x), lc.mesh.g.N, c->gv_b.as<double>(),
                      c->gv_part.as<double>(), done, s);
  }
  // up
  for (int li = L - 2; li >= 1; --li) {
    Level& l = c->lev[li];
    Level& cr = c->lev[li + 1];
    const T* invd = make_op<T>(c, li).invd;
    launch_prolong_smooth<T, T, T>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(),
                                   l.ntiles, cr.mesh.g, vec<T>(cr.x), invd, vec<T>(l.r),
                                   vec<T>(l.rp), omega, vec<T>(l.x), false, (const T*)nullptr,
                                   nullptr, done, s);
  }
  {
    Level& l = c->lev[0];
    Level& cr = c->lev[1];
    const T* invd = make_op<T>(c, 0).invd;
    KTime t(c, 4);
    launch_prolong_smooth<T, double, T>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(),
                                        l.ntiles, cr.mesh.g, vec<T>(cr.x), invd, rin, vec<T>(l.rp),
                                        omega, z, true, zprev, partial, done, s);
  }
  return (int)pointwise_blocks(L1(c).ntiles);
}

// buffers of the symmetric cycle on a level with nu sweeps: R_1 in rp, R_s
// (s >= 2) in wk[s - 2], Q_0 in wk[nu - 1], Q_1 in the buffer of R_nu (free
// once Q_0 exists), Q_s (s >= 2) in wk[nu + s - 2]
template <typename T>
T* Rbuf(Level& l, int s) { return s == 1 ? vec<T>(l.rp) : vec<T>(l.wk[s - 2]); }
template <typename T>
T* Qbuf(Level& l, int s) {
  if (s == 0) return vec<T>(l.wk[l.nu - 1]);
  if (s == 1) return Rbuf<T>(l, l.nu);
  return vec<T>(l.wk[l.nu + s - 2]);
}

// the PCG's residual update carried by the cycle's first level-1 sweep:
// r = r_old - alpha q
struct ResidUpdate {
  const double* r_old = nullptr;
  const double* q = nullptr;
};

// Symmetric V(nu, nu) cycle (no reference counterpart): nu pre-smoothing
// sweeps R_s = R_{s-1} - A(w_{s-1} R_{s-1}/d) from the level's input R_0, the
// coarse correction e = P x_c, the residual recomputed as Q_0 = R_nu - A e
// (A x_pre = R_0 - R_nu), nu post-smoothing sweeps Q_s from it with the same
// weights, and x = sum_{s<nu} w_s (R_s + Q_s)/d + e.  Unlike the stale
// post-smoothing of kernel/Vcycle.m:27-33 this is a symmetric preconditioner:
// the sweeps of a level are polynomials in D^-1 A, which commute, so the
// post-smoother is the A-adjoint of the pre-smoother for any weights.  The
// weights are the Jacobi weight, or the reciprocal roots of the degree-nu
// Chebyshev polynomial on the level's eigenvalue bounds.  With ru, the first
// level-1 sweep also forms rin = ru->r_old - alpha ru->q (the PCG's residual
// update) and the test on |rin| follows it.
template <typename T>
int vcycle_sym(gmgs_ctx* c, const double* rin, T* z, const T* zprev, double* partial, const int* done,
               const ResidUpdate* ru) {
  hipStream_t s = c->stream;
  const int L = c->L;
  int nz = (int)pointwise_blocks(L1(c).ntiles);
  for (int li = 0; li < L - 1; ++li) {
    Level& l = c->lev[li];
    const int nu = l.nu;
    {
      KTime t(c, li == 0 ? 2 : -1);
      if (li == 0 && ru) {
        CycleOut<T> co{};
        co.sc = c->sc.as<PcgScalars>();
        co.qv = ru->q;
        co.rn = const_cast<double*>(rin);
        const bool own = !l.coltilesR.empty();
        const int nr = launch_apply_zm<T, double>(make_op<T>(c, 0, true),
                                                  own ? l.coltilesR.as<ColTile>() : l.coltiles.as<ColTile>(),
                                                  own ? l.ncolR : l.ncol, own ? l.col_txR : l.col_tx,
                                                  own ? l.col_zcR : l.col_zc, AP_SMOOTH_R, ru->r_old,
                                                  (T)l.om[0], nullptr, Rbuf<T>(l, 1), partial, done, s, &co);
        // |r| and the convergence test (CG_solver.m:35-39); the rest of the
        // iteration returns at once once it is met
        launch_finalize(3, partial, nr, c->sc.as<PcgScalars>(), c->hist_rel.as<double>(),
                        c->hist_t.as<unsigned long long>(), done, s);
      } else if (li == 0) {
        apply_level<T, double>(c, 0, AP_SMOOTH, rin, (T)l.om[0], Rbuf<T>(l, 1), nullptr, nullptr, done,
                               nullptr, true);
      }
      else apply_level<T, T>(c, li, AP_SMOOTH, vec<T>(l.r), (T)l.om[0], Rbuf<T>(l, 1), nullptr, nullptr, done,
                             nullptr, true);
      for (int k = 1; k < nu; ++k)
        apply_level<T, T>(c, li, AP_SMOOTH, Rbuf<T>(l, k), (T)l.om[k], Rbuf<T>(l, k + 1), nullptr, nullptr,
                          done, nullptr, true);
    }
    Level& nx = c->lev[li + 1];
    KTime t(c, li == 0 ? 3 : -1);
    launch_restrict<T>(nx.mesh.g, nx.mesh.flags.as<uint32_t>(), nx.tiles.as<TileRec>(), nx.ntiles,
                       l.mesh.g, Rbuf<T>(l, nu), vec<T>(nx.r), done, s);
  }
  {
    Level& lc = c->lev[L - 1];
    coarsest_apply<T>(c->Kinv.as<double>(), c->nfp, c->nfree, c->free_code.as<long long>(),
                      vec<T>(lc.r), vec<T>(lc.x), lc.mesh.g.N, c->gv_b.as<double>(),
                      c->gv_part.as<double>(), done, s);
  }
  for (int li = L - 2; li >= 0; --li) {
    Level& l = c->lev[li];
    Level& cr = c->lev[li + 1];
    const int nu = l.nu;
    KTime t(c, li == 0 ? 4 : -1);
    const Grid& g = l.mesh.g;
    const uint32_t* fl = l.mesh.flags.as<uint32_t>();
    const TileRec* tl = l.tiles.as<TileRec>();
    const T* invd = make_op<T>(c, li, true).invd;
    if (li == 0 && nu == 1 && l.zm) {
      // V(1, 1): the residual kernel writes the cycle's sum itself (no Q_0,
      // no separate sum) and, unless asked not to, forms e = P x_c itself
      // while it stages it (e is then never stored)
      CycleOut<T> co{rin, nullptr, nullptr, zprev, z, (T)0};
      co.gc = cr.mesh.g;
      if constexpr (std::is_same<T, float>::value) {
        co.ufast = l.ufast_invd && !c->simple_kernels ? 1 : 0;
        for (int k = 0; k < 3; ++k) co.ucinvd[k] = l.ucinvd[k];
      }
      if (!l.fuse_prolong)
        launch_prolong<T>(g, fl, tl, l.ntiles, cr.mesh.g, vec<T>(cr.x), vec<T>(l.e), true, done, s);
      nz = launch_apply_zm<T, T>(make_op<T>(c, 0, true), l.coltiles.as<ColTile>(), l.ncol, l.col_tx, l.col_zc,
                                 l.fuse_prolong ? AP_RESID_SUM_P : AP_RESID_SUM,
                                 l.fuse_prolong ? vec<T>(cr.x) : vec<T>(l.e), (T)l.om[0], Rbuf<T>(l, 1),
                                 nullptr, partial, done, s, &co);
      continue;
    }
    launch_prolong<T>(g, fl, tl, l.ntiles, cr.mesh.g, vec<T>(cr.x), vec<T>(l.e), true, done, s);
    apply_level<T, T>(c, li, AP_RESID, vec<T>(l.e), (T)0, Qbuf<T>(l, 0), nullptr, nullptr, done, Rbuf<T>(l, nu),
                      true);
    if (li == 0 && nu == 2 && l.zm) {
      // the last sweep writes the cycle's sum itself (no Q_1, no separate sum)
      const CycleOut<T> co{rin, Rbuf<T>(l, 1), vec<T>(l.e), zprev, z, (T)l.om[1]};
      nz = launch_apply_zm<T, T>(make_op<T>(c, 0, true), l.coltiles.as<ColTile>(), l.ncol, l.col_tx, l.col_zc,
                                 AP_SMOOTH_SUM, Qbuf<T>(l, 0), (T)l.om[0], nullptr, nullptr, partial, done, s,
                                 &co);
      continue;
    }
    for (int k = 1; k < nu; ++k)
      apply_level<T, T>(c, li, AP_SMOOTH, Qbuf<T>(l, k - 1), (T)l.om[k - 1], Qbuf<T>(l, k), nullptr, nullptr,
                        done, nullptr, true);
    CycleSum<T> cs{};
    cs.nu = nu;
    for (int k = 0; k < nu; ++k) {
      cs.w[k] = (T)l.om[k];
      cs.R[k] = k ? Rbuf<T>(l, k) : nullptr;
      cs.Q[k] = Qbuf<T>(l, k);
    }
    if (li == 0)
      launch_combine<T, double, T>(g, fl, tl, l.ntiles, invd, cs, rin, vec<T>(l.e), z, true, zprev, partial,
                                   done, s);
    else
      launch_combine<T, T, T>(g, fl, tl, l.ntiles, invd, cs, vec<T>(l.r), vec<T>(l.e), vec<T>(l.x), false,
                              (const T*)nullptr, nullptr, done, s);
  }
  return nz;
}

// z = M^-1 rin; returns the number n of r'z partials (r'zprev at [n, 2 n))
template <typename T>
int run_vcycle(gmgs_ctx* c, const double* rin, T* z, const T* zprev, double* partial, const int* done,
               const ResidUpdate* ru = nullptr) {
  if (c->sym) return vcycle_sym<T>(c, rin, z, zprev, partial, done, ru);
  if (ru) throw std::logic_error("residual update in the cycle: symmetric cycle only");
  return vcycle<T>(c, rin, z, zprev, partial, done);
}

// ------------------------------------------------------------------ PCG
// p = z + beta p_old, q = A p (into c->q), x += alpha p_old and the p'q
// partials in one level-1 z-march; returns the number of partials
int launch_matvec_p(gmgs_ctx* c, const double* p_old, double* p, double* x, const float* z, double* part,
                    const int* done, bool with_x) {
  Level& l = L1(c);
  CycleOut<double> co{};
  co.zf = z;
  co.pn = p;
  co.x = x;
  co.sc = c->sc.as<PcgScalars>();
  const LevelOp<double> op = make_op<double>(c, 0);
  const bool own = !l.coltiles64.empty();
  return launch_apply_zm<double, double>(op, own ? l.coltiles64.as<ColTile>() : l.coltiles.as<ColTile>(),
                                         own ? l.ncol64 : l.ncol, own ? l.col_tx64 : l.col_tx,
                                         own ? l.col_zc64 : l.col_zc, with_x ? AP_AX_P : AP_AX_P_NOX, p_old,
                                         0.0, nullptr,
                                         c->q.as<double>(), part, done, c->stream, &co);
}

template <typename T>
void pcg_iteration(gmgs_ctx* c, int k) {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "cycle precision");
  hipStream_t s = c->stream;
  Level& l = L1(c);
  const Grid& g = l.mesh.g;
  const uint32_t* fl = l.mesh.flags.as<uint32_t>();
  const TileRec* tl = l.tiles.as<TileRec>();
  PcgScalars* sc = c->sc.as<PcgScalars>();
  const int* done = &sc->done;
  double* part = c->part.as<double>();
  const int np = (int)c->nparts;
  c->cur_iter = k;
  KTime whole(c, 5);
  // Jacobi step on (x, r) (CG_solver.m:18-19): r_a = r - A(omega r/d) here,
  // x += omega r/d in the update of x below (same sum, r is unchanged until
  // then).  Without it the iteration works on r itself.
  const double* ra = c->no_jacobi ? c->r.as<double>() : c->ra.as<double>();
  if (!c->no_jacobi) {
    KTime t(c, 1);
    apply_level<double, double>(c, 0, AP_SMOOTH, c->r.as<double>(), c->prm.jacobi_weight,
                                c->ra.as<double>(), nullptr, nullptr, done);
  }
  // z = V(r) with r'z and r'z_prev partials (CG_solver.m:21)
  T* z = c->z[k & 1].as<T>();
  const T* zprev = c->flexible ? c->z[(k + 1) & 1].as<T>() : nullptr;
  const bool fres = fused_residual(c);
  ResidUpdate ru{};
  if (fres) {
    // r of the previous iteration in one buffer, this iteration's in the
    // other (r_a is not used without the Jacobi step)
    ru.r_old = (k & 1) ? c->ra.as<double>() : c->r.as<double>();
    ru.q = c->q.as<double>();
    ra = (k & 1) ? c->r.as<double>() : c->ra.as<double>();
  }
  const int nz = run_vcycle<T>(c, ra, z, zprev, part, done, fres ? &ru : nullptr);
  launch_finalize(1, part, nz, sc, nullptr, nullptr, done, s);
  int npq;
  double* p = c->p.as<double>();
  if constexpr (std::is_same<T, float>::value) {
    if (fused_direction(c)) {
      // p = z + beta p_old, q = A p and p'q in one z-march (CG_solver.m:23-31);
      // p alternates between two buffers since tiles read their neighbours'
      // p_old.  The x updates of the last two iterations, x += alpha_prev p_2
      // and x += alpha p_old (p_2 is still in the buffer the new p goes to),
      // ride in the even iterations' z-march, so x is read and written every
      // other iteration
      p = (k & 1) ? c->p2.as<double>() : c->p.as<double>();
      const double* p_old = (k & 1) ? c->p.as<double>() : c->p2.as<double>();
      {
        KTime t(c, GMGS_K_MATVEC64_P);
        npq = launch_matvec_p(c, p_old, p, c->x.as<double>(), z, part, done, (k & 1) == 0);
      }
      launch_finalize(2, part, npq, sc, nullptr, nullptr, done, s);
      if (fres) return;   // r = r - alpha q: in the next iteration's first sweep
      launch_update_xr(g, fl, tl, l.ntiles, nullptr, p, ra, c->q.as<double>(), c->r.as<double>(), nullptr,
                       c->prm.jacobi_weight, sc, part, done, s);
      launch_finalize(3, part, np, sc, c->hist_rel.as<double>(), c->hist_t.as<unsigned long long>(),
                      done, s);
      return;
    }
  }
  // p = z + beta p (CG_solver.m:23-28)
  // without the Jacobi step, x += alpha p of the previous iteration rides in
  // this pass over p (read once per iteration); the last one follows the loop
  launch_update_p<T>(g, fl, tl, l.ntiles, z, p, sc, done, s, c->no_jacobi ? c->x.as<double>() : nullptr);
  // q = A p, p'q (CG_solver.m:29-31)
  {
    KTime t(c, 0);
    npq = apply_level<double, double>(c, 0, AP_AX, p, 0.0, c->q.as<double>(), nullptr, part, done);
  }
  launch_finalize(2, part, npq, sc, nullptr, nullptr, done, s);
  // x += alpha p; r = r - alpha q; ||r|| (CG_solver.m:32-39)
  launch_update_xr(g, fl, tl, l.ntiles, c->no_jacobi ? nullptr : c->x.as<double>(), p, ra,
                   c->q.as<double>(),
                   c->r.as<double>(), c->no_jacobi ? nullptr : l.invd64.as<double>(),
                   c->prm.jacobi_weight, sc, part, done, s);
  launch_finalize(3, part, np, sc, c->hist_rel.as<double>(), c->hist_t.as<unsigned long long>(),
                  done, s);
}

void pcg_init(gmgs_ctx* c, double tol) {
  hipStream_t s = c->stream;
  Level& l = L1(c);
  PcgScalars h{};
  h.tol = tol;
  h.max_it = c->prm.max_it;
  h.flexible = c->flexible ? 1 : 0;
  GMGS_HIP(hipMemcpyAsync(c->sc.as<void>(), &h, sizeof(h), hipMemcpyHostToDevice, s));
  // x = p = z = 0 and r1 = b - AtX(0) = b (CG_solver.m:11-12) on the solid
  // nodes, with the partials of b.b; every other entry of these vectors is a
  // void or halo entry, zero since allocation and never written
  const Grid& g = l.mesh.g;
  const uint32_t* fl = l.mesh.flags.as<uint32_t>();
  const TileRec* tl = l.tiles.as<TileRec>();
  // With the direction update fused into the matvec, iteration 0 reads
  // neither p2 nor x (its p_old is not read, and x is first read by the third
  // iteration's update, which takes it as 0), so neither is written here; p
  // and z0 are written before any read; z1 enters only r'z_prev, which the
  // first iteration's beta = 0 does not use.
  const bool fused = fused_direction(c);
  if (c->mixed)
    launch_pcg_start<float>(g, fl, tl, l.ntiles, c->b.as<double>(), fused ? nullptr : c->x.as<double>(),
                            fused ? nullptr : c->p.as<double>(),
                            fused || c->p2.empty() ? nullptr : c->p2.as<double>(),
                            c->r.as<double>(), fused ? nullptr : c->z[0].as<float>(),
                            fused ? nullptr : c->z[1].as<float>(), c->part.as<double>(), s);
  else
    launch_pcg_start<double>(g, fl, tl, l.ntiles, c->b.as<double>(), c->x.as<double>(), c->p.as<double>(),
                             c->p2.empty() ? nullptr : c->p2.as<double>(), c->r.as<double>(),
                             c->z[0].as<double>(), c->z[1].as<double>(), c->part.as<double>(), s);
  launch_finalize(0, c->part.as<double>(), (int)c->nparts, c->sc.as<PcgScalars>(), nullptr, nullptr,
                  nullptr, s);
}

struct LoopResult {
  int its = 0, converged = 0;
  double relres = 0, seconds = 0;
};

// the x update of the last iteration, deferred into an update of p that the
// finished loop no longer runs (see pcg_iteration)
void finish_deferred_x(gmgs_ctx* c, hipStream_t s) {
  if (!c->no_jacobi) return;
  Level& l = L1(c);
  const bool fused = fused_direction(c);
  launch_add_alpha_p(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles, c->x.as<double>(),
                     c->p.as<double>(), fused ? c->p2.as<double>() : nullptr, c->sc.as<PcgScalars>(), s, fused);
}

// Batch graphs (gmgs_batch_solve): one iteration pair of each context of a
// list, forked from and joined into one stream, plus a gather of their done
// flags; kept for the list until one of its contexts changes.
//
// Contexts whose iteration pairs make the same sequence of kernel launches
// (the same kernels in the same order, any grid sizes) form a group: each
// step of the group is ONE launch over all its contexts (k_multi), whose
// workgroups each do the work of one context's own launch, so every context's
// results are bitwise those of its own solve.  A context whose pair holds a
// kernel that cannot be merged keeps its own launches on a branch of its own.
struct BatchGraph {
  hipGraphExec_t ge = nullptr;
  DevBuf dptrs, dflags;   // the contexts' done flags (device pointers) and their gather
  std::vector<DevBuf> keep;   // argument structs and block maps of the merged launches
  int groups = 0;             // branches of the graph
  long long launches = 0;     // kernel launches per iteration pair
};

// The launches of one iteration pair of c, recorded (inside a stream capture
// that is then discarded, so nothing runs).  False when the pair holds a
// node that did not go through launch() (it could not be merged).
bool record_iteration_pair(gmgs_ctx* c, Recorder& rec) {
  GMGS_HIP(hipStreamSynchronize(c->stream));
  struct Install {
    explicit Install(Recorder* r) { current_recorder() = r; }
    ~Install() { current_recorder() = nullptr; }
  } install(&rec);
  hipGraph_t g = nullptr;
  GMGS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  try {
    for (int k = 0; k < 2; ++k) {
      if (c->mixed) pcg_iteration<float>(c, k);
      else pcg_iteration<double>(c, k);
    }
  } catch (...) {
    if (hipStreamEndCapture(c->stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    throw;
  }
  GMGS_HIP(hipStreamEndCapture(c->stream, &g));
  size_t nn = 0;
  hipError_t e = hipGraphGetNodes(g, nullptr, &nn);
  std::vector<hipGraphNode_t> nodes(nn);
  if (e == hipSuccess && nn) e = hipGraphGetNodes(g, nodes.data(), &nn);
  size_t kernels = 0;
  bool other = false;
  for (size_t i = 0; e == hipSuccess && i < nn; ++i) {
    hipGraphNodeType t;
    e = hipGraphNodeGetType(nodes[i], &t);
    if (t == hipGraphNodeTypeKernel) ++kernels;
    else other = true;
  }
  (void)hipGraphDestroy(g);
  GMGS_HIP(e);
  return !other && kernels == rec.recs.size();
}

bool same_launches(const Recorder& a, const Recorder& b) {
  if (a.recs.size() != b.recs.size()) return false;
  for (size_t i = 0; i < a.recs.size(); ++i)
    if (*a.recs[i].type != *b.recs[i].type) return false;
  return true;
}
using BatchKey = std::vector<gmgs_ctx*>;

std::map<BatchKey, BatchGraph>& batch_cache() {
  static std::map<BatchKey, BatchGraph> m;
  return m;
}

std::mutex& batch_mutex() {
  static std::mutex m;
  return m;
}

void invalidate_graph(gmgs_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  std::lock_guard<std::mutex> lk(batch_mutex());
  auto& bc = batch_cache();
  for (auto it = bc.begin(); it != bc.end();) {
    if (std::find(it->first.begin(), it->first.end(), c) != it->first.end()) {
      if (it->second.ge) (void)hipGraphExecDestroy(it->second.ge);
      it = bc.erase(it);
    } else {
      ++it;
    }
  }
}

// capture iterations k = 0, 1 once; their kernels read every scalar from
// device memory, so the pair replays for every later (even, odd) pair
void capture_iteration_pair(gmgs_ctx* c) {
  hipGraph_t g = nullptr;
  GMGS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  try {
    for (int k = 0; k < 2; ++k) {
      if (c->mixed) pcg_iteration<float>(c, k);
      else pcg_iteration<double>(c, k);
    }
  } catch (...) {
    // leave the stream out of capture mode before the error propagates
    if (hipStreamEndCapture(c->stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    throw;
  }
  GMGS_HIP(hipStreamEndCapture(c->stream, &g));
  const hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    GMGS_HIP(e);
  }
}

LoopResult pcg_run(gmgs_ctx* c) {
  hipStream_t s = c->stream;
  if (!c->pinned) {
    GMGS_HIP(hipHostMalloc((void**)&c->pinned, 2 * sizeof(PcgScalars), hipHostMallocDefault));
    GMGS_HIP(hipEventCreateWithFlags(&c->poll_ev[0], hipEventDisableTiming));
    GMGS_HIP(hipEventCreateWithFlags(&c->poll_ev[1], hipEventDisableTiming));
  }
  const bool graph = c->use_graphs && !c->ktime;
  if (graph && !c->gexec) capture_iteration_pair(c);
  HipEvent e0, e1;
  GMGS_HIP(hipEventRecord(e0, s));
  const int maxk = iteration_cap(c);
  // the scalars' relres is that of iteration its, or of its - 1 when the
  // residual is tested in the iteration after
  const int rel_lag = fused_residual(c) ? 1 : 0;
  int k = 0;   // iterations launched
  PcgScalars h{};
  // queue n iterations (even with graphs: one replay is an iteration pair) and
  // a read-back of the scalars into pinned slot sl
  auto launch = [&](int n, int sl) -> bool {
    if (k >= maxk) return false;
    n = std::min(n, maxk - k + (graph ? 1 : 0));
    if (graph) {
      for (int i = 0; i < n; i += 2, k += 2) GMGS_HIP(hipGraphLaunch(c->gexec, s));
    } else {
      for (int i = 0; i < n; ++i, ++k) {
        if (c->mixed) pcg_iteration<float>(c, k);
        else pcg_iteration<double>(c, k);
      }
    }
    GMGS_HIP(hipMemcpyAsync(&c->pinned[sl], c->sc.as<void>(), sizeof(PcgScalars), hipMemcpyDeviceToHost,
                            s));
    GMGS_HIP(hipEventRecord(c->poll_ev[sl], s));
    return true;
  };
  // Two chunks are always queued: while the device runs the newer one the
  // host reads the older one's scalars and sizes the next chunk from the
  // residual reduction rate seen so far, so that only a couple of iterations
  // run (and exit at once) past convergence.
  bool qa = launch(2, 0), qb = launch(2, 1);
  int a = 0;
  double prev_rel = -1.0;
  int prev_its = 0;
  while (qa) {
    GMGS_HIP(hipEventSynchronize(c->poll_ev[a]));
    const PcgScalars ha = c->pinned[a];
    if (ha.done) break;
    int next = prev_rel > 0.0 ? 16 : 4;   // no rate yet: a short chunk; no progress: a long one
    if (prev_rel > 0.0 && ha.its > prev_its && ha.relres > 0.0 && ha.relres < prev_rel) {
      const double rate = std::pow(ha.relres / prev_rel, 1.0 / (ha.its - prev_its));
      const double need = std::log(ha.tol / ha.relres) / std::log(rate);   // after iteration ha.its - lag
      const int queued = k - (ha.its - rel_lag);                            // launched, not yet read
      const double want = std::ceil(need) - queued;
      next = want < 2.0 ? 2 : (want > 16.0 ? 16 : (int)want);
      next += next & 1;
    }
    prev_rel = ha.relres;
    prev_its = ha.its;
    const bool launched = launch(next, a);
    a ^= 1;
    qa = qb;
    qb = launched;
  }
  finish_deferred_x(c, s);
  GMGS_HIP(hipEventRecord(e1, s));
  GMGS_HIP(hipEventSynchronize(e1));
  GMGS_HIP(hipMemcpy(&h, c->sc.as<void>(), sizeof(h), hipMemcpyDeviceToHost));
  float ms = 0;
  GMGS_HIP(hipEventElapsedTime(&ms, e0, e1));
  LoopResult r;
  r.its = h.its;
  r.converged = h.converged;
  r.relres = h.relres;
  r.seconds = ms * 1e-3;
  if (!(h.relres == h.relres)) throw Status(GMGS_ERR_BREAKDOWN, "NaN in the PCG recursion");
  return r;
}

double dot_host(gmgs_ctx* c, const double* a, const double* b) {
  Level& l = L1(c);
  launch_dot(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles, a, b,
             c->part.as<double>(), c->stream);
  std::vector<double> h = download<double>(c->part, c->nparts, c->stream);
  double sum = 0.0;
  for (double v : h) sum += v;
  return sum;
}

// ------------------------------------------------------- smoother bounds
// Chebyshev sweep weights of the symmetric cycle (no reference counterpart)
constexpr int kLanczosSteps = 20;       // CG steps of the lambda_max estimate
constexpr double kChebSafety = 1.1;     // upper bound = kChebSafety * estimate
constexpr double kChebLower = 0.05;     // lower bound = kChebLower * upper bound

// largest eigenvalue of the symmetric tridiagonal (diag d, off-diagonal e) by
// bisection on Sturm sequence counts
double tridiag_max_eig(const std::vector<double>& d, const std::vector<double>& e) {
  const size_t n = d.size();
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < n; ++i) {
    const double rad = (i ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - rad);
    hi = std::max(hi, d[i] + rad);
  }
  // number of eigenvalues below x
  auto below = [&](double x) {
    size_t cnt = 0;
    double q = 1.0;
    for (size_t i = 0; i < n; ++i) {
      q = d[i] - x - (i ? e[i - 1] * e[i - 1] / q : 0.0);
      if (q == 0.0) q = -1e-300;
      if (q < 0.0) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) == n) hi = mid;
    else lo = mid;
  }
  return hi;
}

// lambda_max(D^-1 A) on the free DOFs of level li (cycle precision): the top
// Ritz value of kLanczosSteps Jacobi-preconditioned CG iterations from a fixed
// pseudo-random start.  The CG coefficients give the Lanczos matrix:
// T_kk = 1/alpha_k + beta_{k-1}/alpha_{k-1}, T_k,k+1 = sqrt(beta_k)/alpha_k.
// Uses the level's cycle work vectors; 0 when the level has no free DOF.
template <typename T>
double estimate_lmax(gmgs_ctx* c, int li) {
  hipStream_t s = c->stream;
  Level& l = c->lev[li];
  const Grid& g = l.mesh.g;
  const uint32_t* fl = l.mesh.flags.as<uint32_t>();
  const TileRec* tl = l.tiles.as<TileRec>();
  const T* invd = make_op<T>(c, li, true).invd;
  T* r = vec<T>(l.e);
  T* z = vec<T>(l.rp);
  T* p = vec<T>(l.wk[0]);
  T* q = vec<T>(l.wk[1]);
  double* part = c->part.as<double>();
  auto total = [&](long long np) {
    const std::vector<double> h = download<double>(c->part, (size_t)np, s);
    double t = 0.0;
    for (double v : h) t += v;
    return t;
  };
  const long long npw = pointwise_blocks(l.ntiles);
  launch_seed<T>(g, fl, tl, l.ntiles, r, s);
  launch_jcg_step<T>(g, fl, tl, l.ntiles, invd, (const T*)nullptr, 0.0, r, z, part, s);
  double rz = total(npw);
  if (!(rz > 0.0)) return 0.0;
  const double rz0 = rz;
  launch_xpby<T>(g, fl, tl, l.ntiles, z, 0.0, p, s);
  std::vector<double> al, be;
  for (int k = 0; k < kLanczosSteps; ++k) {
    const int np = apply_level<T, T>(c, li, AP_AX, p, (T)0, q, nullptr, part, nullptr, nullptr, true);
    const double pq = total(np);
    if (!(pq > 0.0)) break;
    const double alpha = rz / pq;
    launch_jcg_step<T>(g, fl, tl, l.ntiles, invd, q, alpha, r, z, part, s);
    const double rzn = total(npw);
    al.push_back(alpha);
    be.push_back(rzn / rz);
    if (!(rzn > 1e-24 * rz0)) break;
    launch_xpby<T>(g, fl, tl, l.ntiles, z, rzn / rz, p, s);
    rz = rzn;
  }
  if (al.empty()) throw Status(GMGS_ERR_BREAKDOWN, "smoother bound estimate: operator not positive");
  const size_t n = al.size();
  std::vector<double> d(n), e(n > 1 ? n - 1 : 0);
  for (size_t i = 0; i < n; ++i) {
    d[i] = 1.0 / al[i] + (i ? be[i - 1] / al[i - 1] : 0.0);
    if (i + 1 < n) e[i] = std::sqrt(be[i]) / al[i];
  }
  return tridiag_max_eig(d, e);
}

// sweep weights of every smoothed level: the Jacobi weight, or with
// Chebyshev smoothing w_i = 1/(theta + delta cos((2i-1) pi/(2 nu))), the
// reciprocal roots of the degree-nu Chebyshev polynomial shifted to
// [kChebLower b, b], b = kChebSafety lambda_max, theta and delta the centre
// and half-width of that interval
void smoother_weights(gmgs_ctx* c) {
  const double pi = 3.14159265358979323846;
  for (int li = 0; li < c->L - 1; ++li) {
    Level& l = c->lev[li];
    for (int k = 0; k < kMaxSweeps; ++k) l.om[k] = k < l.nu ? c->prm.jacobi_weight : 0.0;
    l.lmax = 0.0;
    if (!(c->sym && c->cheb)) continue;
    const double lm = c->mixed ? estimate_lmax<float>(c, li) : estimate_lmax<double>(c, li);
    if (!(lm > 0.0)) continue;
    l.lmax = lm;
    const double b = kChebSafety * lm, a = kChebLower * b;
    const double theta = 0.5 * (b + a), delta = 0.5 * (b - a);
    for (int i = 1; i <= l.nu; ++i)
      l.om[i - 1] = 1.0 / (theta + delta * std::cos((2 * i - 1) * pi / (2 * l.nu)));
  }
  GMGS_HIP(hipStreamSynchronize(c->stream));
}

void set_bc_impl(gmgs_ctx* c, const int32_t* fixed, int64_t k, const double* loads, int64_t m) {
  hipStream_t s = c->stream;
  // the BC-dependent state (F_, fixed flags, coarsest factor, smoother
  // weights) is rebuilt in place: until this call completes the context has
  // no valid BCs, so a failure part-way leaves solves refused (GMGS_ERR_STATE)
  // instead of mixing old buffers with the new free-DOF count
  c->bc = false;
  c->have_solution = false;
  invalidate_graph(c);
  Level& l = L1(c);
  const long long nn = l.mesh.nn;
  if (m <= 0 || !loads) throw Status(GMGS_ERR_NO_LOADS, "No Loads! Unable to Perform Static FEA!");
  if (k <= 0 || !fixed) throw Status(GMGS_ERR_NO_CONSTRAINTS, "No Constraints! Unable to Perform Static FEA!");
  // F_(loadingCond_(:,1),:) = loadingCond_(:,2:4): the last duplicate row wins
  c->F.assign(3 * nn, 0.0);
  for (int64_t i = 0; i < m; ++i) {
    const double id = loads[i];
    if (!(id >= 1 && id <= (double)nn && std::floor(id) == id))
      throw Status(GMGS_ERR_NODE_ID, "load row " + std::to_string(i + 1) + ": node id out of range");
    for (int cc = 0; cc < 3; ++cc) {
      const double v = loads[i + m * (1 + cc)];
      if (!std::isfinite(v)) throw Status(GMGS_ERR_NONFINITE, "non-finite load value");
      c->F[3 * ((long long)id - 1) + cc] = v;
    }
  }
  for (int64_t i = 0; i < k; ++i)
    if (fixed[i] < 1 || fixed[i] > nn)
      throw Status(GMGS_ERR_NODE_ID, "fixed node " + std::to_string(i + 1) + ": id out of range");
  DevBuf dF;
  upload(dF, c->F.data(), c->F.size(), s);
  const long long n3 = 3 * l.mesh.g.N;
  launch_fill<double>(c->b.as<double>(), 0.0, n3, s);
  launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, dF.as<double>(), c->b.as<double>(),
                         l.mesh.g.N, s);
  // fixed flags on every level (ApplyBoundaryCondition.m:14-20, SetupSolver.m:246-254)
  DevBuf dfix;
  upload(dfix, fixed, (size_t)k, s);
  for (int li = 0; li < c->L; ++li)
    clear_fixed(c->lev[li].mesh.g, c->lev[li].mesh.flags.as<uint32_t>(), s);
  set_fixed_nodes(l.mesh.node_pidx.as<long long>(), dfix.as<int>(), k, l.mesh.flags.as<uint32_t>(), s);
  for (int li = 1; li < c->L; ++li) {
    Level& cl = c->lev[li];
    Level& fl = c->lev[li - 1];
    restrict_fixed(cl.mesh.g, cl.tiles.as<TileRec>(), cl.ntiles, cl.mesh.flags.as<uint32_t>(),
                   fl.mesh.g, fl.mesh.flags.as<uint32_t>(), s);
  }
  // free-DOF inverse diagonals of the symmetric cycle (cycle precision)
  if (c->sym) {
    for (int li = 0; li < c->L - 1; ++li) {
      Level& lv = c->lev[li];
      const long long n3 = 3 * lv.mesh.g.N;
      const uint32_t* fg = lv.mesh.flags.as<uint32_t>();
      if (c->mixed) {
        lv.invdm.alloc_zero(sizeof(float) * n3, s);
        launch_mask_fixed<float>(lv.mesh.g, fg, lv.tiles.as<TileRec>(), lv.ntiles, lv.invd32.as<float>(),
                                 lv.invdm.as<float>(), s);
      } else {
        lv.invdm.alloc_zero(sizeof(double) * n3, s);
        launch_mask_fixed<double>(lv.mesh.g, fg, lv.tiles.as<TileRec>(), lv.ntiles, lv.invd64.as<double>(),
                                  lv.invdm.as<double>(), s);
      }
    }
  }
  GMGS_HIP(hipStreamSynchronize(s));
  // coarsest: assemble K_ff, factor, invert (SetupSolver.m:257-259)
  Level& lc = c->lev[c->L - 1];
  c->nfree = coarsest_free_dofs(lc.mesh.g, lc.mesh.node_pidx.as<long long>(), lc.mesh.nn,
                                lc.mesh.flags.as<uint32_t>(), c->free_code, c->dofmap, s);
  c->nfp = std::max<long long>(DB, (c->nfree + DB - 1) / DB * DB);
  c->A.alloc(sizeof(double) * c->nfp * c->nfp);
  assemble_coarsest(lc.mesh.g, lc.mesh.flags.as<uint32_t>(), lc.eop.as<int>(), lc.Ke64.as<double>(),
                    lc.Ks64.empty() ? nullptr : lc.Ks64.as<double>(), c->free_code.as<long long>(),
                    c->dofmap.as<int>(), c->nfree, c->nfp, c->A.as<double>(), s);
  const int info = cholesky(c->A.as<double>(), c->nfp, s);
  if (info != 0)
    throw Status(GMGS_ERR_NOT_SPD, "coarsest-level stiffness matrix is not positive definite (pivot "
                                       + std::to_string(info) + ")");
  c->W.alloc(sizeof(double) * c->nfp * c->nfp);
  c->Kinv.alloc(sizeof(double) * c->nfp * c->nfp);
  inverse_from_cholesky(c->A.as<double>(), c->W.as<double>(), c->Kinv.as<double>(), c->nfp, s);
  c->W.release();
  c->gv_b.alloc_zero(sizeof(double) * c->nfp, s);
  c->gv_part.alloc_zero(sizeof(double) * coarsest_partial_elems(c->nfp), s);
  GMGS_HIP(hipStreamSynchronize(s));
  smoother_weights(c);
  c->bc = true;
  c->have_solution = false;
  if (c->bc_fixed.data() != fixed) {
    c->bc_fixed.assign(fixed, fixed + k);
    c->bc_loads.assign(loads, loads + 4 * m);
    c->bc_m = m;
  }
}

thread_local std::string g_create_error;

int fail(gmgs_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

template <typename F>
int guarded(gmgs_ctx* c, F&& f) {
  try {
    return f();
  } catch (const Status& e) {
    return fail(c, e.code, e.what());
  } catch (const HipError& e) {
    return fail(c, GMGS_ERR_DEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return fail(c, GMGS_ERR_DEVICE, "host out of memory");
  } catch (const std::exception& e) {
    return fail(c, GMGS_ERR_DEVICE, e.what());
  }
}

long long level_dofs(gmgs_ctx* c, int li) { return 3 * c->lev[li].mesh.nn; }

void check_level(gmgs_ctx* c, int level) {
  if (level < 1 || level > c->L) throw Status(GMGS_ERR_ARGUMENT, "level out of range");
}

}  // namespace

// AtX of one level on device-resident compact vectors, with the two padded
// level vectors it needs taken from caller-provided device scratch (the
// entry point that allocates nothing passes its file-scope device array):
// scratch(device, &bytes) returns that device's scratch address and size.
// Nothing here allocates: the operator data belong to the context, x and y to
// the caller.
namespace gmgs {
int atx_device_scratch(gmgs_ctx* c, int level, const double* x, double* y,
                       void* (*scratch)(int device, size_t* bytes)) {
  if (!c || c->lev.empty() || !x || !y || !scratch) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    check_level(c, level);
    if (!c->bc) throw Status(GMGS_ERR_STATE, "boundary conditions not set (fixed DOFs unknown)");
    hipStream_t s = c->stream;
    Level& l = c->lev[level - 1];
    const long long nn = l.mesh.nn, N = l.mesh.g.N;
    size_t have = 0;
    double* base = static_cast<double*>(scratch(c->device, &have));
    if (!base) throw Status(GMGS_ERR_DEVICE, "device scratch array not found on this device");
    const size_t need = sizeof(double) * 6 * (size_t)N;
    if (need > have)
      throw Status(GMGS_ERR_ARGUMENT, "level " + std::to_string(level) + " needs " + std::to_string(need)
                                          + " bytes of scratch (two padded level vectors), more than the "
                                          + std::to_string(have) + " available");
    double* din = base;
    double* dout = base + 3 * N;
    // void and halo entries of the input must read as zero; the output is
    // read back only at solid nodes, which the operator writes
    launch_fill<double>(din, 0.0, 3 * N, s);
    launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, x, din, N, s);
    apply_level<double, double>(c, level - 1, AP_AX, din, 0.0, dout, nullptr, nullptr, nullptr);
    launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), nn, dout, N, y, s);
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}
}  // namespace gmgs

// ====================================================================== C ABI
extern "C" {

void gmgs_default_params(gmgs_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->youngs_modulus = 1.0;                 // GlobalVariables.m:8
  p->poisson_ratio = 0.3;                  // GlobalVariables.m:9
  p->tol = 1.0e-3;                         // GlobalVariables.m:3
  p->jacobi_weight = 0.4;                  // GlobalVariables.m:5
  p->coarsest_resolution_control = 1.0e4;  // GlobalVariables.m:23
  p->max_it = 20000;                       // GlobalVariables.m:4
  p->num_levels = 0;
  p->precision = GMGS_PREC_FP64;
  p->flags = 0;
  p->device = 0;
}

const char* gmgs_status_string(int st) {
  switch (st) {
    case GMGS_OK: return "ok";
    case GMGS_ERR_ARGUMENT: return "invalid argument";
    case GMGS_ERR_RESOLUTION: return "insufficient resolution";
    case GMGS_ERR_LEVELS: return "fewer than two levels";
    case GMGS_ERR_NO_LOADS: return "no loads";
    case GMGS_ERR_NO_CONSTRAINTS: return "no constraints";
    case GMGS_ERR_NODE_ID: return "node id out of range";
    case GMGS_ERR_NONFINITE: return "non-finite input";
    case GMGS_ERR_NOT_SPD: return "coarsest matrix not positive definite";
    case GMGS_ERR_NOT_CONVERGED: return "maximum iterations exceeded";
    case GMGS_ERR_STATE: return "invalid call order";
    case GMGS_ERR_DEVICE: return "device error";
    case GMGS_ERR_BREAKDOWN: return "PCG breakdown";
    default: return "unknown status";
  }
}

int gmgs_create(const gmgs_params* p, const int32_t* vox, int64_t ny, int64_t nx, int64_t nz,
                const double* bbox, gmgs_ctx** out) {
  if (!out) return GMGS_ERR_ARGUMENT;
  *out = nullptr;
  if (!p || !vox || ny <= 0 || nx <= 0 || nz <= 0) return GMGS_ERR_ARGUMENT;
  auto c = std::make_unique<gmgs_ctx>();
  int st = guarded(c.get(), [&]() -> int {
    c->prm = *p;
    if (!(p->poisson_ratio > -1.0 && p->poisson_ratio < 0.5) || !(p->youngs_modulus > 0) ||
        !(p->tol > 0) || p->max_it < 0 || !(p->jacobi_weight > 0) || p->num_levels < 0 ||
        (p->precision != GMGS_PREC_FP64 && p->precision != GMGS_PREC_MIXED) ||
        p->vcycle_sweeps < 0 || p->vcycle_sweeps > kMaxSweeps || p->vcycle_sweeps_coarse < 0 ||
        p->vcycle_sweeps_coarse > kMaxSweeps || p->vcycle_sweeps_level2 < 0 ||
        p->vcycle_sweeps_level2 > kMaxSweeps)
      throw Status(GMGS_ERR_ARGUMENT, "invalid solver parameter");
    c->device = p->device;
    GMGS_HIP(hipSetDevice(c->device));
    GMGS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) == hipSuccess && khz > 0)
      c->clock_khz = khz;
    c->mixed = p->precision == GMGS_PREC_MIXED;
    c->flexible = c->mixed ? !(p->flags & GMGS_FLAG_FR_BETA) : (p->flags & GMGS_FLAG_PR_BETA) != 0;
    c->exact = (p->flags & GMGS_FLAG_EXACT_GALERKIN) != 0;
    c->tables = (p->flags & GMGS_FLAG_NO_INDEX_TABLES) == 0;
    c->simple_kernels = (p->flags & GMGS_FLAG_SIMPLE_KERNELS) != 0;
    c->unfused_prolong = (p->flags & GMGS_FLAG_UNFUSED_PROLONG) != 0;
    c->unfused_resid = (p->flags & GMGS_FLAG_UNFUSED_RESIDUAL) != 0;
    c->autotune = (p->flags & GMGS_FLAG_AUTOTUNE) != 0;
    c->use_graphs = (p->flags & GMGS_FLAG_NO_GRAPHS) == 0;
    c->sym = (p->flags & GMGS_FLAG_SYM_VCYCLE) != 0;
    c->sweeps = p->vcycle_sweeps == 0 ? 2 : p->vcycle_sweeps;
    c->sweeps_coarse = p->vcycle_sweeps_coarse == 0 ? c->sweeps : p->vcycle_sweeps_coarse;
    c->sweeps_l2 = p->vcycle_sweeps_level2 == 0 ? c->sweeps_coarse : p->vcycle_sweeps_level2;
    c->no_jacobi = (p->flags & GMGS_FLAG_NO_JACOBI_STEP) != 0;
    c->cheb = (p->flags & GMGS_FLAG_CHEBYSHEV) != 0;
    c->prefer_as = (p->flags & GMGS_FLAG_ASSEMBLED_COARSE) != 0;
    c->prefer_split = (p->flags & GMGS_FLAG_SPLIT_COARSE) != 0;
    if (c->cheb && !c->sym)
      throw Status(GMGS_ERR_ARGUMENT, "Chebyshev smoothing needs the symmetric cycle");
    c->nely = (int)ny;
    c->nelx = (int)nx;
    c->nelz = (int)nz;
    if (bbox) {
      for (int i = 0; i < 6; ++i) {
        if (!std::isfinite(bbox[i])) throw Status(GMGS_ERR_NONFINITE, "non-finite bounding box");
        c->bbox[i] = bbox[i];
      }
    } else {
      c->bbox[3] = (double)nx; c->bbox[4] = (double)ny; c->bbox[5] = (double)nz;
    }
    for (int d = 0; d < 3; ++d)
      if (!(c->bbox[3 + d] > c->bbox[d])) throw Status(GMGS_ERR_ARGUMENT, "empty bounding box");
    build_mesh(c.get(), vox);
    for (int li = 0; li < c->L; ++li)
      c->lev[li].nu = li == 0 ? c->sweeps : (li == 1 ? c->sweeps_l2 : c->sweeps_coarse);
    double probe[8][3];
    probe_coords(c.get(), probe);
    element_stiffness(p->youngs_modulus, p->poisson_ratio, probe, c->Ke0);
    if (!stress_operator(p->youngs_modulus, p->poisson_ratio, probe, c->M))
      throw Status(GMGS_ERR_ARGUMENT, "singular Gauss-point interpolation matrix");
    double W[8][8][8];
    child_weights(W);
    set_galerkin_weights(W, c->stream);
    for (int li = 0; li < c->L; ++li) upload_level_constants(c.get(), li);
    upload(c->stressM, c->M, 48 * 24, c->stream);
    build_operators(c.get());
    alloc_vectors(c.get());
    for (Level& l : c->lev)
      for (int k = 0; k < kMaxSweeps; ++k) l.om[k] = k < l.nu ? p->jacobi_weight : 0.0;
    if (c->autotune) autotune_levels(c.get());
    return GMGS_OK;
  });
  if (st != GMGS_OK) {
    g_create_error = c->err;      // reachable through gmgs_last_error(NULL)
    gmgs_destroy(c.release());
    return st;
  }
  *out = c.release();
  return GMGS_OK;
}

int gmgs_set_bc(gmgs_ctx* c, const int32_t* fixed, int64_t k, const double* loads, int64_t m) {
  if (!c || c->lev.empty()) return GMGS_ERR_STATE;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    set_bc_impl(c, fixed, k, loads, m);
    return GMGS_OK;
  });
}

int gmgs_set_modulus(gmgs_ctx* c, const double* modulus, int64_t n) {
  if (!c || c->lev.empty()) return GMGS_ERR_STATE;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    Level& l = L1(c);
    hipStream_t s = c->stream;
    if (modulus == nullptr || n == 0) {
      c->simp = false;
    } else {
      if (n != l.mesh.ne) throw Status(GMGS_ERR_ARGUMENT, "modulus length must equal numElements");
      for (int64_t i = 0; i < n; ++i)
        if (!(modulus[i] > 0) || !std::isfinite(modulus[i]))
          throw Status(GMGS_ERR_NONFINITE, "modulus values must be finite and positive");
      DevBuf dm;
      upload(dm, modulus, (size_t)n, s);
      c->emod64.alloc_zero(sizeof(double) * l.mesh.g.N, s);
      // scatter compact element values to their cells
      std::vector<long long> ep = download<long long>(l.mesh.elem_pidx, n, s);
      std::vector<double> dense((size_t)l.mesh.g.N, 0.0);
      for (int64_t i = 0; i < n; ++i) dense[ep[i]] = modulus[i];
      upload(c->emod64, dense.data(), dense.size(), s);
      if (c->mixed) to_f32(c->emod64, c->emod32, l.mesh.g.N, s);
      c->simp = true;
    }
    invalidate_graph(c);
    // the operators are rebuilt in place: BCs count as unset until the
    // BC-dependent set-up has been redone on them
    const bool had_bc = c->bc;
    c->bc = false;
    c->have_solution = false;
    build_operators(c);
    if (had_bc) {
      // the coarse operators changed: redo the BC-dependent set-up
      std::vector<int32_t> fx = c->bc_fixed;
      std::vector<double> ld = c->bc_loads;
      set_bc_impl(c, fx.data(), (int64_t)fx.size(), ld.data(), c->bc_m);
      c->bc_fixed = fx;
      c->bc_loads = ld;
    }
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

static int finish_solve(gmgs_ctx* c, const LoopResult& r, gmgs_solve_info* info, double* hist,
                        int64_t hist_cap);

static int solve_into(gmgs_ctx* c, gmgs_solve_info* info, double* hist, int64_t hist_cap) {
  pcg_init(c, c->prm.tol);
  LoopResult r = pcg_run(c);
  if (c->ktime) kt_collect(c, r.its);
  return finish_solve(c, r, info, hist, hist_cap);
}

static int finish_solve(gmgs_ctx* c, const LoopResult& r, gmgs_solve_info* info, double* hist,
                        int64_t hist_cap) {
  hipStream_t s = c->stream;
  Level& l = L1(c);
  // true FP64 residual at exit: tmp = b - A x and the partials of |tmp|^2,
  // in one z-march where level 1 has it
  long long np = c->nparts;
  if (l.zm && !l.split) {
    np = apply_level<double, double>(c, 0, AP_RESID, c->x.as<double>(), 0.0, c->tmp.as<double>(), nullptr,
                                     c->part.as<double>(), nullptr, c->b.as<double>());
  } else {
    apply_level<double, double>(c, 0, AP_AX, c->x.as<double>(), 0.0, c->tmp.as<double>(), nullptr, nullptr,
                                nullptr);
    launch_resid_sq(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles, c->b.as<double>(),
                    c->tmp.as<double>(), c->part.as<double>(), s);
  }
  std::vector<double> parts((size_t)np);
  PcgScalars h{};
  GMGS_HIP(hipMemcpyAsync(parts.data(), c->part.as<void>(), sizeof(double) * parts.size(),
                          hipMemcpyDeviceToHost, s));
  GMGS_HIP(hipMemcpyAsync(&h, c->sc.as<void>(), sizeof(h), hipMemcpyDeviceToHost, s));
  GMGS_HIP(hipStreamSynchronize(s));
  double rr = 0.0;
  for (double v : parts) rr += v;
  const double tr = std::sqrt(rr) / h.normB;   // normB = |b| (pcg_init)
  const long long nh = r.converged ? r.its - 1 : r.its;
  if (info) {
    info->its = r.its;
    info->converged = r.converged;
    info->relres = r.relres;
    info->true_relres = tr;
    info->solve_seconds = r.seconds;
    info->hist_len = nh;
  }
  if (hist && hist_cap > 0 && nh > 0) {
    std::vector<double> hr = download<double>(c->hist_rel, (size_t)nh, s);
    std::vector<unsigned long long> ht = download<unsigned long long>(c->hist_t, (size_t)nh, s);
    for (long long i = 0; i < nh && i < hist_cap; ++i) {
      hist[3 * i + 0] = (double)(i + 1);
      hist[3 * i + 1] = (double)(ht[i] - h.t0) / (c->clock_khz * 1e3);
      hist[3 * i + 2] = hr[i];
    }
  }
  c->have_solution = true;
  return r.converged ? GMGS_OK : GMGS_ERR_NOT_CONVERGED;
}

int gmgs_solve(gmgs_ctx* c, double* U, gmgs_solve_info* info, double* hist, int64_t hist_cap) {
  if (!c || c->lev.empty()) return GMGS_ERR_STATE;
  if (!c->bc) return fail(c, GMGS_ERR_STATE, "boundary conditions not set");
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    const int st = solve_into(c, info, hist, hist_cap);
    if (U) {
      Level& l = L1(c);
      DevBuf comp;
      comp.alloc(sizeof(double) * 3 * l.mesh.nn);
      launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), l.mesh.nn, c->x.as<double>(),
                                    l.mesh.g.N, comp.as<double>(), c->stream);
      GMGS_HIP(hipMemcpyAsync(U, comp.as<void>(), sizeof(double) * 3 * l.mesh.nn,
                              hipMemcpyDeviceToHost, c->stream));
      GMGS_HIP(hipStreamSynchronize(c->stream));
    }
    if (st == GMGS_ERR_NOT_CONVERGED) c->err = "Exceed the maximum iterate numbers";
    return st;
  });
}

static void copy_solution(gmgs_ctx* c, double* U) {
  Level& l = L1(c);
  DevBuf comp;
  comp.alloc(sizeof(double) * 3 * l.mesh.nn);
  launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), l.mesh.nn, c->x.as<double>(),
                                l.mesh.g.N, comp.as<double>(), c->stream);
  GMGS_HIP(hipMemcpyAsync(U, comp.as<void>(), sizeof(double) * 3 * l.mesh.nn, hipMemcpyDeviceToHost,
                          c->stream));
  GMGS_HIP(hipStreamSynchronize(c->stream));
}

int gmgs_batch_solve(gmgs_ctx** cs, int n, double** U, gmgs_solve_info* infos) {
  if (!cs || n <= 0) return GMGS_ERR_ARGUMENT;
  for (int i = 0; i < n; ++i) {
    if (!cs[i] || cs[i]->lev.empty()) return GMGS_ERR_ARGUMENT;
    if (!cs[i]->bc) return fail(cs[i], GMGS_ERR_STATE, "boundary conditions not set");
    if (cs[i]->device != cs[0]->device) return fail(cs[i], GMGS_ERR_ARGUMENT, "contexts on different devices");
  }
  {
    std::vector<gmgs_ctx*> u(cs, cs + n);
    std::sort(u.begin(), u.end());
    if (std::adjacent_find(u.begin(), u.end()) != u.end())
      return fail(cs[0], GMGS_ERR_ARGUMENT, "a context appears twice in the batch");
  }
  gmgs_ctx* c0 = cs[0];
  return guarded(c0, [&]() -> int {
    GMGS_HIP(hipSetDevice(c0->device));
    // owned here and released on every exit path
    HipStream s0;
    std::vector<std::unique_ptr<HipEvent>> evs;
    std::vector<hipEvent_t> ev;
    for (int i = 0; i < n + 2; ++i) {
      evs.push_back(std::make_unique<HipEvent>(hipEventDisableTiming));
      ev.push_back(*evs.back());
    }
    PinnedBuf<int> pinned((size_t)n);
    int maxk = 0;
    for (int i = 0; i < n; ++i) {
      pcg_init(cs[i], cs[i]->prm.tol);
      GMGS_HIP(hipEventRecord(ev[i], cs[i]->stream));
      GMGS_HIP(hipStreamWaitEvent(s0, ev[i], 0));
      maxk = std::max(maxk, iteration_cap(cs[i]));
    }
    // the graph of one iteration pair of every context of 'act' (merged
    // launches per group, fork/join across groups), from the cache or built now
    auto graph_for = [&](const BatchKey& act) -> BatchGraph& {
      std::lock_guard<std::mutex> lk(batch_mutex());
      auto it = batch_cache().find(act);
      if (it != batch_cache().end()) return it->second;
      const int m = (int)act.size();
      // record every context's pair and group equal sequences
      std::vector<Recorder> recs(m);
      std::vector<int> mergeable(m);
      for (int i = 0; i < m; ++i) mergeable[i] = record_iteration_pair(act[i], recs[i]) ? 1 : 0;
      std::vector<std::vector<int>> groups;
      for (int i = 0; i < m; ++i) {
        bool placed = false;
        for (auto& gr : groups)
          if (mergeable[i] && mergeable[gr[0]] && same_launches(recs[gr[0]], recs[i])) {
            gr.push_back(i);
            placed = true;
            break;
          }
        if (!placed) groups.push_back({i});
      }
      BatchGraph bg;
      std::vector<const int*> hp(m);
      for (int i = 0; i < m; ++i) hp[i] = &act[i]->sc.as<PcgScalars>()->done;
      upload(bg.dptrs, hp.data(), (size_t)m, s0);
      bg.dflags.alloc_zero(sizeof(int) * m, s0);
      // argument structs and block maps of every merged step, before the capture
      std::vector<std::vector<LaunchRec::Multi>> steps(groups.size());
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const auto& gr = groups[gi];
        if (!mergeable[gr[0]]) continue;
        const size_t ns = recs[gr[0]].recs.size();
        for (size_t k = 0; k < ns; ++k) {
          std::vector<const LaunchRec*> rk;
          for (int i : gr) rk.push_back(&recs[i].recs[k]);
          steps[gi].push_back(rk[0]->prepare(rk, bg.keep, s0));
        }
      }
      GMGS_HIP(hipStreamSynchronize(s0));
      hipGraph_t g = nullptr;
      GMGS_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
      GMGS_HIP(hipEventRecord(ev[n], s0));
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const auto& gr = groups[gi];
        gmgs_ctx* c = act[gr[0]];
        GMGS_HIP(hipStreamWaitEvent(c->stream, ev[n], 0));
        if (mergeable[gr[0]]) {
          for (const LaunchRec::Multi& mu : steps[gi])
            if (mu.total > 0) {
              mu.go(mu, c->stream);
              ++bg.launches;
            }
        } else {
          for (int k = 0; k < 2; ++k) {
            if (c->mixed) pcg_iteration<float>(c, k);
            else pcg_iteration<double>(c, k);
          }
          bg.launches += (long long)recs[gr[0]].recs.size();
        }
        GMGS_HIP(hipEventRecord(ev[gi % n], c->stream));
        GMGS_HIP(hipStreamWaitEvent(s0, ev[gi % n], 0));
      }
      launch_gather_flags(bg.dptrs.as<const int*>(), m, bg.dflags.as<int>(), s0);
      GMGS_HIP(hipStreamEndCapture(s0, &g));
      const hipError_t e = hipGraphInstantiate(&bg.ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        bg.ge = nullptr;
        GMGS_HIP(e);
      }
      bg.groups = (int)groups.size();
      return batch_cache()[act] = std::move(bg);
    };
    BatchKey act(cs, cs + n);
    BatchGraph* bg = &graph_for(act);
    GMGS_HIP(hipStreamSynchronize(s0));
    HipEvent t0, t1;
    GMGS_HIP(hipEventRecord(t0, s0));
    int k = 0;
    int chunk = 2;
    while (k < maxk) {
      for (int i = 0; i < chunk && k < maxk; i += 2, k += 2) GMGS_HIP(hipGraphLaunch(bg->ge, s0));
      GMGS_HIP(hipMemcpyAsync(pinned.get(), bg->dflags.as<int>(), sizeof(int) * act.size(), hipMemcpyDeviceToHost,
                              s0));
      GMGS_HIP(hipEventRecord(ev[n + 1], s0));
      GMGS_HIP(hipEventSynchronize(ev[n + 1]));
      BatchKey left;
      for (size_t i = 0; i < act.size(); ++i)
        if (!pinned[i]) left.push_back(act[i]);
      if (left.empty()) break;
      // once half of the graph's problems have converged, continue on a graph
      // of the others (their kernels would only return at once)
      if (2 * left.size() <= act.size()) {
        act = left;
        bg = &graph_for(act);
      }
      chunk = std::min(chunk * 2, 4);
    }
    for (int i = 0; i < n; ++i) finish_deferred_x(cs[i], s0);
    GMGS_HIP(hipEventRecord(t1, s0));
    GMGS_HIP(hipEventSynchronize(t1));
    float ms = 0;
    GMGS_HIP(hipEventElapsedTime(&ms, t0, t1));
    int status = GMGS_OK;
    for (int i = 0; i < n; ++i) {
      gmgs_ctx* c = cs[i];
      PcgScalars h{};
      GMGS_HIP(hipMemcpy(&h, c->sc.as<void>(), sizeof(h), hipMemcpyDeviceToHost));
      LoopResult r;
      r.its = h.its;
      r.converged = h.converged;
      r.relres = h.relres;
      r.seconds = ms * 1e-3;
      const int st = finish_solve(c, r, infos ? &infos[i] : nullptr, nullptr, 0);
      if (st != GMGS_OK && status == GMGS_OK) status = st;
      if (U && U[i]) copy_solution(c, U[i]);
    }
    return status;
  });
}

int gmgs_stress(gmgs_ctx* c, const double* U, double* S) {
  if (!c || c->lev.empty() || !S) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    Level& l = L1(c);
    const double* u = c->x.as<double>();
    if (U) {
      DevBuf dU;
      upload(dU, U, (size_t)(3 * l.mesh.nn), s);
      launch_fill<double>(c->tmp.as<double>(), 0.0, 3 * l.mesh.g.N, s);
      launch_scatter_compact(l.mesh.node_pidx.as<long long>(), l.mesh.nn, dU.as<double>(),
                             c->tmp.as<double>(), l.mesh.g.N, s);
      u = c->tmp.as<double>();
      GMGS_HIP(hipStreamSynchronize(s));
    } else if (!c->have_solution) {
      throw Status(GMGS_ERR_STATE, "No Deformation Field Available, Please Compute Deformation First");
    }
    DevBuf dS;
    dS.alloc(sizeof(double) * 6 * l.mesh.nn);
    launch_stress(l.mesh.g, l.mesh.node_pidx.as<long long>(), l.mesh.nn, l.mesh.flags.as<uint32_t>(),
                  c->stressM.as<double>(), c->simp ? c->emod64.as<double>() : nullptr, u,
                  dS.as<double>(), s);
    GMGS_HIP(hipMemcpyAsync(S, dS.as<void>(), sizeof(double) * 6 * l.mesh.nn, hipMemcpyDeviceToHost, s));
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

int gmgs_adjoint(gmgs_ctx* c, const double* g, const double* U, double* lambda, double* dJdE,
                 gmgs_solve_info* info) {
  if (!c || c->lev.empty() || !g) return GMGS_ERR_ARGUMENT;
  if (!c->bc) return fail(c, GMGS_ERR_STATE, "boundary conditions not set");
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    Level& l = L1(c);
    const long long nn = l.mesh.nn, N = l.mesh.g.N;
    for (long long i = 0; i < 3 * nn; ++i)
      if (!std::isfinite(g[i])) throw Status(GMGS_ERR_NONFINITE, "non-finite objective gradient g");
    // forward displacement (dense) into tmp2
    if (U) {
      DevBuf dU;
      upload(dU, U, (size_t)(3 * nn), s);
      launch_fill<double>(c->tmp2.as<double>(), 0.0, 3 * N, s);
      launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, dU.as<double>(),
                             c->tmp2.as<double>(), N, s);
    } else {
      if (!c->have_solution) throw Status(GMGS_ERR_STATE, "no forward solution");
      GMGS_HIP(hipMemcpyAsync(c->tmp2.as<double>(), c->x.as<double>(), sizeof(double) * 3 * N,
                              hipMemcpyDeviceToDevice, s));
    }
    // adjoint solve K lambda = g (fixed DOFs of g ignored): reuse the MG-PCG with b := g
    DevBuf dg, saved;
    saved.alloc(sizeof(double) * 3 * N);
    GMGS_HIP(hipMemcpyAsync(saved.as<double>(), c->b.as<double>(), sizeof(double) * 3 * N,
                            hipMemcpyDeviceToDevice, s));
    // zero g on fixed DOFs so the adjoint system is the free-DOF one
    std::vector<double> gz(g, g + 3 * nn);
    {
      std::vector<uint32_t> fl = download<uint32_t>(l.mesh.flags, (size_t)N, s);
      std::vector<long long> np = download<long long>(l.mesh.node_pidx, (size_t)nn, s);
      for (long long k = 0; k < nn; ++k)
        for (int cc = 0; cc < 3; ++cc)
          if ((fl[np[k]] >> (F_FIX_SHIFT + cc)) & 1u) gz[3 * k + cc] = 0.0;
    }
    upload(dg, gz.data(), gz.size(), s);
    // b holds g until this scope ends, on every exit path: a thrown solve
    // error must not leave later solves working on K u = g
    struct RestoreLoads {
      gmgs_ctx* c;
      const DevBuf& saved;
      long long n;
      ~RestoreLoads() {
        (void)hipMemcpyAsync(c->b.as<double>(), saved.as<double>(), sizeof(double) * n,
                             hipMemcpyDeviceToDevice, c->stream);
        (void)hipStreamSynchronize(c->stream);
      }
    };
    c->have_solution = false;
    int st;
    {
      RestoreLoads restore{c, saved, 3 * N};
      launch_fill<double>(c->b.as<double>(), 0.0, 3 * N, s);
      launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, dg.as<double>(), c->b.as<double>(), N, s);
      st = solve_into(c, info, nullptr, 0);
    }
    // the adjoint solution is in x (not a solution of K u = F)
    c->have_solution = false;
    if (lambda) {
      DevBuf comp;
      comp.alloc(sizeof(double) * 3 * nn);
      launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), nn, c->x.as<double>(), N,
                                    comp.as<double>(), s);
      GMGS_HIP(hipMemcpyAsync(lambda, comp.as<void>(), sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
    }
    if (dJdE) {
      DevBuf dd;
      dd.alloc(sizeof(double) * l.mesh.ne);
      launch_element_energy(l.mesh.g, l.mesh.elem_pidx.as<long long>(), l.mesh.ne, L1(c).Ke64.as<double>(),
                            c->x.as<double>(), c->tmp2.as<double>(), dd.as<double>(), s);
      GMGS_HIP(hipMemcpyAsync(dJdE, dd.as<void>(), sizeof(double) * l.mesh.ne, hipMemcpyDeviceToHost, s));
    }
    GMGS_HIP(hipStreamSynchronize(s));
    return st;
  });
}

// ---------------------------------------------------------------- operator handles
int gmgs_atx(gmgs_ctx* c, int level, const double* x, double* y) {
  if (!c || c->lev.empty() || !x || !y) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    check_level(c, level);
    if (!c->bc) throw Status(GMGS_ERR_STATE, "boundary conditions not set (fixed DOFs unknown)");
    hipStream_t s = c->stream;
    const int li = level - 1;
    Level& l = c->lev[li];
    const long long nn = l.mesh.nn, N = l.mesh.g.N;
    DevBuf dx, din, dout, comp;
    upload(dx, x, (size_t)(3 * nn), s);
    din.alloc_zero(sizeof(double) * 3 * N, s);
    dout.alloc_zero(sizeof(double) * 3 * N, s);
    launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, dx.as<double>(), din.as<double>(), N, s);
    apply_level<double, double>(c, li, AP_AX, din.as<double>(), 0.0, dout.as<double>(), nullptr,
                                nullptr, nullptr);
    comp.alloc(sizeof(double) * 3 * nn);
    launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), nn, dout.as<double>(), N,
                                  comp.as<double>(), s);
    GMGS_HIP(hipMemcpyAsync(y, comp.as<void>(), sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

int gmgs_smooth(gmgs_ctx* c, int level, const double* r, double* dx, double* r_out) {
  if (!c || c->lev.empty() || !r) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    check_level(c, level);
    if (level == c->L) throw Status(GMGS_ERR_ARGUMENT, "no smoother on the coarsest level");
    if (!c->bc) throw Status(GMGS_ERR_STATE, "boundary conditions not set");
    hipStream_t s = c->stream;
    const int li = level - 1;
    Level& l = c->lev[li];
    const long long nn = l.mesh.nn, N = l.mesh.g.N;
    DevBuf dr, din, dout, dacc, comp;
    upload(dr, r, (size_t)(3 * nn), s);
    din.alloc_zero(sizeof(double) * 3 * N, s);
    dout.alloc_zero(sizeof(double) * 3 * N, s);
    dacc.alloc_zero(sizeof(double) * 3 * N, s);
    launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, dr.as<double>(), din.as<double>(), N, s);
    apply_level<double, double>(c, li, AP_SMOOTH, din.as<double>(), c->prm.jacobi_weight,
                                dout.as<double>(), dacc.as<double>(), nullptr, nullptr);
    comp.alloc(sizeof(double) * 3 * nn);
    if (dx) {
      launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), nn, dacc.as<double>(), N,
                                    comp.as<double>(), s);
      GMGS_HIP(hipMemcpyAsync(dx, comp.as<void>(), sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
      GMGS_HIP(hipStreamSynchronize(s));
    }
    if (r_out) {
      launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), nn, dout.as<double>(), N,
                                    comp.as<double>(), s);
      GMGS_HIP(hipMemcpyAsync(r_out, comp.as<void>(), sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
    }
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

int gmgs_restrict(gmgs_ctx* c, int level, const double* rf, double* rc) {
  if (!c || c->lev.empty() || !rf || !rc) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    check_level(c, level);
    if (level < 2) throw Status(GMGS_ERR_ARGUMENT, "restriction targets levels >= 2");
    hipStream_t s = c->stream;
    Level& f = c->lev[level - 2];
    Level& l = c->lev[level - 1];
    DevBuf dr, din, dout, comp;
    upload(dr, rf, (size_t)(3 * f.mesh.nn), s);
    din.alloc_zero(sizeof(double) * 3 * f.mesh.g.N, s);
    dout.alloc_zero(sizeof(double) * 3 * l.mesh.g.N, s);
    launch_scatter_compact(f.mesh.node_pidx.as<long long>(), f.mesh.nn, dr.as<double>(), din.as<double>(),
                           f.mesh.g.N, s);
    launch_restrict<double>(l.mesh.g, l.mesh.flags.as<uint32_t>(), l.tiles.as<TileRec>(), l.ntiles,
                            f.mesh.g, din.as<double>(), dout.as<double>(), nullptr, s);
    comp.alloc(sizeof(double) * 3 * l.mesh.nn);
    launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), l.mesh.nn, dout.as<double>(),
                                  l.mesh.g.N, comp.as<double>(), s);
    GMGS_HIP(hipMemcpyAsync(rc, comp.as<void>(), sizeof(double) * 3 * l.mesh.nn, hipMemcpyDeviceToHost, s));
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

int gmgs_interpolate(gmgs_ctx* c, int level, const double* xc, double* xf) {
  if (!c || c->lev.empty() || !xc || !xf) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    check_level(c, level);
    if (level < 2) throw Status(GMGS_ERR_ARGUMENT, "interpolation starts at levels >= 2");
    hipStream_t s = c->stream;
    Level& f = c->lev[level - 2];
    Level& l = c->lev[level - 1];
    DevBuf dx, din, dout, zero, comp;
    upload(dx, xc, (size_t)(3 * l.mesh.nn), s);
    din.alloc_zero(sizeof(double) * 3 * l.mesh.g.N, s);
    dout.alloc_zero(sizeof(double) * 3 * f.mesh.g.N, s);
    zero.alloc_zero(sizeof(double) * 3 * f.mesh.g.N, s);
    launch_scatter_compact(l.mesh.node_pidx.as<long long>(), l.mesh.nn, dx.as<double>(), din.as<double>(),
                           l.mesh.g.N, s);
    launch_prolong_smooth<double, double, double>(
        f.mesh.g, f.mesh.flags.as<uint32_t>(), f.tiles.as<TileRec>(), f.ntiles, l.mesh.g,
        din.as<double>(), zero.as<double>(), zero.as<double>(), zero.as<double>(), 0.0,
        dout.as<double>(), false, (const double*)nullptr, nullptr, nullptr, s);
    comp.alloc(sizeof(double) * 3 * f.mesh.nn);
    launch_gather_compact<double>(f.mesh.node_pidx.as<long long>(), f.mesh.nn, dout.as<double>(),
                                  f.mesh.g.N, comp.as<double>(), s);
    GMGS_HIP(hipMemcpyAsync(xf, comp.as<void>(), sizeof(double) * 3 * f.mesh.nn, hipMemcpyDeviceToHost, s));
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

int gmgs_vcycle(gmgs_ctx* c, const double* r, double* z) {
  if (!c || c->lev.empty() || !r || !z) return GMGS_ERR_ARGUMENT;
  if (!c->bc) return fail(c, GMGS_ERR_STATE, "boundary conditions not set");
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    Level& l = L1(c);
    const long long nn = l.mesh.nn, N = l.mesh.g.N;
    DevBuf dr, comp;
    upload(dr, r, (size_t)(3 * nn), s);
    launch_fill<double>(c->tmp.as<double>(), 0.0, 3 * N, s);
    launch_scatter_compact(l.mesh.node_pidx.as<long long>(), nn, dr.as<double>(), c->tmp.as<double>(), N, s);
    comp.alloc(sizeof(double) * 3 * nn);
    if (c->mixed) {
      run_vcycle<float>(c, c->tmp.as<double>(), c->z[0].as<float>(), nullptr, nullptr, nullptr);
      launch_gather_compact<float>(l.mesh.node_pidx.as<long long>(), nn, c->z[0].as<float>(), N,
                                   comp.as<double>(), s);
    } else {
      run_vcycle<double>(c, c->tmp.as<double>(), c->z[0].as<double>(), nullptr, nullptr, nullptr);
      launch_gather_compact<double>(l.mesh.node_pidx.as<long long>(), nn, c->z[0].as<double>(), N,
                                    comp.as<double>(), s);
    }
    GMGS_HIP(hipMemcpyAsync(z, comp.as<void>(), sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
    GMGS_HIP(hipStreamSynchronize(s));
    return GMGS_OK;
  });
}

// ---------------------------------------------------------------- queries
int gmgs_query_size(gmgs_ctx* c, int key, int level, int64_t* count) {
  if (!c || c->lev.empty() || !count) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    if (key != GMGS_Q_NUM_LEVELS && key != GMGS_Q_KE0 && key != GMGS_Q_BBOX && key != GMGS_Q_MAX_IT)
      check_level(c, level);
    const Level* l = (level >= 1 && level <= c->L) ? &c->lev[level - 1] : nullptr;
    const Grid* g = l ? &l->mesh.g : nullptr;
    int64_t n = 0;
    switch (key) {
      case GMGS_Q_NUM_LEVELS: case GMGS_Q_NUM_ELEMENTS: case GMGS_Q_NUM_NODES:
      case GMGS_Q_INTERMEDIATE_NUM_NODES: case GMGS_Q_STORING_STATE: case GMGS_Q_NUM_TILES:
      case GMGS_Q_NUM_FREE_COARSEST: case GMGS_Q_MAX_IT:
        n = 1; break;
      case GMGS_Q_TILE_CONFIG: n = 5; break;
      case GMGS_Q_RES: n = 3; break;
      case GMGS_Q_ELE_MAP_BACK: n = l->mesh.ne; break;
      case GMGS_Q_ELE_MAP_FORWARD: n = (int64_t)g->ey * g->ex * g->ez; break;
      case GMGS_Q_NOD_MAP_BACK: case GMGS_Q_NUM_NOD2ELES: n = l->mesh.nn; break;
      case GMGS_Q_NOD_MAP_FORWARD: n = (int64_t)g->ny * g->nx * g->nz; break;
      case GMGS_Q_ENODMAT: case GMGS_Q_ELEMENT_UPWARD_MAP: n = 8 * l->mesh.ne; break;
      case GMGS_Q_NODES_ON_BOUNDARY: n = l->mesh.n_nob; break;
      case GMGS_Q_ELEMENTS_ON_BOUNDARY: n = l->mesh.n_eob; break;
      case GMGS_Q_TRANSFER_MAT: n = 27 * l->mesh.ne; break;
      case GMGS_Q_TRANSFER_MAT_COEFFI: n = l->mesh.intermediateNumNodes; break;
      case GMGS_Q_SOLID_NODE_MAP_C2F: n = level >= 2 ? c->lev[level - 2].mesh.nn : 0; break;
      case GMGS_Q_UNIQUE_ELEMENTS: n = l->nuniq; break;
      case GMGS_Q_KS: n = 576 * l->mesh.ne; break;
      case GMGS_Q_DIAGK: case GMGS_Q_F: n = 3 * l->mesh.nn; break;
      case GMGS_Q_FIXED_DOFS: case GMGS_Q_FREE_DOFS: {
        if (!c->bc) throw Status(GMGS_ERR_STATE, "boundary conditions not set");
        if (key == GMGS_Q_FIXED_DOFS && level == 1) {   // fixingCond_ order, duplicates kept
          n = 3 * (int64_t)c->bc_fixed.size();
          break;
        }
        std::vector<uint32_t> fl = download<uint32_t>(l->mesh.flags, (size_t)g->N, c->stream);
        std::vector<long long> np = download<long long>(l->mesh.node_pidx, (size_t)l->mesh.nn, c->stream);
        int64_t nf = 0;
        for (long long k = 0; k < l->mesh.nn; ++k) nf += __builtin_popcount((fl[np[k]] >> F_FIX_SHIFT) & 7u);
        n = key == GMGS_Q_FIXED_DOFS ? nf : 3 * l->mesh.nn - nf;
        break;
      }
      case GMGS_Q_KE0: n = 576; break;
      case GMGS_Q_BBOX: n = 6; break;
      case GMGS_Q_SMOOTHER:
        if (level >= c->L) throw Status(GMGS_ERR_ARGUMENT, "the coarsest level is not smoothed");
        if (!c->bc) throw Status(GMGS_ERR_STATE, "boundary conditions not set");
        n = 1 + kMaxSweeps;
        break;
      default: throw Status(GMGS_ERR_ARGUMENT, "unknown query key");
    }
    *count = n;
    return GMGS_OK;
  });
}

int gmgs_query(gmgs_ctx* c, int key, int level, void* out, int64_t count) {
  if (!c || c->lev.empty() || !out) return GMGS_ERR_ARGUMENT;
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    int64_t need = 0;
    int st = gmgs_query_size(c, key, level, &need);
    if (st != GMGS_OK) return st;
    if (count < need) throw Status(GMGS_ERR_ARGUMENT, "output buffer too small");
    hipStream_t s = c->stream;
    const Level* l = (level >= 1 && level <= c->L) ? &c->lev[level - 1] : nullptr;
    auto i64 = static_cast<int64_t*>(out);
    auto i32 = static_cast<int32_t*>(out);
    auto f64 = static_cast<double*>(out);
    auto copy_dev = [&](const DevBuf& b, size_t bytes) {
      if (b.empty()) throw Status(GMGS_ERR_STATE, "table not built (GMGS_FLAG_NO_INDEX_TABLES?)");
      if (bytes) GMGS_HIP(hipMemcpyAsync(out, b.as<void>(), bytes, hipMemcpyDeviceToHost, s));
      GMGS_HIP(hipStreamSynchronize(s));
    };
    switch (key) {
      case GMGS_Q_NUM_LEVELS: i64[0] = c->L; break;
      case GMGS_Q_RES: i64[0] = l->mesh.g.ex; i64[1] = l->mesh.g.ey; i64[2] = l->mesh.g.ez; break;
      case GMGS_Q_NUM_ELEMENTS: i64[0] = l->mesh.ne; break;
      case GMGS_Q_NUM_NODES: i64[0] = l->mesh.nn; break;
      case GMGS_Q_INTERMEDIATE_NUM_NODES: i64[0] = l->mesh.intermediateNumNodes; break;
      case GMGS_Q_STORING_STATE: i64[0] = level == 1 ? 0 : l->state; break;
      case GMGS_Q_NUM_TILES: i64[0] = l->ntiles; break;
      case GMGS_Q_NUM_FREE_COARSEST: i64[0] = c->nfree; break;
      case GMGS_Q_MAX_IT: i64[0] = c->prm.max_it; break;
      case GMGS_Q_TILE_CONFIG:
        // (4: level 1 z-marching, its V(1, 1) residual kernel forming e = P x_c)
        i64[0] = l->split ? 3 : (l->zm ? (level == 1 && v11_fused(c) ? 4 : 1) : (l->as ? 2 : 0));
        i64[1] = l->col_tx;
        i64[2] = l->col_zc;
        // the FP64 matvec of the PCG runs on column tiles of its own when tuned
        i64[3] = l->coltiles64.empty() ? l->col_tx : l->col_tx64;
        i64[4] = l->coltiles64.empty() ? l->col_zc : l->col_zc64;
        break;
      case GMGS_Q_ELE_MAP_BACK: copy_dev(l->mesh.eleMapBack, 4 * need); break;
      case GMGS_Q_ELE_MAP_FORWARD: copy_dev(l->mesh.eleMapForward, 4 * need); break;
      case GMGS_Q_NOD_MAP_BACK: copy_dev(l->mesh.nodMapBack, 4 * need); break;
      case GMGS_Q_NOD_MAP_FORWARD: copy_dev(l->mesh.nodMapForward, 4 * need); break;
      case GMGS_Q_ENODMAT: copy_dev(l->mesh.eNodMat, 4 * need); break;
      case GMGS_Q_NUM_NOD2ELES: copy_dev(l->mesh.numNod2Eles, 4 * need); break;
      case GMGS_Q_NODES_ON_BOUNDARY: copy_dev(l->mesh.nodesOnBoundary, 4 * need); break;
      case GMGS_Q_ELEMENTS_ON_BOUNDARY: copy_dev(l->mesh.elementsOnBoundary, 4 * need); break;
      case GMGS_Q_ELEMENT_UPWARD_MAP: copy_dev(l->mesh.elementUpwardMap, 4 * need); break;
      case GMGS_Q_TRANSFER_MAT: copy_dev(l->mesh.transferMat, 4 * need); break;
      case GMGS_Q_TRANSFER_MAT_COEFFI: copy_dev(l->mesh.transferMatCoeffi, 8 * need); break;
      case GMGS_Q_SOLID_NODE_MAP_C2F: copy_dev(l->mesh.solidNodeMapCoarser2Finer, 4 * need); break;
      case GMGS_Q_UNIQUE_ELEMENTS: if (need) copy_dev(l->uniq, 4 * need); break;
      case GMGS_Q_KS: {
        const long long ne = l->mesh.ne;
        if (level == 1) {
          for (long long k = 0; k < ne; ++k) std::memcpy(f64 + 576 * k, c->Ke0, sizeof(double) * 576);
          break;
        }
        std::vector<long long> ep = download<long long>(l->mesh.elem_pidx, (size_t)ne, s);
        std::vector<int> eop = download<int>(l->eop, (size_t)l->mesh.g.N, s);
        std::vector<double> Ks = download<double>(l->Ks64, l->Ks64.bytes() / sizeof(double), s);
        std::vector<double> Ke = download<double>(l->Ke64, 576, s);
        for (long long k = 0; k < ne; ++k) {
          const int e = eop[ep[k]];
          const double* src = e == 0 ? Ke.data() : Ks.data() + (size_t)(e - 1) * 576;
          std::memcpy(f64 + 576 * k, src, sizeof(double) * 576);
        }
        break;
      }
      case GMGS_Q_DIAGK: {
        if (l->invd64.empty()) throw Status(GMGS_ERR_ARGUMENT, "no diagonal on the coarsest level");
        std::vector<double> iv = download<double>(l->invd64, (size_t)(3 * l->mesh.g.N), s);
        std::vector<long long> np = download<long long>(l->mesh.node_pidx, (size_t)l->mesh.nn, s);
        for (long long k = 0; k < l->mesh.nn; ++k)
          for (int cc = 0; cc < 3; ++cc) f64[3 * k + cc] = 1.0 / iv[cc * l->mesh.g.N + np[k]];
        break;
      }
      case GMGS_Q_FIXED_DOFS: case GMGS_Q_FREE_DOFS: {
        if (key == GMGS_Q_FIXED_DOFS && level == 1) {
          // fixedDOFs = reshape([3f-2 3f-1 3f]', [], 1) over fixingCond_
          // (ApplyBoundaryCondition.m:15-17): interleaved, in input order
          for (size_t k = 0; k < c->bc_fixed.size(); ++k)
            for (int cc = 0; cc < 3; ++cc) i32[3 * k + cc] = 3 * c->bc_fixed[k] - 2 + cc;
          break;
        }
        std::vector<uint32_t> fl = download<uint32_t>(l->mesh.flags, (size_t)l->mesh.g.N, s);
        std::vector<long long> np = download<long long>(l->mesh.node_pidx, (size_t)l->mesh.nn, s);
        int64_t j = 0;
        for (long long k = 0; k < l->mesh.nn; ++k)
          for (int cc = 0; cc < 3; ++cc) {
            const bool fx = (fl[np[k]] >> (F_FIX_SHIFT + cc)) & 1u;
            if (fx == (key == GMGS_Q_FIXED_DOFS)) i32[j++] = (int32_t)(3 * k + cc + 1);
          }
        break;
      }
      case GMGS_Q_KE0: std::memcpy(f64, c->Ke0, sizeof(double) * 576); break;
      case GMGS_Q_BBOX: std::memcpy(f64, c->bbox, sizeof(double) * 6); break;
      case GMGS_Q_SMOOTHER:
        f64[0] = l->lmax;
        for (int k = 0; k < kMaxSweeps; ++k) f64[1 + k] = l->om[k];
        break;
      case GMGS_Q_F:
        if (!c->bc || level != 1) throw Status(GMGS_ERR_STATE, "F_ exists on level 1 after gmgs_set_bc");
        std::memcpy(f64, c->F.data(), sizeof(double) * c->F.size());
        break;
      default: throw Status(GMGS_ERR_ARGUMENT, "unknown query key");
    }
    return GMGS_OK;
  });
}

int gmgs_copy_probe(int device, int64_t nbytes, int reps, double* avg_ms) {
  if (nbytes < 16 || reps < 1 || !avg_ms) return GMGS_ERR_ARGUMENT;
  return guarded(nullptr, [&]() -> int {
    GMGS_HIP(hipSetDevice(device));
    HipStream s;
    DevBuf a, b;
    a.alloc_zero((size_t)nbytes, s);
    b.alloc_zero((size_t)nbytes, s);
    const long long nd = nbytes / 8;
    launch_copy_probe(a.as<double>(), b.as<double>(), nd, s);
    HipEvent e0, e1;
    GMGS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) launch_copy_probe(a.as<double>(), b.as<double>(), nd, s);
    GMGS_HIP(hipEventRecord(e1, s));
    GMGS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GMGS_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    return GMGS_OK;
  });
}

int gmgs_kernel_timing(gmgs_ctx* c, int enable) {
  if (!c || c->lev.empty()) return GMGS_ERR_ARGUMENT;
  c->ktime = enable != 0;
  c->kpairs.clear();
  c->kpool_used = 0;
  for (int i = 0; i < 8; ++i) { c->ktot[i] = 0; c->kcnt[i] = 0; }
  return GMGS_OK;
}

int gmgs_kernel_timing_read(gmgs_ctx* c, int kernel, double* avg_ms, int64_t* count) {
  if (!c || kernel < 0 || kernel > GMGS_K_MATVEC64_P || !avg_ms || !count) return GMGS_ERR_ARGUMENT;
  *count = c->kcnt[kernel];
  *avg_ms = c->kcnt[kernel] ? c->ktot[kernel] / (double)c->kcnt[kernel] : 0.0;
  return GMGS_OK;
}

int gmgs_time_operator(gmgs_ctx* c, int kernel, int reps, double* avg_ms) {
  if (!c || c->lev.empty() || reps < 1 || !avg_ms) return GMGS_ERR_ARGUMENT;
  if (kernel != GMGS_K_MATVEC64 && kernel != GMGS_K_JACOBI64 && kernel != GMGS_K_MATVEC64_P)
    return GMGS_ERR_ARGUMENT;
  if (!c->bc) return fail(c, GMGS_ERR_STATE, "boundary conditions not set");
  if (kernel == GMGS_K_MATVEC64_P && !fused_direction(c))
    return fail(c, GMGS_ERR_STATE, "this solver does not fuse the direction update into its matvec");
  return guarded(c, [&]() -> int {
    GMGS_HIP(hipSetDevice(c->device));
    int nl = 0;
    auto launch = [&]() {
      if (kernel == GMGS_K_MATVEC64)
        apply_level<double, double>(c, 0, AP_AX, c->p.as<double>(), 0.0, c->q.as<double>(), nullptr,
                                    c->part.as<double>(), nullptr);
      else if (kernel == GMGS_K_MATVEC64_P)
        // alternating as in the iteration: with the x updates, then without
        launch_matvec_p(c, c->tmp.as<double>(), c->tmp2.as<double>(), c->ra.as<double>(),
                        c->z[0].as<float>(), c->part.as<double>(), nullptr, (nl++ & 1) == 0);
      else
        apply_level<double, double>(c, 0, AP_SMOOTH, c->r.as<double>(), c->prm.jacobi_weight,
                                    c->tmp.as<double>(), nullptr, nullptr, nullptr);
    };
    launch();
    launch();
    nl = 0;
    HipEvent e0, e1;
    GMGS_HIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < reps; ++i) launch();
    GMGS_HIP(hipEventRecord(e1, c->stream));
    GMGS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GMGS_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    return GMGS_OK;
  });
}

const char* gmgs_last_error(const gmgs_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

void gmgs_destroy(gmgs_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  invalidate_graph(c);
  c->lev.clear();
  for (hipEvent_t e : c->kpool) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->poll_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  hipStream_t s = c->stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
}

}  // extern "C"
