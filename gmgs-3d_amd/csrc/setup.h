// Solver set-up on the device: active tiles, boundary-condition flags,
// Galerkin coarse element matrices, Jacobi diagonals.
#pragma once
#include "common.h"
#include "ops.h"

namespace gmgs {

// list of 16x4x1 tiles holding at least one solid node; returns count
long long build_tiles(const Grid& g, const uint32_t* flags, DevBuf& tiles, hipStream_t s);

// clear fixed bits (and F_FAST) of a level
void clear_fixed(const Grid& g, uint32_t* flags, hipStream_t s);
// level 1: fix all 3 components of the listed compact nodes (1-based ids)
void set_fixed_nodes(const long long* node_pidx, const int* ids, long long k, uint32_t* flags,
                     hipStream_t s);
// coarse fixed bits: OR over the 27 fine nodes around 2*c (SetupSolver.m:246-254)
void restrict_fixed(const Grid& gc, const TileRec* tiles_c, long long nt, uint32_t* flags_c,
                    const Grid& gf, const uint32_t* flags_f, hipStream_t s);

// mark F_FAST: all 8 incident cells present and universal (eop == 0, or level 1)
void mark_fast(const Grid& g, const TileRec* tiles, long long nt, uint32_t* flags, const int* eop,
               bool allow, hipStream_t s);

// child presence mask (elementUpwardMap column order) of each compact coarse element
void child_masks(const long long* elem_pidx_c, long long ne, const Grid& gc, const Grid& gf,
                 const uint8_t* esolid_f, int* mask, hipStream_t s);

// Galerkin products Ks = sum_j P_slot(j)' K_j P_slot(j) (SetupSolver.m:28-31,57-58)
// one job per coarse cell in 'cells' (padded index on gc); children read from
// the finer level: level-1 fine (eop_f == null) uses esolid_f and Ke_f (times
// emod_f when given); otherwise eop_f selects Ke_f / Ks_f / absent.
void galerkin_cells(const long long* cells, long long njobs, const Grid& gc, const Grid& gf,
                    const uint8_t* esolid_f, const int* eop_f, const double* Ke_f,
                    const double* Ks_f, const double* emod_f, int exact, double* Ks_out,
                    hipStream_t s);
// level-2 table: entry m (m = 1..255) for child mask m of level-1 universal children
void galerkin_mask_table(const double* Ke_f, int exact, double* table_out, hipStream_t s);

// B3 test (SetupSolver.m:96-101) for level 3: 1 if some present child is not
// universal (exact=1 also counts absent children)
void unique_flags_l3(const long long* elem_pidx_c, long long ne, const Grid& gc, const Grid& gf,
                     const int* eop_f, int exact, int* flag, hipStream_t s);

// eop for compact elements: value[k] written at cell elem_pidx[k]
void scatter_eop(const long long* elem_pidx, long long ne, const int* value, int* eop,
                 hipStream_t s);

// inverse Jacobi diagonal (SetupSolver.m:204-244), FP64, SoA 3N, 0 off solid nodes
// the inverse diagonal shared by every F_FAST node of a level (all eight
// elements present and universal), per component: true and out[] when every
// F_FAST node holds the same float bits, false otherwise (or with none)
bool fast_inverse_diagonal(const Grid& g, const TileRec* tiles, long long nt, const uint32_t* flags,
                           const float* invd, float out[3], hipStream_t s);

void inverse_diagonal(const Grid& g, const TileRec* tiles, long long nt, const uint32_t* flags,
                      const int* eop, const double* Ke, const double* Ks, const double* emod,
                      double* invd, hipStream_t s);

void set_galerkin_weights(const double W[8][8][8], hipStream_t s);

}  // namespace gmgs
