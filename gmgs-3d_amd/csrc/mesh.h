// Mesh hierarchy on the device: dense per-level grids plus the reference's
// own index arrays (MATLAB layouts, int32, 1-based), rebuilt bit-exactly.
#pragma once
#include "common.h"

namespace gmgs {

struct LevelMesh {
  Grid g{};
  int level = 1;                 // 1 = finest
  long long ne = 0, nn = 0;      // numElements, numNodes
  DevBuf esolid;                 // uint8 per padded cell (cell stored at its min-corner pidx)
  DevBuf flags;                  // uint32 per padded node (element mask, fixed bits, fast bit)
  // reference index arrays (int32 unless noted)
  DevBuf eleMapBack;             // ne, 1-based full cell ids (MATLAB order)
  DevBuf eleMapForward;          // ey*ex*ez, compact 1-based or 0
  DevBuf nodMapBack;             // nn, level 1: full node ids; level>=2: finest-grid ids
  DevBuf nodFull;                // nn, this level's own full node ids (1-based)
  DevBuf nodMapForward;          // ny*nx*nz, compact 1-based or 0
  DevBuf eNodMat;                // ne x 8 column-major, compact 1-based
  DevBuf numNod2Eles;            // nn
  DevBuf nodesOnBoundary;        // 1-based compact
  DevBuf elementsOnBoundary;     // 1-based compact
  long long n_nob = 0, n_eob = 0;
  DevBuf node_pidx;              // int64 nn: compact node -> padded index
  DevBuf elem_pidx;              // int64 ne: compact element -> padded cell index
  // level >= 2
  DevBuf elementUpwardMap;       // ne x 8 column-major, finer compact ids, 0 = void
  DevBuf transferMat;            // 27 x ne column-major, 1-based intermediate ids
  DevBuf transferMatCoeffi;      // double, intermediateNumNodes
  DevBuf solidNodeMapCoarser2Finer;  // finer nn, 1-based intermediate ids
  long long intermediateNumNodes = 0;
};

// exclusive scan of int32 flags; returns the total
long long exclusive_scan(const int* in, int* out, long long n, hipStream_t s);

// Level 1 from the (possibly smaller) voxel array vox (nely x nelx x nelz,
// column-major, 1 = solid); the level grid is already padded.
void build_level1(LevelMesh& L, const int* d_vox, int nely, int nelx, int nelz,
                  bool index_tables, hipStream_t s);

// Level l >= 2 from its finer level.
void build_coarse(LevelMesh& C, const LevelMesh& F, const Grid& finest, bool index_tables,
                  hipStream_t s);

}  // namespace gmgs
