// z-marching level operator (see ops_zm.hip).
#pragma once
#include "common.h"
#include "ops.h"

namespace gmgs {

struct ColTile {
  int y0, x0, z0, pad;   // padded coordinates; y0 % 16 == 0, x0 % TX == 0, z0 >= 1
};

// column tiles (16 x tx x zc nodes) holding at least one solid node
long long build_coltiles(const Grid& g, const uint32_t* flags, int tx, int zc, DevBuf& tiles,
                         hipStream_t s);

// same contract as launch_apply (one partial per block -> ntiles partials)
template <typename T, typename TR>
void launch_apply_zm(const LevelOp<T>& op, const ColTile* tiles, long long ntiles, int tx, int zc,
                     int mode, const TR* in, T omega, T* out, TR* xacc, double* partial,
                     const int* done, hipStream_t s);

}  // namespace gmgs
