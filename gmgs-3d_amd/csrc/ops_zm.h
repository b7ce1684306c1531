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

// per padded column (y + py x) the z range [lo, hi] of its solid nodes packed
// as lo | hi << 16; an empty column holds lo = 0xffff, hi = 0
void build_colrange(const Grid& g, const uint32_t* flags, DevBuf& colz, hipStream_t s);

// same contract as launch_apply without the x += u accumulation, plus the
// RESID mode (out = in2 - A in; with partials, of |out|^2) and, on level 1
// (no stored elements), the
// SMOOTH_SUM mode (co: its outputs; partials r0.z at [0, n) and r0.zprev at
// [n, 2 n)); returns the number n of dot partials written (one per tile).
// fast_only: only the F_FAST nodes (universal stencil, no element
// corrections) are computed and written; launch_apply_as_rows does the rest
template <typename T, typename TR>
int launch_apply_zm(const LevelOp<T>& op, const ColTile* tiles, long long ntiles, int tx, int zc,
                    int mode, const TR* in, T omega, const T* in2, T* out, double* partial,
                    const int* done, hipStream_t s, const CycleOut<T>* co = nullptr,
                    bool fast_only = false);

}  // namespace gmgs
