// Coarsest-level direct solve (kernel/SetupSolver.m:257-259, Vcycle.m:21-23):
// assembled FP64 K_ff, blocked Cholesky, explicit inverse, one GEMV per cycle.
#pragma once
#include "common.h"

namespace gmgs {

constexpr int DB = 64;   // dense tile

// free DOFs of the coarsest level in the reference's order (compact node,
// then component); returns count.  code[i] = c*N + pidx, dofmap over 3N.
long long coarsest_free_dofs(const Grid& g, const long long* node_pidx, long long nn,
                             const uint32_t* flags, DevBuf& code, DevBuf& dofmap, hipStream_t s);

// A (nfp x nfp row-major) = K_ff assembled row by row; identity on padding
void assemble_coarsest(const Grid& g, const uint32_t* flags, const int* eop, const double* Ke,
                       const double* Ks, const long long* code, const int* dofmap, long long nfree,
                       long long nfp, double* A, hipStream_t s);

// in place lower Cholesky; returns 0 or the 1-based first pivot that is not
// above 1e-12 x that row's diagonal before factoring (singular or indefinite)
int cholesky(double* A, long long nfp, hipStream_t s);

// Kinv = inv(A) from the factor L in A (uses W as workspace for L^-1)
void inverse_from_cholesky(const double* L, double* W, double* Kinv, long long nfp, hipStream_t s);

// x_L(free) = Kinv * r_L(free), x_L(fixed) = 0; bbuf holds nfp doubles,
// partial coarsest_partial_elems(nfp)
long long coarsest_partial_elems(long long nfp);
template <typename T>
void coarsest_apply(const double* Kinv, long long nfp, long long nfree, const long long* code,
                    const T* r, T* x, long long N, double* bbuf, double* partial, const int* done,
                    hipStream_t s);

}  // namespace gmgs
