// Host-side finite-element constants (see fe_host.cpp).
#pragma once
#include "common.h"

namespace gmgs {

// Isotropic Hooke matrix, common/ElementElasticityMatrix.m:11-24.
void hooke(double E, double nu, double S[6][6]);

// Ke0 (24x24 row-major, DOF order [x1 y1 z1 ... z8]) by 2x2x2 Gauss
// quadrature on the probe element, common/SetupStiffnessMatrix.m:4-22.
void element_stiffness(double E, double nu, const double probe[8][3], double* Ke);

// Nodal stress operator M = OTP*D*B (48x24 row-major; row 6n+a = stress
// component a [xx yy zz yz zx xy] at local node n), GetCartesianStress.m:17-36.
bool stress_operator(double E, double nu, const double probe[8][3], double* M);

// Span-2 trilinear weights inNODE (27x8).
void transfer_weights(double inNODE[27][8]);

// Per child slot s (SubEleNodMat.m:4-11 order) the 8x8 weight block
// W[s][i][k] = inNODE(SubEleNodMat(s,i), k).
void child_weights(double W[8][8][8]);

// Interior 27-point stencil assembled from 8 copies of Ke.
void interior_stencil(const double* Ke, double S[27][3][3]);

// The 30 class coefficients of a brick stencil (reflection symmetry in each
// axis): cls[0..23] diagonal a_c[|dy| + 2|dx| + 4|dz|] for c = x, y, z;
// cls[24..25] G_xy[|dz|], cls[26..27] G_xz[|dy|], cls[28..29] G_yz[|dx|] with
// S_cd(o) = sign(o_c) sign(o_d) G_cd(|o_e|).  Returns false when some entry
// of S deviates from that model by more than tol * max|S|.
bool stencil_classes(const double S[27][3][3], double cls[30], double tol);

// Canonical row block of a box element matrix, Cc[k][i][j] = K[3 L(0) + i][3 L(k) + j]
// for corner codes k = ro + 2 xo + 4 zo (L = hex_local).  Reflecting the box
// maps every block onto one of these: K[3 L(a) + i][3 L(b) + j] =
// s_i(a) s_j(a) Cc[a ^ b][i][j], s_c(a) = -1 when corner a lies on the far side
// of the axis carrying component c (x: bit 1, y: bit 0, z: bit 2).  Returns
// false when K deviates from that by more than tol * max|K|.
bool corner_blocks(const double* K, double Cc[72], double tol);

}  // namespace gmgs
