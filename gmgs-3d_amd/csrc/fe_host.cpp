// Host-side finite-element constants: Ke0, the nodal stress operator, the
// span-2 transfer weights and the interior 27-point stencil.
#include "fe_host.h"

#include <cmath>
#include <cstring>

namespace gmgs {

namespace {
const double G = 0.577350269189626;   // common/GaussianIntegral.m:11-16 (literal)
const double GS[8] = {-G, G, G, -G, -G, G, G, -G};
const double GT[8] = {-G, -G, G, G, -G, -G, G, G};
const double GP[8] = {-G, -G, -G, -G, G, G, G, G};
const double SGS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
const double SGT[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
const double SGP[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// common/DeShapeFunction.m:3-22 at Gauss point g: d[0..2][n] = dN_n/d(s,t,p)
void de_shape(int g, double d[3][8]) {
  for (int n = 0; n < 8; ++n) {
    d[0][n] = 0.125 * SGS[n] * (1 + SGT[n] * GT[g]) * (1 + SGP[n] * GP[g]);
    d[1][n] = 0.125 * SGT[n] * (1 + SGS[n] * GS[g]) * (1 + SGP[n] * GP[g]);
    d[2][n] = 0.125 * SGP[n] * (1 + SGS[n] * GS[g]) * (1 + SGT[n] * GT[g]);
  }
}

// common/ShapeFunction.m:15-23
double shape(int n, double s, double t, double p) {
  return 0.125 * (1 + SGS[n] * s) * (1 + SGT[n] * t) * (1 + SGP[n] * p);
}

void inv3(const double a[3][3], double r[3][3], double* det) {
  double d = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
           - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
           + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  *det = d;
  r[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / d;
  r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / d;
  r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / d;
  r[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / d;
  r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / d;
  r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / d;
  r[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / d;
  r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / d;
  r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / d;
}

// B (48x24) of common/ElementStrainMatrix.m with per-Gauss-point inverse
// Jacobians of the probe element; returns detJ per Gauss point.
void strain_matrix(const double probe[8][3], double B[48][24], double detJ[8]) {
  std::memset(B, 0, sizeof(double) * 48 * 24);
  for (int g = 0; g < 8; ++g) {
    double d[3][8];
    de_shape(g, d);
    double J[3][3], iJ[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int n = 0; n < 8; ++n) s += d[i][n] * probe[n][j];
        J[i][j] = s;
      }
    inv3(J, iJ, &detJ[g]);
    double der[3][8];
    for (int i = 0; i < 3; ++i)
      for (int n = 0; n < 8; ++n) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += iJ[i][k] * d[k][n];
        der[i][n] = s;
      }
    const int r = 6 * g;
    for (int n = 0; n < 8; ++n) {
      const int c = 3 * n;
      double dx = der[0][n], dy = der[1][n], dz = der[2][n];
      B[r + 0][c + 0] = dx;
      B[r + 1][c + 1] = dy;
      B[r + 2][c + 2] = dz;
      B[r + 3][c + 1] = dz; B[r + 3][c + 2] = dy;
      B[r + 4][c + 0] = dz; B[r + 4][c + 2] = dx;
      B[r + 5][c + 0] = dy; B[r + 5][c + 1] = dx;
    }
  }
}

// invert an n x n matrix (Gauss-Jordan, partial pivoting)
bool invert(int n, const double* a, double* out) {
  double m[16][32];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j)
      m[i][j] = j < n ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(m[i][c]) > std::fabs(m[piv][c])) piv = i;
    if (m[piv][c] == 0.0) return false;
    if (piv != c)
      for (int j = 0; j < 2 * n; ++j) std::swap(m[c][j], m[piv][j]);
    double inv = 1.0 / m[c][c];
    for (int j = 0; j < 2 * n; ++j) m[c][j] *= inv;
    for (int i = 0; i < n; ++i) {
      if (i == c) continue;
      double f = m[i][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * n; ++j) m[i][j] -= f * m[c][j];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[i * n + j] = m[i][n + j];
  return true;
}
}  // namespace

void hooke(double E, double nu, double S[6][6]) {
  std::memset(S, 0, sizeof(double) * 36);
  const double c1 = (1 + nu) * (1 - 2 * nu), c2 = 2 * (1 + nu);
  S[0][0] = S[1][1] = S[2][2] = E * (1 - nu) / c1;
  S[0][1] = S[0][2] = S[1][0] = S[1][2] = S[2][0] = S[2][1] = E * nu / c1;
  S[3][3] = S[4][4] = S[5][5] = E / c2;
}

void element_stiffness(double E, double nu, const double probe[8][3], double* Ke) {
  double B[48][24], detJ[8], S[6][6];
  strain_matrix(probe, B, detJ);
  hooke(E, nu, S);
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) Ke[i * 24 + j] = 0.0;
  for (int g = 0; g < 8; ++g) {
    const double w = 1.0 * detJ[g];
    double DB[6][24];
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 24; ++j) {
        double s = 0;
        for (int b = 0; b < 6; ++b) s += S[a][b] * w * B[6 * g + b][j];
        DB[a][j] = s;
      }
    for (int i = 0; i < 24; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0;
        for (int a = 0; a < 6; ++a) s += B[6 * g + a][i] * DB[a][j];
        Ke[i * 24 + j] += s;
      }
  }
  // exact symmetry (the triple product is symmetric in exact arithmetic)
  for (int i = 0; i < 24; ++i)
    for (int j = i + 1; j < 24; ++j) {
      double v = 0.5 * (Ke[i * 24 + j] + Ke[j * 24 + i]);
      Ke[i * 24 + j] = Ke[j * 24 + i] = v;
    }
}

bool stress_operator(double E, double nu, const double probe[8][3], double* M) {
  // M = OTP * D * B with OTP = inv(kron(N, I6)) = kron(inv(N), I6)
  // (kernel/GetCartesianStress.m:17-36, 44-73).
  double B[48][24], detJ[8], S[6][6];
  strain_matrix(probe, B, detJ);
  hooke(E, nu, S);
  double N[64], Ni[64];
  for (int g = 0; g < 8; ++g)
    for (int n = 0; n < 8; ++n) N[g * 8 + n] = shape(n, GS[g], GT[g], GP[g]);
  if (!invert(8, N, Ni)) return false;
  double DB[48][24];
  for (int g = 0; g < 8; ++g)
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 24; ++j) {
        double s = 0;
        for (int b = 0; b < 6; ++b) s += S[a][b] * B[6 * g + b][j];
        DB[6 * g + a][j] = s;
      }
  for (int n = 0; n < 8; ++n)
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 24; ++j) {
        double s = 0;
        for (int g = 0; g < 8; ++g) s += Ni[n * 8 + g] * DB[6 * g + a][j];
        M[(6 * n + a) * 24 + j] = s;
      }
  return true;
}

void transfer_weights(double inNODE[27][8]) {
  // Operator4MultiGridRestrictionAndInterpolation.m:5-31 from its definition:
  // fine node j = a + 3b + 9c (a = row offset 0..2); coarse corner k at
  // offsets 2*(hex_ro, hex_xo, hex_zo).
  const int ro[8] = {1, 1, 0, 0, 1, 1, 0, 0}, xo[8] = {0, 1, 1, 0, 0, 1, 1, 0},
            zo[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 8; ++k)
          inNODE[a + 3 * b + 9 * c][k] = (1 - std::abs(a - 2 * ro[k]) / 2.0)
                                        * (1 - std::abs(b - 2 * xo[k]) / 2.0)
                                        * (1 - std::abs(c - 2 * zo[k]) / 2.0);
}

void child_weights(double W[8][8][8]) {
  // W[s][i][k]: weight of coarse corner k at local node i of child slot s,
  // where slot s follows SubEleNodMat.m:4-11 (counter-clockwise children).
  double in[27][8];
  transfer_weights(in);
  const int cy[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cx[8] = {0, 0, 1, 1, 0, 0, 1, 1},
            cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  const int ro[8] = {1, 1, 0, 0, 1, 1, 0, 0}, xo[8] = {0, 1, 1, 0, 0, 1, 1, 0},
            zo[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  for (int s = 0; s < 8; ++s)
    for (int i = 0; i < 8; ++i) {
      int j = (cy[s] + ro[i]) + 3 * (cx[s] + xo[i]) + 9 * (cz[s] + zo[i]);
      for (int k = 0; k < 8; ++k) W[s][i][k] = in[j][k];
    }
}

void interior_stencil(const double* Ke, double S[27][3][3]) {
  // S[o][c][d]: coefficient of u_d at node n+o in row c of node n, summed
  // over the 8 elements around n.  o = (dy+1) + 3(dx+1) + 9(dz+1).
  std::memset(S, 0, sizeof(double) * 27 * 9);
  for (int p = 0; p < 8; ++p) {
    const int pyb = p & 1, pxb = (p >> 1) & 1, pzb = (p >> 2) & 1;
    const int ln = node_local_in(p);
    for (int m = 0; m < 8; ++m) {
      const int dy = pyb - 1 + hex_ro(m), dx = pxb - 1 + hex_xo(m), dz = pzb - 1 + hex_zo(m);
      const int o = (dy + 1) + 3 * (dx + 1) + 9 * (dz + 1);
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) S[o][c][d] += Ke[(3 * ln + c) * 24 + 3 * m + d];
    }
  }
}

bool stencil_classes(const double S[27][3][3], double cls[30], double tol) {
  auto at = [&](int dy, int dx, int dz, int c, int d) {
    return S[(dy + 1) + 3 * (dx + 1) + 9 * (dz + 1)][c][d];
  };
  for (int c = 0; c < 3; ++c)
    for (int q = 0; q < 8; ++q) cls[8 * c + q] = at(q & 1, (q >> 1) & 1, q >> 2, c, c);
  for (int q = 0; q < 2; ++q) {
    cls[24 + q] = at(1, 1, q, 0, 1);   // x-y, |dz| = q
    cls[26 + q] = at(q, 1, 1, 0, 2);   // x-z, |dy| = q
    cls[28 + q] = at(1, q, 1, 1, 2);   // y-z, |dx| = q
  }
  double mx = 0.0;
  for (int o = 0; o < 27; ++o)
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) mx = std::max(mx, std::fabs(S[o][c][d]));
  auto sg = [](int v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); };
  for (int dz = -1; dz <= 1; ++dz)
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy) {
        const int off[3] = {dx, dy, dz};   // axis of component c: x <-> dx, y <-> dy, z <-> dz
        for (int c = 0; c < 3; ++c)
          for (int d = 0; d < 3; ++d) {
            double model;
            if (c == d) {
              model = cls[8 * c + std::abs(dy) + 2 * std::abs(dx) + 4 * std::abs(dz)];
            } else {
              const int e = 3 - c - d;
              const int pair = (c + d == 1) ? 24 : (c + d == 2) ? 26 : 28;
              model = sg(off[c]) * sg(off[d]) * cls[pair + std::abs(off[e])];
            }
            if (std::fabs(model - at(dy, dx, dz, c, d)) > tol * mx) return false;
          }
      }
  return true;
}

bool corner_blocks(const double* K, double Cc[72], double tol) {
  double kmax = 0.0;
  for (int i = 0; i < 576; ++i) kmax = std::max(kmax, std::fabs(K[i]));
  for (int k = 0; k < 8; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Cc[(k * 3 + i) * 3 + j] = K[(3 * hex_local(0) + i) * 24 + 3 * hex_local(k) + j];
  auto sgn = [](int a, int c) {
    const int bit = c == 0 ? 1 : (c == 1 ? 0 : 2);
    return ((a >> bit) & 1) ? -1.0 : 1.0;
  };
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double ref = sgn(a, i) * sgn(a, j) * Cc[((a ^ b) * 3 + i) * 3 + j];
          if (std::fabs(K[(3 * hex_local(a) + i) * 24 + 3 * hex_local(b) + j] - ref) > tol * kmax) return false;
        }
  return true;
}

}  // namespace gmgs
