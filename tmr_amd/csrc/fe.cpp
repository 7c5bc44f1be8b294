#include "fe.h"

#include <cmath>
#include <cstring>

namespace tmrx {

void isotropic_C(double E, double nu, double C[36]) {
  std::memset(C, 0, 36 * sizeof(double));
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[6 * i + j] = lam + (i == j ? 2.0 * mu : 0.0);
  for (int i = 3; i < 6; i++) C[6 * i + i] = mu;
}

void hex8_box_B(double hx, double hy, double hz, const double pt[3], double B[144],
                double N[8]) {
  std::memset(B, 0, 144 * sizeof(double));
  const double h[3] = {hx, hy, hz};
  for (int a = 0; a < 8; a++) {
    const int ijk[3] = {a & 1, (a >> 1) & 1, a >> 2};
    double f[3], df[3];
    for (int d = 0; d < 3; d++) {
      double s = ijk[d] ? 1.0 : -1.0;
      f[d] = 0.5 * (1.0 + s * pt[d]);
      df[d] = 0.5 * s * (2.0 / h[d]);  // d/dx = (2/h) d/dxi
    }
    double dNdx = df[0] * f[1] * f[2];
    double dNdy = f[0] * df[1] * f[2];
    double dNdz = f[0] * f[1] * df[2];
    if (N) N[a] = f[0] * f[1] * f[2];
    double *b0 = &B[0 * 24 + 3 * a], *b1 = &B[1 * 24 + 3 * a], *b2 = &B[2 * 24 + 3 * a];
    double *b3 = &B[3 * 24 + 3 * a], *b4 = &B[4 * 24 + 3 * a], *b5 = &B[5 * 24 + 3 * a];
    b0[0] = dNdx;
    b1[1] = dNdy;
    b2[2] = dNdz;
    b3[1] = dNdz;  // gyz = dv/dz + dw/dy
    b3[2] = dNdy;
    b4[0] = dNdz;  // gxz = du/dz + dw/dx
    b4[2] = dNdx;
    b5[0] = dNdy;  // gxy = du/dy + dv/dx
    b5[1] = dNdx;
  }
}

void hex8_box_stiffness_gp(double hx, double hy, double hz, const double C[36], int q,
                           double K[576]) {
  std::memset(K, 0, 576 * sizeof(double));
  const double g = 1.0 / std::sqrt(3.0);
  const double detJ = hx * hy * hz / 8.0;
  double B[144], CB[144];
  const double pt[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
  hex8_box_B(hx, hy, hz, pt, B, nullptr);
  for (int i = 0; i < 6; i++)
    for (int k = 0; k < 24; k++) {
      double s = 0.0;
      for (int j = 0; j < 6; j++) s += C[6 * i + j] * B[24 * j + k];
      CB[24 * i + k] = s;
    }
  for (int r = 0; r < 24; r++)
    for (int c = 0; c < 24; c++) {
      double s = 0.0;
      for (int i = 0; i < 6; i++) s += B[24 * i + r] * CB[24 * i + c];
      K[24 * r + c] = detJ * s;
    }
}

void hex8_box_stiffness(double hx, double hy, double hz, const double C[36], double K[576]) {
  std::memset(K, 0, 576 * sizeof(double));
  double Kq[576];
  for (int q = 0; q < 8; q++) {
    hex8_box_stiffness_gp(hx, hy, hz, C, q, Kq);
    for (int i = 0; i < 576; i++) K[i] += Kq[i];
  }
}

void hex8_box_grad_table(double hx, double hy, double hz, double G[192]) {
  const double g = 1.0 / std::sqrt(3.0);
  double B[144];
  for (int q = 0; q < 8; q++) {
    const double pt[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
    hex8_box_B(hx, hy, hz, pt, B, nullptr);
    for (int a = 0; a < 8; a++)
      for (int d = 0; d < 3; d++) G[24 * q + 3 * a + d] = B[24 * d + 3 * a + d];
  }
}

}  // namespace tmrx
