// Host-side hex8 element quantities for axis-aligned box elements.
#pragma once

namespace tmrx {

// Isotropic linear elasticity, 6x6 Voigt matrix with engineering shear
// strains, ordering [exx, eyy, ezz, gyz, gxz, gxy] (TACS 3D convention used
// by TMROctConstitutive, src/topology/TMROctConstitutive.cpp:476-493).
void isotropic_C(double E, double nu, double C[36]);

// 24x24 stiffness of a trilinear hex with edge lengths hx, hy, hz, 2x2x2
// Gauss quadrature (exact for a box); dof 3*a + d, local node
// a = i + 2j + 4k (x fastest, src/TMROctForest.cpp:4709-4711).
void hex8_box_stiffness(double hx, double hy, double hz, const double C[36], double K[576]);

// Contribution of Gauss point q (bit d of q set: coordinate d at +1/sqrt(3))
// to that stiffness: the 8 of them sum to hex8_box_stiffness.
void hex8_box_stiffness_gp(double hx, double hy, double hz, const double C[36], int q,
                           double K[576]);

// Shape-function gradients at the 8 Gauss points of a box element:
// G[24 q + 3 a + d] = dN_a / dx_d at Gauss point q.
void hex8_box_grad_table(double hx, double hy, double hz, double G[192]);

// Strain-displacement matrix (6x24) at the parametric point (xi,eta,zeta) in
// [-1,1]^3 of a box element, and the shape functions N[8].
void hex8_box_B(double hx, double hy, double hz, const double pt[3], double B[144],
                double N[8]);

}  // namespace tmrx
