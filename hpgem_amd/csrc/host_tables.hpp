// Host-side tables of the library: Gauss-Legendre rules, the 1D hierarchical
// functions behind hpGEM's square/cube DG bases, the map between hpGEM's basis
// order and the library's tensor order, the 1D reference operators the tensor
// kernels apply, and the structured mesh generator.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hgdg {

enum Shape { SHAPE_SQUARE = 2, SHAPE_CUBE = 4, SHAPE_TETRAHEDRON = 5 };

int shape_dim(int shape);
int shape_vertices(int shape);

// n-point Gauss-Legendre rule on [-1, 1] (nodes ascending).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);

// 1D functions psi_0 = (1-t)/2, psi_1 = (1+t)/2 and the bubbles
// psi_{k+2} = (1-t^2)/4 * L_k(t), where L_k is hpGEM's Lobatto kernel
// (kernel/FE/helperFunctions.cpp:49-108). The bubbles are the normalised
// integrated Legendre polynomials, evaluated here as
// (P_{k+2} - P_k) / sqrt(2 (2k + 3)) with derivative sqrt((2k+3)/2) P_{k+1}.
double psi(int a, double t);
double dpsi(int a, double t);

// For the square (dim 2) and cube (dim 3) DG sets of order p: perm[i] is the
// tensor index a + n1*b + n1^2*c (n1 = p+1) of hpGEM basis function i, whose
// value is psi_a(x) psi_b(y) psi_c(z). The sequence follows the construction
// loops of createDGBasisFunctionSet2DH1Square (BasisFunctions2DH1Conforming
// Square.cpp:164-184) and createDGBasisFunctionSet3DH1Cube
// (BasisFunctions3DH1ConformingCube.cpp:434-528); every mirroring sign in
// those sets is +1, so the map carries no signs.
std::vector<int> tensor_permutation(int dim, int p);

// 1D reference operators for order p (n1 = p+1), row-major n1 x n1:
//   M1(a,b) = int psi_a psi_b,  G1(a,b) = int psi_b psi_a' (derivative on the
//   test index a),  K(a,b) = int psi_a psi_b',
//   D = M1^{-1} G1 (weak derivative), Dr = M1^{-1} K (exact derivative of the
//   interpolant: D = -Dr + Mi(:,1) e_1^T - Mi(:,0) e_0^T), Mi = M1^{-1}.
// Integrals use an exact Gauss rule; everything is formed in long double.
struct TensorOps {
    int n1 = 0;
    std::vector<double> M1, G1, D, Dr, Mi;
};
TensorOps tensor_ops(int p);

// Tables of the Gauss-point form of an order-p tensor element, used where the
// element map is not affine or a neighbour's face is not aligned. The element
// rule hpGEM selects for order 2p+1 is the n1 = p+1 point Gauss rule per
// direction, so B(q,a) = psi_a(x_q) is square and invertible and the element
// mass matrix B^T diag(w |det J|) B (B = B1 (x) B1 (x) B1) has the exact
// inverse B^{-1} diag(1 / (w |det J|)) B^{-T}. All n1 x n1 tables row-major:
//   x, w     the n1 Gauss points (ascending) and weights;
//   B, Binv  psi_a(x_q) [q][a] and its inverse [a][q];
//   C        C(q,m) = l_m'(x_q), l_m the Lagrange polynomials of the points:
//            the derivative of the interpolant of point values (C = G B^{-1},
//            G(q,a) = psi_a'(x_q));
//   lam0/1   l_q(-1), l_q(+1): B^{-T} e_0, B^{-T} e_1 (a trace of the
//            hierarchical set at x = -1 / +1 is the coefficient with index 0 / 1);
//   x2, w2   the n1+1 point rule (error norms, rule order 2p+2), B2 = psi_a at
//            its points [a][q].
// Formed in long double.
struct NodalOps {
    int n1 = 0;
    std::vector<double> x, w, B, Binv, C, lam0, lam1, x2, w2, B2;
    // all of the above in this order, as one table for the device
    std::vector<double> packed() const;
};
NodalOps nodal_ops(int p);

// Structured mesh in hpGEM Preprocessor order
// (applications/Preprocessor/lib/hpgem.cpp:100-235).
struct StructuredMesh {
    int shape = 0, dim = 0, nv = 0;
    std::vector<double> coords;
    std::vector<int64_t> nodes;
};
StructuredMesh structured_mesh(int dim, const int32_t* n, const double* lo, const double* hi,
                               const int32_t* periodic, bool simplices);

}  // namespace hgdg
