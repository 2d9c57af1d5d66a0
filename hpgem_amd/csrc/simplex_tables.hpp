// Host-side tables for tetrahedral elements: hpGEM's hierarchical H1 DG basis
// on the reference tetrahedron, exact reference operators, and the face
// tables of the trace/lifting formulation used by the simplex kernels.
#pragma once

#include <array>
#include <vector>

namespace hgdg {

constexpr int kMaxSimplexOrder = 4;
constexpr int kFaceBasis = 15;  // orthonormal polynomials of degree <= 4 on the reference triangle

inline int tet_nb(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
inline int tri_nb(int p) { return (p + 1) * (p + 2) / 2; }

// hpGEM's reference tetrahedron (kernel/Geometry/ReferenceTetrahedron.cpp):
// nodes (0,0,0), (1,0,0), (0,1,0), (0,0,1); faces and edges as node lists.
extern const int kTetFaceNodes[4][3];
extern const int kTetEdgeNodes[6][2];

// Values and reference gradients of hpGEM's tetrahedral DG set of order p
// (createDGBasisFunctionSet3DH1Tetrahedron, BasisFunctions3DH1Conforming
// Tetrahedron.cpp:366-409) at one reference point: vertex functions lambda_i,
// edge functions lambda_a lambda_b L_j(lambda_a - lambda_b), face and interior
// functions with products of Lobatto kernels of barycentric differences.
void tet_basis(int p, const long double* x, long double* val, long double* grad /*[n][3]*/);

// Face map of local face f: reference triangle point -> reference tet point
// (MappingToRefTriangleToTetrahedron0..3).
void tet_face_map(int f, const long double* t, long double* x);

// Orientation maps of the reference triangle (MappingToRefTriangleToTriangle
// 0..5) and the orientation index picked from the node lists of the two sides
// (ReferenceTriangle::getCodim0MappingIndex).
void tri_orientation_map(int o, const long double* t, long double* out);
int tri_orientation_index(const long long* left, const long long* right);

struct SimplexOps {
    int p = 0, nb = 0;
    std::vector<double> Dr;    // [3][nb][nb]: reference derivative d/dxi_e, exact in P_p
    std::vector<double> Mref;  // [nb][nb]
    std::vector<double> Minv;  // [nb][nb]: M_ref^{-1}, inverted in long double
    // face tables per local face f and orientation o (trace points mapped
    // through orientation o before the face map):
    std::array<std::vector<int>, 4> cols;            // nonzero trace columns of face f
    std::array<std::array<std::vector<double>, 6>, 4> proj;  // [kFaceBasis][ncol]: face-basis coefficients of traces
    std::array<std::array<std::vector<double>, 6>, 4> lift;  // [nb][kFaceBasis]: M_ref^{-1} P^T
};

// Built in long double with exact collapsed Gauss rules.
SimplexOps simplex_ops(int p);

// Orientation maps acting on face-basis coefficients: for each orientation o
// the degree blocks (sizes 1, 2, .., 5; degree d at d(d+1)(2d+1)/6) of
// R_o(m, m') = int chi_m chi_m' o g_o, [6][kFaceRotBlock] row-major per block.
// A face polynomial h with coefficients c has h o g_o = sum_m (R_o c)_m chi_m.
constexpr int kFaceRotBlock = 55;
std::vector<double> face_orientation_blocks();

// hpGEM's Ainsworth-Coyle H(curl) DG set on the reference tetrahedron
// (createDGBasisFunctionSet3DAinsworthCoyle, BasisFunctions3DAinsworthCoyle
// .cpp:417-505) for orders 1 and 2: per degree k = 0..p the six edge
// functions on edges (0,1), (0,2), (0,3), (2,3), (1,3), (1,2), then for
// k >= 2 the twelve edge-based face functions; values and curls, [n][3] each.
int hcurl_tet_size(int p);
void hcurl_tet_basis(int p, const long double* x, long double* val, long double* curl);

// Collapsed Gauss rule on the reference tetrahedron exact for total degree k.
void tet_rule(int k, std::vector<long double>& pts /*[n][3]*/, std::vector<long double>& w);

}  // namespace hgdg
