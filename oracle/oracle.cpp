// ============================================================================
// ORACLE — test infrastructure, not product code.
//
// A plain FP64 CPU restatement of hpGEM's explicit DG operator, written to act
// as the checker for the GPU path. Only the test suite, the smoke check and the
// CPU-baseline leg of the benchmark load it. Nothing in the shipped library
// links, loads or calls it.
//
// It follows the reference's own algorithm, function by function:
//   * Gauss rules (kernel/Integration/QuadratureRules/*): tensor Gauss-Legendre
//     for line/square/cube (the reference's literal tables agree with these to
//     <=1e-15); simplex rules are supplied as data (see orc_set_simplex_rule),
//     pinned by the closed-form monomial moments of the reference simplex.
//   * H1 hierarchical DG bases in the reference's own ordering
//     (kernel/FE/BasisFunctions2DH1ConformingSquare.cpp:164,
//      BasisFunctions3DH1ConformingCube.cpp:434,
//      BasisFunctions3DH1ConformingTetrahedron.cpp:366,
//      Lobatto kernels kernel/FE/helperFunctions.cpp:49-108).
//   * Reference shapes, face maps, orientation maps
//     (kernel/Geometry/Reference{Square,Cube,Tetrahedron}.cpp,
//      Mappings/MappingToRef{LineToSquare,SquareToCube,TriangleToTetrahedron,
//      SquareToSquare,TriangleToTriangle,LineToLine}.cpp).
//   * Multilinear / affine element maps and Jacobians
//     (Mappings/MappingToPhysHypercubeLinear.cpp:163-304,
//      MappingToPhysSimplexLinear_Impl.h:51-100).
//   * Face normal = wedge(J_elem * J_face) * sign(det J) * outward sign
//     (kernel/Geometry/FaceGeometry.h:412-459, SmallMatrix_impl.h:198-220).
//   * Face search and orientation index (kernel/Base/MeshManipulator_Impl.h:
//     1860-1950, Face.cpp:70-108, Reference*::getCodim0MappingIndex).
//   * Element / face integration  sum_q w_q |det J| f  and  sum_q w_q |n| f
//     (kernel/Integration/ElementIntegral_Impl.h:96-151, FaceIntegral_Impl.h:
//     67-127).
//   * HpgemAPILinear: precomputed element and face matrices applied by dense
//     mat-vec, then a per-element LU solve with the (re-factorised) mass matrix
//     on every call (kernel/Base/HpgemAPILinear_Impl.h:131-156,250-288,
//     360-434; kernel/LinearAlgebra/MiddleSizeMatrix.cpp:723-745).
//   * Upwind advection integrands (tests/self/Basic/080AdvectionLinear_
//     SelfTest.cpp:108-170) and the acoustic wave integrands
//     (applications/ExampleMultipleVariableProblem/AcousticWave_Impl.h:
//     169-378).
//   * The RK stage loop of HpgemAPISimplified::computeOneTimeStep
//     (kernel/Base/HpgemAPISimplified_Impl.h:599-620,788-820) with the RK4
//     tableau of kernel/Base/TimeIntegration/RK4Methods.cpp:45-63.
//   * L2 projection of the initial condition and the L2 error
//     (HpgemAPISimplified_Impl.h:164-401).
//   * The structured mesh generator of applications/Preprocessor/lib/
//     hpgem.cpp:100-235 (node order, periodic coordinate copies, 5-tet split
//     with alternating rotation).
// ============================================================================
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
// Working precision: double (the checker of the FP64 path, rounding like the
// reference) or long double (the same algorithm in extended precision, used
// to measure how far the FP64 evaluations sit from the exact operator).
typedef ORACLE_REAL Real;
const Real kPi = Real(3.141592653589793238462643383279502884L);

enum Shape { LINE = 1, SQUARE = 2, TRIANGLE = 3, CUBE = 4, TETRAHEDRON = 5 };

int shape_dim(int s) {
    switch (s) {
        case LINE: return 1;
        case SQUARE: case TRIANGLE: return 2;
        case CUBE: case TETRAHEDRON: return 3;
    }
    throw std::runtime_error("unknown shape");
}
bool is_simplex(int s) { return s == TRIANGLE || s == TETRAHEDRON; }
int shape_nodes(int s) {
    switch (s) {
        case LINE: return 2;
        case SQUARE: return 4;
        case TRIANGLE: return 3;
        case CUBE: return 8;
        case TETRAHEDRON: return 4;
    }
    throw std::runtime_error("unknown shape");
}
int face_shape(int s) {
    switch (s) {
        case SQUARE: case TRIANGLE: return LINE;
        case CUBE: return SQUARE;
        case TETRAHEDRON: return TRIANGLE;
    }
    throw std::runtime_error("shape has no faces in this oracle");
}

// ------------------------------------------------------------------ rules
struct Rule {
    int dim = 0;
    int order = 0;
    std::vector<Real> pts;  // [n][dim]
    std::vector<Real> w;
    int n() const { return int(w.size()); }
};

// n-point Gauss-Legendre on [-1,1] by Newton iteration in long double.
void gauss_legendre(int n, std::vector<Real>& x, std::vector<Real>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; ++i) {
        long double t = std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1, p1 = t;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = t; p0 = 1; }
            dp = n * (t * p1 - p0) / (t * t - 1);
            long double dt = p1 / dp;
            t -= dt;
            if (std::fabs((Real)dt) < 1e-22) break;
        }
        long double p0 = 1, p1 = t;
        for (int k = 2; k <= n; ++k) {
            long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (t * p1 - p0) / (t * t - 1);
        x[n - 1 - i] = Real(t);
        w[n - 1 - i] = Real(2.0L / ((1 - t * t) * dp * dp));
    }
    if (n == 1) { x[0] = 0.0; w[0] = 2.0; }
}

// Simplex rules are data: the checker loads them from the reference-built
// fixture table (kernel/Integration/QuadratureRules/GaussQuadratureRulesFor
// {Triangle,Tetrahedron}.cpp, selected as in AllGaussQuadratureRules.cpp:139).
std::map<std::pair<int, int>, Rule> g_simplex_rules;  // (shape, order) -> rule

// AllGaussQuadratureRules::getRule: the lowest-order registered rule with
// order >= requested. The reference registers tensor rules of order 1,3,..,11
// for line/square/cube.
Rule get_rule(int shape, int order) {
    if (!is_simplex(shape)) {
        int n = std::max(1, (order + 2) / 2);
        if (n > 6) throw std::runtime_error("no tensor rule of that order");
        std::vector<Real> x, w;
        gauss_legendre(n, x, w);
        Rule r;
        r.dim = shape_dim(shape);
        r.order = 2 * n - 1;
        int np = 1;
        for (int d = 0; d < r.dim; ++d) np *= n;
        for (int q = 0; q < np; ++q) {
            int rem = q;
            Real wt = 1;
            for (int d = 0; d < r.dim; ++d) {
                int i = rem % n;
                rem /= n;
                r.pts.push_back(x[i]);
                wt *= w[i];
            }
            r.w.push_back(wt);
        }
        return r;
    }
    const Rule* best = nullptr;
    for (auto& kv : g_simplex_rules) {
        if (kv.first.first != shape || kv.second.order < order) continue;
        if (!best || kv.second.order < best->order) best = &kv.second;
    }
    if (!best) throw std::runtime_error("no simplex rule loaded for that order");
    return *best;
}

// ------------------------------------------------------------ Lobatto
// kernel/FE/helperFunctions.cpp:49-108 (values to degree 4, derivatives to 3).
Real lobatto(int k, Real x) {
    switch (k) {
        case 0: return std::sqrt(Real(3) / 2) * -2.;
        case 1: return std::sqrt(Real(5) / 2) * x * -2.;
        case 2: return std::sqrt(Real(7) / 2) * (5 * x * x - 1) / -2.;
        case 3: return std::sqrt(Real(9) / 2) * (7 * x * x - 3) * x / -2.;
        case 4: return std::sqrt(Real(11) / 2) * ((21 * x * x - 14) * x * x + 1) / -4.;
    }
    throw std::runtime_error("Lobatto degree out of range");
}
Real lobatto_d(int k, Real x) {
    switch (k) {
        case 0: return 0;
        case 1: return -2 * std::sqrt(Real(5) / 2);
        case 2: return -2 * std::sqrt(Real(7) / 2) * (2.5 * x);
        case 3: return -2 * std::sqrt(Real(9) / 2) * (5.25 * x * x - .75);
    }
    throw std::runtime_error("Lobatto derivative degree out of range");
}

// ------------------------------------------------------------ bases
// A tensor-shape basis function is c * prod_d f_d(x_d) with each factor
// either linear (1 + s t) or a bubble (1 - t^2) L_k(m t). Every square and
// cube function of the reference has this form; the constructor arguments
// below follow the reference classes one for one.
struct Factor {
    int bubble = 0;  // 0: (1 + s t), 1: (1 - t^2) L_k(m t)
    int s = 0;       // linear slope sign
    int k = 0;       // Lobatto degree
    int m = 1;       // mirroring
    Real val(Real t) const {
        return bubble ? (1 - t) * (1 + t) * lobatto(k, m * t) : 1 + s * t;
    }
    Real der(Real t) const {
        return bubble ? -2 * t * lobatto(k, m * t) +
                            (1 - t) * (1 + t) * lobatto_d(k, m * t) * m
                      : Real(s);
    }
};
struct TensorFn {
    Real scale = 1;
    std::array<Factor, 3> f;
};
// A simplex basis function: product of barycentrics and Lobatto kernels of
// barycentric differences.
struct SimplexFn {
    std::vector<int> lam;                        // barycentric indices multiplied
    std::vector<std::array<int, 3>> kern;        // (degree, node a, node b): L_deg(la - lb)
};

struct Basis {
    int shape = 0;
    int p = 0;
    std::vector<TensorFn> tensor;
    std::vector<SimplexFn> simplex;
    int size() const { return int(is_simplex(shape) ? simplex.size() : tensor.size()); }
};

Factor lin(int s) { Factor f; f.bubble = 0; f.s = s; return f; }
Factor bub(int k, int m) { Factor f; f.bubble = 1; f.k = k; f.m = m; return f; }

// reference node coordinate sign of node n along direction d (x fastest)
int node_sign(int n, int d) { return ((n >> d) & 1) * 2 - 1; }

// kernel/Geometry/ReferenceCube.cpp:52-58
const int CUBE_FACE_NODES[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                   {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};
const int CUBE_EDGE_NODES[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                    {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
// kernel/Geometry/ReferenceSquare.cpp (faces 0..3)
const int SQUARE_FACE_NODES[4][2] = {{0, 1}, {0, 2}, {1, 3}, {2, 3}};
// kernel/Geometry/ReferenceTetrahedron.cpp:56-60
const int TET_FACE_NODES[4][3] = {{0, 3, 2}, {0, 1, 3}, {0, 2, 1}, {1, 2, 3}};
const int TET_EDGE_NODES[6][2] = {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {1, 3}, {1, 2}};
// kernel/Geometry/ReferenceTriangle.cpp (faces)
const int TRI_FACE_NODES[3][2] = {{0, 1}, {0, 2}, {1, 2}};

// BasisFunction2DFaceSquare_0/1 (BasisFunctions2DH1ConformingSquare.cpp:67-120)
TensorFn square_edge(int n0, int n1, int k) {
    TensorFn fn;
    fn.scale = 1. / 8.;
    int m = n0 > n1 ? -1 : 1;
    int e = (n0 + n1 < 3) ? -1 : 1;
    if ((n0 + n1) % 2 == 1) {  // horizontal edge: bubble in x
        fn.f = {bub(k, m), lin(e), lin(0)};
    } else {
        fn.f = {lin(e), bub(k, m), lin(0)};
    }
    return fn;
}

Basis make_square(int p) {
    // createDGBasisFunctionSet2DH1Square (BasisFunctions2DH1ConformingSquare.cpp:164-184)
    Basis b;
    b.shape = SQUARE;
    b.p = p;
    if (p < 1) throw std::runtime_error("p>=1 required");
    for (int i = 0; i < 4; ++i) {
        TensorFn fn;
        fn.scale = 1. / 4.;
        fn.f = {lin(node_sign(i, 0)), lin(node_sign(i, 1)), lin(0)};
        b.tensor.push_back(fn);
    }
    for (int i = 0; i + 2 <= p; ++i) {
        b.tensor.push_back(square_edge(0, 1, i));
        b.tensor.push_back(square_edge(2, 3, i));
        b.tensor.push_back(square_edge(0, 2, i));
        b.tensor.push_back(square_edge(1, 3, i));
        for (int j = 0; j + 2 <= p; ++j) {
            TensorFn fn;
            fn.scale = 1. / 16.;
            fn.f = {bub(i, 1), bub(j, 1), lin(0)};
            b.tensor.push_back(fn);
        }
    }
    return b;
}

// BasisFunction3DEdgeCube_{0,1,2} (BasisFunctions3DH1ConformingCube.cpp:83-238)
TensorFn cube_edge(int dir, int n0, int n1, int k) {
    TensorFn fn;
    fn.scale = 1. / 16.;
    int m = n0 < n1 ? 1 : -1;
    for (int d = 0; d < 3; ++d)
        fn.f[d] = (d == dir) ? bub(k, m) : lin(node_sign(n0, d));
    return fn;
}
// BasisFunction3DFaceCube_{0,1,2} (:240-398): dir = the constant coordinate
TensorFn cube_face(int dir, int n0, int n1, int n2, int k_lo, int k_hi) {
    TensorFn fn;
    fn.scale = 1. / 32.;
    int m1 = n0 < n1 ? 1 : -1;
    int m2 = n0 < n2 ? 1 : -1;
    int d1 = dir == 0 ? 1 : 0;
    int d2 = dir == 2 ? 1 : 2;
    fn.f[dir] = lin(node_sign(n0, dir));
    fn.f[d1] = bub(k_lo, m1);
    fn.f[d2] = bub(k_hi, m2);
    return fn;
}
TensorFn cube_interior(int i, int j, int k) {
    TensorFn fn;
    fn.scale = 1. / 64.;
    fn.f = {bub(i, 1), bub(j, 1), bub(k, 1)};
    return fn;
}

Basis make_cube(int p) {
    // createDGBasisFunctionSet3DH1Cube (BasisFunctions3DH1ConformingCube.cpp:434-528)
    Basis b;
    b.shape = CUBE;
    b.p = p;
    if (p < 1) throw std::runtime_error("p>=1 required");
    for (int i = 0; i < 8; ++i) {
        TensorFn fn;
        fn.scale = 1. / 8.;
        fn.f = {lin(node_sign(i, 0)), lin(node_sign(i, 1)), lin(node_sign(i, 2))};
        b.tensor.push_back(fn);
    }
    auto faces = [&](int i, int j) {
        b.tensor.push_back(cube_face(2, 0, 1, 2, i, j));
        b.tensor.push_back(cube_face(1, 0, 1, 4, i, j));
        b.tensor.push_back(cube_face(0, 0, 2, 4, i, j));
        b.tensor.push_back(cube_face(0, 1, 3, 5, i, j));
        b.tensor.push_back(cube_face(1, 2, 3, 6, i, j));
        b.tensor.push_back(cube_face(2, 4, 5, 6, i, j));
    };
    for (int j = 0; j + 2 <= p; ++j) {
        for (int i = 0; i < 4; ++i)
            b.tensor.push_back(cube_edge(0, CUBE_EDGE_NODES[i][0], CUBE_EDGE_NODES[i][1], j));
        for (int i = 4; i < 8; ++i)
            b.tensor.push_back(cube_edge(1, CUBE_EDGE_NODES[i][0], CUBE_EDGE_NODES[i][1], j));
        for (int i = 8; i < 12; ++i)
            b.tensor.push_back(cube_edge(2, CUBE_EDGE_NODES[i][0], CUBE_EDGE_NODES[i][1], j));
        faces(j, j);
        b.tensor.push_back(cube_interior(j, j, j));
        for (int i = 0; i < j; ++i) {
            faces(i, j);
            faces(j, i);
            b.tensor.push_back(cube_interior(i, i, j));
            b.tensor.push_back(cube_interior(i, j, i));
            b.tensor.push_back(cube_interior(j, i, i));
            b.tensor.push_back(cube_interior(j, j, i));
            b.tensor.push_back(cube_interior(j, i, j));
            b.tensor.push_back(cube_interior(i, j, j));
            for (int k = 0; k < i; ++k) {
                b.tensor.push_back(cube_interior(i, j, k));
                b.tensor.push_back(cube_interior(i, k, j));
                b.tensor.push_back(cube_interior(j, i, k));
                b.tensor.push_back(cube_interior(j, k, i));
                b.tensor.push_back(cube_interior(k, i, j));
                b.tensor.push_back(cube_interior(k, j, i));
            }
        }
    }
    return b;
}

Basis make_tetrahedron(int p) {
    // createDGBasisFunctionSet3DH1Tetrahedron
    // (BasisFunctions3DH1ConformingTetrahedron.cpp:366-409)
    Basis b;
    b.shape = TETRAHEDRON;
    b.p = p;
    if (p < 1) throw std::runtime_error("p>=1 required");
    for (int i = 0; i < 4; ++i) b.simplex.push_back(SimplexFn{{i}, {}});
    for (int j = 0; j + 2 <= p; ++j) {
        for (int i = 0; i < 6; ++i) {
            int a = TET_EDGE_NODES[i][0], c = TET_EDGE_NODES[i][1];
            b.simplex.push_back(SimplexFn{{a, c}, {{j, a, c}}});
        }
        if (j > 0) {
            for (int i = 0; i < 4; ++i) {
                const int* n = TET_FACE_NODES[i];
                for (int k = 0; k + 1 <= j; ++k)
                    b.simplex.push_back(SimplexFn{{n[0], n[1], n[2]},
                                                  {{j - k - 1, n[0], n[1]}, {k, n[1], n[2]}}});
            }
        }
        if (j > 1) {
            for (int i = 0; i + 2 <= j; ++i)
                for (int k = 0; (i + k) + 2 <= j; ++k)
                    b.simplex.push_back(SimplexFn{{0, 1, 2, 3},
                                                  {{i, 0, 1}, {j - i - k - 2, 1, 2}, {k, 2, 3}}});
        }
    }
    return b;
}

Basis make_basis(int shape, int p) {
    switch (shape) {
        case SQUARE: return make_square(p);
        case CUBE: return make_cube(p);
        case TETRAHEDRON: return make_tetrahedron(p);
    }
    throw std::runtime_error("basis not available for this shape");
}

// barycentric coordinates and their gradients (helperFunctions.cpp:110-164)
void barycentric(int dim, const Real* x, Real* lam, Real (*dlam)[3]) {
    lam[0] = 1;
    for (int d = 0; d < dim; ++d) lam[0] -= x[d];
    for (int d = 0; d < 3; ++d) dlam[0][d] = d < dim ? -1 : 0;
    for (int n = 1; n <= dim; ++n) {
        lam[n] = x[n - 1];
        for (int d = 0; d < 3; ++d) dlam[n][d] = (d == n - 1) ? 1 : 0;
    }
}

void eval_basis(const Basis& b, const Real* x, Real* val, Real* grad) {
    int dim = shape_dim(b.shape);
    if (!is_simplex(b.shape)) {
        for (int i = 0; i < b.size(); ++i) {
            const TensorFn& fn = b.tensor[i];
            Real v[3], dv[3];
            for (int d = 0; d < 3; ++d) {
                v[d] = d < dim ? fn.f[d].val(x[d]) : 1.0;
                dv[d] = d < dim ? fn.f[d].der(x[d]) : 0.0;
            }
            val[i] = fn.scale * v[0] * v[1] * v[2];
            for (int d = 0; d < dim; ++d) {
                Real g = fn.scale * dv[d];
                for (int e = 0; e < dim; ++e)
                    if (e != d) g *= v[e];
                grad[i * dim + d] = g;
            }
        }
        return;
    }
    Real lam[4], dlam[4][3];
    barycentric(dim, x, lam, dlam);
    for (int i = 0; i < b.size(); ++i) {
        const SimplexFn& fn = b.simplex[i];
        // factors: lambdas then kernels; product rule
        std::vector<Real> fv, fd[3];
        for (int l : fn.lam) {
            fv.push_back(lam[l]);
            for (int d = 0; d < 3; ++d) fd[d].push_back(dlam[l][d]);
        }
        for (auto& kn : fn.kern) {
            Real s = lam[kn[1]] - lam[kn[2]];
            fv.push_back(lobatto(kn[0], s));
            Real ld = lobatto_d(kn[0], s);
            for (int d = 0; d < 3; ++d) fd[d].push_back(ld * (dlam[kn[1]][d] - dlam[kn[2]][d]));
        }
        Real v = 1;
        for (Real f : fv) v *= f;
        val[i] = v;
        for (int d = 0; d < dim; ++d) {
            Real g = 0;
            for (size_t a = 0; a < fv.size(); ++a) {
                Real t = fd[d][a];
                for (size_t c = 0; c < fv.size(); ++c)
                    if (c != a) t *= fv[c];
                g += t;
            }
            grad[i * dim + d] = g;
        }
    }
}

// ------------------------------------------------------------ reference maps
// Codim-1 maps face reference point -> element reference point and the outward
// normal sign (Mappings/MappingToRef{LineToSquare,SquareToCube,
// TriangleToTetrahedron}.cpp).
void face_map(int shape, int face, const Real* pf, Real* pe, Real* jac /*[dim][dim-1]*/,
              Real* sign) {
    switch (shape) {
        case SQUARE: {
            static const Real S[4] = {-1, 1, -1, 1};
            Real t = pf[0];
            Real J[2] = {0, 0};
            switch (face) {
                case 0: pe[0] = t; pe[1] = -1; J[0] = 1; break;
                case 1: pe[0] = -1; pe[1] = t; J[1] = 1; break;
                case 2: pe[0] = 1; pe[1] = t; J[1] = 1; break;
                case 3: pe[0] = t; pe[1] = 1; J[0] = 1; break;
            }
            if (jac) { jac[0] = J[0]; jac[1] = J[1]; }
            if (sign) *sign = S[face];
            return;
        }
        case CUBE: {
            static const Real S[6] = {-1, 1, -1, 1, -1, 1};
            Real a = pf[0], b = pf[1];
            Real J[3][2] = {{0, 0}, {0, 0}, {0, 0}};
            switch (face) {
                case 0: pe[0] = a; pe[1] = b; pe[2] = -1; J[0][0] = 1; J[1][1] = 1; break;
                case 1: pe[0] = a; pe[1] = -1; pe[2] = b; J[0][0] = 1; J[2][1] = 1; break;
                case 2: pe[0] = -1; pe[1] = a; pe[2] = b; J[1][0] = 1; J[2][1] = 1; break;
                case 3: pe[0] = 1; pe[1] = a; pe[2] = b; J[1][0] = 1; J[2][1] = 1; break;
                case 4: pe[0] = a; pe[1] = 1; pe[2] = b; J[0][0] = 1; J[2][1] = 1; break;
                case 5: pe[0] = a; pe[1] = b; pe[2] = 1; J[0][0] = 1; J[1][1] = 1; break;
            }
            if (jac)
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 2; ++c) jac[r * 2 + c] = J[r][c];
            if (sign) *sign = S[face];
            return;
        }
        case TETRAHEDRON: {
            // MappingToRefTriangleToTetrahedron0..3: the face's nodes, in the
            // order of TET_FACE_NODES, are the images of triangle nodes
            // (0,0),(1,0),(0,1); the map is affine.
            static const Real V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
            const int* n = TET_FACE_NODES[face];
            Real a = pf[0], b = pf[1];
            for (int d = 0; d < 3; ++d) {
                pe[d] = V[n[0]][d] + a * (V[n[1]][d] - V[n[0]][d]) + b * (V[n[2]][d] - V[n[0]][d]);
                if (jac) {
                    jac[d * 2 + 0] = V[n[1]][d] - V[n[0]][d];
                    jac[d * 2 + 1] = V[n[2]][d] - V[n[0]][d];
                }
            }
            if (sign) *sign = 1.0;
            return;
        }
    }
    throw std::runtime_error("face_map: unsupported shape");
}

// Codim-0 orientation maps of the face shape (MappingToRefLineToLine,
// MappingToRefSquareToSquare0..7, MappingToRefTriangleToTriangle0..5).
void orientation_map(int fshape, int idx, const Real* pin, Real* pout) {
    switch (fshape) {
        case LINE:
            pout[0] = idx == 0 ? pin[0] : -pin[0];
            return;
        case SQUARE: {
            Real x = pin[0], y = pin[1];
            switch (idx) {
                case 0: pout[0] = x; pout[1] = y; return;
                case 1: pout[0] = -y; pout[1] = x; return;
                case 2: pout[0] = -x; pout[1] = -y; return;
                case 3: pout[0] = y; pout[1] = -x; return;
                case 4: pout[0] = x; pout[1] = -y; return;
                case 5: pout[0] = -x; pout[1] = y; return;
                case 6: pout[0] = -y; pout[1] = -x; return;
                case 7: pout[0] = y; pout[1] = x; return;
            }
            break;
        }
        case TRIANGLE: {
            Real x = pin[0], y = pin[1];
            switch (idx) {
                case 0: pout[0] = x; pout[1] = y; return;
                case 1: pout[0] = y; pout[1] = x; return;
                case 2: pout[0] = 1 - x - y; pout[1] = x; return;
                case 3: pout[0] = 1 - x - y; pout[1] = y; return;
                case 4: pout[0] = x; pout[1] = 1 - x - y; return;
                case 5: pout[0] = y; pout[1] = 1 - x - y; return;
            }
            break;
        }
    }
    throw std::runtime_error("orientation_map: bad index");
}

// getCodim0MappingIndex of the face shape (ReferenceLine/Square/Triangle.cpp).
int orientation_index(int fshape, const std::vector<int64_t>& l1, const std::vector<int64_t>& l2) {
    switch (fshape) {
        case LINE:
            if (l1[0] == l2[0]) return 0;
            return 1;
        case SQUARE:
            if (l1[0] == l2[0]) {
                if (l1[1] == l2[1]) return 0;
                if (l1[3] == l2[3]) return 7;
            } else if (l1[0] == l2[1]) {
                if (l1[1] == l2[0]) return 5;
                if (l1[1] == l2[3]) return 3;
            } else if (l1[0] == l2[2]) {
                if (l1[2] == l2[0]) return 4;
                if (l1[2] == l2[3]) return 1;
            } else {
                if (l1[1] == l2[1]) return 6;
                if (l1[1] == l2[2]) return 2;
            }
            throw std::runtime_error("impossible square node order");
        case TRIANGLE:
            if (l1[0] == l2[0]) return l1[1] == l2[1] ? 0 : 1;
            if (l1[0] == l2[1]) return l1[1] == l2[2] ? 2 : 3;
            return l1[1] == l2[1] ? 4 : 5;
    }
    throw std::runtime_error("orientation_index: bad shape");
}

const int* face_nodes(int shape, int f) {
    switch (shape) {
        case SQUARE: return SQUARE_FACE_NODES[f];
        case CUBE: return CUBE_FACE_NODES[f];
        case TETRAHEDRON: return TET_FACE_NODES[f];
        case TRIANGLE: return TRI_FACE_NODES[f];
    }
    throw std::runtime_error("face_nodes: bad shape");
}
int n_faces(int shape) {
    switch (shape) {
        case SQUARE: return 4;
        case CUBE: return 6;
        case TETRAHEDRON: return 4;
        case TRIANGLE: return 3;
    }
    throw std::runtime_error("n_faces: bad shape");
}

// ------------------------------------------------------------ geometry
// Reference -> physical map of one element and its Jacobian.
// Hypercubes: multilinear map through the vertices in reference node order
// (MappingToPhysHypercubeLinear.cpp). Simplices: affine map
// x = X0 + sum_d (X_{d+1} - X0) xi_d (MappingToPhysSimplexLinear_Impl.h).
void elem_map(int shape, const Real* X /*[nv][dim]*/, const Real* xi, Real* x, Real* J /*[dim][dim]*/) {
    int dim = shape_dim(shape);
    int nv = shape_nodes(shape);
    for (int r = 0; r < dim; ++r) {
        x[r] = 0;
        for (int c = 0; c < dim; ++c) J[r * dim + c] = 0;
    }
    if (is_simplex(shape)) {
        for (int r = 0; r < dim; ++r) {
            x[r] = X[r];
            for (int c = 0; c < dim; ++c) {
                Real e = X[(c + 1) * dim + r] - X[r];
                x[r] += e * xi[c];
                J[r * dim + c] = e;
            }
        }
        return;
    }
    for (int v = 0; v < nv; ++v) {
        Real N = 1;
        Real dN[3];
        for (int d = 0; d < dim; ++d) N *= (1 + node_sign(v, d) * xi[d]) / 2;
        for (int c = 0; c < dim; ++c) {
            Real g = node_sign(v, c) / 2.0;
            for (int d = 0; d < dim; ++d)
                if (d != c) g *= (1 + node_sign(v, d) * xi[d]) / 2;
            dN[c] = g;
        }
        for (int r = 0; r < dim; ++r) {
            x[r] += N * X[v * dim + r];
            for (int c = 0; c < dim; ++c) J[r * dim + c] += X[v * dim + r] * dN[c];
        }
    }
}

Real det(int dim, const Real* J) {
    switch (dim) {
        case 1: return J[0];
        case 2: return J[0] * J[3] - J[1] * J[2];
        case 3:
            return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                   J[2] * (J[3] * J[7] - J[4] * J[6]);
    }
    return 0;
}

// grad_phys = J^{-T} grad_ref by solving J^T g = gref (H1ConformingTransformation.h:64-69)
void solve_small(int n, Real* A /*row-major n*n*/, Real* b) {
    // Gaussian elimination with partial pivoting
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > std::fabs(A[piv * n + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
            std::swap(b[k], b[piv]);
        }
        for (int i = k + 1; i < n; ++i) {
            Real f = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        Real s = b[k];
        for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
        b[k] = s / A[k * n + k];
    }
}

// ------------------------------------------------------------ dense LU solve
// MiddleSizeMatrix::solve -> dgesv on a copy (MiddleSizeMatrix.cpp:723-745):
// LU with partial pivoting, then forward/back substitution. A is column-major.
void lu_solve_copy(int n, const std::vector<Real>& Acm, Real* b) {
    std::vector<Real> A(Acm);
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        Real best = std::fabs(A[k + k * n]);
        for (int i = k + 1; i < n; ++i) {
            Real v = std::fabs(A[i + k * n]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[k + j * n], A[p + j * n]);
        Real inv = 1.0 / A[k + k * n];
        for (int i = k + 1; i < n; ++i) A[i + k * n] *= inv;
        for (int j = k + 1; j < n; ++j) {
            Real akj = A[k + j * n];
            if (akj != 0.0)
                for (int i = k + 1; i < n; ++i) A[i + j * n] -= A[i + k * n] * akj;
        }
    }
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int j = 0; j < n; ++j) {
        Real bj = b[j];
        for (int i = j + 1; i < n; ++i) b[i] -= A[i + j * n] * bj;
    }
    for (int j = n - 1; j >= 0; --j) {
        b[j] /= A[j + j * n];
        Real bj = b[j];
        for (int i = 0; i < j; ++i) b[i] -= A[i + j * n] * bj;
    }
}

// ------------------------------------------------------------ problem
enum OpKind { OP_NONE = 0, OP_ADVECTION = 1, OP_ACOUSTIC = 2 };

struct FaceRec {
    int64_t L = -1, R = -1;
    int fL = 0, fR = 0;
    int orient = 0;
    bool internal = false;
};

struct Problem {
    int shape = 0, dim = 0, nv = 0, nfields = 1;
    int64_t nE = 0;
    std::vector<Real> X;        // [nE][nv][dim]
    std::vector<int64_t> nodes;   // [nE][nv]
    std::vector<int> order;       // [nE]
    std::vector<int64_t> offset;  // [nE+1] DOF offsets (nB * nfields per element)
    std::vector<FaceRec> faces;
    std::map<int, Basis> bases;   // by p
    int op = OP_NONE;
    Real a[3] = {0, 0, 0};
    std::vector<Real> cinv;     // acoustic: per element
    // assembled (HpgemAPILinear) data
    std::vector<std::vector<Real>> M;      // per element, column-major (nB*nf)^2
    std::vector<std::vector<Real>> S;      // per element stiffness, column-major
    std::vector<std::array<std::vector<Real>, 4>> F;  // per face: LL, LR, RL, RR column-major
    std::vector<std::vector<std::pair<int64_t, int>>> elem_faces;  // (face, side) per element, all-core path
    bool assembled = false;

    const Basis& basis(int p) {
        auto it = bases.find(p);
        if (it == bases.end()) it = bases.emplace(p, make_basis(shape, p)).first;
        return it->second;
    }
    int nB(int64_t e) { return basis(order[e]).size(); }
    const Real* Xe(int64_t e) const { return &X[size_t(e) * nv * dim]; }
};

// Face search of MeshManipulator::readMesh (MeshManipulator_Impl.h:1860-1950):
// elements in order, local faces in order; a face not yet attached to the
// element is matched against the elements sharing all of its nodes.
void build_faces(Problem& P) {
    int nf = n_faces(P.shape);
    int fshape = face_shape(P.shape);
    int nfn = shape_nodes(fshape);
    std::map<int64_t, std::vector<int64_t>> node_elems;
    for (int64_t e = 0; e < P.nE; ++e)
        for (int v = 0; v < P.nv; ++v) {
            auto& lst = node_elems[P.nodes[e * P.nv + v]];
            if (lst.empty() || lst.back() != e) lst.push_back(e);
        }
    std::vector<char> attached(size_t(P.nE) * nf, 0);
    for (int64_t e = 0; e < P.nE; ++e) {
        for (int i = 0; i < nf; ++i) {
            if (attached[e * nf + i]) continue;
            const int* fn = face_nodes(P.shape, i);
            std::vector<int64_t> cand = node_elems[P.nodes[e * P.nv + fn[0]]];
            for (int j = 1; j < nfn; ++j) {
                const auto& nxt = node_elems[P.nodes[e * P.nv + fn[j]]];
                std::vector<int64_t> tmp;
                std::set_intersection(cand.begin(), cand.end(), nxt.begin(), nxt.end(),
                                      std::back_inserter(tmp));
                cand.swap(tmp);
            }
            if (cand.size() != 1 && cand.size() != 2)
                throw std::runtime_error("face bounded by an impossible number of elements");
            FaceRec f;
            f.L = e;
            f.fL = i;
            attached[e * nf + i] = 1;
            if (cand.size() == 2) {
                int64_t other = cand[0] == e ? cand[1] : cand[0];
                std::vector<int64_t> local;
                for (int j = 0; j < nfn; ++j) local.push_back(P.nodes[e * P.nv + fn[j]]);
                bool found = false;
                for (int j = 0; j < nf; ++j) {
                    const int* on = face_nodes(P.shape, j);
                    bool match = true;
                    for (int k = 0; k < nfn; ++k)
                        if (std::find(local.begin(), local.end(), P.nodes[other * P.nv + on[k]]) == local.end())
                            match = false;
                    if (match) {
                        f.R = other;
                        f.fR = j;
                        found = true;
                    }
                }
                if (!found) throw std::runtime_error("no matching face in the opposing element");
                attached[other * nf + f.fR] = 1;
                f.internal = true;
                std::vector<int64_t> ln, rn;
                const int* rnod = face_nodes(P.shape, f.fR);
                for (int k = 0; k < nfn; ++k) {
                    ln.push_back(P.nodes[e * P.nv + fn[k]]);
                    rn.push_back(P.nodes[other * P.nv + rnod[k]]);
                }
                f.orient = orientation_index(fshape, ln, rn);
            }
            P.faces.push_back(f);
        }
    }
}

// Evaluation context at one face quadrature point (PhysicalFace).
struct FacePoint {
    Real w;                 // quadrature weight * |n|
    Real nunit[3];          // unit normal, left -> right
    Real x[3];              // physical point (left side)
    std::vector<Real> phiL, phiR;
};

void face_point(Problem& P, const FaceRec& f, const Real* pf, Real wq, FacePoint& out) {
    int dim = P.dim;
    Real peL[3], jf[6], sgn;
    face_map(P.shape, f.fL, pf, peL, jf, &sgn);
    Real x[3], J[9];
    elem_map(P.shape, P.Xe(f.L), peL, x, J);
    // j3 = J_elem * J_face  (dim x dim-1)
    Real j3[6] = {0, 0, 0, 0, 0, 0};
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim - 1; ++c)
            for (int k = 0; k < dim; ++k) j3[r * (dim - 1) + c] += J[r * dim + k] * jf[k * (dim - 1) + c];
    Real n[3] = {0, 0, 0};
    if (dim == 2) {
        n[0] = -j3[1 * 1 + 0];
        n[1] = j3[0];
    } else {
        // SmallMatrix::computeWedgeStuffVector, 3 rows x 2 columns
        auto m = [&](int r, int c) { return j3[r * 2 + c]; };
        n[0] = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
        n[1] = m(0, 1) * m(2, 0) - m(0, 0) * m(2, 1);
        n[2] = m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1);
    }
    Real dJ = det(dim, J);
    Real s = (dJ > 0 ? 1.0 : -1.0) * sgn;
    Real nn = 0;
    for (int d = 0; d < dim; ++d) {
        n[d] *= s;
        nn += n[d] * n[d];
    }
    nn = std::sqrt(nn);
    out.w = wq * nn;
    for (int d = 0; d < 3; ++d) {
        out.nunit[d] = d < dim ? n[d] / nn : 0;
        out.x[d] = d < dim ? x[d] : 0;
    }
    const Basis& bL = P.basis(P.order[f.L]);
    out.phiL.assign(bL.size(), 0);
    std::vector<Real> g(size_t(bL.size()) * dim);
    eval_basis(bL, peL, out.phiL.data(), g.data());
    out.phiR.clear();
    if (f.internal) {
        Real pf2[2], peR[3];
        orientation_map(face_shape(P.shape), f.orient, pf, pf2);
        face_map(P.shape, f.fR, pf2, peR, nullptr, nullptr);
        const Basis& bR = P.basis(P.order[f.R]);
        out.phiR.assign(bR.size(), 0);
        std::vector<Real> gR(size_t(bR.size()) * dim);
        eval_basis(bR, peR, out.phiR.data(), gR.data());
    }
}

// Element quadrature order 2p+1 (Element.cpp:89-98); face order: the order of
// the larger element rule (Face.cpp:149-168).
int element_rule_order(Problem& P, int64_t e) { return get_rule(P.shape, 2 * P.order[e] + 1).order; }
int face_rule_order(Problem& P, const FaceRec& f) {
    int o = element_rule_order(P, f.L);
    if (f.internal) o = std::max(o, element_rule_order(P, f.R));
    return o;
}

// Element-point evaluation context (PhysicalElement).
struct ElemPoint {
    Real w;  // weight * |det J|
    Real x[3];
    std::vector<Real> phi, dphi;  // dphi physical [nB][dim]
};

void elem_point(Problem& P, int64_t e, const Real* xi, Real wq, ElemPoint& out) {
    int dim = P.dim;
    Real J[9];
    elem_map(P.shape, P.Xe(e), xi, out.x, J);
    Real dJ = det(dim, J);
    out.w = wq * std::fabs(dJ);
    const Basis& b = P.basis(P.order[e]);
    int n = b.size();
    out.phi.assign(n, 0);
    std::vector<Real> gref(size_t(n) * dim);
    eval_basis(b, xi, out.phi.data(), gref.data());
    out.dphi.assign(size_t(n) * dim, 0);
    for (int i = 0; i < n; ++i) {
        Real A[9], rhs[3];
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c) A[r * dim + c] = J[c * dim + r];  // J^T
        for (int d = 0; d < dim; ++d) rhs[d] = gref[i * dim + d];
        solve_small(dim, A, rhs);
        for (int d = 0; d < dim; ++d) out.dphi[i * dim + d] = rhs[d];
    }
}

// ------------------------------------------------------------ assembly
// HpgemAPILinear::createMassMatrices / createStiffnessMatrices with the
// advection integrands of 080AdvectionLinear_SelfTest.cpp:108-170 and the
// acoustic integrands of AcousticWaveLinear_Impl.h:102-283.
void assemble(Problem& P) {
    int dim = P.dim;
    int nfld = P.nfields;
    P.M.assign(P.nE, {});
    P.S.assign(P.nE, {});
    // every basis in use exists before the loops (read-only inside them; the
    // CPU-baseline build runs the loops on several threads)
    for (int64_t e = 0; e < P.nE; ++e) P.basis(P.order[e]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t e = 0; e < P.nE; ++e) {
        int nB = P.nB(e);
        int N = nB * nfld;
        std::vector<Real>& M = P.M[e];
        std::vector<Real>& S = P.S[e];
        M.assign(size_t(N) * N, 0);
        S.assign(size_t(N) * N, 0);
        Rule r = get_rule(P.shape, 2 * P.order[e] + 1);
        ElemPoint ep;
        for (int q = 0; q < r.n(); ++q) {
            elem_point(P, e, &r.pts[q * dim], r.w[q], ep);
            for (int v = 0; v < nfld; ++v) {
                Real mscale = 1.0;
                if (P.op == OP_ACOUSTIC && v > 0) mscale = P.cinv[e];
                for (int j = 0; j < nB; ++j)
                    for (int i = 0; i < nB; ++i)
                        M[(v * nB + i) + size_t(v * nB + j) * N] += ep.w * mscale * ep.phi[i] * ep.phi[j];
            }
            if (P.op == OP_ADVECTION) {
                // result(j,i) = phi_i * (a . grad phi_j): row j = test, column i = trial
                for (int i = 0; i < nB; ++i)
                    for (int j = 0; j < nB; ++j) {
                        Real adg = 0;
                        for (int d = 0; d < dim; ++d) adg += P.a[d] * ep.dphi[j * dim + d];
                        S[j + size_t(i) * N] += ep.w * ep.phi[i] * adg;
                    }
            } else if (P.op == OP_ACOUSTIC) {
                // fields: 0 = pressure-like v, 1..dim = s components (strong form)
                // dv/dt: phi_i * div s ; ds_d/dt: phi_i * d_d v
                for (int i = 0; i < nB; ++i)
                    for (int j = 0; j < nB; ++j)
                        for (int d = 0; d < dim; ++d) {
                            S[(0 * nB + i) + size_t((1 + d) * nB + j) * N] += ep.w * ep.phi[i] * ep.dphi[j * dim + d];
                            S[((1 + d) * nB + i) + size_t(0 * nB + j) * N] += ep.w * ep.phi[i] * ep.dphi[j * dim + d];
                        }
            }
        }
    }
    P.F.assign(P.faces.size(), {});
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t fi = 0; fi < P.faces.size(); ++fi) {
        const FaceRec& f = P.faces[fi];
        int nBL = P.nB(f.L) * nfld, nBR = f.internal ? P.nB(f.R) * nfld : 0;
        auto& blk = P.F[fi];
        blk[0].assign(size_t(nBL) * nBL, 0);
        blk[1].assign(size_t(nBL) * nBR, 0);
        blk[2].assign(size_t(nBR) * nBL, 0);
        blk[3].assign(size_t(nBR) * nBR, 0);
        Rule r = get_rule(face_shape(P.shape), face_rule_order(P, f));
        FacePoint fp;
        for (int q = 0; q < r.n(); ++q) {
            face_point(P, f, &r.pts[q * (dim - 1)], r.w[q], fp);
            if (P.op == OP_ADVECTION) {
                Real A = 0;
                for (int d = 0; d < dim; ++d) A += P.a[d] * fp.nunit[d];
                int nL = int(fp.phiL.size()), nR = int(fp.phiR.size());
                // integrandVal(j, i) = -(a . n_j phi_j) phi_i for trial i on the
                // upwind side; both sides halved when a.n vanishes
                // (080AdvectionLinear_SelfTest.cpp:125-170).
                auto add = [&](int bj, int rows, int j, int bi, int i, Real v) {
                    blk[bj * 2 + bi][j + size_t(i) * rows] += fp.w * v;
                };
                for (int sj = 0; sj < (f.internal ? 2 : 1); ++sj) {
                    const std::vector<Real>& pj = sj == 0 ? fp.phiL : fp.phiR;
                    Real sgnj = sj == 0 ? 1.0 : -1.0;
                    int rows = sj == 0 ? nL : nR;
                    for (int si = 0; si < (f.internal ? 2 : 1); ++si) {
                        const std::vector<Real>& pi = si == 0 ? fp.phiL : fp.phiR;
                        Real factor;
                        if (A > 1e-12 && si == 0) factor = 1.0;
                        else if (A < -1e-12 && si == 1) factor = 1.0;
                        else if (std::fabs(A) < 1e-12) factor = 0.5;
                        else continue;
                        for (size_t j = 0; j < pj.size(); ++j)
                            for (size_t i = 0; i < pi.size(); ++i)
                                add(sj, rows, int(j), si, int(i), -(A * sgnj * pj[j]) * pi[i] * factor);
                    }
                }
            } else if (P.op == OP_ACOUSTIC) {
                // central flux (AcousticWaveLinear_Impl.h:215-283): the normal is
                // the left element's outward normal, negated when the TRIAL
                // function lives on the right element; internal faces carry
                // -1/2 phi_i n_u.s_u (v equation) and -1/2 phi_i n_u,d v_u (s_d
                // equations) for both test sides; a reflecting wall on boundary
                // faces carries -phi_i (n . s) for v and nothing for s.
                int nBe_L = int(fp.phiL.size());
                int nBe_R = int(fp.phiR.size());
                const Real* n = fp.nunit;
                auto addb = [&](int blkid, int rows, int row, int col, Real v) {
                    blk[blkid][row + size_t(col) * rows] += fp.w * v;
                };
                if (!f.internal) {
                    int rows = nBe_L * nfld;
                    for (int i = 0; i < nBe_L; ++i)
                        for (int j = 0; j < nBe_L; ++j)
                            for (int d = 0; d < dim; ++d)
                                addb(0, rows, 0 * nBe_L + i, (1 + d) * nBe_L + j,
                                     -fp.phiL[i] * n[d] * fp.phiL[j]);
                } else {
                    // test side t (0 = L, 1 = R), trial side u
                    for (int t = 0; t < 2; ++t) {
                        const std::vector<Real>& pt = t == 0 ? fp.phiL : fp.phiR;
                        int nt = t == 0 ? nBe_L : nBe_R;
                        for (int u = 0; u < 2; ++u) {
                            const std::vector<Real>& pu = u == 0 ? fp.phiL : fp.phiR;
                            int nu = u == 0 ? nBe_L : nBe_R;
                            Real su = u == 0 ? 1.0 : -1.0;  // normal sign of the trial side
                            int blkid = t * 2 + u;
                            int rows = nt * nfld;
                            for (int i = 0; i < nt; ++i)
                                for (int j = 0; j < nu; ++j)
                                    for (int d = 0; d < dim; ++d) {
                                        Real v = -0.5 * pt[i] * su * n[d] * pu[j];
                                        addb(blkid, rows, 0 * nt + i, (1 + d) * nu + j, v);
                                        addb(blkid, rows, (1 + d) * nt + i, 0 * nu + j, v);
                                    }
                        }
                    }
                }
            }
        }
    }
    P.assembled = true;
}

// getLinearCombinationOfVectors + multiplyStiffnessMatrices
// (HpgemAPILinear::computeRightHandSide, HpgemAPILinear_Impl.h:360-434,
// 485-498): r = S w with w the stage combination; with solve = true, then the
// per-element mass solve (solveMassMatrixEquations, :143-156) as
// computeTimeDerivative does.
void time_derivative(Problem& P, int nin, const Real* const* vecs, const Real* coef, Real* out, bool solve = true) {
    if (!P.assembled) assemble(P);
    int64_t ndof = P.offset[P.nE];
    std::vector<Real> w(ndof);
    // first vector scaled, then axpy for the rest (HpgemAPISimplified_Impl.h:599-620)
    for (int64_t i = 0; i < ndof; ++i) w[i] = vecs[0][i] * coef[0];
    for (int k = 1; k < nin; ++k)
        for (int64_t i = 0; i < ndof; ++i) w[i] += coef[k] * vecs[k][i];
    auto gemv_add = [](int rows, int cols, const std::vector<Real>& A, const Real* x, Real* y, bool overwrite) {
        std::vector<Real> t(rows, 0.0);
        for (int c = 0; c < cols; ++c) {
            Real xc = x[c];
            for (int r = 0; r < rows; ++r) t[r] += A[r + size_t(c) * rows] * xc;
        }
        for (int r = 0; r < rows; ++r) y[r] = overwrite ? t[r] : y[r] + t[r];
    };
    for (int64_t e = 0; e < P.nE; ++e) {
        int N = P.nB(e) * P.nfields;
        gemv_add(N, N, P.S[e], &w[P.offset[e]], &out[P.offset[e]], true);
    }
    for (size_t fi = 0; fi < P.faces.size(); ++fi) {
        const FaceRec& f = P.faces[fi];
        int NL = P.nB(f.L) * P.nfields;
        if (f.internal) {
            int NR = P.nB(f.R) * P.nfields;
            gemv_add(NL, NL, P.F[fi][0], &w[P.offset[f.L]], &out[P.offset[f.L]], false);
            gemv_add(NL, NR, P.F[fi][1], &w[P.offset[f.R]], &out[P.offset[f.L]], false);
            gemv_add(NR, NL, P.F[fi][2], &w[P.offset[f.L]], &out[P.offset[f.R]], false);
            gemv_add(NR, NR, P.F[fi][3], &w[P.offset[f.R]], &out[P.offset[f.R]], false);
        } else {
            gemv_add(NL, NL, P.F[fi][0], &w[P.offset[f.L]], &out[P.offset[f.L]], false);
        }
    }
    if (!solve) return;
    for (int64_t e = 0; e < P.nE; ++e) {
        int N = P.nB(e) * P.nfields;
        lu_solve_copy(N, P.M[e], &out[P.offset[e]]);
    }
}

// The same operator with the loops spread over threads, for the all-core CPU
// baseline only (the reference runs one thread per MPI rank; this stands in
// for its multi-rank scaling): each element gathers the face blocks that act
// on it (its own side's rows of every face it lies on), so the threads share
// no output; the element's sum then runs in a different order than the
// face-ordered serial loop (same terms, different rounding).
void time_derivative_par(Problem& P, int nin, const Real* const* vecs, const Real* coef, Real* out) {
    if (!P.assembled) assemble(P);
    const int64_t ndof = P.offset[P.nE];
    std::vector<Real> w(ndof);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ndof; ++i) {
        Real v = vecs[0][i] * coef[0];
        for (int k = 1; k < nin; ++k) v += coef[k] * vecs[k][i];
        w[i] = v;
    }
    if (P.elem_faces.empty()) {  // (face, side) pairs per element, built once
        P.elem_faces.assign(P.nE, {});
        for (size_t fi = 0; fi < P.faces.size(); ++fi) {
            P.elem_faces[P.faces[fi].L].push_back({int64_t(fi), 0});
            if (P.faces[fi].internal) P.elem_faces[P.faces[fi].R].push_back({int64_t(fi), 1});
        }
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t e = 0; e < P.nE; ++e) {
        const int N = P.nB(e) * P.nfields;
        std::vector<Real> t(N, 0.0);
        auto gemv = [&](int cols, const std::vector<Real>& A, const Real* x) {
            for (int c = 0; c < cols; ++c) {
                const Real xc = x[c];
                for (int r = 0; r < N; ++r) t[r] += A[r + size_t(c) * N] * xc;
            }
        };
        gemv(N, P.S[e], &w[P.offset[e]]);
        for (const auto& fs : P.elem_faces[e]) {
            const FaceRec& f = P.faces[fs.first];
            const int NL = P.nB(f.L) * P.nfields;
            const int NR = f.internal ? P.nB(f.R) * P.nfields : 0;
            if (fs.second == 0) {
                gemv(NL, P.F[fs.first][0], &w[P.offset[f.L]]);
                if (f.internal) gemv(NR, P.F[fs.first][1], &w[P.offset[f.R]]);
            } else {
                gemv(NL, P.F[fs.first][2], &w[P.offset[f.L]]);
                gemv(NR, P.F[fs.first][3], &w[P.offset[f.R]]);
            }
        }
        lu_solve_copy(N, P.M[e], t.data());
        std::copy(t.begin(), t.end(), out + P.offset[e]);
    }
}

// HpgemAPILinear::solveMassMatrixEquations (HpgemAPILinear_Impl.h:143-156):
// every element's block solved in place with its mass matrix by LU.
void solve_mass(Problem& P, Real* v) {
    if (!P.assembled) assemble(P);
    for (int64_t e = 0; e < P.nE; ++e) lu_solve_copy(P.nB(e) * P.nfields, P.M[e], &v[P.offset[e]]);
}

// RK tableau: classic RK4 (RK4Methods.cpp:45-63)
const Real RK4_A[4][3] = {{0, 0, 0}, {0.5, 0, 0}, {0.0, 0.5, 0}, {0.0, 0.0, 1.0}};
const Real RK4_B[4] = {Real(1) / 6, Real(1) / 3, Real(1) / 3, Real(1) / 6};

// One step of an explicit tableau with s stages (a row-major s x s, strictly
// lower), exactly as computeOneTimeStep: stage inputs u + dt sum_j a_ij k_j,
// then u += dt b_j k_j one stage at a time.
void rk_step(Problem& P, Real* u, Real dt, int s, const Real* a, const Real* b) {
    int64_t ndof = P.offset[P.nE];
    std::vector<std::vector<Real>> k(s, std::vector<Real>(ndof));
    for (int i = 0; i < s; ++i) {
        std::vector<const Real*> vecs{u};
        std::vector<Real> coef{1.0};
        for (int j = 0; j < i; ++j) {
            vecs.push_back(k[j].data());
            coef.push_back(dt * a[i * s + j]);
        }
        time_derivative(P, int(vecs.size()), vecs.data(), coef.data(), k[i].data());
    }
    for (int j = 0; j < s; ++j) {
        Real c = dt * b[j];
        for (int64_t i = 0; i < ndof; ++i) u[i] += c * k[j][i];
    }
}

void rk4_step(Problem& P, Real* u, Real dt) {
    int64_t ndof = P.offset[P.nE];
    std::vector<std::vector<Real>> k(4, std::vector<Real>(ndof));
    for (int s = 0; s < 4; ++s) {
        std::vector<const Real*> vecs{u};
        std::vector<Real> coef{1.0};
        for (int j = 0; j < s; ++j) {
            vecs.push_back(k[j].data());
            coef.push_back(dt * RK4_A[s][j]);
        }
        time_derivative(P, int(vecs.size()), vecs.data(), coef.data(), k[s].data());
    }
    for (int s = 0; s < 4; ++s) {
        Real c = dt * RK4_B[s];
        for (int64_t i = 0; i < ndof; ++i) u[i] += c * k[s][i];
    }
}

// Initial conditions: sum of sine modes  f(x) = sum_m A_m sin(2 pi k_m . x)
struct Modes {
    int n = 0;
    const Real* amp = nullptr;
    const Real* k = nullptr;  // [n][3]
    Real eval(const Real* x, int dim, const Real* shift) const {
        Real s = 0;
        for (int m = 0; m < n; ++m) {
            Real arg = 0;
            for (int d = 0; d < dim; ++d) arg += k[m * 3 + d] * (x[d] - (shift ? shift[d] : 0.0));
            s += amp[m] * std::sin(2 * kPi * arg);
        }
        return s;
    }
};

// integrateInitialSolution + solveMassMatrixEquations (HpgemAPISimplified_Impl.h:227-291,751-758);
// every field receives the same scalar function scaled by field_scale[v].
void project(Problem& P, const Modes& m, const Real* field_scale, Real* u) {
    if (!P.assembled) assemble(P);
    int dim = P.dim;
    for (int64_t e = 0; e < P.nE; ++e) {
        int nB = P.nB(e);
        int N = nB * P.nfields;
        std::vector<Real> r(N, 0.0);
        Rule rule = get_rule(P.shape, 2 * P.order[e] + 1);
        ElemPoint ep;
        for (int q = 0; q < rule.n(); ++q) {
            elem_point(P, e, &rule.pts[q * dim], rule.w[q], ep);
            Real f = m.eval(ep.x, dim, nullptr);
            for (int v = 0; v < P.nfields; ++v)
                for (int i = 0; i < nB; ++i) r[v * nB + i] += ep.w * ep.phi[i] * f * field_scale[v];
        }
        // Mass matrix without the acoustic weighting: the initial solution is a
        // plain L2 projection of each field.
        std::vector<Real> M(size_t(N) * N, 0.0);
        for (int q = 0; q < rule.n(); ++q) {
            elem_point(P, e, &rule.pts[q * dim], rule.w[q], ep);
            for (int v = 0; v < P.nfields; ++v)
                for (int j = 0; j < nB; ++j)
                    for (int i = 0; i < nB; ++i)
                        M[(v * nB + i) + size_t(v * nB + j) * N] += ep.w * ep.phi[i] * ep.phi[j];
        }
        lu_solve_copy(N, M, r.data());
        std::copy(r.begin(), r.end(), u + P.offset[e]);
    }
}

// computeTotalError with rule order 2p+2 (HpgemAPISimplified_Impl.h:339-401),
// error against the shifted initial modes u0(x - a t) for field 0.
Real l2_error(Problem& P, const Real* u, const Modes& m, Real t) {
    int dim = P.dim;
    Real total = 0;
    Real shift[3];
    for (int d = 0; d < 3; ++d) shift[d] = P.a[d] * t;
    for (int64_t e = 0; e < P.nE; ++e) {
        int nB = P.nB(e);
        Rule rule = get_rule(P.shape, 2 * P.order[e] + 2);
        ElemPoint ep;
        for (int q = 0; q < rule.n(); ++q) {
            elem_point(P, e, &rule.pts[q * dim], rule.w[q], ep);
            Real num = 0;
            for (int j = 0; j < nB; ++j) num += ep.phi[j] * u[P.offset[e] + j];
            Real err = m.eval(ep.x, dim, shift) - num;
            total += ep.w * err * err;
        }
    }
    return std::sqrt(total);
}

// computeIntegrandTotalError / integrateErrorAtElement (HpgemAPISimplified_Impl.h:
// 293-366, rule order 2p+2) split by field: err2[v] = sum_e int_e |f_v(x - shift)
// - u_h,v|^2, where field v's exact function is sum_m amp[v][m] sin(2 pi k_m .
// (x - shift)). The caller combines the fields (plain sum, or the acoustic
// example's weighting of AcousticWaveLinear_Impl.h:285-312).
void l2_error_fields(Problem& P, const Real* u, int n_modes, const Real* amp, const Real* k, const Real* shift,
                     Real* err2) {
    int dim = P.dim;
    for (int v = 0; v < P.nfields; ++v) err2[v] = 0;
    for (int64_t e = 0; e < P.nE; ++e) {
        int nB = P.nB(e);
        Rule rule = get_rule(P.shape, 2 * P.order[e] + 2);
        ElemPoint ep;
        for (int q = 0; q < rule.n(); ++q) {
            elem_point(P, e, &rule.pts[q * dim], rule.w[q], ep);
            for (int v = 0; v < P.nfields; ++v) {
                Modes m;
                m.n = n_modes;
                m.amp = amp + size_t(v) * n_modes;
                m.k = k;
                Real num = 0;
                for (int j = 0; j < nB; ++j) num += ep.phi[j] * u[P.offset[e] + v * nB + j];
                Real err = m.eval(ep.x, dim, shift) - num;
                err2[v] += ep.w * err * err;
            }
        }
    }
}

// ------------------------------------------------------------ mesh generator
// StructuredReader (applications/Preprocessor/lib/hpgem.cpp:100-235).
struct MeshOut {
    int shape = 0, dim = 0, nv = 0;
    std::vector<Real> X;
    std::vector<int64_t> nodes;
};

MeshOut structured_mesh(int dim, const int* n, const Real* lo, const Real* hi, const int* periodic,
                        int simplices) {
    MeshOut out;
    out.dim = dim;
    std::vector<int> nn(dim);
    for (int d = 0; d < dim; ++d) nn[d] = n[d] + (periodic[d] ? 0 : 1);
    std::vector<int64_t> mult(dim, 1);
    for (int d = 1; d < dim; ++d) mult[d] = mult[d - 1] * nn[d - 1];
    std::vector<std::vector<int>> corners;
    std::vector<int> rot = {2, 0, 3, 1, 6, 4, 7, 5};
    if (simplices) {
        if (dim == 2) corners = {{0, 1, 2}, {1, 3, 2}};
        else corners = {{1, 2, 4, 7}, {0, 1, 2, 4}, {1, 3, 2, 7}, {1, 4, 5, 7}, {2, 6, 4, 7}};
    }
    int nvc = 1 << dim;
    out.shape = simplices ? (dim == 2 ? TRIANGLE : TETRAHEDRON) : (dim == 2 ? SQUARE : CUBE);
    out.nv = simplices ? dim + 1 : nvc;
    std::vector<int> li(dim, 0);
    int64_t total = 1;
    for (int d = 0; d < dim; ++d) total *= n[d];
    if (simplices) total *= int64_t(corners.size());
    size_t tri = 0;
    for (int64_t el = 0; el < total; ++el) {
        // hypercube vertices: bit d of the local index = +1 step along d
        std::vector<int64_t> vnode(nvc);
        std::vector<Real> vx(size_t(nvc) * dim);
        for (int j = 0; j < nvc; ++j) {
            int64_t node = 0;
            for (int d = 0; d < dim; ++d) {
                int idx = li[d] + ((j >> d) & 1);
                Real xc = lo[d] + idx * (hi[d] - lo[d]) / n[d];
                if (idx == nn[d]) idx = 0;  // periodic wrap onto the copy coordinate
                node += idx * mult[d];
                vx[j * dim + d] = xc;
            }
            vnode[j] = node;
        }
        if (simplices) {
            for (int c : corners[tri]) {
                out.nodes.push_back(vnode[c]);
                for (int d = 0; d < dim; ++d) out.X.push_back(vx[c * dim + d]);
            }
            if (++tri == corners.size()) tri = 0;
        } else {
            for (int j = 0; j < nvc; ++j) {
                out.nodes.push_back(vnode[j]);
                for (int d = 0; d < dim; ++d) out.X.push_back(vx[j * dim + d]);
            }
        }
        if (tri == 0) {
            for (int d = 0; d < dim; ++d) {
                for (auto& cs : corners)
                    for (int& k : cs) k = rot[k];
                if (++li[d] < n[d]) break;
                for (auto& cs : corners)
                    for (int& k : cs)
                        if (li[d] % 2 == 1) k = rot[k];
                li[d] = 0;
            }
        }
    }
    return out;
}

thread_local std::string g_err;

}  // namespace orc

using namespace orc;

#define ORC_TRY(...)                       \
    try {                                  \
        __VA_ARGS__;                       \
        return 0;                          \
    } catch (const std::exception& ex) {   \
        g_err = ex.what();                 \
        return 1;                          \
    }

// ------------------------------------------------------------ Maxwell IP (DG-Max)
// Interior-penalty DG curl-curl operator of DG-Max
// (applications/DG-Max/Algorithms/DGMaxDiscretization.cpp:193-260 element
// matrices, :344-410 face matrices) with hpGEM's Ainsworth-Coyle H(curl) set
// (kernel/FE/BasisFunctions3DAinsworthCoyle.cpp) and the covariant transform
// of HCurlConformingTransformation (kernel/Base/HCurlConformingTransformation.h):
// phi = J^{-T} phi_ref, curl phi = J curl_ref / det J.
namespace mx {

Real bary(int node, const Real* p) { return node == 0 ? 1 - p[0] - p[1] - p[2] : p[node - 1]; }
void bary_grad(int node, Real* g) {
    for (int d = 0; d < 3; ++d) g[d] = node == 0 ? -1 : (d == node - 1 ? 1 : 0);
}
void outer(const Real* a, const Real* b, Real* r) {  // OuterProduct: a x b
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
// LegendrePolynomial / LegendrePolynomialDerivative (FE/helperFunctions.cpp:110-137)
Real legendre(int n, Real x) {
    if (n == 0) return 1;
    if (n == 1) return x;
    return (2.0 * n - 1) / n * x * legendre(n - 1, x) - (n - 1.0) / n * legendre(n - 2, x);
}
Real legendre_d(int n, Real x) {
    if (n == 0) return 0;
    if (n == 1) return 1;
    return (2.0 * n - 1) / n * (x * legendre_d(n - 1, x) + legendre(n - 1, x)) - (n - 1.0) / n * legendre_d(n - 2, x);
}

// BasisCurlEdgeAinsworthCoyle(deg, o, i)::eval / evalCurl (:80-165)
void edge(int deg, int o, int i, const Real* p, Real* val, Real* curl) {
    Real go[3], gi[3];
    bary_grad(o, go);
    bary_grad(i, gi);
    const Real vi = bary(i, p), vo = bary(o, p);
    Real ret[3], dum[3];
    for (int d = 0; d < 3; ++d) {
        ret[d] = go[d] * vi;
        dum[d] = gi[d] * vo;
    }
    if (deg == 0) {
        for (int d = 0; d < 3; ++d) val[d] = ret[d] - dum[d];
        outer(gi, go, curl);
        for (int d = 0; d < 3; ++d) curl[d] *= 2;
        return;
    }
    if (deg == 1) {
        for (int d = 0; d < 3; ++d) {
            val[d] = ret[d] + dum[d];
            curl[d] = 0;
        }
        return;
    }
    const Real s = vi - vo;
    Real phi0[3], phi1[3];
    for (int d = 0; d < 3; ++d) {
        phi0[d] = ret[d] - dum[d];
        phi1[d] = phi0[d] + 2 * dum[d];
    }
    const Real ca = -Real(deg - 1) / deg * legendre(deg - 2, s);
    const Real cb = Real(2 * deg - 1) / deg * legendre(deg - 1, s);
    for (int d = 0; d < 3; ++d) val[d] = ca * phi0[d] + cb * phi1[d];
    Real gs[3], c0[3], x1[3], x0[3];
    for (int d = 0; d < 3; ++d) gs[d] = gi[d] - go[d];
    outer(gi, go, c0);
    outer(gs, phi1, x1);
    outer(gs, phi0, x0);
    for (int d = 0; d < 3; ++d)
        curl[d] = ca * 2 * c0[d] + Real(2 * deg - 1) / deg * legendre_d(deg - 1, s) * x1[d] -
                  Real(deg - 1) / deg * legendre_d(deg - 2, s) * x0[d];
}

// BasisCurlEdgeFaceAinsworthCoyle(deg, opposing, special)::eval / evalCurl (:167-215)
void edge_face(int deg, int opp, int c, const Real* p, Real* val, Real* curl) {
    int a = 0;
    while (c == a || opp == a) ++a;
    int b = a + 1;
    while (c == b || opp == b) ++b;
    Real ga[3], gb[3], gc[3];
    bary_grad(a, ga);
    bary_grad(b, gb);
    bary_grad(c, gc);
    const Real va = bary(a, p), vb = bary(b, p);
    const Real P = legendre(deg - 2, vb - va), dP = legendre_d(deg - 2, vb - va);
    for (int d = 0; d < 3; ++d) val[d] = gc[d] * va * vb * P;
    Real r1[3], r2[3];
    outer(ga, gc, r1);
    outer(gb, gc, r2);
    for (int d = 0; d < 3; ++d) curl[d] = r1[d] * (vb * P - va * vb * dP) + r2[d] * (va * P + va * vb * dP);
}

int size(int p) { return 6 * (p + 1) + (p >= 2 ? 12 * (p - 1) : 0); }

// createDGBasisFunctionSet3DAinsworthCoyle (:417-450), orders 1 and 2
void eval(int p, const Real* x, Real* val, Real* curl) {
    if (p < 1 || p > 2) throw std::runtime_error("Ainsworth-Coyle set restated for orders 1 and 2");
    static const int E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {1, 3}, {1, 2}};
    static const int F[12][2] = {{3, 0}, {3, 1}, {3, 2}, {2, 0}, {2, 1}, {2, 3},
                                 {1, 0}, {1, 2}, {1, 3}, {0, 1}, {0, 2}, {0, 3}};
    int n = 0;
    for (int k = 0; k <= p; ++k)
        for (auto& e : E) {
            edge(k, e[0], e[1], x, val + 3 * n, curl + 3 * n);
            ++n;
        }
    for (int k = 2; k <= p; ++k)
        for (auto& f : F) {
            edge_face(k, f[0], f[1], x, val + 3 * n, curl + 3 * n);
            ++n;
        }
}

// physical values / curls at a reference point of an affine tetrahedron
void physical(int p, const Real* J, const Real* xi, std::vector<Real>& val, std::vector<Real>& curl) {
    const int n = size(p);
    std::vector<Real> vr(size_t(3 * n)), cr(size_t(3 * n));
    eval(p, xi, vr.data(), cr.data());
    const Real dJ = det(3, J);
    val.assign(size_t(3 * n), 0);
    curl.assign(size_t(3 * n), 0);
    for (int i = 0; i < n; ++i) {
        Real A[9], rhs[3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) A[r * 3 + c] = J[c * 3 + r];  // J^T
        for (int d = 0; d < 3; ++d) rhs[d] = vr[i * 3 + d];
        solve_small(3, A, rhs);
        for (int d = 0; d < 3; ++d) {
            val[i * 3 + d] = rhs[d];
            Real s = 0;
            for (int k = 0; k < 3; ++k) s += J[d * 3 + k] * cr[i * 3 + k];
            curl[i * 3 + d] = s / dJ;
        }
    }
}

struct Out {
    int n = 0;
    std::vector<std::vector<Real>> S, M;                 // per element, column-major n x n
    std::vector<std::array<std::vector<Real>, 4>> F;     // per face: (test side, trial side) blocks
};

void assemble(Problem& P, int p, const Real* eps, Real stab, Out& out) {
    if (P.shape != TETRAHEDRON) throw std::runtime_error("Maxwell assembly on tetrahedra");
    const int n = size(p);
    out.n = n;
    out.S.assign(P.nE, std::vector<Real>(size_t(n) * n, 0));
    out.M.assign(P.nE, std::vector<Real>(size_t(n) * n, 0));
    std::vector<Real> v, c;
    for (int64_t e = 0; e < P.nE; ++e) {
        Rule r = get_rule(TETRAHEDRON, 2 * p + 1);
        for (int q = 0; q < r.n(); ++q) {
            Real x[3], J[9];
            elem_map(TETRAHEDRON, P.Xe(e), &r.pts[q * 3], x, J);
            const Real w = r.w[q] * std::fabs(det(3, J));
            physical(p, J, &r.pts[q * 3], v, c);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    Real cc = 0, vv = 0;
                    for (int d = 0; d < 3; ++d) {
                        cc += c[i * 3 + d] * c[j * 3 + d];
                        vv += v[i * 3 + d] * v[j * 3 + d];
                    }
                    out.S[e][j + size_t(i) * n] += w * cc;
                    out.M[e][j + size_t(i) * n] += w * eps[e] * vv;
                }
        }
    }
    out.F.assign(P.faces.size(), {});
    for (size_t fi = 0; fi < P.faces.size(); ++fi) {
        const FaceRec& f = P.faces[fi];
        const int ns = f.internal ? 2 : 1;
        for (int b = 0; b < 4; ++b) out.F[fi][b].assign((b == 0 || f.internal) ? size_t(n) * n : 0, 0);
        // face diameter: largest distance between the face's vertices (FaceGeometry.cpp:176-194)
        const int* fn = face_nodes(TETRAHEDRON, f.fL);
        Real diam2 = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = a + 1; b < 3; ++b) {
                Real d2 = 0;
                for (int d = 0; d < 3; ++d) {
                    const Real dd = P.Xe(f.L)[fn[a] * 3 + d] - P.Xe(f.L)[fn[b] * 3 + d];
                    d2 += dd * dd;
                }
                diam2 = std::max(diam2, d2);
            }
        const Real local_stab = stab / std::sqrt(diam2);
        const Real factor = f.internal ? -0.5 : -1.0;
        Rule r = get_rule(TRIANGLE, 2 * p + 1);
        for (int q = 0; q < r.n(); ++q) {
            FacePoint fp;
            face_point(P, f, &r.pts[q * 2], r.w[q], fp);  // weight * |n| and the unit normal
            std::vector<Real> val[2], curl[2];
            Real peL[3], xL[3], JL[9];
            face_map(TETRAHEDRON, f.fL, &r.pts[q * 2], peL, nullptr, nullptr);
            elem_map(TETRAHEDRON, P.Xe(f.L), peL, xL, JL);
            physical(p, JL, peL, val[0], curl[0]);
            if (f.internal) {
                Real pf2[2], peR[3], xR[3], JR[9];
                orientation_map(TRIANGLE, f.orient, &r.pts[q * 2], pf2);
                face_map(TETRAHEDRON, f.fR, pf2, peR, nullptr, nullptr);
                elem_map(TETRAHEDRON, P.Xe(f.R), peR, xR, JR);
                physical(p, JR, peR, val[1], curl[1]);
            }
            // unit normal x phi, negated for the right element's functions
            // (PhysicalFace_Impl.h:454-489)
            std::vector<Real> T[2];
            for (int sd = 0; sd < ns; ++sd) {
                T[sd].assign(size_t(3 * n), 0);
                for (int i = 0; i < n; ++i) {
                    outer(fp.nunit, &val[sd][i * 3], &T[sd][i * 3]);
                    if (sd == 1)
                        for (int d = 0; d < 3; ++d) T[sd][i * 3 + d] *= -1;
                }
            }
            for (int su = 0; su < ns; ++su)      // side of trial function i
                for (int st = 0; st < ns; ++st)  // side of test function j
                    for (int i = 0; i < n; ++i)
                        for (int j = 0; j < n; ++j) {
                            Real ct = 0, tc = 0, tt = 0;
                            for (int d = 0; d < 3; ++d) {
                                ct += curl[su][i * 3 + d] * T[st][j * 3 + d];
                                tc += T[su][i * 3 + d] * curl[st][j * 3 + d];
                                tt += T[su][i * 3 + d] * T[st][j * 3 + d];
                            }
                            out.F[fi][st * 2 + su][j + size_t(i) * n] += fp.w * (factor * (ct + tc) + local_stab * tt);
                        }
        }
    }
}

}  // namespace mx

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Simplex rule data (shape TRIANGLE/TETRAHEDRON), npts points of dim coords.
int orc_set_simplex_rule(int shape, int order, int npts, const double* pts, const double* w) {
    ORC_TRY({
        Rule r;
        r.dim = shape_dim(shape);
        r.order = order;
        r.pts.assign(pts, pts + size_t(npts) * r.dim);
        r.w.assign(w, w + npts);
        g_simplex_rules[{shape, order}] = r;
    })
}

// Quadrature rule selected for a requested order. Two-phase: pass pts = NULL
// to query npts.
int orc_rule(int shape, int order, int* npts, int* actual_order, double* pts, double* w) {
    ORC_TRY({
        Rule r = get_rule(shape, order);
        *npts = r.n();
        *actual_order = r.order;
        if (pts) {
            std::copy(r.pts.begin(), r.pts.end(), pts);
            std::copy(r.w.begin(), r.w.end(), w);
        }
    })
}

int orc_basis_size(int shape, int p) {
    try {
        return make_basis(shape, p).size();
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Values [n] and reference gradients [n][dim] of the DG basis at one point.
int orc_basis_eval(int shape, int p, const double* pt, double* vals, double* grads) {
    ORC_TRY({
        Basis b = make_basis(shape, p);
        const int dim = shape_dim(shape);
        std::vector<Real> x(pt, pt + dim), v(b.size()), g(size_t(b.size()) * dim);
        eval_basis(b, x.data(), v.data(), g.data());
        std::copy(v.begin(), v.end(), vals);
        std::copy(g.begin(), g.end(), grads);
    })
}

int orc_face_map(int shape, int face, const double* pin, double* pout, double* sign) {
    ORC_TRY({
        const int dim = shape_dim(shape);
        Real pi[3] = {0, 0, 0}, po[3] = {0, 0, 0}, sg = 0;
        for (int d = 0; d + 1 < dim; ++d) pi[d] = pin[d];
        face_map(shape, face, pi, po, nullptr, &sg);
        for (int d = 0; d < dim; ++d) pout[d] = double(po[d]);
        *sign = double(sg);
    })
}

int orc_orientation_map(int fshape, int idx, const double* pin, double* pout) {
    ORC_TRY({
        const int dim = shape_dim(fshape);
        Real pi[3] = {0, 0, 0}, po[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) pi[d] = pin[d];
        orientation_map(fshape, idx, pi, po);
        for (int d = 0; d < dim; ++d) pout[d] = double(po[d]);
    })
}

int orc_orientation_index(int fshape, int n, const int64_t* l1, const int64_t* l2) {
    try {
        return orientation_index(fshape, std::vector<int64_t>(l1, l1 + n), std::vector<int64_t>(l2, l2 + n));
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// Structured mesh. Two-phase: call with X = NULL to get n_elem and nv.
int orc_structured_mesh(int dim, const int* n, const double* lo, const double* hi, const int* periodic,
                        int simplices, int64_t* n_elem, int* nv, int* shape, double* X, int64_t* nodes) {
    ORC_TRY({
        std::vector<Real> rlo(lo, lo + dim), rhi(hi, hi + dim);
        MeshOut m = structured_mesh(dim, n, rlo.data(), rhi.data(), periodic, simplices);
        *nv = m.nv;
        *shape = m.shape;
        *n_elem = int64_t(m.nodes.size() / m.nv);
        if (X) {
            std::copy(m.X.begin(), m.X.end(), X);
            std::copy(m.nodes.begin(), m.nodes.end(), nodes);
        }
    })
}

void* orc_problem_create(int shape, int64_t n_elem, const double* X, const int64_t* nodes, const int* orders,
                         int n_fields) {
    try {
        Problem* P = new Problem;
        P->shape = shape;
        P->dim = shape_dim(shape);
        P->nv = shape_nodes(shape);
        P->nE = n_elem;
        P->nfields = n_fields;
        P->X.assign(X, X + size_t(n_elem) * P->nv * P->dim);
        P->nodes.assign(nodes, nodes + size_t(n_elem) * P->nv);
        P->order.assign(orders, orders + n_elem);
        P->offset.assign(n_elem + 1, 0);
        for (int64_t e = 0; e < n_elem; ++e) P->offset[e + 1] = P->offset[e] + int64_t(P->nB(e)) * n_fields;
        build_faces(*P);
        return P;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}

void orc_problem_destroy(void* h) { delete static_cast<Problem*>(h); }

int orc_problem_info(void* h, int64_t* n_faces, int64_t* n_internal, int64_t* n_dof) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        *n_faces = int64_t(P->faces.size());
        int64_t ni = 0;
        for (auto& f : P->faces) ni += f.internal;
        *n_internal = ni;
        *n_dof = P->offset[P->nE];
    })
}

// Face table: per face L, R, fL, fR, orientation (R = -1 on boundary faces).
int orc_faces(void* h, int64_t* out /*[nF][5]*/) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        for (size_t i = 0; i < P->faces.size(); ++i) {
            const FaceRec& f = P->faces[i];
            out[i * 5 + 0] = f.L;
            out[i * 5 + 1] = f.R;
            out[i * 5 + 2] = f.fL;
            out[i * 5 + 3] = f.fR;
            out[i * 5 + 4] = f.orient;
        }
    })
}

int orc_set_advection(void* h, const double* a) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        P->op = OP_ADVECTION;
        for (int d = 0; d < 3; ++d) P->a[d] = d < P->dim ? a[d] : 0.0;
        P->assembled = false;
    })
}

int orc_set_acoustic(void* h, const double* cinv) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        if (P->nfields != P->dim + 1) throw std::runtime_error("acoustic needs dim+1 fields");
        P->op = OP_ACOUSTIC;
        P->cinv.assign(cinv, cinv + P->nE);
        P->assembled = false;
    })
}

int orc_assemble(void* h) { ORC_TRY({ assemble(*static_cast<Problem*>(h)); }) }

int orc_time_derivative(void* h, int n_in, const double* const* vecs, const double* coef, double* out) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        const size_t n = size_t(P->offset[P->nE]);
        std::vector<std::vector<Real>> in(n_in);
        std::vector<const Real*> ptrs(n_in);
        std::vector<Real> c(coef, coef + n_in), o(n);
        for (int i = 0; i < n_in; ++i) {
            in[i].assign(vecs[i], vecs[i] + n);
            ptrs[i] = in[i].data();
        }
        time_derivative(*P, n_in, ptrs.data(), c.data(), o.data());
        std::copy(o.begin(), o.end(), out);
    })
}

int orc_time_derivative_par(void* h, int n_in, const double* const* vecs, const double* coef, double* out) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        const size_t n = size_t(P->offset[P->nE]);
        std::vector<std::vector<Real>> in(n_in);
        std::vector<const Real*> ptrs(n_in);
        std::vector<Real> c(coef, coef + n_in), o(n);
        for (int i = 0; i < n_in; ++i) {
            in[i].assign(vecs[i], vecs[i] + n);
            ptrs[i] = in[i].data();
        }
        time_derivative_par(*P, n_in, ptrs.data(), c.data(), o.data());
        std::copy(o.begin(), o.end(), out);
    })
}

int orc_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_right_hand_side(void* h, int n_in, const double* const* vecs, const double* coef, double* out) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        const size_t n = size_t(P->offset[P->nE]);
        std::vector<std::vector<Real>> in(n_in);
        std::vector<const Real*> ptrs(n_in);
        std::vector<Real> c(coef, coef + n_in), o(n);
        for (int i = 0; i < n_in; ++i) {
            in[i].assign(vecs[i], vecs[i] + n);
            ptrs[i] = in[i].data();
        }
        time_derivative(*P, n_in, ptrs.data(), c.data(), o.data(), false);
        std::copy(o.begin(), o.end(), out);
    })
}

int orc_solve_mass(void* h, double* v) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> w(v, v + P->offset[P->nE]);
        solve_mass(*P, w.data());
        std::copy(w.begin(), w.end(), v);
    })
}

int orc_rk4(void* h, double* u, double dt, int n_steps) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> ur(u, u + P->offset[P->nE]);
        for (int s = 0; s < n_steps; ++s) rk4_step(*P, ur.data(), Real(dt));
        std::copy(ur.begin(), ur.end(), u);
    })
}

int orc_rk_steps(void* h, double* u, double dt, int n_steps, int stages, const double* a, const double* b) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> ur(u, u + P->offset[P->nE]);
        std::vector<Real> ar(a, a + stages * stages), br(b, b + stages);
        for (int s = 0; s < n_steps; ++s) rk_step(*P, ur.data(), Real(dt), stages, ar.data(), br.data());
        std::copy(ur.begin(), ur.end(), u);
    })
}

int orc_project_modes(void* h, int n_modes, const double* amp, const double* k, const double* field_scale,
                      double* u) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> ar(amp, amp + n_modes), kr(k, k + 3 * n_modes), fs(field_scale, field_scale + P->nfields);
        std::vector<Real> ur(size_t(P->offset[P->nE]));
        Modes m;
        m.n = n_modes;
        m.amp = ar.data();
        m.k = kr.data();
        project(*P, m, fs.data(), ur.data());
        std::copy(ur.begin(), ur.end(), u);
    })
}

int orc_l2_error_modes(void* h, const double* u, int n_modes, const double* amp, const double* k, double t,
                       double* err) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> ar(amp, amp + n_modes), kr(k, k + 3 * n_modes), ur(u, u + P->offset[P->nE]);
        Modes m;
        m.n = n_modes;
        m.amp = ar.data();
        m.k = kr.data();
        *err = double(l2_error(*P, ur.data(), m, Real(t)));
    })
}

int orc_l2_error_fields(void* h, const double* u, int n_modes, const double* amp, const double* k,
                        const double* shift, double* err2) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> ar(amp, amp + size_t(P->nfields) * n_modes), kr(k, k + 3 * n_modes),
            ur(u, u + P->offset[P->nE]), e2(P->nfields);
        Real sh[3] = {shift[0], shift[1], shift[2]};
        l2_error_fields(*P, ur.data(), n_modes, ar.data(), kr.data(), sh, e2.data());
        for (int v = 0; v < P->nfields; ++v) err2[v] = double(e2[v]);
    })
}

int orc_hcurl_eval(int p, const double* pt, double* val, double* curl) {
    ORC_TRY({
        const int n = mx::size(p);
        Real x[3] = {pt[0], pt[1], pt[2]};
        std::vector<Real> v(size_t(3 * n)), c(size_t(3 * n));
        mx::eval(p, x, v.data(), c.data());
        for (int i = 0; i < 3 * n; ++i) {
            val[i] = double(v[i]);
            curl[i] = double(c[i]);
        }
    })
}

// DG-Max element (S: curl-curl, M: permittivity mass) and face matrices,
// column-major n x n each; F per face: blocks (test side, trial side) LL, LR,
// RL, RR (boundary faces: LL only, the rest left zero).
int orc_maxwell_assemble(void* h, int p, const double* eps, double stab, double* S, double* M, double* F) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        std::vector<Real> er(eps, eps + P->nE);
        mx::Out o;
        mx::assemble(*P, p, er.data(), Real(stab), o);
        const size_t nn = size_t(o.n) * o.n;
        for (int64_t e = 0; e < P->nE; ++e)
            for (size_t k = 0; k < nn; ++k) {
                S[e * nn + k] = double(o.S[e][k]);
                M[e * nn + k] = double(o.M[e][k]);
            }
        for (size_t fi = 0; fi < o.F.size(); ++fi)
            for (int b = 0; b < 4; ++b)
                for (size_t k = 0; k < nn; ++k) F[(fi * 4 + b) * nn + k] = k < o.F[fi][b].size() ? double(o.F[fi][b][k]) : 0.0;
    })
}

// Element mass matrix of element e (column-major, (nB*nf)^2) for checks.
int orc_mass_matrix(void* h, int64_t e, double* out) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        if (!P->assembled) assemble(*P);
        std::copy(P->M[e].begin(), P->M[e].end(), out);
    })
}

// Element stiffness S_e (column-major n x n, n = nB(e) * nf) and the four
// blocks of face f's matrix (LL, LR, RL, RR, each column-major with rows =
// test side, columns = trial side; boundary faces: LL only), as
// HpgemAPILinear holds them after createStiffnessMatrices: for checks of a
// device path that takes the matrices.
int orc_element_stiffness(void* h, int64_t e, double* out) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        if (!P->assembled) assemble(*P);
        for (size_t k = 0; k < P->S[e].size(); ++k) out[k] = double(P->S[e][k]);
    })
}

int orc_face_blocks(void* h, int64_t f, double* out) {
    ORC_TRY({
        Problem* P = static_cast<Problem*>(h);
        if (!P->assembled) assemble(*P);
        size_t o = 0;
        for (int b = 0; b < 4; ++b)
            for (Real v : P->F[f][b]) out[o++] = double(v);
    })
}

}  // extern "C"
