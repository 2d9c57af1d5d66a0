#include "host_tables.hpp"

#include <cmath>
#include <stdexcept>

namespace hgdg {

int shape_dim(int shape) {
    switch (shape) {
        case SHAPE_SQUARE: return 2;
        case SHAPE_CUBE: case SHAPE_TETRAHEDRON: return 3;
    }
    throw std::invalid_argument("unknown shape " + std::to_string(shape));
}

int shape_vertices(int shape) {
    switch (shape) {
        case SHAPE_SQUARE: return 4;
        case SHAPE_CUBE: return 8;
        case SHAPE_TETRAHEDRON: return 4;
    }
    throw std::invalid_argument("unknown shape " + std::to_string(shape));
}

namespace {

// Legendre P_n and P_{n-1} at t by the three-term recurrence.
void legendre_pair(int n, long double t, long double& pn, long double& pn1) {
    long double a = 1, b = t;
    if (n == 0) { pn = 1; pn1 = 0; return; }
    for (int k = 2; k <= n; ++k) {
        long double c = ((2 * k - 1) * t * b - (k - 1) * a) / k;
        a = b;
        b = c;
    }
    pn = b;
    pn1 = a;
}

long double legendre(int n, long double t) {
    long double pn, pn1;
    legendre_pair(n, t, pn, pn1);
    return pn;
}

}  // namespace

void gauss_legendre_ld(int n, std::vector<long double>& x, std::vector<long double>& w) {
    x.assign(n, 0.0L);
    w.assign(n, 0.0L);
    if (n == 1) { w[0] = 2.0L; return; }
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; ++i) {
        long double t = -std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double pn = 0, pn1 = 0, dp = 1;
        for (int it = 0; it < 60; ++it) {
            legendre_pair(n, t, pn, pn1);
            dp = n * (t * pn - pn1) / (t * t - 1);
            const long double dt = pn / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-19L) break;
        }
        legendre_pair(n, t, pn, pn1);
        dp = n * (t * pn - pn1) / (t * t - 1);
        x[i] = t;
        w[i] = 2.0L / ((1 - t * t) * dp * dp);
    }
}

void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    if (n == 1) { w[0] = 2.0; return; }
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; ++i) {
        long double t = -std::cos(pi * (i + 0.75L) / (n + 0.5L));
        long double pn = 0, pn1 = 0, dp = 1;
        for (int it = 0; it < 60; ++it) {
            legendre_pair(n, t, pn, pn1);
            dp = n * (t * pn - pn1) / (t * t - 1);
            long double dt = pn / dp;
            t -= dt;
            if (std::fabs(static_cast<double>(dt)) < 1e-21) break;
        }
        legendre_pair(n, t, pn, pn1);
        dp = n * (t * pn - pn1) / (t * t - 1);
        x[i] = static_cast<double>(t);
        w[i] = static_cast<double>(2.0L / ((1 - t * t) * dp * dp));
    }
}

long double psi_ld(int a, long double t) {
    if (a == 0) return (1 - t) / 2;
    if (a == 1) return (1 + t) / 2;
    const int k = a - 2;
    return (legendre(k + 2, t) - legendre(k, t)) / std::sqrt(2.0L * (2 * k + 3));
}

long double dpsi_ld(int a, long double t) {
    if (a == 0) return -0.5L;
    if (a == 1) return 0.5L;
    const int k = a - 2;
    return std::sqrt((2.0L * k + 3) / 2) * legendre(k + 1, t);
}

double psi(int a, double t) {
    if (a == 0) return (1 - t) / 2;
    if (a == 1) return (1 + t) / 2;
    int k = a - 2;
    long double v = (legendre(k + 2, t) - legendre(k, t)) / std::sqrt(2.0L * (2 * k + 3));
    return static_cast<double>(v);
}

double dpsi(int a, double t) {
    if (a == 0) return -0.5;
    if (a == 1) return 0.5;
    int k = a - 2;
    return static_cast<double>(std::sqrt((2.0L * k + 3) / 2) * legendre(k + 1, t));
}

std::vector<int> tensor_permutation(int dim, int p) {
    if (p < 1) throw std::invalid_argument("order must be >= 1");
    const int n1 = p + 1;
    std::vector<int> perm;
    auto t2 = [&](int a, int b) { perm.push_back(a + n1 * b); };
    auto t3 = [&](int a, int b, int c) { perm.push_back(a + n1 * (b + n1 * c)); };
    if (dim == 2) {
        for (int v = 0; v < 4; ++v) t2(v & 1, v >> 1);
        for (int i = 0; i + 2 <= p; ++i) {
            t2(i + 2, 0);  // edge (0,1): bubble along x at y = -1
            t2(i + 2, 1);  // edge (2,3)
            t2(0, i + 2);  // edge (0,2): bubble along y at x = -1
            t2(1, i + 2);  // edge (1,3)
            for (int j = 0; j + 2 <= p; ++j) t2(i + 2, j + 2);
        }
        return perm;
    }
    if (dim != 3) throw std::invalid_argument("tensor permutation needs dim 2 or 3");
    for (int v = 0; v < 8; ++v) t3(v & 1, (v >> 1) & 1, v >> 2);
    // six face families of one (first, second) bubble-degree pair, in the
    // order z=-1, y=-1, x=-1, x=+1, y=+1, z=+1; the two bubble directions are
    // the free coordinates in increasing order.
    auto faces = [&](int i, int j) {
        t3(i + 2, j + 2, 0);
        t3(i + 2, 0, j + 2);
        t3(0, i + 2, j + 2);
        t3(1, i + 2, j + 2);
        t3(i + 2, 1, j + 2);
        t3(i + 2, j + 2, 1);
    };
    auto inner = [&](int i, int j, int k) { t3(i + 2, j + 2, k + 2); };
    for (int j = 0; j + 2 <= p; ++j) {
        // edges along x at (y,z) = (-,-), (+,-), (-,+), (+,+); then along y
        // at (x,z), then along z at (x,y), each in that sign order
        for (int q = 0; q < 4; ++q) t3(j + 2, q & 1, q >> 1);
        for (int q = 0; q < 4; ++q) t3(q & 1, j + 2, q >> 1);
        for (int q = 0; q < 4; ++q) t3(q & 1, q >> 1, j + 2);
        faces(j, j);
        inner(j, j, j);
        for (int i = 0; i < j; ++i) {
            faces(i, j);
            faces(j, i);
            inner(i, i, j);
            inner(i, j, i);
            inner(j, i, i);
            inner(j, j, i);
            inner(j, i, j);
            inner(i, j, j);
            for (int k = 0; k < i; ++k) {
                inner(i, j, k);
                inner(i, k, j);
                inner(j, i, k);
                inner(j, k, i);
                inner(k, i, j);
                inner(k, j, i);
            }
        }
    }
    return perm;
}

namespace {

std::vector<long double> invert(int n, const std::vector<long double>& A) {
    std::vector<long double> M(A), I(size_t(n) * n, 0.0L);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[i * n + k]) > std::fabs(M[piv * n + k])) piv = i;
        for (int j = 0; j < n; ++j) {
            std::swap(M[k * n + j], M[piv * n + j]);
            std::swap(I[k * n + j], I[piv * n + j]);
        }
        long double d = M[k * n + k];
        for (int j = 0; j < n; ++j) {
            M[k * n + j] /= d;
            I[k * n + j] /= d;
        }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            long double f = M[i * n + k];
            if (f == 0) continue;
            for (int j = 0; j < n; ++j) {
                M[i * n + j] -= f * M[k * n + j];
                I[i * n + j] -= f * I[k * n + j];
            }
        }
    }
    return I;
}

}  // namespace

TensorOps tensor_ops(int p) {
    TensorOps ops;
    const int n = p + 1;
    ops.n1 = n;
    std::vector<long double> x, w;
    gauss_legendre_ld(n + 1, x, w);  // exact for degree 2n+1 >= 2p
    std::vector<long double> M(size_t(n) * n, 0), G(size_t(n) * n, 0), K(size_t(n) * n, 0);
    for (size_t q = 0; q < x.size(); ++q)
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) {
                const long double wq = w[q];
                M[a * n + b] += wq * psi_ld(a, x[q]) * psi_ld(b, x[q]);
                G[a * n + b] += wq * psi_ld(b, x[q]) * dpsi_ld(a, x[q]);
                K[a * n + b] += wq * psi_ld(a, x[q]) * dpsi_ld(b, x[q]);
            }
    std::vector<long double> Mi = invert(n, M);
    ops.M1.resize(size_t(n) * n);
    ops.G1.resize(size_t(n) * n);
    ops.Mi.resize(size_t(n) * n);
    ops.D.resize(size_t(n) * n);
    ops.Dr.resize(size_t(n) * n);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            long double d = 0, dr = 0;
            for (int c = 0; c < n; ++c) {
                d += Mi[a * n + c] * G[c * n + b];
                dr += Mi[a * n + c] * K[c * n + b];
            }
            ops.M1[a * n + b] = static_cast<double>(M[a * n + b]);
            ops.G1[a * n + b] = static_cast<double>(G[a * n + b]);
            ops.Mi[a * n + b] = static_cast<double>(Mi[a * n + b]);
            ops.D[a * n + b] = static_cast<double>(d);
            ops.Dr[a * n + b] = static_cast<double>(dr);
        }
    return ops;
}

NodalOps nodal_ops(int p) {
    NodalOps o;
    const int n = p + 1;
    o.n1 = n;
    std::vector<long double> x, w, x2, w2;
    gauss_legendre_ld(n, x, w);
    gauss_legendre_ld(n + 1, x2, w2);
    // barycentric weights of the points: l_m'(x_q) = (bw_m / bw_q) / (x_q - x_m)
    // for q != m, and sum_{j != m} 1 / (x_m - x_j) on the diagonal
    std::vector<long double> bw(n, 1.0L);
    for (int m = 0; m < n; ++m)
        for (int j = 0; j < n; ++j)
            if (j != m) bw[m] /= (x[m] - x[j]);
    std::vector<long double> B(size_t(n) * n);
    for (int q = 0; q < n; ++q)
        for (int a = 0; a < n; ++a) B[q * n + a] = psi_ld(a, x[q]);
    const std::vector<long double> Bi = invert(n, B);
    auto lagrange = [&](int m, long double t) {
        long double v = 1;
        for (int j = 0; j < n; ++j)
            if (j != m) v *= (t - x[j]) / (x[m] - x[j]);
        return v;
    };
    for (int q = 0; q < n; ++q) {
        o.x.push_back(static_cast<double>(x[q]));
        o.w.push_back(static_cast<double>(w[q]));
        o.lam0.push_back(static_cast<double>(lagrange(q, -1.0L)));
        o.lam1.push_back(static_cast<double>(lagrange(q, 1.0L)));
    }
    for (int i = 0; i < n * n; ++i) {
        o.B.push_back(static_cast<double>(B[i]));
        o.Binv.push_back(static_cast<double>(Bi[i]));
    }
    for (int q = 0; q < n; ++q)
        for (int m = 0; m < n; ++m) {
            long double d = 0;
            if (q == m) {
                for (int j = 0; j < n; ++j)
                    if (j != m) d += 1.0L / (x[m] - x[j]);
            } else {
                d = (bw[m] / bw[q]) / (x[q] - x[m]);
            }
            o.C.push_back(static_cast<double>(d));
        }
    for (int q = 0; q <= n; ++q) {
        o.x2.push_back(static_cast<double>(x2[q]));
        o.w2.push_back(static_cast<double>(w2[q]));
    }
    for (int a = 0; a < n; ++a)
        for (int q = 0; q <= n; ++q) o.B2.push_back(static_cast<double>(psi_ld(a, x2[q])));
    return o;
}

std::vector<double> NodalOps::packed() const {
    std::vector<double> t;
    for (const std::vector<double>* v : {&x, &w, &B, &Binv, &C, &lam0, &lam1, &x2, &w2, &B2})
        t.insert(t.end(), v->begin(), v->end());
    return t;
}

StructuredMesh structured_mesh(int dim, const int32_t* n, const double* lo, const double* hi,
                               const int32_t* periodic, bool simplices) {
    if (dim != 2 && dim != 3) throw std::invalid_argument("structured mesh needs dim 2 or 3");
    StructuredMesh m;
    m.dim = dim;
    std::vector<int64_t> nn(dim), stride(dim, 1);
    for (int d = 0; d < dim; ++d) {
        if (n[d] < 1) throw std::invalid_argument("cell counts must be positive");
        nn[d] = n[d] + (periodic[d] ? 0 : 1);
        if (d > 0) stride[d] = stride[d - 1] * nn[d - 1];
    }
    const int ncorner = 1 << dim;
    // Simplex split of one cube (local corner bit d = +1 step along d) and the
    // quarter-turn relabelling applied between neighbouring cubes, as the
    // Preprocessor's StructuredReader generates them.
    std::vector<std::vector<int>> split;
    if (simplices) {
        if (dim == 2) split = {{0, 1, 2}, {1, 3, 2}};
        else split = {{1, 2, 4, 7}, {0, 1, 2, 4}, {1, 3, 2, 7}, {1, 4, 5, 7}, {2, 6, 4, 7}};
    }
    static const int turn[8] = {2, 0, 3, 1, 6, 4, 7, 5};
    m.shape = simplices ? (dim == 3 ? SHAPE_TETRAHEDRON : 0) : (dim == 3 ? SHAPE_CUBE : SHAPE_SQUARE);
    if (m.shape == 0) throw std::invalid_argument("2D simplex meshes are not supported");
    m.nv = simplices ? dim + 1 : ncorner;
    int64_t cells = 1;
    for (int d = 0; d < dim; ++d) cells *= n[d];
    std::vector<int> cell(dim, 0);
    std::vector<int64_t> cnode(ncorner);
    std::vector<double> cx(size_t(ncorner) * dim);
    for (int64_t c = 0; c < cells; ++c) {
        for (int j = 0; j < ncorner; ++j) {
            int64_t id = 0;
            for (int d = 0; d < dim; ++d) {
                int64_t i = cell[d] + ((j >> d) & 1);
                cx[j * dim + d] = lo[d] + double(i) * (hi[d] - lo[d]) / n[d];
                if (i == nn[d]) i = 0;  // periodic image of the first node layer
                id += i * stride[d];
            }
            cnode[j] = id;
        }
        const size_t pieces = simplices ? split.size() : 1;
        for (size_t s = 0; s < pieces; ++s) {
            if (simplices) {
                for (int v : split[s]) {
                    m.nodes.push_back(cnode[v]);
                    for (int d = 0; d < dim; ++d) m.coords.push_back(cx[v * dim + d]);
                }
            } else {
                for (int v = 0; v < ncorner; ++v) {
                    m.nodes.push_back(cnode[v]);
                    for (int d = 0; d < dim; ++d) m.coords.push_back(cx[v * dim + d]);
                }
            }
        }
        // advance the cell counter; every step relabels the split by a
        // quarter turn, and wrapping an odd-length row adds one more
        for (int d = 0; d < dim; ++d) {
            for (auto& tet : split)
                for (int& v : tet) v = turn[v];
            if (++cell[d] < n[d]) break;
            if (cell[d] % 2 == 1)
                for (auto& tet : split)
                    for (int& v : tet) v = turn[v];
            cell[d] = 0;
        }
    }
    return m;
}

}  // namespace hgdg
