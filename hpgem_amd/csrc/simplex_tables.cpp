#include "simplex_tables.hpp"

#include <cmath>
#include <stdexcept>

namespace hgdg {

const int kTetFaceNodes[4][3] = {{0, 3, 2}, {0, 1, 3}, {0, 2, 1}, {1, 2, 3}};
const int kTetEdgeNodes[6][2] = {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {1, 3}, {1, 2}};

namespace {

// P_n, P_n', P_n'' by the differentiated three-term recurrence.
void legendre3(int n, long double x, long double& p, long double& dp, long double& ddp) {
    long double p0 = 1, d0 = 0, dd0 = 0;
    if (n == 0) { p = 1; dp = 0; ddp = 0; return; }
    long double p1 = x, d1 = 1, dd1 = 0;
    for (int k = 2; k <= n; ++k) {
        const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        const long double d2 = ((2 * k - 1) * (p1 + x * d1) - (k - 1) * d0) / k;
        const long double dd2 = ((2 * k - 1) * (2 * d1 + x * dd1) - (k - 1) * dd0) / k;
        p0 = p1; d0 = d1; dd0 = dd1;
        p1 = p2; d1 = d2; dd1 = dd2;
    }
    p = p1; dp = d1; ddp = dd1;
}

// hpGEM's Lobatto kernel L_k (kernel/FE/helperFunctions.cpp:49-108): with the
// bubble (1-x^2)/4 L_k(x) equal to the normalised integrated Legendre
// polynomial, L_k(x) = -4 / ((k+1)(k+2)) sqrt((2k+3)/2) P'_{k+1}(x).
void lobatto(int k, long double x, long double& v, long double& dv) {
    long double p, dp, ddp;
    legendre3(k + 1, x, p, dp, ddp);
    const long double c = -4.0L / ((k + 1) * (k + 2)) * std::sqrt((2.0L * k + 3) / 2);
    v = c * dp;
    dv = c * ddp;
}

struct Term {
    int nl = 0;
    int lam[4] = {0, 0, 0, 0};
    int nk = 0;
    int kern[3][3] = {{0}};  // (degree, node a, node b): L_deg(lambda_a - lambda_b)
};

std::vector<Term> tet_terms(int p) {
    std::vector<Term> t;
    for (int i = 0; i < 4; ++i) {
        Term v;
        v.nl = 1;
        v.lam[0] = i;
        t.push_back(v);
    }
    for (int j = 0; j + 2 <= p; ++j) {
        for (int i = 0; i < 6; ++i) {
            Term e;
            e.nl = 2;
            e.lam[0] = kTetEdgeNodes[i][0];
            e.lam[1] = kTetEdgeNodes[i][1];
            e.nk = 1;
            e.kern[0][0] = j;
            e.kern[0][1] = kTetEdgeNodes[i][0];
            e.kern[0][2] = kTetEdgeNodes[i][1];
            t.push_back(e);
        }
        if (j > 0)
            for (int i = 0; i < 4; ++i) {
                const int* n = kTetFaceNodes[i];
                for (int k = 0; k + 1 <= j; ++k) {
                    Term f;
                    f.nl = 3;
                    f.lam[0] = n[0];
                    f.lam[1] = n[1];
                    f.lam[2] = n[2];
                    f.nk = 2;
                    f.kern[0][0] = j - k - 1; f.kern[0][1] = n[0]; f.kern[0][2] = n[1];
                    f.kern[1][0] = k;         f.kern[1][1] = n[1]; f.kern[1][2] = n[2];
                    t.push_back(f);
                }
            }
        if (j > 1)
            for (int i = 0; i + 2 <= j; ++i)
                for (int k = 0; i + k + 2 <= j; ++k) {
                    Term b;
                    b.nl = 4;
                    for (int q = 0; q < 4; ++q) b.lam[q] = q;
                    b.nk = 3;
                    b.kern[0][0] = i;             b.kern[0][1] = 0; b.kern[0][2] = 1;
                    b.kern[1][0] = j - i - k - 2; b.kern[1][1] = 1; b.kern[1][2] = 2;
                    b.kern[2][0] = k;             b.kern[2][1] = 2; b.kern[2][2] = 3;
                    t.push_back(b);
                }
    }
    return t;
}

}  // namespace

void tet_basis(int p, const long double* x, long double* val, long double* grad) {
    if (p < 1 || p > kMaxSimplexOrder) throw std::invalid_argument("tetrahedral order must be in 1..4");
    const long double lam[4] = {1 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
    static const long double dlam[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const std::vector<Term> terms = tet_terms(p);
    for (size_t i = 0; i < terms.size(); ++i) {
        const Term& t = terms[i];
        long double f[7], df[7][3];
        int nf = 0;
        for (int q = 0; q < t.nl; ++q, ++nf) {
            f[nf] = lam[t.lam[q]];
            for (int d = 0; d < 3; ++d) df[nf][d] = dlam[t.lam[q]][d];
        }
        for (int q = 0; q < t.nk; ++q, ++nf) {
            const int a = t.kern[q][1], b = t.kern[q][2];
            long double v, dv;
            lobatto(t.kern[q][0], lam[a] - lam[b], v, dv);
            f[nf] = v;
            for (int d = 0; d < 3; ++d) df[nf][d] = dv * (dlam[a][d] - dlam[b][d]);
        }
        long double v = 1;
        for (int q = 0; q < nf; ++q) v *= f[q];
        val[i] = v;
        for (int d = 0; d < 3; ++d) {
            long double g = 0;
            for (int q = 0; q < nf; ++q) {
                long double term = df[q][d];
                for (int r = 0; r < nf; ++r)
                    if (r != q) term *= f[r];
                g += term;
            }
            grad[i * 3 + d] = g;
        }
    }
}

void tet_face_map(int f, const long double* t, long double* x) {
    static const long double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int* n = kTetFaceNodes[f];
    for (int d = 0; d < 3; ++d) x[d] = V[n[0]][d] + t[0] * (V[n[1]][d] - V[n[0]][d]) + t[1] * (V[n[2]][d] - V[n[0]][d]);
}

void tri_orientation_map(int o, const long double* t, long double* out) {
    const long double x = t[0], y = t[1];
    switch (o) {
        case 0: out[0] = x; out[1] = y; return;
        case 1: out[0] = y; out[1] = x; return;
        case 2: out[0] = 1 - x - y; out[1] = x; return;
        case 3: out[0] = 1 - x - y; out[1] = y; return;
        case 4: out[0] = x; out[1] = 1 - x - y; return;
        case 5: out[0] = y; out[1] = 1 - x - y; return;
    }
    throw std::invalid_argument("triangle orientation index out of range");
}

int tri_orientation_index(const long long* l, const long long* r) {
    if (l[0] == r[0]) return l[1] == r[1] ? 0 : 1;
    if (l[0] == r[1]) return l[1] == r[2] ? 2 : 3;
    return l[1] == r[1] ? 4 : 5;
}

namespace {

void gauss01(int n, std::vector<long double>& x, std::vector<long double>& w) {
    // Gauss-Legendre on [0,1] by Newton iteration on P_n
    x.assign(n, 0);
    w.assign(n, 0);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; ++i) {
        long double t = -std::cos(pi * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 80; ++it) {
            long double p, dp, ddp;
            legendre3(n, t, p, dp, ddp);
            const long double dt = p / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-20L) break;
        }
        long double p, dp, ddp;
        legendre3(n, t, p, dp, ddp);
        x[i] = (t + 1) / 2;
        w[i] = 1.0L / ((1 - t * t) * dp * dp);  // (2 / ((1-t^2) P'^2)) / 2
    }
}

void tri_rule(int k, std::vector<long double>& pts, std::vector<long double>& w) {
    const int n = (k + 3) / 2;
    std::vector<long double> x, wx;
    gauss01(n, x, wx);
    pts.clear();
    w.clear();
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            pts.push_back(x[a]);
            pts.push_back((1 - x[a]) * x[b]);
            w.push_back(wx[a] * wx[b] * (1 - x[a]));
        }
}

std::vector<long double> invert(int n, const std::vector<long double>& A) {
    std::vector<long double> M(A), I(size_t(n) * n, 0.0L);
    for (int i = 0; i < n; ++i) I[size_t(i) * n + i] = 1;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[size_t(i) * n + k]) > std::fabs(M[size_t(piv) * n + k])) piv = i;
        for (int j = 0; j < n; ++j) {
            std::swap(M[size_t(k) * n + j], M[size_t(piv) * n + j]);
            std::swap(I[size_t(k) * n + j], I[size_t(piv) * n + j]);
        }
        const long double d = M[size_t(k) * n + k];
        for (int j = 0; j < n; ++j) {
            M[size_t(k) * n + j] /= d;
            I[size_t(k) * n + j] /= d;
        }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const long double f = M[size_t(i) * n + k];
            if (f == 0) continue;
            for (int j = 0; j < n; ++j) {
                M[size_t(i) * n + j] -= f * M[size_t(k) * n + j];
                I[size_t(i) * n + j] -= f * I[size_t(k) * n + j];
            }
        }
    }
    return I;
}

// Orthonormal basis of P_4 on the reference triangle (Gram-Schmidt of the
// monomials x^i y^j in degree order), evaluated at a point.
struct FaceBasis {
    long double coef[kFaceBasis][kFaceBasis];  // chi_m = sum_q coef[m][q] mono_q
    int ex[kFaceBasis], ey[kFaceBasis];
    FaceBasis() {
        int q = 0;
        for (int d = 0; d <= 4; ++d)
            for (int j = 0; j <= d; ++j) {
                ex[q] = d - j;
                ey[q] = j;
                ++q;
            }
        std::vector<long double> pts, w;
        tri_rule(8, pts, w);
        auto mono = [&](int m, long double x, long double y) { return std::pow(x, ex[m]) * std::pow(y, ey[m]); };
        auto inner = [&](const long double* a, const long double* b) {
            long double s = 0;
            for (size_t r = 0; r < w.size(); ++r) {
                long double va = 0, vb = 0;
                for (int m = 0; m < kFaceBasis; ++m) {
                    va += a[m] * mono(m, pts[2 * r], pts[2 * r + 1]);
                    vb += b[m] * mono(m, pts[2 * r], pts[2 * r + 1]);
                }
                s += w[r] * va * vb;
            }
            return s;
        };
        for (int m = 0; m < kFaceBasis; ++m) {
            long double v[kFaceBasis] = {0};
            v[m] = 1;
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < m; ++k) {
                    const long double c = inner(v, coef[k]);
                    for (int r = 0; r < kFaceBasis; ++r) v[r] -= c * coef[k][r];
                }
            const long double nrm = std::sqrt(inner(v, v));
            for (int r = 0; r < kFaceBasis; ++r) coef[m][r] = v[r] / nrm;
        }
    }
    void eval(long double x, long double y, long double* out) const {
        long double mono[kFaceBasis];
        for (int m = 0; m < kFaceBasis; ++m) mono[m] = std::pow(x, ex[m]) * std::pow(y, ey[m]);
        for (int m = 0; m < kFaceBasis; ++m) {
            long double s = 0;
            for (int r = 0; r < kFaceBasis; ++r) s += coef[m][r] * mono[r];
            out[m] = s;
        }
    }
};

}  // namespace

void tet_rule(int k, std::vector<long double>& pts, std::vector<long double>& w) {
    const int n = (k + 4) / 2;
    std::vector<long double> x, wx;
    gauss01(n, x, wx);
    pts.clear();
    w.clear();
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            for (int c = 0; c < n; ++c) {
                const long double u = x[a], v = x[b], t = x[c];
                pts.push_back(u);
                pts.push_back((1 - u) * v);
                pts.push_back((1 - u) * (1 - v) * t);
                w.push_back(wx[a] * wx[b] * wx[c] * (1 - u) * (1 - u) * (1 - v));
            }
}

namespace {

// Barycentric coordinate lambda_v of the reference tetrahedron and its
// (constant) gradient: lambda_0 = 1 - x - y - z, lambda_k = x_{k-1}.
long double bary(int v, const long double* x) { return v == 0 ? 1 - x[0] - x[1] - x[2] : x[v - 1]; }
void bary_grad(int v, long double* g) {
    for (int d = 0; d < 3; ++d) g[d] = v == 0 ? -1.0L : (d == v - 1 ? 1.0L : 0.0L);
}
void cross(const long double* a, const long double* b, long double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Edge function of degree k on edge (o, i): with phi0 = l_i grad l_o - l_o grad l_i
// (Whitney) and phi1 = l_i grad l_o + l_o grad l_i = grad(l_o l_i),
// k = 0: phi0, k = 1: phi1, k >= 2: -(k-1)/k P_{k-2}(s) phi0 + (2k-1)/k P_{k-1}(s) phi1, s = l_i - l_o.
void ac_edge(int k, int o, int i, const long double* x, long double* val, long double* curl) {
    long double go[3], gi[3], phi0[3], phi1[3], gio[3];
    bary_grad(o, go);
    bary_grad(i, gi);
    const long double lo = bary(o, x), li = bary(i, x);
    for (int d = 0; d < 3; ++d) {
        phi0[d] = li * go[d] - lo * gi[d];
        phi1[d] = li * go[d] + lo * gi[d];
    }
    cross(gi, go, gio);  // curl phi0 = 2 grad l_i x grad l_o, curl phi1 = 0
    if (k == 0) {
        for (int d = 0; d < 3; ++d) {
            val[d] = phi0[d];
            curl[d] = 2 * gio[d];
        }
        return;
    }
    if (k == 1) {
        for (int d = 0; d < 3; ++d) {
            val[d] = phi1[d];
            curl[d] = 0;
        }
        return;
    }
    long double pa, dpa, dda, pb, dpb, ddb;
    legendre3(k - 2, li - lo, pa, dpa, dda);
    legendre3(k - 1, li - lo, pb, dpb, ddb);
    const long double ca = -(k - 1.0L) / k, cb = (2.0L * k - 1) / k;
    long double gs[3], x0[3], x1[3];
    for (int d = 0; d < 3; ++d) gs[d] = gi[d] - go[d];
    cross(gs, phi0, x0);
    cross(gs, phi1, x1);
    // curl(f phi) = grad f x phi + f curl phi
    for (int d = 0; d < 3; ++d) {
        val[d] = ca * pa * phi0[d] + cb * pb * phi1[d];
        curl[d] = ca * (dpa * x0[d] + pa * 2 * gio[d]) + cb * dpb * x1[d];
    }
}

// Edge-based face function of degree k >= 2 on the face opposite vertex v,
// attached to vertex c: with (a, b) the face's other two vertices (a < b),
// grad l_c l_a l_b P_{k-2}(l_b - l_a).
void ac_edge_face(int k, int v, int c, const long double* x, long double* val, long double* curl) {
    int a = 0;
    while (a == c || a == v) ++a;
    int b = a + 1;
    while (b == c || b == v) ++b;
    long double ga[3], gb[3], gc[3];
    bary_grad(a, ga);
    bary_grad(b, gb);
    bary_grad(c, gc);
    const long double la = bary(a, x), lb = bary(b, x);
    long double P, dP, ddP;
    legendre3(k - 2, lb - la, P, dP, ddP);
    long double ac_[3], bc[3];
    cross(ga, gc, ac_);
    cross(gb, gc, bc);
    // f = l_a l_b P(l_b - l_a): grad f = (l_b P - l_a l_b P') grad l_a + (l_a P + l_a l_b P') grad l_b
    const long double fa = lb * P - la * lb * dP, fb = la * P + la * lb * dP;
    for (int d = 0; d < 3; ++d) {
        val[d] = gc[d] * la * lb * P;
        curl[d] = fa * ac_[d] + fb * bc[d];
    }
}

}  // namespace

int hcurl_tet_size(int p) { return p < 1 ? 0 : 6 * (p + 1) + (p >= 2 ? 12 * (p - 1) : 0); }

void hcurl_tet_basis(int p, const long double* x, long double* val, long double* curl) {
    if (p < 1 || p > 2) throw std::invalid_argument("the H(curl) tetrahedral set is built for orders 1 and 2");
    static const int edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {2, 3}, {1, 3}, {1, 2}};
    static const int faces[12][2] = {{3, 0}, {3, 1}, {3, 2}, {2, 0}, {2, 1}, {2, 3},
                                     {1, 0}, {1, 2}, {1, 3}, {0, 1}, {0, 2}, {0, 3}};
    int n = 0;
    for (int k = 0; k <= p; ++k)
        for (const auto& ed : edges) {
            ac_edge(k, ed[0], ed[1], x, val + 3 * n, curl + 3 * n);
            ++n;
        }
    for (int k = 2; k <= p; ++k)
        for (const auto& fc : faces) {
            ac_edge_face(k, fc[0], fc[1], x, val + 3 * n, curl + 3 * n);
            ++n;
        }
}

SimplexOps simplex_ops(int p) {
    if (p < 1 || p > kMaxSimplexOrder) throw std::invalid_argument("tetrahedral order must be in 1..4");
    SimplexOps ops;
    ops.p = p;
    const int nb = tet_nb(p);
    ops.nb = nb;
    std::vector<long double> pts, w;
    tet_rule(2 * p, pts, w);
    std::vector<long double> M(size_t(nb) * nb, 0), K(3 * size_t(nb) * nb, 0);
    std::vector<long double> val(nb), grad(size_t(nb) * 3);
    for (size_t q = 0; q < w.size(); ++q) {
        tet_basis(p, &pts[3 * q], val.data(), grad.data());
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                M[size_t(i) * nb + j] += w[q] * val[i] * val[j];
                for (int e = 0; e < 3; ++e) K[(size_t(e) * nb + i) * nb + j] += w[q] * val[i] * grad[j * 3 + e];
            }
    }
    const std::vector<long double> Mi = invert(nb, M);
    ops.Mref.assign(M.begin(), M.end());
    ops.Minv.assign(Mi.begin(), Mi.end());
    ops.Dr.assign(3 * size_t(nb) * nb, 0.0);
    for (int e = 0; e < 3; ++e)
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                long double s = 0;
                for (int c = 0; c < nb; ++c) s += Mi[size_t(i) * nb + c] * K[(size_t(e) * nb + c) * nb + j];
                ops.Dr[(size_t(e) * nb + i) * nb + j] = double(s);
            }
    // face tables
    static const FaceBasis fb;
    std::vector<long double> tp, tw;
    tri_rule(2 * kMaxSimplexOrder, tp, tw);
    for (int f = 0; f < 4; ++f) {
        // columns whose trace on face f does not vanish
        std::vector<long double> nrm(nb, 0);
        for (size_t q = 0; q < tw.size(); ++q) {
            long double x[3];
            tet_face_map(f, &tp[2 * q], x);
            tet_basis(p, x, val.data(), grad.data());
            for (int j = 0; j < nb; ++j) nrm[j] += tw[q] * val[j] * val[j];
        }
        ops.cols[f].clear();
        for (int j = 0; j < nb; ++j)
            if (nrm[j] > 1e-24L) ops.cols[f].push_back(j);
        const int nc = int(ops.cols[f].size());
        if (nc > kFaceBasis) throw std::logic_error("face trace space larger than the face basis");
        for (int o = 0; o < 6; ++o) {
            std::vector<long double> P(size_t(kFaceBasis) * nc, 0);
            for (size_t q = 0; q < tw.size(); ++q) {
                long double to[2], x[3], chi[kFaceBasis];
                tri_orientation_map(o, &tp[2 * q], to);
                tet_face_map(f, to, x);
                tet_basis(p, x, val.data(), grad.data());
                fb.eval(tp[2 * q], tp[2 * q + 1], chi);
                for (int m = 0; m < kFaceBasis; ++m)
                    for (int c = 0; c < nc; ++c) P[size_t(m) * nc + c] += tw[q] * chi[m] * val[ops.cols[f][c]];
            }
            ops.proj[f][o].assign(P.begin(), P.end());
            // lift = M^{-1} P^T restricted to the trace columns
            std::vector<double> L(size_t(nb) * kFaceBasis, 0.0);
            for (int i = 0; i < nb; ++i)
                for (int m = 0; m < kFaceBasis; ++m) {
                    long double s = 0;
                    for (int c = 0; c < nc; ++c) s += Mi[size_t(i) * nb + ops.cols[f][c]] * P[size_t(m) * nc + c];
                    L[size_t(i) * kFaceBasis + m] = double(s);
                }
            ops.lift[f][o] = L;
        }
    }
    return ops;
}

std::vector<double> face_orientation_blocks() {
    // R_o(m, m') = int chi_m(t) chi_m'(g_o(t)) dt over the reference triangle
    // (rule of degree 8: exact). g_o is affine and maps the triangle onto
    // itself with |det| = 1, so f -> f o g_o is orthogonal in L2 and keeps
    // every P_d: R_o is block diagonal by degree (checked here).
    static const FaceBasis fb;
    std::vector<long double> tp, tw;
    tri_rule(2 * kMaxSimplexOrder, tp, tw);
    std::vector<double> out;
    out.reserve(6 * size_t(kFaceRotBlock));
    for (int o = 0; o < 6; ++o) {
        long double R[kFaceBasis][kFaceBasis] = {};
        for (size_t q = 0; q < tw.size(); ++q) {
            long double to[2], a[kFaceBasis], b[kFaceBasis];
            tri_orientation_map(o, &tp[2 * q], to);
            fb.eval(tp[2 * q], tp[2 * q + 1], a);
            fb.eval(to[0], to[1], b);
            for (int m = 0; m < kFaceBasis; ++m)
                for (int n = 0; n < kFaceBasis; ++n) R[m][n] += tw[q] * a[m] * b[n];
        }
        auto degree = [](int m) {
            int d = 0;
            while ((d + 1) * (d + 2) / 2 <= m) ++d;
            return d;
        };
        for (int m = 0; m < kFaceBasis; ++m)
            for (int n = 0; n < kFaceBasis; ++n)
                if (degree(m) != degree(n) && fabsl(R[m][n]) > 1e-15L)
                    throw std::logic_error("face orientation map mixes polynomial degrees");
        for (int d = 0; d <= kMaxSimplexOrder; ++d) {
            const int m0 = d * (d + 1) / 2;
            for (int a = 0; a <= d; ++a)
                for (int b = 0; b <= d; ++b) out.push_back(double(R[m0 + a][m0 + b]));
        }
    }
    return out;
}

}  // namespace hgdg
