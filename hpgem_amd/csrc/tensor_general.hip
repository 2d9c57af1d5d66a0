// Quadrilaterals and hexahedra that the sum-factorised tensor kernels of
// tensor.hip do not take: elements whose map is not affine (the bilinear /
// trilinear map of MappingToPhysHypercubeLinear, kernel/Geometry/Mappings/
// MappingToPhysHypercubeLinear.cpp:222-304) and neighbours that meet across a
// face rotated against them or through another local face (FaceGeometry.cpp:
// 148-155, the face maps of ReferenceSquare/ReferenceCube).
//
// Gauss-point form. hpGEM integrates an order-p element with its rule of
// order 2p+1, the (p+1)-point Gauss rule per direction, so with B = B1 (x) B1
// (x) B1, B1(q,a) = psi_a(x_q) square and invertible, the mass matrix it
// factorises on every call is M_e = B^T W B (W = diag(w_q |det J(x_q)|)) and
// has the exact inverse B^{-1} W^{-1} B^{-T}. With U = B u the point values:
//   advection  k = B^{-1} W^{-1} [ sum_r C_r^T (W ahat_r U) + sum_f lam_f (x) c_f ],
//              ahat = J^{-1} a at each point, C_r = C along direction r
//              (C = derivative of the interpolant of point values);
//   acoustic   k_v = B^{-1} [ div s + W^{-1} sum_f lam_f (x) c_f^v ],
//              k_s = B^{-1} [ grad v + W^{-1} sum_f lam_f (x) c_f^s ] / c_inv,
// where c_f holds, at the face's Gauss points, the flux integrand times the
// face weight and area element, and lam_f = B1^{-T} e_side (Lagrange
// polynomials of the points at +-1) lifts it to the element's points. This
// is hpGEM's operator with hpGEM's quadrature (element and face integrals of
// HpgemAPILinear::createStiffnessMatrices with the integrands of
// 080AdvectionLinear_SelfTest.cpp:108-170 and AcousticWaveLinear_Impl.h:
// 102-283, mass solve of HpgemAPILinear_Impl.h:143-156), evaluated in another
// order: results agree with it to rounding.
//
// Faces: hpGEM evaluates every face integrand at the face rule's points p
// (face coordinates of the left element L, the lower-numbered one), takes
// L's trace at L's point of p and the right element's trace at R's point of
// map(p), where map is the square's orientation map that the reference picks
// from the two faces' node lists; normal and area element come from L's map at
// p. Each element here loops over its own face points and pairs them the same
// way; for faces where it is R, it evaluates L's normal from L's vertices at
// L's point, so both sides use the same normal, as in hpGEM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <unordered_map>
#include <vector>

#include "context.hpp"

namespace hgdg {

namespace {

constexpr int kGenTarget = 256;  // threads per workgroup aimed at

__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int DIM, int N1>
struct GenShape {
    static constexpr int NP = ipow(N1, DIM);       // points = coefficients per field
    static constexpr int NFP = ipow(N1, DIM - 1);  // points per face
    static constexpr int NV = 1 << DIM;
    static constexpr int NFACE = 2 * DIM;
    static constexpr int EPB = NP >= kGenTarget ? 1 : kGenTarget / NP;
    static constexpr int THREADS = EPB * NP;
};

// Offsets in the packed NodalOps table of order N1 - 1.
template <int N1>
struct NodalTab {
    static constexpr int X = 0, W = N1, B = 2 * N1, BINV = B + N1 * N1, C = BINV + N1 * N1, LAM0 = C + N1 * N1,
                         LAM1 = LAM0 + N1, X2 = LAM1 + N1, W2 = X2 + N1 + 1, B2 = W2 + N1 + 1,
                         SIZE = B2 + N1 * (N1 + 1);
};

// Local faces of the reference square (0: y = -1, 1: x = -1, 2: x = 1, 3: y = 1)
// and cube (0: z = -1, 1: y = -1, 2: x = -1, 3: x = 1, 4: y = 1, 5: z = 1); a
// face's coordinates are the element coordinates other than its axis, in
// increasing order (MappingToRefLineToSquare, MappingToRefSquareToCube).
__host__ __device__ constexpr int face_axis(int dim, int f) {
    return dim == 2 ? ((f == 1 || f == 2) ? 0 : 1) : ((f == 2 || f == 3) ? 0 : (f == 1 || f == 4) ? 1 : 2);
}
__host__ __device__ constexpr int face_side(int dim, int f) { return dim == 2 ? (f >= 2 ? 1 : 0) : (f >= 3 ? 1 : 0); }
// the k-th face coordinate axis (k = 0, 1)
__host__ __device__ constexpr int face_coord_axis(int dim, int f, int k) {
    return face_axis(dim, f) <= k ? k + 1 : k;
}
// Sign that turns the wedge product of the face tangents into the outward
// normal for a positively oriented element (the same sign per face in both
// shapes: odd faces +1, even faces -1).
__host__ __device__ constexpr double face_sign(int f) { return (f & 1) ? 1.0 : -1.0; }

// Point x(xi) and Jacobian J[r][d] = dx_r / dxi_d of the multilinear map of
// the element with vertices X[v][r] (vertex v at the corner whose coordinate
// d is -1 / +1 by bit d of v).
template <int DIM>
__device__ inline void multilinear(const double* X, const double* xi, double (&J)[DIM][DIM], double* x) {
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
        if (x) x[r] = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) J[r][d] = 0;
    }
#pragma unroll
    for (int v = 0; v < (1 << DIM); ++v) {
        double f[DIM], g[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const bool hi = (v >> d) & 1;
            f[d] = hi ? 0.5 * (1 + xi[d]) : 0.5 * (1 - xi[d]);
            g[d] = hi ? 0.5 : -0.5;
        }
        double prod = 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) prod *= f[d];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double dd = g[d];
#pragma unroll
            for (int k = 0; k < DIM; ++k)
                if (k != d) dd *= f[k];
#pragma unroll
            for (int r = 0; r < DIM; ++r) J[r][d] += X[v * DIM + r] * dd;
        }
        if (x) {
#pragma unroll
            for (int r = 0; r < DIM; ++r) x[r] += X[v * DIM + r] * prod;
        }
    }
}

template <int DIM>
__device__ inline double det_inv(const double (&J)[DIM][DIM], double (&Ji)[DIM][DIM]) {
    if constexpr (DIM == 2) {
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Ji[0][0] = J[1][1] / det;
        Ji[0][1] = -J[0][1] / det;
        Ji[1][0] = -J[1][0] / det;
        Ji[1][1] = J[0][0] / det;
        return det;
    } else {
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                Ji[r][c] = (J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1]) / det;
            }
        return det;
    }
}

// Outward normal scaled by the area element at face point (qa, qb) of face f
// of the element with vertices X (hpGEM's PhysicalFace normal: wedge product
// of the mapped face tangents, signed per face and by the orientation of the
// element map).
template <int DIM, int N1>
__device__ inline void face_normal(const double* X, int f, int qa, int qb, const double* tab, double (&n)[DIM]) {
    double xi[DIM];
    xi[face_axis(DIM, f)] = face_side(DIM, f) ? 1.0 : -1.0;
    xi[face_coord_axis(DIM, f, 0)] = tab[NodalTab<N1>::X + qa];
    if constexpr (DIM == 3) xi[face_coord_axis(DIM, f, 1)] = tab[NodalTab<N1>::X + qb];
    double J[DIM][DIM], Ji[DIM][DIM];
    multilinear<DIM>(X, xi, J, nullptr);
    const double det = det_inv<DIM>(J, Ji);
    const double s = face_sign(f) * (det > 0 ? 1.0 : -1.0);
    const int a0 = face_coord_axis(DIM, f, 0);
    if constexpr (DIM == 2) {
        n[0] = -J[1][a0] * s;
        n[1] = J[0][a0] * s;
    } else {
        const int a1 = face_coord_axis(DIM, f, 1);
        n[0] = (J[1][a0] * J[2][a1] - J[2][a0] * J[1][a1]) * s;
        n[1] = (J[2][a0] * J[0][a1] - J[0][a0] * J[2][a1]) * s;
        n[2] = (J[0][a0] * J[1][a1] - J[1][a0] * J[0][a1]) * s;
    }
}

// sum_m M(i, m) in[t with its index along `ax` set to m], i = t's index along
// ax, M row-major N1 x N1 (or its transpose); `along` does it for NF fields of
// NP values each, into out[f][t].
template <int N1, bool TRANSPOSE>
__device__ inline double along_val(const double* M, int ax, int t, const double* in) {
    const int st = ax == 0 ? 1 : ax == 1 ? N1 : N1 * N1;
    const int i = (t / st) % N1;
    const int t0 = t - i * st;
    double s = 0;
#pragma unroll
    for (int m = 0; m < N1; ++m) s += (TRANSPOSE ? M[m * N1 + i] : M[i * N1 + m]) * in[t0 + m * st];
    return s;
}

template <int N1, int NP, int NF, bool TRANSPOSE>
__device__ inline void along(const double* M, int ax, int t, const double* in, double* out) {
#pragma unroll
    for (int f = 0; f < NF; ++f) out[f * NP + t] = along_val<N1, TRANSPOSE>(M, ax, t, in + f * NP);
}

// Per (element, face): the neighbour (or -1) and packed data: bits 0-2 the
// neighbour's local face, bits 3-5 the map of this element's face point
// indices onto the neighbour's (bit 3: swap the two face coordinates first,
// bits 4/5: reverse the first / second), bit 6: this element is the face's
// left element.
__host__ __device__ __forceinline__ void partner_point(int code, int ka, int kb, int n1, int& pa, int& pb) {
    pa = (code & 1) ? kb : ka;
    pb = (code & 1) ? ka : kb;
    if (code & 2) pa = n1 - 1 - pa;
    if (code & 4) pb = n1 - 1 - pb;
}

struct GenParams {
    int64_t n_elem;
    StageIO io;
    const double* X;       // [E][NV][DIM] vertex coordinates
    const int32_t* nbr;    // [NFACE][E]
    const int32_t* info;   // [NFACE][E]
    const double* tab;     // NodalOps of the order, packed
    const double* cinv;    // [E] (acoustic)
    double a[3];           // advection velocity
};

template <int DIM, int N1, bool ACOUSTIC>
struct GenLds {
    using S = GenShape<DIM, N1>;
    static constexpr int NF = ACOUSTIC ? DIM + 1 : 1;
    // per element, in doubles
    static constexpr int XV = 0;                                    // [NV][DIM] own vertices
    static constexpr int XN = XV + S::NV * DIM;                     // [NFACE][NV][DIM] left neighbours' vertices
    static constexpr int U = XN + S::NFACE * S::NV * DIM;           // [NF][NP] stage input (coefficients)
    static constexpr int A = U + NF * S::NP;                        // [NF][NP] work
    static constexpr int B = A + NF * S::NP;                        // [NF][NP] work
    static constexpr int F = B + NF * S::NP;                        // [DIM][NP] advection fluxes at the points
    static constexpr int NB = F + (ACOUSTIC ? 0 : DIM * S::NP);     // [NFACE][NF][NFP] neighbours' face coefficients
    static constexpr int FC = NB + S::NFACE * NF * S::NFP;          // [NFACE][NF][NFP] face integrands
    static constexpr int PER = FC + S::NFACE * NF * S::NFP;
};

template <int DIM, int N1, bool ACOUSTIC>
__global__ __launch_bounds__(kGenTarget) void tensor_general_rhs(const GenParams p) {
    using S = GenShape<DIM, N1>;
    using L = GenLds<DIM, N1, ACOUSTIC>;
    using T = NodalTab<N1>;
    constexpr int NP = S::NP, NFP = S::NFP, NV = S::NV, NFACE = S::NFACE, NF = L::NF;
    __shared__ double lds[S::EPB * L::PER];
    __shared__ double tab[T::SIZE];
    __shared__ int32_t s_nbr[S::EPB][NFACE], s_info[S::EPB][NFACE];
    const int tid = threadIdx.x;
    const int le = tid / NP;
    const int t = tid % NP;
    const int64_t e = int64_t(blockIdx.x) * S::EPB + le;
    const bool ok = e < p.n_elem;
    const int64_t E = p.n_elem;
    double* sX = lds + le * L::PER + L::XV;
    double* sXn = lds + le * L::PER + L::XN;
    double* su = lds + le * L::PER + L::U;
    double* sA = lds + le * L::PER + L::A;
    double* sB = lds + le * L::PER + L::B;
    double* sF = lds + le * L::PER + L::F;
    double* snb = lds + le * L::PER + L::NB;
    double* sc = lds + le * L::PER + L::FC;
    for (int k = tid; k < T::SIZE; k += blockDim.x) tab[k] = p.tab[k];

    // ---- stage input, own vertices, faces; neighbours' face coefficients and,
    // where this element is the right one, the left neighbour's vertices
    if (ok) {
        for (int k = t; k < NV * DIM; k += NP) sX[k] = p.X[e * NV * DIM + k];  // NP may be below NV * DIM
        for (int k = t; k < NFACE; k += NP) {
            s_nbr[le][k] = p.nbr[k * E + e];
            s_info[le][k] = p.info[k * E + e];
        }
        const int64_t base = e * NF * NP;
#pragma unroll
        for (int fld = 0; fld < NF; ++fld) su[fld * NP + t] = combine(p.io, base + fld * NP + t);
        constexpr int PERF = NV * DIM + NF * NFP;
        for (int it = t; it < NFACE * PERF; it += NP) {
            const int f = it / PERF;
            int r = it % PERF;
            const int32_t nb = p.nbr[f * E + e];
            if (nb < 0) continue;
            const int32_t info = p.info[f * E + e];
            if (r < NV * DIM) {
                if (!((info >> 6) & 1)) sXn[f * NV * DIM + r] = p.X[int64_t(nb) * NV * DIM + r];
                continue;
            }
            r -= NV * DIM;
            const int fld = r / NFP, k = r % NFP;
            const int fn = info & 7;
            const int ka = k % N1, kb = k / N1;
            int idx = face_side(DIM, fn) * ipow(N1, face_axis(DIM, fn)) + ka * ipow(N1, face_coord_axis(DIM, fn, 0));
            if constexpr (DIM == 3) idx += kb * ipow(N1, face_coord_axis(DIM, fn, 1));
            snb[(f * NF + fld) * NFP + k] = combine(p.io, int64_t(nb) * NF * NP + fld * NP + idx);
        }
    }
    __syncthreads();

    // ---- point values U = B u (three passes through LDS) and the geometry at
    // this thread's point
    const int ii[3] = {t % N1, (t / N1) % N1, DIM == 3 ? t / (N1 * N1) : 0};
    double Wq, Ji[DIM][DIM];
    {
        double xi[DIM], J[DIM][DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = tab[T::X + ii[d]];
        multilinear<DIM>(sX, xi, J, nullptr);
        const double det = det_inv<DIM>(J, Ji);
        Wq = fabs(det);
#pragma unroll
        for (int d = 0; d < DIM; ++d) Wq *= tab[T::W + ii[d]];
    }
    along<N1, NP, NF, false>(tab + T::B, 0, t, su, sA);
    __syncthreads();
    along<N1, NP, NF, false>(tab + T::B, 1, t, sA, sB);
    __syncthreads();
    const double* Up = sB;
    if constexpr (DIM == 3) {
        along<N1, NP, NF, false>(tab + T::B, 2, t, sB, sA);
        __syncthreads();
        Up = sA;
    }

    // ---- volume terms, as point residuals (W times the integrand's point value)
    double res[NF];
    if constexpr (!ACOUSTIC) {
        double ah[DIM];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            ah[r] = 0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) ah[r] += Ji[r][d] * p.a[d];
        }
        const double u = Up[t];
#pragma unroll
        for (int r = 0; r < DIM; ++r) sF[r * NP + t] = Wq * ah[r] * u;
        __syncthreads();
        double v = 0;
#pragma unroll
        for (int r = 0; r < DIM; ++r) v += along_val<N1, true>(tab + T::C, r, t, sF + r * NP);
        res[0] = v;
    } else {
        // reference-coordinate derivatives of every field at this point
        double g[NF][DIM];
#pragma unroll
        for (int fld = 0; fld < NF; ++fld)
#pragma unroll
            for (int r = 0; r < DIM; ++r) g[fld][r] = along_val<N1, false>(tab + T::C, r, t, Up + fld * NP);
        double div = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int r = 0; r < DIM; ++r) div += Ji[r][d] * g[1 + d][r];
        res[0] = Wq * div;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            double gv = 0;
#pragma unroll
            for (int r = 0; r < DIM; ++r) gv += Ji[r][d] * g[0][r];
            res[1 + d] = Wq * gv;
        }
    }

    // ---- face integrands at the face points: items (face, point)
    if (ok) {
        for (int it = t; it < NFACE * NFP; it += NP) {
            const int f = it / NFP, k = it % NFP;
            const int ax = face_axis(DIM, f), side = face_side(DIM, f);
            const int a0 = face_coord_axis(DIM, f, 0), a1 = DIM == 3 ? face_coord_axis(DIM, f, 1) : 0;
            const int ka = k % N1, kb = DIM == 3 ? k / N1 : 0;
            const int stf = ipow(N1, 0) * (ax == 0 ? 1 : ax == 1 ? N1 : N1 * N1);
            const int st0 = a0 == 0 ? 1 : a0 == 1 ? N1 : N1 * N1;
            const int st1 = a1 == 0 ? 1 : a1 == 1 ? N1 : N1 * N1;
            // own trace: the coefficients with index `side` along ax, at (ka, kb)
            double own[NF], nbv[NF];
#pragma unroll
            for (int fld = 0; fld < NF; ++fld) {
                double s = 0;
#pragma unroll
                for (int b = 0; b < (DIM == 3 ? N1 : 1); ++b) {
                    double sa = 0;
#pragma unroll
                    for (int a = 0; a < N1; ++a) sa += tab[T::B + ka * N1 + a] * su[fld * NP + side * stf + a * st0 + b * st1];
                    s += DIM == 3 ? tab[T::B + kb * N1 + b] * sa : sa;
                }
                own[fld] = s;
            }
            const int32_t nb = s_nbr[le][f];
            const int32_t info = s_info[le][f];
            const bool inner = nb >= 0;
            const bool left = !inner || ((info >> 6) & 1);
            int pa = ka, pb = kb;
            if (inner) {
                partner_point((info >> 3) & 7, ka, kb, N1, pa, pb);
#pragma unroll
                for (int fld = 0; fld < NF; ++fld) {
                    const double* cf = snb + (f * NF + fld) * NFP;
                    double s = 0;
#pragma unroll
                    for (int b = 0; b < (DIM == 3 ? N1 : 1); ++b) {
                        double sa = 0;
#pragma unroll
                        for (int a = 0; a < N1; ++a) sa += tab[T::B + pa * N1 + a] * cf[a + N1 * b];
                        s += DIM == 3 ? tab[T::B + pb * N1 + b] * sa : sa;
                    }
                    nbv[fld] = s;
                }
            }
            // the left element's normal and area element at its point
            double n[DIM];
            int qa = ka, qb = kb;
            if (left) {
                face_normal<DIM, N1>(sX, f, ka, kb, tab, n);
            } else {
                qa = pa;
                qb = pb;
                face_normal<DIM, N1>(sXn + f * NV * DIM, info & 7, pa, pb, tab, n);
            }
            double area = 0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) area += n[d] * n[d];
            area = sqrt(area);
            double wf = tab[T::W + qa];
            if constexpr (DIM == 3) wf *= tab[T::W + qb];
            wf *= area;
            double nu[DIM];  // the left element's unit normal
#pragma unroll
            for (int d = 0; d < DIM; ++d) nu[d] = n[d] / area;
            const double sgn = left ? 1.0 : -1.0;  // own normal = sgn * nu
            double* out = sc + f * NF * NFP + k;
            if constexpr (!ACOUSTIC) {
                double A = 0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) A += p.a[d] * nu[d];
                const double uL = left ? own[0] : nbv[0];
                const double uR = left ? (inner ? nbv[0] : 0.0) : own[0];
                double up = 0;
                if (A > 1e-12) up = uL;
                else if (A < -1e-12) up = inner ? uR : 0.0;
                else if (fabs(A) < 1e-12) up = inner ? 0.5 * uL + 0.5 * uR : 0.5 * uL;
                out[0] = -(A * sgn) * up * wf;
            } else {
                if (inner) {
                    double jn = 0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) jn += nu[d] * (own[1 + d] - nbv[1 + d]);
                    // -1/2 n_own . (s_own - s_nbr) and -1/2 n_own (v_own - v_nbr),
                    // n_own = sgn * nu (both sides see L's normal, negated on R)
                    out[0] = -0.5 * (sgn * jn) * wf;
                    const double jv = own[0] - nbv[0];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) out[(1 + d) * NFP] = -0.5 * (sgn * nu[d]) * jv * wf;
                } else {
                    double ns = 0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) ns += nu[d] * own[1 + d];
                    out[0] = -ns * wf;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) out[(1 + d) * NFP] = 0.0;
                }
            }
        }
    }
    __syncthreads();

    // ---- lift the face integrands to the points, divide by the point weight
    double ci = 1.0;
    if constexpr (ACOUSTIC) {
        if (ok) ci = 1.0 / p.cinv[e];
    }
#pragma unroll
    for (int f = 0; f < NFACE; ++f) {
        const int ax = face_axis(DIM, f);
        const int a0 = face_coord_axis(DIM, f, 0), a1 = DIM == 3 ? face_coord_axis(DIM, f, 1) : 0;
        const int k = ii[a0] + (DIM == 3 ? N1 * ii[a1] : 0);
        const double lam = tab[(face_side(DIM, f) ? T::LAM1 : T::LAM0) + ii[ax]];
#pragma unroll
        for (int fld = 0; fld < NF; ++fld) res[fld] += lam * sc[(f * NF + fld) * NFP + k];
    }
#pragma unroll
    for (int fld = 0; fld < NF; ++fld) sA[fld * NP + t] = res[fld] / Wq * (fld > 0 ? ci : 1.0);
    __syncthreads();

    // ---- back to coefficients: k = B^{-1} (point values)
    along<N1, NP, NF, false>(tab + T::BINV, 0, t, sA, sB);
    __syncthreads();
    const double* Kp = sA;
    along<N1, NP, NF, false>(tab + T::BINV, 1, t, sB, sA);
    if constexpr (DIM == 3) {
        __syncthreads();
        along<N1, NP, NF, false>(tab + T::BINV, 2, t, sA, sB);
        Kp = sB;
    }
    if (ok) {
        const int64_t base = e * NF * NP;
        double kept[NF];  // the values as stored
#pragma unroll
        for (int fld = 0; fld < NF; ++fld) {
            const double v = Kp[fld * NP + t];
            if (p.io.out_f32) {
                const float vf = float(v);
                static_cast<float*>(p.io.out)[base + fld * NP + t] = vf;
                kept[fld] = double(vf);
            } else {
                static_cast<double*>(p.io.out)[base + fld * NP + t] = v;
                kept[fld] = v;
            }
        }
        if (p.io.upd.u != nullptr) {  // the RK update fused into a step's last stage (StageIO::upd)
#pragma unroll
            for (int fld = 0; fld < NF; ++fld) {
                const double own[1] = {kept[fld]};
                rk_update_run<1>(p.io.upd, base + fld * NP + t, own);
            }
        }
    }
}

// v = M v or M^{-1} v in place (FP64), M = B^T W B per element and field
// (c_inv W on the acoustic s-fields).
struct GenMassParams {
    int64_t n_elem;
    int32_t nf, inverse;
    double* v;
    const double* X;
    const double* tab;
    const double* cinv;  // NULL: advection
};

template <int DIM, int N1>
__global__ __launch_bounds__(kGenTarget) void tensor_general_mass(const GenMassParams p) {
    using S = GenShape<DIM, N1>;
    using T = NodalTab<N1>;
    constexpr int NP = S::NP, NV = S::NV;
    __shared__ double sXs[S::EPB][NV * DIM];
    __shared__ double sbuf[S::EPB][2][NP];
    __shared__ double tab[T::SIZE];
    const int tid = threadIdx.x, le = tid / NP, t = tid % NP;
    const int64_t e = int64_t(blockIdx.x) * S::EPB + le;
    const bool ok = e < p.n_elem;
    for (int k = tid; k < T::SIZE; k += blockDim.x) tab[k] = p.tab[k];
    if (ok)
        for (int k = t; k < NV * DIM; k += NP) sXs[le][k] = p.X[e * NV * DIM + k];
    __syncthreads();
    const int ii[3] = {t % N1, (t / N1) % N1, DIM == 3 ? t / (N1 * N1) : 0};
    double Wq;
    {
        double xi[DIM], J[DIM][DIM], Ji[DIM][DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = tab[T::X + ii[d]];
        multilinear<DIM>(sXs[le], xi, J, nullptr);
        Wq = fabs(det_inv<DIM>(J, Ji));
#pragma unroll
        for (int d = 0; d < DIM; ++d) Wq *= tab[T::W + ii[d]];
    }
    double* b0 = sbuf[le][0];
    double* b1 = sbuf[le][1];
    for (int fld = 0; fld < p.nf; ++fld) {
        const int64_t o = (e * p.nf + fld) * NP + t;
        const double scale = (p.cinv && fld > 0 && ok) ? p.cinv[e] : 1.0;
        b0[t] = ok ? p.v[o] : 0.0;
        __syncthreads();
        // M: B^T (W (B x)); M^{-1}: B^{-1} (W^{-1} (B^{-T} x))
        const double* M1 = tab + (p.inverse ? T::BINV : T::B);
        double* src = b0;
        double* dst = b1;
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            if (p.inverse) along<N1, NP, 1, true>(M1, ax, t, src, dst);
            else along<N1, NP, 1, false>(M1, ax, t, src, dst);
            __syncthreads();
            double* tmp = src;
            src = dst;
            dst = tmp;
        }
        src[t] = p.inverse ? src[t] / (Wq * scale) : src[t] * (Wq * scale);
        __syncthreads();
#pragma unroll
        for (int ax = 0; ax < DIM; ++ax) {
            if (p.inverse) along<N1, NP, 1, false>(M1, ax, t, src, dst);
            else along<N1, NP, 1, true>(M1, ax, t, src, dst);
            __syncthreads();
            double* tmp = src;
            src = dst;
            dst = tmp;
        }
        if (ok) p.v[o] = src[t];
        __syncthreads();
    }
}

// L2 projection of sine modes or a Gaussian pulse with the element rule (order
// 2p+1): M^{-1} int f phi = B^{-1} (f at the points).
struct GenProjParams {
    int64_t n_elem;
    int32_t nf, field;
    void* out;
    int32_t out_f32;
    int32_t n_modes;
    const double* modes;  // [n_modes][4]: amp, k
    int32_t gaussian;
    double g[5];          // centre, width, amplitude
    const double* X;
    const double* tab;
};

__device__ inline double source_value(const GenProjParams& p, const double* x, int dim) {
    if (p.gaussian) {
        double r2 = 0;
        for (int d = 0; d < dim; ++d) r2 += (x[d] - p.g[d]) * (x[d] - p.g[d]);
        return p.g[4] * exp(-r2 / (p.g[3] * p.g[3]));
    }
    double v = 0;
    for (int m = 0; m < p.n_modes; ++m) {
        double arg = 0;
        for (int d = 0; d < dim; ++d) arg += p.modes[4 * m + 1 + d] * x[d];
        v += p.modes[4 * m] * sin(2 * M_PI * arg);
    }
    return v;
}

template <int DIM, int N1>
__global__ __launch_bounds__(kGenTarget) void tensor_general_project(const GenProjParams p) {
    using S = GenShape<DIM, N1>;
    using T = NodalTab<N1>;
    constexpr int NP = S::NP, NV = S::NV;
    __shared__ double sXs[S::EPB][NV * DIM];
    __shared__ double sbuf[S::EPB][2][NP];
    __shared__ double tab[T::SIZE];
    const int tid = threadIdx.x, le = tid / NP, t = tid % NP;
    const int64_t e = int64_t(blockIdx.x) * S::EPB + le;
    const bool ok = e < p.n_elem;
    for (int k = tid; k < T::SIZE; k += blockDim.x) tab[k] = p.tab[k];
    if (ok)
        for (int k = t; k < NV * DIM; k += NP) sXs[le][k] = p.X[e * NV * DIM + k];
    __syncthreads();
    const int ii[3] = {t % N1, (t / N1) % N1, DIM == 3 ? t / (N1 * N1) : 0};
    double xi[DIM], J[DIM][DIM], x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xi[d] = tab[T::X + ii[d]];
    multilinear<DIM>(sXs[le], xi, J, x);
    sbuf[le][0][t] = ok ? source_value(p, x, DIM) : 0.0;
    __syncthreads();
    double* src = sbuf[le][0];
    double* dst = sbuf[le][1];
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
        along<N1, NP, 1, false>(tab + T::BINV, ax, t, src, dst);
        __syncthreads();
        double* tmp = src;
        src = dst;
        dst = tmp;
    }
    if (!ok) return;
    for (int fld = 0; fld < p.nf; ++fld) {
        if (p.field >= 0 && fld != p.field) continue;
        const int64_t o = (e * p.nf + fld) * NP + t;
        if (p.out_f32) static_cast<float*>(p.out)[o] = float(src[t]);
        else static_cast<double*>(p.out)[o] = src[t];
    }
}

// computeTotalError (rule 2p+2, HpgemAPISimplified_Impl.h:339-401) and
// computeMaxError (the element rule, :403-530) against shifted sine modes,
// or the mass-weighted L2 of a difference x - y (modes = 0, y set): one
// thread per element, per-block partial sums [blocks][2] (sum, max).
struct GenNormParams {
    int64_t n_elem;
    int32_t nf, field;
    const void* u;
    int32_t u_f32;
    const void* y;        // difference mode: the vector subtracted
    int32_t y_f32;
    int32_t n_modes;
    const double* modes;
    double shift[3];
    const double* X;
    const double* tab;
    double* partial;
};

template <int DIM, int N1>
__global__ __launch_bounds__(kBlock) void tensor_general_norm(const GenNormParams p) {
    using S = GenShape<DIM, N1>;
    using T = NodalTab<N1>;
    constexpr int NP = S::NP, NV = S::NV;
    __shared__ double s_red[2][kBlock];
    const int64_t e = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    double sum = 0, mx = 0;
    auto coef = [&](const void* v, int32_t f32, int64_t i) {
        return f32 ? double(static_cast<const float*>(v)[i]) : static_cast<const double*>(v)[i];
    };
    if (e < p.n_elem) {
        const double* X = p.X + e * NV * DIM;
        for (int fld = 0; fld < p.nf; ++fld) {
            if (p.field >= 0 && fld != p.field) continue;
            const int64_t base = (e * p.nf + fld) * NP;
            if (p.y) {
                for (int i = 0; i < NP; ++i)
                    mx = fmax(mx, fabs(coef(p.u, p.u_f32, base + i) - coef(p.y, p.y_f32, base + i)));
            }
            // pass 0: the L2 integral (difference: the element rule, so that it
            // is x^T M x; modes: rule 2p+2); pass 1: max error at the element rule's points
            for (int pass = 0; pass < (p.y ? 1 : 2); ++pass) {
                const bool l2rule = pass == 0 && !p.y;
                const int NQ = l2rule ? N1 + 1 : N1;
                const double* xq = p.tab + (l2rule ? T::X2 : T::X);
                const double* wq = p.tab + (l2rule ? T::W2 : T::W);
                const int nz = DIM == 3 ? NQ : 1;
                for (int qz = 0; qz < nz; ++qz)
                    for (int qy = 0; qy < NQ; ++qy)
                        for (int qx = 0; qx < NQ; ++qx) {
                            const int q[3] = {qx, qy, qz};
                            auto psi = [&](int a, int qq) {
                                return l2rule ? p.tab[T::B2 + a * (N1 + 1) + qq] : p.tab[T::B + qq * N1 + a];
                            };
                            double val = 0;
                            for (int c = 0; c < (DIM == 3 ? N1 : 1); ++c) {
                                double vy = 0;
                                for (int b = 0; b < N1; ++b) {
                                    double vx = 0;
                                    for (int a = 0; a < N1; ++a) {
                                        const int64_t i = base + a + N1 * (b + N1 * c);
                                        double cf = coef(p.u, p.u_f32, i);
                                        if (p.y) cf -= coef(p.y, p.y_f32, i);
                                        vx += cf * psi(a, qx);
                                    }
                                    vy += vx * psi(b, qy);
                                }
                                val += DIM == 3 ? vy * psi(c, qz) : vy;
                            }
                            double xi[DIM], J[DIM][DIM], Ji[DIM][DIM], x[DIM];
                            for (int d = 0; d < DIM; ++d) xi[d] = xq[q[d]];
                            multilinear<DIM>(X, xi, J, x);
                            double err = val;
                            if (!p.y) {
                                double ex = 0;
                                for (int m = 0; m < p.n_modes; ++m) {
                                    double arg = 0;
                                    for (int d = 0; d < DIM; ++d) arg += p.modes[4 * m + 1 + d] * (x[d] - p.shift[d]);
                                    ex += p.modes[4 * m] * sin(2 * M_PI * arg);
                                }
                                err = ex - val;
                            }
                            if (pass == 0) {
                                double w = fabs(det_inv<DIM>(J, Ji));
                                for (int d = 0; d < DIM; ++d) w *= wq[q[d]];
                                sum += w * err * err;
                            } else {
                                mx = fmax(mx, fabs(err));
                            }
                        }
            }
        }
    }
    s_red[0][threadIdx.x] = sum;
    s_red[1][threadIdx.x] = mx;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h /= 2) {
        if (int(threadIdx.x) < h) {
            s_red[0][threadIdx.x] += s_red[0][threadIdx.x + h];
            s_red[1][threadIdx.x] = fmax(s_red[1][threadIdx.x], s_red[1][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        p.partial[2 * blockIdx.x] = s_red[0][0];
        p.partial[2 * blockIdx.x + 1] = s_red[1][0];
    }
}

template <typename F>
void dispatch_general(int dim, int n1, F&& f) {
    auto go = [&](auto D) {
        constexpr int DIMc = decltype(D)::value;
        switch (n1) {
            case 2: f(D, std::integral_constant<int, 2>{}); return;
            case 3: f(D, std::integral_constant<int, 3>{}); return;
            case 4: f(D, std::integral_constant<int, 4>{}); return;
            case 5: f(D, std::integral_constant<int, 5>{}); return;
            case 6: f(D, std::integral_constant<int, 6>{}); return;
        }
        (void)DIMc;
        throw Failure(HGDG_ERR_UNSUPPORTED, "no general tensor kernel for this order");
    };
    if (dim == 2) go(std::integral_constant<int, 2>{});
    else go(std::integral_constant<int, 3>{});
}

// Face orientation as hpGEM pairs the face points: the map of the left
// element's face coordinates onto the right element's is the square symmetry
// that sends the left face's corners onto the right face's corners with the
// same node ids, except that for the two quarter turns hpGEM's index lookup
// (ReferenceSquare::getCodim0MappingIndex) selects the opposite quarter turn
// (MappingToRefSquareToSquare1/3): the map actually applied is the inverse of
// the corner map. Returned as the index code of partner_point (swap, reverse
// first, reverse second).
int left_to_right_code(int dim, const int64_t* lnodes, const int64_t* rnodes) {
    if (dim == 2) return lnodes[0] == rnodes[0] ? 0 : 2;  // a line: identity or reversed
    // corner k of a face has face coordinates (2 (k & 1) - 1, 2 (k >> 1) - 1)
    for (int code = 0; code < 8; ++code) {
        bool all = true;
        for (int k = 0; k < 4 && all; ++k) {
            int pa, pb;
            partner_point(code, k & 1, k >> 1, 2, pa, pb);
            all = lnodes[k] == rnodes[pa + 2 * pb];
        }
        if (!all) continue;
        const bool swap = code & 1, ra = code & 2, rb = code & 4;
        if (swap && ra != rb) return 1 | (ra ? 4 : 0) | (rb ? 2 : 0);  // quarter turn: the opposite one
        return code;
    }
    throw Failure(HGDG_ERR_ARG, "a shared face whose corners do not match");
}

// Inverse of a partner_point code (maps the right element's face point indices back).
int inverse_code(int code) {
    if (!(code & 1)) return code;
    return 1 | ((code & 2) ? 4 : 0) | ((code & 4) ? 2 : 0);
}

}  // namespace

void tensor_general_upload(hgdg_ctx* c, const hgdg_mesh_desc* m) {
    const int dim = shape_dim(m->shape);
    const int64_t ne = m->n_elements;
    const int nv = 1 << dim, nfaces = 2 * dim, fverts = 1 << (dim - 1);
    TensorMesh& tm = c->tm;
    tm.general = true;
    tm.structured = false;
    tm.gX.assign(m->coords, m->coords + size_t(ne) * nv * dim);
    tm.gnbr.assign(size_t(nfaces) * ne, -1);
    tm.ginfo.assign(size_t(nfaces) * ne, 0);
    // every element must map with a nonzero Jacobian determinant at its points
    // (checked at the element rule's points on the host)
    const NodalOps ops = nodal_ops(tm.p);
    for (int64_t e = 0; e < ne; ++e) {
        const double* X = m->coords + size_t(e) * nv * dim;
        const int npts = dim == 3 ? tm.n1 * tm.n1 * tm.n1 : tm.n1 * tm.n1;
        for (int q = 0; q < npts; ++q) {
            const int ii[3] = {q % tm.n1, (q / tm.n1) % tm.n1, q / (tm.n1 * tm.n1)};
            double J[3][3] = {{0}};
            for (int v = 0; v < nv; ++v)
                for (int d = 0; d < dim; ++d) {
                    double dd = ((v >> d) & 1) ? 0.5 : -0.5;
                    for (int k = 0; k < dim; ++k)
                        if (k != d) dd *= ((v >> k) & 1) ? 0.5 * (1 + ops.x[ii[k]]) : 0.5 * (1 - ops.x[ii[k]]);
                    for (int r = 0; r < dim; ++r) J[r][d] += X[v * dim + r] * dd;
                }
            const double det = dim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                        : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                              J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            require(det != 0.0 && std::isfinite(det), HGDG_ERR_ARG, "degenerate element " + std::to_string(e));
        }
    }
    // faces: hpGEM's face search visits elements in order, so the first
    // element to reach a face is its left element; a face's node list follows
    // the face's corner order (corner k: bit 0 of k along the first face
    // coordinate, bit 1 along the second)
    auto face_node = [&](int64_t e, int f, int k) {
        const int ax = face_axis(dim, f);
        int v = face_side(dim, f) << ax;
        v |= (k & 1) << face_coord_axis(dim, f, 0);
        if (dim == 3) v |= ((k >> 1) & 1) << face_coord_axis(dim, f, 1);
        return m->node_ids[size_t(e) * nv + v];
    };
    std::unordered_map<std::string, std::pair<int64_t, int>> open;
    open.reserve(size_t(ne) * nfaces);
    int64_t n_internal = 0, n_total = 0;
    for (int64_t e = 0; e < ne; ++e) {
        for (int f = 0; f < nfaces; ++f) {
            std::array<int64_t, 4> key{};
            for (int k = 0; k < fverts; ++k) key[k] = face_node(e, f, k);
            std::sort(key.begin(), key.begin() + fverts);
            std::string kstr(reinterpret_cast<const char*>(key.data()), sizeof(int64_t) * fverts);
            auto it = open.find(kstr);
            if (it == open.end()) {
                open.emplace(std::move(kstr), std::make_pair(e, f));
                continue;
            }
            const int64_t l = it->second.first;
            const int fl = it->second.second;
            require(l != e, HGDG_ERR_UNSUPPORTED, "element " + std::to_string(e) + " meets itself across a face");
            require(tm.gnbr[size_t(fl) * ne + l] < 0, HGDG_ERR_ARG, "a face is bounded by more than two elements");
            int64_t ln[4], rn[4];
            for (int k = 0; k < fverts; ++k) {
                ln[k] = face_node(l, fl, k);
                rn[k] = face_node(e, f, k);
            }
            const int code = left_to_right_code(dim, ln, rn);
            tm.gnbr[size_t(fl) * ne + l] = int32_t(e);
            tm.ginfo[size_t(fl) * ne + l] = f | (code << 3) | (1 << 6);
            tm.gnbr[size_t(f) * ne + e] = int32_t(l);
            tm.ginfo[size_t(f) * ne + e] = fl | (inverse_code(code) << 3);
            open.erase(it);
            ++n_internal;
            ++n_total;
        }
    }
    n_total += int64_t(open.size());
    c->n_faces = n_total;
    c->n_internal = n_internal;
    c->n_fields = 1;
    c->dof = ne * tm.nb;
    tm.d_gX.upload(tm.gX, c->stream);
    std::vector<int32_t> fi(tm.gnbr);
    fi.insert(fi.end(), tm.ginfo.begin(), tm.ginfo.end());
    tm.d_gface.upload(fi, c->stream);
    tm.d_gtab.upload(ops.packed(), c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
}

void tensor_general_operator_set(hgdg_ctx* c, const hgdg_op_desc* op) {
    const int64_t ne = c->n_elem;
    if (op->kind == HGDG_OP_ACOUSTIC_CENTRAL) {
        c->tm.d_gcinv.upload(std::vector<double>(op->c_inv, op->c_inv + ne), c->stream);
        c->n_fields = c->dim + 1;
    } else {
        c->n_fields = 1;
    }
    c->dof = ne * c->tm.nb * c->n_fields;
    HGDG_HIP(hipStreamSynchronize(c->stream));
}

void tensor_general_rhs(hgdg_ctx* c, const StageIO& io) {
    GenParams p{};
    p.n_elem = c->n_elem;
    p.io = io;
    p.X = c->tm.d_gX.ptr;
    p.nbr = c->tm.d_gface.ptr;
    p.info = c->tm.d_gface.ptr + size_t(2 * c->dim) * c->n_elem;
    p.tab = c->tm.d_gtab.ptr;
    const bool acoustic = c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL;
    p.cinv = acoustic ? c->tm.d_gcinv.ptr : nullptr;
    for (int d = 0; d < 3; ++d) p.a[d] = d < c->dim ? c->op.velocity[d] : 0.0;
    dispatch_general(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = GenShape<DIMc, N1c>;
        const int64_t grid = (c->n_elem + S::EPB - 1) / S::EPB;
        if (acoustic)
            hipLaunchKernelGGL((tensor_general_rhs<DIMc, N1c, true>), dim3(unsigned(grid)), dim3(S::THREADS), 0,
                               c->stream, p);
        else
            hipLaunchKernelGGL((tensor_general_rhs<DIMc, N1c, false>), dim3(unsigned(grid)), dim3(S::THREADS), 0,
                               c->stream, p);
    });
    HGDG_HIP(hipGetLastError());
}

void tensor_general_mass_apply(hgdg_ctx* c, Vec& v, bool inverse) {
    require(!v.is_f32, HGDG_ERR_UNSUPPORTED, "mass apply/solve needs a vector stored in FP64");
    GenMassParams p{};
    p.n_elem = c->n_elem;
    p.nf = c->n_fields;
    p.inverse = inverse ? 1 : 0;
    p.v = v.f64->ptr;
    p.X = c->tm.d_gX.ptr;
    p.tab = c->tm.d_gtab.ptr;
    p.cinv = c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL ? c->tm.d_gcinv.ptr : nullptr;
    dispatch_general(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = GenShape<DIMc, N1c>;
        const int64_t grid = (c->n_elem + S::EPB - 1) / S::EPB;
        hipLaunchKernelGGL((tensor_general_mass<DIMc, N1c>), dim3(unsigned(grid)), dim3(S::THREADS), 0, c->stream, p);
    });
    HGDG_HIP(hipGetLastError());
}

void tensor_general_project(hgdg_ctx* c, Vec& v, int field, int n_modes, const double* amp, const double* k,
                            const double* gauss) {
    std::vector<double> modes;
    for (int m = 0; m < n_modes; ++m) {
        modes.push_back(amp[m]);
        for (int d = 0; d < 3; ++d) modes.push_back(k[3 * m + d]);
    }
    modes.push_back(0.0);  // never empty
    DevBuf<double> d_modes;
    d_modes.upload(modes, c->stream);
    GenProjParams p{};
    p.n_elem = c->n_elem;
    p.nf = c->n_fields;
    p.field = field;
    p.out = v.ptr();
    p.out_f32 = v.is_f32;
    p.n_modes = n_modes;
    p.modes = d_modes.ptr;
    if (gauss) {
        p.gaussian = 1;
        for (int q = 0; q < 5; ++q) p.g[q] = gauss[q];
    }
    p.X = c->tm.d_gX.ptr;
    p.tab = c->tm.d_gtab.ptr;
    dispatch_general(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = GenShape<DIMc, N1c>;
        const int64_t grid = (c->n_elem + S::EPB - 1) / S::EPB;
        hipLaunchKernelGGL((tensor_general_project<DIMc, N1c>), dim3(unsigned(grid)), dim3(S::THREADS), 0, c->stream,
                           p);
    });
    HGDG_HIP(hipGetLastError());
    HGDG_HIP(hipStreamSynchronize(c->stream));  // before the temporary modes go
}

namespace {

void general_norm(hgdg_ctx* c, GenNormParams& p, double* l2, double* max_abs) {
    p.n_elem = c->n_elem;
    p.nf = c->n_fields;
    p.X = c->tm.d_gX.ptr;
    p.tab = c->tm.d_gtab.ptr;
    const int64_t grid = (c->n_elem + kBlock - 1) / kBlock;
    if (c->d_partial.n < size_t(2 * grid)) c->d_partial.alloc(size_t(2 * grid));
    p.partial = c->d_partial.ptr;
    dispatch_general(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        hipLaunchKernelGGL((tensor_general_norm<DIMc, N1c>), dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
    });
    HGDG_HIP(hipGetLastError());
    std::vector<double> h(size_t(2 * grid));
    HGDG_HIP(hipMemcpyAsync(h.data(), c->d_partial.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
    double s = 0, m = 0;
    for (int64_t i = 0; i < grid; ++i) {
        s += h[2 * i];
        m = std::max(m, h[2 * i + 1]);
    }
    *l2 = std::sqrt(std::max(s, 0.0));
    *max_abs = m;
}

}  // namespace

void tensor_general_error_modes(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp,
                                const double* k, const double* shift, double* l2, double* max_abs) {
    std::vector<double> modes;
    for (int m = 0; m < n_modes; ++m) {
        modes.push_back(amp[m]);
        for (int d = 0; d < 3; ++d) modes.push_back(k[3 * m + d]);
    }
    modes.push_back(0.0);
    DevBuf<double> d_modes;
    d_modes.upload(modes, c->stream);
    GenNormParams p{};
    p.field = field;
    p.u = v.ptr();
    p.u_f32 = v.is_f32;
    p.n_modes = n_modes;
    p.modes = d_modes.ptr;
    for (int d = 0; d < 3; ++d) p.shift[d] = shift ? shift[d] : 0.0;
    general_norm(c, p, l2, max_abs);
}

void tensor_general_state_diff(hgdg_ctx* c, const Vec& x, const Vec& y, double* l2, double* max_abs) {
    GenNormParams p{};
    p.field = -1;
    p.u = x.ptr();
    p.u_f32 = x.is_f32;
    p.y = y.ptr();
    p.y_f32 = y.is_f32;
    general_norm(c, p, l2, max_abs);
}

}  // namespace hgdg
