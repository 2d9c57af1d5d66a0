// Interior-penalty DG assembly of the Maxwell curl-curl operator on
// tetrahedra (DG-Max: applications/DG-Max/Algorithms/DGMaxDiscretization.cpp
// :193-260 element matrices, :344-410 face matrices) with hpGEM's
// Ainsworth-Coyle H(curl) set and the covariant transform
// (kernel/Base/HCurlConformingTransformation.h): phi = J^{-T} phi_ref,
// curl phi = J curl_ref / det J.
//
// Output: the stiffness K = sum_e S_e + sum_faces F as a block-CSR matrix
// (one block row per element, mesh order; blocks n x n, row = test function
// of the row element, column = trial function of the column element; the
// column blocks of a row in ascending element order), and the block-diagonal
// permittivity mass M.
//
// One workgroup per element writes its whole block row, so no two
// workgroups write the same block:
//  * S_e = (1/|det J|) sum_ab (J^T J)_ab Kref^{ab},
//    M_e = eps |det J| sum_ab (J^{-1} J^{-T})_ab Nref^{ab}, with the nine
//    reference matrices Kref^{ab}_ij = int curl_a phi_i curl_b phi_j and
//    Nref^{ab}_ij = int phi_i,a phi_j,b built once on the host;
//  * per face, at the 16 points of the symmetric order-8 face rule (exact for
//    every face term at p <= 2): T = n x phi with the element's own outward
//    normal (the neighbour's functions get -n x phi, as the reference's right
//    side does with the left normal), C = curl phi, and
//    block(i, j) += w [factor (C_j.T_i + T_j.C_i) + stab/diam T_j.T_i],
//    factor = -1/2 inside, -1 on the boundary, diam the face's longest edge;
//    the self term goes into the diagonal block, the neighbour term into the
//    neighbour's block.
#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "context.hpp"

namespace hgdg {

namespace {

struct RefTetRule {
    int order, n;
    const double* pts;
    const double* w;
};
#include "tet_rules.inc"


struct MaxwellParams {
    int64_t n_elem;
    int32_t n;               // functions per element
    const double* geo;       // per device element: J [9], J^{-1} [9], det J, eps (20 doubles)
    const double* face;      // per device element and local face: unit normal [3], w scale (area), stab/diam, factor
    const int32_t* nbr;      // [4][E] neighbour device element or -1
    const int32_t* info;     // [4][E] packed face info (local face of the neighbour, orientations)
    const int64_t* slot;     // per device element: diagonal block, then the block of each face (-1: boundary)
    const double* Kref;      // [9][n][n]
    const double* Nref;      // [9][n][n]
    const double* FV;        // [4 faces][6 orientations][kFaceQ][n][3] reference values at face points
    const double* FC;        // same for curls
    double face_w[kFaceQ];
    double* K;               // [n_blocks][n][n]
    double* M;               // [mesh element][n][n]
    const int64_t* mesh_of;  // device element -> mesh element
};

__device__ __forceinline__ int mx_lf(int v) { return v & 3; }
__device__ __forceinline__ int mx_oown(int v) { return (v >> 2) & 7; }
__device__ __forceinline__ int mx_onbr(int v) { return (v >> 5) & 7; }

typedef double mx_double4 __attribute__((ext_vector_type(4)));

// Face values of one side in LDS: Xt[k][i], k = 6 q + d, d < 3 the component
// d of T = n x phi_i at face point q, d >= 3 the component d - 3 of curl
// phi_i; i padded to 32 with zeros.
constexpr int kMxK = 6 * kFaceQ;  // 96
constexpr int kMxPad = 32;

// B operand of the face GEMM block(i, j) = sum_k Xt[k][i] Y[j][k]: Y[j][k] =
// ws_q (factor C_j + ls T_j)_d for k = 6q + d, d < 3, and ws_q factor T_j,d-3
// for d >= 3 (ws_q = face weight x area scale), from the side's face values.
__device__ __forceinline__ double mx_b(const double* xt, int k, int j, const double* wq, double factor, double ls) {
    const int q = k / 6, d = k % 6;
    if (d < 3) return wq[q] * (factor * xt[(6 * q + 3 + d) * kMxPad + j] + ls * xt[(6 * q + d) * kMxPad + j]);
    return wq[q] * factor * xt[(6 * q + d - 3) * kMxPad + j];
}

__global__ __launch_bounds__(kBlock) void maxwell_assemble_kernel(const MaxwellParams p) {
    __shared__ double s_geo[2][20];
    __shared__ double s_x[2][kMxK * kMxPad];  // side (own, neighbour): face values, Xt layout
    __shared__ double s_face[8];
    __shared__ double s_wq[kFaceQ];           // face weight x area scale
    const int tid = threadIdx.x;
    const int lane = tid % 64, wave = tid / 64;
    const int64_t e = blockIdx.x;
    const int n = p.n;
    const int nn = n * n;
    if (tid < 20) s_geo[0][tid] = p.geo[e * 20 + tid];
    for (int k = tid; k < 2 * kMxK * kMxPad; k += kBlock) (&s_x[0][0])[k] = 0.0;  // zero padding columns once
    __syncthreads();
    const double* J = s_geo[0];
    const double* Ji = s_geo[0] + 9;
    const double det = s_geo[0][18], eps = s_geo[0][19];
    double G[9], H[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double g = 0, h = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                g += J[k * 3 + a] * J[k * 3 + b];    // (J^T J)_ab
                h += Ji[a * 3 + k] * Ji[b * 3 + k];  // (J^{-1} J^{-T})_ab
            }
            G[a * 3 + b] = g / fabs(det);
            H[a * 3 + b] = h * eps * fabs(det);
        }
    // the permittivity mass, all threads over its entries
    const int64_t me = p.mesh_of[e];
    for (int k = tid; k < nn; k += kBlock) {
        double m = 0;
#pragma unroll
        for (int ab = 0; ab < 9; ++ab) m += H[ab] * p.Nref[ab * nn + k];
        p.M[me * nn + k] = m;
    }
    // this wave's 16 x 16 tile of the 32 x 32 (padded) block: entries
    // (row0 + 4 r + lane / 16, col0 + lane % 16), r = 0..3, as the MFMA holds them
    const int tiles = (n + 15) / 16;
    const int ti = wave / 2, tj = wave % 2;
    const bool tile_on = ti < tiles && tj < tiles;
    const int row0 = 16 * ti, col0 = 16 * tj;
    const int col = col0 + lane % 16;
    mx_double4 acc_d = {0.0, 0.0, 0.0, 0.0};  // the diagonal block: S_e, then the self face terms
    if (tile_on)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = row0 + 4 * r + lane / 16;
            if (row < n && col < n) {
                double s = 0;
#pragma unroll
                for (int ab = 0; ab < 9; ++ab) s += G[ab] * p.Kref[ab * nn + row * n + col];
                acc_d[r] = s;
            }
        }
    for (int f = 0; f < 4; ++f) {
        const int64_t nb = p.nbr[f * p.n_elem + e];
        const int inf = p.info[f * p.n_elem + e];
        __syncthreads();  // the previous face's values are consumed
        if (tid < 8) s_face[tid] = p.face[(e * 4 + f) * 8 + tid];
        if (nb >= 0 && tid < 20) s_geo[1][tid] = p.geo[nb * 20 + tid];
        __syncthreads();
        const double* nrm = s_face;
        if (tid < kFaceQ) s_wq[tid] = p.face_w[tid] * s_face[3];
        // T and C at the face points, both sides (items: side, point q, function i)
        for (int it = tid; it < 2 * kFaceQ * n; it += kBlock) {
            const int side = it / (kFaceQ * n);
            if (side == 1 && nb < 0) continue;
            const int q = (it / n) % kFaceQ, i = it % n;
            const double* g = s_geo[side];
            const int face = side == 0 ? f : mx_lf(inf);
            const int o = side == 0 ? mx_oown(inf) : mx_onbr(inf);
            const double* v = p.FV + (((face * 6 + o) * kFaceQ + q) * n + i) * 3;
            const double* c = p.FC + (((face * 6 + o) * kFaceQ + q) * n + i) * 3;
            double ph[3], cu[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                ph[d] = g[9 + 0 * 3 + d] * v[0] + g[9 + 1 * 3 + d] * v[1] + g[9 + 2 * 3 + d] * v[2];  // J^{-T} v
                cu[d] = (g[d * 3 + 0] * c[0] + g[d * 3 + 1] * c[1] + g[d * 3 + 2] * c[2]) / g[18];
            }
            const double sg = side == 0 ? 1.0 : -1.0;
            double* xt = s_x[side] + (6 * q) * kMxPad + i;
            xt[0 * kMxPad] = sg * (nrm[1] * ph[2] - nrm[2] * ph[1]);
            xt[1 * kMxPad] = sg * (nrm[2] * ph[0] - nrm[0] * ph[2]);
            xt[2 * kMxPad] = sg * (nrm[0] * ph[1] - nrm[1] * ph[0]);
            xt[3 * kMxPad] = cu[0];
            xt[4 * kMxPad] = cu[1];
            xt[5 * kMxPad] = cu[2];
        }
        __syncthreads();
        if (!tile_on) continue;
        const double ls = s_face[4], factor = s_face[5];
        // self term: acc_d += X_own^T Y_own over the 96 face values
        {
            const double* xa = s_x[0];
#pragma unroll 4
            for (int s4 = 0; s4 < kMxK / 4; ++s4) {
                const int k = 4 * s4 + lane / 16;
                const double a = xa[k * kMxPad + row0 + lane % 16];
                const double b = mx_b(xa, k, col0 + lane % 16, s_wq, factor, ls);
                acc_d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc_d, 0, 0, 0);
            }
        }
        if (nb >= 0) {  // neighbour term: its block X_own^T Y_nbr
            const double* xa = s_x[0];
            const double* xb = s_x[1];
            mx_double4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
            for (int s4 = 0; s4 < kMxK / 4; ++s4) {
                const int k = 4 * s4 + lane / 16;
                const double a = xa[k * kMxPad + row0 + lane % 16];
                const double b = mx_b(xb, k, col0 + lane % 16, s_wq, factor, ls);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            }
            double* blk = p.K + p.slot[e * 5 + 1 + f] * nn;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + 4 * r + lane / 16;
                if (row < n && col < n) blk[row * n + col] = acc[r];
            }
        }
    }
    if (!tile_on) return;
    double* dblk = p.K + p.slot[e * 5] * nn;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = row0 + 4 * r + lane / 16;
        if (row < n && col < n) dblk[row * n + col] = acc_d[r];
    }
}

// Gauss-Jordan inverse (long double), small dense matrices
void invert3(const double* A, double* inv, double* det_out) {
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                       A[2] * (A[3] * A[7] - A[4] * A[6]);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            const int r1 = (k + 1) % 3, r2 = (k + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
            inv[r * 3 + k] = (A[r1 * 3 + c1] * A[r2 * 3 + c2] - A[r1 * 3 + c2] * A[r2 * 3 + c1]) / det;
        }
    *det_out = det;
}

}  // namespace

// Host side: tables, geometry, block pattern; then one launch.
void maxwell_assemble(hgdg_ctx* c, int order, const double* eps, double stab, int64_t* n_blocks, int32_t* block_size,
                      int64_t* row_ptr, int64_t* col_idx, double* K, double* M) {
    require(c->has_mesh && c->simplex(), HGDG_ERR_STATE, "upload a tetrahedral mesh first");
    require(order == 1 || order == 2, HGDG_ERR_UNSUPPORTED, "the Maxwell assembly is built for orders 1 and 2");
    require(std::isfinite(stab) && stab >= 0, HGDG_ERR_ARG, "penalty must be finite and non-negative");
    SimplexMesh& sm = c->sm;
    const int64_t ne = c->n_elem;
    const int n = hcurl_tet_size(order);
    *block_size = n;
    // block pattern in mesh order: each row lists itself and its face neighbours, ascending
    std::vector<int64_t> rp(size_t(ne) + 1, 0), ci;
    std::vector<int64_t> slot(size_t(ne) * 5, -1);  // device element: diagonal, faces 0..3
    for (int64_t m = 0; m < ne; ++m) {
        const int64_t d = sm.dev_of[m];
        std::vector<std::pair<int64_t, int>> cols{{m, -1}};
        for (int f = 0; f < 4; ++f) {
            const int32_t nb = sm.nbr_dev[size_t(f) * ne + d];
            if (nb >= 0) cols.push_back({sm.mesh_of[nb], f});
        }
        std::sort(cols.begin(), cols.end());
        for (size_t k = 1; k < cols.size(); ++k)
            require(cols[k].first != cols[k - 1].first, HGDG_ERR_UNSUPPORTED,
                    "two faces of an element meet the same neighbour");
        for (const auto& cf : cols) {
            slot[size_t(d) * 5 + 1 + cf.second] = int64_t(ci.size());
            ci.push_back(cf.first);
        }
        rp[size_t(m) + 1] = int64_t(ci.size());
    }
    *n_blocks = int64_t(ci.size());
    if (!row_ptr) return;
    require(col_idx && K && M && eps, HGDG_ERR_ARG, "NULL output array");
    std::copy(rp.begin(), rp.end(), row_ptr);
    std::copy(ci.begin(), ci.end(), col_idx);

    // geometry per device element: J, J^{-1}, det J, eps
    std::vector<double> geo(size_t(ne) * 20), face(size_t(ne) * 4 * 8, 0.0);
    static const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int64_t d = 0; d < ne; ++d) {
        const double* Jd = &sm.jac[size_t(d) * 9];
        double inv[9], det;
        invert3(Jd, inv, &det);
        double* g = &geo[size_t(d) * 20];
        std::copy(Jd, Jd + 9, g);
        std::copy(inv, inv + 9, g + 9);
        g[18] = det;
        const double ev = eps[sm.mesh_of[d]];
        require(std::isfinite(ev) && ev > 0, HGDG_ERR_ARG, "permittivity must be positive");
        g[19] = ev;
        for (int f = 0; f < 4; ++f) {
            double* fc = &face[(size_t(d) * 4 + f) * 8];
            for (int r = 0; r < 3; ++r) fc[r] = sm.normal[(size_t(r) * 4 + f) * ne + d];
            fc[3] = sm.sigma[size_t(f) * ne + d] * sm.absdet[d];  // area element of the reference-triangle map
            double diam2 = 0;
            const int* fn = kTetFaceNodes[f];
            for (int a = 0; a < 3; ++a)
                for (int b = a + 1; b < 3; ++b) {
                    double d2 = 0;
                    for (int r = 0; r < 3; ++r) {
                        double dx = 0;
                        for (int k = 0; k < 3; ++k) dx += Jd[r * 3 + k] * (V[fn[a]][k] - V[fn[b]][k]);
                        d2 += dx * dx;
                    }
                    diam2 = std::max(diam2, d2);
                }
            fc[4] = stab / std::sqrt(diam2);
            fc[5] = sm.nbr_dev[size_t(f) * ne + d] >= 0 ? -0.5 : -1.0;
        }
    }
    // reference tables
    std::vector<double> Kref(size_t(9) * n * n, 0.0), Nref(size_t(9) * n * n, 0.0);
    {
        std::vector<long double> pts, w;
        tet_rule(2 * order + 1, pts, w);
        std::vector<long double> v(size_t(3) * n), cu(size_t(3) * n);
        std::vector<long double> ks(size_t(9) * n * n, 0.0L), ns(size_t(9) * n * n, 0.0L);
        for (size_t q = 0; q < w.size(); ++q) {
            hcurl_tet_basis(order, &pts[q * 3], v.data(), cu.data());
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int i = 0; i < n; ++i)
                        for (int j = 0; j < n; ++j) {
                            ks[((a * 3 + b) * n + i) * n + j] += w[q] * cu[i * 3 + a] * cu[j * 3 + b];
                            ns[((a * 3 + b) * n + i) * n + j] += w[q] * v[i * 3 + a] * v[j * 3 + b];
                        }
        }
        for (size_t k = 0; k < ks.size(); ++k) {
            Kref[k] = double(ks[k]);
            Nref[k] = double(ns[k]);
        }
    }
    // values and curls at the face rule points, per local face and orientation
    std::vector<double> FV(size_t(4) * 6 * kFaceQ * n * 3), FC(FV.size());
    {
        std::vector<long double> v(size_t(3) * n), cu(size_t(3) * n);
        for (int f = 0; f < 4; ++f)
            for (int o = 0; o < 6; ++o)
                for (int q = 0; q < kFaceQ; ++q) {
                    const long double tp[2] = {kFaceRulePts[q][0], kFaceRulePts[q][1]};
                    long double to[2], x[3];
                    tri_orientation_map(o, tp, to);
                    tet_face_map(f, to, x);
                    hcurl_tet_basis(order, x, v.data(), cu.data());
                    const size_t base = (((size_t(f) * 6 + o) * kFaceQ + q) * n) * 3;
                    for (int k = 0; k < 3 * n; ++k) {
                        FV[base + k] = double(v[k]);
                        FC[base + k] = double(cu[k]);
                    }
                }
    }
    std::vector<int64_t> mesh_of(sm.mesh_of.begin(), sm.mesh_of.end());
    DevBuf<double> d_geo, d_face, d_Kref, d_Nref, d_FV, d_FC, d_K, d_M;
    DevBuf<int64_t> d_slot, d_mesh_of;
    d_geo.upload(geo, c->stream);
    d_face.upload(face, c->stream);
    d_Kref.upload(Kref, c->stream);
    d_Nref.upload(Nref, c->stream);
    d_FV.upload(FV, c->stream);
    d_FC.upload(FC, c->stream);
    d_slot.upload(slot, c->stream);
    d_mesh_of.upload(mesh_of, c->stream);
    d_K.alloc(size_t(*n_blocks) * n * n);
    d_M.alloc(size_t(ne) * n * n);
    MaxwellParams p{};
    p.n_elem = ne;
    p.n = n;
    p.geo = d_geo.ptr;
    p.face = d_face.ptr;
    p.nbr = sm.d_face.ptr;
    p.info = sm.d_face.ptr + 4 * ne;
    p.slot = d_slot.ptr;
    p.Kref = d_Kref.ptr;
    p.Nref = d_Nref.ptr;
    p.FV = d_FV.ptr;
    p.FC = d_FC.ptr;
    for (int q = 0; q < kFaceQ; ++q) p.face_w[q] = kFaceRuleW[q];
    p.K = d_K.ptr;
    p.M = d_M.ptr;
    p.mesh_of = d_mesh_of.ptr;
    EventPair ev;
    hipEvent_t t0 = ev.e[0], t1 = ev.e[1];
    HGDG_HIP(hipEventRecord(t0, c->stream));
    hipLaunchKernelGGL(maxwell_assemble_kernel, dim3(unsigned(ne)), dim3(kBlock), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
    HGDG_HIP(hipEventRecord(t1, c->stream));
    HGDG_HIP(hipMemcpyAsync(K, d_K.ptr, size_t(*n_blocks) * n * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipMemcpyAsync(M, d_M.ptr, size_t(ne) * n * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
    float ms = 0;
    HGDG_HIP(hipEventElapsedTime(&ms, t0, t1));
    c->last_kernel_ms = double(ms);
}

}  // namespace hgdg
