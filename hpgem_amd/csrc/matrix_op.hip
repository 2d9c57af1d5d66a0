// Linear operators given by their matrices: hpGEM's precomputed path. An
// application of HpgemAPILinear builds, once, every element's mass and
// stiffness matrix and every face's stiffness matrix from its own integrands
// (computeIntegrandMassMatrix, computeIntegrandStiffnessMatrixAtElement /
// AtFace through ElementIntegral / FaceIntegral: createMassMatrices and
// createStiffnessMatrices, HpgemAPILinear_Impl.h:131-139, 250-288), and each
// RHS evaluation then multiplies them (multiplyStiffnessMatrices, :360-434)
// and solves the mass matrices (solveMassMatrixEquations, :143-156). Any
// linear integrand a user writes reaches the device this way: the matrices
// are handed over once (hgdg_operator_set_matrices) and every RHS, mass solve
// and RK step runs here.
//
// Layout: vectors in hpGEM's order (element by element, iVB = iV nB + iB
// inside an element), no permutation. Per element row, the blocks it
// multiplies: its diagonal block D_e = S_e + the (test e, trial e) blocks of
// its faces, summed once on the host, then one block per (face, neighbour)
// pair, each column-major n_e x n_trial in one pool. The mass matrices'
// inverses are formed once on the host in long double (hpGEM factorises M_e
// on every call; the two agree to the conditioning of M_e).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "context.hpp"

namespace hgdg {

namespace {

constexpr int kMatThreads = 256;
constexpr int kMatMaxRows = 4 * kMatThreads;  // rows per element a workgroup takes

struct MatParams {
    int64_t n_elem;
    StageIO io;
    const int64_t* elem_off;  // [E + 1]
    const int64_t* row_ptr;   // [E + 1] block entries of each element row
    const int64_t* blk_off;   // [entries] offset of the block in pool
    const int32_t* blk_col;   // [entries] trial element
    const double* pool;
    const int64_t* sq_off;    // [E] offset of element e's n_e x n_e matrices in minv / mass
    const double* minv;       // NULL: no mass solve
};

// r_e = sum over the element row's blocks of B w_trial (w = the stage
// combination), then k_e = M_e^{-1} r_e when minv is set. One workgroup per
// element; rows over threads, trial values through LDS.
__global__ __launch_bounds__(kMatThreads) void matrix_rhs(const MatParams p) {
    extern __shared__ double sh[];
    const int64_t e = blockIdx.x;
    const int64_t o = p.elem_off[e];
    const int n = int(p.elem_off[e + 1] - o);
    double acc[4] = {0, 0, 0, 0};
    for (int64_t k = p.row_ptr[e]; k < p.row_ptr[e + 1]; ++k) {
        const int64_t t = p.blk_col[k];
        const int64_t ot = p.elem_off[t];
        const int nt = int(p.elem_off[t + 1] - ot);
        for (int j = threadIdx.x; j < nt; j += kMatThreads) sh[j] = combine(p.io, ot + j);
        __syncthreads();
        const double* B = p.pool + p.blk_off[k];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int i = threadIdx.x + s * kMatThreads;
            if (i >= n) break;
            double a = acc[s];
            for (int j = 0; j < nt; ++j) a += B[i + int64_t(n) * j] * sh[j];
            acc[s] = a;
        }
        __syncthreads();
    }
    if (p.minv) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int i = threadIdx.x + s * kMatThreads;
            if (i < n) sh[i] = acc[s];
        }
        __syncthreads();
        const double* Mi = p.minv + p.sq_off[e];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int i = threadIdx.x + s * kMatThreads;
            if (i >= n) break;
            double a = 0;
            for (int j = 0; j < n; ++j) a += Mi[i + int64_t(n) * j] * sh[j];
            acc[s] = a;
        }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int i = threadIdx.x + s * kMatThreads;
        if (i >= n) break;
        double kept;  // the value as stored
        if (p.io.out_f32) {
            const float v = float(acc[s]);
            static_cast<float*>(p.io.out)[o + i] = v;
            kept = double(v);
        } else {
            static_cast<double*>(p.io.out)[o + i] = acc[s];
            kept = acc[s];
        }
        if (p.io.upd.u != nullptr) {  // the RK update fused into a step's last stage (StageIO::upd)
            const double own[1] = {kept};
            rk_update_run<1>(p.io.upd, o + i, own);
        }
    }
}

// v_e = X_e v_e in place (X = M or M^{-1}), FP64.
__global__ __launch_bounds__(kMatThreads) void matrix_apply(const int64_t* elem_off, const int64_t* sq_off,
                                                            const double* X, double* v) {
    extern __shared__ double sh[];
    const int64_t e = blockIdx.x;
    const int64_t o = elem_off[e];
    const int n = int(elem_off[e + 1] - o);
    for (int j = threadIdx.x; j < n; j += kMatThreads) sh[j] = v[o + j];
    __syncthreads();
    const double* A = X + sq_off[e];
    for (int i = threadIdx.x; i < n; i += kMatThreads) {
        double a = 0;
        for (int j = 0; j < n; ++j) a += A[i + int64_t(n) * j] * sh[j];
        v[o + i] = a;
    }
}

// Per element: d^T M_e d and max |d|, d = x_e - y_e; one partial pair per element.
__global__ __launch_bounds__(kMatThreads) void matrix_diff(const int64_t* elem_off, const int64_t* sq_off,
                                                           const double* M, const void* x, int32_t x_f32,
                                                           const void* y, int32_t y_f32, double* partial) {
    extern __shared__ double sh[];
    __shared__ double red[2][kMatThreads];
    const int64_t e = blockIdx.x;
    const int64_t o = elem_off[e];
    const int n = int(elem_off[e + 1] - o);
    double mx = 0;
    for (int j = threadIdx.x; j < n; j += kMatThreads) {
        const double a = x_f32 ? double(static_cast<const float*>(x)[o + j]) : static_cast<const double*>(x)[o + j];
        const double b = y_f32 ? double(static_cast<const float*>(y)[o + j]) : static_cast<const double*>(y)[o + j];
        sh[j] = a - b;
        mx = fmax(mx, fabs(a - b));
    }
    __syncthreads();
    const double* A = M + sq_off[e];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += kMatThreads) {
        double a = 0;
        for (int j = 0; j < n; ++j) a += A[i + int64_t(n) * j] * sh[j];
        s += sh[i] * a;
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = mx;
    __syncthreads();
    for (int h = kMatThreads / 2; h > 0; h /= 2) {
        if (int(threadIdx.x) < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * e] = red[0][0];
        partial[2 * e + 1] = red[1][0];
    }
}

// Inverse of a column-major n x n matrix in long double (Gauss-Jordan,
// partial pivoting).
std::vector<double> inverse_ld(int n, const double* A) {
    std::vector<long double> M(size_t(n) * n), I(size_t(n) * n, 0.0L);
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) M[size_t(r) * n + c] = A[r + size_t(n) * c];  // row-major copy
    for (int i = 0; i < n; ++i) I[size_t(i) * n + i] = 1;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[size_t(i) * n + k]) > std::fabs(M[size_t(piv) * n + k])) piv = i;
        require(M[size_t(piv) * n + k] != 0.0L, HGDG_ERR_ARG, "a mass matrix is singular");
        if (piv != k)
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
    std::vector<double> out(size_t(n) * n);
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) out[r + size_t(n) * c] = static_cast<double>(I[size_t(r) * n + c]);
    return out;
}

}  // namespace

void matrix_operator_set(hgdg_ctx* c, const hgdg_matrix_desc* d, int precision) {
    require(d != nullptr && d->block_size && d->mass && d->stiffness, HGDG_ERR_ARG, "incomplete matrix description");
    require(precision == HGDG_PREC_FP64 || precision == HGDG_PREC_MIXED_FP32, HGDG_ERR_ARG, "unknown precision");
    require(d->n_elements > 0 && d->n_faces >= 0, HGDG_ERR_ARG, "element and face counts");
    require(d->n_faces == 0 || (d->face_elements && d->face_blocks), HGDG_ERR_ARG, "faces without their data");
    const int64_t ne = d->n_elements;
    MatrixOp m;
    m.elem_off.assign(size_t(ne) + 1, 0);
    m.sq_off.assign(size_t(ne), 0);
    int64_t sq = 0;
    for (int64_t e = 0; e < ne; ++e) {
        const int32_t n = d->block_size[e];
        require(n >= 1 && n <= kMatMaxRows, HGDG_ERR_UNSUPPORTED,
                "element blocks must have 1.." + std::to_string(kMatMaxRows) + " unknowns");
        m.elem_off[size_t(e) + 1] = m.elem_off[size_t(e)] + n;
        m.sq_off[size_t(e)] = sq;
        sq += int64_t(n) * n;
        m.max_rows = std::max(m.max_rows, int(n));
    }
    // element rows: the diagonal block first (S_e + the faces' own blocks, in
    // face order), then one block per internal face towards the neighbour
    std::vector<std::vector<double>> diag(static_cast<size_t>(ne));
    for (int64_t e = 0; e < ne; ++e) {
        const int n = d->block_size[e];
        diag[size_t(e)].assign(d->stiffness + m.sq_off[size_t(e)], d->stiffness + m.sq_off[size_t(e)] + int64_t(n) * n);
    }
    struct Cross {
        int64_t row, col;  // test element, trial element
        size_t at;         // offset of the block (column-major n_row x n_col) in the face data
    };
    std::vector<Cross> cross;
    size_t fo = 0;
    for (int64_t f = 0; f < d->n_faces; ++f) {
        const int64_t L = d->face_elements[2 * f], R = d->face_elements[2 * f + 1];
        require(L >= 0 && L < ne && R >= -1 && R < ne && R != L, HGDG_ERR_ARG, "bad face " + std::to_string(f));
        const int nL = d->block_size[L], nR = R >= 0 ? d->block_size[R] : 0;
        // the face's blocks, each column-major: LL (nL x nL), LR (nL x nR), RL (nR x nL), RR (nR x nR)
        const double* LL = d->face_blocks + fo;
        for (int64_t k = 0; k < int64_t(nL) * nL; ++k) diag[size_t(L)][size_t(k)] += LL[k];
        if (R >= 0) {
            const size_t lr = fo + size_t(nL) * nL, rl = lr + size_t(nL) * nR, rr = rl + size_t(nR) * nL;
            const double* RR = d->face_blocks + rr;
            for (int64_t k = 0; k < int64_t(nR) * nR; ++k) diag[size_t(R)][size_t(k)] += RR[k];
            cross.push_back({L, R, lr});
            cross.push_back({R, L, rl});
            fo = rr + size_t(nR) * nR;
        } else {
            fo += size_t(nL) * nL;
        }
    }
    std::stable_sort(cross.begin(), cross.end(), [](const Cross& a, const Cross& b) { return a.row < b.row; });
    std::vector<double> pool;
    m.row_ptr.assign(size_t(ne) + 1, 0);
    size_t ci = 0;
    for (int64_t e = 0; e < ne; ++e) {
        m.blk_off.push_back(int64_t(pool.size()));
        m.blk_col.push_back(int32_t(e));
        pool.insert(pool.end(), diag[size_t(e)].begin(), diag[size_t(e)].end());
        diag[size_t(e)].clear();
        diag[size_t(e)].shrink_to_fit();
        for (; ci < cross.size() && cross[ci].row == e; ++ci) {
            const Cross& x = cross[ci];
            const int64_t n = d->block_size[e], nt = d->block_size[x.col];
            m.blk_off.push_back(int64_t(pool.size()));
            m.blk_col.push_back(int32_t(x.col));
            pool.insert(pool.end(), d->face_blocks + x.at, d->face_blocks + x.at + n * nt);
        }
        m.row_ptr[size_t(e) + 1] = int64_t(m.blk_off.size());
    }
    std::vector<double> minv(static_cast<size_t>(sq));
    for (int64_t e = 0; e < ne; ++e) {
        const int n = d->block_size[e];
        const std::vector<double> inv = inverse_ld(n, d->mass + m.sq_off[size_t(e)]);
        std::copy(inv.begin(), inv.end(), minv.begin() + m.sq_off[size_t(e)]);
    }
    c->drop_graph();
    c->vecs.clear();
    m.d_elem_off.upload(m.elem_off, c->stream);
    m.d_sq_off.upload(m.sq_off, c->stream);
    m.d_row_ptr.upload(m.row_ptr, c->stream);
    m.d_blk_off.upload(m.blk_off, c->stream);
    m.d_blk_col.upload(m.blk_col, c->stream);
    m.d_pool.upload(pool, c->stream);
    m.d_minv.upload(minv, c->stream);
    m.d_mass.upload(std::vector<double>(d->mass, d->mass + sq), c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
    m.on = true;
    c->mat = std::move(m);
    c->n_elem = ne;
    c->n_faces = d->n_faces;
    c->n_internal = int64_t(cross.size() / 2);
    c->n_fields = std::max(1, int(d->n_fields));
    c->dof = c->mat.elem_off.back();
    c->has_mesh = true;
    c->op = hgdg_op_desc{};
    c->op.kind = HGDG_OP_MATRIX;
    c->op.precision = precision;
    c->has_op = true;
}

void matrix_rhs(hgdg_ctx* c, const StageIO& io, bool solve) {
    MatParams p{};
    p.n_elem = c->n_elem;
    p.io = io;
    p.elem_off = c->mat.d_elem_off.ptr;
    p.row_ptr = c->mat.d_row_ptr.ptr;
    p.blk_off = c->mat.d_blk_off.ptr;
    p.blk_col = c->mat.d_blk_col.ptr;
    p.pool = c->mat.d_pool.ptr;
    p.sq_off = c->mat.d_sq_off.ptr;
    p.minv = solve ? c->mat.d_minv.ptr : nullptr;
    // LDS: the largest trial block (any element's size) and the row values
    const size_t lds = sizeof(double) * size_t(c->mat.max_rows);
    hipLaunchKernelGGL(matrix_rhs, dim3(unsigned(c->n_elem)), dim3(kMatThreads), lds, c->stream, p);
    HGDG_HIP(hipGetLastError());
}

void matrix_mass_apply(hgdg_ctx* c, Vec& v, bool inverse) {
    require(!v.is_f32, HGDG_ERR_UNSUPPORTED, "mass apply/solve needs a vector stored in FP64");
    const size_t lds = sizeof(double) * size_t(c->mat.max_rows);
    hipLaunchKernelGGL(matrix_apply, dim3(unsigned(c->n_elem)), dim3(kMatThreads), lds, c->stream,
                       c->mat.d_elem_off.ptr, c->mat.d_sq_off.ptr,
                       inverse ? c->mat.d_minv.ptr : c->mat.d_mass.ptr, v.f64->ptr);
    HGDG_HIP(hipGetLastError());
}

void matrix_state_diff(hgdg_ctx* c, const Vec& x, const Vec& y, double* l2, double* max_abs) {
    if (c->d_partial.n < size_t(2 * c->n_elem)) c->d_partial.alloc(size_t(2 * c->n_elem));
    const size_t lds = sizeof(double) * size_t(c->mat.max_rows);
    hipLaunchKernelGGL(matrix_diff, dim3(unsigned(c->n_elem)), dim3(kMatThreads), lds, c->stream,
                       c->mat.d_elem_off.ptr, c->mat.d_sq_off.ptr, c->mat.d_mass.ptr, x.ptr(), int32_t(x.is_f32),
                       y.ptr(), int32_t(y.is_f32), c->d_partial.ptr);
    HGDG_HIP(hipGetLastError());
    std::vector<double> h(size_t(2 * c->n_elem));
    HGDG_HIP(hipMemcpyAsync(h.data(), c->d_partial.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
    double s = 0, m = 0;
    for (int64_t i = 0; i < c->n_elem; ++i) {
        s += h[size_t(2 * i)];
        m = std::max(m, h[size_t(2 * i + 1)]);
    }
    *l2 = std::sqrt(std::max(s, 0.0));
    *max_abs = m;
}

}  // namespace hgdg
