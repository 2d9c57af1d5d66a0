// hgdg: MI355X (gfx950) explicit DG operator behind a C ABI.
//
// Tensor elements (squares, cubes). On an affine element with constant
// advection velocity a, hpGEM's stiffness, face and mass integrals
// (kernel/Base/HpgemAPILinear_Impl.h:131-139,250-288 with the integrands of
// tests/self/Basic/080AdvectionLinear_SelfTest.cpp:108-170) factor into 1D
// operators: with the basis psi_a(x) psi_b(y) psi_c(z), M = |J| M1 x M1 x M1,
// and after the block-diagonal inverse mass apply the weak form
//
//   k = sum_d ahat_d (D along d) w - sum_{faces (d,s)} (2s-1) ahat_d Mi(:,s) (x) T_{d,s}
//
// (ahat = J^{-1} a, D = M1^{-1} G1, Mi = M1^{-1}, T_{d,s} the upwind trace
// coefficients on face x_d = 2s-1) is evaluated in its integrated-by-parts
// ("lifting") form, which is the same number up to rounding:
//
//   k = -sum_d ahat_d (Dr along d) w - sum_{faces} (2s-1) ahat_d Mi(:,s) (x) (T_{d,s} - own_{d,s})
//
// with Dr = M1^{-1} K the exact 1D derivative matrix (K(a,b) = int psi_a psi_b')
// and own_{d,s} the element's own trace. Outflow faces drop out, and the face
// terms only carry jumps, which are formed in FP64 from the FP64 stage input:
// no large terms cancel, so the FP32 variant keeps ~1e-7 relative accuracy.
// The 1D operators are built once on the host in long double by exact Gauss
// quadrature; each element needs its stage input, `dim` numbers of geometry
// and the traces of its inflow neighbours. One thread owns one coefficient;
// an element's stage input sits in LDS for the 1D contractions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hgdg.h"
#include "host_tables.hpp"

namespace hgdg {

thread_local std::string g_error;

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HGDG_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t err_ = (call);                                                             \
        if (err_ != hipSuccess)                                                               \
            throw ::hgdg::Failure(HGDG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_)); \
    } while (0)

constexpr int kMaxIn = 8;
constexpr int kBlock = 256;

// Per-face upwind modes, 3 bits per tensor face f = 2d + s.
enum : uint32_t { MODE_NONE = 0, MODE_OWN = 1, MODE_NBR = 2, MODE_AVG = 3, MODE_HALF_OWN = 4 };

struct RhsParams {
    int64_t n_elem;
    const void* in[kMaxIn];
    double coef[kMaxIn];
    uint32_t in_f32;  // bit i set: input i is stored in FP32
    int32_t n_in;
    void* out;
    int32_t out_f32;
    const double* ahat;      // [dim][n_elem]
    const int32_t* nbr;      // [2 dim][n_elem]
    const uint32_t* mode;    // [n_elem]
    const double* Dr;        // [n1][n1] exact derivative matrix
    const double* Mi;        // [n1][n1] inverse 1D mass
};

__device__ __forceinline__ double load_in(const RhsParams& p, int i, int64_t idx) {
    return ((p.in_f32 >> i) & 1u) ? double(static_cast<const float*>(p.in[i])[idx])
                                  : static_cast<const double*>(p.in[i])[idx];
}

// Stage input sum_i coef_i vec_i, first term scaled then accumulated in
// order (HpgemAPISimplified::getLinearCombinationOfVectors).
__device__ __forceinline__ double combine(const RhsParams& p, int64_t idx) {
    double s = p.coef[0] * load_in(p, 0, idx);
#pragma unroll
    for (int i = 1; i < kMaxIn; ++i)
        if (i < p.n_in) s += p.coef[i] * load_in(p, i, idx);
    return s;
}

template <int DIM, int N1>
struct Tensor {
    static constexpr int NB = N1 * N1 * (DIM == 3 ? N1 : 1);
    static constexpr int L = NB <= 4 ? 4 : NB <= 8 ? 8 : NB <= 16 ? 16 : NB <= 32 ? 32 : NB <= 64 ? 64 : NB <= 128 ? 128 : 256;
    static constexpr int EPB = kBlock / L;
};

// Residual accumulation: FP64 plainly; FP32 with a compensated (TwoSum)
// running sum for the face contributions.
template <typename T>
struct Acc {
    T s;
    __device__ Acc() : s(T(0)) {}
    __device__ explicit Acc(T v) : s(v) {}
    __device__ void add(T v) { s += v; }
    __device__ T value() const { return s; }
};
template <>
struct Acc<float> {
    float s, c;
    __device__ Acc() : s(0.f), c(0.f) {}
    __device__ explicit Acc(float v) : s(v), c(0.f) {}
    __device__ void add(float v) {
        float t = s + v;
        float bp = t - s;
        c += (s - (t - bp)) + (v - bp);
        s = t;
    }
    __device__ float value() const { return s + c; }
};

// One thread owns one x-pencil of an element (the N1 coefficients with fixed
// y/z indices): the x-contraction and the x-face terms stay in registers, the
// y/z contractions read the other pencils of the element from LDS, and the
// y/z inflow traces arrive as whole pencils of the neighbour (vector loads).
template <int DIM, int N1>
struct Pencil {
    static constexpr int NP = DIM == 3 ? N1 * N1 : N1;  // pencils per element
    static constexpr int NB = NP * N1;                   // coefficients per element
    static constexpr int EPB = kBlock / NP;              // elements per block
};

// x[a] = vector i at [base, base + N1), as double
template <int N1>
__device__ __forceinline__ void load_pencil(const RhsParams& p, int i, int64_t base, double (&x)[N1]) {
    if ((p.in_f32 >> i) & 1u) {
        const float* f = static_cast<const float*>(p.in[i]) + base;
        if constexpr (N1 % 4 == 0) {
#pragma unroll
            for (int a = 0; a < N1; a += 4) {
                const float4 v = *reinterpret_cast<const float4*>(f + a);
                x[a] = v.x; x[a + 1] = v.y; x[a + 2] = v.z; x[a + 3] = v.w;
            }
        } else if constexpr (N1 % 2 == 0) {
#pragma unroll
            for (int a = 0; a < N1; a += 2) {
                const float2 v = *reinterpret_cast<const float2*>(f + a);
                x[a] = v.x; x[a + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int a = 0; a < N1; ++a) x[a] = f[a];
        }
    } else {
        const double* d = static_cast<const double*>(p.in[i]) + base;
        if constexpr (N1 % 2 == 0) {
#pragma unroll
            for (int a = 0; a < N1; a += 2) {
                const double2 v = *reinterpret_cast<const double2*>(d + a);
                x[a] = v.x; x[a + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int a = 0; a < N1; ++a) x[a] = d[a];
        }
    }
}

// Stage input pencil sum_i coef_i vec_i (getLinearCombinationOfVectors).
template <int N1>
__device__ __forceinline__ void combine_pencil(const RhsParams& p, int64_t base, double (&w)[N1]) {
    double x[N1];
    load_pencil<N1>(p, 0, base, x);
#pragma unroll
    for (int a = 0; a < N1; ++a) w[a] = p.coef[0] * x[a];
#pragma unroll
    for (int i = 1; i < kMaxIn; ++i) {
        if (i >= p.n_in) break;
        load_pencil<N1>(p, i, base, x);
#pragma unroll
        for (int a = 0; a < N1; ++a) w[a] += p.coef[i] * x[a];
    }
}

__device__ __forceinline__ double face_jump(uint32_t m, double own, double other) {
    if (m == MODE_NONE) return -own;
    if (m == MODE_HALF_OWN) return -0.5 * own;
    return (m == MODE_NBR) ? other - own : 0.5 * (other - own);
}

template <int DIM, int N1, typename T>
__global__ __launch_bounds__(kBlock) void tensor_advection_rhs(const RhsParams p) {
    using S = Pencil<DIM, N1>;
    __shared__ double s_w[S::EPB * S::NB];  // FP64 stage input of the block's elements
    __shared__ T s_D[N1 * N1];
    __shared__ T s_Mi[2 * N1];  // s_Mi[s * N1 + a] = Mi(a, s)
    const int tid = threadIdx.x;
    if (tid < N1 * N1) s_D[tid] = T(p.Dr[tid]);
    if (tid < 2 * N1) s_Mi[tid] = T(p.Mi[(tid % N1) * N1 + tid / N1]);
    const int le = tid / S::NP;
    const int pc = tid % S::NP;
    const int64_t e = int64_t(blockIdx.x) * S::EPB + le;
    const bool active = (le < S::EPB) && (e < p.n_elem);
    const int ib = pc % N1;                    // y index of this pencil
    const int ic = DIM == 3 ? pc / N1 : 0;     // z index of this pencil
    double* sw = s_w + le * S::NB;
    double w[N1];
    if (active) {
        combine_pencil<N1>(p, e * S::NB + pc * N1, w);
#pragma unroll
        for (int a = 0; a < N1; ++a) sw[pc * N1 + a] = w[a];
    }
    __syncthreads();
    if (!active) return;

    T ah[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) ah[d] = T(p.ahat[d * p.n_elem + e]);

    // volume (strong form): -sum_d ahat_d (Dr along d) w
    T vol[N1];
#pragma unroll
    for (int a = 0; a < N1; ++a) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) s += s_D[a * N1 + j] * T(w[j]);
        vol[a] = -ah[0] * s;
    }
#pragma unroll
    for (int j = 0; j < N1; ++j) {
        const T cy = -ah[1] * s_D[ib * N1 + j];
        const double* row = sw + (j + N1 * ic) * N1;  // pencil (j, ic)
#pragma unroll
        for (int a = 0; a < N1; ++a) vol[a] += cy * T(row[a]);
    }
    if constexpr (DIM == 3) {
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            const T cz = -ah[2] * s_D[ic * N1 + j];
            const double* row = sw + (ib + N1 * j) * N1;  // pencil (ib, j)
#pragma unroll
            for (int a = 0; a < N1; ++a) vol[a] += cz * T(row[a]);
        }
    }
    Acc<T> acc[N1] = {};
#pragma unroll
    for (int a = 0; a < N1; ++a) acc[a] = Acc<T>(vol[a]);

    // faces: -(2s-1) ahat_d Mi(idx_d, s) (T_{d,s} - own_{d,s}); jumps in FP64
    const uint32_t md = p.mode[e];
#pragma unroll
    for (int s = 0; s < 2; ++s) {  // x faces: one trace value per pencil
        const uint32_t m = (md >> (3 * s)) & 7u;
        if (m == MODE_OWN) continue;
        double other = 0.0;
        if (m == MODE_NBR || m == MODE_AVG) {
            const int64_t nb = p.nbr[s * p.n_elem + e];
            double v[1];
            const int64_t idx = nb * S::NB + pc * N1 + (1 - s);
            v[0] = p.coef[0] * load_in(p, 0, idx);
            for (int i = 1; i < kMaxIn; ++i) {
                if (i >= p.n_in) break;
                v[0] += p.coef[i] * load_in(p, i, idx);
            }
            other = v[0];
        }
        const T jump = T(face_jump(m, w[s], other));
        const T sc = (s ? T(-1) : T(1)) * ah[0] * jump;
#pragma unroll
        for (int a = 0; a < N1; ++a) acc[a].add(sc * s_Mi[s * N1 + a]);
    }
#pragma unroll
    for (int d = 1; d < DIM; ++d) {  // y (and z) faces: a whole pencil of traces
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int f = 2 * d + s;
            const uint32_t m = (md >> (3 * f)) & 7u;
            if (m == MODE_OWN) continue;
            const int own_pc = d == 1 ? s + N1 * ic : ib + N1 * s;
            const int nbr_pc = d == 1 ? (1 - s) + N1 * ic : ib + N1 * (1 - s);
            double other[N1];
            if (m == MODE_NBR || m == MODE_AVG) {
                const int64_t nb = p.nbr[f * p.n_elem + e];
                combine_pencil<N1>(p, nb * S::NB + nbr_pc * N1, other);
            } else {
#pragma unroll
                for (int a = 0; a < N1; ++a) other[a] = 0.0;
            }
            const int my = d == 1 ? ib : ic;
            const T sc = (s ? T(-1) : T(1)) * ah[d] * s_Mi[s * N1 + my];
#pragma unroll
            for (int a = 0; a < N1; ++a) acc[a].add(sc * T(face_jump(m, sw[own_pc * N1 + a], other[a])));
        }
    }
    const int64_t o = e * S::NB + pc * N1;
    if (p.out_f32) {
        float* out = static_cast<float*>(p.out) + o;
#pragma unroll
        for (int a = 0; a < N1; ++a) out[a] = float(acc[a].value());
    } else {
        double* out = static_cast<double*>(p.out) + o;
#pragma unroll
        for (int a = 0; a < N1; ++a) out[a] = double(acc[a].value());
    }
}

// u[i] += sum_j c_j k_j[i], terms applied one after another in FP64
// (HpgemAPISimplified::scaleAndAddVector per stage).
struct AxpyParams {
    int64_t n;
    double* y;
    const void* x[kMaxIn];
    double c[kMaxIn];
    uint32_t x_f32;
    int32_t n_x;
};

__global__ __launch_bounds__(kBlock) void multi_axpy(const AxpyParams p) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= p.n) return;
    double y = p.y[i];
#pragma unroll
    for (int j = 0; j < kMaxIn; ++j) {
        if (j >= p.n_x) break;
        const double x = ((p.x_f32 >> j) & 1u) ? double(static_cast<const float*>(p.x[j])[i])
                                                : static_cast<const double*>(p.x[j])[i];
        y += p.c[j] * x;
    }
    p.y[i] = y;
}

// dst = src, 16 bytes per thread (plain streaming copy)
__global__ __launch_bounds__(kBlock) void copy_words(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

// Per-element mass-norm of x - y: |J| dx^T (M1 x M1 x M1) dx, and max |dx|;
// one partial per block.
struct DiffParams {
    int64_t n_elem;
    const void* x;
    const void* y;
    int32_t x_f32, y_f32;
    const double* absdet;  // [n_elem]
    const double* M1;      // [n1][n1]
    double* partial;       // [grid][2]
};

template <int DIM, int N1>
__global__ __launch_bounds__(kBlock) void tensor_diff_norm(const DiffParams p) {
    using S = Tensor<DIM, N1>;
    __shared__ double s_a[S::EPB * S::NB];
    __shared__ double s_b[S::EPB * S::NB];
    __shared__ double s_M[N1 * N1];
    __shared__ double s_red[2][kBlock];
    const int tid = threadIdx.x;
    if (tid < N1 * N1) s_M[tid] = p.M1[tid];
    const int le = tid / S::L;
    const int dof = tid % S::L;
    const int64_t e = int64_t(blockIdx.x) * S::EPB + le;
    const bool active = (e < p.n_elem) && (dof < S::NB);
    double dx = 0;
    if (active) {
        const int64_t i = e * S::NB + dof;
        const double xv = p.x_f32 ? double(static_cast<const float*>(p.x)[i]) : static_cast<const double*>(p.x)[i];
        const double yv = p.y_f32 ? double(static_cast<const float*>(p.y)[i]) : static_cast<const double*>(p.y)[i];
        dx = xv - yv;
        s_a[le * S::NB + dof] = dx;
    }
    __syncthreads();
    const int idx[3] = {dof % N1, (dof / N1) % N1, DIM == 3 ? dof / (N1 * N1) : 0};
    constexpr int stride[3] = {1, N1, N1 * N1};
    double* src = s_a;
    double* dst = s_b;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        if (active) {
            const int row = dof - idx[d] * stride[d];
            double s = 0;
            for (int j = 0; j < N1; ++j) s += s_M[idx[d] * N1 + j] * src[le * S::NB + row + j * stride[d]];
            dst[le * S::NB + dof] = s;
        }
        __syncthreads();
        double* t = src;
        src = dst;
        dst = t;
    }
    double v = 0, m = 0;
    if (active) {
        v = dx * src[le * S::NB + dof] * p.absdet[e];
        m = fabs(dx);
    }
    s_red[0][tid] = v;
    s_red[1][tid] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            s_red[0][tid] += s_red[0][tid + w];
            s_red[1][tid] = fmax(s_red[1][tid], s_red[1][tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        p.partial[2 * blockIdx.x] = s_red[0][0];
        p.partial[2 * blockIdx.x + 1] = s_red[1][0];
    }
}

// ---------------------------------------------------------------- dispatch
template <typename F>
void dispatch_tensor(int dim, int n1, F&& f) {
#define HGDG_CASE(D, N) \
    if (dim == D && n1 == N) return f(std::integral_constant<int, D>{}, std::integral_constant<int, N>{});
    HGDG_CASE(2, 2) HGDG_CASE(2, 3) HGDG_CASE(2, 4) HGDG_CASE(2, 5) HGDG_CASE(2, 6)
    HGDG_CASE(3, 2) HGDG_CASE(3, 3) HGDG_CASE(3, 4) HGDG_CASE(3, 5) HGDG_CASE(3, 6)
#undef HGDG_CASE
    throw Failure(HGDG_ERR_UNSUPPORTED, "no tensor kernel for this dimension and order");
}

// ---------------------------------------------------------------- context
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        reset();
        if (count == 0) return;
        hipError_t err = hipMalloc(&ptr, count * sizeof(T));
        if (err != hipSuccess) {
            ptr = nullptr;
            throw Failure(HGDG_ERR_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                               " bytes failed: " + hipGetErrorString(err));
        }
        n = count;
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (!h.empty()) HGDG_HIP(hipMemcpyAsync(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct Vec {
    std::unique_ptr<DevBuf<double>> f64;
    std::unique_ptr<DevBuf<float>> f32;
    bool is_f32 = false;
    void* ptr() const { return is_f32 ? static_cast<void*>(f32->ptr) : static_cast<void*>(f64->ptr); }
};

struct GraphKey {
    int solution;
    std::vector<int> aux;
    std::vector<double> a, b;
    double dt;
    bool operator==(const GraphKey& o) const {
        return solution == o.solution && aux == o.aux && a == o.a && b == o.b && dt == o.dt;
    }
};

}  // namespace hgdg

struct hgdg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh
    bool has_mesh = false;
    int shape = 0, dim = 0, p = 0, n1 = 0, nb = 0, n_fields = 1;
    int64_t n_elem = 0, n_faces = 0, n_internal = 0;
    std::vector<int> perm;              // reference basis index -> tensor index
    std::vector<double> Jinv;           // [e][dim][dim]
    std::vector<double> X0, E;          // [e][dim] first vertex, [e][dim][dim] edge vectors
    std::vector<double> absdet;         // [e]
    std::vector<int32_t> nbr;           // [2 dim][e]
    hgdg::TensorOps ops;
    hgdg::DevBuf<int32_t> d_nbr;
    hgdg::DevBuf<double> d_absdet, d_ops;  // d_ops: Dr then Mi then M1
    // operator
    bool has_op = false;
    hgdg_op_desc op{};
    hgdg::DevBuf<double> d_ahat;
    hgdg::DevBuf<uint32_t> d_mode;
    // state
    std::vector<hgdg::Vec> vecs;
    hgdg::DevBuf<double> d_partial;
    // graph cache
    bool graph_valid = false;
    hgdg::GraphKey graph_key{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    int64_t n_dof() const { return n_elem * nb * n_fields; }
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_valid = false;
    }
};

namespace hgdg {

void require(bool ok, int code, const std::string& msg) {
    if (!ok) throw Failure(code, msg);
}

const double* Dr_ptr(hgdg_ctx* c) { return c->d_ops.ptr; }
const double* Mi_ptr(hgdg_ctx* c) { return c->d_ops.ptr + c->n1 * c->n1; }
const double* M1_ptr(hgdg_ctx* c) { return c->d_ops.ptr + 2 * c->n1 * c->n1; }

// Tensor face f = 2d + s <-> hpGEM local face number
// (ReferenceCube.cpp:52-54, ReferenceSquare.cpp faces 0..3).
int tensor_face_of_ref(int dim, int ref_face) {
    static const int cube[6] = {2 * 2 + 0, 2 * 1 + 0, 2 * 0 + 0, 2 * 0 + 1, 2 * 1 + 1, 2 * 2 + 1};
    static const int square[4] = {2 * 1 + 0, 2 * 0 + 0, 2 * 0 + 1, 2 * 1 + 1};
    return dim == 3 ? cube[ref_face] : square[ref_face];
}

void mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m) {
    require(m != nullptr, HGDG_ERR_ARG, "mesh descriptor is NULL");
    require(m->shape == SHAPE_SQUARE || m->shape == SHAPE_CUBE || m->shape == SHAPE_TETRAHEDRON, HGDG_ERR_ARG,
            "unknown element shape");
    require(m->shape != SHAPE_TETRAHEDRON, HGDG_ERR_UNSUPPORTED, "tetrahedral meshes are not supported yet");
    const int dim = shape_dim(m->shape);
    require(m->dim == dim, HGDG_ERR_ARG, "dim does not match the element shape");
    require(m->n_elements > 0 && m->n_elements < (int64_t(1) << 31), HGDG_ERR_ARG, "element count out of range");
    require(m->coords && m->node_ids, HGDG_ERR_ARG, "coords and node_ids are required");
    const int64_t ne = m->n_elements;
    int p = m->order ? m->order[0] : m->uniform_order;
    if (m->order)
        for (int64_t e = 0; e < ne; ++e)
            require(m->order[e] == p, HGDG_ERR_UNSUPPORTED,
                    "tensor elements need one polynomial order per mesh");
    require(p >= 1 && p <= 5, HGDG_ERR_UNSUPPORTED, "polynomial order must be in 1..5");
    const int nv = shape_vertices(m->shape);

    c->drop_graph();
    c->has_mesh = false;
    c->has_op = false;
    c->vecs.clear();
    c->shape = m->shape;
    c->dim = dim;
    c->p = p;
    c->n1 = p + 1;
    c->nb = dim == 3 ? c->n1 * c->n1 * c->n1 : c->n1 * c->n1;
    c->n_fields = 1;
    c->n_elem = ne;
    c->perm = tensor_permutation(dim, p);

    // geometry: affine hypercubes only; J column d = (X_{2^d} - X_0) / 2
    c->Jinv.assign(size_t(ne) * dim * dim, 0.0);
    c->absdet.assign(ne, 0.0);
    c->X0.assign(size_t(ne) * dim, 0.0);
    c->E.assign(size_t(ne) * dim * dim, 0.0);
    for (int64_t e = 0; e < ne; ++e) {
        const double* X = m->coords + size_t(e) * nv * dim;
        double E[3][3] = {{0}};
        double scale = 0;
        for (int d = 0; d < dim; ++d)
            for (int r = 0; r < dim; ++r) {
                E[d][r] = X[(1 << d) * dim + r] - X[r];
                scale = std::max(scale, std::fabs(E[d][r]));
            }
        for (int d = 0; d < dim; ++d) {
            c->X0[size_t(e) * dim + d] = X[d];
            for (int r = 0; r < dim; ++r) c->E[(size_t(e) * dim + d) * dim + r] = E[d][r];
        }
        for (int v = 0; v < nv; ++v)
            for (int r = 0; r < dim; ++r) {
                double pred = X[r];
                for (int d = 0; d < dim; ++d)
                    if ((v >> d) & 1) pred += E[d][r];
                require(std::fabs(pred - X[v * dim + r]) <= 1e-12 * scale, HGDG_ERR_UNSUPPORTED,
                        "element " + std::to_string(e) + " is not affine (parallelogram/parallelepiped)");
            }
        double J[3][3] = {{0}};  // J[r][c] = d x_r / d xi_c
        for (int r = 0; r < dim; ++r)
            for (int col = 0; col < dim; ++col) J[r][col] = E[col][r] / 2;
        double det, inv[3][3];
        if (dim == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            inv[0][0] = J[1][1] / det;
            inv[0][1] = -J[0][1] / det;
            inv[1][0] = -J[1][0] / det;
            inv[1][1] = J[0][0] / det;
        } else {
            det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            for (int r = 0; r < 3; ++r)
                for (int col = 0; col < 3; ++col) {
                    const int r1 = (col + 1) % 3, r2 = (col + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                    inv[r][col] = (J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1]) / det;
                }
        }
        require(det != 0.0 && std::isfinite(det), HGDG_ERR_ARG, "degenerate element " + std::to_string(e));
        c->absdet[e] = std::fabs(det);
        for (int r = 0; r < dim; ++r)
            for (int col = 0; col < dim; ++col) c->Jinv[(size_t(e) * dim + r) * dim + col] = inv[r][col];
    }

    // faces: match sorted node tuples; the lower element index is hpGEM's
    // left element (faces are created while elements are visited in order).
    const int nfaces = 2 * dim;
    const int fverts = 1 << (dim - 1);
    c->nbr.assign(size_t(nfaces) * ne, -1);
    std::unordered_map<std::string, std::pair<int64_t, int>> open;
    open.reserve(size_t(ne) * nfaces);
    int64_t n_internal = 0, n_total = 0;
    auto face_vertex = [&](int f, int q) {
        // local vertex of tensor face f = 2d + s: bit d = s, the other bits from q in order
        const int d = f / 2, s = f % 2;
        int v = 0, bit = 0;
        for (int k = 0; k < dim; ++k) {
            if (k == d) {
                v |= s << k;
            } else {
                v |= ((q >> bit) & 1) << k;
                ++bit;
            }
        }
        return v;
    };
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t* nid = m->node_ids + size_t(e) * nv;
        for (int f = 0; f < nfaces; ++f) {
            std::array<int64_t, 4> key{};
            for (int q = 0; q < fverts; ++q) key[q] = nid[face_vertex(f, q)];
            std::sort(key.begin(), key.begin() + fverts);
            std::string k(reinterpret_cast<const char*>(key.data()), sizeof(int64_t) * fverts);
            auto it = open.find(k);
            if (it == open.end()) {
                open.emplace(std::move(k), std::make_pair(e, f));
                continue;
            }
            const int64_t o = it->second.first;
            const int fo = it->second.second;
            require(o != e, HGDG_ERR_UNSUPPORTED, "element " + std::to_string(e) + " meets itself across a face");
            require(c->nbr[size_t(fo) * ne + o] < 0, HGDG_ERR_ARG, "a face is bounded by more than two elements");
            // tensor kernels need the neighbour's opposite face in the same
            // orientation: my face (d, s) against its (d, 1-s), vertex by vertex
            require(fo / 2 == f / 2 && fo % 2 != f % 2, HGDG_ERR_UNSUPPORTED,
                    "neighbouring elements are not aligned (different local face directions)");
            const int64_t* onid = m->node_ids + size_t(o) * nv;
            for (int q = 0; q < fverts; ++q)
                require(onid[face_vertex(fo, q)] == nid[face_vertex(f, q)], HGDG_ERR_UNSUPPORTED,
                        "neighbouring elements are not aligned (rotated face)");
            c->nbr[size_t(f) * ne + e] = int32_t(o);
            c->nbr[size_t(fo) * ne + o] = int32_t(e);
            open.erase(it);
            ++n_internal;
            ++n_total;
        }
    }
    n_total += int64_t(open.size());
    c->n_faces = n_total;
    c->n_internal = n_internal;

    c->ops = tensor_ops(p);
    std::vector<double> opsv;
    opsv.insert(opsv.end(), c->ops.Dr.begin(), c->ops.Dr.end());
    opsv.insert(opsv.end(), c->ops.Mi.begin(), c->ops.Mi.end());
    opsv.insert(opsv.end(), c->ops.M1.begin(), c->ops.M1.end());
    c->d_ops.upload(opsv, c->stream);
    c->d_nbr.upload(c->nbr, c->stream);
    c->d_absdet.upload(c->absdet, c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
    c->has_mesh = true;
}

void operator_set(hgdg_ctx* c, const hgdg_op_desc* op) {
    require(c->has_mesh, HGDG_ERR_STATE, "upload a mesh before setting the operator");
    require(op != nullptr, HGDG_ERR_ARG, "operator descriptor is NULL");
    require(op->kind == HGDG_OP_ADVECTION_UPWIND || op->kind == HGDG_OP_ACOUSTIC_CENTRAL, HGDG_ERR_ARG,
            "unknown operator kind");
    require(op->kind == HGDG_OP_ADVECTION_UPWIND, HGDG_ERR_UNSUPPORTED,
            "the acoustic operator is not available on tensor elements");
    require(op->precision == HGDG_PREC_FP64 || op->precision == HGDG_PREC_MIXED_FP32, HGDG_ERR_ARG,
            "unknown precision");
    const int dim = c->dim;
    const int64_t ne = c->n_elem;
    std::vector<double> ahat(size_t(dim) * ne);
    std::vector<double> rownorm(size_t(dim) * ne);
    for (int64_t e = 0; e < ne; ++e)
        for (int d = 0; d < dim; ++d) {
            double s = 0, n2 = 0;
            for (int k = 0; k < dim; ++k) {
                const double v = c->Jinv[(size_t(e) * dim + d) * dim + k];
                s += v * op->velocity[k];
                n2 += v * v;
            }
            ahat[size_t(d) * ne + e] = s;
            rownorm[size_t(d) * ne + e] = std::sqrt(n2);
        }
    // a . n_unit on face (d, s) of element e: outward normal (2s-1) J^{-T} e_d
    auto a_dot_n = [&](int64_t e, int f) {
        const int d = f / 2, s = f % 2;
        return (2 * s - 1) * ahat[size_t(d) * ne + e] / rownorm[size_t(d) * ne + e];
    };
    std::vector<uint32_t> mode(ne, 0);
    for (int64_t e = 0; e < ne; ++e) {
        uint32_t word = 0;
        for (int f = 0; f < 2 * dim; ++f) {
            const int32_t o = c->nbr[size_t(f) * ne + e];
            uint32_t m;
            if (o < 0) {
                // boundary: only the left (own) trial functions exist
                const double A = a_dot_n(e, f);
                m = A > 1e-12 ? MODE_OWN : (A < -1e-12 ? MODE_NONE : MODE_HALF_OWN);
            } else {
                // the left element's normal decides the upwind side
                const bool left = e < o;
                const double A = left ? a_dot_n(e, f) : a_dot_n(o, f ^ 1);
                if (A > 1e-12) m = left ? MODE_OWN : MODE_NBR;
                else if (A < -1e-12) m = left ? MODE_NBR : MODE_OWN;
                else m = MODE_AVG;
            }
            word |= m << (3 * f);
        }
        mode[e] = word;
    }
    c->d_ahat.upload(ahat, c->stream);
    c->d_mode.upload(mode, c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
    c->op = *op;
    c->op.c_inv = nullptr;
    c->has_op = true;
    c->drop_graph();
}

Vec& vec_at(hgdg_ctx* c, int id) {
    require(id >= 0 && id < int(c->vecs.size()), HGDG_ERR_ARG, "vector id " + std::to_string(id) + " out of range");
    return c->vecs[id];
}

void set_storage(hgdg_ctx* c, Vec& v, bool f32) {
    if (v.is_f32 == f32 && (f32 ? v.f32 != nullptr : v.f64 != nullptr)) return;
    const size_t n = size_t(c->n_dof());
    if (f32) {
        v.f32.reset(new DevBuf<float>);
        v.f32->alloc(n);
        HGDG_HIP(hipMemsetAsync(v.f32->ptr, 0, n * sizeof(float), c->stream));
        v.f64.reset();
    } else {
        v.f64.reset(new DevBuf<double>);
        v.f64->alloc(n);
        HGDG_HIP(hipMemsetAsync(v.f64->ptr, 0, n * sizeof(double), c->stream));
        v.f32.reset();
    }
    v.is_f32 = f32;
}

void launch_rhs(hgdg_ctx* c, const int* in_ids, const double* coeffs, int n_in, int out_id) {
    require(c->has_op, HGDG_ERR_STATE, "set an operator before computing time derivatives");
    require(n_in >= 1 && n_in <= kMaxIn, HGDG_ERR_ARG, "between 1 and 8 input vectors");
    RhsParams p{};
    p.n_elem = c->n_elem;
    p.n_in = n_in;
    for (int i = 0; i < n_in; ++i) {
        require(in_ids[i] != out_id, HGDG_ERR_ARG, "the output vector must differ from the inputs");
        Vec& v = vec_at(c, in_ids[i]);
        p.in[i] = v.ptr();
        p.coef[i] = coeffs[i];
        if (v.is_f32) p.in_f32 |= 1u << i;
    }
    Vec& out = vec_at(c, out_id);
    p.out = out.ptr();
    p.out_f32 = out.is_f32 ? 1 : 0;
    p.ahat = c->d_ahat.ptr;
    p.nbr = c->d_nbr.ptr;
    p.mode = c->d_mode.ptr;
    p.Dr = Dr_ptr(c);
    p.Mi = Mi_ptr(c);
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    dispatch_tensor(c->dim, c->n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = Pencil<DIMc, N1c>;
        const int64_t grid = (c->n_elem + S::EPB - 1) / S::EPB;
        if (mixed)
            hipLaunchKernelGGL((tensor_advection_rhs<DIMc, N1c, float>), dim3(unsigned(grid)), dim3(kBlock), 0,
                               c->stream, p);
        else
            hipLaunchKernelGGL((tensor_advection_rhs<DIMc, N1c, double>), dim3(unsigned(grid)), dim3(kBlock), 0,
                               c->stream, p);
    });
    HGDG_HIP(hipGetLastError());
}

void launch_axpy(hgdg_ctx* c, int to, const std::vector<int>& add, const std::vector<double>& scale) {
    require(add.size() <= size_t(kMaxIn), HGDG_ERR_ARG, "at most 8 terms per update");
    Vec& y = vec_at(c, to);
    require(!y.is_f32, HGDG_ERR_ARG, "the updated vector must be stored in FP64");
    AxpyParams p{};
    p.n = c->n_dof();
    p.y = y.f64->ptr;
    p.n_x = int32_t(add.size());
    for (size_t j = 0; j < add.size(); ++j) {
        Vec& x = vec_at(c, add[j]);
        p.x[j] = x.ptr();
        p.c[j] = scale[j];
        if (x.is_f32) p.x_f32 |= 1u << j;
    }
    const int64_t grid = (p.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(multi_axpy, dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
}

void rk_steps(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b, double dt,
              int n_steps) {
    require(c->has_op, HGDG_ERR_STATE, "set an operator before time stepping");
    require(stages >= 1 && stages <= kMaxIn - 1, HGDG_ERR_ARG, "1..7 stages supported");
    require(n_steps >= 0, HGDG_ERR_ARG, "negative step count");
    Vec& u = vec_at(c, solution);
    require(!u.is_f32, HGDG_ERR_ARG, "the solution vector must be stored in FP64");
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    for (int i = 0; i < stages; ++i) {
        require(aux[i] != solution, HGDG_ERR_ARG, "stage vectors must differ from the solution");
        for (int j = 0; j < i; ++j) require(aux[i] != aux[j], HGDG_ERR_ARG, "stage vectors must be distinct");
        set_storage(c, vec_at(c, aux[i]), mixed);
    }
    GraphKey key{solution, std::vector<int>(aux, aux + stages), std::vector<double>(a, a + stages * stages),
                 std::vector<double>(b, b + stages), dt};
    if (!(c->graph_valid && c->graph_key == key)) {
        c->drop_graph();
        HGDG_HIP(hipStreamSynchronize(c->stream));
        HGDG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int i = 0; i < stages; ++i) {
                std::vector<int> ids{solution};
                std::vector<double> co{1.0};
                // zero tableau entries contribute exactly nothing; they are skipped
                for (int j = 0; j < i; ++j)
                    if (a[i * stages + j] != 0.0) {
                        ids.push_back(aux[j]);
                        co.push_back(dt * a[i * stages + j]);
                    }
                launch_rhs(c, ids.data(), co.data(), int(ids.size()), aux[i]);
            }
            std::vector<int> add;
            std::vector<double> sc;
            for (int j = 0; j < stages; ++j)
                if (b[j] != 0.0) {
                    add.push_back(aux[j]);
                    sc.push_back(dt * b[j]);
                }
            if (!add.empty()) launch_axpy(c, solution, add, sc);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HGDG_HIP(hipStreamEndCapture(c->stream, &c->graph));
        HGDG_HIP(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->graph_key = key;
        c->graph_valid = true;
    }
    for (int s = 0; s < n_steps; ++s) HGDG_HIP(hipGraphLaunch(c->graph_exec, c->stream));
}

void state_diff(hgdg_ctx* c, int x, int y, double* l2, double* max_abs) {
    require(c->has_mesh, HGDG_ERR_STATE, "no mesh");
    Vec& vx = vec_at(c, x);
    Vec& vy = vec_at(c, y);
    DiffParams p{};
    p.n_elem = c->n_elem;
    p.x = vx.ptr();
    p.y = vy.ptr();
    p.x_f32 = vx.is_f32;
    p.y_f32 = vy.is_f32;
    p.absdet = c->d_absdet.ptr;
    p.M1 = M1_ptr(c);
    int64_t grid = 0;
    dispatch_tensor(c->dim, c->n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = Tensor<DIMc, N1c>;
        grid = (c->n_elem + S::EPB - 1) / S::EPB;
        if (c->d_partial.n < size_t(2 * grid)) c->d_partial.alloc(size_t(2 * grid));
        p.partial = c->d_partial.ptr;
        hipLaunchKernelGGL((tensor_diff_norm<DIMc, N1c>), dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
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

// L2 projection of f(x) = sum_m amp_m sin(2 pi k_m . x) onto the DG space
// (HpgemAPISimplified::setInitialSolution = integrateInitialSolution +
// solveMassMatrixEquations, HpgemAPISimplified_Impl.h:227-291,751-758), with
// the element rule of order 2p+1 (Element.cpp:89-98): on an affine tensor
// element the coefficients are (Mi B^T W)^{(x) dim} applied to f at the
// Gauss points, B(q, a) = psi_a(x_q).
void project_modes(hgdg_ctx* c, int vec, int n_modes, const double* amp, const double* kv) {
    require(c->has_mesh, HGDG_ERR_STATE, "no mesh");
    Vec& v = vec_at(c, vec);
    const int n1 = c->n1, dim = c->dim, nb = c->nb;
    std::vector<double> xq, wq;
    gauss_legendre(n1, xq, wq);
    // P(a, q) = sum_b Mi(a, b) psi_b(x_q) w_q
    std::vector<double> P(size_t(n1) * n1, 0.0);
    for (int a = 0; a < n1; ++a)
        for (int q = 0; q < n1; ++q) {
            double s = 0;
            for (int b = 0; b < n1; ++b) s += c->ops.Mi[a * n1 + b] * psi(b, xq[q]);
            P[a * n1 + q] = s * wq[q];
        }
    const int nq = nb;  // n1^dim points
    std::vector<double> out(size_t(c->n_dof()));
    std::vector<double> F(nq), T1(nq), T2(nq);
    for (int64_t e = 0; e < c->n_elem; ++e) {
        const double* X0 = c->X0.data() + size_t(e) * dim;
        const double* E = c->E.data() + size_t(e) * dim * dim;
        for (int q = 0; q < nq; ++q) {
            int r = q;
            double x[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) x[d] = X0[d];
            for (int d = 0; d < dim; ++d) {
                const double t = (xq[r % n1] + 1) / 2;
                r /= n1;
                for (int k = 0; k < dim; ++k) x[k] += t * E[d * dim + k];
            }
            double f = 0;
            for (int m = 0; m < n_modes; ++m) {
                double arg = 0;
                for (int d = 0; d < dim; ++d) arg += kv[m * 3 + d] * x[d];
                f += amp[m] * std::sin(2 * M_PI * arg);
            }
            F[q] = f;
        }
        // apply P along each direction (sum factorisation)
        std::vector<double>* src = &F;
        std::vector<double>* dst = &T1;
        int stride = 1;
        for (int d = 0; d < dim; ++d) {
            for (int i = 0; i < nq; ++i) {
                const int id = (i / stride) % n1;
                const int base = i - id * stride;
                double s = 0;
                for (int q = 0; q < n1; ++q) s += P[id * n1 + q] * (*src)[base + q * stride];
                (*dst)[i] = s;
            }
            src = dst;
            dst = (dst == &T1) ? &T2 : &T1;
            stride *= n1;
        }
        std::copy(src->begin(), src->end(), out.begin() + size_t(e) * nb);
    }
    if (v.is_f32) {
        std::vector<float> f(out.begin(), out.end());
        HGDG_HIP(hipMemcpyAsync(v.f32->ptr, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HGDG_HIP(hipStreamSynchronize(c->stream));
    } else {
        HGDG_HIP(hipMemcpyAsync(v.f64->ptr, out.data(), out.size() * sizeof(double), hipMemcpyHostToDevice,
                                c->stream));
        HGDG_HIP(hipStreamSynchronize(c->stream));
    }
}

int run(const std::function<void()>& f) {
    try {
        f();
        return HGDG_OK;
    } catch (const Failure& ex) {
        g_error = ex.what();
        return ex.code;
    } catch (const std::invalid_argument& ex) {
        g_error = ex.what();
        return HGDG_ERR_ARG;
    } catch (const std::bad_alloc&) {
        g_error = "host allocation failed";
        return HGDG_ERR_MEMORY;
    } catch (const std::exception& ex) {
        g_error = ex.what();
        return HGDG_ERR_ARG;
    }
}

}  // namespace hgdg

using hgdg::require;
using hgdg::run;

extern "C" {

const char* hgdg_last_error(void) { return hgdg::g_error.c_str(); }

int hgdg_ctx_create(hgdg_ctx** out, int device) {
    return run([&] {
        require(out != nullptr, HGDG_ERR_ARG, "out is NULL");
        *out = nullptr;
        int count = 0;
        HGDG_HIP(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, HGDG_ERR_ARG, "no HIP device " + std::to_string(device));
        HGDG_HIP(hipSetDevice(device));
        std::unique_ptr<hgdg_ctx> c(new hgdg_ctx);
        c->device = device;
        HGDG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        *out = c.release();
    });
}

int hgdg_ctx_destroy(hgdg_ctx* ctx) {
    return run([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->drop_graph();
        ctx->vecs.clear();
        hipStream_t s = ctx->stream;
        delete ctx;
        if (s) (void)hipStreamDestroy(s);
    });
}

int hgdg_mesh_upload(hgdg_ctx* ctx, const hgdg_mesh_desc* mesh) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::mesh_upload(ctx, mesh);
    });
}

int hgdg_mesh_info(hgdg_ctx* ctx, int64_t* n_elements, int64_t* n_faces, int64_t* n_internal_faces,
                   int64_t* n_dof) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(ctx->has_mesh, HGDG_ERR_STATE, "no mesh uploaded");
        if (n_elements) *n_elements = ctx->n_elem;
        if (n_faces) *n_faces = ctx->n_faces;
        if (n_internal_faces) *n_internal_faces = ctx->n_internal;
        if (n_dof) *n_dof = ctx->n_dof();
    });
}

int hgdg_operator_set(hgdg_ctx* ctx, const hgdg_op_desc* op) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::operator_set(ctx, op);
    });
}

int hgdg_state_alloc(hgdg_ctx* ctx, int n_vectors) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(ctx->has_mesh, HGDG_ERR_STATE, "upload a mesh before allocating state");
        require(n_vectors >= 0 && n_vectors <= 64, HGDG_ERR_ARG, "0..64 vectors");
        HGDG_HIP(hipSetDevice(ctx->device));
        ctx->drop_graph();
        ctx->vecs.clear();
        ctx->vecs.resize(n_vectors);
        for (auto& v : ctx->vecs) hgdg::set_storage(ctx, v, false);
        HGDG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int hgdg_state_set(hgdg_ctx* ctx, int vec, const double* host) {
    return run([&] {
        require(ctx != nullptr && host != nullptr, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        const int64_t n = ctx->n_dof();
        const int nb = ctx->nb;
        std::vector<double> dev(size_t(n), 0.0);
        for (int64_t e = 0; e < ctx->n_elem; ++e)
            for (int i = 0; i < nb; ++i) dev[size_t(e) * nb + ctx->perm[i]] = host[size_t(e) * nb + i];
        if (v.is_f32) {
            std::vector<float> f(dev.begin(), dev.end());
            HGDG_HIP(hipMemcpyAsync(v.f32->ptr, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice,
                                    ctx->stream));
            HGDG_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            HGDG_HIP(hipMemcpyAsync(v.f64->ptr, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
            HGDG_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int hgdg_state_get(hgdg_ctx* ctx, int vec, double* host) {
    return run([&] {
        require(ctx != nullptr && host != nullptr, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        const int64_t n = ctx->n_dof();
        const int nb = ctx->nb;
        std::vector<double> dev(static_cast<size_t>(n));
        if (v.is_f32) {
            std::vector<float> f(static_cast<size_t>(n));
            HGDG_HIP(hipMemcpyAsync(f.data(), v.f32->ptr, f.size() * sizeof(float), hipMemcpyDeviceToHost,
                                    ctx->stream));
            HGDG_HIP(hipStreamSynchronize(ctx->stream));
            std::copy(f.begin(), f.end(), dev.begin());
        } else {
            HGDG_HIP(hipMemcpyAsync(dev.data(), v.f64->ptr, dev.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
            HGDG_HIP(hipStreamSynchronize(ctx->stream));
        }
        for (int64_t e = 0; e < ctx->n_elem; ++e)
            for (int i = 0; i < nb; ++i) host[size_t(e) * nb + i] = dev[size_t(e) * nb + ctx->perm[i]];
    });
}

int hgdg_state_set_storage(hgdg_ctx* ctx, int vec, int precision) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(precision == HGDG_PREC_FP64 || precision == HGDG_PREC_MIXED_FP32, HGDG_ERR_ARG, "unknown precision");
        HGDG_HIP(hipSetDevice(ctx->device));
        ctx->drop_graph();
        hgdg::set_storage(ctx, hgdg::vec_at(ctx, vec), precision == HGDG_PREC_MIXED_FP32);
        HGDG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int hgdg_time_derivative(hgdg_ctx* ctx, const int* in_ids, const double* coeffs, int n_in, int out_id,
                         double time) {
    (void)time;  // the operators are autonomous
    return run([&] {
        require(ctx != nullptr && in_ids != nullptr && coeffs != nullptr, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::launch_rhs(ctx, in_ids, coeffs, n_in, out_id);
    });
}

int hgdg_scale_and_add(hgdg_ctx* ctx, int to, int add, double scale) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::launch_axpy(ctx, to, {add}, {scale});
    });
}

int hgdg_rk_steps(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* a, const double* b,
                  double dt, int n_steps) {
    return run([&] {
        require(ctx != nullptr && aux_ids && a && b, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::rk_steps(ctx, solution, aux_ids, stages, a, b, dt, n_steps);
    });
}

int hgdg_state_diff(hgdg_ctx* ctx, int x, int y, double* l2, double* max_abs) {
    return run([&] {
        require(ctx != nullptr && l2 && max_abs, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::state_diff(ctx, x, y, l2, max_abs);
    });
}

int hgdg_project_modes(hgdg_ctx* ctx, int vec, int n_modes, const double* amp, const double* k) {
    return run([&] {
        require(ctx != nullptr && (n_modes == 0 || (amp && k)), HGDG_ERR_ARG, "NULL argument");
        require(n_modes >= 0, HGDG_ERR_ARG, "negative mode count");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::project_modes(ctx, vec, n_modes, amp, k);
    });
}

int hgdg_copy(hgdg_ctx* ctx, int dst, int src) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(dst != src, HGDG_ERR_ARG, "copy onto itself");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::Vec& d = hgdg::vec_at(ctx, dst);
        hgdg::Vec& sv = hgdg::vec_at(ctx, src);
        require(d.is_f32 == sv.is_f32, HGDG_ERR_ARG, "copy between vectors of different storage");
        const int64_t bytes = ctx->n_dof() * (sv.is_f32 ? 4 : 8);
        const int64_t n16 = bytes / 16;
        const int64_t tail = bytes - n16 * 16;
        if (n16 > 0) {
            const int64_t grid = (n16 + hgdg::kBlock - 1) / hgdg::kBlock;
            hipLaunchKernelGGL(hgdg::copy_words, dim3(unsigned(grid)), dim3(hgdg::kBlock), 0, ctx->stream,
                               static_cast<const uint4*>(sv.ptr()), static_cast<uint4*>(d.ptr()), n16);
            HGDG_HIP(hipGetLastError());
        }
        if (tail > 0)
            HGDG_HIP(hipMemcpyAsync(static_cast<char*>(d.ptr()) + n16 * 16,
                                    static_cast<const char*>(sv.ptr()) + n16 * 16, size_t(tail),
                                    hipMemcpyDeviceToDevice, ctx->stream));
    });
}

int hgdg_sync(hgdg_ctx* ctx) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        HGDG_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int hgdg_get_stream(hgdg_ctx* ctx, void** stream) {
    return run([&] {
        require(ctx != nullptr && stream != nullptr, HGDG_ERR_ARG, "NULL argument");
        *stream = static_cast<void*>(ctx->stream);
    });
}

int hgdg_basis_size(int shape, int p) {
    if (p < 1) return -1;
    if (shape == hgdg::SHAPE_SQUARE) return (p + 1) * (p + 1);
    if (shape == hgdg::SHAPE_CUBE) return (p + 1) * (p + 1) * (p + 1);
    if (shape == hgdg::SHAPE_TETRAHEDRON) return (p + 1) * (p + 2) * (p + 3) / 6;
    return -1;
}

int hgdg_basis_eval(int shape, int p, const double* point, double* values) {
    return run([&] {
        require(point && values, HGDG_ERR_ARG, "NULL argument");
        require(shape == hgdg::SHAPE_SQUARE || shape == hgdg::SHAPE_CUBE, HGDG_ERR_UNSUPPORTED,
                "basis evaluation is available for squares and cubes");
        require(p >= 1 && p <= 5, HGDG_ERR_UNSUPPORTED, "order must be in 1..5");
        const int dim = hgdg::shape_dim(shape);
        const int n1 = p + 1;
        std::vector<int> perm = hgdg::tensor_permutation(dim, p);
        for (size_t i = 0; i < perm.size(); ++i) {
            int t = perm[i];
            double v = 1;
            for (int d = 0; d < dim; ++d) {
                v *= hgdg::psi(t % n1, point[d]);
                t /= n1;
            }
            values[i] = v;
        }
    });
}

int hgdg_structured_mesh(int dim, const int32_t* n, const double* lo, const double* hi, const int32_t* periodic,
                         int simplices, int64_t* n_elements, int32_t* n_vertices, int32_t* shape, double* coords,
                         int64_t* node_ids) {
    return run([&] {
        require(n && lo && hi && periodic && n_elements && n_vertices && shape, HGDG_ERR_ARG, "NULL argument");
        hgdg::StructuredMesh m = hgdg::structured_mesh(dim, n, lo, hi, periodic, simplices != 0);
        *n_vertices = m.nv;
        *shape = m.shape;
        *n_elements = int64_t(m.nodes.size() / size_t(m.nv));
        if (coords) {
            require(node_ids != nullptr, HGDG_ERR_ARG, "node_ids is NULL");
            std::copy(m.coords.begin(), m.coords.end(), coords);
            std::copy(m.nodes.begin(), m.nodes.end(), node_ids);
        }
    });
}

}  // extern "C"
