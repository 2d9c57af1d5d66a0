// hgdg: MI355X (gfx950) explicit DG operator behind a C ABI.
//
// This file holds what every element family shares: the context, state
// vectors and their storage precision, host <-> device transfers in hpGEM's
// layout, the RK update and copy kernels, hipGraph time stepping, and the C
// entry points of include/hgdg.h. The right-hand-side kernels live in
// tensor.hip (squares, cubes) and simplex.hip (tetrahedra).
#include <algorithm>
#include <fstream>
#include <sstream>
#include <cmath>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "context.hpp"
#include "vtk_io.hpp"
#include "mesh_io.hpp"

namespace hgdg {

thread_local std::string g_error;

void require(bool ok, int code, const std::string& msg) {
    if (!ok) throw Failure(code, msg);
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

// y[i] + sum_j c_j x_j[i], summed in term order; y and the first four terms
// are loaded before any is summed (their loads in flight together).
template <typename A>
__device__ __forceinline__ double axpy_value(const A& p, int64_t i) {
    auto term = [&](int j) {
        return ((p.x_f32 >> j) & 1u) ? double(static_cast<const float*>(p.x[j])[i])
                                     : static_cast<const double*>(p.x[j])[i];
    };
    double y = p.y[i];
    double x[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < p.n_x) x[j] = term(j);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < p.n_x) y += p.c[j] * x[j];
#pragma unroll
    for (int j = 4; j < kMaxIn; ++j) {
        if (j >= p.n_x) break;
        y += p.c[j] * term(j);
    }
    return y;
}

__global__ __launch_bounds__(kBlock) void multi_axpy(const AxpyParams p) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= p.n) return;
    p.y[i] = axpy_value(p, i);
}

// dst = src, 16 bytes per thread (plain streaming copy)
__global__ __launch_bounds__(kBlock) void copy_words(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n16) dst[i] = src[i];
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

void upload_vector(hgdg_ctx* c, Vec& v, const std::vector<double>& dev) {
    require(dev.size() == size_t(c->n_dof()), HGDG_ERR_ARG, "vector size mismatch");
    if (v.is_f32) {
        std::vector<float> f(dev.begin(), dev.end());
        HGDG_HIP(hipMemcpyAsync(v.f32->ptr, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HGDG_HIP(hipStreamSynchronize(c->stream));
    } else {
        HGDG_HIP(hipMemcpyAsync(v.f64->ptr, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice,
                                c->stream));
        HGDG_HIP(hipStreamSynchronize(c->stream));
    }
}

void download_vector(hgdg_ctx* c, Vec& v, std::vector<double>& dev) {
    const size_t n = size_t(c->n_dof());
    dev.assign(n, 0.0);
    if (v.is_f32) {
        std::vector<float> f(n);
        HGDG_HIP(hipMemcpyAsync(f.data(), v.f32->ptr, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HGDG_HIP(hipStreamSynchronize(c->stream));
        std::copy(f.begin(), f.end(), dev.begin());
    } else {
        HGDG_HIP(hipMemcpyAsync(dev.data(), v.f64->ptr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HGDG_HIP(hipStreamSynchronize(c->stream));
    }
}

void mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m) {
    require(m != nullptr, HGDG_ERR_ARG, "mesh descriptor is NULL");
    require(m->shape >= HGDG_SHAPE_LINE && m->shape <= HGDG_SHAPE_PYRAMID, HGDG_ERR_ARG, "unknown element shape");
    require(m->shape == SHAPE_SQUARE || m->shape == SHAPE_CUBE || m->shape == SHAPE_TETRAHEDRON,
            HGDG_ERR_UNSUPPORTED, "the device operators take squares, cubes and tetrahedra only");
    const int dim = shape_dim(m->shape);
    require(m->dim == dim, HGDG_ERR_ARG, "dim does not match the element shape");
    require(m->n_elements > 0 && m->n_elements < (int64_t(1) << 31), HGDG_ERR_ARG, "element count out of range");
    require(m->coords && m->node_ids, HGDG_ERR_ARG, "coords and node_ids are required");
    c->drop_graph();
    c->has_mesh = false;
    c->has_op = false;
    c->vecs.clear();
    c->shape = m->shape;
    c->dim = dim;
    c->n_elem = m->n_elements;
    c->tm = TensorMesh();
    c->sm = SimplexMesh();
    c->mat = MatrixOp();
    if (c->simplex()) simplex_mesh_upload(c, m);
    else tensor_mesh_upload(c, m);
    c->has_mesh = true;
}

void operator_set(hgdg_ctx* c, const hgdg_op_desc* op) {
    require(c->has_mesh && !c->matrix(), HGDG_ERR_STATE, "upload a mesh before setting the operator");
    require(op != nullptr, HGDG_ERR_ARG, "operator descriptor is NULL");
    require(op->kind == HGDG_OP_ADVECTION_UPWIND || op->kind == HGDG_OP_ACOUSTIC_CENTRAL, HGDG_ERR_ARG,
            "unknown operator kind");
    require(op->precision == HGDG_PREC_FP64 || op->precision == HGDG_PREC_MIXED_FP32, HGDG_ERR_ARG,
            "unknown precision");
    if (op->kind == HGDG_OP_ACOUSTIC_CENTRAL) {
        require(op->c_inv != nullptr, HGDG_ERR_ARG, "the acoustic operator needs c_inv per element");
        c->cinv.assign(op->c_inv, op->c_inv + c->n_elem);
        for (double v : c->cinv) require(v > 0 && std::isfinite(v), HGDG_ERR_ARG, "c_inv must be positive");
    } else {
        c->cinv.clear();
    }
    const int old_fields = c->n_fields;
    if (c->simplex()) simplex_operator_set(c, op);
    else tensor_operator_set(c, op);
    c->op = *op;
    c->op.c_inv = nullptr;
    c->has_op = true;
    c->drop_graph();
    if (c->simplex()) simplex_tune(c);
    else tensor_tune(c);
    if (old_fields != c->n_fields) c->vecs.clear();  // the field count changed the vector size
}

StageIO make_io(hgdg_ctx* c, const int* in_ids, const double* coeffs, int n_in, int out_id) {
    require(c->has_op, HGDG_ERR_STATE, "set an operator before computing time derivatives");
    require(n_in >= 1 && n_in <= kMaxIn, HGDG_ERR_ARG, "between 1 and 8 input vectors");
    StageIO io{};
    io.n_in = n_in;
    for (int i = 0; i < n_in; ++i) {
        require(in_ids[i] != out_id, HGDG_ERR_ARG, "the output vector must differ from the inputs");
        Vec& v = vec_at(c, in_ids[i]);
        io.in[i] = v.ptr();
        io.coef[i] = coeffs[i];
        if (v.is_f32) io.in_f32 |= 1u << i;
    }
    Vec& out = vec_at(c, out_id);
    io.out = out.ptr();
    io.out_f32 = out.is_f32 ? 1 : 0;
    return io;
}

void dispatch_rhs(hgdg_ctx* c, const StageIO& io) {
    if (c->matrix()) matrix_rhs(c, io, true);
    else if (c->simplex()) simplex_rhs(c, io);
    else tensor_rhs(c, io);
}

void launch_rhs(hgdg_ctx* c, const int* in_ids, const double* coeffs, int n_in, int out_id) {
    dispatch_rhs(c, make_io(c, in_ids, coeffs, n_in, out_id));
}

// Whether the operator's RHS kernel performs a fused RK update (StageIO::upd).
bool rhs_fuses_update(const hgdg_ctx* c) {
    if (c->matrix() || c->simplex()) return true;  // (tetrahedra run the acoustic operator only)
    return c->op.kind == HGDG_OP_ADVECTION_UPWIND || c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL;
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

// One step of an explicit tableau enqueued on the context's stream: each
// stage's input combination, RHS and inverse mass in one launch (zero tableau
// entries contribute exactly nothing and are skipped), then the update.
void enqueue_rk_step(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b,
                     double dt) {
    for (int i = 0; i < stages; ++i) {
        std::vector<int> ids{solution};
        std::vector<double> co{1.0};
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
}

// One step as enqueue_rk_step performs it, on the solution held at u_cur,
// with the update fused into the last stage's kernel and written to u_next
// (u_cur stays as it was: the last stage's neighbours still read it). The
// FP64 operations are those of enqueue_rk_step, so the result is bitwise the
// same.
void enqueue_rk_step_fused(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b,
                           double dt, double* u_cur, double* u_next) {
    for (int i = 0; i < stages; ++i) {
        std::vector<int> ids{solution};
        std::vector<double> co{1.0};
        for (int j = 0; j < i; ++j)
            if (a[i * stages + j] != 0.0) {
                ids.push_back(aux[j]);
                co.push_back(dt * a[i * stages + j]);
            }
        StageIO io = make_io(c, ids.data(), co.data(), int(ids.size()), aux[i]);
        io.in[0] = u_cur;
        if (i == stages - 1) {
            RkUpdate& q = io.upd;
            q.u = u_cur;
            q.y_new = u_next;
            q.self = -1;
            for (int j = 0; j < stages; ++j) {
                if (b[j] == 0.0) continue;
                Vec& x = vec_at(c, aux[j]);
                if (j == stages - 1) q.self = q.n_x;
                q.x[q.n_x] = x.ptr();
                q.c[q.n_x] = dt * b[j];
                if (x.is_f32) q.x_f32 |= 1u << q.n_x;
                ++q.n_x;
            }
        }
        dispatch_rhs(c, io);
    }
}

void check_rk_args(hgdg_ctx* c, int solution, const int* aux, int stages) {
    require(c->has_op, HGDG_ERR_STATE, "set an operator before time stepping");
    require(stages >= 1 && stages <= kMaxIn - 1, HGDG_ERR_ARG, "1..7 stages supported");
    Vec& u = vec_at(c, solution);
    require(!u.is_f32, HGDG_ERR_ARG, "the solution vector must be stored in FP64");
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    for (int i = 0; i < stages; ++i) {
        require(aux[i] != solution, HGDG_ERR_ARG, "stage vectors must differ from the solution");
        for (int j = 0; j < i; ++j) require(aux[i] != aux[j], HGDG_ERR_ARG, "stage vectors must be distinct");
        set_storage(c, vec_at(c, aux[i]), mixed);
    }
}

void rk_steps(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b, double dt,
              int n_steps) {
    require(n_steps >= 0, HGDG_ERR_ARG, "negative step count");
    check_rk_args(c, solution, aux, stages);
    if (tensor_rk_tiny(c, solution, aux, stages, a, b, dt, n_steps)) return;
    GraphKey key{solution, std::vector<int>(aux, aux + stages), std::vector<double>(a, a + stages * stages),
                 std::vector<double>(b, b + stages), dt};
    if (c->rk_fused != 0 && rhs_fuses_update(c) && n_steps >= 2) {
        // pairs of steps with the update fused into each last stage: the
        // solution goes to the context's spare vector and back
        double* u0 = vec_at(c, solution).f64->ptr;
        if (c->u_spare.n != size_t(c->n_dof())) c->u_spare.alloc(size_t(c->n_dof()));
        if (!(c->graph2_valid && c->graph2_key == key && c->graph2_u == u0)) {
            c->drop_graph2();
            HGDG_HIP(hipStreamSynchronize(c->stream));
            HGDG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_rk_step_fused(c, solution, aux, stages, a, b, dt, u0, c->u_spare.ptr);
                enqueue_rk_step_fused(c, solution, aux, stages, a, b, dt, c->u_spare.ptr, u0);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HGDG_HIP(hipStreamEndCapture(c->stream, &c->graph2));
            HGDG_HIP(hipGraphInstantiate(&c->graph2_exec, c->graph2, nullptr, nullptr, 0));
            c->graph2_key = key;
            c->graph2_u = u0;
            c->graph2_valid = true;
        }
        for (int s = 0; s < n_steps / 2; ++s) HGDG_HIP(hipGraphLaunch(c->graph2_exec, c->stream));
        c->rk_pair_launches += n_steps / 2;
        n_steps %= 2;
        if (n_steps == 0) return;
    }
    if (!(c->graph_valid && c->graph_key == key)) {
        c->drop_graph();
        HGDG_HIP(hipStreamSynchronize(c->stream));
        HGDG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_rk_step(c, solution, aux, stages, a, b, dt);
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

// Error estimate and solution size of the last step
// (HpgemAPISimplified::computeErrorAndNormOfUpdate, kernel/Base/
// HpgemAPISimplified_Impl.h:718-745): err2 = sum over all coefficients of
// (sum_j dt e_j k_j)^2, norm = sum over elements of the Euclidean norm of the
// element's solution coefficients. One wavefront per element.
struct StepNormParams {
    int64_t n_elem;
    const int64_t* off;  // [n_elem + 1] element DOF offsets, or NULL
    int64_t block;       // DOF per element when off is NULL
    const double* u;
    const void* k[kMaxIn];
    double c[kMaxIn];
    uint32_t k_f32;
    int32_t n_k;
    double* partial;     // [blocks][2]
};

__global__ __launch_bounds__(kBlock) void step_error_norm(const StepNormParams p) {
    constexpr int W = 64;  // wavefront size on CDNA
    __shared__ double s_part[2][kBlock / W];
    const int lane = threadIdx.x % W, wave = threadIdx.x / W;
    const int64_t e = int64_t(blockIdx.x) * (kBlock / W) + wave;
    double err2 = 0, u2 = 0;
    if (e < p.n_elem) {
        const int64_t lo = p.off ? p.off[e] : e * p.block;
        const int64_t hi = p.off ? p.off[e + 1] : lo + p.block;
        for (int64_t i = lo + lane; i < hi; i += W) {
            const double v = p.u[i];
            u2 += v * v;
            double d = 0;
            for (int j = 0; j < p.n_k; ++j)
                d += p.c[j] * (((p.k_f32 >> j) & 1u) ? double(static_cast<const float*>(p.k[j])[i])
                                                      : static_cast<const double*>(p.k[j])[i]);
            err2 += d * d;
        }
    }
    for (int o = W / 2; o > 0; o /= 2) {
        err2 += __shfl_xor(err2, o, W);
        u2 += __shfl_xor(u2, o, W);
    }
    if (lane == 0) {
        s_part[0][wave] = err2;
        s_part[1][wave] = sqrt(u2);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < kBlock / W; ++w) {
            a += s_part[0][w];
            b += s_part[1][w];
        }
        p.partial[2 * blockIdx.x] = a;
        p.partial[2 * blockIdx.x + 1] = b;
    }
}

void step_error_and_norm(hgdg_ctx* c, int solution, const int* aux, int stages, const double* err_coef, double dt,
                         double* error, double* norm) {
    StepNormParams p{};
    p.n_elem = c->n_elem;
    p.off = c->matrix() ? c->mat.d_elem_off.ptr : c->simplex() ? c->sm.d_elem_off.ptr : nullptr;
    p.block = c->dof / c->n_elem;
    p.u = vec_at(c, solution).f64->ptr;
    for (int j = 0; j < stages; ++j) {
        if (err_coef[j] == 0.0) continue;
        Vec& k = vec_at(c, aux[j]);
        p.k[p.n_k] = k.ptr();
        p.c[p.n_k] = dt * err_coef[j];
        if (k.is_f32) p.k_f32 |= 1u << p.n_k;
        ++p.n_k;
    }
    const int64_t per = kBlock / 64;
    const int64_t grid = (c->n_elem + per - 1) / per;
    if (c->d_partial.n < size_t(2 * grid)) c->d_partial.alloc(size_t(2 * grid));
    p.partial = c->d_partial.ptr;
    hipLaunchKernelGGL(step_error_norm, dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
    std::vector<double> h(size_t(2 * grid));
    HGDG_HIP(hipMemcpyAsync(h.data(), c->d_partial.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
    double e2 = 0, nrm = 0;
    for (int64_t i = 0; i < grid; ++i) {
        e2 += h[2 * i];
        nrm += h[2 * i + 1];
    }
    *error = std::sqrt(e2);
    *norm = nrm;
}

// ---- the step error without allocation -------------------------------------
// hgdg_step_error keeps its per-block partial sums in the context's buffer,
// which it allocates and grows on demand, and reduces them on the host in a
// std::vector. hgdg_step_error_noalloc runs the same per-element kernel with
// the partials in `scratch`, a file-scope device array of fixed size, then a
// one-block kernel reduces them into its last two entries and 16 bytes come
// back to the caller's stack. Each GPU holds its own copy of the array (at its
// own address, looked up once per device when a context is created on it);
// every call on one GPU shares it, so calls that overlap on one GPU race.
constexpr int64_t kScratchBlocks = 65536;  // partial pairs: meshes of at most 4 * 65536 elements
constexpr int64_t kScratchMaxElements = 4 * kScratchBlocks;
static __device__ __align__(256) double scratch[2 * kScratchBlocks + 2];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

// The current device's address of `scratch`, looked up once per device.
void setup_scratch(int device) {
    require(device >= 0 && device < kMaxDevices, HGDG_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device] != nullptr) return;
    void* ptr = nullptr;
    HGDG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, HGDG_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch[device] = static_cast<double*>(ptr);
}

// out[0] = sum of partial[2 i], out[1] = sum of partial[2 i + 1], i < n.
__global__ __launch_bounds__(kBlock) void reduce_pairs(const double* partial, int64_t n, double* out) {
    __shared__ double s[2][kBlock];
    const int t = threadIdx.x;
    double a = 0, b = 0;
    for (int64_t i = t; i < n; i += kBlock) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    s[0][t] = a;
    s[1][t] = b;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] += s[0][t + w];
            s[1][t] += s[1][t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = s[0][0];
        out[1] = s[1][0];
    }
}

// Failure of a call that must not allocate on its success path: the message
// is built only here.
int fail_noalloc(int code, const char* msg) {
    g_error = msg;
    return code;
}

int step_error_noalloc(hgdg_ctx* c, int solution, const int* aux, int stages, const double* err_coef, double dt,
                       double* error, double* norm) {
    if (c == nullptr || aux == nullptr || err_coef == nullptr || error == nullptr || norm == nullptr)
        return fail_noalloc(HGDG_ERR_ARG, "NULL argument");
    if (stages < 1 || stages > kMaxIn) return fail_noalloc(HGDG_ERR_ARG, "1..8 stage vectors");
    if (!c->has_op) return fail_noalloc(HGDG_ERR_STATE, "set an operator first");
    const int nv = int(c->vecs.size());
    if (solution < 0 || solution >= nv) return fail_noalloc(HGDG_ERR_ARG, "solution vector id out of range");
    for (int j = 0; j < stages; ++j)
        if (aux[j] < 0 || aux[j] >= nv) return fail_noalloc(HGDG_ERR_ARG, "stage vector id out of range");
    if (c->vecs[solution].is_f32) return fail_noalloc(HGDG_ERR_ARG, "the solution vector must be stored in FP64");
    if (c->n_elem > kScratchMaxElements)
        return fail_noalloc(HGDG_ERR_UNSUPPORTED,
                            "hgdg_step_error_noalloc: at most 262144 elements (its fixed scratch array holds "
                            "65536 partial pairs); use hgdg_step_error for larger meshes");
    if (c->device < 0 || c->device >= kMaxDevices || g_scratch[c->device] == nullptr)
        return fail_noalloc(HGDG_ERR_STATE, "no scratch address for this device (create the context with "
                                            "hgdg_ctx_create)");
    if (hipSetDevice(c->device) != hipSuccess) return fail_noalloc(HGDG_ERR_HIP, "hipSetDevice failed");
    double* ptr = g_scratch[c->device];
    StepNormParams p{};
    p.n_elem = c->n_elem;
    p.off = c->matrix() ? c->mat.d_elem_off.ptr : c->simplex() ? c->sm.d_elem_off.ptr : nullptr;
    p.block = c->n_elem > 0 ? c->dof / c->n_elem : 0;
    p.u = c->vecs[solution].f64->ptr;
    for (int j = 0; j < stages; ++j) {
        if (err_coef[j] == 0.0) continue;
        const Vec& k = c->vecs[aux[j]];
        p.k[p.n_k] = k.ptr();
        p.c[p.n_k] = dt * err_coef[j];
        if (k.is_f32) p.k_f32 |= 1u << p.n_k;
        ++p.n_k;
    }
    const int64_t per = kBlock / 64;
    const int64_t grid = (c->n_elem + per - 1) / per;
    p.partial = ptr;
    double h[2] = {0.0, 0.0};
    if (grid > 0) {
        hipLaunchKernelGGL(step_error_norm, dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
        hipLaunchKernelGGL(reduce_pairs, dim3(1), dim3(kBlock), 0, c->stream, ptr, grid, ptr + 2 * kScratchBlocks);
        if (hipGetLastError() != hipSuccess) return fail_noalloc(HGDG_ERR_HIP, "kernel launch failed");
        if (hipMemcpyAsync(h, ptr + 2 * kScratchBlocks, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail_noalloc(HGDG_ERR_HIP, "reading back the reduced sums failed");
    }
    *error = std::sqrt(h[0]);
    *norm = h[1];
    return HGDG_OK;
}

// HpgemAPISimplified::computeOneTimeStep(time, maximumRelativeError, dtMax)
// (kernel/Base/HpgemAPISimplified_Impl.h:822-861): steps with the current
// estimate (capped by dt_max), rewinds and shrinks dt while the embedded
// error exceeds max_rel_error times the solution size, then updates the
// estimate. The estimate is kept per context (the reference keeps one
// function-static value for the whole process).
void rk_adaptive_step(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b,
                      const double* err_coef, int order, double max_rel_error, double dt_max, double dt_initial,
                      double* time, double* dt_used, double* error_out, double* norm_out) {
    require(order >= 1, HGDG_ERR_ARG, "tableau order must be positive");
    require(max_rel_error > 0 && dt_max > 0 && dt_initial > 0, HGDG_ERR_ARG,
            "tolerance, dt_max and the initial dt must be positive");
    check_rk_args(c, solution, aux, stages);
    if (!(c->dt_estimate > 0)) c->dt_estimate = dt_initial;
    double dt = std::min(c->dt_estimate, dt_max);
    double err = std::numeric_limits<double>::infinity(), nrm = 0;
    for (int attempt = 0; max_rel_error * nrm < err || std::isinf(err); ++attempt) {
        require(attempt < 64, HGDG_ERR_ARG, "adaptive step: no acceptable step size after 64 attempts");
        enqueue_rk_step(c, solution, aux, stages, a, b, dt);
        step_error_and_norm(c, solution, aux, stages, err_coef, dt, &err, &nrm);
        if (max_rel_error * nrm < err || std::isinf(err)) {
            for (int j = 0; j < stages; ++j)
                if (b[j] != 0.0) launch_axpy(c, solution, {aux[j]}, {-dt * b[j]});
            dt = 0.8 * dt * std::pow(max_rel_error / err * nrm, 1.0 / order);
        }
    }
    c->dt_estimate = 0.9 * dt * std::pow(max_rel_error / err * nrm, 1.0 / order);
    *time += dt;
    if (dt_used) *dt_used = dt;
    if (error_out) *error_out = err;
    if (norm_out) *norm_out = nrm;
}

void state_diff(hgdg_ctx* c, int x, int y, double* l2, double* max_abs) {
    require(c->has_mesh, HGDG_ERR_STATE, "no mesh");
    Vec& vx = vec_at(c, x);
    Vec& vy = vec_at(c, y);
    if (c->matrix()) matrix_state_diff(c, vx, vy, l2, max_abs);
    else if (c->simplex()) simplex_state_diff(c, vx, vy, l2, max_abs);
    else tensor_state_diff(c, vx, vy, l2, max_abs);
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
        hgdg::setup_scratch(device);
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

int hgdg_set_advection_tiles(hgdg_ctx* ctx, int tiles) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(tiles == 0 || tiles == 1 || tiles == 2 || tiles == 4 || tiles == 8, HGDG_ERR_ARG,
                "tiles must be 0 (timed choice), 1, 2, 4 or 8");
        HGDG_HIP(hipSetDevice(ctx->device));
        const int before = ctx->adv_tiles_fixed;
        ctx->adv_tiles_fixed = tiles;
        if (ctx->has_op && !ctx->simplex()) {
            ctx->drop_graph();
            try {
                hgdg::tensor_tune(ctx);
            } catch (...) {
                ctx->adv_tiles_fixed = before;  // keep the previous choice in force
                throw;
            }
        }
    });
}

int hgdg_get_advection_tiles(hgdg_ctx* ctx, int* tiles, double* ms) {
    return run([&] {
        require(ctx != nullptr && tiles != nullptr, HGDG_ERR_ARG, "ctx and tiles must be non-NULL");
        *tiles = ctx->adv_tiles;
        if (ms)
            for (int i = 0; i < 4; ++i) ms[i] = ctx->adv_tile_ms[i];
    });
}

int hgdg_set_rk_fused(hgdg_ctx* ctx, int enable) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(enable == 0 || enable == 1, HGDG_ERR_ARG, "enable must be 0 or 1");
        ctx->rk_fused = enable;
    });
}

int hgdg_get_rk_fused(hgdg_ctx* ctx, int* enable, int64_t* launches) {
    return run([&] {
        require(ctx != nullptr && enable != nullptr, HGDG_ERR_ARG, "ctx and enable must be non-NULL");
        *enable = ctx->rk_fused;
        if (launches) *launches = ctx->rk_fused_launches;
    });
}

int hgdg_operator_set_matrices(hgdg_ctx* ctx, const hgdg_matrix_desc* desc, int precision) {
    return run([&] {
        require(ctx != nullptr && desc != nullptr, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::matrix_operator_set(ctx, desc, precision);
    });
}

int hgdg_set_tensor_general(hgdg_ctx* ctx, int force) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(force == 0 || force == 1, HGDG_ERR_ARG, "force must be 0 or 1");
        ctx->tensor_general_forced = force;
    });
}

int hgdg_get_tensor_general(hgdg_ctx* ctx, int* force, int* in_use) {
    return run([&] {
        require(ctx != nullptr && force != nullptr && in_use != nullptr, HGDG_ERR_ARG, "NULL argument");
        *force = ctx->tensor_general_forced;
        *in_use = ctx->has_mesh && !ctx->simplex() && ctx->tm.general ? 1 : 0;
    });
}

int hgdg_set_simplex_threads(hgdg_ctx* ctx, const int32_t* threads) {
    return run([&] {
        require(ctx != nullptr && threads != nullptr, HGDG_ERR_ARG, "ctx and threads must be non-NULL");
        for (int g = 0; g < 4; ++g)
            require(threads[g] == 0 || threads[g] == hgdg::simplex_threads(g + 1, 0) ||
                        threads[g] == hgdg::simplex_threads(g + 1, 1),
                    HGDG_ERR_ARG,
                    "threads per element of order " + std::to_string(g + 1) + " must be 0 (timed choice), " +
                        std::to_string(hgdg::simplex_threads(g + 1, 0)) + " or " +
                        std::to_string(hgdg::simplex_threads(g + 1, 1)));
        HGDG_HIP(hipSetDevice(ctx->device));
        int32_t before[4];
        std::copy(ctx->sm.tpe_fixed, ctx->sm.tpe_fixed + 4, before);
        std::copy(threads, threads + 4, ctx->sm.tpe_fixed);
        if (ctx->has_op && ctx->simplex()) {
            ctx->drop_graph();
            try {
                hgdg::simplex_tune(ctx);
            } catch (...) {
                std::copy(before, before + 4, ctx->sm.tpe_fixed);  // keep the previous choice in force
                throw;
            }
        }
    });
}

int hgdg_get_simplex_threads(hgdg_ctx* ctx, int32_t* threads, double* ms) {
    return run([&] {
        require(ctx != nullptr && threads != nullptr, HGDG_ERR_ARG, "ctx and threads must be non-NULL");
        for (int g = 0; g < 4; ++g) {
            threads[g] = hgdg::simplex_threads(g + 1, ctx->sm.tpe_variant[g]);
            if (ms) {
                ms[2 * g] = ctx->sm.tpe_ms[g][0];
                ms[2 * g + 1] = ctx->sm.tpe_ms[g][1];
            }
        }
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
        std::vector<double> dev;
        if (ctx->matrix()) dev.assign(host, host + ctx->n_dof());
        else if (ctx->simplex()) hgdg::simplex_to_device(ctx, host, dev);
        else hgdg::tensor_to_device(ctx, host, dev);
        hgdg::upload_vector(ctx, v, dev);
    });
}

int hgdg_state_get(hgdg_ctx* ctx, int vec, double* host) {
    return run([&] {
        require(ctx != nullptr && host != nullptr, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        std::vector<double> dev;
        hgdg::download_vector(ctx, v, dev);
        if (ctx->matrix()) std::copy(dev.begin(), dev.end(), host);
        else if (ctx->simplex()) hgdg::simplex_from_device(ctx, dev, host);
        else hgdg::tensor_from_device(ctx, dev, host);
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

int hgdg_right_hand_side(hgdg_ctx* ctx, const int* in_ids, const double* coeffs, int n_in, int out_id,
                         double time) {
    (void)time;  // the operators are autonomous
    return run([&] {
        require(ctx != nullptr && in_ids != nullptr && coeffs != nullptr, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::Vec& out = hgdg::vec_at(ctx, out_id);
        require(!out.is_f32, HGDG_ERR_UNSUPPORTED, "the right-hand side is written to a vector stored in FP64");
        if (ctx->matrix()) {  // the stiffness product alone, as multiplyStiffnessMatrices
            hgdg::matrix_rhs(ctx, hgdg::make_io(ctx, in_ids, coeffs, n_in, out_id), false);
            return;
        }
        // r = M k: the fused kernel's k, multiplied back by the block-diagonal mass
        hgdg::launch_rhs(ctx, in_ids, coeffs, n_in, out_id);
        if (ctx->simplex()) hgdg::simplex_mass_apply(ctx, out, false);
        else hgdg::tensor_mass_apply(ctx, out, false);
    });
}

int hgdg_solve_mass(hgdg_ctx* ctx, int vec) {
    return run([&] {
        require(ctx != nullptr, HGDG_ERR_ARG, "ctx is NULL");
        require(ctx->has_op, HGDG_ERR_STATE, "set an operator before solving with its mass matrix");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        if (ctx->matrix()) hgdg::matrix_mass_apply(ctx, v, true);
        else if (ctx->simplex()) hgdg::simplex_mass_apply(ctx, v, true);
        else hgdg::tensor_mass_apply(ctx, v, true);
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

int hgdg_rk_adaptive_step(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* a,
                          const double* b, const double* err_coef, int order, double max_rel_error, double dt_max,
                          double dt_initial, double* time, double* dt_used, double* error, double* norm) {
    return run([&] {
        require(ctx && aux_ids && a && b && err_coef && time, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::rk_adaptive_step(ctx, solution, aux_ids, stages, a, b, err_coef, order, max_rel_error, dt_max,
                               dt_initial, time, dt_used, error, norm);
    });
}

int hgdg_step_error(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* err_coef, double dt,
                    double* error, double* norm) {
    return run([&] {
        require(ctx && aux_ids && err_coef && error && norm, HGDG_ERR_ARG, "NULL argument");
        require(stages >= 1 && stages <= hgdg::kMaxIn, HGDG_ERR_ARG, "1..8 stage vectors");
        HGDG_HIP(hipSetDevice(ctx->device));
        require(ctx->has_op, HGDG_ERR_STATE, "set an operator first");
        require(!hgdg::vec_at(ctx, solution).is_f32, HGDG_ERR_ARG, "the solution vector must be stored in FP64");
        hgdg::step_error_and_norm(ctx, solution, aux_ids, stages, err_coef, dt, error, norm);
    });
}

int hgdg_step_error_noalloc(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* err_coef,
                            double dt, double* error, double* norm) {
    return hgdg::step_error_noalloc(ctx, solution, aux_ids, stages, err_coef, dt, error, norm);
}

int hgdg_project_modes(hgdg_ctx* ctx, int vec, int n_modes, const double* amp, const double* k) {
    return hgdg_project_field_modes(ctx, vec, -1, n_modes, amp, k);
}

int hgdg_project_field_modes(hgdg_ctx* ctx, int vec, int field, int n_modes, const double* amp, const double* k) {
    return run([&] {
        require(ctx != nullptr && (n_modes == 0 || (amp && k)), HGDG_ERR_ARG, "NULL argument");
        require(n_modes >= 0, HGDG_ERR_ARG, "negative mode count");
        HGDG_HIP(hipSetDevice(ctx->device));
        require(ctx->has_mesh, HGDG_ERR_STATE, "no mesh");
        require(!ctx->matrix(), HGDG_ERR_UNSUPPORTED,
                "a matrix operator has no geometry: project and measure errors with the caller's integrators");
        require(field >= -1 && field < ctx->n_fields, HGDG_ERR_ARG, "field out of range");
        hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        if (ctx->simplex()) hgdg::simplex_project(ctx, v, field, n_modes, amp, k);
        else hgdg::tensor_project_device(ctx, v, field, n_modes, amp, k);
    });
}

int hgdg_error_modes(hgdg_ctx* ctx, int vec, int field, int n_modes, const double* amp, const double* k,
                     const double* shift, double* l2, double* max_abs) {
    return run([&] {
        require(ctx != nullptr && l2 && max_abs && (n_modes == 0 || (amp && k)), HGDG_ERR_ARG, "NULL argument");
        require(n_modes >= 0, HGDG_ERR_ARG, "negative mode count");
        HGDG_HIP(hipSetDevice(ctx->device));
        require(ctx->has_mesh, HGDG_ERR_STATE, "no mesh");
        require(!ctx->matrix(), HGDG_ERR_UNSUPPORTED,
                "a matrix operator has no geometry: project and measure errors with the caller's integrators");
        require(field >= -1 && field < ctx->n_fields, HGDG_ERR_ARG, "field out of range");
        const double zero[3] = {0, 0, 0};
        const hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        if (ctx->simplex()) hgdg::simplex_error_modes(ctx, v, field, n_modes, amp, k, shift ? shift : zero, l2, max_abs);
        else hgdg::tensor_error_modes(ctx, v, field, n_modes, amp, k, shift ? shift : zero, l2, max_abs);
    });
}

int hgdg_project_gaussian(hgdg_ctx* ctx, int vec, int field, const double* center, double width, double amplitude) {
    return run([&] {
        require(ctx != nullptr && center != nullptr, HGDG_ERR_ARG, "NULL argument");
        require(width > 0 && std::isfinite(width) && std::isfinite(amplitude), HGDG_ERR_ARG,
                "the pulse needs a positive width");
        HGDG_HIP(hipSetDevice(ctx->device));
        require(ctx->has_mesh, HGDG_ERR_STATE, "no mesh");
        require(!ctx->matrix(), HGDG_ERR_UNSUPPORTED,
                "a matrix operator has no geometry: project and measure errors with the caller's integrators");
        require(field >= -1 && field < ctx->n_fields, HGDG_ERR_ARG, "field out of range");
        hgdg::Vec& v = hgdg::vec_at(ctx, vec);
        const double g[5] = {center[0], center[1], ctx->dim == 3 ? center[2] : 0.0, width, amplitude};
        if (ctx->simplex()) hgdg::simplex_project(ctx, v, field, 0, nullptr, nullptr, g);
        else hgdg::tensor_project_device(ctx, v, field, 0, nullptr, nullptr, g);
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
        require(shape == hgdg::SHAPE_SQUARE || shape == hgdg::SHAPE_CUBE || shape == hgdg::SHAPE_TETRAHEDRON,
                HGDG_ERR_UNSUPPORTED, "basis evaluation is available for squares, cubes and tetrahedra");
        if (shape == hgdg::SHAPE_TETRAHEDRON) {
            require(p >= 1 && p <= hgdg::kMaxSimplexOrder, HGDG_ERR_UNSUPPORTED, "order must be in 1..4");
            const int n = hgdg::tet_nb(p);
            std::vector<long double> v(n), g(3 * size_t(n));
            const long double x[3] = {point[0], point[1], point[2]};
            hgdg::tet_basis(p, x, v.data(), g.data());
            for (int i = 0; i < n; ++i) values[i] = double(v[i]);
            return;
        }
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

int hgdg_maxwell_assemble(hgdg_ctx* ctx, const hgdg_maxwell_desc* d, int64_t* n_blocks, int32_t* block_size,
                          int64_t* row_ptr, int64_t* col_idx, double* K, double* M, double* kernel_ms) {
    return run([&] {
        require(ctx && d && n_blocks && block_size, HGDG_ERR_ARG, "NULL argument");
        HGDG_HIP(hipSetDevice(ctx->device));
        hgdg::maxwell_assemble(ctx, d->order, d->epsilon, d->stab, n_blocks, block_size, row_ptr, col_idx, K, M);
        if (kernel_ms) *kernel_ms = row_ptr ? ctx->last_kernel_ms : 0.0;
    });
}

int hgdg_hcurl_basis_size(int p) { return hgdg::hcurl_tet_size(p); }

int hgdg_hcurl_basis_eval(int p, const double* point, double* values, double* curls) {
    return run([&] {
        require(point && values && curls, HGDG_ERR_ARG, "NULL argument");
        require(p >= 1 && p <= 2, HGDG_ERR_UNSUPPORTED, "H(curl) tetrahedral orders 1 and 2");
        const int n = hgdg::hcurl_tet_size(p);
        const long double x[3] = {point[0], point[1], point[2]};
        std::vector<long double> v(size_t(n) * 3), c(size_t(n) * 3);
        hgdg::hcurl_tet_basis(p, x, v.data(), c.data());
        for (int i = 0; i < 3 * n; ++i) {
            values[i] = double(v[i]);
            curls[i] = double(c[i]);
        }
    });
}

int hgdg_mesh_read(const char* path, hgdg_mesh_file_info* info, double* coords, int64_t* node_ids, int32_t* zones,
                   int64_t* face_elements, int32_t* face_local, char* zone_names) {
    return run([&] {
        require(path && info, HGDG_ERR_ARG, "NULL argument");
        const hgdg::MeshFile m = hgdg::read_mesh_file(path);
        std::string names;
        for (size_t z = 0; z < m.zone_names.size(); ++z) names += (z ? "\n" : "") + m.zone_names[z];
        *info = hgdg_mesh_file_info{};
        info->version = m.version;
        info->dim = m.dim;
        info->shape = m.shape;
        info->n_vertices = m.n_vertices;
        info->n_nodes = m.n_nodes;
        info->n_elements = m.n_elements;
        info->n_faces = m.n_faces;
        info->n_boundary_faces = m.n_boundary_faces;
        info->n_edges = m.n_edges;
        info->n_zones = int32_t(m.zone_names.size());
        info->zone_name_bytes = int32_t(names.size() + 1);
        if (coords) std::copy(m.coords.begin(), m.coords.end(), coords);
        if (node_ids) std::copy(m.node_ids.begin(), m.node_ids.end(), node_ids);
        if (zones) std::copy(m.zones.begin(), m.zones.end(), zones);
        if (face_elements) std::copy(m.face_elements.begin(), m.face_elements.end(), face_elements);
        if (face_local) std::copy(m.face_local.begin(), m.face_local.end(), face_local);
        if (zone_names) std::copy(names.c_str(), names.c_str() + names.size() + 1, zone_names);
    });
}

int hgdg_vtk_write(const char* base_name, int shape, int64_t n_elements, const double* coords, const int32_t* orders,
                   int uniform_order, int n_fields, const double* state, int64_t state_len, const char* field_names) {
    return run([&] {
        require(base_name && coords && state && n_elements >= 0 && n_fields >= 1, HGDG_ERR_ARG,
                "NULL argument or bad count");
        require(shape == hgdg::SHAPE_SQUARE || shape == hgdg::SHAPE_CUBE || shape == hgdg::SHAPE_TETRAHEDRON,
                HGDG_ERR_UNSUPPORTED, "VTK output is available for squares, cubes and tetrahedra");
        std::vector<int32_t> ord(orders ? orders : &uniform_order, orders ? orders + n_elements : &uniform_order + 1);
        if (!orders) ord.assign(size_t(n_elements), uniform_order);
        std::vector<std::string> names;
        if (field_names && *field_names) {
            std::istringstream in{std::string(field_names)};
            std::string name;
            while (std::getline(in, name)) names.push_back(name);
        } else {
            for (int f = 0; f < n_fields; ++f) names.push_back("field" + std::to_string(f));
        }
        require(int(names.size()) == n_fields, HGDG_ERR_ARG, "one field name per field");
        int64_t need = 0;
        for (int32_t p : ord) {
            const int q = std::max(1, std::min(int(p), 5));
            need += int64_t(n_fields) * (shape == hgdg::SHAPE_TETRAHEDRON ? hgdg::tet_nb(std::min(q, 4))
                                         : shape == hgdg::SHAPE_CUBE ? (q + 1) * (q + 1) * (q + 1) : (q + 1) * (q + 1));
        }
        require(state_len == need, HGDG_ERR_ARG,
                "state length " + std::to_string(state_len) + " does not match the fields and orders (" +
                    std::to_string(need) + ")");
        try {
            hgdg::write_vtk(base_name, shape, n_elements, coords, ord.data(), n_fields, state, names);
        } catch (const std::runtime_error& ex) {
            require(false, HGDG_ERR_ARG, ex.what());
        }
    });
}

int hgdg_mesh_write(const char* path, int shape, int64_t n_elements, const double* coords, const int64_t* node_ids,
                    const int32_t* zones, const char* zone_names) {
    return run([&] {
        require(path && coords && node_ids && n_elements >= 0, HGDG_ERR_ARG, "NULL argument or negative count");
        hgdg::MeshFile m;
        static const int dims[8] = {0, 1, 2, 2, 3, 3, 3, 3}, verts[8] = {0, 2, 4, 3, 8, 4, 6, 5};
        require(shape >= 1 && shape <= 7, HGDG_ERR_ARG, "unknown shape");
        m.shape = shape;
        m.dim = dims[shape];
        m.n_vertices = verts[shape];
        m.n_elements = n_elements;
        const size_t nv = size_t(n_elements) * m.n_vertices;
        m.coords.assign(coords, coords + nv * m.dim);
        m.node_ids.assign(node_ids, node_ids + nv);
        if (zones) m.zones.assign(zones, zones + n_elements);
        if (zone_names && *zone_names) {
            std::string all(zone_names), name;
            std::istringstream in(all);
            while (std::getline(in, name)) m.zone_names.push_back(name);
        }
        int32_t max_zone = -1;
        for (int32_t z : m.zones) {
            require(z >= 0, HGDG_ERR_ARG, "negative zone index");
            max_zone = std::max(max_zone, z);
        }
        require(max_zone < int32_t(std::max<size_t>(m.zone_names.size(), 1)), HGDG_ERR_ARG,
                "zone index without a zone name");
        const std::string text = hgdg::format_mesh_file(m);
        std::ofstream out(path, std::ios::binary);
        require(bool(out), HGDG_ERR_ARG, std::string("cannot open ") + path);
        out << text;
        out.close();
        require(bool(out), HGDG_ERR_ARG, std::string("cannot write ") + path);
    });
}

}  // extern "C"
