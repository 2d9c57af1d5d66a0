// Shared pieces of the hgdg library: error type, HIP checks, stage-input
// loads, compensated accumulation, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hgdg.h"

namespace hgdg {

extern thread_local std::string g_error;

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

// Two HIP events for timing, destroyed on every exit (also when a HIP call
// between creation and the end of the scope throws).
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    EventPair() {
        HGDG_HIP(hipEventCreate(&e[0]));
        const hipError_t err = hipEventCreate(&e[1]);
        if (err != hipSuccess) {
            (void)hipEventDestroy(e[0]);
            throw Failure(HGDG_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(err));
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
};

constexpr int kMaxIn = 8;
constexpr int kBlock = 256;

// MI355X hands consecutive workgroups of a launch to its 8 XCDs in turn, and
// each XCD has its own L2. The element kernels read the traces of
// neighbouring elements, which sit close by in element order; renumbering the
// blocks so that XCD x runs one contiguous range of logical blocks keeps
// those reads inside one L2. A bijection of [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t n) {
    const int64_t q = n / kXcds, r = n % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}


// Moves inside groups of four lanes through DPP quad permutations (a VALU
// operand modifier) instead of ds_bpermute (an LDS round trip). QP is the
// quad_perm control: bits [2l+1:2l] name the source lane of lane l; every lane
// of the quad must be active.
template <int QP>
__device__ __forceinline__ double quad_perm(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), QP, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), QP, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
constexpr int kQuadXor1 = 0xB1;  // lanes (1, 0, 3, 2)
constexpr int kQuadXor2 = 0x4E;  // lanes (2, 3, 0, 1)
__host__ __device__ constexpr int quad_bcast(int l) { return l * 0x55; }  // lanes (l, l, l, l)

// The RK update of a step, fused into its last stage's kernel:
// y_new = u + sum_j c_j x_j, the terms added one after another in FP64 (the
// operations of multi_axpy, so the result is bitwise the same), where term
// `self` is the stage derivative this kernel has just computed, taken as
// stored (rounded to the output's storage type), and the others are read from
// memory. y_new is another vector than u: the kernel's neighbours still read u.
struct RkUpdate {
    const double* u;  // NULL: no update
    double* y_new;
    const void* x[kMaxIn];
    double c[kMaxIn];
    uint32_t x_f32;  // bit j set: term j is stored in FP32
    int32_t n_x;
    int32_t self;    // the term that is this kernel's output, or -1
};

// Inputs and output of one stage evaluation: out = f(sum_i coef_i in_i).
struct StageIO {
    const void* in[kMaxIn];
    double coef[kMaxIn];
    uint32_t in_f32;  // bit i set: input i is stored in FP32
    int32_t n_in;
    void* out;
    int32_t out_f32;
    RkUpdate upd;     // fused RK update (upd.u == NULL: none)
};

// N consecutive values as doubles (N even: 16-byte vector loads, src being
// aligned to N values).
template <int N, typename T>
__device__ __forceinline__ void load_run(const T* src, double (&x)[N]) {
    if constexpr (N % 2 == 0 && sizeof(T) == 8) {
#pragma unroll
        for (int a = 0; a < N; a += 2) {
            const double2 t = *reinterpret_cast<const double2*>(src + a);
            x[a] = t.x;
            x[a + 1] = t.y;
        }
    } else if constexpr (N % 2 == 0) {
#pragma unroll
        for (int a = 0; a < N; a += 2) {
            const float2 t = *reinterpret_cast<const float2*>(src + a);
            x[a] = t.x;
            x[a + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int a = 0; a < N; ++a) x[a] = double(src[a]);
    }
}

// The fused update of N consecutive coefficients at o (a multiple of N), with
// `own` the values of this kernel's output there (as stored).
// The solution and the first kRkPre terms are loaded before any is summed
// (their loads in flight together), further terms one by one; the sum runs
// in term order as written.
constexpr int kRkPre = 4;

template <int N>
__device__ __forceinline__ void rk_term(const RkUpdate& q, int j, int64_t o, const double (&own)[N], double (&x)[N]) {
    if (j == q.self) {
#pragma unroll
        for (int a = 0; a < N; ++a) x[a] = own[a];
    } else if ((q.x_f32 >> j) & 1u) {
        load_run<N>(static_cast<const float*>(q.x[j]) + o, x);
    } else {
        load_run<N>(static_cast<const double*>(q.x[j]) + o, x);
    }
}

template <int N>
__device__ __forceinline__ void rk_update_run(const RkUpdate& q, int64_t o, const double (&own)[N]) {
    double y[N];
    load_run<N>(q.u + o, y);
    double xs[kRkPre][N];
#pragma unroll
    for (int j = 0; j < kRkPre; ++j)
        if (j < q.n_x) rk_term<N>(q, j, o, own, xs[j]);
#pragma unroll
    for (int j = 0; j < kRkPre; ++j)
        if (j < q.n_x) {
#pragma unroll
            for (int a = 0; a < N; ++a) y[a] += q.c[j] * xs[j][a];
        }
#pragma unroll
    for (int j = kRkPre; j < kMaxIn; ++j) {
        if (j >= q.n_x) break;
        double x[N];
        rk_term<N>(q, j, o, own, x);
#pragma unroll
        for (int a = 0; a < N; ++a) y[a] += q.c[j] * x[a];
    }
    if constexpr (N % 2 == 0) {
#pragma unroll
        for (int a = 0; a < N; a += 2) *reinterpret_cast<double2*>(q.y_new + o + a) = make_double2(y[a], y[a + 1]);
    } else {
#pragma unroll
        for (int a = 0; a < N; ++a) q.y_new[o + a] = y[a];
    }
}

__device__ __forceinline__ double load_in(const StageIO& p, int i, int64_t idx) {
    return ((p.in_f32 >> i) & 1u) ? double(static_cast<const float*>(p.in[i])[idx])
                                  : static_cast<const double*>(p.in[i])[idx];
}

// Stage input sum_i coef_i vec_i, first term scaled then accumulated in
// order (HpgemAPISimplified::getLinearCombinationOfVectors).
// The first two inputs (an RK stage's u and k) are loaded before either is
// used, so their loads are in flight together; more inputs follow one by one.
// The arithmetic and its order are those of the plain sum.
__device__ __forceinline__ double combine(const StageIO& p, int64_t idx) {
    const double x0 = load_in(p, 0, idx);
    double x1 = 0;
    if (p.n_in > 1) x1 = load_in(p, 1, idx);
    double s = p.coef[0] * x0;
    if (p.n_in > 1) s += p.coef[1] * x1;
#pragma unroll
    for (int i = 2; i < kMaxIn; ++i)
        if (i < p.n_in) s += p.coef[i] * load_in(p, i, idx);
    return s;
}

// combine() at K indices with the same rounding, input vector by input
// vector: the K loads of one input are independent and issue together, so a
// thread waits once per input instead of once per value.
template <int K>
__device__ __forceinline__ void load_n(const StageIO& p, int i, const int64_t (&idx)[K], double (&x)[K]) {
    if ((p.in_f32 >> i) & 1u) {
        const float* v = static_cast<const float*>(p.in[i]);
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = double(v[idx[k]]);
    } else {
        const double* v = static_cast<const double*>(p.in[i]);
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = v[idx[k]];
    }
}

template <int K>
__device__ __forceinline__ void combine_n(const StageIO& p, const int64_t (&idx)[K], double (&s)[K]) {
    double x0[K], x1[K];
    load_n(p, 0, idx, x0);
    if (p.n_in > 1) load_n(p, 1, idx, x1);
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = p.coef[0] * x0[k];
    if (p.n_in > 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] += p.coef[1] * x1[k];
    }
#pragma unroll
    for (int i = 2; i < kMaxIn; ++i) {
        if (i >= p.n_in) break;
        load_n(p, i, idx, x0);
#pragma unroll
        for (int k = 0; k < K; ++k) s[k] += p.coef[i] * x0[k];
    }
}

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

void require(bool ok, int code, const std::string& msg);

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n) {
        o.ptr = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            ptr = o.ptr;
            n = o.n;
            o.ptr = nullptr;
            o.n = 0;
        }
        return *this;
    }
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
