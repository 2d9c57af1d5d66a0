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
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>

#include "context.hpp"

namespace hgdg {

// Per-face upwind modes, 3 bits per tensor face f = 2d + s.
enum : uint32_t { MODE_NONE = 0, MODE_OWN = 1, MODE_NBR = 2, MODE_AVG = 3, MODE_HALF_OWN = 4 };

struct RhsParams {
    int64_t n_elem;
    StageIO io;
    const double* ahat;    // [dim][n_elem]
    const int32_t* nbr;    // [2 dim][n_elem]
    const uint32_t* mode;  // [n_elem]
    const double* Dr;      // [n1][n1] exact derivative matrix
    const double* Mi;      // [n1][n1] inverse 1D mass
};

template <int DIM, int N1>
struct Tensor {
    static constexpr int NB = N1 * N1 * (DIM == 3 ? N1 : 1);
    static constexpr int L = NB <= 4 ? 4 : NB <= 8 ? 8 : NB <= 16 ? 16 : NB <= 32 ? 32 : NB <= 64 ? 64 : NB <= 128 ? 128 : 256;
    static constexpr int EPB = kBlock / L;
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
__device__ __forceinline__ void load_pencil(const StageIO& p, int i, int64_t base, double (&x)[N1]) {
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
// The first two inputs are loaded before either is used (their loads are in
// flight together); the arithmetic and its order are those of the plain sum.
template <int N1>
__device__ __forceinline__ void combine_pencil(const StageIO& p, int64_t base, double (&w)[N1]) {
    double x[N1], y[N1];
    load_pencil<N1>(p, 0, base, x);
    if (p.n_in > 1) load_pencil<N1>(p, 1, base, y);
#pragma unroll
    for (int a = 0; a < N1; ++a) w[a] = p.coef[0] * x[a];
    if (p.n_in > 1) {
#pragma unroll
        for (int a = 0; a < N1; ++a) w[a] += p.coef[1] * y[a];
    }
#pragma unroll
    for (int i = 2; i < kMaxIn; ++i) {
        if (i >= p.n_in) break;
        load_pencil<N1>(p, i, base, x);
#pragma unroll
        for (int a = 0; a < N1; ++a) w[a] += p.coef[i] * x[a];
    }
}

// Stage input at the traces of all 2 DIM faces of a pencil: x faces (f < 2)
// one value at at[f], y/z faces the pencil [at[f], at[f] + N1). PAIR: the
// first two input vectors' loads all issue before any value is used (twice
// the registers); otherwise, and for further inputs, one input vector at a
// time. Same rounding as combine().
template <int DIM, int N1>
__device__ __forceinline__ void load_traces(const StageIO& p, int i, const int64_t (&at)[2 * DIM],
                                            double (&x)[2 * DIM][N1]) {
    if ((p.in_f32 >> i) & 1u) {
        const float* v = static_cast<const float*>(p.in[i]);
#pragma unroll
        for (int f = 0; f < 2; ++f) x[f][0] = v[at[f]];
#pragma unroll
        for (int f = 2; f < 2 * DIM; ++f) {
            const float* q = v + at[f];
            if constexpr (N1 % 4 == 0) {
#pragma unroll
                for (int a = 0; a < N1; a += 4) {
                    const float4 t = *reinterpret_cast<const float4*>(q + a);
                    x[f][a] = t.x; x[f][a + 1] = t.y; x[f][a + 2] = t.z; x[f][a + 3] = t.w;
                }
            } else if constexpr (N1 % 2 == 0) {
#pragma unroll
                for (int a = 0; a < N1; a += 2) {
                    const float2 t = *reinterpret_cast<const float2*>(q + a);
                    x[f][a] = t.x; x[f][a + 1] = t.y;
                }
            } else {
#pragma unroll
                for (int a = 0; a < N1; ++a) x[f][a] = q[a];
            }
        }
    } else {
        const double* v = static_cast<const double*>(p.in[i]);
#pragma unroll
        for (int f = 0; f < 2; ++f) x[f][0] = v[at[f]];
#pragma unroll
        for (int f = 2; f < 2 * DIM; ++f) {
            const double* q = v + at[f];
            if constexpr (N1 % 2 == 0) {
#pragma unroll
                for (int a = 0; a < N1; a += 2) {
                    const double2 t = *reinterpret_cast<const double2*>(q + a);
                    x[f][a] = t.x; x[f][a + 1] = t.y;
                }
            } else {
#pragma unroll
                for (int a = 0; a < N1; ++a) x[f][a] = q[a];
            }
        }
    }
}

template <int DIM, int N1, bool PAIR = true>
__device__ __forceinline__ void combine_traces(const StageIO& p, const int64_t (&at)[2 * DIM],
                                               double (&w)[2 * DIM][N1]) {
    double x[2 * DIM][N1];
    load_traces<DIM, N1>(p, 0, at, x);
    if constexpr (PAIR) {
        double y[2 * DIM][N1];
        if (p.n_in > 1) load_traces<DIM, N1>(p, 1, at, y);
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
            for (int a = 0; a < (f < 2 ? 1 : N1); ++a) w[f][a] = p.coef[0] * x[f][a];
        if (p.n_in > 1) {
#pragma unroll
            for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
                for (int a = 0; a < (f < 2 ? 1 : N1); ++a) w[f][a] += p.coef[1] * y[f][a];
        }
    } else {
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
            for (int a = 0; a < (f < 2 ? 1 : N1); ++a) w[f][a] = p.coef[0] * x[f][a];
        if (p.n_in > 1) {
            load_traces<DIM, N1>(p, 1, at, x);
#pragma unroll
            for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
                for (int a = 0; a < (f < 2 ? 1 : N1); ++a) w[f][a] += p.coef[1] * x[f][a];
        }
    }
#pragma unroll
    for (int i = 2; i < kMaxIn; ++i) {
        if (i >= p.n_in) break;
        load_traces<DIM, N1>(p, i, at, x);
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f)
#pragma unroll
            for (int a = 0; a < (f < 2 ? 1 : N1); ++a) w[f][a] += p.coef[i] * x[f][a];
    }
}

__device__ __forceinline__ double face_jump(uint32_t m, double own, double other) {
    if (m == MODE_NONE) return -own;
    if (m == MODE_HALF_OWN) return -0.5 * own;
    return (m == MODE_NBR) ? other - own : 0.5 * (other - own);
}

// Per-element data of the face terms: upwind modes, scaled velocity, and the
// neighbour of every face. It is loaded together with the stage input, so
// after the block's barrier only the neighbours' traces are still to come.
template <int DIM>
struct ElemMeta {
    uint32_t md;
    double ah[DIM];
    int32_t nb[2 * DIM];
};

template <int DIM, int N1>
__device__ __forceinline__ bool advection_load(const RhsParams& p, int64_t blk, double (&w)[N1], ElemMeta<DIM>& m) {
    using S = Pencil<DIM, N1>;
    const int le = threadIdx.x / S::NP, pc = threadIdx.x % S::NP;
    const int64_t e = blk * S::EPB + le;
    const bool active = (le < S::EPB) && (e < p.n_elem);
    if (active) {
        m.md = p.mode[e];
#pragma unroll
        for (int d = 0; d < DIM; ++d) m.ah[d] = p.ahat[d * p.n_elem + e];
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) m.nb[f] = p.nbr[f * p.n_elem + e];
        combine_pencil<N1>(p.io, e * S::NB + pc * N1, w);
    }
    return active;
}

// The arithmetic of one pencil once the block's stage input is in LDS (sw:
// this element's coefficients; w: this thread's pencil, also in registers):
// volume, upwind face jumps in FP64, lift, store. The neighbours' traces come
// from xo(s) (x face s, this pencil's value) and yz(d, s, f, nbr_pc, other)
// (y/z face: the neighbour's pencil nbr_pc), called only for faces whose mode
// needs the neighbour; every kernel variant shares this code, so they agree
// bit for bit.
template <int DIM, int N1, typename T, bool UPD, typename XO, typename YZ>
__device__ __forceinline__ void advection_compute(const RhsParams& p, int64_t e, int pc, const double (&w)[N1],
                                                  uint32_t md, const double* ahd, const double* sw, const T* s_D,
                                                  const T* s_Mi, XO xo, YZ yz) {
    using S = Pencil<DIM, N1>;
    const int ib = pc % N1;                    // y index of this pencil
    const int ic = DIM == 3 ? pc / N1 : 0;     // z index of this pencil
    T ah[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) ah[d] = T(ahd[d]);

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
#pragma unroll
    for (int s = 0; s < 2; ++s) {  // x faces: one trace value per pencil
        const uint32_t m = (md >> (3 * s)) & 7u;
        if (m == MODE_OWN) continue;
        const double other = (m == MODE_NBR || m == MODE_AVG) ? xo(s) : 0.0;
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
                yz(d, s, f, nbr_pc, other);
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
    double kv[N1];
    if (p.io.out_f32) {
        float* out = static_cast<float*>(p.io.out) + o;
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            const float v = float(acc[a].value());
            out[a] = v;
            kv[a] = double(v);
        }
    } else {
        double* out = static_cast<double*>(p.io.out) + o;
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            const double v = double(acc[a].value());
            out[a] = v;
            kv[a] = v;
        }
    }
    if constexpr (UPD) rk_update_run<N1>(p.io.upd, o, kv);
}

// General meshes: neighbour ids from the per-element face data; a neighbour
// inside the block is read from LDS, any other from HBM.
template <int DIM, int N1, typename T, bool UPD = false>
__device__ __forceinline__ void advection_finish(const RhsParams& p, int64_t blk, const double (&w)[N1],
                                                 const ElemMeta<DIM>& meta, const double* s_w, const T* s_D,
                                                 const T* s_Mi) {
    using S = Pencil<DIM, N1>;
    const int tid = threadIdx.x;
    const int le = tid / S::NP;
    const int pc = tid % S::NP;
    const int64_t e = blk * S::EPB + le;
    const bool active = (le < S::EPB) && (e < p.n_elem);
    if (!active) return;
    auto xo = [&](int s) -> double {
        const int64_t nb = meta.nb[s];
        const int64_t lnb = nb - blk * S::EPB;
        if (lnb >= 0 && lnb < S::EPB) return s_w[lnb * S::NB + pc * N1 + (1 - s)];  // neighbour in this block
        return combine(p.io, nb * S::NB + pc * N1 + (1 - s));
    };
    auto yz = [&](int, int, int f, int nbr_pc, double (&other)[N1]) {
        const int64_t nb = meta.nb[f];
        const int64_t lnb = nb - blk * S::EPB;
        if (lnb >= 0 && lnb < S::EPB) {
#pragma unroll
            for (int a = 0; a < N1; ++a) other[a] = s_w[lnb * S::NB + nbr_pc * N1 + a];
        } else {
            combine_pencil<N1>(p.io, nb * S::NB + nbr_pc * N1, other);
        }
    };
    advection_compute<DIM, N1, T, UPD>(p, e, pc, w, meta.md, meta.ah, s_w + le * S::NB, s_D, s_Mi, xo, yz);
}

template <int DIM, int N1, typename T>
__device__ __forceinline__ void advection_tables(const RhsParams& p, T* s_D, T* s_Mi) {
    const int tid = threadIdx.x;
    if (tid < N1 * N1) s_D[tid] = T(p.Dr[tid]);
    if (tid < 2 * N1) s_Mi[tid] = T(p.Mi[(tid % N1) * N1 + tid / N1]);  // s_Mi[s * N1 + a] = Mi(a, s)
}

template <int DIM, int N1, typename T, bool UPD = false>
__device__ __forceinline__ void advection_block(const RhsParams& p, int64_t blk) {
    using S = Pencil<DIM, N1>;
    __shared__ double s_w[S::EPB * S::NB];  // FP64 stage input of the block's elements
    __shared__ T s_D[N1 * N1];
    __shared__ T s_Mi[2 * N1];
    advection_tables<DIM, N1, T>(p, s_D, s_Mi);
    double w[N1];
    ElemMeta<DIM> meta;
    if (advection_load<DIM, N1>(p, blk, w, meta)) {
        const int le = threadIdx.x / S::NP, pc = threadIdx.x % S::NP;
#pragma unroll
        for (int a = 0; a < N1; ++a) s_w[le * S::NB + pc * N1 + a] = w[a];
    }
    __syncthreads();
    advection_finish<DIM, N1, T, UPD>(p, blk, w, meta, s_w, s_D, s_Mi);
}

// Whole RK steps of a small advection mesh in one workgroup: the solution,
// the stage input, every stage derivative and the per-element face data stay
// in LDS for all n_steps steps, with workgroup barriers between the phases,
// so one launch replaces 5 n_steps launches (meshes this small are bound by
// launch latency, not by HBM). The arithmetic is the graph path's, operation
// for operation: the stage input as combine() forms it, the RHS through
// advection_compute with the neighbour traces read from LDS, the update as
// multi_axpy forms it; results are bitwise the same (tested).
// The one workgroup: 16 wavefronts where the kernel fits in 128 registers,
// else 8 (256 registers).
template <int DIM, int N1>
constexpr int rk_tiny_block() {
    return (DIM == 2 ? N1 <= 4 : N1 <= 3) ? 1024 : 512;
}

struct RkTinyParams {
    RhsParams rhs;          // geometry, tables; io unused
    double* u;              // [n_dof] solution (read, then written after the last step)
    void* k[kMaxIn];        // stage vectors (written with the last step's derivatives)
    int32_t stages;
    int32_t n_steps;
    double ca[kMaxIn * kMaxIn];  // dt a_ij, row i
    double cb[kMaxIn];           // dt b_j
    uint32_t amask[kMaxIn];      // bit j of row i: a_ij != 0 (the graph path skips zero entries)
    uint32_t bmask;              // bit j: b_j != 0
};

template <int DIM, int N1, typename T>
__global__ __launch_bounds__((rk_tiny_block<DIM, N1>())) void tensor_advection_rk_tiny(const RkTinyParams p) {
    constexpr int kTinyBlock = rk_tiny_block<DIM, N1>();
    using S = Pencil<DIM, N1>;
    extern __shared__ __align__(16) double dyn[];
    __shared__ T s_D[N1 * N1];
    __shared__ T s_Mi[2 * N1];
    const int64_t E = p.rhs.n_elem, n = E * S::NB;
    const int st = p.stages;
    double* su = dyn;                                  // [n] solution
    double* sw = su + n;                               // [n] stage input
    uint32_t* smd = reinterpret_cast<uint32_t*>(sw + n);  // [E] upwind modes
    int32_t* snb = reinterpret_cast<int32_t*>(smd + E);   // [2 DIM][E] neighbours
    double* sah = reinterpret_cast<double*>(snb + 2 * DIM * E + ((E * (1 + 2 * DIM)) & 1));  // [DIM][E]
    T* sk = reinterpret_cast<T*>(sah + DIM * E);       // [stages][n] stage derivatives
    advection_tables<DIM, N1, T>(p.rhs, s_D, s_Mi);
    for (int64_t i = threadIdx.x; i < n; i += kTinyBlock) su[i] = p.u[i];
    for (int64_t e = threadIdx.x; e < E; e += kTinyBlock) {
        smd[e] = p.rhs.mode[e];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sah[d * E + e] = p.rhs.ahat[d * E + e];
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) snb[f * E + e] = p.rhs.nbr[f * E + e];
    }
    __syncthreads();
    RhsParams pp = p.rhs;
    pp.io.out_f32 = sizeof(T) == 4 ? 1 : 0;
    for (int step = 0; step < p.n_steps; ++step) {
        for (int i = 0; i < st; ++i) {
            // stage input: 1 * u, then + (dt a_ij) k_j in order of j
            for (int64_t q = threadIdx.x; q < n; q += kTinyBlock) {
                double v = 1.0 * su[q];
                for (int j = 0; j < i; ++j)
                    if ((p.amask[i] >> j) & 1u) v += p.ca[i * kMaxIn + j] * double(sk[int64_t(j) * n + q]);
                sw[q] = v;
            }
            __syncthreads();
            pp.io.out = sk + int64_t(i) * n;
            for (int64_t q = threadIdx.x; q < E * S::NP; q += kTinyBlock) {
                const int64_t e = q / S::NP;
                const int pc = int(q % S::NP);
                double w[N1];
#pragma unroll
                for (int a = 0; a < N1; ++a) w[a] = sw[e * S::NB + pc * N1 + a];
                double ah[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) ah[d] = sah[d * E + e];
                auto xo = [&](int s) -> double { return sw[int64_t(snb[s * E + e]) * S::NB + pc * N1 + (1 - s)]; };
                auto yz = [&](int, int, int f, int nbr_pc, double (&other)[N1]) {
                    const int64_t nb = snb[f * E + e];
#pragma unroll
                    for (int a = 0; a < N1; ++a) other[a] = sw[nb * S::NB + nbr_pc * N1 + a];
                };
                advection_compute<DIM, N1, T, false>(pp, e, pc, w, smd[e], ah, sw + e * S::NB, s_D, s_Mi, xo, yz);
            }
            __syncthreads();
        }
        // update: u + (dt b_j) k_j in order of j
        for (int64_t q = threadIdx.x; q < n; q += kTinyBlock) {
            double y = su[q];
            for (int j = 0; j < st; ++j)
                if ((p.bmask >> j) & 1u) y += p.cb[j] * double(sk[int64_t(j) * n + q]);
            su[q] = y;
        }
        __syncthreads();
    }
    for (int64_t q = threadIdx.x; q < n; q += kTinyBlock) {
        p.u[q] = su[q];
        for (int j = 0; j < st; ++j) static_cast<T*>(p.k[j])[q] = sk[int64_t(j) * n + q];
    }
}

// Minimum waves per SIMD the compiler is asked to keep for the one-tile
// kernel: 3D p=3 in FP64 needs a few registers beyond the 72 of seven waves
// otherwise, and fits them without spilling.
template <int DIM, int N1, typename T>
constexpr int adv_min_waves() {
    return (DIM == 3 && N1 == 4 && sizeof(T) == 8) ? 7 : 1;
}

// UPD: the RK update fused into this (last) stage, StageIO::upd.
template <int DIM, int N1, typename T, bool UPD = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(adv_min_waves<DIM, N1, T>())))
void tensor_advection_rhs(const RhsParams p) {
    advection_block<DIM, N1, T, UPD>(p, xcd_block(blockIdx.x, gridDim.x));
}

// The same work with each block taking `tiles` consecutive tiles of EPB
// elements: while a tile is computed, the next tile's stage input is already
// being loaded into registers, so the loads of one tile overlap the
// arithmetic and neighbour reads of the previous one, and a launch can be
// sized to fill the GPU in one round. Per element the arithmetic is exactly
// that of tensor_advection_rhs.
template <int DIM, int N1, typename T>
__global__ __launch_bounds__(kBlock) void tensor_advection_rhs_tiled(const RhsParams p, int tiles) {
    using S = Pencil<DIM, N1>;
    __shared__ double s_w[S::EPB * S::NB];
    __shared__ T s_D[N1 * N1];
    __shared__ T s_Mi[2 * N1];
    advection_tables<DIM, N1, T>(p, s_D, s_Mi);
    const int64_t n_tiles = (p.n_elem + S::EPB - 1) / S::EPB;
    const int64_t t0 = xcd_block(blockIdx.x, gridDim.x) * tiles;
    const int64_t t1 = t0 + tiles < n_tiles ? t0 + tiles : n_tiles;
    const int le = threadIdx.x / S::NP, pc = threadIdx.x % S::NP;
    double wn[N1];
    ElemMeta<DIM> mn;
    bool act_n = t0 < t1 && advection_load<DIM, N1>(p, t0, wn, mn);
    for (int64_t t = t0; t < t1; ++t) {
        double w[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) w[a] = wn[a];
        const ElemMeta<DIM> meta = mn;
        const bool act = act_n;
        __syncthreads();  // the previous tile's readers of s_w are done
        if (act)
#pragma unroll
            for (int a = 0; a < N1; ++a) s_w[le * S::NB + pc * N1 + a] = w[a];
        __syncthreads();
        act_n = t + 1 < t1 && advection_load<DIM, N1>(p, t + 1, wn, mn);
        advection_finish<DIM, N1, T>(p, t, w, meta, s_w, s_D, s_Mi);
    }
}

// Block b of a batched launch belongs to problem s with off[s] <= b < off[s+1].
__device__ __forceinline__ int find_problem(const int64_t* off, int n, int64_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Many independent problems of one kind in one launch: every block runs the
// exact code of the single-problem kernel on its own problem's data (UPD:
// with each problem's fused RK update, StageIO::upd).
template <int DIM, int N1, typename T, bool UPD = false>
__global__ __launch_bounds__(kBlock) void tensor_advection_rhs_batched(const RhsParams* ps, const int64_t* off,
                                                                       int n) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int s = find_problem(off, n, b);
    advection_block<DIM, N1, T, UPD>(ps[s], b - off[s]);
}

// Acoustic wave system on tensor elements (AcousticWaveLinear_Impl.h:102-283):
// fields v, s_1..s_dim stored field-major per element. With g_r = sum_d
// (J^{-1})_{r,d} s_d,
//   k_v   = sum_r (Dr along r) g_r + sum_faces Mi(:,s) (x) C_v
//   k_s_d = (1/c_inv) [ sum_r (J^{-1})_{r,d} (Dr along r) v + sum_faces Mi(:,s) (x) C_d ]
// where, on face (d_f, s) with w_k = (2s-1) (J^{-1})_{d_f,k} (= outward normal
// times face area over element volume), C_v = -1/2 w.(s_own - s_nbr) and
// C_d = -1/2 w_d (v_own - v_nbr) on internal faces, C_v = -w.s_own and C_d = 0
// on reflecting walls. Jumps are formed in FP64.
struct AcousticParams {
    int64_t n_elem;
    StageIO io;
    const double* jinv;   // [dim*dim][n_elem]: (J^{-1})_{r,d} at [(r dim + d) n_elem + e]
    const double* cinv;   // [n_elem]
    const int32_t* nbr;   // [2 dim][n_elem], -1 on walls
    const double* Dr;
    const double* Mi;
};

// Four lanes per pencil, one per field (v, s_1 .. s_3; in 2D the fourth lane
// idles), so a lane carries one field's N1 values. Both equations then
// run the same instructions:
//   volume: acc = sum_r scale_r (Dr along r) src_r, with src_r = g_r =
//     sum_d (J^{-1})_{r,d} s_d and scale_r = 1 on the v lane, src_r = v and
//     scale_r = (J^{-1})_{r,d} on the s_d lane;
//   faces: each lane forms its own field's jump; the v lane gets
//     sum_k w_k [s_k] by a reduction over the pencil's four lanes, the s
//     lanes get [v] from the v lane, both by lane shuffles within the quad.
constexpr int kQuad2DMinN1 = 3;  // 2D orders whose pencils take the four-lane layout: N1 >= this

template <int DIM, int N1>
struct AcousticShape {
    static constexpr int NF = DIM + 1;
    static constexpr int NP = Pencil<DIM, N1>::NP;
    static constexpr int NB = Pencil<DIM, N1>::NB;
    static constexpr bool QUAD = DIM == 3 || N1 >= kQuad2DMinN1;
    static constexpr int TPE = QUAD ? 4 * NP : NP;  // threads per element
    static constexpr int EPB = kBlock / TPE;
};

template <int DIM, int N1, typename T, bool UPD>
__device__ __forceinline__ void acoustic_block_quad(const AcousticParams& p, int64_t blk) {
    using S = AcousticShape<DIM, N1>;
    constexpr int NF = S::NF;
    constexpr int NB = S::NB;
    __shared__ double s_w[S::EPB * NF * NB];      // FP64 stage input
    __shared__ T s_g[S::EPB * (DIM + 1) * NB];    // v, then g_r, in T
    __shared__ double s_J[S::EPB * DIM * DIM];    // J^{-1}
    __shared__ T s_D[N1 * N1];
    __shared__ T s_Mi[2 * N1];
    const int tid = threadIdx.x;
    if (tid < N1 * N1) s_D[tid] = T(p.Dr[tid]);
    if (tid < 2 * N1) s_Mi[tid] = T(p.Mi[(tid % N1) * N1 + tid / N1]);
    const int le = tid / S::TPE;
    const int pc = (tid % S::TPE) / 4;
    const int fld = tid % 4;
    const int64_t e = blk * S::EPB + le;
    const bool active = (le < S::EPB) && (e < p.n_elem);
    const bool lane = active && fld < NF;
    const int ib = pc % N1;
    const int ic = DIM == 3 ? pc / N1 : 0;
    double* sw = s_w + le * NF * NB;
    T* sg = s_g + le * (DIM + 1) * NB;
    double* sJ = s_J + le * DIM * DIM;
    const int64_t base = e * NF * NB;
    const int d = fld - 1;  // s lanes: the component

    // ---- stage input, v and g_r into LDS
    double w[N1];
#pragma unroll
    for (int a = 0; a < N1; ++a) w[a] = 0.0;
    if (lane) {
        combine_pencil<N1>(p.io, base + fld * NB + pc * N1, w);
#pragma unroll
        for (int a = 0; a < N1; ++a) sw[fld * NB + pc * N1 + a] = w[a];
        if (pc == 0 && fld == 0)
#pragma unroll
            for (int k = 0; k < DIM * DIM; ++k) sJ[k] = p.jinv[k * p.n_elem + e];
    }
    {
        // lane 1 + r of the quad forms g_r at the pencil's nodes, lane 0 writes v
        double Jr[DIM];
        const int r = fld - 1;
#pragma unroll
        for (int k = 0; k < DIM; ++k) Jr[k] = (active && r >= 0 && r < DIM) ? p.jinv[(r * DIM + k) * p.n_elem + e] : 0.0;
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            double g = 0;
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                const double wk = k == 0 ? quad_perm<quad_bcast(1)>(w[a])
                                : k == 1 ? quad_perm<quad_bcast(2)>(w[a]) : quad_perm<quad_bcast(3)>(w[a]);
                g += Jr[k] * wk;
            }
            if (lane) sg[(r + 1) * NB + pc * N1 + a] = fld == 0 ? T(w[a]) : T(g);
        }
    }
    const double cinv = active ? p.cinv[e] : 1.0;
    int32_t nbi[2 * DIM];  // the face neighbours, loaded with the stage input
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) nbi[f] = active ? p.nbr[f * p.n_elem + e] : -1;
    __syncthreads();

    // ---- volume
    Acc<T> acc[N1];
    if (lane) {
#pragma unroll
        for (int a = 0; a < N1; ++a) acc[a] = Acc<T>(T(0));
        T vol[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) vol[a] = T(0);
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            const T* src = fld == 0 ? sg + (1 + r) * NB : sg;
            const T scale = fld == 0 ? T(1) : T(sJ[r * DIM + d]);
            T t[N1];
            if (r == 0) {  // along this pencil
#pragma unroll
                for (int a = 0; a < N1; ++a) {
                    T sv = T(0);
#pragma unroll
                    for (int j = 0; j < N1; ++j) sv += s_D[a * N1 + j] * src[pc * N1 + j];
                    t[a] = sv;
                }
            } else {  // across the element's other pencils
                const int my = r == 1 ? ib : ic;
#pragma unroll
                for (int a = 0; a < N1; ++a) t[a] = T(0);
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const int opc = r == 1 ? j + N1 * ic : ib + N1 * j;
                    const T dr = s_D[my * N1 + j];
#pragma unroll
                    for (int a = 0; a < N1; ++a) t[a] += dr * src[opc * N1 + a];
                }
            }
#pragma unroll
            for (int a = 0; a < N1; ++a) vol[a] += scale * t[a];
        }
#pragma unroll
        for (int a = 0; a < N1; ++a) acc[a] = Acc<T>(vol[a]);
    }

    // ---- faces: x faces one value per pencil, y/z faces a whole pencil.
    // The neighbours' traces of all faces first, every global load issued
    // before any is used (a neighbour in the block is read from LDS; a face
    // that loads nothing points at the thread's own pencil). At p = 3
    // (N1 = 4) the registers that holds cost more than the shorter wait
    // gains (measured in both precisions), so there each face loads its own.
    constexpr bool kHoist = N1 != 4;
    double oth[2 * DIM][N1];
    if constexpr (kHoist) {
        int64_t at[2 * DIM];
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) {
            const int df = f / 2, s = f % 2;
            const int nbr_pc = df == 0 ? pc : df == 1 ? (1 - s) + N1 * ic : ib + N1 * (1 - s);
            const int noff = df == 0 ? 1 - s : 0;
            const int64_t lnb = int64_t(nbi[f]) - blk * S::EPB;
            const bool glob = lane && nbi[f] >= 0 && !(lnb >= 0 && lnb < S::EPB);
            at[f] = glob ? int64_t(nbi[f]) * NF * NB + fld * NB + nbr_pc * N1 + noff : base + fld * NB + pc * N1;
        }
        // (at order 4 in FP32 holding two inputs' traces costs an occupancy step)
        constexpr bool kPair = !(N1 >= 5 && std::is_same<T, float>::value);
        if (lane) combine_traces<DIM, N1, kPair>(p.io, at, oth);  // (an inactive thread's own pencil may lie past the end)
#pragma unroll
        for (int f = 0; f < 2 * DIM; ++f) {
            const int df = f / 2, s = f % 2;
            const int nbr_pc = df == 0 ? pc : df == 1 ? (1 - s) + N1 * ic : ib + N1 * (1 - s);
            const int noff = df == 0 ? 1 - s : 0;
            const int64_t lnb = int64_t(nbi[f]) - blk * S::EPB;
            if (lane && nbi[f] >= 0 && lnb >= 0 && lnb < S::EPB) {
#pragma unroll
                for (int a = 0; a < (df == 0 ? 1 : N1); ++a)
                    oth[f][a] = s_w[(lnb * NF + fld) * NB + nbr_pc * N1 + noff + a];
            }
        }
    }
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
        const int df = f / 2, s = f % 2;
        const bool inner = nbi[f] >= 0;
        const int own_pc = df == 0 ? pc : df == 1 ? s + N1 * ic : ib + N1 * s;
        const int na = df == 0 ? 1 : N1;  // trace values per pencil
        const int off = df == 0 ? s : 0;  // own node of an x face
        double jmp[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) jmp[a] = 0.0;
        if constexpr (!kHoist) {
            if (lane && inner) {
                const int nbr_pc = df == 0 ? pc : df == 1 ? (1 - s) + N1 * ic : ib + N1 * (1 - s);
                const int noff = df == 0 ? 1 - s : 0;
                const int64_t lnb = int64_t(nbi[f]) - blk * S::EPB;
                if (lnb >= 0 && lnb < S::EPB) {
#pragma unroll
                    for (int a = 0; a < N1; ++a)
                        if (a < na) oth[f][a] = s_w[(lnb * NF + fld) * NB + nbr_pc * N1 + noff + a];
                } else if (df == 0) {
                    oth[f][0] = combine(p.io, int64_t(nbi[f]) * NF * NB + fld * NB + nbr_pc * N1 + noff);
                } else {
                    combine_pencil<N1>(p.io, int64_t(nbi[f]) * NF * NB + fld * NB + nbr_pc * N1, oth[f]);
                }
            }
        }
        if (lane) {
#pragma unroll
            for (int a = 0; a < N1; ++a)
                if (a < na) jmp[a] = sw[fld * NB + own_pc * N1 + off + a] - (inner ? oth[f][a] : 0.0);
        }
        const double wd = (2 * s - 1) * (lane && fld > 0 ? sJ[df * DIM + d] : 0.0);  // w_d on the s_d lane
        const int my = df == 1 ? ib : ic;
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            if (a >= na) break;
            // v lane: sum_k w_k [s_k] over the quad; s lanes: [v] from the v lane
            double x = wd * jmp[a];
            x += quad_perm<kQuadXor1>(x);
            x += quad_perm<kQuadXor2>(x);
            const double jv = quad_perm<quad_bcast(0)>(jmp[a]);
            double c;
            if (fld == 0) c = inner ? -0.5 * x : -x;
            else c = inner ? wd * (-0.5 * jv) : 0.0;
            if (lane) {
                if (df == 0) {
#pragma unroll
                    for (int b = 0; b < N1; ++b) acc[b].add(s_Mi[s * N1 + b] * T(c));
                } else {
                    acc[a].add(s_Mi[s * N1 + my] * T(c));
                }
            }
        }
    }
    if (!lane) return;
    const T ci = T(1.0 / cinv);
    const int64_t o = base + fld * NB + pc * N1;
    double kv[N1];  // the values as stored
    if (p.io.out_f32) {
        float* out = static_cast<float*>(p.io.out);
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            const float v = float(fld == 0 ? acc[a].value() : ci * acc[a].value());
            out[o + a] = v;
            kv[a] = double(v);
        }
    } else {
        double* out = static_cast<double*>(p.io.out);
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            const double v = double(fld == 0 ? acc[a].value() : ci * acc[a].value());
            out[o + a] = v;
            kv[a] = v;
        }
    }
    if constexpr (UPD) rk_update_run<N1>(p.io.upd, o, kv);
}

// In 2D at order 1 (three fields, little work per pencil) one thread per
// pencil does all fields; from order 2 on the four-lane layout above is
// faster (the one-thread form needs up to 256 registers there: one wave per
// SIMD at orders 4 and 5).
template <int DIM, int N1, typename T, bool UPD>
__device__ __forceinline__ void acoustic_block_pencil(const AcousticParams& p, int64_t blk) {
    using S = Pencil<DIM, N1>;
    static_assert(AcousticShape<DIM, N1>::EPB == S::EPB, "launch geometry");
    constexpr int NF = DIM + 1;
    constexpr int NB = S::NB;
    __shared__ double s_w[S::EPB * NF * NB];
    __shared__ T s_D[N1 * N1];
    __shared__ T s_Mi[2 * N1];
    const int tid = threadIdx.x;
    if (tid < N1 * N1) s_D[tid] = T(p.Dr[tid]);
    if (tid < 2 * N1) s_Mi[tid] = T(p.Mi[(tid % N1) * N1 + tid / N1]);
    const int le = tid / S::NP;
    const int pc = tid % S::NP;
    const int64_t e = blk * S::EPB + le;
    const bool active = (le < S::EPB) && (e < p.n_elem);
    const int ib = pc % N1;
    const int ic = DIM == 3 ? pc / N1 : 0;
    double* sw = s_w + le * NF * NB;
    const int64_t base = e * NF * NB;
    double w[NF][N1];
    if (active) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            combine_pencil<N1>(p.io, base + f * NB + pc * N1, w[f]);
#pragma unroll
            for (int a = 0; a < N1; ++a) sw[f * NB + pc * N1 + a] = w[f][a];
        }
    }
    __syncthreads();
    if (!active) return;

    double Ji[DIM][DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int d = 0; d < DIM; ++d) Ji[r][d] = p.jinv[(r * DIM + d) * p.n_elem + e];
    const double cinv = p.cinv[e];

    // g_r of a pencil from its s values
    auto g_of = [&](const double* sv_base, int r, int a) {
        double g = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) g += Ji[r][d] * sv_base[(1 + d) * NB + a];
        return g;
    };
    T kv[N1], t[DIM][N1];
#pragma unroll
    for (int a = 0; a < N1; ++a) {
        kv[a] = T(0);
#pragma unroll
        for (int r = 0; r < DIM; ++r) t[r][a] = T(0);
    }
    // x direction (own pencil, registers)
#pragma unroll
    for (int a = 0; a < N1; ++a) {
        T sg = T(0), sv = T(0);
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            double g0 = 0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) g0 += Ji[0][d] * w[1 + d][j];
            sg += s_D[a * N1 + j] * T(g0);
            sv += s_D[a * N1 + j] * T(w[0][j]);
        }
        kv[a] += sg;
        t[0][a] += sv;
    }
    // y (and z) directions from the element's other pencils
#pragma unroll
    for (int r = 1; r < DIM; ++r) {
        const int my = r == 1 ? ib : ic;
#pragma unroll
        for (int j = 0; j < N1; ++j) {
            const int opc = r == 1 ? j + N1 * ic : ib + N1 * j;
            const double* ob = sw + opc * N1;
            const T dr = s_D[my * N1 + j];
#pragma unroll
            for (int a = 0; a < N1; ++a) {
                kv[a] += dr * T(g_of(ob, r, a));
                t[r][a] += dr * T(ob[a]);
            }
        }
    }
    Acc<T> av[N1], as[DIM][N1];
#pragma unroll
    for (int a = 0; a < N1; ++a) {
        av[a] = Acc<T>(kv[a]);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            T v = T(0);
#pragma unroll
            for (int r = 0; r < DIM; ++r) v += T(Ji[r][d]) * t[r][a];
            as[d][a] = Acc<T>(v);
        }
    }

    // faces
#pragma unroll
    for (int df = 0; df < DIM; ++df) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int f = 2 * df + s;
            const int64_t nb = p.nbr[f * p.n_elem + e];
            double wk[DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) wk[k] = (2 * s - 1) * Ji[df][k];
            if (df == 0) {
                // one trace value per field for this pencil
                double own[NF], oth[NF];
#pragma unroll
                for (int q = 0; q < NF; ++q) own[q] = w[q][s];
                if (nb >= 0) {
                    const int64_t lnb = nb - blk * S::EPB;
                    if (lnb >= 0 && lnb < S::EPB) {  // neighbour in this block
#pragma unroll
                        for (int q = 0; q < NF; ++q) oth[q] = s_w[(lnb * NF + q) * NB + pc * N1 + (1 - s)];
                    } else {
#pragma unroll
                        for (int q = 0; q < NF; ++q) oth[q] = combine(p.io, nb * NF * NB + q * NB + pc * N1 + (1 - s));
                    }
                }
                double cv = 0, cd = 0;
                if (nb >= 0) {
#pragma unroll
                    for (int k = 0; k < DIM; ++k) cv += wk[k] * (own[1 + k] - oth[1 + k]);
                    cv *= -0.5;
                    cd = -0.5 * (own[0] - oth[0]);
                } else {
#pragma unroll
                    for (int k = 0; k < DIM; ++k) cv -= wk[k] * own[1 + k];
                }
#pragma unroll
                for (int a = 0; a < N1; ++a) {
                    av[a].add(s_Mi[s * N1 + a] * T(cv));
                    if (nb >= 0)
#pragma unroll
                        for (int d = 0; d < DIM; ++d) as[d][a].add(s_Mi[s * N1 + a] * T(wk[d] * cd));
                }
            } else {
                const int own_pc = df == 1 ? s + N1 * ic : ib + N1 * s;
                const int nbr_pc = df == 1 ? (1 - s) + N1 * ic : ib + N1 * (1 - s);
                const int my = df == 1 ? ib : ic;
                double oth[NF][N1];
                if (nb >= 0) {
                    const int64_t lnb = nb - blk * S::EPB;
                    if (lnb >= 0 && lnb < S::EPB) {
#pragma unroll
                        for (int q = 0; q < NF; ++q)
#pragma unroll
                            for (int a = 0; a < N1; ++a) oth[q][a] = s_w[(lnb * NF + q) * NB + nbr_pc * N1 + a];
                    } else {
#pragma unroll
                        for (int q = 0; q < NF; ++q)
                            combine_pencil<N1>(p.io, nb * NF * NB + q * NB + nbr_pc * N1, oth[q]);
                    }
                }
                const T m = s_Mi[s * N1 + my];
#pragma unroll
                for (int a = 0; a < N1; ++a) {
                    double cv = 0, cd = 0;
                    if (nb >= 0) {
#pragma unroll
                        for (int k = 0; k < DIM; ++k) cv += wk[k] * (sw[(1 + k) * NB + own_pc * N1 + a] - oth[1 + k][a]);
                        cv *= -0.5;
                        cd = -0.5 * (sw[own_pc * N1 + a] - oth[0][a]);
                    } else {
#pragma unroll
                        for (int k = 0; k < DIM; ++k) cv -= wk[k] * sw[(1 + k) * NB + own_pc * N1 + a];
                    }
                    av[a].add(m * T(cv));
                    if (nb >= 0)
#pragma unroll
                        for (int d = 0; d < DIM; ++d) as[d][a].add(m * T(wk[d] * cd));
                }
            }
        }
    }
    const int64_t o = base + pc * N1;
    const T ci = T(1.0 / cinv);
    double ks_out[NF][N1];  // the values as stored
    if (p.io.out_f32) {
        float* out = static_cast<float*>(p.io.out);
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            const float v = float(av[a].value());
            out[o + a] = v;
            ks_out[0][a] = double(v);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const float vs = float(ci * as[d][a].value());
                out[o + (1 + d) * NB + a] = vs;
                ks_out[1 + d][a] = double(vs);
            }
        }
    } else {
        double* out = static_cast<double*>(p.io.out);
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            const double v = double(av[a].value());
            out[o + a] = v;
            ks_out[0][a] = v;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const double vs = double(ci * as[d][a].value());
                out[o + (1 + d) * NB + a] = vs;
                ks_out[1 + d][a] = vs;
            }
        }
    }
    if constexpr (UPD) {
#pragma unroll
        for (int f = 0; f < NF; ++f) rk_update_run<N1>(p.io.upd, o + f * NB, ks_out[f]);
    }
}

// UPD: the RK update fused into this (last) stage, StageIO::upd.
template <int DIM, int N1, typename T, bool UPD = false>
__device__ __forceinline__ void acoustic_block(const AcousticParams& p, int64_t blk) {
    if constexpr (AcousticShape<DIM, N1>::QUAD) acoustic_block_quad<DIM, N1, T, UPD>(p, blk);
    else acoustic_block_pencil<DIM, N1, T, UPD>(p, blk);
}

template <int DIM, int N1, typename T, bool UPD = false>
__global__ __launch_bounds__(kBlock) void tensor_acoustic_rhs(const AcousticParams p) {
    acoustic_block<DIM, N1, T, UPD>(p, xcd_block(blockIdx.x, gridDim.x));
}

template <int DIM, int N1, typename T, bool UPD = false>
__global__ __launch_bounds__(kBlock) void tensor_acoustic_rhs_batched(const AcousticParams* ps, const int64_t* off,
                                                                      int n) {
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);
    const int s = find_problem(off, n, b);
    acoustic_block<DIM, N1, T, UPD>(ps[s], b - off[s]);
}

// L2 projection of f(x) = sum_m amp_m sin(2 pi k_m . x) on affine tensor
// elements (HpgemAPISimplified::setInitialSolution with the element rule of
// order 2p+1): the coefficients are (Mi B^T W) applied along each direction
// to f at the n1^dim Gauss points, P(a, q) = sum_b Mi(a, b) psi_b(x_q) w_q.
// One thread per x-line of quadrature points; the y/z passes go through LDS.
struct ProjParams {
    int64_t n_elem;
    int32_t nf, field;          // fields per element; target field (-1: all)
    void* out;
    int32_t out_f32;
    int32_t n_modes;
    const double* modes;        // [n_modes][4]: amp, k_x, k_y, k_z
    int32_t gaussian;           // 1: f = amplitude exp(-|x - centre|^2 / width^2) instead of modes
    double g[5];                // centre x, y, z, width, amplitude
    const double* X0;           // [dim][n_elem] (generic meshes)
    const double* E;            // [dim*dim][n_elem] (generic meshes)
    int32_t structured;
    int32_t cells[3];
    const double* P;            // [n1][n1]
    const double* xq;           // [n1]
};

template <int DIM, int N1>
__global__ __launch_bounds__(kBlock) void tensor_project(const ProjParams p) {
    using S = Pencil<DIM, N1>;
    __shared__ double s_a[S::EPB * S::NB];
    __shared__ double s_b[S::EPB * S::NB];
    const int tid = threadIdx.x;
    const int le = tid / S::NP;
    const int pc = tid % S::NP;
    const int64_t e = int64_t(blockIdx.x) * S::EPB + le;
    const bool active = (le < S::EPB) && (e < p.n_elem);
    const int i1 = pc % N1;
    const int i2 = DIM == 3 ? pc / N1 : 0;
    double X0[3] = {0, 0, 0}, E[3][3] = {{0}};
    if (active) {
        if (p.structured) {
            int64_t idx[3] = {e % p.cells[0], (e / p.cells[0]) % p.cells[1],
                              DIM == 3 ? e / (int64_t(p.cells[0]) * p.cells[1]) : 0};
            for (int d = 0; d < DIM; ++d) {
                X0[d] = 0.0 + double(idx[d]) * (1.0 - 0.0) / p.cells[d];
                E[d][d] = (0.0 + double(idx[d] + 1) * (1.0 - 0.0) / p.cells[d]) - X0[d];
            }
        } else {
            for (int d = 0; d < DIM; ++d) {
                X0[d] = p.X0[d * p.n_elem + e];
                for (int k = 0; k < DIM; ++k) E[d][k] = p.E[(d * DIM + k) * p.n_elem + e];
            }
        }
    }
    double* A = s_a + le * S::NB;
    double* B = s_b + le * S::NB;
    if (active) {
        double f[N1];
#pragma unroll
        for (int q = 0; q < N1; ++q) {
            const int qi[3] = {q, i1, i2};
            double x[3] = {X0[0], X0[1], X0[2]};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const double t = (p.xq[qi[d]] + 1) / 2;
#pragma unroll
                for (int k = 0; k < DIM; ++k) x[k] += t * E[d][k];
            }
            double v = 0;
            if (p.gaussian) {
                double r2 = 0;
#pragma unroll
                for (int d = 0; d < DIM; ++d) r2 += (x[d] - p.g[d]) * (x[d] - p.g[d]);
                v = p.g[4] * exp(-r2 / (p.g[3] * p.g[3]));
            } else {
                for (int m = 0; m < p.n_modes; ++m) {
                    double arg = 0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) arg += p.modes[4 * m + 1 + d] * x[d];
                    v += p.modes[4 * m] * sin(2 * M_PI * arg);
                }
            }
            f[q] = v;
        }
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            double t = 0;
#pragma unroll
            for (int q = 0; q < N1; ++q) t += p.P[a * N1 + q] * f[q];
            A[pc * N1 + a] = t;  // (a, q_y, q_z)
        }
    }
    __syncthreads();
    if (active) {  // y pass: this thread makes (., b = i1, q_z = i2)
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            double t = 0;
#pragma unroll
            for (int q = 0; q < N1; ++q) t += p.P[i1 * N1 + q] * A[(q + N1 * i2) * N1 + a];
            B[pc * N1 + a] = t;
        }
    }
    __syncthreads();
    if (!active) return;
    double coef[N1];
    if (DIM == 3) {  // z pass: (., b, c) from (., b, q_z)
#pragma unroll
        for (int a = 0; a < N1; ++a) {
            double t = 0;
#pragma unroll
            for (int q = 0; q < N1; ++q) t += p.P[i2 * N1 + q] * B[(i1 + N1 * q) * N1 + a];
            coef[a] = t;
        }
    } else {
#pragma unroll
        for (int a = 0; a < N1; ++a) coef[a] = B[pc * N1 + a];
    }
    for (int f = 0; f < p.nf; ++f) {
        if (p.field >= 0 && f != p.field) continue;
        const int64_t o = (e * p.nf + f) * S::NB + pc * N1;
        if (p.out_f32) {
#pragma unroll
            for (int a = 0; a < N1; ++a) static_cast<float*>(p.out)[o + a] = float(coef[a]);
        } else {
#pragma unroll
            for (int a = 0; a < N1; ++a) static_cast<double*>(p.out)[o + a] = coef[a];
        }
    }
}

// Per-element mass-norm of x - y: |J| dx^T (M1 x M1 x M1) dx, and max |dx|;
// one partial per block.
struct DiffParams {
    int64_t n_elem;
    const void* x;
    const void* y;
    int32_t x_f32, y_f32;
    const double* absdet;  // [n_elem / fields]
    int32_t fields;        // consecutive blocks of one element
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
        v = dx * src[le * S::NB + dof] * p.absdet[e / p.fields];
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

// The block-diagonal mass matrix (HpgemAPILinear::createMassMatrices, with
// the acoustic weighting of AcousticWaveLinear_Impl.h:102-129) applied to a
// vector, or solved against it, in place: per element and field the 1D factor
// A1 = M1 (apply) or Mi = M1^{-1} (solve) along every direction, then times
// |det J| w_f (apply) or divided by it (solve); w_f = c_inv on the s fields of
// the acoustic system, else 1.
struct MassParams {
    int64_t n_blocks;      // elements x fields, field-major inside an element
    int32_t fields;
    int32_t inverse;
    double* v;
    const double* absdet;  // [n_elem]
    const double* cinv;    // [n_elem], or NULL
    const double* A1;      // [n1][n1]
};

template <int DIM, int N1>
__global__ __launch_bounds__(kBlock) void tensor_mass(const MassParams p) {
    using S = Tensor<DIM, N1>;
    __shared__ double s_a[S::EPB * S::NB];
    __shared__ double s_b[S::EPB * S::NB];
    __shared__ double s_M[N1 * N1];
    const int tid = threadIdx.x;
    if (tid < N1 * N1) s_M[tid] = p.A1[tid];
    const int le = tid / S::L;
    const int dof = tid % S::L;
    const int64_t b = int64_t(blockIdx.x) * S::EPB + le;
    const bool active = (b < p.n_blocks) && (dof < S::NB);
    if (active) s_a[le * S::NB + dof] = p.v[b * S::NB + dof];
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
    if (active) {
        const int64_t e = b / p.fields;
        double w = p.absdet[e];
        if (p.cinv != nullptr && b % p.fields > 0) w *= p.cinv[e];
        const double x = src[le * S::NB + dof];
        p.v[b * S::NB + dof] = p.inverse ? x / w : x * w;
    }
}

template <typename F>
void dispatch_tensor(int dim, int n1, F&& f) {
#define HGDG_CASE(D, N) \
    if (dim == D && n1 == N) return f(std::integral_constant<int, D>{}, std::integral_constant<int, N>{});
    HGDG_CASE(2, 2) HGDG_CASE(2, 3) HGDG_CASE(2, 4) HGDG_CASE(2, 5) HGDG_CASE(2, 6)
    HGDG_CASE(3, 2) HGDG_CASE(3, 3) HGDG_CASE(3, 4) HGDG_CASE(3, 5) HGDG_CASE(3, 6)
#undef HGDG_CASE
    throw Failure(HGDG_ERR_UNSUPPORTED, "no tensor kernel for this dimension and order");
}

// Error of one field (or every field) against f(x - shift), f = sum_m amp_m
// sin(2 pi k_m . x) (HpgemAPISimplified::computeTotalError and computeMaxError,
// kernel/Base/HpgemAPISimplified_Impl.h:293-530): the squared L2 error with
// the (p+2)-point Gauss rule per direction (order 2p+2) and the largest
// pointwise error at the element rule's (p+1)^dim points (order 2p+1). One
// thread per element; per-block partial sums and maxima.
struct ErrParams {
    int64_t n_elem;
    int32_t nf, field;
    const void* u;
    int32_t u_f32;
    int32_t n_modes;
    const double* modes;   // [n_modes][4]: amp, k
    double shift[3];
    const double* X0;      // [dim][n_elem]
    const double* E;       // [dim*dim][n_elem]
    const double* absdet;  // [n_elem]
    const double* tab;     // psi at the L2 points [N1][N1+1], their weights and points [N1+1] each,
                           // psi at the element rule points [N1][N1], those points [N1]
    double* partial;       // [blocks][2]
};

template <int DIM, int N1>
__global__ __launch_bounds__(kBlock) void tensor_error(const ErrParams p) {
    constexpr int NB = N1 * N1 * (DIM == 3 ? N1 : 1);
    __shared__ double s_red[2][kBlock];
    const int64_t e = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    double sum = 0, mx = 0;
    if (e < p.n_elem) {
        double X0[3] = {0, 0, 0}, E[3][3] = {{0}};
        for (int d = 0; d < DIM; ++d) {
            X0[d] = p.X0[d * p.n_elem + e];
            for (int k = 0; k < DIM; ++k) E[d][k] = p.E[(d * DIM + k) * p.n_elem + e];
        }
        const double det = p.absdet[e];
        for (int fld = 0; fld < p.nf; ++fld) {
            if (p.field >= 0 && fld != p.field) continue;
            const int64_t base = (e * p.nf + fld) * NB;
            for (int pass = 0; pass < 2; ++pass) {
                const int NQ = pass == 0 ? N1 + 1 : N1;
                const double* psi = pass == 0 ? p.tab : p.tab + N1 * (N1 + 1) + 2 * (N1 + 1);
                const double* wq = p.tab + N1 * (N1 + 1);
                const double* xq = pass == 0 ? wq + (N1 + 1) : psi + N1 * N1;
                const int nz = DIM == 3 ? NQ : 1;
                for (int qz = 0; qz < nz; ++qz)
                    for (int qy = 0; qy < NQ; ++qy)
                        for (int qx = 0; qx < NQ; ++qx) {
                            double val = 0;
                            for (int c = 0; c < (DIM == 3 ? N1 : 1); ++c) {
                                double vy = 0;
                                for (int b = 0; b < N1; ++b) {
                                    double vx = 0;
                                    for (int a = 0; a < N1; ++a) {
                                        const int64_t i = base + a + N1 * (b + N1 * c);
                                        const double cf = p.u_f32 ? double(static_cast<const float*>(p.u)[i])
                                                                  : static_cast<const double*>(p.u)[i];
                                        vx += cf * psi[a * NQ + qx];
                                    }
                                    vy += vx * psi[b * NQ + qy];
                                }
                                val += DIM == 3 ? vy * psi[c * NQ + qz] : vy;
                            }
                            const int q[3] = {qx, qy, qz};
                            double x[3] = {X0[0], X0[1], X0[2]};
                            for (int d = 0; d < DIM; ++d) {
                                const double t = (xq[q[d]] + 1) / 2;
                                for (int k = 0; k < DIM; ++k) x[k] += t * E[d][k];
                            }
                            double ex = 0;
                            for (int m = 0; m < p.n_modes; ++m) {
                                double arg = 0;
                                for (int d = 0; d < DIM; ++d) arg += p.modes[4 * m + 1 + d] * (x[d] - p.shift[d]);
                                ex += p.modes[4 * m] * sin(2 * M_PI * arg);
                            }
                            const double err = ex - val;
                            if (pass == 0) {
                                double w = det;
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

// ---------------------------------------------------------------- host side
namespace {
const double* Dr_ptr(hgdg_ctx* c) { return c->tm.d_ops.ptr; }
const double* Mi_ptr(hgdg_ctx* c) { return c->tm.d_ops.ptr + c->tm.n1 * c->tm.n1; }
const double* M1_ptr(hgdg_ctx* c) { return c->tm.d_ops.ptr + 2 * c->tm.n1 * c->tm.n1; }
}  // namespace

// J (column d = E_d / 2), |det J| and J^{-1} of affine element e from its edge
// vectors E[d] = X_{2^d} - X_0; shared by the generic and the structured setup
// so both produce the same numbers.
void affine_geometry(hgdg_ctx* c, int64_t e, const double (&E)[3][3]) {
    const int dim = c->dim;
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
    c->tm.absdet[e] = std::fabs(det);
    for (int r = 0; r < dim; ++r)
        for (int col = 0; col < dim; ++col) c->tm.Jinv[(size_t(e) * dim + r) * dim + col] = inv[r][col];
}

void finish_tensor_upload(hgdg_ctx* c) {
    c->n_fields = 1;
    c->dof = c->n_elem * c->tm.nb;
    c->tm.ops = tensor_ops(c->tm.p);
    std::vector<double> opsv;
    opsv.insert(opsv.end(), c->tm.ops.Dr.begin(), c->tm.ops.Dr.end());
    opsv.insert(opsv.end(), c->tm.ops.Mi.begin(), c->tm.ops.Mi.end());
    opsv.insert(opsv.end(), c->tm.ops.M1.begin(), c->tm.ops.M1.end());
    c->tm.d_ops.upload(opsv, c->stream);
    c->tm.d_nbr.upload(c->tm.nbr, c->stream);
    c->tm.d_absdet.upload(c->tm.absdet, c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
}

// The periodic box [0,1]^dim with n[d] cells, as structured_mesh builds it
// (x fastest, vertex order x fastest, coordinates lo + i (hi - lo) / n), set
// up without the generic face search: neighbours are known from the indices.
void tensor_structured_upload(hgdg_ctx* c, int dim, const int32_t* n, int p) {
    require(dim == 2 || dim == 3, HGDG_ERR_ARG, "dim must be 2 or 3");
    require(p >= 1 && p <= 5, HGDG_ERR_UNSUPPORTED, "polynomial order must be in 1..5");
    int64_t ne = 1;
    for (int d = 0; d < dim; ++d) {
        require(n[d] >= 3, HGDG_ERR_ARG, "periodic boxes need at least 3 cells per direction");
        ne *= n[d];
    }
    c->shape = dim == 3 ? SHAPE_CUBE : SHAPE_SQUARE;
    c->dim = dim;
    c->n_elem = ne;
    c->tm.p = p;
    c->tm.n1 = p + 1;
    c->tm.nb = dim == 3 ? c->tm.n1 * c->tm.n1 * c->tm.n1 : c->tm.n1 * c->tm.n1;
    c->tm.perm = tensor_permutation(dim, p);
    c->tm.structured = true;
    for (int d = 0; d < 3; ++d) c->tm.cells[d] = d < dim ? n[d] : 1;
    c->tm.Jinv.assign(size_t(ne) * dim * dim, 0.0);
    c->tm.absdet.assign(ne, 0.0);
    c->tm.nbr.assign(size_t(2 * dim) * ne, -1);
    int64_t stride[3] = {1, n[0], dim == 3 ? int64_t(n[0]) * n[1] : 0};
    for (int64_t e = 0; e < ne; ++e) {
        int64_t idx[3] = {e % n[0], (e / n[0]) % n[1], dim == 3 ? e / (int64_t(n[0]) * n[1]) : 0};
        double E[3][3] = {{0}};
        for (int d = 0; d < dim; ++d)
            E[d][d] = (0.0 + double(idx[d] + 1) * (1.0 - 0.0) / n[d]) - (0.0 + double(idx[d]) * (1.0 - 0.0) / n[d]);
        affine_geometry(c, e, E);
        for (int d = 0; d < dim; ++d) {
            const int64_t lo_n = idx[d] == 0 ? e + (n[d] - 1) * stride[d] : e - stride[d];
            const int64_t hi_n = idx[d] == n[d] - 1 ? e - (n[d] - 1) * stride[d] : e + stride[d];
            c->tm.nbr[size_t(2 * d) * ne + e] = int32_t(lo_n);
            c->tm.nbr[size_t(2 * d + 1) * ne + e] = int32_t(hi_n);
        }
    }
    c->n_faces = int64_t(dim) * ne;
    c->n_internal = c->n_faces;
    finish_tensor_upload(c);
}

void tensor_mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m) {
    const int dim = shape_dim(m->shape);
    const int64_t ne = m->n_elements;
    int p = m->order ? m->order[0] : m->uniform_order;
    if (m->order)
        for (int64_t e = 0; e < ne; ++e)
            require(m->order[e] == p, HGDG_ERR_UNSUPPORTED,
                    "tensor elements need one polynomial order per mesh");
    require(p >= 1 && p <= 5, HGDG_ERR_UNSUPPORTED, "polynomial order must be in 1..5");
    const int nv = shape_vertices(m->shape);

    c->tm.p = p;
    c->tm.n1 = p + 1;
    c->tm.nb = dim == 3 ? c->tm.n1 * c->tm.n1 * c->tm.n1 : c->tm.n1 * c->tm.n1;
    c->tm.perm = tensor_permutation(dim, p);
    if (c->tensor_general_forced) {
        tensor_general_upload(c, m);
        return;
    }

    // geometry: affine hypercubes here (J column d = (X_{2^d} - X_0) / 2);
    // any other element sends the mesh to the Gauss-point path
    c->tm.Jinv.assign(size_t(ne) * dim * dim, 0.0);
    c->tm.absdet.assign(ne, 0.0);
    c->tm.X0.assign(size_t(ne) * dim, 0.0);
    c->tm.E.assign(size_t(ne) * dim * dim, 0.0);
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
            c->tm.X0[size_t(e) * dim + d] = X[d];
            for (int r = 0; r < dim; ++r) c->tm.E[(size_t(e) * dim + d) * dim + r] = E[d][r];
        }
        for (int v = 0; v < nv; ++v)
            for (int r = 0; r < dim; ++r) {
                double pred = X[r];
                for (int d = 0; d < dim; ++d)
                    if ((v >> d) & 1) pred += E[d][r];
                if (!(std::fabs(pred - X[v * dim + r]) <= 1e-12 * scale)) {
                    tensor_general_upload(c, m);
                    return;
                }
            }
        affine_geometry(c, e, E);
    }

    // faces: match sorted node tuples; the lower element index is hpGEM's
    // left element (faces are created while elements are visited in order).
    const int nfaces = 2 * dim;
    const int fverts = 1 << (dim - 1);
    c->tm.nbr.assign(size_t(nfaces) * ne, -1);
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
            require(c->tm.nbr[size_t(fo) * ne + o] < 0, HGDG_ERR_ARG, "a face is bounded by more than two elements");
            // the sum-factorised kernels need the neighbour's opposite face in
            // the same orientation: my face (d, s) against its (d, 1-s), vertex
            // by vertex; any other neighbour sends the mesh to the Gauss-point path
            bool aligned = fo / 2 == f / 2 && fo % 2 != f % 2;
            const int64_t* onid = m->node_ids + size_t(o) * nv;
            for (int q = 0; q < fverts && aligned; ++q) aligned = onid[face_vertex(fo, q)] == nid[face_vertex(f, q)];
            if (!aligned) {
                tensor_general_upload(c, m);
                return;
            }
            c->tm.nbr[size_t(f) * ne + e] = int32_t(o);
            c->tm.nbr[size_t(fo) * ne + o] = int32_t(e);
            open.erase(it);
            ++n_internal;
            ++n_total;
        }
    }
    n_total += int64_t(open.size());
    c->n_faces = n_total;
    c->n_internal = n_internal;
    finish_tensor_upload(c);
}

void tensor_operator_set(hgdg_ctx* c, const hgdg_op_desc* op) {
    if (c->tm.general) {
        tensor_general_operator_set(c, op);
        return;
    }
    const int dim = c->dim;
    const int64_t ne = c->n_elem;
    if (op->kind == HGDG_OP_ACOUSTIC_CENTRAL) {
        std::vector<double> geo(size_t(dim) * dim * ne + ne);
        for (int64_t e = 0; e < ne; ++e) {
            for (int r = 0; r < dim; ++r)
                for (int d = 0; d < dim; ++d)
                    geo[(size_t(r) * dim + d) * ne + e] = c->tm.Jinv[(size_t(e) * dim + r) * dim + d];
            geo[size_t(dim) * dim * ne + e] = c->cinv[e];
        }
        c->tm.d_geo.upload(geo, c->stream);
        c->n_fields = dim + 1;
        c->dof = ne * c->tm.nb * c->n_fields;
        HGDG_HIP(hipStreamSynchronize(c->stream));
        return;
    }
    c->n_fields = 1;
    c->dof = ne * c->tm.nb;
    std::vector<double> ahat(size_t(dim) * ne);
    std::vector<double> rownorm(size_t(dim) * ne);
    for (int64_t e = 0; e < ne; ++e)
        for (int d = 0; d < dim; ++d) {
            double s = 0, n2 = 0;
            for (int k = 0; k < dim; ++k) {
                const double v = c->tm.Jinv[(size_t(e) * dim + d) * dim + k];
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
            const int32_t o = c->tm.nbr[size_t(f) * ne + e];
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
    c->tm.d_geo.upload(ahat, c->stream);
    c->tm.d_mode.upload(mode, c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
}

static RhsParams advection_params(hgdg_ctx* c, const StageIO& io) {
    RhsParams p{};
    p.n_elem = c->n_elem;
    p.io = io;
    p.ahat = c->tm.d_geo.ptr;
    p.nbr = c->tm.d_nbr.ptr;
    p.mode = c->tm.d_mode.ptr;
    p.Dr = Dr_ptr(c);
    p.Mi = Mi_ptr(c);
    return p;
}

static void launch_advection(hgdg_ctx* c, const RhsParams& p, int tiles) {
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    dispatch_tensor(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = Pencil<DIMc, N1c>;
        const int64_t n_tiles = (c->n_elem + S::EPB - 1) / S::EPB;
        if (p.io.upd.u != nullptr) {  // the fused RK update runs in the one-tile kernel
            if (mixed)
                hipLaunchKernelGGL((tensor_advection_rhs<DIMc, N1c, float, true>), dim3(unsigned(n_tiles)),
                                   dim3(kBlock), 0, c->stream, p);
            else
                hipLaunchKernelGGL((tensor_advection_rhs<DIMc, N1c, double, true>), dim3(unsigned(n_tiles)),
                                   dim3(kBlock), 0, c->stream, p);
            return;
        }
        if (tiles > 1) {
            const int64_t grid = (n_tiles + tiles - 1) / tiles;
            if (mixed)
                hipLaunchKernelGGL((tensor_advection_rhs_tiled<DIMc, N1c, float>), dim3(unsigned(grid)), dim3(kBlock),
                                   0, c->stream, p, tiles);
            else
                hipLaunchKernelGGL((tensor_advection_rhs_tiled<DIMc, N1c, double>), dim3(unsigned(grid)),
                                   dim3(kBlock), 0, c->stream, p, tiles);
            return;
        }
        if (mixed)
            hipLaunchKernelGGL((tensor_advection_rhs<DIMc, N1c, float>), dim3(unsigned(n_tiles)), dim3(kBlock), 0,
                               c->stream, p);
        else
            hipLaunchKernelGGL((tensor_advection_rhs<DIMc, N1c, double>), dim3(unsigned(n_tiles)), dim3(kBlock), 0,
                               c->stream, p);
    });
    HGDG_HIP(hipGetLastError());
}

void tensor_tune(hgdg_ctx* c) {
    for (double& t : c->adv_tile_ms) t = 0;
    if (c->op.kind != HGDG_OP_ADVECTION_UPWIND || c->tm.general) return;
    if (c->adv_tiles_fixed > 0) {
        c->adv_tiles = c->adv_tiles_fixed;
        return;
    }
    // One RK stage's shape: u + a k -> k', with the stage vectors in the
    // storage type of the operator's precision. Values do not change the
    // timing (no data-dependent branches), so the inputs are zero.
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    const size_t n = size_t(c->n_dof());
    DevBuf<double> u, k, out;
    u.alloc(n);
    k.alloc(n);
    out.alloc(n);
    HGDG_HIP(hipMemsetAsync(u.ptr, 0, n * sizeof(double), c->stream));
    HGDG_HIP(hipMemsetAsync(k.ptr, 0, n * sizeof(double), c->stream));
    StageIO io{};
    io.n_in = 2;
    io.in[0] = u.ptr;
    io.coef[0] = 1.0;
    io.in[1] = k.ptr;
    io.coef[1] = 0.5;
    io.in_f32 = mixed ? 2u : 0u;
    io.out = out.ptr;
    io.out_f32 = mixed ? 1 : 0;
    const RhsParams p = advection_params(c, io);
    EventPair events;
    hipEvent_t* ev = events.e;
    constexpr int kCand[4] = {1, 2, 4, 8};
    constexpr int kReps = 20;
    constexpr int n_cand = 4;
    double best = 0;
    int best_tiles = 1;
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < n_cand; ++i) {
            for (int r = 0; r < 3; ++r) launch_advection(c, p, kCand[i]);
            HGDG_HIP(hipEventRecord(ev[0], c->stream));
            for (int r = 0; r < kReps; ++r) launch_advection(c, p, kCand[i]);
            HGDG_HIP(hipEventRecord(ev[1], c->stream));
            HGDG_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            HGDG_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            const double t = double(ms) / kReps;
            if (pass == 0 || t < c->adv_tile_ms[i]) c->adv_tile_ms[i] = t;
        }
    for (int i = 0; i < n_cand; ++i)
        if (i == 0 || c->adv_tile_ms[i] < best) {
            best = c->adv_tile_ms[i];
            best_tiles = kCand[i];
        }
    HGDG_HIP(hipStreamSynchronize(c->stream));
    c->adv_tiles = best_tiles;
}

// Largest dynamic LDS the fused RK kernel asks for (MI355X: 160 KB per
// workgroup; the kernel's tables take a few hundred bytes of static LDS).
constexpr size_t kRkTinyMaxLds = 150 * 1024;
constexpr int kRkTinyChunk = 500;  // steps per launch

bool tensor_rk_tiny(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b,
                    double dt, int n_steps) {
    if (c->rk_fused == 0 || c->simplex() || c->tm.general || c->op.kind != HGDG_OP_ADVECTION_UPWIND || n_steps < 1)
        return false;
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    const int64_t E = c->n_elem, n = c->n_dof();
    const size_t tb = mixed ? 4 : 8;
    const size_t bytes = size_t(16 * n) + size_t(4 * E * (1 + 2 * c->dim) + 4) + size_t(8 * c->dim * E) +
                         size_t(stages) * size_t(n) * tb;
    if (bytes > kRkTinyMaxLds) return false;
    RkTinyParams p{};
    StageIO io{};
    p.rhs = advection_params(c, io);
    Vec& u = vec_at(c, solution);
    p.u = u.f64->ptr;
    for (int j = 0; j < stages; ++j) {
        Vec& k = vec_at(c, aux[j]);
        require(k.is_f32 == mixed, HGDG_ERR_ARG, "stage vector storage does not match the operator's precision");
        p.k[j] = k.ptr();
    }
    p.stages = stages;
    p.n_steps = n_steps;
    for (int i = 0; i < stages; ++i)
        for (int j = 0; j < i; ++j) {
            p.ca[i * kMaxIn + j] = dt * a[i * stages + j];
            if (a[i * stages + j] != 0.0) p.amask[i] |= 1u << j;
        }
    for (int j = 0; j < stages; ++j) {
        p.cb[j] = dt * b[j];
        if (b[j] != 0.0) p.bmask |= 1u << j;
    }
    bool launched = false;
    dispatch_tensor(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        const void* fn = mixed ? reinterpret_cast<const void*>(&tensor_advection_rk_tiny<DIMc, N1c, float>)
                               : reinterpret_cast<const void*>(&tensor_advection_rk_tiny<DIMc, N1c, double>);
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)) != hipSuccess) {
            (void)hipGetLastError();  // this device cannot give the LDS: the graph path runs instead
            return;
        }
        constexpr unsigned kThreads = unsigned(rk_tiny_block<DIMc, N1c>());
        // at most kRkTinyChunk steps per launch, so that no launch runs for
        // long; the solution goes through HBM between launches exactly (FP64)
        for (int done = 0; done < n_steps; done += kRkTinyChunk) {
            RkTinyParams q = p;
            q.n_steps = std::min(kRkTinyChunk, n_steps - done);
            if (mixed)
                hipLaunchKernelGGL((tensor_advection_rk_tiny<DIMc, N1c, float>), dim3(1), dim3(kThreads), bytes,
                                   c->stream, q);
            else
                hipLaunchKernelGGL((tensor_advection_rk_tiny<DIMc, N1c, double>), dim3(1), dim3(kThreads), bytes,
                                   c->stream, q);
            HGDG_HIP(hipGetLastError());
            ++c->rk_fused_launches;
        }
        launched = true;
    });
    return launched;
}

void tensor_rhs(hgdg_ctx* c, const StageIO& io) {
    if (c->tm.general) {
        tensor_general_rhs(c, io);
        return;
    }
    const bool mixed = c->op.precision == HGDG_PREC_MIXED_FP32;
    if (c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL) {
        AcousticParams a{};
        a.n_elem = c->n_elem;
        a.io = io;
        a.jinv = c->tm.d_geo.ptr;
        a.cinv = c->tm.d_geo.ptr + size_t(c->dim) * c->dim * c->n_elem;
        a.nbr = c->tm.d_nbr.ptr;
        a.Dr = Dr_ptr(c);
        a.Mi = Mi_ptr(c);
        dispatch_tensor(c->dim, c->tm.n1, [&](auto D, auto N) {
            constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
            using S = AcousticShape<DIMc, N1c>;
            const int64_t grid = (c->n_elem + S::EPB - 1) / S::EPB;
            const bool upd = io.upd.u != nullptr;  // the fused RK update of a step's last stage
            if (mixed && upd)
                hipLaunchKernelGGL((tensor_acoustic_rhs<DIMc, N1c, float, true>), dim3(unsigned(grid)), dim3(kBlock),
                                   0, c->stream, a);
            else if (mixed)
                hipLaunchKernelGGL((tensor_acoustic_rhs<DIMc, N1c, float>), dim3(unsigned(grid)), dim3(kBlock), 0,
                                   c->stream, a);
            else if (upd)
                hipLaunchKernelGGL((tensor_acoustic_rhs<DIMc, N1c, double, true>), dim3(unsigned(grid)),
                                   dim3(kBlock), 0, c->stream, a);
            else
                hipLaunchKernelGGL((tensor_acoustic_rhs<DIMc, N1c, double>), dim3(unsigned(grid)), dim3(kBlock), 0,
                                   c->stream, a);
        });
        HGDG_HIP(hipGetLastError());
        return;
    }
    launch_advection(c, advection_params(c, io), c->adv_tiles);
}

void tensor_state_diff(hgdg_ctx* c, const Vec& vx, const Vec& vy, double* l2, double* max_abs) {
    if (c->tm.general) {
        tensor_general_state_diff(c, vx, vy, l2, max_abs);
        return;
    }
    DiffParams p{};
    p.n_elem = c->n_elem * c->n_fields;
    p.fields = c->n_fields;
    p.x = vx.ptr();
    p.y = vy.ptr();
    p.x_f32 = vx.is_f32;
    p.y_f32 = vy.is_f32;
    p.absdet = c->tm.d_absdet.ptr;
    p.M1 = M1_ptr(c);
    int64_t grid = 0;
    dispatch_tensor(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = Tensor<DIMc, N1c>;
        grid = (p.n_elem + S::EPB - 1) / S::EPB;
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

void tensor_mass_apply(hgdg_ctx* c, Vec& v, bool inverse) {
    if (c->tm.general) {
        tensor_general_mass_apply(c, v, inverse);
        return;
    }
    require(!v.is_f32, HGDG_ERR_UNSUPPORTED, "mass apply/solve needs a vector stored in FP64");
    MassParams p{};
    p.n_blocks = c->n_elem * c->n_fields;
    p.fields = c->n_fields;
    p.inverse = inverse ? 1 : 0;
    p.v = v.f64->ptr;
    p.absdet = c->tm.d_absdet.ptr;
    p.cinv = c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL ? c->tm.d_geo.ptr + size_t(c->dim) * c->dim * c->n_elem : nullptr;
    p.A1 = inverse ? Mi_ptr(c) : M1_ptr(c);
    dispatch_tensor(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = Tensor<DIMc, N1c>;
        const int64_t grid = (p.n_blocks + S::EPB - 1) / S::EPB;
        hipLaunchKernelGGL((tensor_mass<DIMc, N1c>), dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
    });
    HGDG_HIP(hipGetLastError());
}

void tensor_error_modes(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp, const double* k,
                        const double* shift, double* l2, double* max_abs) {
    if (c->tm.general) {
        tensor_general_error_modes(c, v, field, n_modes, amp, k, shift, l2, max_abs);
        return;
    }
    const int n1 = c->tm.n1, dim = c->dim;
    const int64_t ne = c->n_elem;
    std::vector<double> x2, w2, x1, w1;
    gauss_legendre(n1 + 1, x2, w2);
    gauss_legendre(n1, x1, w1);
    std::vector<double> tab;
    for (int a = 0; a < n1; ++a)
        for (int q = 0; q <= n1; ++q) tab.push_back(psi(a, x2[q]));
    tab.insert(tab.end(), w2.begin(), w2.end());
    tab.insert(tab.end(), x2.begin(), x2.end());
    for (int a = 0; a < n1; ++a)
        for (int q = 0; q < n1; ++q) tab.push_back(psi(a, x1[q]));
    tab.insert(tab.end(), x1.begin(), x1.end());
    for (int m = 0; m < n_modes; ++m) {
        tab.push_back(amp[m]);
        for (int d = 0; d < 3; ++d) tab.push_back(k[3 * m + d]);
    }
    // element origins and edge vectors, SoA
    std::vector<double> x0(size_t(dim) * ne), ed(size_t(dim) * dim * ne);
    for (int64_t e = 0; e < ne; ++e) {
        if (c->tm.structured) {
            const int64_t idx[3] = {e % c->tm.cells[0], (e / c->tm.cells[0]) % c->tm.cells[1],
                                    dim == 3 ? e / (int64_t(c->tm.cells[0]) * c->tm.cells[1]) : 0};
            for (int d = 0; d < dim; ++d) {
                x0[size_t(d) * ne + e] = double(idx[d]) / c->tm.cells[d];
                for (int r = 0; r < dim; ++r)
                    ed[(size_t(d) * dim + r) * ne + e] = r == d ? double(idx[d] + 1) / c->tm.cells[d] - double(idx[d]) / c->tm.cells[d] : 0.0;
            }
        } else {
            for (int d = 0; d < dim; ++d) {
                x0[size_t(d) * ne + e] = c->tm.X0[size_t(e) * dim + d];
                for (int r = 0; r < dim; ++r) ed[(size_t(d) * dim + r) * ne + e] = c->tm.E[(size_t(e) * dim + d) * dim + r];
            }
        }
    }
    DevBuf<double> d_tab, d_x0, d_e;
    d_tab.upload(tab, c->stream);
    d_x0.upload(x0, c->stream);
    d_e.upload(ed, c->stream);
    ErrParams p{};
    p.n_elem = ne;
    p.nf = c->n_fields;
    p.field = field;
    p.u = v.ptr();
    p.u_f32 = v.is_f32;
    p.n_modes = n_modes;
    p.modes = d_tab.ptr + (tab.size() - size_t(4) * n_modes);
    for (int d = 0; d < 3; ++d) p.shift[d] = shift ? shift[d] : 0.0;
    p.X0 = d_x0.ptr;
    p.E = d_e.ptr;
    p.absdet = c->tm.d_absdet.ptr;
    p.tab = d_tab.ptr;
    const int64_t grid = (ne + kBlock - 1) / kBlock;
    if (c->d_partial.n < size_t(2 * grid)) c->d_partial.alloc(size_t(2 * grid));
    p.partial = c->d_partial.ptr;
    dispatch_tensor(dim, n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        hipLaunchKernelGGL((tensor_error<DIMc, N1c>), dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
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

void tensor_to_device(hgdg_ctx* c, const double* host, std::vector<double>& dev) {
    const int nb = c->tm.nb;
    const int64_t blocks = c->n_elem * c->n_fields;  // field-major blocks of nb per element
    dev.assign(size_t(c->n_dof()), 0.0);
    for (int64_t b = 0; b < blocks; ++b)
        for (int i = 0; i < nb; ++i) dev[size_t(b) * nb + c->tm.perm[i]] = host[size_t(b) * nb + i];
}

void tensor_from_device(hgdg_ctx* c, const std::vector<double>& dev, double* host) {
    const int nb = c->tm.nb;
    const int64_t blocks = c->n_elem * c->n_fields;
    for (int64_t b = 0; b < blocks; ++b)
        for (int i = 0; i < nb; ++i) host[size_t(b) * nb + i] = dev[size_t(b) * nb + c->tm.perm[i]];
}

// ---------------------------------------------------------------- batched launches
// Problems of one family (kind, dim, order) advance together: their per-problem
// parameter blocks go into `params` (device, written by the caller beforehand
// through `host_staging`), and one launch covers all their elements.
size_t tensor_batch_param_bytes(const hgdg_ctx* c) {
    return c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL ? sizeof(AcousticParams) : sizeof(RhsParams);
}

int64_t tensor_batch_blocks(const hgdg_ctx* c) {
    int64_t blocks = 0;
    dispatch_tensor(c->dim, c->tm.n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        const int64_t epb = c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL ? AcousticShape<DIMc, N1c>::EPB
                                                                   : Pencil<DIMc, N1c>::EPB;
        blocks = (c->n_elem + epb - 1) / epb;
    });
    return blocks;
}

void tensor_batch_param(hgdg_ctx* c, const StageIO& io, void* out) {
    if (c->op.kind == HGDG_OP_ACOUSTIC_CENTRAL) {
        AcousticParams a{};
        a.n_elem = c->n_elem;
        a.io = io;
        a.jinv = c->tm.d_geo.ptr;
        a.cinv = c->tm.d_geo.ptr + size_t(c->dim) * c->dim * c->n_elem;
        a.nbr = c->tm.d_nbr.ptr;
        a.Dr = Dr_ptr(c);
        a.Mi = Mi_ptr(c);
        std::memcpy(out, &a, sizeof(a));
    } else {
        RhsParams r{};
        r.n_elem = c->n_elem;
        r.io = io;
        r.ahat = c->tm.d_geo.ptr;
        r.nbr = c->tm.d_nbr.ptr;
        r.mode = c->tm.d_mode.ptr;
        r.Dr = Dr_ptr(c);
        r.Mi = Mi_ptr(c);
        std::memcpy(out, &r, sizeof(r));
    }
}

void tensor_batch_launch(hipStream_t stream, int kind, int dim, int n1, bool mixed, bool upd, const void* params,
                         const int64_t* off, int n, int64_t blocks) {
    if (n == 0 || blocks == 0) return;
    dispatch_tensor(dim, n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        const dim3 grid{unsigned(blocks), 1, 1}, block{unsigned(kBlock), 1, 1};
        auto go = [&](auto T_, auto U_) {
            using Tc = decltype(T_);
            constexpr bool Uc = decltype(U_)::value;
            if (kind == HGDG_OP_ACOUSTIC_CENTRAL)
                hipLaunchKernelGGL((tensor_acoustic_rhs_batched<DIMc, N1c, Tc, Uc>), grid, block, 0, stream,
                                   static_cast<const AcousticParams*>(params), off, n);
            else
                hipLaunchKernelGGL((tensor_advection_rhs_batched<DIMc, N1c, Tc, Uc>), grid, block, 0, stream,
                                   static_cast<const RhsParams*>(params), off, n);
        };
        if (mixed && upd) go(float{}, std::true_type{});
        else if (mixed) go(float{}, std::false_type{});
        else if (upd) go(double{}, std::true_type{});
        else go(double{}, std::false_type{});
    });
    HGDG_HIP(hipGetLastError());
}

void tensor_project_device(hgdg_ctx* c, Vec& v, int field, int n_modes, const double* amp, const double* k,
                           const double* gauss) {
    if (c->tm.general) {
        tensor_general_project(c, v, field, n_modes, amp, k, gauss);
        return;
    }
    const int n1 = c->tm.n1, dim = c->dim;
    std::vector<double> xq, wq;
    gauss_legendre(n1, xq, wq);
    std::vector<double> P(size_t(n1) * n1, 0.0);
    for (int a = 0; a < n1; ++a)
        for (int q = 0; q < n1; ++q) {
            double s = 0;
            for (int b = 0; b < n1; ++b) s += c->tm.ops.Mi[a * n1 + b] * psi(b, xq[q]);
            P[a * n1 + q] = s * wq[q];
        }
    std::vector<double> tab(P);
    tab.insert(tab.end(), xq.begin(), xq.end());
    for (int m = 0; m < n_modes; ++m) {
        tab.push_back(amp[m]);
        for (int d = 0; d < 3; ++d) tab.push_back(k[3 * m + d]);
    }
    DevBuf<double> d_tab;
    d_tab.upload(tab, c->stream);
    DevBuf<double> d_x0, d_e;
    ProjParams pp{};
    pp.n_elem = c->n_elem;
    pp.nf = c->n_fields;
    pp.field = field;
    pp.out = v.ptr();
    pp.out_f32 = v.is_f32;
    pp.n_modes = n_modes;
    if (gauss) {
        pp.gaussian = 1;
        for (int q = 0; q < 5; ++q) pp.g[q] = gauss[q];
    }
    pp.P = d_tab.ptr;
    pp.xq = d_tab.ptr + size_t(n1) * n1;
    pp.modes = d_tab.ptr + size_t(n1) * n1 + n1;
    pp.structured = c->tm.structured;
    for (int d = 0; d < 3; ++d) pp.cells[d] = c->tm.cells[d];
    if (!c->tm.structured) {
        const int64_t ne = c->n_elem;
        std::vector<double> x0(size_t(dim) * ne), ed(size_t(dim) * dim * ne);
        for (int64_t e = 0; e < ne; ++e)
            for (int d = 0; d < dim; ++d) {
                x0[size_t(d) * ne + e] = c->tm.X0[size_t(e) * dim + d];
                for (int r = 0; r < dim; ++r) ed[(size_t(d) * dim + r) * ne + e] = c->tm.E[(size_t(e) * dim + d) * dim + r];
            }
        d_x0.upload(x0, c->stream);
        d_e.upload(ed, c->stream);
        pp.X0 = d_x0.ptr;
        pp.E = d_e.ptr;
    }
    dispatch_tensor(dim, n1, [&](auto D, auto N) {
        constexpr int DIMc = decltype(D)::value, N1c = decltype(N)::value;
        using S = Pencil<DIMc, N1c>;
        const int64_t grid = (c->n_elem + S::EPB - 1) / S::EPB;
        hipLaunchKernelGGL((tensor_project<DIMc, N1c>), dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, pp);
    });
    HGDG_HIP(hipGetLastError());
    HGDG_HIP(hipStreamSynchronize(c->stream));  // before the temporary tables go
}

}  // namespace hgdg
