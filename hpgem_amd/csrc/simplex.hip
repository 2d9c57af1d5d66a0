// Tetrahedral elements with per-element order p = 1..4 (hp meshes): the
// first-order acoustic wave system of hpGEM's AcousticWaveLinear
// (applications/ExampleMultipleVariableProblem/AcousticWaveLinear_Impl.h:
// 102-283): fields v and s_1..s_3, mass blocks M (v) and c_inv M (s),
//   r_v   = int phi grad.s  - 1/2 int_f phi n_L.(s_L - s_R)   (internal faces)
//   r_s_d = int phi d_d v   - 1/2 int_f phi n_L,d (v_L - v_R)
// and on reflecting walls r_v gets - int_f phi n.s and r_s nothing.
//
// On an affine tetrahedron the inverse-mass apply turns the volume terms into
// exact reference derivative matrices Dr_e (d/dxi_e in P_p) combined with
// J^{-1}, and each face term into a lifting of the flux's face-polynomial
// coefficients: face traces are projected (exactly: everything is a
// polynomial of degree <= 4) onto an orthonormal basis of P_4 on the reference
// triangle, the flux is formed coefficient by coefficient, and
// Lift = M_ref^{-1} P^T scaled by area/volume brings it back. Both sides of a
// face are evaluated in the left element's face coordinates, the right side
// through hpGEM's orientation map (FaceGeometry.h:372-393). Elements are
// grouped by order on the device; one launch covers all groups, each block
// working on elements of one order.
#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>

#include "context.hpp"

namespace hgdg {

namespace {

// int32 table index per order (p-1): offsets into the FP64 (and the FP32)
// table arrays and, per local face, the offset of its list of face columns
// (the basis functions whose trace on that face does not vanish). Face terms
// work with an orthonormal hierarchical basis chi_m of P_4 on the reference
// triangle (degree by degree): the trace tables give the chi-coefficients of
// each face column's trace when the face is parametrised through orientation
// o, the lift tables map flux coefficients back to the element.
constexpr int kTixDrT = 0;           // [3][nb][nb]: DrT[r][j][i] = Dr_r(i, j)
constexpr int kTixM = 1;             // [nb][nb]: M_ref
constexpr int kTixCols = 2;          // + f
constexpr int kTixProj = 6;          // [4 f][6 o][ncol][15 m]: chi_m-coefficients of the face columns' traces
constexpr int kTixLift = 7;          // [4 f][ncol m][nb i]: (M_ref^{-1} int_f phi chi_m)_i, own face coordinates
constexpr int kTixProjId = 8;        // [4 f][ncol c][ncol m]: the trace table at the identity orientation, m < ncol
constexpr int kTixDrP = 9;           // orders >= 3: [3][4 ceil(nb/4) k][16 ceil(nb/16) i] DrT padded with zeros
constexpr int kTixStride = 10;
constexpr int kTixRel = 4 * kTixStride;  // [6 own][6 neighbour]: orientation of the neighbour's face seen from ours
constexpr int kTixIdent = kTixRel + 36;  // the identity orientation
constexpr int kTixRot = kTixIdent + 1;   // [6 o][55]: the orientation maps on face-basis coefficients, by degree block
constexpr int kMaxCols = 15;         // face columns of an order-4 element
constexpr int kTixMax = kTixRot + 1 + 4 * 4 * kMaxCols;  // whole index table

// packed per-face info
__host__ __device__ constexpr int info_lf(int v) { return v & 3; }
__host__ __device__ constexpr int info_oown(int v) { return (v >> 2) & 7; }
__host__ __device__ constexpr int info_onbr(int v) { return (v >> 5) & 7; }
__host__ __device__ constexpr int info_pn(int v) { return (v >> 8) & 7; }
// the orientation that maps the neighbour's face coordinates onto ours,
// o_nbr o o_own^{-1} (bits set once the orientation group table is built)
__host__ __device__ constexpr int info_rel(int v) { return (v >> 11) & 7; }
constexpr int info_pack(int lf, int oown, int onbr, int pn) { return lf | (oown << 2) | (onbr << 5) | (pn << 8); }

struct RefTetRule {
    int order, n;
    const double* pts;  // [n][3] on the reference tetrahedron
    const double* w;
};
#include "tet_rules.inc"
static_assert(kFaceQ == 16, "the table index reserves 16 face points");

}  // namespace

struct SimplexParams {
    StageIO io;
    int64_t n_elem;
    int64_t group_start[5];
    int64_t group_dof[5];
    int32_t group_block[5];
    const double* jinv;    // [9][E]: (J^{-1})_{e,d} at [(3e + d) E + elem]
    const double* cinv;    // [E]
    const double* sigma;   // [4][E]: face area ratio / |det J|
    const double* normal;  // [3][4][E]: outward unit normal
    const int32_t* nbr;    // [4][E]: neighbour device element, -1 on walls
    const int32_t* info;   // [4][E]
    const double* tab;
    const float* tab32;    // the same tables in FP32 (mixed path), same offsets
    const int32_t* tix;
    int32_t n_tix;
    const int4* blocks;    // per block: order, first device element, element count, threads-per-element variant (informative)
    double2* trace;        // face traces of the stage input: [element][4 f][ncol(p) m] (n.s, v), own face coordinates
    int64_t group_tr[5];   // first trace of each order group
};

__host__ __device__ constexpr int simplex_nb(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
__host__ __device__ constexpr int simplex_ncol(int p) { return (p + 1) * (p + 2) / 2; }
// Threads per element of the RHS kernel, two candidates per order (a row
// phase needs at least nb threads): variant 0 = 8 at order 1 (32 elements per
// block), 12 at order 2 (21), 32 at order 3 (8), nb = 35 at order 4 (7);
// variant 1 = 16 at order 1, 32 at order 2, 64 (one element per wavefront)
// above. Small orders gain from many elements per block (fewer blocks, each
// with its barriers and table staging). The library times both per order
// when the operator is set and keeps the faster; results do not depend on
// the choice (every row and item is computed by the same instructions
// whichever thread takes it).
__host__ __device__ constexpr int simplex_tpe(int p) { return p == 1 ? 8 : p == 2 ? 12 : p == 3 ? 32 : simplex_nb(p); }
__host__ __device__ constexpr int simplex_tpe_alt(int p) { return p == 1 ? 16 : p == 2 ? 32 : 64; }
__host__ __device__ constexpr int simplex_tpe_of(int p, int variant) { return variant ? simplex_tpe_alt(p) : simplex_tpe(p); }

template <int P, int TPE_ = simplex_tpe(P), int TB = 8>
struct SimplexShape {
    static_assert(TPE_ >= simplex_nb(P) && TPE_ <= 256, "a row phase needs one thread per basis function");
    static constexpr int NB = simplex_nb(P);
    static constexpr int NCOL = simplex_ncol(P);
    static constexpr int TPE = TPE_;  // threads per element
    static constexpr int EPB = kBlock / TPE;
    // doubles taken by n values of the arithmetic type (TB bytes), rounded
    // up to an even count (regions start double2-aligned)
    static constexpr int tw(int n) { return (n * TB + 15) / 16 * 2; }
    // LDS carve-up of the RHS kernel, in doubles. FM and SF live through the
    // whole kernel; the work area W holds the volume data (G, DR) and, at
    // order 4 in FP64, afterwards one face's lift table; LIFT follows.
    static constexpr int FM = 0;                            // [EPB][4 faces][4]: unit normal, sigma
    static constexpr int SF = FM + EPB * 16;                // [EPB][4 faces][NCOL] (T pairs): flux coefficients
    static constexpr int W = SF + tw(EPB * 8 * NCOL);
    static constexpr int G = W;                             // [EPB][4][NB] (T): v, then g_r = sum_d (J^{-1})_{r,d} s_d
    // DrT: read through the caches by the matrix-core tiles (orders 3 and 4,
    // from a zero-padded copy of the table); otherwise staged in LDS
    static constexpr bool DR_GLOBAL = P >= 3;
    static constexpr int DR = G + tw(EPB * 4 * NB);         // [3][NB][NB] (T): DrT
    static constexpr int VOL_END = DR + (DR_GLOBAL ? 0 : tw(3 * NB * NB));
    static constexpr bool LIFT_LDS = P < 4;                 // lift table staged in LDS for the whole kernel
    // order 4: the lift tables pass through the work area once the volume
    // terms are done, LIFT_CHUNK faces at a time
    static constexpr bool LIFT_ONE = !LIFT_LDS;
    static constexpr int LIFT_CHUNK = 2;                    // faces per pass through the work area
    static constexpr int LIFT = VOL_END;                    // [4 f][NCOL m][NB i] (T), the whole kernel
    static constexpr int W_END = LIFT_ONE && W + tw(LIFT_CHUNK * NCOL * NB) > VOL_END ? W + tw(LIFT_CHUNK * NCOL * NB) : VOL_END;
    static constexpr int TOTAL = LIFT_LDS ? LIFT + tw(4 * NCOL * NB) : W_END;
    static_assert(SF % 2 == 0 && W % 2 == 0 && DR % 2 == 0 && LIFT % 2 == 0, "double2 alignment");
};

// Face-basis degree of coefficient m < 15 (degree d holds m in
// [d(d+1)/2, (d+1)(d+2)/2)).
__host__ __device__ constexpr int face_degree(int m) { return (m >= 1) + (m >= 3) + (m >= 6) + (m >= 10); }

typedef double mfma_d4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mfma_d4 mfma_step(double a, double b, mfma_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mfma_f4 mfma_step(float a, float b, mfma_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T>
struct FluxPair {
    T v, s;  // flux coefficient of the v equation, of the s equations
};

// ---- face traces (the first kernel of a right-hand side)
// Every element writes, for each local face f, the chi_m-coefficients
// (m < (p+1)(p+2)/2, degree <= p) of n_f.s and of v on that face in its own
// face coordinates, n_f its own outward unit normal, into the trace buffer;
// the element's neighbours read them in the RHS kernel. One block takes EPB
// elements of one order, TPE threads each: stage input, normals and the
// identity-orientation trace table into LDS; items (face, column): n.s and v
// at the face columns; items (face, m): their projections.
template <int P>
struct TraceShape {
    static constexpr int NB = simplex_nb(P), NC = simplex_ncol(P), TPE = simplex_tpe(P), EPB = kBlock / TPE;
    static constexpr int FM = 0;                    // [EPB][4 f][3] unit normals
    static constexpr int U = FM + EPB * 12;         // [EPB][4][NB] stage input
    static constexpr int SO = U + EPB * 4 * NB;     // [EPB][4 f][NC] (double2): n.s and v at the face columns
    static constexpr int PR = SO + EPB * 8 * NC;    // [4 f][NC c][NC m]: identity-orientation trace table
    static constexpr int CL = PR + 4 * NC * NC;     // [4 f][NC] (int32, as NC * 2 doubles): the face column lists
    static constexpr int TOTAL = CL + 2 * NC;
    static_assert(SO % 2 == 0, "double2 alignment");
};

constexpr int trace_lds_doubles() {
    constexpr int t[4] = {TraceShape<1>::TOTAL, TraceShape<2>::TOTAL, TraceShape<3>::TOTAL, TraceShape<4>::TOTAL};
    int m = 0;
    for (int k = 0; k < 4; ++k) m = t[k] > m ? t[k] : m;
    return m;
}

template <int P>
__device__ __forceinline__ void simplex_trace_block(const SimplexParams& p, int64_t first, int count, double* lds) {
    using S = TraceShape<P>;
    constexpr int NB = S::NB, NC = S::NC;
    const int tid = threadIdx.x;
    const int le = tid / S::TPE;
    const int i = tid % S::TPE;
    const int64_t e = first + le;
    const bool elem_ok = le < count;
    const int64_t E = p.n_elem;
    const int32_t* tx = p.tix + (P - 1) * kTixStride;
    double* fm = lds + S::FM + le * 12;
    double* su = lds + S::U + le * 4 * NB;
    if (elem_ok) {
        for (int k = i; k < 12; k += S::TPE) fm[k] = p.normal[((k % 3) * 4 + k / 3) * E + e];  // fm[3 f + c]
        if (i < NB) {
            const int64_t base = p.group_dof[P - 1] + (e - p.group_start[P - 1]) * 4 * NB;
            const int64_t at[4] = {base + i, base + NB + i, base + 2 * NB + i, base + 3 * NB + i};
            double sv[4];
            combine_n(p.io, at, sv);
#pragma unroll
            for (int f = 0; f < 4; ++f) su[f * NB + i] = sv[f];
        }
    }
    {
        const double* gP = p.tab + tx[kTixProjId];
        double* sP = lds + S::PR;
#pragma unroll 4
        for (int k = tid; k < 4 * NC * NC; k += kBlock) sP[k] = gP[k];
        int32_t* sc = reinterpret_cast<int32_t*>(lds + S::CL);
        for (int k = tid; k < 4 * NC; k += kBlock) sc[k] = p.tix[tx[kTixCols + k / NC] + k % NC];
    }
    __syncthreads();
    double2* so = reinterpret_cast<double2*>(lds + S::SO) + le * 4 * NC;
    if (elem_ok) {
        for (int it = i; it < 4 * NC; it += S::TPE) {
            const int f = it / NC;
            const int col = reinterpret_cast<const int32_t*>(lds + S::CL)[it];
            so[it] = make_double2(fm[f * 3] * su[NB + col] + fm[f * 3 + 1] * su[2 * NB + col] +
                                      fm[f * 3 + 2] * su[3 * NB + col],
                                  su[col]);
        }
    }
    __syncthreads();
    if (elem_ok) {
        double2* out = p.trace + p.group_tr[P - 1] + (e - p.group_start[P - 1]) * 4 * NC;
        for (int it = i; it < 4 * NC; it += S::TPE) {
            const int f = it / NC, m = it % NC;
            const double* col = lds + S::PR + f * NC * NC + m;
            const double2* x = so + f * NC;
            double a = 0, b = 0;
#pragma unroll 5
            for (int c = 0; c < NC; ++c) {
                const double r = col[c * NC];
                const double2 v = x[c];
                a += r * v.x;
                b += r * v.y;
            }
            out[it] = make_double2(a, b);
        }
    }
}

// Blocks in order-group order: group g takes blocks [group_block[g],
// group_block[g + 1]) of 256 / simplex_tpe(g + 1) elements each.
__global__ __launch_bounds__(kBlock) void simplex_trace(const SimplexParams p) {
    __shared__ __align__(16) double lds[trace_lds_doubles()];
    const int b = int(blockIdx.x);
    int g = 0;
    while (g < 3 && b >= p.group_block[g + 1]) ++g;
    const int epb = kBlock / simplex_tpe(g + 1);
    const int64_t first = p.group_start[g] + int64_t(b - p.group_block[g]) * epb;
    const int64_t left = p.group_start[g + 1] - first;
    const int count = left < epb ? int(left) : epb;
    switch (g) {
        case 0: simplex_trace_block<1>(p, first, count, lds); break;
        case 1: simplex_trace_block<2>(p, first, count, lds); break;
        case 2: simplex_trace_block<3>(p, first, count, lds); break;
        default: simplex_trace_block<4>(p, first, count, lds); break;
    }
}

// ---- the right-hand side (the second kernel)
// Tetrahedral elements of order P: a block takes EPB consecutive device
// elements of that order, TPE threads per element, in phases separated by
// barriers:
//  1. the stage input (all four fields) as v and g_r = sum_d (J^{-1})_{r,d}
//     s_d per basis function, the faces' normals and area ratios, and the
//     order's derivative and lift tables into LDS; items (face, m): the flux
//     coefficients from the element's own traces and its neighbours' (in the
//     neighbour's face coordinates, brought into ours by the orientation
//     map's degree block, R_o; the neighbour's n.s is with its own normal,
//     the negative of ours), central on internal faces, the wall flux on
//     reflecting walls;
//  2. threads of a row: the volume terms;
//  3. threads of a row: the lift of the flux coefficients.
// Every face term is a polynomial of degree <= 4 in the face coordinates and
// chi spans P_4, so the coefficients represent it exactly; only chi_m of
// degree <= p reach the lift (chi_m of higher degree are orthogonal to the
// element's traces), so a face of an order-p element carries (p+1)(p+2)/2
// coefficients, and the orientation maps keep degrees apart, so a neighbour of
// order p_n contributes to degrees <= min(p, p_n) only. T = double: FP64;
// T = float: the volume contractions and the lifts in FP32 on FP32 tables,
// compensated per-row sums, traces and flux coefficients in FP64.
template <int P, int TPE, typename T>
__device__ __forceinline__ void simplex_block(const SimplexParams& p, int64_t first, int count, double* lds) {
    using S = SimplexShape<P, TPE, int(sizeof(T))>;
    constexpr int NB = S::NB;
    constexpr int NC = S::NCOL;
    const int tid = threadIdx.x;
    const int le = tid / S::TPE;
    const int i = tid % S::TPE;
    const int64_t e = first + le;
    const bool elem_ok = le < count;
    const bool row_ok = elem_ok && (i < NB);
    const int64_t E = p.n_elem;
    const int64_t base = p.group_dof[P - 1] + (e - p.group_start[P - 1]) * 4 * NB;
    double* fm = lds + S::FM + le * 16;
    FluxPair<T>* sf = reinterpret_cast<FluxPair<T>*>(lds + S::SF) + le * 4 * NC;
    T* sg = reinterpret_cast<T*>(lds + S::G) + le * 4 * NB;
    T* sDr = reinterpret_cast<T*>(lds + S::DR);
    const T* tabT = std::is_same<T, float>::value ? reinterpret_cast<const T*>(p.tab32)
                                                  : reinterpret_cast<const T*>(p.tab);

    // ---- phase 1: geometry, stage input as v and g, flux coefficients, tables
    double cinv = 1.0;
    if (elem_ok) {
        cinv = p.cinv[e];
        for (int k = i; k < 16; k += S::TPE) {
            const int f = k / 4, c = k % 4;
            fm[k] = c < 3 ? p.normal[(c * 4 + f) * E + e] : p.sigma[f * E + e];
        }
    }
    if (row_ok) {
        double Ji[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) Ji[q] = p.jinv[q * E + e];
        const int64_t at[4] = {base + i, base + NB + i, base + 2 * NB + i, base + 3 * NB + i};
        double sv[4];
        combine_n(p.io, at, sv);
        sg[i] = T(sv[0]);
#pragma unroll
        for (int r = 0; r < 3; ++r) sg[(1 + r) * NB + i] = T(Ji[r * 3] * sv[1] + Ji[r * 3 + 1] * sv[2] + Ji[r * 3 + 2] * sv[3]);
    }
    if (elem_ok) {
        const double2* own = p.trace + p.group_tr[P - 1] + (e - p.group_start[P - 1]) * 4 * NC;
        const double* rot = p.tab + p.tix[kTixRot];
        // every item's own trace and face data first, then the neighbours'
        // coefficients (their addresses need the face data)
        constexpr int NIT = (4 * NC + S::TPE - 1) / S::TPE;  // items per thread
        int nb_[NIT], inf_[NIT];
        double2 a_[NIT];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int it = i + k * S::TPE;
            nb_[k] = -1; inf_[k] = 0; a_[k] = make_double2(0, 0);
            if (it < 4 * NC) {
                const int f = it / NC;
                nb_[k] = p.nbr[f * E + e];
                inf_[k] = p.info[f * E + e];
                a_[k] = own[it];
            }
        }
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const int it = i + k * S::TPE;
            if (it >= 4 * NC) continue;
            const int m = it % NC;
            const int nbr = nb_[k];
            const double2 a = a_[k];
            double Fv, Fs;
            if (nbr >= 0) {
                const int inf = inf_[k];
                const int pn = info_pn(inf);
                const int d = face_degree(m);
                double bns = 0, bv = 0;
                if (d <= pn) {
                    const int o = info_rel(inf);
                    const int m0 = d * (d + 1) / 2, ncn = simplex_ncol(pn);
                    const double* R = rot + o * 55 + d * (d + 1) * (2 * d + 1) / 6 + (m - m0) * (d + 1);
                    const double2* x = p.trace + p.group_tr[pn - 1] + (nbr - p.group_start[pn - 1]) * 4 * ncn +
                                       info_lf(inf) * ncn + m0;
                    double rr[5];
                    double2 xv[5];
#pragma unroll
                    for (int b = 0; b < 5; ++b)
                        if (b <= d) {
                            rr[b] = R[b];
                            xv[b] = x[b];
                        }
#pragma unroll
                    for (int b = 0; b < 5; ++b)
                        if (b <= d) {
                            bns += rr[b] * xv[b].x;
                            bv += rr[b] * xv[b].y;
                        }
                }
                Fv = -0.5 * (a.x + bns);  // n.(s_own - s_nbr), the neighbour's traces with n_nbr = -n
                Fs = -0.5 * (a.y - bv);
            } else {
                Fv = -a.x;
                Fs = 0;
            }
            sf[it] = FluxPair<T>{T(Fv), T(Fs)};
        }
    }
    const T* gDr = tabT + p.tix[(P - 1) * kTixStride + kTixDrT];
    if constexpr (!S::DR_GLOBAL) {
#pragma unroll 4
        for (int k = tid; k < 3 * NB * NB; k += kBlock) sDr[k] = gDr[k];
    }
    const T* gLift = tabT + p.tix[(P - 1) * kTixStride + kTixLift];
    T* sLift = reinterpret_cast<T*>(lds + S::LIFT);
    if constexpr (S::LIFT_LDS) {
#pragma unroll 4
        for (int k = tid; k < 4 * NC * NB; k += kBlock) sLift[k] = gLift[k];
    }
    __syncthreads();

    // ---- phase 2: volume, k_v = sum_r Dr_r g_r ; k_s_d = (1/c_inv) sum_r (J^{-1})_{r,d} Dr_r v
    // Orders 3 and 4 on the matrix cores: per direction r, Dr_r times the
    // block's columns [g_r of elements 0..7 | v of elements 0..7] in
    // 16 x 16 x 4 tiles, one row tile per wavefront (A(i, k) from lane
    // i + 16 k, B(k, j) from lane j + 16 k; output q of lane l holds
    // D(4 q + l / 16, l % 16) for v_mfma_f64_16x16x4f64 and
    // D(4 (l / 16) + q, l % 16) for v_mfma_f32_16x16x4f32). In FP64 each
    // output accumulates its products in k order, one fused multiply-add
    // after another, as the row loop below does: the results are bitwise
    // those of the VALU form (checked on the 24^3 x 5 workload); in FP32 the
    // 100-step RK4 run of that workload gives the row loop's error against
    // FP64 to every digit. Orders 1 and 2: the row loop.
    Acc<T> kv, ks[3];
    // order 4: the first lift chunk is loaded into registers while the
    // volume terms finish, and stored to LDS once they are done
    constexpr int LPRE = S::LIFT_ONE ? (S::LIFT_CHUNK * NC * NB + kBlock - 1) / kBlock : 1;
    T lpre[LPRE];
    if constexpr (P >= 3) {
        static_assert(S::EPB <= 8, "16 columns");
        static_assert(S::DR_GLOBAL, "the tiles read the padded table through the caches");
        constexpr int NTM = (NB + 15) / 16, KS = (NB + 3) / 4;
        constexpr int NBP = 16 * NTM, KP = 4 * KS;  // the padded table: [3][KP][NBP]
        using V4 = typename std::conditional<std::is_same<T, double>::value, mfma_d4, mfma_f4>::type;
        const int lane = tid % 64, wave = tid / 64;
        const int col = lane % 16, kk = lane / 16;
        const int ce = col % 8, cv = col / 8;
        T* sG0 = reinterpret_cast<T*>(lds + S::G);
        V4 acc[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[r] = V4{T(0), T(0), T(0), T(0)};
        if (wave < NTM) {
            // per direction the A operands first (zeros outside nb x nb come
            // from the table, so every load is unconditional), then the
            // accumulation chain
            const T* Dr = tabT + p.tix[(P - 1) * kTixStride + kTixDrP] + kk * NBP +
                          16 * wave + col;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                T av[KS];
#pragma unroll
                for (int s = 0; s < KS; ++s) av[s] = Dr[(r * KP + 4 * s) * NBP];
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    const int k = 4 * s + kk;
                    const T bv = (ce < count && k < NB) ? sG0[ce * 4 * NB + (cv ? 0 : 1 + r) * NB + k] : T(0);
                    acc[r] = mfma_step(av[s], bv, acc[r]);
                }
            }
        }
        if constexpr (S::LIFT_ONE) {
#pragma unroll
            for (int j = 0; j < LPRE; ++j) {
                const int k = tid + j * kBlock;
                lpre[j] = k < S::LIFT_CHUNK * NC * NB ? gLift[k] : T(0);
            }
        }
        __syncthreads();  // every wave has read the inputs
        if (wave < NTM) {
#pragma unroll
            for (int ri = 0; ri < 4; ++ri) {
                const int row = 16 * wave + (std::is_same<T, double>::value ? 4 * ri + lane / 16 : 4 * (lane / 16) + ri);
                if (row < NB && ce < count) {
                    T* o = sG0 + ce * 4 * NB;
                    if (cv == 0) o[row] = acc[0][ri] + acc[1][ri] + acc[2][ri];
                    else {
#pragma unroll
                        for (int r = 0; r < 3; ++r) o[(1 + r) * NB + row] = acc[r][ri];
                    }
                }
            }
        }
        __syncthreads();
        if (row_ok) {
            T Jt[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) Jt[q] = T(p.jinv[q * E + e]);
            kv = Acc<T>(sg[i]);
            const T ci = T(1.0 / cinv);
            const T bb[3] = {sg[NB + i], sg[2 * NB + i], sg[3 * NB + i]};
#pragma unroll
            for (int d = 0; d < 3; ++d)
                ks[d] = Acc<T>((Jt[d] * bb[0] + Jt[3 + d] * bb[1] + Jt[6 + d] * bb[2]) * ci);
        }
    } else {
        T a[3] = {T(0), T(0), T(0)}, b[3] = {T(0), T(0), T(0)};
        if (row_ok) {
#pragma unroll 4
            for (int j = 0; j < NB; ++j) {
                const T v = sg[j];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const T d = sDr[(r * NB + j) * NB + i];
                    a[r] += d * sg[(1 + r) * NB + j];
                    b[r] += d * v;
                }
            }
        }
        if (row_ok) {
            T Jt[9];  // J^{-1} again (not kept in registers across the barrier)
#pragma unroll
            for (int q = 0; q < 9; ++q) Jt[q] = T(p.jinv[q * E + e]);
            kv = Acc<T>(a[0] + a[1] + a[2]);
            const T ci = T(1.0 / cinv);
#pragma unroll
            for (int d = 0; d < 3; ++d)
                ks[d] = Acc<T>((Jt[d] * b[0] + Jt[3 + d] * b[1] + Jt[6 + d] * b[2]) * ci);
        }
    }

    // ---- phase 3: k += sigma_f sum_m (M_ref^{-1} int_f phi chi_m)_i F_m
    // (order 4: the lift table passes through the work area, dead once the
    // volume terms are done, LIFT_CHUNK faces at a time)
    T* sL1 = reinterpret_cast<T*>(lds + S::W);
    constexpr int LC = S::LIFT_CHUNK;
    if constexpr (S::LIFT_ONE) {
        __syncthreads();  // the volume data give way to the first faces' lift tables
#pragma unroll
        for (int j = 0; j < LPRE; ++j) {
            const int k = tid + j * kBlock;
            if (k < LC * NC * NB) sL1[k] = lpre[j];
        }
        __syncthreads();
    }
    const T ci = T(1.0 / cinv);
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
        const T* lift;
        if constexpr (S::LIFT_LDS) {
            lift = sLift + f * NC * NB + i;
        } else if constexpr (!S::LIFT_ONE) {
            lift = gLift + f * NC * NB + i;
        } else {
            if (f > 0 && f % LC == 0) {
                __syncthreads();  // the block is done with the previous faces
                for (int k = tid; k < LC * NC * NB; k += kBlock) sL1[k] = gLift[f * NC * NB + k];
                __syncthreads();
            }
            lift = sL1 + (f % LC) * NC * NB + i;
        }
        if (row_ok) {
            const FluxPair<T>* fl = sf + f * NC;
            T lv = T(0), ls = T(0);
#pragma unroll 5
            for (int m = 0; m < NC; ++m) {
                const T l = lift[m * NB];
                const FluxPair<T> x = fl[m];
                lv += l * x.v;
                ls += l * x.s;
            }
            const T sig = T(fm[f * 4 + 3]);
            kv.add(sig * lv);
            const T c = sig * ls * ci;
            ks[0].add(c * T(fm[f * 4]));
            ks[1].add(c * T(fm[f * 4 + 1]));
            ks[2].add(c * T(fm[f * 4 + 2]));
        }
    }
    if (row_ok) {
        const double vals[4] = {double(kv.value()), double(ks[0].value()), double(ks[1].value()),
                                double(ks[2].value())};
        double kept[4];  // the values as stored
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int64_t o = base + f * NB + i;
            if (p.io.out_f32) {
                const float v = float(vals[f]);
                static_cast<float*>(p.io.out)[o] = v;
                kept[f] = double(v);
            } else {
                static_cast<double*>(p.io.out)[o] = vals[f];
                kept[f] = vals[f];
            }
        }
        if (p.io.upd.u != nullptr) {  // the RK update fused into a step's last stage (StageIO::upd)
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const double own[1] = {kept[f]};
                rk_update_run<1>(p.io.upd, base + f * NB + i, own);
            }
        }
    }
}

template <int TB>
constexpr int simplex_lds_doubles() {
    constexpr int t[8] = {SimplexShape<1, simplex_tpe(1), TB>::TOTAL,     SimplexShape<2, simplex_tpe(2), TB>::TOTAL,
                          SimplexShape<3, simplex_tpe(3), TB>::TOTAL,     SimplexShape<4, simplex_tpe(4), TB>::TOTAL,
                          SimplexShape<1, simplex_tpe_alt(1), TB>::TOTAL, SimplexShape<2, simplex_tpe_alt(2), TB>::TOTAL,
                          SimplexShape<3, simplex_tpe_alt(3), TB>::TOTAL, SimplexShape<4, simplex_tpe_alt(4), TB>::TOTAL};
    int m = 0;
    for (int k = 0; k < 8; ++k) m = t[k] > m ? t[k] : m;
    return m;
}

// All orders in one launch. Blocks follow a schedule built on the host: the
// mesh is cut into slabs of consecutive elements and, slab after slab, each
// order's elements of the slab fill blocks of that order. With the XCD
// renumbering an XCD works through neighbouring slabs, so the face columns an
// element gathers from neighbours of every order are read while they are
// still in that XCD's L2. The LDS carve-up of the largest order leaves room
// for five blocks per CU; the register budget is set to match.
template <typename T, int V>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void simplex_acoustic_rhs(const SimplexParams p) {
    __shared__ __align__(16) double lds[simplex_lds_doubles<int(sizeof(T))>()];  // regions start at even doubles
    const int4 bl = p.blocks[xcd_block(blockIdx.x, gridDim.x)];
    switch (bl.x) {  // the same order for the whole block; bit p-1 of V: the order's threads-per-element variant
        case 1: simplex_block<1, simplex_tpe_of(1, V & 1), T>(p, bl.y, bl.z, lds); break;
        case 2: simplex_block<2, simplex_tpe_of(2, (V >> 1) & 1), T>(p, bl.y, bl.z, lds); break;
        case 3: simplex_block<3, simplex_tpe_of(3, (V >> 2) & 1), T>(p, bl.y, bl.z, lds); break;
        default: simplex_block<4, simplex_tpe_of(4, (V >> 3) & 1), T>(p, bl.y, bl.z, lds); break;
    }
}

// Per-element |det J| sum_f dx_f^T M_ref dx_f (plain L2 of all fields) and
// max |dx|, one partial per block.
struct SimplexDiffParams {
    int64_t n_elem;
    int64_t group_start[5];
    int64_t group_dof[5];
    int32_t group_block[5];
    const void* x;
    const void* y;
    int32_t x_f32, y_f32;
    const double* absdet;
    const double* tab;
    const int32_t* tix;
    double* partial;
};

template <int P>
__device__ void simplex_diff_block(const SimplexDiffParams& p, int blk, double* sd, double* red) {
    using S = SimplexShape<P>;
    constexpr int NB = S::NB;
    const int tid = threadIdx.x;
    const int le = tid / S::TPE;
    const int i = tid % S::TPE;
    const int64_t e = p.group_start[P - 1] + int64_t(blk) * S::EPB + le;
    const bool row_ok = (le < S::EPB) && (e < p.group_start[P]) && (i < NB);
    const int64_t base = p.group_dof[P - 1] + (e - p.group_start[P - 1]) * 4 * NB;
    double* d = sd + le * 4 * NB;
    double v = 0, m = 0;
    if (row_ok)
        for (int f = 0; f < 4; ++f) {
            const int64_t k = base + f * NB + i;
            const double xv = p.x_f32 ? double(static_cast<const float*>(p.x)[k]) : static_cast<const double*>(p.x)[k];
            const double yv = p.y_f32 ? double(static_cast<const float*>(p.y)[k]) : static_cast<const double*>(p.y)[k];
            d[f * NB + i] = xv - yv;
            m = fmax(m, fabs(xv - yv));
        }
    __syncthreads();
    if (row_ok) {
        const double* M = p.tab + p.tix[(P - 1) * kTixStride + kTixM] + i * NB;
        for (int f = 0; f < 4; ++f) {
            double s = 0;
            for (int j = 0; j < NB; ++j) s += M[j] * d[f * NB + j];
            v += s * d[f * NB + i];
        }
        v *= p.absdet[e];
    }
    red[tid] = v;
    red[kBlock + tid] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[tid] += red[tid + w];
            red[kBlock + tid] = fmax(red[kBlock + tid], red[kBlock + tid + w]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void simplex_diff_norm(const SimplexDiffParams p) {
    constexpr int n1 = SimplexShape<1>::EPB * 4 * SimplexShape<1>::NB, n2 = SimplexShape<2>::EPB * 4 * SimplexShape<2>::NB;
    constexpr int n3 = SimplexShape<3>::EPB * 4 * SimplexShape<3>::NB, n4 = SimplexShape<4>::EPB * 4 * SimplexShape<4>::NB;
    constexpr int n12 = n1 > n2 ? n1 : n2, n34 = n3 > n4 ? n3 : n4;
    __shared__ double sd[n12 > n34 ? n12 : n34];  // [EPB][4][NB] of the largest order group
    __shared__ double red[2 * kBlock];
    const int b = blockIdx.x;
    if (b < p.group_block[1]) simplex_diff_block<1>(p, b - p.group_block[0], sd, red);
    else if (b < p.group_block[2]) simplex_diff_block<2>(p, b - p.group_block[1], sd, red);
    else if (b < p.group_block[3]) simplex_diff_block<3>(p, b - p.group_block[2], sd, red);
    else simplex_diff_block<4>(p, b - p.group_block[3], sd, red);
    if (threadIdx.x == 0) {
        p.partial[2 * b] = red[0];
        p.partial[2 * b + 1] = red[kBlock];
    }
}

// The block-diagonal mass matrix of the acoustic system
// (AcousticWaveLinear_Impl.h:102-129: |det J| M_ref on v, c_inv |det J| M_ref
// on each s_d) applied to a vector or solved against it, in place. One
// wavefront per element; the element's four fields are staged in LDS.
struct SimplexMassParams {
    const int64_t* elem_off;  // [n_elem + 1]
    const int32_t* order;     // [n_elem]
    const double* tab;        // per order: M_ref, then M_ref^{-1}
    int32_t tab_off[2][4];    // [apply / solve][order - 1]
    int32_t inverse;
    double* v;
    const double* absdet;     // [n_elem]
    const double* cinv;       // [n_elem]
};

__global__ __launch_bounds__(64) void simplex_mass(const SimplexMassParams p) {
    __shared__ double s[4 * SimplexShape<4>::NB];
    const int64_t e = blockIdx.x;
    const int P = p.order[e];
    const int nb = (P + 1) * (P + 2) * (P + 3) / 6;
    const int n = 4 * nb;
    const int64_t off = p.elem_off[e];
    for (int i = threadIdx.x; i < n; i += 64) s[i] = p.v[off + i];
    __syncthreads();
    const double* A = p.tab + p.tab_off[p.inverse][P - 1];
    const double w0 = p.absdet[e];
    const double ws = w0 * p.cinv[e];
    for (int r = threadIdx.x; r < n; r += 64) {
        const int f = r / nb, i = r - f * nb;
        double acc = 0;
        for (int j = 0; j < nb; ++j) acc += A[i * nb + j] * s[f * nb + j];
        const double w = f > 0 ? ws : w0;
        p.v[off + r] = p.inverse ? acc / w : acc * w;
    }
}

// Quadrature tables of the on-device projection and error norms, per order
// (index p - 1), as offsets into one table array: the element rule of order
// 2p+1 (points; the projection operator R = M_ref^{-1} [w_q phi_i(x_q)], [nb][nq];
// the basis at its points, [nq][nb], for the maximum error) and the rule of
// order 2p+2 (points, weights and the basis at its points, for the L2 error),
// all hpGEM's tetrahedral rules.
struct SimplexQuadTab {
    int32_t nb[4], nq_p[4], nq_e[4];
    int32_t pts_p[4], R[4], phi_p[4], pts_e[4], w_e[4], phi_e[4];
};

struct SimplexFnParams {
    const double* tab;
    SimplexQuadTab t;
    const double* xj;         // [E][12]: vertex 0, then J row-major (device order)
    const int64_t* elem_off;  // [E + 1]
    const int32_t* order;     // [E]
    int32_t n_fields, field;  // field < 0: every field
    int32_t n_modes;
    const double* modes;      // [n_modes][4]: amplitude, k
    int32_t gaussian;         // 1: f = g[4] exp(-|x - g[0..2]|^2 / g[3]^2) instead of the modes
    double g[5];
    double shift[3];          // error: f(x - shift)
    void* u;                  // projection: output; error: input
    int32_t u_f32;
    const double* absdet;
    double* partial;          // error: [E][2] (squared L2, max)
};

__device__ __forceinline__ void tet_point(const double* xj, const double* xi, double (&x)[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) x[r] = xj[r] + xj[3 + r * 3] * xi[0] + xj[3 + r * 3 + 1] * xi[1] + xj[3 + r * 3 + 2] * xi[2];
}

__device__ __forceinline__ double tet_fn(const SimplexFnParams& p, const double (&x)[3]) {
    if (p.gaussian) {
        const double r2 = (x[0] - p.g[0]) * (x[0] - p.g[0]) + (x[1] - p.g[1]) * (x[1] - p.g[1]) +
                          (x[2] - p.g[2]) * (x[2] - p.g[2]);
        return p.g[4] * exp(-r2 / (p.g[3] * p.g[3]));
    }
    double f = 0;
    for (int m = 0; m < p.n_modes; ++m) {
        const double* md = p.modes + 4 * m;
        f += md[0] * sin(2 * M_PI * (md[1] * (x[0] - p.shift[0]) + md[2] * (x[1] - p.shift[1]) +
                                     md[3] * (x[2] - p.shift[2])));
    }
    return f;
}

// L2 projection (HpgemAPISimplified::setInitialSolution with hpGEM's element
// rule of order 2p+1): one wavefront per element, lanes over points, then
// over basis functions.
__global__ __launch_bounds__(64) void simplex_project(const SimplexFnParams p) {
    __shared__ double fq[64];
    const int64_t e = blockIdx.x;
    const int P = p.order[e], nb = p.t.nb[P - 1], nq = p.t.nq_p[P - 1];
    const double* xj = p.xj + 12 * e;
    const int l = threadIdx.x;
    if (l < nq) {
        double x[3];
        tet_point(xj, p.tab + p.t.pts_p[P - 1] + 3 * l, x);
        fq[l] = tet_fn(p, x);
    }
    __syncthreads();
    if (l < nb) {
        const double* R = p.tab + p.t.R[P - 1] + l * nq;
        double s = 0;
        for (int q = 0; q < nq; ++q) s += R[q] * fq[q];
        const int64_t o = p.elem_off[e];
        for (int f = 0; f < p.n_fields; ++f) {
            if (p.field >= 0 && f != p.field) continue;
            if (p.u_f32) static_cast<float*>(p.u)[o + f * nb + l] = float(s);
            else static_cast<double*>(p.u)[o + f * nb + l] = s;
        }
    }
}

// computeTotalError (rule 2p+2) and computeMaxError (the points of the rule
// 2p+1) of one field or all, per element: partial[2e] = |det J| sum_q w_q
// err^2, partial[2e+1] = max |err|.
__global__ __launch_bounds__(64) void simplex_error(const SimplexFnParams p) {
    __shared__ double cf[4 * SimplexShape<4>::NB];
    const int64_t e = blockIdx.x;
    const int P = p.order[e], nb = p.t.nb[P - 1];
    const double* xj = p.xj + 12 * e;
    const int64_t o = p.elem_off[e];
    const int l = threadIdx.x;
    for (int k = l; k < p.n_fields * nb; k += 64)
        cf[k] = p.u_f32 ? double(static_cast<const float*>(p.u)[o + k]) : static_cast<const double*>(p.u)[o + k];
    __syncthreads();
    double s2 = 0, mx = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const int nq = pass == 0 ? p.t.nq_e[P - 1] : p.t.nq_p[P - 1];
        const double* pts = p.tab + (pass == 0 ? p.t.pts_e[P - 1] : p.t.pts_p[P - 1]);
        const double* phi = p.tab + (pass == 0 ? p.t.phi_e[P - 1] : p.t.phi_p[P - 1]);
        for (int q = l; q < nq; q += 64) {
            double x[3];
            tet_point(xj, pts + 3 * q, x);
            const double ex = tet_fn(p, x);
            for (int f = 0; f < p.n_fields; ++f) {
                if (p.field >= 0 && f != p.field) continue;
                double val = 0;
                for (int i = 0; i < nb; ++i) val += cf[f * nb + i] * phi[q * nb + i];
                const double err = ex - val;
                if (pass == 0) s2 += p.tab[p.t.w_e[P - 1] + q] * err * err;
                else mx = fmax(mx, fabs(err));
            }
        }
    }
    for (int w = 32; w > 0; w /= 2) {
        s2 += __shfl_xor(s2, w, 64);
        mx = fmax(mx, __shfl_xor(mx, w, 64));
    }
    if (l == 0) {
        p.partial[2 * e] = s2 * p.absdet[e];
        p.partial[2 * e + 1] = mx;
    }
}

// ---------------------------------------------------------------- host side
namespace {

int epb_of(int p) { return kBlock / simplex_tpe(p); }

void group_blocks(const SimplexMesh& sm, int32_t* gb) {
    gb[0] = 0;
    for (int g = 0; g < 4; ++g) {
        const int64_t n = sm.group_start[g + 1] - sm.group_start[g];
        gb[g + 1] = gb[g] + int32_t((n + epb_of(g + 1) - 1) / epb_of(g + 1));
    }
}

}  // namespace

namespace {

// Processing order of the tetrahedra (the order of the RHS kernel's block
// schedule, and of the device layout inside each order group): the elements'
// centroids are cut into 4 x 4 columns across the two shorter extents of the
// mesh's bounding box and run through column by column, along the longest
// extent inside a column. A face neighbour is then either in the same column,
// at most one column layer (a few hundred elements) away in processing order,
// so its face columns are still in the XCD's L2 when they are gathered, or
// across a column side; in mesh order (x fastest) every neighbour along the
// slowest axis was a whole mesh layer back, beyond the L2.
std::vector<int64_t> locality_rank(const hgdg_mesh_desc* m) {
    const int64_t ne = m->n_elements;
    std::vector<std::array<double, 3>> ctr(static_cast<size_t>(ne));
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t e = 0; e < ne; ++e)
        for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int v = 0; v < 4; ++v) s += m->coords[(size_t(e) * 4 + v) * 3 + r];
            ctr[size_t(e)][r] = s / 4;
            lo[r] = std::min(lo[r], s / 4);
            hi[r] = std::max(hi[r], s / 4);
        }
    int ax[3] = {0, 1, 2};  // ax[2]: the longest extent (the last of equal ones)
    std::stable_sort(ax, ax + 3, [&](int a, int b) { return hi[a] - lo[a] < hi[b] - lo[b]; });
    constexpr int kCols = 4;
    auto cell = [&](int64_t e, int r) {
        const double ext = hi[r] - lo[r];
        if (!(ext > 0)) return 0;
        return std::min(kCols - 1, int((ctr[size_t(e)][r] - lo[r]) / ext * kCols));
    };
    std::vector<int64_t> idx(static_cast<size_t>(ne));
    for (int64_t e = 0; e < ne; ++e) idx[size_t(e)] = e;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        const int ca = cell(a, ax[1]) * kCols + cell(a, ax[0]), cb = cell(b, ax[1]) * kCols + cell(b, ax[0]);
        if (ca != cb) return ca < cb;
        for (int k = 2; k >= 0; --k)
            if (ctr[size_t(a)][ax[k]] != ctr[size_t(b)][ax[k]]) return ctr[size_t(a)][ax[k]] < ctr[size_t(b)][ax[k]];
        return a < b;
    });
    std::vector<int64_t> rank(static_cast<size_t>(ne));
    for (int64_t k = 0; k < ne; ++k) rank[size_t(idx[size_t(k)])] = k;
    return rank;
}

}  // namespace

void simplex_mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m) {
    SimplexMesh& sm = c->sm;
    const int64_t ne = m->n_elements;
    std::vector<int32_t> order(ne);
    for (int64_t e = 0; e < ne; ++e) {
        order[e] = m->order ? m->order[e] : m->uniform_order;
        require(order[e] >= 1 && order[e] <= kMaxSimplexOrder, HGDG_ERR_UNSUPPORTED,
                "tetrahedral orders must be in 1..4");
    }
    // group by order; inside a group, elements follow the locality rank
    int64_t counts[4] = {0, 0, 0, 0};
    for (int64_t e = 0; e < ne; ++e) ++counts[order[e] - 1];
    sm.group_start[0] = 0;
    for (int g = 0; g < 4; ++g) sm.group_start[g + 1] = sm.group_start[g] + counts[g];
    sm.dev_of.assign(ne, 0);
    sm.mesh_of.assign(ne, 0);
    sm.order.assign(ne, 0);
    sm.rank = locality_rank(m);
    std::vector<int64_t> by_rank(ne);
    for (int64_t e = 0; e < ne; ++e) by_rank[size_t(sm.rank[e])] = e;
    int64_t fill[4] = {sm.group_start[0], sm.group_start[1], sm.group_start[2], sm.group_start[3]};
    for (int64_t e : by_rank) {
        const int64_t d = fill[order[e] - 1]++;
        sm.dev_of[e] = d;
        sm.mesh_of[d] = e;
        sm.order[d] = order[e];
    }
    constexpr int nf = 4;  // fields of the acoustic system
    sm.group_dof[0] = 0;
    for (int g = 0; g < 4; ++g) sm.group_dof[g + 1] = sm.group_dof[g] + counts[g] * nf * tet_nb(g + 1);
    sm.dof_off.assign(ne, 0);
    for (int64_t d = 0; d < ne; ++d) {
        const int g = sm.order[d] - 1;
        sm.dof_off[d] = sm.group_dof[g] + (d - sm.group_start[g]) * nf * tet_nb(g + 1);
    }
    c->n_fields = nf;
    c->dof = sm.group_dof[4];
    sm.group_tr[0] = 0;
    for (int g = 0; g < 4; ++g) sm.group_tr[g + 1] = sm.group_tr[g] + counts[g] * 4 * simplex_ncol(g + 1);
    sm.d_trace.alloc(size_t(2 * sm.group_tr[4]));
    std::vector<int64_t> elem_off(sm.dof_off);  // lives until the synchronize at the end
    elem_off.push_back(c->dof);
    sm.d_elem_off.upload(elem_off, c->stream);

    // geometry (device order): J column k = X_{k+1} - X_0
    sm.jinv.assign(size_t(ne) * 9, 0.0);
    sm.absdet.assign(ne, 0.0);
    std::vector<double> jinv_soa(9 * size_t(ne)), detJ(ne);
    std::vector<std::array<double, 9>> Jm(ne);
    sm.x0.assign(size_t(ne) * 3, 0.0);
    sm.jac.assign(size_t(ne) * 9, 0.0);
    for (int64_t d = 0; d < ne; ++d) {
        const double* X = m->coords + size_t(sm.mesh_of[d]) * 12;
        for (int r = 0; r < 3; ++r) sm.x0[size_t(d) * 3 + r] = X[r];
        double J[3][3];
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) J[r][k] = X[(k + 1) * 3 + r] - X[r];
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        require(det != 0.0 && std::isfinite(det), HGDG_ERR_ARG, "degenerate tetrahedron");
        double inv[3][3];
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) {
                const int r1 = (k + 1) % 3, r2 = (k + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
                inv[r][k] = (J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1]) / det;
            }
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) {
                jinv_soa[(r * 3 + k) * size_t(ne) + d] = inv[r][k];
                Jm[d][r * 3 + k] = J[r][k];
                sm.jac[size_t(d) * 9 + r * 3 + k] = J[r][k];
            }
        detJ[d] = det;
        sm.absdet[d] = std::fabs(det);
    }

    // faces: sorted node triples; the lower mesh index is hpGEM's left element
    sm.nbr_dev.assign(4 * size_t(ne), -1);
    sm.info.assign(4 * size_t(ne), 0);
    sm.sigma.assign(4 * size_t(ne), 0.0);
    sm.normal.assign(12 * size_t(ne), 0.0);
    struct Open { int64_t e; int f; };
    std::unordered_map<std::string, Open> open;
    open.reserve(size_t(ne) * 2);
    int64_t n_internal = 0, n_total = 0;
    static const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t* nid = m->node_ids + size_t(e) * 4;
        const int64_t d = sm.dev_of[e];
        for (int f = 0; f < 4; ++f) {
            // outward normal: sign(det J) (J c1) x (J c2), c = face edges in
            // reference coordinates (FaceGeometry.h:412-459)
            const int* fn = kTetFaceNodes[f];
            double a[3], b[3];
            for (int r = 0; r < 3; ++r) {
                a[r] = 0;
                b[r] = 0;
                for (int k = 0; k < 3; ++k) {
                    a[r] += Jm[d][r * 3 + k] * (V[fn[1]][k] - V[fn[0]][k]);
                    b[r] += Jm[d][r * 3 + k] * (V[fn[2]][k] - V[fn[0]][k]);
                }
            }
            double nrm[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            const double sg = detJ[d] > 0 ? 1.0 : -1.0;
            const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            for (int r = 0; r < 3; ++r) sm.normal[(size_t(r) * 4 + f) * ne + d] = sg * nrm[r] / len;
            sm.sigma[size_t(f) * ne + d] = len / sm.absdet[d];
            sm.info[size_t(f) * ne + d] = info_pack(0, 0, 0, 0);

            std::array<int64_t, 3> key{nid[fn[0]], nid[fn[1]], nid[fn[2]]};
            std::sort(key.begin(), key.end());
            std::string k(reinterpret_cast<const char*>(key.data()), sizeof(key));
            auto it = open.find(k);
            if (it == open.end()) {
                open.emplace(std::move(k), Open{e, f});
                continue;
            }
            const int64_t L = it->second.e;
            const int fL = it->second.f;
            require(L != e, HGDG_ERR_ARG, "tetrahedron meets itself across a face");
            const int64_t* lid = m->node_ids + size_t(L) * 4;
            long long ln[3], rn[3];
            for (int q = 0; q < 3; ++q) {
                ln[q] = lid[kTetFaceNodes[fL][q]];
                rn[q] = nid[fn[q]];
            }
            const int o = tri_orientation_index(ln, rn);
            const int64_t dL = sm.dev_of[L];
            require(sm.nbr_dev[size_t(fL) * ne + dL] < 0, HGDG_ERR_ARG, "a face is bounded by more than two elements");
            sm.nbr_dev[size_t(fL) * ne + dL] = int32_t(d);
            sm.nbr_dev[size_t(f) * ne + d] = int32_t(dL);
            sm.info[size_t(fL) * ne + dL] = info_pack(f, 0, o, sm.order[d]);
            sm.info[size_t(f) * ne + d] = info_pack(fL, o, 0, sm.order[dL]);
            open.erase(it);
            ++n_internal;
            ++n_total;
        }
    }
    n_total += int64_t(open.size());
    c->n_faces = n_total;
    c->n_internal = n_internal;

    // reference tables of every order (see the kernel): DrT, M_ref, the face
    // columns, the chi-coefficients of the face columns' traces per local face
    // and orientation, and the lift in the element's own face coordinates
    int perm[6][kFaceQ];  // face point q after orientation o of the reference triangle
    for (int o = 0; o < 6; ++o)
        for (int qq = 0; qq < kFaceQ; ++qq) {
            const long double tp[2] = {kFaceRulePts[qq][0], kFaceRulePts[qq][1]};
            long double m[2];
            tri_orientation_map(o, tp, m);
            int hit = -1;
            for (int k = 0; k < kFaceQ; ++k)
                if (fabsl(m[0] - kFaceRulePts[k][0]) < 1e-12L && fabsl(m[1] - kFaceRulePts[k][1]) < 1e-12L) hit = k;
            require(hit >= 0, HGDG_ERR_ARG, "face rule is not symmetric under the orientation maps");
            perm[o][qq] = hit;
        }
    // Both sides of a face reach a physical point through their orientation:
    // own coordinates o_own(t), neighbour's o_nbr(t). Seen from the element,
    // the neighbour's face coordinates are o_nbr(o_own^{-1}(s)), itself one
    // of the six orientation maps (they form a group), identified here by
    // its action on the (symmetric) face rule points.
    std::vector<int32_t> tix(size_t(kTixRot) + 1, 0);
    int ident = -1;
    for (int o = 0; o < 6 && ident < 0; ++o) {
        bool id = true;
        for (int qq = 0; qq < kFaceQ; ++qq) id = id && perm[o][qq] == qq;
        if (id) ident = o;
    }
    require(ident >= 0, HGDG_ERR_ARG, "no identity among the face orientation maps");
    tix[kTixIdent] = ident;
    for (int oo = 0; oo < 6; ++oo) {
        int inv[kFaceQ];
        for (int qq = 0; qq < kFaceQ; ++qq) inv[perm[oo][qq]] = qq;
        for (int on = 0; on < 6; ++on) {
            int rel = -1;
            for (int o = 0; o < 6 && rel < 0; ++o) {
                bool same = true;
                for (int r = 0; r < kFaceQ; ++r) same = same && perm[o][r] == perm[on][inv[r]];
                if (same) rel = o;
            }
            require(rel >= 0, HGDG_ERR_ARG, "face orientation maps do not compose");
            tix[size_t(kTixRel) + oo * 6 + on] = rel;
        }
    }
    for (size_t k = 0; k < sm.info.size(); ++k)
        if (sm.nbr_dev[k] >= 0)
            sm.info[k] |= tix[size_t(kTixRel) + info_oown(sm.info[k]) * 6 + info_onbr(sm.info[k])] << 11;
    sm.ops.clear();
    for (int q = 1; q <= kMaxSimplexOrder; ++q) sm.ops.push_back(simplex_ops(q));
    std::vector<double> tab;
    std::vector<int32_t> cols;
    for (int q = 1; q <= kMaxSimplexOrder; ++q) {
        const SimplexOps& ops = sm.ops[q - 1];
        const int nb = ops.nb, nc = simplex_ncol(q);
        int32_t* t = tix.data() + (q - 1) * kTixStride;
        t[kTixDrT] = int32_t(tab.size());
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nb; ++i) tab.push_back(ops.Dr[(size_t(r) * nb + i) * nb + j]);
        t[kTixM] = int32_t(tab.size());
        tab.insert(tab.end(), ops.Mref.begin(), ops.Mref.end());
        for (int f = 0; f < 4; ++f) {
            require(int(ops.cols[f].size()) == nc, HGDG_ERR_ARG, "unexpected face column count");
            t[kTixCols + f] = int32_t(cols.size());  // relative, fixed below
            cols.insert(cols.end(), ops.cols[f].begin(), ops.cols[f].end());
        }
        t[kTixDrP] = int32_t(tab.size());  // [r][k < 4 ceil(nb/4)][i < 16 ceil(nb/16)], zeros outside nb x nb
        if (q >= 3) {
            const int kp = 4 * ((nb + 3) / 4), ip = 16 * ((nb + 15) / 16);
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < kp; ++k)
                    for (int i = 0; i < ip; ++i)
                        tab.push_back(k < nb && i < nb ? ops.Dr[(size_t(r) * nb + i) * nb + k] : 0.0);
        }
        t[kTixProjId] = int32_t(tab.size());  // [f][c < nc][m < nc]
        for (int f = 0; f < 4; ++f)
            for (int c = 0; c < nc; ++c)
                for (int m = 0; m < nc; ++m) tab.push_back(ops.proj[f][ident][size_t(m) * nc + c]);
        t[kTixProj] = int32_t(tab.size());  // [f][o][c < nc][m < kFaceBasis]
        for (int f = 0; f < 4; ++f)
            for (int o = 0; o < 6; ++o) {
                require(ops.proj[f][o].size() == size_t(kFaceBasis) * nc, HGDG_ERR_ARG, "face trace table size");
                for (int c = 0; c < nc; ++c)
                    for (int m = 0; m < kFaceBasis; ++m) tab.push_back(ops.proj[f][o][size_t(m) * nc + c]);
            }
        t[kTixLift] = int32_t(tab.size());  // [f][m < nc][i]: only chi_m of degree <= q meet the element's traces
        for (int f = 0; f < 4; ++f)
            for (int m = 0; m < nc; ++m)
                for (int i = 0; i < nb; ++i) tab.push_back(ops.lift[f][ident][size_t(i) * kFaceBasis + m]);
    }
    {
        const std::vector<double> rot = face_orientation_blocks();
        require(rot.size() == 6 * size_t(kFaceRotBlock) && kFaceRotBlock == 55, HGDG_ERR_ARG, "orientation block table size");
        tix[kTixRot] = int32_t(tab.size());
        tab.insert(tab.end(), rot.begin(), rot.end());
    }
    const int32_t col_base = int32_t(tix.size());
    for (int q = 0; q < kMaxSimplexOrder; ++q)
        for (int f = 0; f < 4; ++f) tix[q * kTixStride + kTixCols + f] += col_base;
    tix.insert(tix.end(), cols.begin(), cols.end());
    require(tix.size() <= size_t(kTixMax), HGDG_ERR_ARG, "table index larger than its LDS copy");

    for (int g = 0; g < 4; ++g) {
        sm.tpe_variant[g] = sm.tpe_fixed[g] == simplex_tpe_alt(g + 1) ? 1 : 0;
        sm.tpe_ms[g][0] = sm.tpe_ms[g][1] = 0;
    }
    simplex_schedule(c);

    sm.d_tables.upload(tab, c->stream);
    std::vector<double> mass;  // lives until the synchronize below
    for (int q = 1; q <= kMaxSimplexOrder; ++q) mass.insert(mass.end(), sm.ops[q - 1].Mref.begin(), sm.ops[q - 1].Mref.end());
    for (int q = 1; q <= kMaxSimplexOrder; ++q) mass.insert(mass.end(), sm.ops[q - 1].Minv.begin(), sm.ops[q - 1].Minv.end());
    sm.d_mass.upload(mass, c->stream);
    sm.d_order.upload(sm.order, c->stream);
    const std::vector<float> tab32(tab.begin(), tab.end());  // lives until the synchronize below
    sm.d_tables32.upload(tab32, c->stream);
    sm.d_tab_index.upload(tix, c->stream);
    sm.d_absdet.upload(sm.absdet, c->stream);
    std::vector<double> geo(jinv_soa);
    geo.resize(geo.size() + size_t(ne), 1.0);  // c_inv, set with the operator
    geo.insert(geo.end(), sm.sigma.begin(), sm.sigma.end());
    geo.insert(geo.end(), sm.normal.begin(), sm.normal.end());
    sm.d_geo.upload(geo, c->stream);
    std::vector<int32_t> face(sm.nbr_dev);
    face.insert(face.end(), sm.info.begin(), sm.info.end());
    sm.d_face.upload(face, c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));
}

void simplex_operator_set(hgdg_ctx* c, const hgdg_op_desc* op) {
    require(op->kind == HGDG_OP_ACOUSTIC_CENTRAL, HGDG_ERR_UNSUPPORTED,
            "tetrahedral meshes run the acoustic operator");
    SimplexMesh& sm = c->sm;
    const int64_t ne = c->n_elem;
    std::vector<double> cinv(ne);
    for (int64_t d = 0; d < ne; ++d) cinv[d] = op->c_inv[sm.mesh_of[d]];
    HGDG_HIP(hipMemcpyAsync(sm.d_geo.ptr + 9 * ne, cinv.data(), size_t(ne) * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
}

int simplex_threads(int order, int variant) { return simplex_tpe_of(order, variant); }

// Block schedule: the mesh is cut into slabs of consecutive elements and, slab
// after slab, each order's elements of the slab fill blocks of that order's
// size (256 / threads per element of the order's variant in use).
void simplex_schedule(hgdg_ctx* c) {
    SimplexMesh& sm = c->sm;
    const int64_t ne = c->n_elem;
    const int64_t slab = std::max<int64_t>(1024, (ne + 255) / 256);  // in locality ranks
    int epb[4];
    for (int g = 0; g < 4; ++g) epb[g] = kBlock / simplex_tpe_of(g + 1, sm.tpe_variant[g]);
    std::vector<int32_t> blocks;
    int64_t cursor[4] = {sm.group_start[0], sm.group_start[1], sm.group_start[2], sm.group_start[3]};
    for (int64_t lo = 0; lo < ne; lo += slab) {
        const int64_t hi = std::min(ne, lo + slab);
        for (int g = 0; g < 4; ++g) {
            // device elements of order g+1 with locality rank below hi (rank order inside a group)
            int64_t end = cursor[g];
            while (end < sm.group_start[g + 1] && sm.rank[size_t(sm.mesh_of[end])] < hi) ++end;
            for (int64_t b = cursor[g]; b < end; b += epb[g]) {
                blocks.push_back(g + 1);
                blocks.push_back(int32_t(b));
                blocks.push_back(int32_t(std::min<int64_t>(epb[g], end - b)));
                blocks.push_back(sm.tpe_variant[g]);
            }
            cursor[g] = end;
        }
    }
    sm.n_blocks = int64_t(blocks.size() / 4);
    sm.d_blocks.upload(blocks, c->stream);
    HGDG_HIP(hipStreamSynchronize(c->stream));  // `blocks` goes out of scope
}

namespace {

SimplexParams simplex_params(hgdg_ctx* c, const StageIO& io) {
    SimplexMesh& sm = c->sm;
    const int64_t ne = c->n_elem;
    SimplexParams p{};
    p.io = io;
    p.n_elem = ne;
    for (int g = 0; g < 5; ++g) {
        p.group_start[g] = sm.group_start[g];
        p.group_dof[g] = sm.group_dof[g];
    }
    group_blocks(sm, p.group_block);
    p.jinv = sm.d_geo.ptr;
    p.cinv = sm.d_geo.ptr + 9 * ne;
    p.sigma = sm.d_geo.ptr + 10 * ne;
    p.normal = sm.d_geo.ptr + 14 * ne;
    p.nbr = sm.d_face.ptr;
    p.info = sm.d_face.ptr + 4 * ne;
    p.tab = sm.d_tables.ptr;
    p.tab32 = sm.d_tables32.ptr;
    p.tix = sm.d_tab_index.ptr;
    p.n_tix = int32_t(sm.d_tab_index.n);
    p.blocks = reinterpret_cast<const int4*>(sm.d_blocks.ptr);
    p.trace = reinterpret_cast<double2*>(sm.d_trace.ptr);
    for (int g = 0; g < 5; ++g) p.group_tr[g] = sm.group_tr[g];
    return p;
}

using SimplexKernel = void (*)(SimplexParams);

template <typename T, int... V>
constexpr std::array<SimplexKernel, sizeof...(V)> simplex_kernels(std::integer_sequence<int, V...>) {
    return {&simplex_acoustic_rhs<T, V>...};
}

void launch_simplex(hgdg_ctx* c, const SimplexParams& p) {
    // one kernel per combination of the orders' threads-per-element variants
    static const auto k64 = simplex_kernels<double>(std::make_integer_sequence<int, 16>{});
    static const auto k32 = simplex_kernels<float>(std::make_integer_sequence<int, 16>{});
    const SimplexMesh& sm = c->sm;
    int v = 0;
    for (int g = 0; g < 4; ++g) v |= (sm.tpe_variant[g] & 1) << g;
    const int grid = int(sm.n_blocks);
    if (grid > 0) {
        // the face traces of every element, then the right-hand side
        hipLaunchKernelGGL(simplex_trace, dim3(unsigned(p.group_block[4])), dim3(kBlock), 0, c->stream, p);
        const SimplexKernel k = c->op.precision == HGDG_PREC_MIXED_FP32 ? k32[size_t(v)] : k64[size_t(v)];
        hipLaunchKernelGGL(k, dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
    }
    HGDG_HIP(hipGetLastError());
}

}  // namespace

void simplex_rhs(hgdg_ctx* c, const StageIO& io) { launch_simplex(c, simplex_params(c, io)); }

// Threads per element, order by order: both candidates of an order are timed
// on the whole mesh (the other orders at their current choice) as one RK
// stage (u + a k -> k' on zero vectors of the operator's storage types; the
// kernel has no data-dependent branches), and the faster is kept. Orders with
// a fixed choice or without elements are not timed.
void simplex_tune(hgdg_ctx* c) {
    SimplexMesh& sm = c->sm;
    for (int g = 0; g < 4; ++g) {
        sm.tpe_variant[g] = sm.tpe_fixed[g] == simplex_tpe_alt(g + 1) ? 1 : 0;
        sm.tpe_ms[g][0] = sm.tpe_ms[g][1] = 0;
    }
    bool any = false;
    for (int g = 0; g < 4; ++g) any = any || (sm.tpe_fixed[g] == 0 && sm.group_start[g + 1] > sm.group_start[g]);
    if (!any) {
        simplex_schedule(c);
        return;
    }
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
    EventPair ev;
    constexpr int kReps = 10;
    for (int g = 0; g < 4; ++g) {
        if (sm.tpe_fixed[g] != 0 || sm.group_start[g + 1] == sm.group_start[g]) continue;
        for (int pass = 0; pass < 2; ++pass)
            for (int v = 0; v < 2; ++v) {
                sm.tpe_variant[g] = v;
                simplex_schedule(c);
                const SimplexParams p = simplex_params(c, io);
                for (int r = 0; r < 2; ++r) launch_simplex(c, p);
                HGDG_HIP(hipEventRecord(ev.e[0], c->stream));
                for (int r = 0; r < kReps; ++r) launch_simplex(c, p);
                HGDG_HIP(hipEventRecord(ev.e[1], c->stream));
                HGDG_HIP(hipEventSynchronize(ev.e[1]));
                float ms = 0;
                HGDG_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
                const double t = double(ms) / kReps;
                if (pass == 0 || t < sm.tpe_ms[g][v]) sm.tpe_ms[g][v] = t;
            }
        sm.tpe_variant[g] = sm.tpe_ms[g][1] < sm.tpe_ms[g][0] ? 1 : 0;
    }
    simplex_schedule(c);
}

void simplex_state_diff(hgdg_ctx* c, const Vec& vx, const Vec& vy, double* l2, double* max_abs) {
    SimplexMesh& sm = c->sm;
    SimplexDiffParams p{};
    p.n_elem = c->n_elem;
    for (int g = 0; g < 5; ++g) {
        p.group_start[g] = sm.group_start[g];
        p.group_dof[g] = sm.group_dof[g];
    }
    group_blocks(sm, p.group_block);
    p.x = vx.ptr();
    p.y = vy.ptr();
    p.x_f32 = vx.is_f32;
    p.y_f32 = vy.is_f32;
    p.absdet = sm.d_absdet.ptr;
    p.tab = sm.d_tables.ptr;
    p.tix = sm.d_tab_index.ptr;
    const int grid = p.group_block[4];
    if (c->d_partial.n < size_t(2 * grid)) c->d_partial.alloc(size_t(2 * grid));
    p.partial = c->d_partial.ptr;
    hipLaunchKernelGGL(simplex_diff_norm, dim3(unsigned(grid)), dim3(kBlock), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
    std::vector<double> h(size_t(2 * grid));
    HGDG_HIP(hipMemcpyAsync(h.data(), c->d_partial.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
    double s = 0, mx = 0;
    for (int b = 0; b < grid; ++b) {
        s += h[2 * b];
        mx = std::max(mx, h[2 * b + 1]);
    }
    *l2 = std::sqrt(std::max(s, 0.0));
    *max_abs = mx;
}

void simplex_mass_apply(hgdg_ctx* c, Vec& v, bool inverse) {
    require(!v.is_f32, HGDG_ERR_UNSUPPORTED, "mass apply/solve needs a vector stored in FP64");
    require(c->n_fields == 4, HGDG_ERR_STATE, "the tetrahedral mass matrix belongs to the acoustic operator");
    SimplexMesh& sm = c->sm;
    SimplexMassParams p{};
    int32_t at = 0;
    for (int k = 0; k < 2; ++k)
        for (int q = 1; q <= kMaxSimplexOrder; ++q) {
            p.tab_off[k][q - 1] = at;
            at += tet_nb(q) * tet_nb(q);
        }
    require(sm.d_mass.n == size_t(at), HGDG_ERR_STATE, "set the operator before the mass matrix is used");
    p.elem_off = sm.d_elem_off.ptr;
    p.order = sm.d_order.ptr;
    p.tab = sm.d_mass.ptr;
    p.inverse = inverse ? 1 : 0;
    p.v = v.f64->ptr;
    p.absdet = sm.d_absdet.ptr;
    p.cinv = sm.d_geo.ptr + size_t(9) * c->n_elem;
    if (c->n_elem == 0) return;
    hipLaunchKernelGGL(simplex_mass, dim3(unsigned(c->n_elem)), dim3(64), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
}

namespace {

// hpGEM's rule for a requested order (AllGaussQuadratureRules::getRule).
void ref_tet_rule(int requested, std::vector<long double>& pts, std::vector<long double>& w) {
    require(requested >= 0 && requested <= kRefTetRuleMaxOrder, HGDG_ERR_UNSUPPORTED,
            "no tetrahedral rule of order " + std::to_string(requested));
    const RefTetRule& r = kRefTetRules[requested];
    pts.assign(r.pts, r.pts + 3 * r.n);
    w.assign(r.w, r.w + r.n);
}

// L2 projector of order p on the reference tetrahedron with hpGEM's element
// rule of order 2p+1 (as setInitialSolution integrates): R[i][q] = sum_j
// (M_ref^{-1})_{ij} w_q phi_j(xi_q), so that the coefficients of the
// projection of f on an affine element are sum_q R[i][q] f(x(xi_q)).
struct TetProjector {
    int nb = 0, nq = 0;
    std::vector<double> pts;  // [nq][3]
    std::vector<double> R;    // [nb][nq]
};

TetProjector tet_projector(int p) {
    TetProjector t;
    t.nb = tet_nb(p);
    std::vector<long double> pts, w;
    ref_tet_rule(2 * p + 1, pts, w);
    t.nq = int(w.size());
    const int nb = t.nb, nq = t.nq;
    std::vector<long double> phi(size_t(nq) * nb), grad(size_t(nb) * 3), M(size_t(nb) * nb, 0.0L);
    for (int q = 0; q < nq; ++q) {
        tet_basis(p, &pts[size_t(q) * 3], &phi[size_t(q) * nb], grad.data());
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) M[size_t(i) * nb + j] += w[q] * phi[size_t(q) * nb + i] * phi[size_t(q) * nb + j];
    }
    // Gauss-Jordan inverse of the SPD reference mass matrix (long double)
    std::vector<long double> inv(size_t(nb) * nb, 0.0L);
    for (int i = 0; i < nb; ++i) inv[size_t(i) * nb + i] = 1.0L;
    for (int c = 0; c < nb; ++c) {
        int piv = c;
        for (int r = c + 1; r < nb; ++r)
            if (fabsl(M[size_t(r) * nb + c]) > fabsl(M[size_t(piv) * nb + c])) piv = r;
        for (int k = 0; k < nb; ++k) {
            std::swap(M[size_t(c) * nb + k], M[size_t(piv) * nb + k]);
            std::swap(inv[size_t(c) * nb + k], inv[size_t(piv) * nb + k]);
        }
        const long double d = M[size_t(c) * nb + c];
        for (int k = 0; k < nb; ++k) {
            M[size_t(c) * nb + k] /= d;
            inv[size_t(c) * nb + k] /= d;
        }
        for (int r = 0; r < nb; ++r) {
            if (r == c) continue;
            const long double f = M[size_t(r) * nb + c];
            if (f == 0.0L) continue;
            for (int k = 0; k < nb; ++k) {
                M[size_t(r) * nb + k] -= f * M[size_t(c) * nb + k];
                inv[size_t(r) * nb + k] -= f * inv[size_t(c) * nb + k];
            }
        }
    }
    t.pts.assign(pts.begin(), pts.end());
    t.R.assign(size_t(nb) * nq, 0.0);
    for (int i = 0; i < nb; ++i)
        for (int q = 0; q < nq; ++q) {
            long double s = 0;
            for (int j = 0; j < nb; ++j) s += inv[size_t(i) * nb + j] * w[q] * phi[size_t(q) * nb + j];
            t.R[size_t(i) * nq + q] = double(s);
        }
    return t;
}

}  // namespace

// L2 projection of sine modes or a Gaussian pulse (see tensor_project_device)
// into one field or all fields of a tetrahedral vector, computed on the host
// element by element and uploaded; the other fields keep their values.
namespace {

// The quadrature tables of the on-device projection and error norms, and the
// elements' affine maps, uploaded on first use.
SimplexQuadTab simplex_quad_tables(hgdg_ctx* c) {
    SimplexMesh& sm = c->sm;
    SimplexQuadTab t{};
    if (sm.qtab_index.empty()) {
        std::vector<double> tab;
        std::vector<int32_t> ix(40, 0);
        for (int p = 1; p <= kMaxSimplexOrder; ++p) {
            const int g = p - 1, nb = tet_nb(p);
            const TetProjector P = tet_projector(p);
            ix[g] = nb;
            ix[4 + g] = P.nq;
            ix[12 + g] = int32_t(tab.size());
            tab.insert(tab.end(), P.pts.begin(), P.pts.end());
            ix[16 + g] = int32_t(tab.size());
            tab.insert(tab.end(), P.R.begin(), P.R.end());
            std::vector<long double> phi(static_cast<size_t>(nb)), grad(static_cast<size_t>(nb) * 3);
            ix[20 + g] = int32_t(tab.size());
            for (int q = 0; q < P.nq; ++q) {
                const long double xi[3] = {P.pts[3 * q], P.pts[3 * q + 1], P.pts[3 * q + 2]};
                tet_basis(p, xi, phi.data(), grad.data());
                for (int i = 0; i < nb; ++i) tab.push_back(double(phi[i]));
            }
            std::vector<long double> pts, w;
            ref_tet_rule(2 * p + 2, pts, w);
            const int nq = int(w.size());
            ix[8 + g] = nq;
            ix[24 + g] = int32_t(tab.size());
            tab.insert(tab.end(), pts.begin(), pts.end());
            ix[28 + g] = int32_t(tab.size());
            tab.insert(tab.end(), w.begin(), w.end());
            ix[32 + g] = int32_t(tab.size());
            for (int q = 0; q < nq; ++q) {
                tet_basis(p, &pts[size_t(q) * 3], phi.data(), grad.data());
                for (int i = 0; i < nb; ++i) tab.push_back(double(phi[i]));
            }
            require(P.nq <= 64, HGDG_ERR_UNSUPPORTED, "projection rule with more than 64 points");
        }
        std::vector<double> xj(size_t(c->n_elem) * 12);
        for (int64_t d = 0; d < c->n_elem; ++d) {
            for (int r = 0; r < 3; ++r) xj[size_t(d) * 12 + r] = sm.x0[size_t(d) * 3 + r];
            for (int r = 0; r < 9; ++r) xj[size_t(d) * 12 + 3 + r] = sm.jac[size_t(d) * 9 + r];
        }
        sm.d_qtab.upload(tab, c->stream);
        sm.d_xj.upload(xj, c->stream);
        HGDG_HIP(hipStreamSynchronize(c->stream));  // the host vectors go out of scope
        sm.qtab_index = ix;
    }
    const std::vector<int32_t>& ix = sm.qtab_index;
    for (int g = 0; g < 4; ++g) {
        t.nb[g] = ix[g];
        t.nq_p[g] = ix[4 + g];
        t.nq_e[g] = ix[8 + g];
        t.pts_p[g] = ix[12 + g];
        t.R[g] = ix[16 + g];
        t.phi_p[g] = ix[20 + g];
        t.pts_e[g] = ix[24 + g];
        t.w_e[g] = ix[28 + g];
        t.phi_e[g] = ix[32 + g];
    }
    return t;
}

SimplexFnParams simplex_fn_params(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp,
                                  const double* k, const double* gauss, DevBuf<double>& d_modes) {
    SimplexMesh& sm = c->sm;
    SimplexFnParams p{};
    p.t = simplex_quad_tables(c);
    p.tab = sm.d_qtab.ptr;
    p.xj = sm.d_xj.ptr;
    p.elem_off = sm.d_elem_off.ptr;
    p.order = sm.d_order.ptr;
    p.n_fields = c->n_fields;
    p.field = field;
    p.u = v.ptr();
    p.u_f32 = v.is_f32;
    p.absdet = sm.d_absdet.ptr;
    if (gauss) {
        p.gaussian = 1;
        for (int q = 0; q < 5; ++q) p.g[q] = gauss[q];
    } else if (n_modes > 0) {
        std::vector<double> md(size_t(4) * n_modes);
        for (int m = 0; m < n_modes; ++m) {
            md[4 * m] = amp[m];
            for (int d = 0; d < 3; ++d) md[4 * m + 1 + d] = k[3 * m + d];
        }
        d_modes.upload(md, c->stream);
        HGDG_HIP(hipStreamSynchronize(c->stream));
        p.n_modes = n_modes;
        p.modes = d_modes.ptr;
    }
    return p;
}

}  // namespace

// L2 projection of sine modes or a Gaussian pulse, on the device.
void simplex_project(hgdg_ctx* c, Vec& v, int field, int n_modes, const double* amp, const double* k,
                     const double* gauss) {
    require(c->sm.d_order.n == size_t(c->n_elem), HGDG_ERR_STATE, "set the operator before projecting");
    DevBuf<double> d_modes;
    const SimplexFnParams p = simplex_fn_params(c, v, field, n_modes, amp, k, gauss, d_modes);
    if (c->n_elem > 0) hipLaunchKernelGGL(simplex_project, dim3(unsigned(c->n_elem)), dim3(64), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
    HGDG_HIP(hipStreamSynchronize(c->stream));  // before the modes table goes
}

// Error against f(x - shift) for sine modes f, as tensor_error_modes: L2 with
// hpGEM's tetrahedral rule of order 2p+2, maximum over the points of the
// element rule (order 2p+1), on the device; the per-element sums are added
// on the host in element order.
void simplex_error_modes(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp, const double* k,
                         const double* shift, double* l2, double* max_abs) {
    require(c->sm.d_order.n == size_t(c->n_elem), HGDG_ERR_STATE, "set the operator before computing errors");
    DevBuf<double> d_modes;
    SimplexFnParams p = simplex_fn_params(c, v, field, n_modes, amp, k, nullptr, d_modes);
    for (int d = 0; d < 3; ++d) p.shift[d] = shift[d];
    const int64_t ne = c->n_elem;
    if (c->d_partial.n < size_t(2 * ne)) c->d_partial.alloc(size_t(2 * ne));
    p.partial = c->d_partial.ptr;
    if (ne > 0) hipLaunchKernelGGL(simplex_error, dim3(unsigned(ne)), dim3(64), 0, c->stream, p);
    HGDG_HIP(hipGetLastError());
    std::vector<double> h(size_t(2 * ne));
    HGDG_HIP(hipMemcpyAsync(h.data(), c->d_partial.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HGDG_HIP(hipStreamSynchronize(c->stream));
    double sum = 0, mx = 0;
    for (int64_t d = 0; d < ne; ++d) {
        sum += h[2 * d];
        mx = std::max(mx, h[2 * d + 1]);
    }
    *l2 = std::sqrt(sum);
    *max_abs = mx;
}

void simplex_to_device(hgdg_ctx* c, const double* host, std::vector<double>& dev) {
    const SimplexMesh& sm = c->sm;
    dev.assign(size_t(c->n_dof()), 0.0);
    int64_t off = 0;
    for (int64_t e = 0; e < c->n_elem; ++e) {
        const int64_t d = sm.dev_of[e];
        const int64_t n = int64_t(c->n_fields) * tet_nb(sm.order[d]);
        std::copy(host + off, host + off + n, dev.begin() + sm.dof_off[d]);
        off += n;
    }
}

void simplex_from_device(hgdg_ctx* c, const std::vector<double>& dev, double* host) {
    const SimplexMesh& sm = c->sm;
    int64_t off = 0;
    for (int64_t e = 0; e < c->n_elem; ++e) {
        const int64_t d = sm.dev_of[e];
        const int64_t n = int64_t(c->n_fields) * tet_nb(sm.order[d]);
        std::copy(dev.begin() + sm.dof_off[d], dev.begin() + sm.dof_off[d] + n, host + off);
        off += n;
    }
}

}  // namespace hgdg
