// The library context (hgdg_ctx) and the functions the element families
// implement: tensor elements (squares, cubes) in tensor.hip, tetrahedra in
// simplex.hip, shared state/time stepping and the C ABI in hgdg.hip.
#pragma once

#include <vector>

#include "common.hpp"
#include "host_tables.hpp"
#include "simplex_tables.hpp"

namespace hgdg {


// Squares and cubes of one order p; state in tensor order per element.
struct TensorMesh {
    int p = 0, n1 = 0, nb = 0;
    std::vector<int> perm;          // hpGEM basis index -> tensor index
    std::vector<double> Jinv;       // [e][dim][dim]
    std::vector<double> X0, E;      // [e][dim] first vertex, [e][dim][dim] edge vectors
    std::vector<double> absdet;     // [e]
    std::vector<int32_t> nbr;       // [2 dim][e]
    TensorOps ops;
    DevBuf<int32_t> d_nbr;
    DevBuf<double> d_absdet, d_ops;  // d_ops: Dr, Mi, M1 (n1 x n1 each)
    DevBuf<double> d_geo;            // advection: ahat [dim][e]; acoustic: J^{-1} [dim*dim][e], then c_inv [e]
    DevBuf<uint32_t> d_mode;         // advection: upwind mode, 3 bits per face
    bool structured = false;         // periodic unit box built by tensor_structured_upload
    int32_t cells[3] = {1, 1, 1};
    // Gauss-point path (tensor_general.hip): meshes with non-affine elements
    // or unaligned neighbours, or any mesh when the caller asks for it
    bool general = false;
    std::vector<double> gX;          // [e][2^dim][dim] vertex coordinates
    std::vector<int32_t> gnbr;       // [2 dim][e] neighbour per local face (hpGEM numbering) or -1
    std::vector<int32_t> ginfo;      // [2 dim][e] neighbour's face, point map, left flag
    DevBuf<double> d_gX, d_gtab, d_gcinv;
    DevBuf<int32_t> d_gface;         // gnbr then ginfo
};

// A linear operator given by its matrices (matrix_op.hip): vectors in hpGEM's
// layout, per element row a diagonal block and one block per face neighbour.
struct MatrixOp {
    bool on = false;
    int max_rows = 0;
    std::vector<int64_t> elem_off, sq_off, row_ptr, blk_off;
    std::vector<int32_t> blk_col;
    DevBuf<int64_t> d_elem_off, d_sq_off, d_row_ptr, d_blk_off;
    DevBuf<int32_t> d_blk_col;
    DevBuf<double> d_pool, d_minv, d_mass;
};

// Tetrahedra with per-element order 1..4, grouped by order on the device.
struct SimplexMesh {
    int64_t group_start[5] = {0, 0, 0, 0, 0};    // device elements of order g+1: [start[g], start[g+1])
    int64_t group_dof[5] = {0, 0, 0, 0, 0};      // first DOF of each group
    int64_t group_tr[5] = {0, 0, 0, 0, 0};       // first face trace of each group (4 (p+1)(p+2)/2 per element)
    DevBuf<double> d_trace;                      // face traces of the stage input, double2 per coefficient
    std::vector<int64_t> dev_of;                 // mesh element -> device element
    std::vector<int64_t> mesh_of;                // device element -> mesh element
    std::vector<int64_t> rank;                   // mesh element -> processing (locality) rank
    std::vector<int32_t> order;                  // per device element
    std::vector<int64_t> dof_off;                // per device element: first DOF
    std::vector<SimplexOps> ops;                 // per order (index p-1), built on demand
    DevBuf<double> d_tables;                     // Dr, proj, lift of every order
    DevBuf<float> d_tables32;                    // the same in FP32 (mixed path)
    DevBuf<int32_t> d_blocks;                    // RHS block schedule: [n_blocks][4] order, first, count, variant
    int64_t n_blocks = 0;
    int32_t tpe_variant[4] = {0, 0, 0, 0};       // per order: threads-per-element candidate in use (0 or 1)
    int32_t tpe_fixed[4] = {0, 0, 0, 0};         // per order: 0 = timed choice, else the fixed threads per element
    double tpe_ms[4][2] = {};                    // per order and candidate: timed RHS launch (ms), 0 if not timed
    DevBuf<int32_t> d_tab_index;                 // offsets and column lists into d_tables
    DevBuf<double> d_mass;                       // per order: M_ref [nb][nb], then M_ref^{-1}
    DevBuf<double> d_qtab;                       // projection / error-norm quadrature tables (built on first use)
    std::vector<int32_t> qtab_index;             // their offsets and sizes per order
    DevBuf<double> d_xj;                         // [E][12]: vertex 0 and J (row-major), device order
    DevBuf<int32_t> d_order;                     // per device element: its order
    DevBuf<int64_t> d_elem_off;                  // [E + 1] first DOF of each device element
    DevBuf<double> d_geo;                        // J^{-1} [9][E], then c_inv [E], sigma [4][E], normal [3][4][E]
    DevBuf<int32_t> d_face;                      // [4][E]: neighbour device element or -1, then packed info [4][E]
    DevBuf<double> d_absdet;                     // [E]
    std::vector<double> absdet;
    std::vector<double> jinv;                    // [E][9]
    std::vector<double> X;                       // mesh order [E][4][3]
    std::vector<int32_t> nbr_dev;                // [4][E]
    std::vector<int32_t> info;                   // [4][E]
    std::vector<double> sigma, normal;           // [4][E], [3][4][E]
    std::vector<double> x0, jac;                 // device order: vertex 0 [E][3], J [E][3][3] (row-major)
};

}  // namespace hgdg

struct hgdg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool has_mesh = false;
    bool has_op = false;
    int shape = 0, dim = 0, n_fields = 1;
    int64_t n_elem = 0, n_faces = 0, n_internal = 0, dof = 0;
    hgdg_op_desc op{};
    std::vector<double> cinv;  // mesh order
    hgdg::TensorMesh tm;
    hgdg::SimplexMesh sm;
    hgdg::MatrixOp mat;
    std::vector<hgdg::Vec> vecs;
    hgdg::DevBuf<double> d_partial;
    bool graph_valid = false;
    hgdg::GraphKey graph_key{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // two steps with the RK update fused into each last stage, ping-ponging
    // the solution through u_spare (captured for one solution buffer)
    bool graph2_valid = false;
    hgdg::GraphKey graph2_key{};
    const double* graph2_u = nullptr;
    hipGraph_t graph2 = nullptr;
    hipGraphExec_t graph2_exec = nullptr;
    hgdg::DevBuf<double> u_spare;
    int64_t rk_pair_launches = 0;
    double dt_estimate = 0;  // adaptive stepping: next step size (0 = not started)
    double last_kernel_ms = 0;  // duration of the last assembly kernel (HIP events)
    // tensor advection: tiles of elements each workgroup takes (1 = one tile,
    // no prefetch); chosen by timing at operator set-up unless fixed by the caller
    int adv_tiles = 1;
    int adv_tiles_fixed = 0;
    // 1: RK steps of meshes that fit in LDS in one single-workgroup launch, and
    // elsewhere (tensor advection) the RK update fused into each step's last
    // stage; 0: always the graph of separate launches
    int rk_fused = 1;
    // 1: tensor meshes uploaded from now on take the Gauss-point path even when
    // every element is affine and every neighbour aligned
    int tensor_general_forced = 0;
    int64_t rk_fused_launches = 0;
    double adv_tile_ms[4] = {0, 0, 0, 0};  // timed candidates 1, 2, 4, 8 (0 = not timed)

    bool simplex() const { return shape == HGDG_SHAPE_TETRAHEDRON && !mat.on; }
    bool matrix() const { return mat.on; }
    int64_t n_dof() const { return dof; }
    void drop_graph2() {
        if (graph2_exec) (void)hipGraphExecDestroy(graph2_exec);
        if (graph2) (void)hipGraphDestroy(graph2);
        graph2_exec = nullptr;
        graph2 = nullptr;
        graph2_valid = false;
    }
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_valid = false;
        drop_graph2();
    }
};

namespace hgdg {

Vec& vec_at(hgdg_ctx* c, int id);

// tensor.hip
void tensor_mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m);
void tensor_structured_upload(hgdg_ctx* c, int dim, const int32_t* n, int p);
// L2 projection of f into one field (field = -1: every field) of vector v:
// f = sum_m amp_m sin(2 pi k_m . x), or, with gauss = {cx, cy, cz, width,
// amplitude}, f = amplitude exp(-|x - c|^2 / width^2).
void tensor_project_device(hgdg_ctx* c, Vec& v, int field, int n_modes, const double* amp, const double* k,
                           const double* gauss = nullptr);
void tensor_operator_set(hgdg_ctx* c, const hgdg_op_desc* op);
void tensor_rhs(hgdg_ctx* c, const StageIO& io);
// n_steps RK steps in one single-workgroup launch when the mesh is small
// enough for LDS (tensor advection); false: not eligible, nothing launched
bool tensor_rk_tiny(hgdg_ctx* c, int solution, const int* aux, int stages, const double* a, const double* b,
                    double dt, int n_steps);
// Chooses c->adv_tiles for the advection operator by timing each candidate
// on scratch vectors of the mesh's size (or takes c->adv_tiles_fixed).
void tensor_tune(hgdg_ctx* c);
void tensor_state_diff(hgdg_ctx* c, const Vec& x, const Vec& y, double* l2, double* max_abs);
// v = M v (inverse = false) or v = M^{-1} v (inverse = true), in place, FP64
void tensor_mass_apply(hgdg_ctx* c, Vec& v, bool inverse);
void tensor_error_modes(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp, const double* k,
                        const double* shift, double* l2, double* max_abs);
void tensor_to_device(hgdg_ctx* c, const double* host, std::vector<double>& dev);
void tensor_from_device(hgdg_ctx* c, const std::vector<double>& dev, double* host);

// tensor_general.hip: the Gauss-point path (non-affine elements, unaligned faces)
void tensor_general_upload(hgdg_ctx* c, const hgdg_mesh_desc* m);
void tensor_general_operator_set(hgdg_ctx* c, const hgdg_op_desc* op);
void tensor_general_rhs(hgdg_ctx* c, const StageIO& io);
void tensor_general_mass_apply(hgdg_ctx* c, Vec& v, bool inverse);
void tensor_general_project(hgdg_ctx* c, Vec& v, int field, int n_modes, const double* amp, const double* k,
                            const double* gauss);
void tensor_general_error_modes(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp,
                                const double* k, const double* shift, double* l2, double* max_abs);
void tensor_general_state_diff(hgdg_ctx* c, const Vec& x, const Vec& y, double* l2, double* max_abs);

// batched tensor launches (serving)
size_t tensor_batch_param_bytes(const hgdg_ctx* c);
int64_t tensor_batch_blocks(const hgdg_ctx* c);
void tensor_batch_param(hgdg_ctx* c, const StageIO& io, void* out);
void tensor_batch_launch(hipStream_t stream, int kind, int dim, int n1, bool mixed, bool upd, const void* params,
                         const int64_t* off, int n, int64_t blocks);

// hgdg.hip
void set_storage(hgdg_ctx* c, Vec& v, bool f32);
void mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m);
void operator_set(hgdg_ctx* c, const hgdg_op_desc* op);
void upload_vector(hgdg_ctx* c, Vec& v, const std::vector<double>& dev);
void download_vector(hgdg_ctx* c, Vec& v, std::vector<double>& dev);
void state_diff(hgdg_ctx* c, int x, int y, double* l2, double* max_abs);

// matrix_op.hip
void matrix_operator_set(hgdg_ctx* c, const hgdg_matrix_desc* d, int precision);
void matrix_rhs(hgdg_ctx* c, const StageIO& io, bool solve);
void matrix_mass_apply(hgdg_ctx* c, Vec& v, bool inverse);
void matrix_state_diff(hgdg_ctx* c, const Vec& x, const Vec& y, double* l2, double* max_abs);

// maxwell.hip
void maxwell_assemble(hgdg_ctx* c, int order, const double* eps, double stab, int64_t* n_blocks, int32_t* block_size,
                      int64_t* row_ptr, int64_t* col_idx, double* K, double* M);

// simplex.hip
void simplex_mesh_upload(hgdg_ctx* c, const hgdg_mesh_desc* m);
void simplex_operator_set(hgdg_ctx* c, const hgdg_op_desc* op);
void simplex_schedule(hgdg_ctx* c);
void simplex_tune(hgdg_ctx* c);
int simplex_threads(int order, int variant);
void simplex_rhs(hgdg_ctx* c, const StageIO& io);
void simplex_state_diff(hgdg_ctx* c, const Vec& x, const Vec& y, double* l2, double* max_abs);
void simplex_mass_apply(hgdg_ctx* c, Vec& v, bool inverse);
void simplex_error_modes(hgdg_ctx* c, const Vec& v, int field, int n_modes, const double* amp, const double* k,
                         const double* shift, double* l2, double* max_abs);
void simplex_project(hgdg_ctx* c, Vec& v, int field, int n_modes, const double* amp, const double* k,
                     const double* gauss = nullptr);
void simplex_to_device(hgdg_ctx* c, const double* host, std::vector<double>& dev);
void simplex_from_device(hgdg_ctx* c, const std::vector<double>& dev, double* host);

}  // namespace hgdg
