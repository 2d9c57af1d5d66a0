// Serving loop: many independent explicit DG solves on one GPU with
// continuous batching.
//
// A solve is a structured periodic square/cube mesh with one operator
// (upwind advection or the acoustic system), one order, sine-mode initial
// data and a number of RK4 steps. A preparer thread admits queued solves in
// arrival order as soon as their HBM footprint fits the budget (and a slot is
// free), builds them (mesh, faces, tables, projected initial state) and hands
// them to the worker. The worker advances every resident solve by one RK4
// step per tick: each stage is one batched launch per (operator, dim, order)
// family that covers all resident solves of that family, and the last
// stage's launches also apply each solve's RK update (the solution
// ping-pongs between two vectors, as in the single-problem fused steps). Solves join at the next tick after they are
// built and leave when their step count is reached, freeing memory for the
// queue. Every element of every solve runs the exact code of the
// single-problem kernel on its own data, so a result does not depend on what
// else was resident.
#include <atomic>
#include <algorithm>
#include <array>
#include <chrono>
#include <functional>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "context.hpp"

namespace hgdg {

namespace {

double now_s() {
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

// RK4 (kernel/Base/TimeIntegration/RK4Methods.cpp:45-63)
constexpr double kA[4] = {0.0, 0.5, 0.5, 1.0};  // a_{i,i-1}; all other a_ij vanish
constexpr double kB[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};

struct Request {
    int64_t ticket = 0;
    hgdg_solve_desc d{};
    std::vector<double> amp, k;
    double t_submit = 0;
};

struct Solve {
    Request req;
    hgdg_ctx* ctx = nullptr;
    int steps_done = 0;
    int cur = 0;  // where the solution is: 0 vecs[0], 1 ctx->u_spare
    int64_t bytes = 0;
    double t_start = 0;
};

struct Result {
    int status = 0;  // 0 queued, 1 resident, 2 done, 3 failed
    std::vector<double> state;
    double norm = 0;
    double t_submit = 0, t_start = 0, t_end = 0;
    int64_t n_dof = 0;
    int n_steps = 0;
    std::string error;
};

int64_t footprint(const hgdg_solve_desc& d, bool mixed) {
    int64_t cells = 1;
    for (int i = 0; i < d.dim; ++i) cells *= d.cells[i];
    int64_t nb = 1;
    for (int i = 0; i < d.dim; ++i) nb *= d.order + 1;
    const int nf = d.kind == HGDG_OP_ACOUSTIC_CENTRAL ? d.dim + 1 : 1;
    const int64_t dof = cells * nb * nf;
    // solution and its ping-pong copy in FP64, four stage vectors
    return dof * (2 * 8 + 4 * (mixed ? 4 : 8)) + cells * (8 * (d.dim * d.dim + 2) + 4 * 2 * d.dim + 8);
}

// A solve's context runs on a stream of the server's (the preparer's while it
// is built, the worker's once resident), which outlives it.
void destroy_ctx(hgdg_ctx* c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    c->drop_graph();
    c->vecs.clear();
    delete c;
}

}  // namespace

}  // namespace hgdg

struct hgdg_server {
    int device = 0;
    bool mixed = false;
    int64_t budget = 0;
    int max_active = 0;
    hipStream_t stream = nullptr;       // the worker's: every resident solve's steps
    hipStream_t prep_stream = nullptr;  // the preparer's: building solves
    std::mutex mu;
    std::condition_variable cv;
    // Serialises the two threads' allocation-heavy HIP work: the preparer
    // builds a solve (hipMalloc/hipFree, uploads, launches on prep_stream)
    // while holding it; the worker holds it while it re-stages (parameter
    // blob reallocation, stream capture, graph instantiation) and while it
    // retires solves (downloads, hipFree). hipFree waits for the whole device,
    // so without this lock it could run while the worker's stream is in
    // capture; the steady-state ticks (one graph launch and a stream
    // synchronisation) take no lock and overlap with building.
    std::mutex hip_mu;
    bool stop = false;
    std::deque<hgdg::Request> pending;
    std::deque<hgdg::Solve*> ready;
    std::vector<hgdg::Solve*> active;
    std::map<int64_t, hgdg::Result> results;
    int64_t next_ticket = 0;
    int64_t used = 0, peak = 0;
    int64_t building = 0;  // solves reserved but not yet resident
    double t_first_submit = -1, t_last_done = 0;
    double dof_rhs = 0;  // sum over completed solves of DOF x RHS evaluations
    std::thread preparer, worker;
    // device staging for batched launch parameters; they stay valid while the
    // set of resident solves is the same (the launches of one RK step do not
    // change from step to step)
    hgdg::DevBuf<char> d_blob;
    char* h_blob = nullptr;
    size_t h_blob_bytes = 0;
    struct Launch {
        int kind, dim, n1;
        size_t params, off;
        int n;
        int64_t blocks;
        int phase;  // the step graph it belongs to
        bool upd;   // last stage: with the RK update
    };
    std::vector<int64_t> staged;  // tickets of the solves the device blob describes (never reused, unlike addresses)
    // One RK4 step of the staged solves per parity, captured once per
    // resident set: graph k takes each solve's solution from where it was at
    // staging (Solve::cur) when k = 0, from the other vector when k = 1, and
    // writes the new solution to the other one; the ticks alternate them.
    hipGraph_t step_graph[2] = {nullptr, nullptr};
    hipGraphExec_t step_exec[2] = {nullptr, nullptr};
    int phase = 0;  // the graph of the next step
    void drop_step_graph() {
        for (int k = 0; k < 2; ++k) {
            if (step_exec[k]) (void)hipGraphExecDestroy(step_exec[k]);
            if (step_graph[k]) (void)hipGraphDestroy(step_graph[k]);
            step_exec[k] = nullptr;
            step_graph[k] = nullptr;
        }
    }
    std::vector<Launch> launches;
    // Per-family kernel times: with `profile` set, each tick issues the staged
    // launches one by one between HIP events instead of replaying the step
    // graph (the same launches, so the same results) and accumulates the
    // events' elapsed times per (kind, dim, n1) family; the RK updates run
    // inside each family's last-stage launch, so they are part of its time.
    std::atomic<bool> profile{false};
    std::vector<hipEvent_t> events;  // 2 per staged launch
    std::map<std::array<int, 3>, std::pair<double, int64_t>> family_ms;  // guarded by mu
    double update_ms = 0;  // separate update passes: none (reported as 0)
    // worker accounting (guarded by mu): ticks, re-stagings and their wall time
    int64_t n_ticks = 0, n_restages = 0;
    double restage_s = 0, tick_s = 0;
    void drop_events() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        events.clear();
    }

    void prepare_loop();
    void worker_loop();
    void stage(std::vector<hgdg::Solve*>& solves);
    void tick(std::vector<hgdg::Solve*>& solves, int n_steps);
};

namespace hgdg {
namespace {

hgdg_ctx* build_solve(int device, bool mixed, const Request& r, hipStream_t stream) {
    const hgdg_solve_desc& d = r.d;
    std::unique_ptr<hgdg_ctx> c(new hgdg_ctx);
    c->device = device;
    c->stream = stream;
    const int32_t n[3] = {d.cells[0], d.cells[1], d.cells[2]};
    c->shape = d.dim == 3 ? SHAPE_CUBE : SHAPE_SQUARE;
    tensor_structured_upload(c.get(), d.dim, n, d.order);
    c->has_mesh = true;
    hgdg_op_desc op{};
    op.kind = d.kind;
    op.precision = mixed ? HGDG_PREC_MIXED_FP32 : HGDG_PREC_FP64;
    std::vector<double> cinv;
    if (d.kind == HGDG_OP_ACOUSTIC_CENTRAL) {
        cinv.assign(size_t(c->n_elem), 1.0 / (d.wave_speed * d.wave_speed));
        op.c_inv = cinv.data();
    } else {
        for (int i = 0; i < 3; ++i) op.velocity[i] = d.velocity[i];
    }
    c->adv_tiles_fixed = 1;  // requests run in the batched kernels; no launch-shape timing per request
    operator_set(c.get(), &op);
    c->vecs.resize(5);
    set_storage(c.get(), c->vecs[0], false);
    for (int i = 1; i < 5; ++i) set_storage(c.get(), c->vecs[i], mixed);
    c->u_spare.alloc(size_t(c->n_dof()));  // the solution's ping-pong copy
    HGDG_HIP(hipMemsetAsync(c->vecs[0].f64->ptr, 0, size_t(c->n_dof()) * sizeof(double), c->stream));
    tensor_project_device(c.get(), c->vecs[0], d.kind == HGDG_OP_ACOUSTIC_CENTRAL ? 0 : -1, d.n_modes, r.amp.data(),
                          r.k.data());
    HGDG_HIP(hipStreamSynchronize(c->stream));
    return c.release();
}

}  // namespace
}  // namespace hgdg

void hgdg_server::prepare_loop() {
    (void)hipSetDevice(device);
    for (;;) {
        hgdg::Request r;
        int64_t need = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !pending.empty(); });
            if (stop) return;
            const hgdg::Request& front = pending.front();
            need = hgdg::footprint(front.d, mixed);
            // admission by footprint and slots, in arrival order
            cv.wait(lk, [&] {
                const int64_t resident = int64_t(active.size() + ready.size()) + building;
                return stop || ((used + need <= budget || resident == 0) && resident < max_active);
            });
            if (stop) return;
            r = std::move(pending.front());
            pending.pop_front();
            used += need;
            peak = std::max(peak, used);
            ++building;
        }
        hgdg::Solve* s = new hgdg::Solve;
        s->req = std::move(r);
        s->bytes = need;
        try {
            std::lock_guard<std::mutex> hk(hip_mu);
            s->ctx = hgdg::build_solve(device, mixed, s->req, prep_stream);
        } catch (const std::exception& ex) {
            std::lock_guard<std::mutex> lk(mu);
            auto& res = results[s->req.ticket];
            res.status = 3;
            res.error = ex.what();
            res.t_end = hgdg::now_s();
            used -= need;
            --building;
            delete s;
            cv.notify_all();
            continue;
        }
        std::lock_guard<std::mutex> lk(mu);
        --building;
        ready.push_back(s);
        cv.notify_all();
    }
}

void hgdg_server::stage(std::vector<hgdg::Solve*>& solves) {
    using namespace hgdg;
    // group by kernel family
    std::map<std::array<int, 3>, std::vector<Solve*>> groups;
    for (Solve* s : solves) groups[{s->ctx->op.kind, s->ctx->dim, s->ctx->tm.n1}].push_back(s);
    // lay out the parameter blob: per graph, stage and family [params][offsets]
    launches.clear();
    std::vector<char> blob;
    auto align = [&](size_t a) { blob.resize((blob.size() + a - 1) / a * a); };
    for (int k = 0; k < 2; ++k) {
        for (int stage = 0; stage < 4; ++stage) {
            for (auto& kv : groups) {
                const auto& g = kv.second;
                const size_t pb = tensor_batch_param_bytes(g[0]->ctx);
                Launch L{kv.first[0], kv.first[1], kv.first[2], 0, 0, int(g.size()), 0, k, stage == 3};
                align(64);
                L.params = blob.size();
                blob.resize(blob.size() + pb * g.size());
                std::vector<int64_t> off(g.size() + 1, 0);
                for (size_t i = 0; i < g.size(); ++i) {
                    hgdg_ctx* c = g[i]->ctx;
                    const bool spare = (g[i]->cur ^ k) != 0;  // the step's solution is in u_spare
                    double* u_cur = spare ? c->u_spare.ptr : c->vecs[0].f64->ptr;
                    double* u_next = spare ? c->vecs[0].f64->ptr : c->u_spare.ptr;
                    StageIO io{};
                    io.n_in = stage == 0 ? 1 : 2;
                    io.in[0] = u_cur;
                    io.coef[0] = 1.0;
                    if (stage > 0) {
                        io.in[1] = c->vecs[stage].ptr();
                        io.coef[1] = g[i]->req.d.dt * kA[stage];
                        if (c->vecs[stage].is_f32) io.in_f32 |= 2u;
                    }
                    io.out = c->vecs[1 + stage].ptr();
                    io.out_f32 = c->vecs[1 + stage].is_f32;
                    if (stage == 3) {  // u_next = u_cur + dt sum_j b_j k_j, k_4 from the kernel's registers
                        RkUpdate& q = io.upd;
                        q.u = u_cur;
                        q.y_new = u_next;
                        q.n_x = 4;
                        q.self = 3;
                        for (int j = 0; j < 4; ++j) {
                            q.x[j] = c->vecs[1 + j].ptr();
                            q.c[j] = g[i]->req.d.dt * kB[j];
                            if (c->vecs[1 + j].is_f32) q.x_f32 |= 1u << j;
                        }
                    }
                    tensor_batch_param(c, io, blob.data() + L.params + pb * i);
                    off[i + 1] = off[i] + tensor_batch_blocks(c);
                }
                L.blocks = off.back();
                align(64);
                L.off = blob.size();
                blob.resize(blob.size() + sizeof(int64_t) * off.size());
                std::memcpy(blob.data() + L.off, off.data(), sizeof(int64_t) * off.size());
                launches.push_back(L);
            }
        }
    }

    staged.clear();  // nothing is staged while the blob is replaced
    if (d_blob.n < blob.size()) d_blob.alloc(blob.size() * 2);
    if (h_blob_bytes < blob.size()) {
        if (h_blob) (void)hipHostFree(h_blob);
        h_blob_bytes = blob.size() * 2;
        HGDG_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_blob), h_blob_bytes));
    }
    std::memcpy(h_blob, blob.data(), blob.size());
    HGDG_HIP(hipMemcpyAsync(d_blob.ptr, h_blob, blob.size(), hipMemcpyHostToDevice, stream));
    // each parity's step as a graph: a tick is then one graph launch
    drop_step_graph();
    for (int k = 0; k < 2; ++k) {
        HGDG_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            for (const Launch& L : launches)
                if (L.phase == k)
                    tensor_batch_launch(stream, L.kind, L.dim, L.n1, mixed, L.upd, d_blob.ptr + L.params,
                                        reinterpret_cast<const int64_t*>(d_blob.ptr + L.off), L.n, L.blocks);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HGDG_HIP(hipStreamEndCapture(stream, &step_graph[k]));
        HGDG_HIP(hipGraphInstantiate(&step_exec[k], step_graph[k], nullptr, nullptr, 0));
    }
    phase = 0;
    for (Solve* sv : solves) staged.push_back(sv->req.ticket);
}

// n_steps RK4 steps of every resident solve (the captured steps replayed,
// alternating parity), then one synchronisation.
void hgdg_server::tick(std::vector<hgdg::Solve*>& solves, int n_steps) {
    using namespace hgdg;
    bool same = solves.size() == staged.size();
    for (size_t i = 0; same && i < solves.size(); ++i) same = solves[i]->req.ticket == staged[i];
    const bool timed = profile.load();
    const double t0 = now_s();
    if (!same || (timed && events.size() < 2 * launches.size())) {
        std::lock_guard<std::mutex> hk(hip_mu);
        if (!same) {
            stage(solves);
            const double dt = now_s() - t0;
            std::lock_guard<std::mutex> lk(mu);
            ++n_restages;
            restage_s += dt;
        }
        while (timed && events.size() < 2 * launches.size()) {
            hipEvent_t e;
            HGDG_HIP(hipEventCreate(&e));
            events.push_back(e);
        }
    }
    auto advance = [&] {  // the step just enqueued moved every solution to its other vector
        phase ^= 1;
        for (Solve* s : solves) s->cur ^= 1;
    };
    if (!timed) {
        for (int step = 0; step < n_steps; ++step) {
            HGDG_HIP(hipGraphLaunch(step_exec[phase], stream));
            advance();
        }
        HGDG_HIP(hipStreamSynchronize(stream));
        std::lock_guard<std::mutex> lk(mu);
        n_ticks += n_steps;
        tick_s += now_s() - t0;
        return;
    }
    for (int step = 0; step < n_steps; ++step) {
        for (size_t i = 0; i < launches.size(); ++i) {
            const Launch& L = launches[i];
            if (L.phase != phase) continue;
            HGDG_HIP(hipEventRecord(events[2 * i], stream));
            tensor_batch_launch(stream, L.kind, L.dim, L.n1, mixed, L.upd, d_blob.ptr + L.params,
                                reinterpret_cast<const int64_t*>(d_blob.ptr + L.off), L.n, L.blocks);
            HGDG_HIP(hipEventRecord(events[2 * i + 1], stream));
        }
        HGDG_HIP(hipStreamSynchronize(stream));
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < launches.size(); ++i) {
                if (launches[i].phase != phase) continue;
                float ms = 0;
                HGDG_HIP(hipEventElapsedTime(&ms, events[2 * i], events[2 * i + 1]));
                auto& f = family_ms[{launches[i].kind, launches[i].dim, launches[i].n1}];
                f.first += ms;
                ++f.second;
            }
            ++n_ticks;
        }
        advance();
    }
    std::lock_guard<std::mutex> lk(mu);
    tick_s += now_s() - t0;
}

void hgdg_server::worker_loop() {
    (void)hipSetDevice(device);
    std::vector<hgdg::Solve*> mine;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !ready.empty() || !mine.empty(); });
            if (stop) break;
            const double t = hgdg::now_s();
            while (!ready.empty()) {
                hgdg::Solve* s = ready.front();
                ready.pop_front();
                s->t_start = t;
                s->ctx->stream = stream;  // resident: the worker's stream from now on
                results[s->req.ticket].status = 1;
                results[s->req.ticket].t_start = t;
                mine.push_back(s);
                active.push_back(s);
            }
        }
        std::vector<hgdg::Solve*> running;
        std::vector<hgdg::Solve*> finished;
        for (hgdg::Solve* s : mine) {
            if (s->steps_done < s->req.d.n_steps) running.push_back(s);
            else finished.push_back(s);
        }
        std::string err;
        if (!running.empty()) {
            // steps between synchronisations: more than one saves host round
            // trips but delays retiring and joining; measured on the serving
            // stream, the GPU is the bottleneck and one step per sync gives
            // the lowest latencies
            constexpr int kStepsPerSync = 1;
            int n_steps = kStepsPerSync;
            for (hgdg::Solve* s : running) n_steps = std::min(n_steps, int(s->req.d.n_steps - s->steps_done));
            try {
                tick(running, n_steps);
                for (hgdg::Solve* s : running) {
                    s->steps_done += n_steps;
                    if (s->steps_done >= s->req.d.n_steps) finished.push_back(s);
                }
            } catch (const std::exception& ex) {
                err = ex.what();
                finished = mine;  // a failed tick ends every resident solve
                std::lock_guard<std::mutex> hk(hip_mu);
                (void)hipStreamSynchronize(stream);  // nothing of the old graph may still run
                staged.clear();
                drop_step_graph();
            }
        }
        for (hgdg::Solve* s : finished) {
            std::unique_lock<std::mutex> hk(hip_mu);
            std::vector<double> state;
            double norm = 0;
            int64_t n_dof = 0;
            std::string serr = err;
            if (serr.empty()) {
                try {
                    hgdg_ctx* c = s->ctx;
                    n_dof = c->n_dof();
                    std::vector<double> dev;
                    if (s->cur)  // the last step left the solution in the ping-pong copy
                        HGDG_HIP(hipMemcpyAsync(c->vecs[0].f64->ptr, c->u_spare.ptr, size_t(n_dof) * sizeof(double),
                                                hipMemcpyDeviceToDevice, c->stream));
                    hgdg::download_vector(c, c->vecs[0], dev);
                    for (double v : dev) norm += v * v;
                    norm = std::sqrt(norm);
                    if (s->req.d.keep_state) {
                        state.resize(size_t(n_dof));
                        hgdg::tensor_from_device(c, dev, state.data());
                    }
                } catch (const std::exception& ex) {
                    serr = ex.what();
                }
            }
            hgdg::destroy_ctx(s->ctx);
            s->ctx = nullptr;
            hk.unlock();
            std::lock_guard<std::mutex> lk(mu);
            auto& res = results[s->req.ticket];
            res.t_end = hgdg::now_s();
            res.n_dof = n_dof;
            res.n_steps = s->steps_done;
            if (serr.empty()) {
                res.status = 2;
                res.state = std::move(state);
                res.norm = norm;
                dof_rhs += double(n_dof) * 4.0 * double(s->steps_done);
            } else {
                res.status = 3;
                res.error = serr;
            }
            t_last_done = std::max(t_last_done, res.t_end);
            used -= s->bytes;
            active.erase(std::remove(active.begin(), active.end(), s), active.end());
            mine.erase(std::remove(mine.begin(), mine.end(), s), mine.end());
            delete s;
        }
        if (!finished.empty()) {
            std::lock_guard<std::mutex> lk(mu);
            cv.notify_all();
        }
    }
    std::lock_guard<std::mutex> hk(hip_mu);
    for (hgdg::Solve* s : mine) {
        hgdg::destroy_ctx(s->ctx);
        delete s;
    }
}

using hgdg::require;

namespace hgdg {
int run(const std::function<void()>& f);
}

extern "C" {

int hgdg_server_create(hgdg_server** out, int device, int precision, int64_t hbm_budget_bytes, int max_active) {
    return hgdg::run([&] {
        require(out != nullptr, HGDG_ERR_ARG, "out is NULL");
        require(precision == HGDG_PREC_FP64 || precision == HGDG_PREC_MIXED_FP32, HGDG_ERR_ARG, "unknown precision");
        require(hbm_budget_bytes > 0 && max_active > 0, HGDG_ERR_ARG, "budget and slots must be positive");
        int count = 0;
        HGDG_HIP(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, HGDG_ERR_ARG, "no HIP device " + std::to_string(device));
        HGDG_HIP(hipSetDevice(device));
        std::unique_ptr<hgdg_server> s(new hgdg_server);
        s->device = device;
        s->mixed = precision == HGDG_PREC_MIXED_FP32;
        s->budget = hbm_budget_bytes;
        s->max_active = max_active;
        HGDG_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HGDG_HIP(hipStreamCreateWithFlags(&s->prep_stream, hipStreamNonBlocking));
        hgdg_server* raw = s.get();
        s->preparer = std::thread([raw] { raw->prepare_loop(); });
        s->worker = std::thread([raw] { raw->worker_loop(); });
        *out = s.release();
    });
}

int hgdg_server_submit(hgdg_server* s, const hgdg_solve_desc* d, int64_t* ticket) {
    return hgdg::run([&] {
        require(s != nullptr && d != nullptr && ticket != nullptr, HGDG_ERR_ARG, "NULL argument");
        require(d->dim == 2 || d->dim == 3, HGDG_ERR_ARG, "dim must be 2 or 3");
        require(d->kind == HGDG_OP_ADVECTION_UPWIND || d->kind == HGDG_OP_ACOUSTIC_CENTRAL, HGDG_ERR_ARG,
                "unknown operator kind");
        require(d->order >= 1 && d->order <= 5, HGDG_ERR_UNSUPPORTED, "order must be in 1..5");
        for (int i = 0; i < d->dim; ++i) require(d->cells[i] >= 3, HGDG_ERR_ARG, "at least 3 cells per direction");
        require(d->n_steps >= 0 && d->dt > 0, HGDG_ERR_ARG, "need dt > 0 and n_steps >= 0");
        require(d->n_modes >= 0 && (d->n_modes == 0 || (d->amp && d->k)), HGDG_ERR_ARG, "bad initial modes");
        require(d->kind != HGDG_OP_ACOUSTIC_CENTRAL || d->wave_speed > 0, HGDG_ERR_ARG, "wave speed must be positive");
        hgdg::Request r;
        r.d = *d;
        r.amp.assign(d->amp, d->amp + d->n_modes);
        r.k.assign(d->k, d->k + 3 * size_t(d->n_modes));
        r.d.amp = nullptr;
        r.d.k = nullptr;
        r.t_submit = hgdg::now_s();
        std::lock_guard<std::mutex> lk(s->mu);
        r.ticket = s->next_ticket++;
        if (s->t_first_submit < 0) s->t_first_submit = r.t_submit;
        auto& res = s->results[r.ticket];
        res.status = 0;
        res.t_submit = r.t_submit;
        *ticket = r.ticket;
        s->pending.push_back(std::move(r));
        s->cv.notify_all();
    });
}

int hgdg_server_wait(hgdg_server* s, int64_t ticket, double timeout_s, int* done) {
    return hgdg::run([&] {
        require(s != nullptr && done != nullptr, HGDG_ERR_ARG, "NULL argument");
        std::unique_lock<std::mutex> lk(s->mu);
        auto finished = [&] {
            if (ticket >= 0) {
                auto it = s->results.find(ticket);
                require(it != s->results.end(), HGDG_ERR_ARG, "unknown ticket");
                return it->second.status >= 2;
            }
            for (auto& kv : s->results)
                if (kv.second.status < 2) return false;
            return true;
        };
        *done = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), finished) ? 1 : 0;
    });
}

int hgdg_server_result(hgdg_server* s, int64_t ticket, int64_t* n_dof, double* state, double* norm, double* latency_s) {
    return hgdg::run([&] {
        require(s != nullptr, HGDG_ERR_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->results.find(ticket);
        require(it != s->results.end(), HGDG_ERR_ARG, "unknown ticket");
        const hgdg::Result& r = it->second;
        require(r.status != 3, HGDG_ERR_STATE, "solve failed: " + r.error);
        require(r.status == 2, HGDG_ERR_STATE, "solve not finished");
        if (n_dof) *n_dof = r.n_dof;
        if (norm) *norm = r.norm;
        if (latency_s) *latency_s = r.t_end - r.t_submit;
        if (state) {
            require(!r.state.empty() || r.n_dof == 0, HGDG_ERR_STATE, "the final state was not kept (keep_state = 0)");
            std::copy(r.state.begin(), r.state.end(), state);
        }
    });
}

int hgdg_server_get_stats(hgdg_server* s, hgdg_server_stats* out) {
    return hgdg::run([&] {
        require(s != nullptr && out != nullptr, HGDG_ERR_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        std::vector<double> lat;
        hgdg_server_stats st{};
        for (auto& kv : s->results) {
            ++st.submitted;
            if (kv.second.status == 2) {
                ++st.completed;
                lat.push_back(kv.second.t_end - kv.second.t_submit);
            } else if (kv.second.status == 3) {
                ++st.failed;
            } else if (kv.second.status == 1) {
                ++st.active;
            }
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) {
            if (lat.empty()) return 0.0;
            // nearest-rank percentile
            size_t idx = size_t(std::ceil(q * double(lat.size()))) - 1;
            return lat[std::min(idx, lat.size() - 1)];
        };
        st.latency_p50_s = pct(0.50);
        st.latency_p99_s = pct(0.99);
        double sum = 0;
        for (double v : lat) sum += v;
        st.latency_mean_s = lat.empty() ? 0 : sum / double(lat.size());
        st.wall_s = s->t_first_submit < 0 ? 0 : s->t_last_done - s->t_first_submit;
        st.dof_rhs_per_s = st.wall_s > 0 ? s->dof_rhs / st.wall_s : 0;
        st.peak_bytes = s->peak;
        st.ticks = s->n_ticks;
        st.restages = s->n_restages;
        st.restage_s = s->restage_s;
        st.tick_s = s->tick_s;
        *out = st;
    });
}

int hgdg_server_set_profiling(hgdg_server* s, int enable) {
    return hgdg::run([&] {
        require(s != nullptr, HGDG_ERR_ARG, "NULL argument");
        require(enable == 0 || enable == 1, HGDG_ERR_ARG, "enable must be 0 or 1");
        s->profile.store(enable == 1);
    });
}

int hgdg_server_family_times(hgdg_server* s, int capacity, int32_t* kind, int32_t* dim, int32_t* order,
                             double* kernel_ms, int64_t* launches, int* n_families, double* update_ms) {
    return hgdg::run([&] {
        require(s != nullptr && n_families != nullptr, HGDG_ERR_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        *n_families = int(s->family_ms.size());
        if (update_ms) *update_ms = s->update_ms;
        int i = 0;
        for (const auto& kv : s->family_ms) {
            if (i >= capacity) break;
            if (kind) kind[i] = kv.first[0];
            if (dim) dim[i] = kv.first[1];
            if (order) order[i] = kv.first[2] - 1;
            if (kernel_ms) kernel_ms[i] = kv.second.first;
            if (launches) launches[i] = kv.second.second;
            ++i;
        }
    });
}

int hgdg_server_destroy(hgdg_server* s) {
    return hgdg::run([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->cv.notify_all();
        }
        if (s->preparer.joinable()) s->preparer.join();
        if (s->worker.joinable()) s->worker.join();
        for (hgdg::Solve* r : s->ready) {
            hgdg::destroy_ctx(r->ctx);
            delete r;
        }
        (void)hipSetDevice(s->device);
        s->drop_step_graph();
        s->drop_events();
        if (s->h_blob) (void)hipHostFree(s->h_blob);
        s->d_blob.reset();
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (s->prep_stream) (void)hipStreamDestroy(s->prep_stream);
        delete s;
    });
}

}  // extern "C"
