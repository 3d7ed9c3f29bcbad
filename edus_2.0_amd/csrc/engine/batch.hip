// Batches of contexts advanced together (edus_batch_* in include/edus_gpu.h).
//
// No counterpart in the reference, which runs one simulation per process.
// Members with the same shape (nb, grid, layout, precision, exchange,
// filledbands) form a group: one Engine whose per-k arrays stack the members
// along the plane axis, so every kernel launch advances the whole group and
// small instances share the chip instead of queueing behind each other.  Each
// group has its own stream; the fields of all members for a chunk of steps are
// written to device tables up front (E at the three RK4 stage times and the
// gradient coefficients c_j / N_j from each member's own lattice).
//
// Graph mode captures S steps of every group -- each group on its own stream,
// forked from and joined back to the batch stream -- into one hipGraph, which
// is then replayed once per S steps after the next S steps' field tables are
// copied in.  As with direct launches, gradient axes and E.xi terms whose
// field is zero throughout a call are skipped; skipping or not gives the same
// numbers (a zero field adds exact zeros).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "engine_internal.hpp"

#define HIP_CHECK EDUS_HIP_CHECK

namespace edus {

namespace {

// Members that can share one stacked engine: same shape and arithmetic, so
// that each ends exactly where it ends alone -- nb, grid, layout, precision,
// exchange, filled bands, gradient space (and, for the k-space gradient, the
// same stencil, which depends on the lattice), commutator engine,
// spectral-gradient form, the storage the commutator reads H0 and xi from and
// whether the exchange term's forward row pass comes from the gradient's (the
// last two forms agree only to rounding).
bool same_group(const Engine& a, const Engine& b)
{
    if (a.nb != b.nb || a.N0 != b.N0 || a.N1 != b.N1 || a.layout != b.layout || a.prec != b.prec ||
        a.coulomb != b.coulomb || a.nf != b.nf || a.kgrad != b.kgrad || a.comm_kind != b.comm_kind ||
        a.grad_fused != b.grad_fused || a.tri != b.tri || a.hf_fuse != b.hf_fuse)
        return false;
    if (!a.kgrad) return true;
    if (a.kst.size() != b.kst.size()) return false;
    for (size_t i = 0; i < a.kst.size(); ++i)
        if (std::memcmp(&a.kst[i], &b.kst[i], sizeof(KStencilPoint)) != 0) return false;
    return true;
}

}  // namespace

Batch::Batch(int nmem, edus_ctx* const* ctxs) : n(nmem)
{
    if (n < 1) throw EngineError(EDUS_ERR_ARG, "a batch needs at least one member");
    std::vector<std::vector<const Engine*>> gm;
    std::vector<std::vector<int>> gidx;
    where.resize(n);
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i] || !ctxs[i]->e) throw EngineError(EDUS_ERR_ARG, "null batch member");
        const Engine& e = *ctxs[i]->e;
        if (e.B != 1) throw EngineError(EDUS_ERR_ARG, "batch members must be single contexts");
        ctx.push_back(ctxs[i]);
        int g = -1;
        for (size_t q = 0; q < gm.size() && g < 0; ++q)
            if (same_group(*gm[q][0], e)) g = int(q);
        if (g < 0) {
            g = int(gm.size());
            gm.emplace_back();
            gidx.emplace_back();
        }
        where[i] = {g, int(gm[g].size())};
        gm[g].push_back(&e);
        gidx[g].push_back(i);
    }
    for (size_t g = 0; g < gm.size(); ++g) {
        BatchGroup bg;
        bg.own = std::make_unique<Engine>(gm[g]);
        bg.e = bg.own.get();
        bg.members = gidx[g];
        groups.push_back(std::move(bg));
    }
    HIP_CHECK(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    joins.resize(groups.size());
    for (auto& ev : joins) HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
}

Batch::Batch(Engine* self) : n(1)
{
    BatchGroup bg;
    bg.e = self;
    bg.members = {0};
    groups.push_back(std::move(bg));
    where = {{0, 0}};
    HIP_CHECK(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    joins.resize(1);
    HIP_CHECK(hipEventCreateWithFlags(&joins[0], hipEventDisableTiming));
}

Batch::Batch(std::unique_ptr<Engine> group) : n(group->B)
{
    BatchGroup bg;
    bg.own = std::move(group);
    bg.e = bg.own.get();
    for (int i = 0; i < n; ++i) {
        bg.members.push_back(i);
        where.push_back({0, i});
    }
    groups.push_back(std::move(bg));
    HIP_CHECK(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    joins.resize(1);
    HIP_CHECK(hipEventCreateWithFlags(&joins[0], hipEventDisableTiming));
}

double2* Batch::state(int i)
{
    const BatchGroup& g = groups[where[i].first];
    return g.e->y + size_t(where[i].second) * g.e->inst_elems();
}

void delete_batch(Batch* b) { delete b; }

void Engine::rk4_steps(double dt, int nsteps, const double* E)
{
    if (graph_steps <= 0 || B != 1) {
        for (int st = 0; st < nsteps; ++st) rk4_step(dt, E + 9 * size_t(st));
        return;
    }
    if (!self_batch) self_batch = new Batch(this);
    self_batch->graph_steps = graph_steps;
    self_batch->rk4_steps(dt, nsteps, E);  // E[s][j][1][3]: the same layout
}

Batch::~Batch()
{
    if (bs) (void)hipStreamSynchronize(bs);
    for (auto& g : groups) {
        if (g.e) (void)hipStreamSynchronize(g.e->s);
        if (g.host_tab) (void)hipHostFree(g.host_tab);
    }
    drop_graph();
    if (fork) (void)hipEventDestroy(fork);
    for (auto& ev : joins)
        if (ev) (void)hipEventDestroy(ev);
    if (bs) (void)hipStreamDestroy(bs);
}

void Batch::drop_graph()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_S = 0;
}

void Batch::sync_all()
{
    HIP_CHECK(hipStreamSynchronize(bs));
    for (auto& g : groups) HIP_CHECK(hipStreamSynchronize(g.e->s));
}

// Host tables of group g for all nsteps of the caller's field array
// E[((s * 3 + j) * n + member) * 3 + d]:
//   E part     [s][j][slot][3]
//   coef part  [s][j][slot][2]  c_a / N_a with c_a = sum_i A[i][a] E_i
// and the flags of gradient axes / E.xi terms that are non-zero somewhere.
void Batch::fill_tables(BatchGroup& g, int nsteps, const double* E, bool flags_axis[2], bool flags_xi[3])
{
    Engine& e = *g.e;
    const int B = e.B;
    const size_t ne = size_t(nsteps) * 3 * B * 3, nc = size_t(nsteps) * 3 * B * 2;
    if (g.host_cap < ne + nc) {
        if (g.host_tab) HIP_CHECK(hipHostFree(g.host_tab));
        g.host_tab = nullptr;
        g.host_cap = 0;
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&g.host_tab), (ne + nc) * sizeof(double)));
        g.host_cap = ne + nc;
    }
    g.host_steps = size_t(nsteps);
    double* te = g.host_tab;
    double* tc = g.host_tab + ne;
    for (int a = 0; a < 2; ++a) flags_axis[a] = false;
    for (int d = 0; d < 3; ++d) flags_xi[d] = false;
    for (int s = 0; s < nsteps; ++s)
        for (int j = 0; j < 3; ++j)
            for (int b = 0; b < B; ++b) {
                const double* src = E + ((size_t(s) * 3 + j) * n + g.members[b]) * 3;
                double* dst = te + ((size_t(s) * 3 + j) * B + b) * 3;
                for (int d = 0; d < 3; ++d) {
                    dst[d] = src[d];
                    flags_xi[d] = flags_xi[d] || src[d] != 0.0;
                }
                const auto& A = e.Ainst[b];
                for (int a = 0; a < 2; ++a) {
                    const double c =
                        (A[0 * 3 + a] * src[0] + A[1 * 3 + a] * src[1] + A[2 * 3 + a] * src[2]) / (a == 0 ? e.N0 : e.N1);
                    tc[((size_t(s) * 3 + j) * B + b) * 2 + a] = c;
                    flags_axis[a] = flags_axis[a] || c != 0.0;
                }
            }
}

// host steps [h0, h0 + cnt) -> device table steps [0, cnt)
void Batch::upload_tables(BatchGroup& g, size_t h0, int cnt, hipStream_t st)
{
    Engine& e = *g.e;
    const size_t B = size_t(e.B);
    const double* te = g.host_tab;
    const double* tc = g.host_tab + g.host_steps * 3 * B * 3;
    HIP_CHECK(hipMemcpyAsync(e.E_dev, te + h0 * 3 * B * 3, size_t(cnt) * 3 * B * 3 * sizeof(double),
                             hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(e.coef_dev, tc + h0 * 3 * B * 2, size_t(cnt) * 3 * B * 2 * sizeof(double),
                             hipMemcpyHostToDevice, st));
}

void Batch::fork_from(hipStream_t origin)
{
    HIP_CHECK(hipEventRecord(fork, origin));
    for (auto& g : groups) HIP_CHECK(hipStreamWaitEvent(g.e->s, fork, 0));
}

void Batch::join_to(hipStream_t origin)
{
    for (size_t i = 0; i < groups.size(); ++i) {
        HIP_CHECK(hipEventRecord(joins[i], groups[i].e->s));
        HIP_CHECK(hipStreamWaitEvent(origin, joins[i], 0));
    }
}

// One graph holding S RK4 steps of every group, table steps 0 .. S-1; the
// gradient axes and E.xi terms in `flags` are the ones the graph evaluates.
void Batch::capture(double dt, int S, const std::vector<std::array<bool, 5>>& flags)
{
    drop_graph();
    std::vector<bool> timing(groups.size());
    for (size_t i = 0; i < groups.size(); ++i) {
        groups[i].e->ensure_tables(S);
        timing[i] = groups[i].e->timing;  // no per-kernel events inside a graph
        groups[i].e->timing = false;
    }
    HIP_CHECK(hipStreamSynchronize(bs));
    for (auto& g : groups) HIP_CHECK(hipStreamSynchronize(g.e->s));
    HIP_CHECK(hipStreamBeginCapture(bs, hipStreamCaptureModeThreadLocal));
    try {
        fork_from(bs);
        for (int st = 0; st < S; ++st)
            for (size_t gi = 0; gi < groups.size(); ++gi) {
                const bool ua[2] = {flags[gi][0], flags[gi][1]};
                const bool ux[3] = {flags[gi][2], flags[gi][3], flags[gi][4]};
                groups[gi].e->rk4_step_tab(dt, st, ua, ux);
            }
        join_to(bs);
    } catch (...) {
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(bs, &broken);
        if (broken) (void)hipGraphDestroy(broken);
        for (size_t i = 0; i < groups.size(); ++i) groups[i].e->timing = timing[i];
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(bs, &graph));
    HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (size_t i = 0; i < groups.size(); ++i) groups[i].e->timing = timing[i];
    graph_dt = dt;
    graph_S = S;
    graph_flags = flags;
    for (auto& g : groups) g.cap_E = g.e->E_dev;
}

void Batch::rk4_steps(double dt, int nsteps, const double* E)
{
    if (nsteps <= 0) return;
    std::vector<std::array<bool, 5>> flags(groups.size());
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        bool fa[2], fx[3];
        fill_tables(groups[gi], nsteps, E, fa, fx);
        flags[gi] = {fa[0], fa[1], fx[0], fx[1], fx[2]};
    }
    // work queued on the groups' own streams (a state copied in, earlier
    // direct steps) comes before anything on the batch stream
    join_to(bs);
    int done = 0;
    const int S = graph_steps;
    if (S > 0 && nsteps >= S) {
        // the graph evaluates the gradient axes and E.xi terms that are non-zero
        // in this call; a call that needs others (e.g. the pulse switching on)
        // captures again
        bool stale = exec == nullptr || graph_S != S || graph_dt != dt || graph_flags != flags;
        for (auto& g : groups) stale = stale || g.cap_E != g.e->E_dev;
        if (stale) capture(dt, S, flags);
        for (; done + S <= nsteps; done += S) {
            // the graph reads table steps 0..S-1: copy this chunk's fields in
            // behind the previous replay on the same stream
            for (auto& g : groups) upload_tables(g, size_t(done), S, bs);
            HIP_CHECK(hipGraphLaunch(exec, bs));
        }
    }
    if (done < nsteps) {
        // direct launches, each group on its own stream
        const int rem = nsteps - done;
        fork_from(bs);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            BatchGroup& g = groups[gi];
            g.e->ensure_tables(rem);
            upload_tables(g, size_t(done), rem, g.e->s);
        }
        for (int st = 0; st < rem; ++st)
            for (size_t gi = 0; gi < groups.size(); ++gi) {
                const bool ua[2] = {flags[gi][0], flags[gi][1]};
                const bool ux[3] = {flags[gi][2], flags[gi][3], flags[gi][4]};
                groups[gi].e->rk4_step_tab(dt, st, ua, ux);
            }
        join_to(bs);
    }
    for (auto& g : groups) g.e->state_dirty = true;  // replays wrote states no host call saw
    // the pinned staging is rewritten by the next call
    sync_all();
    HIP_CHECK(hipGetLastError());
}

void Batch::copy_state(int i, bool to_member)
{
    if (i < 0 || i >= n) throw EngineError(EDUS_ERR_ARG, "batch member index out of range");
    BatchGroup& g = groups[where[i].first];
    Engine& m = *ctx[i]->e;
    const size_t bytes = g.e->inst_elems() * sizeof(double2);
    double2* slot = g.e->y + size_t(where[i].second) * g.e->inst_elems();
    sync_all();
    HIP_CHECK(hipStreamSynchronize(m.s));
    (to_member ? m : *g.e).check_final_state();  // the state being replaced is checked first
    if (to_member)
        HIP_CHECK(hipMemcpyAsync(m.y, slot, bytes, hipMemcpyDeviceToDevice, g.e->s));
    else
        HIP_CHECK(hipMemcpyAsync(slot, m.y, bytes, hipMemcpyDeviceToDevice, g.e->s));
    HIP_CHECK(hipStreamSynchronize(g.e->s));
}

void Batch::check(double* trace_dev, double* herm_dev)
{
    sync_all();
    double mt = 0.0, mh = 0.0;
    for (auto& g : groups) {
        double v[2];
        g.e->read_monitor(v);
        mt = std::max(mt, v[0]);
        mh = std::max(mh, v[1]);
    }
    if (trace_dev) *trace_dev = mt;
    if (herm_dev) *herm_dev = mh;
}

}  // namespace edus

struct edus_batch {
    edus::Batch* b;
};

using edus::guarded;

extern "C" {

edus_batch* edus_batch_create(int n, edus_ctx* const* members)
{
    if (!members || n < 1) {
        edus::set_error("edus_batch_create: no members");
        return nullptr;
    }
    edus_batch* b = nullptr;
    const int rc = guarded([&] { b = new edus_batch{new edus::Batch(n, members)}; });
    return rc == EDUS_OK ? b : nullptr;
}

int edus_batch_destroy(edus_batch* b)
{
    if (!b) return EDUS_OK;
    const int rc = guarded([&] { delete b->b; });
    delete b;
    return rc;
}

int edus_batch_groups(edus_batch* b, int* ngroups, int* member_group)
{
    EDUS_REQUIRE(b);
    if (ngroups) *ngroups = int(b->b->groups.size());
    if (member_group)
        for (int i = 0; i < b->b->n; ++i) member_group[i] = b->b->where[i].first;
    return EDUS_OK;
}

int edus_batch_graph(edus_batch* b, int steps_per_graph)
{
    EDUS_REQUIRE(b);
    if (steps_per_graph < 0) {
        edus::set_error("steps_per_graph < 0");
        return EDUS_ERR_ARG;
    }
    b->b->graph_steps = steps_per_graph;
    return EDUS_OK;
}

int edus_batch_rk4_steps(edus_batch* b, double dt, int nsteps, const double* E)
{
    EDUS_REQUIRE(b && (nsteps <= 0 || E));
    if (nsteps < 0) {
        edus::set_error("nsteps < 0");
        return EDUS_ERR_ARG;
    }
    return guarded([&] { b->b->rk4_steps(dt, nsteps, E); });
}

int edus_batch_store(edus_batch* b, int i)
{
    EDUS_REQUIRE(b);
    return guarded([&] { b->b->copy_state(i, true); });
}

int edus_batch_load(edus_batch* b, int i)
{
    EDUS_REQUIRE(b);
    return guarded([&] { b->b->copy_state(i, false); });
}

int edus_batch_observables(edus_batch* b, int i, double* pop_bloch, double* pop_wannier, double* vel)
{
    EDUS_REQUIRE(b);
    return guarded([&] {
        if (i < 0 || i >= b->b->n) throw edus::EngineError(EDUS_ERR_ARG, "batch member index out of range");
        b->b->sync_all();
        auto& g = b->b->groups[b->b->where[i].first];
        g.e->observables(pop_bloch, pop_wannier, vel, b->b->where[i].second);
    });
}

int edus_batch_check(edus_batch* b, double* trace_dev, double* herm_dev)
{
    EDUS_REQUIRE(b);
    return guarded([&] { b->b->check(trace_dev, herm_dev); });
}

int edus_batch_synchronize(edus_batch* b)
{
    EDUS_REQUIRE(b);
    return guarded([&] { b->b->sync_all(); });
}

}  // extern "C"
