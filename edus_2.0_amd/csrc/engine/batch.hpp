// Batches of engine instances advanced together (see batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <utility>
#include <vector>

#include "engine_internal.hpp"

namespace edus {

struct BatchGroup {
    Engine* e = nullptr;
    std::unique_ptr<Engine> own;  // null when the group borrows a context's engine
    std::vector<int> members;    // batch member index of each stacked instance
    double* host_tab = nullptr;  // pinned staging: E then coef tables of one call
    size_t host_cap = 0;         // doubles
    size_t host_steps = 0;       // steps held by the staging of the current call
    const double* cap_E = nullptr;  // device tables the graph was captured with
};

class Batch {
  public:
    Batch(int n, edus_ctx* const* ctxs);
    explicit Batch(Engine* self);  // one member, the engine itself (graph replay of a context)
    explicit Batch(std::unique_ptr<Engine> group);  // one stacked group, member i = instance i
    double2* state(int i);  // device state of member i
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    int n;
    std::vector<edus_ctx*> ctx;
    std::vector<BatchGroup> groups;
    std::vector<std::pair<int, int>> where;  // member -> (group, slot)
    hipStream_t bs = nullptr;                // origin stream of graph launches
    int graph_steps = 0;

    void rk4_steps(double dt, int nsteps, const double* E);
    void copy_state(int i, bool to_member);
    void check(double* trace_dev, double* herm_dev);
    void sync_all();
    // forget the captured graph (the members' kernels or tiles changed),
    // after the replays in flight have finished
    void invalidate()
    {
        sync_all();
        drop_graph();
    }

  private:
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_dt = 0.0;
    int graph_S = 0;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;

    void fill_tables(BatchGroup& g, int nsteps, const double* E, bool flags_axis[2], bool flags_xi[3]);
    void upload_tables(BatchGroup& g, size_t h0, int cnt, hipStream_t st);
    // flags[g] = {axis 0, axis 1, xi x, xi y, xi z} of group g used by the graph
    void capture(double dt, int S, const std::vector<std::array<bool, 5>>& flags);
    std::vector<std::array<bool, 5>> graph_flags;
    void fork_from(hipStream_t origin);
    void join_to(hipStream_t origin);
    void drop_graph();
};

}  // namespace edus
