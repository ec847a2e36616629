// Taskflow<K> / Threadpool drop-in over the tasktorrent_amd C-ABI (header-only C++17).
//
// Mirrors the reference's parametrized-task-graph API
//   tasktorrent/src/taskflow.hpp:40-297          (Taskflow<K>: set_task, set_fulfill,
//                                                set_indegree, set_mapping, set_binding,
//                                                set_name, set_priority, fulfill_promise, name)
//   tasktorrent/src/threadpool_shared.hpp:79-126 (Threadpool: size, join, set_logger)
// with the same names, argument meaning and fluent setters, for K = int,
// std::array<int,2>, std::array<int,3> (or any hashable K).
//
// Difference that follows from running on the GPU: a task body is a device task
// descriptor (kind/precision/matrix slot/tile coordinates), returned by the
// set_task callback, instead of a host lambda. fulfill_promise calls made by the
// set_fulfill callbacks are recorded on the host when the graph is compiled
// (first join()), so the out-edges of a task may depend on its index but not on
// computed data -- the case of the dense Cholesky PTG. The compiled graph lives in
// HBM; the persistent executor decrements the dependency counters with device
// atomics. Later join()s relaunch it. Errors that the reference turns into an
// assert (count < 0, taskflow.hpp:278-282) or a hang (a task whose promises are
// never all fulfilled) are reported by join() as TTAD_ERR_GRAPH with the task name.
#pragma once
#include <array>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../tasktorrent_amd.h"

namespace ttad {

// What a task body is on the device (replaces f_run).
struct DeviceTask {
    int kind = TTAD_GEMM;
    int prec = TTAD_F64;
    int mat = 0;
    int i = 0, j = 0, k = 0;
};

// FNV-1a over the 4N bytes of std::array<int,N> (tasktorrent/src/hashes.hpp:27-57).
template <class K> struct HashKey {
    size_t operator()(const K& k) const { return std::hash<K>{}(k); }
};
template <size_t N> struct HashKey<std::array<int, N>> {
    size_t operator()(const std::array<int, N>& t) const {
        uint32_t h = 0x811C9DC5u;
        const unsigned char* p = reinterpret_cast<const unsigned char*>(t.data());
        for (size_t b = 0; b < 4 * N; b++) h = (p[b] ^ h) * 0x01000193u;
        return h;
    }
};

class Threadpool;

class TaskflowBase {
  public:
    virtual ~TaskflowBase() = default;
    virtual void expand(int local) = 0;              // run f_fulfill of a discovered task
    virtual DeviceTask descriptor(int local) = 0;    // run f_task
    virtual int indegree(int local) = 0;
    virtual double priority(int local) = 0;
    virtual std::string task_name(int local) = 0;
    virtual int mapping_of(int local) = 0;
    virtual bool binding_of(int local) = 0;
};

class Threadpool {
  public:
    // n_workers <= 0: one persistent workgroup per resident slot of the device.
    explicit Threadpool(ttad_context* ctx, int n_workers = 0, int verb = 0, std::string basename = "Wk_")
        : ctx_(ctx), n_workers_(n_workers), verb_(verb), basename_(std::move(basename)) {}
    ~Threadpool() {
        if (graph_) ttad_graph_destroy(graph_);
    }
    Threadpool(const Threadpool&) = delete;
    Threadpool& operator=(const Threadpool&) = delete;

    int size() const { return n_workers_; }
    // Matrix slot used by DeviceTask::mat.
    void bind(int slot, ttad_matrix* M) {
        if ((int)mats_.size() <= slot) mats_.resize(slot + 1, nullptr);
        mats_[slot] = M;
    }
    void set_trace(bool on) { trace_ = on; }

    // Compiles the graph (first call), seeds, runs it to completion on the device.
    int join(double timeout_s = 0.0, ttad_graph_stats* stats = nullptr) {
        int rc = compiled_ ? TTAD_OK : compile();
        if (rc != TTAD_OK) return rc;
        for (size_t s = 0; s < mats_.size(); s++)
            if (mats_[s] && (rc = ttad_graph_bind(graph_, (int)s, mats_[s])) != TTAD_OK) return rc;
        if ((rc = ttad_graph_launch(graph_, n_workers_, trace_ ? 1 : 0, nullptr)) != TTAD_OK) return rc;
        return ttad_graph_wait(graph_, timeout_s, stats);
    }
    // Compile without running (introspection, benchmarking of the launch alone).
    int compile() {
        if (compiled_) return TTAD_OK;
        while (!work_.empty()) {
            const int id = work_.front();
            work_.pop_front();
            current_ = id;
            nodes_[id].flow->expand(nodes_[id].local);
            current_ = -1;
            if (!error_.empty()) break;
        }
        if (error_.empty())
            for (auto& nd : nodes_)
                if (nd.fulfilled != nd.indegree) {
                    error_ = "task " + nd.flow->task_name(nd.local) + " received " + std::to_string(nd.fulfilled) +
                             " of " + std::to_string(nd.indegree) + " promises (would never run)";
                    break;
                }
        if (!error_.empty()) {
            if (verb_ > 0) fprintf(stderr, "ttad::Threadpool: %s\n", error_.c_str());
            return TTAD_ERR_GRAPH;
        }
        const int n = (int)nodes_.size();
        // Mapping and binding (threadpool_shared.cpp:200-224): the reference pins a task to the
        // thread `mapping` names, and a bound task runs only on that thread, so that bound tasks of
        // one thread never overlap (its accumulate mode binds ACCU tasks, 2d_cholesky.cpp:596-601).
        // Device workers are not threads: every worker serves every queue. A mapping is therefore only
        // checked (a negative thread is an error, as the reference's assert), and binding is accepted
        // only for accumulate-mode updates (TTAD_GEMM_ACC / TTAD_SYRK_ACC), whose additions into a tile
        // are already serialized by the tile's locks; any other bound task is refused, not ignored.
        for (int t = 0; t < n && error_.empty(); t++) {
            const Node& nd = nodes_[t];
            if (nd.flow->mapping_of(nd.local) < 0) {
                error_ = "mapping < 0 for " + nd.flow->task_name(nd.local);
                if (verb_ > 0) fprintf(stderr, "ttad::Threadpool: %s\n", error_.c_str());
                return TTAD_ERR_INVALID;
            }
            if (nd.flow->binding_of(nd.local)) {
                const int kind = nd.flow->descriptor(nd.local).kind;
                if (kind != TTAD_GEMM_ACC && kind != TTAD_SYRK_ACC) {
                    error_ = "binding = true for " + nd.flow->task_name(nd.local) +
                             ": device workers are not threads; only accumulate-mode updates may be bound";
                    if (verb_ > 0) fprintf(stderr, "ttad::Threadpool: %s\n", error_.c_str());
                    return TTAD_ERR_INVALID;
                }
            }
        }
        desc_.assign(n, ttad_task_desc{});
        indeg_.assign(n, 0);
        out_ptr_.assign(n + 1, 0);
        out_idx_.clear();
        prio_.assign(n, 0.0);
        seeds_.assign(n, 0);
        for (int t = 0; t < n; t++) {
            const DeviceTask d = nodes_[t].flow->descriptor(nodes_[t].local);
            desc_[t] = ttad_task_desc{d.kind, d.prec, d.mat, d.i, d.j, d.k};
            indeg_[t] = nodes_[t].indegree;
            prio_[t] = nodes_[t].flow->priority(nodes_[t].local);
            seeds_[t] = nodes_[t].host_seeds;
            out_ptr_[t + 1] = out_ptr_[t] + (int32_t)nodes_[t].out.size();
            for (int s : nodes_[t].out) out_idx_.push_back(s);
        }
        if (ctx_) {  // no context: a dry compile that only records the graph (introspection)
            int rc = ttad_graph_create(ctx_, n, desc_.data(), indeg_.data(), out_ptr_.data(),
                                       out_idx_.empty() ? nullptr : out_idx_.data(), prio_.data(), &graph_);
            if (rc != TTAD_OK) return rc;
            for (int t = 0; t < n; t++)
                for (int s = 0; s < seeds_[t]; s++)
                    if ((rc = ttad_graph_seed(graph_, t)) != TTAD_OK) return rc;
        }
        compiled_ = true;
        return TTAD_OK;
    }

    ttad_graph* graph() const { return graph_; }
    // The compiled graph in CSR form (valid after compile()).
    const std::vector<ttad_task_desc>& task_descs() const { return desc_; }
    const std::vector<int32_t>& indegrees() const { return indeg_; }
    const std::vector<int32_t>& out_ptr() const { return out_ptr_; }
    const std::vector<int32_t>& out_idx() const { return out_idx_; }
    const std::vector<double>& priorities() const { return prio_; }
    const std::vector<int32_t>& host_seeds() const { return seeds_; }
    // Hands the compiled graph to the caller (who then owns and destroys it).
    ttad_graph* release_graph() {
        ttad_graph* g = graph_;
        graph_ = nullptr;
        return g;
    }
    int n_tasks() const { return (int)nodes_.size(); }
    std::string task_name(int id) const { return nodes_[id].flow->task_name(nodes_[id].local); }
    const std::vector<int>& successors(int id) const { return nodes_[id].out; }
    const std::string& error() const { return error_; }
    const std::string& basename() const { return basename_; }
    int write_trace(const char* path) { return graph_ ? ttad_graph_write_trace(graph_, path, basename_.c_str()) : TTAD_ERR_INVALID; }

  private:
    template <class K> friend class Taskflow;
    struct Node {
        TaskflowBase* flow;
        int local;
        int indegree;
        int fulfilled = 0;
        int host_seeds = 0;
        std::vector<int> out;
    };
    ttad_context* ctx_;
    int n_workers_;
    int verb_;
    std::string basename_;
    bool trace_ = false;
    bool compiled_ = false;
    ttad_graph* graph_ = nullptr;
    std::vector<ttad_matrix*> mats_;
    std::vector<Node> nodes_;
    std::deque<int> work_;
    int current_ = -1;  // node whose f_fulfill is being expanded; -1 = host
    std::string error_;
    std::vector<ttad_task_desc> desc_;
    std::vector<int32_t> indeg_, out_ptr_, out_idx_, seeds_;
    std::vector<double> prio_;

    int add_node(TaskflowBase* f, int local, int indegree) {
        nodes_.push_back(Node{f, local, indegree});
        return (int)nodes_.size() - 1;
    }
    // One promise on node id (taskflow.hpp:241-296, recorded instead of executed).
    void promise(int id) {
        if (compiled_) {
            error_ = "fulfill_promise after the graph was compiled";
            return;
        }
        Node& nd = nodes_[id];
        if (current_ < 0) nd.host_seeds++;
        else nodes_[current_].out.push_back(id);
        if (++nd.fulfilled > nd.indegree) {
            error_ = "count < 0 for " + nd.flow->task_name(nd.local);
            return;
        }
        if (nd.fulfilled == nd.indegree) work_.push_back(id);
    }
};

template <class K> class Taskflow : public TaskflowBase {
  public:
    Taskflow(Threadpool* tp, int verb = 0) : tp_(tp), verb_(verb) {}

    Taskflow& set_task(std::function<DeviceTask(K)> f) { f_task_ = std::move(f); return *this; }
    Taskflow& set_fulfill(std::function<void(K)> f) { f_fulfill_ = std::move(f); return *this; }
    Taskflow& set_mapping(std::function<int(K)> f) { f_mapping_ = std::move(f); return *this; }
    Taskflow& set_binding(std::function<bool(K)> f) { f_binding_ = std::move(f); return *this; }
    Taskflow& set_indegree(std::function<int(K)> f) { f_indegree_ = std::move(f); return *this; }
    Taskflow& set_name(std::function<std::string(K)> f) { f_name_ = std::move(f); return *this; }
    Taskflow& set_priority(std::function<double(K)> f) { f_prio_ = std::move(f); return *this; }
    std::string name(K k) { return f_name_(k); }

    // Host seed (no task running) or an out-edge of the task being expanded.
    void fulfill_promise(K k) {
        auto it = ids_.find(k);
        int id;
        if (it == ids_.end()) {
            const int local = (int)keys_.size();
            keys_.push_back(k);
            const int indeg = f_indegree_(k);
            if (indeg < 1) {
                tp_->error_ = "indegree < 1 for " + f_name_(k);
                return;
            }
            id = tp_->add_node(this, local, indeg);
            ids_.emplace(k, id);
            node_of_.push_back(id);
        } else {
            id = it->second;
        }
        tp_->promise(id);
    }

    // TaskflowBase
    void expand(int local) override { f_fulfill_(keys_[local]); }
    DeviceTask descriptor(int local) override { return f_task_(keys_[local]); }
    int indegree(int local) override { return f_indegree_(keys_[local]); }
    double priority(int local) override { return f_prio_(keys_[local]); }
    std::string task_name(int local) override { return f_name_(keys_[local]); }
    // Mapping/binding choose a CPU thread in the reference; Threadpool::compile checks them
    // (the device has no threads to pin to).
    int mapping(K k) { return f_mapping_(k); }
    bool binding(K k) { return f_binding_(k); }
    int mapping_of(int local) override { return f_mapping_(keys_[local]); }
    bool binding_of(int local) override { return f_binding_(keys_[local]); }

  private:
    Threadpool* tp_;
    int verb_;
    std::unordered_map<K, int, HashKey<K>> ids_;
    std::vector<K> keys_;
    std::vector<int> node_of_;
    std::function<DeviceTask(K)> f_task_ = [](K) { return DeviceTask{}; };
    std::function<void(K)> f_fulfill_ = [](K) {};
    std::function<int(K)> f_mapping_ = [](K) { return 0; };
    std::function<bool(K)> f_binding_ = [](K) { return false; };
    std::function<int(K)> f_indegree_ = [](K) { return 1; };
    std::function<std::string(K)> f_name_ = [](K) { return std::string("_"); };
    std::function<double(K)> f_prio_ = [](K) { return 0.0; };
};

}  // namespace ttad
