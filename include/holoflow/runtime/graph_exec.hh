// Execution side of the holoflow runtime, HIP-typed.
// Mirrors src/holoflow/include/holoflow/runtime/graph_exec.hh:34-266: the compiled graph
// (NodePlan, EdgePlan, GraphPlan), execution resources (memory blocks, storages, tensor
// descriptors, streams, tasks), sections, and the Scheduler that runs each section on a
// thread of its own: acquire owned inputs -> pop async consumers -> run sync nodes in
// topological order -> synchronise the section's stream -> push async producers -> release
// owned outputs (graph_exec.cc:320-433).
//
// Differences by design: the graph is a plain adjacency list; streams are HIP streams of
// default priority; a section that throws stops the scheduler and its message is kept for
// wait()/error() instead of aborting the process; step() runs every ready section once on
// the calling thread, in section dependency order, for deterministic drivers and tests.
// The UI event router of the reference (holoflow_event) is not part of this build.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "curaii/hip.hh"
#include "holoflow/core/graph_spec.hh"
#include "holoflow/core/tasks.hh"
#include "holoflow/core/tensor.hh"

namespace holoflow::runtime {

struct NodePlan {
  core::NodeSpec spec;
  core::InferResult infer;
  std::vector<int> in_tids;   // tensor id of each input
  std::vector<int> out_tids;  // tensor id of each output
};

struct EdgePlan {
  core::EdgeSpec spec;
  core::TDesc desc;
  int tid = -1;
};

struct GraphPlan {
  struct Edge {
    size_t src = 0, dst = 0;
    EdgePlan plan;
  };
  std::vector<NodePlan> nodes;
  std::vector<Edge> edges;
  std::vector<std::vector<size_t>> in_edges, out_edges;  // edge indices per node
  std::vector<size_t> topo;                             // nodes in topological order
};

struct MemoryBlock {
  core::MemLoc mem_loc = core::MemLoc::Device;
  size_t size_bytes = 0;
  curaii::unique_host_ptr<std::byte> h_data;    // pinned host memory
  curaii::unique_device_ptr<std::byte> d_data;  // device memory
  void *get();
};

struct ExecResources {
  std::map<size_t, MemoryBlock> memory_blocks;                  // storage id -> block
  std::map<size_t, std::unique_ptr<core::Storage>> storages;    // storage id -> storage
  std::map<size_t, core::TDesc> tensor_descs;                   // tensor id -> descriptor
  std::map<size_t, size_t> tid_to_sid;                          // tensor id -> storage id
  std::map<std::string, std::unique_ptr<core::IOStorageAccess>> node_storage_adapters;
  std::map<size_t, std::unique_ptr<curaii::HipStream>> streams;  // section id -> stream
  std::map<std::string, std::unique_ptr<core::ITask>> tasks;    // node name -> task
};

struct Section {
  int id = 0;
  std::string name;
  hipStream_t stream = nullptr;
  std::vector<size_t> sync_topo;   // synchronous nodes in topological order
  std::vector<size_t> async_cons;  // asynchronous nodes this section pops from
  std::vector<size_t> async_prod;  // asynchronous nodes this section pushes into
};

struct NodeMetrics {
  double average_duration_ms = 0.0;
  double runs_per_second = 0.0;
  double host_throughput_bytes_per_second = 0.0;
  double device_throughput_bytes_per_second = 0.0;
  uint64_t sample_count = 0;
};

class Scheduler {
 public:
  Scheduler(const GraphPlan &graph, const std::vector<Section> &sections, ExecResources &res);
  ~Scheduler();
  Scheduler(const Scheduler &) = delete;
  Scheduler &operator=(const Scheduler &) = delete;

  void start();
  void request_stop();
  // Joins the section threads; rethrows the first section error as std::runtime_error.
  void wait();
  bool is_running() const { return running_.load(); }
  bool stop_requested() const { return stop_.load(); }

  // Runs every section whose inputs are ready once, in dependency order, on this thread;
  // returns how many sections ran.  Not allowed while the threads run.
  int step();

  std::map<std::string, NodeMetrics> metrics() const;

 private:
  struct NodeRt {
    core::ITask *task = nullptr;
    core::ISyncTask *sync = nullptr;
    core::IAsyncTask *async = nullptr;
    std::vector<core::TView> in_views, out_views;
    // Once mode: an async consumer popped by a section iteration that stopped before it ran
    // (a later consumer was not ready); its output views stay valid and unreleased, and the
    // next iteration uses them instead of popping again
    bool held = false;
  };
  struct Acc {
    std::atomic<uint64_t> duration_ns{0}, runs{0}, host_bytes{0}, device_bytes{0};
  };
  enum class Mode { Blocking, Once };

  void run_section_loop(int section_id);
  // One iteration of a section; false when (Once mode) an input was not ready.
  bool run_section_once(const Section &sec, Mode mode);
  bool acquire_owned_inputs(size_t v, Mode mode);
  void release_owned_outputs(size_t v);
  core::OpResult run_sync(size_t v);
  core::OpResult run_async_cons(size_t v, Mode mode);
  core::OpResult run_async_prod(size_t v, Mode mode);
  void record(size_t v, uint64_t ns, const std::vector<core::TView> &a, const std::vector<core::TView> *b);
  void fail(const std::string &msg);

  const GraphPlan &graph_;
  const std::vector<Section> &sections_;
  ExecResources &res_;
  std::vector<core::TView> tviews_;  // per tensor id, sharing the compiler's Storage objects
  std::vector<NodeRt> rts_;
  std::vector<std::unique_ptr<Acc>> acc_;
  std::vector<int> section_order_;
  std::atomic<bool> running_{false}, stop_{false};
  std::atomic<bool> cancelled_{false};
  std::vector<std::thread> threads_;
  mutable std::mutex err_mu_;
  std::string error_;
  std::chrono::steady_clock::time_point t_start_;
};

}  // namespace holoflow::runtime
