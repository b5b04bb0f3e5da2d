// holoflow runtime: Compiler (GraphSpec -> plan, buffers, sections, streams, tasks) and
// Scheduler (section threads), plus the C ABI of include/hfx_graph.h.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <numeric>
#include <set>
#include <unordered_map>
#include <unordered_set>

#include "hfx.h"
#include "hfx_graph.h"
#include "hfx_nodes.h"
#include "holoflow/runtime/compiler.hh"

namespace hfx {
void set_last_error(const std::string &msg);  // libhfx's thread-local last error (api.hip)
}

namespace holoflow::runtime {

using core::MemLoc;
using core::OpResult;
using core::TaskKind;

void *MemoryBlock::get() {
  void *p = mem_loc == MemLoc::Host ? static_cast<void *>(h_data.get()) : static_cast<void *>(d_data.get());
  if (!p) throw std::logic_error("MemoryBlock pointer is null");
  return p;
}

// ===================================================================== Compiler
namespace {

class TaskStorageAdapter : public core::IOStorageAccess {
 public:
  TaskStorageAdapter(std::vector<int> in, std::vector<int> out, ExecResources &res)
      : in_(std::move(in)), out_(std::move(out)), res_(res) {}
  core::Storage &owned_input_storage(size_t i) override {
    if (i >= in_.size()) throw std::out_of_range("Input index out of range in TaskStorageAdapter");
    return *res_.storages.at(res_.tid_to_sid.at((size_t)in_[i]));
  }
  core::Storage &owned_output_storage(size_t i) override {
    if (i >= out_.size()) throw std::out_of_range("Output index out of range in TaskStorageAdapter");
    return *res_.storages.at(res_.tid_to_sid.at((size_t)out_[i]));
  }

 private:
  std::vector<int> in_, out_;
  ExecResources &res_;
};

struct Pass {
  const core::Registry &reg;
  const core::GraphSpec &spec;
  CompilerOutput *prev;
  CompilerOutput &out;
  std::unordered_map<std::string, int> node_section;

  GraphPlan &g() { return out.graph; }
  ExecResources &res() { return out.resources; }
  bool is_async(size_t v) { return g().nodes[v].infer.kind == TaskKind::Async; }

  void validate_spec() {
    std::unordered_set<std::string> names, dsts;
    for (const auto &n : spec.nodes) {
      if (!names.insert(n.name).second) throw CompilerException("Duplicate node name: '" + n.name + "'");
      if (!reg.sync_factory(n.kind) && !reg.async_factory(n.kind))
        throw CompilerException("Unknown node kind '" + n.kind + "'");
    }
    for (const auto &e : spec.edges) {
      if (e.src >= spec.nodes.size() || e.dst >= spec.nodes.size()) throw CompilerException("Edge endpoint out of range");
      const std::string label = spec.nodes[e.dst].name + ":" + std::to_string(e.spec.in_idx);
      if (!dsts.insert(label).second) throw CompilerException("Multiple edges targeting: " + label);
    }
  }

  void build_graph_structure() {
    auto &G = g();
    for (const auto &n : spec.nodes) G.nodes.push_back(NodePlan{n, {}, {}, {}});
    G.in_edges.assign(G.nodes.size(), {});
    G.out_edges.assign(G.nodes.size(), {});
    for (const auto &e : spec.edges) {
      G.out_edges[e.src].push_back(G.edges.size());
      G.in_edges[e.dst].push_back(G.edges.size());
      G.edges.push_back(GraphPlan::Edge{e.src, e.dst, EdgePlan{e.spec, {}, -1}});
    }
    // Kahn's algorithm, smallest index first among the ready nodes (deterministic)
    std::vector<size_t> indeg(G.nodes.size());
    for (const auto &e : G.edges) ++indeg[e.dst];
    std::set<size_t> ready;
    for (size_t v = 0; v < G.nodes.size(); ++v)
      if (!indeg[v]) ready.insert(v);
    while (!ready.empty()) {
      const size_t v = *ready.begin();
      ready.erase(ready.begin());
      G.topo.push_back(v);
      for (size_t ei : G.out_edges[v])
        if (--indeg[G.edges[ei].dst] == 0) ready.insert(G.edges[ei].dst);
    }
    if (G.topo.size() != G.nodes.size()) throw CompilerException("Graph contains a cycle (loop), which is not allowed.");
  }

  void run_type_inference() {
    auto &G = g();
    for (size_t v : G.topo) {
      auto &node = G.nodes[v];
      std::vector<core::TDesc> in(G.in_edges[v].size());
      for (size_t ei : G.in_edges[v]) {
        const auto &ep = G.edges[ei].plan;
        if (ep.spec.in_idx < 0 || (size_t)ep.spec.in_idx >= in.size())
          throw CompilerException("Input index out of bounds");
        in[(size_t)ep.spec.in_idx] = ep.desc;
      }
      try {
        node.infer = reg.factory(node.spec.kind).infer(in, node.spec.settings);
      } catch (const std::exception &ex) {
        throw CompilerException("Node '" + node.spec.name + "': " + ex.what());
      }
      for (size_t ei : G.out_edges[v]) {
        auto &ep = G.edges[ei].plan;
        if (ep.spec.out_idx < 0 || (size_t)ep.spec.out_idx >= node.infer.output_descs.size())
          throw CompilerException("Output index out of bounds");
        ep.desc = node.infer.output_descs[(size_t)ep.spec.out_idx];
      }
    }
  }

  void assign_tensor_ids() {
    auto &G = g();
    int next = 0;
    for (size_t v : G.topo) {
      auto &node = G.nodes[v];
      node.in_tids.assign(node.infer.input_descs.size(), -1);
      for (size_t ei : G.in_edges[v]) {
        const auto &ep = G.edges[ei].plan;
        if ((size_t)ep.spec.in_idx < node.in_tids.size()) node.in_tids[(size_t)ep.spec.in_idx] = ep.tid;
        res().tensor_descs[(size_t)ep.tid] = ep.desc;
      }
      for (size_t i = 0; i < node.in_tids.size(); ++i)
        if (node.in_tids[i] < 0)
          throw CompilerException("Node '" + node.spec.name + "': input " + std::to_string(i) + " is not connected");
      node.out_tids.assign(node.infer.output_descs.size(), -1);
      for (size_t i = 0; i < node.out_tids.size(); ++i) {
        const int tid = next++;
        node.out_tids[i] = tid;
        res().tensor_descs[(size_t)tid] = node.infer.output_descs[i];
        for (size_t ei : G.out_edges[v])
          if (G.edges[ei].plan.spec.out_idx == (int)i) G.edges[ei].plan.tid = tid;
      }
    }
  }

  void assign_storage_ids() {
    auto &G = g();
    auto &t2s = res().tid_to_sid;
    size_t next = 0;
    for (size_t v : G.topo) {
      const auto &node = G.nodes[v];
      for (size_t o = 0; o < node.out_tids.size(); ++o) {
        long long sid = -1;
        for (const auto &ip : node.infer.in_place) {
          if (ip.out_idx != (int)o) continue;
          const int in_tid = node.in_tids.at((size_t)ip.in_idx);
          auto it = t2s.find((size_t)in_tid);
          if (it == t2s.end())
            throw CompilerException("Node '" + node.spec.name + "': In-place input TID " + std::to_string(in_tid) +
                                    " has no Storage ID assigned.");
          sid = (long long)it->second;
          break;
        }
        t2s[(size_t)node.out_tids[o]] = sid < 0 ? next++ : (size_t)sid;
      }
    }
  }

  void verify_buffer_consistency() {
    std::map<size_t, int> owners;
    for (const auto &node : g().nodes) {
      for (size_t i = 0; i < node.infer.owned_inputs.size(); ++i)
        if (node.infer.owned_inputs[i]) ++owners[res().tid_to_sid.at((size_t)node.in_tids[i])];
      for (size_t i = 0; i < node.infer.owned_outputs.size(); ++i)
        if (node.infer.owned_outputs[i]) ++owners[res().tid_to_sid.at((size_t)node.out_tids[i])];
    }
    for (const auto &[sid, n] : owners)
      if (n > 1) throw CompilerException("Storage ID " + std::to_string(sid) + " has multiple owners");
  }

  void allocate_buffers() {
    auto &R = res();
    std::unordered_set<size_t> user_managed;
    for (const auto &node : g().nodes) {
      for (size_t i = 0; i < node.infer.owned_outputs.size(); ++i)
        if (node.infer.owned_outputs[i]) user_managed.insert(R.tid_to_sid.at((size_t)node.out_tids[i]));
      for (size_t i = 0; i < node.infer.owned_inputs.size(); ++i)
        if (node.infer.owned_inputs[i]) user_managed.insert(R.tid_to_sid.at((size_t)node.in_tids[i]));
    }
    std::map<size_t, size_t> rep;  // storage id -> first (producing) tensor id
    for (const auto &[tid, sid] : R.tid_to_sid)
      if (!rep.count(sid)) rep[sid] = tid;
    std::multimap<std::pair<int, size_t>, MemoryBlock> pool;  // reusable blocks of the previous graph
    if (prev)
      for (auto &[sid, b] : prev->resources.memory_blocks)
        pool.emplace(std::make_pair((int)b.mem_loc, b.size_bytes), std::move(b));
    for (const auto &[sid, tid] : rep) {
      const auto &d = R.tensor_descs.at(tid);
      auto st = std::make_unique<core::Storage>();
      st->mem_loc = d.mem_loc;
      st->bytes = d.num_bytes();
      st->ptr = nullptr;
      if (!user_managed.count(sid)) {
        MemoryBlock b;
        auto it = pool.find(std::make_pair((int)d.mem_loc, d.num_bytes()));
        if (it != pool.end()) {
          b = std::move(it->second);
          pool.erase(it);
        } else {
          b.mem_loc = d.mem_loc;
          b.size_bytes = d.num_bytes();
          if (d.mem_loc == MemLoc::Host)
            b.h_data = curaii::make_unique_host_ptr<std::byte>(b.size_bytes);
          else
            b.d_data = curaii::make_unique_device_ptr<std::byte>(b.size_bytes);
        }
        st->ptr = static_cast<std::byte *>(b.get());
        R.memory_blocks.emplace(sid, std::move(b));
      }
      R.storages.emplace(sid, std::move(st));
    }
  }

  void create_storage_adapters() {
    for (const auto &node : g().nodes)
      res().node_storage_adapters.emplace(node.spec.name,
                                          std::make_unique<TaskStorageAdapter>(node.in_tids, node.out_tids, res()));
  }

  // Checked once the kinds are known, before anything is allocated (the reference checks it
  // while partitioning, compiler.cc:771-785).
  void check_async_neighbours() {
    auto &G = g();
    for (const auto &e : G.edges)
      if (is_async(e.src) && is_async(e.dst))
        throw CompilerException("Consecutive asynchronous nodes '" + G.nodes[e.src].spec.name + "' and '" +
                                G.nodes[e.dst].spec.name + "' are not supported");
  }

  void partition_sections() {
    auto &G = g();
    const size_t n = G.nodes.size();
    std::vector<size_t> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](size_t i) {
      while (i != parent[i]) i = parent[i] = parent[parent[i]];
      return i;
    };
    auto unite = [&](size_t a, size_t b) { parent[find(a)] = find(b); };
    for (const auto &e : G.edges)
      if (!is_async(e.src) && !is_async(e.dst)) unite(e.src, e.dst);
    for (size_t v = 0; v < n; ++v) {
      if (!is_async(v)) continue;
      std::vector<size_t> preds, succs;
      for (size_t ei : G.in_edges[v])
        if (!is_async(G.edges[ei].src)) preds.push_back(G.edges[ei].src);
      for (size_t ei : G.out_edges[v])
        if (!is_async(G.edges[ei].dst)) succs.push_back(G.edges[ei].dst);
      for (size_t i = 1; i < preds.size(); ++i) unite(preds[0], preds[i]);
      for (size_t i = 1; i < succs.size(); ++i) unite(succs[0], succs[i]);
    }
    std::map<size_t, int> root_sec;
    auto section_of = [&](size_t v) {
      const size_t r = find(v);
      auto it = root_sec.find(r);
      if (it != root_sec.end()) return it->second;
      Section s;
      s.id = (int)out.sections.size();
      s.name = "section-" + std::to_string(s.id);
      out.sections.push_back(s);
      root_sec[r] = s.id;
      return s.id;
    };
    for (size_t v : G.topo) {
      if (is_async(v)) continue;
      const int s = section_of(v);
      out.sections[(size_t)s].sync_topo.push_back(v);
      node_section[G.nodes[v].spec.name] = s;
    }
    for (size_t v = 0; v < n; ++v) {
      if (!is_async(v)) continue;
      std::set<int> cons, prod;
      for (size_t ei : G.out_edges[v])
        if (!is_async(G.edges[ei].dst)) cons.insert(section_of(G.edges[ei].dst));
      for (size_t ei : G.in_edges[v])
        if (!is_async(G.edges[ei].src)) prod.insert(section_of(G.edges[ei].src));
      for (int s : cons) out.sections[(size_t)s].async_cons.push_back(v);
      for (int s : prod) out.sections[(size_t)s].async_prod.push_back(v);
    }
  }

  void assign_streams() {
    auto &streams = res().streams;
    std::deque<std::unique_ptr<curaii::HipStream>> old;
    if (prev)
      for (auto &[id, s] : prev->resources.streams) old.push_back(std::move(s));
    for (auto &sec : out.sections) {
      std::unique_ptr<curaii::HipStream> s;
      if (!old.empty()) {
        s = std::move(old.front());
        old.pop_front();
      } else {
        s = std::make_unique<curaii::HipStream>();
      }
      sec.stream = s->get();
      streams.emplace((size_t)sec.id, std::move(s));
    }
  }

  hipStream_t section_stream(size_t v) {
    auto it = node_section.find(g().nodes[v].spec.name);
    return it == node_section.end() ? nullptr : out.sections[(size_t)it->second].stream;
  }

  // The previous graph's task of the same name and kind, if any (compiler.cc:960-1040).
  std::unique_ptr<core::ITask> take_prev(const NodePlan &np) {
    if (!prev) return nullptr;
    bool same_kind = false;
    for (const auto &pn : prev->graph.nodes)
      if (pn.spec.name == np.spec.name) same_kind = pn.spec.kind == np.spec.kind;
    auto it = prev->resources.tasks.find(np.spec.name);
    if (!same_kind || it == prev->resources.tasks.end()) return nullptr;
    return std::move(it->second);
  }

  void instantiate_tasks() {
    auto &G = g();
    for (size_t v = 0; v < G.nodes.size(); ++v) {
      const auto &np = G.nodes[v];
      std::unique_ptr<core::ITask> old = take_prev(np);
      std::unique_ptr<core::ITask> task;
      try {
        if (!is_async(v)) {
          const core::SyncCreateCtx ctx{section_stream(v)};
          const auto *f = reg.sync_factory(np.spec.kind);
          std::unique_ptr<core::ISyncTask> o(dynamic_cast<core::ISyncTask *>(old.get()));
          if (o) (void)old.release();
          task = o ? f->update(std::move(o), np.infer.input_descs, np.spec.settings, ctx)
                   : f->create(np.infer.input_descs, np.spec.settings, ctx);
          if (ctx.stream) HIP_CHECK(hipStreamSynchronize(ctx.stream));
        } else {
          core::AsyncCreateCtx ctx{};
          for (size_t ei : G.in_edges[v])
            if (!is_async(G.edges[ei].src)) {
              ctx.producer_stream = section_stream(G.edges[ei].src);
              break;
            }
          for (size_t ei : G.out_edges[v])
            if (!is_async(G.edges[ei].dst)) {
              ctx.consumer_stream = section_stream(G.edges[ei].dst);
              break;
            }
          const auto *f = reg.async_factory(np.spec.kind);
          std::unique_ptr<core::IAsyncTask> o(dynamic_cast<core::IAsyncTask *>(old.get()));
          if (o) (void)old.release();
          task = o ? f->update(std::move(o), np.infer.input_descs, np.spec.settings, ctx)
                   : f->create(np.infer.input_descs, np.spec.settings, ctx);
          if (ctx.producer_stream) HIP_CHECK(hipStreamSynchronize(ctx.producer_stream));
          if (ctx.consumer_stream) HIP_CHECK(hipStreamSynchronize(ctx.consumer_stream));
        }
      } catch (const std::invalid_argument &ex) {
        throw CompilerException("Node '" + np.spec.name + "': " + ex.what());
      }
      res().tasks.emplace(np.spec.name, std::move(task));
    }
  }

  void bind_tasks() {
    for (const auto &np : g().nodes)
      res().tasks.at(np.spec.name)->bind_storage_access(res().node_storage_adapters.at(np.spec.name).get());
  }
};

}  // namespace

std::unique_ptr<CompilerOutput> Compiler::compile(const core::GraphSpec &spec, std::unique_ptr<CompilerOutput> prev) {
  auto out = std::make_unique<CompilerOutput>();
  Pass p{registry_, spec, prev.get(), *out, {}};
  p.validate_spec();
  p.build_graph_structure();
  p.run_type_inference();
  p.check_async_neighbours();
  p.assign_tensor_ids();
  p.assign_storage_ids();
  p.verify_buffer_consistency();
  p.allocate_buffers();
  p.create_storage_adapters();
  p.partition_sections();
  p.assign_streams();
  p.instantiate_tasks();
  p.bind_tasks();
  return out;
}

// ===================================================================== Scheduler
Scheduler::Scheduler(const GraphPlan &graph, const std::vector<Section> &sections, ExecResources &res)
    : graph_(graph), sections_(sections), res_(res) {
  size_t n_t = 0;
  for (const auto &kv : res_.tensor_descs) n_t = std::max(n_t, kv.first + 1);
  tviews_.resize(n_t);
  for (const auto &[tid, d] : res_.tensor_descs) tviews_[tid] = core::TView{d, res_.storages.at(res_.tid_to_sid.at(tid)).get()};
  rts_.resize(graph_.nodes.size());
  for (size_t v = 0; v < graph_.nodes.size(); ++v) {
    const auto &np = graph_.nodes[v];
    auto &rt = rts_[v];
    rt.task = res_.tasks.at(np.spec.name).get();
    rt.sync = dynamic_cast<core::ISyncTask *>(rt.task);
    rt.async = dynamic_cast<core::IAsyncTask *>(rt.task);
    for (int t : np.in_tids) rt.in_views.push_back(tviews_.at((size_t)t));
    for (int t : np.out_tids) rt.out_views.push_back(tviews_.at((size_t)t));
    acc_.push_back(std::make_unique<Acc>());
  }
  // section dependency order: producer sections of an async node before its consumers
  const size_t ns = sections_.size();
  std::vector<std::set<int>> succ(ns);
  std::vector<int> indeg(ns, 0);
  std::map<size_t, std::vector<int>> prod_of;
  for (const auto &s : sections_)
    for (size_t v : s.async_prod) prod_of[v].push_back(s.id);
  for (const auto &s : sections_)
    for (size_t v : s.async_cons)
      for (int p : prod_of[v])
        if (p != s.id && succ[(size_t)p].insert(s.id).second) ++indeg[(size_t)s.id];
  std::set<int> ready;
  for (size_t i = 0; i < ns; ++i)
    if (!indeg[i]) ready.insert((int)i);
  while (!ready.empty()) {
    const int s = *ready.begin();
    ready.erase(ready.begin());
    section_order_.push_back(s);
    for (int t : succ[(size_t)s])
      if (--indeg[(size_t)t] == 0) ready.insert(t);
  }
  for (size_t i = 0; i < ns; ++i)  // sections in a dependency cycle run in id order
    if (std::find(section_order_.begin(), section_order_.end(), (int)i) == section_order_.end())
      section_order_.push_back((int)i);
}

Scheduler::~Scheduler() {
  request_stop();
  for (auto &t : threads_)
    if (t.joinable()) t.join();
}

void Scheduler::fail(const std::string &msg) {
  {
    std::lock_guard<std::mutex> lk(err_mu_);
    if (error_.empty()) error_ = msg;
  }
  stop_.store(true);
}

void Scheduler::record(size_t v, uint64_t ns, const std::vector<core::TView> &a, const std::vector<core::TView> *b) {
  uint64_t host = 0, dev = 0;
  auto add = [&](const std::vector<core::TView> &vs) {
    for (const auto &t : vs) {
      if (!t.storage) continue;
      (t.storage->mem_loc == MemLoc::Host ? host : dev) += t.desc.num_bytes();
    }
  };
  add(a);
  if (b) add(*b);
  auto &x = *acc_[v];
  x.duration_ns += ns;
  x.runs += 1;
  x.host_bytes += host;
  x.device_bytes += dev;
}

bool Scheduler::acquire_owned_inputs(size_t v, Mode mode) {
  const auto &mask = graph_.nodes[v].infer.owned_inputs;
  for (size_t i = 0; i < mask.size(); ++i) {
    if (!mask[i]) continue;
    auto view = rts_[v].task->acquire_input((int)i);
    while (!view) {
      if (mode == Mode::Once || stop_.load()) return false;
      std::this_thread::yield();
      view = rts_[v].task->acquire_input((int)i);
    }
  }
  return true;
}

void Scheduler::release_owned_outputs(size_t v) {
  const auto &mask = graph_.nodes[v].infer.owned_outputs;
  for (size_t i = 0; i < mask.size(); ++i)
    if (mask[i]) rts_[v].task->release_output((int)i);
}

OpResult Scheduler::run_sync(size_t v) {
  auto &rt = rts_[v];
  core::SyncCtx ctx{rt.in_views, rt.out_views, &cancelled_};
  const auto t0 = std::chrono::steady_clock::now();
  const OpResult r = rt.sync->execute(ctx);
  const auto ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
  if (r == OpResult::Ok)
    record(v, ns, rt.in_views, &rt.out_views);
  else if (r == OpResult::NotReady)
    fail("synchronous task '" + graph_.nodes[v].spec.name + "' returned NotReady, which is not allowed");
  else
    stop_.store(true);  // Cancelled or Eof ends the run
  return r;
}

OpResult Scheduler::run_async_cons(size_t v, Mode mode) {
  auto &rt = rts_[v];
  core::AsyncPopCtx ctx{rt.out_views, &cancelled_};
  const auto t0 = std::chrono::steady_clock::now();
  OpResult r = rt.async->try_pop(ctx);
  while (r == OpResult::NotReady && mode == Mode::Blocking) {
    if (stop_.load()) return OpResult::Cancelled;
    std::this_thread::yield();
    r = rt.async->try_pop(ctx);
  }
  const auto ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
  if (r == OpResult::Ok)
    record(v, ns, rt.out_views, nullptr);
  else if (r != OpResult::NotReady)
    stop_.store(true);
  return r;
}

OpResult Scheduler::run_async_prod(size_t v, Mode mode) {
  auto &rt = rts_[v];
  core::AsyncPushCtx ctx{rt.in_views, &cancelled_};
  const auto t0 = std::chrono::steady_clock::now();
  OpResult r = rt.async->try_push(ctx);
  while (r == OpResult::NotReady && mode == Mode::Blocking) {
    if (stop_.load()) return OpResult::Cancelled;
    std::this_thread::yield();
    r = rt.async->try_push(ctx);
  }
  const auto ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
  if (r == OpResult::Ok)
    record(v, ns, rt.in_views, nullptr);
  else if (r == OpResult::NotReady)
    fail("asynchronous task '" + graph_.nodes[v].spec.name + "' could not take its input");
  else
    stop_.store(true);
  return r;
}

bool Scheduler::run_section_once(const Section &sec, Mode mode) {
  auto where = [&](size_t v, const char *phase) {
    return "Exception in node '" + graph_.nodes[v].spec.name + "'\n  phase: " + phase + "\n  section: " + sec.name +
           " (id=" + std::to_string(sec.id) + ")\n  what: ";
  };
  // 1. owned inputs (ring slots the section's nodes write into)
  for (size_t v : sec.sync_topo) try {
      if (!acquire_owned_inputs(v, mode)) return false;
    } catch (const std::exception &e) {
      throw std::runtime_error(where(v, "acquire_owned_inputs") + e.what());
    }
  for (size_t v : sec.async_prod) try {
      if (!acquire_owned_inputs(v, mode)) return false;
    } catch (const std::exception &e) {
      throw std::runtime_error(where(v, "acquire_owned_inputs") + e.what());
    }
  if (stop_.load()) return false;
  std::vector<size_t> produced;
  // 2. pop the async nodes feeding this section
  for (size_t i = 0; i < sec.async_cons.size(); ++i) {
    const size_t v = sec.async_cons[i];
    OpResult r = OpResult::Ok;
    if (!rts_[v].held) {
      try {
        r = run_async_cons(v, mode);
      } catch (const std::exception &e) {
        throw std::runtime_error(where(v, "execute_async_consumer") + e.what());
      }
    }
    if (r == OpResult::NotReady) {
      // Once mode: skip the section.  Releasing the consumers already popped would consume
      // their windows (a queue's release advances its read index); they are held instead, and
      // the next iteration runs on the same windows.
      for (size_t w : produced) rts_[w].held = true;
      return false;
    }
    if (r != OpResult::Ok) return false;
    rts_[v].held = false;
    produced.push_back(v);
  }
  // 3. synchronous nodes in topological order, then the section's stream
  for (size_t v : sec.sync_topo) {
    OpResult r;
    try {
      r = run_sync(v);
    } catch (const std::exception &e) {
      throw std::runtime_error(where(v, "execute_sync") + e.what());
    }
    if (r != OpResult::Ok) return false;
    produced.push_back(v);
  }
  if (sec.stream) HIP_CHECK(hipStreamSynchronize(sec.stream));
  // 4. push into the async nodes this section feeds
  for (size_t v : sec.async_prod) try {
      if (run_async_prod(v, mode) != OpResult::Ok) return false;
    } catch (const std::exception &e) {
      throw std::runtime_error(where(v, "execute_async_producer") + e.what());
    }
  // 5. release what this iteration produced
  for (size_t v : produced) try {
      release_owned_outputs(v);
    } catch (const std::exception &e) {
      throw std::runtime_error(where(v, "release_owned_outputs") + e.what());
    }
  return true;
}

void Scheduler::run_section_loop(int id) {
  const Section &sec = sections_.at((size_t)id);
  try {
    while (!stop_.load()) (void)run_section_once(sec, Mode::Blocking);
  } catch (const std::exception &e) {
    fail(e.what());
  } catch (...) {
    fail("section " + sec.name + ": non-standard exception");
  }
}

void Scheduler::start() {
  if (running_.exchange(true)) throw std::logic_error("scheduler already running");
  stop_.store(false);
  cancelled_.store(false);
  t_start_ = std::chrono::steady_clock::now();
  for (const auto &s : sections_) threads_.emplace_back([this, id = s.id] { run_section_loop(id); });
}

void Scheduler::request_stop() {
  stop_.store(true);
  cancelled_.store(true);
}

void Scheduler::wait() {
  for (auto &t : threads_)
    if (t.joinable()) t.join();
  threads_.clear();
  running_.store(false);
  std::lock_guard<std::mutex> lk(err_mu_);
  if (!error_.empty()) throw std::runtime_error(error_);
}

int Scheduler::step() {
  if (running_.load()) throw std::logic_error("step() while the section threads run");
  if (t_start_ == std::chrono::steady_clock::time_point{}) t_start_ = std::chrono::steady_clock::now();
  stop_.store(false);
  int ran = 0;
  for (int id : section_order_) {
    try {
      if (run_section_once(sections_.at((size_t)id), Mode::Once)) ++ran;
    } catch (const std::exception &e) {
      fail(e.what());
      throw;
    }
    if (stop_.load()) break;
  }
  return ran;
}

std::map<std::string, NodeMetrics> Scheduler::metrics() const {
  std::map<std::string, NodeMetrics> m;
  const double el = t_start_ == std::chrono::steady_clock::time_point{}
                        ? 0.0
                        : std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start_).count();
  for (size_t v = 0; v < graph_.nodes.size(); ++v) {
    const auto &a = *acc_[v];
    NodeMetrics x;
    x.sample_count = a.runs.load();
    if (x.sample_count) x.average_duration_ms = (double)a.duration_ns.load() / 1e6 / (double)x.sample_count;
    if (el > 0) {
      x.runs_per_second = (double)x.sample_count / el;
      x.host_throughput_bytes_per_second = (double)a.host_bytes.load() / el;
      x.device_throughput_bytes_per_second = (double)a.device_bytes.load() / el;
    }
    m[graph_.nodes[v].spec.name] = x;
  }
  return m;
}

}  // namespace holoflow::runtime

// ===================================================================== C ABI
using namespace holoflow;

struct hfx_graph {
  core::GraphSpec spec;
  std::unique_ptr<runtime::CompilerOutput> out;
  std::unique_ptr<runtime::Scheduler> sched;
};

namespace {

const core::Registry &graph_registry() {
  static core::Registry *r = [] {
    auto *reg = new core::Registry();
    core::register_hip_nodes(*reg);
    return reg;
  }();
  return *r;
}

template <typename F> hfx_status guarded(F &&f) {
  try {
    f();
    hfx::set_last_error("");
    return HFX_OK;
  } catch (const std::logic_error &e) {
    hfx::set_last_error(e.what());
    return HFX_E_INVALID;
  } catch (const runtime::CompilerException &e) {
    hfx::set_last_error(e.what());
    return HFX_E_INVALID;
  } catch (const core::GraphSpecError &e) {
    hfx::set_last_error(e.what());
    return HFX_E_INVALID;
  } catch (const curaii::HfxError &e) {
    hfx::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    hfx::set_last_error(e.what());
    return HFX_E_HIP;
  }
}

hfx_status copy_text(const std::string &s, char *buf, int64_t cap) {
  if (!buf || cap <= (int64_t)s.size()) {
    hfx::set_last_error("output buffer too small (" + std::to_string(s.size() + 1) + " bytes needed)");
    return HFX_E_INVALID;
  }
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return HFX_OK;
}

void stop_and_join(hfx_graph *g) {
  if (!g->sched) return;
  g->sched->request_stop();
  try {
    g->sched->wait();
  } catch (...) {
  }
}

}  // namespace

extern "C" {

hfx_status hfx_graph_compile(const char *graph_json, hfx_graph *prev, hfx_graph **out) {
  if (!out) return HFX_E_INVALID;
  *out = nullptr;
  std::unique_ptr<hfx_graph> old(prev);
  return guarded([&] {
    auto g = std::make_unique<hfx_graph>();
    g->spec = core::graph_from_json(core::Json::parse(graph_json ? graph_json : ""));
    std::unique_ptr<runtime::CompilerOutput> prev_out;
    if (old) {
      stop_and_join(old.get());
      old->sched.reset();
      prev_out = std::move(old->out);
    }
    g->out = runtime::Compiler(graph_registry()).compile(g->spec, std::move(prev_out));
    g->sched = std::make_unique<runtime::Scheduler>(g->out->graph, g->out->sections, g->out->resources);
    *out = g.release();
  });
}

hfx_status hfx_graph_destroy(hfx_graph *g) {
  if (!g) return HFX_OK;
  stop_and_join(g);
  g->sched.reset();
  delete g;
  return HFX_OK;
}

hfx_status hfx_graph_describe(const hfx_graph *g, char *buf, int64_t cap) {
  if (!g) return HFX_E_INVALID;
  std::string s;
  hfx_status rc = guarded([&] {
    const auto &G = g->out->graph;
    const auto &R = g->out->resources;
    core::Json j = core::Json::object();
    core::Json nodes = core::Json::object();
    std::map<size_t, int> sec_of;
    for (const auto &sec : g->out->sections)
      for (size_t v : sec.sync_topo) sec_of[v] = sec.id;
    for (size_t v = 0; v < G.nodes.size(); ++v) {
      const auto &np = G.nodes[v];
      core::Json n = core::Json::object();
      n["kind"] = core::Json(np.spec.kind);
      n["async"] = core::Json(np.infer.kind == core::TaskKind::Async);
      n["section"] = core::Json(sec_of.count(v) ? sec_of[v] : -1);
      core::Json outs = core::Json::array(), sids = core::Json::array();
      for (int t : np.out_tids) {
        outs.push_back(core::to_json(R.tensor_descs.at((size_t)t)));
        sids.push_back(core::Json((long long)R.tid_to_sid.at((size_t)t)));
      }
      n["outputs"] = outs;
      n["storage_ids"] = sids;
      nodes[np.spec.name] = n;
    }
    core::Json secs = core::Json::array();
    auto names = [&](const std::vector<size_t> &vs) {
      core::Json a = core::Json::array();
      for (size_t v : vs) a.push_back(core::Json(G.nodes[v].spec.name));
      return a;
    };
    for (const auto &sec : g->out->sections) {
      core::Json o = core::Json::object();
      o["id"] = core::Json(sec.id);
      o["sync"] = names(sec.sync_topo);
      o["async_cons"] = names(sec.async_cons);
      o["async_prod"] = names(sec.async_prod);
      secs.push_back(o);
    }
    size_t dev = 0, host = 0;
    for (const auto &[sid, b] : R.memory_blocks) (b.mem_loc == core::MemLoc::Host ? host : dev) += b.size_bytes;
    j["nodes"] = nodes;
    j["sections"] = secs;
    j["device_bytes"] = core::Json((unsigned long long)dev);
    j["host_bytes"] = core::Json((unsigned long long)host);
    j["storages"] = core::Json((unsigned long long)R.storages.size());
    s = j.dump();
  });
  return rc != HFX_OK ? rc : copy_text(s, buf, cap);
}

hfx_status hfx_graph_step(hfx_graph *g, int32_t *sections_run) {
  if (!g) return HFX_E_INVALID;
  return guarded([&] {
    const int n = g->sched->step();
    if (sections_run) *sections_run = n;
  });
}

hfx_status hfx_graph_start(hfx_graph *g) {
  if (!g) return HFX_E_INVALID;
  return guarded([&] { g->sched->start(); });
}

hfx_status hfx_graph_stop(hfx_graph *g) {
  if (!g) return HFX_E_INVALID;
  return guarded([&] {
    g->sched->request_stop();
    g->sched->wait();
  });
}

hfx_status hfx_graph_metrics(const hfx_graph *g, char *buf, int64_t cap) {
  if (!g) return HFX_E_INVALID;
  std::string s;
  hfx_status rc = guarded([&] {
    core::Json j = core::Json::object();
    for (const auto &[name, m] : g->sched->metrics()) {
      core::Json o = core::Json::object();
      o["average_duration_ms"] = core::Json(m.average_duration_ms);
      o["runs_per_second"] = core::Json(m.runs_per_second);
      o["host_throughput_bytes_per_second"] = core::Json(m.host_throughput_bytes_per_second);
      o["device_throughput_bytes_per_second"] = core::Json(m.device_throughput_bytes_per_second);
      o["sample_count"] = core::Json((unsigned long long)m.sample_count);
      j[name] = o;
    }
    s = j.dump();
  });
  return rc != HFX_OK ? rc : copy_text(s, buf, cap);
}

hfx_status hfx_graph_output(const hfx_graph *g, const char *node, int32_t out_idx, void **data, char *desc_json,
                            int64_t cap) {
  if (!g || !node || !data) return HFX_E_INVALID;
  std::string s;
  hfx_status rc = guarded([&] {
    const auto &G = g->out->graph;
    const auto &R = g->out->resources;
    for (const auto &np : G.nodes) {
      if (np.spec.name != node) continue;
      if (out_idx < 0 || (size_t)out_idx >= np.out_tids.size()) throw std::out_of_range("output index out of range");
      const size_t tid = (size_t)np.out_tids[(size_t)out_idx];
      const auto &st = *R.storages.at(R.tid_to_sid.at(tid));
      *data = st.ptr;
      s = core::to_json(R.tensor_descs.at(tid)).dump();
      return;
    }
    throw std::invalid_argument(std::string("unknown node '") + node + "'");
  });
  if (rc != HFX_OK) return rc;
  return desc_json ? copy_text(s, desc_json, cap) : HFX_OK;
}

hfx_status hfx_graph_dot(const char *graph_json, char *buf, int64_t cap) {
  std::string s;
  hfx_status rc = guarded([&] { s = core::to_dot(core::graph_from_json(core::Json::parse(graph_json ? graph_json : ""))); });
  return rc != HFX_OK ? rc : copy_text(s, buf, cap);
}

}  // extern "C"
