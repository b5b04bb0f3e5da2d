// Compiler of a GraphSpec into an executable graph, HIP-typed.
// Mirrors src/holoflow/include/holoflow/runtime/compiler.hh:30-66 and the passes of
// compiler.cc:288-1124: validate the spec (unique names, registered kinds, one edge per
// input) -> build the plan -> type inference in topological order (factory.infer) -> tensor
// ids -> storage ids (in-place outputs share their input's storage) -> one owner per
// storage -> allocate buffers (storages owned by a task are left to it; blocks of the
// previous compilation with the same memory location and size are reused) -> storage
// adapters -> sections (union-find over sync-sync edges and the sync neighbours of each
// async node) -> one stream per section (reused) -> create, or update through the factory
// when the previous graph had a node of the same name and kind -> bind storage access.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "holoflow/core/graph_spec.hh"
#include "holoflow/core/registry.hh"
#include "holoflow/runtime/graph_exec.hh"

namespace holoflow::runtime {

class CompilerException : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

struct CompilerOutput {
  GraphPlan graph;
  std::vector<Section> sections;
  ExecResources resources;
};

class Compiler {
 public:
  explicit Compiler(const core::Registry &registry) : registry_(registry) {}
  // `prev` (may be null) is consumed: its memory blocks, streams and tasks are reused.
  std::unique_ptr<CompilerOutput> compile(const core::GraphSpec &spec, std::unique_ptr<CompilerOutput> prev = nullptr);

 private:
  const core::Registry &registry_;
};

}  // namespace holoflow::runtime
