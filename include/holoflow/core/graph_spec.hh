// Graph specification of a pipeline: named nodes (kind + settings) and indexed edges.
// Mirrors src/holoflow/include/holoflow/core/graph_spec.hh:21-68 (NodeSpec, EdgeSpec,
// GraphSpec, to_json / from_json / to_dot) and its JSON format (graph_spec.cc:66-170):
//   {"nodes": {"<name>": {"type": "<kind>", "params": {...}, "debug": false?}, ...},
//    "edges": [{"from": "<name>", "to": "<name>", "out": <int>, "in": <int>}, ...]}
// Vertices are numbered in sorted node-name order, edges in array order.  The graph is a
// plain adjacency list (vectors) instead of a boost graph.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "holoflow/core/json.hh"

namespace holoflow::core {

struct NodeSpec {
  std::string name;
  std::string kind;
  Json settings = Json::object();
  bool debug = true;
};

struct EdgeSpec {
  int out_idx = 0;
  int in_idx = 0;
};

struct GraphSpec {
  struct Edge {
    size_t src = 0, dst = 0;
    EdgeSpec spec;
  };
  std::vector<NodeSpec> nodes;
  std::vector<Edge> edges;

  size_t add_node(NodeSpec n) {
    nodes.push_back(std::move(n));
    return nodes.size() - 1;
  }
  void add_edge(size_t src, size_t dst, EdgeSpec e) { edges.push_back(Edge{src, dst, e}); }
};

// Malformed graph JSON: a std::runtime_error("GraphSpec JSON: ...") as the reference raises.
class GraphSpecError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

Json to_json(const GraphSpec &g);
GraphSpec graph_from_json(const Json &j);
std::string to_dot(const GraphSpec &g);

}  // namespace holoflow::core
