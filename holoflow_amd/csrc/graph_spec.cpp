// GraphSpec JSON / dot (include/holoflow/core/graph_spec.hh).
#include "holoflow/core/graph_spec.hh"

#include <cmath>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace holoflow::core {

namespace {

[[noreturn]] void json_error(const std::string &msg) { throw GraphSpecError("GraphSpec JSON: " + msg); }

void require(bool cond, const std::string &msg) {
  if (!cond) json_error(msg);
}

bool has_key(const Json &j, const char *key) { return j.is_object() && j.contains(key) && !j.at(key).is_null(); }

bool is_int(const Json &j) { return j.is_number() && j.number() == std::floor(j.number()); }

std::string escape_label(const std::string &s) {
  std::string out;
  for (char c : s) {
    if (c == '\\')
      out += "\\\\";
    else if (c == '"')
      out += "\\\"";
    else if (c == '\n')
      out += "\\n";
    else if (c != '\r')
      out += c;
  }
  return out;
}

}  // namespace

Json to_json(const GraphSpec &g) {
  Json j = Json::object();
  Json nodes = Json::object();
  for (const auto &n : g.nodes) {
    require(!n.name.empty(), "node has empty name");
    Json o = Json::object();
    o["type"] = Json(n.kind);
    o["params"] = n.settings.is_null() ? Json::object() : n.settings;
    if (!n.debug) o["debug"] = Json(false);
    nodes[n.name] = o;
  }
  Json edges = Json::array();
  for (const auto &e : g.edges) {
    const auto &s = g.nodes.at(e.src).name, &d = g.nodes.at(e.dst).name;
    require(!s.empty() && !d.empty(), "edge connects unnamed node(s)");
    Json o = Json::object();
    o["from"] = Json(s);
    o["to"] = Json(d);
    o["out"] = Json(e.spec.out_idx);
    o["in"] = Json(e.spec.in_idx);
    edges.push_back(o);
  }
  j["nodes"] = nodes;
  j["edges"] = edges;
  return j;
}

GraphSpec graph_from_json(const Json &j) {
  require(j.is_object(), "root must be an object");
  require(has_key(j, "nodes"), "missing required key 'nodes'");
  require(j.at("nodes").is_object(), "'nodes' must be an object");
  const Json no_edges = Json::array();
  const Json &jedges = j.contains("edges") ? j.at("edges") : no_edges;
  require(jedges.is_array(), "'edges' must be an array if present");

  GraphSpec g;
  std::unordered_map<std::string, size_t> index;
  for (const auto &[name, node] : j.at("nodes").members()) {  // std::map: sorted names
    require(!name.empty(), "node key (name) must not be empty");
    require(node.is_object(), "node '" + name + "' must be an object");
    require(has_key(node, "type"), "node '" + name + "': missing 'type'");
    require(node.at("type").is_string(), "node '" + name + "': 'type' must be a string");
    require(node.contains("params"), "node '" + name + "': missing 'params'");
    NodeSpec spec;
    spec.name = name;
    spec.kind = node.at("type").str();
    spec.settings = node.at("params").is_null() ? Json::object() : node.at("params");
    if (node.contains("debug")) {
      require(node.at("debug").is_bool(), "node '" + name + "': 'debug' must be boolean");
      spec.debug = node.at("debug").boolean();
    }
    const bool inserted = index.emplace(name, g.add_node(std::move(spec))).second;
    require(inserted, "duplicate node name '" + name + "'");
  }
  for (size_t i = 0; i < jedges.size(); ++i) {
    const Json &e = jedges[i];
    const std::string at = "edge[" + std::to_string(i) + "]";
    require(e.is_object(), at + " must be an object");
    require(has_key(e, "from"), at + ": missing 'from'");
    require(has_key(e, "to"), at + ": missing 'to'");
    require(e.at("from").is_string(), at + ": 'from' must be string");
    require(e.at("to").is_string(), at + ": 'to' must be string");
    const auto f = index.find(e.at("from").str()), t = index.find(e.at("to").str());
    require(f != index.end(), at + ": unknown 'from' node '" + e.at("from").str() + "'");
    require(t != index.end(), at + ": unknown 'to' node '" + e.at("to").str() + "'");
    require(e.contains("out") && is_int(e.at("out")), at + ": 'out' must be int");
    require(e.contains("in") && is_int(e.at("in")), at + ": 'in' must be int");
    g.add_edge(f->second, t->second, EdgeSpec{(int)e.at("out").integer(), (int)e.at("in").integer()});
  }
  return g;
}

std::string to_dot(const GraphSpec &g) {
  std::ostringstream ss;
  ss << "digraph holoflow_graph {\n  rankdir=LR;\n  node [shape=box, fontname=\"Helvetica\"];\n"
        "  edge [fontname=\"Helvetica\"];\n\n";
  for (size_t v = 0; v < g.nodes.size(); ++v) {
    const auto &n = g.nodes[v];
    std::string label = n.name.empty() ? "(unnamed)" : n.name;
    if (!n.kind.empty()) label += "\n(" + n.kind + ")";
    if (n.debug && n.settings.is_object() && n.settings.size() > 0) label += "\n" + n.settings.dump();
    ss << "  v" << v << " [label=\"" << escape_label(label) << "\"];\n";
  }
  ss << "\n";
  for (const auto &e : g.edges)
    ss << "  v" << e.src << " -> v" << e.dst << " [taillabel=\"" << e.spec.out_idx << "\" headlabel=\""
       << e.spec.in_idx << "\"];\n";
  ss << "}\n";
  return ss.str();
}

}  // namespace holoflow::core
