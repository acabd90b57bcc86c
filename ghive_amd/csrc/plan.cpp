// Query plans: Hive's EXPLAIN (formatted) and EXPLAIN EXTENDED texts of one query, parsed into the
// operator tree of every Tez vertex with each operator's resolved descriptor -- the host-side
// front end of the reference (gpu_impl/src/Parser/Parser.cu:58-677) and of its operators'
// parseExtended (JoinOperator.cu:21-167, GroupByOperator.cu:16-99, SinkOperator.cu:17-96).
//
// The concise plan gives the tree (indentation; "<-" marks an input edge), operator ids and the
// short attributes (Output:[...], predicate:, Conds:, aggregations:/keys:, PartitionCols:). The
// extended plan gives, per vertex, the operators from the table scan (or the reduce input) up to
// the vertex's sink, each with its long attributes (expressions:, condition map:, keys:,
// outputColumnNames:, key expressions:, sort order:); block k counted from the sink belongs to
// the operator at depth k of the vertex's chain. Column names resolve to positions in the child
// operator's output columns, as the reference does; a reduce sink lists its key columns first.
// Differences from the reference, each following Hive: KEY._colN / VALUE._colN take the whole
// number N (the reference's greedy regex keeps only the last digit); a VALUE column of a
// group-by over a reduce input is offset by that sink's key count at parse time (the reference
// reads the key count of the child's runtime result).
#include <algorithm>
#include <array>
#include <climits>
#include <cstring>
#include <functional>
#include <map>
#include <regex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ghive/ghive.h"
#include "common.hpp"
#include "plan_internal.hpp"

namespace {

struct Op {
  std::string id;      // FIL_25, RS_30, MAPJOIN_27, TS_0, Union 2, ...
  std::string kind;    // TableScan Filter Select MapJoin MergeJoin GroupBy ReduceSink FileSink Limit Union PTF
  std::string vertex;  // the vertex the operator runs in
  int padding = 0;
  std::vector<int> children;
  int parent = -1;
  std::vector<std::string> outputs;           // output column names (a sink: keys first)
  std::vector<std::string> original_outputs;  // a sink: before the keys-first reordering
  std::string table;                          // TableScan: db@table
  std::string predicate;                      // Filter
  std::string conds;                          // joins (concise Conds:)
  std::vector<std::string> agg_strs, key_strs, partition_cols;
  int limit = -1;
  std::string extended;  // the linked EXPLAIN EXTENDED block
  // resolved descriptors
  std::vector<std::string> expressions;  // Select: one expression per output column
  std::vector<std::string> join_inputs;  // join: input operator ids in table order
  std::vector<std::vector<int>> join_cols;
  std::vector<std::vector<int>> join_types;  // {type (GHIVE join type), left, right}
  std::vector<std::vector<int>> maintain_cols;
  std::vector<int> key_offsets;              // GroupBy keys / sink key expressions
  std::vector<std::vector<int>> agg_descs;   // GroupBy: {AggregationType, column}
  std::vector<std::string> key_expressions;  // sink
  std::string sort_orders;                   // sink: '+' / '-' per key
  std::string gby_mode;
  bool require_sort = false;                 // the plan's root sink: the final ORDER BY
};

std::string trim(const std::string &s) {
  const size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

std::vector<std::string> split_lines(const std::string &s) {
  std::vector<std::string> out;
  std::string line;
  std::istringstream in(s);
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    out.push_back(line);
  }
  return out;
}

// Splits on `sep` at nesting depth 0 of (), [] and quotes; trims and unquotes each part.
std::vector<std::string> split_list(const std::string &s, char sep = ',') {
  std::vector<std::string> out;
  std::string cur;
  int depth = 0;
  bool quoted = false;
  for (char c : s) {
    if (c == '"') quoted = !quoted;
    if (!quoted) {
      if (c == '(' || c == '[') ++depth;
      if (c == ')' || c == ']') --depth;
      if (c == sep && depth == 0) {
        out.push_back(cur);
        cur.clear();
        continue;
      }
    }
    cur += c;
  }
  if (!trim(cur).empty() || !out.empty()) out.push_back(cur);
  for (auto &p : out) {
    p = trim(p);
    if (p.size() >= 2 && p.front() == '"' && p.back() == '"') p = p.substr(1, p.size() - 2);
  }
  return out;
}

// "expr (type: t)" -> "expr"
std::string strip_type(const std::string &s) {
  static const std::regex re(R"(^(.*) \(type: [^()]*\)$)");
  std::smatch m;
  const std::string t = trim(s);
  return std::regex_match(t, m, re) ? trim(m[1].str()) : t;
}

int index_of(const std::vector<std::string> &v, const std::string &x) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == x) return (int)i;
  return -1;
}

// Minimal JSON reader for {"op": {"tag": ["col", ...], ...}, ...} (Gson of the Java side's
// Map<String, Map<Byte, List<String>>>).
struct Json {
  const char *p;
  std::string err;
  void ws() {
    while (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r') ++p;
  }
  bool str(std::string &out) {
    ws();
    if (*p != '"') return false;
    ++p;
    out.clear();
    while (*p && *p != '"') {
      if (*p == '\\' && p[1]) ++p;
      out += *p++;
    }
    if (*p != '"') return false;
    ++p;
    return true;
  }
  bool lit(char c) {
    ws();
    if (*p != c) return false;
    ++p;
    return true;
  }
  bool map3(std::map<std::string, std::map<int, std::vector<std::string>>> &out) {
    if (!lit('{')) return false;
    if (lit('}')) return true;
    do {
      std::string op;
      if (!str(op) || !lit(':') || !lit('{')) return false;
      if (!lit('}')) {
        do {
          std::string tag;
          if (!str(tag) || !lit(':') || !lit('[')) return false;
          std::vector<std::string> cols;
          if (!lit(']')) {
            do {
              std::string c;
              if (!str(c)) return false;
              cols.push_back(c);
            } while (lit(','));
            if (!lit(']')) return false;
          }
          out[op][atoi(tag.c_str())] = cols;
        } while (lit(','));
        if (!lit('}')) return false;
      }
    } while (lit(','));
    return lit('}');
  }
};

}  // namespace

struct ghive_plan {
  std::vector<Op> ops;
  std::map<std::string, int> by_id;
  std::map<std::string, int> vertex_root;  // vertex -> its sink
  std::vector<std::string> vertices;       // in order of appearance
  int root = -1;
  std::map<std::string, std::map<int, std::vector<std::string>>> maintain;
};

namespace {

int add_op(ghive_plan &P, Op op) {
  P.ops.push_back(std::move(op));
  const int i = (int)P.ops.size() - 1;
  P.by_id[P.ops[i].id] = i;
  return i;
}

bool is_sink(const Op &o) { return o.kind == "ReduceSink" || o.kind == "FileSink" || o.kind == "Union"; }

// --- the concise plan: the operator tree ----------------------------------------------------
int parse_tree(ghive_plan &P, const std::string &text, std::string &err) {
  static const std::regex vertex_re(R"(^(Map|Reducer) (\d+)( \[[A-Z_]+\])?( vectorized)?( llap)?$)");
  static const std::regex id_re(R"(\[([A-Z]+_\d+)\])");
  static const std::regex union_re(R"(^(Union \d+))");
  static const std::regex refer_vertex_re(R"(Please refer to the previous ((Map|Reducer) \d+))");
  static const std::regex refer_op_re(R"(Please refer to the previous .* \[([A-Z]+_\d+)\])");
  static const std::regex output_re(R"(Output:\[(.*?)\](,|$))");
  const std::vector<std::string> lines = split_lines(text);
  size_t i = 0;
  // the tree starts at the first indented vertex header (after "Stage-N")
  for (; i < lines.size(); ++i) {
    const std::string t = trim(lines[i]);
    if (lines[i].find_first_not_of(' ') > 0 && std::regex_match(t, vertex_re)) break;
  }
  if (i == lines.size()) {
    err = "plan: no vertex in the EXPLAIN text";
    return GHIVE_ERR_INVALID;
  }
  std::vector<int> stack;
  for (; i < lines.size(); ++i) {
    const std::string &raw = lines[i];
    std::string t = trim(raw);
    if (t.empty()) continue;
    int padding = (int)raw.find_first_not_of(' ');
    if (t.rfind("<-", 0) == 0) {
      padding += 2;
      t = trim(t.substr(2));
    }
    std::smatch m;
    int node = -1;
    if (std::regex_search(t, m, refer_op_re)) {
      auto it = P.by_id.find(m[1].str());
      if (it == P.by_id.end()) {
        err = "plan: reference to unknown operator " + m[1].str();
        return GHIVE_ERR_INVALID;
      }
      node = it->second;
    } else if (std::regex_search(t, m, refer_vertex_re)) {
      auto it = P.vertex_root.find(m[1].str());
      if (it == P.vertex_root.end()) {
        err = "plan: reference to unknown vertex " + m[1].str();
        return GHIVE_ERR_INVALID;
      }
      node = it->second;
    } else if (std::regex_match(t, m, vertex_re)) {
      const std::string vertex = m[1].str() + " " + m[2].str();
      if (i + 1 >= lines.size()) break;
      const std::string s = trim(lines[++i]);
      std::smatch ms;
      if (!std::regex_search(s, ms, id_re)) {
        err = "plan: vertex " + vertex + " has no sink operator";
        return GHIVE_ERR_INVALID;
      }
      Op o;
      o.id = ms[1].str();
      o.kind = o.id.rfind("FS_", 0) == 0 ? "FileSink" : "ReduceSink";
      o.vertex = vertex;
      if (i + 1 < lines.size()) {
        const std::string nx = trim(lines[i + 1]);
        if (nx.rfind("PartitionCols:", 0) == 0) {
          o.partition_cols = split_list(nx.substr(14));
          ++i;
        }
      }
      node = add_op(P, o);
      P.vertex_root[vertex] = node;
      P.vertices.push_back(vertex);
    } else if (std::regex_search(t, m, union_re)) {
      Op o;
      o.id = m[1].str();
      o.kind = "Union";
      o.outputs = {"_col0"};
      node = add_op(P, o);
    } else if (std::regex_search(t, m, id_re)) {
      Op o;
      o.id = m[1].str();
      const std::string attr = i + 1 < lines.size() ? trim(lines[i + 1]) : std::string();
      std::smatch ma;
      const std::string pre = o.id.substr(0, o.id.find('_'));
      if (pre == "TS") {
        o.kind = "TableScan";
        static const std::regex ts_re(R"(^(\S+?),)");
        if (std::regex_search(attr, ma, ts_re)) o.table = ma[1].str();
        if (std::regex_search(attr, ma, output_re)) o.outputs = split_list(ma[1].str());
        ++i;
      } else if (pre == "FIL") {
        o.kind = "Filter";
        if (attr.rfind("predicate:", 0) == 0) {
          o.predicate = trim(attr.substr(10));
          ++i;
        }
      } else if (pre == "SEL") {
        o.kind = "Select";
        if (std::regex_search(attr, ma, output_re)) {
          o.outputs = split_list(ma[1].str());
          ++i;
        }
      } else if (pre == "GBY") {
        o.kind = "GroupBy";
        static const std::regex aggs_re(R"(aggregations:\[(.*?)\](,|$))");
        static const std::regex keys_re(R"(keys:(.*)$)");
        if (std::regex_search(attr, ma, output_re)) o.outputs = split_list(ma[1].str());
        if (std::regex_search(attr, ma, aggs_re)) o.agg_strs = split_list(ma[1].str());
        if (std::regex_search(attr, ma, keys_re)) o.key_strs = split_list(ma[1].str());
        ++i;
      } else if (pre == "MAPJOIN" || pre == "MERGEJOIN") {
        o.kind = pre == "MAPJOIN" ? "MapJoin" : "MergeJoin";
        static const std::regex conds_re(R"(Conds:(.*?),(HybridGraceHashJoin|Output))");
        if (std::regex_search(attr, ma, conds_re)) o.conds = ma[1].str();
        if (std::regex_search(attr, ma, output_re)) o.outputs = split_list(ma[1].str());
        ++i;
      } else if (pre == "LIM") {
        o.kind = "Limit";
        static const std::regex lim_re(R"(Number of rows:(\d+))");
        if (std::regex_search(attr, ma, lim_re)) {
          o.limit = std::stoi(ma[1].str());
          ++i;
        }
      } else if (pre == "PTF") {
        o.kind = "PTF";
        ++i;
      } else if (pre == "RS" || pre == "FS") {  // a sink met without its vertex header
        o.kind = pre == "FS" ? "FileSink" : "ReduceSink";
      } else {
        o.kind = pre;
      }
      node = add_op(P, o);
    } else {
      continue;  // an attribute line of an operator already handled, or text we do not use
    }
    // attach by indentation
    while (!stack.empty() && P.ops[stack.back()].padding >= padding) stack.pop_back();
    if (!stack.empty()) {
      const int parent = stack.back();
      P.ops[parent].children.push_back(node);
      if (P.ops[node].parent < 0) P.ops[node].parent = parent;
    } else if (P.root < 0) {
      P.root = node;
    }
    // a referenced node keeps its first position; only new nodes become the stack top
    P.ops[node].padding = padding;
    stack.push_back(node);
  }
  if (P.root < 0) {
    err = "plan: empty operator tree";
    return GHIVE_ERR_INVALID;
  }
  return GHIVE_OK;
}

// vertex of every operator: the vertex of the nearest sink above it within its own chain
void assign_vertices(ghive_plan &P, int node, const std::string &vertex, std::set<int> &seen) {
  if (seen.count(node)) return;
  seen.insert(node);
  Op &o = P.ops[node];
  if (o.vertex.empty() || !is_sink(o)) {
    if (o.vertex.empty()) o.vertex = vertex;
  }
  for (int c : o.children) {
    const Op &ch = P.ops[c];
    if (ch.kind == "ReduceSink" || ch.kind == "FileSink")
      assign_vertices(P, c, ch.vertex, seen);
    else
      assign_vertices(P, c, o.vertex, seen);
  }
}

// output columns that pass-through operators take from their first child (sinks, filters,
// limits), bottom-up
void propagate_outputs(ghive_plan &P, int node, std::set<int> &done) {
  if (done.count(node)) return;
  done.insert(node);
  for (int c : P.ops[node].children) propagate_outputs(P, c, done);
  Op &o = P.ops[node];
  if ((o.kind == "ReduceSink" || o.kind == "FileSink" || o.kind == "Filter" || o.kind == "Limit") &&
      o.outputs.empty() && !o.children.empty())
    o.outputs = P.ops[o.children[0]].outputs;
}

// --- the extended plan: per-vertex operator blocks ------------------------------------------
std::map<std::string, std::vector<std::string>> extended_blocks(const std::string &text) {
  static const std::regex vertex_re(R"(^(Map|Reducer) \d+$)");
  static const std::regex op_re(R"(^[a-zA-Z]+( [a-zA-Z]+)? Operator$)");
  std::map<std::string, std::vector<std::string>> out;
  const std::vector<std::string> lines = split_lines(text);
  int vpad = -1;
  std::string vertex;
  std::string cur;
  bool in_op = false;
  auto flush = [&]() {
    if (in_op && !vertex.empty()) out[vertex].push_back(cur);
    cur.clear();
    in_op = false;
  };
  for (const std::string &raw : lines) {
    const std::string t = trim(raw);
    const int pad = (int)raw.find_first_not_of(' ');
    if (t == "Vertices:") {
      vpad = pad;
      continue;
    }
    if (vpad < 0) continue;
    if (!t.empty() && pad <= vpad) {  // the Vertices section ended
      flush();
      vertex.clear();
      vpad = -1;
      continue;
    }
    if (pad == vpad + 2 && std::regex_match(t, vertex_re)) {
      flush();
      vertex = t;
      continue;
    }
    if (std::regex_match(t, op_re) || t == "TableScan" || t == "Limit") {
      flush();
      in_op = true;
    }
    if (in_op) cur += raw + "\n";
  }
  flush();
  return out;
}

void link_extended(ghive_plan &P, const std::map<std::string, std::vector<std::string>> &blocks) {
  for (const auto &v : P.vertex_root) {
    auto it = blocks.find(v.first);
    if (it == blocks.end()) continue;
    const std::vector<std::string> &b = it->second;
    // the vertex's chain: from its sink down through the children that are not sinks
    int node = v.second, depth = 0;
    while (node >= 0 && depth < (int)b.size()) {
      P.ops[node].extended = b[b.size() - 1 - depth];
      int next = -1;
      for (int c : P.ops[node].children)
        if (!is_sink(P.ops[c]) && next < 0) next = c;
      node = next;
      ++depth;
    }
  }
}

std::vector<std::string> attr_list(const std::string &block, const std::string &name) {
  for (const std::string &l : split_lines(block)) {
    const std::string t = trim(l);
    if (t.rfind(name + ":", 0) == 0) return split_list(t.substr(name.size() + 1));
  }
  return {};
}

std::string attr_value(const std::string &block, const std::string &name) {
  for (const std::string &l : split_lines(block)) {
    const std::string t = trim(l);
    if (t.rfind(name + ":", 0) == 0) return trim(t.substr(name.size() + 1));
  }
  return std::string();
}

// --- EXPLAIN EXTENDED alone: the operator tree ------------------------------------------------
// The extended text holds the whole plan by itself: "Edges:" names each vertex's inputs with the
// edge kind, and each vertex's "Map/Reduce Operator Tree:" nests every operator under the one that
// feeds it. The tree built here is the one parse_tree builds from the formatted plan, with
// operator ids numbered in order of appearance (TS_0, FIL_1, ...). A vertex's first operator
// reads the vertex's SIMPLE / CUSTOM_SIMPLE / ONE_TO_ONE inputs (a merge join: every one, placed by
// the tag of its sink); a map join's other inputs are the BROADCAST vertices its "input vertices:"
// lines place. A table scan's outputs are the table's columns (its "columns" property) that its
// vertex's operators name.
int parse_extended_tree(ghive_plan &P, const std::string &text, std::string &err) {
  static const std::regex edge_re(R"(^((Map|Reducer) \d+) <- (.*)$)");
  static const std::regex edge_in_re(R"(((Map|Reducer|Union) \d+) \(([A-Z_]+)\))");
  static const std::regex vertex_re(R"(^(Map|Reducer) \d+$)");
  static const std::regex op_re(R"(^([a-zA-Z]+( [a-zA-Z]+)?) Operator$)");
  static const std::regex name_re(R"(^name: (\w+)\.(\w+)$)");
  static const std::regex cols_re(R"(^columns (\S+)$)");
  static const std::regex ident_re(R"([A-Za-z_][A-Za-z0-9_]*)");
  static const std::regex input_vertex_re(R"(^(\d+) ((Map|Reducer) \d+)$)");
  const std::vector<std::string> lines = split_lines(text);
  auto pad_of = [](const std::string &raw) { return (int)raw.find_first_not_of(' '); };
  // edges: vertex -> (input vertex, edge kind)
  std::map<std::string, std::vector<std::pair<std::string, std::string>>> edges;
  // vertices in order, each with its lines
  std::vector<std::pair<std::string, std::vector<std::string>>> vertices;
  int epad = -1, vpad = -1;
  for (const std::string &raw : lines) {
    const std::string t = trim(raw);
    if (t.empty()) continue;
    const int pad = pad_of(raw);
    if (t == "Edges:") {
      epad = pad;
      continue;
    }
    if (epad >= 0) {
      std::smatch m;
      if (pad > epad && std::regex_match(t, m, edge_re)) {
        const std::string v = m[1].str(), rest = m[3].str();
        for (auto it = std::sregex_iterator(rest.begin(), rest.end(), edge_in_re); it != std::sregex_iterator(); ++it) {
          if ((*it)[2].str() == "Union") {
            err = "plan: edge " + v + " <- " + (*it)[1].str() + ": UNION vertices are not supported";
            return GHIVE_ERR_UNSUPPORTED;
          }
          edges[v].push_back({(*it)[1].str(), (*it)[3].str()});
        }
        continue;
      }
      epad = -1;
    }
    if (t == "Vertices:") {
      vpad = pad;
      continue;
    }
    if (vpad < 0) continue;
    if (pad <= vpad) {
      vpad = -1;
      continue;
    }
    if (pad == vpad + 2 && std::regex_match(t, vertex_re)) {
      vertices.push_back({t, {}});
      continue;
    }
    if (!vertices.empty()) vertices.back().second.push_back(raw);
  }
  if (vertices.empty()) {
    err = "plan: no vertex in the EXPLAIN EXTENDED text";
    return GHIVE_ERR_INVALID;
  }
  struct VOp {
    int node;
    int pad;
    int up;  // the operator that feeds it inside the vertex, else -1
  };
  std::map<std::string, std::vector<VOp>> vops;
  std::map<std::string, std::vector<int>> sinks;
  int counter = 0;
  for (const auto &v : vertices) {
    const std::string &vertex = v.first;
    const std::vector<std::string> &vl = v.second;
    // the table's name and columns (map vertices), from the Partition section
    std::string table;
    std::vector<std::string> table_cols;
    for (const std::string &raw : vl) {
      const std::string t = trim(raw);
      std::smatch m;
      if (table.empty() && std::regex_match(t, m, name_re)) table = m[1].str() + "@" + m[2].str();
      if (table_cols.empty() && std::regex_match(t, m, cols_re)) table_cols = split_list(m[1].str());
    }
    int tpad = -1;
    std::vector<VOp> &ops = vops[vertex];
    std::vector<int> stack;  // indexes into ops
    std::string all_text;    // the attribute text of every operator of the vertex
    for (const std::string &raw : vl) {
      const std::string t = trim(raw);
      if (t.empty()) continue;
      const int pad = pad_of(raw);
      if (t == "Map Operator Tree:" || t == "Reduce Operator Tree:") {
        tpad = pad;
        continue;
      }
      if (tpad < 0) continue;
      if (pad <= tpad) {  // the operator tree ended ("Execution mode:", "Path -> Alias:", ...)
        tpad = -1;
        continue;
      }
      std::smatch m;
      std::string kind, prefix;
      if (t == "TableScan") kind = "TableScan", prefix = "TS";
      else if (t == "Limit") kind = "Limit", prefix = "LIM";
      else if (std::regex_match(t, m, op_re)) {
        const std::string k = m[1].str();
        if (k == "Filter") kind = "Filter", prefix = "FIL";
        else if (k == "Select") kind = "Select", prefix = "SEL";
        else if (k == "Map Join") kind = "MapJoin", prefix = "MAPJOIN";
        else if (k == "Merge Join") kind = "MergeJoin", prefix = "MERGEJOIN";
        else if (k == "Group By") kind = "GroupBy", prefix = "GBY";
        else if (k == "Reduce Output") kind = "ReduceSink", prefix = "RS";
        else if (k == "File Output") kind = "FileSink", prefix = "FS";
        else {
          kind = k;
          kind.erase(std::remove(kind.begin(), kind.end(), ' '), kind.end());
          prefix = kind;
          std::transform(prefix.begin(), prefix.end(), prefix.begin(), ::toupper);
        }
      }
      if (kind.empty()) {  // an attribute line of the current operator
        if (!ops.empty()) {
          P.ops[ops.back().node].extended += raw + "\n";
          all_text += t + "\n";
        }
        continue;
      }
      while (!stack.empty() && ops[stack.back()].pad >= pad) stack.pop_back();
      Op o;
      o.id = prefix + "_" + std::to_string(counter++);
      o.kind = kind;
      o.vertex = vertex;
      o.extended = raw + "\n";
      if (kind == "TableScan") o.table = table;
      const int node = add_op(P, o);
      ops.push_back({node, pad, stack.empty() ? -1 : ops[stack.back()].node});
      stack.push_back((int)ops.size() - 1);
      if (kind == "ReduceSink" || kind == "FileSink") sinks[vertex].push_back(node);
    }
    if (sinks[vertex].empty()) {
      err = "plan: vertex " + vertex + " has no sink operator";
      return GHIVE_ERR_INVALID;
    }
    if (sinks[vertex].size() > 1) {  // one result per vertex run, as GPUProcess returns
      err = "plan: vertex " + vertex + " has " + std::to_string(sinks[vertex].size()) +
            " sink operators (one result per vertex is supported)";
      return GHIVE_ERR_UNSUPPORTED;
    }
    P.vertex_root[vertex] = sinks[vertex].back();
    P.vertices.push_back(vertex);
    // short attributes from the extended blocks
    for (const VOp &vo : ops) {
      Op &o = P.ops[vo.node];
      if (o.kind == "TableScan") {
        std::set<std::string> named;
        for (auto it = std::sregex_iterator(all_text.begin(), all_text.end(), ident_re); it != std::sregex_iterator(); ++it)
          named.insert(it->str());
        for (const std::string &c : table_cols)
          if (named.count(c)) o.outputs.push_back(c);
      } else if (o.kind == "Select" || o.kind == "GroupBy" || o.kind == "MapJoin" || o.kind == "MergeJoin") {
        o.outputs = attr_list(o.extended, "outputColumnNames");
        if (o.kind == "GroupBy") {
          o.agg_strs = attr_list(o.extended, "aggregations");
          o.key_strs = attr_list(o.extended, "keys");
        }
      } else if (o.kind == "Filter") {
        o.predicate = strip_type(attr_value(o.extended, "predicate"));
      } else if (o.kind == "ReduceSink") {
        for (const std::string &c : attr_list(o.extended, "Map-reduce partition columns"))
          o.partition_cols.push_back(strip_type(c));
      } else if (o.kind == "Limit") {
        const std::string n = attr_value(o.extended, "Number of rows");
        if (!n.empty()) o.limit = std::stoi(n);
      }
    }
  }
  // inputs: in-vertex upstream operators, input vertices' sinks
  auto sink_of = [&](const std::string &v) -> int {
    const auto it = P.vertex_root.find(v);
    return it == P.vertex_root.end() ? -1 : it->second;
  };
  auto link = [&](int parent, int child) {
    P.ops[parent].children.push_back(child);
    if (P.ops[child].parent < 0) P.ops[child].parent = parent;
  };
  for (const auto &v : vertices) {
    const std::string &vertex = v.first;
    std::vector<std::pair<int, int>> tagged;  // (tag, sink) of the vertex's shuffled inputs
    for (const auto &e : edges[vertex]) {
      if (e.second == "BROADCAST_EDGE") continue;
      const int s = sink_of(e.first);
      if (s < 0) {
        err = "plan: " + vertex + " reads unknown vertex " + e.first;
        return GHIVE_ERR_INVALID;
      }
      const std::string tag = attr_value(P.ops[s].extended, "tag");
      tagged.push_back({tag.empty() ? 0 : std::stoi(tag), s});
    }
    std::stable_sort(tagged.begin(), tagged.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    for (const VOp &vo : vops[vertex]) {
      Op &o = P.ops[vo.node];
      if (o.kind == "TableScan") continue;
      std::vector<int> ins;
      if (o.kind == "MapJoin") {
        // positions named by "input vertices:", the rest from inside the vertex (or its input)
        std::map<int, int> at;
        bool in_list = false;
        for (const std::string &l : split_lines(o.extended)) {
          const std::string t = trim(l);
          std::smatch m;
          if (t == "input vertices:") {
            in_list = true;
            continue;
          }
          if (in_list && std::regex_match(t, m, input_vertex_re)) {
            const int s = sink_of(m[2].str());
            if (s < 0) {
              err = "plan: " + o.id + " reads unknown vertex " + m[2].str();
              return GHIVE_ERR_INVALID;
            }
            at[std::stoi(m[1].str())] = s;
            continue;
          }
          in_list = false;
        }
        const int big = vo.up >= 0 ? vo.up : (tagged.empty() ? -1 : tagged.front().second);
        int n = 2;
        for (const auto &a : at) n = std::max(n, a.first + 1);
        for (int pos = 0; pos < n; ++pos) {
          const auto it = at.find(pos);
          ins.push_back(it != at.end() ? it->second : big);
        }
      } else if (vo.up >= 0) {
        ins.push_back(vo.up);
      } else {
        for (const auto &t : tagged) ins.push_back(t.second);
      }
      for (int c : ins) {
        if (c < 0) {
          err = "plan: " + o.id + " in " + vertex + " has no input";
          return GHIVE_ERR_INVALID;
        }
        link(vo.node, c);
      }
      if (o.kind == "MapJoin" || o.kind == "MergeJoin") {  // the input operators, in table order
        std::string conds;
        for (size_t i = 0; i < ins.size(); ++i) conds += (i ? "=" : "") + P.ops[ins[i]].id + "._col";
        o.conds = conds;
      }
    }
  }
  // the root: the one file sink
  for (size_t i = 0; i < P.ops.size(); ++i)
    if (P.ops[i].kind == "FileSink") P.root = (int)i;
  if (P.root < 0) {
    err = "plan: no File Output Operator in the EXPLAIN EXTENDED text";
    return GHIVE_ERR_INVALID;
  }
  return GHIVE_OK;
}

// --- parseExtended per operator -------------------------------------------------------------
void resolve_sink(ghive_plan &P, Op &o) {
  if (!o.original_outputs.empty() || o.kind == "Union") return;  // done already
  o.original_outputs = o.outputs;
  if (o.extended.empty()) return;
  const std::vector<std::string> child_out = o.children.empty() ? o.outputs : P.ops[o.children[0]].outputs;
  for (const std::string &e : attr_list(o.extended, "key expressions")) {
    const std::string x = strip_type(e);
    o.key_expressions.push_back(x);
    const int k = index_of(child_out, x);
    if (k >= 0) o.key_offsets.push_back(k);
  }
  const std::string so = attr_value(o.extended, "sort order");
  o.sort_orders = so;
  std::vector<std::string> reordered = o.key_expressions;
  for (const std::string &c : o.outputs)
    if (index_of(o.key_expressions, c) < 0) reordered.push_back(c);
  o.outputs = reordered;
}

int agg_type_of(const std::string &name) {
  if (name == "sum") return GHIVE_AGG_SUM;
  if (name == "max") return GHIVE_AGG_MAX;
  if (name == "min") return GHIVE_AGG_MIN;
  if (name == "avg") return GHIVE_AGG_AVG;
  if (name == "count") return GHIVE_AGG_CNT;
  return GHIVE_AGG_UNKNOWN;
}

void resolve_groupby(ghive_plan &P, Op &o) {
  static const std::regex key_re(R"(^KEY\._col(\d+)$)"), value_re(R"(^VALUE\._col(\d+)$)");
  static const std::regex agg_re(R"(^(\w+)\((.*)\)$)");
  if (o.children.empty()) return;
  const Op &child = P.ops[o.children[0]];
  const bool reduce_input = child.kind == "ReduceSink" || child.kind == "Union";
  const int child_keys = child.kind == "ReduceSink" ? (int)child.key_expressions.size() : (int)o.key_strs.size();
  auto col_of = [&](const std::string &name, int value_base) -> int {
    std::smatch m;
    if (reduce_input && std::regex_match(name, m, key_re)) return std::stoi(m[1].str());
    if (reduce_input && std::regex_match(name, m, value_re)) return std::stoi(m[1].str()) + value_base;
    return index_of(child.outputs, name);
  };
  for (const std::string &k : o.key_strs) {
    const int c = col_of(strip_type(k), (int)o.key_strs.size());
    if (c >= 0) o.key_offsets.push_back(c);
  }
  for (const std::string &a : o.agg_strs) {
    std::smatch m;
    if (!std::regex_match(a, m, agg_re)) continue;
    const std::string arg = trim(m[2].str());
    int c = arg.empty() ? -1 : col_of(arg, child_keys);
    if (c < 0) c = 0;  // count() / an argument that is not a column: column 0, as the reference
    o.agg_descs.push_back({agg_type_of(m[1].str()), c});
  }
  o.gby_mode = attr_value(o.extended, "mode");
}

void resolve_select(Op &o) {
  if (o.extended.empty()) return;
  o.expressions.clear();
  for (const std::string &e : attr_list(o.extended, "expressions")) o.expressions.push_back(strip_type(e));
  const std::vector<std::string> names = attr_list(o.extended, "outputColumnNames");
  if (!names.empty()) o.outputs = names;
}

void resolve_join(ghive_plan &P, Op &o) {
  static const std::regex name_re(R"([A-Z]+_\d+)");
  static const std::regex type_re(R"((Inner|Left Outer|Left Semi|Right Outer|Full Outer) Join (\d+) to (\d+))");
  static const std::regex key_line_re(R"(^(\d+) (.+)$)");
  // input tables in the order the conditions name them
  for (auto it = std::sregex_iterator(o.conds.begin(), o.conds.end(), name_re); it != std::sregex_iterator(); ++it) {
    const std::string n = it->str();
    if (index_of(o.join_inputs, n) >= 0) break;
    o.join_inputs.push_back(n);
  }
  std::vector<int> inputs;
  for (const std::string &n : o.join_inputs) {
    int found = -1;
    for (int c : o.children)
      if (P.ops[c].id == n) found = c;
    inputs.push_back(found);
  }
  o.join_cols.assign(inputs.size(), {});
  o.maintain_cols.assign(inputs.size(), {});
  bool in_keys = false;
  for (const std::string &l : split_lines(o.extended)) {
    const std::string t = trim(l);
    std::smatch m;
    if (t.rfind("outputColumnNames:", 0) == 0) {
      o.outputs = split_list(t.substr(18));
      break;
    }
    if (std::regex_search(t, m, type_re)) {
      const std::string k = m[1].str();
      const int type = k == "Inner" ? GHIVE_INNER_JOIN
                       : k == "Left Outer" ? GHIVE_LEFT_OUTER_JOIN
                       : k == "Left Semi" ? GHIVE_LEFT_SEMI_JOIN
                       : k == "Right Outer" ? GHIVE_RIGHT_OUTER_JOIN
                                            : GHIVE_FULL_OUTER_JOIN;
      o.join_types.push_back({type, std::stoi(m[2].str()), std::stoi(m[3].str())});
      continue;
    }
    if (t == "keys:") {
      in_keys = true;
      continue;
    }
    if (in_keys && std::regex_match(t, m, key_line_re)) {
      const int tbl = std::stoi(m[1].str());
      if (tbl < 0 || tbl >= (int)inputs.size() || inputs[tbl] < 0) continue;
      for (const std::string &c : split_list(m[2].str())) {
        const int k = index_of(P.ops[inputs[tbl]].outputs, strip_type(c));
        if (k >= 0) o.join_cols[tbl].push_back(k);
      }
    }
  }
  // maintained columns from the Java side's map (JoinOperator.cu:130-162)
  auto mit = P.maintain.find(o.id);
  if (mit == P.maintain.end()) return;
  static const std::regex value_re(R"(^VALUE\._col(\d+)$)");
  for (const auto &tag : mit->second) {
    if (tag.first < 0 || tag.first >= (int)inputs.size() || inputs[tag.first] < 0) continue;
    const Op &in = P.ops[inputs[tag.first]];
    for (const std::string &c : tag.second) {
      int k = -1;
      std::smatch m;
      if (o.kind == "MergeJoin" && std::regex_match(c, m, value_re)) {
        // a reduce-side input: VALUE._colN is the sink's column _col(N+1) (its key is _col0)
        k = index_of(in.original_outputs, "_col" + std::to_string(std::stoi(m[1].str()) + 1));
      } else {
        k = index_of(in.outputs, c);
      }
      if (k >= 0) o.maintain_cols[tag.first].push_back(k);
    }
  }
}

void resolve(ghive_plan &P, int node, std::set<int> &done) {
  if (done.count(node)) return;
  done.insert(node);
  for (int c : P.ops[node].children) resolve(P, c, done);
  Op &o = P.ops[node];
  if (o.kind == "ReduceSink" || o.kind == "FileSink") resolve_sink(P, o);
  else if (o.kind == "GroupBy") resolve_groupby(P, o);
  else if (o.kind == "Select") resolve_select(o);
  else if (o.kind == "MapJoin" || o.kind == "MergeJoin") resolve_join(P, o);
  else if (o.kind == "Filter") {
    const std::string pred = attr_value(o.extended, "predicate");
    if (!pred.empty() && o.predicate.empty()) o.predicate = strip_type(pred);
  }
  // pass-through operators above a resolved child take its (possibly reordered) outputs
  if ((o.kind == "Filter" || o.kind == "Limit") && !o.children.empty()) o.outputs = P.ops[o.children[0]].outputs;
}

// --- JSON description ---------------------------------------------------------------------
std::string jstr(const std::string &s) {
  std::string r = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') r += '\\';
    if (c == '\n') {
      r += "\\n";
      continue;
    }
    r += c;
  }
  return r + "\"";
}
std::string jstrs(const std::vector<std::string> &v) {
  std::string r = "[";
  for (size_t i = 0; i < v.size(); ++i) r += (i ? "," : "") + jstr(v[i]);
  return r + "]";
}
std::string jints(const std::vector<int> &v) {
  std::string r = "[";
  for (size_t i = 0; i < v.size(); ++i) r += (i ? "," : "") + std::to_string(v[i]);
  return r + "]";
}
std::string jints2(const std::vector<std::vector<int>> &v) {
  std::string r = "[";
  for (size_t i = 0; i < v.size(); ++i) r += (i ? "," : "") + jints(v[i]);
  return r + "]";
}

// Whether GroupBy `node` runs as one fused fragment (the plan alone decides the shape).
bool fusable(const ghive_plan &P, int node, std::string &why);

std::string describe(const ghive_plan &P) {
  std::string r = "{\"root\":" + jstr(P.root >= 0 ? P.ops[P.root].id : "") + ",\"vertices\":{";
  bool first = true;
  for (const std::string &v : P.vertices) {
    r += (first ? "" : ",") + jstr(v) + ":" + jstr(P.ops[P.vertex_root.at(v)].id);
    first = false;
  }
  r += "},\"operators\":{";
  for (size_t i = 0; i < P.ops.size(); ++i) {
    const Op &o = P.ops[i];
    std::vector<std::string> kids;
    for (int c : o.children) kids.push_back(P.ops[c].id);
    r += (i ? "," : "") + jstr(o.id) + ":{\"kind\":" + jstr(o.kind) + ",\"vertex\":" + jstr(o.vertex) +
         ",\"children\":" + jstrs(kids) + ",\"outputs\":" + jstrs(o.outputs);
    if (o.kind == "TableScan") r += ",\"table\":" + jstr(o.table);
    if (o.kind == "Filter") r += ",\"predicate\":" + jstr(o.predicate);
    if (o.kind == "Select") r += ",\"expressions\":" + jstrs(o.expressions);
    if (o.kind == "Limit") r += ",\"limit\":" + std::to_string(o.limit);
    if (o.kind == "MapJoin" || o.kind == "MergeJoin")
      r += ",\"conds\":" + jstr(o.conds) + ",\"join_inputs\":" + jstrs(o.join_inputs) +
           ",\"join_cols\":" + jints2(o.join_cols) + ",\"join_types\":" + jints2(o.join_types) +
           ",\"maintain_cols\":" + jints2(o.maintain_cols);
    if (o.kind == "GroupBy") {
      std::string why;
      const bool f = fusable(P, (int)i, why);
      r += ",\"keys\":" + jints(o.key_offsets) + ",\"aggs\":" + jints2(o.agg_descs) + ",\"mode\":" + jstr(o.gby_mode) +
           ",\"fused_form\":" + (f ? "true" : "false") + ",\"not_fused\":" + jstr(f ? "" : why);
    }
    if (o.kind == "ReduceSink" || o.kind == "FileSink")
      r += ",\"key_expressions\":" + jstrs(o.key_expressions) + ",\"key_offsets\":" + jints(o.key_offsets) +
           ",\"sort_orders\":" + jstr(o.sort_orders) + ",\"partition_cols\":" + jstrs(o.partition_cols) +
           ",\"original_outputs\":" + jstrs(o.original_outputs) +
           ",\"require_sort\":" + (o.require_sort ? "true" : "false");
    r += ",\"linked\":" + std::string(o.extended.empty() ? "false" : "true") + "}";
  }
  return r + "}}";
}

void set_err(char *err, uint64_t cap, const std::string &msg) {
  if (err && cap) {
    const size_t n = msg.size() < cap - 1 ? msg.size() : (size_t)cap - 1;
    memcpy(err, msg.data(), n);
    err[n] = 0;
  }
}

}  // namespace

namespace {

// ---- execution of one vertex ------------------------------------------------------------------
// Operator::execute over a vertex's chain (execute() of gpu_impl/src/Operator/), each operator
// on the library's entry points: TableScan binds an input table by name, Filter parses its
// predicate text (ghive_parse_predicate) and compacts (ghive_filter_compact), Select aliases
// columns or computes binary arithmetic (ghive_select_arith), a join runs ghive_join and
// materialises its maintained columns (ghive_join_result), GroupBy runs ghive_groupby, the sink
// orders its columns keys first and, as the plan's root, performs the final ORDER BY (ghive_sort
// + ghive_gather) and LIMIT.
// The columns each input of a join keeps, in output order (indexes into in_names[t]): from the
// Java side's maintain-values map, else by Hive's naming -- a join's outputs are numbered _colN by
// position in the concatenation of its inputs' columns, so input 0's columns keep their names and
// input 1's come renumbered after them. Input 0 maps by name; the rest map to input 1 by the offset
// that puts the last of them on input 1's last non-key column (exact when that column is kept, as
// in every SSB plan).
int join_keep(const Op &o, const std::vector<std::vector<std::string>> &in_names,
              std::vector<std::vector<uint32_t>> &keep, std::string &err) {
  const size_t nt = in_names.size();
  keep.assign(nt, {});
  bool from_map = false;
  for (size_t t = 0; t < nt && t < o.maintain_cols.size(); ++t)
    for (int c : o.maintain_cols[t]) {
      keep[t].push_back((uint32_t)c);
      from_map = true;
    }
  if (from_map) return GHIVE_OK;
  static const std::regex col_re(R"(^_col(\d+)$)");
  std::vector<std::pair<long, std::string>> rest;
  for (const std::string &n : o.outputs) {
    const int c = index_of(in_names[0], n);
    if (c >= 0) {
      keep[0].push_back((uint32_t)c);
      continue;
    }
    std::smatch m;
    if (nt != 2 || !std::regex_match(n, m, col_re)) {
      err = "plan: " + o.id + ": output '" + n + "' is not a column of input 0; pass the maintain-values map";
      return GHIVE_ERR_UNSUPPORTED;
    }
    rest.push_back({std::stol(m[1].str()), n});
  }
  if (rest.empty()) return GHIVE_OK;
  long jmax = -1;
  for (size_t c = 0; c < in_names[1].size(); ++c) {
    std::smatch m;
    if (o.join_cols.size() > 1 && !o.join_cols[1].empty() && (int)c != o.join_cols[1][0] &&
        std::regex_match(in_names[1][c], m, col_re))
      jmax = std::max(jmax, std::stol(m[1].str()));
  }
  const long offset = rest.back().first - jmax;
  for (const auto &r : rest) {
    const int c = jmax < 0 ? -1 : index_of(in_names[1], "_col" + std::to_string(r.first - offset));
    if (c < 0) {
      err = "plan: " + o.id + ": cannot place output '" + r.second + "' in input 1; pass the maintain-values map";
      return GHIVE_ERR_UNSUPPORTED;
    }
    keep[1].push_back((uint32_t)c);
  }
  return GHIVE_OK;
}

// ---- the fused form of a map vertex ------------------------------------------------------------
// A map vertex of the CS1 shape (TableScan -> [Filter] -> [Select] -> MapJoin+ -> [Select] ->
// GroupBy -> sink) needs no intermediate table: every column the group-by reads is a fact column,
// a product of two, or an attribute of a joined dimension, so the whole chain is one
// ghive_fragment_execute. The recognizer follows the column names through the chain symbolically.
struct Sym {
  enum Kind { kNone, kFact, kDim, kMul } kind = kNone;
  int col = -1;        // kFact: fact column (position in the scan's outputs); kDim: dimension column
  int join = -1;       // kDim: the join (0 = nearest the scan)
  int a = -1, b = -1;  // kMul: fact columns
};

struct FuseJoin {
  const Op *op = nullptr;
  const Op *dim = nullptr;  // the broadcast sink of the dimension's vertex
  int fact_col = -1;        // probed fact column
  int dim_key = -1;         // key column of the dimension table
  int dim_pay = -1;         // the attribute the join carries (a group key), -1: none
};

struct FuseAgg {
  int type = 0;  // GHIVE_AGG_*
  Sym arg;       // kNone: count()
};

struct FuseShape {
  const Op *scan = nullptr;
  const Op *filter = nullptr;
  std::vector<std::string> filter_names;  // the filter's input columns
  std::vector<Sym> filter_syms;
  std::vector<FuseJoin> joins;
  std::vector<Sym> keys;
  std::vector<FuseAgg> aggs;
};

// "(x * y)" -> x, y
bool split_product(const std::string &e0, std::string &x, std::string &y) {
  const std::string e = trim(e0);
  if (e.size() < 7 || e.front() != '(' || e.back() != ')') return false;
  const std::string in = e.substr(1, e.size() - 2);
  const size_t p = in.find(" * ");
  if (p == std::string::npos || in.find(" * ", p + 1) != std::string::npos || in.find('(') != std::string::npos)
    return false;
  x = trim(in.substr(0, p));
  y = trim(in.substr(p + 3));
  return !x.empty() && !y.empty();
}

// Whether the GroupBy `gby` of `vertex` has the fused shape; on success S describes it, else
// `why` says what is missing. Needs the plan only (no data).
bool fuse_shape(const ghive_plan &P, const std::string &vertex, int gby, FuseShape &S, std::string &why) {
  const Op &g = P.ops[gby];
  if (g.kind != "GroupBy" || g.vertex != vertex) return why = "not a group-by of this vertex", false;
  if (!g.gby_mode.empty() && g.gby_mode != "hash") return why = "group-by mode " + g.gby_mode, false;
  if (g.children.empty()) return why = "group-by without input", false;
  std::vector<int> chain;  // from the group-by's input down to the scan
  for (int node = g.children[0]; node >= 0;) {
    const Op &o = P.ops[node];
    if (o.vertex != vertex) return why = "the group-by reads another vertex's output", false;
    chain.push_back(node);
    if (o.kind == "TableScan") break;
    if (o.kind == "MapJoin") {
      if (o.join_inputs.empty()) return why = o.id + " has no inputs", false;
      const auto it = P.by_id.find(o.join_inputs[0]);
      node = it == P.by_id.end() ? -1 : it->second;
    } else if (o.kind == "Select" || o.kind == "Filter") {
      node = o.children.empty() ? -1 : o.children[0];
    } else {
      return why = o.kind + " " + o.id + " in the chain", false;
    }
    if (node < 0) return why = "broken operator chain at " + o.id, false;
  }
  if (chain.empty() || P.ops[chain.back()].kind != "TableScan") return why = "no table scan below the group-by", false;
  std::vector<std::string> names;
  std::vector<Sym> syms;
  bool joined = false;
  for (auto it = chain.rbegin(); it != chain.rend(); ++it) {
    const Op &o = P.ops[*it];
    if (o.kind == "TableScan") {
      S.scan = &o;
      names = o.outputs;
      syms.assign(names.size(), Sym());
      for (size_t c = 0; c < names.size(); ++c) {
        syms[c].kind = Sym::kFact;
        syms[c].col = (int)c;
      }
    } else if (o.kind == "Filter") {
      if (joined || S.filter) return why = "a filter above a join", false;
      S.filter = &o;
      S.filter_names = names;
      S.filter_syms = syms;
    } else if (o.kind == "Select") {
      const std::vector<std::string> &ex = o.expressions.empty() ? o.outputs : o.expressions;
      if (ex.size() != o.outputs.size()) return why = o.id + ": expressions and outputs differ", false;
      std::vector<Sym> ns(ex.size());
      for (size_t i = 0; i < ex.size(); ++i) {
        const int k = index_of(names, trim(ex[i]));
        std::string x, y;
        if (k >= 0) {
          ns[i] = syms[k];
        } else if (split_product(ex[i], x, y)) {
          const int kx = index_of(names, x), ky = index_of(names, y);
          if (kx >= 0 && ky >= 0 && syms[kx].kind == Sym::kFact && syms[ky].kind == Sym::kFact) {
            ns[i].kind = Sym::kMul;
            ns[i].a = syms[kx].col;
            ns[i].b = syms[ky].col;
          }
        }
      }
      names = o.outputs;
      syms = ns;
    } else {  // MapJoin
      if (o.join_inputs.size() != 2) return why = o.id + ": not a two-input join", false;
      if (o.join_types.size() != 1 || o.join_types[0][0] != GHIVE_INNER_JOIN || o.join_types[0][1] != 0 ||
          o.join_types[0][2] != 1)
        return why = o.id + ": not an inner join of input 0 to input 1", false;
      const auto dit = P.by_id.find(o.join_inputs[1]);
      if (dit == P.by_id.end()) return why = o.id + ": unknown input", false;
      const Op &dim = P.ops[dit->second];
      if (dim.kind != "ReduceSink" || dim.vertex == vertex)
        return why = o.id + ": input 1 is not another vertex's broadcast", false;
      if (o.join_cols.size() != 2 || o.join_cols[0].size() != 1 || o.join_cols[1].size() != 1)
        return why = o.id + ": not a single-column key", false;
      const int pk = o.join_cols[0][0];
      if (pk < 0 || pk >= (int)syms.size() || syms[pk].kind != Sym::kFact)
        return why = o.id + ": the probe key is not a fact column", false;
      FuseJoin J;
      J.op = &o;
      J.dim = &dim;
      J.fact_col = syms[pk].col;
      J.dim_key = o.join_cols[1][0];
      std::vector<std::vector<uint32_t>> keep;
      std::string e;
      if (join_keep(o, {names, dim.outputs}, keep, e) != GHIVE_OK) return why = e, false;
      if (keep[0].size() + keep[1].size() != o.outputs.size()) return why = o.id + ": outputs not placed", false;
      std::vector<Sym> ns;
      for (uint32_t c : keep[0]) ns.push_back(c < syms.size() ? syms[c] : Sym());
      for (uint32_t c : keep[1]) {
        Sym d;
        d.kind = Sym::kDim;
        d.join = (int)S.joins.size();
        d.col = (int)c;
        ns.push_back(d);
      }
      S.joins.push_back(J);
      names = o.outputs;
      syms = ns;
      joined = true;
    }
  }
  if (S.joins.empty()) return why = "no map join below the group-by", false;
  if (g.key_offsets.size() != g.key_strs.size()) return why = "unresolved group-by keys", false;
  if (g.key_offsets.size() > 2) return why = "more than 2 group-by keys", false;
  for (int k : g.key_offsets) {
    if (k < 0 || k >= (int)syms.size()) return why = "group-by key beyond its input", false;
    const Sym &s = syms[k];
    if (s.kind == Sym::kDim) {
      int &pay = S.joins[s.join].dim_pay;
      if (pay >= 0 && pay != s.col) return why = "two group-by keys from one dimension", false;
      pay = s.col;
    } else if (s.kind != Sym::kFact) {
      return why = "a group-by key that is neither a fact column nor a dimension attribute", false;
    }
    S.keys.push_back(s);
  }
  static const std::regex agg_re(R"(^(\w+)\((.*)\)$)");
  if (g.agg_descs.size() != g.agg_strs.size()) return why = "unresolved aggregates", false;
  for (size_t a = 0; a < g.agg_strs.size(); ++a) {
    std::smatch m;
    if (!std::regex_match(g.agg_strs[a], m, agg_re)) return why = "aggregate " + g.agg_strs[a], false;
    FuseAgg A;
    A.type = g.agg_descs[a][0];
    if (A.type != GHIVE_AGG_SUM && A.type != GHIVE_AGG_CNT && A.type != GHIVE_AGG_MIN && A.type != GHIVE_AGG_MAX)
      return why = "aggregate " + g.agg_strs[a], false;
    if (!trim(m[2].str()).empty()) {
      const int c = g.agg_descs[a][1];
      if (c < 0 || c >= (int)syms.size() || (syms[c].kind != Sym::kFact && syms[c].kind != Sym::kMul))
        return why = "aggregate over " + g.agg_strs[a] + ": not a fact column or a product of two", false;
      A.arg = syms[c];
    } else if (A.type != GHIVE_AGG_CNT) {
      return why = "aggregate " + g.agg_strs[a], false;
    }
    S.aggs.push_back(A);
  }
  if (S.aggs.empty()) return why = "a group-by without aggregates", false;
  if (S.keys.size() + S.aggs.size() != g.outputs.size()) return why = "keys and aggregates do not match the outputs", false;
  return true;
}

bool fusable(const ghive_plan &P, int node, std::string &why) {
  FuseShape S;
  return fuse_shape(P, P.ops[node].vertex, node, S, why);
}

struct ExecTable {
  std::vector<std::string> names;
  std::vector<ghive_column> cols;
  std::vector<int> logical;  // from_input: the column's type in the plan (LONG for a 32-bit staged copy)
  uint64_t rows = 0;
  int reduce_keys = -1;  // a reduce input: the number of key columns leading it
  ghive_table view() const { return ghive_table{cols.data(), (uint32_t)cols.size(), 0, rows}; }
  int find(const std::string &n) const { return index_of(names, n); }
};

struct Executor {
  ghive_ctx *ctx;
  const ghive_plan &P;
  const std::string vertex;
  std::map<std::string, const ghive_plan_input *> inputs;
  std::vector<void *> owned;  // device buffers made by this execution
  std::string err;
  // input name -> per column: 1 when the staged column holds LONG values as 32-bit integers
  const std::map<std::string, std::vector<char>> *narrowed = nullptr;

  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  // A STRING column gets its (start,end) pairs; the operator that fills it points data at the blob.
  int alloc_col(int type, uint64_t rows, ghive_column &c) {
    memset(&c, 0, sizeof c);
    c.type = type;
    c.row_num = rows;
    const uint64_t w = (type == GHIVE_LONG || type == GHIVE_DOUBLE || type == GHIVE_STRING) ? 8 : 4;
    void *p = nullptr;
    const int st = ghive_malloc(ctx, (rows ? rows : 1) * w, &p);
    if (st != GHIVE_OK) return fail(st, std::string("plan: ") + ghive_last_error(ctx));
    owned.push_back(p);
    if (type == GHIVE_STRING) c.aux = p;
    else c.data = p;
    return GHIVE_OK;
  }
  int op_fail(int st) { return fail(st, std::string("plan: ") + ghive_last_error(ctx)); }
  int op_error(int st, const Op &o) {
    return fail(st, "plan: " + o.id + ": " + ghive_last_error(ctx));
  }

  // Binds input `name` (else `alt`, when given). Without column names an input's columns are
  // positional: `positional` names them when it has one name per column (a TableScan's outputs, a
  // reduce sink's keys-first outputs -- the order GTables delivers them in), else _col0.._colN-1.
  // A LONG column staged as 32-bit integers reaches the fused kernel as it is (keep_narrow) and every
  // other operator as LONG again.
  int from_input(const std::string &name, const std::string &alt, const std::vector<std::string> &want,
                 const std::vector<std::string> &positional, ExecTable &t, bool keep_narrow = false) {
    auto it = inputs.find(name);
    if (it == inputs.end() && !alt.empty()) it = inputs.find(alt);
    if (it == inputs.end()) return fail(GHIVE_ERR_INVALID, "plan: no input named '" + name + "'");
    const std::vector<char> *nflags = nullptr;
    if (narrowed) {
      const auto nit = narrowed->find(it->first);
      if (nit != narrowed->end()) nflags = &nit->second;
    }
    const ghive_plan_input &in = *it->second;
    if (!in.table) return fail(GHIVE_ERR_INVALID, "plan: input '" + name + "' has no table");
    std::vector<std::string> names;
    for (uint32_t c = 0; c < in.table->ncols; ++c) {
      if (in.column_names && in.column_names[c]) names.push_back(in.column_names[c]);
      else if (!in.column_names && positional.size() == in.table->ncols) names.push_back(positional[c]);
      else names.push_back("_col" + std::to_string(c));
    }
    t.rows = in.table->row_num;
    const std::vector<std::string> &order = want.empty() ? names : want;
    for (const std::string &n : order) {
      const int c = index_of(names, n);
      if (c < 0) return fail(GHIVE_ERR_INVALID, "plan: input '" + name + "' has no column '" + n + "'");
      t.names.push_back(n);
      ghive_column col = in.table->columns[c];
      const bool narrow = nflags && c < (int)nflags->size() && (*nflags)[c] && col.type == GHIVE_INT;
      t.logical.push_back(narrow ? GHIVE_LONG : col.type);
      if (narrow && !keep_narrow) {
        ghive_column wide;
        int st = alloc_col(GHIVE_LONG, t.rows, wide);
        if (st != GHIVE_OK) return st;
        if ((st = ghive_select_cast(ctx, &col, GHIVE_LONG, t.rows, &wide)) != GHIVE_OK) return op_fail(st);
        wide.row_num = t.rows;
        col = wide;
      }
      t.cols.push_back(col);
    }
    return GHIVE_OK;
  }

  // one operand of an arithmetic expression: a column of t, or a numeric literal (1 row)
  int operand(const std::string &e, const ExecTable &t, ghive_column &c) {
    const int k = t.find(e);
    if (k >= 0) {
      c = t.cols[k];
      return GHIVE_OK;
    }
    static const std::regex lit_re(R"(^(-?\d+(\.\d*)?([eE][-+]?\d+)?)([DdLlSsYy]?)$)");
    std::smatch m;
    if (!std::regex_match(e, m, lit_re)) return fail(GHIVE_ERR_UNSUPPORTED, "plan: operand '" + e + "'");
    const std::string suffix = m[4].str();
    const bool fp = !suffix.empty() ? (suffix == "D" || suffix == "d") : (m[2].matched || m[3].matched);
    const bool lng = suffix == "L" || suffix == "l";
    int st = alloc_col(fp ? GHIVE_DOUBLE : (lng ? GHIVE_LONG : GHIVE_INT), 1, c);
    if (st != GHIVE_OK) return st;
    if (fp) {
      const double v = std::stod(m[1].str());
      return ghive_memcpy(ctx, c.data, &v, 8, 0);
    }
    if (lng) {
      const int64_t v = std::stoll(m[1].str());
      return ghive_memcpy(ctx, c.data, &v, 8, 0);
    }
    const int32_t v = (int32_t)std::stol(m[1].str());
    return ghive_memcpy(ctx, c.data, &v, 4, 0);
  }

  // Select expression: a column, KEY.reducesinkkeyN / VALUE._colN of a reduce input, or
  // (a op b) with op in + - * /, operands columns or literals, nested
  int expression(const std::string &e0, const ExecTable &t, ghive_column &out) {
    const std::string e = trim(e0);
    const int k = t.find(e);
    if (k >= 0) {
      out = t.cols[k];
      return GHIVE_OK;
    }
    std::smatch m;
    static const std::regex key_re(R"(^KEY\.reducesinkkey(\d+)$)"), val_re(R"(^VALUE\._col(\d+)$)");
    if (std::regex_match(e, m, key_re)) {
      const int c = std::stoi(m[1].str());
      if (c >= (int)t.cols.size()) return fail(GHIVE_ERR_INVALID, "plan: " + e + " beyond the reduce input");
      out = t.cols[c];
      return GHIVE_OK;
    }
    if (e.size() > 2 && e.front() == '(' && e.back() == ')') {
      const std::string in = e.substr(1, e.size() - 2);
      int depth = 0;
      for (size_t i = 0; i < in.size(); ++i) {  // the top-level binary operator " op "
        if (in[i] == '(') ++depth;
        if (in[i] == ')') --depth;
        if (depth == 0 && i + 2 < in.size() && in[i] == ' ' && in[i + 2] == ' ' &&
            (in[i + 1] == '+' || in[i + 1] == '-' || in[i + 1] == '*' || in[i + 1] == '/')) {
          ghive_column a, b;
          int st = sub_operand(in.substr(0, i), t, a);
          if (st != GHIVE_OK) return st;
          st = sub_operand(in.substr(i + 3), t, b);
          if (st != GHIVE_OK) return st;
          const int otype = (in[i + 1] == '/' || a.type == GHIVE_DOUBLE || b.type == GHIVE_DOUBLE ||
                             a.type == GHIVE_FLOAT || b.type == GHIVE_FLOAT)
                                ? GHIVE_DOUBLE
                                : (a.type == GHIVE_INT && b.type == GHIVE_INT ? GHIVE_INT : GHIVE_LONG);
          st = alloc_col(otype, t.rows, out);
          if (st != GHIVE_OK) return st;
          st = ghive_select_arith(ctx, &a, &b, in[i + 1], t.rows, &out);
          return st == GHIVE_OK ? GHIVE_OK : fail(st, std::string("plan: ") + ghive_last_error(ctx));
        }
      }
      return expression(in, t, out);  // redundant parentheses
    }
    static const std::regex call_re(R"(^(\w+)\((.*)\)$)");
    if (std::regex_match(e, m, call_re)) {  // UDFToInteger / UDFToLong / UDFToDouble / UDFToFloat, substr
      const std::string fn = m[1].str();
      const std::vector<std::string> args = split_list(m[2].str());
      int to = -1;
      if (fn == "UDFToInteger") to = GHIVE_INT;
      if (fn == "UDFToLong") to = GHIVE_LONG;
      if (fn == "UDFToDouble") to = GHIVE_DOUBLE;
      if (fn == "UDFToFloat") to = GHIVE_FLOAT;
      if (to >= 0 && args.size() == 1) {
        ghive_column x;
        int st = expression(args[0], t, x);
        if (st != GHIVE_OK) return st;
        if ((st = alloc_col(to, t.rows, out)) != GHIVE_OK) return st;
        st = ghive_select_cast(ctx, &x, to, t.rows, &out);
        return st == GHIVE_OK ? GHIVE_OK : op_fail(st);
      }
      if ((fn == "substr" || fn == "substring") && (args.size() == 2 || args.size() == 3)) {
        ghive_column x;
        int st = expression(args[0], t, x);
        if (st != GHIVE_OK) return st;
        static const std::regex int_re(R"(^-?\d+$)");
        if (!std::regex_match(args[1], int_re) || (args.size() == 3 && !std::regex_match(args[2], int_re)))
          return fail(GHIVE_ERR_UNSUPPORTED, "plan: substr with non-constant bounds: " + e);
        const int32_t pos = (int32_t)std::stol(args[1]);
        const int32_t len = args.size() == 3 ? (int32_t)std::stol(args[2]) : INT32_MAX;
        if ((st = alloc_col(GHIVE_STRING, t.rows, out)) != GHIVE_OK) return st;
        st = ghive_select_substr(ctx, &x, pos, len, t.rows, &out);
        return st == GHIVE_OK ? GHIVE_OK : op_fail(st);
      }
      return fail(GHIVE_ERR_UNSUPPORTED, "plan: function " + fn);
    }
    if (e.size() >= 2 && e.front() == '\'' && e.back() == '\'') {  // a string constant, on every row
      const std::string lit = e.substr(1, e.size() - 2);
      int st = alloc_col(GHIVE_STRING, t.rows, out);
      if (st != GHIVE_OK) return st;
      void *blob = nullptr;
      if ((st = ghive_malloc(ctx, lit.size() + 1, &blob)) != GHIVE_OK) return op_fail(st);
      owned.push_back(blob);
      if ((st = ghive_memcpy(ctx, blob, lit.c_str(), lit.size() + 1, 0)) != GHIVE_OK) return op_fail(st);
      std::vector<int32_t> pairs(2 * t.rows);
      for (uint64_t r = 0; r < t.rows; ++r) {
        pairs[2 * r] = 0;
        pairs[2 * r + 1] = (int32_t)lit.size();
      }
      if (t.rows && (st = ghive_memcpy(ctx, out.aux, pairs.data(), 8 * t.rows, 0)) != GHIVE_OK) return op_fail(st);
      out.data = blob;
      out.char_size = lit.size();
      return GHIVE_OK;
    }
    static const std::regex num_re(R"(^-?\d+(\.\d*)?([eE][-+]?\d+)?[DdLlSsYy]?$)");
    if (std::regex_match(e, num_re)) {  // a numeric constant, on every row
      ghive_column one;
      int st = operand(e, t, one);
      if (st != GHIVE_OK) return st;
      if ((st = alloc_col(one.type, t.rows, out)) != GHIVE_OK) return st;
      const size_t w = (one.type == GHIVE_LONG || one.type == GHIVE_DOUBLE) ? 8 : 4;
      std::vector<uint8_t> v(w), host(w * t.rows);
      if ((st = ghive_memcpy(ctx, v.data(), one.data, w, 1)) != GHIVE_OK) return op_fail(st);
      for (uint64_t r = 0; r < t.rows; ++r) memcpy(host.data() + r * w, v.data(), w);
      if (t.rows && (st = ghive_memcpy(ctx, out.data, host.data(), host.size(), 0)) != GHIVE_OK) return op_fail(st);
      return GHIVE_OK;
    }
    if (std::regex_match(e, m, val_re)) {  // values follow the keys of a reduce input
      const int c = std::stoi(m[1].str()) + t.reduce_keys;
      if (t.reduce_keys < 0 || c >= (int)t.cols.size())
        return fail(GHIVE_ERR_INVALID, "plan: " + e + " does not name a column of the reduce input");
      out = t.cols[c];
      return GHIVE_OK;
    }
    return fail(GHIVE_ERR_UNSUPPORTED, "plan: expression '" + e + "'");
  }
  int sub_operand(const std::string &e, const ExecTable &t, ghive_column &c) {
    static const std::regex num_re(R"(^-?\d.*$)");
    const std::string x = trim(e);
    return std::regex_match(x, num_re) ? operand(x, t, c) : expression(x, t, c);
  }

  // Releases the buffers this execution made after position `mark` (a fused attempt given up).
  void release_from(size_t mark) {
    for (size_t i = mark; i < owned.size(); ++i) ghive_free(ctx, owned[i]);
    owned.resize(mark);
  }

  // GroupBy `node` and the chain below it (shape S) as one ghive_fragment_execute. `done` stays
  // false, with the reason in `why`, when the data rule the fused form out (column types, values
  // beyond 32 bits, dimension keys that are not unique); the caller then runs the operator chain.
  int fused(int node, const FuseShape &S, ExecTable &out, bool &done, std::string &why) {
    done = false;
    const Op &g = P.ops[node];
    const size_t mark = owned.size();
    std::vector<ghive_hash_table *> tables;
    auto give_up = [&](const std::string &w) {
      for (ghive_hash_table *h : tables) ghive_hash_table_free(ctx, h);
      release_from(mark);
      why = w;
      return GHIVE_OK;
    };
    ExecTable fact;
    int st = from_input(S.scan->table, vertex, S.scan->outputs, S.scan->outputs, fact, true);
    if (st != GHIVE_OK) return st;
    std::vector<ExecTable> dims(S.joins.size());
    for (size_t j = 0; j < S.joins.size(); ++j)
      if ((st = from_input(S.joins[j].dim->vertex, "", {}, S.joins[j].dim->outputs, dims[j], true)) != GHIVE_OK)
        return st;
    // 32-bit copies of the LONG columns the kernel streams (checked: every value must fit)
    struct Narrow {
      const int64_t *src;
      uint64_t n;
      int32_t *dst;
    };
    std::vector<Narrow> narrows;
    std::string bad;
    auto as_i32 = [&](const ghive_column &c, uint64_t rows, ghive_column &v) -> int {
      if (c.type == GHIVE_INT && ((uintptr_t)c.data & 15) == 0) {
        v = c;
        return GHIVE_OK;
      }
      if (c.type != GHIVE_INT && c.type != GHIVE_LONG) {
        bad = "a joined or aggregated column is neither INT nor LONG";
        return GHIVE_OK;
      }
      ghive_column n;
      int s = alloc_col(GHIVE_INT, rows, n);
      if (s != GHIVE_OK) return s;
      if (c.type == GHIVE_INT) {  // an unaligned INT column: a copy the kernel can stream
        if (rows && (s = ghive_memcpy(ctx, n.data, c.data, 4 * rows, 2)) != GHIVE_OK) return op_fail(s);
      } else {
        narrows.push_back({static_cast<const int64_t *>(c.data), rows, static_cast<int32_t *>(n.data)});
      }
      v = n;
      return GHIVE_OK;
    };
    std::vector<ghive_column> fcols = fact.cols;  // the scan's columns, then the 32-bit copies
    std::map<int, uint32_t> i32_of;               // scan column -> its INT column in fcols
    auto fact_i32 = [&](int c, uint32_t &idx) -> int {
      auto it = i32_of.find(c);
      if (it != i32_of.end()) {
        idx = it->second;
        return GHIVE_OK;
      }
      ghive_column v;
      const int s = as_i32(fact.cols[c], fact.rows, v);
      if (s != GHIVE_OK || !bad.empty()) return s;
      if (v.data == fact.cols[c].data) {
        idx = (uint32_t)c;
      } else {
        idx = (uint32_t)fcols.size();
        fcols.push_back(v);
      }
      i32_of[c] = idx;
      return GHIVE_OK;
    };
    std::vector<ghive_fragment_join> fj(S.joins.size());
    std::vector<ghive_column> dkey(S.joins.size()), dpay(S.joins.size());
    for (size_t j = 0; j < S.joins.size() && bad.empty(); ++j) {
      const FuseJoin &J = S.joins[j];
      if (J.dim_key >= (int)dims[j].cols.size() || J.dim_pay >= (int)dims[j].cols.size())
        return give_up("a dimension input has too few columns");
      if ((st = fact_i32(J.fact_col, fj[j].fact_key_col)) != GHIVE_OK) return st;
      if ((st = as_i32(dims[j].cols[J.dim_key], dims[j].rows, dkey[j])) != GHIVE_OK) return st;
      if (J.dim_pay >= 0 && (st = as_i32(dims[j].cols[J.dim_pay], dims[j].rows, dpay[j])) != GHIVE_OK) return st;
    }
    // aggregates: (type, expression, inputs) and the output type the operator chain would give
    std::vector<ghive_fragment_agg> fa(S.aggs.size());
    std::vector<int> agg_type(S.aggs.size());
    for (size_t a = 0; a < S.aggs.size() && bad.empty(); ++a) {
      const FuseAgg &A = S.aggs[a];
      ghive_fragment_agg &E = fa[a];
      E.type = A.type;
      if (A.arg.kind == Sym::kNone) {
        E.expr = GHIVE_EXPR_ONE;
        agg_type[a] = GHIVE_LONG;
        continue;
      }
      const int ta = fact.logical[A.arg.kind == Sym::kMul ? A.arg.a : A.arg.col];
      const int tb = A.arg.kind == Sym::kMul ? fact.logical[A.arg.b] : ta;
      E.expr = A.arg.kind == Sym::kMul ? GHIVE_EXPR_MUL : GHIVE_EXPR_COL;
      if (ta == GHIVE_DOUBLE && tb == GHIVE_DOUBLE) {
        E.a = (uint32_t)(A.arg.kind == Sym::kMul ? A.arg.a : A.arg.col);
        E.b = (uint32_t)(A.arg.kind == Sym::kMul ? A.arg.b : A.arg.col);
        agg_type[a] = A.type == GHIVE_AGG_CNT ? GHIVE_LONG : GHIVE_DOUBLE;
        continue;
      }
      if (!ghive::is_int_type(ta) || !ghive::is_int_type(tb)) return give_up("an aggregate over a column that is neither INT, LONG nor DOUBLE");
      // a product of two INT columns wraps at 32 bits in Hive; the fused kernel multiplies in 64
      if (A.arg.kind == Sym::kMul && ta == GHIVE_INT && tb == GHIVE_INT) return give_up("INT * INT product");
      if (A.arg.kind == Sym::kMul) {
        if ((st = fact_i32(A.arg.a, E.a)) != GHIVE_OK || (st = fact_i32(A.arg.b, E.b)) != GHIVE_OK) return st;
      } else if ((st = fact_i32(A.arg.col, E.a)) != GHIVE_OK) {
        return st;
      }
      agg_type[a] = (A.type == GHIVE_AGG_MIN || A.type == GHIVE_AGG_MAX) && A.arg.kind == Sym::kFact ? ta : GHIVE_LONG;
    }
    std::vector<ghive_fragment_key> fk(S.keys.size());
    std::vector<int> key_type(S.keys.size());
    for (size_t k = 0; k < S.keys.size() && bad.empty(); ++k) {
      const Sym &s = S.keys[k];
      if (s.kind == Sym::kDim) {
        fk[k] = ghive_fragment_key{GHIVE_SRC_JOIN, (uint32_t)s.join};
        key_type[k] = dims[s.join].logical[s.col];
      } else {
        fk[k].source = GHIVE_SRC_FACT;
        if ((st = fact_i32(s.col, fk[k].index)) != GHIVE_OK) return st;
        key_type[k] = fact.logical[s.col];
      }
    }
    if (!bad.empty()) return give_up(bad);
    if (!narrows.empty()) {  // all range checks read back with one wait
      void *flags = nullptr;
      if ((st = ghive_malloc(ctx, 8 * narrows.size(), &flags)) != GHIVE_OK) return op_fail(st);
      owned.push_back(flags);
      uint64_t *fl = static_cast<uint64_t *>(flags);
      if ((st = ghive::fill_u64(ctx, fl, 0, narrows.size())) != GHIVE_OK) return op_fail(st);
      for (size_t i = 0; i < narrows.size(); ++i)
        if ((st = ghive::narrow_long(ctx, narrows[i].src, narrows[i].n, narrows[i].dst, fl + i)) != GHIVE_OK)
          return op_fail(st);
      std::vector<uint64_t> host(narrows.size());
      if ((st = ghive_memcpy(ctx, host.data(), flags, 8 * narrows.size(), 1)) != GHIVE_OK) return op_fail(st);
      for (uint64_t f : host)
        if (f) return give_up("a LONG column holds values beyond 32 bits");
    }
    // dimension tables below the regional-build size: built by one call (one fill launch, one
    // statistics launch for all of them); a larger one takes the regional form, one by one
    bool batch = S.joins.size() >= 1 && S.joins.size() <= GHIVE_MAX_FILTERED_BUILDS;
    for (size_t j = 0; j < S.joins.size() && batch; ++j)
      batch = ctx->tune.hash_regional_rows <= 0 || dims[j].rows < (uint64_t)ctx->tune.hash_regional_rows;
    if (batch) {
      std::vector<std::array<ghive_column, 2>> dcols(S.joins.size());
      std::vector<ghive_table> dt(S.joins.size());
      std::vector<ghive_filtered_build> jobs(S.joins.size());
      std::vector<ghive_hash_table *> hs(S.joins.size(), nullptr);
      for (size_t j = 0; j < S.joins.size(); ++j) {
        const bool pay = S.joins[j].dim_pay >= 0;
        dcols[j][0] = dkey[j];
        if (pay) dcols[j][1] = dpay[j];
        for (uint32_t c = 0; c < (pay ? 2u : 1u); ++c) dcols[j][c].row_num = dims[j].rows;
        dt[j] = ghive_table{dcols[j].data(), pay ? 2u : 1u, 0, dims[j].rows};
        memset(&jobs[j], 0, sizeof jobs[j]);
        jobs[j].table = &dt[j];
        jobs[j].key_col = 0;
        jobs[j].payload_col = pay ? 1 : -1;
        jobs[j].load_factor = 0.5f;
      }
      if ((st = ghive_hash_build_filtered(ctx, jobs.data(), (uint32_t)jobs.size(), hs.data())) != GHIVE_OK) {
        if (st == GHIVE_ERR_UNSUPPORTED) return give_up(ghive_last_error(ctx));
        return op_fail(st);
      }
      for (size_t j = 0; j < S.joins.size(); ++j) {
        tables.push_back(hs[j]);
        fj[j].table = hs[j];
      }
    }
    for (size_t j = 0; j < S.joins.size() && !batch; ++j) {
      ghive_hash_table *h = nullptr;
      st = ghive::hash_build_deferred(ctx, &dkey[j], nullptr, dims[j].rows, S.joins[j].dim_pay >= 0 ? &dpay[j] : nullptr,
                                      0.5f, &h);
      if (st != GHIVE_OK) {
        if (st == GHIVE_ERR_UNSUPPORTED) return give_up(ghive_last_error(ctx));
        for (ghive_hash_table *t : tables) ghive_hash_table_free(ctx, t);
        return op_fail(st);
      }
      tables.push_back(h);
      fj[j].table = h;
    }
    // the fact predicate, on the scan's columns; IS NOT NULL of a probed key is implied by the join
    std::vector<ghive_pred> pred;
    uint32_t proot = 0;
    if (S.filter && !S.filter->predicate.empty()) {
      std::vector<const char *> cn;
      for (auto &n : S.filter_names) cn.push_back(n.c_str());
      std::vector<ghive_pred> nodes(256);
      uint32_t nn = 0, root = 0;
      char perr[256] = {0};
      st = ghive_parse_predicate(S.filter->predicate.c_str(), cn.data(), (uint32_t)cn.size(), nodes.data(),
                                 (uint32_t)nodes.size(), &nn, &root, perr, sizeof perr);
      if (st != GHIVE_OK) {
        for (ghive_hash_table *t : tables) ghive_hash_table_free(ctx, t);
        return fail(st, "plan: " + S.filter->id + ": " + perr);
      }
      std::set<uint32_t> probed;
      for (const FuseJoin &J : S.joins) probed.insert((uint32_t)J.fact_col);
      for (uint32_t i = 0; i < nn; ++i)
        for (int c = 0; c < nodes[i].pred_col_num && c < 2; ++c) {
          const uint32_t x = nodes[i].filter_col[c];
          if (x >= S.filter_syms.size() || S.filter_syms[x].kind != Sym::kFact)
            return give_up("the filter reads a computed column");
          nodes[i].filter_col[c] = (uint32_t)S.filter_syms[x].col;
        }
      std::vector<int> seen(nn, 0);
      std::function<int(uint32_t)> keep = [&](uint32_t i) -> int {  // new index; -1: always true
        if (i >= nn || seen[i]++ > 0) return -2;                       // malformed or shared node
        ghive_pred p = nodes[i];
        if (p.pred_col_num == 0 && (p.mode == GHIVE_FILTER_AND || p.mode == GHIVE_FILTER_OR)) {
          const int a = keep((uint32_t)p.children[0]), b = keep((uint32_t)p.children[1]);
          if (a == -2 || b == -2) return -2;
          if (p.mode == GHIVE_FILTER_AND) {
            if (a < 0) return b;
            if (b < 0) return a;
          } else if (a < 0 || b < 0) {
            return -1;
          }
          p.children[0] = a;
          p.children[1] = b;
        } else if (p.mode == GHIVE_FILTER_NOT_NULL && p.pred_col_num == 1 && probed.count(p.filter_col[0])) {
          return -1;
        }
        pred.push_back(p);
        return (int)pred.size() - 1;
      };
      const int r = keep(root);
      if (r == -2) return give_up("an unexpected predicate tree");
      if (r >= 0) proot = (uint32_t)r;
      else pred.clear();
    }
    ghive_fragment f;
    memset(&f, 0, sizeof f);
    f.fact = ghive_table{fcols.data(), (uint32_t)fcols.size(), 0, fact.rows};
    f.filter = pred.empty() ? nullptr : pred.data();
    f.filter_nnodes = (uint32_t)pred.size();
    f.filter_root = proot;
    f.joins = fj.data();
    f.njoins = (uint32_t)fj.size();
    f.keys = fk.data();
    f.nkeys = (uint32_t)fk.size();
    f.aggs = fa.data();
    f.naggs = (uint32_t)fa.size();
    const size_t nout = fk.size() + fa.size();
    uint64_t cap = std::max<uint64_t>(1, std::min<uint64_t>(fact.rows, 1u << 16)), groups = 0;
    std::vector<ghive_column> oc(nout);
    for (int attempt = 0;; ++attempt) {
      for (size_t c = 0; c < nout; ++c) {
        if ((st = alloc_col(c < fk.size() ? GHIVE_INT : GHIVE_LONG, cap, oc[c])) != GHIVE_OK) {
          for (ghive_hash_table *t : tables) ghive_hash_table_free(ctx, t);
          return st;
        }
      }
      st = ghive_fragment_execute(ctx, &f, oc.data(), cap, &groups);
      if (st == GHIVE_ERR_CAPACITY && attempt == 0 && groups > cap) {
        cap = groups;
        continue;
      }
      break;
    }
    ghive_stage_times times{};
    (void)ghive_last_stage_times(ctx, &times);
    for (ghive_hash_table *t : tables) ghive_hash_table_free(ctx, t);
    tables.clear();
    if (st == GHIVE_ERR_UNSUPPORTED) return give_up(ghive_last_error(ctx));
    if (st != GHIVE_OK) return op_error(st, g);
    // the operator chain's output types: LONG group keys, INT MIN/MAX of an INT column
    out.cols.clear();
    for (size_t c = 0; c < nout; ++c) {
      ghive_column x = oc[c];
      x.row_num = groups;
      const int want = c < fk.size() ? key_type[c] : agg_type[c - fk.size()];
      if (want != x.type) {
        ghive_column y;
        if ((st = alloc_col(want, groups, y)) != GHIVE_OK) return st;
        if ((st = ghive_select_cast(ctx, &x, want, groups, &y)) != GHIVE_OK) return op_error(st, g);
        y.row_num = groups;
        x = y;
      }
      out.cols.push_back(x);
    }
    out.rows = groups;
    out.names = g.outputs;
    ghive_plan_exec_info &I = ctx->plan_info;
    I.fused = 1;
    I.joins = (uint32_t)S.joins.size();
    uint32_t staged32 = 0;  // columns that arrived as 32-bit copies from the staging
    for (size_t c = 0; c < fact.cols.size(); ++c) staged32 += fact.logical[c] == GHIVE_LONG && fact.cols[c].type == GHIVE_INT;
    for (const ExecTable &d : dims)
      for (size_t c = 0; c < d.cols.size(); ++c) staged32 += d.logical[c] == GHIVE_LONG && d.cols[c].type == GHIVE_INT;
    I.narrowed_columns = (uint32_t)narrows.size() + staged32;
    I.has_filter = pred.empty() ? 0 : 1;
    I.fused_kernel_ms = times.main_kernel_ms;
    I.not_fused[0] = 0;
    done = true;
    return GHIVE_OK;
  }

  int run(int node, ExecTable &out) {
    const Op &o = P.ops[node];
    if (o.kind == "GroupBy" && o.vertex == vertex && ctx->tune.plan_fuse) {
      FuseShape S;
      std::string why;
      if (fuse_shape(P, vertex, node, S, why)) {
        bool done = false;
        const int st = fused(node, S, out, done, why);
        if (st != GHIVE_OK || done) return st;
      }
      snprintf(ctx->plan_info.not_fused, sizeof ctx->plan_info.not_fused, "%s: %s", o.id.c_str(), why.c_str());
    }
    if ((o.kind == "ReduceSink" || o.kind == "FileSink" || o.kind == "Union") && o.vertex != vertex) {
      const int st = from_input(o.vertex, "", {}, o.outputs, out);  // the result of another vertex
      if (o.kind == "ReduceSink") out.reduce_keys = (int)o.key_expressions.size();
      return st;
    }
    // a table scan reads "db@table", or -- as GTables names it -- this vertex's own input
    if (o.kind == "TableScan") return from_input(o.table, vertex, o.outputs, o.outputs, out);
    if (o.children.empty()) return fail(GHIVE_ERR_INVALID, "plan: " + o.id + " has no input");
    if (o.kind == "MapJoin" || o.kind == "MergeJoin") return join(o, out);
    ExecTable in;
    int st = run(o.children[0], in);
    if (st != GHIVE_OK) return st;
    if (o.kind == "Filter") {
      std::vector<const char *> names;
      for (auto &n : in.names) names.push_back(n.c_str());
      std::vector<ghive_pred> nodes(256);
      uint32_t nn = 0, root = 0;
      char perr[256] = {0};
      st = ghive_parse_predicate(o.predicate.c_str(), names.data(), (uint32_t)names.size(), nodes.data(),
                                 (uint32_t)nodes.size(), &nn, &root, perr, sizeof perr);
      if (st != GHIVE_OK) return fail(st, "plan: " + o.id + ": " + perr);
      out.names = in.names;
      out.cols.resize(in.cols.size());
      for (size_t c = 0; c < in.cols.size(); ++c)
        if ((st = alloc_col(in.cols[c].type, in.rows, out.cols[c])) != GHIVE_OK) return st;
      ghive_table tv = in.view();
      uint64_t cnt = 0;
      st = ghive_filter_compact(ctx, &tv, nodes.data(), nn, root, out.cols.data(), &cnt);
      if (st != GHIVE_OK) return op_error(st, o);
      out.rows = cnt;
      for (auto &c : out.cols) c.row_num = cnt;
      return GHIVE_OK;
    }
    if (o.kind == "Select") {
      out.rows = in.rows;
      out.names = o.outputs;
      const std::vector<std::string> &ex = o.expressions.empty() ? o.outputs : o.expressions;
      if (ex.size() != o.outputs.size())
        return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": expressions and outputs differ in number");
      for (const std::string &e : ex) {
        ghive_column c;
        if ((st = expression(e, in, c)) != GHIVE_OK) return st;
        out.cols.push_back(c);
      }
      return GHIVE_OK;
    }
    if (o.kind == "GroupBy") {
      std::vector<uint32_t> keys(o.key_offsets.begin(), o.key_offsets.end());
      std::vector<ghive_agg> aggs;
      for (const auto &a : o.agg_descs) aggs.push_back(ghive_agg{a[0], (uint32_t)a[1]});
      for (uint32_t k : keys)
        if (k >= in.cols.size()) return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": key column beyond the input");
      if (keys.size() + aggs.size() != o.outputs.size())
        return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": keys and aggregates do not match its outputs");
      const uint64_t cap = in.rows ? in.rows : 1;
      out.cols.resize(keys.size() + aggs.size());
      for (size_t k = 0; k < keys.size(); ++k)
        if ((st = alloc_col(in.cols[keys[k]].type, cap, out.cols[k])) != GHIVE_OK) return st;
      for (size_t a = 0; a < aggs.size(); ++a) {
        if (aggs[a].index >= in.cols.size()) return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": aggregate column");
        const int ot = ghive_agg_output_type(aggs[a].type, in.cols[aggs[a].index].type);
        if (ot < 0) return fail(GHIVE_ERR_UNSUPPORTED, "plan: " + o.id + ": aggregate type");
        if ((st = alloc_col(ot, cap, out.cols[keys.size() + a])) != GHIVE_OK) return st;
      }
      ghive_table tv = in.view();
      uint64_t groups = 0;
      st = ghive_groupby(ctx, &tv, nullptr, in.rows, keys.data(), (uint32_t)keys.size(), aggs.data(),
                         (uint32_t)aggs.size(), out.cols.data(), cap, &groups);
      if (st != GHIVE_OK) return op_error(st, o);
      out.rows = groups;
      out.names = o.outputs;
      return GHIVE_OK;
    }
    if (o.kind == "Limit") {
      out = in;
      // In the plan's root vertex with a final ORDER BY, LIMIT keeps the first rows of the sorted
      // result: ghive_plan_execute applies it after the sort, so nothing is cut here. (The
      // reference's LimitOperator.cu:11 truncates before the sink sorts, which keeps an arbitrary
      // subset of unsorted input such as group-by output.)
      const auto vr = P.vertex_root.find(vertex);
      const bool after_sort = vr != P.vertex_root.end() && vr->second == P.root && P.ops[P.root].require_sort &&
                              !P.ops[P.root].sort_orders.empty();
      if (!after_sort && o.limit >= 0 && (uint64_t)o.limit < out.rows) out.rows = (uint64_t)o.limit;
      for (auto &c : out.cols) c.row_num = out.rows;
      return GHIVE_OK;
    }
    if (o.kind == "ReduceSink" || o.kind == "FileSink") {
      // the vertex's own sink: its columns keys first
      out.rows = in.rows;
      for (const std::string &n : o.outputs) {
        const int c = in.find(n);
        if (c < 0) return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": no column '" + n + "' below it");
        out.names.push_back(n);
        out.cols.push_back(in.cols[c]);
      }
      if (o.outputs.empty()) out = in;
      return GHIVE_OK;
    }
    return fail(GHIVE_ERR_UNSUPPORTED, "plan: " + o.kind + " operator " + o.id);
  }

  int join(const Op &o, ExecTable &out) {
    const size_t nt = o.join_inputs.size();
    if (nt < 2) return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": fewer than two inputs");
    std::vector<ExecTable> ts(nt);
    for (size_t t = 0; t < nt; ++t) {
      const auto it = P.by_id.find(o.join_inputs[t]);
      if (it == P.by_id.end()) return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": unknown input");
      int st = run(it->second, ts[t]);
      if (st != GHIVE_OK) return st;
      if (o.join_cols[t].size() != 1)
        return fail(GHIVE_ERR_UNSUPPORTED, "plan: " + o.id + ": join on more than one key column");
    }
    std::vector<ghive_table> views;
    for (auto &t : ts) views.push_back(t.view());
    std::vector<uint32_t> jc;
    for (size_t t = 0; t < nt; ++t) jc.push_back((uint32_t)o.join_cols[t][0]);
    std::vector<ghive_join_cond> conds;
    for (const auto &c : o.join_types) conds.push_back(ghive_join_cond{c[0], (uint32_t)c[1], (uint32_t)c[2]});
    uint64_t card = 0;
    int st = ghive_join(ctx, views.data(), (uint32_t)nt, jc.data(), conds.data(), (uint32_t)conds.size(), nullptr, 0,
                        &card);
    if (st != GHIVE_OK && st != GHIVE_ERR_CAPACITY) return op_error(st, o);
    size_t width = nt;
    for (const auto &c : conds)
      if (c.type == GHIVE_LEFT_SEMI_JOIN) --width;
    void *offs = nullptr;
    if ((st = ghive_malloc(ctx, (card ? card : 1) * width * 4, &offs)) != GHIVE_OK) return op_error(st, o);
    owned.push_back(offs);
    uint64_t card2 = 0;
    st = ghive_join(ctx, views.data(), (uint32_t)nt, jc.data(), conds.data(), (uint32_t)conds.size(),
                    static_cast<int32_t *>(offs), card, &card2);
    if (st != GHIVE_OK) return op_error(st, o);
    // the maintained columns: from the Java side's map, else the join outputs found by name
    std::vector<std::vector<std::string>> in_names;
    for (auto &t : ts) in_names.push_back(t.names);
    std::vector<std::vector<uint32_t>> keep;
    if ((st = join_keep(o, in_names, keep, err)) != GHIVE_OK) return st;
    std::vector<const uint32_t *> kp;
    std::vector<uint32_t> kc;
    size_t ncols = 0;
    for (auto &k : keep) {
      kp.push_back(k.data());
      kc.push_back((uint32_t)k.size());
      ncols += k.size();
    }
    if (ncols == 0) ncols = 1;  // the reference keeps column 0 of the first input
    std::vector<ghive_column> cols(ncols);
    // output types follow the gathered columns: allocate per the plan's own order
    {
      size_t oi = 0;
      size_t first = 0;
      while (first < nt && keep[first].empty() && first < nt) ++first;
      for (size_t t = 0; t < nt; ++t)
        for (uint32_t c : keep[t])
          if ((st = alloc_col(ts[t].cols[c].type, card, cols[oi++])) != GHIVE_OK) return st;
      if (oi == 0 && (st = alloc_col(ts[0].cols[0].type, card, cols[0])) != GHIVE_OK) return st;
    }
    uint32_t got = 0;
    st = ghive_join_result(ctx, views.data(), (uint32_t)nt, conds.data(), (uint32_t)conds.size(), kp.data(),
                           kc.data(), static_cast<int32_t *>(offs), card, cols.data(), (uint32_t)cols.size(), &got);
    if (st != GHIVE_OK) return op_error(st, o);
    out.rows = card;
    out.cols.assign(cols.begin(), cols.begin() + got);
    out.names = o.outputs;
    if (out.names.size() != out.cols.size())
      return fail(GHIVE_ERR_INVALID, "plan: " + o.id + ": the maintained columns do not match its outputs");
    return GHIVE_OK;
  }

  // the plan's root: ORDER BY on the reduce keys (the vertex's Select expressions that read
  // KEY.reducesinkkeyN, else the leading columns), then LIMIT
  int finish_root(int root_node, ExecTable &t) {
    const Op &o = P.ops[root_node];
    if (!o.require_sort || o.sort_orders.empty() || t.rows == 0) return GHIVE_OK;
    const size_t nk = o.sort_orders.size();
    std::vector<uint32_t> keys(nk, UINT32_MAX);
    static const std::regex key_re(R"(^KEY\.reducesinkkey(\d+)$)");
    for (int node = root_node; node >= 0;) {
      const Op &x = P.ops[node];
      if (x.kind == "Select") {
        for (size_t k = 0; k < x.expressions.size() && k < x.outputs.size(); ++k) {
          std::smatch m;
          if (std::regex_match(x.expressions[k], m, key_re)) {
            const size_t j = (size_t)std::stoi(m[1].str());
            const int c = t.find(x.outputs[k]);
            if (j < nk && c >= 0) keys[j] = (uint32_t)c;
          }
        }
        break;
      }
      int next = -1;
      for (int c : x.children)
        if (P.ops[c].vertex == vertex && next < 0) next = c;
      node = next;
    }
    for (size_t j = 0; j < nk; ++j)
      if (keys[j] == UINT32_MAX) keys[j] = (uint32_t)j;
    void *perm = nullptr;
    int st = ghive_malloc(ctx, t.rows * 4, &perm);
    if (st != GHIVE_OK) return fail(st, std::string("plan: ") + ghive_last_error(ctx));
    owned.push_back(perm);
    ghive_table tv = t.view();
    st = ghive_sort(ctx, &tv, keys.data(), o.sort_orders.c_str(), (uint32_t)nk, static_cast<uint32_t *>(perm));
    if (st != GHIVE_OK) return op_error(st, o);
    std::vector<ghive_column> sorted(t.cols.size());
    for (size_t c = 0; c < t.cols.size(); ++c)
      if ((st = alloc_col(t.cols[c].type, t.rows, sorted[c])) != GHIVE_OK) return st;
    st = ghive_gather(ctx, &tv, static_cast<uint32_t *>(perm), t.rows, sorted.data());
    if (st != GHIVE_OK) return op_error(st, o);
    t.cols = sorted;
    return GHIVE_OK;
  }
};

}  // namespace

struct ghive_result {
  std::vector<std::string> names;
  std::vector<ghive_column> cols;
  uint64_t rows = 0;
  std::vector<void *> owned;
};

extern "C" {

int ghive_plan_parse(const char *plan, const char *extended, const char *maintain_json, char *err, uint64_t errcap,
                     ghive_plan **out) {
  const bool extended_only = (!plan || !*plan) && extended && *extended;
  if ((!plan && !extended_only) || !out) {
    set_err(err, errcap, "plan: NULL argument");
    return GHIVE_ERR_INVALID;
  }
  *out = nullptr;
  auto *P = new ghive_plan();
  std::string e;
  try {
    if (maintain_json && *maintain_json) {
      Json j{maintain_json, {}};
      if (!j.map3(P->maintain)) {
        delete P;
        set_err(err, errcap, "plan: malformed maintain-values JSON");
        return GHIVE_ERR_INVALID;
      }
    }
    int st = extended_only ? parse_extended_tree(*P, extended, e) : parse_tree(*P, plan, e);
    if (st != GHIVE_OK) {
      delete P;
      set_err(err, errcap, e);
      return st;
    }
    std::set<int> seen;
    assign_vertices(*P, P->root, P->ops[P->root].vertex, seen);
    std::set<int> done;
    propagate_outputs(*P, P->root, done);
    if (extended && !extended_only) link_extended(*P, extended_blocks(extended));
    std::set<int> resolved;
    resolve(*P, P->root, resolved);
    // the root sink performs the final ORDER BY with the order of the sink that feeds its vertex
    Op &root = P->ops[P->root];
    for (int node = P->root; node >= 0;) {
      int next = -1;
      for (int c : P->ops[node].children) {
        if (P->ops[c].kind == "ReduceSink") {
          root.require_sort = true;
          root.sort_orders = P->ops[c].sort_orders;
          next = -1;
          break;
        }
        if (next < 0) next = c;
      }
      if (root.require_sort) break;
      node = next;
    }
  } catch (const std::exception &ex) {
    delete P;
    set_err(err, errcap, std::string("plan: ") + ex.what());
    return GHIVE_ERR_INVALID;
  }
  *out = P;
  return GHIVE_OK;
}

int ghive_plan_describe(const ghive_plan *plan, char *buf, uint64_t cap, uint64_t *len) {
  if (!plan || !len) return GHIVE_ERR_INVALID;
  const std::string s = describe(*plan);
  *len = s.size();
  if (!buf || cap < s.size() + 1) return GHIVE_ERR_CAPACITY;
  memcpy(buf, s.c_str(), s.size() + 1);
  return GHIVE_OK;
}

void ghive_plan_free(ghive_plan *plan) { delete plan; }

}  // extern "C"

extern "C" {

int ghive_plan_execute(ghive_ctx *ctx, const ghive_plan *plan, const char *vertex, const ghive_plan_input *inputs,
                       uint32_t ninputs, char *err, uint64_t errcap, ghive_result **out) {
  return ghive::plan_execute_staged(ctx, plan, vertex, inputs, ninputs, nullptr, err, errcap, out);
}

}  // extern "C"

namespace ghive {

void plan_narrow_candidates(const ghive_plan *plan, const char *vertex,
                            std::map<std::string, std::set<uint32_t>> &out) {
  out.clear();
  if (!plan || !vertex) return;
  for (size_t i = 0; i < plan->ops.size(); ++i) {
    const Op &o = plan->ops[i];
    if (o.kind != "GroupBy" || o.vertex != vertex) continue;
    FuseShape S;
    std::string why;
    if (!fuse_shape(*plan, vertex, (int)i, S, why)) continue;
    // the scanned table arrives named by the vertex being run (GTables) or by its table name
    std::set<uint32_t> fact;
    for (const FuseJoin &J : S.joins) fact.insert((uint32_t)J.fact_col);
    for (const Sym &k : S.keys)
      if (k.kind == Sym::kFact) fact.insert((uint32_t)k.col);
    for (const FuseAgg &A : S.aggs) {
      if (A.arg.kind == Sym::kFact) fact.insert((uint32_t)A.arg.col);
      if (A.arg.kind == Sym::kMul) {
        fact.insert((uint32_t)A.arg.a);
        fact.insert((uint32_t)A.arg.b);
      }
    }
    out[vertex].insert(fact.begin(), fact.end());
    out[S.scan->table].insert(fact.begin(), fact.end());
    for (const FuseJoin &J : S.joins) {
      out[J.dim->vertex].insert((uint32_t)J.dim_key);
      if (J.dim_pay >= 0) out[J.dim->vertex].insert((uint32_t)J.dim_pay);
    }
  }
}

int plan_execute_staged(ghive_ctx *ctx, const ghive_plan *plan, const char *vertex, const ghive_plan_input *inputs,
                        uint32_t ninputs, const std::map<std::string, std::vector<char>> *narrowed, char *err,
                        uint64_t errcap, ghive_result **out) {
  if (!ctx || !plan || !vertex || !out || (ninputs && !inputs)) {
    set_err(err, errcap, "plan_execute: NULL argument");
    return GHIVE_ERR_INVALID;
  }
  *out = nullptr;
  auto vit = plan->vertex_root.find(vertex);
  if (vit == plan->vertex_root.end()) {
    set_err(err, errcap, std::string("plan_execute: no vertex '") + vertex + "'");
    return GHIVE_ERR_INVALID;
  }
  memset(&ctx->plan_info, 0, sizeof ctx->plan_info);
  Executor X{ctx, *plan, vertex, {}, {}, {}};
  X.narrowed = narrowed;
  for (uint32_t i = 0; i < ninputs; ++i)
    if (inputs[i].name) X.inputs[inputs[i].name] = &inputs[i];
  ExecTable t;
  int st;
  try {
    st = X.run(vit->second, t);
    if (st == GHIVE_OK && vit->second == plan->root) {
      st = X.finish_root(vit->second, t);
      // a LIMIT below the root applies after the ORDER BY
      for (int node = plan->root; st == GHIVE_OK && node >= 0;) {
        const Op &o = plan->ops[node];
        if (o.kind == "Limit" && o.limit >= 0 && (uint64_t)o.limit < t.rows) t.rows = (uint64_t)o.limit;
        int next = -1;
        for (int c : o.children)
          if (plan->ops[c].vertex == vertex && next < 0) next = c;
        node = next;
      }
    }
  } catch (const std::exception &ex) {
    st = GHIVE_ERR_INVALID;
    X.err = std::string("plan_execute: ") + ex.what();
  }
  if (st == GHIVE_OK) (void)ghive_sync(ctx);
  if (st != GHIVE_OK) {
    for (void *p : X.owned) ghive_free(ctx, p);
    set_err(err, errcap, X.err);
    return st;
  }
  auto *R = new ghive_result();
  R->names = t.names;
  R->cols = t.cols;
  R->rows = t.rows;
  for (auto &c : R->cols) c.row_num = t.rows;
  R->owned = X.owned;  // intermediates included; released together
  *out = R;
  return GHIVE_OK;
}

}  // namespace ghive

extern "C" {

int ghive_result_info(const ghive_result *r, uint64_t *rows, uint32_t *ncols) {
  if (!r) return GHIVE_ERR_INVALID;
  if (rows) *rows = r->rows;
  if (ncols) *ncols = (uint32_t)r->cols.size();
  return GHIVE_OK;
}

int ghive_result_column(const ghive_result *r, uint32_t i, ghive_column *col, const char **name) {
  if (!r || i >= r->cols.size()) return GHIVE_ERR_INVALID;
  if (col) *col = r->cols[i];
  if (name) *name = r->names[i].c_str();
  return GHIVE_OK;
}

void ghive_result_free(ghive_ctx *ctx, ghive_result *r) {
  if (!r) return;
  for (void *p : r->owned) ghive_free(ctx, p);
  delete r;
}

}  // extern "C"
