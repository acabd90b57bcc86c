// Host side of predicates: Hive EXPLAIN predicate text -> ghive_pred tree (the job of
// FilterOperator::generatePredicate, gpu_impl/src/Operator/FilterOperator.cu:227-487), the
// FilterPredicate::toString(0) dump (FilterPredicate.cu:670-710), and the flattening of a tree
// into the DevPred program the kernels evaluate.
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "pred_host.hpp"

namespace ghive {
namespace {

struct Parser {
  std::vector<std::string> tok;
  size_t pos = 0;
  const char *const *names;
  uint32_t ncols;
  std::vector<ghive_pred> nodes;
  std::string err;

  static bool is_op(const std::string &t) { return t == "and" || t == "or"; }

  bool parse_literal(const std::string &s, int &dtype, int64_t &l, double &d, int32_t &i, std::string &str) {
    auto all = [&](size_t b, size_t e, bool dot) {
      if (b >= e) return false;
      bool digit = false;
      for (size_t k = b; k < e; ++k) {
        if (s[k] >= '0' && s[k] <= '9') digit = true;
        else if (!(dot && s[k] == '.')) return false;
      }
      return digit;
    };
    size_t b = (!s.empty() && s[0] == '-') ? 1 : 0;
    if (s.size() > 1 && s.back() == 'D' && all(b, s.size() - 1, true)) {
      dtype = 1;
      d = std::stod(s.substr(0, s.size() - 1));
      return true;
    }
    if (s.size() > 1 && s.back() == 'L' && all(b, s.size() - 1, false)) {
      dtype = 0;
      l = std::stoll(s.substr(0, s.size() - 1));
      return true;
    }
    if (all(b, s.size(), false)) {
      long long v = std::stoll(s);
      if (v >= INT32_MIN && v <= INT32_MAX) {
        dtype = 2;
        i = (int32_t)v;
      } else {  // too wide for the reference's int literal: keep the value as a long
        dtype = 0;
        l = v;
      }
      return true;
    }
    if (s.size() >= 2 && s.front() == '\'' && s.back() == '\'') {  // '...' as the regex '(.*)'
      dtype = 3;
      str = s.substr(1, s.size() - 2);
      return true;
    }
    return false;
  }

  int col_index(const std::string &name) {
    for (uint32_t k = 0; k < ncols; ++k)
      if (names[k] && name == names[k]) return (int)k;
    return -1;
  }

  // One comparison such as "lo_quantity < 25.0D" or "x BETWEEN 1 AND 3".
  int basic(const std::string &s) {
    ghive_pred p;
    memset(&p, 0, sizeof p);
    p.children[0] = p.children[1] = -1;
    p.pred_col_num = 1;
    std::string col, v1, v2;
    static const char *kNotNull = " is not null";
    size_t L = strlen(kNotNull);
    // positions of `op` outside '...' literals
    auto unquoted = [&](const char *op, bool last, size_t from) -> size_t {
      size_t found = std::string::npos;
      const size_t L2 = strlen(op);
      bool q = false;
      for (size_t k = 0; k < s.size(); ++k) {
        if (s[k] == '\'') q = !q;
        if (q || k < from || s.compare(k, L2, op) != 0) continue;
        found = k;
        if (!last) break;
      }
      return found;
    };
    auto split2 = [&](const char *op) -> bool {  // greedy: the last occurrence, like (.*) OP (.*)
      size_t at = unquoted(op, true, 0);
      if (at == std::string::npos || at == 0) return false;
      col = s.substr(0, at);
      v1 = s.substr(at + strlen(op));
      return !v1.empty();
    };
    auto between = [&](const char *kw) -> bool {
      size_t at = unquoted(kw, false, 0);
      if (at == std::string::npos || at == 0) return false;
      size_t an = unquoted(" AND ", false, at + strlen(kw));
      if (an == std::string::npos) return false;
      col = s.substr(0, at);
      v1 = s.substr(at + strlen(kw), an - at - strlen(kw));
      v2 = s.substr(an + 5);
      return true;
    };
    int nparam = 1;
    if (s.size() > L && s.compare(s.size() - L, L, kNotNull) == 0) {
      p.mode = GHIVE_FILTER_NOT_NULL;
      col = s.substr(0, s.size() - L);
      nparam = 0;
    } else if (between(" NOT BETWEEN ")) {  // checked before BETWEEN, which would shadow it
      p.mode = GHIVE_FILTER_NOT_RANGE;
      nparam = 2;
    } else if (between(" BETWEEN ")) {
      p.mode = GHIVE_FILTER_EQ_RANGE;
      nparam = 2;
    } else if (split2(" = ")) {
      p.mode = GHIVE_FILTER_EQ;
    } else if (split2(" <= ")) {
      p.mode = GHIVE_FILTER_LE;
    } else if (split2(" < ")) {
      p.mode = GHIVE_FILTER_LT;
    } else if (split2(" >= ")) {
      p.mode = GHIVE_FILTER_GE;
    } else if (split2(" > ")) {
      p.mode = GHIVE_FILTER_GT;
    } else if (split2(" <> ")) {
      p.mode = GHIVE_FILTER_NOT_EQ;
    } else {
      err = "unsupported predicate: '" + s + "'";
      return -1;
    }
    // UDFToDouble(x), UDFToLong(x), ...: a cast of a numeric column, which the comparison of
    // values already performs (Hive writes it around a column compared with another type's literal)
    for (const char *cast : {"UDFToDouble(", "UDFToFloat(", "UDFToLong(", "UDFToInteger("}) {
      const size_t L3 = strlen(cast);
      if (col.size() > L3 + 1 && col.compare(0, L3, cast) == 0 && col.back() == ')') col = col.substr(L3, col.size() - L3 - 1);
    }
    int c = col_index(col);
    if (c < 0) {
      err = "column '" + col + "' is not an output column of the operator";
      return -1;
    }
    p.filter_col[0] = (uint32_t)c;
    p.param_num = nparam;
    if (nparam == 0) {
      p.data_type = 4;
    } else {
      const std::string *vs[2] = {&v1, &v2};
      int first = -1;
      for (int k = 0; k < nparam; ++k) {
        int dt;
        int64_t l = 0;
        double d = 0;
        int32_t iv = 0;
        std::string str;
        if (!parse_literal(*vs[k], dt, l, d, iv, str)) {
          err = "unsupported literal '" + *vs[k] + "'";
          return -1;
        }
        if (dt == 3) {
          if (str.size() > (size_t)GHIVE_PRED_STR_MAX) {
            err = "string literal longer than " + std::to_string(GHIVE_PRED_STR_MAX) + " bytes: " + *vs[k];
            return -1;
          }
          p.string_sizes[k] = (int32_t)str.size();
          memcpy(p.string_params[k], str.data(), str.size());
        }
        if (first >= 0 && dt != first) {
          // mixed int/long literals of one BETWEEN: widen both to long
          if ((first == 0 || first == 2) && (dt == 0 || dt == 2)) {
            if (first == 2) p.long_params[0] = p.int_params[0];
            if (dt == 2) l = iv;
            dt = 0;
            first = 0;
          } else {
            err = "literal types of '" + s + "' do not match";
            return -1;
          }
        }
        first = dt;
        p.data_type = dt;
        p.long_params[k] = l;
        p.double_params[k] = d;
        p.int_params[k] = iv;
      }
    }
    nodes.push_back(p);
    return (int)nodes.size() - 1;
  }

  int item() {
    if (pos >= tok.size()) {
      err = "unexpected end of predicate";
      return -1;
    }
    if (tok[pos] == "(") {
      ++pos;
      int g = group();
      if (g < 0) return -1;
      if (pos >= tok.size() || tok[pos] != ")") {
        err = "missing ')'";
        return -1;
      }
      ++pos;
      return g;
    }
    std::string s;
    while (pos < tok.size() && tok[pos] != ")" && tok[pos] != "(" && !is_op(tok[pos])) {
      if (!s.empty()) s += " ";
      s += tok[pos++];
    }
    if (s.empty()) {
      err = "empty predicate";
      return -1;
    }
    return basic(s);
  }

  // A chain X1 op1 X2 op2 ... Xk folds from the right: acc = Xk; acc = op_j(acc, X_j) for
  // j = k-1..1 -- the tree shape GHive's stack parser builds (children[0] = the later part).
  int group() {
    std::vector<int> items;
    std::vector<int> ops;
    int it = item();
    if (it < 0) return -1;
    items.push_back(it);
    while (pos < tok.size() && is_op(tok[pos])) {
      ops.push_back(tok[pos] == "and" ? GHIVE_FILTER_AND : GHIVE_FILTER_OR);
      ++pos;
      it = item();
      if (it < 0) return -1;
      items.push_back(it);
    }
    int acc = items.back();
    for (int j = (int)items.size() - 2; j >= 0; --j) {
      ghive_pred p;
      memset(&p, 0, sizeof p);
      p.mode = ops[j];
      p.pred_col_num = 0;
      p.param_num = 0;
      p.data_type = 4;
      p.children[0] = acc;
      p.children[1] = items[j];
      nodes.push_back(p);
      acc = (int)nodes.size() - 1;
    }
    return acc;
  }
};

void dump(const ghive_pred *nodes, uint32_t nnodes, uint32_t idx, int level, std::string &out, int guard) {
  if (idx >= nnodes || guard > 64) return;
  const ghive_pred &p = nodes[idx];
  for (int i = 0; i < level; ++i) out += "  ";
  out += "predicate: " + std::to_string(p.mode) + "; ";
  bool logic = p.mode == GHIVE_FILTER_AND || p.mode == GHIVE_FILTER_OR;
  out += logic ? std::string("filterCol: x;") : "filterCol: " + std::to_string(p.filter_col[0]) + ";";
  out += "paramNum: " + std::to_string(p.param_num) + ";";
  out += "params: ";
  if (p.param_num == 0) {
    out += "x";
  } else {
    for (int i = 0; i < p.param_num && i < 2; ++i) {
      if (p.data_type == 0) out += std::to_string((long long)p.long_params[i]) + " ";
      else if (p.data_type == 1) out += std::to_string(p.double_params[i]) + " ";
      else if (p.data_type == 2) out += std::to_string(p.int_params[i]) + " ";
      else if (p.data_type == 3) {
        const int len = p.string_sizes[i] < 0 ? 0 : (p.string_sizes[i] > GHIVE_PRED_STR_MAX ? GHIVE_PRED_STR_MAX : p.string_sizes[i]);
        out += std::string(p.string_params[i], (size_t)len) + " ";
      }
    }
  }
  out += ";\n";
  if (logic && p.children[0] >= 0 && p.children[1] >= 0) {
    dump(nodes, nnodes, (uint32_t)p.children[0], level + 1, out, guard + 1);
    dump(nodes, nnodes, (uint32_t)p.children[1], level + 1, out, guard + 1);
  }
}

}  // namespace

int flatten_pred(ghive_ctx *ctx, const ghive_table *t, const ghive_pred *nodes, uint32_t nnodes,
                 uint32_t root, DevPred &out) {
  memset(&out, 0, sizeof out);
  if (!nodes || nnodes == 0 || root >= nnodes)
    return set_error(ctx, GHIVE_ERR_INVALID, "predicate: no nodes or bad root");
  std::vector<int> state(nnodes, 0), order;
  // iterative post-order with cycle detection
  std::vector<std::pair<uint32_t, int>> st{{root, 0}};
  while (!st.empty()) {
    const uint32_t idx = st.back().first;
    const int phase = st.back().second;
    const ghive_pred &p = nodes[idx];
    bool logic = p.mode == GHIVE_FILTER_AND || p.mode == GHIVE_FILTER_OR;
    if (phase == 0) {
      if (state[idx] == 1)
        return set_error(ctx, GHIVE_ERR_INVALID, "predicate: node %u is part of a cycle", idx);
      state[idx] = 1;
      st.back().second = 1;
      if (logic) {
        for (int c = 1; c >= 0; --c) {
          int ch = p.children[c];
          if (ch < 0 || (uint32_t)ch >= nnodes)
            return set_error(ctx, GHIVE_ERR_INVALID, "predicate: node %u has bad child %d", idx, ch);
          if (state[ch] == 1)
            return set_error(ctx, GHIVE_ERR_INVALID, "predicate: node %u is part of a cycle", (uint32_t)ch);
          if (state[ch] == 2) continue;  // shared sub-tree, already scheduled
          st.push_back({(uint32_t)ch, 0});
        }
      }
    } else {
      state[idx] = 2;
      order.push_back((int)idx);
      st.pop_back();
    }
  }
  if (order.size() > (size_t)kMaxPredNodes)
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "predicate: %zu nodes (max %d)", order.size(), kMaxPredNodes);
  std::vector<int> pos_of(nnodes, -1);
  std::vector<uint32_t> slot_col;
  auto slot = [&](uint32_t col) -> int {
    for (size_t k = 0; k < slot_col.size(); ++k)
      if (slot_col[k] == col) return (int)k;
    if ((int)slot_col.size() == kMaxPredCols) return -1;
    slot_col.push_back(col);
    return (int)slot_col.size() - 1;
  };
  int lit_used = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    const ghive_pred &p = nodes[order[i]];
    DevPredNode &n = out.nodes[i];
    pos_of[order[i]] = (int)i;
    n.mode = (int8_t)p.mode;
    if (p.mode == GHIVE_FILTER_AND || p.mode == GHIVE_FILTER_OR) {
      n.op = p.mode == GHIVE_FILTER_AND ? kAnd : kOr;
      n.a = (int8_t)pos_of[p.children[0]];
      n.b = (int8_t)pos_of[p.children[1]];
      continue;
    }
    if (p.mode < 0 || p.mode > GHIVE_FILTER_NOT_NULL)
      return set_error(ctx, GHIVE_ERR_INVALID, "predicate: mode %d", p.mode);
    if (p.filter_col[0] >= t->ncols)
      return set_error(ctx, GHIVE_ERR_INVALID, "predicate: column %u of %u", p.filter_col[0], t->ncols);
    const ghive_column &c0 = t->columns[p.filter_col[0]];
    const bool str0 = c0.type == GHIVE_STRING;
    int s0 = slot(p.filter_col[0]);
    if (s0 < 0) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "predicate: more than %d columns", kMaxPredCols);
    n.a = (int8_t)s0;
    if (p.mode == GHIVE_FILTER_NOT_NULL) {
      n.op = str0 ? kNotNullStr : kNotNull;
      continue;
    }
    if (p.pred_col_num == 2) {
      if (p.filter_col[1] >= t->ncols)
        return set_error(ctx, GHIVE_ERR_INVALID, "predicate: column %u of %u", p.filter_col[1], t->ncols);
      const ghive_column &c1 = t->columns[p.filter_col[1]];
      if (str0 != (c1.type == GHIVE_STRING))
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "predicate: a STRING column against a numeric column");
      int s1 = slot(p.filter_col[1]);
      if (s1 < 0) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "predicate: more than %d columns", kMaxPredCols);
      n.b = (int8_t)s1;
      n.op = str0 ? kColColStr : (is_int_type(c0.type) && is_int_type(c1.type)) ? kColColInt : kColColDbl;
      if (p.mode > GHIVE_FILTER_NOT_EQ || p.mode == GHIVE_FILTER_RANGE || p.mode == GHIVE_FILTER_EQ_RANGE)
        return set_error(ctx, GHIVE_ERR_INVALID, "predicate: mode %d needs literals", p.mode);
      continue;
    }
    if (p.pred_col_num != 1)
      return set_error(ctx, GHIVE_ERR_INVALID, "predicate: pred_col_num %d", p.pred_col_num);
    int dt = p.data_type;
    if (str0 || dt == 3) {
      if (!str0 || dt != 3)
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "predicate: a %s column against a %s literal",
                         str0 ? "STRING" : "numeric", dt == 3 ? "string" : "numeric");
      const bool two = p.mode == GHIVE_FILTER_RANGE || p.mode == GHIVE_FILTER_EQ_RANGE ||
                       p.mode == GHIVE_FILTER_NOT_RANGE || p.mode == GHIVE_FILTER_NOT_EQ_RANGE;
      if (p.param_num != (two ? 2 : 1))
        return set_error(ctx, GHIVE_ERR_INVALID, "predicate: mode %d with %d literals", p.mode, p.param_num);
      int64_t lit[2] = {0, 0};
      for (int k = 0; k < p.param_num; ++k) {
        const int len = p.string_sizes[k];
        if (len < 0 || len > GHIVE_PRED_STR_MAX)
          return set_error(ctx, GHIVE_ERR_INVALID, "predicate: string literal of %d bytes", len);
        if (lit_used + len > kPredLitBytes)
          return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "predicate: string literals above %d bytes", kPredLitBytes);
        memcpy(out.lit + lit_used, p.string_params[k], (size_t)len);
        lit[k] = ((int64_t)len << 32) | (int64_t)lit_used;
        lit_used += len;
      }
      n.op = kLeafStr;
      n.ia = lit[0];
      n.ib = lit[1];
      n.pad = p.param_num;
      continue;
    }
    if (dt < 0 || dt > 2) return set_error(ctx, GHIVE_ERR_INVALID, "predicate: data_type %d", dt);
    if (is_int_type(c0.type) && dt != 1) {
      n.op = kLeafInt;
      n.ia = dt == 0 ? p.long_params[0] : p.int_params[0];
      n.ib = dt == 0 ? p.long_params[1] : p.int_params[1];
    } else {
      n.op = kLeafDbl;
      if (dt == 1) { n.da = p.double_params[0]; n.db = p.double_params[1]; }
      else if (dt == 0) { n.da = (double)p.long_params[0]; n.db = (double)p.long_params[1]; }
      else { n.da = p.int_params[0]; n.db = p.int_params[1]; }
    }
  }
  out.nnodes = (int32_t)order.size();
  out.ncols = (int32_t)slot_col.size();
  for (size_t k = 0; k < slot_col.size(); ++k) {
    out.col_data[k] = t->columns[slot_col[k]].data;
    out.col_aux[k] = static_cast<const int32_t *>(t->columns[slot_col[k]].aux);
    out.col_type[k] = t->columns[slot_col[k]].type;
  }
  return GHIVE_OK;
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_parse_predicate(const char *text, const char *const *col_names, uint32_t ncols,
                          ghive_pred *nodes, uint32_t capacity, uint32_t *nnodes, uint32_t *root,
                          char *err, size_t err_cap) {
  auto fail = [&](const std::string &m) {
    if (err && err_cap) snprintf(err, err_cap, "%s", m.c_str());
    return GHIVE_ERR_INVALID;
  };
  if (!text || !nodes || !nnodes || !root) return fail("NULL argument");
  Parser P;
  P.names = col_names;
  P.ncols = ncols;
  std::string cur;
  for (const char *c = text; *c; ++c) {
    if (*c == '\'') {  // a '...' literal is one token, spaces and parentheses included
      const char *e = strchr(c + 1, '\'');
      if (!e) return fail("unterminated string literal");
      cur.append(c, (size_t)(e - c + 1));
      c = e;
      continue;
    }
    if (*c == '(' && !cur.empty()) {  // f(x): a function call is part of its token
      int depth = 0;
      for (; *c; ++c) {
        cur += *c;
        if (*c == '(') ++depth;
        if (*c == ')' && --depth == 0) break;
      }
      if (!*c) return fail("missing ')'");
      continue;
    }
    if (*c == ' ' || *c == '(' || *c == ')') {
      if (!cur.empty()) P.tok.push_back(cur);
      cur.clear();
      if (*c != ' ') P.tok.push_back(std::string(1, *c));
    } else {
      cur += *c;
    }
  }
  if (!cur.empty()) P.tok.push_back(cur);
  int r = P.group();
  if (r < 0) return fail(P.err);
  if (P.pos != P.tok.size()) return fail("unexpected '" + P.tok[P.pos] + "'");
  if (P.nodes.size() > capacity) {
    *nnodes = (uint32_t)P.nodes.size();
    return GHIVE_ERR_CAPACITY;
  }
  memcpy(nodes, P.nodes.data(), sizeof(ghive_pred) * P.nodes.size());
  *nnodes = (uint32_t)P.nodes.size();
  *root = (uint32_t)r;
  if (err && err_cap) err[0] = 0;
  return GHIVE_OK;
}

size_t ghive_predicate_to_string(const ghive_pred *nodes, uint32_t nnodes, uint32_t root, char *buf,
                                 size_t cap) {
  std::string s;
  if (nodes) dump(nodes, nnodes, root, 0, s, 0);
  if (buf && cap) snprintf(buf, cap, "%s", s.c_str());
  return s.size() + 1;
}

}  // extern "C"
