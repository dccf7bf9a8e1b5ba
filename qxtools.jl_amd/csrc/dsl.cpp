// Parser and validator of the .qx compute-graph DSL.
//
// Format: docs/src/users_guide.md:95-165 (load / view / ncon / output / save),
// produced by build_compute_graph (src/compute_graph/compute_graph.jl:15-98) via
// QXContexts' DSL writer.  Semantics enforced here:
//   * view keeps the rank and sets the viewed leg to size 1 (users_guide.md:107-122);
//   * ncon labels fall into batch H = a&b&c, contracted K = (a&b)-c, free-A M and
//     free-B N; a label in one input only but not in C (a sum-out) and a label
//     repeated within one tensor are rejected; label "0" is an empty list (scalar).
#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <set>
#include <sstream>

#include "qx_internal.hpp"

namespace qx {
namespace {

std::vector<std::string> split_ws(const std::string &s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

[[noreturn]] void perr(int line, const std::string &msg) {
  throw Error(QX_ERR_PARSE, "line " + std::to_string(line) + ": " + msg);
}

int64_t to_int(const std::string &s, int line, const char *what) {
  if (s.empty()) perr(line, std::string("empty ") + what);
  char *end = nullptr;
  errno = 0;
  long long v = std::strtoll(s.c_str(), &end, 10);
  if (errno != 0 || end == s.c_str() || *end != '\0') perr(line, std::string("bad ") + what + " '" + s + "'");
  return (int64_t)v;
}

std::vector<int64_t> int_list(const std::string &s, int line, const char *what) {
  std::vector<int64_t> out;
  size_t i = 0;
  while (i <= s.size()) {
    size_t j = s.find(',', i);
    if (j == std::string::npos) j = s.size();
    out.push_back(to_int(s.substr(i, j - i), line, what));
    i = j + 1;
  }
  return out;
}

std::vector<int> labels(const std::string &s, int line) {
  if (s == "0") return {};
  std::vector<int> out;
  for (int64_t v : int_list(s, line, "label")) {
    if (v <= 0 || v > (1 << 30)) perr(line, "labels are positive integers");
    out.push_back((int)v);
  }
  return out;
}

}  // namespace

Program parse_program(const char *text, size_t len) {
  if (!text) throw Error(QX_ERR_ARGUMENT, "dsl text is null");
  Program p;
  std::string all(text, len);
  std::istringstream is(all);
  std::string raw;
  int line = 0;
  std::map<std::string, int> var_first;
  while (std::getline(is, raw)) {
    ++line;
    size_t a = raw.find_first_not_of(" \t\r");
    if (a == std::string::npos || raw[a] == '#') continue;
    auto f = split_ws(raw);
    Node n;
    n.line = line;
    const std::string &op = f[0];
    auto need = [&](size_t k) {
      if (f.size() != k) perr(line, op + " expects " + std::to_string(k - 1) + " fields");
    };
    if (op == "load") {
      if (f.size() != 4 && f.size() != 3) perr(line, "load expects name, key, dims");
      n.op = Op::Load;
      n.name = f[1];
      n.key = f[2];
      if (f.size() == 4) n.dims = int_list(f[3], line, "dim");
      for (int64_t d : n.dims)
        if (d <= 0) perr(line, "dims are positive");
    } else if (op == "output") {
      need(4);
      n.op = Op::Output;
      n.name = f[1];
      n.qubit = (int)to_int(f[2], line, "qubit");
      int64_t d = to_int(f[3], line, "dim");
      if (n.qubit <= 0) perr(line, "qubits are numbered from 1");
      if (d != 2) perr(line, "output tensors have dim 2");
      n.dims = {d};
    } else if (op == "view") {
      need(6);
      n.op = Op::View;
      n.name = f[1];
      n.src = f[2];
      n.var = f[3];
      n.pos = (int)to_int(f[4], line, "leg") - 1;
      n.vdim = to_int(f[5], line, "dim");
      if (n.pos < 0) perr(line, "view legs are numbered from 1");
      if (n.vdim <= 0) perr(line, "view dim is positive");
    } else if (op == "ncon") {
      need(7);
      n.op = Op::Ncon;
      n.name = f[1];
      n.clab = labels(f[2], line);
      n.a = f[3];
      n.alab = labels(f[4], line);
      n.b = f[5];
      n.blab = labels(f[6], line);
    } else if (op == "save") {
      need(3);
      n.op = Op::Save;
      n.name = f[1];
      n.src = f[2];
    } else {
      perr(line, "unknown instruction '" + op + "'");
    }
    if (n.op != Op::Save) {
      if (p.by_name.count(n.name)) perr(line, "tensor '" + n.name + "' defined twice");
      p.by_name[n.name] = (int)p.nodes.size();
    } else {
      if (p.save >= 0) perr(line, "more than one save");
    }
    p.nodes.push_back(std::move(n));
    if (p.nodes.back().op == Op::Save) p.save = (int)p.nodes.size() - 1;
    if (p.nodes.back().op == Op::View) var_first.emplace(p.nodes.back().var, (int)var_first.size());
  }
  if (p.save < 0) throw Error(QX_ERR_PARSE, "program has no save instruction");

  // slice variables: numeric order when all are v<k>, else order of first use
  std::vector<std::string> vars;
  for (auto &kv : var_first) vars.push_back(kv.first);
  bool numeric = std::all_of(vars.begin(), vars.end(), [](const std::string &v) {
    return v.size() > 1 && v[0] == 'v' && v.find_first_not_of("0123456789", 1) == std::string::npos;
  });
  if (numeric)
    std::sort(vars.begin(), vars.end(), [](const std::string &x, const std::string &y) {
      return std::atol(x.c_str() + 1) < std::atol(y.c_str() + 1);
    });
  else
    std::sort(vars.begin(), vars.end(),
              [&](const std::string &x, const std::string &y) { return var_first[x] < var_first[y]; });
  p.vars = vars;
  p.var_dims.assign(vars.size(), 0);

  auto find = [&](const std::string &nm, int line) {
    auto it = p.by_name.find(nm);
    if (it == p.by_name.end()) perr(line, "tensor '" + nm + "' used before definition");
    return it->second;
  };

  for (size_t i = 0; i < p.nodes.size(); ++i) {
    Node &n = p.nodes[i];
    switch (n.op) {
      case Op::Load: {
        n.shape = n.dims;
        int64_t e = 1;
        for (int64_t d : n.dims) {
          if (d > (int64_t(1) << 40) / std::max<int64_t>(e, 1)) perr(n.line, "tensor too large");
          e *= d;
        }
        auto it = p.key_elems.find(n.key);
        if (it == p.key_elems.end()) {
          p.key_elems[n.key] = e;
          p.keys.push_back(n.key);
        } else if (it->second != e) {
          perr(n.line, "data key '" + n.key + "' loaded with different sizes");
        }
        break;
      }
      case Op::Output:
        n.shape = n.dims;
        n.dep = true;
        p.num_chars = std::max(p.num_chars, n.qubit);
        p.outputs.push_back((int)i);
        break;
      case Op::View: {
        n.ia = find(n.src, n.line);
        if (n.ia >= (int)i) perr(n.line, "view of a later tensor");
        const Node &s = p.nodes[n.ia];
        if (n.pos >= (int)s.shape.size()) perr(n.line, "view leg beyond the tensor's rank");
        if (s.shape[n.pos] != n.vdim)
          perr(n.line, "view dim " + std::to_string(n.vdim) + " differs from leg dim " +
                           std::to_string(s.shape[n.pos]));
        n.shape = s.shape;
        n.shape[n.pos] = 1;
        n.dep = s.dep;
        n.sdep = true;
        n.var_id = (int)(std::find(p.vars.begin(), p.vars.end(), n.var) - p.vars.begin());
        int64_t &vd = p.var_dims[n.var_id];
        if (vd == 0) vd = n.vdim;
        else if (vd != n.vdim) perr(n.line, "slice variable '" + n.var + "' used with two dims");
        break;
      }
      case Op::Ncon: {
        n.ia = find(n.a, n.line);
        n.ib = find(n.b, n.line);
        if (n.ia == n.ib) perr(n.line, "ncon of a tensor with itself");
        if (n.ia >= (int)i || n.ib >= (int)i) perr(n.line, "ncon of a later tensor");
        const Node &A = p.nodes[n.ia];
        const Node &B = p.nodes[n.ib];
        if (A.shape.size() != n.alab.size() || B.shape.size() != n.blab.size())
          throw Error(QX_ERR_LABELS, "line " + std::to_string(n.line) + ": label count differs from tensor rank");
        auto uniq = [&](const std::vector<int> &l) {
          std::set<int> s(l.begin(), l.end());
          if (s.size() != l.size())
            throw Error(QX_ERR_LABELS, "line " + std::to_string(n.line) + ": repeated label within one tensor");
        };
        uniq(n.alab);
        uniq(n.blab);
        uniq(n.clab);
        std::map<int, int64_t> dim;
        for (size_t j = 0; j < n.alab.size(); ++j) dim[n.alab[j]] = A.shape[j];
        for (size_t j = 0; j < n.blab.size(); ++j) {
          auto it = dim.find(n.blab[j]);
          if (it != dim.end() && it->second != B.shape[j])
            throw Error(QX_ERR_LABELS, "line " + std::to_string(n.line) + ": label " +
                                           std::to_string(n.blab[j]) + " has different dims in A and B");
          dim[n.blab[j]] = B.shape[j];
        }
        std::set<int> as(n.alab.begin(), n.alab.end()), bs(n.blab.begin(), n.blab.end()),
            cs(n.clab.begin(), n.clab.end());
        for (int l : n.clab)
          if (!as.count(l) && !bs.count(l))
            throw Error(QX_ERR_LABELS, "line " + std::to_string(n.line) + ": output label " +
                                           std::to_string(l) + " is in neither input");
        for (int l : n.alab)
          if (!bs.count(l) && !cs.count(l))
            throw Error(QX_ERR_LABELS, "line " + std::to_string(n.line) + ": label " + std::to_string(l) +
                                           " is summed out of A alone");
        for (int l : n.blab)
          if (!as.count(l) && !cs.count(l))
            throw Error(QX_ERR_LABELS, "line " + std::to_string(n.line) + ": label " + std::to_string(l) +
                                           " is summed out of B alone");
        n.shape.clear();
        int64_t e = 1;
        for (int l : n.clab) {
          n.shape.push_back(dim[l]);
          if (dim[l] > (int64_t(1) << 48) / std::max<int64_t>(e, 1)) perr(n.line, "tensor too large");
          e *= dim[l];
        }
        n.dep = A.dep || B.dep;
        n.sdep = A.sdep || B.sdep;
        break;
      }
      case Op::Save: {
        n.ia = find(n.src, n.line);
        if (n.ia >= (int)i) perr(n.line, "save of a later tensor");
        n.shape = p.nodes[n.ia].shape;
        n.dep = p.nodes[n.ia].dep;
        n.sdep = p.nodes[n.ia].sdep;
        break;
      }
    }
  }
  return p;
}

}  // namespace qx
