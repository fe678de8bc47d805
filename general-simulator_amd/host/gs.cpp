// Script host: field I/O, expression evaluation, the command interpreter and the
// GS engine (system registry + step loop).
//
// Reference behaviour followed (General-Simulator):
//   script normalisation        src/input.cu:133-213
//   command dispatch            src/input.cu:14-108
//   readhere                    src/input.cu:216-242, src/gs.cu:125-133
//   {expr}, `shell`, :x=expr     src/variable.cu:66-214
//   set / read / write / sys    src/gs.cu:57-165
//   run, info, dump             src/gs.cu:20-55, 168-212
#include "gs.h"

#include <sys/stat.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace gs {

// =========================================================================== fields
static long long flat(const std::vector<int>& dims, std::initializer_list<int> idx) {
  long long o = 0;
  int k = 0;
  for (int i : idx) o = o * dims[k++] + i;
  return o;
}
double& Field::at(std::initializer_list<int> idx) { return data[(size_t)flat(dims, idx)]; }

int Field::write_text(const std::string& file) const {
  std::ofstream os(file);
  if (!os) return Code_ERR;
  os << dims.size() << " ";
  for (int d : dims) os << d << " ";
  os << "\n";
  for (double v : data) os << (float)v << "\n";
  return os ? Code_NORMAL : Code_ERR;
}

int Field::write_binary(const std::string& file) const {
  std::ofstream os(file, std::ios::binary);
  if (!os) return Code_ERR;
  int r = (int)dims.size();
  os.write((const char*)&r, sizeof(int));
  os.write((const char*)dims.data(), r * sizeof(int));
  std::vector<float> f(data.begin(), data.end());
  os.write((const char*)f.data(), f.size() * sizeof(float));
  return os ? Code_NORMAL : Code_ERR;
}

int Field::read_text(const std::string& file) {
  std::ifstream is(file);
  if (!is) {
    log_line("Error: \"" + file + "\" not found!");
    return Code_ERR;
  }
  int r = 0;
  is >> r;
  if (r <= 0 || r > 8) return Code_ERR;
  std::vector<int> d(r);
  for (int i = 0; i < r; ++i) is >> d[i];
  init(d);
  for (auto& v : data) is >> v;
  return is ? Code_NORMAL : Code_ERR;
}

int Field::read_binary(const std::string& file) {
  std::ifstream is(file, std::ios::binary);
  if (!is) {
    log_line("Error: \"" + file + "\" not found!");
    return Code_ERR;
  }
  int r = 0;
  is.read((char*)&r, sizeof(int));
  if (r <= 0 || r > 8) return Code_ERR;
  std::vector<int> d(r);
  is.read((char*)d.data(), r * sizeof(int));
  init(d);
  std::vector<float> f((size_t)count());
  is.read((char*)f.data(), f.size() * sizeof(float));
  for (size_t i = 0; i < f.size(); ++i) data[i] = f[i];
  return is ? Code_NORMAL : Code_ERR;
}

// =========================================================================== registry / vars / log
std::map<std::string, Factory>& registry() {
  static std::map<std::string, Factory> r;
  return r;
}

bool var_values(const VarMap& v, const std::string& key, std::vector<std::string>& out) {
  out.clear();
  auto it = v.find(key);
  if (it == v.end()) return false;
  std::istringstream ss(it->second);
  std::string w;
  while (ss >> w) out.push_back(w);
  return !out.empty();
}

template <typename T>
int var_get(const VarMap& v, const std::string& key, T* dst, int count) {
  std::vector<std::string> w;
  if (!var_values(v, key, w)) return 0;
  int k = 0;
  for (; k < count && k < (int)w.size(); ++k) {
    std::istringstream ss(w[k]);
    ss >> dst[k];
  }
  return k;
}
template int var_get<double>(const VarMap&, const std::string&, double*, int);
template int var_get<int>(const VarMap&, const std::string&, int*, int);
template int var_get<long long>(const VarMap&, const std::string&, long long*, int);
template int var_get<unsigned long long>(const VarMap&, const std::string&, unsigned long long*, int);

static std::ofstream* g_logfile = nullptr;
static bool g_echo = true;
void log_line(const std::string& s) {
  std::cout << s << "\n";
  std::cout.flush();
  if (g_logfile) (*g_logfile) << s << "\n";
}
void log_open(const std::string& file) {
  static std::ofstream f;
  f.open(file);
  if (f) g_logfile = &f;
}
void log_echo(bool on) { g_echo = on; }
bool log_echo_on() { return g_echo; }

std::string num_to_string(double v) {
  std::ostringstream os;
  os << v;
  return os.str();
}

// =========================================================================== expressions
// Recursive descent over double values; operators by increasing binding:
//   =  +=  -=  *=  /=   (right assoc; assigns script variables, $name assigns set keys)
//   == != <= >= < >
//   + -
//   * /
//   ^ (integer power as the reference's Real^int, src/variable.cu:194)
//   unary -, parentheses, numbers, names
// Each value carries its text as the reference's ExprTree does (src/variable.cu:139-209): a
// variable or a number literal keeps its text verbatim (so `:t=`date +%s`` stores all ten
// digits and {t} prints them), an operation's text is its double printed with the stream's
// default six significant digits, and `=` stores the right-hand side's text.
struct EVal {
  double v = NAN;
  std::string t;
};

static EVal ev_num(double v) { return EVal{v, num_to_string(v)}; }

struct Expr {
  const std::string& s;
  size_t p = 0;
  VarMap& vars;
  VarMap& gvars;
  std::string err;
  Expr(const std::string& str, VarMap& v, VarMap& g) : s(str), vars(v), gvars(g) {}

  void ws() { while (p < s.size() && std::isspace((unsigned char)s[p])) ++p; }
  bool eat(const char* op) {
    ws();
    size_t n = std::char_traits<char>::length(op);
    if (s.compare(p, n, op) == 0) { p += n; return true; }
    return false;
  }
  std::string name() {
    ws();
    size_t b = p;
    if (p < s.size() && s[p] == '$') ++p;
    while (p < s.size() && (std::isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '.')) ++p;
    return s.substr(b, p - b);
  }
  EVal lookup(const std::string& nm) {
    auto it = vars.find(nm);
    std::string key = nm;
    if (it == vars.end()) {
      if (!key.empty() && key[0] == '$') key.erase(0, 1);
      it = gvars.find(key);
      if (it == gvars.end()) {
        err = "\"" + nm + "\" unknown";
        return EVal{};
      }
    }
    std::istringstream ss(it->second);
    EVal r;
    ss >> r.v;
    if (ss.fail()) r.v = NAN;
    std::istringstream ws_(it->second);
    ws_ >> r.t;  // the first word, as the reference's `>>=` reads it
    return r;
  }
  EVal primary() {
    ws();
    if (p >= s.size()) { err = "incomplete expression"; return EVal{}; }
    if (s[p] == '(') {
      ++p;
      EVal v = assign();
      if (!eat(")")) err = "missing )";
      return v;
    }
    if (std::isdigit((unsigned char)s[p]) || s[p] == '.') {
      char* end = nullptr;
      const size_t b = p;
      double v = std::strtod(s.c_str() + p, &end);
      p = end - s.c_str();
      return EVal{v, s.substr(b, p - b)};
    }
    std::string nm = name();
    if (nm.empty()) { err = std::string("unexpected '") + s[p] + "'"; ++p; return EVal{}; }
    return lookup(nm);
  }
  EVal unary() {
    if (eat("-")) return ev_num(-unary().v);
    if (eat("+")) return unary();
    return primary();
  }
  EVal power() {
    EVal a = unary();
    if (eat("^")) {
      double b = power().v;
      int e = (int)b;
      double r = 1.0;
      for (int i = 0; i < std::abs(e); ++i) r *= a.v;
      return ev_num(e < 0 ? 1.0 / r : r);
    }
    return a;
  }
  EVal mul() {
    EVal a = power();
    for (;;) {
      if (eat("*")) a = ev_num(a.v * power().v);
      else if (eat("/")) { double b = power().v; a = (b == 0) ? EVal{NAN, "nan"} : ev_num(a.v / b); }
      else return a;
    }
  }
  EVal add() {
    EVal a = mul();
    for (;;) {
      ws();
      if (p + 1 < s.size() && (s[p] == '+' || s[p] == '-') && s[p + 1] == '=') return a;
      if (eat("+")) a = ev_num(a.v + mul().v);
      else if (eat("-")) a = ev_num(a.v - mul().v);
      else return a;
    }
  }
  EVal cmp() {
    EVal a = add();
    if (eat("==")) return ev_num(a.v == add().v);
    if (eat("!=")) return ev_num(a.v != add().v);
    if (eat("<=")) return ev_num(a.v <= add().v);
    if (eat(">=")) return ev_num(a.v >= add().v);
    if (eat("<")) return ev_num(a.v < add().v);
    if (eat(">")) return ev_num(a.v > add().v);
    return a;
  }
  EVal assign() {
    size_t save = p;
    std::string nm = name();
    ws();
    if (!nm.empty() && !std::isdigit((unsigned char)nm[0]) && p < s.size()) {
      const char* ops[] = {"+=", "-=", "*=", "/="};
      for (const char* op : ops)
        if (eat(op)) {
          double cur = lookup(nm).v, r = assign().v;
          double v = op[0] == '+' ? cur + r : op[0] == '-' ? cur - r : op[0] == '*' ? cur * r : cur / r;
          EVal out = ev_num(v);
          store(nm, out.t);
          return out;
        }
      if (s[p] == '=' && (p + 1 >= s.size() || s[p + 1] != '=')) {
        ++p;
        EVal v = assign();
        store(nm, v.t);
        return v;
      }
    }
    p = save;
    return cmp();
  }
  void store(const std::string& nm, const std::string& text) {
    if (!nm.empty() && nm[0] == '$') gvars[nm.substr(1)] = text;
    else vars[nm] = text;
  }
};

int eval_expr(const std::string& text, VarMap& vars, VarMap& gvars, double* out, std::string* out_text) {
  Expr e(text, vars, gvars);
  EVal v = e.assign();
  e.ws();
  if (e.err.empty() && e.p != text.size()) e.err = "trailing characters in \"" + text + "\"";
  if (!e.err.empty()) {
    log_line("Error: " + e.err);
    return Code_ERR;
  }
  if (out) *out = v.v;
  if (out_text) *out_text = v.t;
  return Code_NORMAL;
}

// =========================================================================== GS engine
GS::GS() = default;
GS::~GS() = default;

int GS::SetSys(const std::string& name) {
  // hand the shared fields over between systems (one set of arrays in the reference)
  static const char* const shared[] = {"eta", "concentration"};
  if (!current_.empty() && current_ != name)
    for (const char* f : shared)
      if (int e = sys_[current_]->Pull(f)) return e;
  if (sys_.count(name)) {
    const bool handover = current_ != name && inited_[name];
    current_ = name;
    if (handover)
      for (const char* f : shared)
        if (int e = sys_[name]->Push(f)) return e;
    return Code_NORMAL;
  }
  auto it = registry().find(name);
  if (it == registry().end()) {
    log_line("Error: unknown system \"" + name + "\"");
    return Code_ERR;
  }
  sys_[name] = it->second();
  inited_[name] = false;
  current_ = name;
  return Code_NORMAL;
}

int GS::Set(const std::string& rest) {
  std::istringstream ss(rest);
  std::string key;
  ss >> key;
  std::string val;
  std::getline(ss, val);
  size_t b = val.find_first_not_of(' ');
  val = (b == std::string::npos) ? "" : val.substr(b);
  if (key == "binary") binary_ = (val.rfind("on", 0) == 0);
  else if (key == "log") log_echo(val.rfind("off", 0) != 0);
  else Vars[key] = val;
  return Code_NORMAL;
}

// `link target alias` (src/gs.cu:110-115): the string variable `alias` becomes a name for
// `target`. As in the reference, an unknown target leaves the variables unchanged and the
// command still succeeds; here a warning is logged.
int GS::Link(const std::string& rest) {
  std::istringstream ss(rest);
  std::string target, alias;
  ss >> target >> alias;
  if (target.empty() || alias.empty()) {
    log_line("Error: link needs a target and an alias");
    return Code_ERR;
  }
  if (Vars.link(target, alias)) log_line("Warning: link target \"" + target + "\" has no value; not linked");
  return Code_NORMAL;
}

int GS::Read(const std::string& rest) {
  std::istringstream ss(rest);
  std::string var, file;
  ss >> var >> file;
  return binary_ ? Datas[var].read_binary(file) : Datas[var].read_text(file);
}

int GS::ReadHere(const std::string& name, const std::vector<double>& vals) {
  if (vals.empty()) return Code_ERR;
  int r = (int)vals[0];
  if (r <= 0 || (int)vals.size() < 1 + r) return Code_ERR;
  std::vector<int> d(r);
  for (int i = 0; i < r; ++i) d[i] = (int)vals[1 + i];
  Field& f = Datas[name];
  f.init(d);
  if ((long long)vals.size() < 1 + r + f.count()) {
    log_line("Error: readhere " + name + " expects " + std::to_string(f.count()) + " values");
    return Code_ERR;
  }
  for (long long i = 0; i < f.count(); ++i) f.data[(size_t)i] = vals[(size_t)(1 + r + i)];
  return Code_NORMAL;
}

int GS::WriteField(const std::string& name, const std::string& file) {
  if (!current_.empty()) sys_[current_]->Pull(name);
  auto it = Datas.find(name);
  if (it == Datas.end() || it->second.empty()) {
    log_line("Error: data " + name + " does not exist!");
    return Code_ERR;
  }
  return binary_ ? it->second.write_binary(file) : it->second.write_text(file);
}

int GS::Write(const std::string& rest) {
  std::istringstream ss(rest);
  std::string name, file;
  while (ss >> name) {
    if (!(ss >> file)) file = name + ".data";
    if (int e = WriteField(name, file)) return e;
  }
  return Code_NORMAL;
}

int GS::WriteHere(const std::string& rest) {
  std::istringstream ss(rest);
  std::string name;
  while (ss >> name) {
    if (!current_.empty()) sys_[current_]->Pull(name);
    auto it = Datas.find(name);
    if (it != Datas.end() && !it->second.empty()) {
      std::ostringstream os;
      os << it->second.dims.size() << " ";
      for (int d : it->second.dims) os << d << " ";
      os << "\n";
      for (double v : it->second.data) os << v << "\n";
      log_line(os.str());
    } else if (Vars.count(name)) {
      log_line(Vars[name]);
    } else {
      log_line("Error: unknown data \"" + name + "\"");
    }
  }
  return Code_NORMAL;
}

int GS::SetInfo(const std::string& rest) {
  std::istringstream ss(rest);
  ss >> info_count_;
  std::string w;
  info_cols_.clear();
  while (ss >> w) info_cols_.push_back(w);
  return Code_NORMAL;
}

int GS::SetDump(const std::string& rest) {
  std::istringstream ss(rest);
  ss >> dump_dir_ >> dump_count_;
  std::string w;
  dump_fields_.clear();
  while (ss >> w) dump_fields_.push_back(w);
  mkdir(dump_dir_.c_str(), 0755);
  return Code_NORMAL;
}

int GS::InfoOut() {
  std::string line = "Step  " + std::to_string(CurrentStep) + " \t";
  for (auto& c : info_cols_) line += sys_[current_]->Get(c) + " \t";
  log_line(line);
  return Code_NORMAL;
}

int GS::DumpOut() {
  for (auto& f : dump_fields_) {
    std::string file = dump_dir_ + "/" + f + "." + std::to_string(CurrentStep) + ".data";
    if (WriteField(f, file)) log_line("Error: data " + f + " does not exist!");
  }
  return Code_NORMAL;
}

int GS::Run(const std::string& rest) {
  if (current_.empty()) {
    log_line("Error: system not set");
    return Code_ERR;
  }
  std::istringstream ss(rest);
  int total = 1;
  ss >> total;
  if (total <= 0) return Code_ERR;
  Dynamics* d = sys_[current_].get();
  d->Datas = &Datas;
  d->Vars = &Vars;
  if (!inited_[current_]) {
    // the reference ignores Initialize's status (src/gs.cu:180); a failure aborts here
    if (d->Initialize() != Code_NORMAL) {
      log_line("Error: system " + current_ + " failed to initialise");
      return Code_ERR;
    }
    inited_[current_] = true;
    CurrentStep = 0;
    long long ahead = 0;
    if (var_get(Vars, "ahead_steps", &ahead, 1)) CurrentStep = ahead;
  }
  const int info_iv = info_count_ == 0 ? total + 1 : std::max(1, total / info_count_);
  const int dump_iv = dump_count_ == 0 ? total + 1 : std::max(1, total / dump_count_);
  {
    std::string h = "Info \t\t";
    for (auto& c : info_cols_) h += c + "\t";
    log_line(h);
  }
  for (int i = 1; i <= total;) {
    // run up to the next step that needs the host (info, dump or the last step)
    int next = total;
    for (int j = i; j <= total; ++j)
      if (j % info_iv == 0 || j % dump_iv == 0) { next = j; break; }
    const int batch = d->FixIsInert() ? next - i + 1 : 1;
    for (int k = 0; k < batch; ++k) {
      if (int e = d->Fix((double)(i + k) / total)) return e;
    }
    int e = batch == 1 ? d->Calculate() : d->CalculateMany(batch);
    if (e) {
      log_line("Error: step failed at step " + std::to_string(CurrentStep + 1));
      return Code_ERR;
    }
    CurrentStep += batch;
    i += batch;
    const int done = i - 1;
    if (done % info_iv == 0 || done == total) InfoOut();
    if (done % dump_iv == 0 || done == total) DumpOut();
  }
  return Code_NORMAL;
}

int GS::RunFunc(const std::string& func) {
  if (current_.empty()) {
    log_line("Error: no system set, runfunc commands not available");
    return Code_ERR;
  }
  Dynamics* d = sys_[current_].get();
  d->Datas = &Datas;
  d->Vars = &Vars;
  if (!inited_[current_]) {
    log_line("Warning: Initialization function called before \"run\"");
    if (d->Initialize() != Code_NORMAL) return Code_ERR;
    inited_[current_] = true;
  }
  return d->RunFunc(func);
}

// =========================================================================== interpreter
std::string Interpreter::standardize(std::string script) {
  script += "\n";
  for (auto& c : script)
    if (c == '\t' || c == '\r') c = ' ';
  for (auto& c : script)
    if (c == ';') c = '\n';
  // comments run to the end of the line
  size_t p1;
  while ((p1 = script.find('#')) != std::string::npos) {
    size_t p2 = script.find('\n', p1);
    script.erase(p1, p2 - p1);
  }
  // no blanks inside (...) expressions
  std::string out;
  int depth = 0;
  for (char c : script) {
    if (c == '(') ++depth;
    if (c == ')') --depth;
    if (depth > 0 && (c == ' ' || c == '\n')) continue;
    out += c;
  }
  // no blanks inside {...} either (one expression per brace pair)
  std::string out2;
  depth = 0;
  for (char c : out) {
    if (c == '{') ++depth;
    if (c == '}') --depth;
    if (depth > 0 && (c == ' ' || c == '\n')) continue;
    out2 += c;
  }
  for (auto& c : out2)
    if (c == ',') c = ' ';
  return out2;
}

int Interpreter::substitute(std::string& s) {
  // `cmd` -> its output
  size_t l;
  while ((l = s.find('`')) != std::string::npos) {
    size_t r = s.find('`', l + 1);
    if (r == std::string::npos) {
      log_line("Error: uncomplete expression");
      return Code_ERR;
    }
    std::string cmd = s.substr(l + 1, r - l - 1), res;
    if (shell_capture(cmd, res)) return Code_ERR;
    s.replace(l, r - l + 1, res);
  }
  // {expr} -> value (innermost first)
  while ((l = s.rfind('{')) != std::string::npos) {
    size_t r = s.find('}', l);
    if (r == std::string::npos) {
      log_line("Error: uncomplete expression");
      return Code_ERR;
    }
    std::string t;
    if (eval_expr(s.substr(l + 1, r - l - 1), vars_, gs_.Vars, nullptr, &t)) return Code_ERR;
    s.replace(l, r - l + 1, t);
  }
  return Code_NORMAL;
}

int Interpreter::shell_capture(const std::string& cmd, std::string& out) {
  out.clear();
  FILE* f = popen((cmd + " 2>&1").c_str(), "r");
  if (!f) return Code_ERR;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, n);
  pclose(f);
  while (!out.empty() && (out.back() == '\n')) out.pop_back();
  return Code_NORMAL;
}

static bool next_line(std::string& script, std::string& line) {
  if (script.empty()) return false;
  size_t n = script.find('\n');
  if (n == std::string::npos) n = script.size();
  line = script.substr(0, n);
  script.erase(0, n + 1);
  return true;
}

int Interpreter::readhere(const std::string& rest, std::string& script) {
  std::istringstream hs(rest);
  std::string name;
  hs >> name;
  std::vector<double> vals;
  long long need = -1;
  int rank = -1;
  std::string line;
  auto take = [&](std::istringstream& ss) {
    std::string w;
    while (ss >> w) {
      vals.push_back(std::strtod(w.c_str(), nullptr));
      if (rank < 0) rank = (int)vals[0];
      if (rank > 0 && (int)vals.size() == 1 + rank) {
        need = 1;
        for (int i = 0; i < rank; ++i) need *= (long long)vals[1 + i];
        need += 1 + rank;
      }
    }
  };
  {
    std::string tail;
    std::getline(hs, tail);
    std::istringstream ts(tail);
    take(ts);
  }
  while ((need < 0 || (long long)vals.size() < need) && next_line(script, line)) {
    if (substitute(line)) return Code_ERR;
    std::istringstream ls(line);
    take(ls);
  }
  return gs_.ReadHere(name, vals);
}

// The lines of a loop / while / if block up to its matching `end` (nested blocks count),
// with the `else` part of an if split off (src/input.cu:254-336). Consumes them from script.
static int split_block(std::string& script, std::string& body, std::string* other) {
  int depth = 1;
  bool in_else = false;
  std::string line;
  body.clear();
  if (other) other->clear();
  while (next_line(script, line)) {
    std::istringstream ls(line);
    std::string w;
    ls >> w;
    if (w == "loop" || w == "while" || w == "if") {
      ++depth;
    } else if (w == "end") {
      if (--depth == 0) return Code_NORMAL;
    } else if (w == "else" && depth == 1 && other) {
      in_else = true;
      continue;
    }
    (in_else ? *other : body) += line + "\n";
  }
  return Code_INPUT_UNCOMPLETE;
}

// loop N / while expr / if expr [else] ... end (src/input.cu:269-336)
int Interpreter::control(const std::string& cmd, const std::string& rest, std::string& script) {
  std::string body, other;
  if (split_block(script, body, cmd == "if" ? &other : nullptr) != Code_NORMAL) {
    log_line("Error: " + cmd + " without end");
    return Code_INPUT_UNCOMPLETE;
  }
  auto value = [&](double* v) -> int {
    std::string e = rest;
    if (substitute(e)) return Code_ERR;
    return eval_expr(e, vars_, gs_.Vars, v);
  };
  double v = 0.0;
  if (cmd == "loop") {
    if (value(&v)) return Code_ERR;
    for (long i = 0; i < (long)v; ++i) {
      std::string b = body;
      const int e = exec(b);
      if (e != Code_NORMAL) return e;
    }
    return Code_NORMAL;
  }
  if (cmd == "while") {
    for (;;) {
      if (value(&v)) return Code_ERR;
      if (!(v > 0)) return Code_NORMAL;
      std::string b = body;
      const int e = exec(b);
      if (e != Code_NORMAL) return e;
    }
  }
  if (value(&v)) return Code_ERR;
  return exec(v > 0 ? body : other);
}

int Interpreter::run(std::string script) {
  script = standardize(script);
  return exec(script);
}

int Interpreter::exec(std::string& script) {
  std::string line;
  while (next_line(script, line)) {
    size_t b = line.find_first_not_of(' ');
    if (b == std::string::npos) continue;
    line = line.substr(b, line.find_last_not_of(' ') - b + 1);
    std::string raw = line;
    std::istringstream ls(line);
    std::string cmd;
    ls >> cmd;
    std::string rest;
    std::getline(ls, rest);
    if (!rest.empty() && rest[0] == ' ') rest.erase(0, 1);
    if (!cmd.empty() && cmd[0] == ':') {
      rest = line.substr(1);  // the whole line after the colon, blanks kept
      cmd = "expr";
    }
    if (cmd == "loop" || cmd == "while" || cmd == "if") {
      if (log_echo_on()) log_line(">>>" + raw);
      const int e = control(cmd, rest, script);
      if (e == Code_QUIT) return Code_QUIT;
      if (e != Code_NORMAL) {
        log_line("Error: error in command \"" + raw + "\"");
        return Code_ERR;
      }
      continue;
    }
    // every line but print has `cmd` and {expr} substituted before dispatch, expressions
    // (`:x=...` lines) included (src/input.cu:119-120, HEAD side)
    if (cmd != "print") {
      if (substitute(rest)) return Code_ERR;
    }
    if (log_echo_on()) log_line(">>>" + raw);
    int err = Code_NORMAL;
    if (cmd == "expr") {
      double v;
      err = eval_expr(rest, vars_, gs_.Vars, &v);
    } else if (cmd == "shell") {
      std::string out;
      err = shell_capture(rest, out);
      if (!out.empty()) log_line(out);
    } else if (cmd == "log" || cmd == "logon") {
      log_echo(true);
    } else if (cmd == "logoff") {
      log_echo(false);
    } else if (cmd == "break") {
      break;
    } else if (cmd == "quit") {
      return Code_QUIT;
    } else if (cmd == "info") {
      err = gs_.SetInfo(rest);
    } else if (cmd == "dump") {
      err = gs_.SetDump(rest);
    } else if (cmd == "set") {
      err = gs_.Set(rest);
    } else if (cmd == "read") {
      err = gs_.Read(rest);
    } else if (cmd == "link") {
      err = gs_.Link(rest);
    } else if (cmd == "readhere") {
      err = readhere(rest, script);
    } else if (cmd == "write") {
      err = gs_.Write(rest);
    } else if (cmd == "writehere") {
      err = gs_.WriteHere(rest);
    } else if (cmd == "print") {
      std::string r = rest;
      if (substitute(r)) return Code_ERR;
      log_line(r);
    } else if (cmd == "sys") {
      std::istringstream ss(rest);
      std::string nm;
      ss >> nm;
      err = gs_.SetSys(nm);
    } else if (cmd == "run") {
      err = gs_.Run(rest);
    } else if (cmd == "device") {
      std::istringstream ss(rest);
      ss >> device_;
      gs_.Vars["device"] = std::to_string(device_);
    } else if (cmd == "runfunc") {
      err = gs_.RunFunc(rest);
    } else {
      err = gs_.RunFunc(cmd + " " + rest);
      if (err == Code_COMMAND_UNKNOW) {
        log_line("Error: " + raw + " unknown");
        return Code_ERR;
      }
    }
    if (err == Code_QUIT) return Code_QUIT;
    if (err != Code_NORMAL) {
      log_line("Error: error in command \"" + raw + "\"");
      return Code_ERR;
    }
  }
  return Code_NORMAL;
}

}  // namespace gs
