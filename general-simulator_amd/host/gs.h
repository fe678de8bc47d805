// GS engine and script interpreter of the host (see gs.cpp for the reference
// behaviour each part follows).
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dynamics.h"

namespace gs {

// Evaluates `text`; *out gets the value and *out_text (may be null) its text (a variable's stored
// text or a literal verbatim, an operation's value with six significant digits).
int eval_expr(const std::string& text, VarMap& vars, VarMap& gvars, double* out, std::string* out_text = nullptr);
std::string num_to_string(double v);
void log_open(const std::string& file);
void log_echo(bool on);
bool log_echo_on();

class GS {
 public:
  GS();
  ~GS();
  FieldMap Datas;
  VarMap Vars;
  long long CurrentStep = 0;

  int SetSys(const std::string& name);
  int Set(const std::string& rest);
  int Read(const std::string& rest);
  int Link(const std::string& rest);
  int ReadHere(const std::string& name, const std::vector<double>& vals);
  int Write(const std::string& rest);
  int WriteHere(const std::string& rest);
  int SetInfo(const std::string& rest);
  int SetDump(const std::string& rest);
  int Run(const std::string& rest);
  int RunFunc(const std::string& func);
  int InfoOut();
  int DumpOut();

 private:
  int WriteField(const std::string& name, const std::string& file);
  std::map<std::string, std::unique_ptr<Dynamics>> sys_;
  std::map<std::string, bool> inited_;
  std::string current_;
  bool binary_ = false;
  int info_count_ = 0, dump_count_ = 0;
  std::vector<std::string> info_cols_, dump_fields_;
  std::string dump_dir_ = ".";
};

class Interpreter {
 public:
  explicit Interpreter(GS& g) : gs_(g) {}
  int run(std::string script);
  static std::string standardize(std::string script);
  int substitute(std::string& s);
  VarMap& vars() { return vars_; }

 private:
  int exec(std::string& script);  // runs standardized lines, consuming script
  int control(const std::string& cmd, const std::string& rest, std::string& script);
  int readhere(const std::string& rest, std::string& script);
  int shell_capture(const std::string& cmd, std::string& out);
  GS& gs_;
  VarMap vars_;
  int device_ = 0;
};

}  // namespace gs
