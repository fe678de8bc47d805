// Physics plugin interface of the script host.
//
// Same five methods and injected maps as General-Simulator's Dynamics
// (src/dynamics.h:9-25): Initialize, Calculate, RunFunc, Fix(progress), Get(column),
// plus nx,ny,nz,dx,dy,dz and pointers to the named-field and variable maps.
// Pull() is an addition: it copies a device-resident field back into the field
// map before GS writes it (the reference calls Data::DeviceToHost there,
// src/gs.cu:46-48).
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <string>

#include "field.h"

namespace gs {

// return codes mirror src/pub.h:17-22
enum Code {
  Code_NORMAL = 0,
  Code_ERR = -1,
  Code_BREAK = -2,
  Code_QUIT = -3,
  Code_COMMAND_UNKNOW = -4,
  Code_INPUT_UNCOMPLETE = -5
};

class Dynamics {
 public:
  virtual ~Dynamics() = default;
  int nx = 0, ny = 0, nz = 0;
  double dx = 1, dy = 1, dz = 1;
  FieldMap* Datas = nullptr;
  VarMap* Vars = nullptr;
  virtual int Initialize() = 0;
  virtual int Calculate() = 0;
  virtual int RunFunc(const std::string& func) = 0;
  virtual int Fix(double progress) = 0;
  virtual std::string Get(const std::string& col) = 0;
  virtual int Pull(const std::string& /*name*/) { return Code_NORMAL; }
  // Push() is the reverse of Pull(): it uploads a field of the map into the plugin's device copy
  // (the reference's systems share one set of device arrays; here each GPU plugin holds its own).
  virtual int Push(const std::string& /*name*/) { return Code_NORMAL; }
  // Several steps at once when nothing between them needs the host (Fix is a no-op
  // for the physics); the default calls Calculate n times.
  virtual int CalculateMany(int n) {
    for (int i = 0; i < n; ++i)
      if (int e = Calculate()) return e;
    return Code_NORMAL;
  }
  // true when Fix(progress) cannot change the step (lets GS batch steps)
  virtual bool FixIsInert() const { return false; }
};

// name -> factory registry (replaces the generated GS_SYS_DEFINE block, src/gs.cu:71-79)
using Factory = std::function<std::unique_ptr<Dynamics>()>;
std::map<std::string, Factory>& registry();
struct Register {
  Register(const std::string& name, Factory f) { registry()[name] = std::move(f); }
};

// helpers for Vars: read whitespace-separated values of a key
bool var_values(const VarMap& v, const std::string& key, std::vector<std::string>& out);
template <typename T>
int var_get(const VarMap& v, const std::string& key, T* dst, int count);
void log_line(const std::string& s);

}  // namespace gs
