// Phase-field plugins of the script host, running on the GPU through the C ABI.
//
// sys mart        -> Dynamics_mart        (General-Simulator src/dynamics_mart.cu)
// sys precipitate -> Dynamics_precipitate (General-Simulator src/dynamics_precipitate.cu)
// Keys keep the reference names and defaults where the physics is the same; keys
// the reference does not read are stored by its `set` unchanged (src/gs.cu:106),
// so scripts stay valid for both programs.
#include <cmath>
#include <sstream>

#include "dynamics.h"
#include "gsk.h"

namespace gs {
namespace {

std::string fmt(double v) {
  std::ostringstream os;
  os << v;
  return os.str();
}

// C[i][j][k][l] of a cubic crystal: the reference's cijkl (src/dynamics_mart.cu:80-89)
void cubic_modulus(double c11, double c12, double c44, double* C) {
  for (int i = 0; i < 81; ++i) C[i] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double v = 0.0;
          if (i == j && k == l) v = (i == k) ? c11 : c12;
          else if ((i == k && j == l) || (i == l && j == k)) v = c44;
          C[((i * 3 + j) * 3 + k) * 3 + l] = v;
        }
}

// Common part of both plugins: grid, handle, fields.
class GpuDynamics : public Dynamics {
 public:
  ~GpuDynamics() override {
    if (sim_) gsk_destroy(sim_);
  }
  int RunFunc(const std::string& /*func*/) override { return Code_COMMAND_UNKNOW; }
  // the reference martensite rule reads the temperature every step; everything else ignores Fix
  bool FixIsInert() const override { return !(ref_ && cfg_.physics == GSK_PHYS_MART); }

  // "set fix temperature Ts Te": linear temperature route (src/dynamics_mart.cu:421-441)
  int Fix(double progress) override {
    std::vector<std::string> w;
    if (!var_values(*Vars, "fix", w)) return Code_NORMAL;
    for (size_t i = 0; i < w.size();) {
      if (w[i] == "temperature" && i + 2 < w.size()) {
        double st = std::stod(w[i + 1]), et = std::stod(w[i + 2]);
        (*Vars)["temperature"] = fmt(st + progress * (et - st));
        i += 3;
      } else if (w[i] == "pressure") {
        i += 1;
      } else {
        log_line("Error: fix style " + w[i] + " does not find!");
        return Code_NORMAL;
      }
    }
    return Code_NORMAL;
  }

  int Calculate() override {
    if (ref_ && cfg_.physics == GSK_PHYS_MART) {  // (*Vars)["temperature"] >>= Temperature (:354)
      double T;
      if (var_get(*Vars, "temperature", &T, 1) && check(gsk_set_param(sim_, "temperature", T))) return Code_ERR;
    }
    return step(1);
  }
  int CalculateMany(int n) override {
    if (!FixIsInert()) {
      for (int i = 0; i < n; ++i)
        if (Calculate() != Code_NORMAL) return Code_ERR;
      return Code_NORMAL;
    }
    return step(n);
  }

 protected:
  gsk_sim* sim_ = nullptr;
  gsk_config cfg_{};
  int nvar_ = 0;
  bool ref_ = false;  // `set rule reference`: the plugin's own update rule

  // `set rule reference|spectral`. A script that names no rule gets the plugin's own update rule,
  // as the reference program runs it; `set rule spectral` selects the semi-implicit spectral rule.
  bool read_rule(const char* plugin) {
    std::vector<std::string> w;
    const bool spectral = var_values(*Vars, "rule", w) && w[0] == "spectral";
    if (!spectral && !(w.empty() || w[0] == "reference")) log_line("Warning: unknown rule \"" + w[0] + "\", using reference");
    log_line(std::string("Info: ") + plugin + " runs the " + (spectral ? "semi-implicit spectral" : "reference") +
             " update rule" + (spectral ? "" : " (set rule spectral selects the semi-implicit spectral rule)"));
    return !spectral;
  }

  // One warning per key of `keys` the script set that the chosen rule does not read.
  void warn_unread(std::initializer_list<const char*> keys, const char* plugin, const char* rule) {
    for (const char* k : keys)
      if (Vars->count(k))
        log_line(std::string("Warning: ") + plugin + " ignores \"set " + k + "\" under the " + rule + " rule");
  }

  // optional field from the map (e.g. defect, dislocationstress)
  int upload_optional(const std::string& name, long long want) {
    auto it = Datas->find(name);
    if (it == Datas->end() || it->second.empty()) return Code_NORMAL;
    if (it->second.count() != want) {
      log_line("Error: field " + name + " has " + std::to_string(it->second.count()) + " values, expected " +
               std::to_string(want));
      return Code_ERR;
    }
    return check(gsk_set_field(sim_, name.c_str(), it->second.data.data(), GSK_F64));
  }

  int step(int n) {
    if (gsk_step(sim_, n) != GSK_OK) {
      log_line(std::string("Error: ") + gsk_last_error());
      return Code_ERR;
    }
    return Code_NORMAL;
  }

  int read_grid() {
    std::vector<std::string> w;
    if (var_values(*Vars, "gridsize", w) && w.size() >= 3) {
      nx = std::stoi(w[0]); ny = std::stoi(w[1]); nz = std::stoi(w[2]);
      if (w.size() >= 6) { dx = std::stod(w[3]); dy = std::stod(w[4]); dz = std::stod(w[5]); }
    }
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0) || !(dy > 0) || !(dz > 0)) {
      log_line("Error: gridsize needs positive edges and spacings");
      return Code_ERR;
    }
    return Code_NORMAL;
  }

  bool cube() const { return nx == ny && ny == nz && dx == dy && dy == dz; }

  void read_common() {
    // a box grid (edges or spacings not all equal) runs the library's box path
    cfg_.n = nx;
    cfg_.ny = ny;
    cfg_.nz = nz;
    cfg_.dx = dx;
    cfg_.dy = dy;
    cfg_.dz = dz;
    double v;
    if (var_get(*Vars, "noiseamp", &v, 1)) cfg_.noise = v;
    unsigned long long seed;
    if (var_get(*Vars, "seed", &seed, 1)) cfg_.seed = seed;
    int dev;
    if (var_get(*Vars, "device", &dev, 1)) cfg_.device = dev;
    std::vector<std::string> w;
    if (var_values(*Vars, "wavevector", w) && w[0] == "reference") cfg_.kconv = GSK_K_REFERENCE;
  }

  int create() {
    if (gsk_create(&cfg_, &sim_) != GSK_OK) {
      log_line(std::string("Error: ") + gsk_last_error());
      return Code_ERR;
    }
    long long ahead = 0;
    if (var_get(*Vars, "ahead_steps", &ahead, 1)) gsk_set_step(sim_, ahead);
    return Code_NORMAL;
  }

  int check(int rc) {
    if (rc != GSK_OK) {
      log_line(std::string("Error: ") + gsk_last_error());
      return Code_ERR;
    }
    return Code_NORMAL;
  }

  // Upload a named field from the map, or initialise it.
  // `init<name> uniform lo hi [seed]` draws Philox uniforms; otherwise `fill` is used.
  int load_field(const std::string& name, const std::string& key, std::vector<int> dims, double fill) {
    auto it = Datas->find(name);
    const long long n3 = (long long)nx * ny * nz;
    if (it != Datas->end() && !it->second.empty()) {
      long long cnt = it->second.count();
      long long want = 1;
      for (int d : dims) want *= d;
      if (cnt != want) {
        log_line("Error: field " + name + " has " + std::to_string(cnt) + " values, expected " + std::to_string(want));
        return Code_ERR;
      }
      return check(gsk_set_field(sim_, name.c_str(), it->second.data.data(), GSK_F64));
    }
    std::vector<std::string> w;
    if (var_values(*Vars, key, w) && w[0] == "uniform" && w.size() >= 3) {
      unsigned long long seed = w.size() >= 4 ? std::stoull(w[3]) : 0ull;
      if (check(gsk_init_uniform(sim_, name.c_str(), std::stod(w[1]), std::stod(w[2]), seed))) return Code_ERR;
    } else {
      long long cnt = 1;
      for (int d : dims) cnt *= d;
      std::vector<double> f((size_t)cnt, fill);
      if (check(gsk_set_field(sim_, name.c_str(), f.data(), GSK_F64))) return Code_ERR;
    }
    (void)n3;
    Field& fd = (*Datas)[name];
    fd.init(dims);
    return check(gsk_get_field(sim_, name.c_str(), fd.data.data(), GSK_F64));
  }

  int Push(const std::string& name) override {
    if (!sim_) return Code_NORMAL;
    const bool ok = name == "eta" || (name == "concentration" && cfg_.physics == GSK_PHYS_PREC);
    auto it = Datas->find(name);
    if (!ok || it == Datas->end() || it->second.empty()) return Code_NORMAL;
    const long long want = (long long)(name == "eta" ? nvar_ : 1) * nx * ny * nz;
    if (it->second.count() != want) return Code_NORMAL;  // not this system's field
    return check(gsk_set_field(sim_, name.c_str(), it->second.data.data(), GSK_F64));
  }

  int pull(const std::string& name, std::vector<int> dims) {
    Field& f = (*Datas)[name];
    f.init(dims);
    return check(gsk_get_field(sim_, name.c_str(), f.data.data(), GSK_F64));
  }

  std::string energy_col(const std::string& col) {
    double out[16];
    if (gsk_reduce(sim_, GSK_RED_ENERGY, out) != GSK_OK) return "nan";
    if (col == "energy") return fmt(out[3]);
    if (col == "chemical_energy") return fmt(out[0]);
    if (col == "gradient_energy") return fmt(out[1]);
    return fmt(out[2]);
  }
};

// ------------------------------------------------------------------------------ martensite
// Semi-implicit Allen-Cahn with Landau f = A/2 sum eta^2 - B/3 sum eta^3 + C/4 (sum eta^2)^2
// and Khachaturyan elasticity on eta^2 (the HEAD contraction of src/dynamics_mart.cu:270-303
// with the variational factor 2 of src/dynamics_precipitate.cu:193).
class Dynamics_mart : public GpuDynamics {
 public:
  int Initialize() override {
    if (read_grid()) return Code_ERR;
    ref_ = read_rule("mart");
    if (ref_) {
      gsk_config_reference(&cfg_, GSK_PHYS_MART);
      warn_unread({"kappa", "mobility", "precision"}, "mart", "reference");
    } else {
      gsk_config_default(&cfg_, GSK_PHYS_MART);
      warn_unread({"transitiontemperature", "weightgradient", "weightchemical", "weightdislocation", "weightnoise",
                   "chemical_form", "fix"}, "mart", "spectral");
      for (const char* f : {"defect", "dislocationstress"}) {
        auto it = Datas->find(f);
        if (it != Datas->end() && !it->second.empty())
          log_line(std::string("Warning: mart ignores the field ") + f + " under the spectral rule");
      }
    }
    read_common();
    auto vt = Datas->find("varianttensor");
    if (vt == Datas->end() || vt->second.empty() || vt->second.dims.size() != 3 || vt->second.dims[1] != 3 ||
        vt->second.dims[2] != 3) {
      log_line("Error: variants' strain tensor does not set while initialize dynamics");  // src/dynamics_mart.cu:44-47
      return Code_ERR;
    }
    nvar_ = vt->second.dims[0];
    cfg_.nvar = nvar_;
    var_get(*Vars, "deltatime", &cfg_.dt, 1);
    double abc[3] = {cfg_.A, cfg_.B, cfg_.C};
    var_get(*Vars, "coefficient", abc, 3);
    cfg_.A = abc[0]; cfg_.B = abc[1]; cfg_.C = abc[2];
    var_get(*Vars, "kappa", &cfg_.kappa, 1);
    var_get(*Vars, "mobility", &cfg_.L, 1);
    var_get(*Vars, "weightelastic", &cfg_.w_el, 1);
    std::vector<std::string> w;
    if (!ref_ && var_values(*Vars, "precision", w) && w[0] == "fp64") cfg_.path = GSK_PATH_PLAIN64;
    if (ref_) {  // src/dynamics_mart.cu:18-41
      var_get(*Vars, "transitiontemperature", &cfg_.T0, 1);
      var_get(*Vars, "weightgradient", &cfg_.w_grad, 1);
      var_get(*Vars, "weightchemical", &cfg_.w_chem, 1);
      var_get(*Vars, "weightdislocation", &cfg_.w_dis, 1);
      var_get(*Vars, "weightnoise", &cfg_.noise, 1);
      var_get(*Vars, "temperature", &cfg_.temperature, 1);
      if (var_values(*Vars, "chemical_form", w) && w[0] == "master") cfg_.chem_form = 1;
    }
    double C[81];
    auto md = Datas->find("modulus");
    if (md != Datas->end() && md->second.count() == 81) {
      for (int i = 0; i < 81; ++i) C[i] = md->second.data[i];
    } else {
      double c[4] = {3.5, 1.5, 1.0, 1.0};  // src/dynamics_mart.cu:80
      int k = var_get(*Vars, "cubicmodulus", c, 4);
      if (k < 4) c[3] = 1.0;
      cubic_modulus(c[0] * c[3], c[1] * c[3], c[2] * c[3], C);
    }
    if (create()) return Code_ERR;
    if (check(gsk_build_btensor(sim_, C, vt->second.data.data(), nvar_))) return Code_ERR;
    if (ref_) {  // src/dynamics_mart.cu:117-142
      const long long n3 = (long long)nx * ny * nz;
      if (upload_optional("defect", n3) || upload_optional("dislocationstress", 6 * n3)) return Code_ERR;
    }
    return load_field("eta", "initeta", {nvar_, nx, ny, nz}, 0.0);
  }

  std::string Get(const std::string& col) override {
    if (col == "temperature") {
      auto it = Vars->find("temperature");
      return it == Vars->end() ? "nan" : it->second;
    }
    if (col == "eta" || col == "eta_average") {
      double m[16];
      if (gsk_reduce(sim_, GSK_RED_MEAN_ETA, m) != GSK_OK) return "nan";
      double s = 0;
      for (int v = 0; v < nvar_; ++v) s += m[v];
      return fmt(s / nvar_);
    }
    if (col == "energy" || col == "chemical_energy" || col == "gradient_energy" || col == "elastic_energy")
      return energy_col(col);
    if (col == "step") return fmt((double)gsk_get_step(sim_));
    // means of the force terms of the last step (src/dynamics_mart.cu:446-451), reference rule
    static const char* const forces[] = {"gradient", "chemical", "elastic", "dislocation"};
    for (int j = 0; j < 4; ++j)
      if (col == forces[j]) {
        double m[16];
        if (!ref_ || gsk_reduce(sim_, GSK_RED_MART_FORCES, m) != GSK_OK) return "nan";
        return fmt(m[j]);
      }
    return "nan";
  }

  int Pull(const std::string& name) override {
    if (name == "eta" && sim_) return pull("eta", {nvar_, nx, ny, nz});
    return Code_NORMAL;
  }
};

// ------------------------------------------------------------------------------ precipitate
// Cahn-Hilliard for c (k^4 implicit, optional stabilisation S k^2) and Allen-Cahn for eta
// (src/dynamics_precipitate.cu:118-275) with the dilatational c-sourced elasticity.
class Dynamics_precipitate : public GpuDynamics {
 public:
  int Initialize() override {
    if (read_grid()) return Code_ERR;
    ref_ = read_rule("precipitate");
    if (ref_) {
      gsk_config_reference(&cfg_, GSK_PHYS_PREC);
      warn_unread({"stabilization", "dilatation", "precision"}, "precipitate", "reference");
    } else {
      gsk_config_default(&cfg_, GSK_PHYS_PREC);
      warn_unread({"weightnoise"}, "precipitate", "spectral");
    }
    read_common();
    auto vt = Datas->find("varianttensor");
    if (vt == Datas->end() || vt->second.empty()) {
      log_line("Error: variants' strain tensor deoos not set while initialize dynamics");  // :32-35
      return Code_ERR;
    }
    nvar_ = vt->second.dims[0];
    cfg_.nvar = nvar_;
    var_get(*Vars, "deltatime", &cfg_.dt, 1);
    double A[7];
    if (var_get(*Vars, "coefficient", A, 7) == 7)
      for (int i = 0; i < 7; ++i) cfg_.a[i + 1] = A[i];
    var_get(*Vars, "arfi", &cfg_.kappa, 1);
    var_get(*Vars, "beta", &cfg_.kappa_c, 1);
    var_get(*Vars, "meta", &cfg_.M, 1);
    var_get(*Vars, "lpp", &cfg_.L, 1);
    var_get(*Vars, "xi", &cfg_.xi, 1);
    var_get(*Vars, "stabilization", &cfg_.S, 1);
    double cc[2] = {cfg_.c1, cfg_.c2};
    var_get(*Vars, "concentration", cc, 2);
    cfg_.c1 = cc[0]; cfg_.c2 = cc[1];
    std::vector<std::string> w;
    if (ref_) {
      // weightnoise w gives noise 1e-4 w on c and eta (src/dynamics_precipitate.cu:30, :203, :244)
      double wn = 1.0;
      var_get(*Vars, "weightnoise", &wn, 1);
      cfg_.noise = 1e-4 * wn;
      if (var_values(*Vars, "elastic_source", w) && w[0] != "eta2") {
        log_line("Error: the reference rule uses the eta^2 elastic source");
        return Code_ERR;
      }
    } else if (var_values(*Vars, "elastic_source", w) && w[0] != "c") {
      log_line("Error: elastic_source " + w[0] + " is not supported by the spectral rule (only c)");
      return Code_ERR;
    }
    if (!ref_ && var_values(*Vars, "precision", w) && w[0] == "fp64") cfg_.path = GSK_PATH_PLAIN64;
    double mod[3] = {3.5, 1.5, 1.0};  // src/dynamics_precipitate.cu:55-56
    var_get(*Vars, "modulus", mod, 3);
    double C[81];
    cubic_modulus(mod[0], mod[1], mod[2], C);
    double d0 = 0.02;
    var_get(*Vars, "dilatation", &d0, 1);
    double eps[9] = {d0, 0, 0, 0, d0, 0, 0, 0, d0};
    if (create()) return Code_ERR;
    if (ref_) {  // eta^2 source with the variant tensors (src/dynamics_precipitate.cu:57-67)
      if (vt->second.count() != (long long)nvar_ * 9) {
        log_line("Error: varianttensor must hold nvar 3x3 tensors");
        return Code_ERR;
      }
      if (check(gsk_build_btensor(sim_, C, vt->second.data.data(), nvar_))) return Code_ERR;
    } else if (check(gsk_build_btensor(sim_, C, eps, 1))) {
      return Code_ERR;
    }
    if (load_field("concentration", "initconcentration", {nx, ny, nz}, cfg_.c1 + (cfg_.c2 - cfg_.c1) / 3))
      return Code_ERR;  // default c (src/dynamics_precipitate.cu:46)
    return load_field("eta", "initeta", {nvar_, nx, ny, nz}, 0.0);
  }

  std::string Get(const std::string& col) override {
    if (col == "temperature") {
      auto it = Vars->find("temperature");
      return it == Vars->end() ? "nan" : it->second;
    }
    if (col == "eta_average" || col == "eta") {
      double m[16];
      if (gsk_reduce(sim_, GSK_RED_MEAN_ETA, m) != GSK_OK) return "nan";
      double s = 0;
      for (int v = 0; v < nvar_; ++v) s += m[v];
      return fmt(s / nvar_);
    }
    if (col == "mass") {
      double m[16];
      if (gsk_reduce(sim_, GSK_RED_MASS, m) != GSK_OK) return "nan";
      std::ostringstream os;
      os.precision(17);
      os << m[0];
      return os.str();
    }
    if (col == "energy" || col == "chemical_energy" || col == "gradient_energy" || col == "elastic_energy")
      return energy_col(col);
    if (col == "step") return fmt((double)gsk_get_step(sim_));
    return "nan";
  }

  int Pull(const std::string& name) override {
    if (!sim_) return Code_NORMAL;
    if (name == "eta") return pull("eta", {nvar_, nx, ny, nz});
    if (name == "concentration") return pull("concentration", {nx, ny, nz});
    return Code_NORMAL;
  }
};

// ------------------------------------------------------------------------------ stress
// Elastic stress of the order parameters (src/dynamics_stress.cu:19-194), the input of the
// martensite plugin's dislocation term: `stress` (6,n,n,n; components 00,01,02,11,12,22)
// and `defect` (n,n,n; 1 where any |eta_v| > 0.9), computed on the GPU by gsk_stress.
class Dynamics_stress : public GpuDynamics {
 public:
  int Initialize() override {
    if (read_grid()) return Code_ERR;
    if (!cube()) {
      log_line("Error: the stress plugin takes a periodic n x n x n grid with one spacing");
      return Code_ERR;
    }
    gsk_config_default(&cfg_, GSK_PHYS_MART);
    read_common();
    var_get(*Vars, "xi", &xi_, 1);  // default 4000 (:25)
    std::vector<std::string> w;
    if (var_values(*Vars, "stress_source", w) && w[0] == "eta") source_ = 1;  // master side of :143-147
    auto vt = Datas->find("varianttensor");
    if (vt == Datas->end() || vt->second.empty() || vt->second.count() % 9 != 0) {
      log_line("Error: variants' strain tensor does not set while initialize dynamics");  // :27-30
      return Code_ERR;
    }
    nvar_ = (int)(vt->second.count() / 9);
    eps_ = vt->second.data;
    auto md = Datas->find("modulus");
    if (md != Datas->end() && md->second.count() == 81) {
      C_ = md->second.data;
    } else {  // the plugin's own default modulus (:44-64)
      C_.assign(81, 0.0);
      auto c = [&](int i, int j, int k, int l, double v) { C_[((i * 3 + j) * 3 + k) * 3 + l] = v; };
      c(0, 0, 0, 0, 5.39); c(1, 1, 1, 1, 5.39); c(2, 2, 2, 2, 5.22);
      c(0, 0, 1, 1, 3.39); c(1, 1, 0, 0, 3.39);
      c(1, 1, 2, 2, 3.56); c(2, 2, 1, 1, 3.56); c(2, 2, 0, 0, 3.56); c(0, 0, 2, 2, 3.56);
      for (auto p : {std::make_pair(0, 1), std::make_pair(1, 0)})
        for (auto q : {std::make_pair(0, 1), std::make_pair(1, 0)}) c(p.first, p.second, q.first, q.second, 0.6);
      for (auto p : {std::make_pair(1, 2), std::make_pair(2, 1)})
        for (auto q : {std::make_pair(1, 2), std::make_pair(2, 1)}) c(p.first, p.second, q.first, q.second, 0.77);
      for (auto p : {std::make_pair(0, 2), std::make_pair(2, 0)})
        for (auto q : {std::make_pair(0, 2), std::make_pair(2, 0)}) c(p.first, p.second, q.first, q.second, 0.77);
    }
    Field& eta = (*Datas)["eta"];
    if (eta.empty()) eta.init({nvar_, nx, ny, nz});  // zero when not read (:36-41)
    if (eta.count() != (long long)nvar_ * nx * ny * nz) {
      log_line("Error: eta does not match gridsize and the variant count");
      return Code_ERR;
    }
    return Code_NORMAL;
  }

  int Calculate() override {
    const Field& eta = (*Datas)["eta"];
    std::vector<float> e(eta.data.begin(), eta.data.end());
    const long long n3 = (long long)nx * ny * nz;
    std::vector<float> st((size_t)(6 * n3)), df((size_t)n3);
    if (check(gsk_stress(nx, dx, cfg_.kconv, C_.data(), eps_.data(), nvar_, xi_, source_, e.data(), st.data(),
                         df.data(), cfg_.device)))
      return Code_ERR;
    Field& s = (*Datas)["stress"];
    s.init({6, nx, ny, nz});
    for (size_t i = 0; i < st.size(); ++i) s.data[i] = st[i];
    Field& d = (*Datas)["defect"];
    d.init({nx, ny, nz});
    for (size_t i = 0; i < df.size(); ++i) d.data[i] = df[i];
    return Code_NORMAL;
  }
  int CalculateMany(int) override { return Calculate(); }  // the result does not change with steps

  std::string Get(const std::string&) override { return "nan"; }
  int Pull(const std::string&) override { return Code_NORMAL; }

 private:
  double xi_ = 4000.0;
  int source_ = 2;  // eta^2 (HEAD side)
  std::vector<double> eps_, C_;
};

// ------------------------------------------------------------------------------ cores
// Initial microstructure of spherical nuclei (src/dynamics_cores.cu:11-195): eta_v = 1 and
// c = c2 inside each core, c = c1 elsewhere; periodic images along each axis. The reference
// draws positions and variants with libc random(); here Philox4x32-10 at counter
// (core, attempt, call, 3) keyed by `seed` draws them (x, y, z, variant from words 0-3).
class Dynamics_cores : public GpuDynamics {
 public:
  int Initialize() override {
    if (Vars->find("gridsize") == Vars->end()) {  // :14 defaults
      nx = ny = nz = 16;
      dx = dy = dz = 0.1;
    } else if (read_grid()) {
      return Code_ERR;
    }
    var_get(*Vars, "variantn", &V_, 1);
    var_get(*Vars, "coresn", &cores_, 1);
    var_get(*Vars, "radius", &radius_, 1);
    double cc[2] = {c1_, c2_};
    var_get(*Vars, "concentration", cc, 2);
    c1_ = cc[0];
    c2_ = cc[1];
    std::vector<std::string> w;
    if (var_values(*Vars, "method", w)) method_ = w[0];
    var_get(*Vars, "seed", &seed_, 1);
    if (V_ < 1 || cores_ < 0 || radius_ < 0) {
      log_line("Error: cores needs variantn >= 1, coresn >= 0, radius >= 0");
      return Code_ERR;
    }
    (*Datas)["eta"].init({V_, nx, ny, nz});
    (*Datas)["concentration"].init({nx, ny, nz}, c1_);
    return Code_NORMAL;
  }

  int Calculate() override {
    Field& eta = (*Datas)["eta"];
    Field& c = (*Datas)["concentration"];
    mark_.assign((size_t)nx * ny * nz, 0);
    if (method_ == "random") {
      eta.init({V_, nx, ny, nz});
      c.init({nx, ny, nz}, c1_);
      for (int cn = 0; cn < cores_; ++cn) {
        unsigned int o[4] = {0, 0, 0, 0};
        for (unsigned int t = 0; t < 100000; ++t) {
          const unsigned int ctr[4] = {(unsigned)cn, t, (unsigned)calls_, 3u};
          gsk_philox4x32(ctr, seed_, o);
          if (!mark_[((size_t)(o[0] % nx) * ny + o[1] % ny) * nz + o[2] % nz]) break;
        }
        place(eta, c, (int)(o[0] % nx), (int)(o[1] % ny), (int)(o[2] % nz), (int)(o[3] % V_));
      }
    } else if (method_ == "regular") {  // :86-115, variant cn % V
      for (int cn = 0; cn < cores_; ++cn) place(eta, c, nx / 2, ny / 2, nz / cores_ / 2 + nx / cores_ * cn, cn % V_);
    } else if (method_ == "regular2d" || method_ == "regular3d") {  // :116-190
      const int d3 = method_ == "regular3d";
      int id = 0;
      for (int cn1 = 0; cn1 < cores_; ++cn1)
        for (int cn2 = 0; cn2 < cores_; ++cn2)
          for (int cn3 = 0; cn3 < (d3 ? cores_ : 1); ++cn3, ++id) {
            const int ox = d3 ? nx / cores_ / 2 + nx / cores_ * cn3 : nx / 2;
            const int oy = ny / cores_ / 2 + ny / cores_ * cn2;
            const int oz = nz / cores_ / 2 + nx / cores_ * cn1;
            unsigned int o[4];
            const unsigned int ctr[4] = {(unsigned)id, 0u, (unsigned)calls_, 3u};
            gsk_philox4x32(ctr, seed_, o);
            place(eta, c, ox, oy, oz, (int)(o[3] % V_));
          }
    } else {
      log_line("Error: cores method " + method_ + " is unknown");
      return Code_ERR;
    }
    ++calls_;
    return Code_NORMAL;
  }
  int CalculateMany(int n) override {
    for (int i = 0; i < n; ++i)
      if (Calculate() != Code_NORMAL) return Code_ERR;
    return Code_NORMAL;
  }

  std::string Get(const std::string&) override { return "nan"; }
  int Pull(const std::string&) override { return Code_NORMAL; }

 private:
  int V_ = 4, cores_ = 5, radius_ = 5;
  double c1_ = 0.2, c2_ = 0.44;
  std::string method_ = "random";
  unsigned long long seed_ = 0;
  int calls_ = 0;
  std::vector<char> mark_;

  // a sphere of radius r about (ox, oy, oz) and its images shifted by one period along
  // one axis (the seven tests of the reference)
  void place(Field& eta, Field& c, int ox, int oy, int oz, int v) {
    const long long r2 = (long long)radius_ * radius_;
    auto in = [&](long long a, long long b, long long d) { return a * a + b * b + d * d <= r2; };
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const long long a = i - ox, b = j - oy, d = k - oz;
          if (in(a, b, d) || in(a + nx, b, d) || in(a - nx, b, d) || in(a, b + ny, d) || in(a, b - ny, d) ||
              in(a, b, d + nz) || in(a, b, d - nz)) {
            const size_t idx = ((size_t)i * ny + j) * nz + k;
            mark_[idx] = 1;
            c.data[idx] = c2_;
            eta.data[(size_t)v * nx * ny * nz + idx] = 1.0;
          }
        }
  }
};

Register reg_mart("mart", [] { return std::unique_ptr<Dynamics>(new Dynamics_mart); });
Register reg_stress("stress", [] { return std::unique_ptr<Dynamics>(new Dynamics_stress); });
Register reg_cores("cores", [] { return std::unique_ptr<Dynamics>(new Dynamics_cores); });
Register reg_prec("precipitate", [] { return std::unique_ptr<Dynamics>(new Dynamics_precipitate); });

}  // namespace
}  // namespace gs
