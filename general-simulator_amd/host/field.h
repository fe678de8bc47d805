// Named N-d fields on the host and their on-disk formats.
//
// Mirrors Data<Real> of General-Simulator as the scripts see it
// (include/datamain.th:113-192): a row-major array with a rank and dimensions,
// and the two file formats of include/datamain.th:338-457:
//   text   : "rank d1 .. dn " on the first line, then one value per line
//   binary : int32 rank, int32 dims[rank], float32 values
#pragma once
#include <map>
#include <string>
#include <vector>

namespace gs {

struct Field {
  std::vector<int> dims;      // dims[0..rank-1]
  std::vector<double> data;   // row-major, last index fastest
  bool empty() const { return dims.empty(); }
  long long count() const {
    long long c = 1;
    for (int d : dims) c *= d;
    return dims.empty() ? 0 : c;
  }
  void init(const std::vector<int>& d, double v = 0.0) {
    dims = d;
    data.assign((size_t)count(), v);
  }
  double& at(std::initializer_list<int> idx);
  int write_text(const std::string& file) const;
  int write_binary(const std::string& file) const;
  int read_text(const std::string& file);
  int read_binary(const std::string& file);
};

using FieldMap = std::map<std::string, Field>;

// String variables of the script host (`set key value`, `:x = expr`), with the aliases of
// the `link target alias` command (include/datamap.th:54-100, 126-139): once linked, the
// alias reads and writes its target's value and its own value is dropped. Chains resolve
// transitively; a cycle stops after 64 hops at the last name reached.
class VarMap {
 public:
  using Map = std::map<std::string, std::string>;
  using iterator = Map::iterator;
  using const_iterator = Map::const_iterator;

  std::string resolve(const std::string& k) const {
    std::string cur = k;
    for (int hop = 0; hop < 64; ++hop) {
      auto it = links_.find(cur);
      if (it == links_.end()) break;
      cur = it->second;
    }
    return cur;
  }
  std::string& operator[](const std::string& k) { return m_[resolve(k)]; }
  iterator find(const std::string& k) { return m_.find(resolve(k)); }
  const_iterator find(const std::string& k) const { return m_.find(resolve(k)); }
  size_t count(const std::string& k) const { return m_.count(resolve(k)); }
  iterator begin() { return m_.begin(); }
  iterator end() { return m_.end(); }
  const_iterator begin() const { return m_.begin(); }
  const_iterator end() const { return m_.end(); }
  // alias -> target; -1 when the target has no value (the reference's Map::link refuses too)
  int link(const std::string& target, const std::string& alias) {
    if (!count(target) || resolve(target) == alias) return -1;
    m_.erase(alias);
    links_[alias] = target;
    return 0;
  }
  int unlink(const std::string& alias) { return links_.erase(alias) ? 0 : -1; }
  bool linked(const std::string& k) const { return links_.count(k) != 0; }

 private:
  Map m_;
  std::map<std::string, std::string> links_;
};

}  // namespace gs
