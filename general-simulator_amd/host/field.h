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
using VarMap = std::map<std::string, std::string>;

}  // namespace gs
