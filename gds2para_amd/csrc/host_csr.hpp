// host_csr.hpp -- host-side CSR used to build and upload structured operators.
#pragma once
#include <cstdint>
#include <vector>

namespace gp {
struct HostCsr {
  int64_t nr = 0, nc = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
};
}  // namespace gp
