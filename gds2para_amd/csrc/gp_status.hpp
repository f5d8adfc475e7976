// gp_status.hpp -- thread-local error text behind gp_last_error().
#pragma once
#include <string>
#include "gds2para.h"

namespace gp {
void set_error(const std::string &msg);
inline gp_status fail(gp_status st, const std::string &msg) {
  set_error(msg);
  return st;
}
}  // namespace gp
