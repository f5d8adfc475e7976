// status.cpp -- thread-local storage for gp_last_error().
#include "gp_status.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace gp {
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace gp

extern "C" const char *gp_last_error(void) { return g_last_error.c_str(); }
