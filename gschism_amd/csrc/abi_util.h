// abi_util.h -- status/last-error plumbing for the extern "C" entry points.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "host_mesh.h"
#include "../../include/gsh.h"

namespace gsh {
void set_last_error(const std::string& msg);
void clear_last_error();
}  // namespace gsh

// Runs `body`; converts exceptions into a status code and a thread-local message.
#define GSH_API_BODY(...)                                                      \
  try {                                                                        \
    gsh::clear_last_error();                                                   \
    __VA_ARGS__;                                                               \
    return GSH_OK;                                                             \
  } catch (const gsh::Error& e) {                                              \
    gsh::set_last_error(e.what());                                             \
    return e.code;                                                             \
  } catch (const std::bad_alloc&) {                                            \
    gsh::set_last_error("host allocation failed");                             \
    return GSH_E_OOM;                                                          \
  } catch (const std::exception& e) {                                          \
    gsh::set_last_error(e.what());                                             \
    return GSH_E_ARG;                                                          \
  }

#define GSH_REQUIRE(cond, code, msg)                                           \
  do {                                                                         \
    if (!(cond)) throw gsh::Error((code), (msg));                              \
  } while (0)
