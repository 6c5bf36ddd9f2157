// Every C entry point of libcvchip runs its body through guarded(): no C++
// exception (a failed host allocation of a std::vector, a std::string, new)
// crosses the C ABI into the caller.  A failed host allocation returns
// CVCHIP_ERR_OOM, any other exception CVCHIP_ERR_STATE, each with a message
// for cvchip_last_error().  Host-only: included by the HIP and the C++ sources.
#pragma once

#include <exception>
#include <new>

#include "../../include/cvchip.h"

namespace cvchip {

void set_error(const char *fmt, ...);

template <typename F>
int guarded(const char *entry, F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    set_error("%s: host memory allocation failed", entry);
    return CVCHIP_ERR_OOM;
  } catch (const std::exception &e) {
    set_error("%s: %s", entry, e.what());
    return CVCHIP_ERR_STATE;
  } catch (...) {
    set_error("%s: unexpected exception", entry);
    return CVCHIP_ERR_STATE;
  }
}

}  // namespace cvchip
