// Shared helpers of libinching_hip.so (host side).
#pragma once
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/inching.h"

namespace inch {

void set_error(const char *fmt, ...);

}  // namespace inch

#define INCH_FAIL(code, ...)            \
  do {                                  \
    ::inch::set_error(__VA_ARGS__);     \
    return (code);                      \
  } while (0)

#define INCH_REQUIRE(cond, ...)                                  \
  do {                                                           \
    if (!(cond)) INCH_FAIL(INCH_ERR_INVALID, __VA_ARGS__);       \
  } while (0)
