// Shared helpers of the hippymfem_amd library: the error slot and HIP checks.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/hippymfem_amd.h"

namespace hm {

// largest element dof count of the per-thread element arrays
constexpr int HM_MAXND = 64;

void set_error(const char* fmt, ...);

}  // namespace hm

#define HM_FAIL(code, ...)              \
  do {                                  \
    ::hm::set_error(__VA_ARGS__);       \
    return (code);                      \
  } while (0)

#define HM_REQUIRE(cond, ...)                             \
  do {                                                    \
    if (!(cond)) HM_FAIL(HM_ERR_ARG, __VA_ARGS__);        \
  } while (0)
