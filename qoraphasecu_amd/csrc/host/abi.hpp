// C-ABI status codes -> the C++ API's exceptions.
#pragma once

#include <stdexcept>
#include <string>

#include "qpch.h"

namespace qpcu::detail {

inline void qpch_check(int status) {
  if (status == QPCH_OK) return;
  const std::string msg = qpch_last_error();
  if (status == QPCH_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

}  // namespace qpcu::detail
