// capi_util.hpp -- exception-to-status translation shared by the extern "C"
// entry points of libdft_core.so (message in dft_core_last_error()).
#pragma once

#include <functional>

namespace dftc {
// Runs f; returns DFT_OK, or DFT_EINVAL for std::invalid_argument and
// DFT_ENUMERIC for any other exception, keeping its message.
int guarded_call(const std::function<void()>& f);
}  // namespace dftc
