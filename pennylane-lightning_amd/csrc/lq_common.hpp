// Common host-side definitions: error type and message format, HIP checks.
//
// Errors mirror the reference's PL_ABORT / LightningException
// (core/utils/Error.hpp:23-140): the message reads
//   "[file][Line:n][Method:f]: Error in PennyLane Lightning: <msg>"
// and crosses the C-ABI as a non-zero status plus lq_last_error().
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace lqamd {

using cd = std::complex<double>;

class LightningException : public std::exception {
  public:
    explicit LightningException(std::string msg) : msg_(std::move(msg)) {}
    const char *what() const noexcept override { return msg_.c_str(); }

  private:
    std::string msg_;
};

[[noreturn]] void abort_with(const char *msg, const char *file, int line, const char *func);

} // namespace lqamd

#define LQ_ABORT(msg) ::lqamd::abort_with((msg), __FILE__, __LINE__, __func__)
#define LQ_ABORT_IF(cond, msg)                                                 \
    do {                                                                       \
        if (cond) LQ_ABORT(msg);                                               \
    } while (0)
#define LQ_ABORT_IF_NOT(cond, msg) LQ_ABORT_IF(!(cond), msg)
