// honei/util/exception.hh -- exception types of the operator API
// (util/exception.hh:78-170, la/vector_error.hh, la/matrix_error.hh).
#pragma once

#include <stdexcept>
#include <string>

#include "honei_hip.h"

namespace honei {

class Exception : public std::runtime_error {
  public:
    explicit Exception(const std::string &m) : std::runtime_error(m) {}
};
class InternalError : public Exception {
  public:
    explicit InternalError(const std::string &m) : Exception("Internal error: " + m) {}
};
class ExternalError : public Exception {
  public:
    ExternalError(const std::string &lib, const std::string &m) : Exception(lib + " error: " + m) {}
};
class VectorSizeDoesNotMatch : public Exception {
  public:
    VectorSizeDoesNotMatch(unsigned long a, unsigned long b)
        : Exception("Vector size '" + std::to_string(a) + "' does not match expected size '" + std::to_string(b) +
                    "'") {}
};
class MatrixRowsDoNotMatch : public Exception {
  public:
    MatrixRowsDoNotMatch(unsigned long a, unsigned long b)
        : Exception("Matrix rows '" + std::to_string(a) + "' do not match expected '" + std::to_string(b) + "'") {}
};
class NotConverged : public Exception {
  public:
    explicit NotConverged(const std::string &m) : Exception("not converged: " + m) {}
};

// Map a C-ABI status to the exception types above.
inline void hip_check(int status, const char *what) {
    if (status == HONEI_HIP_SUCCESS) return;
    const std::string msg = std::string(what) + ": " + honei_hip_last_error();
    switch (status) {
        case HONEI_HIP_ERROR_SIZE_MISMATCH: throw Exception(msg);
        case HONEI_HIP_ERROR_NOT_CONVERGED: throw NotConverged(msg);
        case HONEI_HIP_ERROR_INVALID_ARGUMENT: throw InternalError(msg);
        default: throw ExternalError("HIP", msg);
    }
}

}  // namespace honei
