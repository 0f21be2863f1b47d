// errors.hpp - error type of the library and its message format.
//
// Codes and the message layout follow the reference
// (src/util/VshipExceptions.hpp:12-113):
//   VshipException\n<Name>: <advice>\n - At line <L> of <file>\n[Detail: <d>\n]
// Nothing thrown here crosses the C ABI; api.cpp turns it into a return code.
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "../../include/VshipAPI.h"

namespace vship {

const char* error_text(Vship_Exception code);

class Error {
public:
    Error(Vship_Exception type, const char* file, int line, std::string detail = "")
        : type_(type), file_(file), line_(line), detail_(std::move(detail)) {}
    Vship_Exception type() const { return type_; }
    std::string message() const;

private:
    Vship_Exception type_;
    std::string file_;
    int line_;
    std::string detail_;
};

#define VSHIP_THROW(code, detail) throw ::vship::Error((code), __FILE__, __LINE__, (detail))

#define VSHIP_HIP_CHECK(expr)                                                        \
    do {                                                                             \
        hipError_t vship_err_ = (expr);                                              \
        if (vship_err_ != hipSuccess)                                                \
            throw ::vship::Error(Vship_HIPError, __FILE__, __LINE__,                 \
                                 hipGetErrorString(vship_err_));                     \
    } while (0)

}  // namespace vship
