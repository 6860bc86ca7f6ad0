// qbx_common.h -- error type shared by the translation units of libqbx_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>

namespace qbx {
struct QbxError : std::runtime_error {
    int code;
    QbxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
}  // namespace qbx

#define QBX_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw qbx::QbxError(QBX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
