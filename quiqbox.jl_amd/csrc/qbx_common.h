// qbx_common.h -- error type shared by the translation units of libqbx_hip.so.
#pragma once
#include "qbx.h"

#include <hip/hip_runtime.h>
#include <new>
#include <stdexcept>
#include <string>

namespace qbx {
struct QbxError : std::runtime_error {
    int code;
    QbxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// message of the calling thread's last failed call (qbx_last_error)
inline thread_local std::string g_last_error;

// Run f, turning exceptions into the C-ABI's negative return codes.
template <typename F> int guarded(F&& f)
{
    try {
        f();
        return QBX_OK;
    } catch (const QbxError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return QBX_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return QBX_EINTERNAL;
    }
}
}  // namespace qbx

#define QBX_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw qbx::QbxError(QBX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
