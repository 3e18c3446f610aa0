// Per-thread error message behind supra_last_error() and the fail() helper
// every entry point uses to return a SUPRA_ERR_* code with a message.
#pragma once

#include <cstdarg>
#include <cstdio>

namespace supra_amd {

char* error_buffer();   // thread-local, 512 bytes (status.cpp)

inline void clear_error() { error_buffer()[0] = '\0'; }

#if defined(__GNUC__)
__attribute__((format(printf, 2, 3)))
#endif
inline int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace supra_amd
