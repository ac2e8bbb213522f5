// Shared host-side helpers: error reporting across the C ABI.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/crushrs.h"

namespace crush {

// Thrown inside the library; converted to a status code at the ABI edge.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

void set_last_error(const std::string& msg);
void clear_last_error();

// Run f, mapping exceptions to status codes and the thread's last error.
template <class F>
int32_t guarded(F&& f) {
    try {
        clear_last_error();
        f();
        return CRUSH_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return CRUSH_E_OOM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return CRUSH_E_INPUT;
    }
}

}  // namespace crush
