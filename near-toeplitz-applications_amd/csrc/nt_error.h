// Per-thread error message shared by every C-ABI entry point (nt_last_error).
#pragma once

namespace nt {

// Store a printf-style message for nt_last_error() and return `code`
// (NT_OK clears the message).
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace nt
