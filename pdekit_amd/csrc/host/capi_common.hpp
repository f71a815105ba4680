// Shared error reporting of the C ABI (thread-local last error message).
#pragma once

#include <string>

namespace pdrd {

extern thread_local std::string g_last_error;
int set_error(int code, const std::string &msg);

}  // namespace pdrd
