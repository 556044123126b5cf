// Host-side error reporting shared by every translation unit.
#pragma once
#include <string>

namespace tcx {

// Records msg as this thread's last error (tcx_last_error) and returns code.
int fail(int code, const std::string& msg);
// Clears this thread's last error.
void clear_error();

}  // namespace tcx
