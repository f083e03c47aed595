// Shared helpers of the host-side translation units (not part of the ABI).
#pragma once

#include <string>

namespace bm3d {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

}  // namespace bm3d
