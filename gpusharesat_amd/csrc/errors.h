// Thread-local last-error message shared by every C entry point of the library
// (read back with gss_last_error()).
#pragma once

namespace gss {
// Records msg as the calling thread's last error and returns code.
int record_error(int code, const char *msg);
}  // namespace gss
