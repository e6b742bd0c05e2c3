// Thread-local last-error message shared by every C entry point of the library
// (read back with gss_last_error()).
#pragma once

namespace gss {
// Records msg as the calling thread's last error and returns code.
int record_error(int code, const char *msg);
// Writes "what took ms" to stderr, with the seconds since the library's
// first such check, when GPUSHARESAT_LOG_STALLS_MS is set and ms reaches it
// (host calls that can hold a serving thread or an engine's allocator
// thread: step submits and waits, record buffer allocations and releases,
// clause bucket growth, step graph captures).
void log_stall(const char *what, double ms);
}  // namespace gss
