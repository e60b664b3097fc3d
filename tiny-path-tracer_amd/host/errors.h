// Per-thread last-error string behind tpt_last_error().
#ifndef TPT_HOST_ERRORS_H
#define TPT_HOST_ERRORS_H

#include <string>

namespace tpt {
// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string &msg);
}  // namespace tpt

#endif  // TPT_HOST_ERRORS_H
