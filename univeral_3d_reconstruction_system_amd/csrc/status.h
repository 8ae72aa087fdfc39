// Status codes and the thread-local error message of the C ABI, for code with no HIP
// dependency (the file readers and writers); common.h adds the HIP error check.
#pragma once
#include <string>

#include "../../include/u3d_hip.h"

namespace u3d {

void setError(const std::string &msg); // the message u3d_last_error() returns

#define U3D_CHECK(cond, code, msg)                                          \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::u3d::setError(msg);                                           \
            return code;                                                    \
        }                                                                   \
    } while (0)

#define U3D_TRY(call)                                                       \
    do {                                                                    \
        int rc_ = (call);                                                   \
        if (rc_ != U3D_OK)                                                  \
            return rc_;                                                     \
    } while (0)

} // namespace u3d
