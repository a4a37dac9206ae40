// Error reporting and version of libdosehip.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "dh_common.h"

namespace dh {

static thread_local std::string t_last_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_last_error = buf;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return DH_ERR_HIP;
    }
    return DH_OK;
}

}  // namespace dh

extern "C" int dh_version(void) { return 100; }

extern "C" int dh_last_error(char *buf, size_t len) {
    const std::string &m = dh::t_last_error;
    if (buf && len) {
        size_t n = m.size() < len - 1 ? m.size() : len - 1;
        memcpy(buf, m.data(), n);
        buf[n] = '\0';
    }
    return (int)m.size();
}

// Quad RLSP volume plus its table of 4x4x4 cell blocks (value and run radius)
// and two ints of scratch per block.
extern "C" int64_t dh_padded_len(dh_grid g) {
    return (int64_t)4 * g.nx * g.ny * (g.nz + 1) +
           (int64_t)4 * ((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4);
}
