// ugks_runtime.hpp -- block handle and error plumbing shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ugks_device.hpp"

// The opaque block handle of the C ABI: the per-block constants (with the
// device copies of the velocity tables they point to).
struct ugks_block {
    ugks::Consts c;
    void *dev_tables;   // one allocation holding quad, weight and the mirrors
    int device;
};

namespace ugks_rt {
int fail(int code, const char *msg);
int hip_fail(hipError_t e, const char *where);
int check_launch(const char *name);
int copy_small(void *dst_host, const void *src_dev, size_t bytes, hipStream_t s);
}  // namespace ugks_rt
