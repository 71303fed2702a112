// rtapi.hpp — helpers shared by the tiler classes (not part of the reference's API).
#pragma once

#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "rt_hip.h"

namespace rtapi {

// A HIP context for `device` owned by a shared_ptr (tilers stay copyable, as the
// reference's handle-holding classes are).
std::shared_ptr<rt_ctx> make_ctx(int device);

// File name part of a path.
std::string basename(const std::string& path);

// TinyMT's default parameter set (tinymt32def.clh:48-50) for n generators.
void load_default_params(rt_ctx* ctx, int64_t n);

}  // namespace rtapi
