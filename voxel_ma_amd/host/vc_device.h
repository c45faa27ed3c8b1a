// vc_device.h -- the process-wide device context used by the C++ layer
// (one vc_ctx on the device named by VC_DEVICE, default 0).
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/vc_api.h"

namespace vcdev
{
vc_ctx* context();                 // created on first use; throws on failure
std::string last_error();          // message of the last failed call
void check(int rc, const char* what); // throws std::runtime_error on rc < 0
} // namespace vcdev
