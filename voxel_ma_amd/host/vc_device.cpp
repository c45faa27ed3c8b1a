#include "vc_device.h"

#include <cstdlib>
#include <mutex>

namespace vcdev
{
namespace
{
vc_ctx* g_ctx = nullptr;
std::once_flag g_once;
int g_rc = 0;
} // namespace

vc_ctx* context()
{
    std::call_once(g_once, [] {
        int dev = 0;
        if (const char* d = std::getenv("VC_DEVICE"))
            dev = std::atoi(d);
        g_rc = vc_ctx_create(dev, &g_ctx);
    });
    if (!g_ctx)
        throw std::runtime_error("vc_ctx_create failed (" + std::to_string(g_rc) +
                                 "): no usable MI355X device");
    return g_ctx;
}

std::string last_error()
{
    if (!g_ctx)
        return "no context";
    char buf[1024];
    vc_last_error(g_ctx, buf, sizeof(buf));
    return buf;
}

void check(int rc, const char* what)
{
    if (rc < 0)
        throw std::runtime_error(std::string(what) + " failed (" +
                                 std::to_string(rc) + "): " + last_error());
}
} // namespace vcdev
