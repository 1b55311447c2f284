// kifi/detail.h -- the one device context the drop-in classes share, and the
// error policy.  The reference asserts on bad input and has no other error
// channel (Integrator.cpp:35-36, Volume.cpp:22-26); these classes abort with the
// library's message on any non-OK status, in every build type.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "kifi.h"   // the C ABI (include/kifi.h)

namespace kifi {
namespace detail {

// Device of the shared context: KIFI_DEVICE in the environment, else 0.
inline kifi_ctx *context()
{
    static kifi_ctx *ctx = [] {
        const char *env = std::getenv("KIFI_DEVICE");
        kifi_ctx *c = nullptr;
        if (kifi_ctx_create(env ? std::atoi(env) : 0, &c) != KIFI_OK) {
            std::fprintf(stderr, "kifi: no HIP device context (kifi_ctx_create failed)\n");
            std::abort();
        }
        return c;
    }();
    return ctx;
}

inline void check(kifi_status s, const char *what)
{
    if (s != KIFI_OK) {
        std::fprintf(stderr, "kifi: %s failed (status %d): %s\n", what, (int)s, kifi_last_error(context()));
        std::abort();
    }
}

} // namespace detail
} // namespace kifi
