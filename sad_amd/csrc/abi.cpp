/*
 * abi.cpp - status reporting of the C ABI: every entry point returns an int
 * code and records a message for sadh_last_error() on the calling thread.
 */
#include <hip/hip_runtime.h>
#include <string>

#include "sad_device.h"

namespace sadh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int launch_status(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return SADH_OK;
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e;
}

}  // namespace sadh

extern "C" const char* sadh_last_error(void) { return sadh::g_last_error.c_str(); }

extern "C" const char* sadh_version(void) { return "sad_hip 0.1 (gfx950)"; }
