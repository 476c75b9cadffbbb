// Library-level C ABI: version, error strings, device selection.
#include <cstring>

#include "aq_internal.hpp"

namespace aq {

std::string& last_error_ref()
{
    static thread_local std::string msg;
    return msg;
}

int set_error(int code, const std::string& msg)
{
    last_error_ref() = msg;
    return code;
}

}  // namespace aq

using namespace aq;

extern "C" {

AQ_API const char* aq_version(void) { return "aquagpusph_amd 0.1.0 (gfx950)"; }

AQ_API const char* aq_strerror(int code)
{
    switch (code) {
    case AQ_OK:
        return "success";
    case AQ_EINVAL:
        return "invalid argument";
    case AQ_EHIP:
        return "HIP runtime error";
    case AQ_ENOMEM:
        return "out of memory or capacity exceeded";
    case AQ_EPARSE:
        return "parse error";
    case AQ_ENOTFOUND:
        return "not found";
    case AQ_EASSERT:
        return "assertion failed";
    case AQ_EUNSUP:
        return "unsupported";
    default:
        return "unknown error";
    }
}

AQ_API const char* aq_last_error(void) { return last_error_ref().c_str(); }

AQ_API int aq_set_device(int device)
{
    return guard([&] { AQ_HIP(hipSetDevice(device)); });
}

}  // extern "C"
