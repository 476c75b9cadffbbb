// Internal helpers shared by the HIP kernels and the C ABI launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/aquagpusph_amd.h"

namespace aq {

typedef uint32_t usize;

// ---- error reporting across the C ABI ------------------------------------
int set_error(int code, const std::string& msg);
std::string& last_error_ref();

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

#define AQ_HIP(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            throw ::aq::Error(AQ_EHIP, std::string(#call " failed: ") +        \
                                           hipGetErrorString(e_));             \
    } while (0)

#define AQ_REQUIRE(cond, msg)                                                  \
    do {                                                                       \
        if (!(cond)) throw ::aq::Error(AQ_EINVAL, (msg));                      \
    } while (0)

// Wrap a C ABI body: translate exceptions into error codes.
template <class F> inline int guard(F f, bool check_launch = true)
{
    try {
        f();
        if (!check_launch) return AQ_OK;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return set_error(AQ_EHIP, std::string("kernel launch failed: ") +
                                          hipGetErrorString(e));
        return AQ_OK;
    } catch (const Error& e) {
        return set_error(e.code(), e.what());
    } catch (const std::bad_alloc&) {
        return set_error(AQ_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return set_error(AQ_EINVAL, e.what());
    }
}

inline hipStream_t S(void* s) { return (hipStream_t)s; }

inline unsigned blocks_for(size_t n, unsigned bs)
{
    return (unsigned)((n + bs - 1) / bs);
}

inline int check_dims(int dims)
{
    if (dims != 2 && dims != 3)
        throw Error(AQ_EINVAL, "dims must be 2 or 3");
    return dims;
}

// Largest float s with fl(fl(sqrt(s)) / H) < 2 (the reference acceptance
// test q < SUPPORT with q = length(r_ij)/H).  sqrt and division are
// correctly rounded and monotone, so q(s) is non-decreasing in s and the
// test "q < 2" is exactly "s <= s_max" for every non-NaN s.
float support_threshold(float H);

}  // namespace aq
