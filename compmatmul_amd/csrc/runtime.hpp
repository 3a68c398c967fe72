// Services of the C-ABI runtime (api.cpp) shared with the batch engine
// (batch.cpp): the library lock, device/stream, error string, cached
// twiddle tables and the per-kernel timing wrapper.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <mutex>
#include <string>

namespace cm {
namespace rt {

std::recursive_mutex& mutex();
int init();                                   // device 0 and the library stream
hipStream_t stream();
int error(int code, const std::string& msg);  // sets cm_last_error(), returns code
int twiddles(int lb, int f64, const void** out);
int fft_t2_table(const void** out);
// runs `launch`, bracketed by events when timing is on (not while capturing)
int timed(const char* name, const std::function<hipError_t()>& launch);
// the same, with the events on `stream`, the stream the launch runs on
int timed_on(hipStream_t stream, const char* name, const std::function<hipError_t()>& launch);

}  // namespace rt
}  // namespace cm
