// Library-level state of the C ABI: version, per-thread error text, device
// probe.
#include <hip/hip_runtime.h>

#include <string>

#include "oslmi.h"
#include "oslmi_group.h"
#include "oslmi_internal.h"

namespace oslmi {

static thread_local std::string t_last_error;

int set_error(int code, const std::string& msg)
{
    t_last_error = msg;
    return code;
}

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(OSLMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return OSLMI_OK;
}

int check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        return set_error(OSLMI_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return OSLMI_OK;
}

}  // namespace oslmi

using namespace oslmi;

extern "C" const char* oslmi_version(void) { return "oslmi 0.1 (gfx950)"; }

extern "C" const char* oslmi_last_error(void) { return t_last_error.c_str(); }

extern "C" int oslmi_init(void)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1)
        return set_error(OSLMI_ERR_HIP, "oslmi_init: no HIP device");
    int dev = 0;
    if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))
        return rc;
    hipDeviceProp_t prop;
    if (int rc = check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties"))
        return rc;
    std::string arch(prop.gcnArchName);
    if (arch.rfind("gfx950", 0) != 0)
        return set_error(OSLMI_ERR_UNSUPPORTED, "oslmi_init: device is " + arch + ", library is built for gfx950");
    // per-device set-up: the colour system's tables in device memory and
    // the address of the no-alloc workspace on this device
    const ColorSysDev* cs = nullptr;
    if (int rc = device_colorsys(&cs))
        return rc;
    return noalloc_scratch_init();
}
