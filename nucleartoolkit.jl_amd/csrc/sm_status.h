// Exception -> C status mapping shared by every C ABI translation unit.
#pragma once
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/smgpu.h"
#include "sm_trl.h"

namespace smgpu {

void set_last_error(const std::string& msg);   // thread-local text for smgpu_last_error

// Status of the exception being handled (call only inside a catch block).
inline int32_t status_of_current_exception(std::string& msg) {
    try {
        throw;
    } catch (const Canceled& e) {
        msg = e.what();
        return SMGPU_ECANCELED;
    } catch (const std::invalid_argument& e) {
        msg = e.what();
        return SMGPU_EINVAL;
    } catch (const std::bad_alloc&) {
        msg = "out of host memory";
        return SMGPU_ENOMEM;
    } catch (const std::runtime_error& e) {
        msg = e.what();
        if (msg.rfind("HIP error out of memory", 0) == 0 || msg.find("hipErrorOutOfMemory") != std::string::npos)
            return SMGPU_ENOMEM;
        if (msg.rfind("HIP error", 0) == 0) return SMGPU_EDEVICE;
        if (msg.rfind("cannot open", 0) == 0) return SMGPU_EIO;
        if (msg.rfind("TRL: iteration limit", 0) == 0) return SMGPU_ENOTCONV;
        if (msg.rfind("does not fit", 0) == 0) return SMGPU_ENOMEM;
        return SMGPU_EINVAL;
    } catch (...) {
        msg = "unknown error";
        return SMGPU_EINVAL;
    }
}

template <typename F> int32_t guard(F&& f) {
    try {
        return f();
    } catch (...) {
        std::string m;
        int32_t c = status_of_current_exception(m);
        set_last_error(m);
        return c;
    }
}

}  // namespace smgpu
