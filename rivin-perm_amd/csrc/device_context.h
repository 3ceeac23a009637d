// Per-device state of the library: one stream at default priority, grow-only
// device buffers, and a mutex that serialises the plain entry points on that
// device.  Error messages are kept per calling thread.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <string>

namespace phip {

void set_error(const std::string& msg);
const std::string& last_error();

#define PHIP_HIP_TRY(expr)                                                                    \
    do {                                                                                      \
        hipError_t phip_err_ = (expr);                                                        \
        if (phip_err_ != hipSuccess) {                                                        \
            ::phip::set_error(std::string(#expr) + " failed: " + hipGetErrorString(phip_err_)); \
            return -1;                                                                        \
        }                                                                                     \
    } while (0)

class DeviceContext {
   public:
    static constexpr int kSlots = 56;

    // Context of the calling thread's current device, created on first use.
    // Returns nullptr (with the error set) when no device is usable.
    static DeviceContext* current();

    std::mutex& mutex() { return mu_; }
    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }
    int compute_units() const { return cus_; }

    // Device buffer for a slot, at least `bytes` long (contents undefined).
    void* buffer(int slot, size_t bytes);
    // Pinned host staging buffer for a slot.
    void* host_buffer(int slot, size_t bytes);

    // Optional event timing of the Glynn kernel launches.
    bool profiling = false;
    double profiled_ms = 0;
    uint64_t profiled_launches = 0;
    hipEvent_t prof_start = nullptr, prof_stop = nullptr;

   private:
    explicit DeviceContext(int device);
    int device_ = -1;
    int cus_ = 0;
    hipStream_t stream_ = nullptr;
    std::mutex mu_;
    void* dev_[kSlots] = {};
    size_t dev_cap_[kSlots] = {};
    void* host_[kSlots] = {};
    size_t host_cap_[kSlots] = {};
};

}  // namespace phip
