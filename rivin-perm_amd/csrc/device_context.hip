#include "device_context.h"

#include <memory>
#include <vector>

namespace phip {

namespace {
thread_local std::string g_last_error;
std::mutex g_registry_mu;
std::vector<std::unique_ptr<DeviceContext>>* g_registry = nullptr;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

DeviceContext::DeviceContext(int device) : device_(device) {}

DeviceContext* DeviceContext::current() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        set_error(std::string("no HIP device available: ") + hipGetErrorString(e));
        return nullptr;
    }
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("hipGetDevice failed: ") + hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_registry_mu);
    if (!g_registry) g_registry = new std::vector<std::unique_ptr<DeviceContext>>(count);
    if (dev < 0 || dev >= (int)g_registry->size()) {
        set_error("current device index out of range");
        return nullptr;
    }
    auto& slot = (*g_registry)[dev];
    if (!slot) {
        std::unique_ptr<DeviceContext> ctx(new DeviceContext(dev));
        e = hipStreamCreateWithFlags(&ctx->stream_, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamCreateWithFlags failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ctx->cus_ = prop.multiProcessorCount;
        slot = std::move(ctx);
    }
    return slot.get();
}

void* DeviceContext::buffer(int slot, size_t bytes) {
    if (slot < 0 || slot >= kSlots) {
        set_error("internal: device buffer slot out of range");
        return nullptr;
    }
    if (bytes == 0) bytes = 1;
    if (dev_cap_[slot] < bytes) {
        if (dev_[slot]) {
            hipStreamSynchronize(stream_);
            hipFree(dev_[slot]);
            dev_[slot] = nullptr;
            dev_cap_[slot] = 0;
        }
        size_t cap = 256;
        while (cap < bytes) cap <<= 1;
        hipError_t e = hipMalloc(&dev_[slot], cap);
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
            dev_[slot] = nullptr;
            return nullptr;
        }
        dev_cap_[slot] = cap;
    }
    return dev_[slot];
}

void* DeviceContext::host_buffer(int slot, size_t bytes) {
    if (slot < 0 || slot >= kSlots) {
        set_error("internal: host buffer slot out of range");
        return nullptr;
    }
    if (bytes == 0) bytes = 1;
    if (host_cap_[slot] < bytes) {
        if (host_[slot]) {
            hipStreamSynchronize(stream_);
            hipHostFree(host_[slot]);
            host_[slot] = nullptr;
            host_cap_[slot] = 0;
        }
        size_t cap = 256;
        while (cap < bytes) cap <<= 1;
        hipError_t e = hipHostMalloc(&host_[slot], cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            set_error(std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
            host_[slot] = nullptr;
            return nullptr;
        }
        host_cap_[slot] = cap;
    }
    return host_[slot];
}

}  // namespace phip
