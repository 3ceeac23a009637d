// Allocation-free variant of phip_permanent_residues for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// It runs the same persistent Glynn engine as the plain entry point, but
// the plan's device image (tables, prime sets, jobs, work items, per-item
// partial sums, class counters, residues) lives in `scratch`, a file-scope
// device array of kScratchBytes that every GPU the library runs on holds
// until the process ends.  Host code never uses the array's name as an
// address: at the first call on each device, with that device current, the
// library asks the runtime for that device's copy (hipGetSymbolAddress) and
// keeps the pointer per device.  Each call builds its plan on the host,
// copies it into the scratch, runs the kernels and copies the residues
// straight into the caller's array: no hipMalloc, hipFree, pinned staging
// or allocating library routine.  The class counters are cleared by the
// plan's own reset kernel and every other scratch word the kernels read is
// written by the call first, so what an earlier call left behind is never
// read.  All calls on one GPU share its scratch: the library serialises
// them with the device's mutex, and two processes (or code that bypasses
// the library) writing the same device's scratch at once would race.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "batch.h"
#include "device_context.h"
#include "launchers.h"
#include "permhip.h"
#include "residues.h"

// 8 MiB: the default plan of one matrix with up to 64 moduli needs about
// 1.6 MiB (2^14 work items of 96 bytes of partials, 6.4 KB tables per
// distinct table).
constexpr size_t kScratchBytes = size_t(8) << 20;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxModuli = 64;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`, looked up once with the device current.
uint8_t* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) {
        phip::set_error("device index out of range for the allocation-free scratch");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            phip::set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
    }
    return g_scratch_ptr[device];
}

size_t carve(size_t& offset, size_t bytes) {
    offset = (offset + 255) & ~size_t(255);
    const size_t at = offset;
    offset += std::max<size_t>(bytes, 1);
    return at;
}

}  // namespace

extern "C" int phip_permanent_residues_noalloc(const int64_t* A, int64_t prime_stride, int n,
                                               const int64_t* moduli, int k, int64_t* out) {
    using namespace phip;
    if (n < kGlynnMinOrder || n > kGlynnMaxOrder) {
        set_error("the allocation-free entry point takes orders 5..46 (use phip_permanent_residues)");
        return -1;
    }
    if (k < 1 || k > kMaxModuli) {
        set_error("the allocation-free entry point takes 1..64 moduli per call");
        return -1;
    }
    if (!A || !moduli || !out) {
        set_error("null argument");
        return -1;
    }
    if (prime_stride != 0 && prime_stride < (int64_t)n * n) {
        set_error("prime_stride must be 0 or at least n * n");
        return -1;
    }
    for (int i = 0; i < k; ++i) {
        if (moduli[i] < 3 || moduli[i] >= (int64_t(1) << 31) || !(moduli[i] & 1)) {
            set_error("the allocation-free entry point takes odd moduli in [3, 2^31) (use phip_permanent_residues)");
            return -1;
        }
    }
    DeviceContext* ctx = DeviceContext::current();
    if (!ctx) return -1;
    std::lock_guard<std::mutex> lock(ctx->mutex());
    PHIP_HIP_TRY(hipSetDevice(ctx->device()));
    uint8_t* base = scratch_on(ctx->device());
    if (!base) return -1;

    std::vector<PlanInput> in;
    in.reserve(k);
    for (int i = 0; i < k; ++i) in.push_back({A + (int64_t)i * prime_stride, n, (uint32_t)moduli[i], 0});
    const Tuning& t = tuning();
    PlanHost h;
    if (build_plan_host(in, std::max(0, t.lanes_log2 - 6), t.min_cbits, h) != 0) return -1;
    if (h.slots.size() != (size_t)k) {
        set_error("internal: plan slots differ from the moduli");
        return -1;
    }

    size_t off = 0;
    const size_t o_jobs = carve(off, h.jobs.size() * sizeof(BatchJob));
    const size_t o_items = carve(off, h.items.size() * sizeof(uint2));
    const size_t o_blob = carve(off, h.blob.size() * sizeof(int32_t));
    const size_t o_p = carve(off, h.set_p.size() * sizeof(int32_t));
    const size_t o_pinv = carve(off, h.set_pinv.size() * sizeof(uint32_t));
    const size_t o_slots = carve(off, h.slots.size() * sizeof(BatchSlot));
    const size_t o_part = carve(off, h.items.size() * kItemSlots * sizeof(uint64_t));
    const size_t o_cnt = carve(off, 32 * sizeof(uint32_t));
    const size_t o_res = carve(off, h.slots.size() * sizeof(int64_t));
    if (off > kScratchBytes) {
        set_error("this plan needs " + std::to_string(off) + " bytes of device scratch; the allocation-free entry " +
                  "point holds " + std::to_string(kScratchBytes) + " (use phip_permanent_residues)");
        return -1;
    }
    if (h.classes.size() > 32) {
        set_error("internal: more plan classes than counters");
        return -1;
    }
    PlanDevice d;
    d.jobs = reinterpret_cast<BatchJob*>(base + o_jobs);
    d.items = reinterpret_cast<uint2*>(base + o_items);
    d.blob = reinterpret_cast<int32_t*>(base + o_blob);
    d.set_p = reinterpret_cast<int32_t*>(base + o_p);
    d.set_pinv = reinterpret_cast<uint32_t*>(base + o_pinv);
    d.slots = reinterpret_cast<BatchSlot*>(base + o_slots);
    d.partials = reinterpret_cast<uint64_t*>(base + o_part);
    d.counters = reinterpret_cast<uint32_t*>(base + o_cnt);
    d.residues = reinterpret_cast<int64_t*>(base + o_res);

    hipStream_t st = ctx->stream();
    PHIP_HIP_TRY(hipMemcpyAsync(d.jobs, h.jobs.data(), h.jobs.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d.items, h.items.data(), h.items.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d.blob, h.blob.data(), h.blob.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(
        hipMemcpyAsync(d.set_p, h.set_p.data(), h.set_p.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(hipMemcpyAsync(d.set_pinv, h.set_pinv.data(), h.set_pinv.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, st));
    PHIP_HIP_TRY(
        hipMemcpyAsync(d.slots, h.slots.data(), h.slots.size() * sizeof(BatchSlot), hipMemcpyHostToDevice, st));
    const int cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    if (enqueue_plan(h, d, st, cus, t.batch_blocks_per_cu, nullptr, nullptr) != 0) return -1;
    PHIP_HIP_TRY(hipMemcpyAsync(out, d.residues, (size_t)k * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    PHIP_HIP_TRY(hipStreamSynchronize(st));
    return 0;
}
