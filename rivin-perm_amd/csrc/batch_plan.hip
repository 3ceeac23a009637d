// Host side of the persistent batch: plan construction (device-resident
// tables, cost-sorted work items), graph capture and replay, and the C ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "batch.h"
#include "device_context.h"
#include "permhip.h"
#include "phip_arith.h"
#include "phip_reduce.h"
#include "residues.h"
#include "tables.h"

namespace phip {

namespace {

__global__ void batch_reset_kernel(uint32_t* counters, int n) {
    const int i = threadIdx.x;
    if (i < n) counters[i] = 0;
}

// One workgroup per job: fold the job's contiguous item sums, reduce, scale.
__global__ __launch_bounds__(256) void batch_finalize_kernel(const BatchJob* __restrict__ jobs,
                                                             const uint64_t* __restrict__ item_partials,
                                                             int64_t* __restrict__ residues) {
    const BatchJob jb = jobs[blockIdx.x];
    const uint64_t p = (uint64_t)jb.p;
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < jb.n_items; i += 256) s += item_partials[jb.first_item + i] % p;
    s = block_reduce_sum_u64<256>(s);
    if (threadIdx.x == 0) residues[blockIdx.x] = (int64_t)(((s % p) * (uint64_t)jb.scale) % p);
}

BatchLaunchFn launcher_for(int group) {
    switch (group) {
        case 0: return launch_batch_g0;
        case 2: return launch_batch_g2;
        case 3: return launch_batch_g3;
        case 4: return launch_batch_g4;
        case 6: return launch_batch_g6;
        default: return nullptr;
    }
}

struct Class {
    int group;
    uint32_t first_item;
    uint32_t n_items;
};

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    ~DevBuf() {
        if (ptr) hipFree(ptr);
    }
    int alloc(size_t count) {
        hipError_t e = hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
            ptr = nullptr;
            return -1;
        }
        return 0;
    }
};

}  // namespace

struct BatchPlan {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::vector<BatchJob> jobs;  // in input (matrix, modulus) order
    std::vector<Class> classes;
    uint32_t n_items = 0;
    double total_states = 0, total_mulmods = 0;
    DevBuf<BatchJob> d_jobs;
    DevBuf<uint2> d_items;
    DevBuf<int32_t> d_blob;
    DevBuf<uint64_t> d_partials;
    DevBuf<uint32_t> d_counters;
    DevBuf<int64_t> d_res;
    int64_t* h_res = nullptr;
    int blocks_per_cu = 2;
    int cus = 256;
    bool use_graph = true;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0;

    ~BatchPlan() {
        if (stream) hipStreamSynchronize(stream);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (h_res) hipHostFree(h_res);
        if (stream) hipStreamDestroy(stream);
    }

    int enqueue(bool timed) {
        hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream, d_counters.ptr, (int)classes.size());
        PHIP_HIP_TRY(hipGetLastError());
        if (timed) PHIP_HIP_TRY(hipEventRecord(ev0, stream));
        const dim3 grid((unsigned)(cus * blocks_per_cu));
        for (size_t c = 0; c < classes.size(); ++c) {
            const Class& cl = classes[c];
            PHIP_HIP_TRY(launcher_for(cl.group)(grid, stream, d_jobs.ptr, d_items.ptr + cl.first_item, cl.n_items,
                                                d_counters.ptr + c, d_blob.ptr, d_partials.ptr + cl.first_item));
        }
        if (timed) PHIP_HIP_TRY(hipEventRecord(ev1, stream));
        hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)jobs.size()), dim3(256), 0, stream, d_jobs.ptr,
                           d_partials.ptr, d_res.ptr);
        PHIP_HIP_TRY(hipGetLastError());
        PHIP_HIP_TRY(hipMemcpyAsync(h_res, d_res.ptr, jobs.size() * sizeof(int64_t), hipMemcpyDeviceToHost, stream));
        return 0;
    }

    bool launched = false;

    // Enqueue one run on the plan's stream without waiting.
    int launch_locked() {
        PHIP_HIP_TRY(hipSetDevice(device));
        if (use_graph) {
            if (!exec) {
                PHIP_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                if (enqueue(false) != 0) {
                    hipGraph_t g = nullptr;
                    hipStreamEndCapture(stream, &g);
                    if (g) hipGraphDestroy(g);
                    return -1;
                }
                PHIP_HIP_TRY(hipStreamEndCapture(stream, &graph));
                PHIP_HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            }
            PHIP_HIP_TRY(hipEventRecord(ev0, stream));
            PHIP_HIP_TRY(hipGraphLaunch(exec, stream));
            PHIP_HIP_TRY(hipEventRecord(ev1, stream));
        } else {
            if (enqueue(true) != 0) return -1;
        }
        launched = true;
        return 0;
    }

    // 1: finished (residues copied to out when not null), 0: still running.
    int finish_locked(int64_t* out, bool block) {
        if (!launched) {
            set_error("batch plan has not been launched");
            return -1;
        }
        if (block) {
            PHIP_HIP_TRY(hipStreamSynchronize(stream));
        } else {
            const hipError_t e = hipStreamQuery(stream);
            if (e == hipErrorNotReady) return 0;
            PHIP_HIP_TRY(e);
        }
        launched = false;
        PHIP_HIP_TRY(hipEventElapsedTime(&last_ms, ev0, ev1));
        if (out) memcpy(out, h_res, jobs.size() * sizeof(int64_t));
        return 1;
    }

    int run(int64_t* out) {
        std::lock_guard<std::mutex> lock(mu);
        if (launched && finish_locked(nullptr, true) < 0) return -1;
        if (launch_locked() != 0) return -1;
        return finish_locked(out, true) < 0 ? -1 : 0;
    }
};

namespace {

int build_plan(BatchPlan& plan, const int64_t* entries, const int32_t* orders, int count, const int64_t* moduli,
               const int32_t* moduli_counts) {
    if (count < 0 || (count > 0 && (!entries || !orders || !moduli || !moduli_counts))) {
        set_error("invalid batch description");
        return -1;
    }
    // Pass 1: validate, pick the row-sum representation of every job.
    struct Pending {
        int matrix;
        uint64_t entry_off;
        int n;
        uint32_t q;
        int group;
        std::vector<int> order;
    };
    std::vector<Pending> pend;
    uint64_t entry_off = 0;
    size_t mod_off = 0;
    for (int m = 0; m < count; ++m) {
        const int n = orders[m];
        if (n < kBatchMinOrder || n > kBatchMaxOrder) {
            set_error("batch orders must lie in [5, 40]; use the single-matrix entry point otherwise");
            return -1;
        }
        if (moduli_counts[m] < 0) {
            set_error("negative modulus count");
            return -1;
        }
        for (int k = 0; k < moduli_counts[m]; ++k) {
            const int64_t q = moduli[mod_off + k];
            if (q < 3 || q >= (INT64_C(1) << 31) || !(q & 1)) {
                set_error("batch moduli must be odd and lie in [3, 2^31)");
                return -1;
            }
            std::vector<int> order;
            const int group = plan_rows(centred_row_bounds(entries + entry_off, n, q), order);
            pend.push_back({m, entry_off, n, (uint32_t)q, group, std::move(order)});
        }
        entry_off += (uint64_t)n * n;
        mod_off += moduli_counts[m];
    }
    // Item size: total states spread over the target item count.
    double total_states = 0;
    for (const auto& p : pend) total_states += std::ldexp(1.0, p.n - 1);
    const double target = std::ldexp(1.0, tuning().batch_items_log2);
    int s = (int)std::lround(std::log2(std::max(1.0, total_states / target)));
    s = std::min(std::max(s, 10), 30);

    // Pass 2: tables, jobs.
    std::vector<int32_t> blob;
    plan.jobs.resize(pend.size());
    plan.total_states = 0;
    plan.total_mulmods = 0;
    for (size_t j = 0; j < pend.size(); ++j) {
        const Pending& pd = pend[j];
        const int n = pd.n;
        const int L = n - 1;
        int c = (L >= 10) ? std::max(4, std::min(s - 6, L - 6)) : 4;
        c = std::min(c, L);
        BatchJob jb;
        jb.data_off = blob.size();
        jb.n = n;
        jb.cbits = c;
        jb.lanes = 1ull << (L - c);
        jb.p = (int32_t)pd.q;
        jb.pinv = inverse_mod_2_32(pd.q);
        jb.group = pd.group;
        const int groups = pd.group == 0 ? n : (n + pd.group - 1) / pd.group;
        const uint64_t q = pd.q;
        jb.scale = (uint32_t)host_mulmod(host_powmod((1ull << 32) % q, (uint64_t)(groups - 1), q),
                                         host_powmod((q + 1) / 2, (uint64_t)L, q), q);
        const size_t words = (size_t)n * glynn_stride(n, pd.group);
        blob.resize(blob.size() + words);
        build_glynn_tables(entries + pd.entry_off, n, pd.q, pd.group, pd.order, blob.data() + jb.data_off);
        plan.jobs[j] = jb;
        plan.total_states += std::ldexp(1.0, L);
        plan.total_mulmods += (double)n * std::ldexp(1.0, L);
    }
    // Items: per class, jobs by descending cost, each job's items contiguous.
    std::vector<uint2> items;
    plan.classes.clear();
    for (int group : {0, 2, 3, 4, 6}) {
        std::vector<uint32_t> order;
        for (uint32_t j = 0; j < plan.jobs.size(); ++j)
            if (plan.jobs[j].group == group) order.push_back(j);
        if (order.empty()) continue;
        auto cost = [&](uint32_t j) {
            const BatchJob& b = plan.jobs[j];
            const int groups = b.group == 0 ? b.n : (b.n + b.group - 1) / b.group;
            return std::ldexp(1.0, b.cbits) * (double)std::min<uint64_t>(b.lanes, 64) * (b.n + 3.0 * groups);
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
        Class cl{group, (uint32_t)items.size(), 0};
        for (uint32_t j : order) {
            BatchJob& b = plan.jobs[j];
            const uint64_t waves = (b.lanes + 63) / 64;
            b.first_item = (uint32_t)items.size();
            b.n_items = (uint32_t)waves;
            for (uint64_t w = 0; w < waves; ++w) items.push_back(make_uint2(j, (uint32_t)w));
        }
        cl.n_items = (uint32_t)items.size() - cl.first_item;
        plan.classes.push_back(cl);
    }
    if (items.size() >= 0xffffffffull) {
        set_error("batch has too many work items");
        return -1;
    }
    plan.n_items = (uint32_t)items.size();

    if (plan.d_jobs.alloc(plan.jobs.size()) || plan.d_items.alloc(items.size()) || plan.d_blob.alloc(blob.size()) ||
        plan.d_partials.alloc(items.size()) || plan.d_counters.alloc(8) || plan.d_res.alloc(plan.jobs.size()))
        return -1;
    PHIP_HIP_TRY(hipHostMalloc((void**)&plan.h_res, std::max<size_t>(plan.jobs.size(), 1) * sizeof(int64_t),
                               hipHostMallocDefault));
    PHIP_HIP_TRY(hipMemcpy(plan.d_jobs.ptr, plan.jobs.data(), plan.jobs.size() * sizeof(BatchJob),
                           hipMemcpyHostToDevice));
    PHIP_HIP_TRY(hipMemcpy(plan.d_items.ptr, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
    PHIP_HIP_TRY(hipMemcpy(plan.d_blob.ptr, blob.data(), blob.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

}  // namespace phip

using phip::BatchPlan;

extern "C" {

int phip_batch_create(const int64_t* entries, const int32_t* orders, int count, const int64_t* moduli,
                      const int32_t* moduli_counts, void** out) {
    if (!out) {
        phip::set_error("output handle pointer is null");
        return -1;
    }
    *out = nullptr;
    phip::DeviceContext* ctx = phip::DeviceContext::current();
    if (!ctx) return -1;
    std::unique_ptr<BatchPlan> plan(new BatchPlan());
    plan->device = ctx->device();
    plan->cus = ctx->compute_units() > 0 ? ctx->compute_units() : 256;
    plan->blocks_per_cu = phip::tuning().batch_blocks_per_cu;
    PHIP_HIP_TRY(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking));
    PHIP_HIP_TRY(hipEventCreate(&plan->ev0));
    PHIP_HIP_TRY(hipEventCreate(&plan->ev1));
    if (phip::build_plan(*plan, entries, orders, count, moduli, moduli_counts) != 0) return -1;
    *out = plan.release();
    return 0;
}

int phip_batch_run(void* handle, int64_t* residues_out) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    if (plan->jobs.empty()) return 0;
    return plan->run(residues_out);
}

int phip_batch_launch(void* handle) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    if (plan->jobs.empty()) return 0;
    std::lock_guard<std::mutex> lock(plan->mu);
    if (plan->launched) {
        phip::set_error("batch plan is already running");
        return -1;
    }
    return plan->launch_locked();
}

int phip_batch_query(void* handle, int64_t* residues_out, int block) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    if (plan->jobs.empty()) return 1;
    std::lock_guard<std::mutex> lock(plan->mu);
    return plan->finish_locked(residues_out, block != 0);
}

int phip_batch_set_option(void* handle, int option, int64_t value) {
    if (!handle) {
        phip::set_error("null batch handle");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    std::lock_guard<std::mutex> lock(plan->mu);
    if (plan->launched) {
        phip::set_error("cannot change options while the plan is running");
        return -1;
    }
    switch (option) {
        case PHIP_BATCH_BLOCKS_PER_CU:
            if (value < 1 || value > 64) {
                phip::set_error("blocks per CU must lie in [1, 64]");
                return -1;
            }
            plan->blocks_per_cu = (int)value;
            break;
        case PHIP_BATCH_USE_GRAPH:
            plan->use_graph = value != 0;
            break;
        default:
            phip::set_error("unknown batch option");
            return -1;
    }
    if (plan->exec) {
        hipGraphExecDestroy(plan->exec);
        plan->exec = nullptr;
    }
    if (plan->graph) {
        hipGraphDestroy(plan->graph);
        plan->graph = nullptr;
    }
    return 0;
}

int phip_batch_stats(void* handle, double* out, int len) {
    if (!handle || !out) {
        phip::set_error("null argument");
        return -1;
    }
    auto* plan = static_cast<BatchPlan*>(handle);
    const double v[] = {(double)plan->last_ms,
                        plan->total_states,
                        plan->total_mulmods,
                        (double)plan->jobs.size(),
                        (double)plan->n_items,
                        (double)plan->classes.size(),
                        (double)(plan->cus * plan->blocks_per_cu * 4),
                        plan->classes.empty() ? -1.0 : (double)plan->classes[0].group};
    const int m = std::min<int>(len, (int)(sizeof(v) / sizeof(v[0])));
    for (int i = 0; i < m; ++i) out[i] = v[i];
    return m;
}

int phip_batch_destroy(void* handle) {
    delete static_cast<BatchPlan*>(handle);
    return 0;
}

}  // extern "C"
