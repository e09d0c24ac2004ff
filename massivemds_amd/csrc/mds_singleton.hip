// Instance registry of the BEAST bridge (include/mds_hip_singleton.h): the
// reference JNI singleton's global vector of engines
// (src/jni/dr_inference_multidimensionalscaling_NativeMDSSingleton.cpp:9-21)
// over this library's C ABI, with status returns instead of exceptions
// crossing the foreign-function boundary.  Host code only.
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mds_hip.h"
#include "mds_hip_singleton.h"

namespace {

// One instance.  Its mutex is held for the whole of every call on it and by
// release, so a release on another thread waits for a call in flight and a
// later call finds the handle gone; the error message is written and read
// under it.  Slots are shared, so a caller that found one keeps it alive.
struct Slot {
    std::mutex m;
    mds_hip_handle h = nullptr;
    std::string error = "no error";
};

std::mutex g_mutex;  // the slot list and the registry's own error
std::vector<std::shared_ptr<Slot>> g_slots;
std::string g_error = "no error";

void set_registry_error(const std::string& message) {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_error = message;
}

// The slot of an instance number, or null (with the registry's error set).
std::shared_ptr<Slot> slot_of(int instance) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (instance < 0 || (size_t)instance >= g_slots.size() || !g_slots[instance]) {
        g_error = "no instance " + std::to_string(instance);
        return nullptr;
    }
    return g_slots[instance];
}

template <typename F>
int call(int instance, F&& f) {
    const std::shared_ptr<Slot> s = slot_of(instance);
    if (!s) return MDS_HIP_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->m);
    if (!s->h) {
        s->error = "instance " + std::to_string(instance) + " was released";
        set_registry_error(s->error);
        return MDS_HIP_ERR_ARGUMENT;
    }
    const int st = f(s->h);
    if (st != MDS_HIP_OK) s->error = mds_hip_last_error(s->h);
    return st;
}

// The returned message is this thread's copy, valid until its next call of
// mds_hip_singleton_last_error.
thread_local std::string t_message;

}  // namespace

extern "C" {

int mds_hip_singleton_initialize(int embeddingDimension, int elementCount, long flags, int device, int threads,
                                 int* instance) {
    (void)threads;  // the reference's CPU thread count; the GPU engine has none
    if (!instance) {
        std::lock_guard<std::mutex> lock(g_mutex);
        g_error = "null instance pointer";
        return MDS_HIP_ERR_ARGUMENT;
    }
    if (!(flags & (MDS_HIP_FLAG_OPENCL | MDS_HIP_FLAG_HIP))) {
        std::lock_guard<std::mutex> lock(g_mutex);
        g_error = "libmds_hip provides the GPU engine only: set the OPENCL (or HIP) flag";
        return MDS_HIP_ERR_ARGUMENT;
    }
    mds_hip_handle h = nullptr;
    const int st = mds_hip_create(embeddingDimension, elementCount, flags, device, &h);
    std::lock_guard<std::mutex> lock(g_mutex);
    if (st != MDS_HIP_OK) {
        g_error = mds_hip_last_error(nullptr);
        return st;
    }
    auto s = std::make_shared<Slot>();
    s->h = h;
    g_slots.push_back(std::move(s));
    *instance = (int)g_slots.size() - 1;
    return MDS_HIP_OK;
}

int mds_hip_singleton_update_locations(int instance, int index, const double* x, size_t length) {
    return call(instance, [&](mds_hip_handle h) { return mds_hip_update_locations(h, index, x, length); });
}

int mds_hip_singleton_get_sum_of_increments(int instance, double* sum) {
    return call(instance, [&](mds_hip_handle h) { return mds_hip_get_sum_of_increments(h, sum); });
}

int mds_hip_singleton_store_state(int instance) {
    return call(instance, [](mds_hip_handle h) { return mds_hip_store_state(h); });
}

int mds_hip_singleton_restore_state(int instance) {
    return call(instance, [](mds_hip_handle h) { return mds_hip_restore_state(h); });
}

int mds_hip_singleton_accept_state(int instance) {
    return call(instance, [](mds_hip_handle h) { return mds_hip_accept_state(h); });
}

int mds_hip_singleton_make_dirty(int instance) {
    return call(instance, [](mds_hip_handle h) { return mds_hip_make_dirty(h); });
}

int mds_hip_singleton_set_pairwise_data(int instance, const double* data, size_t length) {
    return call(instance, [&](mds_hip_handle h) { return mds_hip_set_pairwise_data(h, data, length); });
}

int mds_hip_singleton_get_location_gradient(int instance, double* gradient, size_t length) {
    return call(instance, [&](mds_hip_handle h) { return mds_hip_get_log_likelihood_gradient(h, gradient, length); });
}

int mds_hip_singleton_set_parameters(int instance, const double* parameters, size_t length) {
    return call(instance, [&](mds_hip_handle h) { return mds_hip_set_parameters(h, parameters, length); });
}

int mds_hip_singleton_get_internal_dimension(int instance, int* dimension) {
    return call(instance, [&](mds_hip_handle h) { return mds_hip_get_internal_dimension(h, dimension); });
}

int mds_hip_singleton_instance_count(int* count) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!count) {
        g_error = "null count pointer";
        return MDS_HIP_ERR_ARGUMENT;
    }
    *count = (int)g_slots.size();
    return MDS_HIP_OK;
}

int mds_hip_singleton_release_all(void) {
    std::vector<std::shared_ptr<Slot>> slots;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        slots = g_slots;
    }
    for (auto& s : slots) {
        if (!s) continue;
        std::lock_guard<std::mutex> lock(s->m);  // waits for a call in flight on this instance
        if (s->h) {
            mds_hip_destroy(s->h);
            s->h = nullptr;
        }
    }
    return MDS_HIP_OK;
}

const char* mds_hip_singleton_last_error(int instance) {
    std::shared_ptr<Slot> s;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (instance < 0 || (size_t)instance >= g_slots.size() || !g_slots[instance]) {
            t_message = g_error;
            return t_message.c_str();
        }
        s = g_slots[instance];
    }
    std::lock_guard<std::mutex> lock(s->m);
    t_message = s->error;
    return t_message.c_str();
}

}  // extern "C"
