// The BEAST bridge's instance registry (massivemds_amd/csrc/mds_singleton.hip,
// host code) over a stub engine, under ThreadSanitizer: several threads call
// their own instances while another thread releases every instance.  The
// stub marks a handle destroyed and fails the run when a call reaches a
// destroyed handle or overlaps another call on the same handle; after the
// release every call must return MDS_HIP_ERR_ARGUMENT.  Exit status 0 =
// clean.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "mds_hip.h"
#include "mds_hip_singleton.h"

struct mds_hip_engine {
    std::atomic<int> busy{0};
    std::atomic<bool> destroyed{false};
    double sum = 0.0;
};

static std::atomic<int> g_violations{0};

static int use(mds_hip_handle h) {
    if (h->destroyed.load()) {
        ++g_violations;
        return MDS_HIP_ERR_DEVICE;
    }
    if (h->busy.fetch_add(1) != 0) ++g_violations;  // two calls at once on one instance
    std::this_thread::sleep_for(std::chrono::microseconds(50));
    h->sum += 1.0;
    h->busy.fetch_sub(1);
    return MDS_HIP_OK;
}

extern "C" {
int mds_hip_create(int, int, long, int, mds_hip_handle* out) {
    *out = new mds_hip_engine;
    return MDS_HIP_OK;
}
int mds_hip_destroy(mds_hip_handle h) {
    if (h->busy.load() != 0) ++g_violations;  // destroyed under a call
    h->destroyed = true;  // kept allocated so that a late call is detected, not a crash
    return MDS_HIP_OK;
}
const char* mds_hip_last_error(mds_hip_handle) { return "stub error"; }
int mds_hip_update_locations(mds_hip_handle h, int, const double*, size_t) { return use(h); }
int mds_hip_get_sum_of_increments(mds_hip_handle h, double* out) {
    const int st = use(h);
    *out = h->sum;
    return st;
}
int mds_hip_get_log_likelihood_gradient(mds_hip_handle h, double*, size_t) { return use(h); }
int mds_hip_store_state(mds_hip_handle h) { return use(h); }
int mds_hip_restore_state(mds_hip_handle h) { return use(h); }
int mds_hip_accept_state(mds_hip_handle h) { return use(h); }
int mds_hip_set_pairwise_data(mds_hip_handle h, const double*, size_t) { return use(h); }
int mds_hip_set_parameters(mds_hip_handle h, const double*, size_t) { return use(h) == MDS_HIP_OK ? 7 : 7; }
int mds_hip_make_dirty(mds_hip_handle h) { return use(h); }
int mds_hip_get_internal_dimension(mds_hip_handle h, int* out) {
    *out = 2;
    return use(h);
}
}

int main() {
    const int threads = 6;
    std::vector<int> ids(threads);
    for (int t = 0; t < threads; ++t)
        if (mds_hip_singleton_initialize(2, 10, MDS_HIP_FLAG_OPENCL, 0, 1, &ids[t]) != MDS_HIP_OK) return 1;
    std::atomic<bool> released{false};
    std::atomic<int> badAfterRelease{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            // two threads share each instance: calls on it must be serialised
            const int id = ids[t / 2 * 2];
            for (int k = 0; k < 400; ++k) {
                const bool after = released.load();
                double s = 0.0;
                const int st = mds_hip_singleton_get_sum_of_increments(id, &s);
                if (after && st != MDS_HIP_ERR_ARGUMENT) ++badAfterRelease;
                // a failing call writes the message; other threads read theirs
                const double tau = 1.0;
                (void)mds_hip_singleton_set_parameters(id, &tau, 1);
                std::string msg = mds_hip_singleton_last_error(id);
                if (msg.empty()) ++badAfterRelease;
            }
        });
    std::thread releaser([&] {
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
        mds_hip_singleton_release_all();
        released = true;
    });
    for (auto& th : pool) th.join();
    releaser.join();
    double s = 0.0;
    if (mds_hip_singleton_get_sum_of_increments(ids[0], &s) != MDS_HIP_ERR_ARGUMENT) ++badAfterRelease;
    if (mds_hip_singleton_get_sum_of_increments(99, &s) != MDS_HIP_ERR_ARGUMENT) ++badAfterRelease;
    if (g_violations.load() || badAfterRelease.load()) {
        std::fprintf(stderr, "violations %d, bad results after release %d\n", g_violations.load(),
                     badAfterRelease.load());
        return 1;
    }
    return 0;
}
