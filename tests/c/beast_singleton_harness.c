/*
 * C harness of the BEAST bridge (include/mds_hip_singleton.h): the call
 * sequence BEAST's likelihood makes through the reference JNI singleton
 * (src/jni/dr_inference_multidimensionalscaling_NativeMDSSingleton.cpp:9-101),
 * made through the C registry, since no Java is available to drive the JNI
 * layer itself.  Reads N x N observations and N x P locations (raw doubles),
 * writes the results as "key value ..." lines with 17 significant digits.
 *
 *   beast_singleton_harness <delta.bin> <x.bin> <N> <P> <tau> <out.txt>
 *
 * Exit status: 0, or the first failing call's status (its message on stderr).
 */
#include <stdio.h>
#include <stdlib.h>

#include "mds_hip_singleton.h"

#define CHECK(call, inst)                                                                       \
    do {                                                                                        \
        int st_ = (call);                                                                       \
        if (st_ != MDS_HIP_OK) {                                                                \
            fprintf(stderr, "%s failed (%d): %s\n", #call, st_, mds_hip_singleton_last_error(inst)); \
            return st_;                                                                         \
        }                                                                                       \
    } while (0)

static double* read_doubles(const char* path, size_t count) {
    FILE* f = fopen(path, "rb");
    double* v = (double*)malloc(count * sizeof(double));
    if (!f || !v || fread(v, sizeof(double), count, f) != count) {
        fprintf(stderr, "cannot read %zu doubles from %s\n", count, path);
        exit(9);
    }
    fclose(f);
    return v;
}

static void put(FILE* out, const char* key, const double* v, size_t count) {
    fprintf(out, "%s", key);
    for (size_t k = 0; k < count; ++k) fprintf(out, " %.17g", v[k]);
    fprintf(out, "\n");
}

int main(int argc, char** argv) {
    if (argc != 7) {
        fprintf(stderr, "usage: %s delta.bin x.bin N P tau out.txt\n", argv[0]);
        return 8;
    }
    const int n = atoi(argv[3]), p = atoi(argv[4]);
    const double tau = atof(argv[5]);
    double* delta = read_doubles(argv[1], (size_t)n * n);
    double* x = read_doubles(argv[2], (size_t)n * p);
    double* g = (double*)malloc(sizeof(double) * (size_t)n * p);
    FILE* out = fopen(argv[6], "w");
    if (!out || !g) return 9;

    int before = 0, inst = -1, inst2 = -1;
    CHECK(mds_hip_singleton_instance_count(&before), -1);
    /* flags as BEAST passes them for the GPU engine: OPENCL | LEFT_TRUNCATION */
    CHECK(mds_hip_singleton_initialize(p, n, MDS_HIP_FLAG_OPENCL | MDS_HIP_FLAG_LEFT_TRUNCATION, 0, 1, &inst), -1);
    CHECK(mds_hip_singleton_initialize(p, n, MDS_HIP_FLAG_OPENCL | MDS_HIP_FLAG_FLOAT, 0, 1, &inst2), -1);
    fprintf(out, "instances %d %d %d\n", before, inst, inst2);

    CHECK(mds_hip_singleton_set_pairwise_data(inst, delta, (size_t)n * n), inst);
    CHECK(mds_hip_singleton_set_parameters(inst, &tau, 1), inst);
    CHECK(mds_hip_singleton_update_locations(inst, -1, x, (size_t)n * p), inst);
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    CHECK(mds_hip_singleton_get_sum_of_increments(inst, &s0), inst);
    CHECK(mds_hip_singleton_get_location_gradient(inst, g, (size_t)n * p), inst);
    put(out, "sum0", &s0, 1);
    put(out, "grad0", g, (size_t)n * p);

    /* a rejected single-location move, then an accepted one */
    CHECK(mds_hip_singleton_store_state(inst), inst);
    double row[16];
    for (int k = 0; k < p; ++k) row[k] = x[3 * p + k] + 0.1 * (k + 1);
    CHECK(mds_hip_singleton_update_locations(inst, 3, row, (size_t)p), inst);
    CHECK(mds_hip_singleton_get_sum_of_increments(inst, &s1), inst);
    CHECK(mds_hip_singleton_restore_state(inst), inst);
    CHECK(mds_hip_singleton_get_sum_of_increments(inst, &s2), inst);
    put(out, "sum_moved", &s1, 1);
    put(out, "sum_restored", &s2, 1);
    CHECK(mds_hip_singleton_store_state(inst), inst);
    CHECK(mds_hip_singleton_update_locations(inst, 3, row, (size_t)p), inst);
    CHECK(mds_hip_singleton_get_sum_of_increments(inst, &s3), inst);
    CHECK(mds_hip_singleton_accept_state(inst), inst);
    CHECK(mds_hip_singleton_get_location_gradient(inst, g, (size_t)n * p), inst);
    put(out, "sum_accepted", &s3, 1);
    put(out, "grad_accepted", g, (size_t)n * p);

    int dim = 0;
    CHECK(mds_hip_singleton_get_internal_dimension(inst, &dim), inst);
    fprintf(out, "internal_dimension %d\n", dim);
    double s4 = 0;
    CHECK(mds_hip_singleton_make_dirty(inst), inst);
    CHECK(mds_hip_singleton_get_sum_of_increments(inst, &s4), inst);
    put(out, "sum_dirty", &s4, 1);

    /* errors come back as statuses with messages */
    double dummy = 0;
    fprintf(out, "bad_instance %d\n", mds_hip_singleton_get_sum_of_increments(inst2 + 1000, &dummy));
    fprintf(out, "bad_length %d\n", mds_hip_singleton_update_locations(inst, 3, row, (size_t)p + 1));
    fprintf(out, "cpu_flags %d\n", mds_hip_singleton_initialize(p, n, MDS_HIP_FLAG_LEFT_TRUNCATION, 0, 1, &dim));
    CHECK(mds_hip_singleton_release_all(), -1);
    fprintf(out, "after_release %d\n", mds_hip_singleton_get_sum_of_increments(inst, &dummy));
    fclose(out);
    free(delta);
    free(x);
    free(g);
    return 0;
}
