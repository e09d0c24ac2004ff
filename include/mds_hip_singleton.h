/*
 * Instance registry of the BEAST bridge: the calls of the reference's JNI
 * singleton (src/jni/dr_inference_multidimensionalscaling_NativeMDSSingleton.cpp:9-101,
 * .h:15-98) as plain C, so that the JNI layer (or any other foreign-function
 * binding) is one line per method over libmds_hip.so.  Instances are numbered
 * 0, 1, ... in creation order, as the reference's global vector numbers them,
 * and live until mds_hip_singleton_release_all().
 *
 * Every call returns MDS_HIP_OK or an MDS_HIP_ERR_* status (mds_hip.h);
 * mds_hip_singleton_last_error(instance) gives the message (instance -1: the
 * registry's own errors, such as a bad instance number).  The registry is safe
 * to use from several threads: calls on one instance are serialised (each
 * holds the instance's lock), mds_hip_singleton_release_all waits for calls
 * in flight, and a call on a released instance returns MDS_HIP_ERR_ARGUMENT.
 * mds_hip_singleton_last_error returns the calling thread's copy of the
 * message, valid until that thread calls it again.
 */
#ifndef MDS_HIP_SINGLETON_H
#define MDS_HIP_SINGLETON_H

#include <stddef.h>

#include "mds_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* NativeMDSSingleton.initialize (.cpp:12-21): mds::factory(embeddingDimension,
 * elementCount, flags, device, threads).  The GPU route is selected as the
 * reference factory selects it (MDS_HIP_FLAG_OPENCL, or MDS_HIP_FLAG_HIP);
 * other flag sets are refused (this library has no CPU engine).  *instance
 * receives the new instance number. */
int mds_hip_singleton_initialize(int embeddingDimension, int elementCount, long flags, int device,
                                 int threads, int* instance);
/* updateLocations (.cpp:23-32): index -1 = all N x P values, else one row of P. */
int mds_hip_singleton_update_locations(int instance, int index, const double* x, size_t length);
/* getSumOfIncrements (.cpp:34-38). */
int mds_hip_singleton_get_sum_of_increments(int instance, double* sum);
/* storeState, restoreState, acceptState, makeDirty (.cpp:40-56, 91-95). */
int mds_hip_singleton_store_state(int instance);
int mds_hip_singleton_restore_state(int instance);
int mds_hip_singleton_accept_state(int instance);
int mds_hip_singleton_make_dirty(int instance);
/* setPairwiseData (.cpp:58-67): N x N row-major, NaN = missing. */
int mds_hip_singleton_set_pairwise_data(int instance, const double* data, size_t length);
/* getLocationGradient (.cpp:69-78): the log-likelihood gradient, N x P. */
int mds_hip_singleton_get_location_gradient(int instance, double* gradient, size_t length);
/* setParameters (.cpp:80-89): the precision. */
int mds_hip_singleton_set_parameters(int instance, const double* parameters, size_t length);
/* getInternalDimension (.cpp:97-101). */
int mds_hip_singleton_get_internal_dimension(int instance, int* dimension);
/* Number of instances created so far (including released ones: numbers are not reused). */
int mds_hip_singleton_instance_count(int* count);
/* Releases every instance (the reference keeps them for the process's life). */
int mds_hip_singleton_release_all(void);
const char* mds_hip_singleton_last_error(int instance);

#ifdef __cplusplus
}
#endif

#endif /* MDS_HIP_SINGLETON_H */
