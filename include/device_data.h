/* Drop-in replacement for the reference's include/device_data.h:11-37,
 * implemented by libdc_hip.so on one MI355X (gfx950).
 *
 * Semantics kept from the reference backends (src/device_data.cu,
 * src/device_data.c):
 *  - dc_device_data_init copies pp, pc, qp, qc and the ten precomp_vars arrays
 *    of the process (sizes[0]*sizes[1]*sizes[2] floats each, x fastest) to the
 *    device; the host keeps its arrays.
 *  - pp/pc/qp/qc in the returned struct are device addresses the caller only
 *    passes back (as the from/to array of the halo calls).  vpz and vsv are
 *    NULL: the kernels read the ten precomputed arrays, as the reference's
 *    OpenMP path does (src/openmp_propagate.c:22-27).  The precomp_vars
 *    pointers are device arrays the library only reads; it also keeps copies
 *    of their 4-cell x bands made at init, so they are not to be written
 *    afterwards.
 *  - Every call is ordered on one HIP stream owned by the struct.  Calls that
 *    hand data back to host memory (dc_device_data_get_results,
 *    dc_device_extract_halo_face) return only once it is there; insert returns
 *    once the host buffer has been consumed.
 *  - Errors print a message to stderr and exit(1), as the reference does
 *    (src/device_data.cu:10-16, src/cuda_propagate.cu:63-68). */
#pragma once

#include "dc_process.h"
#include "definitions.h"
#include "precomp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/device_data.h:11-15 (layout unchanged; the library appends its own
 * state after these fields in memory it allocates). */
typedef struct {
  float *pp, *pc, *qp, *qc;
  float *vpz, *vsv;
  dc_precomp_vars precomp_vars;
} dc_device_data;

/* include/device_data.h:17 -- allocate and upload one rank's arrays. */
dc_device_data *dc_device_data_init(dc_process_t *process);

/* include/device_data.h:19 -- release device memory and the struct. */
void dc_device_data_free(dc_device_data *data);

/* include/device_data.h:21 -- copy the latest pc and qc into process->pc and
 * process->qc (src/device_data.cu:107-117). */
void dc_device_data_get_results(dc_process_t *process, dc_device_data *data);

/* include/device_data.h:23 -- pp<->pc, qp<->qc (src/device_data.cu:119-129). */
void dc_device_swap_arrays(dc_device_data *data);

/* include/device_data.h:25 -- pc[index] += source; qc[index] += source. */
void dc_device_add_source(dc_device_data *data, size_t index, float source);

/* include/device_data.h:27-31 -- copy the box [start,end) of from_array (a
 * device array of data, laid out with `sizes`) into the host buffer, x
 * fastest (src/device_data.c:61-76). */
void dc_device_extract_halo_face(dc_device_data *data, float *buffer,
                                 const size_t start_coords[DIMENSIONS],
                                 const size_t end_coords[DIMENSIONS],
                                 const size_t sizes[DIMENSIONS],
                                 const float *from_array);

/* include/device_data.h:33-37 -- the inverse copy, host buffer -> box of
 * to_array (src/device_data.c:78-92). */
void dc_device_insert_halo_face(dc_device_data *data, const float *buffer,
                                const size_t start_coords[DIMENSIONS],
                                const size_t end_coords[DIMENSIONS],
                                const size_t sizes[DIMENSIONS],
                                float *to_array);

#ifdef __cplusplus
}
#endif
