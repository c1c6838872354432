/* Drop-in replacement for the reference's include/dc_process.h:7-21.  The
 * rank descriptor that dc_device_data_init and dc_device_data_get_results
 * receive; same fields in the same order. */
#pragma once

#include <stddef.h>

#include "definitions.h"
#include "precomp.h"

typedef struct {
  int rank;
  int coordinates[DIMENSIONS];
  int neighbours[NEIGHBOURHOOD];
  int topology[DIMENSIONS];
  unsigned int iterations;
  int source_index;
  float dx, dy, dz, dt;
  size_t sizes[DIMENSIONS];
  dc_anisotropy_t anisotropy_vars;
  dc_precomp_vars precomp_vars;
  float *pp, *pc, *qp, *qc;
  char *hostnames; /* num_workers entries of 256 chars (MPI_MAX_PROCESSOR_NAME) */
  size_t num_workers;
} dc_process_t;
