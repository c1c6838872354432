/* Drop-in replacement for the reference's include/propagate.h:13-19.
 *
 * dc_propagate advances every cell of the half-open box [start,end) of one
 * rank's local arrays by one Fletcher TTI time step (pp and qp updated in
 * place from pc, qc and the ten coefficient arrays), as
 * src/openmp_propagate.c:6-32 does.  The box must lie inside
 * [STENCIL, sizes-STENCIL) on every axis (the stencil reads 4 cells around
 * each updated cell); any other box is an error (message + exit(1)).  An
 * empty box is a no-op.  process_coordinates and topology are accepted and
 * unused, as in the reference.  The work is queued on the struct's stream
 * (see device_data.h); ordering with the other entry points is preserved. */
#pragma once

#ifndef DC_PROPAGATE_H
#define DC_PROPAGATE_H

#include "definitions.h"
#include "device_data.h"

#ifdef __cplusplus
extern "C" {
#endif

void dc_propagate(const size_t start_coords[DIMENSIONS],
                  const size_t end_coords[DIMENSIONS],
                  const size_t sizes[DIMENSIONS],
                  const int process_coordinates[DIMENSIONS],
                  const int topology[DIMENSIONS], dc_device_data *data,
                  const float dx, const float dy, const float dz,
                  const float dt);

#ifdef __cplusplus
}
#endif

#endif /* DC_PROPAGATE_H */
