/* Drop-in replacement for the reference's include/precomp.h:3-26: the TTI
 * coefficient and anisotropy structs, field for field in the same order, so
 * the struct layout (and therefore the ABI of dc_process_t and
 * dc_device_data) is unchanged. */
#pragma once

#define SIGMA 0.75
#define MAX_SIGMA 10.0

/* Per-cell Fletcher TTI coefficients (include/precomp.h:6-17). */
typedef struct {
  float *ch1dxx;
  float *ch1dyy;
  float *ch1dzz;
  float *ch1dxy;
  float *ch1dyz;
  float *ch1dxz;
  float *v2px;
  float *v2pz;
  float *v2sz;
  float *v2pn;
} dc_precomp_vars;

/* Per-cell model parameters (include/precomp.h:19-26). */
typedef struct {
  float *theta;
  float *phi;
  float *vsv;
  float *vpz;
  float *epsilon;
  float *delta;
} dc_anisotropy_t;
