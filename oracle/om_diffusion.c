/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Placeholder translation unit for the viscous (diffusion) builders of
 * src/ns/ns_diffusion_module.F90 and ns_vectorial_diffusion_module.F90.
 */
#include "om_internal.h"

int om_diffusion_available(void) { return 0; }
