/*
 * The plain spectral path tracer in FP64 — TEST INFRASTRUCTURE ONLY (see
 * oracle.h). Filled in with the integrator.
 */
#include "oracle.h"
