// The library source integrator.hip, built by g++ for the bounds-checking build.
#include "integrator.hip"
