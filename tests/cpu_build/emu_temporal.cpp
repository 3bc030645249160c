// The library source temporal.hip, built by g++ for the bounds-checking build.
#include "temporal.hip"
