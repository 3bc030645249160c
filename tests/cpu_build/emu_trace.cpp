// The library source trace.hip, built by g++ for the bounds-checking build.
#include "trace.hip"
