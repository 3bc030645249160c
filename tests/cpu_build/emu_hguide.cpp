// The library source hguide.hip, built by g++ for the bounds-checking build.
#include "hguide.hip"
