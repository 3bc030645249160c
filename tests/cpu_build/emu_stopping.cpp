// The library source stopping.hip, built by g++ for the bounds-checking build.
#include "stopping.hip"
