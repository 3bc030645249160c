// The library source guide.hip, built by g++ for the bounds-checking build.
#include "guide.hip"
