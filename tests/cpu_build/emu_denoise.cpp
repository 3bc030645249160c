// The library source denoise.hip, built by g++ for the bounds-checking build.
#include "denoise.hip"
