// The library source serve.hip, built by g++ for the bounds-checking build.
#include "serve.hip"
