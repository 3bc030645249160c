// The library source octree.hip, built by g++ for the bounds-checking build.
#include "octree.hip"
