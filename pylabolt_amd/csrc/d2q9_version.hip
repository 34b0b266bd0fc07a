// d2q9_source_hash(): the digest of the sources and compiler flags this
// library was built from, stamped by the build as -DD2Q9_SOURCE_HASH=h<hex>
// (pylabolt_amd/build.py compiles this file on every build).
#include "../../include/d2q9.h"

#ifndef D2Q9_SOURCE_HASH
#define D2Q9_SOURCE_HASH unknown
#endif
#define D2Q9_STR2(x) #x
#define D2Q9_STR(x) D2Q9_STR2(x)

extern "C" const char* d2q9_source_hash(void) { return D2Q9_STR(D2Q9_SOURCE_HASH); }
