// Build identity of the library: the hash of the sources it was compiled from, set by the build
// recipe. The Python loader compares it with the sources on disk and refuses a library built from
// other sources.
#include <hip/hip_runtime.h>

#ifndef ALDVC_SOURCE_HASH
#define ALDVC_SOURCE_HASH "unknown"
#endif

extern "C" const char* aldvc_source_hash(void) { return ALDVC_SOURCE_HASH; }
