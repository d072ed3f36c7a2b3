// buildinfo.cpp - the build's identity: a SHA-256 over the library's sources
// and compile flags (computed by strelka_amd/build.py and passed as
// STK_SOURCE_HASH) and the offload architecture.
#include "../../include/stk/stk.h"

#ifndef STK_SOURCE_HASH
#define STK_SOURCE_HASH "unknown"
#endif
#ifndef STK_OFFLOAD_ARCH
#define STK_OFFLOAD_ARCH "unknown"
#endif

extern "C" STK_API const char* stk_build_info(void)
{
    return "source-sha256 " STK_SOURCE_HASH "; offload-arch " STK_OFFLOAD_ARCH;
}
