// Op-counting build of the oracle (test infrastructure): the oracle sources
// recompiled with counted scalar types. Exposes per-workload op totals.
#include "counted.h"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>

thread_local uint64_t ocount::g_ops = 0;

#define OREF_COUNTING 1
#define float CFloat
#define uint32_t CU32
#include "../noise.cpp"
#include "../gabor.cpp"
#include "../color.cpp"
#include "../closure.cpp"
#include "../shade.cpp"
#include "../mtx.cpp"
#include "../mtx_api.cpp"
#include "../osoint.cpp"
#undef float
#undef uint32_t

extern "C" {

void ocount_reset(void) { ocount::g_ops = 0; }
unsigned long long ocount_get(void) { return ocount::g_ops; }

}  // extern "C"
