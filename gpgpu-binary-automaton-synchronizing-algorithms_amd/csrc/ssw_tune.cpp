// ssw_tune.cpp -- placeholder until the autotuner lands.
#include "ssw_host.h"
extern "C" {
int ssw_autotune(ssw_context *, int32_t, int32_t, int64_t, uint64_t, const char *) { return SSW_ERR_INTERNAL; }
int ssw_load_tuning(ssw_context *, const char *) { return SSW_ERR_INTERNAL; }
}
