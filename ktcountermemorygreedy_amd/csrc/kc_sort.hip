// placeholder replaced below
#include "kc_internal.h"
extern "C" int kc_sort_records(kc_ctx *, void *, void *, uint64_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, int *)
{
    return KC_ENOTSUP;
}
