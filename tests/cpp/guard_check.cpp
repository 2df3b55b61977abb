// kc_guard: a C-ABI body's allocation or thread-start failure becomes
// KC_ENOMEM instead of an exception crossing the C ABI; ordinary returns,
// including negative statuses and 64-bit values, pass through.
#include <cstdint>
#include <cstdio>
#include <new>
#include <stdexcept>
#include <system_error>
#include <vector>

#include "kc_guard.h"

static int fail(const char *what)
{
    std::printf("FAIL %s\n", what);
    return 1;
}

int main()
{
    if (kc_guard([] { return KC_OK; }) != KC_OK) return fail("ok passes through");
    if (kc_guard([] { return KC_EIO; }) != KC_EIO) return fail("status passes through");
    if (kc_guard([]() -> int64_t { return (int64_t)1 << 40; }) != ((int64_t)1 << 40)) return fail("int64 passes through");
    if (kc_guard([]() -> int { throw std::bad_alloc(); }) != KC_ENOMEM) return fail("bad_alloc");
    if (kc_guard([]() -> int64_t { throw std::bad_alloc(); }) != KC_ENOMEM) return fail("bad_alloc int64");
    if (kc_guard([]() -> int { throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again)); }) !=
        KC_ENOMEM)
        return fail("thread start");
    if (kc_guard([]() -> int {
            std::vector<char> v;
            v.resize(v.max_size() + 1ull);  // std::length_error
            return KC_OK;
        }) != KC_ENOMEM)
        return fail("length_error");
    std::printf("ok\n");
    return 0;
}
