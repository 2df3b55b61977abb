// C-ABI bodies that allocate host memory or start threads run inside
// kc_guard: no C++ exception may cross the C ABI, so a failed allocation
// (std::bad_alloc, std::length_error) or thread start (std::system_error)
// becomes KC_ENOMEM, the status for a failed host or device allocation.
#pragma once

#include <utility>

#include "../../include/kc_api.h"

template <class F>
auto kc_guard(F &&body) noexcept -> decltype(body())
{
    try {
        return std::forward<F>(body)();
    } catch (...) {
        return KC_ENOMEM;
    }
}
