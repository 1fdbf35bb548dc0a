// log.h -- the reference's log line format: every argument followed by a
// space, KINFO(session, ...) prefixed "[ session ] " (src/util/debug.hpp:37-46,
// 195-224). koishi::log_enabled() turns it off for library use.
#pragma once
#include <iostream>
#include <mutex>
#include <sstream>

namespace koishi {

inline bool &log_enabled() {
    static bool on = true;
    return on;
}
inline std::recursive_mutex &log_mutex() {
    static std::recursive_mutex m;
    return m;
}
inline void log_print(std::ostream &) {}
template <class X, class... A>
void log_print(std::ostream &os, const X &x, A &&...a) {
    os << x << " ";
    log_print(os, std::forward<A>(a)...);
}
template <class... A>
void log_line(A &&...a) {
    if (!log_enabled()) return;
    std::lock_guard<std::recursive_mutex> g(log_mutex());
    log_print(std::cout, std::forward<A>(a)...);
    std::cout << std::endl << std::flush;
}

}  // namespace koishi

#define KLOG(...) ::koishi::log_line(__VA_ARGS__)
#define KINFO(session, ...) ::koishi::log_line("[", session, "] ", __VA_ARGS__)
