#pragma once
#include <cstdarg>
#include <cstdio>
namespace megamol::core::utility::log {
class Log {
public:
    void WriteError(const char* fmt, ...) { va_list a; va_start(a, fmt); std::fputs("[error] ", stderr); std::vfprintf(stderr, fmt, a); std::fputc('\n', stderr); va_end(a); }
    void WriteInfo(const char* fmt, ...) { va_list a; va_start(a, fmt); std::vfprintf(stderr, fmt, a); std::fputc('\n', stderr); va_end(a); }
    static Log DefaultLog;
};
inline Log Log::DefaultLog;
}  // namespace megamol::core::utility::log
