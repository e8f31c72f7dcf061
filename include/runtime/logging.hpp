// Leveled logging to stderr with a UTC timestamp; the interface of the
// original project's include/runtime/logging.hpp.
#pragma once

#include <string_view>

namespace hytgraph::runtime
{
    enum class LogLevel
    {
        Debug,
        Info,
        Warning,
        Error,
    };

    void set_log_level(LogLevel level) noexcept;
    void log(LogLevel level, std::string_view message);
} // namespace hytgraph::runtime
