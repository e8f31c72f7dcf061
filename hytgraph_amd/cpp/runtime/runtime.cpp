// Runtime helpers of the C++ API: Config (typed lookups over string
// settings), stderr logging with a level filter, and the experiment Result
// written as JSON (original interfaces: include/runtime/*.hpp).
#include "runtime/config.hpp"
#include "runtime/logging.hpp"
#include "runtime/result.hpp"

#include <charconv>
#include <cmath>
#include <ctime>
#include <fstream>
#include <iostream>
#include <mutex>
#include <stdexcept>

namespace hytgraph::runtime
{
    // ------------------------------------------------------------- Config
    void Config::set(std::string key, std::string value)
    {
        if (key.empty())
            throw std::invalid_argument("configuration key must not be empty");
        values_[std::move(key)] = std::move(value);
    }

    bool Config::contains(const std::string &key) const { return values_.count(key) != 0; }

    const std::string &Config::get(const std::string &key) const
    {
        const auto it = values_.find(key);
        if (it == values_.end())
            throw std::out_of_range("missing configuration key: " + key);
        return it->second;
    }

    std::string Config::get_or(const std::string &key, std::string fallback) const
    {
        const auto it = values_.find(key);
        return it == values_.end() ? fallback : it->second;
    }

    std::int64_t Config::get_int(const std::string &key) const { return std::stoll(get(key)); }

    double Config::get_double(const std::string &key) const { return std::stod(get(key)); }

    bool Config::get_bool(const std::string &key) const
    {
        std::string v = get(key);
        for (char &c : v)
            if (c >= 'A' && c <= 'Z')
                c = static_cast<char>(c - 'A' + 'a');
        for (const char *t : {"true", "1", "yes", "on"})
            if (v == t)
                return true;
        for (const char *f : {"false", "0", "no", "off"})
            if (v == f)
                return false;
        throw std::invalid_argument("invalid boolean configuration value for key: " + key);
    }

    const std::map<std::string, std::string> &Config::values() const noexcept { return values_; }

    // ------------------------------------------------------------ logging
    namespace
    {
        std::mutex g_log_mutex;
        LogLevel g_log_level = LogLevel::Info;

        const char *level_tag(LogLevel level) noexcept
        {
            switch (level)
            {
            case LogLevel::Debug: return "DEBUG";
            case LogLevel::Info: return "INFO";
            case LogLevel::Warning: return "WARN";
            case LogLevel::Error: return "ERROR";
            }
            return "UNKNOWN";
        }
    } // namespace

    void set_log_level(LogLevel level) noexcept
    {
        std::lock_guard<std::mutex> lock(g_log_mutex);
        g_log_level = level;
    }

    void log(LogLevel level, std::string_view message)
    {
        std::lock_guard<std::mutex> lock(g_log_mutex);
        if (static_cast<int>(level) < static_cast<int>(g_log_level))
            return;
        const std::time_t now = std::time(nullptr);
        std::tm utc{};
        gmtime_r(&now, &utc);
        char stamp[32];
        std::strftime(stamp, sizeof(stamp), "%Y-%m-%dT%H:%M:%SZ", &utc);
        std::cerr << '[' << stamp << "] " << level_tag(level) << ' ' << message << '\n';
    }

    // ------------------------------------------------------------- Result
    namespace
    {
        void put_string(std::string &out, const std::string &s)
        {
            out += '"';
            for (const char c : s)
            {
                switch (c)
                {
                case '"': out += "\\\""; break;
                case '\\': out += "\\\\"; break;
                case '\n': out += "\\n"; break;
                case '\r': out += "\\r"; break;
                case '\t': out += "\\t"; break;
                default:
                    if (static_cast<unsigned char>(c) < 0x20)
                    {
                        char buf[8];
                        std::snprintf(buf, sizeof(buf), "\\u%04x", static_cast<unsigned>(c));
                        out += buf;
                    }
                    else
                        out += c;
                }
            }
            out += '"';
        }

        void put_number(std::string &out, double v)
        {
            if (!std::isfinite(v))
            {
                out += "null";
                return;
            }
            char buf[64];
            const auto r = std::to_chars(buf, buf + sizeof(buf), v); // shortest round trip
            out.append(buf, r.ptr);
        }

        void put_field(std::string &out, const char *name, const std::string &value)
        {
            out += "  ";
            put_string(out, name);
            out += ": ";
            put_string(out, value);
            out += ",\n";
        }
    } // namespace

    std::string result_to_json(const Result &result)
    {
        std::string out = "{\n";
        put_field(out, "schema_version", result.schema_version);
        put_field(out, "experiment_id", result.experiment_id);
        put_field(out, "status", result.status);
        put_field(out, "timestamp_utc", result.timestamp_utc);
        put_field(out, "git_commit", result.git_commit);
        out += "  \"config\": {";
        const char *sep = "\n";
        for (const auto &[k, v] : result.config)
        {
            out += sep;
            out += "    ";
            put_string(out, k);
            out += ": ";
            put_string(out, v);
            sep = ",\n";
        }
        out += result.config.empty() ? "},\n" : "\n  },\n";
        out += "  \"measurements\": {";
        sep = "\n";
        for (const auto &[k, v] : result.measurements)
        {
            out += sep;
            out += "    ";
            put_string(out, k);
            out += ": ";
            put_number(out, v);
            sep = ",\n";
        }
        out += result.measurements.empty() ? "}\n" : "\n  }\n";
        out += "}\n";
        return out;
    }

    bool write_result_json(const Result &result, const std::string &path)
    {
        std::ofstream f(path);
        if (!f)
            return false;
        f << result_to_json(result);
        return static_cast<bool>(f);
    }
} // namespace hytgraph::runtime
