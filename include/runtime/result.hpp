// Result - one experiment record as JSON; the interface and field layout of
// the original project's include/runtime/result.hpp and
// docs/result_schema.json (schema_version, experiment_id, status,
// timestamp_utc, git_commit, config: string map, measurements: number map).
#pragma once

#include <map>
#include <string>

namespace hytgraph::runtime
{
    struct Result
    {
        std::string schema_version = "0.1";
        std::string experiment_id;
        std::string status = "success";
        std::string timestamp_utc;
        std::string git_commit = "unknown";
        std::map<std::string, std::string> config;
        std::map<std::string, double> measurements;
    };

    // Measurements are written in the shortest form that reads back to the
    // same double; a non-finite measurement is written as null.
    [[nodiscard]] std::string result_to_json(const Result &result);
    // false when the file cannot be written
    bool write_result_json(const Result &result, const std::string &path);
} // namespace hytgraph::runtime
