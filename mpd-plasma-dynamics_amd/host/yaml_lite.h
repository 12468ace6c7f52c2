// yaml_lite.h -- a small reader for the YAML subset that config.yaml uses
// (block maps and sequences, flow maps/sequences, quoted and plain scalars,
// comments).  It stands in for yaml-cpp 0.8.0 (cmake/Dependencies.cmake:5 of
// the reference), which is not available here; only the node queries the
// config loader needs are provided.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace mpd {

struct YNode {
    enum Kind { kNull, kScalar, kMap, kSeq };
    Kind kind = kNull;
    std::string scalar;
    bool quoted = false;
    std::vector<std::pair<std::string, YNode>> map;
    std::vector<YNode> seq;

    bool IsNull() const { return kind == kNull; }
    bool IsScalar() const { return kind == kScalar; }
    bool IsMap() const { return kind == kMap; }
    bool IsSequence() const { return kind == kSeq; }
    size_t size() const { return kind == kMap ? map.size() : seq.size(); }

    // Child lookup; returns nullptr when absent (or when this is not a map).
    const YNode* Get(const std::string& key) const;

    // Conversions with yaml-cpp's as<T>() meaning; throw std::runtime_error.
    double AsDouble() const;
    int AsInt() const;
    bool AsBool() const;
    const std::string& AsString() const;
};

// Parse a whole document; throws std::runtime_error with a line number.
YNode ParseYaml(const std::string& text);
YNode LoadYamlFile(const std::string& path);

}  // namespace mpd
