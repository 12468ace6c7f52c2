// yaml_lite.cpp -- see yaml_lite.h.
#include "yaml_lite.h"

#include <cctype>
#include <cmath>
#include <cstring>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace mpd {

namespace {

struct Line {
    int indent;
    std::string text;  // without indentation and comment
    int number;        // 1-based
};

[[noreturn]] void Fail(int line, const std::string& msg) {
    throw std::runtime_error("YAML parse error (line " + std::to_string(line) + "): " + msg);
}

std::string StripComment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '\\' && dq && i + 1 < s.size()) ++i;
        else if (c == '#' && !sq && !dq && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
    }
    return s;
}

std::string Trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Position of the ':' that separates a mapping key from its value, or npos.
size_t FindKeyColon(const std::string& s) {
    bool sq = false, dq = false;
    int depth = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (sq || dq) continue;
        else if (c == '{' || c == '[') ++depth;
        else if (c == '}' || c == ']') --depth;
        else if (c == ':' && depth == 0 && (i + 1 == s.size() || std::isspace((unsigned char)s[i + 1]))) return i;
    }
    return std::string::npos;
}

YNode MakeScalar(const std::string& raw, int line) {
    YNode n;
    const std::string t = Trim(raw);
    if (t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL") return n;  // null
    n.kind = YNode::kScalar;
    if (t[0] == '"') {
        if (t.size() < 2 || t.back() != '"') Fail(line, "unterminated double-quoted scalar");
        std::string out;
        for (size_t i = 1; i + 1 < t.size(); ++i) {
            if (t[i] == '\\' && i + 2 < t.size()) {
                const char e = t[++i];
                out += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
            } else {
                out += t[i];
            }
        }
        n.scalar = out;
        n.quoted = true;
    } else if (t[0] == '\'') {
        if (t.size() < 2 || t.back() != '\'') Fail(line, "unterminated single-quoted scalar");
        std::string out;
        for (size_t i = 1; i + 1 < t.size(); ++i) {
            if (t[i] == '\'' && i + 2 < t.size() && t[i + 1] == '\'') ++i;
            out += t[i];
        }
        n.scalar = out;
        n.quoted = true;
    } else {
        n.scalar = t;
    }
    return n;
}

// Flow collections: { k: v, ... } and [ a, b, ... ], possibly nested.
class FlowParser {
public:
    FlowParser(const std::string& s, int line) : s_(s), line_(line) {}
    YNode Parse() {
        YNode n = Value();
        Ws();
        if (p_ != s_.size()) Fail(line_, "trailing characters after flow collection");
        return n;
    }

private:
    void Ws() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }
    YNode Value() {
        Ws();
        if (p_ >= s_.size()) Fail(line_, "unexpected end of flow collection");
        if (s_[p_] == '{') return Map();
        if (s_[p_] == '[') return Seq();
        return Scalar(",}]");
    }
    YNode Scalar(const char* stops) {
        Ws();
        size_t st = p_;
        if (p_ < s_.size() && (s_[p_] == '"' || s_[p_] == '\'')) {
            const char q = s_[p_++];
            while (p_ < s_.size() && s_[p_] != q) {
                if (q == '"' && s_[p_] == '\\') ++p_;
                ++p_;
            }
            if (p_ >= s_.size()) Fail(line_, "unterminated quoted scalar");
            ++p_;
        } else {
            while (p_ < s_.size() && !std::strchr(stops, s_[p_])) {
                if (s_[p_] == ':' && (p_ + 1 == s_.size() || std::isspace((unsigned char)s_[p_ + 1]) ||
                                      std::strchr(",}]", s_[p_ + 1])))
                    break;
                ++p_;
            }
        }
        return MakeScalar(s_.substr(st, p_ - st), line_);
    }
    YNode Map() {
        YNode n;
        n.kind = YNode::kMap;
        ++p_;
        Ws();
        if (p_ < s_.size() && s_[p_] == '}') {
            ++p_;
            return n;
        }
        for (;;) {
            YNode key = Scalar(":,}");
            Ws();
            YNode val;
            if (p_ < s_.size() && s_[p_] == ':') {
                ++p_;
                Ws();
                if (p_ < s_.size() && (s_[p_] == ',' || s_[p_] == '}')) {
                    // null value
                } else {
                    val = Value();
                }
            }
            n.map.emplace_back(key.scalar, val);
            Ws();
            if (p_ < s_.size() && s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return n;
            }
            Fail(line_, "expected ',' or '}' in flow map");
        }
    }
    YNode Seq() {
        YNode n;
        n.kind = YNode::kSeq;
        ++p_;
        Ws();
        if (p_ < s_.size() && s_[p_] == ']') {
            ++p_;
            return n;
        }
        for (;;) {
            n.seq.push_back(Value());
            Ws();
            if (p_ < s_.size() && s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return n;
            }
            Fail(line_, "expected ',' or ']' in flow sequence");
        }
    }
    const std::string& s_;
    int line_;
    size_t p_ = 0;
};

YNode InlineValue(const std::string& text, int line) {
    const std::string t = Trim(text);
    if (!t.empty() && (t[0] == '{' || t[0] == '[')) return FlowParser(t, line).Parse();
    return MakeScalar(t, line);
}

class BlockParser {
public:
    explicit BlockParser(std::vector<Line> lines) : ln_(std::move(lines)) {}

    YNode Document() {
        if (ln_.empty()) return {};
        YNode n = Block(ln_[0].indent);
        if (i_ != ln_.size()) Fail(ln_[i_].number, "unexpected indentation");
        return n;
    }

private:
    YNode Block(int indent) {
        const Line& first = ln_[i_];
        if (first.text.rfind("- ", 0) == 0 || first.text == "-") return Sequence(indent);
        if (FindKeyColon(first.text) != std::string::npos) return Mapping(indent);
        YNode n = InlineValue(first.text, first.number);
        ++i_;
        return n;
    }

    YNode Mapping(int indent) {
        YNode n;
        n.kind = YNode::kMap;
        while (i_ < ln_.size() && ln_[i_].indent == indent) {
            const Line& cur = ln_[i_];
            if (cur.text.rfind("- ", 0) == 0) Fail(cur.number, "sequence item inside a mapping");
            const size_t c = FindKeyColon(cur.text);
            if (c == std::string::npos) Fail(cur.number, "expected 'key: value'");
            std::string key = MakeScalar(cur.text.substr(0, c), cur.number).scalar;
            const std::string rest = Trim(cur.text.substr(c + 1));
            ++i_;
            YNode val;
            if (!rest.empty()) {
                val = InlineValue(rest, cur.number);
            } else if (i_ < ln_.size() && ln_[i_].indent > indent) {
                val = Block(ln_[i_].indent);
            } else if (i_ < ln_.size() && ln_[i_].indent == indent && ln_[i_].text.rfind("- ", 0) == 0) {
                // A block sequence may sit at the key's own indentation.
                val = Sequence(indent);
            }
            for (const auto& kv : n.map)
                if (kv.first == key) Fail(cur.number, "duplicate key '" + key + "'");
            n.map.emplace_back(key, std::move(val));
        }
        if (i_ < ln_.size() && ln_[i_].indent > indent) Fail(ln_[i_].number, "bad indentation");
        return n;
    }

    YNode Sequence(int indent) {
        YNode n;
        n.kind = YNode::kSeq;
        while (i_ < ln_.size() && ln_[i_].indent == indent &&
               (ln_[i_].text.rfind("- ", 0) == 0 || ln_[i_].text == "-")) {
            Line& cur = ln_[i_];
            if (cur.text == "-") {
                ++i_;
                if (i_ < ln_.size() && ln_[i_].indent > indent) n.seq.push_back(Block(ln_[i_].indent));
                else n.seq.emplace_back();
                continue;
            }
            // Turn "- rest" into a line "rest" indented past the dash.
            size_t k = 1;
            while (k < cur.text.size() && cur.text[k] == ' ') ++k;
            cur.indent += (int)k;
            cur.text = cur.text.substr(k);
            n.seq.push_back(Block(cur.indent));
        }
        return n;
    }

    std::vector<Line> ln_;
    size_t i_ = 0;
};

}  // namespace

const YNode* YNode::Get(const std::string& key) const {
    if (kind != kMap) return nullptr;
    for (const auto& kv : map)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

double YNode::AsDouble() const {
    if (kind != kScalar) throw std::runtime_error("YAML: expected a number");
    const std::string& s = scalar;
    if (s == ".inf" || s == ".Inf" || s == "+.inf") return HUGE_VAL;
    if (s == "-.inf" || s == "-.Inf") return -HUGE_VAL;
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != 0) throw std::runtime_error("YAML: bad conversion of '" + s + "' to double");
    return v;
}

int YNode::AsInt() const {
    if (kind != kScalar) throw std::runtime_error("YAML: expected an integer");
    char* end = nullptr;
    const long v = std::strtol(scalar.c_str(), &end, 10);
    if (end == scalar.c_str() || *end != 0) throw std::runtime_error("YAML: bad conversion of '" + scalar + "' to int");
    return (int)v;
}

bool YNode::AsBool() const {
    if (kind != kScalar) throw std::runtime_error("YAML: expected a boolean");
    static const char* yes[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON", "y", "Y"};
    static const char* no[] = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF", "n", "N"};
    for (const char* t : yes)
        if (scalar == t) return true;
    for (const char* t : no)
        if (scalar == t) return false;
    throw std::runtime_error("YAML: bad conversion of '" + scalar + "' to bool");
}

const std::string& YNode::AsString() const {
    if (kind != kScalar) throw std::runtime_error("YAML: expected a scalar string");
    return scalar;
}

YNode ParseYaml(const std::string& text) {
    std::vector<Line> lines;
    std::istringstream in(text);
    std::string raw;
    int no = 0;
    while (std::getline(in, raw)) {
        ++no;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        if (raw.find('\t') != std::string::npos && Trim(StripComment(raw)).size() &&
            raw.find_first_not_of(" \t") != std::string::npos && raw[raw.find_first_not_of(' ')] == '\t')
            Fail(no, "tab indentation is not allowed");
        const std::string s = StripComment(raw);
        const std::string t = Trim(s);
        if (t.empty() || t == "---") continue;
        int ind = 0;
        while (ind < (int)s.size() && s[ind] == ' ') ++ind;
        lines.push_back({ind, t, no});
    }
    return BlockParser(std::move(lines)).Document();
}

YNode LoadYamlFile(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return ParseYaml(ss.str());
}

}  // namespace mpd
