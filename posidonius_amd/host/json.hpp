// json.hpp -- a small JSON document model for the case and snapshot files.
//
// Objects keep their key order and numbers keep their source text, so that a
// document that is read, partly updated and written back differs from the
// input only in the updated values (the reference's files hold 64-bit
// integers, e.g. the snapshot hash, that a double would not round-trip).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace posjson {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    std::string text;   // number source text or string contents
    std::vector<Value> items;
    std::vector<std::pair<std::string, Value>> members;

    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
    bool is_string() const { return kind == String; }
    bool is_number() const { return kind == Number; }

    const Value *find(const std::string &key) const {
        for (const auto &m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
    Value *find(const std::string &key) {
        for (auto &m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
    const Value &at(const std::string &key) const {
        const Value *v = find(key);
        if (!v) throw std::runtime_error("missing key '" + key + "'");
        return *v;
    }
    Value &at(const std::string &key) {
        Value *v = find(key);
        if (!v) throw std::runtime_error("missing key '" + key + "'");
        return *v;
    }
    const Value &operator[](size_t i) const {
        if (kind != Array || i >= items.size()) throw std::runtime_error("array index out of range");
        return items[i];
    }
    Value &operator[](size_t i) {
        if (kind != Array || i >= items.size()) throw std::runtime_error("array index out of range");
        return items[i];
    }
    double num() const {
        if (kind == Null) return std::nan("");   // serde_json writes a non-finite f64 as null
        if (kind != Number) throw std::runtime_error("expected a number");
        return std::strtod(text.c_str(), nullptr);
    }
    long long integer() const {
        if (kind != Number) throw std::runtime_error("expected a number");
        return std::strtoll(text.c_str(), nullptr, 10);
    }
    const std::string &str() const {
        if (kind != String) throw std::runtime_error("expected a string");
        return text;
    }

    static Value number(double x) {
        Value v;
        if (!std::isfinite(x)) return v;   // null, as serde_json writes NaN and infinities
        v.kind = Number;
        char buf[40];
        std::snprintf(buf, sizeof buf, "%.17g", x);
        v.text = buf;
        if (v.text.find_first_of(".eEn") == std::string::npos) v.text += ".0";   // keep it a float for serde
        return v;
    }
    static Value integer_value(long long x) {
        Value v;
        v.kind = Number;
        v.text = std::to_string(x);
        return v;
    }
};

class Parser {
  public:
    explicit Parser(const std::string &s) : s_(s) {}
    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) error("trailing characters");
        return v;
    }

  private:
    const std::string &s_;
    size_t i_ = 0;

    [[noreturn]] void error(const char *what) {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
    }
    char peek() {
        ws();
        if (i_ >= s_.size()) error("unexpected end");
        return s_[i_];
    }
    void expect(char c) {
        if (peek() != c) error("unexpected character");
        ++i_;
    }
    bool literal(const char *w) {
        const size_t n = std::strlen(w);
        if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
        return false;
    }
    Value value() {
        const char c = peek();
        Value v;
        if (c == '{') {
            v.kind = Value::Object;
            ++i_;
            if (peek() == '}') { ++i_; return v; }
            for (;;) {
                std::string k = string_body();
                expect(':');
                v.members.emplace_back(std::move(k), value());
                if (peek() == ',') { ++i_; continue; }
                expect('}');
                return v;
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            ++i_;
            if (peek() == ']') { ++i_; return v; }
            for (;;) {
                v.items.push_back(value());
                if (peek() == ',') { ++i_; continue; }
                expect(']');
                return v;
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.text = string_body();
            return v;
        }
        if (literal("true")) { v.kind = Value::Bool; v.b = true; return v; }
        if (literal("false")) { v.kind = Value::Bool; v.b = false; return v; }
        if (literal("null")) return v;
        const size_t start = i_;
        while (i_ < s_.size() && std::strchr("+-0123456789.eE", s_[i_])) ++i_;
        if (i_ == start) error("unexpected character");
        v.kind = Value::Number;
        v.text = s_.substr(start, i_ - start);
        return v;
    }
    std::string string_body() {
        expect('"');
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) error("bad escape");
                const char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) error("bad escape");
                        const unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                        else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) error("unterminated string");
        ++i_;
        return out;
    }
};

inline Value parse(const std::string &s) { return Parser(s).parse(); }

inline void dump_string(const std::string &s, std::string &out) {
    out += '"';
    for (char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\t': out += "\\t"; break;
            case '\r': out += "\\r"; break;
            default:
                if ((unsigned char)c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)c);
                    out += buf;
                } else {
                    out += c;
                }
        }
    }
    out += '"';
}

inline void dump(const Value &v, std::string &out, int indent = 2, int depth = 0) {
    const std::string pad = indent > 0 ? "\n" + std::string((size_t)(indent * (depth + 1)), ' ') : "";
    const std::string pad_end = indent > 0 ? "\n" + std::string((size_t)(indent * depth), ' ') : "";
    switch (v.kind) {
        case Value::Null: out += "null"; break;
        case Value::Bool: out += v.b ? "true" : "false"; break;
        case Value::Number: out += v.text; break;
        case Value::String: dump_string(v.text, out); break;
        case Value::Array:
            if (v.items.empty()) { out += "[]"; break; }
            out += '[';
            for (size_t i = 0; i < v.items.size(); ++i) {
                if (i) out += ',';
                out += pad;
                dump(v.items[i], out, indent, depth + 1);
            }
            out += pad_end;
            out += ']';
            break;
        case Value::Object:
            if (v.members.empty()) { out += "{}"; break; }
            out += '{';
            for (size_t i = 0; i < v.members.size(); ++i) {
                if (i) out += ',';
                out += pad;
                dump_string(v.members[i].first, out);
                out += indent > 0 ? ": " : ":";
                dump(v.members[i].second, out, indent, depth + 1);
            }
            out += pad_end;
            out += '}';
            break;
    }
}

}  // namespace posjson
