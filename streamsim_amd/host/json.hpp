// json.hpp -- a small JSON reader for facet_engine's project.json input:
// a value tree (null, bool, number, string, array, object with members in
// file order) and the lookups of the reference's ProjectJson helpers
// (models/histream/src/base/projectjson.cpp:155-187): a missing member or
// one of another type gives the fallback, numbers may be given as strings.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace facetcli::json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Value> items;                           // Array
    std::vector<std::pair<std::string, Value>> members;  // Object

    bool isObject() const { return kind == Object; }
    bool isArray() const { return kind == Array; }
    // member of an object (the first of that name), or nullptr
    const Value* get(const std::string& key) const {
        if (kind != Object) return nullptr;
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

namespace detail {
struct Parser {
    const std::string& t;
    size_t i = 0;
    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("Invalid project.json: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < t.size() && (t[i] == ' ' || t[i] == '\t' || t[i] == '\n' || t[i] == '\r')) ++i;
    }
    bool lit(const char* w) {
        size_t k = 0;
        while (w[k] && i + k < t.size() && t[i + k] == w[k]) ++k;
        if (w[k]) return false;
        i += k;
        return true;
    }
    static void utf8(std::string& out, uint32_t c) {
        if (c < 0x80) {
            out += char(c);
        } else if (c < 0x800) {
            out += char(0xC0 | (c >> 6));
            out += char(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            out += char(0xE0 | (c >> 12));
            out += char(0x80 | ((c >> 6) & 0x3F));
            out += char(0x80 | (c & 0x3F));
        } else {
            out += char(0xF0 | (c >> 18));
            out += char(0x80 | ((c >> 12) & 0x3F));
            out += char(0x80 | ((c >> 6) & 0x3F));
            out += char(0x80 | (c & 0x3F));
        }
    }
    uint32_t hex4() {
        if (i + 4 > t.size()) fail("short \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = t[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        if (t[i] != '"') fail("expected a string");
        ++i;
        std::string out;
        while (true) {
            if (i >= t.size()) fail("unterminated string");
            const char c = t[i++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= t.size()) fail("unterminated escape");
            const char e = t[i++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t c1 = hex4();
                    if (c1 >= 0xD800 && c1 < 0xDC00 && i + 1 < t.size() && t[i] == '\\' && t[i + 1] == 'u') {
                        i += 2;
                        const uint32_t c2 = hex4();
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    utf8(out, c1);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    Value value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (i >= t.size()) fail("unexpected end");
        Value v;
        const char c = t[i];
        if (c == '{') {
            v.kind = Value::Object;
            ++i;
            ws();
            if (i < t.size() && t[i] == '}') {
                ++i;
                return v;
            }
            while (true) {
                ws();
                if (i >= t.size()) fail("unterminated object");
                std::string key = str();
                ws();
                if (i >= t.size() || t[i] != ':') fail("expected ':'");
                ++i;
                v.members.emplace_back(std::move(key), value(depth + 1));
                ws();
                if (i < t.size() && t[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < t.size() && t[i] == '}') {
                    ++i;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            ++i;
            ws();
            if (i < t.size() && t[i] == ']') {
                ++i;
                return v;
            }
            while (true) {
                v.items.push_back(value(depth + 1));
                ws();
                if (i < t.size() && t[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < t.size() && t[i] == ']') {
                    ++i;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.s = str();
            return v;
        }
        if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (lit("null")) return v;
        const char* start = t.c_str() + i;
        char* end = nullptr;
        const double x = std::strtod(start, &end);
        if (end == start) fail("unexpected character");
        v.kind = Value::Number;
        v.n = x;
        // integers keep their text for string() of a number member
        v.s.assign(start, size_t(end - start));
        i += size_t(end - start);
        return v;
    }
};
}  // namespace detail

inline Value parse(const std::string& text) {
    detail::Parser p{text};
    Value v = p.value(0);
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

// ProjectJson::string: a string member, another non-null member's text, or
// the fallback
inline std::string str(const Value& obj, const char* key, const std::string& fallback = "") {
    const Value* v = obj.get(key);
    if (!v || v->kind == Value::Null) return fallback;
    if (v->kind == Value::String || v->kind == Value::Number) return v->s;
    if (v->kind == Value::Bool) return v->b ? "true" : "false";
    return fallback;
}

// ProjectJson::number: a finite number member or a number written as a string
inline float num(const Value& obj, const char* key, float fallback) {
    const Value* v = obj.get(key);
    if (!v) return fallback;
    float x = fallback;
    if (v->kind == Value::Number) {
        x = float(v->n);
    } else if (v->kind == Value::String) {
        char* end = nullptr;
        const float y = std::strtof(v->s.c_str(), &end);
        if (end == v->s.c_str()) return fallback;
        x = y;
    } else {
        return fallback;
    }
    return std::isfinite(x) ? x : fallback;
}

inline int integer(const Value& obj, const char* key, int fallback) {
    return int(std::lround(num(obj, key, float(fallback))));
}

// ProjectJson::boolean: a bool, an integer (!= 0), or the strings "true" / "1"
inline bool boolean(const Value& obj, const char* key, bool fallback) {
    const Value* v = obj.get(key);
    if (!v) return fallback;
    if (v->kind == Value::Bool) return v->b;
    if (v->kind == Value::Number && v->n == std::floor(v->n)) return v->n != 0.0;
    if (v->kind == Value::String) return v->s == "true" || v->s == "1";
    return fallback;
}

inline const Value& member(const Value& obj, const char* key, Value::Kind kind) {
    static const Value empty;
    const Value* v = obj.get(key);
    return v && v->kind == kind ? *v : empty;
}

// ProjectJson::numbers: an array of numbers, or numbers in a string
// separated by commas, semicolons or spaces
inline std::vector<float> numbers(const Value* v) {
    std::vector<float> out;
    if (!v) return out;
    if (v->kind == Value::Array) {
        for (const Value& x : v->items)
            if (x.kind == Value::Number) out.push_back(float(x.n));
        return out;
    }
    std::string text = v->kind == Value::String ? v->s : v->kind == Value::Number ? v->s : std::string{};
    for (char& c : text)
        if (c == ';' || c == ' ') c = ',';
    size_t p = 0;
    while (p <= text.size()) {
        const size_t q = text.find(',', p);
        const std::string tok = text.substr(p, q == std::string::npos ? std::string::npos : q - p);
        if (!tok.empty()) {
            char* end = nullptr;
            const float x = std::strtof(tok.c_str(), &end);
            if (end != tok.c_str() && std::isfinite(x)) out.push_back(x);
        }
        if (q == std::string::npos) break;
        p = q + 1;
    }
    return out;
}

}  // namespace facetcli::json
