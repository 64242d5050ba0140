// json.hpp -- small JSON reader / writer for the apps' input files and result
// metadata (the reference uses nlohmann::json: include/run.hpp:51-126,
// include/save.hpp:160-190).
//
// Reading: the full JSON grammar (objects keep their key order).  Writing
// follows nlohmann's dump(4) layout so result files read the same: 4-space
// indentation, "key": value, empty containers as [] / {}, integers as
// integers, doubles in the shortest form that round-trips, printed in fixed
// notation for decimal exponents -4 < n <= 15 (with ".0" on integral values)
// and as d.ddde+XX otherwise.
#pragma once
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hfxh {

struct Json {
    enum Type { Null, Bool, Int, Real, String, Array, Object };
    Type t = Null;
    bool b = false;
    long long i = 0;
    double d = 0.0;
    std::string s;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    Json() = default;
    Json(bool v) : t(Bool), b(v) {}
    Json(int v) : t(Int), i(v) {}
    Json(long long v) : t(Int), i(v) {}
    Json(size_t v) : t(Int), i((long long)v) {}
    Json(double v) : t(Real), d(v) {}
    Json(const char* v) : t(String), s(v) {}
    Json(std::string v) : t(String), s(std::move(v)) {}
    template <class T>
    static Json array(const std::vector<T>& v) {
        Json j;
        j.t = Array;
        for (const auto& x : v) j.arr.push_back(Json(x));
        return j;
    }
    static Json object() {
        Json j;
        j.t = Object;
        return j;
    }

    bool is_object() const { return t == Object; }
    bool is_array() const { return t == Array; }
    bool is_number() const { return t == Int || t == Real; }
    bool is_string() const { return t == String; }
    size_t size() const { return t == Array ? arr.size() : t == Object ? obj.size() : 0; }

    bool contains(const std::string& k) const {
        if (t != Object) return false;
        for (const auto& kv : obj)
            if (kv.first == k) return true;
        return false;
    }
    const Json& at(const std::string& k) const {
        if (t != Object) throw std::runtime_error("JSON value is not an object (looking for key '" + k + "')");
        for (const auto& kv : obj)
            if (kv.first == k) return kv.second;
        throw std::runtime_error("JSON key not found: " + k);
    }
    const Json& at(size_t idx) const {
        if (t != Array) throw std::runtime_error("JSON value is not an array");
        if (idx >= arr.size()) throw std::runtime_error("JSON array index out of range");
        return arr[idx];
    }
    // insertion-ordered assignment (replaces an existing key)
    Json& operator[](const std::string& k) {
        if (t == Null) t = Object;
        if (t != Object) throw std::runtime_error("JSON value is not an object");
        for (auto& kv : obj)
            if (kv.first == k) return kv.second;
        obj.emplace_back(k, Json());
        return obj.back().second;
    }

    double as_double() const {
        if (t == Int) return (double)i;
        if (t == Real) return d;
        throw std::runtime_error("JSON type error: expected a number");
    }
    int as_int() const {
        if (t == Int) return (int)i;
        if (t == Real) return (int)d;
        throw std::runtime_error("JSON type error: expected a number");
    }
    const std::string& as_string() const {
        if (t != String) throw std::runtime_error("JSON type error: expected a string");
        return s;
    }
    bool as_bool() const {
        if (t != Bool) throw std::runtime_error("JSON type error: expected a boolean");
        return b;
    }
    // nlohmann's value(key, default): the default when the key is absent
    double value(const std::string& k, double def) const { return contains(k) ? at(k).as_double() : def; }
    int value(const std::string& k, int def) const { return contains(k) ? at(k).as_int() : def; }
    std::string value(const std::string& k, const char* def) const { return contains(k) ? at(k).as_string() : def; }
};

namespace json_detail {

struct Parser {
    const std::string& src;
    size_t p = 0;

    [[noreturn]] void fail(const std::string& what) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < p && k < src.size(); ++k) {
            if (src[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ", column " +
                                 std::to_string(col) + ": " + what);
    }
    void ws() {
        while (p < src.size() && (src[p] == ' ' || src[p] == '\t' || src[p] == '\n' || src[p] == '\r')) ++p;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (src.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out.push_back((char)cp);
        } else if (cp < 0x800) {
            out.push_back((char)(0xC0 | (cp >> 6)));
            out.push_back((char)(0x80 | (cp & 0x3F)));
        } else if (cp < 0x10000) {
            out.push_back((char)(0xE0 | (cp >> 12)));
            out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            out.push_back((char)(0x80 | (cp & 0x3F)));
        } else {
            out.push_back((char)(0xF0 | (cp >> 18)));
            out.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
            out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            out.push_back((char)(0x80 | (cp & 0x3F)));
        }
    }
    unsigned hex4() {
        if (p + 4 > src.size()) fail("truncated \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = src[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        if (src[p] != '"') fail("expected a string");
        ++p;
        std::string out;
        while (true) {
            if (p >= src.size()) fail("unterminated string");
            const char c = src[p++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (p >= src.size()) fail("unterminated escape");
            const char e = src[p++];
            switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && lit("\\u")) {
                        const unsigned lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    Json number() {
        const size_t b = p;
        if (src[p] == '-') ++p;
        bool real = false;
        while (p < src.size()) {
            const char c = src[p];
            if (c >= '0' && c <= '9') {
                ++p;
            } else if (c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-') {
                real = true;
                ++p;
            } else {
                break;
            }
        }
        const std::string tok = src.substr(b, p - b);
        if (tok.empty() || tok == "-") fail("bad number");
        Json j;
        if (!real) {
            long long v = 0;
            auto r = std::from_chars(tok.data(), tok.data() + tok.size(), v);
            if (r.ec == std::errc() && r.ptr == tok.data() + tok.size()) {
                j.t = Json::Int;
                j.i = v;
                return j;
            }
        }
        char* end = nullptr;
        const double v = std::strtod(tok.c_str(), &end);
        if (end != tok.c_str() + tok.size()) fail("bad number '" + tok + "'");
        j.t = Json::Real;
        j.d = v;
        return j;
    }
    Json value() {
        ws();
        if (p >= src.size()) fail("unexpected end of input");
        const char c = src[p];
        Json j;
        if (c == '{') {
            ++p;
            j.t = Json::Object;
            ws();
            if (p < src.size() && src[p] == '}') {
                ++p;
                return j;
            }
            while (true) {
                ws();
                if (p >= src.size()) fail("unexpected end of input");
                std::string k = str();
                ws();
                if (p >= src.size() || src[p] != ':') fail("expected ':'");
                ++p;
                Json v = value();
                bool replaced = false;
                for (auto& kv : j.obj)
                    if (kv.first == k) {
                        kv.second = std::move(v);
                        replaced = true;
                        break;
                    }
                if (!replaced) j.obj.emplace_back(std::move(k), std::move(v));
                ws();
                if (p < src.size() && src[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < src.size() && src[p] == '}') {
                    ++p;
                    return j;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            j.t = Json::Array;
            ws();
            if (p < src.size() && src[p] == ']') {
                ++p;
                return j;
            }
            while (true) {
                j.arr.push_back(value());
                ws();
                if (p < src.size() && src[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < src.size() && src[p] == ']') {
                    ++p;
                    return j;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') return Json(str());
        if (lit("true")) return Json(true);
        if (lit("false")) return Json(false);
        if (lit("null")) return Json();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail(std::string("unexpected character '") + c + "'");
    }
};

// nlohmann-style double formatting from the shortest round-trip digits.
inline std::string fmt_real(double v) {
    if (!std::isfinite(v)) return "null";
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    std::string sci(buf, r.ptr);   // [-]d[.ddd]e[+-]XX
    std::string out;
    if (sci[0] == '-') {
        out.push_back('-');
        sci.erase(0, 1);
    }
    const size_t epos = sci.find('e');
    const int e10 = std::atoi(sci.c_str() + epos + 1);
    std::string digits;
    for (size_t k = 0; k < epos; ++k)
        if (sci[k] != '.') digits.push_back(sci[k]);
    const int k = (int)digits.size();
    const int n = e10 + 1;   // value = 0.digits x 10^n
    if (k <= n && n <= 15) {
        out += digits + std::string(n - k, '0') + ".0";
    } else if (0 < n && n <= 15) {
        out += digits.substr(0, n) + "." + digits.substr(n);
    } else if (-4 < n && n <= 0) {
        out += "0." + std::string(-n, '0') + digits;
    } else {
        out += digits.substr(0, 1);
        if (k > 1) out += "." + digits.substr(1);
        const int ex = n - 1;
        out += ex < 0 ? "e-" : "e+";
        const int a = ex < 0 ? -ex : ex;
        if (a < 10) out += "0";
        out += std::to_string(a);
    }
    return out;
}

inline void esc(std::string& out, const std::string& s) {
    out.push_back('"');
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof(b), "\\u%04x", c);
                    out += b;
                } else {
                    out.push_back((char)c);
                }
        }
    }
    out.push_back('"');
}

inline void dump(std::string& out, const Json& j, int indent, int level) {
    const std::string pad((size_t)indent * (level + 1), ' '), pad0((size_t)indent * level, ' ');
    switch (j.t) {
        case Json::Null: out += "null"; break;
        case Json::Bool: out += j.b ? "true" : "false"; break;
        case Json::Int: out += std::to_string(j.i); break;
        case Json::Real: out += fmt_real(j.d); break;
        case Json::String: esc(out, j.s); break;
        case Json::Array:
            if (j.arr.empty()) {
                out += "[]";
                break;
            }
            out += "[\n";
            for (size_t k = 0; k < j.arr.size(); ++k) {
                out += pad;
                dump(out, j.arr[k], indent, level + 1);
                out += k + 1 < j.arr.size() ? ",\n" : "\n";
            }
            out += pad0 + "]";
            break;
        case Json::Object:
            if (j.obj.empty()) {
                out += "{}";
                break;
            }
            out += "{\n";
            for (size_t k = 0; k < j.obj.size(); ++k) {
                out += pad;
                esc(out, j.obj[k].first);
                out += ": ";
                dump(out, j.obj[k].second, indent, level + 1);
                out += k + 1 < j.obj.size() ? ",\n" : "\n";
            }
            out += pad0 + "}";
            break;
    }
}

}  // namespace json_detail

inline Json json_parse(const std::string& text) {
    json_detail::Parser ps{text};
    Json j = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.fail("trailing characters");
    return j;
}

inline Json json_read_file(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Could not open JSON file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return json_parse(ss.str());
}

inline std::string json_dump(const Json& j, int indent = 4) {
    std::string out;
    json_detail::dump(out, j, indent, 0);
    return out;
}

}  // namespace hfxh
