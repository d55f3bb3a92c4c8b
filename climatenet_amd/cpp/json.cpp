#include "json.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace json {

Value Value::make_bool(bool b) {
    Value v;
    v.type_ = Type::Bool;
    v.bool_ = b;
    return v;
}
Value Value::make_int(int64_t i) {
    Value v;
    v.type_ = Type::Int;
    v.int_ = i;
    return v;
}
Value Value::make_double(double d) {
    Value v;
    v.type_ = Type::Double;
    v.dbl_ = d;
    return v;
}
Value Value::make_string(std::string s) {
    Value v;
    v.type_ = Type::String;
    v.str_ = std::move(s);
    return v;
}
Value Value::make_array() {
    Value v;
    v.type_ = Type::Array;
    return v;
}
Value Value::make_object() {
    Value v;
    v.type_ = Type::Object;
    return v;
}

bool Value::GetBool() const {
    if (type_ != Type::Bool) throw std::invalid_argument("JSON value is not a boolean");
    return bool_;
}
int Value::GetInt() const {
    if (!IsInt()) throw std::invalid_argument("JSON value is not an integer");
    return (int)int_;
}
unsigned Value::GetUint() const {
    if (!IsUint()) throw std::invalid_argument("JSON value is not an unsigned integer");
    return (unsigned)int_;
}
double Value::GetDouble() const {
    if (type_ == Type::Double) return dbl_;
    if (type_ == Type::Int) return (double)int_;
    throw std::invalid_argument("JSON value is not a number");
}
const std::string& Value::GetString() const {
    if (type_ != Type::String) throw std::invalid_argument("JSON value is not a string");
    return str_;
}
size_t Value::Size() const {
    if (type_ == Type::Array) return arr_.size();
    if (type_ == Type::Object) return obj_.size();
    return 0;
}
const Value& Value::operator[](size_t i) const {
    if (type_ != Type::Array || i >= arr_.size()) throw std::invalid_argument("JSON array index out of range");
    return arr_[i];
}
bool Value::HasMember(const std::string& key) const { return FindMember(key) != nullptr; }
const Value* Value::FindMember(const std::string& key) const {
    if (type_ != Type::Object) return nullptr;
    for (const auto& kv : obj_)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const char* what) const {
        throw std::invalid_argument(std::string("JSON parse error at offset ") + std::to_string(i) + ": " + what);
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool lit(const char* w) {
        size_t k = 0;
        while (w[k]) {
            if (i + k >= s.size() || s[i + k] != w[k]) return false;
            ++k;
        }
        i += k;
        return true;
    }
    void append_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (i + 4 > s.size()) fail("truncated \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        if (s[i] != '"') fail("expected string");
        ++i;
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            char c = s[i++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            char e = s[i++];
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                        i += 2;
                        unsigned lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    append_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    Value number() {
        size_t start = i;
        bool is_double = false;
        if (s[i] == '-') ++i;
        if (i >= s.size() || !std::isdigit((unsigned char)s[i])) fail("bad number");
        if (s[i] == '0') ++i;
        else while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        if (i < s.size() && s[i] == '.') {
            is_double = true;
            ++i;
            if (i >= s.size() || !std::isdigit((unsigned char)s[i])) fail("bad fraction");
            while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            is_double = true;
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
            if (i >= s.size() || !std::isdigit((unsigned char)s[i])) fail("bad exponent");
            while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        }
        std::string tok = s.substr(start, i - start);
        if (!is_double) {
            errno = 0;
            long long v = std::strtoll(tok.c_str(), nullptr, 10);
            if (errno == 0) return Value::make_int(v);
        }
        return Value::make_double(std::strtod(tok.c_str(), nullptr));
    }
    Value value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            ++i;
            Value obj = Value::make_object();
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return obj;
            }
            while (true) {
                ws();
                if (i >= s.size()) fail("unterminated object");
                std::string key = string();
                ws();
                if (i >= s.size() || s[i] != ':') fail("expected ':'");
                ++i;
                Value v = value(depth + 1);
                obj.members().emplace_back(std::move(key), std::move(v));
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    return obj;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i;
            Value arr = Value::make_array();
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return arr;
            }
            while (true) {
                arr.array().push_back(value(depth + 1));
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    return arr;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') return Value::make_string(string());
        if (lit("true")) return Value::make_bool(true);
        if (lit("false")) return Value::make_bool(false);
        if (lit("null")) return Value();
        return number();
    }
};

}  // namespace

Value parse(const std::string& text) {
    Parser p{text};
    Value v = p.value(0);
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

Value parse_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::invalid_argument("Couldn't open file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace json
