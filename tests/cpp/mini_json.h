// A small JSON reader and writer for the C++ test programs: enough for the
// fixture files under tests/golden (objects, arrays, strings without escapes
// beyond \" \\ \/ \n \t, numbers, true / false / null).
#pragma once

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mini_json {

struct Value {
    enum class Type { Null, Bool, Number, String, Array, Object } type = Type::Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool is_null() const { return type == Type::Null; }
    bool has(const std::string& k) const { return type == Type::Object && obj.count(k) && !obj.at(k).is_null(); }
    const Value& operator[](const std::string& k) const {
        if (type != Type::Object) throw std::runtime_error("json: not an object (key " + k + ")");
        auto it = obj.find(k);
        if (it == obj.end()) throw std::runtime_error("json: missing key " + k);
        return it->second;
    }
    const Value& operator[](size_t i) const {
        if (type != Type::Array || i >= arr.size()) throw std::runtime_error("json: bad array index");
        return arr[i];
    }
    size_t size() const { return type == Type::Array ? arr.size() : obj.size(); }
    double d() const {
        if (type != Type::Number) throw std::runtime_error("json: not a number");
        return num;
    }
    float f() const { return (float)d(); }
    uint32_t u() const { return (uint32_t)d(); }
    int i() const { return (int)d(); }
    const std::string& s() const {
        if (type != Type::String) throw std::runtime_error("json: not a string");
        return str;
    }
    bool boolean() const { return type == Type::Bool ? b : false; }
};

class Parser {
   public:
    explicit Parser(const std::string& text) : t_(text) {}
    Value parse() {
        Value v = value();
        ws();
        if (p_ != t_.size()) err("trailing characters");
        return v;
    }

   private:
    const std::string& t_;
    size_t p_ = 0;

    [[noreturn]] void err(const char* what) {
        throw std::runtime_error(std::string("json parse error: ") + what + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) ++p_;
    }
    bool eat(char c) {
        ws();
        if (p_ < t_.size() && t_[p_] == c) {
            ++p_;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) err("unexpected character");
    }
    std::string string_body() {
        expect('"');
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c == '\\') {
                if (p_ >= t_.size()) err("bad escape");
                char e = t_[p_++];
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    default: err("unsupported escape");
                }
            } else {
                out += c;
            }
        }
        if (p_ >= t_.size()) err("unterminated string");
        ++p_;
        return out;
    }
    Value value() {
        ws();
        if (p_ >= t_.size()) err("unexpected end");
        Value v;
        const char c = t_[p_];
        if (c == '{') {
            ++p_;
            v.type = Value::Type::Object;
            if (eat('}')) return v;
            do {
                ws();
                std::string k = string_body();
                expect(':');
                v.obj[k] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            ++p_;
            v.type = Value::Type::Array;
            if (eat(']')) return v;
            do v.arr.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.type = Value::Type::String;
            v.str = string_body();
        } else if (t_.compare(p_, 4, "true") == 0) {
            p_ += 4;
            v.type = Value::Type::Bool;
            v.b = true;
        } else if (t_.compare(p_, 5, "false") == 0) {
            p_ += 5;
            v.type = Value::Type::Bool;
        } else if (t_.compare(p_, 4, "null") == 0) {
            p_ += 4;
        } else {
            const char* start = t_.c_str() + p_;
            char* end = nullptr;
            v.num = std::strtod(start, &end);
            if (end == start) err("bad value");
            v.type = Value::Type::Number;
            p_ += (size_t)(end - start);
        }
        return v;
    }
};

inline Value load(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return Parser(ss.str()).parse();
}

// Writer: numbers with 17 significant digits (doubles round-trip exactly).
inline std::string num(double x) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    return buf;
}
inline std::string quote(const std::string& s) {
    std::string out = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') out += '\\';
        out += c;
    }
    return out + "\"";
}

}  // namespace mini_json
