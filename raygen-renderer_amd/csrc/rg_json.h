// rg_json.h — a small JSON reader for scene files: accepts the JSONC the
// reference's scene files use (// and /* */ comments, trailing commas) and
// keeps object properties in file order.
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rg {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;  // file order

    const Value *get(const std::string &key) const {
        for (auto it = obj.rbegin(); it != obj.rend(); ++it)
            if (it->first == key) return &it->second;
        return nullptr;
    }
    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
    bool is_number() const { return kind == Number; }
    bool is_string() const { return kind == String; }
};

class Parser {
  public:
    explicit Parser(const std::string &text) : s_(text) {}

    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string &s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char *what) {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); k++) line += s_[k] == '\n';
        throw std::runtime_error(std::string("JSON parse error at line ") + std::to_string(line) + ": " + what);
    }

    void ws() {
        for (;;) {
            while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') i_++;
                continue;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                const size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                i_ = e + 2;
                continue;
            }
            break;
        }
    }

    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            i_++;
            return true;
        }
        return false;
    }

    std::string string_lit() {
        if (!eat('"')) fail("expected string");
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\' && i_ < s_.size()) {
                const char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) fail("bad \\u escape");
                        const unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        i_++;
        return out;
    }

    // An object key: a string, or (as in the reference's JavaScript-style
    // scene files, e.g. `{ stacks: 16 }`) a bare identifier.
    std::string key() {
        ws();
        if (i_ < s_.size() && s_[i_] == '"') return string_lit();
        const size_t b = i_;
        auto ident = [](char c, bool first) {
            return std::isalpha((unsigned char)c) || c == '_' || c == '$' || (!first && std::isdigit((unsigned char)c));
        };
        while (i_ < s_.size() && ident(s_[i_], i_ == b)) i_++;
        if (i_ == b) fail("expected string");
        return s_.substr(b, i_ - b);
    }

    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            i_++;
            v.kind = Value::Object;
            for (;;) {
                if (eat('}')) break;
                std::string k = key();
                if (!eat(':')) fail("expected ':'");
                v.obj.emplace_back(std::move(k), value());
                if (eat(',')) continue;
                if (eat('}')) break;
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            i_++;
            v.kind = Value::Array;
            for (;;) {
                if (eat(']')) break;
                v.arr.push_back(value());
                if (eat(',')) continue;
                if (eat(']')) break;
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string_lit();
        } else if (s_.compare(i_, 4, "true") == 0) {
            i_ += 4;
            v.kind = Value::Bool;
            v.b = true;
        } else if (s_.compare(i_, 5, "false") == 0) {
            i_ += 5;
            v.kind = Value::Bool;
        } else if (s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
        } else {
            const char *start = s_.c_str() + i_;
            char *end = nullptr;
            v.num = std::strtod(start, &end);
            if (end == start) fail("unexpected character");
            i_ += (size_t)(end - start);
            v.kind = Value::Number;
        }
        return v;
    }
};

}  // namespace json
}  // namespace rg
