// A small reader for the TOML subset the vehicle files use: comments,
// [table] headers, key = value with numbers, strings, booleans and (nested,
// multi-line) arrays.  Keys are flattened to "table.key".
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace crush {
namespace toml {

struct Value {
    enum Kind { Number, String, Bool, Array } kind = Number;
    double num = 0.0;
    std::string str;
    bool flag = false;
    std::vector<Value> items;
};

class Parser {
  public:
    explicit Parser(const std::string& text) : s_(text) {}

    std::map<std::string, Value> parse() {
        std::map<std::string, Value> out;
        std::string table;
        while (true) {
            skip_ws_nl();
            if (eof()) break;
            char c = peek();
            if (c == '[') {
                ++i_;
                if (peek() == '[') fail(CRUSH_E_INPUT, "toml line %d: arrays of tables are not supported", line_);
                skip_ws();
                table = key_path();
                skip_ws();
                expect(']');
                end_line();
                continue;
            }
            std::string key = key_path();
            skip_ws();
            expect('=');
            skip_ws();
            Value v = value();
            std::string full = table.empty() ? key : table + "." + key;
            if (out.count(full)) fail(CRUSH_E_INPUT, "toml line %d: duplicate key '%s'", line_, full.c_str());
            out[full] = std::move(v);
            end_line();
        }
        return out;
    }

  private:
    const std::string& s_;
    size_t i_ = 0;
    int line_ = 1;

    bool eof() const { return i_ >= s_.size(); }
    char peek() const { return eof() ? '\0' : s_[i_]; }
    void expect(char c) {
        if (peek() != c) fail(CRUSH_E_INPUT, "toml line %d: expected '%c'", line_, c);
        ++i_;
    }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\r')) ++i_;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') ++i_;
    }
    void skip_ws_nl() {
        while (!eof()) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') {
                ++i_;
                ++line_;
            } else {
                break;
            }
        }
    }
    void end_line() {
        skip_ws();
        skip_comment();
        if (!eof() && peek() != '\n') fail(CRUSH_E_INPUT, "toml line %d: trailing characters", line_);
    }
    std::string bare_key() {
        size_t b = i_;
        while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-')) ++i_;
        if (b == i_) fail(CRUSH_E_INPUT, "toml line %d: expected a key", line_);
        return s_.substr(b, i_ - b);
    }
    std::string key_path() {
        std::string k = (peek() == '"') ? string_lit() : bare_key();
        skip_ws();
        while (peek() == '.') {
            ++i_;
            skip_ws();
            k += "." + ((peek() == '"') ? string_lit() : bare_key());
            skip_ws();
        }
        return k;
    }
    std::string string_lit() {
        expect('"');
        std::string r;
        while (!eof() && peek() != '"') {
            char c = s_[i_++];
            if (c == '\n') fail(CRUSH_E_INPUT, "toml line %d: unterminated string", line_);
            if (c == '\\') {
                char e = s_[i_++];
                switch (e) {
                    case 'n': r += '\n'; break;
                    case 't': r += '\t'; break;
                    case '"': r += '"'; break;
                    case '\\': r += '\\'; break;
                    default: fail(CRUSH_E_INPUT, "toml line %d: unsupported escape", line_);
                }
            } else {
                r += c;
            }
        }
        expect('"');
        return r;
    }
    Value value() {
        Value v;
        char c = peek();
        if (c == '"') {
            v.kind = Value::String;
            v.str = string_lit();
        } else if (c == '[') {
            ++i_;
            v.kind = Value::Array;
            while (true) {
                skip_ws_nl();
                if (peek() == ']') {
                    ++i_;
                    break;
                }
                v.items.push_back(value());
                skip_ws_nl();
                if (peek() == ',') {
                    ++i_;
                    continue;
                }
                skip_ws_nl();
                expect(']');
                break;
            }
        } else if (s_.compare(i_, 4, "true") == 0) {
            v.kind = Value::Bool;
            v.flag = true;
            i_ += 4;
        } else if (s_.compare(i_, 5, "false") == 0) {
            v.kind = Value::Bool;
            i_ += 5;
        } else {
            size_t b = i_;
            while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '+' || peek() == '-' ||
                              peek() == '.' || peek() == '_'))
                ++i_;
            std::string t;
            for (size_t k = b; k < i_; ++k)
                if (s_[k] != '_') t += s_[k];
            if (t == "inf" || t == "+inf") {
                v.num = 1.0 / 0.0;
            } else if (t == "-inf") {
                v.num = -1.0 / 0.0;
            } else {
                char* end = nullptr;
                v.num = std::strtod(t.c_str(), &end);
                if (t.empty() || end != t.c_str() + t.size())
                    fail(CRUSH_E_INPUT, "toml line %d: bad value '%s'", line_, t.c_str());
            }
            v.kind = Value::Number;
        }
        return v;
    }
};

inline std::map<std::string, Value> parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace toml
}  // namespace crush
