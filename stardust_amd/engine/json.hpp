// Minimal JSON reader for scene files (objects, arrays, numbers, strings,
// true/false/null). Errors throw std::runtime_error with the byte offset.
#ifndef STARDUST_AMD_ENGINE_JSON_HPP
#define STARDUST_AMD_ENGINE_JSON_HPP

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace STARDUST {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool has(const std::string& k) const { return kind == Object && obj.count(k) != 0; }
    const Value& operator[](const std::string& k) const {
        if (kind != Object) throw std::runtime_error("json: not an object (looking up \"" + k + "\")");
        auto it = obj.find(k);
        if (it == obj.end()) throw std::runtime_error("json: missing key \"" + k + "\"");
        return it->second;
    }
    const Value& operator[](size_t i) const {
        if (kind != Array || i >= arr.size()) throw std::runtime_error("json: array index out of range");
        return arr[i];
    }
    double number() const {
        if (kind != Number) throw std::runtime_error("json: value is not a number");
        return num;
    }
    const std::string& string() const {
        if (kind != String) throw std::runtime_error("json: value is not a string");
        return str;
    }
};

class Parser {
public:
    explicit Parser(const std::string& text) : s(text) {}
    Value parse() {
        Value v = value();
        ws();
        if (i != s.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    bool word(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    std::string string_lit() {
        expect('"');
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                const char e = s[i++];
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    default: fail("unsupported escape");
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    Value value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        Value v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = Value::Object;
            if (eat('}')) return v;
            do {
                ws();
                const std::string k = string_lit();
                expect(':');
                v.obj[k] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            ++i;
            v.kind = Value::Array;
            if (eat(']')) return v;
            do {
                v.arr.push_back(value());
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string_lit();
        } else if (word("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (word("false")) {
            v.kind = Value::Bool;
        } else if (word("null")) {
            v.kind = Value::Null;
        } else {
            const char* b = s.c_str() + i;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("bad value");
            i += (size_t)(e - b);
            v.kind = Value::Number;
        }
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace STARDUST

#endif
