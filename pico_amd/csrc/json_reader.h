// The JSON reader of scene files (host code): the grammar the reference's
// scene loader accepts, jsoncpp's `std::istream >> Json::Value` with the
// default CharReaderBuilder settings (Source/Core/Scene.cpp:43-47; jsoncpp
// json_reader.cpp, OurReader): RFC 8259 plus
//   - // and /* */ comments wherever a token is read (not between a member
//     name and its ':'; a comment after a value lets any following token act
//     as the separator, as jsoncpp's readObject does);
//   - trailing commas in objects and arrays;
//   - a UTF-8 byte-order mark, and any text after the top-level value;
//   - lenient numbers: an optional '+' or '-', digits, an optional fraction
//     and exponent, each part possibly empty ("-" alone is 0, "1." is 1);
//     plain integers are exact, other forms go through std::istream >> double
//     (so "1e" is an error);
//   - raw control characters inside strings; a lone low surrogate is kept,
//     a high surrogate takes the next \u escape as its pair whatever it is,
//     and is an error without one.
// Object members are kept sorted by name (jsoncpp keeps them in a std::map),
// a repeated name keeps the last value.
#pragma once
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace pico_json {

struct Value {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    // a number: a signed or an unsigned 64-bit integer, or a double
    enum NumKind { Int, UInt, Real } nk = Real;
    long long i = 0;
    unsigned long long u = 0;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    // Json::Value::asFloat / asDouble of a number
    float as_float() const { return nk == Int ? (float)i : nk == UInt ? (float)u : (float)num; }
    double as_double() const { return nk == Int ? (double)i : nk == UInt ? (double)u : num; }

    const Value* get(const char* k) const {
        if (kind != Obj) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct Error {
    std::string msg;
};

class Reader {
public:
    Reader(const std::string& text, const std::string& file) : s_(text), file_(file) {}

    Value parse() {
        if (s_.compare(0, 3, "\xEF\xBB\xBF") == 0) i_ = 3;
        Value v;
        value(v, 0);
        return v;  // text after the top-level value is not read
    }

private:
    enum Tok { End, ObjB, ObjE, ArrB, ArrE, Str, Num, True, False, Null, Comma, Colon, Comment, Bad };
    const std::string& s_;
    std::string file_;
    size_t i_ = 0;
    size_t tb_ = 0, te_ = 0;  // the last token's text

    [[noreturn]] void fail(const std::string& what, size_t at) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < at && k < s_.size(); ++k) {
            if (s_[k] == '\n') { ++line; col = 1; } else { ++col; }
        }
        throw Error{file_ + ":" + std::to_string(line) + ":" + std::to_string(col) + ": " + what};
    }
    char peek() const { return i_ < s_.size() ? s_[i_] : '\0'; }
    void spaces() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\r' || s_[i_] == '\n')) ++i_;
    }
    bool match(const char* w) {
        const size_t n = std::strlen(w);
        if (s_.compare(i_, n, w) != 0) return false;
        i_ += n;
        return true;
    }
    // one token, spaces before it skipped (comments are tokens)
    Tok token() {
        spaces();
        tb_ = i_;
        if (i_ >= s_.size()) { te_ = i_; return End; }
        const char c = s_[i_++];
        Tok t = Bad;
        switch (c) {
            case '{': t = ObjB; break;
            case '}': t = ObjE; break;
            case '[': t = ArrB; break;
            case ']': t = ArrE; break;
            case ',': t = Comma; break;
            case ':': t = Colon; break;
            case '"': {
                char d = 0;
                while (i_ < s_.size()) {
                    d = s_[i_++];
                    if (d == '\\') { if (i_ < s_.size()) ++i_; }
                    else if (d == '"') break;
                }
                t = d == '"' && i_ > tb_ + 1 ? Str : Bad;
                break;
            }
            case '/': {
                const char d = peek();
                if (d == '*') {
                    // up to a '*' followed by '/', or to the last character,
                    // which must then be the '/' ("/*/" is a comment)
                    ++i_;
                    while (i_ + 1 < s_.size()) {
                        const char e = s_[i_++];
                        if (e == '*' && s_[i_] == '/') break;
                    }
                    const char last = i_ < s_.size() ? s_[i_++] : '\0';
                    t = last == '/' ? Comment : Bad;
                } else if (d == '/') {
                    while (i_ < s_.size() && s_[i_] != '\n' && s_[i_] != '\r') ++i_;
                    if (i_ < s_.size()) {
                        if (s_[i_] == '\r' && i_ + 1 < s_.size() && s_[i_ + 1] == '\n') ++i_;
                        ++i_;
                    }
                    t = Comment;
                } else {
                    if (i_ < s_.size()) ++i_;
                    t = Bad;
                }
                break;
            }
            case 't': t = match("rue") ? True : Bad; break;
            case 'f': t = match("alse") ? False : Bad; break;
            case 'n': t = match("ull") ? Null : Bad; break;
            case '-':
            case '+':
                if (peek() == 'I') { ++i_; t = Bad; break; }  // Infinity is not allowed
                number_text();
                t = Num;
                break;
            default:
                if (c >= '0' && c <= '9') {
                    number_text();
                    t = Num;
                }
                break;
        }
        te_ = i_;
        return t;
    }
    void number_text() {
        auto digits = [&] { while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_; };
        digits();
        if (peek() == '.') { ++i_; digits(); }
        if (peek() == 'e' || peek() == 'E') {
            ++i_;
            if (peek() == '+' || peek() == '-') ++i_;
            digits();
        }
    }
    Tok token_after_comments() {
        Tok t;
        do t = token(); while (t == Comment);
        return t;
    }
    void value(Value& v, int depth) {
        if (depth >= 1000) fail("nesting too deep", i_);
        const Tok t = token_after_comments();
        switch (t) {
            case ObjB: object(v, depth); return;
            case ArrB: array(v, depth); return;
            case Num: v.kind = Value::Num; number(v); return;
            case Str: v.kind = Value::Str; v.str = string(tb_, te_); return;
            case True: v.kind = Value::Bool; v.b = true; return;
            case False: v.kind = Value::Bool; v.b = false; return;
            case Null: v.kind = Value::Null; return;
            default: fail("a value, object or array was expected", tb_);
        }
    }
    void object(Value& v, int depth) {
        v.kind = Value::Obj;
        for (;;) {
            Tok t = token_after_comments();
            if (t == ObjE) return;  // empty object, or after a trailing comma
            if (t != Str) fail("an object member name or '}' was expected", tb_);
            const std::string name = string(tb_, te_);
            if (token() != Colon) fail("':' was expected after an object member name", tb_);
            Value& m = v.obj[name];
            m = Value{};
            value(m, depth + 1);
            t = token();
            if (t != ObjE && t != Comma && t != Comment) fail("',' or '}' was expected in an object", tb_);
            while (t == Comment) t = token();
            if (t == ObjE) return;
        }
    }
    void array(Value& v, int depth) {
        v.kind = Value::Arr;
        for (;;) {
            spaces();
            if (peek() == ']') { ++i_; return; }  // empty array, or after a trailing comma
            v.arr.emplace_back();
            value(v.arr.back(), depth + 1);
            Tok t = token_after_comments();
            if (t == ArrE) return;
            if (t != Comma) fail("',' or ']' was expected in an array", tb_);
        }
    }
    void number(Value& v) {
        // digits only (after a '-'): an exact 64-bit integer; anything else,
        // and integers out of range, as std::istream >> double reads the token
        size_t k = tb_;
        const bool neg = s_[k] == '-';
        if (neg) ++k;
        bool plain = true;
        unsigned long long u = 0;
        for (size_t j = k; j < te_; ++j) {
            const char c = s_[j];
            if (c < '0' || c > '9') { plain = false; break; }
            const unsigned d = (unsigned)(c - '0');
            const unsigned long long lim = neg ? 922337203685477580ull : 1844674407370955161ull;
            const unsigned last = neg ? 8u : 5u;
            if (u > lim || (u == lim && (j + 1 != te_ || d > last))) { plain = false; break; }
            u = u * 10 + d;
        }
        if (plain) {
            if (neg) {
                v.nk = Value::Int;
                v.i = -(long long)(u / 10) * 10 - (long long)(u % 10);
            } else if (u <= 9223372036854775807ull) {
                v.nk = Value::Int;
                v.i = (long long)u;
            } else {
                v.nk = Value::UInt;
                v.u = u;
            }
            v.num = v.as_double();
            return;
        }
        std::istringstream is(s_.substr(tb_, te_ - tb_));
        double x = 0.0;
        if (!(is >> x)) fail("'" + s_.substr(tb_, te_ - tb_) + "' is not a number", tb_);
        v.nk = Value::Real;
        v.num = x;
    }
    static void put_utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) o += (char)cp;
        else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4(size_t& k, size_t end) const {
        if (end - k < 4) fail("a \\u escape needs four hexadecimal digits", k);
        uint32_t v = 0;
        for (int n = 0; n < 4; ++n) {
            const char h = s_[k++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
            else fail("bad hexadecimal digit in a \\u escape", k - 1);
        }
        return v;
    }
    // the text of a string token [b, e) with its quotes
    std::string string(size_t b, size_t e) const {
        std::string o;
        size_t k = b + 1;
        const size_t end = e - 1;
        while (k != end) {
            const char c = s_[k++];
            if (c != '\\') { o += c; continue; }
            if (k == end) fail("empty escape sequence in a string", k);
            const char x = s_[k++];
            switch (x) {
                case '"': o += '"'; break;
                case '/': o += '/'; break;
                case '\\': o += '\\'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4(k, end);
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        // the next escape completes the pair, whatever its value
                        if (end - k < 6 || s_[k] != '\\' || s_[k + 1] != 'u')
                            fail("a high surrogate needs a second \\u escape", k);
                        k += 2;
                        const uint32_t lo = hex4(k, end);
                        cp = 0x10000 + ((cp & 0x3FF) << 10) + (lo & 0x3FF);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("bad escape sequence in a string", k - 1);
            }
        }
        return o;
    }
};

}  // namespace pico_json
