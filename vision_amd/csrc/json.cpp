#include "json.h"

#include <cstdlib>
#include <cstring>

namespace vs {

namespace {
const Json &null_json() {
    static const Json j;
    return j;
}

class Parser {
public:
    explicit Parser(const std::string &t) : s_(t), p_(0) {}

    Json parse_document() {
        skip_ws();
        Json v = parse_value();
        skip_ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t p_;

    [[noreturn]] void fail(const char *what) const {
        size_t line = 1;
        for (size_t i = 0; i < p_ && i < s_.size(); ++i) line += s_[i] == '\n';
        throw JsonError(std::string("json parse error: ") + what + " at line " + std::to_string(line));
    }

    // whitespace plus `//` and `/* */` comments (comments never start inside
    // strings because strings are consumed by parse_string)
    void skip_ws() {
        while (p_ < s_.size()) {
            char c = s_[p_];
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
                ++p_;
            } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {
                p_ += 2;
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '*') {
                p_ += 2;
                while (p_ + 1 < s_.size() && !(s_[p_] == '*' && s_[p_ + 1] == '/')) ++p_;
                if (p_ + 1 >= s_.size()) fail("unterminated comment");
                p_ += 2;
            } else {
                break;
            }
        }
    }

    Json parse_value() {
        if (p_ >= s_.size()) fail("unexpected end");
        char c = s_[p_];
        if (c == '{') return parse_object();
        if (c == '[') return parse_array();
        if (c == '"') return Json::make_string(parse_string());
        if (c == 't' && s_.compare(p_, 4, "true") == 0) { p_ += 4; return Json::make_bool(true); }
        if (c == 'f' && s_.compare(p_, 5, "false") == 0) { p_ += 5; return Json::make_bool(false); }
        if (c == 'n' && s_.compare(p_, 4, "null") == 0) { p_ += 4; return Json(); }
        if (c == '-' || c == '+' || (c >= '0' && c <= '9') || c == '.') return parse_number();
        fail("unexpected character");
    }

    Json parse_number() {
        const char *begin = s_.c_str() + p_;
        char *end = nullptr;
        double d = std::strtod(begin, &end);
        if (end == begin) fail("bad number");
        p_ += static_cast<size_t>(end - begin);
        return Json::make_number(d);
    }

    static void append_utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }

    std::string parse_string() {
        ++p_;// opening quote
        std::string out;
        while (true) {
            if (p_ >= s_.size()) fail("unterminated string");
            char c = s_[p_++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= s_.size()) fail("bad escape");
            char e = s_[p_++];
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
                    if (p_ + 4 > s_.size()) fail("bad unicode escape");
                    uint32_t cp = static_cast<uint32_t>(std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16));
                    p_ += 4;
                    append_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }

    Json parse_array() {
        ++p_;
        Json arr = Json::make_array();
        skip_ws();
        if (p_ < s_.size() && s_[p_] == ']') { ++p_; return arr; }
        while (true) {
            skip_ws();
            arr.push_back(parse_value());
            skip_ws();
            if (p_ >= s_.size()) fail("unterminated array");
            char c = s_[p_++];
            if (c == ']') break;
            if (c != ',') fail("expected , or ]");
            skip_ws();
            // tolerate a trailing comma before ]
            if (p_ < s_.size() && s_[p_] == ']') { ++p_; break; }
        }
        return arr;
    }

    Json parse_object() {
        ++p_;
        Json obj = Json::make_object();
        skip_ws();
        if (p_ < s_.size() && s_[p_] == '}') { ++p_; return obj; }
        while (true) {
            skip_ws();
            if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
            std::string key = parse_string();
            skip_ws();
            if (p_ >= s_.size() || s_[p_] != ':') fail("expected :");
            ++p_;
            skip_ws();
            obj.set(key, parse_value());
            skip_ws();
            if (p_ >= s_.size()) fail("unterminated object");
            char c = s_[p_++];
            if (c == '}') break;
            if (c != ',') fail("expected , or }");
            skip_ws();
            if (p_ < s_.size() && s_[p_] == '}') { ++p_; break; }
        }
        return obj;
    }
};
}// namespace

double Json::number() const {
    if (type_ != Type::Number) throw JsonError("json: not a number");
    return d_;
}
bool Json::boolean() const {
    if (type_ != Type::Bool) throw JsonError("json: not a bool");
    return b_;
}
const std::string &Json::str() const {
    if (type_ != Type::String) throw JsonError("json: not a string");
    return s_;
}
size_t Json::size() const {
    if (type_ == Type::Array) return arr_.size();
    if (type_ == Type::Object) return obj_.size();
    return 0;
}
const Json &Json::at(size_t i) const {
    if (type_ != Type::Array || i >= arr_.size()) throw JsonError("json: index out of range");
    return arr_[i];
}
const Json &Json::operator[](const std::string &key) const {
    if (type_ != Type::Object) return null_json();
    for (const auto &kv : obj_)
        if (kv.first == key) return kv.second;
    return null_json();
}
bool Json::contains(const std::string &key) const {
    if (type_ != Type::Object) return false;
    for (const auto &kv : obj_)
        if (kv.first == key) return true;
    return false;
}
void Json::set(const std::string &key, Json v) {
    if (type_ != Type::Object) {
        type_ = Type::Object;
        obj_.clear();
    }
    for (auto &kv : obj_) {
        if (kv.first == key) {
            kv.second = std::move(v);
            return;
        }
    }
    obj_.emplace_back(key, std::move(v));
}
std::vector<double> Json::as_numbers() const {
    std::vector<double> out;
    if (is_number()) {
        out.push_back(d_);
    } else if (is_array()) {
        for (const auto &e : arr_) {
            if (e.is_number()) out.push_back(e.d_);
            else if (e.is_array())
                for (const auto &f : e.arr_)
                    if (f.is_number()) out.push_back(f.d_);
        }
    }
    return out;
}

Json Json::parse(const std::string &text) {
    Parser p(text);
    return p.parse_document();
}

}// namespace vs
