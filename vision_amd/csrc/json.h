// Minimal JSON value + parser for Vision scene files.
//
// Vision scene files are JSON with C/C++ comments allowed (`//` and `/* */`),
// stripped before parsing (reference: src/core/json_util.h:12-110).
// This parser accepts those comments directly and keeps object key order.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vs {

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    static Json make_bool(bool b) { Json j; j.type_ = Type::Bool; j.b_ = b; return j; }
    static Json make_number(double d) { Json j; j.type_ = Type::Number; j.d_ = d; return j; }
    static Json make_string(std::string s) { Json j; j.type_ = Type::String; j.s_ = std::move(s); return j; }
    static Json make_array() { Json j; j.type_ = Type::Array; return j; }
    static Json make_object() { Json j; j.type_ = Type::Object; return j; }

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }

    double number() const;
    bool boolean() const;
    const std::string &str() const;
    size_t size() const;
    const Json &at(size_t i) const;
    // object access; missing key -> shared null value
    const Json &operator[](const std::string &key) const;
    bool contains(const std::string &key) const;
    const std::vector<std::pair<std::string, Json>> &items() const { return obj_; }
    const std::vector<Json> &elements() const { return arr_; }

    void push_back(Json v) { arr_.push_back(std::move(v)); }
    void set(const std::string &key, Json v);

    // typed getters with defaults (mirror ParameterSet::as_* in the reference:
    // a missing key yields the default)
    double as_float(double def) const { return is_number() ? d_ : (is_bool() ? (b_ ? 1.0 : 0.0) : def); }
    uint32_t as_uint(uint32_t def) const { return is_number() ? static_cast<uint32_t>(d_) : (is_bool() ? (b_ ? 1u : 0u) : def); }
    int as_int(int def) const { return is_number() ? static_cast<int>(d_) : (is_bool() ? (b_ ? 1 : 0) : def); }
    bool as_bool(bool def) const { return is_bool() ? b_ : (is_number() ? d_ != 0.0 : def); }
    std::string as_string(const std::string &def) const { return is_string() ? s_ : def; }
    std::vector<double> as_numbers() const;

    static Json parse(const std::string &text);

private:
    Type type_{Type::Null};
    bool b_{false};
    double d_{0.0};
    std::string s_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;
};

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}// namespace vs
