// json.h -- a small ordered JSON value for umcx's input contract.
//
// umcx reads its configuration as JSON (src/umcx.cpp:670-671, :495) and keeps key insertion
// order (nlohmann::ordered_json, :57). This DOM keeps the observable behaviour the simulator
// depends on: integers and floats are told apart at parse time (a number with '.', 'e' or 'E'
// is a double from strtod, others are int64/uint64), numbers reach float through double,
// objects keep insertion order, "--json" merges objects recursively and replaces everything
// else, and a JSON pointer assignment creates missing parents.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace umcx {

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Json {
public:
    enum class Type { Null, Bool, Int, UInt, Float, String, Array, Object };

    Json() = default;
    Json(std::nullptr_t) {}
    Json(bool b) : type_(Type::Bool), b_(b) {}
    Json(int v) : type_(Type::Int), i_(v) {}
    Json(long v) : type_(Type::Int), i_(v) {}
    Json(long long v) : type_(Type::Int), i_(v) {}
    Json(unsigned v) : type_(Type::UInt), u_(v) {}
    Json(unsigned long v) : type_(Type::UInt), u_(v) {}
    Json(unsigned long long v) : type_(Type::UInt), u_(v) {}
    Json(double v) : type_(Type::Float), d_(v) {}
    Json(const char* s);
    Json(std::string s);
    Json(const Json& o);
    Json(Json&& o) noexcept;
    Json& operator=(const Json& o);
    Json& operator=(Json&& o) noexcept;
    ~Json();

    static Json array() { Json j; j.type_ = Type::Array; return j; }
    static Json object() { Json j; j.type_ = Type::Object; return j; }
    static Json parse(const std::string& text);

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_object() const { return type_ == Type::Object; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_string() const { return type_ == Type::String; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_number() const { return type_ == Type::Int || type_ == Type::UInt || type_ == Type::Float; }

    // numeric views (bools convert to 0/1); throw JsonError for non-numbers
    double as_double() const;
    float as_float() const { return (float)as_double(); }
    int64_t as_int() const;
    uint64_t as_uint() const;
    const std::string& as_string() const;

    size_t size() const;                       // elements of an array / keys of an object
    bool contains(const std::string& key) const;
    const Json* find(const std::string& key) const;
    Json* find(const std::string& key);
    const Json& at(const std::string& key) const;
    const Json& at(size_t i) const;
    Json& operator[](const std::string& key);  // object: insert null when missing
    Json& operator[](size_t i);
    void push_back(Json v);

    // object iteration in insertion order
    const std::string& key(size_t i) const;
    const Json& val(size_t i) const;
    Json& val(size_t i);

    // value(key, default): missing or null keys give the default
    double value(const std::string& key, double dflt) const;
    int64_t value_int(const std::string& key, int64_t dflt) const;
    std::string value_str(const std::string& key, const std::string& dflt) const;

    void update(const Json& other);                        // recursive object merge
    void set_pointer(const std::string& ptr, Json v);      // "/a/b" = v, creating parents

    std::string dump(int indent = -1) const;

private:
    void dump_to(std::string& out, int indent, int level) const;

    // 24 bytes per value, so that large arrays (label volumes as _ArrayData_) stay compact: the
    // scalar in a union, strings and containers behind one pointer
    struct Heap;
    const Heap& heap() const;   // an empty heap when none is allocated
    Heap& heap_mut();           // allocates on first use

    Type type_ = Type::Null;
    union {
        int64_t i_ = 0;
        uint64_t u_;
        double d_;
        bool b_;
    };
    Heap* h_ = nullptr;         // String: s; Array: vals; Object: keys and vals (parallel)
};

}  // namespace umcx
