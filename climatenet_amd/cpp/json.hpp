// Minimal JSON document model for scene files.
//
// Numbers keep the distinction the reference's parser (rapidjson) makes:
// a literal with a fraction or an exponent is a double, otherwise an integer.
// The reference requires doubles where it reads physical parameters
// (jsonRequireDouble -> IsDouble, include/jsonHelper.hpp:136-139), so "0" is
// rejected there while "0.0" is accepted; this model reproduces that.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace json {

class Value {
public:
    enum class Type { Null, Bool, Int, Double, String, Array, Object };

    Value() = default;
    static Value make_bool(bool b);
    static Value make_int(int64_t i);
    static Value make_double(double d);
    static Value make_string(std::string s);
    static Value make_array();
    static Value make_object();

    Type type() const { return type_; }
    bool IsNull() const { return type_ == Type::Null; }
    bool IsBool() const { return type_ == Type::Bool; }
    bool IsInt() const { return type_ == Type::Int && int_ >= INT32_MIN && int_ <= INT32_MAX; }
    bool IsUint() const { return type_ == Type::Int && int_ >= 0 && int_ <= UINT32_MAX; }
    bool IsDouble() const { return type_ == Type::Double; }
    bool IsNumber() const { return type_ == Type::Int || type_ == Type::Double; }
    bool IsString() const { return type_ == Type::String; }
    bool IsArray() const { return type_ == Type::Array; }
    bool IsObject() const { return type_ == Type::Object; }

    bool GetBool() const;
    int GetInt() const;
    unsigned GetUint() const;
    double GetDouble() const;  // integers convert, like rapidjson's GetDouble
    const std::string& GetString() const;

    size_t Size() const;
    const Value& operator[](size_t i) const;
    const std::vector<Value>& array() const { return arr_; }
    std::vector<Value>& array() { return arr_; }

    bool HasMember(const std::string& key) const;
    const Value* FindMember(const std::string& key) const;  // nullptr if absent
    const std::vector<std::pair<std::string, Value>>& members() const { return obj_; }
    std::vector<std::pair<std::string, Value>>& members() { return obj_; }

private:
    Type type_ = Type::Null;
    bool bool_ = false;
    int64_t int_ = 0;
    double dbl_ = 0.;
    std::string str_;
    std::vector<Value> arr_;
    std::vector<std::pair<std::string, Value>> obj_;
};

// Throws std::invalid_argument with the byte offset on malformed input.
Value parse(const std::string& text);
Value parse_file(const std::string& path);

}  // namespace json
