// json.h -- small JSON value and parser for the scene files.
//
// The reference reads scenes with nlohmann::json 2.1 through its jsel
// reflection (src/util/jsel.hpp); this build parses the same documents with
// its own recursive-descent reader. Objects keep their key order.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace koishi {

class Json {
public:
    enum Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    explicit Json(bool b) : type_(Bool), b_(b) {}
    explicit Json(double d) : type_(Number), n_(d) {}
    explicit Json(std::string s) : type_(String), s_(std::move(s)) {}

    static Json array() {
        Json j;
        j.type_ = Array;
        return j;
    }
    static Json object() {
        Json j;
        j.type_ = Object;
        return j;
    }
    static Json parse(const std::string &text);  // throws std::logic_error

    Type type() const { return type_; }
    bool is_null() const { return type_ == Null; }
    bool is_number() const { return type_ == Number; }
    bool is_string() const { return type_ == String; }
    bool is_array() const { return type_ == Array; }
    bool is_object() const { return type_ == Object; }
    bool is_bool() const { return type_ == Bool; }

    double number() const;
    bool boolean() const;
    const std::string &str() const;
    const std::vector<Json> &items() const;
    const std::vector<std::pair<std::string, Json>> &members() const;

    size_t size() const;
    const Json &operator[](size_t i) const;
    bool has(const std::string &key) const;
    const Json &at(const std::string &key) const;  // throws when missing
    const Json *find(const std::string &key) const;

    void push(Json v) { arr_.push_back(std::move(v)); }
    void set(const std::string &k, Json v);

    std::string dump() const;

private:
    Type type_ = Null;
    bool b_ = false;
    double n_ = 0;
    std::string s_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;
};

const char *json_type_name(Json::Type t);

}  // namespace koishi
