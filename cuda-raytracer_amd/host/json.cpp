// json.cpp -- recursive-descent JSON reader (RFC 8259 subset: no surrogate
// pair decoding beyond the BMP) with line/column error messages.
#include "json.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace koishi {

const char *json_type_name(Json::Type t) {
    switch (t) {
    case Json::Null: return "null";
    case Json::Bool: return "boolean";
    case Json::Number: return "number";
    case Json::String: return "string";
    case Json::Array: return "array";
    case Json::Object: return "object";
    }
    return "?";
}

namespace {
[[noreturn]] void type_error(const char *want, Json::Type got) {
    throw std::logic_error(std::string("json: expected ") + want + ", got " + json_type_name(got));
}

struct Parser {
    const std::string &t;
    size_t i = 0;

    [[noreturn]] void error(const std::string &msg) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i && k < t.size(); ++k) {
            if (t[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        std::ostringstream os;
        os << "json parse error at line " << line << ", column " << col << ": " << msg;
        throw std::logic_error(os.str());
    }
    void ws() {
        while (i < t.size() && (t[i] == ' ' || t[i] == '\t' || t[i] == '\n' || t[i] == '\r')) ++i;
    }
    bool lit(const char *s) {
        size_t n = std::char_traits<char>::length(s);
        if (t.compare(i, n, s) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    Json value(int depth) {
        if (depth > 512) error("nesting too deep");
        ws();
        if (i >= t.size()) error("unexpected end of input");
        char c = t[i];
        if (c == '{') return object(depth);
        if (c == '[') return array(depth);
        if (c == '"') return Json(string());
        if (lit("true")) return Json(true);
        if (lit("false")) return Json(false);
        if (lit("null")) return Json();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        error(std::string("unexpected character '") + c + "'");
    }
    Json number() {
        size_t b = i;
        if (t[i] == '-') ++i;
        if (i >= t.size() || !isdigit((unsigned char)t[i])) error("bad number");
        while (i < t.size() && isdigit((unsigned char)t[i])) ++i;
        if (i < t.size() && t[i] == '.') {
            ++i;
            if (i >= t.size() || !isdigit((unsigned char)t[i])) error("bad number");
            while (i < t.size() && isdigit((unsigned char)t[i])) ++i;
        }
        if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
            ++i;
            if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
            if (i >= t.size() || !isdigit((unsigned char)t[i])) error("bad number");
            while (i < t.size() && isdigit((unsigned char)t[i])) ++i;
        }
        return Json(std::strtod(t.substr(b, i - b).c_str(), nullptr));
    }
    static void utf8(std::string &out, unsigned cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    std::string string() {
        ++i;  // opening quote
        std::string out;
        while (true) {
            if (i >= t.size()) error("unterminated string");
            char c = t[i++];
            if (c == '"') break;
            if (c == '\\') {
                if (i >= t.size()) error("unterminated escape");
                char e = t[i++];
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
                    if (i + 4 > t.size()) error("bad \\u escape");
                    unsigned cp = unsigned(std::strtoul(t.substr(i, 4).c_str(), nullptr, 16));
                    i += 4;
                    utf8(out, cp);
                    break;
                }
                default: error("bad escape");
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    Json array(int depth) {
        ++i;
        Json a = Json::array();
        ws();
        if (i < t.size() && t[i] == ']') {
            ++i;
            return a;
        }
        while (true) {
            a.push(value(depth + 1));
            ws();
            if (i < t.size() && t[i] == ',') {
                ++i;
                continue;
            }
            if (i < t.size() && t[i] == ']') {
                ++i;
                return a;
            }
            error("expected ',' or ']'");
        }
    }
    Json object(int depth) {
        ++i;
        Json o = Json::object();
        ws();
        if (i < t.size() && t[i] == '}') {
            ++i;
            return o;
        }
        while (true) {
            ws();
            if (i >= t.size() || t[i] != '"') error("expected a key string");
            std::string k = string();
            ws();
            if (i >= t.size() || t[i] != ':') error("expected ':'");
            ++i;
            o.set(k, value(depth + 1));
            ws();
            if (i < t.size() && t[i] == ',') {
                ++i;
                continue;
            }
            if (i < t.size() && t[i] == '}') {
                ++i;
                return o;
            }
            error("expected ',' or '}'");
        }
    }
};
}  // namespace

Json Json::parse(const std::string &text) {
    Parser p{text};
    Json v = p.value(0);
    p.ws();
    if (p.i != text.size()) p.error("trailing characters");
    return v;
}

double Json::number() const {
    if (type_ == Bool) return b_ ? 1.0 : 0.0;
    if (type_ != Number) type_error("number", type_);
    return n_;
}
bool Json::boolean() const {
    if (type_ == Number) return n_ != 0;
    if (type_ != Bool) type_error("boolean", type_);
    return b_;
}
const std::string &Json::str() const {
    if (type_ != String) type_error("string", type_);
    return s_;
}
const std::vector<Json> &Json::items() const {
    if (type_ != Array) type_error("array", type_);
    return arr_;
}
const std::vector<std::pair<std::string, Json>> &Json::members() const {
    if (type_ != Object) type_error("object", type_);
    return obj_;
}
size_t Json::size() const {
    if (type_ == Array) return arr_.size();
    if (type_ == Object) return obj_.size();
    return 0;
}
const Json &Json::operator[](size_t i) const {
    const auto &a = items();
    if (i >= a.size()) throw std::logic_error("json: array index out of range");
    return a[i];
}
const Json *Json::find(const std::string &key) const {
    if (type_ != Object) return nullptr;
    for (auto &kv : obj_)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
bool Json::has(const std::string &key) const { return find(key) != nullptr; }
const Json &Json::at(const std::string &key) const {
    const Json *j = find(key);
    if (!j) throw std::logic_error("No such key named \"" + key + "\".");
    return *j;
}
void Json::set(const std::string &k, Json v) {
    if (type_ != Object) type_error("object", type_);
    for (auto &kv : obj_)
        if (kv.first == k) {
            kv.second = std::move(v);
            return;
        }
    obj_.emplace_back(k, std::move(v));
}

namespace {
// A JSON string literal (RFC 8259 section 7): quote, backslash and control
// characters escaped.
std::string quote(const std::string &s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\b': o += "\\b"; break;
        case '\f': o += "\\f"; break;
        case '\n': o += "\\n"; break;
        case '\r': o += "\\r"; break;
        case '\t': o += "\\t"; break;
        default:
            if (c < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", c);
                o += buf;
            } else {
                o += char(c);
            }
        }
    }
    return o + "\"";
}
}  // namespace

std::string Json::dump() const {
    std::ostringstream os;
    switch (type_) {
    case Null: os << "null"; break;
    case Bool: os << (b_ ? "true" : "false"); break;
    case Number: {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.9g", n_);
        os << buf;
        break;
    }
    case String: os << quote(s_); break;
    case Array:
        os << '[';
        for (size_t k = 0; k < arr_.size(); ++k) os << (k ? "," : "") << arr_[k].dump();
        os << ']';
        break;
    case Object:
        os << '{';
        for (size_t k = 0; k < obj_.size(); ++k) os << (k ? "," : "") << quote(obj_[k].first) << ':' << obj_[k].second.dump();
        os << '}';
        break;
    }
    return os.str();
}

}  // namespace koishi
