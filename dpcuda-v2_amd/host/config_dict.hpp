// config_dict.hpp -- ConfigDict (include/utils/config_dict.h:8-37, a wrapper of an
// nlohmann::json object) for the particle, integrator, routine and I/O set-up code: keyed
// values of type double, long, bool, string, array or nested object, written with
// operator[] and brace lists, read with at(key).get<T>() or implicitly, and saved to /
// loaded from JSON files (save, load, load_config_dict).
//
// JSON text follows nlohmann::json's dump(): keys in sorted order, `dump(4)` indents by
// four spaces with ": " after keys, `dump()` is compact; integers print in decimal;
// doubles print as the shortest decimal string that reads back to the same double, laid
// out as nlohmann's to_chars does (plain notation with a ".0" for decimal exponents -4 <
// n <= 15, otherwise d.ddde+XX with at least two exponent digits); NaN and infinities
// print as null.  Numbers written without '.', 'e' or 'E' read back as integers.
#pragma once
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <initializer_list>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

class ConfigValue;

namespace config_detail {
template <class T>
struct is_vector : std::false_type {};
template <class U>
struct is_vector<std::vector<U>> : std::true_type {
    using elem = U;
};
}  // namespace config_detail

class ConfigValue {
   public:
    using Object = std::map<std::string, ConfigValue>;
    using Array = std::vector<ConfigValue>;

    ConfigValue() = default;
    ConfigValue(std::nullptr_t) {}
    ConfigValue(double x) : v_(x) {}
    ConfigValue(int x) : v_(static_cast<long>(x)) {}
    ConfigValue(long x) : v_(x) {}
    ConfigValue(long long x) : v_(static_cast<long>(x)) {}
    ConfigValue(unsigned x) : v_(static_cast<long>(x)) {}
    ConfigValue(unsigned long x) : v_(static_cast<long>(x)) {}
    ConfigValue(bool x) : v_(x) {}
    ConfigValue(const char* x) : v_(std::string(x)) {}
    ConfigValue(std::string x) : v_(std::move(x)) {}
    ConfigValue(const std::filesystem::path& p) : v_(p.string()) {}
    template <class U>
    ConfigValue(const std::vector<U>& xs) : v_(std::make_shared<Array>()) {
        Array& a = *std::get<std::shared_ptr<Array>>(v_);
        for (const U& x : xs) a.emplace_back(x);
    }
    // {{"key", value}, ...} builds a nested object, as in the reference's configs
    ConfigValue(std::initializer_list<std::pair<const std::string, ConfigValue>> kv)
        : v_(std::make_shared<Object>(kv)) {}
    // copies are deep, like the JSON value the reference wraps
    ConfigValue(const ConfigValue& o) : v_(o.v_) { deepen(); }
    ConfigValue(ConfigValue&&) = default;
    ConfigValue& operator=(const ConfigValue& o) {
        if (this != &o) {
            v_ = o.v_;
            deepen();
        }
        return *this;
    }
    ConfigValue& operator=(ConfigValue&&) = default;

    static ConfigValue array() {
        ConfigValue v;
        v.v_ = std::make_shared<Array>();
        return v;
    }
    static ConfigValue object_value() {
        ConfigValue v;
        v.v_ = std::make_shared<Object>();
        return v;
    }

    bool is_object() const { return std::holds_alternative<std::shared_ptr<Object>>(v_); }
    bool is_array() const { return std::holds_alternative<std::shared_ptr<Array>>(v_); }
    bool is_null() const { return std::holds_alternative<std::monostate>(v_); }
    bool is_string() const { return std::holds_alternative<std::string>(v_); }
    bool is_boolean() const { return std::holds_alternative<bool>(v_); }
    bool is_number_integer() const { return std::holds_alternative<long>(v_); }
    bool is_number_float() const { return std::holds_alternative<double>(v_); }
    bool is_number() const { return is_number_integer() || is_number_float(); }

    ConfigValue& operator[](const std::string& key) {
        if (is_null()) v_ = std::make_shared<Object>();
        return object()[key];
    }
    ConfigValue& operator[](const char* key) { return (*this)[std::string(key)]; }
    const ConfigValue& at(const std::string& key) const {
        const Object& o = object();
        auto it = o.find(key);
        if (it == o.end()) throw std::out_of_range("ConfigDict: key '" + key + "' not found");
        return it->second;
    }
    ConfigValue& at(const std::string& key) {
        Object& o = object();
        auto it = o.find(key);
        if (it == o.end()) throw std::out_of_range("ConfigDict: key '" + key + "' not found");
        return it->second;
    }
    const ConfigValue& at(size_t i) const {
        const Array& a = arr();
        if (i >= a.size()) throw std::out_of_range("ConfigDict: array index out of range");
        return a[i];
    }
    bool contains(const std::string& key) const {
        return is_object() && object().count(key) != 0;
    }
    size_t size() const {
        if (is_object()) return object().size();
        if (is_array()) return arr().size();
        return is_null() ? 0 : 1;
    }
    void push_back(const ConfigValue& x) {
        if (is_null()) v_ = std::make_shared<Array>();
        arr().push_back(x);
    }
    const Object& items() const { return object(); }

    // implicit read, as in `double dt = run_config["dt"];`
    template <class T, std::enable_if_t<std::is_arithmetic_v<T> || std::is_same_v<T, std::string>,
                                        int> = 0>
    operator T() const {
        return get<T>();
    }

    // nlohmann's conversions: numbers and booleans convert among the arithmetic types
    // (floating to integral truncates); bool, string, array and object need that type
    template <class T>
    T get() const {
        if constexpr (std::is_same_v<T, std::string>) {
            if (auto p = std::get_if<std::string>(&v_)) return *p;
        } else if constexpr (std::is_same_v<T, std::filesystem::path>) {
            if (auto p = std::get_if<std::string>(&v_)) return std::filesystem::path(*p);
        } else if constexpr (std::is_same_v<T, bool>) {
            if (auto p = std::get_if<bool>(&v_)) return *p;
        } else if constexpr (std::is_arithmetic_v<T>) {
            if (auto p = std::get_if<long>(&v_)) return static_cast<T>(*p);
            if (auto p = std::get_if<double>(&v_)) return static_cast<T>(*p);
            if (auto p = std::get_if<bool>(&v_)) return static_cast<T>(*p);
        } else if constexpr (std::is_same_v<T, ConfigValue>) {
            return *this;
        } else if constexpr (config_detail::is_vector<T>::value) {
            if (is_array()) {
                T out;
                for (const ConfigValue& x : arr())
                    out.push_back(x.template get<typename config_detail::is_vector<T>::elem>());
                return out;
            }
        }
        throw std::invalid_argument("ConfigDict: value has a different type (" + dump() + ")");
    }

    bool operator==(const ConfigValue& o) const { return dump() == o.dump(); }
    bool operator!=(const ConfigValue& o) const { return !(*this == o); }

    // ---- JSON text ----
    std::string dump(int indent = -1) const {
        std::string out;
        write(out, indent, 0);
        return out;
    }
    static ConfigValue parse(const std::string& text) {
        size_t pos = 0;
        ConfigValue v = parse_value(text, pos);
        skip_ws(text, pos);
        if (pos != text.size()) parse_error(text, pos, "trailing characters");
        return v;
    }

    // the layout of nlohmann::json's number_float output (to_chars, format_buffer)
    static std::string format_double(double x) {
        if (!std::isfinite(x)) return "null";
        char buf[64];
        auto r = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::scientific);
        std::string s(buf, r.ptr);
        std::string sign;
        if (s[0] == '-') {
            sign = "-";
            s.erase(0, 1);
        }
        size_t e = s.find('e');
        std::string mant = s.substr(0, e);
        int exp10 = std::atoi(s.c_str() + e + 1);
        std::string digits;
        for (char c : mant)
            if (c != '.') digits += c;
        int k = (int)digits.size();
        int n = exp10 + 1;  // value = 0.d1d2...dk * 10^n
        std::string out;
        if (k <= n && n <= 15) {
            out = digits + std::string(n - k, '0') + ".0";
        } else if (0 < n && n <= 15) {
            out = digits.substr(0, n) + "." + digits.substr(n);
        } else if (-4 < n && n <= 0) {
            out = "0." + std::string(-n, '0') + digits;
        } else {
            out = digits.substr(0, 1);
            if (k > 1) out += "." + digits.substr(1);
            int ex = n - 1;
            out += ex < 0 ? "e-" : "e+";
            int a = ex < 0 ? -ex : ex;
            if (a < 10) out += "0";
            out += std::to_string(a);
        }
        return sign + out;
    }

   private:
    void deepen() {
        if (is_object()) v_ = std::make_shared<Object>(*std::get<std::shared_ptr<Object>>(v_));
        if (is_array()) v_ = std::make_shared<Array>(*std::get<std::shared_ptr<Array>>(v_));
    }
    Object& object() {
        if (!is_object()) throw std::invalid_argument("ConfigDict: value is not an object");
        return *std::get<std::shared_ptr<Object>>(v_);
    }
    const Object& object() const {
        if (!is_object()) throw std::invalid_argument("ConfigDict: value is not an object");
        return *std::get<std::shared_ptr<Object>>(v_);
    }
    Array& arr() {
        if (!is_array()) throw std::invalid_argument("ConfigDict: value is not an array");
        return *std::get<std::shared_ptr<Array>>(v_);
    }
    const Array& arr() const {
        if (!is_array()) throw std::invalid_argument("ConfigDict: value is not an array");
        return *std::get<std::shared_ptr<Array>>(v_);
    }

    static void write_string(std::string& out, const std::string& s) {
        out += '"';
        for (unsigned char c : s) {
            switch (c) {
                case '"': out += "\\\""; break;
                case '\\': out += "\\\\"; break;
                case '\b': out += "\\b"; break;
                case '\f': out += "\\f"; break;
                case '\n': out += "\\n"; break;
                case '\r': out += "\\r"; break;
                case '\t': out += "\\t"; break;
                default:
                    if (c < 0x20) {
                        char b[8];
                        std::snprintf(b, sizeof b, "\\u%04x", c);
                        out += b;
                    } else {
                        out += (char)c;
                    }
            }
        }
        out += '"';
    }
    void write(std::string& out, int indent, int level) const {
        const bool pretty = indent >= 0;
        auto newline = [&](int lv) {
            if (pretty) {
                out += '\n';
                out += std::string((size_t)(indent * lv), ' ');
            }
        };
        if (is_null()) {
            out += "null";
        } else if (auto p = std::get_if<bool>(&v_)) {
            out += *p ? "true" : "false";
        } else if (auto p = std::get_if<long>(&v_)) {
            out += std::to_string(*p);
        } else if (auto p = std::get_if<double>(&v_)) {
            out += format_double(*p);
        } else if (auto p = std::get_if<std::string>(&v_)) {
            write_string(out, *p);
        } else if (is_array()) {
            const Array& a = arr();
            if (a.empty()) {
                out += "[]";
                return;
            }
            out += '[';
            for (size_t i = 0; i < a.size(); ++i) {
                if (i) out += ',';
                newline(level + 1);
                a[i].write(out, indent, level + 1);
            }
            newline(level);
            out += ']';
        } else {
            const Object& o = object();
            if (o.empty()) {
                out += "{}";
                return;
            }
            out += '{';
            bool first = true;
            for (const auto& kv : o) {
                if (!first) out += ',';
                first = false;
                newline(level + 1);
                write_string(out, kv.first);
                out += pretty ? ": " : ":";
                kv.second.write(out, indent, level + 1);
            }
            newline(level);
            out += '}';
        }
    }

    [[noreturn]] static void parse_error(const std::string& t, size_t pos, const char* what) {
        (void)t;
        throw std::invalid_argument(std::string("ConfigDict: JSON parse error at offset ") +
                                    std::to_string(pos) + ": " + what);
    }
    static void skip_ws(const std::string& t, size_t& pos) {
        while (pos < t.size() &&
               (t[pos] == ' ' || t[pos] == '\t' || t[pos] == '\n' || t[pos] == '\r'))
            ++pos;
    }
    static void append_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    static unsigned parse_hex4(const std::string& t, size_t& pos) {
        if (pos + 4 > t.size()) parse_error(t, pos, "short \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = t[pos++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else parse_error(t, pos - 1, "bad \\u escape");
        }
        return v;
    }
    static std::string parse_string(const std::string& t, size_t& pos) {
        if (t[pos] != '"') parse_error(t, pos, "expected '\"'");
        ++pos;
        std::string out;
        while (true) {
            if (pos >= t.size()) parse_error(t, pos, "unterminated string");
            char c = t[pos++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) parse_error(t, pos - 1, "control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (pos >= t.size()) parse_error(t, pos, "unterminated escape");
            char e = t[pos++];
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
                    unsigned cp = parse_hex4(t, pos);
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        if (pos + 2 > t.size() || t[pos] != '\\' || t[pos + 1] != 'u')
                            parse_error(t, pos, "unpaired surrogate");
                        pos += 2;
                        unsigned lo = parse_hex4(t, pos);
                        if (lo < 0xDC00 || lo > 0xDFFF) parse_error(t, pos, "bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    append_utf8(out, cp);
                    break;
                }
                default: parse_error(t, pos - 1, "bad escape");
            }
        }
        return out;
    }
    static ConfigValue parse_value(const std::string& t, size_t& pos) {
        skip_ws(t, pos);
        if (pos >= t.size()) parse_error(t, pos, "unexpected end of input");
        char c = t[pos];
        if (c == '{') {
            ++pos;
            ConfigValue v = object_value();
            skip_ws(t, pos);
            if (pos < t.size() && t[pos] == '}') {
                ++pos;
                return v;
            }
            while (true) {
                skip_ws(t, pos);
                if (pos >= t.size()) parse_error(t, pos, "unterminated object");
                std::string key = parse_string(t, pos);
                skip_ws(t, pos);
                if (pos >= t.size() || t[pos] != ':') parse_error(t, pos, "expected ':'");
                ++pos;
                v[key] = parse_value(t, pos);
                skip_ws(t, pos);
                if (pos < t.size() && t[pos] == ',') {
                    ++pos;
                    continue;
                }
                if (pos < t.size() && t[pos] == '}') {
                    ++pos;
                    return v;
                }
                parse_error(t, pos, "expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++pos;
            ConfigValue v = array();
            skip_ws(t, pos);
            if (pos < t.size() && t[pos] == ']') {
                ++pos;
                return v;
            }
            while (true) {
                v.push_back(parse_value(t, pos));
                skip_ws(t, pos);
                if (pos < t.size() && t[pos] == ',') {
                    ++pos;
                    continue;
                }
                if (pos < t.size() && t[pos] == ']') {
                    ++pos;
                    return v;
                }
                parse_error(t, pos, "expected ',' or ']'");
            }
        }
        if (c == '"') return ConfigValue(parse_string(t, pos));
        if (t.compare(pos, 4, "true") == 0) {
            pos += 4;
            return ConfigValue(true);
        }
        if (t.compare(pos, 5, "false") == 0) {
            pos += 5;
            return ConfigValue(false);
        }
        if (t.compare(pos, 4, "null") == 0) {
            pos += 4;
            return ConfigValue();
        }
        if (c == '-' || (c >= '0' && c <= '9')) {
            size_t start = pos;
            bool is_float = false;
            if (t[pos] == '-') ++pos;
            while (pos < t.size()) {
                char d = t[pos];
                if (d >= '0' && d <= '9') {
                    ++pos;
                } else if (d == '.' || d == 'e' || d == 'E' || d == '+' || d == '-') {
                    is_float = true;
                    ++pos;
                } else {
                    break;
                }
            }
            std::string num = t.substr(start, pos - start);
            const char* b = num.c_str();
            char* end = nullptr;
            if (!is_float) {
                errno = 0;
                long long iv = std::strtoll(b, &end, 10);
                if (*end == '\0' && errno == 0) return ConfigValue(iv);
            }
            double dv = std::strtod(b, &end);
            if (*end != '\0') parse_error(t, start, "bad number");
            return ConfigValue(dv);
        }
        parse_error(t, pos, "unexpected character");
    }

    std::variant<std::monostate, double, long, bool, std::string, std::shared_ptr<Object>,
                 std::shared_ptr<Array>>
        v_;
};

inline std::ostream& operator<<(std::ostream& os, const ConfigValue& v) { return os << v.dump(); }

class ConfigDict : public ConfigValue {
   public:
    ConfigDict() : ConfigValue(std::initializer_list<std::pair<const std::string, ConfigValue>>{}) {}
    ConfigDict(const ConfigValue& v) : ConfigValue(v) {}

    // ConfigDict::save / load (config_dict.h:11-30): dump(4) to the file; parse the file
    void save(const std::string& path) const {
        std::ofstream ofs(path);
        if (!ofs) throw std::runtime_error("Failed to open file for writing: " + path);
        ofs << dump(4);
    }
    void save(const std::filesystem::path& path) const { save(path.string()); }
    void load(const std::string& path) {
        std::ifstream ifs(path);
        if (!ifs) throw std::runtime_error("Failed to open file for reading: " + path);
        std::stringstream ss;
        ss << ifs.rdbuf();
        ConfigValue::operator=(ConfigValue::parse(ss.str()));
    }
    void load(const std::filesystem::path& path) { load(path.string()); }
};

// load_config_dict (config_dict.h:33-37)
inline ConfigDict load_config_dict(const std::filesystem::path& path) {
    ConfigDict config;
    config.load(path);
    return config;
}
