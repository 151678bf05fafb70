// config_dict.hpp -- the subset of ConfigDict (include/utils/config_dict.h, a wrapper of a
// JSON object) that the particle, integrator and routine set-up code uses: keyed values
// of type double, long, bool, string or nested object, written with operator[] and
// brace lists and read with at(key).get<T>().
#pragma once
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>

class ConfigValue {
   public:
    using Object = std::map<std::string, ConfigValue>;

    ConfigValue() = default;
    ConfigValue(double x) : v_(x) {}
    ConfigValue(int x) : v_(static_cast<long>(x)) {}
    ConfigValue(long x) : v_(x) {}
    ConfigValue(long long x) : v_(static_cast<long>(x)) {}
    ConfigValue(bool x) : v_(x) {}
    ConfigValue(const char* x) : v_(std::string(x)) {}
    ConfigValue(std::string x) : v_(std::move(x)) {}
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

    bool is_object() const { return std::holds_alternative<std::shared_ptr<Object>>(v_); }
    bool is_null() const { return std::holds_alternative<std::monostate>(v_); }

    ConfigValue& operator[](const std::string& key) {
        if (is_null()) v_ = std::make_shared<Object>();
        return object()[key];
    }
    const ConfigValue& at(const std::string& key) const {
        const Object& o = object();
        auto it = o.find(key);
        if (it == o.end()) throw std::out_of_range("ConfigDict: key '" + key + "' not found");
        return it->second;
    }
    bool contains(const std::string& key) const {
        return is_object() && object().count(key) != 0;
    }

    // implicit read, as in `double dt = run_config["dt"];`
    template <class T, std::enable_if_t<std::is_arithmetic_v<T> || std::is_same_v<T, std::string>,
                                        int> = 0>
    operator T() const {
        return get<T>();
    }

    template <class T>
    T get() const {
        if constexpr (std::is_same_v<T, std::string>) {
            if (auto p = std::get_if<std::string>(&v_)) return *p;
        } else if constexpr (std::is_same_v<T, bool>) {
            if (auto p = std::get_if<bool>(&v_)) return *p;
        } else if constexpr (std::is_integral_v<T>) {
            if (auto p = std::get_if<long>(&v_)) return static_cast<T>(*p);
        } else if constexpr (std::is_floating_point_v<T>) {
            if (auto p = std::get_if<double>(&v_)) return static_cast<T>(*p);
            if (auto p = std::get_if<long>(&v_)) return static_cast<T>(*p);
        } else if constexpr (std::is_same_v<T, ConfigValue>) {
            return *this;
        }
        throw std::invalid_argument("ConfigDict: value has a different type");
    }

   private:
    void deepen() {
        if (is_object()) v_ = std::make_shared<Object>(*std::get<std::shared_ptr<Object>>(v_));
    }
    Object& object() {
        if (!is_object()) throw std::invalid_argument("ConfigDict: value is not an object");
        return *std::get<std::shared_ptr<Object>>(v_);
    }
    const Object& object() const {
        if (!is_object()) throw std::invalid_argument("ConfigDict: value is not an object");
        return *std::get<std::shared_ptr<Object>>(v_);
    }
    std::variant<std::monostate, double, long, bool, std::string, std::shared_ptr<Object>> v_;
};

class ConfigDict : public ConfigValue {
   public:
    ConfigDict() : ConfigValue(std::initializer_list<std::pair<const std::string, ConfigValue>>{}) {}
    ConfigDict(const ConfigValue& v) : ConfigValue(v) {}
};
