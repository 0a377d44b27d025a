// The subset of p-ranav/argparse's interface that cfdARCO's command line uses:
// named optional arguments with default / implicit values, scan<> conversion,
// `nargs(any)` lists, help text, parse_args() and typed get<>().
#pragma once

#include <any>
#include <cstdlib>
#include <functional>
#include <typeinfo>
#include <map>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace argparse {

enum class nargs_pattern { optional, any, at_least_one };

class ArgumentParser;

class Argument {
public:
    explicit Argument(std::vector<std::string> names) : names_(std::move(names)) {}

    template <class T> Argument& default_value(T v) {
        value_ = to_any(std::move(v));
        return *this;
    }
    template <class T> Argument& implicit_value(T v) {
        implicit_ = to_any(std::move(v));
        return *this;
    }
    Argument& help(std::string h) {
        help_ = std::move(h);
        return *this;
    }
    Argument& nargs(nargs_pattern p) {
        many_ = p != nargs_pattern::optional;
        return *this;
    }
    // 'i' integral, 'g' / 'f' floating point
    template <char Shape, class T> Argument& scan() {
        convert_ = [](const std::string& s) -> std::any {
            if constexpr (std::is_integral<T>::value) return T(std::stoll(s));
            else return T(std::stod(s));
        };
        return *this;
    }

private:
    friend class ArgumentParser;
    template <class T> static std::any to_any(T v) {
        if constexpr (std::is_same<T, const char*>::value) return std::string(v);
        else return v;
    }

    std::vector<std::string> names_;
    std::any value_;
    std::any implicit_;
    std::string help_;
    bool many_ = false;
    std::function<std::any(const std::string&)> convert_;
};

class ArgumentParser {
public:
    explicit ArgumentParser(std::string prog = "") : prog_(std::move(prog)) {}

    template <class... Names> Argument& add_argument(Names... names) {
        args_.push_back(std::make_unique<Argument>(std::vector<std::string>{std::string(names)...}));
        Argument* a = args_.back().get();
        for (const auto& n : a->names_) index_[n] = a;
        return *a;
    }

    void parse_args(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) {
            const std::string tok = argv[i];
            auto it = index_.find(tok);
            if (it == index_.end()) throw std::runtime_error("Unknown argument: " + tok);
            Argument* a = it->second;
            if (a->implicit_.has_value()) {
                a->value_ = a->implicit_;
                continue;
            }
            if (a->many_) {
                std::vector<std::any> items;
                while (i + 1 < argc && index_.find(argv[i + 1]) == index_.end()) items.push_back(convert(a, argv[++i]));
                a->value_ = items;
                continue;
            }
            if (i + 1 >= argc) throw std::runtime_error("Argument " + tok + " needs a value");
            a->value_ = convert(a, argv[++i]);
        }
    }

    template <class T> T get(const std::string& name) const {
        Argument* a = find(name);
        if constexpr (is_vector<T>::value) {
            using E = typename T::value_type;
            if (a->value_.type() == typeid(std::vector<std::any>)) {
                T out;
                for (const auto& x : std::any_cast<const std::vector<std::any>&>(a->value_)) out.push_back(std::any_cast<E>(x));
                return out;
            }
            return std::any_cast<T>(a->value_);
        } else {
            const std::any& v = a->value_;
            if (v.type() == typeid(T)) return std::any_cast<T>(v);
            // default values given as a different arithmetic type (e.g. 1.0 for a float)
            if constexpr (std::is_arithmetic<T>::value) {
                if (v.type() == typeid(double)) return T(std::any_cast<double>(v));
                if (v.type() == typeid(float)) return T(std::any_cast<float>(v));
                if (v.type() == typeid(int)) return T(std::any_cast<int>(v));
                if (v.type() == typeid(long long)) return T(std::any_cast<long long>(v));
                if (v.type() == typeid(bool)) return T(std::any_cast<bool>(v));
            }
            throw std::runtime_error("argument " + name + " has another type");
        }
    }

    friend std::ostream& operator<<(std::ostream& os, const ArgumentParser& p) {
        os << "Usage: " << p.prog_ << " [options]\n";
        for (const auto& a : p.args_) {
            os << "  ";
            for (size_t i = 0; i < a->names_.size(); ++i) os << (i ? ", " : "") << a->names_[i];
            if (!a->help_.empty()) os << "\t" << a->help_;
            os << "\n";
        }
        return os;
    }

private:
    template <class T> struct is_vector : std::false_type {};
    template <class E, class A> struct is_vector<std::vector<E, A>> : std::true_type {};

    Argument* find(const std::string& name) const {
        for (const std::string& cand : {name, "-" + name, "--" + name}) {
            auto it = index_.find(cand);
            if (it != index_.end()) return it->second;
        }
        throw std::runtime_error("No such argument: " + name);
    }
    static std::any convert(Argument* a, const std::string& s) {
        if (a->convert_) return a->convert_(s);
        return s;
    }

    std::string prog_;
    std::vector<std::unique_ptr<Argument>> args_;
    std::map<std::string, Argument*> index_;
};

}  // namespace argparse
