// key=value command-line parameters (the style of Opm::parameter::ParameterGroup).
#ifndef MHS_EXAMPLE_PARAMS_HPP
#define MHS_EXAMPLE_PARAMS_HPP
#include <cstdlib>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>

struct Params {
    std::map<std::string, std::string> kv;
    Params(int argc, char **argv) {
        for (int i = 1; i < argc; ++i) {
            std::string a(argv[i]);
            const auto eq = a.find('=');
            if (eq != std::string::npos) kv[a.substr(0, eq)] = a.substr(eq + 1);
        }
    }
    double get(const std::string &k, double def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : std::atof(it->second.c_str());
    }
    int geti(const std::string &k, int def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : std::atoi(it->second.c_str());
    }
    std::string gets(const std::string &k, const std::string &def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : it->second;
    }
    // ParameterGroup::getDefault
    double getDefault(const std::string &k, double def) const { return get(k, def); }
    int getDefault(const std::string &k, int def) const { return geti(k, def); }
    bool getDefault(const std::string &k, bool def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : (it->second == "1" || it->second == "true");
    }
    // ParameterGroup::has, get<T>, getDefault<T>
    bool has(const std::string &k) const { return kv.count(k) != 0; }
    template <class T>
    T get(const std::string &k) const {
        auto it = kv.find(k);
        if (it == kv.end()) throw std::runtime_error("missing parameter " + k);
        return convert<T>(it->second);
    }
    template <class T>
    T getDefault(const std::string &k, const T &def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : convert<T>(it->second);
    }

private:
    template <class T>
    static T convert(const std::string &v) {
        if constexpr (std::is_same<T, std::string>::value) {
            return v;
        } else if constexpr (std::is_same<T, bool>::value) {
            return v == "1" || v == "true";
        } else {
            std::istringstream is(v);
            T x{};
            if (!(is >> x)) throw std::runtime_error("unreadable parameter value " + v);
            return x;
        }
    }
};
#endif
