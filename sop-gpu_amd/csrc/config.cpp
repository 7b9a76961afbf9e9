// config.cpp -- parameter files in the reference's format (see config.h).
#include "config.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <sstream>

namespace sop {

std::string string_replace(const std::string& s, const std::string& from, const std::string& to) {
    if (from.empty()) return s;
    std::string r(s);
    size_t pos = 0;
    while ((pos = r.find(from, pos)) != std::string::npos) {
        r.replace(pos, from.size(), to);
        pos += to.size();
    }
    return r;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string to_lower(const std::string& s) {
    std::string r(s);
    std::transform(r.begin(), r.end(), r.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return r;
}

void Config::load(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw ConfigError("Can't open parameters file '" + path + "'");
    char buf[4096];
    while (std::fgets(buf, sizeof buf, f)) {
        const char c0 = buf[0];
        if (c0 == '#' || c0 == ' ' || c0 == '\n' || c0 == '\t' || c0 == '\r' || c0 == '\0') continue;
        char* name = std::strtok(buf, " \t");
        char* value = std::strtok(nullptr, "#\n");
        if (!name) continue;
        std::string n = trim(name);
        if (n.empty()) continue;
        values_[n] = value ? trim(value) : std::string();
    }
    std::fclose(f);
}

const std::string& Config::raw(const std::string& name) const {
    auto it = values_.find(name);
    if (it == values_.end()) throw ConfigError("Parameter '" + name + "' not specified");
    return it->second;
}

std::string Config::masked(const std::string& name) const {
    std::string v = raw(name);
    for (const auto& kv : values_) v = string_replace(v, "<" + kv.first + ">", kv.second);
    return v;
}

template <typename T>
static T extract(const std::string& name, const std::string& v) {
    std::istringstream ss(v);
    T out{};
    ss >> out;
    if (ss.fail()) throw ConfigError("Error converting string `" + v + "` to value (parameter '" + name + "')");
    return out;
}

std::string Config::str(const std::string& name) const {
    std::string v = extract<std::string>(name, masked(name));
    used_.push_back({name, v});
    return v;
}
std::string Config::str_or(const std::string& name, const std::string& def) const {
    if (!has(name)) {
        used_.push_back({name, def});
        return def;
    }
    return str(name);
}

long long Config::integer(const std::string& name) const {
    long long v = extract<long long>(name, masked(name));
    used_.push_back({name, std::to_string(v)});
    return v;
}
long long Config::integer_or(const std::string& name, long long def) const {
    if (!has(name)) {
        used_.push_back({name, std::to_string(def)});
        return def;
    }
    return integer(name);
}

double Config::real(const std::string& name) const {
    double v = extract<double>(name, masked(name));
    used_.push_back({name, raw(name)});
    return v;
}
double Config::real_or(const std::string& name, double def) const {
    if (!has(name)) {
        used_.push_back({name, std::to_string(def)});
        return def;
    }
    return real(name);
}

bool Config::boolean(const std::string& name) const {
    const std::string v = to_lower(raw(name));
    bool r;
    if (v == "yes" || v == "true" || v == "on" || v == "y") r = true;
    else if (v == "no" || v == "false" || v == "off" || v == "n") r = false;
    else throw ConfigError("Wrong value '" + v + "' for boolean parameter '" + name + "'");
    used_.push_back({name, r ? "true" : "false"});
    return r;
}
bool Config::boolean_or(const std::string& name, bool def) const { return has(name) ? boolean(name) : def; }

template <typename T>
static std::vector<T> array_of(const std::string& name, const std::string& rawv) {
    const std::string v = to_lower(rawv);
    std::vector<T> r;
    if (v == "empty" || v == "none") return r;
    std::istringstream ss(v);
    std::string tok;
    while (ss >> tok) {
        if (tok == "to") {
            if (r.empty()) throw ConfigError("In parameters of array type, 'to' keyword must be preceeded by number");
            T from = r.back(), to;
            if (!(ss >> to)) throw ConfigError("Parameter '" + name + "': 'to' without an upper bound");
            for (T i = from + 1; i <= to; ++i) r.push_back(i);
        } else {
            r.push_back(extract<T>(name, tok));
        }
    }
    return r;
}

std::vector<int> Config::int_array(const std::string& name) const { return array_of<int>(name, raw(name)); }
std::vector<float> Config::float_array(const std::string& name) const { return array_of<float>(name, raw(name)); }

std::array<float, 3> Config::float3(const std::string& name) const {
    const std::vector<float> v = float_array(name);
    if (v.size() != 3)
        throw ConfigError("Float3 parameter '" + name + "' has " + std::to_string(v.size()) +
                          " components; must have 3.");
    return {v[0], v[1], v[2]};
}
std::array<float, 3> Config::float3_or(const std::string& name, std::array<float, 3> def) const {
    return has(name) ? float3(name) : def;
}

}  // namespace sop
