// config.h -- parameter files in the reference's format.
//
// Syntax (src/IO/configreader.cpp:23-58): one `name value` per line; lines
// starting with '#', space, tab or newline are skipped; the value ends at
// '#' or the end of line and is trimmed.  `<name>` inside a value is
// replaced by that parameter's raw value (getMaskedParameter, :77-86).
// Arrays are whitespace-separated with `a to b` ranges and the words
// `empty`/`none` (:89-118); booleans accept yes/no/on/off/true/false/y/n
// (:122-129); float3 is an array of exactly three numbers (:131-141).
// Numbers and strings are read with stream extraction, so a string value
// is its first whitespace-delimited token (src/Util/mystl.h:49-63).
#pragma once
#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace sop {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Config {
   public:
    void load(const std::string& path);  // throws ConfigError
    void set(const std::string& name, const std::string& value) { values_[name] = value; }
    bool has(const std::string& name) const { return values_.count(name) != 0; }
    const std::string& raw(const std::string& name) const;
    std::string masked(const std::string& name) const;

    // Typed reads; the *_or forms return the default when the parameter is
    // absent (ParameterOptional), the plain forms require it (ParameterMandatory).
    std::string str(const std::string& name) const;
    std::string str_or(const std::string& name, const std::string& def) const;
    long long integer(const std::string& name) const;
    long long integer_or(const std::string& name, long long def) const;
    double real(const std::string& name) const;
    double real_or(const std::string& name, double def) const;
    bool boolean(const std::string& name) const;
    bool boolean_or(const std::string& name, bool def) const;
    std::array<float, 3> float3(const std::string& name) const;
    std::array<float, 3> float3_or(const std::string& name, std::array<float, 3> def) const;
    std::vector<int> int_array(const std::string& name) const;
    std::vector<float> float_array(const std::string& name) const;

    // The log of parameters read, as the reference prints "Using parameter".
    const std::vector<std::pair<std::string, std::string>>& used() const { return used_; }

   private:
    std::map<std::string, std::string> values_;
    mutable std::vector<std::pair<std::string, std::string>> used_;
};

std::string string_replace(const std::string& s, const std::string& from, const std::string& to);
std::string trim(const std::string& s);
std::string to_lower(const std::string& s);

}  // namespace sop
