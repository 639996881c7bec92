// config.hpp -- the keyword configuration file of RIFT (`key = value` lines).
//
// Grammar of ConfigFile (config.h:48-209): `;` starts a comment; a line that
// is empty or holds only spaces is skipped; every other line needs a `=`
// that is not its first non-blank character and a non-space character after
// it; the key is the text before `=` cut at its first space or tab; the value
// is the text after `=` with blanks trimmed; keys are unique.  Typed values
// are read with `operator>>`, so a string value is its first word and a bool
// is 0/1.  Differences: errors are reported as exceptions (`ConfigError`)
// instead of blocking on stdin and exiting, and a missing key returns the
// supplied default with a message on stderr, which is what the reference
// returns after its prompt.
#pragma once

#include <complex>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace rift {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Config {
public:
    explicit Config(const std::string& path);

    bool has(const std::string& key) const { return kv_.count(key) != 0; }
    // keyIsNumber (config.h:176-190): only digits and '.' -- "3e-2" or "-1"
    // are not numbers and are taken as file names by the caller.
    bool is_number(const std::string& key) const;
    const std::string& raw(const std::string& key) const;

    template <class T>
    T get(const std::string& key, const T& dflt = T()) const
    {
        auto it = kv_.find(key);
        if (it == kv_.end()) {
            warn_missing(key);
            return dflt;
        }
        std::istringstream is(it->second);
        T v;
        if (!(is >> v)) throw ConfigError("CONFIG: not a valid value for key '" + key + "'");
        return v;
    }

    const std::map<std::string, std::string>& entries() const { return kv_; }

private:
    void warn_missing(const std::string& key) const;
    std::map<std::string, std::string> kv_;
};

}  // namespace rift
