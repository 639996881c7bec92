// config.cpp -- keyword configuration parser (grammar of config.h:61-167).
#include "config.hpp"

#include <fstream>
#include <iostream>

namespace rift {

namespace {

bool only_spaces(const std::string& s) { return s.find_first_not_of(' ') == std::string::npos; }

std::string trim_left_blanks(const std::string& s)
{
    std::string t = s;
    t.erase(0, t.find_first_not_of("\t "));
    return t;
}

}  // namespace

Config::Config(const std::string& path)
{
    std::ifstream f(path.c_str());
    if (!f) throw ConfigError("CFG: File " + path + " couldn't be found!");
    std::string line;
    size_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (line.empty()) continue;
        std::string t = line;
        size_t sc = t.find(';');
        if (sc != std::string::npos) t.erase(sc);
        if (only_spaces(t)) continue;
        if (t.find('=') == std::string::npos)
            throw ConfigError("CFG: Couldn't find separator on line: " + std::to_string(lineno));
        // validLine (config.h:73-91)
        std::string v = trim_left_blanks(t);
        bool valid = v[0] != '=';
        if (valid) {
            valid = false;
            for (size_t i = v.find('=') + 1; i < v.size(); ++i)
                if (v[i] != ' ') { valid = true; break; }
        }
        if (!valid) throw ConfigError("CFG: Bad format for line: " + std::to_string(lineno));
        // extractKey / extractValue (config.h:93-108)
        size_t sep = v.find('=');
        std::string key = v.substr(0, sep);
        size_t cut = key.find_first_of("\t ");
        if (cut != std::string::npos) key.erase(cut);
        std::string value = v.substr(sep + 1);
        value.erase(0, value.find_first_not_of("\t "));
        size_t last = value.find_last_not_of("\t ");
        value.erase(last == std::string::npos ? 0 : last + 1);
        if (kv_.count(key)) throw ConfigError("CFG: Can only have unique key names!");
        kv_[key] = value;
    }
}

const std::string& Config::raw(const std::string& key) const
{
    auto it = kv_.find(key);
    if (it == kv_.end()) throw ConfigError("CONFIG: missing key '" + key + "'");
    return it->second;
}

bool Config::is_number(const std::string& key) const
{
    auto it = kv_.find(key);
    if (it == kv_.end()) throw ConfigError("CONFIG: missing key '" + key + "'");
    std::istringstream is(it->second);
    std::string tok;
    is >> tok;
    if (tok.empty()) return false;
    for (char c : tok)
        if (!(std::isdigit((unsigned char)c) || c == '.')) return false;
    return true;
}

void Config::warn_missing(const std::string& key) const
{
    std::cerr << "Error! " << key
              << " has no corresponding value, please go to the Config file and add that value"
              << std::endl;
}

}  // namespace rift
