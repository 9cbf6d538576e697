// toml.h -- the subset of TOML the reference's input files use (tables,
// strings, numbers, booleans, nested arrays, '#' comments).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace uqk_host {

struct TomlValue {
    enum Kind { STRING, NUMBER, BOOL, ARRAY } kind = NUMBER;
    std::string str;
    double num = 0.0;
    bool boolean = false;
    std::vector<TomlValue> arr;
};

class Toml {
  public:
    // table -> key -> value ("" is the root table)
    std::map<std::string, std::map<std::string, TomlValue>> tables;

    static Toml parse_file(const std::string& path);
    static Toml parse(const std::string& text);

    const TomlValue* get(const std::string& table, const std::string& key) const;
    bool has(const std::string& table, const std::string& key) const { return get(table, key) != nullptr; }
    std::string get_string(const std::string& table, const std::string& key, const std::string& def) const;
    double get_number(const std::string& table, const std::string& key, double def) const;
    bool get_bool(const std::string& table, const std::string& key, bool def) const;
};

}  // namespace uqk_host
