// toml.cpp -- recursive-descent parser for the TOML subset of toml.h.
#include "toml.h"

#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace uqk_host {

namespace {

struct Cursor {
    const std::string& s;
    size_t i = 0;
    int line = 1;
    explicit Cursor(const std::string& t) : s(t) {}
    bool eof() const { return i >= s.size(); }
    char peek() const { return eof() ? '\0' : s[i]; }
    char get() {
        char c = s[i++];
        if (c == '\n') ++line;
        return c;
    }
    [[noreturn]] void fail(const std::string& msg) const {
        throw std::runtime_error("toml line " + std::to_string(line) + ": " + msg);
    }
    // skip blanks, comments and (inside arrays) newlines
    void skip(bool newlines) {
        for (;;) {
            while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\r' || (newlines && peek() == '\n'))) get();
            if (peek() == '#') {
                while (!eof() && peek() != '\n') get();
                continue;
            }
            break;
        }
    }
};

TomlValue parse_value(Cursor& c);

TomlValue parse_array(Cursor& c) {
    TomlValue v;
    v.kind = TomlValue::ARRAY;
    c.get();  // '['
    c.skip(true);
    while (c.peek() != ']') {
        v.arr.push_back(parse_value(c));
        c.skip(true);
        if (c.peek() == ',') {
            c.get();
            c.skip(true);
        } else if (c.peek() != ']') {
            c.fail("expected ',' or ']' in array");
        }
    }
    c.get();  // ']'
    return v;
}

TomlValue parse_value(Cursor& c) {
    c.skip(false);
    TomlValue v;
    char ch = c.peek();
    if (ch == '[') return parse_array(c);
    if (ch == '"' || ch == '\'') {
        const char q = c.get();
        v.kind = TomlValue::STRING;
        while (!c.eof() && c.peek() != q) {
            char d = c.get();
            if (d == '\\' && q == '"' && !c.eof()) {
                char e = c.get();
                d = e == 'n' ? '\n' : e == 't' ? '\t' : e;
            }
            v.str.push_back(d);
        }
        if (c.eof()) c.fail("unterminated string");
        c.get();
        return v;
    }
    std::string tok;
    while (!c.eof() && !std::isspace((unsigned char)c.peek()) && c.peek() != ',' && c.peek() != ']' && c.peek() != '#')
        tok.push_back(c.get());
    if (tok == "true" || tok == "false") {
        v.kind = TomlValue::BOOL;
        v.boolean = tok == "true";
        return v;
    }
    std::string clean;
    for (char d : tok)
        if (d != '_') clean.push_back(d);
    try {
        size_t pos = 0;
        v.num = std::stod(clean, &pos);
        if (pos != clean.size()) c.fail("bad number '" + tok + "'");
    } catch (const std::invalid_argument&) {
        c.fail("bad value '" + tok + "'");
    }
    v.kind = TomlValue::NUMBER;
    return v;
}

}  // namespace

Toml Toml::parse(const std::string& text) {
    Toml t;
    Cursor c(text);
    std::string table;
    t.tables[table];
    for (;;) {
        c.skip(true);
        if (c.eof()) break;
        if (c.peek() == '[') {
            c.get();
            table.clear();
            while (!c.eof() && c.peek() != ']') table.push_back(c.get());
            if (c.eof()) c.fail("unterminated table header");
            c.get();
            while (!table.empty() && std::isspace((unsigned char)table.back())) table.pop_back();
            while (!table.empty() && std::isspace((unsigned char)table.front())) table.erase(table.begin());
            t.tables[table];
            continue;
        }
        std::string key;
        while (!c.eof() && c.peek() != '=' && c.peek() != '\n') key.push_back(c.get());
        if (c.peek() != '=') c.fail("expected '=' after key");
        c.get();
        while (!key.empty() && std::isspace((unsigned char)key.back())) key.pop_back();
        while (!key.empty() && std::isspace((unsigned char)key.front())) key.erase(key.begin());
        if (key.size() >= 2 && key.front() == '"' && key.back() == '"') key = key.substr(1, key.size() - 2);
        t.tables[table][key] = parse_value(c);
        c.skip(false);
        if (!c.eof() && c.peek() != '\n') c.fail("trailing characters after value of '" + key + "'");
    }
    return t;
}

Toml Toml::parse_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

const TomlValue* Toml::get(const std::string& table, const std::string& key) const {
    auto t = tables.find(table);
    if (t == tables.end()) return nullptr;
    auto k = t->second.find(key);
    return k == t->second.end() ? nullptr : &k->second;
}

std::string Toml::get_string(const std::string& table, const std::string& key, const std::string& def) const {
    const TomlValue* v = get(table, key);
    return v && v->kind == TomlValue::STRING ? v->str : def;
}

double Toml::get_number(const std::string& table, const std::string& key, double def) const {
    const TomlValue* v = get(table, key);
    return v && v->kind == TomlValue::NUMBER ? v->num : def;
}

bool Toml::get_bool(const std::string& table, const std::string& key, bool def) const {
    const TomlValue* v = get(table, key);
    return v && v->kind == TomlValue::BOOL ? v->boolean : def;
}

}  // namespace uqk_host
