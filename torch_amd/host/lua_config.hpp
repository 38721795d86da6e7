// lua_config.hpp -- reader for the subset of Lua that TORCH parameter files
// use (config/torch-config.lua): global assignments of numbers, strings,
// booleans and nested tables, arithmetic (+ - * / ^, unary minus,
// parentheses) and references to earlier globals, with -- and --[[ ]]
// comments.  It stands in for the Lua 5.2 + Selene front end of the
// reference (src/main.cpp:126-208, src/IO/ParseLua.hpp), which is shipped
// only as a prebuilt library.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace tg {

struct LuaValue {
    enum Kind { NIL, NUMBER, STRING, BOOLEAN, TABLE } kind = NIL;
    double num = 0;
    std::string str;
    bool boolean = false;
    std::map<std::string, std::shared_ptr<LuaValue>> fields;
    std::vector<std::shared_ptr<LuaValue>> items;

    // Selector-style lookup: returns NIL for a missing key.
    const LuaValue& operator[](const std::string& key) const;
    bool exists() const { return kind != NIL; }
};

class LuaConfig {
public:
    static LuaConfig parse_file(const std::string& filename);
    static LuaConfig parse_string(const std::string& text);
    const LuaValue& operator[](const std::string& name) const;

private:
    LuaValue globals_;
};

struct LuaError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace tg
