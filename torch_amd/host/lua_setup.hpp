// lua_setup.hpp -- interpreter for the Lua 5.2 subset that TORCH set-up
// files use (config/torch-setup.lua): a chunk of global and local
// assignments, function definitions, if / elseif / else, while and numeric
// for loops, multiple assignment and multiple returns, closures, the
// arithmetic, comparison, logical and concatenation operators, and the math
// library.  Numbers are C doubles and every operation is evaluated in source
// order with the C library's functions (pow for ^ and math.pow, fmod-free %
// as Lua 5.2 defines it), so values equal those of the reference's Lua
// build on the same libm.  It stands in for Torch::setUpLua's Selene/Lua
// state (src/Torch/Torch.cpp:192-228), which is shipped only as a prebuilt
// library.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace tg {
namespace lua {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Expr;
struct Block;
struct Scope;
struct Table;
struct Value;
using Builtin = std::vector<Value> (*)(const std::vector<Value>&);

struct FuncProto {
    std::vector<std::string> params;
    std::shared_ptr<Block> body;
    int line = 0;
};

struct Closure {
    std::shared_ptr<FuncProto> proto;
    std::shared_ptr<Scope> env;
};

struct Value {
    enum Kind { NIL, NUMBER, BOOLEAN, STRING, FUNCTION, BUILTIN, TABLE } kind = NIL;
    double num = 0;
    bool boolean = false;
    std::string str;
    std::shared_ptr<Closure> fn;
    Builtin builtin = nullptr;
    std::shared_ptr<Table> table;

    static Value number(double x);
    bool truthy() const { return !(kind == NIL || (kind == BOOLEAN && !boolean)); }
};

struct Table {
    std::map<std::string, Value> fields;  // string keys
    std::map<double, Value> items;        // numeric keys
};

// A loaded set-up chunk: globals after running the file, and calls of its
// functions with numeric arguments.
class State {
public:
    State();
    ~State();
    void load_file(const std::string& filename);
    void load_string(const std::string& text, const std::string& chunkname = "chunk");
    // global function `name` called with numbers; returns its results
    std::vector<Value> call(const std::string& name, const std::vector<double>& args);
    // global value (NIL when missing)
    Value global(const std::string& name) const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace lua
}  // namespace tg
