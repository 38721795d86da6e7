// lua_config.cpp -- recursive-descent reader of the Lua parameter subset.
#include "lua_config.hpp"

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

namespace tg {

namespace {

const LuaValue& nil_value() {
    static const LuaValue nil;
    return nil;
}

struct Lexer {
    enum Tok { END, NAME, NUMBER, STRING, SYM };
    const std::string* src;
    size_t i = 0;
    int line = 1;
    Tok tok = END;
    std::string text;
    double num = 0;

    explicit Lexer(const std::string& text_) : src(&text_) { next(); }
    const std::string& s() const { return *src; }

    [[noreturn]] void fail(const std::string& msg) const {
        throw LuaError("lua config line " + std::to_string(line) + ": " + msg);
    }

    void skip_space() {
        for (;;) {
            while (i < s().size() && std::isspace((unsigned char)s()[i])) {
                if (s()[i] == '\n') ++line;
                ++i;
            }
            if (i + 1 < s().size() && s()[i] == '-' && s()[i + 1] == '-') {
                i += 2;
                if (i + 1 < s().size() && s()[i] == '[' && s()[i + 1] == '[') {
                    size_t e = s().find("]]", i + 2);
                    for (size_t k = i; k < (e == std::string::npos ? s().size() : e); ++k)
                        if (s()[k] == '\n') ++line;
                    i = (e == std::string::npos) ? s().size() : e + 2;
                } else {
                    while (i < s().size() && s()[i] != '\n') ++i;
                }
                continue;
            }
            break;
        }
    }

    void next() {
        skip_space();
        text.clear();
        if (i >= s().size()) {
            tok = END;
            return;
        }
        char c = s()[i];
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t b = i;
            while (i < s().size() && (std::isalnum((unsigned char)s()[i]) || s()[i] == '_')) ++i;
            text = s().substr(b, i - b);
            tok = NAME;
        } else if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < s().size() && std::isdigit((unsigned char)s()[i + 1]))) {
            size_t used = 0;
            num = std::stod(s().substr(i), &used);
            i += used;
            tok = NUMBER;
        } else if (c == '"' || c == '\'') {
            char q = c;
            ++i;
            while (i < s().size() && s()[i] != q) {
                if (s()[i] == '\\' && i + 1 < s().size()) {
                    ++i;
                    char e = s()[i];
                    text += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
                } else {
                    text += s()[i];
                }
                ++i;
            }
            if (i >= s().size()) fail("unterminated string");
            ++i;
            tok = STRING;
        } else {
            if (c == '.' && i + 1 < s().size() && s()[i + 1] == '.') {
                text = "..";
                i += 2;
            } else {
                text = std::string(1, c);
                ++i;
            }
            tok = SYM;
        }
    }

    bool sym(const char* x) const { return tok == SYM && text == x; }
    void expect(const char* x) {
        if (!sym(x)) fail(std::string("expected '") + x + "' near '" + text + "'");
        next();
    }
};

class Parser {
public:
    Parser(const std::string& src, LuaValue& globals) : lx_(src), g_(globals) {}

    void chunk() {
        while (lx_.tok != Lexer::END) {
            if (lx_.tok == Lexer::NAME && lx_.text == "local") lx_.next();
            if (lx_.tok != Lexer::NAME) lx_.fail("expected a global assignment near '" + lx_.text + "'");
            std::string name = lx_.text;
            lx_.next();
            lx_.expect("=");
            auto v = std::make_shared<LuaValue>(exp());
            g_.kind = LuaValue::TABLE;
            g_.fields[name] = v;
            if (lx_.sym(";") || lx_.sym(",")) lx_.next();
        }
    }

private:
    Lexer lx_;
    LuaValue& g_;

    static double as_num(const LuaValue& v, Lexer& lx) {
        if (v.kind != LuaValue::NUMBER) lx.fail("arithmetic on a non-number");
        return v.num;
    }
    static LuaValue number(double x) {
        LuaValue v;
        v.kind = LuaValue::NUMBER;
        v.num = x;
        return v;
    }

    LuaValue exp() {
        LuaValue a = term();
        while (lx_.sym("+") || lx_.sym("-") || lx_.sym("..")) {
            std::string op = lx_.text;
            lx_.next();
            LuaValue b = term();
            if (op == "..") {
                LuaValue r;
                r.kind = LuaValue::STRING;
                auto str = [](const LuaValue& v) {
                    if (v.kind == LuaValue::STRING) return v.str;
                    std::ostringstream os;
                    os.precision(14);
                    os << v.num;
                    return os.str();
                };
                r.str = str(a) + str(b);
                a = r;
            } else {
                a = number(op == "+" ? as_num(a, lx_) + as_num(b, lx_) : as_num(a, lx_) - as_num(b, lx_));
            }
        }
        return a;
    }
    LuaValue term() {
        LuaValue a = unary();
        while (lx_.sym("*") || lx_.sym("/") || lx_.sym("%")) {
            std::string op = lx_.text;
            lx_.next();
            LuaValue b = unary();
            double x = as_num(a, lx_), y = as_num(b, lx_);
            a = number(op == "*" ? x * y : op == "/" ? x / y : x - std::floor(x / y) * y);
        }
        return a;
    }
    LuaValue unary() {
        if (lx_.sym("-")) {
            lx_.next();
            LuaValue v = unary();
            return number(-as_num(v, lx_));
        }
        if (lx_.tok == Lexer::NAME && lx_.text == "not") {
            lx_.next();
            LuaValue v = unary();
            LuaValue r;
            r.kind = LuaValue::BOOLEAN;
            r.boolean = (v.kind == LuaValue::NIL) || (v.kind == LuaValue::BOOLEAN && !v.boolean);
            return r;
        }
        LuaValue b = primary();
        if (lx_.sym("^")) {  // right associative, binds tighter than unary minus
            lx_.next();
            LuaValue e = unary();
            return number(std::pow(as_num(b, lx_), as_num(e, lx_)));
        }
        return b;
    }
    LuaValue primary() {
        if (lx_.tok == Lexer::NUMBER) {
            double x = lx_.num;
            lx_.next();
            return number(x);
        }
        if (lx_.tok == Lexer::STRING) {
            LuaValue v;
            v.kind = LuaValue::STRING;
            v.str = lx_.text;
            lx_.next();
            return v;
        }
        if (lx_.sym("(")) {
            lx_.next();
            LuaValue v = exp();
            lx_.expect(")");
            return v;
        }
        if (lx_.sym("{")) return table();
        if (lx_.tok == Lexer::NAME) {
            std::string name = lx_.text;
            lx_.next();
            LuaValue v;
            if (name == "true" || name == "false") {
                v.kind = LuaValue::BOOLEAN;
                v.boolean = (name == "true");
                return v;
            }
            if (name == "nil") return v;
            const LuaValue* cur = &g_[name];
            while (lx_.sym(".")) {
                lx_.next();
                if (lx_.tok != Lexer::NAME) lx_.fail("expected a field name");
                cur = &(*cur)[lx_.text];
                lx_.next();
            }
            return *cur;
        }
        lx_.fail("unexpected '" + lx_.text + "'");
    }
    LuaValue table() {
        lx_.expect("{");
        LuaValue t;
        t.kind = LuaValue::TABLE;
        while (!lx_.sym("}")) {
            if (lx_.tok == Lexer::NAME) {
                // lookahead: Name '=' is a keyed field
                Lexer save = lx_;
                std::string name = lx_.text;
                lx_.next();
                if (lx_.sym("=")) {
                    lx_.next();
                    t.fields[name] = std::make_shared<LuaValue>(exp());
                } else {
                    lx_ = save;
                    t.items.push_back(std::make_shared<LuaValue>(exp()));
                }
            } else if (lx_.sym("[")) {
                lx_.next();
                LuaValue k = exp();
                lx_.expect("]");
                lx_.expect("=");
                std::string key = k.kind == LuaValue::STRING ? k.str : std::to_string((long long)k.num);
                t.fields[key] = std::make_shared<LuaValue>(exp());
            } else {
                t.items.push_back(std::make_shared<LuaValue>(exp()));
            }
            if (lx_.sym(",") || lx_.sym(";")) lx_.next();
            else if (!lx_.sym("}")) lx_.fail("expected ',' or '}' near '" + lx_.text + "'");
        }
        lx_.next();
        return t;
    }
};

}  // namespace

const LuaValue& LuaValue::operator[](const std::string& key) const {
    if (kind != TABLE) return nil_value();
    auto it = fields.find(key);
    return it == fields.end() ? nil_value() : *it->second;
}

LuaConfig LuaConfig::parse_string(const std::string& text) {
    LuaConfig c;
    c.globals_.kind = LuaValue::TABLE;
    Parser p(text, c.globals_);
    p.chunk();
    return c;
}

LuaConfig LuaConfig::parse_file(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw LuaError("ParseParameters: could not open lua file: " + filename);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_string(ss.str());
}

const LuaValue& LuaConfig::operator[](const std::string& name) const { return globals_[name]; }

}  // namespace tg
