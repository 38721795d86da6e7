// lua_setup.cpp -- the Lua 5.2 subset interpreter of lua_setup.hpp: a
// lexer, a recursive-descent parser to a small syntax tree, and a
// tree-walking evaluator with lexical scopes.
#include "lua_setup.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tg {
namespace lua {

Value Value::number(double x) {
    Value v;
    v.kind = NUMBER;
    v.num = x;
    return v;
}

namespace {

Value boolean(bool b) {
    Value v;
    v.kind = Value::BOOLEAN;
    v.boolean = b;
    return v;
}

Value string_value(const std::string& s) {
    Value v;
    v.kind = Value::STRING;
    v.str = s;
    return v;
}

// ------------------------------------------------------------------ lexer --
struct Token {
    enum Kind { END, NUM, STR, NAME, OP } kind = END;
    std::string text;
    double num = 0;
    int line = 1;
};

class Lexer {
public:
    Lexer(const std::string& src, const std::string& name) : s_(src), name_(name) { advance(); }
    const Token& tok() const { return t_; }
    void advance() { t_ = scan(); }
    [[noreturn]] void fail(const std::string& msg) const {
        throw Error(name_ + ":" + std::to_string(t_.line) + ": " + msg + " near '" + t_.text + "'");
    }
    bool is_op(const char* op) const { return t_.kind == Token::OP && t_.text == op; }
    bool is_kw(const char* kw) const { return t_.kind == Token::NAME && t_.text == kw; }
    void expect_op(const char* op) {
        if (!is_op(op)) fail(std::string("'") + op + "' expected");
        advance();
    }
    void expect_kw(const char* kw) {
        if (!is_kw(kw)) fail(std::string("'") + kw + "' expected");
        advance();
    }
    std::string name() {
        if (t_.kind != Token::NAME || keyword(t_.text)) fail("name expected");
        std::string n = t_.text;
        advance();
        return n;
    }
    static bool keyword(const std::string& w) {
        static const char* kws[] = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "goto",
                                    "if", "in", "local", "nil", "not", "or", "repeat", "return", "then", "true",
                                    "until", "while"};
        for (const char* k : kws)
            if (w == k) return true;
        return false;
    }

private:
    void skip() {
        for (;;) {
            while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) {
                if (s_[p_] == '\n') ++line_;
                ++p_;
            }
            if (s_.compare(p_, 2, "--") != 0) return;
            p_ += 2;
            int level = long_bracket();
            if (level >= 0) {
                long_string(level);
            } else {
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            }
        }
    }
    // [[ or [=*[ at p_: returns the level and moves past it, else -1
    int long_bracket() {
        if (p_ >= s_.size() || s_[p_] != '[') return -1;
        size_t q = p_ + 1;
        int level = 0;
        while (q < s_.size() && s_[q] == '=') ++level, ++q;
        if (q < s_.size() && s_[q] == '[') {
            p_ = q + 1;
            return level;
        }
        return -1;
    }
    std::string long_string(int level) {
        const std::string close = "]" + std::string(level, '=') + "]";
        size_t e = s_.find(close, p_);
        if (e == std::string::npos) throw Error(name_ + ":" + std::to_string(line_) + ": unfinished long string");
        std::string out = s_.substr(p_, e - p_);
        for (char ch : out)
            if (ch == '\n') ++line_;
        if (!out.empty() && out[0] == '\n') out.erase(0, 1);
        p_ = e + close.size();
        return out;
    }
    Token scan() {
        skip();
        Token t;
        t.line = line_;
        if (p_ >= s_.size()) return t;
        const char c = s_[p_];
        if (std::isdigit((unsigned char)c) || (c == '.' && p_ + 1 < s_.size() && std::isdigit((unsigned char)s_[p_ + 1]))) {
            size_t q = p_;
            if (c == '0' && q + 1 < s_.size() && (s_[q + 1] == 'x' || s_[q + 1] == 'X')) {
                q += 2;
                while (q < s_.size() && (std::isxdigit((unsigned char)s_[q]) || s_[q] == '.')) ++q;
            } else {
                while (q < s_.size() && (std::isdigit((unsigned char)s_[q]) || s_[q] == '.')) ++q;
                if (q < s_.size() && (s_[q] == 'e' || s_[q] == 'E')) {
                    ++q;
                    if (q < s_.size() && (s_[q] == '+' || s_[q] == '-')) ++q;
                    while (q < s_.size() && std::isdigit((unsigned char)s_[q])) ++q;
                }
            }
            t.kind = Token::NUM;
            t.text = s_.substr(p_, q - p_);
            char* end = nullptr;
            t.num = std::strtod(t.text.c_str(), &end);  // lua_str2number
            if (!end || *end) {
                t_ = t;
                fail("malformed number");
            }
            p_ = q;
            return t;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t q = p_;
            while (q < s_.size() && (std::isalnum((unsigned char)s_[q]) || s_[q] == '_')) ++q;
            t.kind = Token::NAME;
            t.text = s_.substr(p_, q - p_);
            p_ = q;
            return t;
        }
        if (c == '"' || c == '\'') {
            std::string out;
            size_t q = p_ + 1;
            while (q < s_.size() && s_[q] != c) {
                char ch = s_[q++];
                if (ch == '\n') throw Error(name_ + ":" + std::to_string(line_) + ": unfinished string");
                if (ch == '\\' && q < s_.size()) {
                    const char e = s_[q++];
                    switch (e) {
                    case 'n': ch = '\n'; break;
                    case 't': ch = '\t'; break;
                    case 'r': ch = '\r'; break;
                    case 'a': ch = '\a'; break;
                    case 'b': ch = '\b'; break;
                    case 'f': ch = '\f'; break;
                    case 'v': ch = '\v'; break;
                    default: ch = e;
                    }
                }
                out.push_back(ch);
            }
            if (q >= s_.size()) throw Error(name_ + ":" + std::to_string(line_) + ": unfinished string");
            p_ = q + 1;
            t.kind = Token::STR;
            t.text = out;
            return t;
        }
        if (c == '[') {
            const int level = long_bracket();
            if (level >= 0) {
                t.kind = Token::STR;
                t.text = long_string(level);
                return t;
            }
        }
        static const char* ops3[] = {"..."};
        static const char* ops2[] = {"==", "~=", "<=", ">=", "..", "::"};
        for (const char* op : ops3)
            if (s_.compare(p_, 3, op) == 0) {
                t.kind = Token::OP;
                t.text = op;
                p_ += 3;
                return t;
            }
        for (const char* op : ops2)
            if (s_.compare(p_, 2, op) == 0) {
                t.kind = Token::OP;
                t.text = op;
                p_ += 2;
                return t;
            }
        if (std::strchr("+-*/%^#<>=(){}[];:,.", c)) {
            t.kind = Token::OP;
            t.text = std::string(1, c);
            ++p_;
            return t;
        }
        t.text = std::string(1, c);
        t_ = t;
        fail("unexpected symbol");
    }

    const std::string& s_;
    std::string name_;
    size_t p_ = 0;
    int line_ = 1;
    Token t_;
};

}  // namespace

// ---------------------------------------------------------- syntax tree --
struct Expr {
    enum Kind { NUM, STR, NIL, TRUE, FALSE, NAME, INDEX, CALL, BIN, AND, OR, UN, FUNC, TABLE, PAREN } kind;
    int line = 0;
    double num = 0;
    std::string str;  // NAME, STR, operator of BIN / UN
    std::shared_ptr<Expr> a, b;
    std::vector<std::shared_ptr<Expr>> list;  // CALL arguments, TABLE values
    std::vector<std::shared_ptr<Expr>> keys;  // TABLE keys (null: positional)
    std::shared_ptr<FuncProto> proto;
};
using ExprP = std::shared_ptr<Expr>;

struct Stmt {
    enum Kind { LOCAL, ASSIGN, CALL, IF, WHILE, REPEAT, NUMFOR, RETURN, BREAK, DO, LOCALFUNC } kind;
    int line = 0;
    std::vector<std::string> names;        // LOCAL, NUMFOR (loop variable), LOCALFUNC
    std::vector<ExprP> targets;            // ASSIGN
    std::vector<ExprP> exprs;              // values, conditions, for start/limit/step
    std::vector<std::shared_ptr<Block>> blocks;
    std::shared_ptr<FuncProto> proto;
};
using StmtP = std::shared_ptr<Stmt>;

struct Block {
    std::vector<StmtP> stmts;
};

struct Scope {
    std::map<std::string, Value> vars;
    std::shared_ptr<Scope> parent;
};

namespace {

ExprP mk(Expr::Kind k, int line) {
    auto e = std::make_shared<Expr>();
    e->kind = k;
    e->line = line;
    return e;
}

class Parser {
public:
    explicit Parser(Lexer& lx) : lx_(lx) {}

    std::shared_ptr<Block> chunk() {
        auto b = block();
        if (lx_.tok().kind != Token::END) lx_.fail("'<eof>' expected");
        return b;
    }

private:
    bool block_end() const {
        return lx_.tok().kind == Token::END || lx_.is_kw("end") || lx_.is_kw("else") || lx_.is_kw("elseif") ||
               lx_.is_kw("until");
    }

    std::shared_ptr<Block> block() {
        auto b = std::make_shared<Block>();
        while (!block_end()) {
            if (lx_.is_op(";")) {
                lx_.advance();
                continue;
            }
            if (lx_.is_kw("return")) {
                auto s = std::make_shared<Stmt>();
                s->kind = Stmt::RETURN;
                s->line = lx_.tok().line;
                lx_.advance();
                if (!block_end() && !lx_.is_op(";")) s->exprs = exprlist();
                if (lx_.is_op(";")) lx_.advance();
                b->stmts.push_back(s);
                if (!block_end()) lx_.fail("'end' expected");
                break;
            }
            b->stmts.push_back(statement());
        }
        return b;
    }

    std::shared_ptr<FuncProto> funcbody(int line) {
        auto f = std::make_shared<FuncProto>();
        f->line = line;
        lx_.expect_op("(");
        if (!lx_.is_op(")")) {
            for (;;) {
                if (lx_.is_op("...")) lx_.fail("varargs are not supported");
                f->params.push_back(lx_.name());
                if (!lx_.is_op(",")) break;
                lx_.advance();
            }
        }
        lx_.expect_op(")");
        f->body = block();
        lx_.expect_kw("end");
        return f;
    }

    StmtP statement() {
        auto s = std::make_shared<Stmt>();
        s->line = lx_.tok().line;
        if (lx_.is_kw("local")) {
            lx_.advance();
            if (lx_.is_kw("function")) {
                lx_.advance();
                s->kind = Stmt::LOCALFUNC;
                s->names.push_back(lx_.name());
                s->proto = funcbody(s->line);
                return s;
            }
            s->kind = Stmt::LOCAL;
            for (;;) {
                s->names.push_back(lx_.name());
                if (!lx_.is_op(",")) break;
                lx_.advance();
            }
            if (lx_.is_op("=")) {
                lx_.advance();
                s->exprs = exprlist();
            }
            return s;
        }
        if (lx_.is_kw("function")) {
            // function a.b.c(...) -> a.b.c = function(...)
            lx_.advance();
            ExprP target = mk(Expr::NAME, s->line);
            target->str = lx_.name();
            while (lx_.is_op(".")) {
                lx_.advance();
                ExprP idx = mk(Expr::INDEX, s->line);
                idx->a = target;
                idx->b = mk(Expr::STR, s->line);
                idx->b->str = lx_.name();
                target = idx;
            }
            if (lx_.is_op(":")) lx_.fail("methods are not supported");
            ExprP fn = mk(Expr::FUNC, s->line);
            fn->proto = funcbody(s->line);
            s->kind = Stmt::ASSIGN;
            s->targets.push_back(target);
            s->exprs.push_back(fn);
            return s;
        }
        if (lx_.is_kw("if")) {
            s->kind = Stmt::IF;
            lx_.advance();
            s->exprs.push_back(expr());
            lx_.expect_kw("then");
            s->blocks.push_back(block());
            while (lx_.is_kw("elseif")) {
                lx_.advance();
                s->exprs.push_back(expr());
                lx_.expect_kw("then");
                s->blocks.push_back(block());
            }
            if (lx_.is_kw("else")) {
                lx_.advance();
                s->blocks.push_back(block());
            }
            lx_.expect_kw("end");
            return s;
        }
        if (lx_.is_kw("while")) {
            s->kind = Stmt::WHILE;
            lx_.advance();
            s->exprs.push_back(expr());
            lx_.expect_kw("do");
            s->blocks.push_back(block());
            lx_.expect_kw("end");
            return s;
        }
        if (lx_.is_kw("repeat")) {
            s->kind = Stmt::REPEAT;
            lx_.advance();
            s->blocks.push_back(block());
            lx_.expect_kw("until");
            s->exprs.push_back(expr());
            return s;
        }
        if (lx_.is_kw("for")) {
            lx_.advance();
            s->kind = Stmt::NUMFOR;
            s->names.push_back(lx_.name());
            if (!lx_.is_op("=")) lx_.fail("only numeric for loops are supported");
            lx_.advance();
            s->exprs.push_back(expr());
            lx_.expect_op(",");
            s->exprs.push_back(expr());
            if (lx_.is_op(",")) {
                lx_.advance();
                s->exprs.push_back(expr());
            }
            lx_.expect_kw("do");
            s->blocks.push_back(block());
            lx_.expect_kw("end");
            return s;
        }
        if (lx_.is_kw("do")) {
            s->kind = Stmt::DO;
            lx_.advance();
            s->blocks.push_back(block());
            lx_.expect_kw("end");
            return s;
        }
        if (lx_.is_kw("break")) {
            s->kind = Stmt::BREAK;
            lx_.advance();
            return s;
        }
        // assignment or call
        ExprP e = suffixedexp();
        if (lx_.is_op("=") || lx_.is_op(",")) {
            s->kind = Stmt::ASSIGN;
            s->targets.push_back(e);
            while (lx_.is_op(",")) {
                lx_.advance();
                s->targets.push_back(suffixedexp());
            }
            lx_.expect_op("=");
            s->exprs = exprlist();
            for (const ExprP& t : s->targets)
                if (t->kind != Expr::NAME && t->kind != Expr::INDEX) lx_.fail("cannot assign");
            return s;
        }
        if (e->kind != Expr::CALL) lx_.fail("syntax error");
        s->kind = Stmt::CALL;
        s->exprs.push_back(e);
        return s;
    }

    std::vector<ExprP> exprlist() {
        std::vector<ExprP> v{expr()};
        while (lx_.is_op(",")) {
            lx_.advance();
            v.push_back(expr());
        }
        return v;
    }

    ExprP primaryexp() {
        const int line = lx_.tok().line;
        if (lx_.is_op("(")) {
            lx_.advance();
            ExprP e = mk(Expr::PAREN, line);
            e->a = expr();
            lx_.expect_op(")");
            return e;
        }
        ExprP e = mk(Expr::NAME, line);
        e->str = lx_.name();
        return e;
    }

    ExprP suffixedexp() {
        ExprP e = primaryexp();
        for (;;) {
            const int line = lx_.tok().line;
            if (lx_.is_op(".")) {
                lx_.advance();
                ExprP idx = mk(Expr::INDEX, line);
                idx->a = e;
                idx->b = mk(Expr::STR, line);
                idx->b->str = lx_.name();
                e = idx;
            } else if (lx_.is_op("[")) {
                lx_.advance();
                ExprP idx = mk(Expr::INDEX, line);
                idx->a = e;
                idx->b = expr();
                lx_.expect_op("]");
                e = idx;
            } else if (lx_.is_op("(") || lx_.tok().kind == Token::STR || lx_.is_op("{")) {
                ExprP call = mk(Expr::CALL, line);
                call->a = e;
                if (lx_.tok().kind == Token::STR) {
                    ExprP s = mk(Expr::STR, line);
                    s->str = lx_.tok().text;
                    lx_.advance();
                    call->list.push_back(s);
                } else if (lx_.is_op("{")) {
                    call->list.push_back(table());
                } else {
                    lx_.advance();
                    if (!lx_.is_op(")")) call->list = exprlist();
                    lx_.expect_op(")");
                }
                e = call;
            } else if (lx_.is_op(":")) {
                lx_.fail("methods are not supported");
            } else {
                return e;
            }
        }
    }

    ExprP table() {
        ExprP t = mk(Expr::TABLE, lx_.tok().line);
        lx_.expect_op("{");
        while (!lx_.is_op("}")) {
            if (lx_.is_op("[")) {
                lx_.advance();
                t->keys.push_back(expr());
                lx_.expect_op("]");
                lx_.expect_op("=");
                t->list.push_back(expr());
            } else if (lx_.tok().kind == Token::NAME && !Lexer::keyword(lx_.tok().text)) {
                // name = value, or a positional expression starting with a name
                ExprP e = expr();
                if (lx_.is_op("=") && e->kind == Expr::NAME) {
                    lx_.advance();
                    ExprP k = mk(Expr::STR, e->line);
                    k->str = e->str;
                    t->keys.push_back(k);
                    t->list.push_back(expr());
                } else {
                    t->keys.push_back(nullptr);
                    t->list.push_back(e);
                }
            } else {
                t->keys.push_back(nullptr);
                t->list.push_back(expr());
            }
            if (lx_.is_op(",") || lx_.is_op(";")) lx_.advance();
            else break;
        }
        lx_.expect_op("}");
        return t;
    }

    ExprP simpleexp() {
        const Token& t = lx_.tok();
        const int line = t.line;
        ExprP e;
        if (t.kind == Token::NUM) {
            e = mk(Expr::NUM, line);
            e->num = t.num;
        } else if (t.kind == Token::STR) {
            e = mk(Expr::STR, line);
            e->str = t.text;
        } else if (lx_.is_kw("nil")) {
            e = mk(Expr::NIL, line);
        } else if (lx_.is_kw("true")) {
            e = mk(Expr::TRUE, line);
        } else if (lx_.is_kw("false")) {
            e = mk(Expr::FALSE, line);
        } else if (lx_.is_kw("function")) {
            lx_.advance();
            e = mk(Expr::FUNC, line);
            e->proto = funcbody(line);
            return e;
        } else if (lx_.is_op("{")) {
            return table();
        } else {
            return suffixedexp();
        }
        lx_.advance();
        return e;
    }

    // binary operator priorities (left, right); unary operators bind at 8
    static bool binop(const Lexer& lx, std::string& op, int& left, int& right) {
        const Token& t = lx.tok();
        static const struct {
            const char* op;
            int l, r;
        } tab[] = {{"+", 6, 6}, {"-", 6, 6}, {"*", 7, 7}, {"/", 7, 7}, {"%", 7, 7}, {"^", 10, 9}, {"..", 5, 4},
                   {"==", 3, 3}, {"~=", 3, 3}, {"<", 3, 3}, {"<=", 3, 3}, {">", 3, 3}, {">=", 3, 3}};
        if (t.kind == Token::NAME && (t.text == "and" || t.text == "or")) {
            op = t.text;
            left = right = (t.text == "and") ? 2 : 1;
            return true;
        }
        if (t.kind != Token::OP) return false;
        for (const auto& e : tab)
            if (t.text == e.op) {
                op = e.op;
                left = e.l;
                right = e.r;
                return true;
            }
        return false;
    }

    ExprP subexpr(int limit) {
        ExprP e;
        const int line = lx_.tok().line;
        if (lx_.is_kw("not") || lx_.is_op("-") || lx_.is_op("#")) {
            const std::string op = lx_.tok().text;
            lx_.advance();
            e = mk(Expr::UN, line);
            e->str = op;
            e->a = subexpr(8);
        } else {
            e = simpleexp();
        }
        std::string op;
        int l, r;
        while (binop(lx_, op, l, r) && l > limit) {
            const int bl = lx_.tok().line;
            lx_.advance();
            ExprP rhs = subexpr(r);
            ExprP b = mk(op == "and" ? Expr::AND : op == "or" ? Expr::OR : Expr::BIN, bl);
            b->str = op;
            b->a = e;
            b->b = rhs;
            e = b;
        }
        return e;
    }

    ExprP expr() { return subexpr(0); }

    Lexer& lx_;
};

// ------------------------------------------------------------ builtins --
double arg_num(const std::vector<Value>& a, size_t i, const char* fn) {
    if (i >= a.size() || a[i].kind != Value::NUMBER)
        throw Error(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fn + "' (number expected)");
    return a[i].num;
}

#define MATH1(NAME, EXPR)                                                      \
    std::vector<Value> m_##NAME(const std::vector<Value>& a) {                \
        const double x = arg_num(a, 0, #NAME);                                 \
        return {Value::number(EXPR)};                                          \
    }
MATH1(abs, std::fabs(x))
MATH1(ceil, std::ceil(x))
MATH1(floor, std::floor(x))
MATH1(sqrt, std::sqrt(x))
MATH1(exp, std::exp(x))
MATH1(log10, std::log10(x))
MATH1(sin, std::sin(x))
MATH1(cos, std::cos(x))
MATH1(tan, std::tan(x))
MATH1(asin, std::asin(x))
MATH1(acos, std::acos(x))
MATH1(atan, std::atan(x))
MATH1(sinh, std::sinh(x))
MATH1(cosh, std::cosh(x))
MATH1(tanh, std::tanh(x))
MATH1(deg, x / (3.1415926535897932384626433832795 / 180.0))
MATH1(rad, x * (3.1415926535897932384626433832795 / 180.0))
#undef MATH1

std::vector<Value> m_pow(const std::vector<Value>& a) {
    return {Value::number(std::pow(arg_num(a, 0, "pow"), arg_num(a, 1, "pow")))};
}
std::vector<Value> m_atan2(const std::vector<Value>& a) {
    return {Value::number(std::atan2(arg_num(a, 0, "atan2"), arg_num(a, 1, "atan2")))};
}
std::vector<Value> m_fmod(const std::vector<Value>& a) {
    return {Value::number(std::fmod(arg_num(a, 0, "fmod"), arg_num(a, 1, "fmod")))};
}
std::vector<Value> m_log(const std::vector<Value>& a) {
    // Lua 5.2 math.log(x [, base])
    const double x = arg_num(a, 0, "log");
    if (a.size() < 2 || a[1].kind == Value::NIL) return {Value::number(std::log(x))};
    const double base = arg_num(a, 1, "log");
    if (base == 10.0) return {Value::number(std::log10(x))};
    return {Value::number(std::log(x) / std::log(base))};
}
std::vector<Value> m_min(const std::vector<Value>& a) {
    double m = arg_num(a, 0, "min");
    for (size_t i = 1; i < a.size(); ++i) {
        const double d = arg_num(a, i, "min");
        if (d < m) m = d;
    }
    return {Value::number(m)};
}
std::vector<Value> m_max(const std::vector<Value>& a) {
    double m = arg_num(a, 0, "max");
    for (size_t i = 1; i < a.size(); ++i) {
        const double d = arg_num(a, i, "max");
        if (d > m) m = d;
    }
    return {Value::number(m)};
}
std::vector<Value> m_random(const std::vector<Value>&) {
    throw Error("math.random: the reference's Lua random sequence cannot be reproduced; use a deterministic set-up");
}
std::vector<Value> b_print(const std::vector<Value>&) { return {}; }
std::vector<Value> b_error(const std::vector<Value>& a) {
    throw Error(a.empty() || a[0].kind != Value::STRING ? std::string("error") : a[0].str);
}
std::vector<Value> b_tonumber(const std::vector<Value>& a) {
    if (a.empty()) return {Value()};
    if (a[0].kind == Value::NUMBER) return {a[0]};
    if (a[0].kind == Value::STRING) {
        char* end = nullptr;
        const double d = std::strtod(a[0].str.c_str(), &end);
        if (end && !*end && !a[0].str.empty()) return {Value::number(d)};
    }
    return {Value()};
}
std::vector<Value> b_type(const std::vector<Value>& a) {
    static const char* names[] = {"nil", "number", "boolean", "string", "function", "function", "table"};
    return {string_value(names[a.empty() ? 0 : a[0].kind])};
}

}  // namespace

// ---------------------------------------------------------- evaluation --
struct State::Impl {
    std::shared_ptr<Table> globals = std::make_shared<Table>();
    std::shared_ptr<Block> chunk;
    std::string name = "chunk";

    enum Flow { NORMAL, BREAK, RETURN };

    [[noreturn]] void fail(int line, const std::string& msg) const {
        throw Error(name + ":" + std::to_string(line) + ": " + msg);
    }

    Value* find(const std::shared_ptr<Scope>& sc, const std::string& n) {
        for (Scope* s = sc.get(); s; s = s->parent.get()) {
            auto it = s->vars.find(n);
            if (it != s->vars.end()) return &it->second;
        }
        return nullptr;
    }

    static std::string tostr(const Value& v) {
        if (v.kind == Value::STRING) return v.str;
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.14g", v.num);  // LUAI_NUMFFORMAT
        return buf;
    }

    double tonum(const Value& v, int line, const char* what) const {
        if (v.kind == Value::NUMBER) return v.num;
        if (v.kind == Value::STRING) {
            char* end = nullptr;
            const double d = std::strtod(v.str.c_str(), &end);
            if (end && !*end && !v.str.empty()) return d;
        }
        fail(line, std::string("attempt to perform arithmetic on a ") +
                       (v.kind == Value::NIL ? "nil" : v.kind == Value::BOOLEAN ? "boolean" : "non-number") +
                       " value" + (what ? std::string(" (") + what + ")" : ""));
    }

    static bool equal(const Value& a, const Value& b) {
        if (a.kind != b.kind) return false;
        switch (a.kind) {
        case Value::NIL: return true;
        case Value::NUMBER: return a.num == b.num;
        case Value::BOOLEAN: return a.boolean == b.boolean;
        case Value::STRING: return a.str == b.str;
        case Value::FUNCTION: return a.fn == b.fn;
        case Value::BUILTIN: return a.builtin == b.builtin;
        case Value::TABLE: return a.table == b.table;
        }
        return false;
    }

    Value index(const Value& obj, const Value& key, int line) {
        if (obj.kind != Value::TABLE) fail(line, "attempt to index a non-table value");
        if (key.kind == Value::STRING) {
            auto it = obj.table->fields.find(key.str);
            return it == obj.table->fields.end() ? Value() : it->second;
        }
        if (key.kind == Value::NUMBER) {
            auto it = obj.table->items.find(key.num);
            return it == obj.table->items.end() ? Value() : it->second;
        }
        return Value();
    }

    void setindex(const Value& obj, const Value& key, const Value& v, int line) {
        if (obj.kind != Value::TABLE) fail(line, "attempt to index a non-table value");
        if (key.kind == Value::STRING) obj.table->fields[key.str] = v;
        else if (key.kind == Value::NUMBER) obj.table->items[key.num] = v;
        else fail(line, "unsupported table key");
    }

    std::vector<Value> call(const Value& f, const std::vector<Value>& args, int line) {
        if (f.kind == Value::BUILTIN) {
            try {
                return f.builtin(args);
            } catch (Error& e) {
                fail(line, e.what());
            }
        }
        if (f.kind != Value::FUNCTION) fail(line, "attempt to call a non-function value");
        auto sc = std::make_shared<Scope>();
        sc->parent = f.fn->env;
        const FuncProto& p = *f.fn->proto;
        for (size_t i = 0; i < p.params.size(); ++i) sc->vars[p.params[i]] = i < args.size() ? args[i] : Value();
        std::vector<Value> ret;
        if (++depth_ > 200) fail(line, "stack overflow");
        exec_block(*p.body, sc, ret, false);
        --depth_;
        return ret;
    }

    // all values of an expression (calls expand), else one value
    std::vector<Value> eval_multi(const ExprP& e, const std::shared_ptr<Scope>& sc) {
        if (e->kind == Expr::CALL) return eval_call(e, sc);
        return {eval(e, sc)};
    }

    std::vector<Value> eval_list(const std::vector<ExprP>& es, const std::shared_ptr<Scope>& sc) {
        std::vector<Value> out;
        for (size_t i = 0; i < es.size(); ++i) {
            if (i + 1 == es.size()) {
                std::vector<Value> m = eval_multi(es[i], sc);
                out.insert(out.end(), m.begin(), m.end());
            } else {
                out.push_back(eval(es[i], sc));
            }
        }
        return out;
    }

    std::vector<Value> eval_call(const ExprP& e, const std::shared_ptr<Scope>& sc) {
        const Value f = eval(e->a, sc);
        return call(f, eval_list(e->list, sc), e->line);
    }

    Value arith(const std::string& op, const Value& x, const Value& y, int line) {
        if (op == "..") {
            if ((x.kind != Value::STRING && x.kind != Value::NUMBER) || (y.kind != Value::STRING && y.kind != Value::NUMBER))
                fail(line, "attempt to concatenate a non-string value");
            return string_value(tostr(x) + tostr(y));
        }
        if (op == "==") return boolean(equal(x, y));
        if (op == "~=") return boolean(!equal(x, y));
        if (op == "<" || op == "<=" || op == ">" || op == ">=") {
            if (x.kind == Value::NUMBER && y.kind == Value::NUMBER) {
                const double a = x.num, b = y.num;
                return boolean(op == "<" ? a < b : op == "<=" ? a <= b : op == ">" ? a > b : a >= b);
            }
            if (x.kind == Value::STRING && y.kind == Value::STRING) {
                const int c = std::strcmp(x.str.c_str(), y.str.c_str());
                return boolean(op == "<" ? c < 0 : op == "<=" ? c <= 0 : op == ">" ? c > 0 : c >= 0);
            }
            fail(line, "attempt to compare two incompatible values");
        }
        const double a = tonum(x, line, nullptr), b = tonum(y, line, nullptr);
        switch (op[0]) {
        case '+': return Value::number(a + b);
        case '-': return Value::number(a - b);
        case '*': return Value::number(a * b);
        case '/': return Value::number(a / b);
        case '%': return Value::number(a - std::floor(a / b) * b);  // luai_nummod
        case '^': return Value::number(std::pow(a, b));              // luai_numpow
        }
        fail(line, "unknown operator " + op);
    }

    Value eval(const ExprP& e, const std::shared_ptr<Scope>& sc) {
        switch (e->kind) {
        case Expr::NUM: return Value::number(e->num);
        case Expr::STR: return string_value(e->str);
        case Expr::NIL: return Value();
        case Expr::TRUE: return boolean(true);
        case Expr::FALSE: return boolean(false);
        case Expr::NAME: {
            if (Value* v = find(sc, e->str)) return *v;
            auto it = globals->fields.find(e->str);
            return it == globals->fields.end() ? Value() : it->second;
        }
        case Expr::INDEX: return index(eval(e->a, sc), eval(e->b, sc), e->line);
        case Expr::CALL: {
            std::vector<Value> r = eval_call(e, sc);
            return r.empty() ? Value() : r[0];
        }
        case Expr::PAREN: return eval(e->a, sc);
        case Expr::AND: {
            Value a = eval(e->a, sc);
            return a.truthy() ? eval(e->b, sc) : a;
        }
        case Expr::OR: {
            Value a = eval(e->a, sc);
            return a.truthy() ? a : eval(e->b, sc);
        }
        case Expr::BIN: {
            const Value a = eval(e->a, sc);
            const Value b = eval(e->b, sc);
            return arith(e->str, a, b, e->line);
        }
        case Expr::UN: {
            const Value a = eval(e->a, sc);
            if (e->str == "not") return boolean(!a.truthy());
            if (e->str == "-") return Value::number(-tonum(a, e->line, nullptr));
            if (a.kind == Value::STRING) return Value::number((double)a.str.size());
            if (a.kind == Value::TABLE) {
                double n = 0;
                while (a.table->items.count(n + 1)) n += 1;
                return Value::number(n);
            }
            fail(e->line, "attempt to get length of a non-table value");
        }
        case Expr::FUNC: {
            Value v;
            v.kind = Value::FUNCTION;
            v.fn = std::make_shared<Closure>();
            v.fn->proto = e->proto;
            v.fn->env = sc;
            return v;
        }
        case Expr::TABLE: {
            Value t;
            t.kind = Value::TABLE;
            t.table = std::make_shared<Table>();
            double pos = 1;
            for (size_t i = 0; i < e->list.size(); ++i) {
                if (e->keys[i]) {
                    setindex(t, eval(e->keys[i], sc), eval(e->list[i], sc), e->line);
                } else if (i + 1 == e->list.size()) {
                    for (const Value& v : eval_multi(e->list[i], sc)) t.table->items[pos++] = v;
                } else {
                    t.table->items[pos++] = eval(e->list[i], sc);
                }
            }
            return t;
        }
        }
        fail(e->line, "bad expression");
    }

    void assign(const ExprP& target, const Value& v, const std::shared_ptr<Scope>& sc) {
        if (target->kind == Expr::NAME) {
            if (Value* slot = find(sc, target->str)) *slot = v;
            else globals->fields[target->str] = v;
        } else {
            setindex(eval(target->a, sc), eval(target->b, sc), v, target->line);
        }
    }

    Flow exec_block(const Block& b, const std::shared_ptr<Scope>& parent, std::vector<Value>& ret, bool new_scope = true) {
        std::shared_ptr<Scope> sc = parent;
        if (new_scope) {
            sc = std::make_shared<Scope>();
            sc->parent = parent;
        }
        for (const StmtP& s : b.stmts) {
            const Flow f = exec(*s, sc, ret);
            if (f != NORMAL) return f;
        }
        return NORMAL;
    }

    Flow exec(const Stmt& s, std::shared_ptr<Scope>& sc, std::vector<Value>& ret) {
        switch (s.kind) {
        case Stmt::LOCAL: {
            std::vector<Value> vals = eval_list(s.exprs, sc);
            // each `local` statement opens the names for the statements after it
            auto inner = std::make_shared<Scope>();
            inner->parent = sc;
            for (size_t i = 0; i < s.names.size(); ++i) inner->vars[s.names[i]] = i < vals.size() ? vals[i] : Value();
            sc = inner;
            return NORMAL;
        }
        case Stmt::LOCALFUNC: {
            auto inner = std::make_shared<Scope>();
            inner->parent = sc;
            sc = inner;
            Value v;
            v.kind = Value::FUNCTION;
            v.fn = std::make_shared<Closure>();
            v.fn->proto = s.proto;
            v.fn->env = sc;  // visible to itself (recursion)
            sc->vars[s.names[0]] = v;
            return NORMAL;
        }
        case Stmt::ASSIGN: {
            std::vector<Value> vals = eval_list(s.exprs, sc);
            for (size_t i = 0; i < s.targets.size(); ++i) assign(s.targets[i], i < vals.size() ? vals[i] : Value(), sc);
            return NORMAL;
        }
        case Stmt::CALL: eval_call(s.exprs[0], sc); return NORMAL;
        case Stmt::IF: {
            for (size_t i = 0; i < s.exprs.size(); ++i)
                if (eval(s.exprs[i], sc).truthy()) return exec_block(*s.blocks[i], sc, ret);
            if (s.blocks.size() > s.exprs.size()) return exec_block(*s.blocks.back(), sc, ret);
            return NORMAL;
        }
        case Stmt::WHILE: {
            while (eval(s.exprs[0], sc).truthy()) {
                const Flow f = exec_block(*s.blocks[0], sc, ret);
                if (f == BREAK) break;
                if (f == RETURN) return f;
            }
            return NORMAL;
        }
        case Stmt::REPEAT: {
            for (;;) {
                auto body = std::make_shared<Scope>();
                body->parent = sc;
                std::shared_ptr<Scope> cur = body;
                Flow f = NORMAL;
                for (const StmtP& st : s.blocks[0]->stmts) {
                    f = exec(*st, cur, ret);
                    if (f != NORMAL) break;
                }
                if (f == BREAK) break;
                if (f == RETURN) return f;
                if (eval(s.exprs[0], cur).truthy()) break;  // `until` sees the body's locals
            }
            return NORMAL;
        }
        case Stmt::NUMFOR: {
            // Lua 5.2 OP_FORPREP / OP_FORLOOP: idx = init - step; loop idx += step
            const double init = tonum(eval(s.exprs[0], sc), s.line, "'for' initial value");
            const double limit = tonum(eval(s.exprs[1], sc), s.line, "'for' limit");
            const double step = s.exprs.size() > 2 ? tonum(eval(s.exprs[2], sc), s.line, "'for' step") : 1.0;
            double idx = init - step;
            for (;;) {
                idx += step;
                if (!(0 < step ? idx <= limit : limit <= idx)) break;
                auto body = std::make_shared<Scope>();
                body->parent = sc;
                body->vars[s.names[0]] = Value::number(idx);
                const Flow f = exec_block(*s.blocks[0], body, ret, false);
                if (f == BREAK) break;
                if (f == RETURN) return f;
            }
            return NORMAL;
        }
        case Stmt::DO: return exec_block(*s.blocks[0], sc, ret);
        case Stmt::BREAK: return BREAK;
        case Stmt::RETURN: ret = eval_list(s.exprs, sc); return RETURN;
        }
        return NORMAL;
    }

    void run(const std::string& text, const std::string& chunkname) {
        name = chunkname;
        Lexer lx(text, chunkname);
        Parser ps(lx);
        chunk = ps.chunk();
        auto top = std::make_shared<Scope>();
        std::vector<Value> ret;
        exec_block(*chunk, top, ret, false);
    }

    int depth_ = 0;
};

State::State() : impl_(new Impl) {
    auto fn = [](Builtin b) {
        Value v;
        v.kind = Value::BUILTIN;
        v.builtin = b;
        return v;
    };
    Value math;
    math.kind = Value::TABLE;
    math.table = std::make_shared<Table>();
    auto& m = math.table->fields;
    m["abs"] = fn(m_abs);
    m["ceil"] = fn(m_ceil);
    m["floor"] = fn(m_floor);
    m["sqrt"] = fn(m_sqrt);
    m["exp"] = fn(m_exp);
    m["log"] = fn(m_log);
    m["log10"] = fn(m_log10);
    m["pow"] = fn(m_pow);
    m["sin"] = fn(m_sin);
    m["cos"] = fn(m_cos);
    m["tan"] = fn(m_tan);
    m["asin"] = fn(m_asin);
    m["acos"] = fn(m_acos);
    m["atan"] = fn(m_atan);
    m["atan2"] = fn(m_atan2);
    m["sinh"] = fn(m_sinh);
    m["cosh"] = fn(m_cosh);
    m["tanh"] = fn(m_tanh);
    m["fmod"] = fn(m_fmod);
    m["deg"] = fn(m_deg);
    m["rad"] = fn(m_rad);
    m["min"] = fn(m_min);
    m["max"] = fn(m_max);
    m["random"] = fn(m_random);
    m["randomseed"] = fn(m_random);
    m["pi"] = Value::number(3.1415926535897932384626433832795);
    m["huge"] = Value::number(HUGE_VAL);
    auto& g = impl_->globals->fields;
    g["math"] = math;
    g["print"] = fn(b_print);
    g["error"] = fn(b_error);
    g["tonumber"] = fn(b_tonumber);
    g["type"] = fn(b_type);
}

State::~State() = default;

void State::load_string(const std::string& text, const std::string& chunkname) { impl_->run(text, chunkname); }

void State::load_file(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw Error("cannot open " + filename);
    std::stringstream ss;
    ss << f.rdbuf();
    impl_->run(ss.str(), filename);
}

std::vector<Value> State::call(const std::string& name, const std::vector<double>& args) {
    auto it = impl_->globals->fields.find(name);
    if (it == impl_->globals->fields.end()) throw Error("attempt to call a nil value (global '" + name + "')");
    std::vector<Value> a;
    for (double d : args) a.push_back(Value::number(d));
    return impl_->call(it->second, a, 0);
}

Value State::global(const std::string& name) const {
    auto it = impl_->globals->fields.find(name);
    return it == impl_->globals->fields.end() ? Value() : it->second;
}

}  // namespace lua
}  // namespace tg
