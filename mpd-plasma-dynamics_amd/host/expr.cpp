// expr.cpp -- see expr.h.
#include "expr.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace mpd {

namespace {

struct Tok {
    enum Kind { kNum, kId, kOp, kEnd } kind;
    double num = 0.0;
    std::string text;
    size_t pos = 0;
};

class Compiler {
public:
    Compiler(const std::string& src, const std::vector<std::string>& names) : src_(src), names_(names) {
        Lex();
    }

    CompiledExpr Run() {
        Add();
        if (Cur().kind != Tok::kEnd) Error("unexpected '" + Cur().text + "'");
        if (out_.code.empty()) Error("empty expression");
        int d = 0, md = 0;
        for (const MpdInstr& in : out_.code) {
            switch (in.op) {
                case MPD_OP_CONST: case MPD_OP_SYM: ++d; break;
                case MPD_OP_NEG: case MPD_OP_F1: break;
                case MPD_OP_CLAMP: d -= 2; break;
                case MPD_OP_MIN: case MPD_OP_MAX: d -= in.arg - 1; break;
                default: --d; break;
            }
            md = std::max(md, d);
        }
        out_.max_stack = md;
        std::sort(out_.symbols_used.begin(), out_.symbols_used.end());
        out_.symbols_used.erase(std::unique(out_.symbols_used.begin(), out_.symbols_used.end()),
                                out_.symbols_used.end());
        return out_;
    }

private:
    [[noreturn]] void Error(const std::string& msg) const {
        throw std::runtime_error("expression \"" + src_ + "\": " + msg);
    }

    void Lex() {
        size_t i = 0;
        while (i < src_.size()) {
            const char c = src_[i];
            if (std::isspace((unsigned char)c)) {
                ++i;
                continue;
            }
            Tok t;
            t.pos = i;
            if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < src_.size() && std::isdigit((unsigned char)src_[i + 1]))) {
                char* end = nullptr;
                t.kind = Tok::kNum;
                t.num = std::strtod(src_.c_str() + i, &end);
                t.text = src_.substr(i, (size_t)(end - (src_.c_str() + i)));
                i = (size_t)(end - src_.c_str());
            } else if (std::isalpha((unsigned char)c) || c == '_') {
                size_t j = i;
                while (j < src_.size() && (std::isalnum((unsigned char)src_[j]) || src_[j] == '_')) ++j;
                t.kind = Tok::kId;
                t.text = src_.substr(i, j - i);
                i = j;
            } else if (std::string("+-*/%^(),").find(c) != std::string::npos) {
                t.kind = Tok::kOp;
                t.text = std::string(1, c);
                ++i;
            } else {
                Error(std::string("unexpected character '") + c + "'");
            }
            toks_.push_back(t);
        }
        Tok e;
        e.kind = Tok::kEnd;
        e.pos = src_.size();
        toks_.push_back(e);
    }

    const Tok& Cur() const { return toks_[p_]; }
    bool IsOp(const char* s) const { return Cur().kind == Tok::kOp && Cur().text == s; }

    void Emit(int op, int arg = 0, double val = 0.0) {
        MpdInstr in;
        in.op = op;
        in.arg = arg;
        in.val = val;
        out_.code.push_back(in);
    }

    void Add() {
        Mul();
        while (IsOp("+") || IsOp("-")) {
            const bool plus = IsOp("+");
            ++p_;
            Mul();
            Emit(plus ? MPD_OP_ADD : MPD_OP_SUB);
        }
    }

    void Mul() {
        Unary();
        while (IsOp("*") || IsOp("/") || IsOp("%")) {
            const char o = Cur().text[0];
            ++p_;
            Unary();
            Emit(o == '*' ? MPD_OP_MUL : (o == '/' ? MPD_OP_DIV : MPD_OP_MOD));
        }
    }

    void Unary() {
        if (IsOp("-")) {
            ++p_;
            Unary();
            Emit(MPD_OP_NEG);
        } else if (IsOp("+")) {
            ++p_;
            Unary();
        } else {
            Power();
        }
    }

    void Power() {
        Primary();
        if (IsOp("^")) {
            ++p_;
            Unary();
            Emit(MPD_OP_POW);
        }
    }

    void Primary() {
        if (++depth_ > 200) Error("nested too deeply");
        const Tok t = Cur();
        if (t.kind == Tok::kNum) {
            ++p_;
            Emit(MPD_OP_CONST, 0, t.num);
        } else if (t.kind == Tok::kOp && t.text == "(") {
            ++p_;
            Add();
            if (!IsOp(")")) Error("expected ')'");
            ++p_;
        } else if (t.kind == Tok::kId) {
            ++p_;
            if (IsOp("(")) {
                ++p_;
                int nargs = 0;
                if (!IsOp(")")) {
                    for (;;) {
                        Add();
                        ++nargs;
                        if (IsOp(",")) {
                            ++p_;
                            continue;
                        }
                        break;
                    }
                }
                if (!IsOp(")")) Error("expected ')' after arguments of " + t.text);
                ++p_;
                Call(t.text, nargs);
            } else if (t.text == "pi") {
                Emit(MPD_OP_CONST, 0, 3.14159265358979323846);
            } else {
                const auto it = std::find(names_.begin(), names_.end(), t.text);
                if (it == names_.end()) Error("unknown symbol '" + t.text + "'");
                const int slot = (int)(it - names_.begin());
                Emit(MPD_OP_SYM, slot);
                out_.symbols_used.push_back(slot);
            }
        } else {
            Error(t.kind == Tok::kEnd ? "unexpected end" : "unexpected '" + t.text + "'");
        }
        --depth_;
    }

    void Call(const std::string& f, int n) {
        static const char* f1[] = {"sin", "cos", "tan", "asin", "acos", "atan", "exp",
                                   "log", "log2", "log10", "sqrt", "abs", "sign"};
        for (int i = 0; i < 13; ++i) {
            if (f == f1[i]) {
                if (n != 1) Error(f + " takes one argument");
                Emit(MPD_OP_F1, i);
                return;
            }
        }
        if (f == "atan2" || f == "pow") {
            if (n != 2) Error(f + " takes two arguments");
            Emit(f == "pow" ? MPD_OP_POW : MPD_OP_ATAN2);
        } else if (f == "clamp") {
            if (n != 3) Error("clamp takes three arguments (lo, x, hi)");
            Emit(MPD_OP_CLAMP);
        } else if (f == "min" || f == "max") {
            if (n < 2) Error(f + " takes at least two arguments");
            Emit(f == "min" ? MPD_OP_MIN : MPD_OP_MAX, n);
        } else {
            Error("unknown function '" + f + "'");
        }
    }

    const std::string& src_;
    const std::vector<std::string>& names_;
    std::vector<Tok> toks_;
    size_t p_ = 0;
    int depth_ = 0;
    CompiledExpr out_;
};

}  // namespace

CompiledExpr CompileExpr(const std::string& src, const std::vector<std::string>& names) {
    CompiledExpr e = Compiler(src, names).Run();
    if (e.max_stack > MPD_MAX_STACK) {
        throw std::runtime_error("expression \"" + src + "\": needs a stack deeper than " +
                                 std::to_string(MPD_MAX_STACK));
    }
    return e;
}

double EvalExpr(const CompiledExpr& e, const double* sym) {
    double st[MPD_MAX_STACK];
    int sp = 0;
    for (const MpdInstr& in : e.code) {
        switch (in.op) {
            case MPD_OP_CONST: st[sp++] = in.val; break;
            case MPD_OP_SYM: st[sp++] = sym[in.arg]; break;
            case MPD_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
            case MPD_OP_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
            case MPD_OP_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
            case MPD_OP_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
            case MPD_OP_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
            case MPD_OP_MOD: --sp; st[sp - 1] = std::fmod(st[sp - 1], st[sp]); break;
            case MPD_OP_POW: --sp; st[sp - 1] = std::pow(st[sp - 1], st[sp]); break;
            case MPD_OP_ATAN2: --sp; st[sp - 1] = std::atan2(st[sp - 1], st[sp]); break;
            case MPD_OP_F1: {
                const double x = st[sp - 1];
                double y;
                switch (in.arg) {
                    case MPD_FN_SIN: y = std::sin(x); break;
                    case MPD_FN_COS: y = std::cos(x); break;
                    case MPD_FN_TAN: y = std::tan(x); break;
                    case MPD_FN_ASIN: y = std::asin(x); break;
                    case MPD_FN_ACOS: y = std::acos(x); break;
                    case MPD_FN_ATAN: y = std::atan(x); break;
                    case MPD_FN_EXP: y = std::exp(x); break;
                    case MPD_FN_LOG: y = std::log(x); break;
                    case MPD_FN_LOG2: y = std::log2(x); break;
                    case MPD_FN_LOG10: y = std::log10(x); break;
                    case MPD_FN_SQRT: y = std::sqrt(x); break;
                    case MPD_FN_ABS: y = std::fabs(x); break;
                    default: y = (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); break;
                }
                st[sp - 1] = y;
                break;
            }
            case MPD_OP_CLAMP: {
                const double lo = st[sp - 3], x = st[sp - 2], hi = st[sp - 1];
                sp -= 2;
                st[sp - 1] = (x < lo) ? lo : ((x > hi) ? hi : x);
                break;
            }
            case MPD_OP_MIN:
            case MPD_OP_MAX: {
                const int n = in.arg;
                double acc = st[sp - n];
                for (int k = 1; k < n; ++k) {
                    const double v = st[sp - n + k];
                    if (in.op == MPD_OP_MIN) acc = (v < acc) ? v : acc;
                    else acc = (acc < v) ? v : acc;
                }
                sp -= n - 1;
                st[sp - 1] = acc;
                break;
            }
            default: break;
        }
    }
    return st[0];
}

}  // namespace mpd
