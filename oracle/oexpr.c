/*
 * oexpr.c -- expression evaluator of the CPU oracle (test infrastructure only).
 *
 * This file belongs to the oracle: only tests, the smoke check and the CPU
 * baseline leg of bench.py load it, and only as the checker.  It is an
 * independent restatement of the formula language the reference compiles with
 * exprtk 0.0.2 (third-party, not vendored: cmake/Dependencies.cmake:6), used at
 * the reference call sites src/bcs/per_field_bc.cpp:162-165 (BC expressions)
 * and src/ics/expression_ic.cpp:155-171 (IC expressions).
 *
 * Language subset (the symbol/function list of include/field_condition.hpp:31-62
 * and include/ics/expression_ic.hpp:38-67):
 *   numbers (strtod), identifiers, ( ), + - * / % ^, unary + -,
 *   sin cos tan asin acos atan atan2 exp log log2 log10 sqrt pow abs sign
 *   min max (two or more arguments) clamp(lo, x, hi).
 * Precedence: ^ binds tighter than unary minus (-a^b == -(a^b)); ^ is right
 * associative; * / % before + -; all other binary operators left associative.
 * exprtk's precedence/associativity of ^ and its literal parser are not pinned
 * by anything the reference ships ("parity unpinned" beyond single-operation
 * expressions, which the tests pin with known answers).
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oexpr.h"

enum { OP_CONST, OP_VAR, OP_NEG, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_POW,
       OP_F1, OP_F2, OP_CLAMP, OP_MIN, OP_MAX };
enum { F_SIN, F_COS, F_TAN, F_ASIN, F_ACOS, F_ATAN, F_EXP, F_LOG, F_LOG2, F_LOG10,
       F_SQRT, F_ABS, F_SIGN, F_ATAN2, F_POW };

typedef struct { int op; int arg; double val; } oinstr;

struct oexpr {
    oinstr* code;
    int n, cap;
    int max_stack;
};

typedef struct {
    const char* s;
    size_t pos;
    const char* const* names;
    int nnames;
    oexpr* e;
    char* err;
    int errlen;
    int failed;
    int depth;
} parser;

static void emit(parser* p, int op, int arg, double val) {
    if (p->failed) return;
    if (p->e->n == p->e->cap) {
        int nc = p->e->cap ? 2 * p->e->cap : 32;
        oinstr* c = (oinstr*)realloc(p->e->code, (size_t)nc * sizeof(oinstr));
        if (!c) { p->failed = 1; return; }
        p->e->code = c;
        p->e->cap = nc;
    }
    p->e->code[p->e->n].op = op;
    p->e->code[p->e->n].arg = arg;
    p->e->code[p->e->n].val = val;
    p->e->n++;
}

static void fail(parser* p, const char* msg) {
    if (!p->failed) {
        snprintf(p->err, (size_t)p->errlen, "%s at offset %zu in \"%s\"", msg, p->pos, p->s);
    }
    p->failed = 1;
}

static void skip_ws(parser* p) {
    while (p->s[p->pos] && isspace((unsigned char)p->s[p->pos])) p->pos++;
}

static int peek(parser* p) {
    skip_ws(p);
    return (unsigned char)p->s[p->pos];
}

static void parse_add(parser* p);
static void parse_unary(parser* p);

static void parse_primary(parser* p) {
    int c = peek(p);
    if (p->failed) return;
    if (++p->depth > 200) { fail(p, "expression nested too deeply"); return; }
    if (c == '(') {
        p->pos++;
        parse_add(p);
        if (peek(p) != ')') { fail(p, "expected ')'"); return; }
        p->pos++;
    } else if (isdigit(c) || c == '.') {
        char* end = NULL;
        double v = strtod(p->s + p->pos, &end);
        if (end == p->s + p->pos) { fail(p, "bad number"); return; }
        p->pos = (size_t)(end - p->s);
        emit(p, OP_CONST, 0, v);
    } else if (isalpha(c) || c == '_') {
        size_t st = p->pos;
        while (isalnum((unsigned char)p->s[p->pos]) || p->s[p->pos] == '_') p->pos++;
        size_t len = p->pos - st;
        char name[64];
        if (len >= sizeof(name)) { fail(p, "identifier too long"); return; }
        memcpy(name, p->s + st, len);
        name[len] = 0;
        if (peek(p) == '(') {
            p->pos++;
            int nargs = 0;
            if (peek(p) != ')') {
                for (;;) {
                    parse_add(p);
                    if (p->failed) return;
                    nargs++;
                    if (peek(p) == ',') { p->pos++; continue; }
                    break;
                }
            }
            if (peek(p) != ')') { fail(p, "expected ')' after arguments"); return; }
            p->pos++;
            static const char* f1n[] = {"sin", "cos", "tan", "asin", "acos", "atan", "exp",
                                        "log", "log2", "log10", "sqrt", "abs", "sign"};
            for (int i = 0; i < 13; ++i) {
                if (strcmp(name, f1n[i]) == 0) {
                    if (nargs != 1) { fail(p, "function takes one argument"); return; }
                    emit(p, OP_F1, i, 0.0);
                    p->depth--;
                    return;
                }
            }
            if (strcmp(name, "atan2") == 0 || strcmp(name, "pow") == 0) {
                if (nargs != 2) { fail(p, "function takes two arguments"); return; }
                emit(p, OP_F2, name[0] == 'a' ? F_ATAN2 : F_POW, 0.0);
            } else if (strcmp(name, "clamp") == 0) {
                if (nargs != 3) { fail(p, "clamp takes three arguments"); return; }
                emit(p, OP_CLAMP, 0, 0.0);
            } else if (strcmp(name, "min") == 0 || strcmp(name, "max") == 0) {
                if (nargs < 2) { fail(p, "min/max take at least two arguments"); return; }
                emit(p, name[1] == 'i' ? OP_MIN : OP_MAX, nargs, 0.0);
            } else {
                fail(p, "unknown function");
                return;
            }
        } else if (strcmp(name, "pi") == 0) {
            emit(p, OP_CONST, 0, 3.14159265358979323846);
        } else {
            int idx = -1;
            for (int i = 0; i < p->nnames; ++i) {
                if (strcmp(p->names[i], name) == 0) { idx = i; break; }
            }
            if (idx < 0) { fail(p, "unknown symbol"); return; }
            emit(p, OP_VAR, idx, 0.0);
        }
    } else {
        fail(p, "unexpected character");
        return;
    }
    p->depth--;
}

static void parse_power(parser* p) {
    parse_primary(p);
    if (p->failed) return;
    if (peek(p) == '^') {
        p->pos++;
        parse_unary(p); /* right associative, allows a^-b */
        emit(p, OP_POW, 0, 0.0);
    }
}

static void parse_unary(parser* p) {
    int c = peek(p);
    if (c == '-') {
        p->pos++;
        parse_unary(p);
        emit(p, OP_NEG, 0, 0.0);
    } else if (c == '+') {
        p->pos++;
        parse_unary(p);
    } else {
        parse_power(p);
    }
}

static void parse_mul(parser* p) {
    parse_unary(p);
    for (;;) {
        if (p->failed) return;
        int c = peek(p);
        if (c == '*' || c == '/' || c == '%') {
            p->pos++;
            parse_unary(p);
            emit(p, c == '*' ? OP_MUL : (c == '/' ? OP_DIV : OP_MOD), 0, 0.0);
        } else {
            return;
        }
    }
}

static void parse_add(parser* p) {
    parse_mul(p);
    for (;;) {
        if (p->failed) return;
        int c = peek(p);
        if (c == '+' || c == '-') {
            p->pos++;
            parse_mul(p);
            emit(p, c == '+' ? OP_ADD : OP_SUB, 0, 0.0);
        } else {
            return;
        }
    }
}

oexpr* oexpr_compile(const char* src, const char* const* names, int nnames, char* err,
                     int errlen) {
    oexpr* e = (oexpr*)calloc(1, sizeof(oexpr));
    if (!e) return NULL;
    parser p = {src, 0, names, nnames, e, err, errlen, 0, 0};
    if (err && errlen > 0) err[0] = 0;
    parse_add(&p);
    if (!p.failed && peek(&p) != 0) fail(&p, "trailing characters");
    if (!p.failed && e->n == 0) fail(&p, "empty expression");
    if (p.failed) {
        oexpr_free(e);
        return NULL;
    }
    /* stack depth */
    int d = 0, md = 0;
    for (int i = 0; i < e->n; ++i) {
        switch (e->code[i].op) {
            case OP_CONST: case OP_VAR: d++; break;
            case OP_NEG: case OP_F1: break;
            case OP_CLAMP: d -= 2; break;
            case OP_MIN: case OP_MAX: d -= e->code[i].arg - 1; break;
            default: d--; break;
        }
        if (d > md) md = d;
    }
    e->max_stack = md;
    return e;
}

void oexpr_free(oexpr* e) {
    if (!e) return;
    free(e->code);
    free(e);
}

int oexpr_uses(const oexpr* e, int var) {
    for (int i = 0; i < e->n; ++i)
        if (e->code[i].op == OP_VAR && e->code[i].arg == var) return 1;
    return 0;
}

double oexpr_eval(const oexpr* e, const double* vars) {
    double st[256];
    int sp = 0;
    for (int i = 0; i < e->n; ++i) {
        const oinstr* in = &e->code[i];
        switch (in->op) {
            case OP_CONST: st[sp++] = in->val; break;
            case OP_VAR: st[sp++] = vars[in->arg]; break;
            case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
            case OP_ADD: sp--; st[sp - 1] = st[sp - 1] + st[sp]; break;
            case OP_SUB: sp--; st[sp - 1] = st[sp - 1] - st[sp]; break;
            case OP_MUL: sp--; st[sp - 1] = st[sp - 1] * st[sp]; break;
            case OP_DIV: sp--; st[sp - 1] = st[sp - 1] / st[sp]; break;
            case OP_MOD: sp--; st[sp - 1] = fmod(st[sp - 1], st[sp]); break;
            case OP_POW: sp--; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
            case OP_F1: {
                double x = st[sp - 1], y;
                switch (in->arg) {
                    case F_SIN: y = sin(x); break;
                    case F_COS: y = cos(x); break;
                    case F_TAN: y = tan(x); break;
                    case F_ASIN: y = asin(x); break;
                    case F_ACOS: y = acos(x); break;
                    case F_ATAN: y = atan(x); break;
                    case F_EXP: y = exp(x); break;
                    case F_LOG: y = log(x); break;
                    case F_LOG2: y = log2(x); break;
                    case F_LOG10: y = log10(x); break;
                    case F_SQRT: y = sqrt(x); break;
                    case F_ABS: y = fabs(x); break;
                    default: y = (x > 0.0) ? 1.0 : ((x < 0.0) ? -1.0 : 0.0); break;
                }
                st[sp - 1] = y;
                break;
            }
            case OP_F2:
                sp--;
                st[sp - 1] = (in->arg == F_ATAN2) ? atan2(st[sp - 1], st[sp]) : pow(st[sp - 1], st[sp]);
                break;
            case OP_CLAMP: {
                double lo = st[sp - 3], x = st[sp - 2], hi = st[sp - 1];
                sp -= 2;
                st[sp - 1] = (x < lo) ? lo : ((x > hi) ? hi : x);
                break;
            }
            case OP_MIN: case OP_MAX: {
                int n = in->arg;
                double acc = st[sp - n];
                for (int k = 1; k < n; ++k) {
                    double v = st[sp - n + k];
                    if (in->op == OP_MIN) acc = (v < acc) ? v : acc;
                    else acc = (acc < v) ? v : acc;
                }
                sp -= n - 1;
                st[sp - 1] = acc;
                break;
            }
        }
    }
    return st[0];
}
