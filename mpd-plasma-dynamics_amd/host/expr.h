// expr.h -- compiler for the boundary/initial-condition formula language.
//
// The reference compiles these formulas with exprtk 0.0.2 (third-party, not
// vendored; call sites src/bcs/per_field_bc.cpp:162-165 and
// src/ics/expression_ic.cpp:168-171).  This compiler accepts the documented
// subset (include/field_condition.hpp:31-62): numbers, symbols, ( ), + - * /
// % ^, unary + -, sin cos tan asin acos atan atan2 exp log log2 log10 sqrt pow
// abs sign min max clamp, and pi.  ^ binds tighter than unary minus and is
// right associative.  Output is stack bytecode (MpdInstr) that both the host
// evaluator below and the device interpreter run.
#pragma once

#include <string>
#include <vector>

#include "../csrc/lf_types.h"

namespace mpd {

struct CompiledExpr {
    std::vector<MpdInstr> code;
    int max_stack = 0;
    std::vector<int> symbols_used;  // slot indices, sorted, unique
};

// Compile src against the symbol names (slot i = names[i]).  Throws
// std::runtime_error naming the expression on a syntax error.
CompiledExpr CompileExpr(const std::string& src, const std::vector<std::string>& names);

// Host evaluation (IEEE binary64, glibc libm).
double EvalExpr(const CompiledExpr& e, const double* sym);

}  // namespace mpd
