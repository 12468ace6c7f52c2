// lf_types.h -- plain-data structures shared by the host library and the
// gfx950 kernels of the Lax-Friedrichs step.  No HIP or C++ library types.
#pragma once

#include <stdint.h>

// Field-condition kinds of include/field_condition.hpp:94-99 (reference).
enum MpdCondType { MPD_NEUMANN = 0, MPD_DIRICHLET = 1, MPD_EXPRESSION = 2, MPD_AXIS_LF = 3 };

// Faces, in the reference's FaceBC::Face order (include/face_bc.hpp:47).
enum MpdFace { MPD_L_LO = 0, MPD_L_HI = 1, MPD_M_LO = 2, MPD_M_HI = 3 };

// Symbols readable by a boundary-condition expression
// (src/bcs/per_field_bc.cpp:82-119), in this fixed slot order.
enum MpdBcSym {
    MPD_BS_GAMMA, MPD_BS_BETA, MPD_BS_HZ0, MPD_BS_R0, MPD_BS_DZ, MPD_BS_DR, MPD_BS_R, MPD_BS_RZ,
    MPD_BS_Z, MPD_BS_RNB, MPD_BS_RZNB, MPD_BS_RHO_NB, MPD_BS_VZ_NB, MPD_BS_VR_NB, MPD_BS_VPHI_NB,
    MPD_BS_E_NB, MPD_BS_HZ_NB, MPD_BS_HR_NB, MPD_BS_HPHI_NB, MPD_BS_RHO, MPD_BS_VZ, MPD_BS_VR,
    MPD_BS_VPHI, MPD_BS_HZ, MPD_BS_HR, MPD_BS_HPHI, MPD_BS_E, MPD_BS_COUNT
};

// Bytecode of a compiled expression: a stack machine over doubles.
enum MpdOp {
    MPD_OP_CONST, MPD_OP_SYM, MPD_OP_NEG, MPD_OP_ADD, MPD_OP_SUB, MPD_OP_MUL, MPD_OP_DIV,
    MPD_OP_MOD, MPD_OP_POW, MPD_OP_F1, MPD_OP_ATAN2, MPD_OP_CLAMP, MPD_OP_MIN, MPD_OP_MAX
};
enum MpdFn1 {
    MPD_FN_SIN, MPD_FN_COS, MPD_FN_TAN, MPD_FN_ASIN, MPD_FN_ACOS, MPD_FN_ATAN, MPD_FN_EXP,
    MPD_FN_LOG, MPD_FN_LOG2, MPD_FN_LOG10, MPD_FN_SQRT, MPD_FN_ABS, MPD_FN_SIGN
};

struct MpdInstr {
    int32_t op;
    int32_t arg;
    double val;
};

#define MPD_MAX_SEGS_PER_FACE 16
#define MPD_MAX_CODE 1024
#define MPD_MAX_STACK 16

// One field's condition.  Expression code lives in MpdBcProgram::code.
struct MpdCond {
    int32_t type;
    int32_t code_off;
    int32_t code_len;
    int32_t pad;
    double value;
};

// One boundary segment with its range already resolved to local indices on
// the face's free axis (sentinels, clipping and the M_LO corner policy of
// src/face_bc.cpp:104-129 applied on the host).  lo > hi means empty.
// Conditions are in the evaluation order rho, v_z, v_r, v_phi, H_z, H_r,
// H_phi, e (src/bcs/per_field_bc.cpp:322-360).
struct MpdSeg {
    int32_t lo, hi;
    int32_t has_axis, has_expr;
    MpdCond c[8];
};

struct MpdBcProgram {
    int32_t nseg[4];
    int32_t ncode;
    int32_t max_stack;  // deepest evaluation stack of the program's expressions
    MpdSeg seg[4][MPD_MAX_SEGS_PER_FACE];
    MpdInstr code[MPD_MAX_CODE];
};

// Device-resident loop state of the time loop (src/main.cpp:191-247).
struct MpdLoopState {
    double t;          // simulated time
    double dt;         // dt variable of the loop (after ComputeDt)
    double dt_use;     // dt of the next step after the T truncation
    double change;     // last SolutionChange value
    double max_speed;  // max wave speed of the state after the last step
    int32_t step;      // steps taken
    int32_t done;      // 1: loop condition false (t >= T, converged, step cap)
    int32_t converged;
    int32_t max_steps;  // <= 0: no cap
    uint64_t speed_bits;  // running max of the wave speed (bit pattern, values >= 0)
    uint32_t ticket;      // completion counter for last-block finalisation
    uint32_t ticket2;
    int32_t checked_step;  // step at which SolutionChange last ran in the loop
    int32_t fallback;      // 1: the candidate max is not provably the max; rescan
    uint32_t range_bits;   // nonzero: a value of the new state is outside split_div_safe's range
    int32_t range_stop;    // 1: the loop idles until the host switches to IEEE division
};
