// bfa_expr.h -- compiled material expressions shared by the host compiler
// (bfa_material.cpp) and the device interpreter (bfa_assembly.hip).
//
// An SVTensor (src/libs/libbuff/SVTensor.cc) is up to nine expressions
// Q[mu][nu](w, x, y, z, r, Theta, Phi, Eps1..3) plus up to three MatProp
// permittivities Eps_k(w); each is compiled to a postfix program over a
// complex stack.  One device buffer holds the header, the ops and the
// complex constants.
#pragma once
#include <cstdint>

namespace bfa {

enum ExprOpCode : int {
  EOP_CONST = 0,   // push consts[arg]
  EOP_VAR,         // push vars[arg]
  EOP_ADD, EOP_SUB, EOP_MUL, EOP_DIV, EOP_POW, EOP_NEG,
  EOP_FN           // apply function arg to the top of the stack
};

enum ExprFn : int {
  EFN_EXP = 0, EFN_LOG, EFN_SQRT, EFN_SIN, EFN_COS, EFN_TAN, EFN_COT, EFN_SEC, EFN_CSC,
  EFN_ASIN, EFN_ACOS, EFN_ATAN, EFN_SINH, EFN_COSH, EFN_TANH, EFN_ABS, EFN_STEP, EFN_COUNT
};

// variable slots of the SVTensor evaluation (SVTensor.cc:319-339)
enum ExprVar : int { EV_W = 0, EV_X, EV_Y, EV_Z, EV_R, EV_THETA, EV_PHI, EV_EPS1, EV_EPS2, EV_EPS3, EV_COUNT };

#define EXPR_SLOTS 12        // 9 tensor components (3*mu+nu), then Eps1..Eps3
#define EXPR_STACK 16

struct ExprOp { int op, arg; };

struct ExprHeader {
  int start[EXPR_SLOTS];     // first op of each program
  int len[EXPR_SLOTS];       // 0: absent
  int nops, nconst;
  int num_mp;                // MatProp permittivities referenced (Eps1..)
  int matprop;               // 1: homogeneous isotropic MatProp (Q = Eps1 everywhere)
  double freq_unit;          // w = Omega * freq_unit (MatProp::FreqUnit)
  double pad;
  // followed by ExprOp ops[nops] and double consts[2 * nconst]
};

}  // namespace bfa
