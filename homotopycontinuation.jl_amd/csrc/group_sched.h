// Slot programs of the lane-group tracker: a tape (ModelKit InstructionSequence) scheduled
// for a group of 16 lanes that share one path.
//
// The tape's instructions are list-scheduled into "slots": a slot holds up to 16
// instructions with the same OpType, one per lane of the group, that depend only on values
// produced in earlier slots.  The priority is the longest path to the end of the tape,
// so independent chains (one per output row, typically) advance side by side and few values
// are live at once.  Values live in LDS registers; a linear scan over the slots assigns them.
// Every instruction computes exactly the operation the tape names on exactly its operands, so
// the results equal a sequential execution of the tape bit for bit.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "hcgpu.h"

namespace hcg {

constexpr int kGroupLanes = 16;

// Operand references: kind in bits 28..31, index in bits 0..27.
enum : int32_t {
    GREF_REG = 0,          // LDS register
    GREF_CONST = 1 << 28,  // entry of the program's constant table
    GREF_PARAM = 2 << 28,  // parameter of a parameter homotopy, formed at t when read
    GREF_ZERO = 3 << 28,   // the value 0 (an output the tape does not assign)
    GREF_INT = 4 << 28,    // an integer immediate (OP_POW_INT's exponent)
    GREF_NONE = 5 << 28,   // no operand / no instruction
};
inline int32_t gref_kind(int32_t r) { return r & (int32_t)0xF0000000; }
inline int32_t gref_index(int32_t r) { return r & 0x0FFFFFFF; }

// One tape of a slot program: where its variables are loaded and where its outputs are read.
struct GTapeIO {
    std::vector<int32_t> var_reg;  // per variable: LDS register it is loaded into (-1: unused)
    std::vector<int32_t> u_ref;    // per output row: operand reference
    std::vector<int32_t> U_ref;    // per (row, column), row-major n x n: operand reference (eval_jac tapes)
};

struct GProgram {
    int nslots = 0;
    int nregs = 0;
    int ninstr = 0;
    std::vector<uint8_t> op;     // nslots
    std::vector<int32_t> arg;    // nslots * 16 * 4 operand references
    std::vector<int32_t> out;    // nslots * 16 result registers (-1: lane idle)
    std::vector<double> consts;  // constant table (re, im)
    std::vector<GTapeIO> io;     // one per tape, in the order given
};

// Schedule one or more tapes (independent programs evaluated together, e.g. G and F of a
// straight-line homotopy) into one slot program.  fixed[k] (may be null) are the fixed
// parameter values of tape k, which become constants; otherwise parameters stay GREF_PARAM.
// n: the system size (U_ref is n x n).  Returns false with *err on malformed tapes.
bool schedule_tapes(const std::vector<const hcg_tape*>& tapes, const std::vector<const double*>& fixed, int n,
                    GProgram* out, std::string* err);

// Check a slot program against its tapes: every instruction once, operands produced in an
// earlier slot, and no register overwritten while a value in it is still to be read
// (simulating the register file slot by slot).  Returns false with *err on the first violation.
bool check_program(const std::vector<const hcg_tape*>& tapes, const std::vector<const double*>& fixed,
                   const GProgram& p, std::string* err);

}  // namespace hcg
