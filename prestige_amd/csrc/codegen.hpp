// codegen.hpp — internal interface between the equation registry
// (equations.cpp, host only) and the device runtime that compiles and runs
// generated gather passes (jit.hip).  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/prestige_hip.h"

namespace prestige {

// Error of the equation front end, converted to a status at the C ABI.
struct GenError {
  prestige_status code;
  std::string msg;
};

// One field parameter of a generated pass kernel, in parameter order.
//   read: the pass reads the field (at [i] or [j]); written: it writes F[i].
//   split: read at [j] and written in the same pass; the kernel takes an
//   input and a separate output array (two parameters, in then out).
struct GenField {
  std::string name;
  bool read = false, written = false, split = false;
};

struct GenPass {
  std::string kernel;            // extern "C" symbol
  std::vector<GenField> fields;  // after the fixed parameters
  // cell loop: fields read at [j], passed after fields as copies in slot order
  std::vector<GenField> slot_fields;
};

struct GenProgram {
  std::string source;  // one translation unit, every pass
  std::vector<GenPass> passes;
};

// HIP source of the gather passes of the fused equations eq_ids[0..n) for
// loop form `loop` (PRESTIGE_LOOP_ALLPAIRS or PRESTIGE_LOOP_CELLS); f32_pos
// selects float4 positions (FP32 contexts) instead of double4.  Throws
// GenError.
GenProgram generate_hip_program(const int32_t* eq_ids, uint32_t n, int32_t loop, bool f32_pos);

// Message returned by prestige_last_error(NULL) for calls without a context.
void set_context_free_error(const std::string& msg);

}  // namespace prestige
