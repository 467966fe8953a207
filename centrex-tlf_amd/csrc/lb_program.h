// Device form of the parameter/coefficient bytecode (see lb_program.cpp).
#pragma once

#include <string>
#include <vector>

#include "lb_device.h"

struct LbProgram {
  std::vector<DevInstr> instr;
  std::vector<int32_t> expr_start;     // n_compounds + K + 1
  std::vector<int32_t> compound_slot;  // n_compounds (hidden ones included)
  std::vector<uint8_t> compound_dyn;
  std::vector<uint8_t> coef_dyn;
  int n_compounds = 0;
  int n_slots = 0;  // original slots + hidden cache slots
  int max_depth = 0;
};

// Returns 0 on success, -1 (with `why`) on a malformed expression.
int lb_build_program(const lb_plan_desc* d, const std::vector<uint8_t>& dyn_slot,
                     const std::vector<uint8_t>& compound_dyn, const std::vector<uint8_t>& coef_dyn,
                     LbProgram* out, std::string& why);
