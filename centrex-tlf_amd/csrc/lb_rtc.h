// Plan-specialised kernels: the time-dependent coefficient program emitted as
// straight-line device code and compiled with the paired integrator at run
// time (hiprtc), so the per-step coefficient evaluation runs without the
// expression machine's dispatch, LDS stack and instruction fetches.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lb_program.h"

// The device function lb_gen_coef(sl, cf, t) for the time-dependent programs
// of `prog` (expressions e with expr_dyn[e], compounds first, then the K
// coefficients), or an empty string with `why` set when an instruction has no
// generated form.
std::string lb_rtc_coef_source(const LbProgram& prog, const std::vector<int32_t>& expr_dyn, int K,
                               std::string& why);

// Compiles the kernel sources with `coef_source` for gfx950 and loads the
// kernel named by `name_expr` (a template-id such as
// "lb_solve_pair_kernel<0, 8, 2>") from the resulting code object.
// Returns 0, or -1 with `why` holding the compiler log or the runtime error.
int lb_rtc_compile(const std::string& coef_source, const std::string& name_expr, hipModule_t* module,
                   hipFunction_t* function, std::string& why);
