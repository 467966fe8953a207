// Library-internal entry points shared between the host sources.
#pragma once

#include <stddef.h>

#include <string>

#include "../../include/centrex_lindblad.h"

// The core of lb_adjoint (scratch == nullptr: library-owned workspace) and of
// lb_adjoint_noalloc (scratch: a device buffer of scratch_bytes that holds
// the work-queue counter and the workspace; nothing is allocated).
int lb_adjoint_launch(const lb_plan* plan, const lb_solve_args* args, void* stream, void* scratch,
                      size_t scratch_bytes);

// Leave `msg` for lb_last_error() on this thread and return `code`.
int lb_set_error(int code, const std::string& msg);
