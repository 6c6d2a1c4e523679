// Internal helpers shared by the host translation units of libjgs2hip.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "jgs2hip.h"

namespace jgs2 {

// Records the calling thread's last error message and returns `status`.
jgs2_status fail(jgs2_status status, const char* format, ...);
void clear_error();

// minimumStaticIpcContactDistance (src/simulation/cpu/ipc-contact.ts:349-391)
// in f64 over positions with 4 floats per vertex.
double min_static_contact_distance(const float* positions4, const jgs2_contact_input& contact);

}  // namespace jgs2
