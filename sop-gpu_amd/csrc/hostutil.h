// hostutil.h -- host-side helpers of the engine.
#pragma once
#include <stdint.h>

#include <string>

namespace sop {

// Seed table of the Hybrid Taus generator: for every bead, four 32-bit words
// drawn from Numerical Recipes' ran2 (src/Integrators/ran2.cpp:51-89) and
// redrawn while below 128 (src/Integrators/ht.cu:40-56).  ran2's shuffle
// state starts from its static initial values on every call, as it does in
// a fresh reference process.
void generate_seeds(int seed, int n, uint32_t* out4);

// The thread's last error, read back by sop_last_error(); fail() records
// it and returns the code.
void set_error(int code, const std::string& msg);
int fail(int code, const std::string& msg);

}  // namespace sop
