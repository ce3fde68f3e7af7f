// bem_prof.h -- hooks of the kernel-time registry (bem_prof.hip)
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/bem3_hip.h"

extern "C" void bem_prof_begin(int kid, double work, hipStream_t st);
extern "C" void bem_prof_begin_b(int kid, double work, double bytes, hipStream_t st);
extern "C" void bem_prof_end(int kid, hipStream_t st);
extern "C" int bem_prof_on(void);
