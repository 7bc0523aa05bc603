// gdb_kernels.cpp — fixed (ahead-of-time compiled) device kernels of the
// C-ABI library. Generated query kernels are compiled at run time; the
// kernels here do not depend on a plan.
#include "gdb/gdb_hip.h"

#include <hip/hip_runtime.h>
