// the request server (server.hip) over the CPU stand-in runtime and the emulated kernels (test infrastructure)
#include <hip/hip_runtime.h>
#include "../../weave-sim_amd/csrc/hip/batch.hpp"
#include "emu_kernels.hpp"
#include "../../weave-sim_amd/csrc/hip/server.hip"
