// Stand-in for the HIP runtime header in the CPU emulation build of the
// batched kernels (test infrastructure): the device facilities those kernels
// use, emulated on the host.
#pragma once
#include "../hip_emu.hpp"
