// Host builds for the CPU tests: <hip/hip_runtime.h> resolves to the shim.
#pragma once
#include "../cpu_hip.h"
