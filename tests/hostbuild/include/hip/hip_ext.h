// Host stand-in: the extended launch (hipExtLaunchKernelGGL) is defined with the other launches.
#pragma once
#include <hip/hip_runtime.h>
