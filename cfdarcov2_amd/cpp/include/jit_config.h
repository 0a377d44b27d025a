// Build configuration of the operator API: the device path is always built.
#pragma once
#ifndef USE_GPU
#define USE_GPU 1
#endif
