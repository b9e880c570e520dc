// common.hpp -- shared definitions of the MI355X (gfx950) FMM library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace vpm {

// Wavefront width of CDNA4.
constexpr int kWave = 64;
// Tree depth limit: 32-bit Morton keys hold 16 levels of (ix, iy) bits.
constexpr int kMaxDepth = 16;
// Tolerance of the reference's geometric list tests ("10e-10" = 1e-9,
// reference src/FMM/treeCell.cpp:5).
constexpr double kListEps = 10e-10;

// Status codes of the C ABI (include/vpmfmm.h).
enum Status : int {
    kOk = 0,
    kBadArg = 1,
    kNonFinite = 2,
    kDegenerate = 3,
    kDepth = 4,
    kNoMem = 5,
    kInternal = 6,
    kDuplicate = 7,
    kHipError = 8,
    kOutsideBoxes = 9,
    kLevel1Leaf = 10,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define VPM_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess)                                                          \
            throw ::vpm::Error(::vpm::kHipError, std::string(#call) + ": " +          \
                                                     hipGetErrorString(_e));           \
    } while (0)

#define VPM_LAUNCH_CHECK() VPM_HIP(hipGetLastError())

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Device-side flags raised by kernels and read back by the host.
enum DevFlag : unsigned {
    kFlagNonFinite = 1u,
    kFlagDepth = 2u,
    kFlagDuplicate = 4u,
    kFlagListOverflow = 8u,
    kFlagRange = 16u,
    kFlagT1Lost = 32u,  // type-1 tree: a particle outside every box of its level
};

}  // namespace vpm
