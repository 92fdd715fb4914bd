//-------------------------------------------------------------------------------
// PSkelDefs.h -- qualifiers, backend switches and the error convention of the
// MI355X drop-in for PSkel's stencil API.
//
// Mirrors include/PSkelDefs.h of the reference (paths relative to the PSkel
// tree): __parallel__ marks a function callable on host and device and forces
// inlining (:48-49, CPU fallbacks :57-63); gpuErrchk reports a failed GPU call
// (:50-56); MIN (:41).
//
// Backend switches (a TU defines them before including PSkel.h):
//   PSKEL_CUDA or PSKEL_GPU   GPU methods on (they run on the MI355X through
//                             libpskel_amd.so; the reference's CUDA name is
//                             accepted so existing sources build unchanged)
//   PSKEL_OMP                 runIterativeCPU uses OpenMP (as in the reference)
//   PSKEL_THROW_ON_ERROR      GPU failures throw PSkel::Error instead of the
//                             reference's print-and-exit
//-------------------------------------------------------------------------------
#ifndef PSKEL_DEFS_H
#define PSKEL_DEFS_H

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#if defined(PSKEL_CUDA) && !defined(PSKEL_GPU)
#define PSKEL_GPU 1
#endif

#define MIN(a, b) ((a) <= (b) ? (a) : (b))

#if defined(__HIPCC__)
// hipcc: user stencil kernels can also be compiled for the device
#include <hip/hip_runtime.h>
#define PSKEL_DEVICE_COMPILER 1
#define __parallel__ __device__ __host__ __forceinline__
#else
#ifndef __device__
#define __device__
#endif
#ifndef __host__
#define __host__
#endif
#ifndef __forceinline__
#define __forceinline__ __attribute__((always_inline)) inline
#endif
#define __parallel__ __forceinline__
#endif

#ifndef __forceinline
#define __forceinline __forceinline__
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define PSKEL_ON_DEVICE 1
#endif

#ifdef PSKEL_GPU
#include "pskel_amd.h"
#endif

namespace PSkel {

// The error a failed GPU call raises in PSKEL_THROW_ON_ERROR mode.
struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

#ifdef PSKEL_GPU
// gpuAssert: report a failed library call.  Default: print and exit with the
// status, like the reference's gpuErrchk(abort = true); with
// PSKEL_THROW_ON_ERROR: throw PSkel::Error.
inline void gpuAssert(pskel_status_t code, const char* file, int line)
{
    if (code == PSKEL_OK) return;
    std::string msg = std::string("GPUassert: ") + pskel_status_string(code) + ": " + pskel_last_error() +
                      " " + file + " " + std::to_string(line);
#ifdef PSKEL_THROW_ON_ERROR
    throw Error((int)code, msg);
#else
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit((int)code);
#endif
}
#define gpuErrchk(ans) ::PSkel::gpuAssert((ans), __FILE__, __LINE__)
#endif

}  // namespace PSkel

#endif
