//-------------------------------------------------------------------------------
// PSkel.h -- umbrella header of the MI355X drop-in for PSkel's stencil API
// (the reference's include/PSkel.h; paths relative to the PSkel tree).
//
// A PSkel program includes this header and builds as before:
//   #define PSKEL_CUDA          // GPU methods (run on the MI355X; PSKEL_GPU also works)
//   #define PSKEL_OMP           // OpenMP for runIterativeCPU (optional)
//   #include "PSkel.h"
// and links libpskel_amd.so.  Compiled with g++ the GPU methods run the stock
// kernels (PSkelKernels.h) and any kernel with a StencilTraits descriptor;
// compiled with hipcc any user stencilKernel also runs on the device.
// Unlike the reference, nothing here needs TBB (the reference includes
// <tbb/scalable_allocator.h> unconditionally, :80), GAlib or PAPI.
//-------------------------------------------------------------------------------
#ifndef PSKEL_H
#define PSKEL_H

#include "PSkelDefs.h"
#include "PSkelArray.h"
#include "PSkelArgs.h"
#include "PSkelMask.h"
#include "PSkelMap.h"
#include "PSkelStencil.h"
#include "PSkelStencilTiling.h"
#include "PSkelKernels.h"

#endif
