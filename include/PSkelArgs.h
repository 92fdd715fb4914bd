//-------------------------------------------------------------------------------
// PSkelArgs.h -- field arguments Args / Args2D / Args3D (drop-in for the
// reference's include/PSkelArgs.h and include/PSkelArgs.hpp; paths relative
// to the PSkel tree).
//
// Same API as the reference: public hostArray / deviceArray and extents,
// getWidth/getHeight/getDepth, and operator() with the host layout
// x + width*y and x + width*(y + height*z) (include/PSkelArgs.hpp:116-123,
// 155-160).  A stencil kernel reads a field with the cell's coordinates, the
// way the reference's kernels do: args.f(h, w) or args.f(h, w, d)
// (include/PSkelStencil.hpp:366-369).
//
// Different on purpose: the reference keeps fields in zero-copy mapped host
// memory, and Args(int) resets the device (include/PSkelArgs.hpp:49-57).  Here
// a float field of a GPU build also has a device copy in HBM, with the axes
// reversed (element (x, y) at x*pitch + y), so a row of cells reads a
// contiguous row of the field.  Every GPU run method uploads the host arrays
// of all live float fields before its sweeps, once per call, and the sweeps
// read HBM.  Fields are inputs: device writes to them are not copied back.
// Like the reference, fields are never freed (the reference declares no
// destructor); hostFree() releases the host array and the device copy.
//-------------------------------------------------------------------------------
#ifndef PSKEL_ARGS_H
#define PSKEL_ARGS_H

#include <cstddef>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "PSkelDefs.h"

namespace PSkel {

namespace detail {

// One live field: its host array and, in GPU builds, its device copy.
struct FieldState {
    int ndim = 0;
    size_t width = 0, height = 1, depth = 1;
    const float* host = nullptr;
#ifdef PSKEL_GPU
    pskel_field_t* field = nullptr;
#endif
};

inline std::vector<FieldState*>& fieldRegistry()
{
    static std::vector<FieldState*> reg;
    return reg;
}

inline std::mutex& fieldMutex()
{
    static std::mutex m;
    return m;
}

#ifdef PSKEL_GPU
// Upload every live float field (called by the GPU run methods).
inline void syncFields()
{
    std::lock_guard<std::mutex> lock(fieldMutex());
    for (FieldState* f : fieldRegistry())
        if (f->field && f->host) gpuErrchk(pskel_field_upload(f->field, f->host));
}

// Create the device copy of a float field; returns its base pointer and pitch.
inline FieldState* newField(int ndim, size_t w, size_t h, size_t d, const float* host, void** dev,
                            size_t* pitch)
{
    auto* st = new FieldState();
    st->ndim = ndim;
    st->width = w;
    st->height = h;
    st->depth = d;
    st->host = host;
    pskel_field_desc_t fd{};
    fd.ndim = ndim;
    fd.dtype = PSKEL_DTYPE_F32;
    fd.width = w;
    fd.height = h;
    fd.depth = d;
    fd.device = 0;
    gpuErrchk(pskel_field_create(&fd, &st->field));
    gpuErrchk(pskel_field_info(st->field, dev, pitch));
    std::lock_guard<std::mutex> lock(fieldMutex());
    fieldRegistry().push_back(st);
    return st;
}

inline void dropField(FieldState* st)
{
    if (!st) return;
    {
        std::lock_guard<std::mutex> lock(fieldMutex());
        auto& reg = fieldRegistry();
        for (size_t i = 0; i < reg.size(); ++i)
            if (reg[i] == st) {
                reg.erase(reg.begin() + (long)i);
                break;
            }
    }
    if (st->field) pskel_field_destroy(st->field);
    delete st;
}
#endif

}  // namespace detail

template <typename T>
class Args {
public:
    T* hostArray = nullptr;
    T* deviceArray = nullptr;
    int width = 0;
    detail::FieldState* state = nullptr;

    Args() {}
    explicit Args(int _width) : width(_width)
    {
        hostArray = new T[(size_t)width]();
#ifdef PSKEL_GPU
        if constexpr (std::is_same_v<T, float>) {
            void* dev = nullptr;
            size_t pitch = 0;
            state = detail::newField(1, (size_t)width, 1, 1, hostArray, &dev, &pitch);
            deviceArray = (T*)dev;
        }
#endif
    }
    __parallel__ int getWidth() const { return width; }
    __parallel__ T& operator()(int x) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return deviceArray[x];
#else
        return hostArray[x];
#endif
    }
    void hostFree()
    {
#ifdef PSKEL_GPU
        detail::dropField(state);
#endif
        state = nullptr;
        delete[] hostArray;
        hostArray = nullptr;
        deviceArray = nullptr;
    }
#ifdef PSKEL_GPU
    pskel_field_t* deviceField() const { return state ? state->field : nullptr; }
#endif
};

template <typename T>
class Args2D {
public:
    T* hostArray = nullptr;
    T* deviceArray = nullptr;
    int width = 0, height = 0;
    int devicePitch = 0;  // elements; device element (x, y) at x*devicePitch + y
    detail::FieldState* state = nullptr;

    Args2D() {}
    Args2D(int _width, int _height) : width(_width), height(_height)
    {
        hostArray = new T[(size_t)width * height]();
#ifdef PSKEL_GPU
        if constexpr (std::is_same_v<T, float>) {
            void* dev = nullptr;
            size_t pitch = 0;
            state = detail::newField(2, (size_t)width, (size_t)height, 1, hostArray, &dev, &pitch);
            deviceArray = (T*)dev;
            devicePitch = (int)pitch;
        }
#endif
    }
    __parallel__ int getWidth() const { return width; }
    __parallel__ int getHeight() const { return height; }
    __parallel__ T& operator()(int x, int y) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return deviceArray[(size_t)x * devicePitch + y];
#else
        return hostArray[(size_t)y * width + x];
#endif
    }
    void hostFree()
    {
#ifdef PSKEL_GPU
        detail::dropField(state);
#endif
        state = nullptr;
        delete[] hostArray;
        hostArray = nullptr;
        deviceArray = nullptr;
    }
#ifdef PSKEL_GPU
    pskel_field_t* deviceField() const { return state ? state->field : nullptr; }
#endif
};

template <typename T>
class Args3D {
public:
    T* hostArray = nullptr;
    T* deviceArray = nullptr;
    int width = 0, height = 0, depth = 0;
    int devicePitch = 0;  // elements; device element (x, y, z) at (x*height + y)*devicePitch + z
    detail::FieldState* state = nullptr;

    Args3D() {}
    Args3D(int _width, int _height, int _depth) : width(_width), height(_height), depth(_depth)
    {
        hostArray = new T[(size_t)width * height * depth]();
#ifdef PSKEL_GPU
        if constexpr (std::is_same_v<T, float>) {
            void* dev = nullptr;
            size_t pitch = 0;
            state = detail::newField(3, (size_t)width, (size_t)height, (size_t)depth, hostArray, &dev, &pitch);
            deviceArray = (T*)dev;
            devicePitch = (int)pitch;
        }
#endif
    }
    __parallel__ int getWidth() const { return width; }
    __parallel__ int getHeight() const { return height; }
    __parallel__ int getDepth() const { return depth; }
    __parallel__ T& operator()(int x, int y, int z) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return deviceArray[((size_t)x * height + y) * devicePitch + z];
#else
        return hostArray[((size_t)z * height + y) * width + x];
#endif
    }
    void hostFree()
    {
#ifdef PSKEL_GPU
        detail::dropField(state);
#endif
        state = nullptr;
        delete[] hostArray;
        hostArray = nullptr;
        deviceArray = nullptr;
    }
#ifdef PSKEL_GPU
    pskel_field_t* deviceField() const { return state ? state->field : nullptr; }
#endif
};

}  // namespace PSkel

#endif
