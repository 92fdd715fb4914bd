//-------------------------------------------------------------------------------
// PSkelArray.h -- Array / Array2D / Array3D handles (drop-in for the reference's
// include/PSkelArray.h and include/PSkelArray.hpp; paths relative to the PSkel tree).
//
// Same contract as the reference:
//   * shallow handles: copying an Array aliases its buffers; the user owns host
//     memory (hostAlloc / hostFree), there is no destructor
//     (include/PSkelArray.h:299-304);
//   * "virtual" dims (width, height, depth) over "real" allocated dims plus
//     offsets, for slices (include/PSkelArray.h:53-96);
//   * Array2D(h, w) = hostArray[(h+hoff)*realWidth + (w+woff)]
//     (include/PSkelArray.hpp:668-673); Array3D(h, w, d) =
//     hostArray[((h+hoff)*realWidth + (w+woff))*realDepth + (d+doff)]
//     (include/PSkelArray.hpp:627-645): d is the contiguous axis.
// On the device a handle points at a resident, row-pitched buffer: Array2D(h, w)
// = deviceArray[h*devicePitch + w], Array3D(h, w, d) =
// deviceArray[(h*width + w)*devicePitch + d].  The library owns device buffers
// (they live in a pskel_grid_t for the duration of a GPU run, or across runs when
// a Stencil keeps its grid resident).
//-------------------------------------------------------------------------------
#ifndef PSKEL_ARRAY_H
#define PSKEL_ARRAY_H

#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "PSkelDefs.h"

namespace PSkel {

template <typename T>
class ArrayBase {
public:
    T* deviceArray;  // resident device buffer (pitched rows), NULL when not resident
    T* hostArray;

protected:
    size_t realWidth, realHeight, realDepth;
    size_t widthOffset, heightOffset, depthOffset;
    size_t width, height, depth;
    size_t devicePitch;  // elements per contiguous device row

    ArrayBase(size_t w, size_t h, size_t d)
        : deviceArray(nullptr), hostArray(nullptr), realWidth(w), realHeight(h), realDepth(d),
          widthOffset(0), heightOffset(0), depthOffset(0), width(w), height(h), depth(d), devicePitch(0)
    {
        if (size() > 0) hostAlloc();
    }

    __forceinline__ T& hostGet(size_t h, size_t w, size_t d) const
    {
        return hostArray[((h + heightOffset) * realWidth + (w + widthOffset)) * realDepth + (d + depthOffset)];
    }

public:
    template <typename U>
    friend class ArrayBase;

    // ---- host memory (include/PSkelArray.hpp:97-260) ----------------------
    void hostAlloc(size_t w, size_t h, size_t d)
    {
        width = realWidth = w;
        height = realHeight = h;
        depth = realDepth = d;
        widthOffset = heightOffset = depthOffset = 0;
        hostArray = nullptr;
        deviceArray = nullptr;
        hostAlloc();
    }

    void hostAlloc()
    {
        if (hostArray == nullptr && size() > 0) {
            // 64-byte aligned like a cache line (the reference uses _mm_malloc(.., 16))
            const size_t bytes = (size() * sizeof(T) + 63) / 64 * 64;
            hostArray = static_cast<T*>(std::aligned_alloc(64, bytes));
            if (!hostArray) throw Error(2, "host allocation failed");
        }
    }

    void hostFree()
    {
        if (hostArray) std::free(hostArray);
        hostArray = nullptr;
    }

    // ---- dims (include/PSkelArray.hpp:262-340) -----------------------------
    __host__ __device__ size_t getWidth() const { return width; }
    __host__ __device__ size_t getHeight() const { return height; }
    __host__ __device__ size_t getDepth() const { return depth; }
    __host__ __device__ size_t getRealHeight() const { return realHeight; }
    __host__ __device__ size_t getWidthOffset() const { return widthOffset; }
    __host__ __device__ size_t getHeightOffset() const { return heightOffset; }
    __host__ __device__ size_t getDepthOffset() const { return depthOffset; }
    __host__ __device__ size_t memSize() const { return size() * sizeof(T); }
    __host__ __device__ size_t size() const { return height * width * depth; }
    __host__ __device__ size_t typeSize() const { return sizeof(T); }
    __host__ __device__ size_t realSize() const { return realHeight * realWidth * realDepth; }
    __host__ __device__ size_t getDevicePitch() const { return devicePitch; }

    // ---- slices and copies (include/PSkelArray.hpp:354-433) -----------------
    template <typename Arrays>
    void hostSlice(Arrays array, size_t wOff, size_t hOff, size_t dOff, size_t w, size_t h, size_t d)
    {
        width = w;
        height = h;
        depth = d;
        widthOffset = array.widthOffset + wOff;
        heightOffset = array.heightOffset + hOff;
        depthOffset = array.depthOffset + dOff;
        realWidth = array.realWidth;
        realHeight = array.realHeight;
        realDepth = array.realDepth;
        hostArray = array.hostArray;
        deviceArray = nullptr;
    }

    template <typename Arrays>
    void hostClone(Arrays array)
    {
        width = realWidth = array.width;
        height = realHeight = array.height;
        depth = realDepth = array.depth;
        widthOffset = heightOffset = depthOffset = 0;
        hostArray = nullptr;
        deviceArray = nullptr;
        hostAlloc();
        hostMemCopy(array);
    }

    template <typename Arrays>
    void hostMemCopy(Arrays array)
    {
        if (array.size() == array.realSize() && size() == realSize()) {
            std::memcpy(hostArray, array.hostArray, size() * sizeof(T));
            return;
        }
        for (size_t i = 0; i < height; ++i)
            for (size_t j = 0; j < width; ++j)
                for (size_t k = 0; k < depth; ++k) hostGet(i, j, k) = array.hostGet(i, j, k);
    }

    // The host array when this handle covers all of it in the packed order (no
    // slice), else NULL: the GPU runners move it without packing.
    T* denseHost() const { return size() == realSize() ? hostArray : nullptr; }

    // Pack the virtual array (a slice or not) into a dense buffer and back.
    void packTo(T* dst) const
    {
        if (size() == realSize()) {
            std::memcpy(dst, hostArray, size() * sizeof(T));
            return;
        }
        size_t n = 0;
        for (size_t i = 0; i < height; ++i)
            for (size_t j = 0; j < width; ++j)
                for (size_t k = 0; k < depth; ++k) dst[n++] = hostGet(i, j, k);
    }

    void unpackFrom(const T* src)
    {
        if (size() == realSize()) {
            std::memcpy(hostArray, src, size() * sizeof(T));
            return;
        }
        size_t n = 0;
        for (size_t i = 0; i < height; ++i)
            for (size_t j = 0; j < width; ++j)
                for (size_t k = 0; k < depth; ++k) hostGet(i, j, k) = src[n++];
    }

    // Point this handle at a resident device buffer (used by the GPU runners).
    void setDeviceView(T* ptr, size_t pitch)
    {
        deviceArray = ptr;
        devicePitch = pitch;
    }

    __host__ __device__ explicit operator bool() const
    {
#ifdef PSKEL_ON_DEVICE
        return deviceArray != nullptr;
#else
        return hostArray != nullptr;
#endif
    }
};

template <typename T>
class Array3D : public ArrayBase<T> {
public:
    static constexpr int kDims = 3;
    Array3D() : ArrayBase<T>(0, 0, 0) {}
    Array3D(size_t width, size_t height, size_t depth) : ArrayBase<T>(width, height, depth) {}

    __host__ __device__ __forceinline__ T& operator()(size_t h, size_t w, size_t d) const
    {
#ifdef PSKEL_ON_DEVICE
        return this->deviceArray[(h * this->width + w) * this->devicePitch + d];
#else
        return this->hostGet(h, w, d);
#endif
    }
};

template <typename T>
class Array2D : public ArrayBase<T> {
public:
    static constexpr int kDims = 2;
    Array2D() : ArrayBase<T>(0, 0, 0) {}
    Array2D(size_t width, size_t height) : ArrayBase<T>(width, height, 1) {}

    __host__ __device__ __forceinline__ T& operator()(size_t h, size_t w) const
    {
#ifdef PSKEL_ON_DEVICE
        return this->deviceArray[h * this->devicePitch + w];
#else
        return this->hostArray[(h + this->heightOffset) * this->realWidth + (w + this->widthOffset)];
#endif
    }
};

template <typename T>
class Array : public ArrayBase<T> {
public:
    static constexpr int kDims = 1;
    Array() : ArrayBase<T>(0, 0, 0) {}
    explicit Array(size_t size) : ArrayBase<T>(size, 1, 1) {}

    __host__ __device__ __forceinline__ T& operator()(size_t w) const
    {
#ifdef PSKEL_ON_DEVICE
        return this->deviceArray[w];
#else
        return this->hostGet(0, w, 0);
#endif
    }
};

}  // namespace PSkel

#endif
