//-------------------------------------------------------------------------------
// PSkelMask.h -- Mask / Mask2D / Mask3D (drop-in for the reference's
// include/PSkelMask.h and include/PSkelMask.hpp; paths relative to the PSkel tree).
//
// A mask is `size` neighbour offsets of `dimension` ints plus one weight per
// neighbour (include/PSkelMask.h:47-121).  set(n, h, w[, d], weight) stores
// neighbour n (include/PSkelMask.hpp:170-180, 266-277); get(n, array, h, w[, d])
// reads it: Mask2D::get has no bounds check (:292-305), Mask3D::get and Mask::get
// return the halo value outside the array (:186-199, :383-395).  getRange() is
// the Chebyshev radius of the offsets unless a range was given (:202-236,
// :308-334).
//
// On the device the offsets and weights are read from a resident copy made
// by deviceAlloc + copyToDevice (the GPU runners do this).
//-------------------------------------------------------------------------------
#ifndef PSKEL_MASK_H
#define PSKEL_MASK_H

#include <cstdlib>
#include <cstring>

#include "PSkelArray.h"
#include "PSkelDefs.h"

namespace PSkel {

template <typename T>
class MaskBase {
protected:
    int* hostMask;
    int* deviceMask;
    T* hostWeight;
    T* deviceWeight;
    T haloValue;

    MaskBase(size_t size_ = 0, size_t dimension_ = 0, T haloVal = T(0), size_t range_ = 0)
        : hostMask(nullptr), deviceMask(nullptr), hostWeight(nullptr), deviceWeight(nullptr),
          haloValue(haloVal), size(size_), dimension(dimension_), range(range_)
    {
        if (size > 0) hostAlloc();
    }

public:
    size_t size, dimension;
    size_t range;

    void hostAlloc()
    {
        if (hostMask == nullptr && hostWeight == nullptr) {
            hostMask = static_cast<int*>(std::calloc(dimension * size, sizeof(int)));
            hostWeight = static_cast<T*>(std::calloc(size, sizeof(T)));
        }
    }

    void hostFree()
    {
        std::free(hostMask);
        std::free(hostWeight);
        hostMask = nullptr;
        hostWeight = nullptr;
    }

#ifdef PSKEL_GPU
    // include/PSkelMask.hpp:63-94
    void deviceAlloc()
    {
        if (!deviceMask) gpuErrchk(pskel_device_alloc(dimension * size * sizeof(int), (void**)&deviceMask));
        if (!deviceWeight) gpuErrchk(pskel_device_alloc(size * sizeof(T), (void**)&deviceWeight));
    }

    void copyToDevice()
    {
        gpuErrchk(pskel_memcpy(deviceMask, hostMask, dimension * size * sizeof(int), 1));
        gpuErrchk(pskel_memcpy(deviceWeight, hostWeight, size * sizeof(T), 1));
    }

    void deviceFree()
    {
        if (deviceMask) pskel_device_free(deviceMask);
        if (deviceWeight) pskel_device_free(deviceWeight);
        deviceMask = nullptr;
        deviceWeight = nullptr;
    }
#endif

    size_t memSize() const { return dimension * size * sizeof(int) + size * sizeof(T); }

    __host__ __device__ T& getWeight(size_t n) const
    {
#ifdef PSKEL_ON_DEVICE
        return deviceWeight[n];
#else
        return hostWeight[n];
#endif
    }

    __host__ __device__ T getHaloValue() const { return haloValue; }

    // Offset k (0 .. dimension-1) of neighbour n, host copy.
    int hostOffset(size_t n, size_t k) const { return hostMask[dimension * n + k]; }

protected:
    __host__ __device__ const int* offsets() const
    {
#ifdef PSKEL_ON_DEVICE
        return deviceMask;
#else
        return hostMask;
#endif
    }

    size_t computeRange() const
    {
        if (range > 0) return range;
        size_t r = 0;
        for (size_t i = 0; i < size * dimension; ++i) {
            const size_t v = (size_t)std::abs(hostMask[i]);
            if (v > r) r = v;
        }
        return r;
    }
};

template <typename T>
class Mask3D : public MaskBase<T> {
public:
    Mask3D(size_t size = 0, T haloVal = T(0), size_t range = 0) : MaskBase<T>(size, 3, haloVal, range) {}

    void set(size_t n, int h, int w, int d, T weight = T(0))
    {
        this->hostMask[3 * n] = h;
        this->hostMask[3 * n + 1] = w;
        this->hostMask[3 * n + 2] = d;
        this->hostWeight[n] = weight;
    }

    template <typename V>
    __host__ __device__ __forceinline__ T get(size_t n, const Array3D<V>& array, size_t h, size_t w, size_t d) const
    {
        const int* m = this->offsets();
        h += m[3 * n];
        w += m[3 * n + 1];
        d += m[3 * n + 2];
        return (w < array.getWidth() && h < array.getHeight() && d < array.getDepth()) ? (T)array(h, w, d)
                                                                                         : this->haloValue;
    }

    size_t getRange()
    {
        this->range = this->computeRange();
        return this->range;
    }
};

template <typename T>
class Mask2D : public MaskBase<T> {
public:
    Mask2D(size_t size = 0, T haloVal = T(0), size_t range = 0) : MaskBase<T>(size, 2, haloVal, range) {}

    Mask2D(size_t size, int array[][2]) : MaskBase<T>(size, 2, T(0), 0)
    {
        for (size_t i = 0; i < size; ++i) set(i, array[i][0], array[i][1], T(1));
    }

    void set(size_t n, int h, int w, T weight = T(0))
    {
        this->hostMask[2 * n] = h;
        this->hostMask[2 * n + 1] = w;
        this->hostWeight[n] = weight;
    }

    // no bounds check, as in the reference
    template <typename V>
    __host__ __device__ __forceinline__ V& get(size_t n, const Array2D<V>& array, size_t h, size_t w) const
    {
        const int* m = this->offsets();
        return array(h + m[2 * n], w + m[2 * n + 1]);
    }

    size_t getRange()
    {
        this->range = this->computeRange();
        return this->range;
    }
};

template <typename T>
class Mask : public MaskBase<T> {
public:
    Mask(size_t size = 0, T haloVal = T(0), size_t range = 0) : MaskBase<T>(size, 1, haloVal, range) {}

    void set(size_t n, int i, T weight = T(0))
    {
        this->hostMask[n] = i;
        this->hostWeight[n] = weight;
    }

    template <typename V>
    __host__ __device__ __forceinline__ T get(size_t n, const Array<V>& array, size_t i) const
    {
        i += this->offsets()[n];
        return (i < array.getWidth()) ? (T)array(i) : this->haloValue;
    }

    size_t getRange()
    {
        this->range = this->computeRange();
        return this->range;
    }
};

}  // namespace PSkel

#endif
