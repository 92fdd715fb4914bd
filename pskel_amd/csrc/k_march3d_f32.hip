// Marching 3D kernels with f32 storage (the template is in k_march3d.hpp).
#define PSKEL_MARCH_INSTANTIATE
#include "k_march3d.hpp"
#include "k_march3d_t2.hpp"

namespace pskel {

hipError_t march3d_f32(const Sweep3D& s, const Points& p, const MarchShape& m, hipStream_t st)
{
    return march_vd<PSKEL_DTYPE_F32>(s, p, m, st);
}

hipError_t march3d_t2_f32(const Sweep3D& s, const Points& p, const March2Shape& m, hipStream_t st)
{
    return march2_dispatch<PSKEL_DTYPE_F32>(s, p, m, st);
}

size_t march3d_t2_f32_blocks(const Sweep3D& s, const March2Shape& m) { return march2_blocks<PSKEL_DTYPE_F32>(s, m); }

}  // namespace pskel
