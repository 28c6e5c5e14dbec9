// decompose.hip -- domain decomposition on the GPU.
//
// Replaces the reference's offline tool (src/domain_decomposition.f90): for
// every region label r, the bounding box of the points labelled r (:38-45),
// the box grown by rmax on every side (:48-53), membership of every point that
// lies strictly inside the grown box (:62-64) and the buffer flag, 0 for a
// point inside the closed box and 1 otherwise (:67-70).  The tool reads its
// coordinates into default reals, so boxes and tests are single precision.
//
// Membership lists keep the input order (the tool writes points in input
// order): per region, one flag pass, an exclusive scan of the flags and one
// scatter pass.
#include "gs_kernels.h"

namespace gs {

// float -> int key whose signed order is the float order (finite values)
__device__ __forceinline__ int fkey(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}

__global__ __launch_bounds__(256) void k_region_box(const float* __restrict__ xyz, const int32_t* __restrict__ reg,
                                                    int64_t n, int nreg, int32_t* __restrict__ keys /*[nreg][6]*/)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = reg[i];
        if (r < 1 || r > nreg) continue;
        int32_t* k = keys + (size_t)(r - 1) * 6;
        for (int c = 0; c < 3; ++c) {
            const int v = fkey(xyz[3 * i + c]);
            atomicMin(&k[c], v);
            atomicMax(&k[3 + c], v);
        }
    }
}

// ext: xmin, ymin, zmin, xmax, ymax, zmax of the box grown by rmax
__global__ __launch_bounds__(256) void k_region_flag(const float* __restrict__ xyz, int64_t n,
                                                     const float* __restrict__ ext, int32_t* __restrict__ flag)
{
    const float e0 = ext[0], e1 = ext[1], e2 = ext[2], e3 = ext[3], e4 = ext[4], e5 = ext[5];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        flag[i] = (x > e0 && x < e3 && y > e1 && y < e4 && z > e2 && z < e5) ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void k_region_scatter(const float* __restrict__ xyz, int64_t n,
                                                        const float* __restrict__ box, const int32_t* __restrict__ flag,
                                                        const int32_t* __restrict__ pos, int32_t* __restrict__ idx,
                                                        int32_t* __restrict__ buf)
{
    const float b0 = box[0], b1 = box[1], b2 = box[2], b3 = box[3], b4 = box[4], b5 = box[5];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const int p = pos[i];
        idx[p] = (int32_t)i;
        buf[p] = (x >= b0 && x <= b3 && y >= b1 && y <= b4 && z >= b2 && z <= b5) ? 0 : 1;
    }
}

static int blocks_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 65535) b = 65535;
    return (int)(b < 1 ? 1 : b);
}

hipError_t launch_region_box(const float* xyz, const int32_t* reg, int64_t n, int nreg, int32_t* keys, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_region_box, dim3(blocks_for(n)), dim3(256), 0, st, xyz, reg, n, nreg, keys);
    return hipGetLastError();
}

hipError_t launch_region_flag(const float* xyz, int64_t n, const float* ext, int32_t* flag, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_region_flag, dim3(blocks_for(n)), dim3(256), 0, st, xyz, n, ext, flag);
    return hipGetLastError();
}

hipError_t launch_region_scatter(const float* xyz, int64_t n, const float* box, const int32_t* flag, const int32_t* pos,
                                 int32_t* idx, int32_t* buf, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_region_scatter, dim3(blocks_for(n)), dim3(256), 0, st, xyz, n, box, flag, pos, idx, buf);
    return hipGetLastError();
}

}  // namespace gs
