// march_gen.hip -- the general kind's (K_GEN) kernel instantiations: every
// built-in formula, the 4D and interpolation hybrids, Julia mode and the DE
// combinations, FP64 only (march_kernels.h).  A translation unit of its own
// so that it compiles beside march.hip.
#include <hip/hip_runtime.h>

#include "march_kernels.h"
#include "mc_kernel.h"

namespace mb3d {

template <class F>
static void with_dem(int dem, F &&f) {
    if (dem == DEM_BOX)
        f(std::integral_constant<int, DEM_BOX>());
    else
        f(std::integral_constant<int, DEM_NUMERIC>());
}

void launch_gen_finish(int grid, hipStream_t s, const DevP &d, const Bufs &b, int all, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_finish64<K_GEN, decltype(D)::value, false, false>), dim3(grid), dim3(256), 0, s, d, b,
                           all, nullptr, 1);
    });
}

void launch_gen_normals(int grid, hipStream_t s, const DevP &d, const Bufs &b, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_normals<K_GEN, decltype(D)::value, false, false>), dim3(grid), dim3(256), 0, s, d, b,
                           nullptr, 1);
    });
}

void launch_gen_normals_ext(int grid, hipStream_t s, const DevP &d, const Bufs &b, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_normals_ext<K_GEN, decltype(D)::value, false>), dim3(grid), dim3(256), 0, s, d, b,
                           nullptr, 1);
    });
}

void launch_gen_shadow(int grid, hipStream_t s, const DevP &d, const Bufs &b, const ShadowJob &j, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_shadow<K_GEN, decltype(D)::value, false>), dim3(grid), dim3(256), 0, s, d, b, j);
    });
}

void launch_gen_deao(int grid, hipStream_t s, const DevP &d, const Bufs &b, const AoRays &r, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_deao<K_GEN, decltype(D)::value, false>), dim3(grid), dim3(256), 0, s, d, b, r);
    });
}

void launch_gen_de_at(int grid, hipStream_t s, const DevP &d, const double *pts, long long n, double *de,
                      int32_t *its) {
    hipLaunchKernelGGL((k_de_at<K_GEN>), dim3(grid), dim3(256), 0, s, d, pts, n, de, its);
}

void launch_gen_rays(int grid, hipStream_t s, const DevP &d, const mb3d_hip_ray_task *tasks, long long n,
                     int depth_sqrt, mb3d_hip_ray_hit *hits, unsigned long long *ctr, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_rays<K_GEN, decltype(D)::value>), dim3(grid), dim3(256), 0, s, d, tasks, n, depth_sqrt,
                           hits, ctr);
    });
}

void launch_gen_ray_post(int grid, hipStream_t s, const DevP &d, const mb3d_hip_ray_task *tasks, long long n,
                         const RayPost &j, mb3d_hip_ray_hit *hits, unsigned long long *ctr, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_ray_post<K_GEN, decltype(D)::value>), dim3(grid), dim3(256), 0, s, d, tasks, n, j, hits,
                           ctr);
    });
}

void launch_gen_mc(int grid, hipStream_t s, const DevP &d, const mb3d_hip_mc_task *tasks, long long n, int secant,
                   mb3d_hip_mc_res *res, unsigned long long *ctr, int dem) {
    with_dem(dem, [&](auto D) {
        hipLaunchKernelGGL((k_mc<K_GEN, decltype(D)::value>), dim3(grid), dim3(256), 0, s, d, tasks, n, secant, res,
                           ctr);
    });
}

}  // namespace mb3d
