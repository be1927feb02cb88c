// Compute_CFL for callers that cannot allocate device memory: code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// st_cfl reduces into partials that the context allocated.  This variant
// takes device pointers and a stream, needs no context, and keeps its
// per-block partial maxima in a file-scope device array (`scratch`) of a
// fixed size (256 KiB per device, held until the process ends).  Its
// device address is looked up once per device with hipGetSymbolAddress
// and kept; host code passes that address, never the array's name.  Every
// call on a device shares that device's array, so calls that overlap on
// one device race.  The maxima are order-independent, so the results are
// the same bits as st_cfl's on the same state.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/st_api.h"
#include "st_engine.hpp"
#include "st_internal.h"

namespace {

constexpr int NT = 256;                                  // threads per block
constexpr int PTS = 8;                                   // points per thread
constexpr long SCRATCH_DOUBLES = 32768;                  // 256 KiB
constexpr long MAX_BLOCKS = SCRATCH_DOUBLES / 2;         // two partials per block
constexpr long MAX_POINTS = MAX_BLOCKS * NT * PTS;       // 33,554,432 grid points

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double *scratch_addr[MAX_DEVICES] = {};
std::mutex scratch_mu;

// The current device's address of `scratch` (looked up on first use).
int scratch_of_device(double **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return st::set_call_error(-3, "st_cfl_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return st::set_call_error(-3, "st_cfl_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return st::set_call_error(-3, "st_cfl_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        scratch_addr[dev] = (double *)p;
    }
    *out = scratch_addr[dev];
    return 0;
}

// max over one block's values of two quantities (4 waves of 64 lanes)
__device__ __forceinline__ void block_max2(double &a, double &b)
{
    __shared__ double sh[2][NT / 64];
    for (int o = 32; o > 0; o >>= 1) {
        const double x = __shfl_xor(a, o), y = __shfl_xor(b, o);
        a = x > a ? x : a;
        b = y > b ? y : b;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wave] = a;
        sh[1][wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < NT / 64; w++) {
            a = sh[0][w] > a ? sh[0][w] : a;
            b = sh[1][w] > b ? sh[1][w] : b;
        }
}

// Block b covers points [b NT PTS, (b+1) NT PTS) of Q(n points, 5 fields);
// part[2b], part[2b+1] = its maxima of the two CFL numbers
// (src/lddrk4_periodic_sf.f:29-56).
__global__ void __launch_bounds__(NT) k_cfl_part(const double *Q, long n, st::Phys ph, double *part)
{
    double m1 = -INFINITY, m2 = -INFINITY;
    const long base = (long)blockIdx.x * NT * PTS;
#pragma unroll
    for (int k = 0; k < PTS; k++) {
        const long i = base + (long)k * NT + threadIdx.x;
        if (i < n) {
            double w1, w2;
            st::cfl_point(ph, Q[i], Q[n + i], Q[2 * n + i], Q[3 * n + i], Q[4 * n + i], w1, w2);
            m1 = w1 > m1 ? w1 : m1;
            m2 = w2 > m2 ? w2 : m2;
        }
    }
    block_max2(m1, m2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = m1;
        part[2 * blockIdx.x + 1] = m2;
    }
}

// One block: out[0..1] = max over nb blocks' partials.
__global__ void __launch_bounds__(NT) k_cfl_fold(const double *part, long nb, double *out)
{
    double m1 = -INFINITY, m2 = -INFINITY;
    for (long b = threadIdx.x; b < nb; b += NT) {
        m1 = part[2 * b] > m1 ? part[2 * b] : m1;
        m2 = part[2 * b + 1] > m2 ? part[2 * b + 1] : m2;
    }
    block_max2(m1, m2);
    if (threadIdx.x == 0) {
        out[0] = m1;
        out[1] = m2;
    }
}

}  // namespace

extern "C" int st_cfl_noalloc(const st_grid *grid, double gamma, const double *Q_dev, double *out_dev, void *stream)
{
    if (!grid || !Q_dev || !out_dev) return st::set_call_error(-1, "st_cfl_noalloc: null argument");
    if (grid->nx < 2 || grid->ny < 2 || grid->nz < 2 || !(gamma > 1.0))
        return st::set_call_error(-1, "st_cfl_noalloc: bad grid %d x %d x %d or gamma %g", grid->nx, grid->ny,
                                  grid->nz, gamma);
    const long n = (long)grid->nx * grid->ny * grid->nz;
    if (n > MAX_POINTS)
        return st::set_call_error(-2, "st_cfl_noalloc: %ld grid points exceed the maximum of %ld (scratch of %ld "
                                      "doubles)", n, MAX_POINTS, SCRATCH_DOUBLES);
    double *part = nullptr;
    const int rc = scratch_of_device(&part);
    if (rc) return rc;
    st::Phys ph = {};
    ph.gamma = gamma;
    ph.gm1 = gamma - 1.0;
    const int nn[3] = {grid->nx, grid->ny, grid->nz};
    const double per[3] = {grid->period_x, grid->period_y, grid->period_z};
    const double start[3] = {grid->x_start, grid->y_start, grid->z_start};
    for (int d = 0; d < 3; d++)
        ph.D[d] = st::grid_spacing(per[d], start[d], nn[d], d == 0 && grid->nonperiodic_x ? 1 : 0);
    const long nb = (n + (long)NT * PTS - 1) / ((long)NT * PTS);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cfl_part, dim3((unsigned)nb), dim3(NT), 0, s, Q_dev, n, ph, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cfl_fold, dim3(1), dim3(NT), 0, s, (const double *)part, nb, out_dev);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return st::set_call_error(-4, "st_cfl_noalloc: launch: %s", hipGetErrorString(e));
    return 0;
}
