/*
 * Scratch memory of the entry points that allocate nothing
 * (fjsph_surrogate_grad_noalloc).
 *
 * `scratch` is a file-scope device array: each GPU that loads the library
 * holds its own copy, at its own address, until the process ends.  In host
 * code the array's name is only a host-side stand-in, never a device
 * address, so the library takes each device's address once, with that
 * device current, from hipGetSymbolAddress (noalloc_init, called by
 * fjsph_create) and hands that pointer to kernels and copies.
 *
 * Every call on one GPU shares that GPU's array and the array keeps the
 * last call's contents: calls that overlap on one GPU race.
 */
#include "device_common.hpp"
#include "engine_internal.hpp"

#include <mutex>

namespace fj {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice; /* the caller makes `device` current */
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
    return hipSuccess;
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_ptr[device];
}

/* Samples into the training kernel's layout: inputs at stride 4, targets at
 * stride 8 (unused components zero), the batch as positions 0..B-1. */
__global__ __launch_bounds__(TPB) void k_mlp_stage(const float* __restrict__ xn, const float* __restrict__ tn,
                                                   int32_t batch, int32_t dim, float* __restrict__ X,
                                                   float* __restrict__ T, int32_t* __restrict__ idx)
{
    const int64_t j = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (j >= batch) return;
    const int no = dim + 2;
#pragma unroll
    for (int a = 0; a < 4; ++a) X[j * 4 + a] = a < dim ? xn[j * dim + a] : 0.f;
#pragma unroll
    for (int a = 0; a < 8; ++a) T[j * 8 + a] = a < no ? tn[j * no + a] : 0.f;
    idx[j] = int32_t(j);
}

/* Mean squared error over the batch and outputs from the summed squares. */
__global__ void k_mlp_loss(const float* __restrict__ sse, double denom, double* __restrict__ loss)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *loss = double(*sse) / denom;
}

void launch_mlp_stage(const float* xn, const float* tn, int32_t batch, int32_t dim, float* X, float* T, int32_t* idx,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_stage, dim3(grid_of(batch)), dim3(TPB), 0, st, xn, tn, batch, dim, X, T, idx);
}

void launch_mlp_loss(const float* sse, double denom, double* loss, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(64), 0, st, sse, denom, loss);
}

}  // namespace fj
