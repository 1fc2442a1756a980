// Wavefront/workgroup max-reductions for non-negative doubles.
//
// A max over non-negative doubles is exact and independent of order, so the
// device result equals a sequential one bit for bit.  Non-negative IEEE
// doubles order like their bit patterns read as unsigned integers, which
// lets one 64-bit atomicMax per workgroup combine partial maxima.
//
// Every max here carries a NaN forward (vmax), where fmax would drop it: a
// norm of a field holding a NaN reads as NaN, and the host reports it
// instead of taking the NaN residual for a converged one.  For finite
// operands vmax returns what fmax returns.
#ifndef VD_REDUCE_H_
#define VD_REDUCE_H_
#include <hip/hip_runtime.h>

namespace vd {

// max(a, b) that returns a NaN operand (either one) instead of dropping it
__host__ __device__ __forceinline__ double vmax(double a, double b) { return (a > b || a != a) ? a : b; }

__device__ inline double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = vmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Workgroup max of v; lane 0 of the first wave ends up with the result and
// performs the atomic.  scratch: LDS of >= 16 doubles.  Every thread of the
// workgroup must call this.
__device__ inline void block_max_atomic(double v, unsigned long long* dst, double* scratch)
{
    v = wave_max(v);
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6;
    const int nw = (nthreads + 63) >> 6;
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    if (wid == 0)
    {
        double w = (lane < nw) ? scratch[lane] : 0.0;
        w = wave_max(w);
        if (lane == 0 && (w > 0.0 || w != w)) atomicMax(dst, (unsigned long long)__double_as_longlong(w));
    }
    __syncthreads();
}

// Workgroup max of v stored (no atomics) into *slot by one thread: one
// partial per workgroup, combined by finalize_max_kernel.  Every thread of
// the workgroup must call this.  Values may be negative (-1e308 is the
// neutral element).
__device__ inline void block_max_store(double v, double* slot, double* scratch)
{
    v = wave_max(v);
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6;
    const int nw = (nthreads + 63) >> 6;
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    if (wid == 0)
    {
        double w = (lane < nw) ? scratch[lane] : -1.0e308;
        w = wave_max(w);
        if (lane == 0) *slot = w;
    }
    __syncthreads();
}

__device__ inline long long flat_block_id()
{
    return (long long)blockIdx.x + (long long)gridDim.x * ((long long)blockIdx.y + (long long)gridDim.y * blockIdx.z);
}

// Wave max of v stored by lane 0 of each wave into slots[flat block id *
// waves_per_block + wave]: one partial per wave and no workgroup barrier,
// for kernels that reduce many values; finalize_max_kernel combines
// gridDim * waves_per_block slots.  Every thread of the wave must call this.
constexpr int kCellWaves = 4;   // waves of a 256-thread cell workgroup
__device__ inline void wave_max_store(double v, double* slots)
{
    v = wave_max(v);
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    if ((tid & 63) == 0) slots[flat_block_id() * kCellWaves + (tid >> 6)] = v;
}

// out[c] = bits of max over partials[c*nblocks .. +nblocks), one workgroup
// per component (gridDim.x = number of components).
__device__ __forceinline__ void finalize_max_block(const double* __restrict__ p, long long nblocks,
                                                   unsigned long long* __restrict__ outc)
{
    __shared__ double scratch[16];
    double m0 = -1.0e308, m1 = -1.0e308, m2 = -1.0e308, m3 = -1.0e308;
    const long long nt = blockDim.x;
    long long q = threadIdx.x;
    for (; q + 3 * nt < nblocks; q += 4 * nt)
    {
        m0 = vmax(m0, p[q]);
        m1 = vmax(m1, p[q + nt]);
        m2 = vmax(m2, p[q + 2 * nt]);
        m3 = vmax(m3, p[q + 3 * nt]);
    }
    for (; q < nblocks; q += nt) m0 = vmax(m0, p[q]);
    double m = vmax(vmax(m0, m1), vmax(m2, m3));
    m = wave_max(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) scratch[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double w = -1.0e308;
        for (int q = 0; q < (int)((blockDim.x + 63) >> 6); ++q) w = vmax(w, scratch[q]);
        *outc = (unsigned long long)__double_as_longlong(w);
    }
}

__global__ void __launch_bounds__(1024) finalize_max_kernel(const double* __restrict__ partials, long long nblocks,
                                                            unsigned long long* __restrict__ out)
{
    const int c = blockIdx.x;
    finalize_max_block(partials + (long long)c * nblocks, nblocks, out + c);
}

// First stage of a split maximum over many partials: block (s, r) takes
// slice s of kMaxSplit of row r (rows [0, n) of pa, then the 2n rows of pb)
// into q[r * kMaxSplit + s]; finalize_max3_kernel then reduces those rows.
// A maximum does not depend on the order, so the result is the same.
constexpr int kMaxSplit = 64;
__global__ void __launch_bounds__(256) max_split3_kernel(const double* __restrict__ pa, const double* __restrict__ pb,
                                                         long long nblocks, int n, double* __restrict__ q)
{
    const int s = blockIdx.x, r = blockIdx.y;
    const double* p = (r < n) ? pa + (long long)r * nblocks : pb + (long long)(r - n) * nblocks;
    const long long per = (nblocks + kMaxSplit - 1) / kMaxSplit;
    const long long b0 = (long long)s * per, b1 = (b0 + per < nblocks) ? b0 + per : nblocks;
    __shared__ double scratch[16];
    double m = -1.0e308;
    for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) m = vmax(m, p[i]);
    m = wave_max(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) scratch[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double w = -1.0e308;
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); ++k) w = vmax(w, scratch[k]);
        q[(long long)r * kMaxSplit + s] = w;
    }
}

// Three finalize_max_kernel launches in one: blocks [0, n) reduce pa into
// oa, blocks [n, 3n) reduce the 2n partial rows of pb into ob and ob2.
__global__ void __launch_bounds__(1024) finalize_max3_kernel(const double* __restrict__ pa,
                                                             const double* __restrict__ pb, long long nblocks, int n,
                                                             unsigned long long* __restrict__ oa,
                                                             unsigned long long* __restrict__ ob,
                                                             unsigned long long* __restrict__ ob2)
{
    const int c = blockIdx.x;
    if (c < n) finalize_max_block(pa + (long long)c * nblocks, nblocks, oa + c);
    else if (c < 2 * n) finalize_max_block(pb + (long long)(c - n) * nblocks, nblocks, ob + (c - n));
    else finalize_max_block(pb + (long long)(c - n) * nblocks, nblocks, ob2 + (c - 2 * n));
}

__host__ __device__ inline double bits_to_double(unsigned long long b)
{
    union
    {
        unsigned long long u;
        double d;
    } x;
    x.u = b;
    return x.d;
}

// ---------------------------------------------------------------------------
// Sums.  Floating-point sums depend on order, so every sum here has a fixed
// order that does not depend on scheduling: a fixed butterfly inside the
// wavefront, waves in index order, workgroups in index order.  Repeated runs
// give identical bits.

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Workgroup sum of v stored into *slot (one partial per workgroup).  Every
// thread of the workgroup must call this.
__device__ inline void block_sum_store(double v, double* slot, double* scratch)
{
    v = wave_sum(v);
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6;
    const int nw = (nthreads + 63) >> 6;
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    if (tid == 0)
    {
        double w = 0.0;
        for (int q = 0; q < nw; ++q) w += scratch[q];
        *slot = w;
    }
    __syncthreads();
}

// out[c] (+)= sum over partials[c*nblocks .. +nblocks), one workgroup per
// component; accumulate != 0 adds to out instead of overwriting it.
__global__ void __launch_bounds__(1024) finalize_sum_kernel(const double* __restrict__ partials, long long nblocks,
                                                            double* __restrict__ out, int accumulate)
{
    __shared__ double scratch[16];
    const int c = blockIdx.x;
    const double* p = partials + (long long)c * nblocks;
    double m = 0.0;
    for (long long q = threadIdx.x; q < nblocks; q += blockDim.x) m += p[q];
    m = wave_sum(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) scratch[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double w = 0.0;
        for (int q = 0; q < (int)((blockDim.x + 63) >> 6); ++q) w += scratch[q];
        out[c] = accumulate ? out[c] + w : w;
    }
}

// Fill n 64-bit words with zero (a kernel, so it can sit in a captured graph).
__global__ void zero_u64_kernel(unsigned long long* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ULL;
}

} // namespace vd
#endif
